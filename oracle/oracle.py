"""ctypes binding of the CPU oracle (oracle/simox_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker or the timed CPU baseline.  The
product package (simox_amd) never imports this module.

All arrays use the oracle's conventions: species ids 1..5 (Al, Ni, Fe, O, H),
3-vectors atom-major (shape (n, 3)), 0-based atom indices.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")
REF_DRIVER = os.path.join(HERE, "_ref", "ref_driver")
NPT_MAX = 500
NNB_MAX = 500
KC = 14.399644154020082
TFM = 10.180505306645389
KB = 8.617343e-5


class Metal(C.Structure):
    _fields_ = [(n, C.c_double) for n in
                ("r_e", "f_e", "rho_e", "rho_s", "alpha", "beta", "A", "B", "kappa", "lambda_")] + [
        ("Fn", C.c_double * 4), ("Fm", C.c_double * 4), ("Fp", C.c_double * 4),
        ("eta", C.c_double), ("Fe", C.c_double)]


class Oxid(C.Structure):
    _fields_ = [(n, C.c_double * 5) for n in ("r_e", "alph", "beta", "A", "B", "kapp", "lamb")] + [
        ("F", (C.c_double * 4) * 5), ("re", C.c_double * 5), ("rmax", C.c_double * 5),
        ("fe", C.c_double), ("gamm", C.c_double), ("nu", C.c_double)]


class Potn(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("xm", "qmin", "qmax", "chi", "J", "xi", "Z")] + [
        ("fafb", (C.c_double * 4) * 5)]


class Params(C.Structure):
    _fields_ = [("eam", Metal * 3), ("ox", Oxid), ("prm", Potn * 5)]


class Sys(C.Structure):
    _fields_ = [("n", C.c_int), ("box", C.c_double * 3), ("ncell", C.c_int * 3),
                ("lcell", C.c_double * 3), ("fft", C.c_int * 3), ("rcut", C.c_double),
                ("rcut2", C.c_double), ("a", C.c_double)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle library not built: run `make -C oracle oracle`")
        _lib = C.CDLL(LIB_PATH)
    return _lib


def _p(a, ct):
    return a.ctypes.data_as(C.POINTER(ct))


def _d(a):
    return _p(a, C.c_double)


def _i(a):
    return _p(a, C.c_int)


def set_spread_rounding(on):
    """True: FP32-rounded PME spreading weights as the reference (quirk Q1)."""
    lib().orc_set_spread_rounding(C.c_int(1 if on else 0))


def set_eam_cutoff(r_eam):
    """Split cutoff: EAM terms only for r < r_eam (None or <= 0: the common Rcut)."""
    lib().orc_set_eam_cutoff(C.c_double(r_eam if r_eam else 0.0))


def params():
    p = Params()
    lib().orc_potential_setup(C.byref(p))
    return p


def sys_(n, box, fft, rcut):
    s = Sys()
    b = (C.c_double * 3)(*[float(v) for v in box])
    f = (C.c_int * 3)(*[int(v) for v in fft])
    lib().orc_sys_init(C.byref(s), C.c_int(n), b, f, C.c_double(rcut))
    return s


class Oracle:
    """One system (box, PME grid, cutoff) and its potential parameters."""

    def __init__(self, n, box, fft, rcut, prm=None):
        self.n = int(n)
        self.box = np.asarray(box, dtype=np.float64)
        self.fft = tuple(int(v) for v in fft)
        self.rcut = float(rcut)
        self.s = sys_(self.n, self.box, self.fft, self.rcut)
        self.p = prm if prm is not None else params()
        self.ncell = tuple(self.s.ncell)
        self.ncall = int(np.prod(self.ncell))
        self.pair = np.zeros((self.ncall, 13), dtype=np.int32)
        lib().orc_cell_hierarchy(self.s.ncell, _i(self.pair))
        self.npt = None
        self.link = None

    # cellsort.f90:53-80
    def cell_sort(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        self.npt = np.zeros(self.ncall, dtype=np.int32)
        self.link = np.zeros((self.ncall, NPT_MAX), dtype=np.int32)
        err = lib().orc_cell_sort(C.byref(self.s), _d(x), _i(self.npt), _i(self.link))
        if err:
            raise RuntimeError("=== Cell particle limit crashed !!! ===")
        return self.npt, self.link

    def force(self, ids, x, q, want_lists=False):
        ids = np.ascontiguousarray(ids, dtype=np.int32)
        x = np.ascontiguousarray(x, dtype=np.float64)
        q = np.ascontiguousarray(q, dtype=np.float64)
        f = np.zeros((self.n, 3))
        en = np.zeros(4)
        nlist = np.zeros(self.n, dtype=np.int32) if want_lists else None
        lst = np.zeros((self.n, NNB_MAX), dtype=np.int32) if want_lists else None
        rho = np.zeros(self.n)
        err = lib().orc_force(C.byref(self.s), C.byref(self.p), _i(ids), _d(x), _d(q),
                              _i(self.npt), _i(self.link), _i(self.pair), _d(f), _d(en),
                              _i(nlist) if want_lists else None,
                              _i(lst) if want_lists else None, _d(rho))
        if err:
            raise RuntimeError("oracle FORCE error %d" % err)
        out = dict(f=f, Epair=en[0], Edir=en[1], Ectip=en[2], Egroup=en[3], rho=rho)
        if want_lists:
            out["nlist"] = nlist
            out["list"] = lst
        return out

    def pme_self(self, ids, q):
        ids = np.ascontiguousarray(ids, dtype=np.int32)
        q = np.ascontiguousarray(q, dtype=np.float64)
        psi = np.zeros(self.fft[::-1])  # [z][y][x]
        lib().orc_pme_self.restype = C.c_double
        e = lib().orc_pme_self(C.byref(self.s), C.byref(self.p), _i(ids), _d(q), _d(psi))
        return e, psi

    def pme(self, ids, x, q, psi, f=None):
        ids = np.ascontiguousarray(ids, dtype=np.int32)
        x = np.ascontiguousarray(x, dtype=np.float64)
        q = np.ascontiguousarray(q, dtype=np.float64)
        f = np.zeros((self.n, 3)) if f is None else np.ascontiguousarray(f, dtype=np.float64)
        lib().orc_pme.restype = C.c_double
        e = lib().orc_pme(C.byref(self.s), _i(ids), _d(x), _d(q), _d(np.ascontiguousarray(psi)), _d(f))
        return e, f

    def phi_mesh(self, x, q, psi):
        x = np.ascontiguousarray(x, dtype=np.float64)
        q = np.ascontiguousarray(q, dtype=np.float64)
        phi = np.zeros(self.n)
        lib().orc_phi_mesh(C.byref(self.s), _d(x), _d(q), _d(np.ascontiguousarray(psi)), _d(phi))
        return phi

    def ctip_matrix(self, ids, x):
        ids = np.ascontiguousarray(ids, dtype=np.int32)
        x = np.ascontiguousarray(x, dtype=np.float64)
        nlist = np.zeros(self.n, dtype=np.int32)
        lst = np.zeros((self.n, NNB_MAX), dtype=np.int32)
        asym = np.zeros((self.n, NNB_MAX))
        sym = np.zeros((self.n, NNB_MAX))
        err = lib().orc_ctip_matrix(C.byref(self.s), C.byref(self.p), _i(ids), _d(x),
                                    _i(self.npt), _i(self.link), _i(self.pair), _i(nlist),
                                    _i(lst), _d(asym), _d(sym))
        if err:
            raise RuntimeError("oracle CTIP matrix error %d" % err)
        return nlist, lst, asym, sym

    def ctip_solve(self, ids, x, q, psi):
        ids = np.ascontiguousarray(ids, dtype=np.int32)
        x = np.ascontiguousarray(x, dtype=np.float64)
        q = np.array(q, dtype=np.float64)
        eself = C.c_double()
        iters = C.c_int()
        nev = C.c_int()
        err = lib().orc_ctip_solve(C.byref(self.s), C.byref(self.p), _i(ids), _d(x),
                                   _i(self.npt), _i(self.link), _i(self.pair),
                                   _d(np.ascontiguousarray(psi)), _d(q), C.byref(eself),
                                   C.byref(iters), C.byref(nev))
        if err:
            raise RuntimeError("oracle CTIP error %d" % err)
        return q, eself.value, iters.value, nev.value

    def ctip_energy(self, ids, x, q, psi):
        lib().orc_ctip_energy.restype = C.c_double
        return lib().orc_ctip_energy(
            C.byref(self.s), C.byref(self.p), _i(np.ascontiguousarray(ids, dtype=np.int32)),
            _d(np.ascontiguousarray(x, dtype=np.float64)), _i(self.npt), _i(self.link),
            _i(self.pair), _d(np.ascontiguousarray(psi)), _d(np.ascontiguousarray(q, dtype=np.float64)))

    def xjz(self, ids):
        kc_const = KC * 2.0 * self.s.a / 1.7724538509055159
        pr = self.p.prm
        ids = np.asarray(ids)
        xjz = np.array([[pr[i - 1].chi, pr[i - 1].J - kc_const, pr[i - 1].Z] for i in ids])
        qmx = np.array([[pr[i - 1].qmin, pr[i - 1].qmax] for i in ids])
        return xjz, qmx

    def dudq(self, ids, nlist, lst, asym, sym, q, phi):
        xjz, qmx = self.xjz(ids)
        g = np.zeros(self.n)
        lib().orc_dudq(C.c_int(self.n), _d(np.ascontiguousarray(xjz)), _d(asym), _d(sym),
                       _i(lst), _i(nlist), _d(np.ascontiguousarray(q, dtype=np.float64)),
                       _d(np.ascontiguousarray(qmx)), _d(np.ascontiguousarray(phi)), _d(g))
        return g

    def vv1(self, ids, x, v, f, dt):
        x = np.array(x, dtype=np.float64)
        v = np.array(v, dtype=np.float64)
        lib().orc_vv1(C.byref(self.s), C.byref(self.p), _i(np.ascontiguousarray(ids, dtype=np.int32)),
                      _d(x), _d(v), _d(np.ascontiguousarray(f, dtype=np.float64)), C.c_double(dt))
        return x, v

    def vv2(self, ids, v, f, dt):
        v = np.array(v, dtype=np.float64)
        lib().orc_vv2.restype = C.c_double
        mv2 = lib().orc_vv2(C.byref(self.s), C.byref(self.p),
                            _i(np.ascontiguousarray(ids, dtype=np.int32)), _d(v),
                            _d(np.ascontiguousarray(f, dtype=np.float64)), C.c_double(dt))
        return v, mv2

    # One REGULAR NVE step in the reference's order (main.f90:41-61), with the
    # cell sort and charge solve every step (freq_sort = 1, CTIP every step).
    def step(self, ids, x, v, f, q, psi, dt, ctip=True):
        x, v = self.vv1(ids, x, v, f, dt)
        self.cell_sort(x)
        eself = None
        if ctip:
            q, eself, _, _ = self.ctip_solve(ids, x, q, psi)
        out = self.force(ids, x, q)
        erec, fr = self.pme(ids, x, q, psi, out["f"])
        v, mv2 = self.vv2(ids, v, fr, dt)
        return dict(x=x, v=v, f=fr, q=q, Eself=eself, Erec=erec, mv2=mv2, Epair=out["Epair"],
                    Edir=out["Edir"], Ectip=out["Ectip"], Egroup=out["Egroup"])


LINE_FN = C.CFUNCTYPE(C.c_double, C.c_void_p, C.c_double)
GRAD_FN = C.CFUNCTYPE(None, C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double))


def brent(f):
    """BRENT_SOLVER's scalar search (orc_brent) on a Python f(alpha):
    returns (alpha, evaluations); raises where the reference STOPs."""
    cb = LINE_FN(lambda ctx, a: float(f(a)))
    alpha = C.c_double()
    nev = C.c_int()
    err = lib().orc_brent(cb, None, C.byref(alpha), C.byref(nev))
    if err:
        raise RuntimeError("Brent solver limit breached (%d evaluations)" % nev.value)
    return alpha.value, nev.value


def ncg(grad, q0, maxit=0):
    """CTIP_SOLVER's nonlinear CG loop (orc_ncg) on a Python grad(q) -> g =
    -dU/dq: returns (q, iterations, gradient evaluations)."""
    n = len(q0)

    def _g(ctx, qp, gp):
        q = np.ctypeslib.as_array(qp, shape=(n,)).copy()
        g = np.asarray(grad(q), dtype=np.float64)
        np.ctypeslib.as_array(gp, shape=(n,))[:] = g

    cb = GRAD_FN(_g)
    q = np.array(q0, dtype=np.float64)
    it = C.c_int()
    nev = C.c_int()
    err = lib().orc_ncg(C.c_int(n), cb, None, _d(q), C.c_int(int(maxit)), C.byref(it),
                        C.byref(nev))
    if err:
        raise RuntimeError("Brent solver limit breached in the CG loop")
    return q, it.value, nev.value


def run_reference(ids, x, v, f, q, box, fft, rcut, dt, dudq_inputs=None, workdir=None,
                  t_given=600.0, alpha=0.1, zshift=0.0, reps=0, oxid=None):
    """Run the reference build (oracle/_ref/ref_driver) on one record; see ref_driver.f90.
    t_given (K) is the Berendsen target; alpha the Langevin friction (run at T = 0);
    zshift moves z before VV_NVE1_REFLECT.  reps > 0 also times that many
    CELL_SORT + FORCE repetitions: out["ref_time"] = (reps, seconds).
    oxid = (n_o2, seed, z_new, q_crits): OXIDE_CHECK per q_crit (out["new_o2"])
    and NEW_O2 after RANDOM_SEED(put = (seed, 1, ...)) (out["o2_x"], "o2_v",
    "o2_id", "o2_q")."""
    import tempfile
    if not os.path.exists(REF_DRIVER):
        raise RuntimeError("reference build absent (oracle/_ref/ref_driver)")
    n = len(ids)
    td = workdir or tempfile.mkdtemp(prefix="simox_ref_")
    fin = os.path.join(td, "in.bin")
    fout = os.path.join(td, "out.bin")
    nnb = NNB_MAX if dudq_inputs is not None else 0
    with open(fin, "wb") as fh:
        np.array([n, *fft, nnb], dtype=np.int32).tofile(fh)
        np.array([*box, rcut, dt, t_given, alpha, zshift], dtype=np.float64).tofile(fh)
        np.asarray(ids, dtype=np.int32).tofile(fh)
        for a in (x, v, f):
            np.asarray(a, dtype=np.float64).reshape(n, 3).tofile(fh)
        np.asarray(q, dtype=np.float64).tofile(fh)
        if dudq_inputs is not None:
            nlist, lst, asym, sym, phi = dudq_inputs
            np.asarray(nlist, dtype=np.int32).tofile(fh)
            # Fortran list(n, 500) is column-major: element (i, j) at i + n*j
            np.asfortranarray(np.asarray(lst, dtype=np.int32) + 1).T.tofile(fh)
            np.asfortranarray(asym).T.tofile(fh)
            np.asfortranarray(sym).T.tofile(fh)
            np.asarray(phi, dtype=np.float64).tofile(fh)
        if oxid is not None:
            n_o2, seed, z_new, q_crits = oxid
            np.array([n_o2, seed, len(q_crits)], dtype=np.int32).tofile(fh)
            np.array([z_new, *q_crits], dtype=np.float64).tofile(fh)
    env = dict(os.environ)
    r = subprocess.run(["bash", "-c", 'ulimit -s unlimited; exec "$0" "$1" "$2" "$3"', REF_DRIVER,
                        fin, fout, str(int(reps))], check=True, env=env, stdout=subprocess.PIPE,
                       text=True)
    ref_time = None
    for line in r.stdout.splitlines():
        if line.startswith("REF_TIME"):
            _, k, sec = line.split()
            ref_time = (int(k), float(sec))
    raw = open(fout, "rb").read()
    off = 0

    def take(dtype, count):
        nonlocal off
        a = np.frombuffer(raw, dtype=dtype, count=count, offset=off)
        off += a.nbytes
        return a.copy()

    ncell = take(np.int32, 3)
    ncall = int(np.prod(ncell))
    out = dict(ncell=ncell)
    out["npt"] = take(np.int32, ncall)
    out["link"] = take(np.int32, ncall * NPT_MAX).reshape(ncall, NPT_MAX) - 1
    out["pair"] = take(np.int32, ncall * 13).reshape(ncall, 13) - 1
    out["f"] = take(np.float64, 3 * n).reshape(n, 3)
    e = take(np.float64, 5)
    out.update(Epair=e[0], Edir=e[1], Ectip=e[2], Egroup=e[3], Eself=e[4])
    out["psi"] = take(np.float64, int(np.prod(fft))).reshape(fft[2], fft[1], fft[0])
    out["x1"] = take(np.float64, 3 * n).reshape(n, 3)
    out["v1"] = take(np.float64, 3 * n).reshape(n, 3)
    out["v2"] = take(np.float64, 3 * n).reshape(n, 3)
    out["mv2"] = take(np.float64, 1)[0]
    out["npt_new"] = take(np.int32, 1)[0]
    out["rdf"] = take(np.int32, 5 * 5 * 201).reshape(201, 5, 5)
    out["vb"] = take(np.float64, 3 * n).reshape(n, 3)
    out["mv2b"] = take(np.float64, 1)[0]
    out["xs"] = take(np.float64, 3 * n).reshape(n, 3)
    out["vs"] = take(np.float64, 3 * n).reshape(n, 3)
    out["vs2"] = take(np.float64, 3 * n).reshape(n, 3)
    out["fs2"] = take(np.float64, 3 * n).reshape(n, 3)
    out["mv2s"] = take(np.float64, 1)[0]
    for k in ("xr", "vr", "xr2", "vr2"):
        out[k] = take(np.float64, 3 * n).reshape(n, 3)
    if dudq_inputs is not None:
        out["g"] = take(np.float64, n)
    if oxid is not None:
        n_o2, _, _, q_crits = oxid
        m = 2 * n_o2
        out["new_o2"] = take(np.int32, len(q_crits))
        out["o2_x"] = take(np.float64, 3 * m).reshape(m, 3)
        out["o2_v"] = take(np.float64, 3 * m).reshape(m, 3)
        out["o2_id"] = take(np.int32, m)
        out["o2_q"] = take(np.float64, m)
    if reps:
        out["ref_time"] = ref_time
    return out
