"""ctypes binding of libsimox_gpu.so (include/simox_gpu.h).

Simox mirrors the reference's external subroutines for the hot path; each
method cites the routine it replaces.  Arrays follow the reference's
conventions (caller atom order, species 1..5, 3-vectors as (n, 3) float64).
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsimox_gpu.so")

MODE_FAITHFUL, MODE_FP64, MODE_MIXED = 0, 1, 2

# names of the entry points declared in include/simox_gpu.h
EXPORTS = [
    "simox_potential_setup", "simox_fafb", "simox_create", "simox_destroy", "simox_last_error",
    "simox_get_geometry", "simox_set_mode", "simox_set_charge_tol", "simox_set_state",
    "simox_get_state", "simox_cell_sort", "simox_get_cells", "simox_force", "simox_get_rho",
    "simox_pme_self", "simox_pme", "simox_phi", "simox_ctip_solve", "simox_dudq",
    "simox_get_neighbors", "simox_vv1", "simox_vv2", "simox_step", "simox_kernel_timing",
    "simox_timing_enable", "simox_timing_select", "simox_set_charge_precond", "simox_set_inner_grid", "simox_inner_entries", "simox_charge_stats", "simox_set_params",
    "simox_autotune", "simox_fft_shapes", "simox_inner_shapes", "simox_tile_shape", "simox_set_eam_cutoff", "simox_set_eam_splines",
    "simox_eam_spline_grad", "simox_param_grad_all", "simox_eam_spline_force_vjp",
    "simox_param_force_grad",
    "simox_set_thermostat", "simox_set_rdf", "simox_get_rdf", "simox_set_reflect",
    "simox_remove_rigid_motion", "simox_param_grad", "simox_batch_step",
    "simox_cell_sort_noalloc", "simox_time_kernel",
]

# simox_kernel_timing categories
T_SPREAD, T_FFT, T_GATHER, T_SPMV, T_FORCE, T_DENSITY = range(6)


class SimoxError(RuntimeError):
    """Raised where the reference would STOP; carries the reference's text."""

    _m = 0  # coefficients per spline function once set_eam_splines ran

    def __init__(self, code, text):
        super().__init__(text)
        self.code = code


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


_lib = None


def load_library(path=LIB_PATH):
    """Load the HIP library; raises if it has not been built (no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError("libsimox_gpu.so is not built: run `python -m simox_amd.build`")
    lib = C.CDLL(path)
    i, d, vp = C.c_int, C.c_double, C.c_void_p
    P = C.POINTER
    lib.simox_last_error.restype = C.c_char_p
    lib.simox_last_error.argtypes = [vp]
    lib.simox_create.argtypes = [i, P(d), P(i), d, i, vp, P(vp)]
    lib.simox_destroy.argtypes = [vp]
    lib.simox_get_geometry.argtypes = [vp, P(i), P(d), P(d)]
    lib.simox_set_mode.argtypes = [vp, i]
    lib.simox_set_charge_tol.argtypes = [vp, d, i]
    lib.simox_set_params.argtypes = [vp, vp]
    lib.simox_autotune.argtypes = [vp]
    lib.simox_set_eam_splines.argtypes = [vp, i, vp, vp, vp]
    lib.simox_eam_spline_grad.argtypes = [vp, vp]
    lib.simox_eam_spline_force_vjp.argtypes = [vp, vp, vp]
    lib.simox_set_eam_cutoff.argtypes = [vp, d]
    lib.simox_set_thermostat.argtypes = [vp, i, d, d, C.c_ulonglong]
    lib.simox_set_rdf.argtypes = [vp, i]
    lib.simox_set_reflect.argtypes = [vp, i]
    lib.simox_param_grad.argtypes = [vp, P(d), vp, P(i)]
    lib.simox_param_grad_all.argtypes = [vp, P(d), vp, P(i)]
    lib.simox_param_force_grad.argtypes = [vp, vp, P(d), P(i)]
    lib.simox_batch_step.argtypes = [P(vp), i, i, d, i, i, vp]
    lib.simox_remove_rigid_motion.argtypes = [vp]
    lib.simox_get_rdf.argtypes = [vp, P(C.c_longlong), i]
    lib.simox_fft_shapes.argtypes = [vp, P(i)]
    lib.simox_inner_shapes.argtypes = [vp, P(i)]
    lib.simox_tile_shape.argtypes = [vp, P(i)]
    lib.simox_set_state.argtypes = [vp, vp, vp, vp, vp, vp]
    lib.simox_get_state.argtypes = [vp, vp, vp, vp, vp]
    lib.simox_cell_sort.argtypes = [vp]
    lib.simox_get_cells.argtypes = [vp, vp, vp, i]
    lib.simox_force.argtypes = [vp, P(d)]
    lib.simox_get_rho.argtypes = [vp, vp]
    lib.simox_pme_self.argtypes = [vp, P(d)]
    lib.simox_pme.argtypes = [vp, P(d)]
    lib.simox_phi.argtypes = [vp, vp]
    lib.simox_ctip_solve.argtypes = [vp, P(d), P(i), P(i)]
    lib.simox_dudq.argtypes = [vp, vp]
    lib.simox_get_neighbors.argtypes = [vp, vp, vp, i]
    lib.simox_vv1.argtypes = [vp, d]
    lib.simox_vv2.argtypes = [vp, d, P(d)]
    lib.simox_step.argtypes = [vp, i, d, i, i, vp]
    lib.simox_kernel_timing.argtypes = [vp, i, P(d), P(C.c_longlong)]
    lib.simox_timing_enable.argtypes = [vp, i]
    lib.simox_timing_select.argtypes = [vp, C.c_uint]
    lib.simox_set_charge_precond.argtypes = [vp, i]
    lib.simox_set_inner_grid.argtypes = [vp, i]
    lib.simox_inner_entries.argtypes = [vp, P(C.c_longlong)]
    lib.simox_charge_stats.argtypes = [vp, P(i), P(i)]
    lib.simox_potential_setup.argtypes = [vp]
    lib.simox_fafb.argtypes = [vp]
    lib.simox_cell_sort_noalloc.argtypes = [i, P(d), d, vp, vp, vp, vp]
    lib.simox_time_kernel.argtypes = [vp, i, i, P(d)]
    _lib = lib
    return lib


def potential_setup():
    """POTENTIAL_SETUP (datafmt.f90:142-262) -> Params."""
    p = Params()
    load_library().simox_potential_setup(C.byref(p))
    return p


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _f64(a, shape=None):
    if a is None:
        return None
    a = np.ascontiguousarray(a, dtype=np.float64)
    if shape is not None:
        a = a.reshape(shape)
    return a


class Simox:
    """One system on the GPU: box, PME grid, cutoff, mode (include/simox_gpu.h)."""

    def __init__(self, n, box, fft, rcut, mode=MODE_FAITHFUL, params=None):
        self.lib = load_library()
        self.n = int(n)
        self.mode = mode
        box = (C.c_double * 3)(*[float(b) for b in box])
        fftc = (C.c_int * 3)(*[int(k) for k in fft])
        h = C.c_void_p()
        rc = self.lib.simox_create(self.n, box, fftc, float(rcut), int(mode),
                                   C.byref(params) if params is not None else None, C.byref(h))
        self.h = h
        self._check(rc)

    @classmethod
    def from_system(cls, s, mode=MODE_FAITHFUL, params=None, with_state=True):
        obj = cls(s.n, s.box, s.fft, s.rcut, mode=mode, params=params)
        if with_state:
            obj.set_state(s.ids, s.x, s.v, None, s.q)
        return obj

    def _check(self, rc):
        if rc != 0:
            txt = self.lib.simox_last_error(self.h).decode() if self.h else "error"
            raise SimoxError(rc, txt)

    def close(self):
        if getattr(self, "h", None):
            self.lib.simox_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def geometry(self):
        nc = (C.c_int * 3)()
        lc = (C.c_double * 3)()
        a = C.c_double()
        self._check(self.lib.simox_get_geometry(self.h, nc, lc, C.byref(a)))
        return tuple(nc), tuple(lc), a.value

    def set_thermostat(self, kind, t_kelvin=0.0, alpha=0.0, seed=1):
        """0 NVE, 1 Berendsen (VV_NVT2_BEREND), 2 Langevin (VV_NVT1/2_STOCH)."""
        self._check(self.lib.simox_set_thermostat(self.h, int(kind), float(t_kelvin),
                                                  float(alpha), int(seed)))

    def param_grad(self, want_lambda=False, with_xi=False):
        """dE/dchi (5), dE/dJ (5) and, with_xi, dE/dxi (5) of E = Eself + Edir +
        Erec + Ectip at the converged charges, through the charge
        equilibration (adjoint solve)."""
        g = (C.c_double * 15)()
        lam = np.zeros(self.n) if want_lambda else None
        it = C.c_int()
        fn = self.lib.simox_param_grad_all if with_xi else self.lib.simox_param_grad
        self._check(fn(self.h, g, _ptr(lam), C.byref(it)))
        out = dict(dchi=np.array(g[0:5]), dJ=np.array(g[5:10]), iters=it.value)
        if with_xi:
            out["dxi"] = np.array(g[10:15])
        if want_lambda:
            out["lambda"] = lam
        return out

    def param_force_grad(self, u=None):
        """d(sum_i u_i . F_i)/d(chi, J, xi) through the converged charges (F the
        forces of FORCE + PME; u None: the current forces)."""
        g = (C.c_double * 15)()
        it = C.c_int()
        ua = None if u is None else np.ascontiguousarray(u, dtype=np.float64).reshape(-1, 3)
        self._check(self.lib.simox_param_force_grad(self.h, _ptr(ua), g, C.byref(it)))
        return dict(dchi=np.array(g[0:5]), dJ=np.array(g[5:10]), dxi=np.array(g[10:15]),
                    iters=it.value)

    def set_reflect(self, on=True):
        """OXIDATION mode first half: VV_NVE1_REFLECT instead of VV_NVE1."""
        self._check(self.lib.simox_set_reflect(self.h, 1 if on else 0))

    def remove_rigid_motion(self):
        """REMOVE_RIGID_MOTION: subtract the mean velocity and the mean z."""
        self._check(self.lib.simox_remove_rigid_motion(self.h))

    def set_rdf(self, on=True):
        """Accumulate FORCE's pair histogram (pair%rdf) on every force call."""
        self._check(self.lib.simox_set_rdf(self.h, 1 if on else 0))

    def rdf(self, reset=False):
        """Histogram [bin-1][sj-1][si-1] (201 x 5 x 5), one count per pair direction."""
        h = np.zeros(201 * 25, dtype=np.int64)
        self._check(self.lib.simox_get_rdf(self.h, h.ctypes.data_as(C.POINTER(C.c_longlong)),
                                           1 if reset else 0))
        return h.reshape(201, 5, 5)

    def set_eam_cutoff(self, r_eam):
        """Split cutoff: EAM terms only for r < r_eam (<= Rcut)."""
        self._check(self.lib.simox_set_eam_cutoff(self.h, float(r_eam)))

    def autotune(self):
        """Re-time the FFT tile shapes on this context (done at creation)."""
        self._check(self.lib.simox_autotune(self.h))

    def fft_shapes(self):
        """{"fp64": (mx, my, mz), "fp32": (mx, my, mz)} FFT tile shapes."""
        a = (C.c_int * 6)()
        self._check(self.lib.simox_fft_shapes(self.h, a))
        return {"fp64": tuple(a[0:3]), "fp32": tuple(a[3:6])}

    def inner_shapes(self):
        """Launch shapes of the MIXED inner solve (tuned at the first solve):
        {"matvec": variant, "gather": tile index, "inner_grid": (x-y threads,
        z lines), "prec_grid": (x-y threads, z lines)}."""
        a = (C.c_int * 6)()
        self._check(self.lib.simox_inner_shapes(self.h, a))
        return {"matvec": a[0], "gather": a[1], "inner_grid": (a[2], a[3]),
                "prec_grid": (a[4], a[5])}

    def tile_shape(self):
        """Cell tile of the neighbour-list build (tuned at the first list
        build): (atoms staged per LDS chunk, threads per workgroup)."""
        a = (C.c_int * 2)()
        self._check(self.lib.simox_tile_shape(self.h, a))
        return (a[0], a[1])

    def set_params(self, params):
        """Replace the potential parameters (a Params from potential_setup())."""
        self._check(self.lib.simox_set_params(self.h, C.byref(params)))

    def set_charge_tol(self, tol, maxit=1000):
        self._check(self.lib.simox_set_charge_tol(self.h, float(tol), int(maxit)))

    def set_charge_precond(self, on=True):
        self._check(self.lib.simox_set_charge_precond(self.h, 1 if on else 0))

    def set_inner_grid(self, on=True):
        """MIXED: the inner solves' reciprocal part on a half-resolution grid."""
        self._check(self.lib.simox_set_inner_grid(self.h, 1 if on else 0))

    def inner_entries(self):
        """Entries of the matrix the inner solves' sparse part reads."""
        e = C.c_longlong()
        self._check(self.lib.simox_inner_entries(self.h, C.byref(e)))
        return e.value

    def charge_stats(self):
        it, rs = C.c_int(), C.c_int()
        self._check(self.lib.simox_charge_stats(self.h, C.byref(it), C.byref(rs)))
        return it.value, rs.value

    # TYPE PTCL q(:) in / out
    def set_state(self, ids=None, x=None, v=None, f=None, q=None):
        n = self.n
        idc = None if ids is None else np.ascontiguousarray(ids, dtype=np.int32)
        xa, va, fa = _f64(x, (n, 3)), _f64(v, (n, 3)), _f64(f, (n, 3))
        qa = _f64(q, (n,))
        self._check(self.lib.simox_set_state(self.h, _ptr(idc), _ptr(xa), _ptr(va), _ptr(fa),
                                             _ptr(qa)))

    def get_state(self):
        n = self.n
        x, v, f, q = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros((n, 3)), np.zeros(n)
        self._check(self.lib.simox_get_state(self.h, _ptr(x), _ptr(v), _ptr(f), _ptr(q)))
        return x, v, f, q

    # CELL_SORT (cellsort.f90:53-80)
    def cell_sort(self):
        self._check(self.lib.simox_cell_sort(self.h))

    def get_cells(self, link_stride=500):
        nc, _, _ = self.geometry()
        ncall = int(np.prod(nc))
        npt = np.zeros(ncall, dtype=np.int32)
        link = np.full((ncall, link_stride), -1, dtype=np.int32)
        self._check(self.lib.simox_get_cells(self.h, _ptr(npt), _ptr(link), link_stride))
        return npt, link

    # FORCE (force.f90:3-361)
    def force(self):
        e = (C.c_double * 4)()
        self._check(self.lib.simox_force(self.h, e))
        return dict(Epair=e[0], Edir=e[1], Ectip=e[2], Egroup=e[3])

    def rho(self):
        r = np.zeros(self.n)
        self._check(self.lib.simox_get_rho(self.h, _ptr(r)))
        return r

    # PME_SELF (force.f90:467-517)
    def pme_self(self):
        e = C.c_double()
        self._check(self.lib.simox_pme_self(self.h, C.byref(e)))
        return e.value

    # PME (force.f90:364-464)
    def pme(self):
        e = C.c_double()
        self._check(self.lib.simox_pme(self.h, C.byref(e)))
        return e.value

    # PHI_MESH x c_pme (ctip.f90:336-396, 170)
    def phi(self):
        p = np.zeros(self.n)
        self._check(self.lib.simox_phi(self.h, _ptr(p)))
        return p

    # CTIP_SOLVER (ctip.f90:3-212)
    def ctip_solve(self):
        e, it, ne = C.c_double(), C.c_int(), C.c_int()
        self._check(self.lib.simox_ctip_solve(self.h, C.byref(e), C.byref(it), C.byref(ne)))
        return dict(Eself=e.value, iters=it.value, n_eval=ne.value)

    # dUdq (ctip.f90:307-333)
    def dudq(self):
        g = np.zeros(self.n)
        self._check(self.lib.simox_dudq(self.h, _ptr(g)))
        return g

    def neighbors(self, stride=500):
        """Counts and (stride > 0) column lists of the CTIP matrix rows."""
        cnt = np.zeros(self.n, dtype=np.int32)
        cols = np.full((self.n, stride), -1, dtype=np.int32) if stride > 0 else None
        self._check(self.lib.simox_get_neighbors(self.h, _ptr(cnt), _ptr(cols), max(stride, 0)))
        return cnt, cols

    # VV_NVE1 / VV_NVE2 (vverlet.f90:74-157)
    def vv1(self, dt):
        self._check(self.lib.simox_vv1(self.h, float(dt)))

    def vv2(self, dt):
        m = C.c_double()
        self._check(self.lib.simox_vv2(self.h, float(dt), C.byref(m)))
        return m.value

    # REGULAR NVE loop body (main.f90:41-61)
    def step(self, nsteps, dt, sort_every=1, ctip_every=1, stats=True):
        st = np.zeros(8) if stats else None
        self._check(self.lib.simox_step(self.h, int(nsteps), float(dt), int(sort_every),
                                        int(ctip_every), _ptr(st)))
        if st is None:
            return None
        keys = ["Eself", "Edir", "Erec", "Ectip", "Epair", "Egroup", "mv2", "Etotal"]
        return dict(zip(keys, st.tolist()))

    def set_eam_splines(self, lo, h, coef):
        """Spline-form EAM (coef [8, m]); coef None: back to the analytic forms."""
        if coef is None:
            self._check(self.lib.simox_set_eam_splines(self.h, 0, None, None, None))
            self._m = 0
            return
        coef = np.ascontiguousarray(coef, dtype=np.float64)
        lo = np.ascontiguousarray(lo, dtype=np.float64)
        h = np.ascontiguousarray(h, dtype=np.float64)
        self._check(self.lib.simox_set_eam_splines(self.h, coef.shape[1], _ptr(lo), _ptr(h),
                                                   _ptr(coef)))
        self._m = coef.shape[1]

    def eam_spline_grad(self):
        """dE/dcoef [8, m] of E = Epair + Egroup at the current positions."""
        g = np.zeros((8, self._m))
        self._check(self.lib.simox_eam_spline_grad(self.h, _ptr(g)))
        return g

    def eam_spline_force_vjp(self, u=None):
        """dS/dcoef [8, m] for S = sum_i u_i . F_i(EAM); u None: the current
        total forces (then dE/dcoef + 2 w this = d(E + w sum |F|^2)/dcoef)."""
        g = np.zeros((8, self._m))
        ua = None if u is None else np.ascontiguousarray(u, dtype=np.float64).reshape(-1, 3)
        self._check(self.lib.simox_eam_spline_force_vjp(self.h, _ptr(ua), _ptr(g)))
        return g

    def timing(self, on=True):
        self._check(self.lib.simox_timing_enable(self.h, 1 if on else 0))

    def timing_select(self, which=None):
        """Time only the categories in `which` (None: all)."""
        mask = 0 if which is None else sum(1 << int(w) for w in which)
        self._check(self.lib.simox_timing_select(self.h, mask))

    def time_kernel(self, which, reps=50):
        """ms per launch of an inner-solve kernel (6 matvec, 10 gather, 11
        inner spread) on the last solve's buffers, or of a pair kernel (7 force,
        8 neighbour list + density, 9 CTIP matrix) on the current positions:
        reps launches back to back."""
        ms = C.c_double()
        self._check(self.lib.simox_time_kernel(self.h, int(which), int(reps), C.byref(ms)))
        return ms.value

    def kernel_time(self, which):
        ms = C.c_double()
        cnt = C.c_longlong()
        self._check(self.lib.simox_kernel_timing(self.h, int(which), C.byref(ms), C.byref(cnt)))
        return ms.value, cnt.value


def cell_sort_noalloc(x, box, rcut, npt, start, link):
    """simox_cell_sort_noalloc on device buffers given as addresses (int):
    x (3n float64), npt (ncell int32), start (ncell + 1 int32), link (n
    int32).  No allocation and no context; returns nothing, raises on a
    nonzero status."""
    lib = load_library()
    n = int(x[1])
    b = (C.c_double * 3)(*[float(v) for v in box])
    rc = lib.simox_cell_sort_noalloc(n, b, float(rcut), C.c_void_p(x[0]), C.c_void_p(npt),
                                     C.c_void_p(start), C.c_void_p(link))
    if rc:
        raise SimoxError(rc, "simox_cell_sort_noalloc failed with status %d" % rc)


def batch_step(sims, nsteps, dt, sort_every=1, ctip_every=1, stats=True):
    """simox_batch_step: the same steps on every Simox of sims in lockstep
    (independent systems; each keeps its own box, grid and stream).  Returns
    a list of stats dicts (as Simox.step) or None."""
    lib = load_library()
    arr = (C.c_void_p * len(sims))(*[s.h.value for s in sims])
    st = np.zeros(8 * len(sims)) if stats else None
    rc = lib.simox_batch_step(arr, len(sims), int(nsteps), float(dt), int(sort_every),
                              int(ctip_every), _ptr(st))
    if rc:
        for s in sims:  # report the replica that holds the message
            txt = lib.simox_last_error(s.h)
            if txt:
                raise SimoxError(rc, txt.decode())
        raise SimoxError(rc, "batch step failed")
    if st is None:
        return None
    keys = ["Eself", "Edir", "Erec", "Ectip", "Epair", "Egroup", "mv2", "Etotal"]
    return [dict(zip(keys, st[8 * k:8 * k + 8].tolist())) for k in range(len(sims))]
