"""Converged (FP64) and mixed-precision (MIXED) modes on the GPU.

FP64 mode: exact PME weights, charges solved to an RMS projected-gradient
tolerance; checked against the exact minimiser of the same charge energy from
the oracle (dense KKT solve with exact weights).  Bars: charges 1e-9, forces
1e-10 relative.
MIXED mode: FP32 pair / PME / sparse arithmetic, FP64 outer residual; checked
against FP64 mode on the same state.  Bars: relative force error
e_F = max_i |dF_i| / RMS_i |F_i| <= 1e-5, charges 1e-6, energies 1e-6 relative.
"""
import os
import sys

import numpy as np
import pytest

from simox_amd import MODE_FP64, MODE_MIXED, Simox
from simox_amd.systems import tiny

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
from test_oracle_math import kkt_minimiser  # noqa: E402

pytestmark = pytest.mark.gpu


def e_F(f, ref):
    rms = np.sqrt((ref ** 2).sum(1).mean())
    return np.sqrt(((f - ref) ** 2).sum(1)).max() / rms


@pytest.fixture(scope="module")
def small():
    return tiny(seed=6, nx=3, ny=3, nz=3, n_o=8, fft=(12, 12, 24), rcut=5.0)


def test_fp64_charges_are_the_exact_minimiser(small, oracle_mod):
    s = small
    g = Simox(s.n, s.box, s.fft, s.rcut, mode=MODE_FP64)
    g.set_charge_tol(1e-12, 4000)
    g.set_state(s.ids, s.x, s.v, None, np.zeros(s.n))
    res = g.ctip_solve()
    q = g.get_state()[3]
    o = oracle_mod.Oracle(s.n, s.box, s.fft, s.rcut)
    o.cell_sort(s.x)
    oracle_mod.set_spread_rounding(False)
    try:
        _, psi = o.pme_self(s.ids, np.zeros(s.n))
        q_kkt = kkt_minimiser(o, s.ids, s.x, psi, q.copy())
    finally:
        oracle_mod.set_spread_rounding(True)
    assert abs(q.sum()) < 1e-10
    assert np.abs(q - q_kkt).max() <= 1e-9
    assert res["iters"] > 0


def test_fp64_forces_match_oracle_exact_weights(small, oracle_mod):
    s = small
    rng = np.random.default_rng(1)
    q = rng.normal(size=s.n) * 0.3
    g = Simox(s.n, s.box, s.fft, s.rcut, mode=MODE_FP64)
    g.set_state(s.ids, s.x, s.v, None, q)
    g.force()
    erec = g.pme()
    f = g.get_state()[2]
    o = oracle_mod.Oracle(s.n, s.box, s.fft, s.rcut)
    o.cell_sort(s.x)
    oracle_mod.set_spread_rounding(False)
    try:
        _, psi = o.pme_self(s.ids, q)
        ref = o.force(s.ids, s.x, q)
        e_o, f_o = o.pme(s.ids, s.x, q, psi, ref["f"])
    finally:
        oracle_mod.set_spread_rounding(True)
    assert np.abs(f - f_o).max() <= 1e-10 * np.abs(f_o).max()
    assert erec == pytest.approx(e_o, rel=1e-10)


@pytest.mark.parametrize("name", ["tiny", "c1"])
def test_mixed_matches_fp64(name):
    from simox_amd.systems import config
    s = tiny(seed=2) if name == "tiny" else config("c1")
    out = {}
    for mode in (MODE_FP64, MODE_MIXED):
        g = Simox(s.n, s.box, s.fft, s.rcut, mode=mode)
        g.set_charge_tol(1e-10, 2000)
        g.set_state(s.ids, s.x, s.v, None, np.zeros(s.n))
        r = g.ctip_solve()
        e = g.force()
        erec = g.pme()
        x, v, f, q = g.get_state()
        out[mode] = dict(f=f, q=q, e=e, erec=erec, eself=r["Eself"])
    a, b = out[MODE_MIXED], out[MODE_FP64]
    assert e_F(a["f"], b["f"]) <= 1e-5
    assert np.abs(a["q"] - b["q"]).max() <= 1e-6
    for k in ("Epair", "Egroup", "Edir", "Ectip"):
        assert a["e"][k] == pytest.approx(b["e"][k], rel=1e-6, abs=1e-6)
    assert a["erec"] == pytest.approx(b["erec"], rel=1e-6)
    assert a["eself"] == pytest.approx(b["eself"], rel=1e-6)


def test_mixed_nve_energy_tracks_fp64():
    """20 NVE steps (CTIP every step) in both modes: same total-energy history
    to 1e-6 relative."""
    s = tiny(seed=2)
    hist = {}
    for mode in (MODE_FP64, MODE_MIXED):
        g = Simox(s.n, s.box, s.fft, s.rcut, mode=mode)
        g.set_charge_tol(1e-10, 2000)
        g.set_state(s.ids, s.x, s.v, None, np.zeros(s.n))
        g.ctip_solve()
        g.force()
        g.pme()
        hist[mode] = [g.step(1, s.dt, 1, 1)["Etotal"] for _ in range(20)]
    a, b = np.array(hist[MODE_MIXED]), np.array(hist[MODE_FP64])
    assert np.abs(a - b).max() <= 1e-6 * np.abs(b).max()


@pytest.mark.parametrize("mode", [MODE_FP64, MODE_MIXED])
def test_steps_are_deterministic(mode):
    """Two fresh contexts, same input, 3 steps with sort + charge solve every
    step: bit-identical positions, velocities, forces, charges and energies."""
    s = tiny(seed=4)
    out = []
    for _ in range(2):
        g = Simox(s.n, s.box, s.fft, s.rcut, mode=mode)
        g.set_charge_tol(1e-10, 2000)
        g.set_state(s.ids, s.x, s.v, None, np.zeros(s.n))
        g.force()
        g.pme()
        st = g.step(3, s.dt, 1, 1)
        out.append((g.get_state(), st))
    for a, b in zip(out[0][0], out[1][0]):
        assert np.array_equal(a, b)
    assert out[0][1] == out[1][1]


@pytest.mark.parametrize("mode", [MODE_FP64, MODE_MIXED])
def test_kernel_timing_leaves_the_state(mode):
    """simox_time_kernel relaunches the pair kernels (7 force, 8 list +
    density, 9 CTIP matrix) and the solve's matvec and gather (6, 10) on the
    current buffers: the following steps are bit-identical to those of a
    context that was not timed."""
    s = tiny(seed=4)
    out = []
    for timed in (False, True):
        g = Simox(s.n, s.box, s.fft, s.rcut, mode=mode)
        g.set_charge_tol(1e-10, 2000)
        g.set_state(s.ids, s.x, s.v, None, np.zeros(s.n))
        g.force()
        g.pme()
        g.step(2, s.dt, 1, 1)
        if timed:
            for w in (7, 8, 9, 6, 10):
                assert g.time_kernel(w, 3) > 0.0
        st = g.step(2, s.dt, 1, 1)
        out.append((g.get_state(), st))
        g.close()
    for a, b in zip(out[0][0], out[1][0]):
        assert np.array_equal(a, b)
    assert out[0][1] == out[1][1]


def test_split_cutoff_matches_oracle(oracle_mod):
    """Build-only split cutoff (EAM 6 A inside Rcut 8 A): FAITHFUL forces and
    energies against the oracle with the same split, 1e-10 relative."""
    from simox_amd import MODE_FAITHFUL
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from test_oracle_math import _split_case
    s = _split_case()
    rng = np.random.default_rng(3)
    q = rng.normal(size=s.n) * 0.2
    q -= q.mean()
    g = Simox(s.n, s.box, s.fft, 8.0, mode=MODE_FAITHFUL)
    g.set_eam_cutoff(6.0)
    g.set_state(s.ids, s.x, s.v, None, q)
    e = g.force()
    f = g.get_state()[2]
    o = oracle_mod.Oracle(s.n, s.box, s.fft, 8.0)
    o.cell_sort(s.x)
    oracle_mod.set_eam_cutoff(6.0)
    try:
        ref = o.force(s.ids, s.x, q)
    finally:
        oracle_mod.set_eam_cutoff(None)
    assert np.abs(f - ref["f"]).max() <= 1e-10 * np.abs(ref["f"]).max()
    for k in ("Epair", "Egroup", "Edir", "Ectip"):
        assert e[k] == pytest.approx(ref[k], rel=1e-10, abs=1e-12)
    with pytest.raises(Exception):
        g.set_eam_cutoff(9.0)


def test_long_range_preconditioner_same_minimiser_fewer_iterations(oracle_mod):
    """FP64 solve with the long-range preconditioner (on the half-resolution
    grid 8x8x16): the same exact minimiser (dense KKT solve of the oracle) to
    1e-9, and fewer iterations than Jacobi from the same start."""
    s = tiny(seed=2)
    q = {}
    iters = {}
    for prec in (0, 1):
        g = Simox(s.n, s.box, s.fft, s.rcut, mode=MODE_FP64)
        g.set_charge_tol(1e-12, 4000)
        g.set_charge_precond(prec)
        g.set_state(s.ids, s.x, s.v, None, np.zeros(s.n))
        iters[prec] = g.ctip_solve()["iters"]
        q[prec] = g.get_state()[3]
    o = oracle_mod.Oracle(s.n, s.box, s.fft, s.rcut)
    o.cell_sort(s.x)
    oracle_mod.set_spread_rounding(False)
    try:
        _, psi = o.pme_self(s.ids, np.zeros(s.n))
        q_kkt = kkt_minimiser(o, s.ids, s.x, psi, q[1].copy())
    finally:
        oracle_mod.set_spread_rounding(True)
    assert np.abs(q[1] - q_kkt).max() <= 1e-9
    assert np.abs(q[1] - q[0]).max() <= 1e-9
    assert iters[1] < iters[0]


def test_mixed_preconditioned_matches_jacobi_c2_shape():
    """A 24x24x8-cell slab (18,432 Al + 1,024 O, grid 96x96x64; coarse grid
    48x48x32): MIXED with its default long-range preconditioner against FP64
    with Jacobi, forces within 1e-5 and charges 1e-6, in fewer iterations."""
    from simox_amd.systems import al_o_slab
    s = al_o_slab(24, 24, 8, 1024, 5, 600.0, (96, 96, 64), 10.0, vacuum=20.0,
                  o_layer=(2.0, 10.0), o_dmin=2.0)
    out = {}
    for mode in (MODE_FP64, MODE_MIXED):
        g = Simox(s.n, s.box, s.fft, s.rcut, mode=mode)
        g.set_charge_tol(1e-10, 5000)
        g.set_state(s.ids, s.x, s.v, None, np.zeros(s.n))
        r = g.ctip_solve()
        g.force()
        g.pme()
        out[mode] = (g.get_state()[2], g.get_state()[3], r["iters"])
    (fa, qa, ia), (fb, qb, ib) = out[MODE_MIXED], out[MODE_FP64]
    assert e_F(fa, fb) <= 1e-5
    assert np.abs(qa - qb).max() <= 1e-6
    assert ia < ib


def test_z512_register_pass(oracle_mod, monkeypatch):
    """Kz = 512 takes the register-resident z pass (k_fft_z512_reg) outside
    FAITHFUL: FP64 forces and reciprocal energy against the oracle (exact
    weights, 1e-10), and against the tiled LDS z pass (SIMOX_FFT_ZREG=0) in
    FP64 (1e-13); MIXED with either pass within 1e-5 of FP64."""
    s = tiny(seed=4, nx=3, ny=3, nz=3, n_o=8, fft=(12, 12, 512), rcut=5.0)
    q = np.random.default_rng(5).normal(size=s.n) * 0.3
    print("z512 register pass, n", s.n, flush=True)

    def run(mode, zreg):
        monkeypatch.setenv("SIMOX_FFT_ZREG", "1" if zreg else "0")
        g = Simox(s.n, s.box, s.fft, s.rcut, mode=mode)
        g.set_state(s.ids, s.x, s.v, None, q)
        g.force()
        erec = g.pme()
        f = g.get_state()[2]
        g.close()
        return erec, f

    e_reg, f_reg = run(MODE_FP64, True)
    e_lds, f_lds = run(MODE_FP64, False)
    assert e_reg == pytest.approx(e_lds, rel=1e-13)
    assert np.abs(f_reg - f_lds).max() <= 1e-13 * np.abs(f_lds).max()
    o = oracle_mod.Oracle(s.n, s.box, s.fft, s.rcut)
    o.cell_sort(s.x)
    oracle_mod.set_spread_rounding(False)
    try:
        _, psi = o.pme_self(s.ids, q)
        ref = o.force(s.ids, s.x, q)
        e_o, f_o = o.pme(s.ids, s.x, q, psi, ref["f"])
    finally:
        oracle_mod.set_spread_rounding(True)
    assert np.abs(f_reg - f_o).max() <= 1e-10 * np.abs(f_o).max()
    assert e_reg == pytest.approx(e_o, rel=1e-10)
    m_reg, fm_reg = run(MODE_MIXED, True)
    m_lds, fm_lds = run(MODE_MIXED, False)
    # MIXED: FP32 grids, the bar of MIXED against FP64 (1e-5 of the RMS force)
    assert m_reg == pytest.approx(m_lds, rel=1e-6)
    assert e_F(fm_reg, f_reg) <= 1e-5 and e_F(fm_lds, f_reg) <= 1e-5
    assert m_reg == pytest.approx(e_reg, rel=1e-5)


def test_z512_register_pass_mixed_solve(monkeypatch):
    """The MIXED charge solve's FP32 PME passes with Kz = 512: the same
    converged charges (1e-9) with the register z pass as with the tiled one."""
    s = tiny(seed=4, nx=3, ny=3, nz=3, n_o=8, fft=(12, 12, 512), rcut=5.0)
    out = {}
    for zreg in (True, False):
        monkeypatch.setenv("SIMOX_FFT_ZREG", "1" if zreg else "0")
        g = Simox(s.n, s.box, s.fft, s.rcut, mode=MODE_MIXED)
        g.set_charge_tol(1e-10, 2000)
        g.set_state(s.ids, s.x, s.v, None, np.zeros(s.n))
        r = g.ctip_solve()
        out[zreg] = (g.get_state()[3], r["Eself"])
        g.close()
    assert np.abs(out[True][0] - out[False][0]).max() <= 1e-9
    assert out[True][1] == pytest.approx(out[False][1], rel=1e-9)


@pytest.mark.parametrize("pack", ["1", "0"])
def test_mixed_coarse_inner_grid_matches_fp64(monkeypatch, pack):
    """MIXED with the inner operator's reciprocal part on the half-resolution
    grid and its sparse part from the compacted matrix (simox_set_inner_grid;
    pack 1: column and 12-bit value in one 32-bit word, 0: FP16 values): the
    restarts use the reference's grid and the FP64 matrix, so the converged
    charges equal FP64's to 1e-9 and the forces to 1e-5 of the RMS force,
    with the coarse operator on and off."""
    from simox_amd.systems import al_o_slab
    monkeypatch.setenv("SIMOX_INNER_PACK", pack)
    s = al_o_slab(24, 24, 8, 1024, 5, 600.0, (96, 96, 64), 10.0, vacuum=20.0,
                  o_layer=(2.0, 10.0), o_dmin=2.0, name="inner")
    print("coarse inner grid, n", s.n, "pack", pack, flush=True)
    out = {}
    for tag in ("fp64", "coarse", "fine"):
        g = Simox(s.n, s.box, s.fft, s.rcut, mode=MODE_FP64 if tag == "fp64" else MODE_MIXED)
        if tag != "fp64":
            g.set_inner_grid(tag == "coarse")
        g.set_charge_tol(1e-10, 5000)
        g.set_state(s.ids, s.x, s.v, None, np.zeros(s.n))
        r = g.ctip_solve()
        if tag == "coarse":
            # the inner matrix leaves out the far entries (|H0_ij| < 1e-3)
            full = int(g.neighbors(stride=0)[0].sum())
            kept = g.inner_entries()
            assert 0 < kept < 0.8 * full, (kept, full)
        g.force()
        g.pme()
        out[tag] = (g.get_state()[3], g.get_state()[2], r["iters"], g.charge_stats()[1])
        g.close()
    for tag in ("coarse", "fine"):
        assert np.abs(out[tag][0] - out["fp64"][0]).max() <= 1e-9, tag
        assert e_F(out[tag][1], out["fp64"][1]) <= 1e-5, tag
    assert out["coarse"][3] >= out["fine"][3]  # more restarts with the coarse operator
    with pytest.raises(Exception):
        g = Simox(s.n, s.box, s.fft, s.rcut, mode=MODE_FP64)
        try:
            g.set_inner_grid(True)
        finally:
            g.close()


@pytest.mark.parametrize("fft,coarse", [((128, 128, 64), True), ((32, 32, 32), False)])
def test_owner_spread_matches_tiled_spread(monkeypatch, fft, coarse):
    """The owner-computes 32-bit spread with its fold (k_spread_own,
    k_spread_fold), taken on the coarse grids when their tiles divide them
    (grid 128 x 128 x 64: the inner operator's and the preconditioner's
    64 x 64 x 32 grids, 32 x 32 x 16 tiles; grid 32^3 with the inner operator
    on the reference grid, whose 16^3 half-resolution grid is too coarse for
    the outer iteration to converge: the preconditioner's 16^3 grid, 16 x 16 x 8
    tiles owning about 9,700 atoms, more than the 8,192 whose charges a tile
    keeps in LDS, so the spread reads them again; one tile along x and y, so
    the fold adds a tile's own overhang across the periodic boundary),
    against the tiled spread over each tile's candidates (SIMOX_SPREAD_OWN=0):
    both solves converge to FP64's charges within 1e-9 and forces within 1e-5
    of the RMS force, and the owner path gives the same bits when repeated
    (integer sums per tile, the fold in a fixed order)."""
    from simox_amd.systems import al_o_slab
    s = al_o_slab(24, 24, 8, 1024, 5, 600.0, fft, 10.0, vacuum=20.0,
                  o_layer=(2.0, 10.0), o_dmin=2.0, name="own")
    print("owner spread, n", s.n, "fft", fft, flush=True)

    def run(mode, own):
        monkeypatch.setenv("SIMOX_SPREAD_OWN", "1" if own else "0")
        g = Simox(s.n, s.box, s.fft, s.rcut, mode=mode)
        if mode == MODE_MIXED:
            g.set_inner_grid(coarse)
        g.set_charge_tol(1e-10, 5000)
        g.set_state(s.ids, s.x, s.v, None, np.zeros(s.n))
        r = g.ctip_solve()
        g.force()
        g.pme()
        _, _, f, q = g.get_state()
        g.close()
        return q, f, r["iters"]

    q64, f64, _ = run(MODE_FP64, True)
    q1, f1, it1 = run(MODE_MIXED, True)
    q0, f0, it0 = run(MODE_MIXED, False)
    q2, f2, _ = run(MODE_MIXED, True)
    print("iterations owner", it1, "tiled", it0, flush=True)
    for q, f in ((q1, f1), (q0, f0)):
        assert np.abs(q - q64).max() <= 1e-9
        assert e_F(f, f64) <= 1e-5
    assert np.array_equal(q1, q2) and np.array_equal(f1, f2)


def test_inner_shapes_change_no_result(monkeypatch):
    """The inner solve's tuned launch shapes (simox_inner_shapes: packed
    matvec variant, record gather tile, plane FFT threads and z lines of both
    coarse grids) give the same charges and forces bit for bit whichever is
    chosen: each fixed in turn (SIMOX_AUTOTUNE=0 and one switch) against the
    tuned run; so do the pair kernels' batch depths (list entries loaded
    ahead, sums in the same order)."""
    from simox_amd.systems import al_o_slab
    s = al_o_slab(24, 24, 8, 1024, 5, 600.0, (96, 96, 64), 10.0, vacuum=20.0,
                  o_layer=(2.0, 10.0), o_dmin=2.0, name="inner")

    def run(env):
        for k in ("SIMOX_AUTOTUNE", "SIMOX_MV_VARIANT", "SIMOX_GATHER_SHAPE", "SIMOX_FFT_XYNT",
                  "SIMOX_FFT_NL", "SIMOX_FORCE_BATCH", "SIMOX_DENS_BATCH", "SIMOX_CTIP_BATCH"):
            monkeypatch.delenv(k, raising=False)
        for k, v in env.items():
            monkeypatch.setenv(k, v)
        g = Simox(s.n, s.box, s.fft, s.rcut, mode=MODE_MIXED)
        g.set_inner_grid(True)
        g.set_charge_tol(1e-10, 5000)
        g.set_state(s.ids, s.x, s.v, None, np.zeros(s.n))
        g.ctip_solve()
        g.force()
        g.pme()
        _, _, f, q = g.get_state()
        sh = g.inner_shapes()
        g.close()
        return q, f, sh

    q0, f0, sh = run({})
    print("tuned inner shapes", sh, flush=True)
    # row blocks past a lowered halo limit gather p from global memory: the
    # halo variants then run both kinds of block, with the same results
    for v in ("8", "10"):
        q, f, _ = run({"SIMOX_AUTOTUNE": "0", "SIMOX_MV_VARIANT": v, "SIMOX_HALO_MAX": "1200"})
        assert np.array_equal(q, q0) and np.array_equal(f, f0), v
    assert sh["matvec"] in range(11) and sh["gather"] in range(6)
    assert sh["inner_grid"][0] in (256, 512, 1024) and sh["inner_grid"][1] in (4, 8, 11, 16)
    cases = [{"SIMOX_MV_VARIANT": str(v)} for v in range(11)]
    cases += [{"SIMOX_GATHER_SHAPE": str(v)} for v in range(6)]
    cases += [{"SIMOX_FFT_XYNT": t} for t in ("256", "512", "1024")]
    cases += [{"SIMOX_FFT_NL": m} for m in ("4", "8", "16")]
    # list entries whose neighbour data each lane loads ahead: the same sums
    # (each lane walks its atom's whole list, so this system covers them)
    cases += [{k: b} for k in ("SIMOX_FORCE_BATCH", "SIMOX_DENS_BATCH", "SIMOX_CTIP_BATCH")
              for b in ("1", "2")]
    for env in cases:
        env = dict(env, SIMOX_AUTOTUNE="0")
        q, f, sh = run(env)
        assert np.array_equal(q, q0) and np.array_equal(f, f0), (env, sh)


def test_z256_register_pass_fp32(monkeypatch):
    """Kz = 256 on the FP32 chain takes the radix-4 register z pass
    (k_fft_z256_reg): the MIXED reciprocal energy with it and with the tiled
    pass (SIMOX_FFT_ZREG=0) agree to 1e-6, and the forces of both stay within
    1e-5 of FP64 (the bar of MIXED; the two FP32 passes round differently)."""
    s = tiny(seed=4, nx=3, ny=3, nz=3, n_o=8, fft=(12, 12, 256), rcut=5.0)
    q = np.random.default_rng(6).normal(size=s.n) * 0.3
    print("z256 register pass, n", s.n, flush=True)

    def run(mode, zreg):
        monkeypatch.setenv("SIMOX_FFT_ZREG", "1" if zreg else "0")
        g = Simox(s.n, s.box, s.fft, s.rcut, mode=mode)
        g.set_state(s.ids, s.x, s.v, None, q)
        g.force()
        erec = g.pme()
        f = g.get_state()[2]
        g.close()
        return erec, f

    e64, f64 = run(MODE_FP64, True)
    e_reg, f_reg = run(MODE_MIXED, True)
    e_lds, f_lds = run(MODE_MIXED, False)
    assert e_reg == pytest.approx(e_lds, rel=1e-6)
    assert e_F(f_reg, f64) <= 1e-5 and e_F(f_lds, f64) <= 1e-5
    assert e_reg == pytest.approx(e64, rel=1e-5)


@pytest.mark.parametrize("fft,nl,nt", [((32, 32, 64), "16", "256"), ((48, 48, 96), "8", "512"),
                                       ((128, 128, 32), "16", "512"), ((64, 64, 256), "16", "256"),
                                       ((96, 96, 128), "4", "1024"), ((128, 128, 64), "11", "256"),
                                       ((128, 128, 256), "8", "1024")])
def test_plane_fft_chain_fp32(monkeypatch, fft, nl, nt):
    """The plane-resident FP32 chain (x and y transformed in one LDS pass per
    direction, z pass on the plane-major spectrum) that the coarse grids of
    the MIXED solve take, switched onto the reference's grid with
    SIMOX_FFT_PLANE=2: the reciprocal energy against the five-pass FP32 chain
    (1e-6 relative; the two pair the real x lines differently, so they round
    differently) and the forces of both against FP64 (1e-5, the bar of
    MIXED)."""
    s = tiny(seed=4, nx=3, ny=3, nz=3, n_o=8, fft=fft, rcut=5.0)
    q = np.random.default_rng(6).normal(size=s.n) * 0.3
    print("plane chain", fft, "nl", nl, "threads", nt, "n", s.n, flush=True)
    monkeypatch.setenv("SIMOX_FFT_NL", nl)
    monkeypatch.setenv("SIMOX_FFT_XYNT", nt)

    def run(mode, plane):
        monkeypatch.setenv("SIMOX_FFT_PLANE", plane)
        g = Simox(s.n, s.box, s.fft, s.rcut, mode=mode)
        g.set_state(s.ids, s.x, s.v, None, q)
        g.force()
        erec = g.pme()
        f = g.get_state()[2]
        g.close()
        return erec, f

    e64, f64 = run(MODE_FP64, "0")
    e_pl, f_pl = run(MODE_MIXED, "2")
    e_5, f_5 = run(MODE_MIXED, "0")
    assert not np.array_equal(f_pl, f_5)  # the plane chain ran
    assert e_pl == pytest.approx(e_5, rel=1e-6)
    assert e_F(f_pl, f64) <= 1e-5 and e_F(f_5, f64) <= 1e-5
    assert e_pl == pytest.approx(e64, rel=1e-5)
