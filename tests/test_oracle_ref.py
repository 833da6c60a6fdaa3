"""The CPU oracle against golden vectors written by the reference build.

tests/golden/*.npz hold outputs of SIMOX's own FORTRAN routines (CELL_SORT,
CELL_HIERARCHY, FORCE, PME_SELF, VV_NVE1, VV_NVE2, dUdq) compiled from the
reference sources (oracle/Makefile, tests/golden/make_golden.py).  tiny_s2 has
Ncell = 2 along x and y, so FORCE's duplicated pairs (quirk Q8) are pinned.
"""
import glob
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURES = sorted(p for p in glob.glob(os.path.join(HERE, "golden", "*.npz"))
                  if not os.path.basename(p).startswith("oxid_"))


def load(path):
    d = dict(np.load(path, allow_pickle=False))
    d["fft"] = tuple(int(k) for k in d["fft"])
    d["rcut"] = float(d["rcut"])
    d["dt"] = float(d["dt"])
    return d


@pytest.fixture(params=FIXTURES, ids=[os.path.basename(p)[:-4] for p in FIXTURES])
def case(request, oracle_mod):
    d = load(request.param)
    o = oracle_mod.Oracle(len(d["ids"]), d["box"], d["fft"], d["rcut"])
    return d, o


def test_fixtures_present():
    assert len(FIXTURES) >= 3


def test_cell_lists_bit_exact(case):
    d, o = case
    npt, link = o.cell_sort(d["x"])
    assert np.array_equal(npt, d["ref_npt"])
    flat = np.concatenate([link[c, :npt[c]] for c in range(len(npt))])
    assert np.array_equal(flat, d["ref_link"])
    assert np.array_equal(o.pair, d["ref_pair"])


def test_force_matches_reference(case):
    d, o = case
    o.cell_sort(d["x"])
    out = o.force(d["ids"], d["x"], d["q"])
    scale = np.abs(d["ref_f"]).max()
    assert np.abs(out["f"] - d["ref_f"]).max() <= 1e-12 * scale
    e = np.array([out["Epair"], out["Edir"], out["Ectip"], out["Egroup"]])
    np.testing.assert_allclose(e, d["ref_E"][:4], rtol=1e-12, atol=1e-12)


def test_pme_self_matches_reference(case):
    d, o = case
    e, psi = o.pme_self(d["ids"], d["q"])
    assert abs(e - d["ref_E"][4]) <= 1e-12 * abs(d["ref_E"][4])
    if "ref_psi" in d:
        np.testing.assert_array_equal(psi, d["ref_psi"])


def test_verlet_matches_reference(case):
    d, o = case
    x1, v1 = o.vv1(d["ids"], d["x"], d["v"], d["f0"], d["dt"])
    assert np.array_equal(x1, d["ref_x1"])
    assert np.array_equal(v1, d["ref_v1"])
    v2, mv2 = o.vv2(d["ids"], v1, d["f0"], d["dt"])
    assert np.array_equal(v2, d["ref_v2"])
    assert mv2 == pytest.approx(float(d["ref_mv2"]), rel=1e-15)


def test_dudq_matches_reference(case):
    d, o = case
    o.cell_sort(d["x"])
    nlist, lst, asym, sym = o.ctip_matrix(d["ids"], d["x"])
    g = o.dudq(d["ids"], nlist, lst, asym, sym, d["q"], d["dudq_phi"])
    np.testing.assert_allclose(g, d["ref_g"], rtol=1e-12, atol=1e-12)


def test_force_neighbor_counts_match_ctip_matrix(case):
    """FORCE and CTIP_SOLVER walk the same cell pairs: identical neighbour lists."""
    d, o = case
    o.cell_sort(d["x"])
    f = o.force(d["ids"], d["x"], d["q"], want_lists=True)
    nlist, lst, _, _ = o.ctip_matrix(d["ids"], d["x"])
    assert np.array_equal(f["nlist"], nlist)
    for i in range(len(nlist)):
        assert np.array_equal(np.sort(f["list"][i, :nlist[i]]), np.sort(lst[i, :nlist[i]]))


def test_neighbor_overflow_is_reported(oracle_mod):
    """At Rcut = 10 A the C1 slab has Ncell = 2 laterally; the duplicated pairs
    (Q8) push atoms past Nnb_max = 500 neighbours.  The reference overruns its
    arrays there (its check tests loop indices, Q10); the oracle reports it."""
    from simox_amd.systems import config
    s = config("c1_rcut10")
    o = oracle_mod.Oracle(s.n, s.box, s.fft, s.rcut)
    o.cell_sort(s.x)
    with pytest.raises(RuntimeError):
        o.force(s.ids, s.x, s.q)


def test_reference_timing_run_keeps_its_outputs(oracle_mod):
    """The CPU baseline's timing run of the reference build (CELL_SORT + FORCE
    repeated after the normal outputs) reports its repetitions and leaves the
    golden outputs as they are."""
    if not os.path.exists(oracle_mod.REF_DRIVER):
        pytest.skip("reference build absent (oracle/_ref/ref_driver)")
    d = load(os.path.join(HERE, "golden", "tiny_s1.npz"))
    args = (d["ids"], d["x"], d["v"], d["f0"], d["q"], d["box"], d["fft"], d["rcut"], d["dt"])
    kw = dict(t_given=float(d["t_given"]), alpha=float(d["alpha"]), zshift=float(d["zshift"]))
    r = oracle_mod.run_reference(*args, reps=3, **kw)
    k, sec = r["ref_time"]
    assert k == 3 and sec > 0.0
    np.testing.assert_array_equal(r["f"], d["ref_f"])
    np.testing.assert_array_equal(r["v2"], d["ref_v2"])
    assert "ref_time" not in oracle_mod.run_reference(*args, **kw)
