"""OXIDATION-mode host routines (simox_amd/fortran/simox_oxid.f90) against
the reference build: OXIDE_CHECK (util.f90:152-188) and NEW_O2
(datafmt.f90:98-139 with RANDOM_POSITION_O2, util.f90:104-149).

tests/golden/oxid_s1.npz holds, for four records of a small Al/O slab (as
built, its last O moved below the top Al, an Al among the last atoms, one to
three new dimers), the new-dimer flag the reference's OXIDE_CHECK returns for
five q_crit values and the dimers its NEW_O2 places after
RANDOM_SEED(put = (seed, 1, 1, ...)) (tests/golden/make_golden.py).  The
product module runs the same cases through tests/oxid_driver.f90, built with
the same Fortran compiler and runtime (so the same random stream): flags
equal, positions and velocities bit for bit.
"""
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
FC = os.environ.get("SIMOX_FC", "/opt/rocm/bin/amdflang")
KB = 8.617343e-5  # eV/K (datafmt.f90:27)


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    if not os.path.exists(FC):
        pytest.skip("no Fortran compiler")
    d = str(tmp_path_factory.mktemp("oxid"))
    for src in (os.path.join(ROOT, "simox_amd", "fortran", "simox_oxid.f90"),
                os.path.join(HERE, "oxid_driver.f90")):
        subprocess.run([FC, "-O2", "-ffp-contract=off", "-c", src, "-module-dir", d], cwd=d,
                       check=True)
    exe = os.path.join(d, "oxid_driver")
    subprocess.run([FC, "-o", exe, "simox_oxid.o", "oxid_driver.o"], cwd=d, check=True)
    return exe, d


def run(driver, ids, x, q, n_o2, seed, z_new, box, t_ev, q_crit):
    exe, d = driver
    fin, fout = os.path.join(d, "in.bin"), os.path.join(d, "out.bin")
    n = len(ids)
    with open(fin, "wb") as fh:
        np.array([n, n_o2, seed, len(q_crit)], dtype=np.int32).tofile(fh)
        np.array([z_new, box[0], box[1], t_ev], dtype=np.float64).tofile(fh)
        np.asarray(q_crit, dtype=np.float64).tofile(fh)
        np.asarray(ids, dtype=np.int32).tofile(fh)
        np.asarray(x, dtype=np.float64).reshape(n, 3).tofile(fh)
        np.asarray(q, dtype=np.float64).tofile(fh)
    subprocess.run([exe, fin, fout], check=True)
    raw = np.fromfile(fout, dtype=np.uint8)
    m = 2 * n_o2
    flags = raw[:4 * len(q_crit)].view(np.int32)
    rest = raw[4 * len(q_crit):].view(np.float64)
    return flags, rest[:3 * m].reshape(m, 3), rest[3 * m:6 * m].reshape(m, 3)


def test_oxide_check_and_new_o2_match_reference(driver):
    g = dict(np.load(os.path.join(HERE, "golden", "oxid_s1.npz"), allow_pickle=False))
    t_ev = float(g["t_given"]) * KB
    seen = set()
    for k in range(int(g["nrec"])):
        n_o2 = int(g["n_o2_%d" % k])
        flags, xo, vo = run(driver, g["ids%d" % k], g["x%d" % k], g["q"], n_o2,
                            int(g["seed%d" % k]), float(g["z_new%d" % k]), g["box"], t_ev,
                            g["q_crit"])
        np.testing.assert_array_equal(flags, g["ref_new_o2_%d" % k])
        np.testing.assert_array_equal(xo, g["ref_o2_x%d" % k])
        np.testing.assert_array_equal(vo, g["ref_o2_v%d" % k])
        assert np.all(g["ref_o2_id%d" % k] == 4) and np.all(g["ref_o2_q%d" % k] == 0.0)
        seen.update(int(v) for v in flags)
    assert seen == {0, 1}  # the cases exercise both outcomes


def test_dimer_geometry(driver):
    """Properties NEW_O2 fixes whatever the random stream: atoms at c +- d
    with d = 2 (sin t cos p, sin t cos p, cos t) (the reference repeats cos p
    in y), centres at z_new and at least sqrt(Lx Ly / n_o2) / 2 apart in x-y."""
    box = np.array([40.0, 40.0, 60.0])
    ids = np.array([1, 4], dtype=np.int32)
    x = np.array([[0.0, 0.0, 1.0], [0.0, 0.0, 5.0]])
    _, xo, vo = run(driver, ids, x, np.zeros(2), 6, 3, 20.0, box, 600 * KB, [0.0])
    a, b = xo[0::2], xo[1::2]
    c = 0.5 * (a + b)
    dd = 0.5 * (a - b)
    # |d_z| = 2 |cos t| <= 2 and d_x^2 = 4 sin^2 t cos^2 p <= 4 - d_z^2
    assert np.all(np.abs(dd[:, 2]) <= 2.0 + 1e-14)
    assert np.all(dd[:, 0] ** 2 <= 4.0 - dd[:, 2] ** 2 + 1e-12)
    np.testing.assert_allclose(c[:, 2], 20.0, rtol=1e-14)
    rmin = np.sqrt(box[0] * box[1] / 6) * 0.5
    d = np.sqrt(((c[:, None, :2] - c[None, :, :2]) ** 2).sum(-1)) + np.eye(6) * 1e9
    assert d.min() >= rmin * (1 - 1e-12)
    # the reference's orientation: x and y offsets equal (cos(phi) twice)
    np.testing.assert_allclose(dd[:, 0], dd[:, 1], rtol=0, atol=1e-13)
    np.testing.assert_allclose(vo[:, 2], -np.sqrt(600 * KB * 3 / 16), rtol=1e-15)
