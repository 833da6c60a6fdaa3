"""Ni/Fe cross pairs on the GPU: METAL_HETERO (eam.f90:38-98) and AL_NI
(eam.f90:101-168) against the reference build (golden alloy_s1: an Al slab
with 60 Ni and 40 Fe substitutions and 16 O, written by the reference's own
FORCE).  AL_NI hands each atom of an Al-Ni pair the density of the
reference's loop orientation (force.f90:60-63), so the densities pin that
orientation too.  Bars as the other FAITHFUL parity tests: forces and energies
1e-10 relative; MIXED against FP64 e_F <= 1e-5.

The device evaluates each pair once from its forward list entry, which
carries FORCE's orientation (simox_amd/csrc/tile.hpp; the list's order and
indexing are checked on the CPU over alloy_s1 by tests/test_tile_cpu.py)."""
import numpy as np
import pytest

from simox_amd import MODE_FAITHFUL, MODE_FP64, MODE_MIXED

from test_gpu_parity import gpu_for, load, oracle_for, relmax

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def alloy():
    return load("alloy_s1")


def test_alloy_force_matches_reference_build(alloy):
    d = alloy
    assert set(np.unique(d["ids"])) == {1, 2, 3, 4}
    g = gpu_for(d, MODE_FAITHFUL)
    e = g.force()
    f = g.get_state()[2]
    assert relmax(f, d["ref_f"]) <= 1e-10
    got = np.array([e["Epair"], e["Edir"], e["Ectip"], e["Egroup"]])
    np.testing.assert_allclose(got, d["ref_E"][:4], rtol=1e-10, atol=1e-10)


def test_alloy_densities_match_oracle(alloy, oracle_mod):
    d = alloy
    g = gpu_for(d, MODE_FAITHFUL)
    g.force()
    rho = g.rho()
    o = oracle_for(oracle_mod, d)
    ref = o.force(d["ids"], d["x"], d["q"])
    np.testing.assert_allclose(rho, ref["rho"], rtol=1e-12, atol=1e-12)


def test_alloy_mixed_matches_fp64(alloy):
    d = alloy
    out = {}
    for mode in (MODE_FP64, MODE_MIXED):
        g = gpu_for(d, mode)
        g.set_charge_tol(1e-10, 3000)
        g.ctip_solve()
        g.force()
        g.pme()
        out[mode] = g.get_state()[2]
    ref = out[MODE_FP64]
    rms = np.sqrt((ref ** 2).sum(1).mean())
    assert np.sqrt(((out[MODE_MIXED] - ref) ** 2).sum(1)).max() / rms <= 1e-5

