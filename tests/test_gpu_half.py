"""The half-shell force pass (SIMOX_FORCE_HALF=1): each pair of FORCE's loop
evaluated once, from its forward list entry (the reference's orientation),
with both atoms' shares summed in 64-bit fixed point (2^40 per eV/A).

Against the default owner-computes pass and the reference build: FAITHFUL
forces within 1e-10 relative of the reference build's (the golden cases,
Ni/Fe included), energies 1e-10; MIXED forces within 1e-5 of the default
pass (max|dF| / RMS|F|); and the fixed-point sums give the same forces to
the last bit when the pass is repeated."""
import os

import numpy as np
import pytest

from simox_amd import MODE_FAITHFUL, MODE_MIXED

from test_gpu_parity import gpu_for, load, relmax

pytestmark = pytest.mark.gpu


def with_half(d, mode, half):
    old = os.environ.get("SIMOX_FORCE_HALF")
    os.environ["SIMOX_FORCE_HALF"] = "1" if half else "0"
    try:
        return gpu_for(d, mode)
    finally:
        if old is None:
            del os.environ["SIMOX_FORCE_HALF"]
        else:
            os.environ["SIMOX_FORCE_HALF"] = old


@pytest.mark.parametrize("case", ["tiny_s1", "tiny_s2", "c1_s1", "alloy_s1"])
def test_half_shell_matches_reference_build(case):
    d = load(case)
    print("case", case, flush=True)
    g = with_half(d, MODE_FAITHFUL, True)
    e = g.force()
    f = g.get_state()[2]
    assert relmax(f, d["ref_f"]) <= 1e-10
    got = np.array([e["Epair"], e["Edir"], e["Ectip"], e["Egroup"]])
    np.testing.assert_allclose(got, d["ref_E"][:4], rtol=1e-10, atol=1e-10)
    g.force()
    assert np.array_equal(g.get_state()[2], f)  # fixed-point sums: same bits
    g.close()


@pytest.mark.parametrize("case", ["c1_s1", "alloy_s1"])
def test_half_shell_mixed_matches_default(case):
    d = load(case)
    out = {}
    for half in (False, True):
        g = with_half(d, MODE_MIXED, half)
        e = g.force()
        out[half] = (g.get_state()[2], e)
        g.close()
    f0, e0 = out[False]
    f1, e1 = out[True]
    rms = np.sqrt((f0 ** 2).sum(1).mean())
    assert np.sqrt(((f1 - f0) ** 2).sum(1)).max() / rms <= 1e-5
    for k in ("Epair", "Edir", "Ectip", "Egroup"):
        assert e1[k] == pytest.approx(e0[k], rel=1e-6, abs=1e-8), k
