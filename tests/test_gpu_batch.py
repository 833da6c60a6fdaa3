"""Batched replicas (build-only): simox_batch_step advances independent
systems in lockstep, with the charge-solve iterations replayed from captured
HIP graphs.  The results must equal stepping each context on its own, bit for
bit, over several steps (so over many graph replays)."""
import numpy as np
import pytest

from simox_amd import MODE_FP64, MODE_MIXED, Simox, batch_step
from simox_amd.systems import tiny

pytestmark = pytest.mark.gpu


def _replicas():
    out = []
    specs = [(1, 4, 4, 4, 16, (16, 16, 32), 6.0, MODE_MIXED),
             (2, 3, 3, 4, 8, (12, 12, 24), 5.0, MODE_FP64),
             (3, 5, 4, 3, 12, (24, 16, 24), 6.0, MODE_MIXED),
             (4, 3, 5, 3, 10, (12, 24, 24), 5.0, MODE_MIXED)]
    for seed, nx, ny, nz, no, fft, rcut, mode in specs:
        s = tiny(seed=seed, nx=nx, ny=ny, nz=nz, n_o=no, fft=fft, rcut=rcut)
        g = Simox(s.n, s.box, s.fft, s.rcut, mode=mode)
        g.set_charge_tol(1e-10, 3000)
        g.set_state(s.ids, s.x, s.v, None, np.zeros(s.n))
        g.force()
        g.pme()
        out.append((s, g))
    return out


def test_batch_equals_single_steps():
    a = _replicas()
    b = _replicas()
    dt = a[0][0].dt
    ref = [g.step(4, dt, 1, 1) for _, g in a]
    got = batch_step([g for _, g in b], 4, dt, 1, 1)
    for (_, ga), (_, gb), ra, rb in zip(a, b, ref, got):
        for u, v in zip(ga.get_state(), gb.get_state()):
            np.testing.assert_array_equal(u, v)
        assert ra == rb
        assert ga.charge_stats() == gb.charge_stats()
    # again, reusing the captured graphs
    ref = [g.step(2, dt, 1, 1) for _, g in a]
    got = batch_step([g for _, g in b], 2, dt, 1, 1)
    assert ref == got
