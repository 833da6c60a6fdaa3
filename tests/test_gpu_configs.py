"""The named workloads at their own sizes on the GPU.

C2 (77,824 atoms, PME 96x96x192): FP64 forces against the oracle at step 0;
then 50 MIXED steps (sort and charge solve every step) with the force, charge
and energy error against the FP64 path on the same positions every 10 steps:
e_F = max_i |dF_i| / RMS_i |F_i| <= 1e-5, charges 1e-9, energies 1e-6.
C3 (1,048,576 atoms, PME 256x256x512): after 10 MIXED steps, e_F <= 1e-5
against FP64 on the same positions, total charge 0, and two runs bit-equal.
C4 (replicas of 410-4,904 atoms): a seeded subset of the 256 replicas at their
real sizes, batched stepping bit-equal to stepping each context alone.
C5 (5,200-atom snapshot, FP64): gradients of E + w sum |F|^2 in chi, J, xi of
Al and O and in the EAM spline coefficients against central differences with
step 1e-6, to 1e-6 relative above the rounding floor of the differences.
"""
import numpy as np
import pytest

from simox_amd import MODE_FP64, MODE_MIXED, Simox, batch_step
from simox_amd.systems import config, replicas, snapshots

pytestmark = pytest.mark.gpu


def e_F(f, ref):
    rms = np.sqrt((ref ** 2).sum(1).mean())
    return np.sqrt(((f - ref) ** 2).sum(1)).max() / rms


def fp64_reference(s, x, q0, tol=1e-10):
    """FP64 forces, charges and energies at positions x (charges re-solved
    from q0 to tol)."""
    d = Simox(s.n, s.box, s.fft, s.rcut, mode=MODE_FP64)
    d.set_charge_tol(tol, 20000)
    d.set_state(s.ids, x, None, None, q0)
    r = d.ctip_solve()
    e = d.force()
    e["Erec"] = d.pme()
    e["Eself"] = r["Eself"]
    _, _, f, q = d.get_state()
    d.close()
    return f, q, e


def mixed_now(g):
    """Forces, charges and energies of the MIXED context at its current
    positions and charges (the charge solve of the last step)."""
    e = g.force()
    e["Erec"] = g.pme()
    _, _, f, q = g.get_state()
    return f, q, e


@pytest.fixture(scope="module")
def c2():
    return config("c2")


def test_c2_fp64_forces_match_oracle(c2, oracle_mod):
    s = c2
    rng = np.random.default_rng(11)
    q = rng.normal(size=s.n) * 0.2
    q -= q.mean()
    g = Simox(s.n, s.box, s.fft, s.rcut, mode=MODE_FP64)
    g.set_state(s.ids, s.x, s.v, None, q)
    e = g.force()
    erec = g.pme()
    f = g.get_state()[2]
    g.close()
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
    for k in ("Epair", "Egroup", "Edir", "Ectip"):
        assert e[k] == pytest.approx(ref[k], rel=1e-10)


def test_c2_mixed_50_steps_against_fp64(c2):
    s = c2
    g = Simox.from_system(s, mode=MODE_MIXED)
    g.set_charge_tol(1e-10, 5000)
    g.ctip_solve()
    g.force()
    g.pme()
    worst = 0.0
    for block in range(6):
        if block:
            g.step(10, s.dt, 1, 1, stats=False)
        x = g.get_state()[0]
        f_m, q_m, e_m = mixed_now(g)
        f_d, q_d, e_d = fp64_reference(s, x, np.zeros(s.n) if block == 0 else q_m)
        err = e_F(f_m, f_d)
        worst = max(worst, err)
        print("C2 step %d: e_F %.2e, charges %.2e" % (10 * block, err, np.abs(q_m - q_d).max()),
              flush=True)
        assert err <= 1e-5, (10 * block, err)
        assert np.abs(q_m - q_d).max() <= 1e-9
        assert abs(q_m.sum()) < 1e-9
        for k in ("Epair", "Egroup", "Edir", "Ectip", "Erec"):
            assert e_m[k] == pytest.approx(e_d[k], rel=1e-6, abs=1e-6), (k, 10 * block)
    g.close()
    assert worst > 0.0


def test_c3_ten_mixed_steps():
    s = config("c3")
    runs = []
    for rep in range(2):
        g = Simox.from_system(s, mode=MODE_MIXED)
        g.set_charge_tol(1e-10, 5000)
        g.force()
        g.pme()
        st = g.step(10, s.dt, 1, 1)
        x, v, f, q = g.get_state()
        runs.append((x, v, f, q, st, g.charge_stats()))
        if rep == 0:
            f_m, q_m, _ = mixed_now(g)
        g.close()
    (x, v, f, q, st, cs), (x2, v2, f2, q2, st2, cs2) = runs
    # determinism: the same ten steps give the same state bit for bit
    for a, b in ((x, x2), (v, v2), (f, f2), (q, q2)):
        np.testing.assert_array_equal(a, b)
    assert st == st2 and cs == cs2
    assert abs(q.sum()) < 1e-8
    f_d, q_d, _ = fp64_reference(s, x, q)
    err = e_F(f_m, f_d)
    print("C3 after 10 steps: e_F %.2e, charges %.2e" % (err, np.abs(q_m - q_d).max()))
    assert err <= 1e-5
    assert np.abs(q_m - q_d).max() <= 1e-8


def test_c4_replicas_batched_equal_single():
    rs = replicas(256)[::8]  # 32 of the 256 replicas, at their own sizes
    assert max(s.n for s in rs) > 3000 and min(s.n for s in rs) < 1500

    def contexts():
        out = []
        for s in rs:
            g = Simox.from_system(s, mode=MODE_MIXED)
            g.set_charge_tol(1e-10, 5000)
            g.force()
            g.pme()
            out.append(g)
        return out

    a, b = contexts(), contexts()
    dt = rs[0].dt
    ref = [g.step(2, dt, 1, 1) for g in a]
    got = batch_step(b, 2, dt, 1, 1)
    assert ref == got
    for ga, gb in zip(a, b):
        for u, w in zip(ga.get_state(), gb.get_state()):
            np.testing.assert_array_equal(u, w)
        assert ga.charge_stats() == gb.charge_stats()
    for g in a + b:
        g.close()


@pytest.fixture(scope="module")
def c5():
    """The first C5 snapshot with its converged FP64 charges."""
    s = snapshots(1)[0]
    g = Simox.from_system(s, mode=MODE_FP64)
    g.set_charge_tol(1e-12, 20000)
    g.ctip_solve()
    s.q = g.get_state()[3]
    g.close()
    return s


def test_c5_ctip_parameter_gradients(c5):
    from simox_amd.autograd import potential_setup, solve_loss, theta_of
    s = c5
    t0 = theta_of(potential_setup(), with_xi=True)
    w = 0.01
    L0, grad = solve_loss(s, t0, w, tol=1e-12)
    h = 1e-6
    for k in (0, 3, 5, 8, 10, 13):  # chi, J, xi of Al and O
        tp, tm = t0.copy(), t0.copy()
        tp[k] += h
        tm[k] -= h
        lp = solve_loss(s, tp, w, tol=1e-12)[0]
        lm = solve_loss(s, tm, w, tol=1e-12)[0]
        fd = (lp - lm) / (2 * h)
        floor = 8 * np.finfo(float).eps * abs(L0) * np.sqrt(s.n) / h
        print("C5 theta[%d]: adjoint %.10e  fd %.10e  floor %.1e" % (k, grad[k], fd, floor))
        assert abs(fd - grad[k]) <= 1e-6 * abs(grad[k]) + floor, (k, grad[k], fd)
    assert np.all(grad[[1, 2, 4, 6, 7, 9, 11, 12, 14]] == 0.0)


def test_c5_spline_coefficient_gradients(c5):
    from simox_amd import eam_spline as es
    s = c5
    lo, hh, coef = es.fit(r_end=10.0, m=50)
    w = 0.01
    L0, grad = es.loss_and_grad(s, lo, hh, coef, w)
    h = 1e-6
    rng = np.random.default_rng(5)
    # the largest gradient of each function and three random coefficients more
    picks = [(f, int(np.argmax(np.abs(grad[f])))) for f in range(8)]
    picks += [(int(f), int(k)) for f, k in zip(rng.integers(0, 8, 16), rng.integers(0, 50, 16))]
    for f, k in picks:
        cp, cm = coef.copy(), coef.copy()
        cp[f, k] += h
        cm[f, k] -= h
        fd = (es.loss_and_grad(s, lo, hh, cp, w)[0] - es.loss_and_grad(s, lo, hh, cm, w)[0]) / (2 * h)
        floor = 8 * np.finfo(float).eps * abs(L0) * np.sqrt(s.n) / h
        assert abs(fd - grad[f, k]) <= 1e-6 * abs(grad[f, k]) + floor, (f, k, grad[f, k], fd)
