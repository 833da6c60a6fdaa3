"""Spline-form EAM and its coefficient gradients on the GPU (build-only).

* energy (Epair + Egroup) against a brute-force numpy evaluation of the same
  spline model (all pairs by minimum image, 3 or more cells per axis so that
  FORCE's pair set is the plain one), 1e-10 relative;
* dE/dcoef against central finite differences of the GPU energy, step 1e-6,
  1e-6 relative above the rounding floor eps |E| / step of the differences;
* forces against -dE/dx by central differences (charges 0, so that only the
  EAM terms act), the same bar;
* torch.autograd.gradcheck of EamSplineEnergy on a few coefficients;
* the fitted spline model stays within 1e-3 of the analytic potential.
"""
import numpy as np
import pytest
import torch

from simox_amd import MODE_FP64, Simox
from simox_amd import eam_spline as es
from simox_amd.systems import tiny

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def system():
    s = tiny(seed=3, nx=5, ny=5, nz=4, n_o=16, fft=(20, 20, 32), rcut=6.0)
    assert (s.box / s.rcut >= 3).all()
    return s


@pytest.fixture(scope="module")
def model(system):
    return es.fit(r_end=system.rcut, m=50)


def numpy_energy(s, lo, h, c):
    n = s.n
    d = s.x[:, None, :] - s.x[None, :, :]
    d -= s.box * np.round(d / s.box)
    r = np.sqrt((d ** 2).sum(-1))
    np.fill_diagonal(r, np.inf)
    ii, jj = np.nonzero(r < s.rcut)
    rr = r[ii, jj]
    si, sj = s.ids[ii], s.ids[jj]
    fpair = np.where((si == 1) & (sj == 1), 3, np.where((si == 4) & (sj == 4), 5, 4))
    fdens = np.where(sj == 1, 0, np.where(si == 1, 1, 2))
    epair = 0.0
    rho = np.zeros(n)
    for f in range(6):
        sel = fpair == f
        epair += 0.5 * es.evaluate(rr[sel], lo[f], h[f], c[f]).sum()
        sel = fdens == f
        np.add.at(rho, ii[sel], es.evaluate(rr[sel], lo[f], h[f], c[f]))
    eg = 0.0
    for sp, f in ((1, 6), (4, 7)):
        sel = s.ids == sp
        eg += es.evaluate(rho[sel], lo[f], h[f], c[f]).sum()
    return epair, eg


def gpu_energy(s, lo, h, c, q=None):
    g = Simox(s.n, s.box, s.fft, s.rcut, mode=MODE_FP64)
    g.set_state(s.ids, s.x, s.v, None, s.q if q is None else q)
    g.set_eam_splines(lo, h, c)
    e = g.force()
    return g, e


def test_spline_energy_matches_numpy(system, model):
    lo, h, c = model
    g, e = gpu_energy(system, lo, h, c)
    ep, eg = numpy_energy(system, lo, h, c)
    assert e["Epair"] == pytest.approx(ep, rel=1e-10)
    assert e["Egroup"] == pytest.approx(eg, rel=1e-10)
    g.close()


def test_spline_model_close_to_analytic(system, model):
    lo, h, c = model
    g = Simox(system.n, system.box, system.fft, system.rcut, mode=MODE_FP64)
    g.set_state(system.ids, system.x, system.v, None, system.q)
    ea = g.force()
    g.set_eam_splines(lo, h, c)
    es_ = g.force()
    g.set_eam_splines(None, None, None)
    eb = g.force()
    assert eb == ea  # switching back restores the analytic forms exactly
    for k in ("Epair", "Egroup"):
        assert es_[k] == pytest.approx(ea[k], rel=1e-3)
    g.close()


def test_spline_gradient_matches_finite_differences(system, model):
    lo, h, c = model
    g, _ = gpu_energy(system, lo, h, c)
    grad = g.eam_spline_grad()
    g.close()
    rng = np.random.default_rng(0)
    picks = []
    for f in range(es.NSPL):  # three coefficients with weight in each function
        nz = np.flatnonzero(np.abs(grad[f]) > 1e-3 * np.abs(grad[f]).max())
        picks += [(f, int(k)) for k in rng.choice(nz, size=min(3, len(nz)), replace=False)]
    step = 1e-6
    for f, k in picks:
        e = []
        for sgn in (1, -1):
            cc = c.copy()
            cc[f, k] += sgn * step
            gg, ee = gpu_energy(system, lo, h, cc)
            gg.close()
            e.append(ee["Epair"] + ee["Egroup"])
        fd = (e[0] - e[1]) / (2 * step)
        # FD floor: the energy is a sum rounded at ~ eps |E|, so the central
        # difference carries ~ eps |E| / step of rounding noise
        floor = np.finfo(float).eps * abs(e[0]) / step
        assert abs(fd - grad[f, k]) <= 1e-6 * abs(grad[f, k]) + floor, (f, k, fd, grad[f, k])


def test_spline_forces_are_minus_energy_gradient(system, model):
    lo, h, c = model
    q0 = np.zeros(system.n)
    g, _ = gpu_energy(system, lo, h, c, q=q0)
    f = g.get_state()[2]
    g.close()
    step = 1e-6
    for i in (0, 37, system.n - 1):
        for d in range(3):
            e = []
            for sgn in (1, -1):
                s2 = system.x.copy()
                s2[i, d] += sgn * step
                gg = Simox(system.n, system.box, system.fft, system.rcut, mode=MODE_FP64)
                gg.set_state(system.ids, s2, system.v, None, q0)
                gg.set_eam_splines(lo, h, c)
                ee = gg.force()
                gg.close()
                e.append(ee["Epair"] + ee["Egroup"])
            fd = -(e[0] - e[1]) / (2 * step)
            floor = np.finfo(float).eps * abs(e[0]) / step
            assert abs(fd - f[i, d]) <= 1e-6 * abs(f[i, d]) + floor, (i, d, fd, f[i, d])


def test_torch_gradcheck_on_some_coefficients(system, model):
    lo, h, c = model
    picks = [(3, 10), (4, 20), (6, 25), (0, 12), (7, 5)]
    base = torch.tensor(c, dtype=torch.float64)

    def energy(v):
        full = base.clone()
        for t, (f, k) in enumerate(picks):
            full[f, k] = v[t]
        return es.EamSplineEnergy.apply(full, system, lo, h)

    v0 = torch.tensor([c[f, k] for f, k in picks], dtype=torch.float64, requires_grad=True)
    assert torch.autograd.gradcheck(energy, (v0,), eps=1e-6, atol=1e-5, rtol=1e-5)


def _charges(system, seed=5):
    q = np.random.default_rng(seed).normal(size=system.n) * 0.2
    return q - q.mean()


def test_force_vjp_matches_finite_differences(system, model):
    """d/dcoef of S = sum_i u_i . F_i(EAM) for a random u against central
    differences of S over GPU forces (fixed charges: the other force terms do
    not depend on the coefficients)."""
    lo, h, c = model
    q = _charges(system)
    u = np.random.default_rng(7).normal(size=(system.n, 3))

    def S(cc):
        g = Simox(system.n, system.box, system.fft, system.rcut, mode=MODE_FP64)
        g.set_state(system.ids, system.x, system.v, None, q)
        g.set_eam_splines(lo, h, cc)
        g.force()
        val = float((u * g.get_state()[2]).sum())
        return val, g

    s0, g = S(c)
    vjp = g.eam_spline_force_vjp(u)
    g.close()
    rng = np.random.default_rng(1)
    step = 1e-6
    for f in range(es.NSPL):
        nz = np.flatnonzero(np.abs(vjp[f]) > 1e-3 * np.abs(vjp[f]).max())
        for k in rng.choice(nz, size=min(3, len(nz)), replace=False):
            vals = []
            for sgn in (1, -1):
                cc = c.copy()
                cc[f, k] += sgn * step
                v, gg = S(cc)
                gg.close()
                vals.append(v)
            fd = (vals[0] - vals[1]) / (2 * step)
            floor = 64 * np.finfo(float).eps * np.abs(u).sum() * 10.0 / step
            assert abs(fd - vjp[f, k]) <= 1e-6 * abs(vjp[f, k]) + floor, (f, k, fd, vjp[f, k])


def test_force_loss_gradient_and_gradcheck(system, model):
    """L = Epair + Egroup + w sum |F|^2 (total forces, fixed charges): the
    library gradient against central differences, and torch gradcheck."""
    lo, h, c = model
    q = _charges(system)
    w = 0.01
    L0, grad = es.loss_and_grad(system, lo, h, c, w, q)
    step = 1e-6
    for f, k in ((3, 20), (4, 15), (0, 22), (1, 18), (6, 24), (7, 8)):
        cp, cm = c.copy(), c.copy()
        cp[f, k] += step
        cm[f, k] -= step
        fd = (es.loss_and_grad(system, lo, h, cp, w, q)[0] -
              es.loss_and_grad(system, lo, h, cm, w, q)[0]) / (2 * step)
        floor = np.finfo(float).eps * abs(L0) / step
        assert abs(fd - grad[f, k]) <= 1e-6 * abs(grad[f, k]) + floor, (f, k, fd, grad[f, k])
    picks = [(3, 10), (1, 12), (6, 20)]
    base = torch.tensor(c, dtype=torch.float64)

    def loss(v):
        full = base.clone()
        for t, (f, k) in enumerate(picks):
            full[f, k] = v[t]
        return es.EamSplineLoss.apply(full, system, lo, h, w, q)

    v0 = torch.tensor([c[f, k] for f, k in picks], dtype=torch.float64, requires_grad=True)
    assert torch.autograd.gradcheck(loss, (v0,), eps=1e-6, atol=1e-5, rtol=1e-5)
