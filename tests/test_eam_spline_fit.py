"""Spline layout and fit of the spline-form EAM (CPU): the basis is a partition
of unity on the valid range, the least-squares fit reproduces the analytic
forms (eam.f90, force.f90:256-344) to 1e-3 of their range."""
import numpy as np

from simox_amd import eam_spline as es


def test_basis_partition_of_unity():
    m, lo, h = 20, 1.5, 0.5
    x = np.linspace(lo, lo + (m - 3) * h - 1e-9, 777)
    B = es.basis(x, lo, h, m)
    np.testing.assert_allclose(B.sum(1), 1.0, rtol=0, atol=1e-14)
    assert (B >= -1e-15).all()
    assert ((B > 0).sum(1) <= 4).all()


def test_basis_reproduces_cubics():
    # cubic B-splines span the cubic polynomials: the fit is exact
    m, lo, h = 16, 0.0, 0.25
    x = np.linspace(lo, lo + (m - 3) * h, 400)
    y = 0.3 - 1.2 * x + 0.5 * x ** 2 - 0.07 * x ** 3
    c = np.linalg.lstsq(es.basis(x, lo, h, m), y, rcond=None)[0]
    np.testing.assert_allclose(es.evaluate(x, lo, h, c), y, atol=1e-12)


def test_fit_reproduces_analytic_forms():
    lo, h, c = es.fit(m=50)
    fns = es.analytic_functions()
    for f in range(es.NSPL):
        x = np.linspace(lo[f], lo[f] + 47 * h[f], 2001)
        y = fns[f](x)
        err = np.abs(es.evaluate(x, lo[f], h[f], c[f]) - y).max()
        assert err <= 1e-3 * np.abs(y).max(), (es.NAMES[f], err)
