"""EAM in spline form (build-only) and gradients of the EAM energy with respect
to the spline coefficients, over the library's C-ABI (simox_set_eam_splines,
simox_eam_spline_grad), with a torch.autograd.Function for fitting.

The reference's EAM is analytic (eam.f90).  Here its eight Al/O functions are
written as uniform cubic B-splines of m coefficients each (default 50):

  0 density of an Al source       f(r) of METAL_HOMO / METAL_OXYGEN (eam.f90:3-35,171-225)
  1 density of an O source into Al   O density with r_e of the Al row (eam.f90:196-222, Q15)
  2 density of an O source into O    O density with r_e of the O row (eam.f90:297-325)
  3 pair term Al-Al  4 Al-O  5 O-O   (eam.f90:3-35, 171-328)
  6 embedding F of Al                (force.f90:256-296)
  7 embedding F of O                 (force.f90:297-344)

`fit` samples the analytic forms (restated below from the potential
parameters) and solves for the coefficients by least squares, so the spline
model starts as the reference's potential; the coefficients are then free
parameters.  torch is only the autograd plumbing.
"""
import ctypes as C

import numpy as np
import torch

from .api import MODE_FP64, Simox, potential_setup

NSPL = 8
NAMES = ("dens_Al", "dens_O_into_Al", "dens_O_into_O", "pair_AlAl", "pair_AlO", "pair_OO",
         "embed_Al", "embed_O")


def _zjw_density(r, re, fe, b, lam):
    x = r / re
    return fe * np.exp(-b * (x - 1.0)) / (1.0 + (x - lam) ** 20)


def _zjw_pair(r, re, a, b, A, B, kap, lam):
    x = r / re
    return (A * np.exp(-a * (x - 1.0)) / (1.0 + (x - kap) ** 20)
            - B * np.exp(-b * (x - 1.0)) / (1.0 + (x - lam) ** 20))


def _poly3(c, t):
    return c[0] + t * (c[1] + t * (c[2] + t * c[3]))


def _embed_metal(m, rho):
    """Metal embedding, three cubic pieces around rho_e then the universal form
    (force.f90:262-296)."""
    rho = np.asarray(rho, dtype=np.float64)
    rn, re, r0 = 0.85 * m.rho_e, m.rho_e, 1.15 * m.rho_e
    out = np.empty_like(rho)
    a = rho < rn
    out[a] = _poly3(list(m.Fn), rho[a] / rn - 1.0)
    b = (~a) & (rho < re)
    out[b] = _poly3(list(m.Fm), rho[b] / re - 1.0)
    c = (~a) & (~b) & (rho < r0)
    out[c] = _poly3(list(m.Fp), rho[c] / re - 1.0)
    d = ~(a | b | c)
    t = (rho[d] / m.rho_s) ** m.eta
    out[d] = m.Fe * (1.0 - np.log(t)) * t
    return out


def _embed_oxygen(o, rho):
    """O embedding: cubic pieces on the rmax intervals (force.f90:297-344)."""
    rho = np.asarray(rho, dtype=np.float64)
    out = np.empty_like(rho)
    k = np.full(rho.shape, 4)
    for s in range(3, -1, -1):
        k[rho < o.rmax[s]] = s
    for s in range(5):
        sel = k == s
        out[sel] = _poly3([o.F[s][j] for j in range(4)], rho[sel] / o.re[s] - 1.0)
    return out


def analytic_functions(params=None):
    """The eight functions of the spline layout, as callables of r or rho."""
    p = potential_setup() if params is None else params
    al, o = p.eam[0], p.ox
    return [
        lambda r: _zjw_density(r, al.r_e, al.f_e, al.beta, al.lambda_),
        lambda r: _zjw_density(r, o.r_e[0], o.fe, o.gamm, o.nu),
        lambda r: _zjw_density(r, o.r_e[3], o.fe, o.gamm, o.nu),
        lambda r: _zjw_pair(r, al.r_e, al.alpha, al.beta, al.A, al.B, al.kappa, al.lambda_),
        lambda r: _zjw_pair(r, o.r_e[0], o.alph[0], o.beta[0], o.A[0], o.B[0], o.kapp[0], o.lamb[0]),
        lambda r: _zjw_pair(r, o.r_e[3], o.alph[3], o.beta[3], o.A[3], o.B[3], o.kapp[3], o.lamb[3]),
        lambda x: _embed_metal(al, x),
        lambda x: _embed_oxygen(o, x),
    ]


def basis(x, lo, h, m):
    """Cubic B-spline basis matrix [len(x), m] on the spline layout (the
    device's spl_basis: u = (x - lo)/h + 1, coefficients int(u)-1 .. int(u)+2)."""
    x = np.asarray(x, dtype=np.float64)
    u = np.clip((x - lo) / h + 1.0, 1.0, m - 2 - 1e-6)
    k0 = u.astype(int)
    t = u - k0
    w = np.stack([(1 - t) ** 3 / 6, (3 * t ** 3 - 6 * t ** 2 + 4) / 6,
                  (-3 * t ** 3 + 3 * t ** 2 + 3 * t + 1) / 6, t ** 3 / 6], 1)
    B = np.zeros((len(x), m))
    for q in range(4):
        np.add.at(B, (np.arange(len(x)), k0 - 1 + q), w[:, q])
    return B


def evaluate(x, lo, h, c):
    return basis(x, lo, h, len(c)) @ c


def fit(r_min=1.5, r_end=10.0, rho_end=None, m=50, params=None, samples=40):
    """Least-squares B-spline coefficients of the analytic functions.
    Radial functions on [r_min, r_end] (r_end: the EAM cutoff), embeddings on
    [0, rho_end] (default 2 x the Al rho_e).  Returns (lo[8], h[8], coef[8, m])."""
    p = potential_setup() if params is None else params
    rho_end = 2.0 * p.eam[0].rho_e if rho_end is None else rho_end
    fns = analytic_functions(p)
    lo = np.zeros(NSPL)
    h = np.zeros(NSPL)
    coef = np.zeros((NSPL, m))
    for f in range(NSPL):
        a, b = (r_min, r_end) if f < 6 else (0.0, rho_end)
        lo[f] = a
        h[f] = (b - a) / (m - 3)
        x = np.linspace(a, b, samples * (m - 3) + 1)
        coef[f] = np.linalg.lstsq(basis(x, a, h[f], m), fns[f](x), rcond=None)[0]
    return lo, h, coef


def spline_system(system, lo, h, coef, mode=MODE_FP64):
    """A context for system with the spline EAM set (charges as in system)."""
    s = system
    g = Simox(s.n, s.box, s.fft, s.rcut, mode=mode)
    g.set_state(s.ids, s.x, s.v, None, s.q)
    g.set_eam_splines(lo, h, coef)
    return g


class EamSplineEnergy(torch.autograd.Function):
    """coef (float64 tensor [8, m]) -> E = Epair + Egroup of a fixed system
    (spline EAM); backward from the library's dE/dcoef."""

    @staticmethod
    def forward(ctx, coef, system, lo, h):
        g = spline_system(system, lo, h, coef.detach().cpu().numpy())
        e = g.force()
        grad = g.eam_spline_grad()
        g.close()
        ctx.save_for_backward(torch.tensor(grad, dtype=coef.dtype))
        return torch.tensor(e["Epair"] + e["Egroup"], dtype=coef.dtype)

    @staticmethod
    def backward(ctx, grad_out):
        (grad,) = ctx.saved_tensors
        return grad_out * grad, None, None, None


def loss_and_grad(system, lo, h, coef, w, q=None, mode=MODE_FP64):
    """L = Epair + Egroup + w sum_i |F_i|^2 at fixed charges (system.q or q),
    F the total forces of FORCE + PME, and dL/dcoef [8, m] =
    dE/dcoef + 2 w (vector-Jacobian product of the EAM forces with F)."""
    s = system
    g = Simox(s.n, s.box, s.fft, s.rcut, mode=mode)
    g.set_state(s.ids, s.x, s.v, None, s.q if q is None else q)
    g.set_eam_splines(lo, h, coef)
    e = g.force()
    g.pme()
    f = g.get_state()[2]
    loss = e["Epair"] + e["Egroup"] + w * float((f ** 2).sum())
    grad = g.eam_spline_grad() + 2.0 * w * g.eam_spline_force_vjp()
    g.close()
    return loss, grad


class EamSplineLoss(torch.autograd.Function):
    """coef (float64 tensor [8, m]) -> L = Epair + Egroup + w sum |F|^2 of a
    fixed system at fixed charges; backward from the library's gradients."""

    @staticmethod
    def forward(ctx, coef, system, lo, h, w, q=None):
        loss, grad = loss_and_grad(system, lo, h, coef.detach().cpu().numpy(), w, q)
        ctx.save_for_backward(torch.tensor(grad, dtype=coef.dtype))
        return torch.tensor(loss, dtype=coef.dtype)

    @staticmethod
    def backward(ctx, grad_out):
        (grad,) = ctx.saved_tensors
        return grad_out * grad, None, None, None, None, None
