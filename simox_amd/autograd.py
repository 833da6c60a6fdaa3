"""CTIP parameter gradients as a torch.autograd.Function, with a thin
gradcheck / fitting harness (build-only; the reference has no derivatives).

E(theta) = Eself + Edir + Erec + Ectip of the MD step at the charges the
converged solve returns, theta = (chi_1..5, J_1..5) for Al, Ni, Fe, O, H, and
optionally (xi_1..5), the Slater exponents (zeta), with the [fa|fb]
coefficients recomputed from them (datafmt.f90:201-220).
The backward pass uses the library's adjoint solve through the charge
equilibration (simox_param_grad); torch is only the autograd plumbing here.
"""
import ctypes as C

import numpy as np
import torch

from .api import MODE_FP64, Simox, load_library, potential_setup


def params_with(theta, base=None):
    """A copy of the potential parameters with chi, J (and, for 15 entries,
    xi with the [fa|fb] coefficients that follow from it) replaced by theta."""
    p = potential_setup() if base is None else _copy(base)
    t = np.asarray(theta, dtype=np.float64)
    for s in range(5):
        p.prm[s].chi = float(t[s])
        p.prm[s].J = float(t[5 + s])
    if len(t) == 15:
        for s in range(5):
            p.prm[s].xi = float(t[10 + s])
        load_library().simox_fafb(C.byref(p))
    return p


def theta_of(params, with_xi=False):
    t = [params.prm[s].chi for s in range(5)] + [params.prm[s].J for s in range(5)]
    if with_xi:
        t += [params.prm[s].xi for s in range(5)]
    return np.array(t)


def _copy(p):
    q = type(p)()
    C.memmove(C.byref(q), C.byref(p), C.sizeof(p))
    return q


def solve_energy(system, theta, mode=MODE_FP64, tol=1e-12, q0=None):
    """(E, context) at theta: charges solved from q0 (default 0) to tol."""
    s = system
    g = Simox(s.n, s.box, s.fft, s.rcut, mode=mode, params=params_with(theta))
    g.set_charge_tol(tol, 20000)
    g.set_state(s.ids, s.x, s.v, None, np.zeros(s.n) if q0 is None else q0)
    r = g.ctip_solve()
    e = g.force()
    erec = g.pme()
    return r["Eself"] + e["Edir"] + erec + e["Ectip"], g


class CtipEnergy(torch.autograd.Function):
    """theta (float64 tensor [10], or [15] with xi) -> E (scalar tensor) for a
    fixed system."""

    @staticmethod
    def forward(ctx, theta, system, mode=MODE_FP64, tol=1e-12):
        th = theta.detach().cpu().numpy()
        E, g = solve_energy(system, th, mode, tol)
        gr = g.param_grad(with_xi=len(th) == 15)
        g.close()
        parts = [gr["dchi"], gr["dJ"]] + ([gr["dxi"]] if len(th) == 15 else [])
        ctx.save_for_backward(torch.tensor(np.concatenate(parts), dtype=theta.dtype))
        return torch.tensor(E, dtype=theta.dtype)

    @staticmethod
    def backward(ctx, grad_out):
        (grad,) = ctx.saved_tensors
        return grad_out * grad, None, None, None


def solve_loss(system, theta, w, mode=MODE_FP64, tol=1e-12):
    """(L, dL/dtheta) for L = E + w sum_i |F_i|^2 at the charges of a
    converged solve, theta = (chi, J, xi) of the five species (15 entries)."""
    E, g = solve_energy(system, theta, mode, tol)
    f = g.get_state()[2]
    L = E + w * float((f ** 2).sum())
    ge = g.param_grad(with_xi=True)
    gf = g.param_force_grad()
    g.close()
    grad = np.concatenate([ge["dchi"], ge["dJ"], ge["dxi"]]) + 2.0 * w * np.concatenate(
        [gf["dchi"], gf["dJ"], gf["dxi"]])
    return L, grad


class CtipLoss(torch.autograd.Function):
    """theta (float64 tensor [15]) -> L = E + w sum |F|^2 for a fixed system."""

    @staticmethod
    def forward(ctx, theta, system, w, mode=MODE_FP64, tol=1e-12):
        L, grad = solve_loss(system, theta.detach().cpu().numpy(), w, mode, tol)
        ctx.save_for_backward(torch.tensor(grad, dtype=theta.dtype))
        return torch.tensor(L, dtype=theta.dtype)

    @staticmethod
    def backward(ctx, grad_out):
        (grad,) = ctx.saved_tensors
        return grad_out * grad, None, None, None, None


def gradcheck(system, theta=None, eps=1e-5, atol=1e-6, rtol=1e-4):
    """torch.autograd.gradcheck of CtipEnergy (central differences of E)."""
    t0 = theta_of(potential_setup()) if theta is None else np.asarray(theta)
    t = torch.tensor(t0, dtype=torch.float64, requires_grad=True)
    return torch.autograd.gradcheck(lambda x: CtipEnergy.apply(x, system), (t,), eps=eps,
                                    atol=atol, rtol=rtol)


def fit(system, target, theta=None, mask=None, steps=10):
    """Gauss-Newton fit of E(theta) to a target energy over the entries of
    theta selected by mask (the gradient comes from CtipEnergy's backward);
    returns (theta, history of (E - target)^2)."""
    t0 = theta_of(potential_setup()) if theta is None else np.asarray(theta, dtype=np.float64)
    m = torch.ones(10, dtype=torch.float64) if mask is None else torch.tensor(mask, dtype=torch.float64)
    t = torch.tensor(t0, dtype=torch.float64, requires_grad=True)
    hist = []
    for _ in range(steps):
        resid = CtipEnergy.apply(t, system) - target
        resid.backward()
        with torch.no_grad():
            g = m * t.grad
            gg = float(g @ g)
            if gg > 0.0:
                t -= float(resid) * g / gg
            t.grad.zero_()
        hist.append(float(resid.detach()) ** 2)
    return t.detach().numpy(), hist
