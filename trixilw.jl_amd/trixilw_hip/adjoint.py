"""Gradients through the LW time loop with the device adjoint.

``adjoint_step`` is one reverse step, λ = λ' + dt (∂rhs/∂u)ᵀ λ' (tlw_adjoint_lw).
``gradient_lw`` runs solve_lwfr's LW loop (compute_dt -> rhs! -> u += dt du,
src/solvers/lwfr.jl:212-230) forward with checkpoints every ``checkpoint_every``
steps, then sweeps back segment by segment: each segment is recomputed from its
checkpoint and reversed with adjoint steps.  The CFL dts of the forward run are
frozen, so the gradient is that of the fixed-dt map u0 -> u_K (the max in
compute_dt is not differentiable); the dts are returned for the same reason.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .semidiscretization import rhs


def _torch():
    import torch
    return torch


def adjoint_step(lam, u, lam_next, semi, t, dt):
    """lam = lam_next + dt (∂rhs/∂u)ᵀ lam_next at the state u (device tensors)."""
    for name, x in (("lam", lam), ("u", u), ("lam_next", lam_next)):
        semi._check_state(x, name)
    semi._bind_stream()
    _lib.check(_lib.lib().tlw_adjoint_lw(semi.handle, C.c_void_p(u.data_ptr()), C.c_void_p(lam_next.data_ptr()),
                                         C.c_void_p(lam.data_ptr()), float(t), float(dt)), semi.handle)


def lw_step(u, semi, t, dt, du):
    """u + dt rhs!(u) with cache.dt = dt, into a new tensor."""
    semi.cache["dt"][0] = dt
    rhs(du, u, semi, t)
    return _torch().add(u, du, alpha=dt)


def kinetic_energy(semi, u):
    """Σ_e Σ_ij w_i w_j |J_e| ½ (ρv1² + ρv2²)/ρ and its gradient with respect to u."""
    torch = _torch()
    w = semi.solver.basis.weights
    W = torch.from_numpy(np.outer(w, w)).to(u.device)
    jac = torch.from_numpy(1.0 / np.asarray(semi.mesh.inverse_jacobian, dtype=np.float64)).to(u.device)
    scale = W[None, :, :] * (jac ** 2)[:, None, None]  # 2-D: |J| = (h/2)²
    rho, m1, m2 = u[..., 0], u[..., 1], u[..., 2]
    ke = 0.5 * (m1 * m1 + m2 * m2) / rho
    J = float((scale * ke).sum())
    g = torch.zeros_like(u)
    g[..., 0] = scale * (-ke / rho)
    g[..., 1] = scale * (m1 / rho)
    g[..., 2] = scale * (m2 / rho)
    return J, g


def gradient_lw(semi, u0, nsteps, cfl, objective=kinetic_energy, checkpoint_every=20, t0=0.0, dts=None):
    """(J(u_K), ∇_{u0} J, u_K, dts) for K = nsteps LW steps from u0.

    ``dts`` (optional) prescribes the steps; by default dt_k = cfl · max_dt(u_k)
    as compute_dt does, and the dts used are returned."""
    torch = _torch()
    if checkpoint_every < 1:
        raise ValueError("checkpoint_every >= 1")
    du = torch.empty_like(u0)
    checkpoints = {}
    dts = list(dts) if dts is not None else []
    times = []
    u, t = u0.clone(), float(t0)
    for k in range(nsteps):
        if k % checkpoint_every == 0:
            checkpoints[k] = u.clone()
        if k >= len(dts):
            dts.append(cfl * semi.max_dt(u))
        times.append(t)
        u = lw_step(u, semi, t, dts[k], du)
        t += dts[k]
    J, lam = objective(semi, u)
    lam_prev = torch.empty_like(lam)
    for seg in sorted(checkpoints, reverse=True):
        end = min(seg + checkpoint_every, nsteps)
        states = [checkpoints[seg]]
        for k in range(seg, end - 1):
            states.append(lw_step(states[-1], semi, times[k], dts[k], du))
        for k in range(end - 1, seg - 1, -1):
            adjoint_step(lam_prev, states[k - seg], lam, semi, times[k], dts[k])
            lam, lam_prev = lam_prev, lam
    return J, lam, u, dts
