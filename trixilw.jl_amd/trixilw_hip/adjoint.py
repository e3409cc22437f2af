"""Gradients through the LW time loop with the device adjoint.

``adjoint_step`` is one reverse step, λ = λ' + dt (∂rhs/∂u)ᵀ λ' (tlw_adjoint_lw).
``gradient_lw`` runs solve_lwfr's LW loop (compute_dt -> rhs! -> u += dt du,
src/solvers/lwfr.jl:212-230) forward on the device with checkpoints every
``checkpoint_every`` steps, then sweeps back segment by segment: each segment is
recomputed from its checkpoint and reversed with adjoint steps.  The CFL dts of the forward run are
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


def adjoint_step_mdrk(lam, u, us, lam_next, semi, t, dt):
    """lam: the adjoint of uⁿ for one MDRK step (tlw_adjoint_mdrk), with us the
    step's stage state uⁿ + dt rhs_mdrk1!(uⁿ) (device tensors)."""
    for name, x in (("lam", lam), ("u", u), ("us", us), ("lam_next", lam_next)):
        semi._check_state(x, name)
    semi._bind_stream()
    _lib.check(_lib.lib().tlw_adjoint_mdrk(semi.handle, C.c_void_p(u.data_ptr()), C.c_void_p(us.data_ptr()),
                                           C.c_void_p(lam_next.data_ptr()), C.c_void_p(lam.data_ptr()), float(t),
                                           float(dt)), semi.handle)


def mdrk_stage_state(u, semi, t, dt, du):
    """us = u + dt rhs_mdrk1!(u) (mdrk.jl:13-17), into a new tensor."""
    from .semidiscretization import rhs_mdrk1

    semi.cache["dt"][0] = dt
    rhs_mdrk1(du, u, semi, t)
    return _torch().add(u, du, alpha=dt)


def mdrk_step(u, semi, t, dt, du):
    """One MDRK step: us, then uⁿ⁺¹ = u + dt rhs_mdrk2!(u, us), into a new tensor."""
    from .semidiscretization import rhs_mdrk2

    us = mdrk_stage_state(u, semi, t, dt, du)
    semi.set_us(us)
    rhs_mdrk2(du, u, semi, t)
    return _torch().add(u, du, alpha=dt)


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


def gaussian_random_field(x, y, period, seed, amplitude=1e-2, modes=8, correlation=0.15):
    """Seeded periodic Gaussian random field at the points (x, y): a sum of
    Fourier modes k = (k1, k2), |k1|, |k2| <= modes, with independent N(0, 1)
    cosine and sine coefficients (numpy PCG64 via default_rng(seed)) weighted
    by the Gaussian spectrum exp(-(2π|k| correlation)² / 2) of a field whose
    correlation length is `correlation` times the period, scaled so that its
    standard deviation over the points is `amplitude`."""
    rng = np.random.default_rng(seed)
    Lx, Ly = period
    ks = np.arange(-modes, modes + 1)
    f = np.zeros(np.shape(x))
    for k1 in ks:
        for k2 in ks:
            a, b = rng.standard_normal(2)
            if k1 == 0 and k2 == 0:
                continue
            w = np.exp(-0.5 * (2.0 * np.pi * correlation) ** 2 * (k1 * k1 + k2 * k2))
            ph = 2.0 * np.pi * (k1 * np.asarray(x) / Lx + k2 * np.asarray(y) / Ly)
            f = f + w * (a * np.cos(ph) + b * np.sin(ph))
    return amplitude * f / np.std(f)


def _forward_device(semi, u0, nsteps, cfl, t0, dts, checkpoint_every):
    """solve_lwfr's LW loop on the device (tlw_solve_init / tlw_advance) from
    u0, nsteps steps, the state copied out every checkpoint_every steps; dts
    prescribed or recorded by the device.  Returns (u_K, dts, checkpoints)."""
    torch = _torch()
    L, h = _lib.lib(), semi.handle
    semi._bind_stream()
    dp = C.POINTER(C.c_double)
    if dts is None:
        _lib.check(L.tlw_prescribe_dts(h, None, 0), h)
        _lib.check(L.tlw_record_dts(h, nsteps), h)
    else:
        d = np.ascontiguousarray(dts[:nsteps], dtype=np.float64)
        _lib.check(L.tlw_record_dts(h, 0), h)
        _lib.check(L.tlw_prescribe_dts(h, d.ctypes.data_as(dp), d.size), h)
    try:
        _lib.check(L.tlw_set_state_dev(h, C.c_void_p(u0.data_ptr())), h)
        _lib.check(L.tlw_solve_init(h, float(t0), float("inf"), float(cfl)), h)
        checkpoints = {}
        it = 0
        while True:
            if checkpoint_every and it % checkpoint_every == 0 and it < nsteps:
                c = torch.empty_like(u0)
                _lib.check(L.tlw_get_state_dev(h, C.c_void_p(c.data_ptr())), h)
                checkpoints[it] = c
            if it >= nsteps:
                break
            k = min(checkpoint_every or nsteps, nsteps - it)
            _lib.check(L.tlw_advance(h, k, None, None, None, None), h)
            it += k
        uK = torch.empty_like(u0)
        _lib.check(L.tlw_get_state_dev(h, C.c_void_p(uK.data_ptr())), h)
        if dts is None:
            rec = np.zeros(nsteps)
            _lib.check(L.tlw_get_dts(h, rec.ctypes.data_as(dp), nsteps), h)
            dts = [float(v) for v in rec]
        torch.cuda.current_stream(u0.device).synchronize()
    finally:
        L.tlw_record_dts(h, 0)
        L.tlw_prescribe_dts(h, None, 0)
    return uK, list(dts[:nsteps]), checkpoints


def gradient_lw(semi, u0, nsteps, cfl, objective=kinetic_energy, checkpoint_every=20, t0=0.0, dts=None):
    """(J(u_K), ∇_{u0} J, u_K, dts) for K = nsteps LW steps from u0 (MDRK steps
    on an MDRK semidiscretization, with tlw_adjoint_mdrk).

    The forward sweep is solve_lwfr's loop on the device (no host
    synchronisation inside a checkpoint segment), keeping the state every
    ``checkpoint_every`` steps; ``dts`` (optional) prescribes the steps, by
    default dt_k = cfl · max_dt(u_k) as compute_dt does on the device, and the
    dts used are returned.  The reverse sweep recomputes each segment from its
    checkpoint (rhs! + update, stream-ordered) and applies the adjoint steps."""
    torch = _torch()
    if checkpoint_every < 1:
        raise ValueError("checkpoint_every >= 1")
    if dts is not None and len(dts) < nsteps:
        raise ValueError("dts must hold nsteps values")
    semi._check_state(u0, "u0")
    u, dts, checkpoints = _forward_device(semi, u0, nsteps, cfl, t0, dts, checkpoint_every)
    times = []
    t = float(t0)
    for k in range(nsteps):  # set_t_and_iter!: t accumulates the dts in order
        times.append(t)
        t += dts[k]
    J, lam = objective(semi, u)
    lam_prev = torch.empty_like(lam)
    du = torch.empty_like(u0)
    mdrk = _is_mdrk(semi)
    step = mdrk_step if mdrk else lw_step
    for seg in sorted(checkpoints, reverse=True):
        end = min(seg + checkpoint_every, nsteps)
        states = [checkpoints[seg]]
        for k in range(seg, end - 1):
            states.append(step(states[-1], semi, times[k], dts[k], du))
        for k in range(end - 1, seg - 1, -1):
            if mdrk:
                us = mdrk_stage_state(states[k - seg], semi, times[k], dts[k], du)
                adjoint_step_mdrk(lam_prev, states[k - seg], us, lam, semi, times[k], dts[k])
            else:
                adjoint_step(lam_prev, states[k - seg], lam, semi, times[k], dts[k])
            lam, lam_prev = lam_prev, lam
    return J, lam, u, dts


def _is_mdrk(semi):
    from .semidiscretization import MDRK

    return isinstance(semi.time_discretization, MDRK)


def objective_lw(semi, u0, nsteps, dts, objective=kinetic_energy, t0=0.0):
    """J(u_K) of the device loop with prescribed dts (no gradient): the
    forward map a finite-difference check differentiates."""
    # the CFL number only enters the first (unused) CFL reduction: every dt is prescribed
    u, _, _ = _forward_device(semi, u0, nsteps, 1.0, t0, dts, 0)
    return objective(semi, u)[0]
