"""Host mirror of TrixiLW's LW/MDRK time loop (src/solvers/lwfr.jl, src/solvers/mdrk.jl).

Two drivers with the reference's semantics:
  * ``perform_step`` replays ``perform_step!`` step by step through the C ABI
    rhs entry points (compute_dt -> set_dt! -> uprev = u -> rhs! ->
    u = uprev + dt du -> set_t_and_iter!), with callbacks and limiters
    applied between steps as ``solve_lwfr`` does (lwfr.jl:233-263);
  * without callbacks (and with at most the positivity limiter),
    ``solve_lwfr`` runs the same loop on the device (tlw_solve_init /
    tlw_advance): dt, the truncation at the final time, t, the limiter and
    the isfinished test all live on the GPU and steps are launched in chunks
    without host synchronisation.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from .limiters import PositivityPreservingLimiterZhangShu
from .semidiscretization import (LW, MDRK, Adaptive, CFLBased, LWUpdate, rhs, rhs_mdrk1, rhs_mdrk2,
                                 get_time_discretization)

RTOL_ISAPPROX = 1.4901161193847656e-08  # sqrt(eps(Float64)), Julia's default for ≈


def isapprox(x: float, y: float) -> bool:
    """Julia's x ≈ y for Float64 scalars (atol = 0, rtol = sqrt(eps))."""
    if x == y:
        return True
    if not (math.isfinite(x) and math.isfinite(y)):
        return False
    return abs(x - y) <= RTOL_ISAPPROX * max(abs(x), abs(y))


@dataclass
class Stats:
    naccept: int = 0
    nreject: int = 0


@dataclass
class LWSolution:
    u: object
    prob: LWUpdate
    t: float
    stats: Stats
    dts: list = field(default_factory=list)


class LWIntegrator:
    """LWIntegrator(lw_update, time_discretization, ...) (lwfr.jl:31-174)."""

    def __init__(self, lw_update: LWUpdate, time_discretization, tolerances=None, time_step_computation=None):
        torch = __import__("torch")
        self.p = lw_update.p
        self.u = lw_update.u0.clone()
        self.uprev = torch.empty_like(self.u)
        self.cache = (lw_update.soln_arrays["du_ode"],)
        self.iter = 0
        self.tspan = lw_update.tspan
        self.t = lw_update.tspan[0]
        self.dt = 0.0
        self.dtpropose = 0.0
        self.dtcache = 0.0
        self.stats = Stats()
        self.alg = time_discretization
        self.tolerances = tolerances or {"abstol": 0.0, "reltol": 0.0}
        self.time_step_computation = time_step_computation
        self.dts = []


def set_dt(integrator: LWIntegrator, dt: float) -> float:
    """set_dt! (lwfr.jl:76-90): truncate at the final time, store in semi.cache.dt."""
    t, Tf = integrator.t, integrator.tspan[1]
    if t + dt > Tf:
        dt = Tf - t
    integrator.p.cache["dt"][0] = dt
    integrator.dtcache = integrator.dtpropose = integrator.dt = dt
    return dt


def set_t_and_iter(integrator: LWIntegrator, dt: float) -> None:
    """set_t_and_iter! (lwfr.jl:92-103)."""
    t, Tf = integrator.t, integrator.tspan[1]
    if t + dt > Tf:
        dt = Tf - t
    integrator.t += dt
    integrator.iter += 1
    integrator.stats.naccept += 1


def isfinished(integrator: LWIntegrator) -> bool:
    return isapprox(integrator.t, integrator.tspan[1])


def compute_dt(semi, mesh, time_step_computation: CFLBased, integrator: LWIntegrator) -> float:
    """compute_dt (lwfr.jl:176-190): cfl_number * Trixi max_dt(u)."""
    return time_step_computation.cfl_number * semi.max_dt(integrator.u)


def perform_step(integrator, limiters, callbacks, lw_update, time_step_computation, time_discretization):
    """perform_step! for LW (lwfr.jl:212-230) and MDRK (mdrk.jl:1-29)."""
    torch = __import__("torch")
    semi = integrator.p
    dt = compute_dt(semi, semi.mesh, time_step_computation, integrator)
    dt = set_dt(integrator, dt)
    u, uprev = integrator.u, integrator.uprev
    du = lw_update.soln_arrays["du_ode"]
    uprev.copy_(u)
    if isinstance(time_discretization, MDRK):
        rhs_mdrk1(du, u, semi, integrator.t)
        us = torch.add(uprev, du, alpha=dt)  # us = uprev + dt du
        semi.set_us(us)
        for lim in limiters:
            lim(integrator.u, integrator, semi, integrator.t)
        rhs_mdrk2(du, u, semi, integrator.t)
        torch.add(uprev, du, alpha=dt, out=u)
    else:
        rhs(du, u, semi, integrator.t)
        torch.add(uprev, du, alpha=dt, out=u)  # update_soln!: u = uprev + dt du
        for lim in limiters:
            lim(integrator.u, integrator, semi, integrator.t)
    integrator.dts.append(dt)
    set_t_and_iter(integrator, dt)


def solve_lwfr(lw_update: LWUpdate, callbacks=(), dt_initial=1e-6, tolerances=None,
               time_step_computation: CFLBased | None = None, limiters=(), device_loop=None, chunk=64,
               max_steps=None):
    """solve_lwfr(lw_update, callbacks, dt_initial, tolerances; time_step_computation, limiters).

    Returns an LWSolution whose ``u`` is the final state (device tensor).
    ``device_loop`` defaults to True when there are no callbacks and every
    limiter is a PositivityPreservingLimiterZhangShu (at most one).
    ``max_steps`` optionally stops the loop early (for fixed-step runs).
    """
    if time_step_computation is None:
        raise ValueError("time_step_computation=CFLBased(cfl) or Adaptive() is required")
    semi = lw_update.p
    if isinstance(time_step_computation, Adaptive):
        # perform_step! for Adaptive (adaptive.jl:95-173) runs on the device only
        if callbacks or limiters or device_loop is False:
            raise NotImplementedError("adaptive steps run in the device loop, without callbacks and limiters")
        tol = tolerances or {}
        tol = tol if isinstance(tol, dict) else tol._asdict()
        integrator = LWIntegrator(lw_update, get_time_discretization(semi.solver), tol, time_step_computation)
        set_dt(integrator, dt_initial)
        semi.set_positivity_limiter(None)
        _device_solve(integrator, time_step_computation, chunk, max_steps, dt_initial=dt_initial)
        return LWSolution(integrator.u, lw_update, integrator.t, integrator.stats, integrator.dts)
    td = get_time_discretization(semi.solver)
    integrator = LWIntegrator(lw_update, td, tolerances, time_step_computation)
    set_dt(integrator, dt_initial)
    for cb in callbacks:
        cb.initialize(integrator)
    for lim in limiters:
        lim(integrator.u, integrator, semi, integrator.t)
    for cb in callbacks:
        if cb.condition(integrator.u, integrator.t, integrator):
            cb.affect(integrator)
    dev_limiters = all(isinstance(lim, PositivityPreservingLimiterZhangShu) for lim in limiters) and len(limiters) <= 1
    if device_loop is None:
        # time-dependent caller-defined boundary data is evaluated by the host for each rhs
        host_bc = getattr(semi, "_bc_table", None) is not None and not semi._bc_table.steady
        device_loop = not callbacks and dev_limiters and not host_bc
    if device_loop:
        if callbacks or not dev_limiters:
            raise NotImplementedError("the device loop runs without callbacks and with at most one positivity limiter")
        semi.set_positivity_limiter(limiters[0].device_thresholds if limiters else None)
        _device_solve(integrator, time_step_computation, chunk, max_steps)
    else:
        while not isfinished(integrator) and (max_steps is None or integrator.iter < max_steps):
            perform_step(integrator, limiters, callbacks, lw_update, time_step_computation, td)
            for cb in callbacks:
                if cb.condition(integrator.u, integrator.t, integrator):
                    cb.affect(integrator)
            for lim in limiters:
                lim(integrator.u, integrator, semi, integrator.t)
    return LWSolution(integrator.u, lw_update, integrator.t, integrator.stats, integrator.dts)


def _device_solve(integrator: LWIntegrator, tsc, chunk: int, max_steps, dt_initial=None):
    torch = __import__("torch")
    semi = integrator.p
    L = _lib.lib()
    h = semi.handle
    semi._bind_stream()
    u_host = np.ascontiguousarray(integrator.u.detach().cpu().numpy())
    dp = C.POINTER(C.c_double)
    _lib.check(L.tlw_set_state(h, u_host.ctypes.data_as(dp), u_host.size), h)
    adaptive = isinstance(tsc, Adaptive)
    if adaptive:
        tol = integrator.tolerances
        _lib.check(L.tlw_solve_init_adaptive(h, integrator.t, integrator.tspan[1], float(dt_initial),
                                             float(tol.get("abstol", 0.0)), float(tol.get("reltol", 0.0))), h)
    else:
        _lib.check(L.tlw_solve_init(h, integrator.t, integrator.tspan[1], tsc.cfl_number), h)
    t, dt = C.c_double(), C.c_double()
    it, fin = C.c_int64(), C.c_int32()
    attempts = 0
    while True:
        k = chunk if max_steps is None else min(chunk, max_steps - it.value)
        if k <= 0:
            break
        _lib.check(L.tlw_advance(h, k, C.byref(t), C.byref(dt), C.byref(it), C.byref(fin)), h)
        attempts += k
        if fin.value:
            break
        if adaptive and attempts > 100 * (it.value + 100):
            raise RuntimeError(f"adaptive steps make no progress (t = {t.value}, dt = {dt.value})")
    if adaptive:
        nrej = C.c_int64()
        _lib.check(L.tlw_rejected_steps(h, C.byref(nrej)), h)
        integrator.stats.nreject = int(nrej.value)
    _lib.check(L.tlw_get_state(h, u_host.ctypes.data_as(dp), u_host.size), h)
    integrator.u = torch.from_numpy(u_host).to(semi.device)
    integrator.t = t.value
    integrator.iter = int(it.value)
    integrator.stats.naccept = int(it.value)
