"""Pins of the oracle's adaptive time stepping (no GPU).

The temporal-error estimate of lw_volume_kernel_4! (u_np1 against u_np1_low,
src/solvers/dgsem_tree/dg_2d.jl:1183-1199, 1251-1265) is checked against the
dense linear closed form; the controller of perform_step! for Adaptive
(src/solvers/adaptive.jl:95-173) through properties it fixes: the growth cap
of 1.5 per accepted step, rejection and recovery from an unstable dt, the
truncation at the final time and the accuracy of the controlled run.
"""
import numpy as np
import pytest

from conftest import rel_linf
from linear_closed_form import LinearLW


def test_temporal_errors_match_closed_form(oracle_mod):
    """u_np1 = u - dt J⁻¹ (Dx F̃x + Dy F̃y) with the strong-form D, element-local,
    and u_np1_low the same with F̃ before the f_tttt term (a·A⁴u/240 for linear
    advection)."""
    nx, ny, h, N = 3, 2, 0.5, 4
    a = (1.0, 0.7)
    abstol, reltol = 1e-5, 1e-3
    o = oracle_mod.Oracle(N, nx, ny, 0.0, 0.0, h, equation="linadv", adv=a, tolerances=(abstol, reltol))
    nodes, w, D, Dh = o.basis()
    ref = LinearLW(nodes, w, D, Dh, nx, ny, h, a=a)
    u = 1.0 + np.random.default_rng(17).standard_normal((nx * ny, N + 1, N + 1))
    dt = 0.02
    o.rhs(u[..., None], 0.0, dt)
    terms = [u]
    for _ in range(4):
        terms.append(ref.A(terms[-1], dt))
    q = [1.0, 1 / 2, 1 / 6, 1 / 24, 1 / 240]
    Q = sum(q[k] * terms[k] for k in range(5))
    Qlow = sum(q[k] * terms[k] for k in range(4))

    def local_update(Qt):
        return u - dt * ref.invJ * (ref.dx(ref.a1 * Qt, D) + ref.dy(ref.a2 * Qt, D))

    hi, lo = local_update(Q), local_update(Qlow)
    d = -dt * ref.invJ * (ref.dx(ref.a1 * terms[4] / 240, D) + ref.dy(ref.a2 * terms[4] / 240, D))
    assert np.max(np.abs((hi - lo) - d)) < 1e-13 * np.max(np.abs(hi))
    r = d / (abstol + reltol * np.maximum(np.abs(hi), np.abs(lo)))
    expected = np.sum(r**2, axis=(1, 2))
    assert np.all(expected > 1e-8)
    assert rel_linf(o.aux("temporal_errors"), expected) < 1e-12


def _vortex(oracle_mod, nx=16, tol=(1e-6, 1e-6)):
    from trixilw_hip.equations import IsentropicVortex

    ic = IsentropicVortex()
    o = oracle_mod.Oracle(4, nx, nx, -10.0, -10.0, 20.0 / nx, tolerances=tol)
    x, y = o.node_coords()
    return o, ic, x, y


def test_controller_growth_cap_and_accuracy(oracle_mod):
    o, ic, x, y = _vortex(oracle_mod)
    u, t, dts, nrej = o.solve_adaptive(ic(x, y, 0.0), 0.0, 1.0, 1e-4, max_attempts=5000)
    assert abs(t - 1.0) < 1e-12
    # from a tiny dt the error estimate is far below 1: every step grows by the cap 1.5
    assert np.all(dts[1:6] == 1.5 * dts[:5]), dts[:6]
    assert dts[-1] <= dts[-2] * 1.5
    # the controlled run is as accurate as the CFL-limited one
    err = np.sqrt(np.mean((u - ic(x, y, t))[..., 0] ** 2))
    u_cfl, t_cfl, _ = o.solve(ic(x, y, 0.0), 0.0, 1.0, 0.98 * 5 * 0.103)
    err_cfl = np.sqrt(np.mean((u_cfl - ic(x, y, t_cfl))[..., 0] ** 2))
    assert err < 2.0 * err_cfl, (err, err_cfl)


def test_rejection_and_recovery_from_unstable_dt(oracle_mod):
    o, ic, x, y = _vortex(oracle_mod)
    u, t, dts, nrej = o.solve_adaptive(ic(x, y, 0.0), 0.0, 0.5, 5.0, max_attempts=5000)
    assert nrej >= 1
    assert abs(t - 0.5) < 1e-12
    assert np.all(np.isfinite(u))
    assert np.sqrt(np.mean((u - ic(x, y, t))[..., 0] ** 2)) < 1e-2
    # the accepted steps stay near the stable range
    assert dts.max() < 0.2


def test_last_step_is_truncated(oracle_mod):
    o, ic, x, y = _vortex(oracle_mod, nx=8)
    tf = 0.123456789
    _, t, dts, _ = o.solve_adaptive(ic(x, y, 0.0), 0.0, tf, 1e-3, max_attempts=5000)
    assert abs(t - tf) < 1e-15
    assert abs(np.sum(dts) - tf) < 1e-13
