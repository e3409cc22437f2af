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


def test_parabolic_temporal_errors_match_closed_form(oracle_mod):
    """Kernel 4 of the hyperbolic-parabolic path (dg_2d_parabolic.jl:1749-1775,
    1841-1855) on linear advection-diffusion, f = a u - ν∇u: with the Taylor
    terms u_k, g_k of the dense closed form, u_np1 - u_np1_low is the local
    update of the f_tttt term (a u_4 - ν g_4)/240 and the element error follows."""
    from linear_closed_form import LinearParabolicLW

    nx, ny, h, N = 3, 2, 0.5, 4
    a, nu = (1.0, 0.7), 0.05
    abstol, reltol = 1e-5, 1e-3
    o = oracle_mod.Oracle(N, nx, ny, 0.0, 0.0, h, equation="linadv", adv=a, parabolic="laplace", mu=nu,
                          tolerances=(abstol, reltol))
    nodes, w, D, Dh = o.basis()
    ref = LinearParabolicLW(nodes, w, D, Dh, nx, ny, h, a=a, nu=nu)
    u = 1.0 + np.random.default_rng(23).standard_normal((nx * ny, N + 1, N + 1))
    dt = 0.02
    o.rhs(u[..., None], 0.0, dt)
    us_, (gx, gy) = [u], ref.br1_gradient(u)
    gxs, gys = [gx], [gy]
    for _ in range(4):
        wk = ref.advance(us_[-1], gxs[-1], gys[-1], dt)
        us_.append(wk)
        wx, wy = ref.local_gradient(wk)
        gxs.append(wx)
        gys.append(wy)
    q = [1.0, 1 / 2, 1 / 6, 1 / 24, 1 / 240]
    Fx = sum(q[k] * (ref.a1 * us_[k] - nu * gxs[k]) for k in range(5))
    Fy = sum(q[k] * (ref.a2 * us_[k] - nu * gys[k]) for k in range(5))
    hi = u - dt * ref.invJ * (ref.dx(Fx, D) + ref.dy(Fy, D))
    d = -dt * ref.invJ * (ref.dx((ref.a1 * us_[4] - nu * gxs[4]) / 240, D) +
                          ref.dy((ref.a2 * us_[4] - nu * gys[4]) / 240, D))
    lo = hi - d
    r = d / (abstol + reltol * np.maximum(np.abs(hi), np.abs(lo)))
    expected = np.sum(r**2, axis=(1, 2))
    assert np.all(expected > 1e-8)
    assert rel_linf(o.aux("temporal_errors"), expected) < 1e-12
    # without the diffusion part the estimate is a different one
    adv = LinearParabolicLW(nodes, w, D, Dh, nx, ny, h, a=a, nu=0.0)
    assert np.max(np.abs(adv.advance(u, gx, gy, dt) - us_[1])) > 1e-3


def test_parabolic_temporal_errors_at_zero_viscosity_equal_euler(oracle_mod):
    """Navier-Stokes with μ = 0 reproduces the Euler kernel 4, its temporal
    errors included."""
    from trixilw_hip.equations import IsentropicVortex

    ic = IsentropicVortex()
    kw = dict(tolerances=(1e-6, 1e-6))
    e = oracle_mod.Oracle(4, 6, 6, -10.0, -10.0, 20.0 / 6, **kw)
    p = oracle_mod.Oracle(4, 6, 6, -10.0, -10.0, 20.0 / 6, parabolic="navier_stokes", mu=0.0, **kw)
    x, y = e.node_coords()
    u = ic(x, y, 0.0) * (1.0 + 1e-3 * np.random.default_rng(4).standard_normal(x.shape + (4,)))
    dt = 0.5 * e.max_dt(u)
    e.rhs(u, 0.0, dt)
    p.rhs(u, 0.0, dt)
    te, tp = e.aux("temporal_errors"), p.aux("temporal_errors")
    assert np.all(te > 0)
    assert rel_linf(tp, te) < 1e-12


def test_parabolic_adaptive_solve_growth_and_accuracy(oracle_mod):
    """The controller on the Navier-Stokes shear flow: from a tiny dt every
    accepted step grows by the cap 1.5; a tighter tolerance takes more steps
    and lands closer to a fine fixed-step run."""
    GAMMA, PR = 1.4, 0.72

    def shear(x, y):
        rho = 1.0 + 0.2 * np.sin(np.pi * x) * np.cos(np.pi * y)
        v1 = 0.3 + 0.1 * np.sin(np.pi * y)
        v2 = -0.2 + 0.1 * np.cos(np.pi * x)
        p = 1.0 + 0.1 * np.cos(np.pi * (x + y))
        return np.stack([rho, rho * v1, rho * v2, p / (GAMMA - 1) + 0.5 * rho * (v1**2 + v2**2)], axis=-1)

    runs = {}
    for tol in (1e-5, 1e-8):
        o = oracle_mod.Oracle(4, 6, 6, -1.0, -1.0, 2.0 / 6, parabolic="navier_stokes", mu=0.01, prandtl=PR,
                              tolerances=(tol, tol))
        x, y = o.node_coords()
        u, t, dts, nrej = o.solve_adaptive(shear(x, y), 0.0, 0.2, 1e-6, max_attempts=20000)
        assert abs(t - 0.2) < 1e-12
        assert np.all(dts[1:6] == 1.5 * dts[:5]), dts[:6]
        runs[tol] = (u, len(dts))
    ref, _, _ = o.solve(shear(x, y), 0.0, 0.2, 0.05)
    e_loose = np.max(np.abs(runs[1e-5][0] - ref))
    e_tight = np.max(np.abs(runs[1e-8][0] - ref))
    assert runs[1e-8][1] > runs[1e-5][1]
    assert e_tight < e_loose
