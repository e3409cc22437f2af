"""Pins of the oracle's hyperbolic-parabolic path (no GPU).

The reference ships no tests or golden vectors for this path.  The oracle's
BR1 gradient, Taylor states of ∇u, split advective/viscous fluxes and
interface flux are pinned by an independent dense formulation of linear
advection-diffusion (tests/linear_closed_form.py) and by convergence to the
exact decaying wave of the reference's advection-diffusion examples
(examples/tree_2d_dgsem/elixir_advection_diffusion*.jl); the Navier-Stokes
viscous flux by the chain rule through (ρ, v1, v2, T); the Euler part by
μ = 0 against the hyperbolic solver; the whole right-hand side by
free-stream preservation and conservation; compute_dt by its definition on a
constant state.
"""
import numpy as np
import pytest

from conftest import rel_linf
from linear_closed_form import LinearParabolicLW


def make(oracle_mod, N, nx, ny=None, L=2.0, x0=-1.0, **kw):
    ny = nx if ny is None else ny
    return oracle_mod.Oracle(N, nx, ny, x0, x0, L / nx, **kw)


# ---------------------------------------------------------------- linear closed form
@pytest.mark.parametrize("N", [3, 4])
def test_advection_diffusion_matches_closed_form(oracle_mod, N):
    nx, ny, h = 3, 4, 0.5
    a, nu = (1.5, 1.0), 0.05
    o = oracle_mod.Oracle(N, nx, ny, 0.0, 0.0, h, equation="linadv", adv=a, parabolic="laplace", mu=nu)
    nodes, w, D, Dh = o.basis()
    ref = LinearParabolicLW(nodes, w, D, Dh, nx, ny, h, a=a, nu=nu)
    rng = np.random.default_rng(4321 + N)
    u = rng.standard_normal((nx * ny, N + 1, N + 1))
    dt = 0.011
    du = o.rhs(u[..., None], 0.0, dt)[..., 0]
    _, du_ref = ref.lw_step(u, dt)
    assert rel_linf(du, du_ref) < 1e-13
    # the BR1 gradient the rhs used
    gx, gy = ref.br1_gradient(u)
    assert rel_linf(o.aux("gradients_x").reshape(u.shape), gx) < 1e-13
    assert rel_linf(o.aux("gradients_y").reshape(u.shape), gy) < 1e-13
    # MDRK: stage 1, then stage 2 from us = u + dt du1
    du1 = o.rhs(u[..., None], 0.0, dt, kind="mdrk1")[..., 0]
    _, du1_ref, du2_ref, _ = ref.mdrk_step(u, dt)
    assert rel_linf(du1, du1_ref) < 1e-13
    o.set_us((u + dt * du1)[..., None])
    du2 = o.rhs(u[..., None], 0.0, dt, kind="mdrk2")[..., 0]
    assert rel_linf(du2, du2_ref) < 1e-13


def test_closed_form_distinguishes_viscous_terms(oracle_mod):
    """The closed form is sensitive to the pieces it pins: dropping the Taylor
    states of ∇u (viscous flux frozen at uⁿ) or the viscous interface average
    changes the step far beyond round-off."""
    nx, ny, h, N = 3, 3, 0.5, 3
    a, nu = (1.5, 1.0), 0.05
    o = oracle_mod.Oracle(N, nx, ny, 0.0, 0.0, h, equation="linadv", adv=a, parabolic="laplace", mu=nu)
    nodes, w, D, Dh = o.basis()
    ref = LinearParabolicLW(nodes, w, D, Dh, nx, ny, h, a=a, nu=nu)
    u = np.random.default_rng(5).standard_normal((nx * ny, N + 1, N + 1))
    dt = 0.02
    du = o.rhs(u[..., None], 0.0, dt)[..., 0]
    # variant 1: element-local gradients of the time derivatives set to zero
    frozen = LinearParabolicLW(nodes, w, D, Dh, nx, ny, h, a=a, nu=nu)
    frozen.local_gradient = lambda v: (0.0 * v, 0.0 * v)
    assert rel_linf(du, frozen.lw_step(u, dt)[1]) > 1e-6
    # variant 2: no diffusion at all
    adv_only = LinearParabolicLW(nodes, w, D, Dh, nx, ny, h, a=a, nu=0.0)
    assert rel_linf(du, adv_only.lw_step(u, dt)[1]) > 1e-4
    assert rel_linf(du, ref.lw_step(u, dt)[1]) < 1e-13


@pytest.mark.parametrize("N,mdrk,eoc_min", [(3, False, 3.6), (4, False, 4.5), (3, True, 3.6)])
def test_advection_diffusion_convergence(oracle_mod, N, mdrk, eoc_min):
    """EOC against the exact solution of the reference's advection-diffusion
    examples: u = 1 + ½ sin(ω(x + y - (a1 + a2) t)) e^{-2νω²t}, ω = π on
    [-1, 1]² (examples/tree_2d_dgsem/elixir_advection_diffusion.jl:25-38)."""
    a, nu = (1.5, 1.0), 0.05
    om = np.pi
    tf = 0.5

    def exact(x, y, t):
        return 1.0 + 0.5 * np.sin(om * (x + y - (a[0] + a[1]) * t)) * np.exp(-2 * nu * om**2 * t)

    errs = []
    for nx in (4, 8, 16):
        o = make(oracle_mod, N, nx, equation="linadv", adv=a, parabolic="laplace", mu=nu)
        x, y = o.node_coords()
        u0 = exact(x, y, 0.0)[..., None]
        u, t, _ = o.solve(u0, 0.0, tf, 0.9 if not mdrk else 0.45, mdrk=mdrk)
        assert abs(t - tf) < 1e-12
        errs.append(np.sqrt(np.mean((u[..., 0] - exact(x, y, t)) ** 2)))
    eoc = np.log(errs[1] / errs[2]) / np.log(2)
    assert eoc > eoc_min, (errs, eoc)


# ---------------------------------------------------------------- Navier-Stokes
def _prim(u, gamma):
    rho = u[0]
    v1, v2 = u[1] / rho, u[2] / rho
    p = (gamma - 1) * (u[3] - 0.5 * (u[1] ** 2 + u[2] ** 2) / rho)
    return np.array([rho, v1, v2, p / rho])


def test_navier_stokes_viscous_flux_chain_rule(oracle_mod):
    """f_v from conservative gradients equals the textbook Navier-Stokes flux
    of the primitive gradients obtained by the chain rule (complex-step
    derivative of u -> (ρ, v1, v2, T = p/ρ))."""
    gamma, mu, Pr = 1.4, 0.013, 0.72
    kappa = gamma / ((gamma - 1) * Pr)
    rng = np.random.default_rng(8)
    for _ in range(20):
        rho = 0.5 + rng.random()
        v = rng.standard_normal(2)
        p = 0.5 + rng.random()
        u = np.array([rho, rho * v[0], rho * v[1], p / (gamma - 1) + 0.5 * rho * (v @ v)])
        gx, gy = rng.standard_normal(4), rng.standard_normal(4)
        fx, fy = oracle_mod.ns_viscous_fluxes(u, gx, gy, gamma, mu, Pr)
        h = 1e-30
        dpx = np.imag(_prim(u + 1j * h * gx, gamma)) / h
        dpy = np.imag(_prim(u + 1j * h * gy, gamma)) / h
        _, v1, v2, _ = _prim(u, gamma)
        t11 = 4 / 3 * dpx[1] - 2 / 3 * dpy[2]
        t22 = 4 / 3 * dpy[2] - 2 / 3 * dpx[1]
        t12 = dpy[1] + dpx[2]
        ex = mu * np.array([0.0, t11, t12, v1 * t11 + v2 * t12 + kappa * dpx[3]])
        ey = mu * np.array([0.0, t12, t22, v1 * t12 + v2 * t22 + kappa * dpy[3]])
        assert np.max(np.abs(fx - ex)) < 1e-13 * (1 + np.max(np.abs(ex)))
        assert np.max(np.abs(fy - ey)) < 1e-13 * (1 + np.max(np.abs(ey)))


def _ns_state(x, y, gamma=1.4):
    rho = 1.0 + 0.2 * np.sin(np.pi * x) * np.cos(np.pi * y)
    v1 = 0.3 + 0.1 * np.sin(np.pi * y)
    v2 = -0.2 + 0.1 * np.cos(np.pi * x)
    p = 1.0 + 0.1 * np.cos(np.pi * (x + y))
    return np.stack([rho, rho * v1, rho * v2, p / (gamma - 1) + 0.5 * rho * (v1**2 + v2**2)], axis=-1)


@pytest.mark.parametrize("N", [3, 4])
def test_navier_stokes_mu_zero_is_euler(oracle_mod, N):
    o = make(oracle_mod, N, 4, parabolic="navier_stokes", mu=0.0)
    e = make(oracle_mod, N, 4)
    x, y = o.node_coords()
    u = _ns_state(x, y)
    dt = 0.003
    assert rel_linf(o.rhs(u, 0.0, dt), e.rhs(u, 0.0, dt)) < 1e-13
    du1 = o.rhs(u, 0.0, dt, kind="mdrk1")
    assert rel_linf(du1, e.rhs(u, 0.0, dt, kind="mdrk1")) < 1e-13
    o.set_us(u + dt * du1)
    e.set_us(u + dt * du1)
    assert rel_linf(o.rhs(u, 0.0, dt, kind="mdrk2"), e.rhs(u, 0.0, dt, kind="mdrk2")) < 1e-13


@pytest.mark.parametrize("N", [3, 4])
@pytest.mark.parametrize("kind", ["lw", "mdrk"])
def test_navier_stokes_free_stream(oracle_mod, N, kind):
    o = make(oracle_mod, N, 4, parabolic="navier_stokes", mu=0.05)
    x, y = o.node_coords()
    u = np.zeros(x.shape + (4,))
    u[...] = [1.3, 0.4, -0.7, 3.1]
    dt = 0.01
    if kind == "lw":
        du = o.rhs(u, 0.0, dt)
    else:
        du = o.rhs(u, 0.0, dt, kind="mdrk1")
        o.set_us(u + dt * du)
        du = o.rhs(u, 0.0, dt, kind="mdrk2")
    assert np.max(np.abs(du)) < 1e-12


@pytest.mark.parametrize("N", [3, 4])
def test_navier_stokes_conservation(oracle_mod, N):
    o = make(oracle_mod, N, 6, parabolic="navier_stokes", mu=0.02)
    x, y = o.node_coords()
    u = _ns_state(x, y)
    dt = 0.5 * o.max_dt(u)
    _, w, _, _ = o.basis()
    W = np.outer(w, w)
    for kind in ("lw", "mdrk1"):
        du = o.rhs(u, 0.0, dt, kind=kind)
        total = np.einsum("ji,ejiv->v", W, du)
        scale = np.einsum("ji,ejiv->v", W, np.abs(du))
        assert np.all(np.abs(total) <= 1e-13 * scale), (kind, total, scale)


@pytest.mark.parametrize("N", [3, 4])
def test_compute_dt_hyperbolic_parabolic(oracle_mod, N):
    """compute_dt = CFL·min(advective max_dt, parabolic max_dt) on a constant
    state, with the parabolic one 1/((λa + λv)·2(2N+1)),
    λv = 2(2N+1)·J⁻²·max(4μ/(3ρ), μκ/(ρPr)) (lwfr.jl:192-210,
    compressible_navier_stokes_2d.jl:204-240)."""
    gamma, Pr = 1.4, 0.72
    kappa = gamma / ((gamma - 1) * Pr)
    rho, v1, v2, p = 1.3, 0.4, -0.7, 2.0
    c = np.sqrt(gamma * p / rho)
    nx = 4
    invJ = 2.0 / (2.0 / nx)
    lam_a = invJ * (abs(v1) + c + abs(v2) + c)
    u0 = np.array([rho, rho * v1, rho * v2, p / (gamma - 1) + 0.5 * rho * (v1**2 + v2**2)])
    for mu in (1e-4, 0.5):
        o = make(oracle_mod, N, nx, parabolic="navier_stokes", mu=mu, prandtl=Pr)
        x, _ = o.node_coords()
        u = np.broadcast_to(u0, x.shape + (4,)).copy()
        lam_v = 2 * (2 * N + 1) * invJ**2 * max(4 * mu / (3 * rho), mu * kappa / (rho * Pr))
        dt_visc = 1.0 / ((lam_a + lam_v) * 2 * (2 * N + 1))
        dt_adv = 2.0 / ((N + 1) * lam_a)
        assert abs(o.max_dt(u) - min(dt_adv, dt_visc)) <= 1e-14 * min(dt_adv, dt_visc)
    # the viscous bound is the active one for large μ, the advective one for small μ
    assert dt_visc < dt_adv


def test_navier_stokes_solve_decays_kinetic_energy(oracle_mod):
    """A viscous shear wave loses kinetic energy and keeps mass."""
    o = make(oracle_mod, 3, 6, parabolic="navier_stokes", mu=0.02)
    x, y = o.node_coords()
    gamma = 1.4
    v1 = 0.1 * np.sin(np.pi * y)
    u = np.stack([np.ones_like(x), v1, np.zeros_like(x), 1.0 / (gamma - 1) + 0.5 * v1**2], axis=-1)
    _, w, _, _ = o.basis()
    W = np.outer(w, w)

    def ke(q):
        return np.einsum("ji,eji->", W, 0.5 * (q[..., 1] ** 2 + q[..., 2] ** 2) / q[..., 0])

    u1, t, dts = o.solve(u, 0.0, 0.5, 0.5)
    assert abs(t - 0.5) < 1e-12 and len(dts) > 5
    assert ke(u1) < 0.95 * ke(u)
    m0, m1 = np.einsum("ji,eji->", W, u[..., 0]), np.einsum("ji,eji->", W, u1[..., 0])
    assert abs(m1 - m0) < 1e-13 * m0
