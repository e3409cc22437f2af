"""Pins of the CPU oracle (no GPU).

The reference ships no tests or golden vectors for this path, so the oracle is
pinned by closed forms, invariants and an independent dense-matrix
formulation of the linear case (tests/linear_closed_form.py).
"""
import numpy as np
import pytest

from conftest import rel_l2, rel_linf
from linear_closed_form import LinearLW


def make(oracle_mod, N, nx, ny=None, L=20.0, x0=-10.0, **kw):
    ny = nx if ny is None else ny
    return oracle_mod.Oracle(N, nx, ny, x0, x0, L / nx, **kw)


# ---------------------------------------------------------------- basis
@pytest.mark.parametrize("N", [3, 4])
def test_gll_closed_forms(oracle_mod, N):
    o = make(oracle_mod, N, 2)
    nodes, w, D, Dh = o.basis()
    if N == 3:
        xs = np.array([-1, -np.sqrt(1 / 5), np.sqrt(1 / 5), 1])
        ws = np.array([1 / 6, 5 / 6, 5 / 6, 1 / 6])
    else:
        xs = np.array([-1, -np.sqrt(3 / 7), 0, np.sqrt(3 / 7), 1])
        ws = np.array([1 / 10, 49 / 90, 32 / 45, 49 / 90, 1 / 10])
    assert np.max(np.abs(nodes - xs)) < 2e-16
    assert np.max(np.abs(w - ws)) < 2e-16


@pytest.mark.parametrize("N", [3, 4])
def test_derivative_matrix_exact_on_polynomials(oracle_mod, N):
    o = make(oracle_mod, N, 2)
    nodes, w, D, Dh = o.basis()
    for deg in range(N + 1):
        p = np.polynomial.Polynomial(np.arange(1, deg + 2, dtype=float))
        assert np.max(np.abs(D @ p(nodes) - p.deriv()(nodes))) < 1e-13
    # SBP: W D + D^T W = diag(-1, 0, ..., 0, 1); D̂ = -W⁻¹ Dᵀ W
    W = np.diag(w)
    B = np.zeros((N + 1, N + 1))
    B[0, 0], B[-1, -1] = -1, 1
    assert np.max(np.abs(W @ D + D.T @ W - B)) < 1e-14
    assert np.max(np.abs(Dh - (-np.linalg.inv(W) @ D.T @ W))) < 1e-14


@pytest.mark.parametrize("N", [3, 4])
def test_host_basis_matches_oracle(oracle_mod, N):
    from trixilw_hip.basis import LobattoLegendreBasis

    o = make(oracle_mod, N, 2)
    nodes, w, D, Dh = o.basis()
    b = LobattoLegendreBasis(N)
    assert np.array_equal(b.nodes, nodes)
    assert np.array_equal(b.weights, w)
    assert np.max(np.abs(b.derivative_matrix - D)) <= 4e-16 * np.max(np.abs(D))
    assert np.max(np.abs(b.derivative_dhat - Dh)) <= 4e-16 * np.max(np.abs(Dh))


# ---------------------------------------------------------------- pointwise Euler
def test_euler_flux_closed_form(oracle_mod):
    g = 1.4
    rho, v1, v2, p = 1.3, 0.7, -0.4, 2.1
    E = p / (g - 1) + 0.5 * rho * (v1 ** 2 + v2 ** 2)
    u = np.array([rho, rho * v1, rho * v2, E])
    f, gg = oracle_mod.euler_fluxes(u, g)
    assert np.allclose(f, [rho * v1, rho * v1 ** 2 + p, rho * v1 * v2, (E + p) * v1], rtol=1e-15, atol=0)
    assert np.allclose(gg, [rho * v2, rho * v1 * v2, rho * v2 ** 2 + p, (E + p) * v2], rtol=1e-15, atol=0)
    # mirror x <-> y with v1 <-> v2 maps f <-> g
    um = np.array([rho, rho * v2, rho * v1, E])
    fm, gm = oracle_mod.euler_fluxes(um, g)
    assert np.allclose(fm[[0, 2, 1, 3]], gg, rtol=1e-15)
    # max_abs_speed_naive = max(|v_n|) + max(c)
    ur = np.array([0.9, 0.9 * -1.2, 0.9 * 0.3, 1.5 / (g - 1) + 0.5 * 0.9 * (1.44 + 0.09)])
    cl, cr = np.sqrt(g * p / rho), np.sqrt(g * 1.5 / 0.9)
    lam = oracle_mod.max_abs_speed_naive(u, ur, 1, g)
    assert abs(lam - (max(abs(v1), 1.2) + max(cl, cr))) < 1e-15 * lam
    lam2 = oracle_mod.max_abs_speed_naive(u, ur, 2, g)
    assert abs(lam2 - (max(abs(v2), 0.3) + max(cl, cr))) < 1e-15 * lam2


# ---------------------------------------------------------------- linear closed form
@pytest.mark.parametrize("N", [3, 4])
@pytest.mark.parametrize("flux_order", [0, 1])
def test_linear_lw_and_mdrk_match_closed_form(oracle_mod, N, flux_order):
    nx, ny, h = 3, 2, 0.5
    a = (1.0, 0.7)
    o = oracle_mod.Oracle(N, nx, ny, 0.0, 0.0, h, equation="linadv", adv=a, flux_order=flux_order)
    nodes, w, D, Dh = o.basis()
    ref = LinearLW(nodes, w, D, Dh, nx, ny, h, a=a, flux_order="tree" if flux_order == 0 else "structured")
    rng = np.random.default_rng(1234 + N)
    u = rng.standard_normal((nx * ny, N + 1, N + 1))
    dt = 0.013
    du = o.rhs(u[..., None], 0.0, dt)[..., 0]
    _, du_ref = ref.lw_step(u, dt)
    assert rel_linf(du, du_ref) < 1e-14
    du1 = o.rhs(u[..., None], 0.0, dt, kind="mdrk1")[..., 0]
    un_ref, du1_ref, du2_ref, us_ref = ref.mdrk_step(u, dt)
    assert rel_linf(du1, du1_ref) < 1e-14
    o.set_us((u + dt * du1)[..., None])
    du2 = o.rhs(u[..., None], 0.0, dt, kind="mdrk2")[..., 0]
    assert rel_linf(du2, du2_ref) < 1e-14


def test_f4_quirk_is_pinned(oracle_mod):
    """The N=4 f_tttt term carries the reference's factor 0.5: a closed form with
    the standard coefficient 1/120 instead of 1/240 must not match."""
    nx, ny, h, N = 2, 2, 0.5, 4
    o = oracle_mod.Oracle(N, nx, ny, 0.0, 0.0, h, equation="linadv", adv=(1.0, 0.5))
    nodes, w, D, Dh = o.basis()
    ref = LinearLW(nodes, w, D, Dh, nx, ny, h, a=(1.0, 0.5))
    u = np.random.default_rng(7).standard_normal((4, 5, 5))
    dt = 0.05
    du = o.rhs(u[..., None], 0.0, dt)[..., 0]
    terms = [u]
    for _ in range(4):
        terms.append(ref.A(terms[-1], dt))
    fact = [1, 1 / 2, 1 / 6, 1 / 24, 1 / 120]
    Ut = sum(fact[k] * terms[k] for k in range(5))
    _, du_std = ref.fr_update(u, Ut, Ut, dt)
    assert rel_linf(du, du_std) > 1e-8


# ---------------------------------------------------------------- Euler invariants
@pytest.mark.parametrize("N", [3, 4])
@pytest.mark.parametrize("kind", ["lw", "mdrk"])
def test_free_stream_preservation(oracle_mod, N, kind):
    from trixilw_hip.equations import ConstantState

    o = make(oracle_mod, N, 4, L=2.0, x0=-1.0)
    x, y = o.node_coords()
    u = ConstantState(1.2, 0.3, -0.5, 0.9)(x, y, 0.0)
    dt = 0.3 * o.max_dt(u)
    if kind == "lw":
        du = o.rhs(u, 0.0, dt)
    else:
        du = o.rhs(u, 0.0, dt, kind="mdrk1")
        o.set_us(u + dt * du)
        du = o.rhs(u, 0.0, dt, kind="mdrk2")
    assert np.max(np.abs(du)) < 1e-12


@pytest.mark.parametrize("N", [3, 4])
def test_discrete_conservation(oracle_mod, N):
    from trixilw_hip.equations import IsentropicVortex

    o = make(oracle_mod, N, 8)
    x, y = o.node_coords()
    rng = np.random.default_rng(3)
    u = IsentropicVortex()(x, y, 0.0) * (1 + 1e-3 * rng.standard_normal(x.shape + (1,)))
    dt = 0.5 * o.max_dt(u)
    du = o.rhs(u, 0.0, dt)
    _, w, _, _ = o.basis()
    W = np.outer(w, w)
    total = np.einsum("ji,ejiv->v", W, du)
    scale = np.einsum("ji,ejiv->v", W, np.abs(du))
    assert np.all(np.abs(total) <= 1e-13 * scale)


def test_mirror_symmetry(oracle_mod):
    """Transposing the state (x <-> y, v1 <-> v2) transposes the residual."""
    from trixilw_hip.equations import IsentropicVortex

    o = make(oracle_mod, 4, 4)
    x, y = o.node_coords()
    rng = np.random.default_rng(11)
    u = IsentropicVortex(center=(0.3, -1.1))(x, y, 0.0)
    u = u * (1 + 1e-2 * rng.standard_normal(u.shape))
    # element (ix, iy) -> (iy, ix); node (i, j) -> (j, i)
    nx = o.nx
    perm = np.array([(e % nx) * nx + e // nx for e in range(o.E)])
    ut = u[perm].transpose(0, 2, 1, 3)[..., [0, 2, 1, 3]]
    dt = 0.4 * o.max_dt(u)
    du = o.rhs(u, 0.0, dt)
    dut = o.rhs(np.ascontiguousarray(ut), 0.0, dt)
    back = dut[perm].transpose(0, 2, 1, 3)[..., [0, 2, 1, 3]]
    assert rel_linf(back, du) < 1e-13


@pytest.mark.parametrize("N,mdrk,eoc_min", [(4, False, 3.6), (3, False, 3.4), (3, True, 3.4)])
def test_isentropic_vortex_convergence(oracle_mod, N, mdrk, eoc_min):
    """EOC of the L2 error at t=1 (the reference's verification method,
    utils/plot_convergence.jl:66-69)."""
    from trixilw_hip.equations import IsentropicVortex

    ic = IsentropicVortex()
    fc = {3: 0.170, 4: 0.103}[N]
    cfl = 0.98 * (N + 1) * fc * (0.5 if mdrk else 1.0)
    errs = []
    for nx in (16, 32, 64):
        o = make(oracle_mod, N, nx)
        x, y = o.node_coords()
        u, t, dts = o.solve(ic(x, y, 0.0), 0.0, 1.0, cfl, mdrk=mdrk)
        assert abs(t - 1.0) < 1e-12
        errs.append(np.sqrt(np.mean((u - ic(x, y, t))[..., 0] ** 2)))
    eoc = np.log(errs[1] / errs[2]) / np.log(2)
    assert eoc > eoc_min, (errs, eoc)


def test_solve_truncates_last_step(oracle_mod):
    from trixilw_hip.equations import IsentropicVortex

    o = make(oracle_mod, 4, 8)
    x, y = o.node_coords()
    u, t, dts = o.solve(IsentropicVortex()(x, y, 0.0), 0.0, 0.37, 0.5)
    assert t == pytest.approx(0.37, rel=0, abs=1e-15)
    assert dts[-1] <= dts[-2]


# ---------------------------------------------------------------- tangent (dual numbers)
def test_dual_tangent_matches_central_differences(oracle_mod):
    """The dual-number tangent of the fixed-dt LW loop (the forward-mode check
    of the device adjoint) against central differences, and its value part
    against the plain double loop."""
    from trixilw_hip.equations import IsentropicVortex

    o = make(oracle_mod, 4, 8)
    x, y = o.node_coords()
    u0 = IsentropicVortex()(x, y, 0.0)
    rng = np.random.default_rng(12)
    v = np.zeros_like(u0)
    v[..., 0] = 1e-2 * rng.standard_normal(x.shape)
    v[..., 3] = 1e-2 * rng.standard_normal(x.shape)
    dts = [0.3 * o.max_dt(u0)] * 6
    uK, duK = o.tangent_lw(u0, v, dts)
    assert np.max(np.abs(uK - o.solve_fixed_dt(u0, dts))) <= 1e-14 * np.max(np.abs(uK))
    h = 1e-6
    fd = (o.solve_fixed_dt(u0 + h * v, dts) - o.solve_fixed_dt(u0 - h * v, dts)) / (2 * h)
    assert rel_linf(duK, fd) < 1e-7


def test_dual_tangent_mdrk_and_frozen_blend_match_central_differences(oracle_mod):
    """The dual-number tangent of the fixed-dt MDRK loop, and of the LW loop
    with shock capturing at frozen blending factors, against central
    differences of the same maps (the value part of the dual loop with the
    same dts and factors); the MDRK value part equals the oracle's CFL solve
    when fed that solve's dts."""
    from trixilw_hip.equations import IsentropicVortex

    o = make(oracle_mod, 3, 8)
    x, y = o.node_coords()
    u0 = IsentropicVortex()(x, y, 0.0)
    rng = np.random.default_rng(13)
    v = np.zeros_like(u0)
    v[..., 0] = 1e-2 * rng.standard_normal(x.shape)
    v[..., 2] = 1e-2 * rng.standard_normal(x.shape)
    u_ref, _, dts = o.solve(u0, 0.0, 1e9, 0.3, max_steps=5, mdrk=True)
    uK, duK = o.tangent(u0, v, dts, mdrk=True)
    assert np.max(np.abs(uK - u_ref)) <= 1e-13 * np.max(np.abs(u_ref))
    h = 1e-6
    fd = (o.tangent(u0 + h * v, 0 * v, dts, mdrk=True)[0] - o.tangent(u0 - h * v, 0 * v, dts, mdrk=True)[0]) / (2 * h)
    assert rel_linf(duK, fd) < 1e-7
    # shock capturing with the blending factors frozen at a blocky pattern
    osc = oracle_mod.Oracle(4, 8, 8, -10.0, -10.0, 2.5, shock_capturing=True, alpha_max=0.5)
    x, y = osc.node_coords()
    # node noise breaks the vortex's mirror symmetry, where the Rusanov speed's max has ties (kinks)
    u0 = IsentropicVortex()(x, y, 0.0) * (1 + 1e-3 * rng.standard_normal(x.shape + (4,)))
    v = np.zeros_like(u0)
    v[..., 0] = 1e-2 * rng.standard_normal(x.shape)
    v[..., 3] = 1e-2 * rng.standard_normal(x.shape)
    alpha = rng.uniform(0.0, 0.5, 64) * (rng.uniform(size=64) < 0.5)
    dts = [0.2 * osc.max_dt(u0)] * 4
    uK, duK = osc.tangent(u0, v, dts, alpha=alpha)
    fd = (osc.tangent(u0 + h * v, 0 * v, dts, alpha=alpha)[0] - osc.tangent(u0 - h * v, 0 * v, dts, alpha=alpha)[0]) / (2 * h)
    assert rel_linf(duK, fd) < 1e-7
    # the frozen factors matter
    assert rel_linf(osc.tangent(u0, v, dts, alpha=0 * alpha)[1], duK) > 1e-4
