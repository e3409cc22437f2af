"""Pins of the oracle's Navier-Stokes convergence test (the manufactured
solution and source of examples/tree_2d_dgsem/elixir_navierstokes_convergence.jl,
the source folded into the hyperbolic-parabolic LW and MDRK kernels as
dg_2d_parabolic.jl:725-1860 and multi_derivative_2d_parabolic.jl:221-711 do;
no GPU).

- The source is the residual ∂t u + ∇·f_a - ∇·f_v of the manufactured
  solution: central differences of the state (time, space) and of fluxes
  built here in numpy from the state and its differenced gradient agree with
  it to the differences' second-order error.
- The manufactured velocity vanishes on the walls y = ±1, so the example's
  no-slip walls hold it there.
- The Taylor weights of the source (S, and dt·s added to each level): on a
  uniform state with a source of t alone every flux derivative vanishes and
  the residual is S itself.  LW kernel 4 (five-point s_tt, calc_source_tt_N4)
  gives the exact average of s over [t, t + dt] for quartics; kernel 3 for
  cubics, a quartic then exceeding it by dt⁴ s⁽⁴⁾/180 (three-point s_tt and
  no s_tttt level); MDRK stage 1 gives s/2 + dt s'/8 (for a cubic, short of
  the average over [t, t + dt/2] by dt² s⁽²⁾/48 + dt³ s⁽³⁾/384); the MDRK step (stage 2
  with s_t at the call's t, as the reference evaluates it) gives s + dt s'/2,
  short of the exact average by dt² s⁽²⁾/6 for a quadratic.
- The example's setup (periodic in x, no-slip adiabatic walls in y, μ 0.01,
  Pr 0.72) converges to the manufactured solution at order ≈ N + 1 in the
  density's L2 error at t = 0.1 on 4², 8², 16² cells, LW and MDRK; a source
  missing from the Taylor levels (or from S) stalls at low order.
"""
import numpy as np
import pytest

G, MU, PR = 1.4, 0.01, 0.72


def np_fluxes(u, ux, uy):
    """f_a - f_v of compressible Navier-Stokes with conservative gradients
    (compressible_navier_stokes_2d.jl's viscous flux, T = p/ρ)."""
    rho, m1, m2, E = np.moveaxis(u, -1, 0)
    v1, v2 = m1 / rho, m2 / rho
    p = (G - 1) * (E - 0.5 * rho * (v1 ** 2 + v2 ** 2))
    T = p / rho

    def prim(d):
        r, a, b, e = np.moveaxis(d, -1, 0)
        w1, w2 = (a - r * v1) / rho, (b - r * v2) / rho
        pd = (G - 1) * (e - 0.5 * r * (v1 ** 2 + v2 ** 2) - rho * (v1 * w1 + v2 * w2))
        return w1, w2, (pd - r * T) / rho

    v1x, v2x, Tx = prim(ux)
    v1y, v2y, Ty = prim(uy)
    kappa = G / ((G - 1) * PR)
    t11, t12, t22 = 4 / 3 * v1x - 2 / 3 * v2y, v1y + v2x, 4 / 3 * v2y - 2 / 3 * v1x
    z = 0 * rho
    fv = MU * np.stack([z, t11, t12, v1 * t11 + v2 * t12 + kappa * Tx], -1)
    gv = MU * np.stack([z, t12, t22, v1 * t12 + v2 * t22 + kappa * Ty], -1)
    fa = np.stack([m1, m1 * v1 + p, m1 * v2, (E + p) * v1], -1)
    ga = np.stack([m2, m2 * v1, m2 * v2 + p, (E + p) * v2], -1)
    return fa - fv, ga - gv


def test_source_is_the_navier_stokes_residual(oracle_mod):
    rng = np.random.default_rng(0)
    X, Y, t = rng.uniform(-1, 1, 40), rng.uniform(-0.95, 0.95, 40), 0.37

    def U(x, y, tt):
        return oracle_mod.ns_convergence(x, y, tt, G, MU, PR)[0]

    def residual(h):
        def flux(x, y):
            ux = (U(x + h, y, t) - U(x - h, y, t)) / (2 * h)
            uy = (U(x, y + h, t) - U(x, y - h, t)) / (2 * h)
            return np_fluxes(U(x, y, t), ux, uy)

        return ((U(X, Y, t + h) - U(X, Y, t - h)) / (2 * h) + (flux(X + h, Y)[0] - flux(X - h, Y)[0]) / (2 * h) +
                (flux(X, Y + h)[1] - flux(X, Y - h)[1]) / (2 * h))

    s = oracle_mod.ns_convergence(X, Y, t, G, MU, PR)[1]
    e2, e1 = np.abs(residual(2e-3) - s).max(), np.abs(residual(1e-3) - s).max()
    assert e1 < 1e-5 * np.abs(s).max()
    assert 3.0 < e2 / e1 < 5.0  # the differences' own second-order error


def test_manufactured_velocity_vanishes_on_the_walls(oracle_mod):
    x = np.linspace(-1, 1, 33)
    for y in (-1.0, 1.0):
        for t in (0.0, 0.3, 0.7):
            u, _ = oracle_mod.ns_convergence(x, np.full_like(x, y), t, G, MU, PR)
            assert np.max(np.abs(u[:, 1:3])) < 1e-15


@pytest.mark.parametrize("N,mdrk,orders", [(3, True, (2.7, 3.5)), (4, False, (4.3, 4.7))])
def test_example_converges_at_order_n_plus_one(oracle_mod, N, mdrk, orders):
    errs = []
    for nx in (4, 8, 16):
        o = oracle_mod.Oracle(N, nx, nx, -1.0, -1.0, 2.0 / nx, gamma=G, periodic=(True, False),
                              bc_kind=oracle_mod.BV_SLIP_WALL, parabolic="navier_stokes", mu=MU, prandtl=PR,
                              ns_walls=[(0.0, 0.0)] * 4, source_terms="navier_stokes_convergence_test")
        x, y = o.node_coords()
        u0 = oracle_mod.ns_convergence(x, y, 0.0, G, MU, PR)[0]
        u1, t, _ = o.solve(u0, 0.0, 0.1, 0.5, mdrk=mdrk)
        assert t == 0.1
        ue = oracle_mod.ns_convergence(x, y, t, G, MU, PR)[0]
        _, w, _, _ = o.basis()
        W = np.outer(w, w)[None] * (1.0 / nx) ** 2
        errs.append(np.sqrt(np.einsum("eji,eji->", W, (u1[..., 0] - ue[..., 0]) ** 2)))
    eoc = [np.log2(errs[k] / errs[k + 1]) for k in range(2)]
    assert eoc[0] > orders[0] and eoc[1] > orders[1], (errs, eoc)


W4 = np.array([1.0, 0.5, -0.3, 2.0])  # a different weight per equation


@pytest.mark.parametrize("N", [3, 4])
def test_source_weights_on_a_uniform_state(oracle_mod, N):
    t, dt = 0.3, 0.01
    for deg in (1, 2, 3, 4):
        a = [0.3, -0.7, 0.5, 0.9, -0.4][:deg + 1]
        P = np.polynomial.Polynomial(a)
        d1p, d2p, d3p, d4p = (P.deriv(k)(t) for k in (1, 2, 3, 4))

        def src(u, x, y, tt, P=P):
            return P(tt) * W4

        o = oracle_mod.Oracle(N, 4, 4, 0.0, 0.0, 0.25, gamma=G, parabolic="navier_stokes", mu=MU, prandtl=PR,
                              source_terms=src)
        x, _ = o.node_coords()
        u = np.empty(x.shape + (4,))
        u[:] = [1.0, 0.1, 0.2, 1 / 0.4 + 0.5 * (0.01 + 0.04)]
        I = P.integ()
        avg = (I(t + dt) - I(t)) / dt
        lw = o.rhs(u, t, dt)
        excess = dt ** 4 * d4p / 180.0 if N == 3 else 0.0
        np.testing.assert_allclose(lw, np.broadcast_to((avg + excess) * W4, u.shape), rtol=1e-12, atol=1e-13)
        d1 = o.rhs(u, t, dt, kind="mdrk1")
        np.testing.assert_allclose(d1, np.broadcast_to((P(t) / 2 + dt * d1p / 8) * W4, u.shape), rtol=1e-12,
                                   atol=1e-13)
        if deg == 3:  # stage 1 is the half-step average to second order
            half = (I(t + dt / 2) - I(t)) / dt
            np.testing.assert_allclose(d1[0, 0, 0], (half - dt ** 2 * d2p / 48 - dt ** 3 * d3p / 384) * W4,
                                       rtol=1e-12)
        o.set_us(u + dt * d1)
        d2 = o.rhs(u, t, dt, kind="mdrk2")
        np.testing.assert_allclose(d2, np.broadcast_to((P(t) + dt * d1p / 2) * W4, u.shape), rtol=1e-12,
                                   atol=1e-13)
        if deg == 2:  # the stage-2 evaluation at t: short of the exact average by dt² s⁽²⁾/6
            np.testing.assert_allclose(d2[0, 0, 0], (avg - dt ** 2 * d2p / 6) * W4, rtol=1e-12)
