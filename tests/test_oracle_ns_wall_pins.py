"""Pins of the oracle's no-slip adiabatic walls of the Navier-Stokes LW path
(BoundaryConditionNavierStokesWall{NoSlip, Adiabatic} with conservative
gradient variables, src/equations/compressible_navier_stokes_2d.jl:81-124,
with the reflecting wall for the hyperbolic part, as in the reference's
lid-driven cavity; no GPU).

- Plane Couette flow (ρ = 1, p = 1, v1 = U y / H, x periodic, wall y = 0 at
  rest, wall y = H moving with U): the gradient, the viscous stress and the
  viscous work are polynomials the scheme represents exactly, so mass and
  momentum do not change and the energy grows by the viscous heating
  μ (U/H)² at every node.
- The BR1 gradient at a wall takes the wall state u_outer = prim2cons(ρ, v_wall,
  p) as its interface value: for a linear state that slips along the wall the
  gradient there is the exact one plus (u_outer - u)(2/h)/w_n (the lifting of
  the full difference; an average would lift half of it).
- Mass is conserved with walls; a fluid at rest stays at rest.
- MDRK with walls (multi_derivative_2d_parabolic.jl:59-218: both stages call
  the hyperbolic wall flux and the wall's viscous flux from the stage's time
  average F̃v): in Couette flow the energy residual of stage 1 is half the
  viscous heating (F̃ = f/2 + f_t/8 with f_t = 0) and that of stage 2 the full
  heating; mass is conserved by both stages; an MDRK solve with walls and an
  LW solve at small steps reach the same state (the walls of the LW path are
  pinned above, and the difference of the two time integrations is O(dt)).
"""
import numpy as np
import pytest

MU, PR, H = 0.01, 0.72, 1.0


def walled(oracle_mod, N, nx, lid=0.3, periodic_x=True):
    return oracle_mod.Oracle(N, nx, nx, 0.0, 0.0, H / nx, periodic=(periodic_x, False),
                             bc_kind=oracle_mod.BV_SLIP_WALL, parabolic="navier_stokes", mu=MU, prandtl=PR,
                             ns_walls=[(0.0, 0.0), (0.0, 0.0), (0.0, 0.0), (lid, 0.0)])


def shear(x, y, a, b):
    rho, p = np.ones_like(x), np.ones_like(x)
    v1 = a + b * y
    return np.stack([rho, rho * v1, 0 * x, p / 0.4 + 0.5 * rho * v1 ** 2], axis=-1)


@pytest.mark.parametrize("N", [3, 4])
def test_couette_flow_heats_at_the_viscous_rate(oracle_mod, N):
    U = 0.3
    o = walled(oracle_mod, N, 4, lid=U)
    x, y = o.node_coords()
    u = shear(x, y, 0.0, U / H)
    du = o.rhs(u, 0.0, 0.5 * o.max_dt(u))
    assert np.max(np.abs(du[..., 0])) < 1e-13
    assert np.max(np.abs(du[..., 1:3])) < 1e-13
    # absolute round-off of the O(J⁻¹ |F|) terms that cancel (J⁻¹ = 8 here)
    assert np.max(np.abs(du[..., 3] - MU * U ** 2 / H ** 2)) < 1e-10


@pytest.mark.parametrize("N", [3, 4])
def test_wall_gradient_lifts_the_full_wall_state(oracle_mod, N):
    nx, a, b, U = 4, 0.05, 0.2, 0.3
    o = walled(oracle_mod, N, nx, lid=U)
    x, y = o.node_coords()
    u = shear(x, y, a, b)
    o.rhs(u, 0.0, 1e-4)
    gy = o.aux("gradients_y").reshape(u.shape)
    _, w, _, _ = o.basis()
    n, h = N + 1, H / nx
    top = [e for e in range(nx * nx) if e // nx == nx - 1]
    bottom = [e for e in range(nx * nx) if e // nx == 0]
    lift = (2.0 / h) / w[-1]
    # momentum component: d(ρ v1)/dy = b, plus the lifted jump to the wall value
    np.testing.assert_allclose(gy[top, n - 1, :, 1], b + (U - (a + b * H)) * lift, rtol=0, atol=1e-11)
    np.testing.assert_allclose(gy[bottom, 0, :, 1], b - (0.0 - a) * lift, rtol=0, atol=1e-11)
    # away from the walls the gradient is the exact one
    np.testing.assert_allclose(gy[top, 1:n - 1, :, 1], b, rtol=0, atol=1e-11)


@pytest.mark.parametrize("N", [3, 4])
def test_walls_conserve_mass_and_keep_rest(oracle_mod, N):
    from trixilw_hip.equations import IsentropicVortex

    o = walled(oracle_mod, N, 4, periodic_x=False)
    x, y = o.node_coords()
    rest = shear(x, y, 0.0, 0.0)
    o_rest = walled(oracle_mod, N, 4, lid=0.0, periodic_x=False)
    assert np.max(np.abs(o_rest.rhs(rest, 0.0, 1e-3))) < 1e-12
    ic = IsentropicVortex(center=(0.5, 0.5), period=(1e9, 1e9), strength=1.0)
    rng = np.random.default_rng(5 + N)
    u = ic(x, y, 0.0) * (1 + 1e-2 * rng.standard_normal(x.shape + (4,)))
    du = o.rhs(u, 0.0, 0.5 * o.max_dt(u))
    _, w, _, _ = o.basis()
    W = np.outer(w, w)[None]
    total = np.einsum("eji,eji->", W, du[..., 0])
    assert abs(total) <= 1e-13 * np.einsum("eji,eji->", W, np.abs(du[..., 0]))


@pytest.mark.parametrize("N", [3, 4])
def test_isothermal_wall_gradient_lifts_the_wall_temperature_state(oracle_mod, N):
    """An isothermal wall's state is prim2cons(ρ, v_wall, ρ T_wall): for a
    linear shear at uniform T = 1 the energy gradient at the lid is the exact
    one plus the lifted jump to the wall state's energy."""
    nx, a, b, U, Tw = 4, 0.05, 0.2, 0.3, 1.5
    o = oracle_mod.Oracle(N, nx, nx, 0.0, 0.0, H / nx, periodic=(True, False), bc_kind=oracle_mod.BV_SLIP_WALL,
                          parabolic="navier_stokes", mu=MU, prandtl=PR,
                          ns_walls=[(0.0, 0.0), (0.0, 0.0), (0.0, 0.0), (U, 0.0)],
                          wall_temperature=[None, None, None, Tw])
    x, y = o.node_coords()
    u = shear(x, y, a, b)
    o.rhs(u, 0.0, 1e-4)
    gy = o.aux("gradients_y").reshape(u.shape)
    _, w, _, _ = o.basis()
    n, h = N + 1, H / nx
    top = [e for e in range(nx * nx) if e // nx == nx - 1]
    v_top = a + b * H
    E_in = 1.0 / 0.4 + 0.5 * v_top ** 2
    E_out = Tw / 0.4 + 0.5 * U ** 2
    dEdy = b * v_top  # d/dy of 0.5 (a + b y)^2 at y = H
    np.testing.assert_allclose(gy[top, n - 1, :, 3], dEdy + (E_out - E_in) * (2.0 / h) / w[-1], rtol=0, atol=1e-10)


@pytest.mark.parametrize("N", [3, 4])
def test_isothermal_walls_pass_the_heat_flux(oracle_mod, N):
    """Fluid at rest with a linear temperature profile between walls held at
    its own end temperatures: the conduction flux is uniform, so with
    isothermal walls (which pass the inner viscous flux) the energy does not
    change, while adiabatic walls block the flux and the wall elements heat or
    cool (a tiny dt keeps the time-derivative terms of the hydrostatic
    imbalance out of the comparison)."""
    nx = 4
    x0 = 0.0
    common = dict(periodic=(True, False), bc_kind=oracle_mod.BV_SLIP_WALL, parabolic="navier_stokes", mu=MU,
                  prandtl=PR, ns_walls=[(0.0, 0.0)] * 4)
    iso = oracle_mod.Oracle(N, nx, nx, x0, 0.0, H / nx, wall_temperature=[None, None, 1.0, 1.2], **common)
    adi = oracle_mod.Oracle(N, nx, nx, x0, 0.0, H / nx, **common)
    x, y = iso.node_coords()
    T = 1.0 + 0.2 * y / H
    rho = np.ones_like(x)
    u = np.stack([rho, 0 * x, 0 * x, rho * T / 0.4], axis=-1)
    dt = 1e-9
    du_iso = iso.rhs(u, 0.0, dt)
    du_adi = adi.rhs(u, 0.0, dt)
    # (2e-8 seen at the wall elements: the dt-sized terms of the hydrostatic imbalance)
    assert np.max(np.abs(du_iso[..., 3])) < 1e-6
    assert np.max(np.abs(du_adi[..., 3])) > 1e-3


@pytest.mark.parametrize("N", [3, 4])
def test_mdrk_couette_stages_heat_at_their_weights(oracle_mod, N):
    U = 0.3
    o = walled(oracle_mod, N, 4, lid=U)
    x, y = o.node_coords()
    u = shear(x, y, 0.0, U / H)
    dt = 0.5 * o.max_dt(u)
    heat = MU * U ** 2 / H ** 2
    d1 = o.rhs(u, 0.0, dt, kind="mdrk1")
    o.set_us(u + dt * d1)
    d2 = o.rhs(u, 0.0, dt, kind="mdrk2")
    for d, weight in ((d1, 0.5), (d2, 1.0)):
        assert np.max(np.abs(d[..., :3])) < 1e-13
        assert np.max(np.abs(d[..., 3] - weight * heat)) < 1e-10


@pytest.mark.parametrize("N", [3, 4])
def test_mdrk_walls_conserve_mass(oracle_mod, N):
    from trixilw_hip.equations import IsentropicVortex

    o = walled(oracle_mod, N, 4, periodic_x=False)
    x, y = o.node_coords()
    ic = IsentropicVortex(center=(0.5, 0.5), period=(1e9, 1e9), strength=1.0)
    rng = np.random.default_rng(11 + N)
    u = ic(x, y, 0.0) * (1 + 1e-2 * rng.standard_normal(x.shape + (4,)))
    _, w, _, _ = o.basis()
    W = np.outer(w, w)[None]
    dt = 0.25 * o.max_dt(u)
    d1 = o.rhs(u, 0.0, dt, kind="mdrk1")
    o.set_us(u + dt * d1)
    d2 = o.rhs(u, 0.0, dt, kind="mdrk2")
    for d in (d1, d2):
        assert abs(np.einsum("eji,eji->", W, d[..., 0])) <= 1e-13 * np.einsum("eji,eji->", W, np.abs(d[..., 0]))
    u1, _, dts = o.solve(u, 0.0, 1e9, 0.2, max_steps=10, mdrk=True)
    assert len(dts) == 10
    m0, m1 = np.einsum("eji,eji->", W, u[..., 0]), np.einsum("eji,eji->", W, u1[..., 0])
    assert abs(m1 - m0) <= 1e-14 * m0


@pytest.mark.parametrize("N", [3, 4])
def test_mdrk_with_walls_agrees_with_lw_at_small_steps(oracle_mod, N):
    """A cavity with a moving lid from a density bump at rest: the MDRK and LW
    solves at CFL 0.025 reach t = 0.05 within 1e-7 (each moves by 6e-5 from
    CFL 0.025 to 0.00625); MDRK without stage 2's wall fluxes breaks this bound
    (stage 1 enters the step at O(dt²): the Couette stages pin it)."""
    o = walled(oracle_mod, N, 4, periodic_x=False)
    x, y = o.node_coords()
    rho = 1.0 + 0.1 * np.sin(np.pi * x) * np.sin(np.pi * y)
    u = np.stack([rho, 0 * x, 0 * x, 1 / 0.4 + 0 * x], axis=-1)
    u_md, t_md, _ = o.solve(u, 0.0, 0.05, 0.025, mdrk=True)
    u_lw, t_lw, _ = o.solve(u, 0.0, 0.05, 0.025, mdrk=False)
    assert t_md == t_lw == 0.05
    assert np.max(np.abs(u_md - u_lw)) < 1e-7
