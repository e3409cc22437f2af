"""GPU parity of the Navier-Stokes LW and MDRK paths with no-slip adiabatic and isothermal walls
(BoundaryConditionNavierStokesWall{NoSlip, Adiabatic | Isothermal}, conservative gradient
variables, with the reflecting wall for the hyperbolic part: the reference's
lid-driven cavity, examples/tree_2d_dgsem/elixir_navierstokes_lid_driven_cavity.jl)
against the CPU oracle: residuals of a perturbed cavity state and of plane
Couette flow, both MDRK stages (multi_derivative_2d_parabolic.jl:59-218),
K-step device-loop states within 1e-12 relative (L2 and L∞), and the
combinations the path rejects.  Inputs are seeded."""
import numpy as np
import pytest

from gpu_helpers import assert_close, to_dev, to_lex, torch_mod

pytestmark = pytest.mark.gpu

GAMMA, PR = 1.4, 0.72
LID = (1.0, 0.0)


def cavity_state(x, y, t=0.0):
    """initial_condition_cavity of the reference example: ρ = 1, v = 0,
    p = 1 / (Ma² γ) with Ma = 0.1."""
    rho = np.ones_like(x)
    p = np.full_like(x, 1.0 / (0.1 ** 2 * GAMMA))
    z = np.zeros_like(x)
    return np.stack([rho, z, z, p / (GAMMA - 1)], axis=-1)


def setup_walls(oracle_mod, N, level, mu=0.01, periodic_x=False, ic=cavity_state, lid=LID, td="lw",
                heat=(None, None, None, None)):
    """``heat``: per direction (x_neg, x_pos, y_neg, y_pos) None for Adiabatic(0) or a
    wall temperature for Isothermal(T)."""
    import trixilw_hip as tlw

    eq = tlw.CompressibleEulerEquations2D(GAMMA)
    eqp = tlw.CompressibleNavierStokesDiffusion2D(eq, mu=mu, Prandtl=PR,
                                                 gradient_variables=tlw.GradientVariablesConservative())
    mesh = tlw.TreeMesh((-1.0, -1.0), (1.0, 1.0), level, periodicity=(periodic_x, False))
    solver = tlw.DGSEM(N, volume_integral=tlw.VolumeIntegralFR(tlw.MDRK() if td == "mdrk" else tlw.LW()))
    def wall(v, T):
        return tlw.BoundaryConditionNavierStokesWall(tlw.NoSlip(lambda x, t, e: v),
                                                     tlw.Adiabatic(0.0) if T is None else tlw.Isothermal(T))
    bc_hyp = {"x_neg": tlw.boundary_condition_slip_wall_vertical, "x_pos": tlw.boundary_condition_slip_wall_vertical,
              "y_neg": tlw.boundary_condition_slip_wall_horizontal,
              "y_pos": tlw.boundary_condition_slip_wall_horizontal}
    bc_par = {"x_neg": wall((0.0, 0.0), heat[0]), "x_pos": wall((0.0, 0.0), heat[1]),
              "y_neg": wall((0.0, 0.0), heat[2]), "y_pos": wall(lid, heat[3])}
    semi = tlw.SemidiscretizationHyperbolicParabolic(mesh, tlw.get_time_discretization(solver), (eq, eqp), ic,
                                                     solver, boundary_conditions=(bc_hyp, bc_par))
    nx = mesh.nx
    o = oracle_mod.Oracle(N, nx, nx, -1.0, -1.0, 2.0 / nx, gamma=GAMMA, periodic=(periodic_x, False),
                          bc_kind=oracle_mod.BV_SLIP_WALL, parabolic="navier_stokes", mu=mu, prandtl=PR,
                          ns_walls=[(0.0, 0.0), (0.0, 0.0), (0.0, 0.0), lid], wall_temperature=list(heat))
    return semi, o, mesh


def perturbed(o, seed, amp=0.05):
    """The cavity state with seeded velocity and density perturbations."""
    x, y = o.node_coords()
    u = cavity_state(x, y)
    rng = np.random.default_rng(seed)
    rho = u[..., 0] * (1 + amp * rng.standard_normal(x.shape))
    v1 = 0.3 * np.sin(np.pi * (x + 1) / 2) * np.sin(np.pi * (y + 1)) + amp * rng.standard_normal(x.shape)
    v2 = 0.2 * np.cos(np.pi * x / 2) * np.sin(np.pi * y) + amp * rng.standard_normal(x.shape)
    p = (GAMMA - 1) * u[..., 3]
    return np.stack([rho, rho * v1, rho * v2, p / (GAMMA - 1) + 0.5 * rho * (v1 ** 2 + v2 ** 2)], axis=-1)


@pytest.mark.parametrize("N", [3, 4])
def test_cavity_rhs_parity(oracle_mod, N):
    import trixilw_hip as tlw

    torch = torch_mod()
    print(f"NS walls cavity rhs N={N}", flush=True)
    semi, o, mesh = setup_walls(oracle_mod, N, 3)
    u_lex = perturbed(o, 30 + N)
    dt = 0.8 * o.max_dt(u_lex)
    du_ref = o.rhs(u_lex, 0.0, dt)
    u = to_dev(mesh, u_lex)
    du = torch.empty_like(u)
    semi.cache["dt"][0] = dt
    tlw.rhs(du, u, semi, 0.0)
    torch.cuda.synchronize()
    assert_close(to_lex(mesh, du), du_ref, what=f"NS walls cavity N={N}")
    semi.close()


@pytest.mark.parametrize("N", [3, 4])
def test_couette_rhs_parity(oracle_mod, N):
    """Plane Couette flow between the resting wall y = -1 and the lid y = 1
    (x periodic): energy grows by μ (U/H)², mass and momentum stay."""
    import trixilw_hip as tlw

    torch = torch_mod()
    print(f"NS walls Couette N={N}", flush=True)
    U, mu = 0.5, 0.02

    def couette(x, y, t=0.0):
        v1 = U * (y + 1.0) / 2.0
        z = np.zeros_like(x)
        return np.stack([z + 1.0, v1, z, z + 1.0 / (GAMMA - 1) + 0.5 * v1 ** 2], axis=-1)

    semi, o, mesh = setup_walls(oracle_mod, N, 2, mu=mu, periodic_x=True, ic=couette, lid=(U, 0.0))
    x, y = o.node_coords()
    u_lex = couette(x, y)
    dt = 0.5 * o.max_dt(u_lex)
    du_ref = o.rhs(u_lex, 0.0, dt)
    u = to_dev(mesh, u_lex)
    du = torch.empty_like(u)
    semi.cache["dt"][0] = dt
    tlw.rhs(du, u, semi, 0.0)
    torch.cuda.synchronize()
    got = to_lex(mesh, du)
    # the residual is the small remainder of O(J⁻¹ |F|) ≈ 10 terms that cancel, so the
    # two computations are compared in absolute terms at that scale
    assert np.max(np.abs(got - du_ref)) < 1e-11
    assert np.max(np.abs(got[..., 3] - mu * (U / 2.0) ** 2)) < 1e-10
    assert np.max(np.abs(got[..., :3])) < 1e-12
    semi.close()


@pytest.mark.parametrize("N,K", [(4, 20), (3, 25)])
def test_cavity_device_loop_k_steps(oracle_mod, N, K):
    import trixilw_hip as tlw

    print(f"NS walls cavity device loop N={N} K={K}", flush=True)
    semi, o, mesh = setup_walls(oracle_mod, N, 3)
    u_lex = perturbed(o, 40 + N, amp=0.02)
    cfl = 0.5
    u_ref, t_ref, dts = o.solve(u_lex, 0.0, 1e9, cfl, max_steps=K)
    lw_update = tlw.semidiscretize(semi, semi.time_discretization, (0.0, float("inf")))
    lw_update.u0.copy_(to_dev(mesh, u_lex))
    sol = tlw.solve_lwfr(lw_update, (), 1e-6, None, time_step_computation=tlw.CFLBased(cfl), max_steps=K)
    assert sol.stats.naccept == K
    assert abs(sol.t - t_ref) <= 1e-13 * t_ref
    assert_close(to_lex(mesh, sol.u), u_ref, what=f"NS walls cavity N={N} after {K} steps")
    semi.close()


def test_ns_walls_reject_unsupported(oracle_mod):
    import trixilw_hip as tlw
    from trixilw_hip import _lib

    torch_mod()
    # an isothermal wall needs a finite positive temperature
    eq0 = tlw.CompressibleEulerEquations2D(GAMMA)
    eqp0 = tlw.CompressibleNavierStokesDiffusion2D(eq0, mu=0.01, Prandtl=PR,
                                                  gradient_variables=tlw.GradientVariablesConservative())
    hot = tlw.BoundaryConditionNavierStokesWall(tlw.NoSlip((0.0, 0.0)), tlw.Isothermal(-1.0))
    bc_hyp = (tlw.boundary_condition_slip_wall_vertical, tlw.boundary_condition_slip_wall_vertical,
              tlw.boundary_condition_slip_wall_horizontal, tlw.boundary_condition_slip_wall_horizontal)
    with pytest.raises(_lib.TLWError) as ei:
        tlw.SemidiscretizationHyperbolicParabolic(tlw.TreeMesh((-1.0, -1.0), (1.0, 1.0), 2, periodicity=False),
                                                  tlw.LW(), (eq0, eqp0), cavity_state,
                                                  tlw.DGSEM(3, volume_integral=tlw.VolumeIntegralFR(tlw.LW())),
                                                  boundary_conditions=(bc_hyp, (hot,) * 4))
    assert ei.value.code == _lib.TLW_EINVAL
    eq = tlw.CompressibleEulerEquations2D(GAMMA)
    eqp = tlw.CompressibleNavierStokesDiffusion2D(eq, mu=0.01, Prandtl=PR,
                                                 gradient_variables=tlw.GradientVariablesConservative())
    solver = tlw.DGSEM(3, volume_integral=tlw.VolumeIntegralFR(tlw.LW()))
    mesh = tlw.TreeMesh((-1.0, -1.0), (1.0, 1.0), 2, periodicity=False)
    with pytest.raises(NotImplementedError):  # a non-periodic mesh without parabolic walls
        tlw.SemidiscretizationHyperbolicParabolic(mesh, tlw.LW(), (eq, eqp), cavity_state, solver,
                                                  boundary_conditions=(tlw.boundary_condition_slip_wall_vertical,
                                                                       None))


# cavity temperature T = p/ρ (the reference's R = 1) is 1/(Ma² γ) ≈ 71.4: walls a little hotter or colder
HOT = (70.0, 73.0, 71.0, 72.5)


@pytest.mark.parametrize("N", [3, 4])
@pytest.mark.parametrize("heat", [HOT, (None, 73.0, 70.5, None)])
def test_isothermal_cavity_rhs_parity(oracle_mod, N, heat):
    """Isothermal walls (compressible_navier_stokes_2d.jl:93-111): the gradient's
    boundary value prim2cons(ρ, v_wall, ρ T_wall) and the inner viscous flux as the
    wall's viscous flux, alone and mixed with adiabatic walls."""
    import trixilw_hip as tlw

    torch = torch_mod()
    print(f"NS isothermal walls rhs N={N} heat={heat}", flush=True)
    semi, o, mesh = setup_walls(oracle_mod, N, 3, heat=heat)
    u_lex = perturbed(o, 60 + N)
    u = to_dev(mesh, u_lex)
    du = torch.empty_like(u)
    dt = 0.4 * o.max_dt(u_lex)
    semi.cache["dt"][0] = dt
    du_ref = o.rhs(u_lex, 0.0, dt)
    tlw.rhs(du, u, semi, 0.0)
    torch.cuda.synchronize()
    assert_close(to_lex(mesh, du), du_ref, what=f"NS isothermal walls N={N}")
    semi.close()


@pytest.mark.parametrize("N,K", [(4, 20), (3, 25)])
def test_isothermal_cavity_device_loop_k_steps(oracle_mod, N, K):
    import trixilw_hip as tlw

    print(f"NS isothermal walls device loop N={N} K={K}", flush=True)
    semi, o, mesh = setup_walls(oracle_mod, N, 3, heat=HOT)
    u_lex = perturbed(o, 70 + N, amp=0.02)
    cfl = 0.5
    u_ref, t_ref, _ = o.solve(u_lex, 0.0, 1e9, cfl, max_steps=K)
    lw_update = tlw.semidiscretize(semi, semi.time_discretization, (0.0, float("inf")))
    lw_update.u0.copy_(to_dev(mesh, u_lex))
    sol = tlw.solve_lwfr(lw_update, (), 1e-6, None, time_step_computation=tlw.CFLBased(cfl), max_steps=K)
    assert sol.stats.naccept == K
    assert abs(sol.t - t_ref) <= 1e-13 * t_ref
    assert_close(to_lex(mesh, sol.u), u_ref, what=f"NS isothermal walls N={N} after {K} steps")
    semi.close()


@pytest.mark.parametrize("N", [3, 4])
@pytest.mark.parametrize("heat", [(None,) * 4, (None, 73.0, 70.5, None)])
def test_cavity_mdrk_stage_parity(oracle_mod, N, heat):
    """Both MDRK stages with walls: stage 1 takes the walls' viscous average
    f_v/2 + f_v,t/8 (and the hyperbolic wall with its 0.5 scaling), stage 2
    f_v + f_v,t/6 + f_v,t(us)/3 with the BR1 gradient of us at the walls."""
    import trixilw_hip as tlw

    torch = torch_mod()
    print(f"NS walls MDRK stages N={N} heat={heat}", flush=True)
    semi, o, mesh = setup_walls(oracle_mod, N, 3, td="mdrk", heat=heat)
    u_lex = perturbed(o, 80 + N)
    dt = 0.4 * o.max_dt(u_lex)
    d1 = o.rhs(u_lex, 0.0, dt, kind="mdrk1")
    us = u_lex + dt * d1
    o.set_us(us)
    d2 = o.rhs(u_lex, 0.0, dt, kind="mdrk2")
    u = to_dev(mesh, u_lex)
    du = torch.empty_like(u)
    semi.cache["dt"][0] = dt
    tlw.rhs_mdrk1(du, u, semi, 0.0)
    torch.cuda.synchronize()
    assert_close(to_lex(mesh, du), d1, what=f"NS walls mdrk1 N={N}")
    semi.set_us(to_dev(mesh, us))
    tlw.rhs_mdrk2(du, u, semi, 0.0)
    torch.cuda.synchronize()
    assert_close(to_lex(mesh, du), d2, what=f"NS walls mdrk2 N={N}")
    semi.close()


@pytest.mark.parametrize("N", [3, 4])
def test_couette_mdrk_stage_parity(oracle_mod, N):
    """Couette flow: stage 1 heats at half the viscous rate, stage 2 at the full
    rate; mass and momentum stay."""
    import trixilw_hip as tlw

    torch = torch_mod()
    print(f"NS walls Couette MDRK N={N}", flush=True)
    U, mu = 0.5, 0.02

    def couette(x, y, t=0.0):
        v1 = U * (y + 1.0) / 2.0
        z = np.zeros_like(x)
        return np.stack([z + 1.0, v1, z, z + 1.0 / (GAMMA - 1) + 0.5 * v1 ** 2], axis=-1)

    semi, o, mesh = setup_walls(oracle_mod, N, 2, mu=mu, periodic_x=True, ic=couette, lid=(U, 0.0), td="mdrk")
    x, y = o.node_coords()
    u_lex = couette(x, y)
    dt = 0.5 * o.max_dt(u_lex)
    u = to_dev(mesh, u_lex)
    du = torch.empty_like(u)
    semi.cache["dt"][0] = dt
    tlw.rhs_mdrk1(du, u, semi, 0.0)
    torch.cuda.synchronize()
    g1 = to_lex(mesh, du)
    semi.set_us(to_dev(mesh, u_lex + dt * g1))
    tlw.rhs_mdrk2(du, u, semi, 0.0)
    torch.cuda.synchronize()
    g2 = to_lex(mesh, du)
    for g, weight in ((g1, 0.5), (g2, 1.0)):
        assert np.max(np.abs(g[..., :3])) < 1e-12
        assert np.max(np.abs(g[..., 3] - weight * mu * (U / 2.0) ** 2)) < 1e-10
    semi.close()


@pytest.mark.parametrize("N,K,heat", [(4, 15, (None,) * 4), (3, 20, HOT)])
def test_cavity_mdrk_device_loop_k_steps(oracle_mod, N, K, heat):
    import trixilw_hip as tlw

    print(f"NS walls MDRK device loop N={N} K={K} heat={heat}", flush=True)
    semi, o, mesh = setup_walls(oracle_mod, N, 3, td="mdrk", heat=heat)
    u_lex = perturbed(o, 90 + N, amp=0.02)
    cfl = 0.25
    u_ref, t_ref, _ = o.solve(u_lex, 0.0, 1e9, cfl, max_steps=K, mdrk=True)
    lw_update = tlw.semidiscretize(semi, semi.time_discretization, (0.0, float("inf")))
    lw_update.u0.copy_(to_dev(mesh, u_lex))
    sol = tlw.solve_lwfr(lw_update, (), 1e-6, None, time_step_computation=tlw.CFLBased(cfl), max_steps=K)
    assert sol.stats.naccept == K
    assert abs(sol.t - t_ref) <= 1e-13 * t_ref
    assert_close(to_lex(mesh, sol.u), u_ref, what=f"NS walls MDRK N={N} after {K} steps")
    semi.close()


@pytest.mark.parametrize("dt_initial", [1e-5, 0.01])
def test_cavity_adaptive_matches_oracle(oracle_mod, dt_initial):
    """Adaptive steps with walls (the reference's Ghia utility runs the cavity
    with Adaptive()): the LW N=4 wall kernel with the temporal error of
    dg_2d_parabolic.jl:1749-1775, 1841-1855; accepted and rejected counts,
    final time and state equal the oracle's solve_adaptive (≤1e-12), from a
    small dt_initial and from an unstable one."""
    import trixilw_hip as tlw

    torch_mod()
    print(f"NS walls adaptive dt0={dt_initial}", flush=True)
    tol = {"abstol": 1e-6, "reltol": 1e-6}
    semi, _, mesh = setup_walls(oracle_mod, 4, 3, heat=(None, 73.0, None, None))
    o = oracle_mod.Oracle(4, mesh.nx, mesh.nx, -1.0, -1.0, 2.0 / mesh.nx, gamma=GAMMA, periodic=(False, False),
                          bc_kind=oracle_mod.BV_SLIP_WALL, parabolic="navier_stokes", mu=0.01, prandtl=PR,
                          ns_walls=[(0.0, 0.0), (0.0, 0.0), (0.0, 0.0), LID],
                          wall_temperature=[None, 73.0, None, None], tolerances=(tol["abstol"], tol["reltol"]))
    u_lex = perturbed(o, 50, amp=0.02)
    tf = 0.02
    u_ref, t_ref, dts, nrej = o.solve_adaptive(u_lex, 0.0, tf, dt_initial, max_attempts=20000)
    lw_update = tlw.semidiscretize(semi, semi.time_discretization, (0.0, tf))
    lw_update.u0.copy_(to_dev(mesh, u_lex))
    sol = tlw.solve_lwfr(lw_update, (), dt_initial, tol, time_step_computation=tlw.Adaptive())
    assert sol.stats.naccept == len(dts) > 3
    assert sol.stats.nreject == nrej
    if dt_initial > 1e-3:
        assert nrej >= 1
    assert abs(sol.t - t_ref) <= 1e-13 * t_ref
    assert_close(to_lex(mesh, sol.u), u_ref, what=f"NS walls adaptive dt0={dt_initial}")
    semi.close()
