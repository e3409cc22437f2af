"""GPU parity of the reference's Navier-Stokes convergence test
(examples/tree_2d_dgsem/elixir_navierstokes_convergence.jl: MDRK, x periodic,
no-slip adiabatic walls at y = ±1 holding the manufactured velocity, μ 0.01,
Pr 0.72, source_terms_navier_stokes_convergence_test folded into the MDRK
stages as multi_derivative_2d_parabolic.jl:221-711 do, and into the LW
kernels 3/4 with walls as dg_2d_parabolic.jl:725-1860 do) against the CPU
oracle: both stages' residuals and the LW residual, K-step device-loop states
within 1e-12 relative, the device's own convergence order on 4², 8², 16² cells
(MDRK N=3, 4 and LW N=4), and the configurations refused (an LW context
without walls, a source whose μ differs from the equations').  Inputs are
seeded."""
import numpy as np
import pytest

from gpu_helpers import assert_close, to_dev, to_lex, torch_mod

pytestmark = pytest.mark.gpu

G, MU, PR = 1.4, 0.01, 0.72


def setup(oracle_mod, N, level, td="mdrk", source=None, walls=True):
    import trixilw_hip as tlw

    eq = tlw.CompressibleEulerEquations2D(G)
    eqp = tlw.CompressibleNavierStokesDiffusion2D(eq, mu=MU, Prandtl=PR,
                                                 gradient_variables=tlw.GradientVariablesConservative())
    solver = tlw.DGSEM(N, volume_integral=tlw.VolumeIntegralFR(tlw.MDRK() if td == "mdrk" else tlw.LW()))
    mesh = tlw.TreeMesh((-1.0, -1.0), (1.0, 1.0), level, periodicity=(True, not walls))
    ic = tlw.initial_condition_navier_stokes_convergence_test
    wall = tlw.BoundaryConditionNavierStokesWall(tlw.NoSlip(lambda x, t, e: ic(x[0], x[1], t, e)[1:3]),
                                                 tlw.Adiabatic(lambda x, t, e: 0.0))
    bc_hyp = {"x_neg": tlw.boundary_condition_periodic, "x_pos": tlw.boundary_condition_periodic,
              "y_neg": tlw.boundary_condition_slip_wall_horizontal,
              "y_pos": tlw.boundary_condition_slip_wall_horizontal}
    bc_par = {"x_neg": tlw.boundary_condition_periodic, "x_pos": tlw.boundary_condition_periodic,
              "y_neg": wall, "y_pos": wall}
    semi = tlw.SemidiscretizationHyperbolicParabolic(
        mesh, tlw.get_time_discretization(solver), (eq, eqp), ic, solver,
        boundary_conditions=(bc_hyp, bc_par) if walls else None,
        source_terms=source if source is not None else tlw.source_terms_navier_stokes_convergence_test)
    nx = mesh.nx
    o = oracle_mod.Oracle(N, nx, nx, -1.0, -1.0, 2.0 / nx, gamma=G, periodic=(True, False),
                          bc_kind=oracle_mod.BV_SLIP_WALL, parabolic="navier_stokes", mu=MU, prandtl=PR,
                          ns_walls=[(0.0, 0.0)] * 4, source_terms="navier_stokes_convergence_test")
    return semi, o, mesh


@pytest.mark.parametrize("N", [3, 4])
def test_ns_convergence_mdrk_stage_parity(oracle_mod, N):
    import trixilw_hip as tlw

    torch = torch_mod()
    print(f"NS convergence MDRK stages N={N}", flush=True)
    semi, o, mesh = setup(oracle_mod, N, 3)
    x, y = o.node_coords()
    t = 0.3
    u_lex = oracle_mod.ns_convergence(x, y, t, G, MU, PR)[0]
    u_lex = u_lex * (1 + 1e-3 * np.random.default_rng(N).standard_normal(u_lex.shape))
    dt = 0.4 * o.max_dt(u_lex)
    d1 = o.rhs(u_lex, t, dt, kind="mdrk1")
    us = u_lex + dt * d1
    o.set_us(us)
    d2 = o.rhs(u_lex, t, dt, kind="mdrk2")
    u = to_dev(mesh, u_lex)
    du = torch.empty_like(u)
    semi.cache["dt"][0] = dt
    tlw.rhs_mdrk1(du, u, semi, t)
    torch.cuda.synchronize()
    assert_close(to_lex(mesh, du), d1, what=f"NS convergence mdrk1 N={N}")
    semi.set_us(to_dev(mesh, us))
    tlw.rhs_mdrk2(du, u, semi, t)
    torch.cuda.synchronize()
    assert_close(to_lex(mesh, du), d2, what=f"NS convergence mdrk2 N={N}")
    semi.close()


@pytest.mark.parametrize("N", [3, 4])
def test_ns_convergence_lw_rhs_parity(oracle_mod, N):
    import trixilw_hip as tlw

    torch = torch_mod()
    print(f"NS convergence LW rhs N={N}", flush=True)
    semi, o, mesh = setup(oracle_mod, N, 3, td="lw")
    x, y = o.node_coords()
    t = 0.3
    u_lex = oracle_mod.ns_convergence(x, y, t, G, MU, PR)[0]
    u_lex = u_lex * (1 + 1e-3 * np.random.default_rng(10 + N).standard_normal(u_lex.shape))
    dt = 0.4 * o.max_dt(u_lex)
    d_ref = o.rhs(u_lex, t, dt)
    u = to_dev(mesh, u_lex)
    du = torch.empty_like(u)
    semi.cache["dt"][0] = dt
    tlw.rhs(du, u, semi, t)
    torch.cuda.synchronize()
    assert_close(to_lex(mesh, du), d_ref, what=f"NS convergence LW N={N}")
    semi.close()


@pytest.mark.parametrize("N,level,K,td", [(3, 3, 30, "mdrk"), (4, 3, 20, "mdrk"), (4, 3, 20, "lw"),
                                          (3, 3, 25, "lw")])
def test_ns_convergence_device_loop_k_steps(oracle_mod, N, level, K, td):
    import trixilw_hip as tlw

    print(f"NS convergence device loop N={N} level={level} K={K} {td}", flush=True)
    semi, o, mesh = setup(oracle_mod, N, level, td=td)
    x, y = o.node_coords()
    u_lex = oracle_mod.ns_convergence(x, y, 0.0, G, MU, PR)[0]
    cfl = 0.5
    u_ref, t_ref, _ = o.solve(u_lex, 0.0, 1e9, cfl, max_steps=K, mdrk=(td == "mdrk"))
    lw_update = tlw.semidiscretize(semi, semi.time_discretization, (0.0, float("inf")))
    lw_update.u0.copy_(to_dev(mesh, u_lex))
    sol = tlw.solve_lwfr(lw_update, (), 1e-6, None, time_step_computation=tlw.CFLBased(cfl), max_steps=K)
    assert sol.stats.naccept == K
    assert abs(sol.t - t_ref) <= 1e-13 * t_ref
    assert_close(to_lex(mesh, sol.u), u_ref, what=f"NS convergence N={N} after {K} steps")
    semi.close()


@pytest.mark.parametrize("N,orders,td", [(3, (2.7, 3.5), "mdrk"), (4, (4.3, 4.7), "mdrk"), (4, (4.3, 4.7), "lw")])
def test_ns_convergence_device_eoc(oracle_mod, N, orders, td):
    """The device solves the example to t = 0.1 on 4², 8², 16² cells: the
    density's L2 error against the manufactured solution falls at order ≈ N + 1,
    and each final state matches the oracle's solve."""
    import trixilw_hip as tlw

    errs = []
    for level in (2, 3, 4):
        print(f"NS convergence EOC N={N} level={level} {td}", flush=True)
        semi, o, mesh = setup(oracle_mod, N, level, td=td)
        x, y = o.node_coords()
        u_lex = oracle_mod.ns_convergence(x, y, 0.0, G, MU, PR)[0]
        lw_update = tlw.semidiscretize(semi, semi.time_discretization, (0.0, 0.1))
        lw_update.u0.copy_(to_dev(mesh, u_lex))
        sol = tlw.solve_lwfr(lw_update, (), 1e-6, None, time_step_computation=tlw.CFLBased(0.5))
        assert abs(sol.t - 0.1) < 1e-14
        got = to_lex(mesh, sol.u)
        u_ref, _, _ = o.solve(u_lex, 0.0, 0.1, 0.5, mdrk=(td == "mdrk"))
        assert_close(got, u_ref, what=f"NS convergence solve to t=0.1 N={N} level={level}")
        ue = oracle_mod.ns_convergence(x, y, 0.1, G, MU, PR)[0]
        _, w, _, _ = o.basis()
        W = np.outer(w, w)[None] * (2.0 / mesh.nx / 2.0) ** 2
        errs.append(np.sqrt(np.einsum("eji,eji->", W, (got[..., 0] - ue[..., 0]) ** 2)))
        semi.close()
    eoc = [np.log2(errs[k] / errs[k + 1]) for k in range(2)]
    print(f"NS convergence EOC N={N}: errors {errs}, orders {eoc}", flush=True)
    assert eoc[0] > orders[0] and eoc[1] > orders[1], (errs, eoc)


def test_ns_convergence_refusals(oracle_mod):
    import trixilw_hip as tlw
    from trixilw_hip import _lib

    torch_mod()
    with pytest.raises(_lib.TLWError) as ei:  # an LW context without walls
        setup(oracle_mod, 3, 2, td="lw", walls=False)
    assert ei.value.code == _lib.TLW_ENOTSUP
    with pytest.raises(NotImplementedError):  # a source whose μ is not the equations'
        setup(oracle_mod, 3, 2, source=tlw.SourceTermsNavierStokesConvergenceTest(mu=0.02))
