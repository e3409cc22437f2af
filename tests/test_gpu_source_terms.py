"""GPU parity of the source terms (source_terms_convergence_test folded into
the LW N=3/4 and MDRK Taylor procedures, dg_2d.jl:89-176) against the oracle:
rhs, both MDRK stages, device-loop steps with and without Dirichlet data, and
the manufactured solution's errors on the device equal to the oracle's.
Inputs are seeded; tolerances 1e-12 relative (L2 and L∞)."""
import numpy as np
import pytest

from gpu_helpers import assert_close, to_dev, to_lex, torch_mod

pytestmark = pytest.mark.gpu


def setup_src(oracle_mod, N, nx, td="lw", periodic=True, seed=0, amp=1e-3, ordering="lexicographic"):
    """[0, 2]² TreeMesh of nx² elements, the convergence test's state (perturbed)
    with its source, periodic or with BoundaryConditionDirichlet of its state."""
    import trixilw_hip as tlw

    eq = tlw.CompressibleEulerEquations2D(1.4)
    level = int(round(np.log2(nx)))
    mesh = tlw.TreeMesh((0.0, 0.0), (2.0, 2.0), level, periodicity=periodic, ordering=ordering)
    solver = tlw.DGSEM(N, volume_integral=tlw.VolumeIntegralFR(tlw.MDRK() if td == "mdrk" else tlw.LW()))
    ic = tlw.initial_condition_convergence_test
    bc = None if periodic else tlw.BoundaryConditionDirichlet(ic)
    semi = tlw.SemidiscretizationHyperbolic(mesh, tlw.get_time_discretization(solver), eq, ic, solver,
                                            boundary_conditions=bc, source_terms=tlw.source_terms_convergence_test)
    okw = {} if periodic else dict(periodic=(False, False), bc_kind=oracle_mod.BV_CONVERGENCE_TEST)
    o = oracle_mod.Oracle(N, nx, nx, 0.0, 0.0, 2.0 / nx, source_terms="convergence_test", **okw)
    x, y = o.node_coords()
    rng = np.random.default_rng(seed)
    u_lex = ic(x, y, 0.0) * (1.0 + amp * rng.standard_normal(x.shape + (4,)))
    return semi, o, u_lex, mesh, solver


@pytest.mark.parametrize("N", [3, 4])
@pytest.mark.parametrize("periodic", [True, False])
def test_rhs_lw_with_source_terms(oracle_mod, N, periodic):
    import trixilw_hip as tlw

    torch = torch_mod()
    semi, o, u_lex, mesh, _ = setup_src(oracle_mod, N, 8, periodic=periodic)
    t = 0.37
    dt = 0.8 * o.max_dt(u_lex)
    du_ref = o.rhs(u_lex, t, dt)
    u = to_dev(mesh, u_lex)
    du = torch.empty_like(u)
    semi.cache["dt"][0] = dt
    tlw.rhs(du, u, semi, t)
    torch.cuda.synchronize()
    assert_close(to_lex(mesh, du), du_ref, what=f"LW N={N} periodic={periodic}")
    # the source is exercised: without it the residual differs
    plain = oracle_mod.Oracle(N, 8, 8, 0.0, 0.0, 0.25, **({} if periodic else dict(
        periodic=(False, False), bc_kind=oracle_mod.BV_CONVERGENCE_TEST)))
    assert np.max(np.abs(plain.rhs(u_lex, t, dt) - du_ref)) > 1e-2 * np.max(np.abs(du_ref))
    semi.close()


@pytest.mark.parametrize("N", [3, 4])
def test_rhs_mdrk_stages_with_source_terms(oracle_mod, N):
    import trixilw_hip as tlw

    torch = torch_mod()
    semi, o, u_lex, mesh, _ = setup_src(oracle_mod, N, 8, td="mdrk", periodic=False)
    t = 0.21
    dt = 0.5 * o.max_dt(u_lex)
    du1_ref = o.rhs(u_lex, t, dt, kind="mdrk1")
    o.set_us(u_lex + dt * du1_ref)
    du2_ref = o.rhs(u_lex, t, dt, kind="mdrk2")
    u = to_dev(mesh, u_lex)
    du = torch.empty_like(u)
    semi.cache["dt"][0] = dt
    tlw.rhs_mdrk1(du, u, semi, t)
    assert_close(to_lex(mesh, du), du1_ref, what=f"mdrk1 N={N}")
    semi.set_us(torch.add(u, du, alpha=dt))
    tlw.rhs_mdrk2(du, u, semi, t)
    torch.cuda.synchronize()
    assert_close(to_lex(mesh, du), du2_ref, what=f"mdrk2 N={N}")
    semi.close()


@pytest.mark.parametrize("ordering", ["morton", "strips"])
@pytest.mark.parametrize("N,td,periodic,K", [(4, "lw", True, 25), (3, "lw", False, 25), (3, "mdrk", True, 15),
                                              (4, "mdrk", False, 15)])
def test_device_loop_with_source_terms(oracle_mod, N, td, periodic, K, ordering):
    """K steps of solve_lwfr's device loop (t from the device state feeds the
    source) against the oracle's loop on 16²."""
    import trixilw_hip as tlw

    semi, o, u_lex, mesh, solver = setup_src(oracle_mod, N, 16, td=td, periodic=periodic, ordering=ordering)
    cfl = tlw.trixi2lw(0.9, solver) * (0.5 if td == "mdrk" else 1.0)
    u_ref, t_ref, dts = o.solve(u_lex, 0.1, 1e9, cfl, max_steps=K, mdrk=(td == "mdrk"))
    upd = tlw.semidiscretize(semi, semi.time_discretization, (0.1, float("inf")))
    upd.u0.copy_(to_dev(mesh, u_lex))
    sol = tlw.solve_lwfr(upd, (), 1e-6, None, time_step_computation=tlw.CFLBased(cfl), max_steps=K)
    assert sol.stats.naccept == K
    assert abs(sol.t - t_ref) <= 1e-13 * t_ref
    assert_close(to_lex(mesh, sol.u), u_ref, what=f"{td} N={N} periodic={periodic} after {K} steps")
    semi.close()


@pytest.mark.parametrize("N,td", [(4, "lw"), (3, "lw"), (3, "mdrk")])
def test_manufactured_solution_on_device(oracle_mod, N, td):
    """The reference's elixir_euler_source_terms.jl setup to t = 0.5 on 8², 16²,
    32²: the device's final states equal the oracle's (≤1e-12), so do their
    errors against the exact solution, and the errors fall at the scheme's
    order."""
    import trixilw_hip as tlw

    errs = []
    for nx in (8, 16, 32):
        semi, o, _, mesh, solver = setup_src(oracle_mod, N, nx, td=td)
        x, y = o.node_coords()
        u0 = tlw.initial_condition_convergence_test(x, y, 0.0)
        cfl = tlw.trixi2lw(0.9, solver) * (0.5 if td == "mdrk" else 1.0)
        u_ref, t_ref, _ = o.solve(u0, 0.0, 0.5, cfl, mdrk=(td == "mdrk"))
        upd = tlw.semidiscretize(semi, semi.time_discretization, (0.0, 0.5))
        upd.u0.copy_(to_dev(mesh, u0))
        sol = tlw.solve_lwfr(upd, (), 1e-6, None, time_step_computation=tlw.CFLBased(cfl))
        assert abs(sol.t - 0.5) <= 1e-13
        exact = tlw.initial_condition_convergence_test(x, y, 0.5)
        got = to_lex(mesh, sol.u)
        assert_close(got, u_ref, what=f"manufactured solution {nx}² at t = 0.5")
        e_dev = np.linalg.norm((got - exact).ravel()) / np.linalg.norm(exact.ravel())
        e_ref = np.linalg.norm((u_ref - exact).ravel()) / np.linalg.norm(exact.ravel())
        assert abs(e_dev - e_ref) <= 1e-12, (nx, e_dev, e_ref)
        errs.append(e_dev)
        semi.close()
    slope = -np.polyfit(np.arange(3), np.log2(errs), 1)[0]
    assert slope > (4.3 if N == 4 else 3.6), errs


def test_source_terms_reject_unsupported(oracle_mod):
    import trixilw_hip as tlw
    from trixilw_hip import _lib

    torch = torch_mod()
    eq = tlw.CompressibleEulerEquations2D(1.4)
    mesh = tlw.TreeMesh((0.0, 0.0), (2.0, 2.0), 2)
    ic = tlw.initial_condition_convergence_test
    src = tlw.source_terms_convergence_test
    ind = tlw.IndicatorHennemannGassner(eq, None, alpha_max=0.5)
    sc = tlw.DGSEM(4, volume_integral=tlw.VolumeIntegralFRShockCapturing(ind, volume_integralFR=tlw.VolumeIntegralFR(tlw.LW())))
    with pytest.raises(_lib.TLWError) as ei:
        tlw.SemidiscretizationHyperbolic(mesh, tlw.LW(), eq, ic, sc, source_terms=src)
    assert ei.value.code == _lib.TLW_ENOTSUP
    eqp = tlw.CompressibleNavierStokesDiffusion2D(eq, mu=0.01, Prandtl=0.72,
                                                 gradient_variables=tlw.GradientVariablesConservative())
    with pytest.raises(_lib.TLWError) as ei:
        tlw.SemidiscretizationHyperbolicParabolic(mesh, tlw.LW(), (eq, eqp), ic, tlw.DGSEM(4), source_terms=src)
    assert ei.value.code == _lib.TLW_ENOTSUP
    with pytest.raises(NotImplementedError):
        tlw.SemidiscretizationHyperbolic(mesh, tlw.LW(), eq, ic, tlw.DGSEM(4), source_terms=lambda u, x, t, e: u)
    semi = tlw.SemidiscretizationHyperbolic(mesh, tlw.LW(), eq, ic, tlw.DGSEM(4), source_terms=src)
    h = semi.handle
    with pytest.raises(_lib.TLWError) as ei:
        _lib.check(_lib.lib().tlw_solve_init_adaptive(h, 0.0, 1.0, 1e-3, 1e-6, 1e-6), h)
    assert ei.value.code == _lib.TLW_ENOTSUP
    x, y = semi.node_coordinates()
    u = torch.from_numpy(np.ascontiguousarray(ic(x, y, 0.0))).cuda()
    with pytest.raises(_lib.TLWError) as ei:
        tlw.adjoint_step(torch.empty_like(u), u, u.clone(), semi, 0.0, 1e-3)
    assert ei.value.code == _lib.TLW_ENOTSUP
    semi.close()
