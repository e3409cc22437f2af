"""GPU parity of the hyperbolic-parabolic path (compressible Navier-Stokes,
conservative gradient variables, periodic TreeMesh) against the CPU oracle:
rhs / rhs_mdrk1 / rhs_mdrk2 residuals, compute_dt and K-step device-loop states
within 1e-12 relative (L2 and L∞).  Inputs are seeded."""
import numpy as np
import pytest

from gpu_helpers import assert_close, to_dev, to_lex, torch_mod

pytestmark = pytest.mark.gpu

GAMMA, PR = 1.4, 0.72


def shear_state(x, y, t=0.0, gamma=GAMMA):
    """Smooth periodic state on [-1, 1]² with shear, density and pressure waves
    (an initial condition: t is not used)."""
    rho = 1.0 + 0.2 * np.sin(np.pi * x) * np.cos(np.pi * y)
    v1 = 0.3 + 0.1 * np.sin(np.pi * y)
    v2 = -0.2 + 0.1 * np.cos(np.pi * x)
    p = 1.0 + 0.1 * np.cos(np.pi * (x + y))
    return np.stack([rho, rho * v1, rho * v2, p / (gamma - 1) + 0.5 * rho * (v1**2 + v2**2)], axis=-1)


def setup_ns(oracle_mod, N, level, td="lw", mu=0.002, keep_aux=False, seed=0, amp=1e-3, ordering="morton"):
    import trixilw_hip as tlw

    eq = tlw.CompressibleEulerEquations2D(GAMMA)
    eqp = tlw.CompressibleNavierStokesDiffusion2D(eq, mu=mu, Prandtl=PR,
                                                 gradient_variables=tlw.GradientVariablesConservative())
    mesh = tlw.TreeMesh((-1.0, -1.0), (1.0, 1.0), level, ordering=ordering)
    solver = tlw.DGSEM(N, volume_integral=tlw.VolumeIntegralFR(tlw.MDRK() if td == "mdrk" else tlw.LW()))
    semi = tlw.SemidiscretizationHyperbolicParabolic(mesh, tlw.get_time_discretization(solver), (eq, eqp),
                                                     shear_state, solver, keep_aux=keep_aux)
    nx = mesh.nx
    o = oracle_mod.Oracle(N, nx, nx, -1.0, -1.0, 2.0 / nx, gamma=GAMMA, parabolic="navier_stokes", mu=mu,
                          prandtl=PR)
    x, y = o.node_coords()
    rng = np.random.default_rng(seed)
    u_lex = shear_state(x, y) * (1.0 + amp * rng.standard_normal(x.shape + (4,)))
    return semi, o, u_lex, mesh


@pytest.mark.parametrize("N", [3, 4])
def test_parabolic_rhs_lw_parity(oracle_mod, N):
    import trixilw_hip as tlw

    torch = torch_mod()
    semi, o, u_lex, mesh = setup_ns(oracle_mod, N, 4, mu=0.01)
    dt = 0.9 * o.max_dt(u_lex)
    du_ref = o.rhs(u_lex, 0.0, dt)
    u = to_dev(mesh, u_lex)
    du = torch.empty_like(u)
    semi.cache["dt"][0] = dt
    tlw.rhs(du, u, semi, 0.0)
    torch.cuda.synchronize()
    assert_close(to_lex(mesh, du), du_ref, what=f"NS LW N={N}")
    # the viscous part is not negligible at this μ: the Euler residual differs
    e = oracle_mod.Oracle(N, mesh.nx, mesh.nx, -1.0, -1.0, 2.0 / mesh.nx, gamma=GAMMA)
    du_euler = e.rhs(u_lex, 0.0, dt)
    assert np.max(np.abs(du_euler - du_ref)) > 1e-3 * np.max(np.abs(du_ref))


@pytest.mark.parametrize("N", [3, 4])
def test_parabolic_rhs_mdrk_parity(oracle_mod, N):
    import trixilw_hip as tlw

    torch = torch_mod()
    semi, o, u_lex, mesh = setup_ns(oracle_mod, N, 4, td="mdrk", mu=0.01)
    dt = 0.5 * o.max_dt(u_lex)
    du1_ref = o.rhs(u_lex, 0.2, dt, kind="mdrk1")
    o.set_us(u_lex + dt * du1_ref)
    du2_ref = o.rhs(u_lex, 0.2, dt, kind="mdrk2")
    u = to_dev(mesh, u_lex)
    du = torch.empty_like(u)
    semi.cache["dt"][0] = dt
    tlw.rhs_mdrk1(du, u, semi, 0.2)
    assert_close(to_lex(mesh, du), du1_ref, what=f"NS mdrk1 N={N}")
    semi.set_us(torch.add(u, du, alpha=dt))
    tlw.rhs_mdrk2(du, u, semi, 0.2)
    torch.cuda.synchronize()
    assert_close(to_lex(mesh, du), du2_ref, what=f"NS mdrk2 N={N}")


@pytest.mark.parametrize("N", [3, 4])
@pytest.mark.parametrize("mu", [1e-5, 0.01])
def test_parabolic_max_dt_parity(oracle_mod, N, mu):
    """compute_dt's min of the advective and parabolic bounds.  The parabolic
    one, 1/((λa + λv)·2(2N+1)), contains the advective speeds too, so it is
    the active one at every μ."""
    semi, o, u_lex, mesh = setup_ns(oracle_mod, N, 4, mu=mu)
    ref = o.max_dt(u_lex)
    got = semi.max_dt(to_dev(mesh, u_lex))
    assert abs(got - ref) <= 1e-14 * ref
    e = oracle_mod.Oracle(N, mesh.nx, mesh.nx, -1.0, -1.0, 2.0 / mesh.nx, gamma=GAMMA)
    adv = e.max_dt(u_lex)
    assert ref < adv


def test_every_tile_matches_oracle(oracle_mod):
    import trixilw_hip as tlw

    torch = torch_mod()
    for N in (3, 4):
        semi, o, u_lex, mesh = setup_ns(oracle_mod, N, 4, mu=0.01, seed=3)
        dt = 0.8 * o.max_dt(u_lex)
        du_ref = o.rhs(u_lex, 0.0, dt)
        u = to_dev(mesh, u_lex)
        du = torch.empty_like(u)
        semi.cache["dt"][0] = dt
        for eb in ((5, 10, 15) if N == 4 else (4, 8, 16)):
            semi.elements_per_workgroup(eb)
            du.fill_(np.nan)
            tlw.rhs(du, u, semi, 0.0)
            torch.cuda.synchronize()
            assert_close(to_lex(mesh, du), du_ref, what=f"NS N={N} tile {eb}")
        semi.close()


@pytest.mark.parametrize("ordering", ["morton", "strips"])
@pytest.mark.parametrize("N,td,K", [(4, "lw", 25), (3, "lw", 25), (3, "mdrk", 20), (4, "mdrk", 15)])
def test_parabolic_device_loop_k_steps(oracle_mod, N, td, K, ordering):
    """K steps of the device loop (viscous CFL bound active) against the
    oracle's solve loop on a 16×16 mesh."""
    import trixilw_hip as tlw

    semi, o, u_lex, mesh = setup_ns(oracle_mod, N, 4, td=td, mu=0.002, ordering=ordering)
    cfl = 0.9 if td == "lw" else 0.45
    u_ref, t_ref, dts = o.solve(u_lex, 0.0, 1e9, cfl, max_steps=K, mdrk=(td == "mdrk"))
    lw_update = tlw.semidiscretize(semi, semi.time_discretization, (0.0, float("inf")))
    lw_update.u0.copy_(to_dev(mesh, u_lex))
    sol = tlw.solve_lwfr(lw_update, (), 1e-6, None, time_step_computation=tlw.CFLBased(cfl), max_steps=K)
    assert sol.stats.naccept == K
    assert abs(sol.t - t_ref) <= 1e-13 * t_ref
    assert_close(to_lex(mesh, sol.u), u_ref, what=f"NS {td} N={N} after {K} steps")


def test_parabolic_host_loop_matches_device_loop(oracle_mod):
    """solve_lwfr's host loop (rhs + compute_dt per step) and the device loop
    take the same steps."""
    import trixilw_hip as tlw

    semi, o, u_lex, mesh = setup_ns(oracle_mod, 3, 4, mu=0.002)
    u_ref, t_ref, _ = o.solve(u_lex, 0.0, 1e9, 0.9, max_steps=10)
    lw_update = tlw.semidiscretize(semi, semi.time_discretization, (0.0, float("inf")))
    lw_update.u0.copy_(to_dev(mesh, u_lex))
    sol = tlw.solve_lwfr(lw_update, (), 1e-6, None, time_step_computation=tlw.CFLBased(0.9), device_loop=False,
                         max_steps=10)
    assert abs(sol.t - t_ref) <= 1e-13 * t_ref
    assert_close(to_lex(mesh, sol.u), u_ref, what="NS host loop after 10 steps")


def test_parabolic_conservation_at_scale():
    """256×256 periodic mesh (1.6 M DOFs): 20 device-loop steps conserve
    mass, momentum and energy and dissipate kinetic energy."""
    import trixilw_hip as tlw

    torch = torch_mod()
    eq = tlw.CompressibleEulerEquations2D(GAMMA)
    eqp = tlw.CompressibleNavierStokesDiffusion2D(eq, mu=1e-4, Prandtl=PR,
                                                 gradient_variables=tlw.GradientVariablesConservative())
    mesh = tlw.TreeMesh((-1.0, -1.0), (1.0, 1.0), 8)
    solver = tlw.DGSEM(4, volume_integral=tlw.VolumeIntegralFR(tlw.LW()))
    semi = tlw.SemidiscretizationHyperbolicParabolic(mesh, tlw.LW(), (eq, eqp), shear_state, solver)
    lw_update = tlw.semidiscretize(semi, tlw.LW(), (0.0, 1e9))
    u0 = lw_update.u0.clone()
    sol = tlw.solve_lwfr(lw_update, (), 1e-6, None, time_step_computation=tlw.CFLBased(0.9), max_steps=20)
    w = torch.from_numpy(np.outer(solver.basis.weights, solver.basis.weights)).cuda()
    m0 = torch.einsum("ji,ejiv->v", w, u0)
    m1 = torch.einsum("ji,ejiv->v", w, sol.u)
    scale = torch.einsum("ji,ejiv->v", w, u0.abs())
    assert torch.all((m1 - m0).abs() <= 1e-13 * scale), (m0, m1)
    assert torch.isfinite(sol.u).all()
    semi.close()


def test_parabolic_rejects_unsupported(oracle_mod):
    import trixilw_hip as tlw
    from trixilw_hip import _lib

    torch = torch_mod()
    semi, o, u_lex, mesh = setup_ns(oracle_mod, 3, 3)
    u = to_dev(mesh, u_lex)
    with pytest.raises(_lib.TLWError) as ei:
        tlw.adjoint_step(torch.empty_like(u), u, u.clone(), semi, 0.0, 1e-3)
    assert ei.value.code == _lib.TLW_ENOTSUP
    eq = tlw.CompressibleEulerEquations2D(GAMMA)
    with pytest.raises(NotImplementedError):
        tlw.CompressibleNavierStokesDiffusion2D(eq, mu=0.01, Prandtl=PR)  # primitive gradient variables
    semi.close()
