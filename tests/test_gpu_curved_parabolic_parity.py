"""GPU parity of compressible Navier-Stokes on the periodic curvilinear mesh
(conservative gradient variables) against the CPU oracle: LW residuals and both
MDRK stages at N=3, 4 on warped meshes, compute_dt, every volume tile, K-step
device-loop states within 1e-12 relative (L2 and L∞), free stream and
conservation at 256² with the host's metric terms, and the combinations the
path rejects.  Parity runs the kernels on the oracle's metric terms.  Inputs
are seeded."""
import numpy as np
import pytest

from gpu_helpers import GAMMA, assert_close, to_dev, to_lex, torch_mod, warp_mapping
from test_gpu_parabolic_parity import shear_state

pytestmark = pytest.mark.gpu

PR, L = 0.72, 2.0


def state(x, y, t=0.0):
    """The smooth shear/density/pressure waves of the TreeMesh tests, moved to
    [0, 2]² (periodic)."""
    return shear_state(x - 1.0, y - 1.0)


def setup_cns(oracle_mod, N, nx, td="lw", mu=0.01, A=0.08, keep_aux=False, seed=0, amp=1e-3, ordering="lexicographic"):
    import trixilw_hip as tlw

    eq = tlw.CompressibleEulerEquations2D(GAMMA)
    eqp = tlw.CompressibleNavierStokesDiffusion2D(eq, mu=mu, Prandtl=PR,
                                                 gradient_variables=tlw.GradientVariablesConservative())
    mesh = tlw.StructuredMesh((nx, nx), warp_mapping(L, A), ordering=ordering)
    solver = tlw.DGSEM(N, volume_integral=tlw.VolumeIntegralFR(tlw.MDRK() if td == "mdrk" else tlw.LW()))
    o = oracle_mod.Oracle(N, nx, nx, 0.0, 0.0, L / nx, gamma=GAMMA, mapping=(oracle_mod.MAP_WARP, (L, A)),
                          parabolic="navier_stokes", mu=mu, prandtl=PR)
    mesh.set_metrics(solver.basis, o.aux("contravariant_vectors"), o.aux("inverse_jacobian"))
    semi = tlw.SemidiscretizationHyperbolicParabolic(mesh, tlw.get_time_discretization(solver), (eq, eqp), state,
                                                     solver, keep_aux=keep_aux)
    x, y = o.node_coords()
    rng = np.random.default_rng(seed)
    u_lex = state(x, y) * (1.0 + amp * rng.standard_normal(x.shape + (4,)))
    return semi, o, u_lex, mesh


@pytest.mark.parametrize("N", [3, 4])
@pytest.mark.parametrize("A", [0.05, 0.12])
def test_curved_ns_rhs_lw_parity(oracle_mod, N, A):
    import trixilw_hip as tlw

    torch = torch_mod()
    print(f"curved NS LW N={N} A={A}", flush=True)
    semi, o, u_lex, mesh = setup_cns(oracle_mod, N, 8, A=A, keep_aux=True)
    dt = 0.9 * o.max_dt(u_lex)
    du_ref = o.rhs(u_lex, 0.0, dt)
    u = to_dev(mesh, u_lex)
    du = torch.empty_like(u)
    semi.cache["dt"][0] = dt
    tlw.rhs(du, u, semi, 0.0)
    torch.cuda.synchronize()
    assert_close(to_lex(mesh, du), du_ref, what=f"curved NS LW N={N} A={A}")
    # the viscous part is exercised: the Euler residual on the same mesh differs
    e = oracle_mod.Oracle(N, 8, 8, 0.0, 0.0, L / 8, gamma=GAMMA, mapping=(oracle_mod.MAP_WARP, (L, A)))
    assert np.max(np.abs(e.rhs(u_lex, 0.0, dt) - du_ref)) > 1e-3 * np.max(np.abs(du_ref))
    semi.close()


@pytest.mark.parametrize("N", [3, 4])
def test_curved_ns_rhs_mdrk_parity(oracle_mod, N):
    import trixilw_hip as tlw

    torch = torch_mod()
    print(f"curved NS MDRK N={N}", flush=True)
    semi, o, u_lex, mesh = setup_cns(oracle_mod, N, 8, td="mdrk")
    dt = 0.5 * o.max_dt(u_lex)
    du1_ref = o.rhs(u_lex, 0.2, dt, kind="mdrk1")
    o.set_us(u_lex + dt * du1_ref)
    du2_ref = o.rhs(u_lex, 0.2, dt, kind="mdrk2")
    u = to_dev(mesh, u_lex)
    du = torch.empty_like(u)
    semi.cache["dt"][0] = dt
    tlw.rhs_mdrk1(du, u, semi, 0.2)
    torch.cuda.synchronize()
    assert_close(to_lex(mesh, du), du1_ref, what=f"curved NS mdrk1 N={N}")
    semi.set_us(torch.add(u, du, alpha=dt))
    tlw.rhs_mdrk2(du, u, semi, 0.2)
    torch.cuda.synchronize()
    assert_close(to_lex(mesh, du), du2_ref, what=f"curved NS mdrk2 N={N}")
    semi.close()


@pytest.mark.parametrize("N", [3, 4])
def test_curved_ns_max_dt_parity(oracle_mod, N):
    print(f"curved NS max_dt N={N}", flush=True)
    for mu in (0.01, 0.0005):
        semi, o, u_lex, mesh = setup_cns(oracle_mod, N, 8, mu=mu)
        got = semi.max_dt(to_dev(mesh, u_lex))
        ref = o.max_dt(u_lex)
        assert abs(got - ref) <= 1e-14 * ref, (mu, got, ref)
        semi.close()


def test_curved_ns_every_tile_matches_oracle(oracle_mod):
    import trixilw_hip as tlw

    torch = torch_mod()
    semi, o, u_lex, mesh = setup_cns(oracle_mod, 4, 8)
    dt = 0.9 * o.max_dt(u_lex)
    du_ref = o.rhs(u_lex, 0.0, dt)
    u = to_dev(mesh, u_lex)
    semi.cache["dt"][0] = dt
    for eb in (5, 10, 15):
        print(f"curved NS tile {eb}", flush=True)
        semi.elements_per_workgroup(eb)
        du = torch.empty_like(u)
        tlw.rhs(du, u, semi, 0.0)
        torch.cuda.synchronize()
        assert_close(to_lex(mesh, du), du_ref, what=f"curved NS tile {eb}")
    semi.close()


@pytest.mark.parametrize("N,td,K", [(4, "lw", 20), (3, "lw", 20), (3, "mdrk", 15), (4, "mdrk", 10)])
def test_curved_ns_device_loop_k_steps(oracle_mod, N, td, K):
    """K steps of the device loop (viscous and curvilinear CFL bounds in the
    surface kernel) against the oracle's solve loop on a warped 8×8 mesh."""
    import trixilw_hip as tlw

    print(f"curved NS device loop N={N} {td} K={K}", flush=True)
    semi, o, u_lex, mesh = setup_cns(oracle_mod, N, 8, td=td)
    cfl = 0.5
    u_ref, t_ref, dts = o.solve(u_lex, 0.0, 1e9, cfl, max_steps=K, mdrk=(td == "mdrk"))
    lw_update = tlw.semidiscretize(semi, semi.time_discretization, (0.0, float("inf")))
    lw_update.u0.copy_(to_dev(mesh, u_lex))
    sol = tlw.solve_lwfr(lw_update, (), 1e-6, None, time_step_computation=tlw.CFLBased(cfl), max_steps=K)
    assert sol.stats.naccept == K
    assert abs(sol.t - t_ref) <= 1e-13 * t_ref
    assert_close(to_lex(mesh, sol.u), u_ref, what=f"curved NS {td} N={N} after {K} steps")
    semi.close()


def test_curved_ns_free_stream_and_conservation_at_scale():
    """256×256 warped mesh (1.6 M DOFs, the host's own metric terms): a
    constant state stays constant to the rounding of the discrete metric
    identities, and 20 device-loop steps of the shear state conserve
    Σ w_i w_j J u."""
    import trixilw_hip as tlw

    torch = torch_mod()
    eq = tlw.CompressibleEulerEquations2D(GAMMA)
    eqp = tlw.CompressibleNavierStokesDiffusion2D(eq, mu=0.002, Prandtl=PR,
                                                 gradient_variables=tlw.GradientVariablesConservative())
    mesh = tlw.StructuredMesh((256, 256), warp_mapping(L, 0.08))
    solver = tlw.DGSEM(4, volume_integral=tlw.VolumeIntegralFR(tlw.LW()))
    const = np.array([1.1, 0.4, -0.3, 2.7])

    def free_stream(x, y, t):
        return np.broadcast_to(const, x.shape + (4,))

    semi = tlw.SemidiscretizationHyperbolicParabolic(mesh, tlw.LW(), (eq, eqp), free_stream, solver)
    lw_update = tlw.semidiscretize(semi, tlw.LW(), (0.0, 1e9))
    sol = tlw.solve_lwfr(lw_update, (), 1e-6, None, time_step_computation=tlw.CFLBased(0.5), max_steps=10)
    c = torch.tensor(const, dtype=torch.float64, device=sol.u.device)
    # the metric identities hold to round-off, scaled by J⁻¹ ≈ (2/h)² = 65536 here
    assert (sol.u - c).abs().max().item() <= 1e-9
    semi.close()
    semi = tlw.SemidiscretizationHyperbolicParabolic(mesh, tlw.LW(), (eq, eqp), state, solver)
    lw_update = tlw.semidiscretize(semi, tlw.LW(), (0.0, 1e9))
    u0 = lw_update.u0.clone()
    sol = tlw.solve_lwfr(lw_update, (), 1e-6, None, time_step_computation=tlw.CFLBased(0.5), max_steps=20)
    assert sol.stats.naccept == 20
    _, ijn = mesh.metrics(solver.basis)
    w = np.outer(solver.basis.weights, solver.basis.weights)[None] / ijn
    wt = torch.from_numpy(np.ascontiguousarray(w)).cuda()
    m0 = torch.einsum("eji,ejiv->v", wt, u0)
    m1 = torch.einsum("eji,ejiv->v", wt, sol.u)
    scale = torch.einsum("eji,ejiv->v", wt, u0.abs())
    # the totals are sums of 1.6 M weighted terms (2.3e-13 of a total 4 seen): conservation
    # is exact up to that summation round-off
    assert torch.all((m1 - m0).abs() <= 1e-12 * scale), (m0, m1)
    assert torch.isfinite(sol.u).all()
    semi.close()


def _tgv_semi(tlw, nx, metrics_from=None):
    """The curved Navier-Stokes configuration's setup: MDRK N=3, Taylor-Green
    vortex (U 1, ρ0 1, Mach 0.1) on [0,2]² warped by amplitude 0.1, μ = 0.001
    (Re 1000), Pr 0.72."""
    eq = tlw.CompressibleEulerEquations2D(GAMMA)
    eqp = tlw.CompressibleNavierStokesDiffusion2D(eq, mu=1e-3, Prandtl=PR,
                                                 gradient_variables=tlw.GradientVariablesConservative())
    mesh = tlw.StructuredMesh((nx, nx), warp_mapping(L, 0.1), ordering="strips")
    solver = tlw.DGSEM(3, volume_integral=tlw.VolumeIntegralFR(tlw.MDRK()))
    if metrics_from is not None:
        mesh.set_metrics(solver.basis, metrics_from.aux("contravariant_vectors"), metrics_from.aux("inverse_jacobian"))
    ic = tlw.TaylorGreenVortex(U=1.0, rho0=1.0, mach=0.1, L=L)
    return tlw.SemidiscretizationHyperbolicParabolic(mesh, tlw.MDRK(), (eq, eqp), ic, solver), mesh, ic, solver


def test_c2_config_tgv_mdrk_parity(oracle_mod):
    """The curved Navier-Stokes configuration's parameters (warp 0.1, Re 1000,
    MDRK N=3, Taylor-Green vortex, CFL 0.5) on 32² for 50 device-loop steps,
    strips element order, against the oracle's solve ≤ 1e-12."""
    import trixilw_hip as tlw

    torch_mod()
    nx, K, cfl = 32, 50, 0.5
    o = oracle_mod.Oracle(3, nx, nx, 0.0, 0.0, L / nx, gamma=GAMMA, mapping=(oracle_mod.MAP_WARP, (L, 0.1)),
                          parabolic="navier_stokes", mu=1e-3, prandtl=PR)
    semi, mesh, ic, solver = _tgv_semi(tlw, nx, metrics_from=o)
    x, y = o.node_coords()
    u_lex = ic(x, y)
    u_ref, t_ref, _ = o.solve(u_lex, 0.0, 1e9, cfl, max_steps=K, mdrk=True)
    upd = tlw.semidiscretize(semi, tlw.MDRK(), (0.0, 1e9))
    upd.u0.copy_(to_dev(mesh, u_lex))
    sol = tlw.solve_lwfr(upd, (), 1e-6, None, time_step_computation=tlw.CFLBased(cfl), max_steps=K)
    assert sol.stats.naccept == K
    assert abs(sol.t - t_ref) <= 1e-13 * t_ref
    assert_close(to_lex(mesh, sol.u), u_ref, what="TGV MDRK 32² x 50 steps")
    semi.close()


def test_c2_config_128x128_300_steps_properties():
    """The configuration at its full size (128², 300 MDRK steps, the host's
    own metric terms, as the benchmark runs it): the state stays finite,
    Σ w_i w_j J u of every variable is conserved to summation round-off, and
    the kinetic energy decays monotonically (viscosity, no forcing)."""
    import trixilw_hip as tlw

    torch = torch_mod()
    semi, mesh, ic, solver = _tgv_semi(tlw, 128)
    _, ijn = mesh.metrics(solver.basis)
    w = np.outer(solver.basis.weights, solver.basis.weights)[None] / ijn
    wt = torch.from_numpy(np.ascontiguousarray(w)).cuda()
    upd = tlw.semidiscretize(semi, tlw.MDRK(), (0.0, 1e9))
    u = upd.u0.clone()
    m0 = torch.einsum("eji,ejiv->v", wt, u)
    scale = torch.einsum("eji,ejiv->v", wt, u.abs())

    def kinetic(v):
        return float(torch.einsum("eji,eji->", wt, 0.5 * (v[..., 1] ** 2 + v[..., 2] ** 2) / v[..., 0]))

    ke = [kinetic(u)]
    t = 0.0
    for chunk in range(6):  # 6 x 50 steps, each chunk a fresh solve from the last state
        upd = tlw.LWUpdate(upd.f, u, upd.soln_arrays, (t, 1e9), semi)
        sol = tlw.solve_lwfr(upd, (), 1e-6, None, time_step_computation=tlw.CFLBased(0.5), max_steps=50)
        assert sol.stats.naccept == 50
        u, t = sol.u, sol.t
        assert torch.isfinite(u).all()
        ke.append(kinetic(u))
    m1 = torch.einsum("eji,ejiv->v", wt, u)
    assert torch.all((m1 - m0).abs() <= 1e-12 * scale), (m0, m1)
    assert all(b < a for a, b in zip(ke, ke[1:])), ke
    # Re 1000 over t ≈ 300 dt: a small but clearly resolved decay
    assert 1e-6 < 1.0 - ke[-1] / ke[0] < 0.5, ke
    semi.close()


def test_curved_ns_rejects_unsupported(oracle_mod):
    import trixilw_hip as tlw
    from trixilw_hip import _lib

    torch_mod()
    eq = tlw.CompressibleEulerEquations2D(GAMMA)
    eqp = tlw.CompressibleNavierStokesDiffusion2D(eq, mu=0.01, Prandtl=PR,
                                                 gradient_variables=tlw.GradientVariablesConservative())
    solver = tlw.DGSEM(3, volume_integral=tlw.VolumeIntegralFR(tlw.LW()))
    closed = tlw.StructuredMesh((4, 4), warp_mapping(L, 0.05), periodicity=False)
    with pytest.raises(_lib.TLWError) as ei:
        tlw.SemidiscretizationHyperbolicParabolic(closed, tlw.LW(), (eq, eqp), state, solver)
    assert ei.value.code in (_lib.TLW_ENOTSUP, _lib.TLW_EINVAL)
    semi, o, u_lex, mesh = setup_cns(oracle_mod, 3, 4)
    with pytest.raises(_lib.TLWError) as ei:
        semi.rebuild_mesh(mesh)
    assert ei.value.code == _lib.TLW_ENOTSUP
    semi.close()


@pytest.mark.parametrize("td", ["lw", "mdrk"])
def test_curved_ns_strips_ordering_parity(oracle_mod, td):
    """The strips element order (two-row strips, the bench's order) gives the
    oracle's residuals and MDRK stages at every tile size."""
    import trixilw_hip as tlw

    torch = torch_mod()
    semi, o, u_lex, mesh = setup_cns(oracle_mod, 3, 8, td=td, ordering="strips")
    dt = 0.5 * o.max_dt(u_lex)
    kind = "mdrk1" if td == "mdrk" else "lw"
    du_ref = o.rhs(u_lex, 0.1, dt, kind=kind)
    u = to_dev(mesh, u_lex)
    semi.cache["dt"][0] = dt
    for eb in (4, 8, 16):
        print(f"curved NS strips {td} tile {eb}", flush=True)
        semi.elements_per_workgroup(eb)
        du = torch.empty_like(u)
        (tlw.rhs_mdrk1 if td == "mdrk" else tlw.rhs)(du, u, semi, 0.1)
        torch.cuda.synchronize()
        assert_close(to_lex(mesh, du), du_ref, what=f"curved NS strips {td} tile {eb}")
    semi.close()
