"""GPU parity of source terms on the curvilinear StructuredMesh (the source
lines of lw_volume_kernel_3!/4! in src/solvers/dgsem_structured/dg_2d.jl:
1379-2105; the reference's elixir_euler_source_terms_parallelogram.jl and
_nonperiodic.jl) against the CPU oracle: the LW residual with Trixi's
source_terms_convergence_test on the parallelogram x = ξ + η, y = η, on a
warped mesh and on a box with the convergence test's Dirichlet data; K-step
device-loop states within 1e-12 relative; both MDRK stages
(structured/multi_derivative_2d.jl:82-445 with their source lines, stage 2's
overwrite of S restated); the device's convergence order on the
parallelogram; a P4estMesh with turned element frames (same volume kernels);
shock capturing with source terms refused.  Metric terms are the oracle's on
both sides."""
import numpy as np
import pytest

from gpu_helpers import assert_close, torch_mod, warp_mapping

pytestmark = pytest.mark.gpu

GAMMA = 1.4


def setup(oracle_mod, N, nx, kind="parallelogram", td="lw"):
    import trixilw_hip as tlw

    eq = tlw.CompressibleEulerEquations2D(GAMMA)
    ic = tlw.initial_condition_convergence_test
    okw, bcs = dict(periodic=(True, True)), None
    if kind == "parallelogram":
        mesh = tlw.StructuredMesh((nx, nx), lambda xi, eta: (xi + eta, eta))
        mp = (oracle_mod.MAP_PARALLELOGRAM, (0.0, 0.0, 0.0, 0.0))
    elif kind == "warp":
        mesh = tlw.StructuredMesh((nx, nx), warp_mapping(2.0, 0.05))
        mp = (oracle_mod.MAP_WARP, (2.0, 0.05))
    else:  # the box [0, 2]² with the manufactured solution as Dirichlet data
        mesh = tlw.StructuredMesh((nx, nx), (0.0, 0.0), (2.0, 2.0), periodicity=False)
        mp = (oracle_mod.MAP_AFFINE, (0.0, 0.0, 2.0, 2.0))
        okw = dict(periodic=(False, False), bc_kind=oracle_mod.BV_CONVERGENCE_TEST)
        bcs = tlw.BoundaryConditionDirichlet(ic)
    solver = tlw.DGSEM(N, volume_integral=tlw.VolumeIntegralFR(tlw.MDRK() if td == "mdrk" else tlw.LW()))
    o = oracle_mod.Oracle(N, nx, nx, 0.0, 0.0, 1.0, gamma=GAMMA, mapping=mp, source_terms="convergence_test", **okw)
    mesh.set_metrics(solver.basis, o.aux("contravariant_vectors"), o.aux("inverse_jacobian"))
    semi = tlw.SemidiscretizationHyperbolic(mesh, tlw.get_time_discretization(solver), eq, ic, solver,
                                            boundary_conditions=bcs, source_terms=tlw.source_terms_convergence_test)
    return semi, o, mesh


def exact(x, y, t):
    phi = 2.0 + 0.1 * np.sin(np.pi * (x + y - t))
    return np.stack([phi, phi, phi, phi * phi], -1)


def dev(u):
    return torch_mod().from_numpy(np.ascontiguousarray(u)).cuda()


@pytest.mark.parametrize("N,kind", [(3, "parallelogram"), (4, "parallelogram"), (4, "warp"), (3, "box"), (4, "box")])
def test_curved_source_rhs_parity(oracle_mod, N, kind):
    import trixilw_hip as tlw

    torch = torch_mod()
    print(f"curved source rhs N={N} {kind}", flush=True)
    semi, o, mesh = setup(oracle_mod, N, 8, kind)
    x, y = o.node_coords()
    t = 0.3
    u_np = exact(x, y, t) * (1 + 1e-3 * np.random.default_rng(N).standard_normal(x.shape + (4,)))
    dt = 0.4 * o.max_dt(u_np)
    ref = o.rhs(u_np, t, dt)
    semi.cache["dt"][0] = dt
    u = dev(u_np)
    du = torch.empty_like(u)
    tlw.rhs(du, u, semi, t)
    torch.cuda.synchronize()
    assert_close(du.cpu().numpy(), ref, what=f"curved source N={N} {kind}")
    semi.close()


@pytest.mark.parametrize("N,kind", [(3, "parallelogram"), (4, "warp"), (4, "box")])
def test_curved_source_mdrk_stage_parity(oracle_mod, N, kind):
    import trixilw_hip as tlw

    torch = torch_mod()
    print(f"curved source MDRK stages N={N} {kind}", flush=True)
    semi, o, mesh = setup(oracle_mod, N, 8, kind, td="mdrk")
    x, y = o.node_coords()
    t = 0.3
    u_np = exact(x, y, t) * (1 + 1e-3 * np.random.default_rng(20 + N).standard_normal(x.shape + (4,)))
    dt = 0.3 * o.max_dt(u_np)
    d1 = o.rhs(u_np, t, dt, kind="mdrk1")
    us = u_np + dt * d1
    o.set_us(us)
    d2 = o.rhs(u_np, t, dt, kind="mdrk2")
    semi.cache["dt"][0] = dt
    u = dev(u_np)
    du = torch.empty_like(u)
    tlw.rhs_mdrk1(du, u, semi, t)
    torch.cuda.synchronize()
    assert_close(du.cpu().numpy(), d1, what=f"curved source mdrk1 N={N} {kind}")
    semi.set_us(dev(us))
    tlw.rhs_mdrk2(du, u, semi, t)
    torch.cuda.synchronize()
    assert_close(du.cpu().numpy(), d2, what=f"curved source mdrk2 N={N} {kind}")
    semi.close()


@pytest.mark.parametrize("N,kind,K,td", [(4, "parallelogram", 20, "lw"), (3, "box", 20, "lw"),
                                         (3, "parallelogram", 15, "mdrk"), (4, "box", 15, "mdrk")])
def test_curved_source_device_loop_k_steps(oracle_mod, N, kind, K, td):
    import trixilw_hip as tlw

    print(f"curved source device loop N={N} {kind} K={K} {td}", flush=True)
    semi, o, mesh = setup(oracle_mod, N, 8, kind, td=td)
    x, y = o.node_coords()
    u_np = exact(x, y, 0.0)
    u_ref, t_ref, _ = o.solve(u_np, 0.0, 1e9, 0.2 if td == "lw" else 0.1, max_steps=K, mdrk=(td == "mdrk"))
    lw_update = tlw.semidiscretize(semi, semi.time_discretization, (0.0, float("inf")))
    lw_update.u0.copy_(dev(u_np))
    sol = tlw.solve_lwfr(lw_update, (), 1e-6, None, time_step_computation=tlw.CFLBased(0.2 if td == "lw" else 0.1),
                         max_steps=K)
    assert sol.stats.naccept == K
    assert abs(sol.t - t_ref) <= 1e-13 * t_ref
    assert_close(sol.u.cpu().numpy(), u_ref, what=f"curved source N={N} {kind} after {K} steps")
    semi.close()


def test_parallelogram_device_eoc(oracle_mod):
    """The reference's parallelogram example (N=4, CFL 0.2) to t = 0.5 on 4², 8²,
    16² cells: each final state matches the oracle's, and the density's L2 error
    falls at order above 4.4 and 4.0."""
    import trixilw_hip as tlw

    errs = []
    for nx in (4, 8, 16):
        print(f"parallelogram EOC {nx}x{nx}", flush=True)
        semi, o, mesh = setup(oracle_mod, 4, nx)
        x, y = o.node_coords()
        u0 = exact(x, y, 0.0)
        lw_update = tlw.semidiscretize(semi, semi.time_discretization, (0.0, 0.5))
        lw_update.u0.copy_(dev(u0))
        sol = tlw.solve_lwfr(lw_update, (), 1e-6, None, time_step_computation=tlw.CFLBased(0.2))
        assert abs(sol.t - 0.5) < 1e-14
        got = sol.u.cpu().numpy()
        u_ref, _, _ = o.solve(u0, 0.0, 0.5, 0.2)
        assert_close(got, u_ref, what=f"parallelogram to t=0.5 {nx}x{nx}")
        W = np.outer(*[o.basis()[1]] * 2)[None] / o.aux("inverse_jacobian").reshape(x.shape)
        errs.append(np.sqrt(np.einsum("eji,eji->", W, (got[..., 0] - exact(x, y, 0.5)[..., 0]) ** 2)))
        semi.close()
    eoc = [np.log2(errs[k] / errs[k + 1]) for k in range(2)]
    print(f"parallelogram EOC: errors {errs}, orders {eoc}", flush=True)
    assert eoc[0] > 4.4 and eoc[1] > 4.0, (errs, eoc)


@pytest.mark.parametrize("N,td", [(3, "lw"), (4, "lw"), (3, "mdrk"), (4, "mdrk")])
def test_p4est_source_parity(oracle_mod, N, td):
    """The convergence test on the warped P4estMesh (frames turned by random
    multiples of 90°): the residual (both stages for MDRK) and 15 device-loop
    steps against the oracle."""
    import trixilw_hip as tlw

    torch = torch_mod()
    print(f"P4est source N={N} {td}", flush=True)
    nx = 8
    turns = np.random.default_rng(30 + N).integers(0, 4, nx * nx)
    eq = tlw.CompressibleEulerEquations2D(GAMMA)
    mesh = tlw.P4estMesh((nx, nx), warp_mapping(2.0, 0.05), turns=turns)
    solver = tlw.DGSEM(N, volume_integral=tlw.VolumeIntegralFR(tlw.MDRK() if td == "mdrk" else tlw.LW()))
    o = oracle_mod.Oracle(N, nx, nx, 0.0, 0.0, 1.0, gamma=GAMMA, mapping=(oracle_mod.MAP_WARP, (2.0, 0.05)),
                          periodic=(True, True), p4est=turns, source_terms="convergence_test")
    mesh.set_metrics(solver.basis, o.aux("contravariant_vectors"), o.aux("inverse_jacobian"))
    semi = tlw.SemidiscretizationHyperbolic(mesh, tlw.get_time_discretization(solver), eq,
                                            tlw.initial_condition_convergence_test, solver,
                                            source_terms=tlw.source_terms_convergence_test)
    x, y = o.node_coords()
    t = 0.3
    u_np = exact(x, y, t) * (1 + 1e-3 * np.random.default_rng(N).standard_normal(x.shape + (4,)))
    dt = 0.3 * o.max_dt(u_np)
    semi.cache["dt"][0] = dt
    u = dev(u_np)
    du = torch.empty_like(u)
    if td == "lw":
        ref = o.rhs(u_np, t, dt)
        tlw.rhs(du, u, semi, t)
        torch.cuda.synchronize()
        assert_close(du.cpu().numpy(), ref, what=f"P4est source LW N={N}")
    else:
        d1 = o.rhs(u_np, t, dt, kind="mdrk1")
        us = u_np + dt * d1
        o.set_us(us)
        d2 = o.rhs(u_np, t, dt, kind="mdrk2")
        tlw.rhs_mdrk1(du, u, semi, t)
        torch.cuda.synchronize()
        assert_close(du.cpu().numpy(), d1, what=f"P4est source mdrk1 N={N}")
        semi.set_us(dev(us))
        tlw.rhs_mdrk2(du, u, semi, t)
        torch.cuda.synchronize()
        assert_close(du.cpu().numpy(), d2, what=f"P4est source mdrk2 N={N}")
    cfl = 0.2 if td == "lw" else 0.1
    u0 = exact(x, y, 0.0)
    u_ref, t_ref, _ = o.solve(u0, 0.0, 1e9, cfl, max_steps=15, mdrk=(td == "mdrk"))
    lw_update = tlw.semidiscretize(semi, semi.time_discretization, (0.0, float("inf")))
    lw_update.u0.copy_(dev(u0))
    sol = tlw.solve_lwfr(lw_update, (), 1e-6, None, time_step_computation=tlw.CFLBased(cfl), max_steps=15)
    assert sol.stats.naccept == 15 and abs(sol.t - t_ref) <= 1e-13 * t_ref
    assert_close(sol.u.cpu().numpy(), u_ref, what=f"P4est source {td} N={N} after 15 steps")
    semi.close()


def test_curved_source_refusals(oracle_mod):
    import trixilw_hip as tlw
    from trixilw_hip import _lib

    torch_mod()
    eq = tlw.CompressibleEulerEquations2D(GAMMA)
    mesh = tlw.StructuredMesh((4, 4), warp_mapping(2.0, 0.05))
    ind = tlw.IndicatorHennemannGassner(eq, None, alpha_max=0.5)
    vi = tlw.VolumeIntegralFRShockCapturing(ind, volume_integralFR=tlw.VolumeIntegralFR(tlw.LW()))
    with pytest.raises(_lib.TLWError) as ei:  # shock capturing with source terms
        tlw.SemidiscretizationHyperbolic(mesh, tlw.LW(), eq, tlw.initial_condition_convergence_test,
                                         tlw.DGSEM(3, volume_integral=vi),
                                         source_terms=tlw.source_terms_convergence_test)
    assert ei.value.code == _lib.TLW_ENOTSUP
