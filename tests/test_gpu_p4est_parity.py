"""GPU parity of the P4estMesh path (p4est/dg_2d.jl:8-393) against the oracle,
on the warped mapping's cells with element frames turned by random multiples
of 90° (the orientations of an unstructured quadrilateral mesh): rhs and both
MDRK stages periodic and with the P4est forms of the time-averaged Dirichlet
value, outflow and the approximate slip wall; K-step device-loop states within
1e-12 relative; free stream and conservation at 128²; the aligned P4est mesh
against the StructuredMesh path; the configurations refused.  Metric terms are
the oracle's on both sides (P4estMesh.set_metrics); inputs are seeded."""
import numpy as np
import pytest

from gpu_helpers import GAMMA, assert_close, torch_mod, warp_mapping

pytestmark = pytest.mark.gpu

L, A = 20.0, 0.1


def setup_p4est(oracle_mod, N, nx, td="lw", bc=None, seed=0, turned=True, amp=1e-3):
    import trixilw_hip as tlw

    eq = tlw.CompressibleEulerEquations2D(GAMMA)
    periodic = bc is None
    ic = tlw.IsentropicVortex(center=(L / 2, L / 2), period=(L, L) if periodic else (1e9, 1e9))
    turns = np.random.default_rng(nx + seed).integers(0, 4, nx * nx) if turned else np.zeros(nx * nx, dtype=int)
    mesh = tlw.P4estMesh((nx, nx), warp_mapping(L, A), periodicity=periodic, turns=turns)
    solver = tlw.DGSEM(N, volume_integral=tlw.VolumeIntegralFR(tlw.MDRK() if td == "mdrk" else tlw.LW()))
    if bc == "vortex":
        cond, okw = ic, dict(bc_kind=ic.kind, bc_par=ic.params)
    elif bc == "outflow":
        cond, okw = tlw.boundary_condition_outflow, dict(bc_kind=oracle_mod.BV_OUTFLOW)
    elif bc == "slip":
        cond, okw = tlw.slip_wall_approximate, dict(bc_kind=oracle_mod.BV_SLIP_WALL_APPROX)
    else:
        cond, okw = None, {}
    o = oracle_mod.Oracle(N, nx, nx, 0.0, 0.0, L / nx, gamma=GAMMA, mapping=(oracle_mod.MAP_WARP, (L, A)),
                          periodic=(periodic, periodic), p4est=turns, **okw)
    itf, _ = o.p4est_connectivity()
    assert sorted(map(tuple, itf)) == sorted(map(tuple, mesh.interfaces))
    mesh.set_metrics(solver.basis, o.aux("contravariant_vectors"), o.aux("inverse_jacobian"))
    semi = tlw.SemidiscretizationHyperbolic(mesh, tlw.get_time_discretization(solver), eq, ic, solver,
                                            boundary_conditions=cond)
    x, y = o.node_coords()
    rng = np.random.default_rng(seed)
    u = ic(x, y, 0.0) * (1.0 + amp * rng.standard_normal(x.shape + (4,)))
    return semi, o, u, mesh


def dev(u):
    return torch_mod().from_numpy(np.ascontiguousarray(u)).cuda()


@pytest.mark.parametrize("N", [3, 4])
@pytest.mark.parametrize("bc", [None, "vortex", "outflow", "slip"])
def test_p4est_rhs_lw_parity(oracle_mod, N, bc):
    import trixilw_hip as tlw

    torch = torch_mod()
    print(f"p4est rhs LW N={N} bc={bc}", flush=True)
    semi, o, u_np, mesh = setup_p4est(oracle_mod, N, 8, bc=bc, seed=N)
    u = dev(u_np)
    du = torch.empty_like(u)
    dt = 0.4 * o.max_dt(u_np)
    semi.cache["dt"][0] = dt
    du_ref = o.rhs(u_np, 0.25, dt)
    tlw.rhs(du, u, semi, 0.25)
    torch.cuda.synchronize()
    assert_close(du.cpu().numpy(), du_ref, what=f"p4est LW N={N} {bc}")
    semi.close()


@pytest.mark.parametrize("N", [3, 4])
@pytest.mark.parametrize("bc", [None, "vortex", "slip"])
def test_p4est_rhs_mdrk_parity(oracle_mod, N, bc):
    import trixilw_hip as tlw

    torch = torch_mod()
    print(f"p4est rhs MDRK N={N} bc={bc}", flush=True)
    semi, o, u_np, mesh = setup_p4est(oracle_mod, N, 8, td="mdrk", bc=bc, seed=5 + N)
    u = dev(u_np)
    du = torch.empty_like(u)
    dt = 0.4 * o.max_dt(u_np)
    semi.cache["dt"][0] = dt
    d1 = o.rhs(u_np, 0.25, dt, kind="mdrk1")
    us = u_np + 0.5 * dt * d1
    o.set_us(us)
    d2 = o.rhs(u_np, 0.25, dt, kind="mdrk2")
    tlw.rhs_mdrk1(du, u, semi, 0.25)
    assert_close(du.cpu().numpy(), d1, what=f"p4est mdrk1 N={N} {bc}")
    semi.set_us(dev(us))
    tlw.rhs_mdrk2(du, u, semi, 0.25)
    torch.cuda.synchronize()
    assert_close(du.cpu().numpy(), d2, what=f"p4est mdrk2 N={N} {bc}")
    semi.close()


@pytest.mark.parametrize("N,nx,td,K,bc", [(4, 8, "lw", 25, None), (3, 16, "mdrk", 20, "vortex"),
                                          (4, 12, "lw", 20, "slip")])
def test_p4est_device_loop_k_steps(oracle_mod, N, nx, td, K, bc):
    import trixilw_hip as tlw

    print(f"p4est device loop N={N} {nx}x{nx} {td} K={K} bc={bc}", flush=True)
    semi, o, u_np, mesh = setup_p4est(oracle_mod, N, nx, td=td, bc=bc)
    cfl = tlw.trixi2lw(0.9, semi.solver) * (0.5 if td == "mdrk" else 1.0)
    u_ref, t_ref, _ = o.solve(u_np, 0.0, 1e9, cfl, max_steps=K, mdrk=(td == "mdrk"))
    lw_update = tlw.semidiscretize(semi, semi.time_discretization, (0.0, float("inf")))
    lw_update.u0.copy_(dev(u_np))
    sol = tlw.solve_lwfr(lw_update, (), 1e-6, None, time_step_computation=tlw.CFLBased(cfl), max_steps=K)
    assert sol.stats.naccept == K
    assert abs(sol.t - t_ref) <= 1e-13 * t_ref
    assert_close(sol.u.cpu().numpy(), u_ref, what=f"p4est {td} N={N} after {K} steps")
    semi.close()


def test_p4est_free_stream_and_conservation_at_scale():
    """128² turned P4est mesh with the host's own metric terms: a uniform flow
    stays uniform (≤1e-11 after 20 steps) and a perturbed vortex keeps
    Σ w_i w_j u / J⁻¹ (≤1e-13 relative) over 20 steps."""
    import trixilw_hip as tlw

    torch = torch_mod()
    print("p4est 128^2 turned, 20 steps", flush=True)
    nx = 128
    turns = np.random.default_rng(3).integers(0, 4, nx * nx)
    mesh = tlw.P4estMesh((nx, nx), warp_mapping(L, A), turns=turns)
    eq = tlw.CompressibleEulerEquations2D(GAMMA)
    ic = tlw.IsentropicVortex(center=(L / 2, L / 2), period=(L, L))
    solver = tlw.DGSEM(4, volume_integral=tlw.VolumeIntegralFR(tlw.LW()))
    semi = tlw.SemidiscretizationHyperbolic(mesh, tlw.get_time_discretization(solver), eq, ic, solver)
    x, y = mesh.node_coordinates(solver.basis.nodes)
    _, ijn = mesh.metrics(solver.basis)
    w = solver.basis.weights
    cfl = tlw.trixi2lw(0.9, semi.solver)
    for kind in ("free_stream", "vortex"):
        if kind == "free_stream":
            u0 = tlw.ConstantState(1.2, 0.3, -0.5, 0.9)(x, y, 0.0)
        else:
            rng = np.random.default_rng(8)
            u0 = ic(x, y, 0.0) * (1 + 1e-2 * rng.standard_normal(x.shape + (1,)))
        lw_update = tlw.semidiscretize(semi, semi.time_discretization, (0.0, float("inf")))
        lw_update.u0.copy_(dev(u0))
        sol = tlw.solve_lwfr(lw_update, (), 1e-6, None, time_step_computation=tlw.CFLBased(cfl), max_steps=20)
        torch.cuda.synchronize()
        u1 = sol.u.cpu().numpy()
        assert sol.stats.naccept == 20 and np.all(np.isfinite(u1))
        if kind == "free_stream":
            assert np.max(np.abs(u1 - u0)) < 1e-11
        else:
            W = np.outer(w, w)[None] / ijn
            m0, m1 = np.einsum("eji,ejiv->v", W, u0), np.einsum("eji,ejiv->v", W, u1)
            assert np.all(np.abs(m1 - m0) <= 1e-13 * np.einsum("eji,ejiv->v", W, np.abs(u0))), (m1 - m0)
    semi.close()


@pytest.mark.parametrize("N", [3, 4])
def test_aligned_p4est_matches_structured_path(oracle_mod, N):
    """The aligned P4est mesh and the StructuredMesh of the same mapping differ only
    in whose contravariant vector a shared face node takes: their device residuals
    agree to round-off."""
    import trixilw_hip as tlw

    from gpu_helpers import setup_curved

    torch = torch_mod()
    semi_p, o, u_np, mesh = setup_p4est(oracle_mod, N, 8, turned=False, seed=2)
    semi_s, o_s, _, mesh_s = setup_curved(oracle_mod, N, 8, 8, A=A)
    u = dev(u_np)
    dt = 0.4 * o.max_dt(u_np)
    for semi in (semi_p, semi_s):
        semi.cache["dt"][0] = dt
    d_p, d_s = torch.empty_like(u), torch.empty_like(u)
    tlw.rhs(d_p, u, semi_p, 0.0)
    tlw.rhs(d_s, u, semi_s, 0.0)
    torch.cuda.synchronize()
    assert_close(d_p.cpu().numpy(), d_s.cpu().numpy(), what=f"aligned P4est vs Structured N={N}")
    semi_p.close()
    semi_s.close()


def test_p4est_refuses_unsupported(oracle_mod):
    import trixilw_hip as tlw
    from trixilw_hip import _lib

    torch_mod()
    eq = tlw.CompressibleEulerEquations2D(GAMMA)
    ic = tlw.IsentropicVortex(center=(L / 2, L / 2), period=(L, L))
    mesh = tlw.P4estMesh((4, 4), warp_mapping(L, A), turns=np.arange(16) % 4)
    ind = tlw.IndicatorHennemannGassner(eq, None, alpha_max=0.5)
    vi = tlw.VolumeIntegralFRShockCapturing(ind, volume_integralFR=tlw.VolumeIntegralFR(tlw.LW()))
    with pytest.raises(_lib.TLWError) as ei:
        tlw.SemidiscretizationHyperbolic(mesh, tlw.LW(), eq, ic, tlw.DGSEM(3, volume_integral=vi))
    assert ei.value.code == _lib.TLW_ENOTSUP
    open_mesh = tlw.P4estMesh((4, 4), warp_mapping(L, A), periodicity=False)
    with pytest.raises(_lib.TLWError) as ei:
        tlw.SemidiscretizationHyperbolic(open_mesh, tlw.LW(), eq, ic, tlw.DGSEM(3),
                                         boundary_conditions={"x_neg": tlw.boundary_condition_slip_wall_vertical,
                                                              "x_pos": tlw.boundary_condition_slip_wall_vertical,
                                                              "y_neg": tlw.boundary_condition_slip_wall_horizontal,
                                                              "y_pos": tlw.boundary_condition_slip_wall_horizontal})
    assert ei.value.code == _lib.TLW_ENOTSUP
