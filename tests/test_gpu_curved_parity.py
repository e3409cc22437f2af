"""GPU parity of the curvilinear StructuredMesh LW and MDRK paths (periodic,
compressible Euler) against the CPU oracle: rhs and both MDRK stages on warped
and affine meshes, the curvilinear max_dt and K-step device-loop states (LW and
MDRK) within 1e-12 relative (L2 and L∞), every
volume tile, the affine mesh against the TreeMesh kernels with the structured
flux order, free-stream preservation and conservation at 256×256, and the
combinations the path rejects.  Inputs are seeded.

Parity runs the kernels on identical inputs: state, dt and the oracle's metric
terms (StructuredMesh.set_metrics).  The host computation of the metric terms
is pinned against the oracle's separately (test_host_structured_mesh.py); the
two differ by ~1e-14 relative, since D is applied to coordinates of size L to
give derivatives of size h/2, and the flux cancellation in the rhs amplifies
that to ~1e-12.  Device-loop runs use LW-stable CFL numbers (Trixi units, the
reference's trixi2lw of 0.9)."""
import numpy as np
import pytest

from gpu_helpers import GAMMA, assert_close, setup_curved, to_dev, to_lex, torch_mod, warp_mapping

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("N", [3, 4])
@pytest.mark.parametrize("kind,A", [("warp", 0.05), ("warp", 0.12), ("affine", 0.0)])
def test_curved_rhs_parity(oracle_mod, N, kind, A):
    import trixilw_hip as tlw

    torch = torch_mod()
    print(f"curved rhs N={N} {kind} A={A}", flush=True)
    semi, o, u_lex, mesh = setup_curved(oracle_mod, N, 8, 6, kind=kind, A=A, keep_aux=True)
    dt = 0.8 * o.max_dt(u_lex)
    du_ref = o.rhs(u_lex, 0.0, dt)
    u = to_dev(mesh, u_lex)
    du = torch.empty_like(u)
    semi.cache["dt"][0] = dt
    tlw.rhs(du, u, semi, 0.0)
    torch.cuda.synchronize()
    assert_close(to_lex(mesh, du), du_ref, what=f"curved LW N={N} {kind} A={A}")
    # Ũ and the Cartesian F̃ of the volume procedure
    U = semi.get_aux("U")[mesh.lexicographic_permutation()]
    F = semi.get_aux("F")[mesh.lexicographic_permutation()]
    assert_close(U, o.aux("U").reshape(U.shape), what="Ũ")
    assert_close(F, o.aux("F").reshape(F.shape), what="F̃")
    semi.close()


@pytest.mark.parametrize("N", [3, 4])
def test_curved_rhs_mdrk_parity(oracle_mod, N):
    """Both MDRK stages (mdrk_kernel_1!/2! for StructuredMesh) with Ũ and F̃."""
    import trixilw_hip as tlw

    torch = torch_mod()
    print(f"curved MDRK rhs N={N}", flush=True)
    semi, o, u_lex, mesh = setup_curved(oracle_mod, N, 8, 6, A=0.1, td="mdrk", keep_aux=True, seed=4)
    dt = 0.4 * o.max_dt(u_lex)
    d1_ref = o.rhs(u_lex, 0.2, dt, kind="mdrk1")
    U1, F1 = o.aux("U"), o.aux("F")
    us = u_lex + dt * d1_ref
    o.set_us(us)
    d2_ref = o.rhs(u_lex, 0.2, dt, kind="mdrk2")
    u = to_dev(mesh, u_lex)
    du = torch.empty_like(u)
    semi.cache["dt"][0] = dt
    tlw.rhs_mdrk1(du, u, semi, 0.2)
    torch.cuda.synchronize()
    assert_close(to_lex(mesh, du), d1_ref, what=f"curved mdrk1 N={N}")
    U = semi.get_aux("U")[mesh.lexicographic_permutation()]
    F = semi.get_aux("F")[mesh.lexicographic_permutation()]
    assert_close(U, U1.reshape(U.shape), what="stage-1 Ũ")
    assert_close(F, F1.reshape(F.shape), what="stage-1 F̃")
    semi.set_us(to_dev(mesh, us))
    tlw.rhs_mdrk2(du, u, semi, 0.2)
    torch.cuda.synchronize()
    assert_close(to_lex(mesh, du), d2_ref, what=f"curved mdrk2 N={N}")
    U = semi.get_aux("U")[mesh.lexicographic_permutation()]
    F = semi.get_aux("F")[mesh.lexicographic_permutation()]
    assert_close(U, o.aux("U").reshape(U.shape), what="stage-2 Ũ")
    assert_close(F, o.aux("F").reshape(F.shape), what="stage-2 F̃")
    semi.close()


@pytest.mark.parametrize("N,nx,K", [(3, 16, 20), (4, 8, 15)])
def test_curved_mdrk_device_loop_k_steps(oracle_mod, N, nx, K):
    import trixilw_hip as tlw

    print(f"curved MDRK device loop N={N} {nx}x{nx} K={K}", flush=True)
    semi, o, u_lex, mesh = setup_curved(oracle_mod, N, nx, nx, A=0.1, td="mdrk")
    cfl = 0.5 * tlw.trixi2lw(0.9, semi.solver)
    u_ref, t_ref, _ = o.solve(u_lex, 0.0, 1e9, cfl, max_steps=K, mdrk=True)
    lw_update = tlw.semidiscretize(semi, semi.time_discretization, (0.0, float("inf")))
    lw_update.u0.copy_(to_dev(mesh, u_lex))
    sol = tlw.solve_lwfr(lw_update, (), 1e-6, None, time_step_computation=tlw.CFLBased(cfl), max_steps=K)
    assert sol.stats.naccept == K
    assert abs(sol.t - t_ref) <= 1e-13 * t_ref
    assert_close(to_lex(mesh, sol.u), u_ref, what=f"curved MDRK N={N} {nx}x{nx} after {K} steps")
    semi.close()


@pytest.mark.parametrize("N", [3, 4])
@pytest.mark.parametrize("td", ["lw", "mdrk"])
def test_curved_dirichlet_rhs_parity(oracle_mod, N, td):
    """Non-periodic warped mesh with the vortex as time-averaged Dirichlet data
    (calc_boundary_flux! for StructuredMesh): rhs, or both MDRK stages."""
    import trixilw_hip as tlw

    torch = torch_mod()
    print(f"curved Dirichlet N={N} {td}", flush=True)
    semi, o, u_lex, mesh = setup_curved(oracle_mod, N, 8, 6, A=0.1, td=td, seed=7, periodic=False)
    u = to_dev(mesh, u_lex)
    du = torch.empty_like(u)
    dt = 0.6 * o.max_dt(u_lex)
    semi.cache["dt"][0] = dt
    if td == "lw":
        du_ref = o.rhs(u_lex, 0.25, dt)
        tlw.rhs(du, u, semi, 0.25)
        torch.cuda.synchronize()
        assert_close(to_lex(mesh, du), du_ref, what=f"curved Dirichlet LW N={N}")
    else:
        d1 = o.rhs(u_lex, 0.25, dt, kind="mdrk1")
        us = u_lex + 0.5 * dt * d1
        o.set_us(us)
        d2 = o.rhs(u_lex, 0.25, dt, kind="mdrk2")
        tlw.rhs_mdrk1(du, u, semi, 0.25)
        assert_close(to_lex(mesh, du), d1, what=f"curved Dirichlet mdrk1 N={N}")
        semi.set_us(to_dev(mesh, us))
        tlw.rhs_mdrk2(du, u, semi, 0.25)
        torch.cuda.synchronize()
        assert_close(to_lex(mesh, du), d2, what=f"curved Dirichlet mdrk2 N={N}")
    semi.close()


@pytest.mark.parametrize("N,nx,td,K", [(4, 8, "lw", 25), (3, 16, "lw", 25), (3, 16, "mdrk", 20)])
def test_curved_dirichlet_device_loop_k_steps(oracle_mod, N, nx, td, K):
    import trixilw_hip as tlw

    print(f"curved Dirichlet device loop N={N} {td} K={K}", flush=True)
    semi, o, u_lex, mesh = setup_curved(oracle_mod, N, nx, nx, A=0.1, td=td, periodic=False)
    cfl = tlw.trixi2lw(0.9, semi.solver) * (0.5 if td == "mdrk" else 1.0)
    u_ref, t_ref, _ = o.solve(u_lex, 0.0, 1e9, cfl, max_steps=K, mdrk=(td == "mdrk"))
    lw_update = tlw.semidiscretize(semi, semi.time_discretization, (0.0, float("inf")))
    lw_update.u0.copy_(to_dev(mesh, u_lex))
    sol = tlw.solve_lwfr(lw_update, (), 1e-6, None, time_step_computation=tlw.CFLBased(cfl), max_steps=K)
    assert sol.stats.naccept == K
    assert abs(sol.t - t_ref) <= 1e-13 * t_ref
    assert_close(to_lex(mesh, sol.u), u_ref, what=f"curved Dirichlet {td} N={N} after {K} steps")
    semi.close()


@pytest.mark.parametrize("N", [3, 4])
def test_curved_max_dt_parity(oracle_mod, N):
    semi, o, u_lex, mesh = setup_curved(oracle_mod, N, 8, 6, A=0.12, amp=1e-2)
    ref = o.max_dt(u_lex)
    got = semi.max_dt(to_dev(mesh, u_lex))
    assert abs(got - ref) <= 1e-14 * ref
    semi.close()


@pytest.mark.parametrize("N", [3, 4])
def test_affine_curved_equals_treemesh_structured_order(oracle_mod, N):
    """The curvilinear kernels on an affine box reproduce the TreeMesh kernels
    run with the structured flux order (metric round-off only)."""
    import trixilw_hip as tlw

    torch = torch_mod()
    L, nx = 20.0, 8
    semi_c, _, u_lex, mesh_c = setup_curved(oracle_mod, N, nx, nx, kind="affine", L=L, seed=5, amp=1e-2)
    eq = tlw.CompressibleEulerEquations2D(GAMMA)
    mesh_t = tlw.TreeMesh((0.0, 0.0), (L, L), 3, ordering="lexicographic")
    solver = tlw.DGSEM(N, volume_integral=tlw.VolumeIntegralFR(tlw.LW()))
    semi_t = tlw.SemidiscretizationHyperbolic(mesh_t, tlw.LW(), eq, None, solver, surface_flux_order="structured")
    u_c, u_t = to_dev(mesh_c, u_lex), to_dev(mesh_t, u_lex)
    dt = 0.8 * semi_t.max_dt(u_t)
    assert abs(semi_c.max_dt(u_c) - semi_t.max_dt(u_t)) <= 1e-14 * semi_t.max_dt(u_t)
    du_c, du_t = torch.empty_like(u_c), torch.empty_like(u_t)
    semi_c.cache["dt"][0] = dt
    semi_t.cache["dt"][0] = dt
    tlw.rhs(du_c, u_c, semi_c, 0.0)
    tlw.rhs(du_t, u_t, semi_t, 0.0)
    torch.cuda.synchronize()
    assert_close(to_lex(mesh_c, du_c), to_lex(mesh_t, du_t), what=f"affine curved vs TreeMesh N={N}")
    semi_c.close()
    semi_t.close()


def test_curved_every_tile_matches_oracle(oracle_mod):
    import trixilw_hip as tlw

    torch = torch_mod()
    for N in (3, 4):
        semi, o, u_lex, mesh = setup_curved(oracle_mod, N, 10, 7, A=0.1, seed=3)
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
            assert_close(to_lex(mesh, du), du_ref, what=f"curved N={N} tile {eb}")
        semi.close()


@pytest.mark.parametrize("N,nx,K", [(4, 8, 25), (3, 16, 25)])
def test_curved_device_loop_k_steps(oracle_mod, N, nx, K):
    """K steps of the device loop (curvilinear CFL reduction in the surface
    kernel) against the oracle's solve loop on a warped mesh (polydeg 3 needs
    16×16 to keep the vortex admissible without a limiter)."""
    import trixilw_hip as tlw

    print(f"curved device loop N={N} {nx}x{nx} K={K}", flush=True)
    semi, o, u_lex, mesh = setup_curved(oracle_mod, N, nx, nx, A=0.1)
    cfl = tlw.trixi2lw(0.9, semi.solver)
    u_ref, t_ref, dts = o.solve(u_lex, 0.0, 1e9, cfl, max_steps=K)
    lw_update = tlw.semidiscretize(semi, semi.time_discretization, (0.0, float("inf")))
    lw_update.u0.copy_(to_dev(mesh, u_lex))
    sol = tlw.solve_lwfr(lw_update, (), 1e-6, None, time_step_computation=tlw.CFLBased(cfl), max_steps=K)
    assert sol.stats.naccept == K
    assert abs(sol.t - t_ref) <= 1e-13 * t_ref
    assert_close(to_lex(mesh, sol.u), u_ref, what=f"curved N={N} {nx}x{nx} after {K} steps")
    semi.close()


def test_curved_host_loop_matches_oracle(oracle_mod):
    import trixilw_hip as tlw

    semi, o, u_lex, mesh = setup_curved(oracle_mod, 4, 8, 8, A=0.1)
    cfl = tlw.trixi2lw(0.9, semi.solver)
    u_ref, t_ref, _ = o.solve(u_lex, 0.0, 1e9, cfl, max_steps=10)
    lw_update = tlw.semidiscretize(semi, semi.time_discretization, (0.0, float("inf")))
    lw_update.u0.copy_(to_dev(mesh, u_lex))
    sol = tlw.solve_lwfr(lw_update, (), 1e-6, None, time_step_computation=tlw.CFLBased(cfl), device_loop=False,
                         max_steps=10)
    assert abs(sol.t - t_ref) <= 1e-13 * t_ref
    assert_close(to_lex(mesh, sol.u), u_ref, what="curved host loop after 10 steps")
    semi.close()


def test_curved_free_stream_and_conservation_at_scale():
    """256×256 warped mesh (1.6 M DOFs, the host's own metric terms): a
    constant state stays constant to the rounding of the discrete metric
    identities, and 20 device-loop steps of the vortex conserve Σ w_i w_j J u.
    The metric identities hold to ~eps |D|² L per element, scaled by J⁻¹ ~
    (2/h)²: the oracle's free-stream rhs at this size is 7.5e-10, 1.8e-12 per
    step; 10 steps stay below 1e-10."""
    import trixilw_hip as tlw

    torch = torch_mod()
    L = 20.0
    eq = tlw.CompressibleEulerEquations2D(GAMMA)
    mesh = tlw.StructuredMesh((256, 256), warp_mapping(L, 0.08))
    solver = tlw.DGSEM(4, volume_integral=tlw.VolumeIntegralFR(tlw.LW()))
    const = np.array([1.1, 0.4, -0.3, 2.7])
    cfl = tlw.trixi2lw(0.9, solver)

    def free_stream(x, y, t):
        return np.broadcast_to(const, x.shape + (4,))

    semi = tlw.SemidiscretizationHyperbolic(mesh, tlw.LW(), eq, free_stream, solver)
    lw_update = tlw.semidiscretize(semi, tlw.LW(), (0.0, 1e9))
    sol = tlw.solve_lwfr(lw_update, (), 1e-6, None, time_step_computation=tlw.CFLBased(cfl), max_steps=10)
    c = torch.tensor(const, dtype=torch.float64, device=sol.u.device)
    assert (sol.u - c).abs().max().item() <= 1e-10
    semi.close()
    ic = tlw.IsentropicVortex(center=(L / 2, L / 2), period=(L, L))
    semi = tlw.SemidiscretizationHyperbolic(mesh, tlw.LW(), eq, ic, solver)
    lw_update = tlw.semidiscretize(semi, tlw.LW(), (0.0, 1e9))
    u0 = lw_update.u0.clone()
    sol = tlw.solve_lwfr(lw_update, (), 1e-6, None, time_step_computation=tlw.CFLBased(cfl), max_steps=20)
    _, ijn = mesh.metrics(solver.basis)
    w = np.outer(solver.basis.weights, solver.basis.weights)[None] / ijn
    wt = torch.from_numpy(np.ascontiguousarray(w)).cuda()
    m0 = torch.einsum("eji,ejiv->v", wt, u0)
    m1 = torch.einsum("eji,ejiv->v", wt, sol.u)
    scale = torch.einsum("eji,ejiv->v", wt, u0.abs())
    assert torch.all((m1 - m0).abs() <= 1e-13 * scale), (m0, m1)
    assert torch.isfinite(sol.u).all()
    semi.close()


def test_curved_rejects_unsupported(oracle_mod):
    import trixilw_hip as tlw
    from trixilw_hip import _lib

    torch = torch_mod()
    semi, o, u_lex, mesh = setup_curved(oracle_mod, 3, 4, 4)
    with pytest.raises(_lib.TLWError) as ei:
        semi.rebuild_mesh(mesh)
    assert ei.value.code == _lib.TLW_ENOTSUP
    u = to_dev(mesh, u_lex)
    with pytest.raises(_lib.TLWError) as ei:
        tlw.adjoint_step(torch.empty_like(u), u, u.clone(), semi, 0.0, 1e-3)
    assert ei.value.code == _lib.TLW_ENOTSUP
    semi.close()
    # adaptive steps take the Euler LW estimate: an MDRK context has none
    semi, o, u_lex, mesh = setup_curved(oracle_mod, 3, 4, 4, td="mdrk")
    lw_update = tlw.semidiscretize(semi, semi.time_discretization, (0.0, 1.0))
    with pytest.raises(_lib.TLWError) as ei:
        tlw.solve_lwfr(lw_update, (), 1e-6, {"abstol": 1e-8, "reltol": 1e-8}, time_step_computation=tlw.Adaptive(),
                       max_steps=2)
    assert ei.value.code == _lib.TLW_ENOTSUP
    semi.close()
    eq = tlw.CompressibleEulerEquations2D(GAMMA)
    ic = tlw.IsentropicVortex()
    # a non-periodic mesh without a boundary value function, and a mirrored mapping (negative Jacobian)
    cases = [
        (tlw.StructuredMesh((4, 4), warp_mapping(20.0, 0.05), periodicity=False), tlw.LW(), _lib.TLW_EINVAL),
        (tlw.StructuredMesh((4, 4), lambda xi, eta: (-xi, eta)), tlw.LW(), _lib.TLW_EINVAL),
    ]
    for mesh_k, td, code in cases:
        solver = tlw.DGSEM(3, volume_integral=tlw.VolumeIntegralFR(td))
        with pytest.raises(_lib.TLWError) as ei:
            tlw.SemidiscretizationHyperbolic(mesh_k, td, eq, ic, solver)
        assert ei.value.code == code


@pytest.mark.parametrize("periodic,sc", [(True, False), (False, False), (True, True)])
def test_curved_strips_ordering_parity(oracle_mod, periodic, sc):
    """The strips element order (two-row strips, the bench's order) on periodic,
    Dirichlet and shock-capturing curvilinear meshes: the oracle's residuals at
    every tile size, and 10 device-loop steps."""
    import trixilw_hip as tlw

    torch = torch_mod()
    semi, o, u_lex, mesh = setup_curved(oracle_mod, 4, 8, 6, periodic=periodic, sc=sc, ordering="strips",
                                        keep_aux=sc)
    if sc:
        from test_gpu_curved_sc_parity import with_jump

        u_lex = with_jump(o, u_lex)
    dt = 0.5 * o.max_dt(u_lex)
    du_ref = o.rhs(u_lex, 0.0, dt)
    u = to_dev(mesh, u_lex)
    semi.cache["dt"][0] = dt
    for eb in (5, 10, 15):
        print(f"curved strips periodic={periodic} sc={sc} tile {eb}", flush=True)
        semi.elements_per_workgroup(eb)
        du = torch.empty_like(u)
        tlw.rhs(du, u, semi, 0.0)
        torch.cuda.synchronize()
        assert_close(to_lex(mesh, du), du_ref, what=f"curved strips periodic={periodic} sc={sc} tile {eb}")
        if sc:
            alpha = semi.get_aux("alpha")[mesh.lexicographic_permutation()]
            assert np.count_nonzero(o.aux("alpha")) > 2
            assert np.max(np.abs(alpha - o.aux("alpha"))) < 1e-13
    K = 10
    u_ref, t_ref, _ = o.solve(u_lex, 0.0, 1e9, 0.5, max_steps=K)
    lw_update = tlw.semidiscretize(semi, semi.time_discretization, (0.0, float("inf")))
    lw_update.u0.copy_(to_dev(mesh, u_lex))
    sol = tlw.solve_lwfr(lw_update, (), 1e-6, None, time_step_computation=tlw.CFLBased(0.5), max_steps=K)
    assert sol.stats.naccept == K
    assert abs(sol.t - t_ref) <= 1e-13 * t_ref
    assert_close(to_lex(mesh, sol.u), u_ref, what=f"curved strips periodic={periodic} sc={sc} after {K} steps")
    semi.close()
