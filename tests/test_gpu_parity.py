"""GPU parity: libtrixilw_hip.so (through the C ABI) against the CPU oracle.

Bar: relative L2 and L∞ differences ≤ 1e-12 (FP64), on the same mesh, state
and dt.  Inputs are seeded.
"""
import numpy as np
import pytest

from gpu_helpers import TOL, assert_close, setup, to_dev, to_lex, torch_mod

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("N", [3, 4])
@pytest.mark.parametrize("level", [2, 5])
def test_rhs_lw_parity(oracle_mod, N, level):
    import trixilw_hip as tlw

    torch = torch_mod()
    semi, o, u_lex, mesh = setup(oracle_mod, N, level, seed=N * 10 + level)
    dt = 0.8 * 0.98 * (N + 1) * tlw.lw_fourier_cfl(N) * o.max_dt(u_lex)
    t = 0.25
    du_ref = o.rhs(u_lex, t, dt)
    u = to_dev(mesh, u_lex)
    du = torch.empty_like(u)
    semi.cache["dt"][0] = dt
    tlw.rhs(du, u, semi, t)
    torch.cuda.synchronize()
    assert_close(to_lex(mesh, du), du_ref, what="du")
    U = semi.get_aux("U")[mesh.lexicographic_permutation()]
    F = semi.get_aux("F")[mesh.lexicographic_permutation()]
    assert_close(U, o.aux("U").reshape(U.shape), what="U")
    assert_close(F, o.aux("F").reshape(F.shape), what="F")


@pytest.mark.parametrize("N", [3, 4])
def test_rhs_mdrk_parity(oracle_mod, N):
    import trixilw_hip as tlw

    torch = torch_mod()
    semi, o, u_lex, mesh = setup(oracle_mod, N, 4, td="mdrk", seed=5 + N)
    dt = 0.4 * o.max_dt(u_lex)
    du1_ref = o.rhs(u_lex, 0.0, dt, kind="mdrk1")
    us_lex = u_lex + dt * du1_ref
    o.set_us(us_lex)
    du2_ref = o.rhs(u_lex, 0.0, dt, kind="mdrk2")
    u = to_dev(mesh, u_lex)
    du = torch.empty_like(u)
    semi.cache["dt"][0] = dt
    tlw.rhs_mdrk1(du, u, semi, 0.0)
    assert_close(to_lex(mesh, du), du1_ref, what="mdrk1")
    semi.set_us(torch.add(u, du, alpha=dt))
    tlw.rhs_mdrk2(du, u, semi, 0.0)
    torch.cuda.synchronize()
    assert_close(to_lex(mesh, du), du2_ref, what="mdrk2")


@pytest.mark.parametrize("N", [3, 4])
def test_max_dt_parity(oracle_mod, N):
    semi, o, u_lex, mesh = setup(oracle_mod, N, 4, seed=2)
    assert abs(semi.max_dt(to_dev(mesh, u_lex)) / o.max_dt(u_lex) - 1) < 1e-14


@pytest.mark.parametrize("td", ["lw", "mdrk"])
def test_dirichlet_boundary_parity(oracle_mod, td):
    import trixilw_hip as tlw

    torch = torch_mod()
    ic = tlw.IsentropicVortex(center=(0.3, 0.2), period=(1e6, 1e6))
    semi, o, u_lex, mesh = setup(oracle_mod, 4, 3, td=td, periodic=False, domain=(-5.0, 5.0), bc=ic, ic=ic,
                                 seed=9)
    dt = 0.4 * o.max_dt(u_lex)
    t = 0.3
    u = to_dev(mesh, u_lex)
    du = torch.empty_like(u)
    semi.cache["dt"][0] = dt
    if td == "lw":
        tlw.rhs(du, u, semi, t)
        assert_close(to_lex(mesh, du), o.rhs(u_lex, t, dt), what="lw dirichlet")
    else:
        du1_ref = o.rhs(u_lex, t, dt, kind="mdrk1")
        tlw.rhs_mdrk1(du, u, semi, t)
        assert_close(to_lex(mesh, du), du1_ref, what="mdrk1 dirichlet")
        o.set_us(u_lex + dt * du1_ref)
        semi.set_us(torch.add(u, du, alpha=dt))
        tlw.rhs_mdrk2(du, u, semi, t)
        assert_close(to_lex(mesh, du), o.rhs(u_lex, t, dt, kind="mdrk2"), what="mdrk2 dirichlet")


def test_riemann_boundary_parity(oracle_mod):
    import trixilw_hip as tlw

    torch = torch_mod()
    ic = tlw.FourQuadrantRiemann()
    semi, o, u_lex, mesh = setup(oracle_mod, 4, 3, periodic=False, domain=(0.0, 1.0), bc=ic, ic=ic, amp=0.0)
    dt = 0.2 * o.max_dt(u_lex)
    u = to_dev(mesh, u_lex)
    du = torch.empty_like(u)
    semi.cache["dt"][0] = dt
    tlw.rhs(du, u, semi, 0.0)
    assert_close(to_lex(mesh, du), o.rhs(u_lex, 0.0, dt), what="riemann")


def test_structured_flux_order_parity(oracle_mod):
    import trixilw_hip as tlw

    torch = torch_mod()
    semi, o, u_lex, mesh = setup(oracle_mod, 4, 3, flux_order="structured", seed=4)
    dt = 0.4 * o.max_dt(u_lex)
    u = to_dev(mesh, u_lex)
    du = torch.empty_like(u)
    semi.cache["dt"][0] = dt
    tlw.rhs(du, u, semi, 0.0)
    assert_close(to_lex(mesh, du), o.rhs(u_lex, 0.0, dt), what="structured order")


@pytest.mark.parametrize("ordering", ["morton", "strips"])
def test_element_ordering_parity(oracle_mod, ordering):
    """Any element order gives the same residual (the bench uses two-row strips,
    whose 10-element runs are 5×2 blocks)."""
    import trixilw_hip as tlw

    torch = torch_mod()
    semi, o, u_lex, mesh = setup(oracle_mod, 4, 4, ordering=ordering, seed=8)
    dt = 0.4 * o.max_dt(u_lex)
    u = to_dev(mesh, u_lex)
    du = torch.empty_like(u)
    semi.cache["dt"][0] = dt
    for tile in (5, 10, 15):
        semi.elements_per_workgroup(tile)
        tlw.rhs(du, u, semi, 0.0)
        torch.cuda.synchronize()
        assert_close(to_lex(mesh, du), o.rhs(u_lex, 0.0, dt), what=f"{ordering} tile {tile}")


@pytest.mark.parametrize("N,td", [(4, "lw"), (3, "lw"), (3, "mdrk"), (4, "mdrk")])
def test_device_loop_k_steps(oracle_mod, N, td):
    """State after K steps of the device-resident loop matches the oracle's loop."""
    import trixilw_hip as tlw

    semi, o, u_lex, mesh = setup(oracle_mod, N, 4, td=td, seed=21, keep_aux=False)
    cfl = 0.98 * tlw.trixi2lw(1.0, semi.solver) * (0.5 if td == "mdrk" else 1.0)
    K = 12
    u_ref, t_ref, dts = o.solve(u_lex, 0.0, 1e9, cfl, max_steps=K, mdrk=(td == "mdrk"))
    # an infinite final time never finishes (Julia: isapprox(t, Inf) is false)
    lw_update = tlw.semidiscretize(semi, semi.time_discretization, (0.0, float("inf")))
    lw_update.u0.copy_(to_dev(mesh, u_lex))
    sol = tlw.solve_lwfr(lw_update, (), 1e-6, None, time_step_computation=tlw.CFLBased(cfl), max_steps=K)
    assert sol.stats.naccept == K
    assert abs(sol.t - t_ref) <= 1e-13 * t_ref
    assert_close(to_lex(mesh, sol.u), u_ref, what=f"{td} N={N} after {K} steps")


def test_host_loop_matches_device_loop(oracle_mod):
    """perform_step! through the rhs entry points and the device loop agree."""
    import trixilw_hip as tlw

    semi, o, u_lex, mesh = setup(oracle_mod, 4, 3, seed=13, keep_aux=False)
    cfl = tlw.trixi2lw(0.98, semi.solver)
    res = []
    for device_loop in (False, True):
        lw_update = tlw.semidiscretize(semi, semi.time_discretization, (0.0, 0.2))
        lw_update.u0.copy_(to_dev(mesh, u_lex))
        sol = tlw.solve_lwfr(lw_update, (), 1e-6, None, time_step_computation=tlw.CFLBased(cfl),
                             device_loop=device_loop)
        res.append((to_lex(mesh, sol.u), sol.t, sol.stats.naccept))
    assert res[0][2] == res[1][2]
    assert res[0][1] == pytest.approx(0.2, abs=1e-14) and res[1][1] == pytest.approx(0.2, abs=1e-14)
    assert_close(res[1][0], res[0][0], what="host vs device loop")


def test_c0_solve_to_t1_matches_oracle(oracle_mod):
    """Config 0: isentropic vortex (ε=5, Mach 0.5), 32×32 periodic TreeMesh, N=4,
    CFL 0.98·LW limit, t ∈ [0, 1]: same step count, same t, state within 1e-12."""
    import trixilw_hip as tlw

    ic = tlw.IsentropicVortex()
    semi, o, u_lex, mesh = setup(oracle_mod, 4, 5, ic=ic, amp=0.0, keep_aux=False)
    cfl = tlw.trixi2lw(0.98, semi.solver)
    u_ref, t_ref, dts = o.solve(u_lex, 0.0, 1.0, cfl)
    lw_update = tlw.semidiscretize(semi, semi.time_discretization, (0.0, 1.0))
    sol = tlw.solve_lwfr(lw_update, (), 1e-6, None, time_step_computation=tlw.CFLBased(cfl))
    assert sol.stats.naccept == len(dts)
    assert sol.t == pytest.approx(1.0, abs=1e-14)
    u = to_lex(mesh, sol.u)
    assert_close(u, u_ref, what="C0 final state")
    x, y = o.node_coords()
    err = np.sqrt(np.mean((u - ic(x, y, sol.t)) ** 2))
    assert err < 1e-3


def test_conservation_and_free_stream_at_scale():
    """Size-independent properties on a 256×256 mesh (1.6 M DOFs): total mass,
    momentum and energy are conserved over 20 steps; a constant state stays constant."""
    import trixilw_hip as tlw

    torch = torch_mod()
    eq = tlw.CompressibleEulerEquations2D(1.4)
    mesh = tlw.TreeMesh((-10, -10), (10, 10), 8)
    solver = tlw.DGSEM(4)
    for ic, check_const in ((tlw.IsentropicVortex(), False), (tlw.ConstantState(), True)):
        semi = tlw.SemidiscretizationHyperbolic(mesh, tlw.LW(), eq, ic, solver)
        lw_update = tlw.semidiscretize(semi, tlw.LW(), (0.0, 1e9))
        u0 = lw_update.u0.clone()
        sol = tlw.solve_lwfr(lw_update, (), 1e-6, None, time_step_computation=tlw.CFLBased(0.5), max_steps=20)
        w = torch.from_numpy(np.outer(solver.basis.weights, solver.basis.weights)).cuda()
        m0 = torch.einsum("ji,ejiv->v", w, u0)
        m1 = torch.einsum("ji,ejiv->v", w, sol.u)
        scale = torch.einsum("ji,ejiv->v", w, u0.abs())
        assert torch.all((m1 - m0).abs() <= 1e-13 * scale), (m0, m1)
        if check_const:
            assert (sol.u - u0).abs().max().item() < 1e-12
        semi.close()


def test_nonadmissible_step_is_reported(oracle_mod):
    """The under-resolved 4×4 vortex reaches p < 0 in its third step (the oracle
    shows the same); the device loop reports it instead of continuing."""
    import trixilw_hip as tlw
    from trixilw_hip._lib import NonAdmissibleState

    semi, o, u_lex, mesh = setup(oracle_mod, 4, 2, ic=tlw.IsentropicVortex(), amp=0.0, keep_aux=False)
    cfl = tlw.trixi2lw(0.98, semi.solver)
    u2, _, _ = o.solve(u_lex, 0.0, 1e9, cfl, max_steps=2)
    u3, _, _ = o.solve(u_lex, 0.0, 1e9, cfl, max_steps=3)
    p = lambda u: 0.4 * (u[..., 3] - 0.5 * (u[..., 1] ** 2 + u[..., 2] ** 2) / u[..., 0])
    assert p(u2).min() > 0 and p(u3).min() < 0
    lw_update = tlw.semidiscretize(semi, semi.time_discretization, (0.0, 1e9))
    sol = tlw.solve_lwfr(lw_update, (), 1e-6, None, time_step_computation=tlw.CFLBased(cfl), max_steps=2)
    assert sol.stats.naccept == 2
    lw_update = tlw.semidiscretize(semi, semi.time_discretization, (0.0, 1e9))
    with pytest.raises(NonAdmissibleState):
        tlw.solve_lwfr(lw_update, (), 1e-6, None, time_step_computation=tlw.CFLBased(cfl), max_steps=3)


def test_nonadmissible_state_is_reported():
    """A state with negative pressure is rejected (the reference raises DomainError)."""
    import trixilw_hip as tlw
    from trixilw_hip._lib import NonAdmissibleState

    eq = tlw.CompressibleEulerEquations2D(1.4)
    mesh = tlw.TreeMesh((-1, -1), (1, 1), 2)
    semi = tlw.SemidiscretizationHyperbolic(mesh, tlw.LW(), eq, tlw.ConstantState(), tlw.DGSEM(4))
    u = semi.compute_coefficients(0.0)
    u[3, 2, 2, 3] = -1.0  # negative energy -> negative pressure
    with pytest.raises(NonAdmissibleState):
        semi.max_dt(u)


@pytest.mark.parametrize("N", [3, 4])
def test_every_tile_matches_oracle(oracle_mod, N):
    """Each compiled volume tile (elements per workgroup) gives the oracle's
    residual, and the autotuner picks one of them."""
    import trixilw_hip as tlw

    torch = torch_mod()
    semi, o, u_lex, mesh = setup(oracle_mod, N, 4, ordering="morton", seed=31 + N)
    dt = 0.5 * o.max_dt(u_lex)
    du_ref = o.rhs(u_lex, 0.0, dt)
    u = to_dev(mesh, u_lex)
    du = torch.empty_like(u)
    semi.cache["dt"][0] = dt
    tiles = (5, 10, 15) if N == 4 else (4, 8, 16)
    for tile in tiles:
        semi.elements_per_workgroup(tile)
        tlw.rhs(du, u, semi, 0.0)
        torch.cuda.synchronize()
        assert_close(to_lex(mesh, du), du_ref, what=f"tile {tile}")
    chosen, ms = semi.autotune(u, reps=1)
    assert chosen in tiles and set(ms) == set(tiles) and all(v > 0 for v in ms.values())
    with pytest.raises(tlw._lib.TLWError):
        semi.elements_per_workgroup(7)
