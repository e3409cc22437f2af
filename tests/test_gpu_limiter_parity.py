"""GPU parity of the Zhang-Shu positivity limiter (tlw_apply_positivity_limiter
and the limited device loop) against the CPU oracle, ≤ 1e-12 relative."""
import numpy as np
import pytest

from gpu_helpers import assert_close, setup, to_dev, to_lex, torch_mod

pytestmark = pytest.mark.gpu
THR = (1e-3, 2e-3)


def violating_state(E, n, seed):
    rng = np.random.default_rng(seed)
    rho = rng.uniform(0.5, 1.0, (E, n, n))
    v1 = rng.uniform(-0.5, 0.5, (E, n, n))
    v2 = rng.uniform(-0.5, 0.5, (E, n, n))
    p = rng.uniform(0.5, 1.0, (E, n, n))
    bad = rng.choice(E, size=E // 4, replace=False)
    for k, e in enumerate(bad):
        i, j = rng.integers(0, n, 2)
        if k % 3 == 0:
            rho[e, j, i] = -0.02
            v1[e, j, i] = v2[e, j, i] = 0.0
        elif k % 3 == 1:
            p[e, j, i] = -0.1
        else:
            rho[e, j, i], p[e, (j + 1) % n, i] = 0.0005, -0.05
    return np.stack([rho, rho * v1, rho * v2, p / 0.4 + 0.5 * rho * (v1 ** 2 + v2 ** 2)], -1)


@pytest.mark.parametrize("N", [3, 4])
def test_limiter_parity(oracle_mod, N):
    import trixilw_hip as tlw

    torch = torch_mod()
    semi, _, _, mesh = setup(oracle_mod, N, 4, domain=(0.0, 1.0))
    nx = mesh.nx
    o = oracle_mod.Oracle(N, nx, nx, 0.0, 0.0, 1.0 / nx, positivity=THR)
    u_lex = violating_state(nx * nx, N + 1, seed=7 + N)
    ref = o.positivity_limiter(u_lex)
    assert np.max(np.abs(ref - u_lex)) > 0.01
    u = to_dev(mesh, u_lex)
    lim = tlw.PositivityPreservingLimiterZhangShu(thresholds=THR, variables=(tlw.density, tlw.pressure))
    lim(u, None, semi, 0.0)
    torch.cuda.synchronize()
    assert_close(to_lex(mesh, u), ref, what=f"limiter N={N}")


@pytest.mark.parametrize("device_loop", [True, False])
def test_limited_riemann_loop_parity(oracle_mod, device_loop):
    """Four-quadrant Riemann problem, shock capturing, 0.98·LW CFL on 32²: the
    limiter first acts at step 52 (without it the state breaks down before
    step 60).  Both sides start from the oracle's step-50 state and take 10
    limited steps; the oracle's unlimited steps from there differ."""
    import trixilw_hip as tlw

    ic = tlw.FourQuadrantRiemann()
    semi, _, u_lex, mesh = setup(oracle_mod, 4, 5, periodic=False, domain=(0.0, 1.0), bc=ic, ic=ic, amp=0.0,
                                 sc=True, keep_aux=False)
    nx = mesh.nx
    thr = (5e-6, 5e-6)

    def oracle(pos):
        return oracle_mod.Oracle(4, nx, nx, 0.0, 0.0, 1.0 / nx, periodic=(False, False), bc_kind=ic.kind,
                                 bc_par=ic.params, shock_capturing=True, alpha_max=0.5, positivity=pos)

    cfl = tlw.trixi2lw(0.98, semi.solver)
    u50, t50, _ = oracle(thr).solve(u_lex, 0.0, 1e9, cfl, max_steps=50)
    K = 10
    u_ref, t_ref, dts = oracle(thr).solve(u50, t50, 1e9, cfl, max_steps=K)
    u_unl, _, _ = oracle(None).solve(u50, t50, 1e9, cfl, max_steps=K)
    assert np.isfinite(u_ref).all()
    assert not np.isfinite(u_unl).all() or np.max(np.abs(u_unl - u_ref)) > 1e-6
    lw_update = tlw.semidiscretize(semi, semi.time_discretization, (t50, float("inf")))
    lw_update.u0.copy_(to_dev(mesh, u50))
    lim = tlw.PositivityPreservingLimiterZhangShu(thresholds=thr, variables=(tlw.density, tlw.pressure))
    sol = tlw.solve_lwfr(lw_update, (), 1e-6, None, time_step_computation=tlw.CFLBased(cfl), limiters=(lim,),
                         max_steps=K, device_loop=device_loop)
    assert sol.stats.naccept == K
    assert abs(sol.t - t_ref) <= 1e-13 * t_ref
    assert_close(to_lex(mesh, sol.u), u_ref, what=f"limited Riemann, device_loop={device_loop}")


@pytest.mark.parametrize("td,K", [("lw", 12), ("mdrk", 8)])
def test_limited_curved_loop_parity(oracle_mod, td, K):
    """The limiter on the curvilinear StructuredMesh (weights only, any
    AbstractMesh{2}): the warped vortex with a density threshold above the
    vortex core's density, so elements are limited every step, device loop
    against the oracle's limited solve (the next dt from the limited state with
    the curvilinear speeds), ≤ 1e-12."""
    import trixilw_hip as tlw
    from gpu_helpers import setup_curved

    semi, o0, u_lex, mesh = setup_curved(oracle_mod, 4, 8, 8, td=td)
    thr = (0.75, 1e-3)
    L = 20.0
    o = oracle_mod.Oracle(4, 8, 8, 0.0, 0.0, L / 8, gamma=1.4, mapping=(oracle_mod.MAP_WARP, (L, 0.05)),
                          positivity=thr)
    cfl = 0.5 * tlw.trixi2lw(1.0, semi.solver) * (0.5 if td == "mdrk" else 1.0)
    u_ref, t_ref, _ = o.solve(u_lex, 0.0, 1e9, cfl, max_steps=K, mdrk=(td == "mdrk"))
    u_unl, _, _ = o0.solve(u_lex, 0.0, 1e9, cfl, max_steps=K, mdrk=(td == "mdrk"))
    assert np.max(np.abs(u_unl - u_ref)) > 1e-4  # the limiter acts
    upd = tlw.semidiscretize(semi, semi.time_discretization, (0.0, float("inf")))
    upd.u0.copy_(to_dev(mesh, u_lex))
    lim = tlw.PositivityPreservingLimiterZhangShu(thresholds=thr, variables=(tlw.density, tlw.pressure))
    sol = tlw.solve_lwfr(upd, (), 1e-6, None, time_step_computation=tlw.CFLBased(cfl), limiters=(lim,),
                         max_steps=K)
    assert sol.stats.naccept == K
    assert abs(sol.t - t_ref) <= 1e-13 * t_ref
    assert_close(to_lex(mesh, sol.u), u_ref, what=f"limited curved {td}")
    semi.close()


@pytest.mark.parametrize("curved", [False, True])
def test_limited_navier_stokes_loop_parity(oracle_mod, curved):
    """The limiter with the hyperbolic-parabolic path (TreeMesh and warped
    mesh): density threshold inside the shear state's density range, device
    loop with the viscous dt bound against the oracle, ≤ 1e-12."""
    import trixilw_hip as tlw

    thr = (0.95, 1e-3)
    K = 10
    if curved:
        from test_gpu_curved_parabolic_parity import setup_cns

        semi, o0, u_lex, mesh = setup_cns(oracle_mod, 4, 8, A=0.08)
        o = oracle_mod.Oracle(4, 8, 8, 0.0, 0.0, 2.0 / 8, gamma=1.4, mapping=(oracle_mod.MAP_WARP, (2.0, 0.08)),
                              parabolic="navier_stokes", mu=0.01, prandtl=0.72, positivity=thr)
    else:
        from test_gpu_parabolic_parity import setup_ns

        semi, o0, u_lex, mesh = setup_ns(oracle_mod, 4, 3, mu=0.002)
        o = oracle_mod.Oracle(4, 8, 8, -1.0, -1.0, 2.0 / 8, gamma=1.4, parabolic="navier_stokes", mu=0.002,
                              prandtl=0.72, positivity=thr)
    cfl = 0.5
    u_ref, t_ref, _ = o.solve(u_lex, 0.0, 1e9, cfl, max_steps=K)
    u_unl, _, _ = o0.solve(u_lex, 0.0, 1e9, cfl, max_steps=K)
    assert np.max(np.abs(u_unl - u_ref)) > 1e-4
    upd = tlw.semidiscretize(semi, semi.time_discretization, (0.0, float("inf")))
    upd.u0.copy_(to_dev(mesh, u_lex))
    lim = tlw.PositivityPreservingLimiterZhangShu(thresholds=thr, variables=(tlw.density, tlw.pressure))
    sol = tlw.solve_lwfr(upd, (), 1e-6, None, time_step_computation=tlw.CFLBased(cfl), limiters=(lim,),
                         max_steps=K)
    assert sol.stats.naccept == K
    assert abs(sol.t - t_ref) <= 1e-13 * t_ref
    assert_close(to_lex(mesh, sol.u), u_ref, what=f"limited Navier-Stokes curved={curved}")
    semi.close()
