"""GPU parity of the adaptive (error-controlled) LW loop against the oracle's
solve_adaptive: same accepted and rejected step counts, final time and state
within 1e-12 relative (L2 and L∞).  Inputs are seeded."""
import numpy as np
import pytest

from gpu_helpers import assert_close, to_dev, to_lex, torch_mod

pytestmark = pytest.mark.gpu

TOL = {"abstol": 1e-6, "reltol": 1e-6}


def vortex_setup(oracle_mod, level=4, N=4, td="lw", sc=False):
    import trixilw_hip as tlw

    eq = tlw.CompressibleEulerEquations2D(1.4)
    mesh = tlw.TreeMesh((-10.0, -10.0), (10.0, 10.0), level)
    vi = tlw.VolumeIntegralFR(tlw.MDRK() if td == "mdrk" else tlw.LW())
    if sc:
        vi = tlw.VolumeIntegralFRShockCapturing(tlw.IndicatorHennemannGassner(eq, None, alpha_max=0.5), vi)
    solver = tlw.DGSEM(N, volume_integral=vi)
    ic = tlw.IsentropicVortex(center=(0.7, -0.4))
    semi = tlw.SemidiscretizationHyperbolic(mesh, tlw.get_time_discretization(solver), eq, ic, solver)
    nx = mesh.nx
    o = oracle_mod.Oracle(N, nx, nx, -10.0, -10.0, 20.0 / nx, tolerances=(TOL["abstol"], TOL["reltol"]),
                          shock_capturing=sc, alpha_max=0.5)
    x, y = o.node_coords()
    rng = np.random.default_rng(5)
    u_lex = ic(x, y, 0.0) * (1.0 + 1e-3 * rng.standard_normal(x.shape + (4,)))
    return semi, o, u_lex, mesh


def run_device(semi, mesh, u_lex, tf, dt_initial, max_steps=None):
    import trixilw_hip as tlw

    lw_update = tlw.semidiscretize(semi, semi.time_discretization, (0.0, tf))
    lw_update.u0.copy_(to_dev(mesh, u_lex))
    return tlw.solve_lwfr(lw_update, (), dt_initial, TOL, time_step_computation=tlw.Adaptive(), max_steps=max_steps)


@pytest.mark.parametrize("dt_initial,tf", [(1e-3, 0.5), (5.0, 0.3)])
def test_adaptive_loop_matches_oracle(oracle_mod, dt_initial, tf):
    """From a small dt (growth at the 1.5 cap, then control) and from an
    unstable one (rejected attempts first)."""
    torch_mod()
    semi, o, u_lex, mesh = vortex_setup(oracle_mod)
    u_ref, t_ref, dts, nrej = o.solve_adaptive(u_lex, 0.0, tf, dt_initial, max_attempts=20000)
    sol = run_device(semi, mesh, u_lex, tf, dt_initial)
    assert sol.stats.naccept == len(dts)
    assert sol.stats.nreject == nrej
    if dt_initial > 1.0:
        assert nrej >= 1
    assert abs(sol.t - t_ref) <= 1e-13 * t_ref
    assert_close(to_lex(mesh, sol.u), u_ref, what=f"adaptive dt0={dt_initial}")


def test_adaptive_with_shock_capturing_matches_oracle(oracle_mod):
    """The estimate comes from the LW kernel also when the volume term is blended."""
    torch_mod()
    semi, o, u_lex, mesh = vortex_setup(oracle_mod, sc=True)
    u_ref, t_ref, dts, nrej = o.solve_adaptive(u_lex, 0.0, 0.2, 1e-3, max_attempts=5000)
    sol = run_device(semi, mesh, u_lex, 0.2, 1e-3)
    assert sol.stats.naccept == len(dts) and sol.stats.nreject == nrej
    assert abs(sol.t - t_ref) <= 1e-13 * t_ref
    assert_close(to_lex(mesh, sol.u), u_ref, what="adaptive with shock capturing")


def test_adaptive_steps_at_scale_conserve(oracle_mod):
    """256² mesh: 30 accepted steps conserve mass, momentum and energy."""
    import trixilw_hip as tlw

    torch = torch_mod()
    semi, _, _, mesh = vortex_setup(oracle_mod, level=8)
    lw_update = tlw.semidiscretize(semi, semi.time_discretization, (0.0, 100.0))
    u0 = lw_update.u0.clone()
    sol = tlw.solve_lwfr(lw_update, (), 1e-3, TOL, time_step_computation=tlw.Adaptive(), max_steps=30)
    assert sol.stats.naccept == 30
    w = torch.from_numpy(np.outer(semi.solver.basis.weights, semi.solver.basis.weights)).cuda()
    m0 = torch.einsum("ji,ejiv->v", w, u0)
    m1 = torch.einsum("ji,ejiv->v", w, sol.u)
    scale = torch.einsum("ji,ejiv->v", w, u0.abs())
    assert torch.all((m1 - m0).abs() <= 1e-13 * scale), (m0, m1)
    semi.close()


def test_adaptive_rejects_unsupported(oracle_mod):
    import trixilw_hip as tlw
    from trixilw_hip import _lib

    torch_mod()
    for N, td in ((3, "lw"), (4, "mdrk")):
        semi, _, u_lex, mesh = vortex_setup(oracle_mod, level=3, N=N, td=td)
        with pytest.raises(_lib.TLWError) as ei:
            run_device(semi, mesh, u_lex, 0.1, 1e-3)
        assert ei.value.code == _lib.TLW_ENOTSUP
        semi.close()
    semi, _, u_lex, mesh = vortex_setup(oracle_mod, level=3)
    semi.set_positivity_limiter((1e-8, 1e-8))
    h = semi.handle
    with pytest.raises(_lib.TLWError) as ei:
        _lib.check(_lib.lib().tlw_solve_init_adaptive(h, 0.0, 1.0, 1e-3, 1e-6, 1e-6), h)
    assert ei.value.code == _lib.TLW_ENOTSUP
    semi.close()


@pytest.mark.parametrize("N,dt_initial", [(4, 1e-3), (3, 1e-3), (4, 5.0)])
def test_adaptive_curved_matches_oracle(oracle_mod, N, dt_initial):
    """Adaptive steps on the warped StructuredMesh: the temporal-error estimate
    of the curvilinear LW kernel (structured dg_2d.jl, N=3 and N=4) drives the
    same controller; accepted and rejected counts, final time and state match
    the oracle's solve_adaptive ≤ 1e-12."""
    import trixilw_hip as tlw
    from gpu_helpers import setup_curved

    torch_mod()
    semi, _, u_lex, mesh = setup_curved(oracle_mod, N, 6, 6)
    L = 20.0
    o = oracle_mod.Oracle(N, 6, 6, 0.0, 0.0, L / 6, gamma=1.4, mapping=(oracle_mod.MAP_WARP, (L, 0.05)),
                          tolerances=(TOL["abstol"], TOL["reltol"]))
    tf = 0.4
    u_ref, t_ref, dts, nrej = o.solve_adaptive(u_lex, 0.0, tf, dt_initial, max_attempts=20000)
    sol = run_device(semi, mesh, u_lex, tf, dt_initial)
    assert sol.stats.naccept == len(dts) > 3
    assert sol.stats.nreject == nrej
    if dt_initial > 1.0:
        assert nrej >= 1
    assert abs(sol.t - t_ref) <= 1e-13 * t_ref
    assert_close(to_lex(mesh, sol.u), u_ref, what=f"curved adaptive N={N} dt0={dt_initial}")
    semi.close()


@pytest.mark.parametrize("dt_initial,tf", [(1e-5, 0.05), (0.5, 0.05)])
def test_adaptive_navier_stokes_matches_oracle(oracle_mod, dt_initial, tf):
    """Adaptive steps of the Navier-Stokes LW N=4 kernel on the periodic
    TreeMesh (its temporal error, dg_2d_parabolic.jl:1841-1855): accepted and
    rejected counts, final time and state equal the oracle's solve_adaptive
    (≤1e-12), from a small dt_initial and from an unstable one."""
    import trixilw_hip as tlw
    from test_gpu_parabolic_parity import setup_ns

    torch_mod()
    semi, o0, u_lex, mesh = setup_ns(oracle_mod, 4, 3, mu=0.01)
    o = oracle_mod.Oracle(4, mesh.nx, mesh.nx, -1.0, -1.0, 2.0 / mesh.nx, gamma=1.4, parabolic="navier_stokes",
                          mu=0.01, prandtl=0.72, tolerances=(TOL["abstol"], TOL["reltol"]))
    u_ref, t_ref, dts, nrej = o.solve_adaptive(u_lex, 0.0, tf, dt_initial, max_attempts=20000)
    sol = run_device(semi, mesh, u_lex, tf, dt_initial)
    assert sol.stats.naccept == len(dts) > 3
    assert sol.stats.nreject == nrej
    if dt_initial > 0.1:
        assert nrej >= 1
    assert abs(sol.t - t_ref) <= 1e-13 * t_ref
    assert_close(to_lex(mesh, sol.u), u_ref, what=f"Navier-Stokes adaptive dt0={dt_initial}")
    semi.close()
