"""Caller-defined Dirichlet data on the device (TLW_BV_TABLE): the host
evaluates a boundary_value_function that is none of the built-in kinds at the
boundary face nodes and uploads it; the device rhs, MDRK stages, host-driven
solve and (steady data) device loop match the oracle, which calls the same
Python function pointwise at its own face nodes and times, ≤ 1e-12."""
import numpy as np
import pytest

from bc_functions import steady_shear, waves
from gpu_helpers import assert_close, setup, setup_curved, to_dev, to_lex, torch_mod

pytestmark = pytest.mark.gpu


def _setup(oracle_mod, N, fn, td="lw", flux_order="tree", steady=False, level=3, ordering="lexicographic"):
    import trixilw_hip as tlw

    bc = tlw.BoundaryConditionDirichlet(fn, steady=steady)
    ic = lambda x, y, t, equations=None: fn((x, y), t)  # noqa: E731
    return setup(oracle_mod, N, level, td=td, periodic=False, domain=(0.0, 1.0), bc=bc, ic=ic, seed=3 + N,
                 flux_order=flux_order, ordering=ordering)


@pytest.mark.parametrize("N", [3, 4])
@pytest.mark.parametrize("flux_order", ["tree", "structured"])
def test_table_rhs_lw(oracle_mod, N, flux_order):
    import trixilw_hip as tlw

    torch = torch_mod()
    semi, o, u_lex, mesh = _setup(oracle_mod, N, waves, flux_order=flux_order, ordering="strips")
    dt, t = 0.4 * o.max_dt(u_lex), 0.37
    u = to_dev(mesh, u_lex)
    du = torch.empty_like(u)
    semi.cache["dt"][0] = dt
    tlw.rhs(du, u, semi, t)
    assert_close(to_lex(mesh, du), o.rhs(u_lex, t, dt), what="lw table")
    # a second call at another time re-tabulates the data
    tlw.rhs(du, u, semi, t + 0.2)
    assert_close(to_lex(mesh, du), o.rhs(u_lex, t + 0.2, dt), what="lw table, second call")
    semi.close()


@pytest.mark.parametrize("N", [3, 4])
def test_table_rhs_mdrk_stages(oracle_mod, N):
    """Stage 1 averages over half the step (the TreeMesh scaling factor), stage 2
    over the whole step."""
    import trixilw_hip as tlw

    torch = torch_mod()
    semi, o, u_lex, mesh = _setup(oracle_mod, N, waves, td="mdrk")
    dt, t = 0.3 * o.max_dt(u_lex), 0.11
    u = to_dev(mesh, u_lex)
    du = torch.empty_like(u)
    semi.cache["dt"][0] = dt
    du1_ref = o.rhs(u_lex, t, dt, kind="mdrk1")
    tlw.rhs_mdrk1(du, u, semi, t)
    assert_close(to_lex(mesh, du), du1_ref, what="mdrk1 table")
    o.set_us(u_lex + dt * du1_ref)
    semi.set_us(torch.add(u, du, alpha=dt))
    tlw.rhs_mdrk2(du, u, semi, t)
    assert_close(to_lex(mesh, du), o.rhs(u_lex, t, dt, kind="mdrk2"), what="mdrk2 table")
    semi.close()


@pytest.mark.parametrize("td", ["lw", "mdrk"])
def test_table_host_loop_solve(oracle_mod, td):
    """solve_lwfr with time-dependent data runs the host loop (the device loop
    cannot evaluate a host function at its own dt) and matches the oracle's
    solve, step sizes included."""
    import trixilw_hip as tlw

    semi, o, u_lex, mesh = _setup(oracle_mod, 4, waves, td=td)
    cfl = 0.5 * tlw.trixi2lw(1.0, semi.solver) * (0.5 if td == "mdrk" else 1.0)
    K = 6
    u_ref, t_ref, dts = o.solve(u_lex, 0.0, 1e9, cfl, max_steps=K, mdrk=(td == "mdrk"))
    upd = tlw.semidiscretize(semi, tlw.get_time_discretization(semi.solver), (0.0, 1e9))
    upd.u0.copy_(to_dev(mesh, u_lex))
    sol = tlw.solve_lwfr(upd, (), 1e-6, None, time_step_computation=tlw.CFLBased(cfl), max_steps=K)
    assert len(sol.dts) == K
    assert np.allclose(sol.dts, dts, rtol=1e-12, atol=0)
    assert_close(to_lex(mesh, sol.u), u_ref, what=f"{td} host loop")
    with pytest.raises(tlw._lib.TLWError):  # per-call data in the device loop
        tlw.solve_lwfr(upd, (), 1e-6, None, time_step_computation=tlw.CFLBased(cfl), max_steps=1,
                       device_loop=True)
    semi.close()


def test_table_steady_device_loop(oracle_mod):
    """Steady data is uploaded once and the device loop runs with it."""
    import trixilw_hip as tlw

    semi, o, u_lex, mesh = _setup(oracle_mod, 4, steady_shear, steady=True)
    cfl = 0.5 * tlw.trixi2lw(1.0, semi.solver)
    K = 10
    u_ref, t_ref, _ = o.solve(u_lex, 0.0, 1e9, cfl, max_steps=K)
    upd = tlw.semidiscretize(semi, tlw.LW(), (0.0, 1e9))
    upd.u0.copy_(to_dev(mesh, u_lex))
    sol = tlw.solve_lwfr(upd, (), 1e-6, None, time_step_computation=tlw.CFLBased(cfl), max_steps=K)
    assert abs(sol.t - t_ref) <= 1e-14 * t_ref
    assert_close(to_lex(mesh, sol.u), u_ref, what="steady table, device loop")
    semi.close()


def test_table_curved_mesh(oracle_mod):
    """The curvilinear boundary path (own face normals, full-step average) with
    a caller-defined function on a warped, non-periodic mesh."""
    import trixilw_hip as tlw

    torch = torch_mod()
    L = 20.0
    ic = tlw.IsentropicVortex(center=(L / 2, L / 2), period=(1e9, 1e9))
    f = lambda x, t, equations=None: ic(x[0], x[1], t + 0.25)  # noqa: E731  (shifted: not the IC)
    semi0, o0, u_lex, mesh = setup_curved(oracle_mod, 4, 4, 4, periodic=False)
    semi0.close()
    o = oracle_mod.Oracle(4, 4, 4, 0.0, 0.0, L / 4, gamma=1.4, mapping=(oracle_mod.MAP_WARP, (L, 0.05)),
                          periodic=(False, False), bc_kind=oracle_mod.BV_CALLBACK,
                          bc_fn=lambda x, y, t: f((x, y), t))
    solver = tlw.DGSEM(4, volume_integral=tlw.VolumeIntegralFR(tlw.LW()))
    semi = tlw.SemidiscretizationHyperbolic(mesh, tlw.LW(), tlw.CompressibleEulerEquations2D(1.4), ic, solver,
                                            boundary_conditions=tlw.BoundaryConditionDirichlet(f))
    dt, t = 0.3 * o.max_dt(u_lex), 0.2
    u = to_dev(mesh, u_lex)
    du = torch.empty_like(u)
    semi.cache["dt"][0] = dt
    tlw.rhs(du, u, semi, t)
    assert_close(to_lex(mesh, du), o.rhs(u_lex, t, dt), what="curved table")
    semi.close()
