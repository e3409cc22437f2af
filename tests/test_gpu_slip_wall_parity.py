"""GPU parity of the reflecting wall (boundary_condition_slip_wall_vertical on
x faces, _horizontal on y faces, src/equations/equations.jl:95-139) on the
TreeMesh against the CPU oracle: LW residuals in both flux orders, both MDRK
stages, shock capturing, K-step device-loop states within 1e-12 relative (L2
and L∞), and the rejection on a curvilinear mesh.  Inputs are seeded."""
import numpy as np
import pytest

from gpu_helpers import assert_close, to_dev, to_lex, torch_mod
from test_oracle_slip_wall_pins import symmetric_state

pytestmark = pytest.mark.gpu


def walls():
    """The reference's NamedTuple (x_neg, x_pos, y_neg, y_pos) of the lid-driven cavity's walls."""
    import trixilw_hip as tlw

    return {"x_neg": tlw.boundary_condition_slip_wall_vertical, "x_pos": tlw.boundary_condition_slip_wall_vertical,
            "y_neg": tlw.boundary_condition_slip_wall_horizontal, "y_pos": tlw.boundary_condition_slip_wall_horizontal}


def setup_walls(oracle_mod, N, level, td="lw", flux_order="tree", sc=False, seed=0, amp=1e-2, keep_aux=False):
    """[0, 1]² with walls on all four sides: the mirror-symmetric state of the
    oracle pins, shifted so that it is not symmetric about the walls, with
    seeded noise."""
    import trixilw_hip as tlw

    eq = tlw.CompressibleEulerEquations2D(1.4)
    mesh = tlw.TreeMesh((0.0, 0.0), (1.0, 1.0), level, periodicity=False)

    def ic(x, y, t=0.0):
        return symmetric_state(x + 0.3, y - 0.2)

    vi = tlw.VolumeIntegralFR(tlw.MDRK() if td == "mdrk" else tlw.LW())
    if sc:
        ind = tlw.IndicatorHennemannGassner(eq, None, alpha_max=0.5, alpha_min=0.001, alpha_smooth=True)
        vi = tlw.VolumeIntegralFRShockCapturing(ind, volume_integralFR=vi)
    solver = tlw.DGSEM(N, volume_integral=vi)
    semi = tlw.SemidiscretizationHyperbolic(mesh, tlw.get_time_discretization(solver), eq, ic, solver,
                                            boundary_conditions=walls(), surface_flux_order=flux_order,
                                            keep_aux=keep_aux)
    nx = mesh.nx
    okw = dict(shock_capturing=True, alpha_max=0.5, alpha_min=0.001, alpha_smooth=True) if sc else {}
    o = oracle_mod.Oracle(N, nx, nx, 0.0, 0.0, 1.0 / nx, periodic=(False, False),
                          flux_order=0 if flux_order == "tree" else 1, bc_kind=oracle_mod.BV_SLIP_WALL, **okw)
    x, y = o.node_coords()
    rng = np.random.default_rng(seed)
    u_lex = ic(x, y) * (1.0 + amp * rng.standard_normal(x.shape + (4,)))
    return semi, o, u_lex, mesh


@pytest.mark.parametrize("N", [3, 4])
@pytest.mark.parametrize("order", ["tree", "structured"])
def test_slip_wall_rhs_lw_parity(oracle_mod, N, order):
    import trixilw_hip as tlw

    torch = torch_mod()
    print(f"slip wall LW N={N} {order}", flush=True)
    semi, o, u_lex, mesh = setup_walls(oracle_mod, N, 3, flux_order=order)
    dt = 0.8 * o.max_dt(u_lex)
    du_ref = o.rhs(u_lex, 0.1, dt)
    u = to_dev(mesh, u_lex)
    du = torch.empty_like(u)
    semi.cache["dt"][0] = dt
    tlw.rhs(du, u, semi, 0.1)
    torch.cuda.synchronize()
    assert_close(to_lex(mesh, du), du_ref, what=f"slip wall LW N={N} {order}")
    semi.close()


@pytest.mark.parametrize("N", [3, 4])
def test_slip_wall_rhs_mdrk_parity(oracle_mod, N):
    import trixilw_hip as tlw

    torch = torch_mod()
    print(f"slip wall MDRK N={N}", flush=True)
    semi, o, u_lex, mesh = setup_walls(oracle_mod, N, 3, td="mdrk")
    dt = 0.4 * o.max_dt(u_lex)
    du1_ref = o.rhs(u_lex, 0.1, dt, kind="mdrk1")
    o.set_us(u_lex + 0.5 * dt * du1_ref)
    du2_ref = o.rhs(u_lex, 0.1, dt, kind="mdrk2")
    u = to_dev(mesh, u_lex)
    du = torch.empty_like(u)
    semi.cache["dt"][0] = dt
    tlw.rhs_mdrk1(du, u, semi, 0.1)
    torch.cuda.synchronize()
    assert_close(to_lex(mesh, du), du1_ref, what=f"slip wall mdrk1 N={N}")
    semi.set_us(torch.add(u, du, alpha=0.5 * dt))
    tlw.rhs_mdrk2(du, u, semi, 0.1)
    torch.cuda.synchronize()
    assert_close(to_lex(mesh, du), du2_ref, what=f"slip wall mdrk2 N={N}")
    semi.close()


@pytest.mark.parametrize("N", [3, 4])
def test_slip_wall_shock_capturing_rhs_parity(oracle_mod, N):
    import trixilw_hip as tlw

    torch = torch_mod()
    print(f"slip wall SC N={N}", flush=True)
    semi, o, u_lex, mesh = setup_walls(oracle_mod, N, 3, sc=True, keep_aux=True)
    x, y = o.node_coords()
    u_lex = u_lex * np.where((x - 0.4) ** 2 + (y - 0.55) ** 2 < 0.06, 1.8, 1.0)[..., None]
    dt = 0.4 * o.max_dt(u_lex)
    du_ref = o.rhs(u_lex, 0.0, dt)
    assert np.count_nonzero(o.aux("alpha")) > 4
    u = to_dev(mesh, u_lex)
    du = torch.empty_like(u)
    semi.cache["dt"][0] = dt
    tlw.rhs(du, u, semi, 0.0)
    torch.cuda.synchronize()
    assert_close(to_lex(mesh, du), du_ref, what=f"slip wall SC N={N}")
    semi.close()


@pytest.mark.parametrize("N,td,K", [(4, "lw", 30), (3, "lw", 30), (3, "mdrk", 20)])
def test_slip_wall_device_loop_k_steps(oracle_mod, N, td, K):
    import trixilw_hip as tlw

    print(f"slip wall device loop N={N} {td} K={K}", flush=True)
    semi, o, u_lex, mesh = setup_walls(oracle_mod, N, 4, td=td)
    cfl = tlw.trixi2lw(0.9, semi.solver) * (0.5 if td == "mdrk" else 1.0)
    u_ref, t_ref, dts = o.solve(u_lex, 0.0, 1e9, cfl, max_steps=K, mdrk=(td == "mdrk"))
    lw_update = tlw.semidiscretize(semi, semi.time_discretization, (0.0, float("inf")))
    lw_update.u0.copy_(to_dev(mesh, u_lex))
    sol = tlw.solve_lwfr(lw_update, (), 1e-6, None, time_step_computation=tlw.CFLBased(cfl), max_steps=K)
    assert sol.stats.naccept == K
    assert abs(sol.t - t_ref) <= 1e-13 * t_ref
    assert_close(to_lex(mesh, sol.u), u_ref, what=f"slip wall {td} N={N} after {K} steps")
    semi.close()


def test_slip_wall_rejected_on_curvilinear_mesh():
    import trixilw_hip as tlw
    from trixilw_hip import _lib

    from gpu_helpers import warp_mapping

    torch_mod()
    eq = tlw.CompressibleEulerEquations2D(1.4)
    mesh = tlw.StructuredMesh((4, 4), warp_mapping(20.0, 0.05), periodicity=False)
    solver = tlw.DGSEM(3, volume_integral=tlw.VolumeIntegralFR(tlw.LW()))
    with pytest.raises(_lib.TLWError) as ei:
        tlw.SemidiscretizationHyperbolic(mesh, tlw.LW(), eq, tlw.IsentropicVortex(), solver,
                                         boundary_conditions=walls())
    assert ei.value.code == _lib.TLW_ENOTSUP
    with pytest.raises(NotImplementedError):  # a wall kind per direction that the device cannot take
        tlw.SemidiscretizationHyperbolic(tlw.TreeMesh((0.0, 0.0), (1.0, 1.0), 2, periodicity=False), tlw.LW(), eq,
                                         tlw.IsentropicVortex(), solver,
                                         boundary_conditions=(tlw.boundary_condition_slip_wall_vertical,) * 4)
