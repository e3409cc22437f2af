"""GPU parity of the two conditions on uⁿ with a normal direction on the
curvilinear StructuredMesh, boundary_condition_outflow and slip_wall_approximate
(src/equations/equations.jl:185-197, 251-260, in the form of their P4est call),
against the oracle: rhs and both MDRK stages (plain and shock-capturing
kernels), K-step device-loop states within 1e-12 relative, the approximate
wall's exact conservation of mass and energy at 128×128 over 100 device-loop
steps, and the meshes the conditions are refused on.  Inputs are seeded; the
warped non-periodic mesh uses the oracle's metric terms on both sides."""
import numpy as np
import pytest

from gpu_helpers import GAMMA, assert_close, setup_curved, to_dev, to_lex, torch_mod

pytestmark = pytest.mark.gpu

KINDS = ["outflow", "slip_wall_approximate"]


def condition(which):
    import trixilw_hip as tlw

    return tlw.boundary_condition_outflow if which == "outflow" else tlw.slip_wall_approximate


def with_jump(o, u_lex, L=20.0, f=1.6):
    """Density and pressure raised by f inside a disk that touches no boundary."""
    x, y = o.node_coords()
    r = np.hypot(x - 0.6 * L, y - 0.45 * L)
    return u_lex * np.where(r < 0.2 * L, f, 1.0)[..., None]


@pytest.mark.parametrize("which", KINDS)
@pytest.mark.parametrize("N,td,sc", [(3, "lw", False), (4, "lw", False), (4, "mdrk", False), (3, "mdrk", False),
                                     (4, "lw", True), (3, "mdrk", True)])
def test_curved_normal_condition_rhs_parity(oracle_mod, which, N, td, sc):
    import trixilw_hip as tlw

    torch = torch_mod()
    print(f"curved {which} N={N} {td} sc={sc}", flush=True)
    kw = dict(sc=True, alpha_smooth=False) if sc else {}
    semi, o, u_lex, mesh = setup_curved(oracle_mod, N, 8, 6, A=0.1, td=td, seed=11, periodic=False,
                                        boundary=condition(which), **kw)
    if sc:
        u_lex = with_jump(o, u_lex)
    u = to_dev(mesh, u_lex)
    du = torch.empty_like(u)
    dt = 0.4 * o.max_dt(u_lex)
    semi.cache["dt"][0] = dt
    if td == "lw":
        du_ref = o.rhs(u_lex, 0.25, dt)
        tlw.rhs(du, u, semi, 0.25)
        torch.cuda.synchronize()
        assert_close(to_lex(mesh, du), du_ref, what=f"curved {which} LW N={N} sc={sc}")
    else:
        d1 = o.rhs(u_lex, 0.25, dt, kind="mdrk1")
        us = u_lex + 0.5 * dt * d1
        o.set_us(us)
        d2 = o.rhs(u_lex, 0.25, dt, kind="mdrk2")
        tlw.rhs_mdrk1(du, u, semi, 0.25)
        assert_close(to_lex(mesh, du), d1, what=f"curved {which} mdrk1 N={N} sc={sc}")
        semi.set_us(to_dev(mesh, us))
        tlw.rhs_mdrk2(du, u, semi, 0.25)
        torch.cuda.synchronize()
        assert_close(to_lex(mesh, du), d2, what=f"curved {which} mdrk2 N={N} sc={sc}")
    semi.close()


@pytest.mark.parametrize("which", KINDS)
@pytest.mark.parametrize("N,nx,td,K", [(4, 8, "lw", 25), (3, 16, "mdrk", 20)])
def test_curved_normal_condition_device_loop_k_steps(oracle_mod, which, N, nx, td, K):
    import trixilw_hip as tlw

    print(f"curved {which} device loop N={N} {nx}x{nx} {td} K={K}", flush=True)
    semi, o, u_lex, mesh = setup_curved(oracle_mod, N, nx, nx, A=0.1, td=td, periodic=False,
                                        boundary=condition(which))
    cfl = tlw.trixi2lw(0.9, semi.solver) * (0.5 if td == "mdrk" else 1.0)
    u_ref, t_ref, _ = o.solve(u_lex, 0.0, 1e9, cfl, max_steps=K, mdrk=(td == "mdrk"))
    lw_update = tlw.semidiscretize(semi, semi.time_discretization, (0.0, float("inf")))
    lw_update.u0.copy_(to_dev(mesh, u_lex))
    sol = tlw.solve_lwfr(lw_update, (), 1e-6, None, time_step_computation=tlw.CFLBased(cfl), max_steps=K)
    assert sol.stats.naccept == K
    assert abs(sol.t - t_ref) <= 1e-13 * t_ref
    assert_close(to_lex(mesh, sol.u), u_ref, what=f"curved {which} {td} N={N} after {K} steps")
    semi.close()


def test_approximate_wall_conserves_mass_and_energy_at_scale(oracle_mod):
    """128×128 warped mesh closed by the approximate slip wall, 100 LW steps of
    the device loop: the wall's mass and energy fluxes vanish, so both totals
    (∑ w_i w_j u / J⁻¹) stay at their initial values to round-off, while the
    momentum changes."""
    import trixilw_hip as tlw

    torch = torch_mod()
    print("curved approximate wall 128x128 100 steps", flush=True)
    semi, o, u_lex, mesh = setup_curved(oracle_mod, 4, 128, 128, A=0.1, periodic=False,
                                        boundary=tlw.slip_wall_approximate)
    w1 = o.basis()[1]
    wq = np.outer(w1, w1)[None] / o.aux("inverse_jacobian").reshape(u_lex.shape[:-1])
    cfl = tlw.trixi2lw(0.9, semi.solver)
    lw_update = tlw.semidiscretize(semi, semi.time_discretization, (0.0, float("inf")))
    lw_update.u0.copy_(to_dev(mesh, u_lex))
    sol = tlw.solve_lwfr(lw_update, (), 1e-6, None, time_step_computation=tlw.CFLBased(cfl), max_steps=100)
    torch.cuda.synchronize()
    assert sol.stats.naccept == 100
    u1 = to_lex(mesh, sol.u)
    m0 = np.einsum("eji,ejiv->v", wq, u_lex)
    m1 = np.einsum("eji,ejiv->v", wq, u1)
    scale = np.einsum("eji,ejiv->v", wq, np.abs(u_lex))
    assert abs(m1[0] - m0[0]) <= 1e-12 * scale[0], (m1[0] - m0[0]) / scale[0]
    assert abs(m1[3] - m0[3]) <= 1e-12 * scale[3], (m1[3] - m0[3]) / scale[3]
    assert np.max(np.abs(m1[1:3] - m0[1:3])) > 1e-6 * scale[0]
    semi.close()


def test_normal_conditions_refused_on_a_treemesh():
    import trixilw_hip as tlw
    from trixilw_hip import _lib

    torch_mod()
    eq = tlw.CompressibleEulerEquations2D(GAMMA)
    ic = tlw.IsentropicVortex()
    for which in KINDS:
        mesh = tlw.TreeMesh((-10.0, -10.0), (10.0, 10.0), 3, periodicity=False)
        solver = tlw.DGSEM(3, volume_integral=tlw.VolumeIntegralFR(tlw.LW()))
        with pytest.raises(_lib.TLWError) as ei:
            tlw.SemidiscretizationHyperbolic(mesh, tlw.LW(), eq, ic, solver, boundary_conditions=condition(which))
        assert ei.value.code == _lib.TLW_ENOTSUP
