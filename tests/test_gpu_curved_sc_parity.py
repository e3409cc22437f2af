"""GPU parity of first-order shock capturing on the curvilinear StructuredMesh
(VolumeIntegralFRShockCapturing with IndicatorHennemannGassner on
density_pressure) against the CPU oracle: blending factors, LW residuals and
both MDRK stages on warped meshes, the admissibility fallback at 1.5× the
stable dt, time-averaged Dirichlet boundaries, K-step device-loop states
within 1e-12 relative (L2 and L∞), conservation at scale, and the
combinations the path rejects.

The subcell FV fluxes use Trixi's accumulated subcell normals; a pure-DG
element's fn_low next to its + faces uses the reference's calc_fn_low! normal
(structured/dg_2d.jl:111-173), and compute_alp reads each face node's own
inverse Jacobian.  Parity runs the kernels on identical inputs (state, dt and
the oracle's metric terms).  Inputs are seeded."""
import numpy as np
import pytest

from gpu_helpers import GAMMA, assert_close, setup_curved, to_dev, to_lex, torch_mod, warp_mapping
from test_gpu_sc_parity import blocky_state

pytestmark = pytest.mark.gpu

L = 20.0


def with_jump(o, u_lex, f=1.6):
    """Density and pressure raised by the factor f inside a disk of radius L/5
    (all conserved variables scaled, the velocity unchanged): the indicator
    fires along the rim, the elements away from it stay pure DG."""
    x, y = o.node_coords()
    r = np.hypot(x - 0.6 * L, y - 0.45 * L)
    return u_lex * np.where(r < 0.2 * L, f, 1.0)[..., None]


def alpha_lex(semi, mesh):
    return semi.get_aux("alpha")[mesh.lexicographic_permutation()]


def check_alpha(semi, o, mesh):
    a_ref = o.aux("alpha")
    assert np.count_nonzero(a_ref) >= 4 and np.count_nonzero(a_ref == 0.0) >= 4  # blended and pure DG
    assert np.max(np.abs(alpha_lex(semi, mesh) - a_ref)) < 1e-13


@pytest.mark.parametrize("N", [3, 4])
@pytest.mark.parametrize("A", [0.05, 0.12])
def test_curved_sc_rhs_parity(oracle_mod, N, A):
    import trixilw_hip as tlw

    torch = torch_mod()
    print(f"curved SC rhs N={N} A={A}", flush=True)
    semi, o, u_lex, mesh = setup_curved(oracle_mod, N, 8, 8, A=A, sc=True, keep_aux=True)
    u_lex = with_jump(o, u_lex)
    dt = 0.4 * o.max_dt(u_lex)
    du_ref = o.rhs(u_lex, 0.0, dt)
    u = to_dev(mesh, u_lex)
    du = torch.empty_like(u)
    semi.cache["dt"][0] = dt
    tlw.rhs(du, u, semi, 0.0)
    torch.cuda.synchronize()
    check_alpha(semi, o, mesh)
    assert_close(to_lex(mesh, du), du_ref, what=f"curved SC LW N={N} A={A}")
    semi.close()


@pytest.mark.parametrize("N", [3, 4])
def test_curved_sc_rhs_mdrk_parity(oracle_mod, N):
    import trixilw_hip as tlw

    torch = torch_mod()
    print(f"curved SC mdrk N={N}", flush=True)
    semi, o, u_lex, mesh = setup_curved(oracle_mod, N, 8, 8, A=0.1, td="mdrk", sc=True, keep_aux=True)
    u_lex = with_jump(o, u_lex)
    dt = 0.3 * o.max_dt(u_lex)
    du1_ref = o.rhs(u_lex, 0.1, dt, kind="mdrk1")
    o.set_us(u_lex + 0.5 * dt * du1_ref)
    du2_ref = o.rhs(u_lex, 0.1, dt, kind="mdrk2")
    u = to_dev(mesh, u_lex)
    du = torch.empty_like(u)
    semi.cache["dt"][0] = dt
    tlw.rhs_mdrk1(du, u, semi, 0.1)
    torch.cuda.synchronize()
    check_alpha(semi, o, mesh)
    assert_close(to_lex(mesh, du), du1_ref, what=f"curved SC mdrk1 N={N}")
    semi.set_us(torch.add(u, du, alpha=0.5 * dt))
    tlw.rhs_mdrk2(du, u, semi, 0.1)
    torch.cuda.synchronize()
    assert_close(to_lex(mesh, du), du2_ref, what=f"curved SC mdrk2 N={N}")
    semi.close()


@pytest.mark.parametrize("N", [3, 4])
def test_curved_sc_blocky_state_with_fallback_parity(oracle_mod, N):
    """Element-wise random states on the warped mesh at 1.5× the stable dt:
    every interface is a discontinuity, the indicator blends most elements and
    compute_alp switches many interfaces to the Rusanov flux."""
    import trixilw_hip as tlw

    torch = torch_mod()
    print(f"curved SC blocky N={N}", flush=True)
    semi, o, _, mesh = setup_curved(oracle_mod, N, 8, 8, A=0.1, sc=True, keep_aux=True)
    u_lex = blocky_state(8, N + 1, seed=80 + N)
    dt = 1.5 * o.max_dt(u_lex)
    du_ref = o.rhs(u_lex, 0.0, dt)
    u = to_dev(mesh, u_lex)
    du = torch.empty_like(u)
    semi.cache["dt"][0] = dt
    tlw.rhs(du, u, semi, 0.0)
    torch.cuda.synchronize()
    assert np.max(np.abs(alpha_lex(semi, mesh) - o.aux("alpha"))) < 1e-13
    assert_close(to_lex(mesh, du), du_ref, what=f"curved SC blocky N={N}")
    semi.close()


@pytest.mark.parametrize("N,td", [(3, "lw"), (4, "lw"), (4, "mdrk")])
def test_curved_sc_dirichlet_rhs_parity(oracle_mod, N, td):
    """Non-periodic warped mesh with the vortex as time-averaged Dirichlet data
    and no alpha smoothing (the reference's smoothing on a StructuredMesh needs
    a periodic mesh): boundary faces take the Dirichlet flux without blending."""
    import trixilw_hip as tlw

    torch = torch_mod()
    print(f"curved SC dirichlet N={N} {td}", flush=True)
    semi, o, u_lex, mesh = setup_curved(oracle_mod, N, 8, 8, A=0.1, td=td, periodic=False, sc=True,
                                        alpha_smooth=False, keep_aux=True)
    u_lex = with_jump(o, u_lex)
    dt = 0.3 * o.max_dt(u_lex)
    u = to_dev(mesh, u_lex)
    du = torch.empty_like(u)
    semi.cache["dt"][0] = dt
    if td == "lw":
        du_ref = o.rhs(u_lex, 0.2, dt)
        tlw.rhs(du, u, semi, 0.2)
        torch.cuda.synchronize()
        check_alpha(semi, o, mesh)
        assert_close(to_lex(mesh, du), du_ref, what=f"curved SC Dirichlet LW N={N}")
    else:
        du1_ref = o.rhs(u_lex, 0.2, dt, kind="mdrk1")
        o.set_us(u_lex + 0.5 * dt * du1_ref)
        du2_ref = o.rhs(u_lex, 0.2, dt, kind="mdrk2")
        tlw.rhs_mdrk1(du, u, semi, 0.2)
        assert_close(to_lex(mesh, du), du1_ref, what=f"curved SC Dirichlet mdrk1 N={N}")
        semi.set_us(torch.add(u, du, alpha=0.5 * dt))
        tlw.rhs_mdrk2(du, u, semi, 0.2)
        torch.cuda.synchronize()
        assert_close(to_lex(mesh, du), du2_ref, what=f"curved SC Dirichlet mdrk2 N={N}")
    semi.close()


@pytest.mark.parametrize("N,nx,td,K", [(4, 8, "lw", 25), (3, 16, "lw", 25), (3, 16, "mdrk", 15)])
def test_curved_sc_device_loop_k_steps(oracle_mod, N, nx, td, K):
    """K steps of the device loop (indicator, blended volume kernel and the
    curvilinear CFL reduction) against the oracle's solve loop."""
    import trixilw_hip as tlw

    print(f"curved SC device loop N={N} {nx}x{nx} {td} K={K}", flush=True)
    semi, o, u_lex, mesh = setup_curved(oracle_mod, N, nx, nx, A=0.1, td=td, sc=True)
    u_lex = with_jump(o, u_lex)
    cfl = tlw.trixi2lw(0.9, semi.solver) * (0.5 if td == "mdrk" else 1.0)
    u_ref, t_ref, dts = o.solve(u_lex, 0.0, 1e9, cfl, max_steps=K, mdrk=(td == "mdrk"))
    lw_update = tlw.semidiscretize(semi, semi.time_discretization, (0.0, float("inf")))
    lw_update.u0.copy_(to_dev(mesh, u_lex))
    sol = tlw.solve_lwfr(lw_update, (), 1e-6, None, time_step_computation=tlw.CFLBased(cfl), max_steps=K)
    assert sol.stats.naccept == K
    assert abs(sol.t - t_ref) <= 1e-13 * t_ref
    assert_close(to_lex(mesh, sol.u), u_ref, what=f"curved SC {td} N={N} {nx}x{nx} after {K} steps")
    semi.close()


def test_curved_sc_conservation_at_scale():
    """256×256 warped mesh (1.6 M DOFs, the host's own metric terms) with a
    density/pressure jump: the indicator fires along the jump and 20 device-loop
    steps conserve Σ w_i w_j J u."""
    import trixilw_hip as tlw

    torch = torch_mod()
    eq = tlw.CompressibleEulerEquations2D(GAMMA)
    mesh = tlw.StructuredMesh((256, 256), warp_mapping(L, 0.08))
    ind = tlw.IndicatorHennemannGassner(eq, None, alpha_max=0.5, alpha_min=0.001, alpha_smooth=True)
    vi = tlw.VolumeIntegralFRShockCapturing(ind, volume_integralFR=tlw.VolumeIntegralFR(tlw.LW()))
    solver = tlw.DGSEM(4, volume_integral=vi)
    vortex = tlw.IsentropicVortex(center=(L / 2, L / 2), period=(L, L))

    def ic(x, y, t):
        r = np.hypot(x - 0.6 * L, y - 0.45 * L)
        return vortex(x, y, t) * np.where(r < 0.2 * L, 1.6, 1.0)[..., None]

    semi = tlw.SemidiscretizationHyperbolic(mesh, tlw.LW(), eq, ic, solver, keep_aux=True)
    lw_update = tlw.semidiscretize(semi, tlw.LW(), (0.0, 1e9))
    u0 = lw_update.u0.clone()
    sol = tlw.solve_lwfr(lw_update, (), 1e-6, None, time_step_computation=tlw.CFLBased(tlw.trixi2lw(0.9, solver)),
                         max_steps=20)
    assert sol.stats.naccept == 20
    alpha = semi.get_aux("alpha")
    assert np.count_nonzero(alpha > 0.01) > 100 and np.count_nonzero(alpha == 0.0) > 0.9 * alpha.size
    _, ijn = mesh.metrics(solver.basis)
    w = np.outer(solver.basis.weights, solver.basis.weights)[None] / ijn
    wt = torch.from_numpy(np.ascontiguousarray(w)).cuda()
    m0 = torch.einsum("eji,ejiv->v", wt, u0)
    m1 = torch.einsum("eji,ejiv->v", wt, sol.u)
    scale = torch.einsum("eji,ejiv->v", wt, u0.abs())
    assert torch.all((m1 - m0).abs() <= 1e-13 * scale), (m0, m1)
    assert torch.isfinite(sol.u).all()
    semi.close()


def test_curved_sc_rejects_unsupported(oracle_mod):
    """MUSCL-Hancock on a curvilinear mesh (the reference rejects it) and alpha
    smoothing on a non-periodic StructuredMesh (the reference asserts a
    periodic mesh) are TLW_ENOTSUP."""
    import trixilw_hip as tlw
    from trixilw_hip import _lib

    torch_mod()
    eq = tlw.CompressibleEulerEquations2D(GAMMA)
    ic = tlw.IsentropicVortex(center=(L / 2, L / 2), period=(1e9, 1e9))
    cases = [
        (tlw.StructuredMesh((4, 4), warp_mapping(L, 0.05)), tlw.MUSCLHancockReconstruction(), True, None),
        (tlw.StructuredMesh((4, 4), warp_mapping(L, 0.05), periodicity=False), tlw.FirstOrderReconstruction(), True,
         ic),
    ]
    for mesh, rec, smooth, bc in cases:
        ind = tlw.IndicatorHennemannGassner(eq, None, alpha_max=0.5, alpha_min=0.001, alpha_smooth=smooth)
        vi = tlw.VolumeIntegralFRShockCapturing(ind, volume_integralFR=tlw.VolumeIntegralFR(tlw.LW()),
                                                reconstruction=rec)
        solver = tlw.DGSEM(3, volume_integral=vi)
        with pytest.raises(_lib.TLWError) as ei:
            tlw.SemidiscretizationHyperbolic(mesh, tlw.LW(), eq, ic, solver, boundary_conditions=bc)
        assert ei.value.code == _lib.TLW_ENOTSUP
