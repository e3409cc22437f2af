"""GPU parity of the shock-capturing path with MUSCLHancockReconstruction
(minmod slopes with β = 2 - α, positivity-scaled face values, the half-step
and Rusanov fluxes between evolved face values; first order on elements at
the domain boundary) against the CPU oracle: residuals of LW and both MDRK
stages, the compute_alp fallback on blocky states, K-step device loops and
conservation at 256², within 1e-12 relative (L2 and L∞).  Inputs are seeded."""
import numpy as np
import pytest

from gpu_helpers import assert_close, setup, to_dev, to_lex, torch_mod

pytestmark = pytest.mark.gpu

MH = "muscl_hancock"


def riemann_setup(oracle_mod, N, level, td="lw", keep_aux=True, ordering="lexicographic"):
    import trixilw_hip as tlw

    ic = tlw.FourQuadrantRiemann()
    return setup(oracle_mod, N, level, td=td, periodic=False, domain=(0.0, 1.0), bc=ic, ic=ic, amp=0.0, sc=True,
                 keep_aux=keep_aux, recon=MH, ordering=ordering)


def blocky_state(nx, n, seed):
    rng = np.random.default_rng(seed)
    E = nx * nx

    def field(lo, hi, rel):
        return rng.uniform(lo, hi, E)[:, None, None] * (1 + rel * rng.standard_normal((E, n, n)))

    rho, p = field(0.2, 1.0, 0.05), field(0.1, 1.0, 0.05)
    v1 = rng.uniform(-1, 1, E)[:, None, None] + 0.05 * rng.standard_normal((E, n, n))
    v2 = rng.uniform(-1, 1, E)[:, None, None] + 0.05 * rng.standard_normal((E, n, n))
    return np.stack([rho, rho * v1, rho * v2, p / 0.4 + 0.5 * rho * (v1 ** 2 + v2 ** 2)], -1)


def alpha_lex(semi, mesh):
    return semi.get_aux("alpha")[mesh.lexicographic_permutation()]


@pytest.mark.parametrize("N", [3, 4])
def test_mh_rhs_lw_riemann_parity(oracle_mod, N):
    import trixilw_hip as tlw

    torch = torch_mod()
    print(f"MH LW Riemann N={N}", flush=True)
    semi, o, u_lex, mesh = riemann_setup(oracle_mod, N, 4)
    dt = 0.3 * o.max_dt(u_lex)
    du_ref = o.rhs(u_lex, 0.0, dt)
    u = to_dev(mesh, u_lex)
    du = torch.empty_like(u)
    semi.cache["dt"][0] = dt
    tlw.rhs(du, u, semi, 0.0)
    torch.cuda.synchronize()
    assert np.max(np.abs(alpha_lex(semi, mesh) - o.aux("alpha"))) < 1e-13
    assert_close(to_lex(mesh, du), du_ref, what=f"MH LW N={N}")
    # the reconstruction acts: first-order FV gives another residual
    fo = type(o)(N, mesh.nx, mesh.nx, 0.0, 0.0, 1.0 / mesh.nx, periodic=(False, False), bc_kind=o._desc.bc_kind,
                 bc_par=list(o._desc.bc_par), shock_capturing=True, alpha_max=0.5)
    assert np.max(np.abs(fo.rhs(u_lex, 0.0, dt) - du_ref)) > 1e-6
    semi.close()


@pytest.mark.parametrize("N", [3, 4])
def test_mh_rhs_mdrk_riemann_parity(oracle_mod, N):
    import trixilw_hip as tlw

    torch = torch_mod()
    semi, o, u_lex, mesh = riemann_setup(oracle_mod, N, 4, td="mdrk")
    dt = 0.2 * o.max_dt(u_lex)
    du1_ref = o.rhs(u_lex, 0.1, dt, kind="mdrk1")
    o.set_us(u_lex + dt * du1_ref)
    du2_ref = o.rhs(u_lex, 0.1, dt, kind="mdrk2")
    u = to_dev(mesh, u_lex)
    du = torch.empty_like(u)
    semi.cache["dt"][0] = dt
    tlw.rhs_mdrk1(du, u, semi, 0.1)
    assert_close(to_lex(mesh, du), du1_ref, what=f"MH mdrk1 N={N}")
    semi.set_us(torch.add(u, du, alpha=dt))
    tlw.rhs_mdrk2(du, u, semi, 0.1)
    torch.cuda.synchronize()
    assert_close(to_lex(mesh, du), du2_ref, what=f"MH mdrk2 N={N}")
    semi.close()


@pytest.mark.parametrize("N", [3, 4])
def test_mh_blocky_state_with_fallback_parity(oracle_mod, N):
    """Periodic element-wise random states (every element reconstructed with
    MUSCL-Hancock where blended) at 1.2× the stable dt, so the compute_alp
    fallback switches interfaces to the Rusanov flux; every tile.  (At 1.5×
    the half-step leaves the admissible set at polydeg 4 and the subcell
    Rusanov speed is the square root of a negative number: NaN here, a
    DomainError in the reference.)"""
    import trixilw_hip as tlw

    torch = torch_mod()
    semi, o, _, mesh = setup(oracle_mod, N, 3, domain=(0.0, 1.0), sc=True, recon=MH)
    u_lex = blocky_state(8, N + 1, seed=50 + N)
    dt = 1.2 * o.max_dt(u_lex)
    du_ref = o.rhs(u_lex, 0.0, dt)
    assert np.isfinite(du_ref).all()
    assert np.count_nonzero(o.aux("alpha")) > 20
    u = to_dev(mesh, u_lex)
    du = torch.empty_like(u)
    semi.cache["dt"][0] = dt
    for eb in ((5, 10, 15) if N == 4 else (4, 8, 16)):
        semi.elements_per_workgroup(eb)
        du.fill_(np.nan)
        tlw.rhs(du, u, semi, 0.0)
        torch.cuda.synchronize()
        assert np.max(np.abs(alpha_lex(semi, mesh) - o.aux("alpha"))) < 1e-13
        assert_close(to_lex(mesh, du), du_ref, what=f"MH blocky N={N} tile {eb}")
    semi.close()


@pytest.mark.parametrize("ordering", ["lexicographic", "strips"])
@pytest.mark.parametrize("N,td,K", [(4, "lw", 30), (3, "lw", 30), (3, "mdrk", 20)])
def test_mh_device_loop_k_steps(oracle_mod, N, td, K, ordering):
    """Four-quadrant Riemann problem on a 32×32 mesh: K steps of the device
    loop against the oracle's solve loop (CFL 0.3·… as the bench runs it)."""
    import trixilw_hip as tlw

    print(f"MH device loop N={N} {td} K={K}", flush=True)
    semi, o, u_lex, mesh = riemann_setup(oracle_mod, N, 5, td=td, keep_aux=False, ordering=ordering)
    cfl = 0.3 * (0.5 if td == "mdrk" else 1.0)
    u_ref, t_ref, dts = o.solve(u_lex, 0.0, 1e9, cfl, max_steps=K, mdrk=(td == "mdrk"))
    lw_update = tlw.semidiscretize(semi, semi.time_discretization, (0.0, float("inf")))
    lw_update.u0.copy_(to_dev(mesh, u_lex))
    sol = tlw.solve_lwfr(lw_update, (), 1e-6, None, time_step_computation=tlw.CFLBased(cfl), max_steps=K)
    assert sol.stats.naccept == K
    assert abs(sol.t - t_ref) <= 1e-13 * t_ref
    assert_close(to_lex(mesh, sol.u), u_ref, what=f"MH {td} N={N} after {K} steps")
    semi.close()


def test_mh_conservation_at_scale():
    """256×256 periodic mesh with a square density/pressure jump: total mass,
    momentum and energy stay conserved over 20 device-loop steps with
    MUSCL-Hancock blending, and the state stays admissible."""
    import trixilw_hip as tlw

    torch = torch_mod()
    eq = tlw.CompressibleEulerEquations2D(1.4)
    mesh = tlw.TreeMesh((0, 0), (1, 1), 8)
    ind = tlw.IndicatorHennemannGassner(eq, None, alpha_max=0.5)
    solver = tlw.DGSEM(4, volume_integral=tlw.VolumeIntegralFRShockCapturing(
        ind, reconstruction=tlw.MUSCLHancockReconstruction()))

    def square(x, y, t):
        inside = (np.abs(x - 0.5) < 0.2) & (np.abs(y - 0.5) < 0.2)
        rho = np.where(inside, 1.0, 0.125)
        p = np.where(inside, 1.0, 0.1)
        v1 = 0.1 + 0 * x
        v2 = -0.05 + 0 * x
        return np.stack([rho, rho * v1, rho * v2, p / 0.4 + 0.5 * rho * (v1 ** 2 + v2 ** 2)], -1)

    semi = tlw.SemidiscretizationHyperbolic(mesh, tlw.LW(), eq, square, solver)
    lw_update = tlw.semidiscretize(semi, tlw.LW(), (0.0, 1e9))
    u0 = lw_update.u0.clone()
    sol = tlw.solve_lwfr(lw_update, (), 1e-6, None, time_step_computation=tlw.CFLBased(0.3), max_steps=20)
    w = torch.from_numpy(np.outer(solver.basis.weights, solver.basis.weights)).cuda()
    m0 = torch.einsum("ji,ejiv->v", w, u0)
    m1 = torch.einsum("ji,ejiv->v", w, sol.u)
    scale = torch.einsum("ji,ejiv->v", w, u0.abs())
    assert torch.all((m1 - m0).abs() <= 1e-13 * scale), (m0, m1)
    assert torch.isfinite(sol.u).all() and (sol.u[..., 0] > 0).all()
    assert (semi.get_aux("alpha") > 0).sum() > 100
    semi.close()
