"""GPU adjoint of the LW step (tlw_adjoint_lw) against the oracle's
dual-number tangent and central differences.

One step: <λ, v> = <λ', du'> with du' the tangent of u' = u + dt rhs!(u) along
v, ≤ 1e-12 relative.  A checkpointed 20-step gradient of the kinetic energy:
against the oracle's tangent (≤ 1e-10) and central differences with h = 1e-6
on the device (≤ 1e-6), dts frozen from the forward run."""
import numpy as np
import pytest

from gpu_helpers import setup, to_dev, to_lex, torch_mod

pytestmark = pytest.mark.gpu


def ke_and_grad_np(u, w, h):
    scale = np.outer(w, w)[None] * (h / 2) ** 2
    rho, m1, m2 = u[..., 0], u[..., 1], u[..., 2]
    ke = 0.5 * (m1 ** 2 + m2 ** 2) / rho
    g = np.zeros_like(u)
    g[..., 0] = scale * (-ke / rho)
    g[..., 1] = scale * m1 / rho
    g[..., 2] = scale * m2 / rho
    return float((scale * ke).sum()), g


@pytest.mark.parametrize("N", [3, 4])
def test_adjoint_step_matches_tangent(oracle_mod, N):
    import trixilw_hip as tlw

    torch = torch_mod()
    semi, o, u_lex, mesh = setup(oracle_mod, N, 3, ordering="morton", seed=50 + N, amp=1e-2)
    dt = 0.5 * o.max_dt(u_lex)
    rng = np.random.default_rng(60 + N)
    v = rng.standard_normal(u_lex.shape)
    lam_next = rng.standard_normal(u_lex.shape)
    _, du1 = o.tangent_lw(u_lex, v, [dt])
    ref = float(np.sum(lam_next * du1))
    u = to_dev(mesh, u_lex)
    lam_n = to_dev(mesh, lam_next)
    lam = torch.empty_like(u)
    tlw.adjoint_step(lam, u, lam_n, semi, 0.0, dt)
    torch.cuda.synchronize()
    got = float(np.sum(to_lex(mesh, lam) * v))
    assert abs(got - ref) <= 1e-12 * abs(ref), (got, ref)
    # the step's own Jacobian part is not the identity alone
    assert abs(float(np.sum(lam_next * v)) - ref) > 1e-3 * abs(ref)


def test_checkpointed_gradient_matches_tangent_and_fd(oracle_mod):
    """d KE(u_K) / d u0 along a density perturbation, 16² vortex, N=4, 20 steps,
    checkpoints every 5."""
    import trixilw_hip as tlw

    torch = torch_mod()
    ic = tlw.IsentropicVortex()
    semi, o, u_lex, mesh = setup(oracle_mod, 4, 4, ic=ic, amp=0.0, keep_aux=False)
    rng = np.random.default_rng(3)
    v = np.zeros_like(u_lex)
    v[..., 0] = 1e-2 * rng.standard_normal(u_lex.shape[:3])
    u0 = to_dev(mesh, u_lex)
    cfl = tlw.trixi2lw(0.98, semi.solver)
    J, grad, uK, dts = tlw.gradient_lw(semi, u0, 20, cfl, checkpoint_every=5)
    torch.cuda.synchronize()
    g = to_lex(mesh, grad)
    dJ_adj = float(np.sum(g * v))
    # forward mode through the oracle with the same dts
    _, w, _, _ = o.basis()
    h_el = 20.0 / mesh.nx
    uK_o, duK = o.tangent_lw(u_lex, v, dts)
    J_o, gJ = ke_and_grad_np(uK_o, w, h_el)
    dJ_tan = float(np.sum(gJ * duK))
    assert abs(J - J_o) <= 1e-12 * abs(J_o)
    assert abs(dJ_adj - dJ_tan) <= 1e-10 * abs(dJ_tan), (dJ_adj, dJ_tan)
    # central differences on the device with the frozen dts
    h = 1e-6
    Jp = tlw.objective_lw(semi, to_dev(mesh, u_lex + h * v), 20, dts)
    Jm = tlw.objective_lw(semi, to_dev(mesh, u_lex - h * v), 20, dts)
    fd = (Jp - Jm) / (2 * h)
    assert abs(dJ_adj - fd) <= 1e-6 * abs(fd), (dJ_adj, fd)


def test_c3_gradient_64x64_200_steps(oracle_mod):
    """The differentiation configuration as stated: the 64² periodic vortex
    (strength 5, Mach 0.5) on [-10,10]², N=4, 200 LW steps at 0.98 of the LW
    CFL limit, checkpoints every 20 steps, an initial density perturbation by
    a seeded Gaussian random field of amplitude 1e-2; d KE(u_200) along a
    second such field: adjoint against the oracle's dual-number tangent
    (≤ 1e-10) and device central differences with h = 1e-6 (≤ 1e-6).  The
    forward run's dts (recorded on the device) match the oracle's CFL dts,
    and J matches the oracle's, ≤ 1e-12."""
    import trixilw_hip as tlw

    torch = torch_mod()
    ic = tlw.IsentropicVortex()
    semi, o, u_lex, mesh = setup(oracle_mod, 4, 6, ic=ic, amp=0.0, keep_aux=False, ordering="strips")
    x, y = o.node_coords()
    u_lex = u_lex.copy()
    u_lex[..., 0] += tlw.gaussian_random_field(x, y, (20.0, 20.0), seed=2025, amplitude=1e-2)
    v = np.zeros_like(u_lex)
    v[..., 0] = tlw.gaussian_random_field(x, y, (20.0, 20.0), seed=2026, amplitude=1.0)
    cfl = tlw.trixi2lw(0.98, semi.solver)
    nsteps = 200
    J, grad, uK, dts = tlw.gradient_lw(semi, to_dev(mesh, u_lex), nsteps, cfl, checkpoint_every=20)
    torch.cuda.synchronize()
    # forward parity: the device loop's CFL dts and final state against the oracle's solve
    u_ref, _, dts_ref = o.solve(u_lex, 0.0, 1e9, cfl, max_steps=nsteps)
    assert len(dts_ref) == nsteps
    assert np.max(np.abs(np.asarray(dts) - np.asarray(dts_ref)) / np.asarray(dts_ref)) <= 1e-12
    uK_lex = to_lex(mesh, uK)
    assert np.max(np.abs(uK_lex - u_ref)) <= 1e-12 * np.max(np.abs(u_ref))
    _, w, _, _ = o.basis()
    J_o, gJ = ke_and_grad_np(u_ref, w, 20.0 / mesh.nx)
    assert abs(J - J_o) <= 1e-12 * abs(J_o)
    # adjoint against the dual-number tangent through the oracle with the same dts
    dJ_adj = float(np.sum(to_lex(mesh, grad) * v))
    uK_t, duK = o.tangent_lw(u_lex, v, dts)
    J_t, gJ_t = ke_and_grad_np(uK_t, w, 20.0 / mesh.nx)
    dJ_tan = float(np.sum(gJ_t * duK))
    assert abs(dJ_adj - dJ_tan) <= 1e-10 * abs(dJ_tan), (dJ_adj, dJ_tan)
    # central differences of the device loop with the frozen dts
    h = 1e-6
    Jp = tlw.objective_lw(semi, to_dev(mesh, u_lex + h * v), nsteps, dts)
    Jm = tlw.objective_lw(semi, to_dev(mesh, u_lex - h * v), nsteps, dts)
    fd = (Jp - Jm) / (2 * h)
    assert abs(dJ_adj - fd) <= 1e-6 * abs(fd), (dJ_adj, fd)
    semi.close()


@pytest.mark.parametrize("N", [3, 4])
def test_adjoint_mdrk_step_matches_tangent(oracle_mod, N):
    """One MDRK step (both stages): <λ, v> = <λ', du'> with du' the oracle's
    dual-number tangent of u' = u + dt R2(u, u + dt R1(u)) along v, ≤ 1e-12."""
    import trixilw_hip as tlw

    torch = torch_mod()
    semi, o, u_lex, mesh = setup(oracle_mod, N, 3, td="mdrk", ordering="strips", seed=70 + N, amp=1e-2)
    dt = 0.4 * o.max_dt(u_lex)
    rng = np.random.default_rng(80 + N)
    v = rng.standard_normal(u_lex.shape)
    lam_next = rng.standard_normal(u_lex.shape)
    _, du1 = o.tangent(u_lex, v, [dt], mdrk=True)
    ref = float(np.sum(lam_next * du1))
    u = to_dev(mesh, u_lex)
    du = torch.empty_like(u)
    us = tlw.adjoint.mdrk_stage_state(u, semi, 0.0, dt, du)
    lam = torch.empty_like(u)
    tlw.adjoint_step_mdrk(lam, u, us, to_dev(mesh, lam_next), semi, 0.0, dt)
    torch.cuda.synchronize()
    got = float(np.sum(to_lex(mesh, lam) * v))
    assert abs(got - ref) <= 1e-12 * abs(ref), (got, ref)
    assert abs(float(np.sum(lam_next * v)) - ref) > 1e-3 * abs(ref)
    semi.close()


def test_checkpointed_gradient_mdrk_matches_tangent_and_fd(oracle_mod):
    """d KE(u_K) / d u0 through 20 MDRK steps (N=3, 16² vortex, checkpoints
    every 5): the adjoint against the oracle's MDRK tangent (≤ 1e-10) and
    device central differences with h = 1e-6 (≤ 1e-6), dts frozen."""
    import trixilw_hip as tlw

    torch = torch_mod()
    ic = tlw.IsentropicVortex()
    semi, o, u_lex, mesh = setup(oracle_mod, 3, 4, td="mdrk", ic=ic, amp=0.0, keep_aux=False)
    x, y = o.node_coords()
    u_lex = u_lex.copy()
    u_lex[..., 0] += tlw.gaussian_random_field(x, y, (20.0, 20.0), seed=7, amplitude=1e-2)
    v = np.zeros_like(u_lex)
    v[..., 0] = tlw.gaussian_random_field(x, y, (20.0, 20.0), seed=8, amplitude=1.0)
    cfl = 0.5 * tlw.trixi2lw(0.98, semi.solver)
    J, grad, uK, dts = tlw.gradient_lw(semi, to_dev(mesh, u_lex), 20, cfl, checkpoint_every=5)
    torch.cuda.synchronize()
    u_ref, _, dts_ref = o.solve(u_lex, 0.0, 1e9, cfl, max_steps=20, mdrk=True)
    assert np.max(np.abs(np.asarray(dts) - dts_ref) / dts_ref) <= 1e-12
    dJ_adj = float(np.sum(to_lex(mesh, grad) * v))
    _, w, _, _ = o.basis()
    uK_t, duK = o.tangent(u_lex, v, dts, mdrk=True)
    J_t, gJ = ke_and_grad_np(uK_t, w, 20.0 / mesh.nx)
    dJ_tan = float(np.sum(gJ * duK))
    assert abs(J - J_t) <= 1e-12 * abs(J_t)
    assert abs(dJ_adj - dJ_tan) <= 1e-10 * abs(dJ_tan), (dJ_adj, dJ_tan)
    h = 1e-6
    Jp = tlw.objective_lw(semi, to_dev(mesh, u_lex + h * v), 20, dts)
    Jm = tlw.objective_lw(semi, to_dev(mesh, u_lex - h * v), 20, dts)
    fd = (Jp - Jm) / (2 * h)
    assert abs(dJ_adj - fd) <= 1e-6 * abs(fd), (dJ_adj, fd)
    semi.close()


def _blended_setup(oracle_mod, N, seed):
    """A periodic [0,1]² 8x8 mesh with a square density/pressure jump (node noise
    breaks mirror symmetry, where the Rusanov speed's max has ties), shock
    capturing with first-order subcell FV on both sides."""
    import trixilw_hip as tlw

    eq = tlw.CompressibleEulerEquations2D(1.4)
    mesh = tlw.TreeMesh((0.0, 0.0), (1.0, 1.0), 3, ordering="strips")
    ind = tlw.IndicatorHennemannGassner(eq, None, alpha_max=0.5, alpha_min=0.001, alpha_smooth=True)
    solver = tlw.DGSEM(N, volume_integral=tlw.VolumeIntegralFRShockCapturing(ind))

    def square(x, y, t=0.0):
        inside = (np.abs(x - 0.5) < 0.22) & (np.abs(y - 0.45) < 0.2)
        rho = np.where(inside, 1.0, 0.4)
        p = np.where(inside, 1.0, 0.3)
        v1, v2 = 0.2 + 0 * x, -0.1 + 0 * x
        return np.stack([rho, rho * v1, rho * v2, p / 0.4 + 0.5 * rho * (v1 ** 2 + v2 ** 2)], -1)

    semi = tlw.SemidiscretizationHyperbolic(mesh, tlw.LW(), eq, square, solver, keep_aux=True)
    o = oracle_mod.Oracle(N, 8, 8, 0.0, 0.0, 1.0 / 8, shock_capturing=True, alpha_max=0.5, alpha_min=0.001,
                          alpha_smooth=True)
    x, y = o.node_coords()
    rng = np.random.default_rng(seed)
    u_lex = square(x, y) * (1.0 + 1e-3 * rng.standard_normal(x.shape + (4,)))
    return semi, o, u_lex, mesh


@pytest.mark.parametrize("N", [3, 4])
def test_adjoint_blended_step_matches_tangent(oracle_mod, N):
    """One shock-capturing LW step (first-order subcell FV, blended interface
    flux), α and the admissibility decision frozen at the state: <λ, v> =
    <λ', du'> against the oracle's dual tangent, whose indicator works on the
    value parts (so α is frozen the same way), ≤ 1e-12; the blend is active."""
    import trixilw_hip as tlw

    torch = torch_mod()
    semi, o, u_lex, mesh = _blended_setup(oracle_mod, N, 90 + N)
    dt = 0.3 * o.max_dt(u_lex)
    u = to_dev(mesh, u_lex)
    du = torch.empty_like(u)
    semi.cache["dt"][0] = dt
    tlw.rhs(du, u, semi, 0.0)
    alpha = semi.get_aux("alpha")
    assert (alpha > 0.05).sum() >= 4 and (alpha < 1e-12).sum() >= 4, alpha
    rng = np.random.default_rng(100 + N)
    v = rng.standard_normal(u_lex.shape)
    lam_next = rng.standard_normal(u_lex.shape)
    _, du1 = o.tangent(u_lex, v, [dt])
    ref = float(np.sum(lam_next * du1))
    lam = torch.empty_like(u)
    tlw.adjoint_step(lam, u, to_dev(mesh, lam_next), semi, 0.0, dt)
    torch.cuda.synchronize()
    got = float(np.sum(to_lex(mesh, lam) * v))
    assert abs(got - ref) <= 1e-12 * abs(ref), (got, ref)
    # the blended terms matter: the frozen-α tangent differs from that of the unblended scheme
    o_plain = oracle_mod.Oracle(N, 8, 8, 0.0, 0.0, 1.0 / 8)
    _, du_plain = o_plain.tangent(u_lex, v, [dt])
    assert abs(float(np.sum(lam_next * du_plain)) - ref) > 1e-6 * abs(ref)
    semi.close()


def test_checkpointed_gradient_blended_matches_tangent(oracle_mod):
    """12 shock-capturing LW steps (N=4, checkpoints every 4) of the blocky
    state: the adjoint gradient of the kinetic energy against the oracle's
    dual tangent with α frozen per step, ≤ 1e-10."""
    import trixilw_hip as tlw

    torch = torch_mod()
    semi, o, u_lex, mesh = _blended_setup(oracle_mod, 4, 5)
    rng = np.random.default_rng(6)
    v = rng.standard_normal(u_lex.shape) * 1e-2
    cfl = 0.3
    J, grad, uK, dts = tlw.gradient_lw(semi, to_dev(mesh, u_lex), 12, cfl, checkpoint_every=4)
    torch.cuda.synchronize()
    dJ_adj = float(np.sum(to_lex(mesh, grad) * v))
    _, w, _, _ = o.basis()
    uK_t, duK = o.tangent(u_lex, v, dts)
    J_t, gJ = ke_and_grad_np(uK_t, w, 1.0 / 8)
    dJ_tan = float(np.sum(gJ * duK))
    assert abs(J - J_t) <= 1e-12 * abs(J_t)
    assert abs(dJ_adj - dJ_tan) <= 1e-10 * abs(dJ_tan), (dJ_adj, dJ_tan)
    semi.close()


def test_adjoint_rejects_unsupported(oracle_mod):
    import trixilw_hip as tlw

    torch = torch_mod()
    ic = tlw.FourQuadrantRiemann()
    semi, o, u_lex, mesh = setup(oracle_mod, 4, 2, periodic=False, domain=(0.0, 1.0), bc=ic, ic=ic, amp=0.0)
    u = to_dev(mesh, u_lex)
    lam = torch.zeros_like(u)
    with pytest.raises(tlw._lib.TLWError):
        tlw.adjoint_step(torch.empty_like(u), u, lam, semi, 0.0, 1e-3)
