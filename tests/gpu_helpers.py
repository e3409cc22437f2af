"""Shared set-up of the GPU parity tests: one device semidiscretization and
the oracle on the same mesh, state and dt, plus element-order conversion."""
import numpy as np

from conftest import rel_l2, rel_linf

TOL = 1e-12


def torch_mod():
    import torch

    assert torch.cuda.is_available(), "GPU tests need a GPU"
    return torch


def setup(oracle_mod, N, level, td="lw", periodic=True, domain=(-10.0, 10.0), bc=None, flux_order="tree",
          ordering="lexicographic", keep_aux=True, seed=0, amp=1e-3, ic=None, sc=False, alpha_max=0.5,
          alpha_smooth=True, recon="first_order"):
    """``sc`` selects VolumeIntegralFRShockCapturing (Hennemann-Gassner on
    density_pressure) on both sides, with first-order subcell FV or, with
    ``recon="muscl_hancock"``, the MUSCL-Hancock reconstruction."""
    import trixilw_hip as tlw

    eq = tlw.CompressibleEulerEquations2D(1.4)
    ic = ic or tlw.IsentropicVortex(center=(0.7, -0.4), period=(domain[1] - domain[0],) * 2)
    mesh = tlw.TreeMesh((domain[0],) * 2, (domain[1],) * 2, level, periodicity=periodic, ordering=ordering)
    vi = tlw.VolumeIntegralFR(tlw.MDRK() if td == "mdrk" else tlw.LW())
    if sc:
        ind = tlw.IndicatorHennemannGassner(eq, None, alpha_max=alpha_max, alpha_min=0.001, alpha_smooth=alpha_smooth)
        rec = tlw.MUSCLHancockReconstruction() if recon == "muscl_hancock" else tlw.FirstOrderReconstruction()
        vi = tlw.VolumeIntegralFRShockCapturing(ind, volume_integralFR=vi, reconstruction=rec)
    solver = tlw.DGSEM(N, volume_integral=vi)
    semi = tlw.SemidiscretizationHyperbolic(mesh, tlw.get_time_discretization(solver), eq, ic, solver,
                                            boundary_conditions=bc, surface_flux_order=flux_order,
                                            keep_aux=keep_aux)
    nx = mesh.nx
    # a caller-defined BoundaryConditionDirichlet reaches the oracle as a pointwise callback
    bc_fn = None
    if isinstance(bc, tlw.BoundaryConditionDirichlet):
        bc_fn = lambda x, y, t, f=bc.boundary_value_function: f((x, y), t, None)  # noqa: E731
    o = oracle_mod.Oracle(N, nx, nx, domain[0], domain[0], (domain[1] - domain[0]) / nx, periodic=(periodic,) * 2,
                          flux_order=0 if flux_order == "tree" else 1, bc_fn=bc_fn,
                          bc_kind=bc.kind if bc is not None else 0, bc_par=bc.params if bc is not None else (),
                          shock_capturing=sc, alpha_max=alpha_max, alpha_min=0.001, alpha_smooth=alpha_smooth,
                          reconstruction=recon if sc else "first_order")
    x, y = o.node_coords()
    rng = np.random.default_rng(seed)
    u_lex = ic(x, y, 0.0) * (1.0 + amp * rng.standard_normal(x.shape + (4,)))
    return semi, o, u_lex, mesh


def to_dev(mesh, u_lex):
    torch = torch_mod()
    perm = mesh.lexicographic_permutation()  # lex index -> element
    u = np.empty_like(u_lex)
    u[perm] = u_lex
    return torch.from_numpy(np.ascontiguousarray(u)).cuda()


def to_lex(mesh, u_dev):
    a = u_dev.detach().cpu().numpy() if hasattr(u_dev, "detach") else u_dev
    return a[mesh.lexicographic_permutation()]


def assert_close(a, b, tol=TOL, what=""):
    l2, li = rel_l2(a, b), rel_linf(a, b)
    assert l2 <= tol and li <= tol, f"{what}: rel L2 {l2:.3e}, rel Linf {li:.3e}"


GAMMA = 1.4


def warp_mapping(L, A):
    """[0, L]² warped by A sin 2πs sin 2πt in both coordinates (s, t the unit
    square coordinates); periodic, the warp vanishes on the domain boundary."""

    def mapping(xi, eta):
        s, t = 0.5 * (xi + 1.0), 0.5 * (eta + 1.0)
        w = A * np.sin(2.0 * np.pi * s) * np.sin(2.0 * np.pi * t)
        return L * (s + w), L * (t + w)

    return mapping


def setup_curved(oracle_mod, N, nx, ny, kind="warp", L=20.0, A=0.05, td="lw", seed=0, amp=1e-3, keep_aux=False,
                 periodic=True, sc=False, alpha_max=0.5, alpha_smooth=True, ordering="lexicographic", boundary=None):
    """Isentropic vortex on a warped or affine StructuredMesh of [0, L]², the
    device semidiscretization and the oracle on the same mesh with the oracle's
    metric terms on both sides.  A non-periodic mesh takes the vortex itself (no
    periodic images) as time-averaged Dirichlet data.  ``sc`` selects
    VolumeIntegralFRShockCapturing (Hennemann-Gassner on density_pressure,
    first-order subcell FV).  ``boundary`` replaces the Dirichlet data of a
    non-periodic mesh by a condition object with a kind code the oracle shares
    (boundary_condition_outflow, slip_wall_approximate)."""
    import trixilw_hip as tlw

    eq = tlw.CompressibleEulerEquations2D(GAMMA)
    period = (L, L * ny / nx if kind == "affine" else L) if periodic else (1e9, 1e9)
    ic = tlw.IsentropicVortex(center=(L / 2, L / 2), period=period)
    if kind == "warp":
        mesh = tlw.StructuredMesh((nx, ny), warp_mapping(L, A), periodicity=periodic, ordering=ordering)
        mp = (oracle_mod.MAP_WARP, (L, A))
    else:
        Ly = L * ny / nx
        mesh = tlw.StructuredMesh((nx, ny), (0.0, 0.0), (L, Ly), periodicity=periodic, ordering=ordering)
        mp = (oracle_mod.MAP_AFFINE, (0.0, 0.0, L, Ly))
    vi = tlw.VolumeIntegralFR(tlw.MDRK() if td == "mdrk" else tlw.LW())
    if sc:
        ind = tlw.IndicatorHennemannGassner(eq, None, alpha_max=alpha_max, alpha_min=0.001, alpha_smooth=alpha_smooth)
        vi = tlw.VolumeIntegralFRShockCapturing(ind, volume_integralFR=vi)
    solver = tlw.DGSEM(N, volume_integral=vi)
    bc = ic if boundary is None else boundary
    okw = {} if periodic else dict(periodic=(False, False), bc_kind=bc.kind, bc_par=bc.params)
    if sc:
        okw.update(shock_capturing=True, alpha_max=alpha_max, alpha_min=0.001, alpha_smooth=alpha_smooth)
    o = oracle_mod.Oracle(N, nx, ny, 0.0, 0.0, L / nx, gamma=GAMMA, mapping=mp, **okw)
    mesh.set_metrics(solver.basis, o.aux("contravariant_vectors"), o.aux("inverse_jacobian"))
    semi = tlw.SemidiscretizationHyperbolic(mesh, tlw.get_time_discretization(solver), eq, ic, solver,
                                            boundary_conditions=None if periodic else bc, keep_aux=keep_aux)
    x, y = o.node_coords()
    rng = np.random.default_rng(seed)
    u_lex = ic(x, y, 0.0) * (1.0 + amp * rng.standard_normal(x.shape + (4,)))
    return semi, o, u_lex, mesh
