"""Pins of the oracle's curvilinear StructuredMesh path (no GPU).

An affine mapping must reproduce the TreeMesh solver run with the
Structured/P4est argument order of the LW Rusanov flux (λ from uⁿ, jump on Ũ,
src/solvers/dgsem_structured/dg_2d.jl:2271): every metric term, the contravariant
fluxes, the normal scaling of the interface flux and the per-node Jacobian
then reduce to the Cartesian ones.  On a warped periodic mesh the discrete
metric identities give free-stream preservation, the flux form gives
conservation (Σ w_i w_j J du = 0), and the isentropic vortex converges at high
order.
"""
import numpy as np
import pytest

from conftest import rel_linf


def vortex(L=20.0):
    from trixilw_hip.equations import IsentropicVortex

    return IsentropicVortex(center=(L / 2, L / 2), period=(L, L))


@pytest.mark.parametrize("N", [3, 4])
def test_affine_mapping_equals_treemesh_structured_order(oracle_mod, N):
    nx, ny, L = 6, 4, 20.0
    h = L / nx
    Ly = h * ny
    ic = vortex(L)
    tree = oracle_mod.Oracle(N, nx, ny, 0.0, 0.0, h, flux_order=1)
    curv = oracle_mod.Oracle(N, nx, ny, 0.0, 0.0, h, mapping=(oracle_mod.MAP_AFFINE, (0.0, 0.0, L, Ly)))
    xt, yt = tree.node_coords()
    xc, yc = curv.node_coords()
    assert np.max(np.abs(xt - xc)) < 1e-13 and np.max(np.abs(yt - yc)) < 1e-13
    rng = np.random.default_rng(21 + N)
    u = ic(xt, yt, 0.0) * (1 + 1e-2 * rng.standard_normal(xt.shape + (4,)))
    assert abs(tree.max_dt(u) - curv.max_dt(u)) < 1e-14 * tree.max_dt(u)
    dt = 0.8 * tree.max_dt(u)
    # the metric terms carry round-off of D applied to the coordinates (~1e-15 relative)
    assert rel_linf(curv.rhs(u, 0.0, dt), tree.rhs(u, 0.0, dt)) < 1e-12
    # the metric terms are the constant scalings of the box
    Ja = curv.aux("contravariant_vectors").reshape(-1, 4)
    assert np.allclose(Ja[:, 0], Ly / ny / 2, rtol=1e-13, atol=0) and np.allclose(Ja[:, 3], h / 2, rtol=1e-13, atol=0)
    assert np.max(np.abs(Ja[:, 1:3])) < 1e-12


@pytest.mark.parametrize("N", [3, 4])
def test_affine_mapping_linear_advection(oracle_mod, N):
    nx, L = 4, 2.0
    tree = oracle_mod.Oracle(N, nx, nx, 0.0, 0.0, L / nx, equation="linadv", adv=(1.0, 0.6), flux_order=1)
    curv = oracle_mod.Oracle(N, nx, nx, 0.0, 0.0, L / nx, equation="linadv", adv=(1.0, 0.6),
                             mapping=(oracle_mod.MAP_AFFINE, (0.0, 0.0, L, L)))
    u = np.random.default_rng(3).standard_normal((nx * nx, N + 1, N + 1, 1))
    dt = 0.05
    assert rel_linf(curv.rhs(u, 0.0, dt), tree.rhs(u, 0.0, dt)) < 1e-13


def warped(oracle_mod, N, nx, L=20.0, A=0.05, **kw):
    return oracle_mod.Oracle(N, nx, nx, 0.0, 0.0, L / nx, mapping=(oracle_mod.MAP_WARP, (L, A)), **kw)


@pytest.mark.parametrize("N", [3, 4])
def test_warped_mesh_metrics_and_free_stream(oracle_mod, N):
    o = warped(oracle_mod, N, 6)
    Ja = o.aux("contravariant_vectors").reshape(-1, 4)
    assert np.max(np.abs(Ja[:, 1])) > 1e-2  # the mesh is really curved
    assert np.all(o.aux("inverse_jacobian") > 0)
    x, _ = o.node_coords()
    u = np.zeros(x.shape + (4,))
    u[...] = [1.1, 0.4, -0.3, 2.7]
    du = o.rhs(u, 0.0, 0.5 * o.max_dt(u))
    assert np.max(np.abs(du)) < 1e-12


@pytest.mark.parametrize("N", [3, 4])
def test_warped_mesh_conservation(oracle_mod, N):
    o = warped(oracle_mod, N, 6)
    x, y = o.node_coords()
    rng = np.random.default_rng(8)
    u = vortex()(x, y, 0.0) * (1 + 1e-3 * rng.standard_normal(x.shape + (4,)))
    du = o.rhs(u, 0.0, 0.5 * o.max_dt(u))
    _, w, _, _ = o.basis()
    W = np.outer(w, w)[None] / o.aux("inverse_jacobian").reshape(x.shape)
    total = np.einsum("eji,ejiv->v", W, du)
    scale = np.einsum("eji,ejiv->v", W, np.abs(du))
    assert np.all(np.abs(total) <= 1e-13 * scale)


@pytest.mark.parametrize("N,eoc_min", [(3, 3.4), (4, 3.8)])
def test_warped_mesh_vortex_convergence(oracle_mod, N, eoc_min):
    ic = vortex()
    errs = []
    for nx in (16, 32, 64):
        o = warped(oracle_mod, N, nx)
        x, y = o.node_coords()
        cfl = 0.9 * (N + 1) * {3: 0.170, 4: 0.103}[N]
        u, t, _ = o.solve(ic(x, y, 0.0), 0.0, 1.0, cfl)
        assert abs(t - 1.0) < 1e-12
        errs.append(np.sqrt(np.mean((u - ic(x, y, t))[..., 0] ** 2)))
    eoc = np.log(errs[1] / errs[2]) / np.log(2)
    assert eoc > eoc_min, (errs, eoc)


@pytest.mark.parametrize("N", [3, 4])
def test_affine_mapping_mdrk_equals_treemesh_structured_order(oracle_mod, N):
    """Both MDRK stages of the curvilinear kernels (mdrk_kernel_1!/2! for
    StructuredMesh) on an affine box equal the TreeMesh MDRK stages with the
    Structured flux order; the stage cache (F2, U2) carries over."""
    nx, ny, L = 6, 4, 20.0
    h = L / nx
    ic = vortex(L)
    tree = oracle_mod.Oracle(N, nx, ny, 0.0, 0.0, h, flux_order=1)
    curv = oracle_mod.Oracle(N, nx, ny, 0.0, 0.0, h, mapping=(oracle_mod.MAP_AFFINE, (0.0, 0.0, L, h * ny)))
    x, y = tree.node_coords()
    rng = np.random.default_rng(31 + N)
    u = ic(x, y, 0.0) * (1 + 1e-2 * rng.standard_normal(x.shape + (4,)))
    dt = 0.4 * tree.max_dt(u)
    d1t, d1c = tree.rhs(u, 0.3, dt, kind="mdrk1"), curv.rhs(u, 0.3, dt, kind="mdrk1")
    assert rel_linf(d1c, d1t) < 1e-12
    assert rel_linf(curv.aux("U"), tree.aux("U")) < 1e-13 and rel_linf(curv.aux("F"), tree.aux("F")) < 1e-12
    us = u + dt * d1t
    tree.set_us(us)
    curv.set_us(us)
    d2t, d2c = tree.rhs(u, 0.3, dt, kind="mdrk2"), curv.rhs(u, 0.3, dt, kind="mdrk2")
    assert rel_linf(d2c, d2t) < 1e-12
    # the stages differ from LW and from each other
    assert rel_linf(d1c, curv.rhs(u, 0.3, dt)) > 1e-3 and rel_linf(d2c, d1c) > 1e-4


@pytest.mark.parametrize("N", [3, 4])
def test_warped_mesh_mdrk_free_stream_and_conservation(oracle_mod, N):
    o = warped(oracle_mod, N, 6)
    x, y = o.node_coords()
    c = np.zeros(x.shape + (4,))
    c[...] = [1.1, 0.4, -0.3, 2.7]
    dt = 0.3 * o.max_dt(c)
    assert np.max(np.abs(o.rhs(c, 0.0, dt, kind="mdrk1"))) < 1e-12
    o.set_us(c)
    assert np.max(np.abs(o.rhs(c, 0.0, dt, kind="mdrk2"))) < 1e-12
    rng = np.random.default_rng(9)
    u = vortex()(x, y, 0.0) * (1 + 1e-3 * rng.standard_normal(x.shape + (4,)))
    _, w, _, _ = o.basis()
    W = np.outer(w, w)[None] / o.aux("inverse_jacobian").reshape(x.shape)
    d1 = o.rhs(u, 0.0, dt, kind="mdrk1")
    o.set_us(u + dt * d1)
    d2 = o.rhs(u, 0.0, dt, kind="mdrk2")
    for du in (d1, d2):
        total = np.einsum("eji,ejiv->v", W, du)
        scale = np.einsum("eji,ejiv->v", W, np.abs(du))
        assert np.all(np.abs(total) <= 1e-13 * scale)


def test_warped_mesh_mdrk_vortex_convergence(oracle_mod):
    """MDRK (two-stage fourth order) at polydeg 3 on the warped mesh."""
    ic = vortex()
    errs = []
    for nx in (16, 32, 64):
        o = warped(oracle_mod, 3, nx)
        x, y = o.node_coords()
        cfl = 0.9 * 4 * 0.170 * 0.5
        u, t, _ = o.solve(ic(x, y, 0.0), 0.0, 1.0, cfl, mdrk=True)
        assert abs(t - 1.0) < 1e-12
        errs.append(np.sqrt(np.mean((u - ic(x, y, t))[..., 0] ** 2)))
    eoc = np.log(errs[1] / errs[2]) / np.log(2)
    assert eoc > 3.4, (errs, eoc)


@pytest.mark.parametrize("N", [3, 4])
def test_affine_mapping_with_dirichlet_boundary_equals_treemesh(oracle_mod, N):
    """A non-periodic affine StructuredMesh with the time-averaged Dirichlet
    boundary (calc_boundary_flux! of structured/dg_2d.jl:2297-2391: the
    element's own contravariant vector as normal, TreeMesh argument order,
    full-step average in both MDRK stages) equals the TreeMesh solver in the
    StructuredMesh conventions (flux_order 2), whose TreeMesh pieces the linear
    closed form pins.  LW and both MDRK stages."""
    from trixilw_hip.equations import IsentropicVortex

    nx, ny, L = 6, 4, 20.0
    h = L / nx
    ic = IsentropicVortex(center=(L / 2, 0.4 * h * ny), period=(1e9, 1e9))
    kw = dict(periodic=(False, False), bc_kind=ic.kind, bc_par=ic.params)
    tree = oracle_mod.Oracle(N, nx, ny, 0.0, 0.0, h, flux_order=2, **kw)
    curv = oracle_mod.Oracle(N, nx, ny, 0.0, 0.0, h, mapping=(oracle_mod.MAP_AFFINE, (0.0, 0.0, L, h * ny)), **kw)
    x, y = tree.node_coords()
    rng = np.random.default_rng(40 + N)
    u = ic(x, y, 0.0) * (1 + 1e-2 * rng.standard_normal(x.shape + (4,)))
    dt = 0.8 * tree.max_dt(u)
    assert rel_linf(curv.rhs(u, 0.3, dt), tree.rhs(u, 0.3, dt)) < 1e-12
    d1t, d1c = tree.rhs(u, 0.3, dt, kind="mdrk1"), curv.rhs(u, 0.3, dt, kind="mdrk1")
    assert rel_linf(d1c, d1t) < 1e-12
    us = u + 0.5 * dt * d1t
    tree.set_us(us)
    curv.set_us(us)
    assert rel_linf(curv.rhs(u, 0.3, dt, kind="mdrk2"), tree.rhs(u, 0.3, dt, kind="mdrk2")) < 1e-12
    # the boundary matters: a periodic mesh gives another residual
    per = oracle_mod.Oracle(N, nx, ny, 0.0, 0.0, h, mapping=(oracle_mod.MAP_AFFINE, (0.0, 0.0, L, h * ny)))
    assert rel_linf(per.rhs(u, 0.3, dt), curv.rhs(u, 0.3, dt)) > 1e-3


@pytest.mark.parametrize("N", [3, 4])
def test_warped_mesh_dirichlet_free_stream(oracle_mod, N):
    """A constant boundary value equal to the constant state keeps it (LW and
    MDRK stage 2): the boundary normal and its sign agree with the interior
    metric terms.  MDRK stage 1 does not: the reference averages the boundary
    data over the full step (structured/multi_derivative_2d.jl:18-20 passes no
    scaling) while the stage-1 element flux is f/2 + f_t/8, so on a boundary
    node F* = ½(f/2 + f)·n = ¾ f·n against f/2 inside; interior elements stay
    at rest."""
    from trixilw_hip.equations import ConstantState

    bc = ConstantState(rho=1.1, v1=0.3, v2=-0.25, p=0.9)
    o = warped(oracle_mod, N, 6, periodic=(False, False), bc_kind=bc.kind, bc_par=bc.params)
    x, y = o.node_coords()
    u = np.ascontiguousarray(bc(x, y, 0.0))
    assert np.ptp(u[..., 1]) == 0.0 and abs(u[0, 0, 0, 1] - 0.33) < 1e-15
    dt = 0.5 * o.max_dt(u)
    assert np.max(np.abs(o.rhs(u, 0.0, dt))) < 1e-12
    d1 = o.rhs(u, 0.0, dt, kind="mdrk1")
    o.set_us(u)
    assert np.max(np.abs(o.rhs(u, 0.0, dt, kind="mdrk2"))) < 1e-12
    edge = np.any(o.neighbors() < 0, axis=1)
    assert np.max(np.abs(d1[~edge])) < 1e-12 and np.max(np.abs(d1[edge])) > 1e-3


@pytest.mark.parametrize("N,eoc_min", [(3, 3.4), (4, 3.8)])
def test_warped_mesh_dirichlet_vortex_convergence(oracle_mod, N, eoc_min):
    """The isentropic vortex with its exact solution as time-averaged Dirichlet
    data on the non-periodic warped mesh converges at high order."""
    from trixilw_hip.equations import IsentropicVortex

    ic = IsentropicVortex(center=(10.0, 10.0), period=(1e9, 1e9))
    errs = []
    for nx in (16, 32, 64):
        o = warped(oracle_mod, N, nx, periodic=(False, False), bc_kind=ic.kind, bc_par=ic.params)
        x, y = o.node_coords()
        cfl = 0.9 * (N + 1) * {3: 0.170, 4: 0.103}[N]
        u, t, _ = o.solve(ic(x, y, 0.0), 0.0, 1.0, cfl)
        assert abs(t - 1.0) < 1e-12
        errs.append(np.sqrt(np.mean((u - ic(x, y, t))[..., 0] ** 2)))
    eoc = np.log(errs[1] / errs[2]) / np.log(2)
    assert eoc > eoc_min, (errs, eoc)


# ---------------------------------------------------------------- shock capturing
def _euler_flux_normal(u, nrm, g=1.4):
    rho = u[..., 0]
    v1, v2 = u[..., 1] / rho, u[..., 2] / rho
    p = (g - 1) * (u[..., 3] - 0.5 * (u[..., 1] * v1 + u[..., 2] * v2))
    vn = v1 * nrm[0] + v2 * nrm[1]
    return np.stack([rho * vn, rho * vn * v1 + p * nrm[0], rho * vn * v2 + p * nrm[1], (u[..., 3] + p) * vn], -1)


def _rusanov_normal(ul, ur, nrm, g=1.4):
    def parts(u):
        rho = u[..., 0]
        v1, v2 = u[..., 1] / rho, u[..., 2] / rho
        p = (g - 1) * (u[..., 3] - 0.5 * (u[..., 1] * v1 + u[..., 2] * v2))
        return np.abs(v1 * nrm[0] + v2 * nrm[1]), np.sqrt(g * p / rho) * np.hypot(*nrm)

    al, cl = parts(ul)
    ar, cr = parts(ur)
    lam = np.maximum(al, ar) + np.maximum(cl, cr)
    return 0.5 * (_euler_flux_normal(ul, nrm, g) + _euler_flux_normal(ur, nrm, g)) - 0.5 * lam * (ur - ul)


@pytest.mark.parametrize("N", [3, 4])
@pytest.mark.parametrize("dtf", [0.4, 1.5])
def test_affine_mapping_shock_capturing_equals_treemesh(oracle_mod, N, dtf):
    """First-order shock capturing on an affine StructuredMesh (subcell normals
    of Trixi's calcflux_fv!, per-node J⁻¹ in compute_alp, the calc_fn_low!
    normals) equals the TreeMesh shock capturing with the Structured order:
    LW and both MDRK stages, blending factors with exact zeros (pure DG), and at
    1.5× the stable dt, where the admissibility fallback switches interfaces."""
    from trixilw_hip.equations import IsentropicVortex

    nx, ny, L = 6, 4, 20.0
    h = L / nx
    ic = IsentropicVortex(center=(L / 2, h * ny / 2), period=(L, h * ny))
    tree = oracle_mod.Oracle(N, nx, ny, 0.0, 0.0, h, flux_order=1, shock_capturing=True)
    curv = oracle_mod.Oracle(N, nx, ny, 0.0, 0.0, h, mapping=(oracle_mod.MAP_AFFINE, (0.0, 0.0, L, h * ny)),
                             shock_capturing=True)
    x, y = tree.node_coords()
    rng = np.random.default_rng(50 + N)
    u = ic(x, y, 0.0) * (1 + 5e-2 * rng.standard_normal(x.shape + (4,)))
    alpha = np.where(rng.random(nx * ny) < 0.3, 0.0, rng.random(nx * ny))
    for o in (tree, curv):
        o.set_alpha(alpha)
    dt = dtf * tree.max_dt(u)
    # the affine metric terms carry the round-off of D applied to the coordinates (~1e-14
    # relative), which the flux cancellation of the larger step amplifies to ~1e-12
    tol = 1e-12 if dtf < 1 else 5e-12
    assert rel_linf(curv.rhs(u, 0.2, dt), tree.rhs(u, 0.2, dt)) < tol
    assert rel_linf(curv.aux("fn_low"), tree.aux("fn_low") * (h / 2)) < tol  # contravariant: |Ja| = h/2
    d1 = tree.rhs(u, 0.2, dt, kind="mdrk1")
    assert rel_linf(curv.rhs(u, 0.2, dt, kind="mdrk1"), d1) < tol
    us = u + 0.5 * dt * d1
    tree.set_us(us)
    curv.set_us(us)
    assert rel_linf(curv.rhs(u, 0.2, dt, kind="mdrk2"), tree.rhs(u, 0.2, dt, kind="mdrk2")) < tol


@pytest.mark.parametrize("N", [3, 4])
def test_warped_mesh_shock_capturing_free_stream_and_conservation(oracle_mod, N):
    """The subcell normals telescope to the element face normals (SBP), so a
    constant state stays at rest at any blending; blended residuals conserve."""
    o = warped(oracle_mod, N, 6, A=0.1, shock_capturing=True)
    x, y = o.node_coords()
    c = np.zeros(x.shape + (4,))
    c[...] = [1.1, 0.4, -0.3, 2.7]
    rng = np.random.default_rng(60 + N)
    for alpha in (np.ones(o.E), rng.random(o.E)):
        o.set_alpha(alpha)
        assert np.max(np.abs(o.rhs(c, 0.0, 0.4 * o.max_dt(c)))) < 1e-12
    u = vortex()(x, y, 0.0) * (1 + 1e-2 * rng.standard_normal(x.shape + (4,)))
    du = o.rhs(u, 0.0, 0.4 * o.max_dt(u))
    _, w, _, _ = o.basis()
    W = np.outer(w, w)[None] / o.aux("inverse_jacobian").reshape(x.shape)
    total = np.einsum("eji,ejiv->v", W, du)
    scale = np.einsum("eji,ejiv->v", W, np.abs(du))
    assert np.all(np.abs(total) <= 1e-13 * scale)


@pytest.mark.parametrize("N", [3, 4])
def test_warped_mesh_fn_low_normals(oracle_mod, N):
    """fn_low next to the + ξ face: blended elements use Trixi's accumulated
    subcell normal Ja¹(1) + Σ_{q<n-1} w_q (D Ja¹)(q); pure-DG elements the
    reference's calc_fn_low! normal, which adds the q = 1 and q = n-1 terms only
    (its loop runs over the tuple (2, nnodes))."""
    n = N + 1
    o = warped(oracle_mod, N, 4, A=0.12, shock_capturing=True)
    x, y = o.node_coords()
    rng = np.random.default_rng(70 + N)
    u = vortex()(x, y, 0.0) * (1 + 1e-2 * rng.standard_normal(x.shape + (4,)))
    alpha = np.where(np.arange(o.E) % 2 == 0, 0.0, 0.6)
    o.set_alpha(alpha)
    o.rhs(u, 0.0, 0.3 * o.max_dt(u))
    fl = o.aux("fn_low").reshape(o.E, 4, n, 4)
    Ja = o.aux("contravariant_vectors").reshape(o.E, n, n, 4)  # [e, j, i, (Ja11, Ja12, Ja21, Ja22)]
    _, w, D, _ = o.basis()
    D = np.asarray(D).reshape(n, n)
    for e in (0, 1):
        for j in range(n):
            ja = Ja[e, j, :, 0:2]  # Ja¹ along the row
            dja = D @ ja  # (D Ja¹)(q)
            qs = (0, n - 2) if alpha[e] == 0.0 else range(n - 1)
            nrm = ja[0] + sum(w[q] * dja[q] for q in qs)
            expect = _rusanov_normal(u[e, j, n - 2], u[e, j, n - 1], nrm)
            np.testing.assert_allclose(fl[e, 1, j], expect, rtol=1e-12, atol=1e-13)
    # the two normals differ on this curved mesh, so the check tells them apart
    ja = Ja[0, 1, :, 0:2]
    dja = D @ ja
    assert np.max(np.abs(sum(w[q] * dja[q] for q in range(1, n - 2)))) > 1e-4


def test_affine_curved_temporal_errors_and_adaptive_solve_equal_treemesh(oracle_mod):
    """The curvilinear kernel's temporal-error estimate (structured dg_2d.jl,
    u_np1 vs u_np1_low without the last stage's term) on an affine mesh equals
    the TreeMesh kernel's at N=4, element by element, and an adaptive solve then
    takes the same steps, accepted and rejected; at N=3 the estimate is the f_ttt
    term's update, nonzero and smaller on a finer mesh."""
    nx, ny, L = 6, 4, 20.0
    h = L / nx
    ic = vortex(L)
    tol = dict(tolerances=(1e-7, 1e-7))
    tree = oracle_mod.Oracle(4, nx, ny, 0.0, 0.0, h, flux_order=1, **tol)
    curv = oracle_mod.Oracle(4, nx, ny, 0.0, 0.0, h, mapping=(oracle_mod.MAP_AFFINE, (0.0, 0.0, L, h * ny)), **tol)
    x, y = tree.node_coords()
    u = ic(x, y, 0.0) * (1 + 1e-2 * np.random.default_rng(5).standard_normal(x.shape + (4,)))
    dt = 0.8 * tree.max_dt(u)
    tree.rhs(u, 0.0, dt)
    curv.rhs(u, 0.0, dt)
    te_t, te_c = tree.aux("temporal_errors"), curv.aux("temporal_errors")
    assert np.max(te_t) > 0 and rel_linf(te_c, te_t) < 1e-9
    ut, tt, dts_t, rt = tree.solve_adaptive(u, 0.0, 0.5, 1e-3)
    uc, tc, dts_c, rc = curv.solve_adaptive(u, 0.0, 0.5, 1e-3)
    assert len(dts_t) == len(dts_c) > 3 and rt == rc and abs(tt - tc) < 1e-13
    assert np.allclose(dts_t, dts_c, rtol=1e-9, atol=0)
    assert rel_linf(uc, ut) < 1e-10
    errs = []
    for m in (4, 8):
        c3 = oracle_mod.Oracle(3, m, m, 0.0, 0.0, L / m, mapping=(oracle_mod.MAP_WARP, (L, 0.05)), **tol)
        xc, yc = c3.node_coords()
        u3 = ic(xc, yc, 0.0)
        c3.rhs(u3, 0.0, 0.5 * c3.max_dt(u3))
        errs.append(np.sum(c3.aux("temporal_errors")))
    assert errs[0] > errs[1] > 0
