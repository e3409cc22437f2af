"""Pins of the oracle's non-conforming TreeMesh (L2 mortars, dg_2d.jl:1456-1646
with Trixi's LobattoLegendreMortarL2 and mortar_fluxes_to_elements!):

- the four mortar operators against an independent numpy construction
  (Legendre-basis L2 projection with 40-point Gauss quadrature on each half,
  Lagrange interpolation), exactness on polynomials and the integral identity
  that makes mortars conservative;
- the geometric connectivity builder against the oracle's own uniform mesh,
  and the general-mesh path against the uniform one on a uniform mesh;
- on meshes with a refined box: free stream, discrete conservation (LW and
  both MDRK stages), symmetry under x <-> y transposition and x mirroring (the
  two orientations and both large sides), and the isentropic vortex's
  convergence order (LW N=3, 4 and MDRK N=3)."""
import numpy as np
import pytest

from conftest import rel_linf

L, X0 = 20.0, -10.0


def refined(oracle_mod, N, level, boxes=(((-5.0, -5.0), (5.0, 5.0)),), periodic=(True, True), **kw):
    import tree_mesh

    leaves = tree_mesh.leaves_refined((X0, X0), L, level, boxes)
    nbr, mort = tree_mesh.connectivity(leaves[:, 0], leaves[:, 1], leaves[:, 2], (X0, X0), L, periodic)
    o = oracle_mod.Oracle(N, 1, 1, 0.0, 0.0, 1.0, tree=(leaves[:, 0], leaves[:, 1], leaves[:, 2], nbr, mort), **kw)
    return o, leaves, nbr, mort


def lagrange_matrix(xs, ys):
    """M[a, k] = ℓ_k(ys[a]) for the Lagrange basis on xs (numpy polyfit-free form)."""
    V = np.vander(xs, increasing=True)
    return np.vander(ys, len(xs), increasing=True) @ np.linalg.inv(V)


@pytest.mark.parametrize("N", [3, 4])
def test_mortar_operators_match_l2_projection(oracle_mod, N):
    o = oracle_mod.Oracle(N, 2, 2, 0.0, 0.0, 1.0)
    nodes, w, _, _ = o.basis()
    fl, fu, rl, ru = o.mortar_matrices()
    np.testing.assert_allclose(fl, lagrange_matrix(nodes, 0.5 * (nodes - 1)), atol=1e-13)
    np.testing.assert_allclose(fu, lagrange_matrix(nodes, 0.5 * (nodes + 1)), atol=1e-13)
    # L2 projection onto degree N of a function given on each half by its GLL values
    g, gw = np.polynomial.legendre.leggauss(40)
    P = np.polynomial.legendre.legvander(g, N)                  # P_m at Gauss nodes of [-1, 1]
    mass = np.diag(2.0 / (2 * np.arange(N + 1) + 1))
    for half, R in ((-1, rl), (1, ru)):
        xs = 0.5 * (g + half)                                   # Gauss nodes of the half
        Ph = np.polynomial.legendre.legvander(xs, N)
        Lh = lagrange_matrix(nodes, g)                          # small-side GLL data at its own Gauss nodes
        rhs = (Ph * (0.5 * gw)[:, None]).T @ Lh                 # ∫_half P_m f dx per GLL value of f
        coef = np.linalg.solve(mass, rhs)
        want = np.polynomial.legendre.legvander(nodes, N) @ coef
        np.testing.assert_allclose(R, want, atol=1e-13)
    rng = np.random.default_rng(4)
    c = rng.standard_normal(N + 1)
    p = lambda x: np.polyval(c, x)  # noqa: E731
    np.testing.assert_allclose(fu @ p(nodes), p(0.5 * (nodes + 1)), atol=1e-13)
    np.testing.assert_allclose(ru @ p(0.5 * (nodes + 1)) + rl @ p(0.5 * (nodes - 1)), p(nodes), atol=1e-13)
    a, b = rng.standard_normal(N + 1), rng.standard_normal(N + 1)
    assert abs(w @ (ru @ a + rl @ b) - 0.5 * (w @ a + w @ b)) < 1e-14


@pytest.mark.parametrize("periodic", [(True, True), (False, True)])
def test_connectivity_of_a_uniform_mesh_matches_the_oracle(oracle_mod, periodic):
    import tree_mesh

    leaves = tree_mesh.leaves_refined((X0, X0), L, 3)
    nbr, mort = tree_mesh.connectivity(leaves[:, 0], leaves[:, 1], leaves[:, 2], (X0, X0), L, periodic)
    assert len(mort) == 0
    o = oracle_mod.Oracle(4, 8, 8, X0, X0, L / 8, periodic=periodic)
    assert np.array_equal(nbr, o.neighbors())


@pytest.mark.parametrize("kind", ["lw", "mdrk"])
def test_general_path_equals_uniform_path(oracle_mod, kind):
    from trixilw_hip.equations import IsentropicVortex

    o, leaves, nbr, mort = refined(oracle_mod, 4, 3, boxes=())
    ou = oracle_mod.Oracle(4, 8, 8, X0, X0, L / 8)
    x, y = ou.node_coords()
    xg, yg = o.node_coords()
    assert np.max(np.abs(x - xg)) < 1e-14 and np.max(np.abs(y - yg)) < 1e-14
    u = IsentropicVortex()(x, y, 0.0)
    dt = 0.5 * ou.max_dt(u)
    assert o.max_dt(u) == ou.max_dt(u)
    if kind == "lw":
        assert np.array_equal(o.rhs(u, 0.0, dt), ou.rhs(u, 0.0, dt))
    else:
        d1 = ou.rhs(u, 0.0, dt, kind="mdrk1")
        assert np.array_equal(o.rhs(u, 0.0, dt, kind="mdrk1"), d1)
        o.set_us(u + 0.5 * dt * d1)
        ou.set_us(u + 0.5 * dt * d1)
        assert np.array_equal(o.rhs(u, 0.0, dt, kind="mdrk2"), ou.rhs(u, 0.0, dt, kind="mdrk2"))


def test_refined_mesh_has_mortars_of_both_orientations_and_sides(oracle_mod):
    _, leaves, nbr, mort = refined(oracle_mod, 3, 3)
    assert len(leaves) == 64 - 16 + 64 and len(mort) == 16
    assert set(mort[:, 3]) == {1, 2} and set(mort[:, 4]) == {1, 2}
    assert np.count_nonzero(nbr == -2) == 16 + 32


@pytest.mark.parametrize("N", [3, 4])
@pytest.mark.parametrize("kind", ["lw", "mdrk"])
def test_refined_free_stream(oracle_mod, N, kind):
    from trixilw_hip.equations import ConstantState

    o, *_ = refined(oracle_mod, N, 2)
    x, y = o.node_coords()
    u = ConstantState(1.2, 0.3, -0.5, 0.9)(x, y, 0.0)
    dt = 0.3 * o.max_dt(u)
    if kind == "lw":
        du = o.rhs(u, 0.0, dt)
    else:
        du = o.rhs(u, 0.0, dt, kind="mdrk1")
        o.set_us(u + dt * du)
        du = o.rhs(u, 0.0, dt, kind="mdrk2")
    assert np.max(np.abs(du)) < 1e-12


@pytest.mark.parametrize("N", [3, 4])
@pytest.mark.parametrize("kind", ["lw", "mdrk1", "mdrk2"])
def test_refined_discrete_conservation(oracle_mod, N, kind):
    """Σ_e Σ_ij w_i w_j du / J⁻¹_e = 0: the large side's projected flux carries
    exactly the two small sides' integrals."""
    from trixilw_hip.equations import IsentropicVortex

    o, leaves, *_ = refined(oracle_mod, N, 3)
    x, y = o.node_coords()
    rng = np.random.default_rng(3)
    u = IsentropicVortex(center=(4.0, -4.5))(x, y, 0.0) * (1 + 1e-2 * rng.standard_normal(x.shape + (1,)))
    dt = 0.5 * o.max_dt(u)
    if kind == "mdrk2":
        o.set_us(u + 0.5 * dt * o.rhs(u, 0.0, dt, kind="mdrk1"))
    du = o.rhs(u, 0.0, dt, kind=kind) if kind != "lw" else o.rhs(u, 0.0, dt)
    _, w, _, _ = o.basis()
    W = np.outer(w, w)[None] * (0.5 * leaves[:, 2])[:, None, None] ** 2
    total = np.einsum("eji,ejiv->v", W, du)
    scale = np.einsum("eji,ejiv->v", W, np.abs(du))
    assert np.all(np.abs(total) <= 1e-13 * scale), total / scale


@pytest.mark.parametrize("sym", ["transpose", "mirror_x"])
def test_refined_symmetries(oracle_mod, sym):
    """A mesh refined off-centre so that mortars of both orientations and large
    sides change roles under the map: transposing (x <-> y, v1 <-> v2), or
    mirroring x (x -> -x, v1 -> -v1) both the mesh and the state maps the
    residual the same way.  Elements are matched by their centres, nodes by
    the map of (i, j)."""
    from trixilw_hip.equations import IsentropicVortex

    box = ((-7.5, -2.5), (2.5, 7.5))
    tbox = {"transpose": ((-2.5, -7.5), (7.5, 2.5)), "mirror_x": ((-2.5, -2.5), (7.5, 7.5))}[sym]
    o, lv, *_ = refined(oracle_mod, 4, 3, boxes=(box,))
    om, lvm, *_ = refined(oracle_mod, 4, 3, boxes=(tbox,))
    x, y = o.node_coords()
    rng = np.random.default_rng(11)
    u = IsentropicVortex(center=(0.3, -1.1))(x, y, 0.0) * (1 + 1e-2 * rng.standard_normal(x.shape + (4,)))
    if sym == "transpose":
        centre = {(round(b, 9), round(a, 9)): e for e, (a, b, _) in enumerate(lv)}
        node_map = lambda a: a.transpose(0, 2, 1, 3)[..., [0, 2, 1, 3]]  # noqa: E731
    else:
        centre = {(round(-a, 9), round(b, 9)): e for e, (a, b, _) in enumerate(lv)}
        node_map = lambda a: a[:, :, ::-1, :] * np.array([1.0, -1.0, 1.0, 1.0])  # noqa: E731
    emap = np.array([centre[(round(a, 9), round(b, 9))] for a, b, _ in lvm])
    um = np.ascontiguousarray(node_map(u[emap]))
    dt = 0.4 * o.max_dt(u)
    du = o.rhs(u, 0.0, dt)
    dum = om.rhs(um, 0.0, dt)
    assert rel_linf(dum, node_map(du[emap])) < 1e-12


@pytest.mark.parametrize("N,mdrk,eoc_min", [(4, False, 3.6), (3, False, 3.3), (3, True, 3.3)])
def test_refined_vortex_convergence(oracle_mod, N, mdrk, eoc_min):
    """EOC of the L2 error at t=1 with the vortex crossing mortars of a refined
    central box (levels 3, 4, 5 plus the box)."""
    from trixilw_hip.equations import IsentropicVortex

    ic = IsentropicVortex()
    fc = {3: 0.170, 4: 0.103}[N]
    cfl = 0.98 * (N + 1) * fc * (0.5 if mdrk else 1.0)
    errs = []
    for level in (3, 4, 5):
        o, leaves, *_ = refined(oracle_mod, N, level, boxes=(((-4.9, -4.9), (2.6, 2.6)),))
        x, y = o.node_coords()
        u, t, dts = o.solve(ic(x, y, 0.0), 0.0, 1.0, cfl, mdrk=mdrk)
        assert abs(t - 1.0) < 1e-12
        wj = (0.5 * leaves[:, 2])[:, None, None] ** 2
        errs.append(np.sqrt(np.sum(wj * (u - ic(x, y, t))[..., 0] ** 2) / np.sum(wj * np.ones_like(x))))
    eoc = np.log(errs[1] / errs[2]) / np.log(2)
    assert eoc > eoc_min, (errs, eoc)


def test_refined_mesh_refuses_unsupported(oracle_mod):
    import tree_mesh

    leaves = tree_mesh.leaves_refined((X0, X0), L, 2, (((-5.0, -5.0), (5.0, 5.0)),))
    nbr, mort = tree_mesh.connectivity(leaves[:, 0], leaves[:, 1], leaves[:, 2], (X0, X0), L)
    tree = (leaves[:, 0], leaves[:, 1], leaves[:, 2], nbr, mort)
    for kw in (dict(shock_capturing=True), dict(parabolic="navier_stokes", mu=0.01)):
        with pytest.raises(ValueError):
            oracle_mod.Oracle(3, 1, 1, 0.0, 0.0, 1.0, tree=tree, **kw)
