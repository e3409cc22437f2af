"""Host-side StructuredMesh (no GPU): node coordinates and metric terms
against the oracle's restatement of Trixi's StructuredMesh cache, the
periodic neighbour table, and the affine mapping against TreeMesh node
coordinates."""
import numpy as np
import pytest


def warp_mapping(L, A):
    def mapping(xi, eta):
        s, t = 0.5 * (xi + 1.0), 0.5 * (eta + 1.0)
        w = A * np.sin(2.0 * np.pi * s) * np.sin(2.0 * np.pi * t)
        return L * (s + w), L * (t + w)

    return mapping


@pytest.mark.parametrize("N", [3, 4])
def test_metrics_match_oracle(oracle_mod, N):
    from trixilw_hip.basis import LobattoLegendreBasis
    from trixilw_hip.mesh import StructuredMesh

    nx, ny, L, A = 7, 5, 20.0, 0.1
    mesh = StructuredMesh((nx, ny), warp_mapping(L, A))
    basis = LobattoLegendreBasis(N)
    o = oracle_mod.Oracle(N, nx, ny, 0.0, 0.0, L / nx, mapping=(oracle_mod.MAP_WARP, (L, A)))
    x, y = mesh.node_coordinates(basis.nodes)
    xo, yo = o.node_coords()
    assert np.max(np.abs(x - xo)) <= 1e-13 * L and np.max(np.abs(y - yo)) <= 1e-13 * L
    Ja, ijn = mesh.metrics(basis)
    Jo = o.aux("contravariant_vectors").reshape(Ja.shape)
    io = o.aux("inverse_jacobian").reshape(ijn.shape)
    assert np.max(np.abs(Ja - Jo)) <= 1e-13 * np.max(np.abs(Jo))
    assert np.max(np.abs(ijn / io - 1.0)) <= 1e-13
    # a curved mesh: the off-diagonal metric terms are not small
    assert np.max(np.abs(Ja[..., 1])) > 1e-2 * np.max(np.abs(Ja[..., 0]))


def test_metric_identities_and_neighbours():
    """Discrete metric identities Σ_i D (Ja¹) + Σ_j D (Ja²) = 0 per component
    (free-stream preservation) and the periodic lexicographic neighbours."""
    from trixilw_hip.basis import LobattoLegendreBasis
    from trixilw_hip.mesh import StructuredMesh

    basis = LobattoLegendreBasis(4)
    mesh = StructuredMesh((6, 4), warp_mapping(3.0, 0.1))
    Ja, _ = mesh.metrics(basis)
    D = basis.derivative_matrix
    for m in range(2):
        div = np.einsum("ik,ejk->eji", D, Ja[..., m]) + np.einsum("jk,eki->eji", D, Ja[..., 2 + m])
        assert np.max(np.abs(div)) < 1e-12
    nbr = mesh.neighbors
    e = 6 * 2 + 5  # cx = 5, cy = 2
    assert list(nbr[e]) == [e - 1, 6 * 2, e - 6, e + 6]
    assert list(nbr[0]) == [5, 1, 6 * 3, 6]
    closed = StructuredMesh((6, 4), warp_mapping(3.0, 0.1), periodicity=False)
    assert list(closed.neighbors[0]) == [-1, 1, -1, 6]


def test_affine_mapping_matches_treemesh_coordinates():
    from trixilw_hip.basis import LobattoLegendreBasis
    from trixilw_hip.mesh import StructuredMesh, TreeMesh

    basis = LobattoLegendreBasis(3)
    tree = TreeMesh((-1.0, -1.0), (3.0, 3.0), 2, ordering="lexicographic")
    curv = StructuredMesh((4, 4), (-1.0, -1.0), (3.0, 3.0))
    xt, yt = tree.node_coordinates(basis.nodes)
    xc, yc = curv.node_coordinates(basis.nodes)
    assert np.max(np.abs(xt - xc)) < 1e-14 and np.max(np.abs(yt - yc)) < 1e-14
    assert np.array_equal(tree.neighbors, curv.neighbors)
    Ja, ijn = curv.metrics(basis)
    assert np.allclose(ijn, 1.0 / (0.5 * 0.5), rtol=1e-13, atol=0)
    assert np.allclose(Ja[..., 0], 0.5, rtol=1e-13) and np.allclose(Ja[..., 3], 0.5, rtol=1e-13)


@pytest.mark.parametrize("nx,ny,periodic", [(6, 4, True), (5, 7, False), (1, 3, True)])
def test_strips_ordering_is_the_lexicographic_mesh_permuted(nx, ny, periodic):
    """The strips-ordered mesh is the lexicographic one with its elements
    renumbered: same cells, neighbours, node coordinates and metric terms
    (host-computed and set from linear-index arrays) through the permutation,
    and runs of 10 elements are 5×2 blocks away from the strip ends."""
    from trixilw_hip.basis import LobattoLegendreBasis
    from trixilw_hip.mesh import StructuredMesh

    m = lambda o: StructuredMesh((nx, ny), warp_mapping(3.0, 0.1), periodicity=periodic, ordering=o)
    lex, st = m("lexicographic"), m("strips")
    p = st.lexicographic_permutation()  # linear index -> element
    assert sorted(p.tolist()) == list(range(nx * ny))
    assert np.array_equal(st.ix[p], lex.ix) and np.array_equal(st.iy[p], lex.iy)
    # a neighbour relation is the same relation under the renumbering
    back = np.where(lex.neighbors >= 0, p[np.maximum(lex.neighbors, 0)], -1)
    assert np.array_equal(st.neighbors[p], back)
    basis = LobattoLegendreBasis(3)
    for a, b in zip(st.node_coordinates(basis.nodes), lex.node_coordinates(basis.nodes)):
        assert np.array_equal(a[p], b)
    Ja, ijn = lex.metrics(basis)
    for a, b in zip(st.metrics(basis), (Ja, ijn)):
        assert np.array_equal(a[p], b)
    st.set_metrics(basis, Ja, ijn)
    for a, b in zip(st.metrics(basis), (Ja, ijn)):
        assert np.array_equal(a[p], b)
    # a round trip in the mesh's own element order is the identity
    own = [a.copy() for a in st.metrics(basis)]
    st.set_metrics(basis, *own, order="element")
    for a, b in zip(st.metrics(basis), own):
        assert np.array_equal(a, b)
    with pytest.raises(ValueError):
        st.set_metrics(basis, *own, order="morton")
    if ny >= 2 and nx >= 5:
        run = np.arange(10)
        assert set(zip(st.ix[run].tolist(), st.iy[run].tolist())) == {(i, j) for i in range(5) for j in range(2)}
    with pytest.raises(ValueError):
        m("hilbert")
