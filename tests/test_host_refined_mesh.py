"""The product's TreeMesh with refined cells (host side, no GPU) against the
oracle's geometric builder: the leaves Trixi's box patches and 2:1 rebalance
give, the face neighbours and the mortar list, the AMR refine/coarsen
operations and their solution-transfer plans (checked by carrying polynomial
fields with the oracle's mortar operators, where refining is exact and
coarsening inverts it)."""
import numpy as np
import pytest

from amr_helpers import transfer

CASES = [
    dict(lo=(-10.0, -10.0), L=20.0, level=3, periodic=(True, True), boxes=[((-5.0, -5.0), (5.0, 5.0))]),
    dict(lo=(0.0, 0.0), L=1.0, level=2, periodic=(False, False), boxes=[((0.0, 0.0), (0.3, 0.3)), ((0.0, 0.0), (0.1, 0.1))]),
    dict(lo=(-1.0, -1.0), L=2.0, level=3, periodic=(True, False), boxes=[((-0.9, -0.2), (0.2, 0.9)), ((-0.6, 0.1), (0.0, 0.6))]),
    dict(lo=(0.0, 0.0), L=4.0, level=2, periodic=(False, True), boxes=[((2.0, 0.0), (4.0, 4.0)), ((3.0, 3.0), (4.0, 4.0)),
                                                                       ((3.5, 3.5), (4.0, 4.0))]),
]


def product_mesh(case):
    from trixilw_hip.mesh import TreeMesh

    lo, L = case["lo"], case["L"]
    patches = [{"type": "box", "coordinates_min": a, "coordinates_max": b} for a, b in case["boxes"]]
    return TreeMesh(lo, (lo[0] + L, lo[1] + L), case["level"], periodicity=case["periodic"],
                    refinement_patches=patches)


def oracle_connectivity(mesh, case):
    import tree_mesh

    return tree_mesh.connectivity(mesh.element_center[:, 0], mesh.element_center[:, 1], mesh.element_length,
                                  case["lo"], case["L"], case["periodic"])


@pytest.mark.parametrize("k", range(len(CASES)))
def test_refined_leaves_and_connectivity_match_the_oracle(k):
    import tree_mesh

    case = CASES[k]
    mesh = product_mesh(case)
    leaves = tree_mesh.leaves_refined(case["lo"], case["L"], case["level"], case["boxes"], periodic=case["periodic"])
    got = {(round(a, 12), round(b, 12), round(h, 12)) for a, b, h in
           zip(mesh.element_center[:, 0], mesh.element_center[:, 1], mesh.element_length)}
    want = {(round(a, 12), round(b, 12), round(h, 12)) for a, b, h in leaves}
    assert got == want
    nbr, mort = oracle_connectivity(mesh, case)
    assert np.array_equal(nbr, mesh.neighbors)
    assert sorted(map(tuple, mort)) == sorted(map(tuple, mesh.mortars))
    assert len(mesh.mortars) > 0
    assert np.allclose(mesh.inverse_jacobian, 2.0 / mesh.element_length)


def test_amr_refine_and_coarsen_round_trip():
    import tree_mesh

    case = CASES[0]
    mesh = product_mesh(case)
    rng = np.random.default_rng(5)
    pick = rng.choice(mesh.n_elements, 9, replace=False)
    fine, plan = mesh.refine(pick)
    # the oracle's view of the same operation: split the picked leaves, rebalance by brute force
    picked = {(round(mesh.element_center[e, 0], 12), round(mesh.element_center[e, 1], 12)) for e in pick}
    assert len(fine.mortars) > 0
    nbr, mort = oracle_connectivity(fine, case)
    assert np.array_equal(nbr, fine.neighbors) and sorted(map(tuple, mort)) == sorted(map(tuple, fine.mortars))
    kinds = plan[:, 0]
    assert np.count_nonzero(kinds == 1) >= 4 * len(pick) and np.count_nonzero(kinds == 2) == 0
    for e in np.nonzero(kinds == 1)[0]:
        p, child = plan[e, 1], plan[e, 2]
        h = mesh.element_length[p]
        off = np.array([(child & 1) - 0.5, (child >> 1) - 0.5]) * 0.5 * h
        assert np.allclose(fine.element_center[e], mesh.element_center[p] + off)
        assert np.isclose(fine.element_length[e], 0.5 * h)
    for e in np.nonzero(kinds == 0)[0]:
        assert np.allclose(fine.element_center[e], mesh.element_center[plan[e, 1]])
    refined_parents = {(round(mesh.element_center[p, 0], 12), round(mesh.element_center[p, 1], 12))
                       for p in plan[kinds == 1, 1]}
    assert picked <= refined_parents
    # coarsening every child of a refined parent restores the parent wherever the balance allows
    back, plan2 = fine.coarsen(np.nonzero(kinds == 1)[0])
    assert np.count_nonzero(plan2[:, 0] == 2) > 0
    nbr, mort = oracle_connectivity(back, case)
    assert np.array_equal(nbr, back.neighbors) and sorted(map(tuple, mort)) == sorted(map(tuple, back.mortars))
    assert back.n_elements <= fine.n_elements - 3 * np.count_nonzero(plan2[:, 0] == 2)
    # coarsening that breaks the 2:1 balance is skipped
    deep, _ = mesh.refine([0])
    deeper, _ = deep.refine([int(np.argmin(deep.element_length))])
    kids = np.nonzero(np.isclose(deeper.element_length, deeper.element_length.max()))[0]
    same, _ = deeper.coarsen(kids)
    nbr, mort = oracle_connectivity(same, case)
    assert np.array_equal(nbr, same.neighbors)


@pytest.mark.parametrize("N", [3, 4])
def test_transfer_plans_carry_polynomials(oracle_mod, N):
    case = CASES[0]
    mesh = product_mesh(case)
    o = oracle_mod.Oracle(N, 2, 2, 0.0, 0.0, 1.0)
    nodes = o.basis()[0]
    fl, fu, rl, ru = o.mortar_matrices()
    x, y = mesh.node_coordinates(nodes)

    def field(x, y):  # degree N in x and in y: carried exactly by refinement
        return np.stack([1 + 0.1 * (x / 10) ** N - 0.2 * (y / 10) ** (N - 1), x * y / 100, (y / 10) ** N,
                         3 + (x / 10) * (y / 10) ** (N - 1)], axis=-1)

    u = field(x, y)
    fine, plan = mesh.refine(np.arange(0, mesh.n_elements, 7))
    xf, yf = fine.node_coordinates(nodes)
    uf = transfer(u, plan, (fl, fu), (rl, ru))
    assert np.max(np.abs(uf - field(xf, yf))) < 1e-12
    back, plan2 = fine.coarsen(np.nonzero(plan[:, 0] == 1)[0])
    ub = transfer(uf, plan2, (fl, fu), (rl, ru))
    xb, yb = back.node_coordinates(nodes)
    assert np.max(np.abs(ub - field(xb, yb))) < 1e-12
    # coarsening conserves the integral of any field
    rng = np.random.default_rng(2)
    r = rng.standard_normal(uf.shape)
    w = o.basis()[1]
    W = lambda m: np.outer(w, w)[None] * (0.5 * m.element_length)[:, None, None] ** 2  # noqa: E731
    rb = transfer(r, plan2, (fl, fu), (rl, ru))
    assert np.allclose(np.einsum("eji,ejiv->v", W(fine), r), np.einsum("eji,ejiv->v", W(back), rb), atol=1e-12)


def test_uniform_mesh_has_no_mortars():
    from trixilw_hip.mesh import TreeMesh

    m = TreeMesh((0.0, 0.0), (1.0, 1.0), 3)
    assert len(m.mortars) == 0 and not m.refined
    with pytest.raises(ValueError):
        product_mesh(CASES[0]).lexicographic_permutation()
