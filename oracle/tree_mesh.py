"""Connectivity of a TreeMesh with refined cells, for the oracle.

TEST INFRASTRUCTURE ONLY: used by tests/ as the checker; never part of the
product.

Trixi's TreeMesh (Trixi.jl src/meshes/tree_mesh.jl with the containers of
src/solvers/dgsem_tree/containers_2d.jl; Trixi is an unpinned dependency of the
reference's Project.toml) keeps leaf cells of a quadtree whose face neighbours
differ by at most one level.  A face between leaves of one level is an
interface; a face of a leaf whose neighbours across it are two leaves of the
next level is an L2 mortar (neighbor_ids lower, upper, large; orientation 1 for
x faces, 2 for y faces; large_side 1 when the large element is on the left or
below).  The reference's LW rhs loops over these containers
(src/solvers/dgsem_tree/dg_2d.jl:28-83, 1456-1646).

This restatement finds every face's neighbours by geometry alone (brute force
over all leaves), independent of how any tree was built, so it serves small
meshes only."""
from __future__ import annotations

import numpy as np


def leaves_refined(coordinates_min, length, level, boxes=(), balance=True, periodic=(True, True)):
    """Leaves (cx, cy, h) of the square [x0, x0 + length]² at the initial
    level, refined once per box in order: every leaf whose centre lies strictly
    inside the box is split in four (Trixi refine_box!, src/meshes/
    abstract_tree.jl), then leaves are split until face neighbours differ by at
    most one level (Trixi's 2:1 rebalance across faces, periodic faces
    included)."""
    x0, y0 = map(float, coordinates_min)
    h0 = float(length) / (1 << level)
    leaves = {(x0 + (i + 0.5) * h0, y0 + (j + 0.5) * h0, h0) for i in range(1 << level) for j in range(1 << level)}

    def split(c):
        cx, cy, h = c
        q = 0.25 * h
        return {(cx - q, cy - q, 0.5 * h), (cx + q, cy - q, 0.5 * h), (cx - q, cy + q, 0.5 * h),
                (cx + q, cy + q, 0.5 * h)}

    for (bx0, by0), (bx1, by1) in boxes:
        hit = {c for c in leaves if bx0 < c[0] < bx1 and by0 < c[1] < by1}
        for c in hit:
            leaves.remove(c)
            leaves |= split(c)
        shifts = [(sx, sy) for sx in ((-length, 0.0, length) if periodic[0] else (0.0,))
                  for sy in ((-length, 0.0, length) if periodic[1] else (0.0,))]
        while balance:
            L = np.array(sorted(leaves))
            bad = set()
            for a in range(len(L)):
                cx0, cy0, h = L[a]
                for sx, sy in shifts:  # the leaf and its periodic images
                    cx, cy = cx0 + sx, cy0 + sy
                    # leaves touching the faces that are more than twice smaller
                    near = (np.abs(L[:, 0] - cx) <= 0.5 * (h + L[:, 2]) + 1e-12 * length) & \
                           (np.abs(L[:, 1] - cy) <= 0.5 * (h + L[:, 2]) + 1e-12 * length)
                    face = near & ((np.abs(np.abs(L[:, 0] - cx) - 0.5 * (h + L[:, 2])) < 1e-12 * length) ^
                                   (np.abs(np.abs(L[:, 1] - cy) - 0.5 * (h + L[:, 2])) < 1e-12 * length))
                    if np.any(face & (L[:, 2] < 0.25 * h * (1 + 1e-9))):
                        bad.add(tuple(L[a]))
            if not bad:
                break
            for c in bad:
                leaves.remove(c)
                leaves |= split(c)
    return np.array(sorted(leaves, key=lambda c: (c[1], c[0])))


def connectivity(cx, cy, h, coordinates_min, length, periodic=(True, True)):
    """Face neighbours [E, 4] (faces -x, +x, -y, +y; -1 domain boundary, -2 on
    a mortar) and mortars [M, 5] (lower, upper, large, orientation, large_side)
    of the leaves with centres (cx, cy) and edge lengths h."""
    cx, cy, h = (np.asarray(a, dtype=np.float64) for a in (cx, cy, h))
    E = len(cx)
    x0, y0 = map(float, coordinates_min)
    L = float(length)
    tol = 1e-10 * L
    nbr = np.full((E, 4), -3, dtype=np.int32)
    mortars = []
    for e in range(E):
        for f in range(4):
            ax = f // 2  # 0: x faces, 1: y faces
            c_n = (cx, cy)[ax]          # normal coordinate
            c_t = (cy, cx)[ax]          # tangential coordinate
            lo = (x0, y0)[ax]
            sgn = 1.0 if f % 2 else -1.0
            pos = c_n[e] + sgn * 0.5 * h[e]
            if abs(pos - (lo + L)) < tol and sgn > 0 or abs(pos - lo) < tol and sgn < 0:
                if not periodic[ax]:
                    nbr[e, f] = -1
                    continue
                pos = lo if sgn > 0 else lo + L
            # leaves whose opposite face lies at pos and overlaps the face's span
            opp = c_n - sgn * 0.5 * h
            overlap = np.minimum(c_t + 0.5 * h, c_t[e] + 0.5 * h[e]) - np.maximum(c_t - 0.5 * h, c_t[e] - 0.5 * h[e])
            cand = np.nonzero((np.abs(opp - pos) < tol) & (overlap > tol))[0]
            if len(cand) == 1 and abs(h[cand[0]] - h[e]) < tol:
                nbr[e, f] = cand[0]
            elif len(cand) == 1 and abs(h[cand[0]] - 2 * h[e]) < tol:
                nbr[e, f] = -2  # small side of a mortar owned by the large neighbour
            elif len(cand) == 2 and all(abs(h[k] - 0.5 * h[e]) < tol for k in cand):
                nbr[e, f] = -2
                lower, upper = sorted(cand, key=lambda k: c_t[k])
                mortars.append((lower, upper, e, ax + 1, 1 if sgn > 0 else 2))
            else:
                raise ValueError(f"face {f} of leaf {e}: not a conforming or 2:1 face")
    assert np.all(nbr >= -2)
    return nbr, np.array(mortars, dtype=np.int32).reshape(-1, 5)
