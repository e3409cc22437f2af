"""2-D TreeMesh (uniform or with refined cells) and curvilinear StructuredMesh on the host.

Mirrors what TrixiLW reads from Trixi's TreeMesh and its containers
(`cache.elements.inverse_jacobian`, `node_coordinates`, `cache.interfaces`,
`cache.boundaries`; src/solvers/dgsem_tree/dg_2d.jl:1269-1454).  The device
only sees a per-element table of the four face neighbours, element centres
and lengths, so any element order works; ``ordering="morton"`` numbers the
leaves in Z-order like Trixi's uniformly refined tree, ``"lexicographic"``
uses e = ix + nx*iy, and ``"strips"`` walks two-row strips column by column
(runs of 10 elements are 5×2 blocks: the volume kernel's 10-element tiles then
have fewer faces that leave the tile, so fewer face records cross HBM).

A TreeMesh with ``refinement_patches`` (or one returned by ``refine`` /
``coarsen``, the AMR operations) holds leaves of several levels, 2:1 across
faces, numbered in Z-order; faces between levels are L2 mortars
(``mortars``: lower, upper, large element, orientation, large side) and carry
the neighbour code TLW_NEIGHBOR_MORTAR (-2).
"""
from __future__ import annotations

import numpy as np


def _morton_index(ix, iy, level):
    z = np.zeros_like(ix)
    for b in range(level):
        z |= ((ix >> b) & 1) << (2 * b)
        z |= ((iy >> b) & 1) << (2 * b + 1)
    return z


class TreeMesh:
    """TreeMesh(coordinates_min, coordinates_max, initial_refinement_level, periodicity;
    refinement_patches): Trixi's TreeMesh constructor with its box patches
    (``{"type": "box", "coordinates_min": ..., "coordinates_max": ...}``): each
    patch refines once every leaf whose centre lies strictly inside the box
    (Trixi refine_box!), and the tree is then refined further until leaves
    sharing a face differ by at most one level (Trixi's rebalance)."""

    mortars = np.zeros((0, 5), dtype=np.int32)

    def __init__(self, coordinates_min, coordinates_max, initial_refinement_level, periodicity=True,
                 ordering="morton", refinement_patches=()):
        xmin, ymin = map(float, coordinates_min)
        xmax, ymax = map(float, coordinates_max)
        length = xmax - xmin
        if not np.isclose(ymax - ymin, length):
            raise ValueError("TreeMesh domains are square")
        if isinstance(periodicity, bool):
            periodicity = (periodicity, periodicity)
        self.periodicity = tuple(bool(p) for p in periodicity)
        self.level = int(initial_refinement_level)
        self.coordinates_min = (xmin, ymin)
        self.coordinates_max = (xmax, ymax)
        self.length = length
        if refinement_patches:
            leaves = {(self.level, i, j) for i in range(1 << self.level) for j in range(1 << self.level)}
            for patch in refinement_patches:
                if patch.get("type", "box") != "box":
                    raise NotImplementedError("refinement patches of type 'box'")
                (bx0, by0), (bx1, by1) = patch["coordinates_min"], patch["coordinates_max"]
                hit = [c for c in leaves if bx0 < self._centre(c)[0] < bx1 and by0 < self._centre(c)[1] < by1]
                leaves = self._balanced(self._split(leaves, hit))
            self._set_leaves(leaves)
            return
        nx = 1 << self.level
        self.nx = self.ny = nx
        self.h = length / nx
        iy, ix = np.meshgrid(np.arange(nx), np.arange(nx), indexing="ij")
        ix, iy = ix.ravel(), iy.ravel()  # lexicographic: k = ix + nx*iy
        if ordering == "morton":
            z = _morton_index(ix, iy, self.level)
            order = np.argsort(z, kind="stable")
        elif ordering == "lexicographic":
            order = np.arange(nx * nx)
        elif ordering == "strips":
            # two-row strips, column by column within a strip: consecutive runs of
            # 10 elements are 5×2 blocks, whose 40 face slots have 14 external faces
            # (Morton runs of 10: 17)
            order = np.lexsort((iy % 2, ix, iy // 2))
        else:
            raise ValueError(f"unknown ordering {ordering!r}")
        self.ordering = ordering
        self.ix, self.iy = ix[order], iy[order]          # cell indices of element e
        self.element_of_cell = np.empty(nx * nx, dtype=np.int64)
        self.element_of_cell[self.ix + nx * self.iy] = np.arange(nx * nx)
        self.n_elements = nx * nx
        E = self.n_elements
        centers = np.empty((E, 2))
        centers[:, 0] = xmin + (self.ix + 0.5) * self.h
        centers[:, 1] = ymin + (self.iy + 0.5) * self.h
        self.element_center = centers
        self.element_length = np.full(E, self.h)
        self.inverse_jacobian = 2.0 / self.element_length
        nbr = np.full((E, 4), -1, dtype=np.int32)
        px, py = self.periodicity
        for face, (dx, dy) in enumerate(((-1, 0), (1, 0), (0, -1), (0, 1))):
            jx, jy = self.ix + dx, self.iy + dy
            if px:
                jx = jx % nx
            if py:
                jy = jy % nx
            ok = (jx >= 0) & (jx < nx) & (jy >= 0) & (jy < nx)
            cell = np.where(ok, jx + nx * jy, 0)
            nbr[:, face] = np.where(ok, self.element_of_cell[cell], -1)
        self.neighbors = nbr

    # ------------------------------------------------------------------ refined cells
    def _centre(self, c):
        lev, i, j = c
        h = self.length / (1 << lev)
        return self.coordinates_min[0] + (i + 0.5) * h, self.coordinates_min[1] + (j + 0.5) * h

    @staticmethod
    def _split(leaves, cells):
        out = set(leaves)
        for lev, i, j in cells:
            out.remove((lev, i, j))
            out |= {(lev + 1, 2 * i + a, 2 * j + b) for a in (0, 1) for b in (0, 1)}
        return out

    def _wrap(self, lev, i, j):
        """The cell (lev, i, j) moved into the domain by periodicity, or None outside it."""
        n = 1 << lev
        if not 0 <= i < n:
            if not self.periodicity[0]:
                return None
            i %= n
        if not 0 <= j < n:
            if not self.periodicity[1]:
                return None
            j %= n
        return lev, i, j

    @staticmethod
    def _covering(leaves, c):
        """The leaf containing cell c (c itself or an ancestor), or None if c is subdivided."""
        lev, i, j = c
        while lev >= 0:
            if (lev, i, j) in leaves:
                return lev, i, j
            lev, i, j = lev - 1, i >> 1, j >> 1
        return None

    def _balanced(self, leaves):
        """Split leaves until face neighbours differ by at most one level."""
        leaves = set(leaves)
        while True:
            coarse = set()
            for lev, i, j in leaves:
                for di, dj in ((-1, 0), (1, 0), (0, -1), (0, 1)):
                    c = self._wrap(lev, i + di, j + dj)
                    if c is None:
                        continue
                    cov = self._covering(leaves, c)
                    if cov is not None and cov[0] < lev - 1:
                        coarse.add(cov)
            if not coarse:
                return leaves
            leaves = self._split(leaves, coarse)

    def _set_leaves(self, leaves):
        """Element arrays, face neighbours and mortars of a set of leaves (lev, i, j)."""
        leaves = set(leaves)
        arr = np.array(sorted(leaves), dtype=np.int64)
        top = int(arr[:, 0].max())
        z = _morton_index(arr[:, 1] << (top - arr[:, 0]), arr[:, 2] << (top - arr[:, 0]), top)
        arr = arr[np.lexsort((arr[:, 0], z))]  # Z-order of the leaves' lower-left corners
        order = [tuple(int(q) for q in c) for c in arr]
        self.leaves = arr
        self.ordering = "morton"
        self.max_level = top
        E = len(order)
        self.n_elements = E
        index = {c: e for e, c in enumerate(order)}
        h = self.length / (1 << self.leaves[:, 0])
        self.element_length = h.astype(np.float64)
        self.element_center = np.stack([self.coordinates_min[0] + (self.leaves[:, 1] + 0.5) * h,
                                        self.coordinates_min[1] + (self.leaves[:, 2] + 0.5) * h], axis=1)
        self.inverse_jacobian = 2.0 / self.element_length
        nbr = np.full((E, 4), -1, dtype=np.int32)
        mortars = []
        for e, (lev, i, j) in enumerate(order):
            for f, (di, dj) in enumerate(((-1, 0), (1, 0), (0, -1), (0, 1))):
                c = self._wrap(lev, i + di, j + dj)
                if c is None:
                    continue
                if c in index:
                    nbr[e, f] = index[c]
                    continue
                cov = self._covering(leaves, c)
                if cov is not None:  # a coarser neighbour: the small side of its mortar
                    if cov[0] != lev - 1:
                        raise ValueError("TreeMesh leaves are not 2:1 balanced")
                    nbr[e, f] = -2
                    continue
                # two finer neighbours: this element is the large side of a mortar
                _, ci, cj = c
                a = 1 if di < 0 else 0           # column of the neighbour's children next to the face
                b = 1 if dj < 0 else 0           # row of them
                if di != 0:
                    kids = [(lev + 1, 2 * ci + a, 2 * cj + q) for q in (0, 1)]
                else:
                    kids = [(lev + 1, 2 * ci + q, 2 * cj + b) for q in (0, 1)]
                if not all(k in index for k in kids):
                    raise ValueError("TreeMesh leaves are not 2:1 balanced")
                nbr[e, f] = -2
                mortars.append((index[kids[0]], index[kids[1]], e, 1 if di != 0 else 2, 1 if (di > 0 or dj > 0) else 2))
        self.neighbors = nbr
        self.mortars = np.array(mortars, dtype=np.int32).reshape(-1, 5)
        self.nx = self.ny = None

    @property
    def refined(self):
        return len(self.mortars) > 0 or hasattr(self, "leaves")

    def _leaf_set(self):
        if hasattr(self, "leaves"):
            return {tuple(int(q) for q in c) for c in self.leaves}
        return {(self.level, int(i), int(j)) for i, j in zip(self.ix, self.iy)}

    def _from_leaves(self, leaves):
        m = object.__new__(TreeMesh)
        m.periodicity, m.level = self.periodicity, self.level
        m.coordinates_min, m.coordinates_max, m.length = self.coordinates_min, self.coordinates_max, self.length
        m._set_leaves(leaves)
        return m

    def refine(self, elements):
        """AMR refinement (Trixi refine!): the TreeMesh with these elements split
        in four, then rebalanced, and the plan that carries a solution over
        (see ``transfer_plan``)."""
        old = [tuple(int(q) for q in c) for c in self._leaf_list()]
        new = self._from_leaves(self._balanced(self._split(set(old), {old[int(e)] for e in elements})))
        return new, transfer_plan(old, new)

    def coarsen(self, elements):
        """AMR coarsening (Trixi coarsen!): every set of four sibling leaves that are
        all listed becomes their parent, unless that breaks the 2:1 balance; the
        TreeMesh and the plan that carries a solution over."""
        old = [tuple(int(q) for q in c) for c in self._leaf_list()]
        leaves = set(old)
        marked = {old[int(e)] for e in elements}
        parents = {(lev - 1, i >> 1, j >> 1) for lev, i, j in marked if lev > 0}
        for p in sorted(parents):
            lev, i, j = p
            kids = {(lev + 1, 2 * i + a, 2 * j + b) for a in (0, 1) for b in (0, 1)}
            if not kids <= marked or not kids <= leaves:
                continue
            trial = (leaves - kids) | {p}
            if self._balanced(trial) == trial:
                leaves = trial
        new = self._from_leaves(leaves)
        return new, transfer_plan(old, new)

    def _leaf_list(self):
        if hasattr(self, "leaves"):
            return [tuple(int(q) for q in c) for c in self.leaves]
        return [(self.level, int(i), int(j)) for i, j in zip(self.ix, self.iy)]

    @property
    def ncells_per_dimension(self):
        return self.nx

    def node_coordinates(self, nodes):
        """x, y of every node, arrays [e, j, i] (Trixi's node_coordinates[:, i, j, e])."""
        half = 0.5 * self.element_length[:, None, None]
        x = self.element_center[:, 0, None, None] + half * nodes[None, None, :]
        y = self.element_center[:, 1, None, None] + half * nodes[None, :, None]
        return np.broadcast_to(x, (self.n_elements, len(nodes), len(nodes))).copy(), \
            np.broadcast_to(y, (self.n_elements, len(nodes), len(nodes))).copy()

    def lexicographic_permutation(self):
        """perm such that u_lex = u[perm] maps this element order to e = ix + nx*iy."""
        if hasattr(self, "leaves"):
            raise ValueError("a TreeMesh with refined cells has no lexicographic order")
        return self.element_of_cell.copy()


def transfer_plan(old_leaves, new_mesh):
    """Per element of new_mesh: (0, e_old, 0, 0, 0) for a kept leaf, (1, parent,
    child, 0, 0) for a child of a refined leaf (child 0 lower left, 1 lower right,
    2 upper left, 3 upper right) and (2, c0, c1, c2, c3) for the parent of four
    coarsened leaves (same child order); the layout tlw_transfer_solution takes."""
    index = {c: e for e, c in enumerate(old_leaves)}
    plan = np.zeros((new_mesh.n_elements, 5), dtype=np.int32)
    for e, c in enumerate(new_mesh.leaves):
        lev, i, j = (int(q) for q in c)
        if (lev, i, j) in index:
            plan[e] = (0, index[(lev, i, j)], 0, 0, 0)
        elif (lev - 1, i >> 1, j >> 1) in index:
            plan[e] = (1, index[(lev - 1, i >> 1, j >> 1)], (i & 1) + 2 * (j & 1), 0, 0)
        else:
            kids = [(lev + 1, 2 * i + a, 2 * j + b) for b in (0, 1) for a in (0, 1)]
            if not all(k in index for k in kids):
                raise ValueError("a new leaf is neither kept, a child nor a parent of old leaves")
            plan[e] = (2, *[index[k] for k in kids])
    return plan


def coordinates2mapping(coordinates_min, coordinates_max):
    """Trixi coordinates2mapping: the affine map of the reference square
    [-1, 1]² onto the box [min, max]."""
    x0, y0 = map(float, coordinates_min)
    x1, y1 = map(float, coordinates_max)

    def mapping(xi, eta):
        return x0 + 0.5 * (xi + 1.0) * (x1 - x0), y0 + 0.5 * (eta + 1.0) * (y1 - y0)

    return mapping


class StructuredMesh:
    """StructuredMesh(cells_per_dimension, mapping; periodicity=True), or
    StructuredMesh(cells_per_dimension, coordinates_min, coordinates_max; periodicity=True).

    ``mapping(xi, eta) -> (x, y)`` maps the global reference square [-1, 1]²
    onto the domain and must accept numpy arrays.  With
    ``ordering="lexicographic"`` (the default) elements are numbered
    e = cx + nx*cy like Trixi's linear indices; ``ordering="strips"`` numbers
    two-row strips column by column within a strip, so that a run of 10
    elements is a 5×2 block (14 of its 40 face slots leave the run, against 22
    for a 10×1 row).  ``lexicographic_permutation()`` maps the linear index
    cx + nx*cy to the element, and ``set_metrics`` takes its arrays in linear
    index order.  The node coordinates and
    metric terms follow Trixi's StructuredMesh cache (restated: the element's
    GLL nodes placed in its cell of the reference square and mapped;
    jacobian_matrix from the derivative matrix applied to the node
    coordinates; contravariant vectors Ja¹ = (y_η, -x_η), Ja² = (-y_ξ, x_ξ);
    inverse_jacobian = 1/(x_ξ y_η - x_η y_ξ)), which TrixiLW's StructuredMesh
    kernels read (src/solvers/dgsem_structured/dg_2d.jl:1379-2391).  A
    non-periodic direction takes the time-averaged Dirichlet condition of the
    semidiscretization's boundary value function.
    """

    curved = True

    def __init__(self, cells_per_dimension, mapping, coordinates_max=None, periodicity=True,
                 ordering="lexicographic"):
        nx, ny = (int(c) for c in cells_per_dimension)
        if nx < 1 or ny < 1:
            raise ValueError("cells_per_dimension must be positive")
        if coordinates_max is not None:
            mapping = coordinates2mapping(mapping, coordinates_max)
        if not callable(mapping):
            raise TypeError("mapping must be callable (or give coordinates_min, coordinates_max)")
        if isinstance(periodicity, bool):
            periodicity = (periodicity, periodicity)
        self.periodicity = tuple(bool(p) for p in periodicity)
        self.mapping = mapping
        self.nx, self.ny = nx, ny
        self.n_elements = nx * ny
        E = self.n_elements
        lin = np.arange(E)
        if ordering == "lexicographic":
            order = lin
        elif ordering == "strips":
            order = np.lexsort((lin // nx % 2, lin % nx, lin // nx // 2))
        else:
            raise ValueError(f"unknown ordering {ordering!r}")
        self.ordering = ordering
        self._order = order  # element -> linear index
        self._element_of = np.empty(E, dtype=np.int64)  # linear index -> element
        self._element_of[order] = lin
        self.ix, self.iy = order % nx, order // nx
        nbr = np.full((E, 4), -1, dtype=np.int32)
        px, py = self.periodicity
        for face, (dx, dy) in enumerate(((-1, 0), (1, 0), (0, -1), (0, 1))):
            jx, jy = self.ix + dx, self.iy + dy
            if px:
                jx = jx % nx
            if py:
                jy = jy % ny
            ok = (jx >= 0) & (jx < nx) & (jy >= 0) & (jy < ny)
            nbr[:, face] = np.where(ok, self._element_of[np.where(ok, jx + nx * jy, 0)], -1)
        self.neighbors = nbr
        self._metrics = {}

    @property
    def ncells_per_dimension(self):
        return (self.nx, self.ny)

    def node_coordinates(self, nodes):
        """x, y of every node, arrays [e, j, i] (Trixi's node_coordinates[:, i, j, e])."""
        nodes = np.asarray(nodes, dtype=np.float64)
        dx, dy = 2.0 / self.nx, 2.0 / self.ny
        ox = -1.0 + self.ix * dx + dx / 2
        oy = -1.0 + self.iy * dy + dy / 2
        xi = ox[:, None, None] + dx / 2 * nodes[None, None, :]
        eta = oy[:, None, None] + dy / 2 * nodes[None, :, None]
        xi, eta = np.broadcast_arrays(xi, eta)
        x, y = self.mapping(xi, eta)
        shape = (self.n_elements, len(nodes), len(nodes))
        return (np.ascontiguousarray(np.broadcast_to(x, shape), dtype=np.float64),
                np.ascontiguousarray(np.broadcast_to(y, shape), dtype=np.float64))

    def metrics(self, basis):
        """(contravariant_vectors [E, j, i, 4] with Ja11, Ja12, Ja21, Ja22 per node,
        inverse_jacobian [E, j, i]) for the basis' nodes and derivative matrix."""
        key = id(basis)
        if key not in self._metrics:
            D = np.asarray(basis.derivative_matrix, dtype=np.float64)
            x, y = self.node_coordinates(basis.nodes)
            x_xi = np.einsum("ik,ejk->eji", D, x)
            y_xi = np.einsum("ik,ejk->eji", D, y)
            x_eta = np.einsum("jk,eki->eji", D, x)
            y_eta = np.einsum("jk,eki->eji", D, y)
            Ja = np.stack([y_eta, -x_eta, -y_xi, x_xi], axis=-1)
            ijn = 1.0 / (x_xi * y_eta - x_eta * y_xi)
            self._metrics[key] = (np.ascontiguousarray(Ja), np.ascontiguousarray(ijn))
        return self._metrics[key]

    def set_metrics(self, basis, contravariant_vectors, inverse_jacobian, order="linear"):
        """Metric terms computed elsewhere (e.g. Trixi's cache.elements of the
        same mesh) for the basis, in the layouts metrics() returns, used in
        place of the host computation.  ``order`` names the element order of
        the given arrays: "linear" (cx + nx*cy, Trixi's StructuredMesh order)
        or "element" (this mesh's element order, as metrics() returns them)."""
        if order not in ("linear", "element"):
            raise ValueError("order must be 'linear' or 'element'")
        n = len(basis.nodes)
        Ja = np.asarray(contravariant_vectors, dtype=np.float64).reshape(self.n_elements, n, n, 4)
        ijn = np.asarray(inverse_jacobian, dtype=np.float64).reshape(self.n_elements, n, n)
        if order == "linear":
            Ja, ijn = Ja[self._order], ijn[self._order]
        self._metrics[id(basis)] = (np.ascontiguousarray(Ja), np.ascontiguousarray(ijn))

    def lexicographic_permutation(self):
        return self._element_of.copy()


class P4estMesh(StructuredMesh):
    """P4estMesh(trees_per_dimension, mapping; periodicity, turns): Trixi's
    P4estMesh of trees_per_dimension cells of the mapped reference square, one
    element per tree (no refinement).  ``turns`` gives each element's local
    frame as a multiple of 90° against its cell (all 0 for the aligned mesh Trixi
    builds from a mapping; other values are the orientations of an unstructured
    quadrilateral mesh, e.g. one read from an Abaqus file).  The node
    coordinates and metric terms are those of the element's own frame; the
    interfaces (primary element and face, secondary element and face, flipped
    node order) are found by matching face end points in the reference square,
    periodic directions wrapped.  Faces are numbered 0..3 = -ξ, +ξ, -η, +η of
    the element's frame; ``neighbors`` holds each interface face's partner and
    -1 on the domain boundary.  The device evaluates the reference's P4est
    surface terms (src/solvers/dgsem_p4est/dg_2d.jl:8-393)."""

    p4est = True

    def __init__(self, trees_per_dimension, mapping, coordinates_max=None, periodicity=True, turns=None):
        super().__init__(trees_per_dimension, mapping, coordinates_max, periodicity=periodicity)
        E = self.n_elements
        self.turns = np.zeros(E, dtype=np.int64) if turns is None else np.asarray(turns, dtype=np.int64) % 4
        if self.turns.shape != (E,):
            raise ValueError(f"turns must have one entry per element ({E})")
        self._connect()

    def _turned(self, a, b):
        """Local (ξ, η) arrays [E, ...] -> the cells' frame."""
        r = self.turns.reshape((-1,) + (1,) * (np.ndim(a) - 1))
        ca = np.select([r == 0, r == 1, r == 2, r == 3], [a, -b, -a, b])
        cb = np.select([r == 0, r == 1, r == 2, r == 3], [b, a, -b, -a])
        return ca, cb

    def _reference_xy(self, xi, eta):
        dx, dy = 2.0 / self.nx, 2.0 / self.ny
        ox = (-1.0 + self.ix * dx + dx / 2).reshape((-1,) + (1,) * (np.ndim(xi) - 1))
        oy = (-1.0 + self.iy * dy + dy / 2).reshape((-1,) + (1,) * (np.ndim(xi) - 1))
        a, b = self._turned(xi, eta)
        return ox + dx / 2 * a, oy + dy / 2 * b

    def node_coordinates(self, nodes):
        """x, y of every node in the element's own frame, arrays [e, j, i]."""
        nodes = np.asarray(nodes, dtype=np.float64)
        n, E = len(nodes), self.n_elements
        xi = np.broadcast_to(nodes[None, None, :], (E, n, n))
        eta = np.broadcast_to(nodes[None, :, None], (E, n, n))
        rx, ry = self._reference_xy(xi, eta)
        x, y = self.mapping(rx, ry)
        return np.ascontiguousarray(x, dtype=np.float64), np.ascontiguousarray(y, dtype=np.float64)

    def _connect(self):
        E = self.n_elements
        # end points of each face (first and last face node, local coordinates)
        ends = {0: ((-1, -1), (-1, 1)), 1: ((1, -1), (1, 1)), 2: ((-1, -1), (1, -1)), 3: ((-1, 1), (1, 1))}
        tol = 1e-9

        def canon(v, periodic):
            if periodic and v > 1.0 - tol:
                v -= 2.0
            return round(v, 8) + 0.0

        table = {}
        nbr = np.full((E, 4), -1, dtype=np.int32)
        for f in range(4):
            (a0, b0), (a1, b1) = ends[f]
            p0 = self._reference_xy(np.full(E, float(a0)), np.full(E, float(b0)))
            p1 = self._reference_xy(np.full(E, float(a1)), np.full(E, float(b1)))
            for e in range(E):
                q0 = (canon(p0[0][e], self.periodicity[0]), canon(p0[1][e], self.periodicity[1]))
                q1 = (canon(p1[0][e], self.periodicity[0]), canon(p1[1][e], self.periodicity[1]))
                table.setdefault(frozenset((q0, q1)), []).append((e, f, q0))
        interfaces = []
        for sides in table.values():
            if len(sides) == 1:
                continue
            if len(sides) != 2:
                raise ValueError("a face is shared by more than two elements")
            (e1, f1, s1), (e2, f2, s2) = sorted(sides)
            nbr[e1, f1], nbr[e2, f2] = e2, e1
            interfaces.append((e1, f1, e2, f2, 0 if s1 == s2 else 1))
        self.neighbors = nbr
        self.interfaces = np.array(sorted(interfaces), dtype=np.int32).reshape(-1, 5)

    def lexicographic_permutation(self):
        return np.arange(self.n_elements)
