"""Uniformly refined 2-D TreeMesh and curvilinear StructuredMesh on the host.

Mirrors what TrixiLW reads from Trixi's TreeMesh and its containers
(`cache.elements.inverse_jacobian`, `node_coordinates`, `cache.interfaces`,
`cache.boundaries`; src/solvers/dgsem_tree/dg_2d.jl:1269-1454).  The device
only sees a per-element table of the four face neighbours, element centres
and lengths, so any element order works; ``ordering="morton"`` numbers the
leaves in Z-order like Trixi's uniformly refined tree, ``"lexicographic"``
uses e = ix + nx*iy, and ``"strips"`` walks two-row strips column by column
(runs of 10 elements are 5×2 blocks: the volume kernel's 10-element tiles then
have fewer faces that leave the tile, so fewer face records cross HBM).
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
    """TreeMesh(coordinates_min, coordinates_max, initial_refinement_level, periodicity)."""

    def __init__(self, coordinates_min, coordinates_max, initial_refinement_level, periodicity=True,
                 ordering="morton"):
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
        return self.element_of_cell.copy()


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
