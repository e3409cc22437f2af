"""Uniformly refined 2-D TreeMesh on the host.

Mirrors what TrixiLW reads from Trixi's TreeMesh and its containers
(`cache.elements.inverse_jacobian`, `node_coordinates`, `cache.interfaces`,
`cache.boundaries`; src/solvers/dgsem_tree/dg_2d.jl:1269-1454).  The device
only sees a per-element table of the four face neighbours, element centres
and lengths, so any element order works; ``ordering="morton"`` numbers the
leaves in Z-order like Trixi's uniformly refined tree, ``"lexicographic"``
uses e = ix + nx*iy.
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
