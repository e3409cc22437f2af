"""Batches of independent semidiscretizations in one device context.

An ensemble of instances (each its own TreeMesh and initial condition, same
polynomial degree, equations and solver) is advanced by one launch per kernel
and step (tlw_create_batch); every instance keeps its own t, CFL dt and
iteration count, exactly as if it ran alone through solve_lwfr.  Steps can be
replayed from a captured HIP graph.  The reference runs one semidiscretization
per solve_lwfr call; this is the batched form of that loop.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .semidiscretization import MDRK, SemidiscretizationHyperbolic, _dptr


def _torch():
    import torch
    return torch


class SemidiscretizationBatch:
    """SemidiscretizationBatch(meshes, time_discretization, equations, initial_conditions, solver).

    ``meshes`` and ``initial_conditions`` are lists of equal length; all
    instances share ``equations``, ``solver`` (basis, volume integral) and the
    boundary treatment (periodic, or the given boundary value function).
    """

    def __init__(self, meshes, time_discretization, equations, initial_conditions, solver, boundary_conditions=None,
                 device=None):
        torch = _torch()
        if len(meshes) != len(initial_conditions) or not meshes:
            raise ValueError("one initial condition per mesh")
        # the solver description is built exactly as for a single instance
        first = SemidiscretizationHyperbolic(meshes[0], time_discretization, equations, initial_conditions[0], solver,
                                             boundary_conditions=boundary_conditions, device=device)
        self._proto = first
        self.meshes = list(meshes)
        self.initial_conditions = list(initial_conditions)
        self.solver = solver
        self.equations = equations
        self.time_discretization = time_discretization
        self.device = first.device
        self.n = first.n
        self._keep = []
        descs = (_lib.MeshDesc * len(meshes))()
        for k, m in enumerate(meshes):
            arrs = first._mesh_arrays(m)
            self._keep.append(arrs)
            descs[k] = _lib.MeshDesc(int(m.n_elements), arrs["nbr"].ctypes.data_as(C.POINTER(C.c_int32)),
                                     _dptr(arrs["invj"]), _dptr(arrs["center"]), _dptr(arrs["length"]))
        b = solver.basis
        self._basis_arrays = (np.ascontiguousarray(b.nodes), np.ascontiguousarray(b.weights),
                              np.asfortranarray(b.derivative_matrix).ravel(order="F").copy(),
                              np.asfortranarray(b.derivative_dhat).ravel(order="F").copy())
        basis = _lib.BasisDesc(b.polydeg, *[_dptr(a) for a in self._basis_arrays])
        h = C.c_void_p()
        rc = _lib.lib().tlw_create_batch(len(meshes), descs, C.byref(basis), C.byref(first._solver_desc), C.byref(h))
        if rc != _lib.TLW_OK:
            raise _lib.TLWError(rc, "tlw_create_batch failed (invalid description or HIP error)")
        self._h = h
        first.close()
        self.sizes = [m.n_elements for m in meshes]
        self.offsets = np.concatenate([[0], np.cumsum(self.sizes)]).astype(int)

    @property
    def handle(self):
        return self._h

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _lib.lib().tlw_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def ndofs(self):
        return sum(self.sizes) * self.n * self.n

    def instance_ndofs(self, k):
        return self.sizes[k] * self.n * self.n

    def compute_coefficients(self, t=0.0):
        """Initial states of all instances, concatenated, as a host array (E_total, n, n, 4)."""
        parts = []
        for m, ic in zip(self.meshes, self.initial_conditions):
            x, y = m.node_coordinates(self.solver.basis.nodes)
            parts.append(np.ascontiguousarray(ic(x, y, t), dtype=np.float64))
        return np.concatenate(parts)

    def split(self, u):
        """Per-instance views of a concatenated state."""
        return [u[self.offsets[k]:self.offsets[k + 1]] for k in range(len(self.sizes))]

    def set_stream(self, stream):
        _lib.check(_lib.lib().tlw_set_stream(self._h, C.c_void_p(stream.cuda_stream)), self._h)

    def solve(self, u0, cfl, nsteps, t0=0.0, t_final=float("inf"), chunk=None, graphs=False, stream=None):
        """solve_lwfr's CFL loop for every instance: nsteps steps (or until every
        instance reaches t_final).  Returns (u, status) with u the concatenated
        final states (host array) and status a dict of per-instance t, dt, iter."""
        torch = _torch()
        L = _lib.lib()
        h = self._h
        if stream is not None:
            self.set_stream(stream)
        elif graphs:
            raise ValueError("graph replay needs an explicit stream")
        u = np.ascontiguousarray(u0, dtype=np.float64)
        dp = C.POINTER(C.c_double)
        _lib.check(L.tlw_enable_graphs(h, int(bool(graphs))), h)
        _lib.check(L.tlw_set_state(h, u.ctypes.data_as(dp), u.size), h)
        _lib.check(L.tlw_solve_init(h, t0, t_final, cfl), h)
        t, dt = C.c_double(), C.c_double()
        it, fin = C.c_int64(), C.c_int32()
        chunk = chunk or nsteps
        done = 0
        while done < nsteps:
            k = min(chunk, nsteps - done)
            _lib.check(L.tlw_advance(h, k, C.byref(t), C.byref(dt), C.byref(it), C.byref(fin)), h)
            done += k
            if fin.value:
                break
        out = np.empty_like(u)
        _lib.check(L.tlw_get_state(h, out.ctypes.data_as(dp), out.size), h)
        return out, self.status()

    def advance(self, nsteps):
        """Advance an initialised batch by nsteps (no host synchronisation inside)."""
        L = _lib.lib()
        t, dt = C.c_double(), C.c_double()
        it, fin = C.c_int64(), C.c_int32()
        _lib.check(L.tlw_advance(self._h, int(nsteps), C.byref(t), C.byref(dt), C.byref(it), C.byref(fin)), self._h)

    def status(self):
        n = len(self.sizes)
        t, dt = np.zeros(n), np.zeros(n)
        it = np.zeros(n, dtype=np.int64)
        fin = np.zeros(n, dtype=np.int32)
        _lib.check(_lib.lib().tlw_batch_status(self._h, n, _dptr(t), _dptr(dt),
                                               it.ctypes.data_as(C.POINTER(C.c_int64)),
                                               fin.ctypes.data_as(C.POINTER(C.c_int32))), self._h)
        return {"t": t, "dt": dt, "iter": it, "finished": fin.astype(bool)}


class SemidiscretizationEnsemble:
    """Uneven instances of two polynomial degrees in one set of launches
    (tlw_create_ensemble): ``batches`` is a SemidiscretizationBatch of polydeg 3
    and one of polydeg 4 (plain LW, periodic); every step of ``solve`` /
    ``advance`` is one volume, one surface and one finalize launch for all
    instances of both, optionally replayed from a captured HIP graph.  The
    batches keep their states and statuses."""

    TILES = ((16, 10), (8, 5))  # (polydeg 3, polydeg 4) elements per volume workgroup, equal block sizes

    def __init__(self, batches):
        if len(batches) != 2:
            raise ValueError("an ensemble joins one polydeg-3 and one polydeg-4 batch")
        self.batches = sorted(batches, key=lambda b: b.n)
        h = C.c_void_p()
        L = _lib.lib()
        rc = L.tlw_create_ensemble(self.batches[0].handle, self.batches[1].handle, C.byref(h))
        if rc != _lib.TLW_OK:
            raise _lib.TLWError(rc, L.tlw_last_error(None).decode())
        self._h = h
        self.set_tiles(*self.TILES[0])

    def _check(self, rc):
        if rc != _lib.TLW_OK:
            msg = _lib.lib().tlw_ensemble_last_error(self._h).decode()
            raise (_lib.NonAdmissibleState if rc == _lib.TLW_EADMISSIBLE else _lib.TLWError)(rc, msg)

    def set_tiles(self, eb3, eb4):
        if (eb3, eb4) not in self.TILES:
            raise ValueError(f"tiles must be one of {self.TILES}")
        L = _lib.lib()
        for b, eb in zip(self.batches, (eb3, eb4)):
            _lib.check(L.tlw_set_elements_per_workgroup(b.handle, int(eb)), b.handle)
        self.tiles = (eb3, eb4)

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _lib.lib().tlw_ensemble_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def ndofs(self):
        return sum(b.ndofs() for b in self.batches)

    def init(self, u0s, cfls, t0=0.0, t_final=float("inf"), stream=None, graphs=False):
        """Loads the batches' initial states (host arrays, polydeg 3 first) and
        runs tlw_ensemble_solve_init with a CFL number per batch."""
        L = _lib.lib()
        if stream is not None:
            self._check(L.tlw_ensemble_set_stream(self._h, C.c_void_p(stream.cuda_stream)))
        elif graphs:
            raise ValueError("graph replay needs an explicit stream")
        self._check(L.tlw_ensemble_enable_graphs(self._h, int(bool(graphs))))
        dp = C.POINTER(C.c_double)
        for b, u0 in zip(self.batches, u0s):
            u = np.ascontiguousarray(u0, dtype=np.float64)
            _lib.check(L.tlw_set_state(b.handle, u.ctypes.data_as(dp), u.size), b.handle)
        self._check(L.tlw_ensemble_solve_init(self._h, float(t0), float(t_final), float(cfls[0]), float(cfls[1])))

    def advance(self, nsteps):
        """nsteps steps of every instance; returns True when all are finished."""
        fin = C.c_int32()
        self._check(_lib.lib().tlw_ensemble_advance(self._h, int(nsteps), C.byref(fin)))
        return bool(fin.value)

    def states(self):
        """The batches' current states (host arrays, polydeg 3 first)."""
        L = _lib.lib()
        out = []
        for b in self.batches:
            u = np.empty((sum(b.sizes), b.n, b.n, 4))
            _lib.check(L.tlw_get_state(b.handle, u.ctypes.data_as(C.POINTER(C.c_double)), u.size), b.handle)
            out.append(u)
        return out

    def solve(self, u0s, cfls, nsteps, stream=None, graphs=False, chunk=None):
        self.init(u0s, cfls, stream=stream, graphs=graphs)
        chunk = chunk or nsteps
        done = 0
        while done < nsteps:
            k = min(chunk, nsteps - done)
            if self.advance(k):
                break
            done += k
        return self.states(), [b.status() for b in self.batches]


def draw_vortex_ensemble(n_instances=64, seed=2024):
    """The batching configuration's draw with numpy default_rng(seed): per
    instance a TreeMesh level in {4, 5, 6, 7} (16² .. 128² elements), polydeg
    in {3, 4}, Mach number U[0.2, 0.8] and vortex strength U[2, 5], in this
    order of draws.  Returns (levels, polydegs, machs, strengths) as arrays."""
    rng = np.random.default_rng(seed)
    levels = rng.integers(4, 8, n_instances)
    degs = rng.choice([3, 4], n_instances)
    machs = rng.uniform(0.2, 0.8, n_instances)
    strengths = rng.uniform(2.0, 5.0, n_instances)
    return levels, degs, machs, strengths
