"""ctypes binding of the CPU oracle (oracle/liboracle_tlw.so).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of ``bench.py`` as the checker; the product path never
imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle_tlw.so"

BV_NONE, BV_VORTEX, BV_CONSTANT, BV_RIEMANN4, BV_SLIP_WALL, BV_CALLBACK, BV_CONVERGENCE_TEST = 0, 1, 2, 3, 4, 5, 6
BV_OUTFLOW, BV_SLIP_WALL_APPROX = 7, 8
SRC_NONE, SRC_CONVERGENCE_TEST, SRC_CALLBACK, SRC_NS_CONVERGENCE = 0, 1, 2, 3
BC_FN = C.CFUNCTYPE(None, C.c_double, C.c_double, C.c_double, C.POINTER(C.c_double))
SRC_FN = C.CFUNCTYPE(None, C.POINTER(C.c_double), C.c_double, C.c_double, C.c_double, C.POINTER(C.c_double))


class _Desc(C.Structure):
    _fields_ = [
        ("polydeg", C.c_int),
        ("nx", C.c_int),
        ("ny", C.c_int),
        ("xmin", C.c_double),
        ("ymin", C.c_double),
        ("h", C.c_double),
        ("periodic_x", C.c_int),
        ("periodic_y", C.c_int),
        ("equation", C.c_int),
        ("gamma", C.c_double),
        ("adv_a1", C.c_double),
        ("adv_a2", C.c_double),
        ("flux_order", C.c_int),
        ("bc_kind", C.c_int),
        ("bc_par", C.c_double * 24),
        ("abstol", C.c_double),
        ("reltol", C.c_double),
        ("nthreads", C.c_int),
        ("shock_capturing", C.c_int),
        ("alpha_max", C.c_double),
        ("alpha_min", C.c_double),
        ("alpha_smooth", C.c_int),
        ("positivity", C.c_int),
        ("pos_threshold", C.c_double * 2),
        ("parabolic", C.c_int),
        ("visc_mu", C.c_double),
        ("prandtl", C.c_double),
        ("curved", C.c_int),
        ("map_kind", C.c_int),
        ("map_par", C.c_double * 4),
        ("reconstruction", C.c_int),
        ("ns_walls", C.c_int),
        ("wall_v", C.c_double * 8),
        ("wall_iso", C.c_int * 4),
        ("wall_T", C.c_double * 4),
        ("bc_fn", C.c_void_p),
        ("source_kind", C.c_int),
        ("src_fn", C.c_void_p),
        ("n_tree", C.c_int),
        ("tree_cx", C.c_void_p),
        ("tree_cy", C.c_void_p),
        ("tree_len", C.c_void_p),
        ("tree_nbr", C.c_void_p),
        ("n_mortars", C.c_int),
        ("mortars", C.c_void_p),
        ("p4est", C.c_int),
        ("rot", C.c_void_p),
    ]


def build(force: bool = False) -> Path:
    """Compile the oracle with its Makefile (g++, OpenMP)."""
    if force or not LIB_PATH.exists():
        subprocess.run(["make", "-C", str(HERE), "-B" if force else "all"], check=True,
                       stdout=subprocess.DEVNULL)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = C.CDLL(str(LIB_PATH))
        dp = C.POINTER(C.c_double)
        L.tlwo_create.restype = C.c_void_p
        L.tlwo_create.argtypes = [C.POINTER(_Desc)]
        L.tlwo_destroy.argtypes = [C.c_void_p]
        L.tlwo_set_threads.argtypes = [C.c_int]
        L.tlwo_max_threads.restype = C.c_int
        L.tlwo_ndofs.restype = C.c_longlong
        L.tlwo_ndofs.argtypes = [C.c_void_p]
        L.tlwo_basis.argtypes = [C.c_void_p, dp, dp, dp, dp]
        L.tlwo_mortar_matrices.argtypes = [C.c_void_p, dp, dp, dp, dp]
        L.tlwo_p4est_connectivity.restype = C.c_int
        L.tlwo_p4est_connectivity.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(C.c_int)]
        L.tlwo_node_coords.argtypes = [C.c_void_p, dp, dp]
        L.tlwo_neighbors.argtypes = [C.c_void_p, C.POINTER(C.c_int)]
        L.tlwo_max_dt.restype = C.c_double
        L.tlwo_max_dt.argtypes = [C.c_void_p, dp]
        for nm in ("tlwo_rhs", "tlwo_rhs_mdrk1", "tlwo_rhs_mdrk2"):
            getattr(L, nm).argtypes = [C.c_void_p, dp, dp, C.c_double, C.c_double]
        L.tlwo_set_us.argtypes = [C.c_void_p, dp]
        L.tlwo_get_aux.restype = C.c_longlong
        L.tlwo_get_aux.argtypes = [C.c_void_p, C.c_int, dp]
        L.tlwo_euler_fluxes.argtypes = [dp, C.c_double, dp, dp]
        L.tlwo_set_alpha.argtypes = [C.c_void_p, dp]
        L.tlwo_positivity_limiter.argtypes = [C.c_void_p, dp]
        L.tlwo_solve_fixed_dt.argtypes = [C.c_void_p, dp, C.c_double, dp, C.c_int]
        L.tlwo_tangent.restype = C.c_int
        L.tlwo_tangent.argtypes = [C.POINTER(_Desc), dp, dp, dp, C.c_int, C.c_int, dp, dp, dp]
        L.tlwo_tangent_lw.restype = C.c_int
        L.tlwo_tangent_lw.argtypes = [C.POINTER(_Desc), dp, dp, dp, C.c_int, dp, dp]
        L.tlwo_inverse_vandermonde.argtypes = [C.c_void_p, dp]
        L.tlwo_max_abs_speed_naive.restype = C.c_double
        L.tlwo_max_abs_speed_naive.argtypes = [dp, dp, C.c_int, C.c_double]
        L.tlwo_ns_viscous_fluxes.argtypes = [dp, dp, dp, C.c_double, C.c_double, C.c_double, dp, dp]
        L.tlwo_ns_convergence.argtypes = [C.c_int, dp, dp, C.c_double, C.c_double, C.c_double, C.c_double, dp, dp]
        L.tlwo_solve_adaptive.restype = C.c_int
        L.tlwo_solve_adaptive.argtypes = [C.c_void_p, dp, C.c_double, C.c_double, C.c_double, C.c_int, dp, dp,
                                          C.POINTER(C.c_longlong)]
        L.tlwo_solve.restype = C.c_int
        L.tlwo_solve.argtypes = [C.c_void_p, dp, C.c_double, C.c_double, C.c_double, C.c_int, C.c_int, dp, dp]
        _lib = L
    return _lib


def _ptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.POINTER(C.c_double))


AUX = {"U": 0, "F": 1, "temporal_errors": 2, "surface_flux_values": 3, "us": 4, "U2": 5, "F2": 6, "alpha": 7,
       "fn_low": 8, "gradients_x": 9, "gradients_y": 10, "Fv": 11, "Fv2": 12, "contravariant_vectors": 13,
       "inverse_jacobian": 14}
MAP_AFFINE, MAP_WARP, MAP_PARALLELOGRAM = 0, 1, 2
PARABOLIC = {None: 0, "laplace": 1, "navier_stokes": 2}


class Oracle:
    """One oracle solver instance on a uniform TreeMesh (elements e = ix + nx*iy)."""

    def __init__(self, polydeg, nx, ny, xmin, ymin, h, periodic=(True, True), equation="euler", gamma=1.4,
                 adv=(1.0, 1.0), flux_order=0, bc_kind=BV_NONE, bc_par=(), tolerances=(0.0, 0.0), nthreads=0,
                 shock_capturing=False, alpha_max=0.5, alpha_min=0.001, alpha_smooth=True, positivity=None,
                 parabolic=None, mu=0.0, prandtl=0.72, mapping=None, reconstruction="first_order",
                 ns_walls=None, wall_temperature=None, bc_fn=None, source_terms=None, tree=None, p4est=None):
        """``positivity``: (threshold_density, threshold_pressure) of a Zhang-Shu
        limiter that ``solve`` applies as a stage limiter, or None.
        ``parabolic``: None, "laplace" (LaplaceDiffusion2D with diffusivity ``mu``,
        linear advection) or "navier_stokes" (CompressibleNavierStokesDiffusion2D,
        dynamic viscosity ``mu``, Prandtl number ``prandtl``, conservative
        gradient variables); periodic meshes only.
        ``ns_walls``: with parabolic="navier_stokes" and bc_kind=BV_SLIP_WALL, no-slip
        adiabatic walls on every boundary face with the wall velocities (v1, v2) of
        x_neg, x_pos, y_neg, y_pos (LW).
        ``mapping``: None (TreeMesh) or (MAP_AFFINE, (x0, y0, Lx, Ly)) /
        (MAP_WARP, (L, A)) for a periodic curvilinear StructuredMesh of nx × ny
        cells (xmin, ymin, h are then unused).
        ``reconstruction``: "first_order" or "muscl_hancock" subcell FV of the
        shock-capturing blend (Euler).
        ``bc_fn``: with bc_kind=BV_CALLBACK, the boundary value function
        f(x, y, t) -> 4 conservative values, called pointwise (Python).
        ``source_terms``: None, "convergence_test" (Trixi's
        source_terms_convergence_test) or a Python function s(u, x, y, t) -> 4
        values, called pointwise (Euler on the TreeMesh); or
        "navier_stokes_convergence_test" (the source of
        elixir_navierstokes_convergence.jl, with parabolic="navier_stokes" on
        the TreeMesh).
        ``tree``: a TreeMesh with refined cells, (cx, cy, h, neighbors [E, 4],
        mortars [M, 5]) as ``tree_mesh.connectivity`` gives them; nx, ny, xmin,
        ymin, h and periodic are then unused and elements keep the given order.
        ``p4est``: with ``mapping``, a P4estMesh of the same nx × ny cells: None for a
        StructuredMesh, True for aligned element frames, or an array of per-element
        frame turns (0..3, ·90°) for the orientations of an unstructured mesh."""
        d = _Desc()
        d.polydeg, d.nx, d.ny = polydeg, nx, ny
        d.xmin, d.ymin, d.h = xmin, ymin, h
        d.periodic_x, d.periodic_y = int(periodic[0]), int(periodic[1])
        d.equation = 0 if equation == "euler" else 1
        d.gamma = gamma
        d.adv_a1, d.adv_a2 = adv
        d.flux_order = flux_order
        d.bc_kind = bc_kind
        for k, v in enumerate(bc_par):
            d.bc_par[k] = v
        d.abstol, d.reltol = tolerances
        d.nthreads = nthreads
        d.shock_capturing = int(shock_capturing)
        d.alpha_max, d.alpha_min = alpha_max, alpha_min
        d.alpha_smooth = int(alpha_smooth)
        d.reconstruction = {"first_order": 0, "muscl_hancock": 1}[reconstruction]
        if positivity is not None:
            d.positivity = 1
            d.pos_threshold[0], d.pos_threshold[1] = positivity
        d.parabolic = PARABOLIC[parabolic]
        if ns_walls is not None:  # (v1, v2) of the walls x_neg, x_pos, y_neg, y_pos
            d.ns_walls = 1
            for k, (a, b) in enumerate(ns_walls):
                d.wall_v[2 * k], d.wall_v[2 * k + 1] = float(a), float(b)
            for k, T in enumerate(wall_temperature or (None,) * 4):  # None: adiabatic
                if T is not None:
                    d.wall_iso[k], d.wall_T[k] = 1, float(T)
        d.visc_mu, d.prandtl = mu, prandtl
        if mapping is not None:
            d.curved = 1
            d.map_kind = int(mapping[0])
            for k, v in enumerate(mapping[1]):
                d.map_par[k] = v
        if bc_fn is not None:
            def _cb(x, y, t, out, f=bc_fn):
                v = f(x, y, t)
                for q in range(4):
                    out[q] = float(v[q])

            self._bc_cb = BC_FN(_cb)  # kept alive with the oracle
            d.bc_fn = C.cast(self._bc_cb, C.c_void_p)
        if source_terms == "convergence_test":
            d.source_kind = SRC_CONVERGENCE_TEST
        elif source_terms == "navier_stokes_convergence_test":
            d.source_kind = SRC_NS_CONVERGENCE
        elif source_terms is not None:
            def _scb(u, x, y, t, out, f=source_terms):
                v = f(np.array([u[0], u[1], u[2], u[3]]), x, y, t)
                for q in range(4):
                    out[q] = float(v[q])

            self._src_cb = SRC_FN(_scb)
            d.source_kind = SRC_CALLBACK
            d.src_fn = C.cast(self._src_cb, C.c_void_p)
        if tree is not None:
            tcx, tcy, th, tnbr, tmort = tree
            self._tree = (np.ascontiguousarray(tcx, dtype=np.float64), np.ascontiguousarray(tcy, dtype=np.float64),
                          np.ascontiguousarray(th, dtype=np.float64), np.ascontiguousarray(tnbr, dtype=np.int32),
                          np.ascontiguousarray(tmort, dtype=np.int32).reshape(-1, 5))
            d.n_tree = len(self._tree[0])
            d.tree_cx, d.tree_cy, d.tree_len, d.tree_nbr = (a.ctypes.data for a in self._tree[:4])
            d.n_mortars = len(self._tree[4])
            d.mortars = self._tree[4].ctypes.data if d.n_mortars else None
            nx, ny = d.n_tree, 1
        if p4est is not None:
            d.p4est = 1
            if p4est is not True:
                self._rot = np.ascontiguousarray(p4est, dtype=np.int32).reshape(-1)
                d.rot = self._rot.ctypes.data
        self._desc = d
        self._h = lib().tlwo_create(C.byref(d))
        if not self._h:
            raise ValueError("unsupported oracle configuration (polydeg 3 and 4; see tlwo_create)")
        self.n = polydeg + 1
        self.nv = 4 if equation == "euler" else 1
        self.E = nx * ny
        self.nx, self.ny = nx, ny

    def __del__(self):
        if getattr(self, "_h", None):
            lib().tlwo_destroy(self._h)
            self._h = None

    @property
    def shape(self):
        return (self.E, self.n, self.n, self.nv)

    def p4est_connectivity(self):
        """P4est interfaces [I, 5] (primary, primary face, secondary, secondary face,
        flipped) and boundary faces [B, 2] (element, face), faces 0-based -ξ, +ξ, -η, +η."""
        nb = C.c_int(0)
        ni = lib().tlwo_p4est_connectivity(self._h, None, None, C.byref(nb))
        if ni < 0:
            raise ValueError("not a P4est oracle")
        itf = np.zeros((ni, 5), dtype=np.int32)
        bd = np.zeros((nb.value, 2), dtype=np.int32)
        lib().tlwo_p4est_connectivity(self._h, itf.ctypes.data if ni else None, bd.ctypes.data if nb.value else None,
                                      C.byref(nb))
        return itf, bd

    def mortar_matrices(self):
        """forward_lower, forward_upper, reverse_lower, reverse_upper ([n, n], row i = output node)."""
        n = self.n
        m = [np.zeros((n, n)) for _ in range(4)]
        lib().tlwo_mortar_matrices(self._h, *[_ptr(a) for a in m])
        return tuple(m)

    def basis(self):
        n = self.n
        nodes, w = np.zeros(n), np.zeros(n)
        D, Dh = np.zeros((n, n)), np.zeros((n, n))
        lib().tlwo_basis(self._h, _ptr(nodes), _ptr(w), _ptr(D), _ptr(Dh))
        return nodes, w, D, Dh

    def node_coords(self):
        x = np.zeros((self.E, self.n, self.n))
        y = np.zeros_like(x)
        lib().tlwo_node_coords(self._h, _ptr(x), _ptr(y))
        return x, y  # [e, j, i]

    def neighbors(self):
        nb = np.zeros((self.E, 4), dtype=np.int32)
        lib().tlwo_neighbors(self._h, nb.ctypes.data_as(C.POINTER(C.c_int)))
        return nb

    def max_dt(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        return lib().tlwo_max_dt(self._h, _ptr(u))

    def rhs(self, u, t, dt, kind="lw"):
        u = np.ascontiguousarray(u, dtype=np.float64)
        du = np.zeros_like(u)
        fn = {"lw": lib().tlwo_rhs, "mdrk1": lib().tlwo_rhs_mdrk1, "mdrk2": lib().tlwo_rhs_mdrk2}[kind]
        fn(self._h, _ptr(u), _ptr(du), t, dt)
        return du

    def set_alpha(self, alpha):
        """Prescribe the blending factors (None returns to the indicator)."""
        if alpha is None:
            lib().tlwo_set_alpha(self._h, None)
        else:
            a = np.ascontiguousarray(alpha, dtype=np.float64)
            lib().tlwo_set_alpha(self._h, _ptr(a))

    def positivity_limiter(self, u):
        """my_limiter_zhang_shu! on (density, pressure) with the thresholds given
        as ``positivity`` at construction; returns the limited copy."""
        out = np.array(u, dtype=np.float64, copy=True, order="C")
        lib().tlwo_positivity_limiter(self._h, _ptr(out))
        return out

    def inverse_vandermonde(self):
        out = np.zeros((self.n, self.n))
        lib().tlwo_inverse_vandermonde(self._h, _ptr(out))
        return out

    def solve_fixed_dt(self, u0, dts, t0=0.0):
        """The LW loop with the given dt per step; returns the final state."""
        u = np.array(u0, dtype=np.float64, copy=True, order="C")
        d = np.ascontiguousarray(dts, dtype=np.float64)
        lib().tlwo_solve_fixed_dt(self._h, _ptr(u), t0, _ptr(d), len(d))
        return u

    def tangent_lw(self, u0, du0, dts):
        """Forward-mode (dual-number) tangent of the fixed-dt LW loop:
        returns (u_K, du_K) for the seed du0."""
        u0 = np.ascontiguousarray(u0, dtype=np.float64)
        du0 = np.ascontiguousarray(du0, dtype=np.float64)
        d = np.ascontiguousarray(dts, dtype=np.float64)
        u, du = np.zeros_like(u0), np.zeros_like(u0)
        rc = lib().tlwo_tangent_lw(C.byref(self._desc), _ptr(u0), _ptr(du0), _ptr(d), len(d), _ptr(u), _ptr(du))
        if rc:
            raise ValueError("tangent needs polydeg 3 or 4")
        return u, du

    def tangent(self, u0, du0, dts, mdrk=False, alpha=None):
        """Forward-mode tangent of the fixed-dt LW or MDRK loop, optionally with
        the blending factors frozen at ``alpha`` [E]: returns (u_K, du_K)."""
        u0 = np.ascontiguousarray(u0, dtype=np.float64)
        du0 = np.ascontiguousarray(du0, dtype=np.float64)
        d = np.ascontiguousarray(dts, dtype=np.float64)
        a = None if alpha is None else np.ascontiguousarray(alpha, dtype=np.float64)
        u, du = np.zeros_like(u0), np.zeros_like(u0)
        rc = lib().tlwo_tangent(C.byref(self._desc), _ptr(u0), _ptr(du0), _ptr(d), len(d), int(mdrk),
                                None if a is None else _ptr(a), _ptr(u), _ptr(du))
        if rc:
            raise ValueError("tangent needs polydeg 3 or 4")
        return u, du

    def set_us(self, us):
        us = np.ascontiguousarray(us, dtype=np.float64)
        lib().tlwo_set_us(self._h, _ptr(us))

    def aux(self, name):
        k = AUX[name]
        m = lib().tlwo_get_aux(self._h, k, None)
        out = np.zeros(m)
        lib().tlwo_get_aux(self._h, k, _ptr(out))
        return out

    def solve(self, u0, t0, tf, cfl, max_steps=10**9, mdrk=False):
        """solve_lwfr with CFLBased stepping; returns (u, t, dts)."""
        u = np.array(u0, dtype=np.float64, copy=True, order="C")
        cap = max_steps if max_steps < 10**6 else 10**6
        dts = np.zeros(cap)
        tout = C.c_double(0.0)
        k = lib().tlwo_solve(self._h, _ptr(u), t0, tf, cfl, int(min(max_steps, cap)), int(mdrk), C.byref(tout),
                             _ptr(dts))
        return u, tout.value, dts[:k]


    def solve_adaptive(self, u0, t0, tf, dt_initial, max_attempts=10**6):
        """solve_lwfr with Adaptive() time stepping (LW; tolerances from the
        constructor); returns (u, t, accepted dts, rejected attempts)."""
        u = np.array(u0, dtype=np.float64, copy=True, order="C")
        dts = np.zeros(max_attempts)
        tout = C.c_double(0.0)
        nrej = C.c_longlong(0)
        k = lib().tlwo_solve_adaptive(self._h, _ptr(u), t0, tf, dt_initial, int(max_attempts), C.byref(tout),
                                      _ptr(dts), C.byref(nrej))
        return u, tout.value, dts[:k], int(nrej.value)


def euler_fluxes(u, gamma=1.4):
    u = np.ascontiguousarray(u, dtype=np.float64)
    f, g = np.zeros(4), np.zeros(4)
    lib().tlwo_euler_fluxes(_ptr(u), gamma, _ptr(f), _ptr(g))
    return f, g


def max_abs_speed_naive(ul, ur, orientation, gamma=1.4):
    ul = np.ascontiguousarray(ul, dtype=np.float64)
    ur = np.ascontiguousarray(ur, dtype=np.float64)
    return lib().tlwo_max_abs_speed_naive(_ptr(ul), _ptr(ur), int(orientation), gamma)


def ns_viscous_fluxes(u, gx, gy, gamma=1.4, mu=1.0, prandtl=0.72):
    """Viscous fluxes (x, y) of CompressibleNavierStokesDiffusion2D with
    conservative gradient variables at one node."""
    u, gx, gy = (np.ascontiguousarray(a, dtype=np.float64) for a in (u, gx, gy))
    fx, fy = np.zeros(4), np.zeros(4)
    lib().tlwo_ns_viscous_fluxes(_ptr(u), _ptr(gx), _ptr(gy), gamma, mu, prandtl, _ptr(fx), _ptr(fy))
    return fx, fy


def set_threads(n: int) -> None:
    lib().tlwo_set_threads(int(n))


def max_threads() -> int:
    return int(lib().tlwo_max_threads())


def ns_convergence(x, y, t, gamma=1.4, mu=0.01, prandtl=0.72):
    """The oracle's Navier-Stokes manufactured solution and its source at the
    points (x, y) and time t: (u [..., 4], s [..., 4])."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(np.broadcast_to(y, x.shape), dtype=np.float64)
    u = np.zeros(x.shape + (4,))
    s = np.zeros(x.shape + (4,))
    lib().tlwo_ns_convergence(x.size, _ptr(x), _ptr(y), float(t), float(gamma), float(mu), float(prandtl), _ptr(u),
                              _ptr(s))
    return u, s
