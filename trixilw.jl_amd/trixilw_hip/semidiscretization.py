"""Host mirror of TrixiLW's solver types and SemidiscretizationHyperbolic.

Names, arguments and semantics follow the reference:
  LW, MDRK, CFLBased                src/basic_types.jl:1-16
  VolumeIntegralFR, get_time_discretization   src/solvers/fr.jl:15-30, 81-84
  lw_fourier_cfl, trixi2lw          src/auxiliary/auxiliary.jl:14-23
  SemidiscretizationHyperbolic      src/semidiscretization/semidiscretization_hyperbolic.jl:58-84
  rhs!, rhs_mdrk1!, rhs_mdrk2!      src/semidiscretization/semidiscretization_hyperbolic.jl:4-56
  semidiscretize                    src/semidiscretization/semidiscretization.jl:4-14
The right-hand side itself runs in libtrixilw_hip.so; state vectors are
torch float64 tensors on the GPU in Trixi's [v, i, j, e] layout, i.e. shape
(E, n, n, 4) in row-major order.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass

import numpy as np

from . import _lib
from .basis import LobattoLegendreBasis
from .equations import CompressibleEulerEquations2D


class LW:
    def __repr__(self):
        return "Lax-Wendroff"


class MDRK:
    def __repr__(self):
        return "Multi-Derivative Runge-Kutta"


@dataclass
class CFLBased:
    cfl_number: float


class Adaptive:
    """Error-controlled time stepping (src/basic_types.jl:10, src/solvers/adaptive.jl)."""

    def __repr__(self):
        return "Adaptive"


@dataclass
class VolumeIntegralFR:
    time_discretization: object


flux_lax_friedrichs = "flux_lax_friedrichs"
density_pressure = "density_pressure"


class FirstOrderReconstruction:
    """Subcell FV reconstruction tag (src/solvers/fr.jl:58)."""


class MUSCLHancockReconstruction:
    """MUSCL-Hancock subcell reconstruction tag (src/solvers/fr.jl:58): minmod
    slopes with β = 2 - α, positivity-scaled face values, a half-step with the
    face fluxes and Rusanov fluxes between the evolved face values
    (calcflux_mh!, src/solvers/dgsem_structured/dg_2d.jl:313-538, on the
    TreeMesh's Cartesian metric).  Elements with a face on the domain
    boundary use first-order subcell FV, as in the reference."""


@dataclass
class IndicatorHennemannGassner:
    """IndicatorHennemannGassner(equations, basis; alpha_max, alpha_min, alpha_smooth, variable)."""

    equations: object = None
    basis: object = None
    alpha_max: float = 0.5
    alpha_min: float = 0.001
    alpha_smooth: bool = True
    variable: str = density_pressure


class VolumeIntegralFRShockCapturing:
    """VolumeIntegralFRShockCapturing(indicator; volume_integralFR, volume_flux_fv, reconstruction)
    (src/solvers/fr.jl:63-86)."""

    def __init__(self, indicator, volume_integralFR=None, volume_flux_fv=flux_lax_friedrichs,
                 reconstruction=None):
        self.indicator = indicator
        self.volume_integralFR = volume_integralFR if volume_integralFR is not None else VolumeIntegralFR(LW())
        self.volume_flux_fv = volume_flux_fv
        self.reconstruction = reconstruction if reconstruction is not None else FirstOrderReconstruction()
        if not isinstance(self.reconstruction, (FirstOrderReconstruction, MUSCLHancockReconstruction)):
            raise NotImplementedError("FirstOrderReconstruction or MUSCLHancockReconstruction")
        if volume_flux_fv != flux_lax_friedrichs or indicator.variable != density_pressure:
            raise NotImplementedError("flux_lax_friedrichs subcell fluxes and density_pressure only")

    @property
    def time_discretization(self):
        return self.volume_integralFR.time_discretization


class DGSEM:
    """DGSEM(polydeg, surface_flux=flux_lax_friedrichs, volume_integral=VolumeIntegralFR(LW()))."""

    def __init__(self, polydeg, surface_flux=flux_lax_friedrichs, volume_integral=None):
        if surface_flux != flux_lax_friedrichs:
            raise NotImplementedError("only flux_lax_friedrichs (Rusanov) is implemented")
        self.basis = LobattoLegendreBasis(polydeg)
        self.surface_flux = surface_flux
        self.volume_integral = volume_integral if volume_integral is not None else VolumeIntegralFR(LW())

    def polydeg(self):
        return self.basis.polydeg

    def nnodes(self):
        return self.basis.polydeg + 1


def get_time_discretization(solver_or_integral):
    vi = solver_or_integral.volume_integral if isinstance(solver_or_integral, DGSEM) else solver_or_integral
    return vi.time_discretization


def lw_fourier_cfl(N: int) -> float:
    return (1.000, 0.333, 0.170, 0.103)[N - 1]


def trixi2lw(cfl_number: float, dg: DGSEM) -> float:
    N = dg.polydeg()
    return cfl_number * (N + 1) * lw_fourier_cfl(N)


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _torch():
    import torch
    return torch


class SemidiscretizationHyperbolic:
    """SemidiscretizationHyperbolic(mesh, time_discretization, equations, initial_condition, solver;
    boundary_conditions=periodic).

    ``boundary_conditions`` is a boundary-value function object (IsentropicVortex,
    ConstantState, FourQuadrantRiemann) used as the time-averaged Dirichlet
    condition on every non-periodic face, or None for a fully periodic mesh.
    ``surface_flux_order`` selects the TreeMesh ("tree") or Structured/P4est
    ("structured") argument order of the LW Rusanov flux.

    On a curvilinear StructuredMesh (LW or MDRK, periodic or with the
    time-averaged Dirichlet boundary) the kernels of
    src/solvers/dgsem_structured/dg_2d.jl apply: contravariant fluxes, per-node
    inverse Jacobians, interface normals and the structured flux order.
    Shock capturing there uses the first-order subcell scheme with Trixi's
    subcell normals (MUSCL-Hancock, and alpha smoothing with boundaries, raise
    TLW_ENOTSUP).
    """

    def __init__(self, mesh, time_discretization, equations, initial_condition, solver, boundary_conditions=None,
                 source_terms=None, surface_flux_order="tree", keep_aux=False, device=None):
        src_kind = _lib.TLW_SRC_NONE
        if source_terms is not None:
            from .equations import SRC_CONVERGENCE_TEST, SRC_NS_CONVERGENCE_TEST

            kind = getattr(source_terms, "kind", None)
            if kind == SRC_CONVERGENCE_TEST:
                src_kind = _lib.TLW_SRC_CONVERGENCE_TEST
            elif kind == SRC_NS_CONVERGENCE_TEST:
                src_kind = _lib.TLW_SRC_NAVIER_STOKES_CONVERGENCE_TEST
            else:
                raise NotImplementedError("the device path implements source_terms_convergence_test and "
                                          "source_terms_navier_stokes_convergence_test")
        self.source_terms = source_terms
        if not isinstance(equations, CompressibleEulerEquations2D):
            raise NotImplementedError("the device path implements CompressibleEulerEquations2D")
        torch = _torch()
        self.mesh = mesh
        self.equations = equations
        self.initial_condition = initial_condition
        self.solver = solver
        self.time_discretization = time_discretization
        self.boundary_conditions = boundary_conditions
        self.device = torch.device(device) if device is not None else torch.device("cuda")
        self.n = solver.nnodes()
        self.cache = {"dt": [math.nan]}
        self.performance_counter = []
        b = solver.basis
        self._nodes = np.ascontiguousarray(b.nodes)
        self._weights = np.ascontiguousarray(b.weights)
        self._D = np.asfortranarray(b.derivative_matrix).ravel(order="F").copy()
        self._Dh = np.asfortranarray(b.derivative_dhat).ravel(order="F").copy()
        basis = _lib.BasisDesc(b.polydeg, _dptr(self._nodes), _dptr(self._weights), _dptr(self._D), _dptr(self._Dh))
        sd = _lib.SolverDesc()
        sd.time_discretization = _lib.TLW_MDRK if isinstance(time_discretization, MDRK) else _lib.TLW_LW
        sd.gamma = equations.gamma
        sd.surface_flux_order = {"tree": _lib.TLW_FLUX_TREE, "structured": _lib.TLW_FLUX_STRUCTURED}[
            surface_flux_order]
        from .equations import boundary_conditions_kind

        bcd = boundary_conditions_kind(boundary_conditions, getattr(mesh, "periodicity", None))
        if bcd is None:
            sd.boundary_kind = _lib.TLW_BV_NONE
        else:
            sd.boundary_kind = int(bcd.kind)
            for k, v in enumerate(bcd.params):
                sd.boundary_params[k] = float(v)
        sd.flags = _lib.TLW_FLAG_KEEP_AUX if keep_aux else 0
        sd.source_terms = src_kind
        vi = solver.volume_integral
        if isinstance(vi, VolumeIntegralFRShockCapturing):
            ind = vi.indicator
            self._vinv = np.asfortranarray(b.inverse_vandermonde_legendre).ravel(order="F").copy()
            sd.shock_capturing = 1
            sd.alpha_max, sd.alpha_min = ind.alpha_max, ind.alpha_min
            sd.alpha_smooth = int(bool(ind.alpha_smooth))
            sd.inverse_vandermonde_host = _dptr(self._vinv)
            sd.reconstruction = (_lib.TLW_RECON_MUSCL_HANCOCK if isinstance(vi.reconstruction, MUSCLHancockReconstruction)
                                 else _lib.TLW_RECON_FIRST_ORDER)
        self._solver_desc = sd
        self._keep = self._mesh_arrays(mesh)
        md = self._mesh_desc()
        h = C.c_void_p()
        torch.cuda.init()
        self._create(md, basis, sd, h)
        self._h = h
        self._setup_boundary_data()

    def _setup_boundary_data(self):
        """Caller-defined Dirichlet data: the boundary face-node coordinates in the
        library's table order, and the steady values uploaded once."""
        from .equations import BoundaryConditionDirichlet, boundary_conditions_kind

        bcd = boundary_conditions_kind(self.boundary_conditions, getattr(self.mesh, "periodicity", None))
        self._bc_table = bcd if isinstance(bcd, BoundaryConditionDirichlet) else None
        if self._bc_table is None:
            return
        L = _lib.lib()
        nb = C.c_int64()
        _lib.check(L.tlw_boundary_nodes(self._h, C.byref(nb), None, 0), self._h)
        xy = np.zeros((nb.value, 2))
        _lib.check(L.tlw_boundary_nodes(self._h, C.byref(nb), _dptr(xy), xy.size), self._h)
        self._bxy = xy
        if self._bc_table.steady and nb.value > 0:
            self._upload_boundary(np.ascontiguousarray(self._bc_table.values(xy[:, 0], xy[:, 1], 0.0,
                                                                             self.equations)), 1)

    def _upload_boundary(self, vals, levels):
        vals = np.ascontiguousarray(vals, dtype=np.float64)
        self._bind_stream()
        _lib.check(_lib.lib().tlw_set_boundary_data(self._h, _dptr(vals), vals.size, int(levels)), self._h)

    def _before_rhs(self, t, dt, stage):
        """Per-call boundary data for the rhs about to run: the function at the
        n GLL time levels of [t, t + s dt] (s = 0.5 for the first MDRK stage on a
        TreeMesh, the reference's scaling_factor) and at t."""
        bc = getattr(self, "_bc_table", None)
        if bc is None or bc.steady or len(self._bxy) == 0:
            return
        nodes = self.solver.basis.nodes
        n = len(nodes)
        half = not getattr(self.mesh, "curved", False) or getattr(self.mesh, "p4est", False)
        s = 0.5 if (stage == 1 and half) else 1.0
        dts = s * dt
        x, y = self._bxy[:, 0], self._bxy[:, 1]
        vals = np.empty((len(x), n + 1, 4))
        for q in range(n):
            vals[:, q] = bc.values(x, y, t + 0.5 * dts * (nodes[q] + 1.0), self.equations)
        vals[:, n] = bc.values(x, y, t, self.equations)
        self._upload_boundary(vals, n + 1)

    def _create(self, md, basis, sd, h):
        if getattr(self.mesh, "p4est", False):
            k = self._keep
            k["interfaces"] = np.ascontiguousarray(self.mesh.interfaces, dtype=np.int32).reshape(-1)
            pd = _lib.P4estDesc(_dptr(k["Ja"]), _dptr(k["ijn"]), _dptr(k["xy"]), int(len(self.mesh.interfaces)),
                                k["interfaces"].ctypes.data_as(C.POINTER(C.c_int32)))
            rc = _lib.lib().tlw_create_p4est(C.byref(md), C.byref(pd), C.byref(basis), C.byref(sd), C.byref(h))
            if rc != _lib.TLW_OK:
                raise _lib.TLWError(rc, "tlw_create_p4est failed (Euler, VolumeIntegralFR, a consistent interface "
                                        "list; boundary faces need a boundary condition)")
            return
        if getattr(self.mesh, "curved", False):
            k = self._keep
            cd = _lib.CurvedDesc(_dptr(k["Ja"]), _dptr(k["ijn"]), _dptr(k["xy"]))
            rc = _lib.lib().tlw_create_curved(C.byref(md), C.byref(basis), C.byref(sd), C.byref(cd), C.byref(h))
            if rc != _lib.TLW_OK:
                raise _lib.TLWError(rc, "tlw_create_curved failed (positive Jacobians and, with boundary "
                                        "faces, a boundary value function required; first-order shock "
                                        "capturing only)")
            return
        if getattr(self.mesh, "refined", False):
            mo = self._mortar_desc()
            rc = _lib.lib().tlw_create_nonconforming(C.byref(md), C.byref(mo), C.byref(basis), C.byref(sd),
                                                     C.byref(h))
            if rc != _lib.TLW_OK:
                raise _lib.TLWError(rc, "tlw_create_nonconforming failed (mortars: Euler, VolumeIntegralFR, "
                                        "a valid 2:1 mortar list)")
            return
        rc = _lib.lib().tlw_create(C.byref(md), C.byref(basis), C.byref(sd), C.byref(h))
        if rc != _lib.TLW_OK:
            raise _lib.TLWError(rc, "tlw_create failed (invalid description or HIP error)")

    def _mortar_desc(self):
        self._keep["mortars"] = np.ascontiguousarray(self.mesh.mortars, dtype=np.int32).reshape(-1)
        return _lib.MortarDesc(int(len(self.mesh.mortars)),
                               self._keep["mortars"].ctypes.data_as(C.POINTER(C.c_int32)))

    def _mesh_arrays(self, mesh):
        if getattr(mesh, "curved", False):
            Ja, ijn = mesh.metrics(self.solver.basis)
            x, y = mesh.node_coordinates(self.solver.basis.nodes)
            return {"nbr": np.ascontiguousarray(mesh.neighbors, dtype=np.int32), "Ja": Ja, "ijn": ijn,
                    "xy": np.ascontiguousarray(np.stack([x, y], axis=-1))}
        return {
            "nbr": np.ascontiguousarray(mesh.neighbors, dtype=np.int32),
            "invj": np.ascontiguousarray(mesh.inverse_jacobian, dtype=np.float64),
            "center": np.ascontiguousarray(mesh.element_center, dtype=np.float64),
            "length": np.ascontiguousarray(mesh.element_length, dtype=np.float64),
        }

    def _mesh_desc(self):
        k = self._keep
        if "Ja" in k:  # curvilinear: the TreeMesh element arrays are not read
            null = C.POINTER(C.c_double)()
            return _lib.MeshDesc(int(self.mesh.n_elements), k["nbr"].ctypes.data_as(C.POINTER(C.c_int32)), null,
                                 null, null)
        return _lib.MeshDesc(int(self.mesh.n_elements), k["nbr"].ctypes.data_as(C.POINTER(C.c_int32)),
                             _dptr(k["invj"]), _dptr(k["center"]), _dptr(k["length"]))

    def rebuild_mesh(self, mesh):
        """Host-side mesh change (AMR) followed by a device re-upload."""
        self.mesh = mesh
        self._keep = self._mesh_arrays(mesh)
        md = self._mesh_desc()
        if getattr(mesh, "refined", False):
            mo = self._mortar_desc()
            _lib.check(_lib.lib().tlw_rebuild_mesh_nonconforming(self._h, C.byref(md), C.byref(mo)), self._h)
        else:
            _lib.check(_lib.lib().tlw_rebuild_mesh(self._h, C.byref(md)), self._h)

    def adapt(self, u, new_mesh, plan):
        """One AMR step of the reference's callback (Trixi's refine!/coarsen! of the
        solution, then src/callbacks/amr.jl:2-208 resize!): the state u (device
        tensor of the current mesh) carried to new_mesh by ``plan`` (as
        TreeMesh.refine/coarsen return it) on the device, then the mesh
        re-uploaded.  Returns the new state tensor."""
        import torch

        plan = np.ascontiguousarray(plan, dtype=np.int32)
        n = self.solver.basis.nnodes()
        old_E = self.mesh.n_elements
        if tuple(u.shape) != (old_E, n, n, 4) or not u.is_cuda or u.dtype != torch.float64:
            raise ValueError(f"u must be a float64 device tensor of shape {(old_E, n, n, 4)}")
        u = u.contiguous()
        unew = torch.empty((new_mesh.n_elements, n, n, 4), dtype=torch.float64, device=u.device)
        torch.cuda.synchronize()
        _lib.check(_lib.lib().tlw_transfer_solution(self._h, C.c_void_p(u.data_ptr()), int(old_E),
                                                    C.c_void_p(unew.data_ptr()), int(new_mesh.n_elements),
                                                    plan.ctypes.data_as(C.POINTER(C.c_int32))), self._h)
        self.rebuild_mesh(new_mesh)
        return unew

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _lib.lib().tlw_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- helpers ---------------------------------------------------------------
    @property
    def handle(self):
        return self._h

    @property
    def state_shape(self):
        return (self.mesh.n_elements, self.n, self.n, 4)

    def ndofs(self):
        return self.mesh.n_elements * self.n * self.n

    def node_coordinates(self):
        return self.mesh.node_coordinates(self.solver.basis.nodes)

    def compute_coefficients(self, t):
        """Initial condition at the nodes as a device tensor (Trixi compute_coefficients)."""
        torch = _torch()
        x, y = self.node_coordinates()
        u = np.ascontiguousarray(self.initial_condition(x, y, t), dtype=np.float64)
        return torch.from_numpy(u).to(self.device)

    def _bind_stream(self):
        torch = _torch()
        s = torch.cuda.current_stream(self.device).cuda_stream
        _lib.check(_lib.lib().tlw_set_stream(self._h, C.c_void_p(s)), self._h)

    def _check_state(self, x, name):
        torch = _torch()
        if not isinstance(x, torch.Tensor) or x.dtype != torch.float64 or not x.is_cuda:
            raise TypeError(f"{name} must be a float64 CUDA tensor")
        if not x.is_contiguous() or x.numel() != self.ndofs() * 4:
            raise ValueError(f"{name} must be contiguous with {self.ndofs() * 4} entries")

    def max_dt(self, u_ode):
        """Trixi max_dt(u, t, mesh, constant_speed=False, equations, solver, cache)."""
        self._check_state(u_ode, "u")
        self._bind_stream()
        out = C.c_double(0.0)
        _lib.check(_lib.lib().tlw_max_dt(self._h, C.c_void_p(u_ode.data_ptr()), C.byref(out)), self._h)
        return out.value

    def us(self):
        """Device pointer of the MDRK stage buffer (element_cache.mdrk_cache._us)."""
        p = C.c_void_p()
        _lib.check(_lib.lib().tlw_us_ptr(self._h, C.byref(p)), self._h)
        return p.value

    def set_us(self, us_ode):
        """Stream-ordered copy of a device tensor into the MDRK stage buffer."""
        self._check_state(us_ode, "us")
        self._bind_stream()
        _lib.check(_lib.lib().tlw_set_us(self._h, C.c_void_p(us_ode.data_ptr())), self._h)

    def set_positivity_limiter(self, thresholds):
        """Zhang-Shu limiter of the device loop: (threshold_density,
        threshold_pressure), or None to turn it off."""
        if thresholds is None:
            _lib.check(_lib.lib().tlw_set_positivity_limiter(self._h, 0, 0.0, 0.0), self._h)
        else:
            _lib.check(_lib.lib().tlw_set_positivity_limiter(self._h, 1, float(thresholds[0]), float(thresholds[1])),
                       self._h)

    def apply_positivity_limiter(self, u_ode, thresholds):
        self._check_state(u_ode, "u")
        self._bind_stream()
        self.set_positivity_limiter(thresholds)
        _lib.check(_lib.lib().tlw_apply_positivity_limiter(self._h, C.c_void_p(u_ode.data_ptr())), self._h)

    def elements_per_workgroup(self, elements=None):
        """Tile of the fused volume kernel: returns it, or sets it when given."""
        if elements is not None:
            _lib.check(_lib.lib().tlw_set_elements_per_workgroup(self._h, int(elements)), self._h)
        out = C.c_int32()
        _lib.check(_lib.lib().tlw_get_elements_per_workgroup(self._h, C.byref(out)), self._h)
        return out.value

    def autotune(self, u0, reps=5):
        """Loads the state ``u0`` (host array or device tensor) into the
        context, times each compiled tile and keeps the fastest.  Returns
        (chosen tile, {tile: ms per rhs evaluation})."""
        u = u0.detach().cpu().numpy() if hasattr(u0, "detach") else np.asarray(u0)
        u = np.ascontiguousarray(u, dtype=np.float64)
        if u.size != self.ndofs() * 4:
            raise ValueError(f"state must have {self.ndofs() * 4} entries")
        self._bind_stream()
        _lib.check(_lib.lib().tlw_set_state(self._h, _dptr(u), u.size), self._h)
        chosen = C.c_int32()
        ms = (C.c_double * 3)()
        _lib.check(_lib.lib().tlw_autotune(self._h, int(reps), C.byref(chosen), ms), self._h)
        tiles = (5, 10, 15) if self.n == 5 else (4, 8, 16)
        return chosen.value, dict(zip(tiles, list(ms)))

    def get_aux(self, which):
        n = self.n
        E = self.mesh.n_elements
        if which == "U":
            out = np.zeros((E, n, n, 4))
            code = _lib.TLW_AUX_U
        elif which == "F":
            out = np.zeros((E, n, n, 2, 4))
            code = _lib.TLW_AUX_F
        elif which == "alpha":
            out = np.zeros(E)
            code = _lib.TLW_AUX_ALPHA
        else:
            raise ValueError(which)
        _lib.check(_lib.lib().tlw_get_aux(self._h, code, _dptr(out), out.size), self._h)
        return out


class SemidiscretizationHyperbolicParabolic(SemidiscretizationHyperbolic):
    """SemidiscretizationHyperbolicParabolic(mesh, time_discretization,
    (equations, equations_parabolic), initial_condition, solver;
    boundary_conditions=(periodic, periodic))
    (src/semidiscretization/semidiscretization_hyperbolic_parabolic.jl:8-118).

    ``equations_parabolic`` is a CompressibleNavierStokesDiffusion2D with
    GradientVariablesConservative.  The mesh is periodic (boundary conditions
    None), or a TreeMesh with walls: ``boundary_conditions=(bc_hyperbolic,
    bc_parabolic)`` with the reflecting walls for the hyperbolic part and four
    BoundaryConditionNavierStokesWall(NoSlip(...), Adiabatic(0.0)) for x_neg,
    x_pos, y_neg, y_pos (the reference's lid-driven cavity; LW).  rhs / rhs_mdrk1 / rhs_mdrk2, max_dt and the device loop
    then include the viscous terms (BR1 gradient, Taylor states of ∇u, central
    viscous interface flux, compute_dt's min of advective and viscous bounds).
    On a periodic curvilinear StructuredMesh the P4est parabolic kernels'
    curvilinear form applies (src/solvers/dgsem_p4est/dg_2d_parabolic.jl).
    """

    def __init__(self, mesh, time_discretization, equations, initial_condition, solver, boundary_conditions=None,
                 source_terms=None, keep_aux=False, device=None):
        from .equations import CompressibleNavierStokesDiffusion2D

        if not isinstance(equations, tuple) or len(equations) != 2:
            raise TypeError("equations must be the tuple (equations, equations_parabolic)")
        eq, eqp = equations
        if not isinstance(eqp, CompressibleNavierStokesDiffusion2D):
            raise NotImplementedError("the device path implements CompressibleNavierStokesDiffusion2D")
        from .equations import BoundaryConditionNavierStokesWall, boundary_condition_periodic

        bcs = boundary_conditions if boundary_conditions is not None else (None, None)
        bc_hyp, bc_par = bcs
        if isinstance(solver.volume_integral, VolumeIntegralFRShockCapturing):
            raise NotImplementedError("the reference's parabolic kernels take VolumeIntegralFR only")
        self.equations_parabolic = eqp
        self._par = _lib.ParabolicDesc(_lib.TLW_PAR_NAVIER_STOKES, eqp.mu, eqp.Pr, _lib.TLW_GRAD_CONSERVATIVE)
        if (getattr(source_terms, "mu", eqp.mu), getattr(source_terms, "prandtl", eqp.Pr)) != (eqp.mu, eqp.Pr):
            raise NotImplementedError("the Navier-Stokes convergence source takes the equations' mu and Prandtl")
        if bc_par is not None:
            # boundary_conditions_parabolic: the NamedTuple (x_neg, x_pos, y_neg, y_pos) of walls;
            # directions in which the mesh is periodic may name boundary_condition_periodic
            seq = [bc_par[k] for k in ("x_neg", "x_pos", "y_neg", "y_pos")] if isinstance(bc_par, dict) else bc_par
            per = tuple(getattr(mesh, "periodicity", (False, False)))
            seq = [BoundaryConditionNavierStokesWall.periodic() if (b is boundary_condition_periodic and per[k // 2])
                   else b for k, b in enumerate(seq)]
            if len(seq) != 4 or not all(isinstance(b, BoundaryConditionNavierStokesWall) for b in seq):
                raise NotImplementedError("parabolic boundary conditions: four BoundaryConditionNavierStokesWall")
            self._par.walls = 1
            for k, b in enumerate(seq):
                vel, temp = b.resolve(_wall_face_points(mesh, k, solver.basis.nodes))
                self._par.wall_velocity[2 * k], self._par.wall_velocity[2 * k + 1] = vel
                if temp is not None:
                    self._par.wall_isothermal[k], self._par.wall_temperature[k] = 1, temp
        elif bc_hyp is not None:
            raise NotImplementedError("the hyperbolic-parabolic device path needs a periodic mesh or walls")
        super().__init__(mesh, time_discretization, eq, initial_condition, solver, boundary_conditions=bc_hyp,
                         source_terms=source_terms, keep_aux=keep_aux, device=device)

    def _create(self, md, basis, sd, h):
        if getattr(self.mesh, "curved", False):
            k = self._keep
            cd = _lib.CurvedDesc(_dptr(k["Ja"]), _dptr(k["ijn"]), _dptr(k["xy"]))
            rc = _lib.lib().tlw_create_curved_parabolic(C.byref(md), C.byref(basis), C.byref(sd), C.byref(cd),
                                                        C.byref(self._par), C.byref(h))
            if rc != _lib.TLW_OK:
                raise _lib.TLWError(rc, "tlw_create_curved_parabolic failed (periodic mesh and positive Jacobians "
                                        "required)")
            return
        rc = _lib.lib().tlw_create_parabolic(C.byref(md), C.byref(basis), C.byref(sd), C.byref(self._par),
                                             C.byref(h))
        if rc != _lib.TLW_OK:
            raise _lib.TLWError(rc, "tlw_create_parabolic failed (invalid description, non-periodic mesh or HIP "
                                    "error)")


def _wall_face_points(mesh, direction, nodes):
    """Coordinates of the face nodes of a TreeMesh's boundary faces in one
    direction (0 x_neg, 1 x_pos, 2 y_neg, 3 y_pos)."""
    nbr = np.asarray(mesh.neighbors).reshape(-1, 4)
    elems = np.nonzero(nbr[:, direction] < 0)[0]
    if len(elems) == 0 or not hasattr(mesh, "element_center"):
        return []
    c = np.asarray(mesh.element_center).reshape(-1, 2)[elems]
    half = 0.5 * np.asarray(mesh.element_length)[elems]
    nodes = np.asarray(nodes)
    side = -1.0 if direction % 2 == 0 else 1.0
    if direction < 2:
        x = np.repeat(c[:, 0] + side * half, len(nodes))
        y = (c[:, 1][:, None] + half[:, None] * nodes[None, :]).ravel()
    else:
        x = (c[:, 0][:, None] + half[:, None] * nodes[None, :]).ravel()
        y = np.repeat(c[:, 1] + side * half, len(nodes))
    return list(zip(x.tolist(), y.tolist()))


def _rhs_call(fn, du_ode, u_ode, semi, t, stage=0):
    semi._check_state(u_ode, "u")
    semi._check_state(du_ode, "du")
    semi._bind_stream()
    dt = semi.cache["dt"][0]
    semi._before_rhs(float(t), float(dt), stage)
    _lib.check(fn(semi.handle, C.c_void_p(u_ode.data_ptr()), C.c_void_p(du_ode.data_ptr()), float(t), float(dt)),
               semi.handle)


def rhs(du_ode, u_ode, semi, t, tolerances=None):
    """rhs!(du_ode, u_ode, semi, t): du = LW-FR residual with dt = semi.cache["dt"][0]."""
    _rhs_call(_lib.lib().tlw_rhs_lw, du_ode, u_ode, semi, t)


def rhs_mdrk1(du_ode, u_ode, semi, t, tolerances=None):
    """rhs_mdrk1!(du_ode, u_ode, semi, t): first MDRK stage."""
    _rhs_call(_lib.lib().tlw_rhs_mdrk1, du_ode, u_ode, semi, t, stage=1)


def rhs_mdrk2(du_ode, u_ode, semi, t, tolerances=None):
    """rhs_mdrk2!(du_ode, u_ode, semi, t): second MDRK stage; reads us from the cache."""
    _rhs_call(_lib.lib().tlw_rhs_mdrk2, du_ode, u_ode, semi, t, stage=2)


@dataclass
class LWUpdate:
    f: object
    u0: object
    soln_arrays: dict
    tspan: tuple
    p: SemidiscretizationHyperbolic


def semidiscretize(semi, time_discretization, tspan):
    """semidiscretize(semi, time_discretization, tspan) -> LWUpdate (semidiscretization.jl:4-14)."""
    u0_ode = semi.compute_coefficients(tspan[0])
    du_ode = _torch().empty_like(u0_ode)
    return LWUpdate(rhs, u0_ode, {"u0_ode": u0_ode, "du_ode": du_ode}, (float(tspan[0]), float(tspan[1])), semi)


def max_dt_noalloc(u, inverse_jacobian, polydeg, gamma, out, stream=None):
    """Trixi max_dt of a TreeMesh Euler state without a context or any device
    allocation (tlw_max_dt_noalloc): u [E, n, n, 4] and inverse_jacobian [E]
    are float64 device tensors, the result lands in out[0] (a float64 device
    tensor) in stream order, NaN for a non-admissible state.  Calls that overlap
    on one GPU share the library's scratch array and race."""
    for name, x in (("u", u), ("inverse_jacobian", inverse_jacobian), ("out", out)):
        if not x.is_cuda or str(x.dtype) != "torch.float64" or not x.is_contiguous():
            raise ValueError(f"{name} must be a contiguous float64 device tensor")
    n_el = int(inverse_jacobian.numel())
    n = int(polydeg) + 1
    if u.numel() != 4 * n * n * n_el:
        raise ValueError("u must hold 4 (polydeg+1)^2 values per element")
    if out.numel() < 1:
        raise ValueError("out needs one element")
    s = stream.cuda_stream if stream is not None else None
    _lib.check(_lib.lib().tlw_max_dt_noalloc(C.c_void_p(u.data_ptr()), C.c_void_p(inverse_jacobian.data_ptr()), n_el,
                                             int(polydeg), float(gamma), C.c_void_p(out.data_ptr()),
                                             C.c_void_p(s) if s else None))
