"""ctypes binding of libtrixilw_hip.so (the C ABI in include/trixilw_hip.h).

The library must exist: there is no CPU fallback.  Build it with
``python trixilw.jl_amd/build.py`` (or ``__graft_entry__.build()``).
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

# TRIXILW_HIP_LIBRARY selects another build of the same ABI (e.g. other compile-time options)
LIB_PATH = Path(os.environ.get("TRIXILW_HIP_LIBRARY") or Path(__file__).resolve().parent / "libtrixilw_hip.so")

TLW_OK, TLW_EINVAL, TLW_EHIP, TLW_EADMISSIBLE, TLW_ENOTSUP = 0, 1, 2, 3, 4
TLW_LW, TLW_MDRK = 0, 1
TLW_FLUX_TREE, TLW_FLUX_STRUCTURED = 0, 1
TLW_BV_NONE, TLW_BV_VORTEX, TLW_BV_CONSTANT, TLW_BV_RIEMANN4, TLW_BV_SLIP_WALL, TLW_BV_TABLE = 0, 1, 2, 3, 4, 5
TLW_BV_CONVERGENCE_TEST = 6
TLW_BV_OUTFLOW, TLW_BV_SLIP_WALL_APPROXIMATE = 7, 8
TLW_SRC_NONE, TLW_SRC_CONVERGENCE_TEST, TLW_SRC_NAVIER_STOKES_CONVERGENCE_TEST = 0, 1, 2
TLW_AUX_U, TLW_AUX_F, TLW_AUX_ALPHA = 0, 1, 2
TLW_FLAG_KEEP_AUX = 1
TLW_PAR_NAVIER_STOKES = 1
TLW_GRAD_CONSERVATIVE, TLW_GRAD_PRIMITIVE = 0, 1

EXPORTED = (
    "tlw_create", "tlw_rebuild_mesh", "tlw_destroy", "tlw_last_error", "tlw_set_stream", "tlw_state_len",
    "tlw_rhs_lw", "tlw_rhs_mdrk1", "tlw_rhs_mdrk2", "tlw_us_ptr", "tlw_set_us", "tlw_max_dt", "tlw_set_state",
    "tlw_get_state", "tlw_state_ptr", "tlw_solve_init", "tlw_advance", "tlw_get_aux",
    "tlw_enable_kernel_timing", "tlw_kernel_times", "tlw_set_elements_per_workgroup",
    "tlw_get_elements_per_workgroup", "tlw_autotune", "tlw_set_positivity_limiter", "tlw_apply_positivity_limiter",
    "tlw_adjoint_lw", "tlw_create_batch", "tlw_n_instances", "tlw_batch_status", "tlw_enable_graphs",
    "tlw_create_parabolic", "tlw_solve_init_adaptive", "tlw_rejected_steps", "tlw_create_curved",
    "tlw_create_curved_parabolic", "tlw_max_dt_noalloc", "tlw_record_dts", "tlw_get_dts", "tlw_prescribe_dts",
    "tlw_set_state_dev", "tlw_get_state_dev", "tlw_boundary_nodes", "tlw_set_boundary_data",
    "tlw_create_ensemble", "tlw_ensemble_set_stream", "tlw_ensemble_enable_graphs", "tlw_ensemble_solve_init",
    "tlw_ensemble_advance", "tlw_ensemble_last_error", "tlw_ensemble_destroy", "tlw_adjoint_mdrk",
    "tlw_create_nonconforming", "tlw_rebuild_mesh_nonconforming", "tlw_transfer_solution", "tlw_create_p4est",
)


TLW_RECON_FIRST_ORDER, TLW_RECON_MUSCL_HANCOCK = 0, 1


TLW_NEIGHBOR_MORTAR = -2


class MortarDesc(C.Structure):
    _fields_ = [
        ("n_mortars", C.c_int32),
        ("mortars_host", C.POINTER(C.c_int32)),
    ]


class MeshDesc(C.Structure):
    _fields_ = [
        ("n_elements", C.c_int32),
        ("neighbors_host", C.POINTER(C.c_int32)),
        ("inverse_jacobian_host", C.POINTER(C.c_double)),
        ("element_center_host", C.POINTER(C.c_double)),
        ("element_length_host", C.POINTER(C.c_double)),
    ]


class BasisDesc(C.Structure):
    _fields_ = [
        ("polydeg", C.c_int32),
        ("nodes_host", C.POINTER(C.c_double)),
        ("weights_host", C.POINTER(C.c_double)),
        ("derivative_matrix_host", C.POINTER(C.c_double)),
        ("derivative_dhat_host", C.POINTER(C.c_double)),
    ]


class SolverDesc(C.Structure):
    _fields_ = [
        ("time_discretization", C.c_int32),
        ("gamma", C.c_double),
        ("surface_flux_order", C.c_int32),
        ("boundary_kind", C.c_int32),
        ("boundary_params", C.c_double * 24),
        ("flags", C.c_int32),
        ("shock_capturing", C.c_int32),
        ("alpha_max", C.c_double),
        ("alpha_min", C.c_double),
        ("alpha_smooth", C.c_int32),
        ("inverse_vandermonde_host", C.POINTER(C.c_double)),
        ("reconstruction", C.c_int32),
        ("source_terms", C.c_int32),
    ]


class CurvedDesc(C.Structure):
    _fields_ = [
        ("contravariant_vectors_host", C.POINTER(C.c_double)),
        ("inverse_jacobian_host", C.POINTER(C.c_double)),
        ("node_coordinates_host", C.POINTER(C.c_double)),
    ]


class P4estDesc(C.Structure):
    _fields_ = [
        ("contravariant_vectors_host", C.POINTER(C.c_double)),
        ("inverse_jacobian_host", C.POINTER(C.c_double)),
        ("node_coordinates_host", C.POINTER(C.c_double)),
        ("n_interfaces", C.c_int32),
        ("interfaces_host", C.POINTER(C.c_int32)),
    ]


class ParabolicDesc(C.Structure):
    _fields_ = [
        ("equations", C.c_int32),
        ("mu", C.c_double),
        ("prandtl", C.c_double),
        ("gradient_variables", C.c_int32),
        ("walls", C.c_int32),
        ("wall_velocity", C.c_double * 8),
        ("wall_isothermal", C.c_int32 * 4),
        ("wall_temperature", C.c_double * 4),
    ]


class TLWError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"libtrixilw_hip error {code}: {msg}")
        self.code = code


class NonAdmissibleState(TLWError):
    """The device found ρ <= 0 or p <= 0 (the reference raises DomainError)."""


_lib = None


def lib():
    """Load libtrixilw_hip.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python trixilw.jl_amd/build.py`")
    # torch bundles its own libamdhip64.so.7; loading torch first makes this
    # library bind to that single HIP runtime (same soname), so streams and
    # device pointers are shared.
    import torch  # noqa: F401
    L = C.CDLL(str(LIB_PATH))
    vp, dp, i32, sz = C.c_void_p, C.POINTER(C.c_double), C.c_int32, C.c_size_t
    L.tlw_create.restype = C.c_int
    L.tlw_create.argtypes = [C.POINTER(MeshDesc), C.POINTER(BasisDesc), C.POINTER(SolverDesc), C.POINTER(vp)]
    L.tlw_create_parabolic.restype = C.c_int
    L.tlw_create_parabolic.argtypes = [C.POINTER(MeshDesc), C.POINTER(BasisDesc), C.POINTER(SolverDesc),
                                       C.POINTER(ParabolicDesc), C.POINTER(vp)]
    L.tlw_create_curved.restype = C.c_int
    L.tlw_create_curved.argtypes = [C.POINTER(MeshDesc), C.POINTER(BasisDesc), C.POINTER(SolverDesc),
                                    C.POINTER(CurvedDesc), C.POINTER(vp)]
    L.tlw_create_curved_parabolic.restype = C.c_int
    L.tlw_create_curved_parabolic.argtypes = [C.POINTER(MeshDesc), C.POINTER(BasisDesc), C.POINTER(SolverDesc),
                                              C.POINTER(CurvedDesc), C.POINTER(ParabolicDesc), C.POINTER(vp)]
    L.tlw_rebuild_mesh.restype = C.c_int
    L.tlw_rebuild_mesh.argtypes = [vp, C.POINTER(MeshDesc)]
    L.tlw_create_nonconforming.restype = C.c_int
    L.tlw_create_nonconforming.argtypes = [C.POINTER(MeshDesc), C.POINTER(MortarDesc), C.POINTER(BasisDesc),
                                           C.POINTER(SolverDesc), C.POINTER(vp)]
    L.tlw_rebuild_mesh_nonconforming.restype = C.c_int
    L.tlw_rebuild_mesh_nonconforming.argtypes = [vp, C.POINTER(MeshDesc), C.POINTER(MortarDesc)]
    L.tlw_create_p4est.restype = C.c_int
    L.tlw_create_p4est.argtypes = [C.POINTER(MeshDesc), C.POINTER(P4estDesc), C.POINTER(BasisDesc),
                                   C.POINTER(SolverDesc), C.POINTER(vp)]
    L.tlw_transfer_solution.restype = C.c_int
    L.tlw_transfer_solution.argtypes = [vp, vp, C.c_int32, vp, C.c_int32, C.POINTER(C.c_int32)]
    L.tlw_destroy.restype = None
    L.tlw_destroy.argtypes = [vp]
    L.tlw_last_error.restype = C.c_char_p
    L.tlw_last_error.argtypes = [vp]
    L.tlw_set_stream.restype = C.c_int
    L.tlw_set_stream.argtypes = [vp, vp]
    L.tlw_state_len.restype = sz
    L.tlw_state_len.argtypes = [vp]
    for nm in ("tlw_rhs_lw", "tlw_rhs_mdrk1", "tlw_rhs_mdrk2"):
        f = getattr(L, nm)
        f.restype = C.c_int
        f.argtypes = [vp, vp, vp, C.c_double, C.c_double]
    L.tlw_us_ptr.restype = C.c_int
    L.tlw_us_ptr.argtypes = [vp, C.POINTER(vp)]
    L.tlw_set_us.restype = C.c_int
    L.tlw_set_us.argtypes = [vp, vp]
    L.tlw_max_dt.restype = C.c_int
    L.tlw_max_dt.argtypes = [vp, vp, dp]
    L.tlw_set_state.restype = C.c_int
    L.tlw_set_state.argtypes = [vp, dp, sz]
    L.tlw_get_state.restype = C.c_int
    L.tlw_get_state.argtypes = [vp, dp, sz]
    L.tlw_state_ptr.restype = C.c_int
    L.tlw_state_ptr.argtypes = [vp, C.POINTER(vp)]
    L.tlw_solve_init.restype = C.c_int
    L.tlw_solve_init.argtypes = [vp, C.c_double, C.c_double, C.c_double]
    L.tlw_solve_init_adaptive.restype = C.c_int
    L.tlw_solve_init_adaptive.argtypes = [vp, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double]
    L.tlw_rejected_steps.restype = C.c_int
    L.tlw_rejected_steps.argtypes = [vp, C.POINTER(C.c_int64)]
    L.tlw_advance.restype = C.c_int
    L.tlw_advance.argtypes = [vp, i32, dp, dp, C.POINTER(C.c_int64), C.POINTER(i32)]
    L.tlw_enable_kernel_timing.restype = C.c_int
    L.tlw_enable_kernel_timing.argtypes = [vp, i32]
    L.tlw_kernel_times.restype = C.c_int
    L.tlw_kernel_times.argtypes = [vp, dp, C.POINTER(C.c_int64), dp, C.POINTER(C.c_int64)]
    L.tlw_get_aux.restype = C.c_int
    L.tlw_get_aux.argtypes = [vp, i32, dp, sz]
    L.tlw_set_elements_per_workgroup.restype = C.c_int
    L.tlw_set_elements_per_workgroup.argtypes = [vp, i32]
    L.tlw_get_elements_per_workgroup.restype = C.c_int
    L.tlw_get_elements_per_workgroup.argtypes = [vp, C.POINTER(i32)]
    L.tlw_set_positivity_limiter.restype = C.c_int
    L.tlw_set_positivity_limiter.argtypes = [vp, i32, C.c_double, C.c_double]
    L.tlw_apply_positivity_limiter.restype = C.c_int
    L.tlw_apply_positivity_limiter.argtypes = [vp, vp]
    L.tlw_create_batch.restype = C.c_int
    L.tlw_create_batch.argtypes = [i32, C.POINTER(MeshDesc), C.POINTER(BasisDesc), C.POINTER(SolverDesc),
                                   C.POINTER(vp)]
    L.tlw_n_instances.restype = C.c_int
    L.tlw_n_instances.argtypes = [vp, C.POINTER(i32)]
    L.tlw_batch_status.restype = C.c_int
    L.tlw_batch_status.argtypes = [vp, i32, dp, dp, C.POINTER(C.c_int64), C.POINTER(i32)]
    L.tlw_enable_graphs.restype = C.c_int
    L.tlw_enable_graphs.argtypes = [vp, i32]
    L.tlw_adjoint_lw.restype = C.c_int
    L.tlw_adjoint_lw.argtypes = [vp, vp, vp, vp, C.c_double, C.c_double]
    L.tlw_autotune.restype = C.c_int
    L.tlw_autotune.argtypes = [vp, i32, C.POINTER(i32), dp]
    L.tlw_adjoint_mdrk.restype = C.c_int
    L.tlw_adjoint_mdrk.argtypes = [vp, vp, vp, vp, vp, C.c_double, C.c_double]
    L.tlw_create_ensemble.restype = C.c_int
    L.tlw_create_ensemble.argtypes = [vp, vp, C.POINTER(vp)]
    L.tlw_ensemble_set_stream.restype = C.c_int
    L.tlw_ensemble_set_stream.argtypes = [vp, vp]
    L.tlw_ensemble_enable_graphs.restype = C.c_int
    L.tlw_ensemble_enable_graphs.argtypes = [vp, i32]
    L.tlw_ensemble_solve_init.restype = C.c_int
    L.tlw_ensemble_solve_init.argtypes = [vp, C.c_double, C.c_double, C.c_double, C.c_double]
    L.tlw_ensemble_advance.restype = C.c_int
    L.tlw_ensemble_advance.argtypes = [vp, i32, C.POINTER(i32)]
    L.tlw_ensemble_last_error.restype = C.c_char_p
    L.tlw_ensemble_last_error.argtypes = [vp]
    L.tlw_ensemble_destroy.restype = None
    L.tlw_ensemble_destroy.argtypes = [vp]
    L.tlw_boundary_nodes.restype = C.c_int
    L.tlw_boundary_nodes.argtypes = [vp, C.POINTER(C.c_int64), dp, sz]
    L.tlw_set_boundary_data.restype = C.c_int
    L.tlw_set_boundary_data.argtypes = [vp, dp, sz, i32]
    L.tlw_record_dts.restype = C.c_int
    L.tlw_record_dts.argtypes = [vp, C.c_int64]
    L.tlw_get_dts.restype = C.c_int
    L.tlw_get_dts.argtypes = [vp, dp, C.c_int64]
    L.tlw_prescribe_dts.restype = C.c_int
    L.tlw_prescribe_dts.argtypes = [vp, dp, C.c_int64]
    L.tlw_set_state_dev.restype = C.c_int
    L.tlw_set_state_dev.argtypes = [vp, vp]
    L.tlw_get_state_dev.restype = C.c_int
    L.tlw_get_state_dev.argtypes = [vp, vp]
    L.tlw_max_dt_noalloc.restype = C.c_int
    L.tlw_max_dt_noalloc.argtypes = [vp, vp, i32, i32, C.c_double, vp, vp]
    _lib = L
    return L


def check(rc: int, handle=None) -> None:
    if rc == TLW_OK:
        return
    # a NULL handle gives the message of the last failed context-free call of this thread
    msg = lib().tlw_last_error(handle).decode()
    if rc == TLW_EADMISSIBLE:
        raise NonAdmissibleState(rc, msg)
    raise TLWError(rc, msg)
