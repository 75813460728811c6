"""Loader for libself_amd.so, the C ABI declared in include/self_amd.h.

The library is built in-tree by ``make -C self_amd`` (see __graft_entry__.build).
There is no fallback: if the library is missing or a call fails, an
exception is raised.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libself_amd.so")

SELF_OK = 0
SELF_NOT_CONVERGED = -1
SELF_ERR_OTHER = -2
SELF_ERR_HIP = -3
SELF_ERR_OOM = -4
SELF_ERR_ARG = -5
SELF_ERR_STATE = -6

GAUSS, GAUSS_LOBATTO = 1, -1
DG, CG = 2000, 2001
NO_NORMAL_FLOW, RADIATION = -100, -101
INFLOW_ONE, INFLOW_TWO = -103, -104
FP64, MIXED = 0, 1
PLAIN, FUSED = 0, 1

dp = C.POINTER(C.c_double)
fp = C.POINTER(C.c_float)
ip = C.POINTER(C.c_int)
szp = C.POINTER(C.c_size_t)

_lib = None


class SelfError(RuntimeError):
    """A C-ABI call returned a negative status."""

    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is not built; run __graft_entry__.build() or make -C self_amd")
        _lib = C.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


def check(rc, allow=()):
    if rc != SELF_OK and rc not in allow:
        msg = lib().self_last_error().decode(errors="replace")
        raise SelfError(rc, msg)
    return rc


def ptr(a, kind=dp):
    if a is None:
        return None
    return a.ctypes.data_as(kind)


class SWDesc(C.Structure):
    _fields_ = [("N", C.c_int), ("nElems", C.c_int), ("nEdges", C.c_int), ("precision", C.c_int),
                ("form", C.c_int), ("g", C.c_double), ("dt", C.c_double), ("linearDrag", C.c_double)]


class Request(C.Structure):
    """self_request of include/self_amd.h"""
    _fields_ = [("kind", C.c_int), ("nX", C.c_int), ("N", C.c_int), ("precision", C.c_int), ("steps", C.c_int),
                ("precond", C.c_int), ("maxIters", C.c_int), ("L", C.c_double), ("g", C.c_double),
                ("f0", C.c_double), ("dt", C.c_double), ("tol", C.c_double), ("in0", dp), ("in1", dp),
                ("in2", dp), ("out", dp)]


class Result(C.Structure):
    """self_request_result of include/self_amd.h"""
    _fields_ = [("status", C.c_int), ("ioerr", C.c_int), ("iterations", C.c_int), ("queue_ms", C.c_double),
                ("run_ms", C.c_double), ("latency_ms", C.c_double), ("bytes", C.c_size_t),
                ("members", C.c_int), ("models", C.c_int)]


_SIGS = {
    "self_last_error": ([], C.c_char_p),
    "self_version": ([], C.c_int),
    "self_init": ([C.c_int], C.c_int),
    "self_device_count": ([ip], C.c_int),
    "self_mem_info": ([szp, szp], C.c_int),
    "self_stream_copy": ([C.c_size_t, C.c_int, dp, dp], C.c_int),
    "self_legendre_quadrature": ([C.c_int, C.c_int, dp, dp], C.c_int),
    "self_barycentric_weights": ([C.c_int, dp, dp], C.c_int),
    "self_lagrange_polynomials": ([C.c_int, dp, dp, C.c_double, dp], C.c_int),
    "self_derivative_matrix": ([C.c_int, dp, dp, dp], C.c_int),
    "self_interpolation_matrix": ([C.c_int, dp, dp, C.c_int, dp, dp], C.c_int),
    "self_nodal2d_build": ([C.c_int, C.c_int, C.c_int, C.c_int] + [dp] * 8, C.c_int),
    "self_quadmesh_counts": ([C.c_int, C.c_int, C.c_int, C.c_int, ip, ip], C.c_int),
    "self_quadmesh_structured": ([C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                  C.c_double, dp, ip] + [dp] * 9, C.c_int),
    "self_sw2d_create": ([C.POINTER(SWDesc), C.POINTER(C.c_void_p)], C.c_int),
    "self_sw2d_destroy": ([C.c_void_p], C.c_int),
    "self_sw2d_set_nodal": ([C.c_void_p] + [dp] * 8, C.c_int),
    "self_sw2d_set_geometry": ([C.c_void_p, ip] + [dp] * 7, C.c_int),
    "self_sw2d_set_field": ([C.c_void_p, C.c_int, dp], C.c_int),
    "self_sw2d_get_field": ([C.c_void_p, C.c_int, dp], C.c_int),
    "self_sw2d_set_params": ([C.c_void_p, C.c_double, C.c_double, C.c_double], C.c_int),
    "self_sw2d_calculate_solution_at_boundaries": ([C.c_void_p], C.c_int),
    "self_sw2d_edge_flux": ([C.c_void_p, C.c_double], C.c_int),
    "self_sw2d_mapped_time_derivative": ([C.c_void_p, C.c_double], C.c_int),
    "self_sw2d_global_time_derivative": ([C.c_void_p, C.c_double], C.c_int),
    "self_sw2d_check_mixed": ([C.c_void_p, C.c_double, dp, ip], C.c_int),
    "self_sw2d_forward_step_rk3": ([C.c_void_p, C.c_double, C.c_int], C.c_int),
    "self_sw2d_forward_step_rk3_timed": ([C.c_void_p, C.c_double, C.c_int, dp], C.c_int),
    "self_sw2d_calculate_bathymetry_gradient": ([C.c_void_p], C.c_int),
    "self_sw2d_kernel_times": ([C.c_void_p, C.c_double, C.c_int, dp, ip], C.c_int),
    "self_sw2d_synchronize": ([C.c_void_p], C.c_int),
    "self_sw2d_device_bytes": ([C.c_void_p, szp], C.c_int),
    "self_linop_dense_create": ([C.c_int, dp, dp, C.POINTER(C.c_void_p)], C.c_int),
    "self_linop_elliptic1d_create": ([C.c_int, C.c_int, C.c_double, C.POINTER(C.c_void_p)], C.c_int),
    "self_linop_elliptic1d_unmask": ([C.c_void_p, dp, dp, dp], C.c_int),
    "self_linop_size": ([C.c_void_p, ip], C.c_int),
    "self_linop_destroy": ([C.c_void_p], C.c_int),
    "self_krylov_solve": ([C.c_void_p, C.c_int, C.c_double, C.c_int, C.c_int, C.c_int, dp, dp, C.c_int, ip, ip],
                          C.c_int),
    "self_cgsem2d_create": ([C.c_int, C.c_int, C.c_int, C.c_int, ip, ip] + [dp] * 5 + [C.POINTER(C.c_void_p)],
                            C.c_int),
    "self_cgsem2d_destroy": ([C.c_void_p], C.c_int),
    "self_cgsem2d_global_dofs": ([C.c_void_p, ip], C.c_int),
    "self_cgsem2d_set_flux_coefficient": ([C.c_void_p, dp], C.c_int),
    "self_cgsem2d_set_source": ([C.c_void_p, dp], C.c_int),
    "self_cgsem2d_set_dirichlet": ([C.c_void_p, dp], C.c_int),
    "self_cgsem2d_set_solution": ([C.c_void_p, dp], C.c_int),
    "self_cgsem2d_get_solution": ([C.c_void_p, dp], C.c_int),
    "self_cgsem2d_matrix_action": ([C.c_void_p, dp, dp], C.c_int),
    "self_cgsem2d_residual": ([C.c_void_p, dp], C.c_int),
    "self_cgsem2d_solve": ([C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_int, C.c_int, C.c_int, ip, ip, dp,
                            C.c_int, dp], C.c_int),
    "self_cgsem2d_solve_batched": ([C.POINTER(C.c_void_p), C.c_int, C.c_int, C.c_double, C.c_int, ip, ip, dp],
                                   C.c_int),
    "self_cgsem2d_mask": ([C.c_void_p, dp], C.c_int),
    "self_cgsem2d_dof_map": ([C.c_void_p, ip], C.c_int),
    "self_cgsem2d_pc_matrix": ([C.c_void_p, ip, ip, dp], C.c_int),
    "self_cgsem2d_set_preconditioner": ([C.c_void_p, dp, dp, C.c_int, C.c_double, C.c_int], C.c_int),
    "self_cgsem2d_apply_preconditioner": ([C.c_void_p, dp, dp, ip], C.c_int),
    "self_cgsem2d_solve_gmres_mixed": ([C.c_void_p, C.c_int, C.c_double, C.c_int, C.c_int, C.c_int,
                                        ip, ip, dp, dp], C.c_int),
    "self_cgsem2d_solve_mixed": ([C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_int, C.c_int, ip, ip, ip, dp, dp],
                                 C.c_int),
    "self_cgsem2d_solve_reliable": ([C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_int, C.c_int, ip, ip, ip, dp,
                                     dp], C.c_int),
    "self_cgsem2d_solve_reliable_fp64": ([C.c_void_p, C.c_int, C.c_double, C.c_double, C.c_int, C.c_int, ip, ip, ip, dp,
                                     dp], C.c_int),
    "self_adv2d_create": ([C.c_int, C.c_int, C.c_int, C.c_int, ip] + [dp] * 14 + [C.c_double, C.c_double,
                                                                          C.POINTER(C.c_void_p)], C.c_int),
    "self_adv2d_destroy": ([C.c_void_p], C.c_int),
    "self_adv2d_set_field": ([C.c_void_p, C.c_int, dp], C.c_int),
    "self_adv2d_get_field": ([C.c_void_p, C.c_int, dp], C.c_int),
    "self_adv2d_global_time_derivative": ([C.c_void_p, C.c_double], C.c_int),
    "self_adv2d_forward_step_rk3": ([C.c_void_p, C.c_double, C.c_int, dp], C.c_int),
    "self_adv2d_synchronize": ([C.c_void_p], C.c_int),
    "self_nodal3d_build": ([C.c_int, C.c_int, C.c_int, dp, dp, dp, dp, dp], C.c_int),
    "self_rolloff_filter1d": ([C.c_int, C.c_int, dp, dp, dp], C.c_int),
    "self_hexmesh_counts": ([C.c_int] * 6 + [ip, ip], C.c_int),
    "self_hexmesh_structured": ([C.c_int] * 7 + [C.c_double] * 4 + [ip, dp, ip] + [dp] * 8, C.c_int),
    "self_adv3d_create": ([C.c_int] * 4 + [ip] + [dp] * 10 + [C.c_int, C.c_double, C.POINTER(C.c_void_p)],
                          C.c_int),
    "self_adv3d_destroy": ([C.c_void_p], C.c_int),
    "self_adv3d_set_field": ([C.c_void_p, C.c_int, dp], C.c_int),
    "self_adv3d_get_field": ([C.c_void_p, C.c_int, dp], C.c_int),
    "self_adv3d_global_time_derivative": ([C.c_void_p, C.c_double, C.c_int], C.c_int),
    "self_adv3d_forward_step_rk3": ([C.c_void_p, C.c_double, C.c_int, dp], C.c_int),
    "self_adv3d_filter_count": ([C.c_void_p, C.c_int, C.POINTER(C.c_uint)], C.c_int),
    "self_adv3d_synchronize": ([C.c_void_p], C.c_int),
    "self_adv3d_set_precision": ([C.c_void_p, C.c_int], C.c_int),
    "self_hpe_create": ([C.c_int] * 4 + [ip] + [dp] * 10 + [C.c_int, C.c_double, C.POINTER(C.c_void_p)], C.c_int),
    "self_hpe_destroy": ([C.c_void_p], C.c_int),
    "self_hpe_set_field": ([C.c_void_p, C.c_int, dp], C.c_int),
    "self_hpe_get_field": ([C.c_void_p, C.c_int, dp], C.c_int),
    "self_hpe_global_time_derivative": ([C.c_void_p], C.c_int),
    "self_hpe_forward_step": ([C.c_void_p, C.c_int, C.c_int, dp], C.c_int),
    "self_hpe_kernel_times": ([C.c_void_p, C.c_int, dp], C.c_int),
    "self_hpe_synchronize": ([C.c_void_p], C.c_int),
    "self_nodal2d_apply_derivative": ([C.c_int, C.c_size_t, dp, dp, dp, dp, dp], C.c_int),
    "self_nodal2d_apply_derivative_dev": ([C.c_int, C.c_size_t, dp, dp, dp, dp, dp, C.c_void_p], C.c_int),
    "self_nodal2d_boundary_interp": ([C.c_int, C.c_size_t] + [dp] * 9, C.c_int),
    "self_nodal2d_boundary_interp_dev": ([C.c_int, C.c_size_t] + [dp] * 9 + [C.c_void_p], C.c_int),
    "self_nodal3d_apply_derivative": ([C.c_int, C.c_size_t] + [dp] * 5, C.c_int),
    "self_nodal3d_apply_derivative_dev": ([C.c_int, C.c_size_t] + [dp] * 5 + [C.c_void_p], C.c_int),
    "self_nodal3d_boundary_interp": ([C.c_int, C.c_size_t] + [dp] * 4, C.c_int),
    "self_nodal3d_boundary_interp_dev": ([C.c_int, C.c_size_t] + [dp] * 4 + [C.c_void_p], C.c_int),
    "self_coarse_to_fine_2d": ([C.c_int, C.c_int, C.c_size_t, dp, dp, C.c_int, C.c_int, dp, dp], C.c_int),
    "self_coarse_to_fine_2d_noalloc": ([C.c_int, C.c_int, C.c_size_t, dp, dp, C.c_int, C.c_int, dp, dp], C.c_int),
    "self_coarse_to_fine_2d_dev": ([C.c_int, C.c_int, C.c_size_t, dp, dp, C.c_int, C.c_int, dp, dp, C.c_void_p],
                                   C.c_int),
    "self_coarse_to_fine_3d": ([C.c_int, C.c_size_t, dp, C.c_int, dp, dp], C.c_int),
    "self_coarse_to_fine_3d_dev": ([C.c_int, C.c_size_t, dp, C.c_int, dp, dp, C.c_void_p], C.c_int),
    "self_sw2d_set_groups": ([C.c_void_p, C.c_int], C.c_int),
    "self_sw2d_set_graph": ([C.c_void_p, C.c_int], C.c_int),
    "self_sw2d_forward_step_rk3_multi": ([C.POINTER(C.c_void_p), C.c_int, C.c_int, dp], C.c_int),
    "self_sw2d_set_variant": ([C.c_void_p, C.c_int], C.c_int),
    "self_sw2d_set_members": ([C.c_void_p, C.c_int], C.c_int),
    "self_sw2d_write_pickup": ([C.c_void_p, C.c_char_p], C.c_int),
    "self_sw2d_read_pickup": ([C.c_void_p, C.c_char_p], C.c_int),
    "self_adv2d_write_pickup": ([C.c_void_p, C.c_char_p], C.c_int),
    "self_adv2d_read_pickup": ([C.c_void_p, C.c_char_p], C.c_int),
    "self_pickup_filename": ([C.c_char_p, C.c_int, C.c_char_p, C.c_int], C.c_int),
    "self_specmesh_counts": ([C.c_char_p, ip, ip, ip, ip], C.c_int),
    "self_quadmesh_scale": ([C.c_int, dp, C.c_int, C.c_int, C.c_double, C.c_double] + [dp] * 11, C.c_int),
    "self_specmesh_read": ([C.c_char_p, C.c_int, dp, ip, ip, ip, dp, dp] + [dp] * 9, C.c_int),
    "self_sw2d_pickup_write": ([C.c_char_p, C.c_int, C.c_int, dp, dp, dp], C.c_int),
    "self_sw2d_pickup_read": ([C.c_char_p, C.c_int, C.c_int, dp, dp, dp], C.c_int),
    "self_adv2d_pickup_write": ([C.c_char_p, C.c_int, C.c_int, C.c_int, dp, dp, dp, dp], C.c_int),
    "self_adv2d_pickup_read": ([C.c_char_p, C.c_int, C.c_int, C.c_int, dp, dp, dp, dp], C.c_int),
    "self_sw2d_set_member_solution": ([C.c_void_p, C.c_int, dp], C.c_int),
    "self_sw2d_get_member_solution": ([C.c_void_p, C.c_int, dp], C.c_int),
    "self_sw2d_set_member_time_step": ([C.c_void_p, C.c_int, C.c_double], C.c_int),
    "self_sw2d_autotune": ([C.c_void_p, C.c_int, ip, dp], C.c_int),
    "self_tune_count": ([ip], C.c_int),
    "self_tune_entry": ([C.c_int, C.c_char_p, C.c_int, ip, ip, dp], C.c_int),
    "self_tune_clear": ([], C.c_int),
    "self_tune_save": ([C.c_char_p], C.c_int),
    "self_tune_load": ([C.c_char_p], C.c_int),
    "self_serve_create": ([C.c_int, C.c_double, C.POINTER(C.c_void_p)], C.c_int),
    "self_serve_submit": ([C.c_void_p, C.POINTER(Request), C.POINTER(C.c_longlong)], C.c_int),
    "self_serve_wait": ([C.c_void_p, C.c_longlong, C.POINTER(Result)], C.c_int),
    "self_serve_stats": ([C.c_void_p, ip, ip, szp, szp], C.c_int),
    "self_serve_destroy": ([C.c_void_p], C.c_int),
}

EXTRA_SIGS: dict = {}


def _declare(L):
    for name, (args, res) in list(_SIGS.items()) + list(EXTRA_SIGS.items()):
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res


def exported_symbols():
    """Names this module binds (used by tests against include/self_amd.h)."""
    return sorted(list(_SIGS) + list(EXTRA_SIGS))


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)
