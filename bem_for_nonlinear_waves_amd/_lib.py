"""ctypes binding of libbemfmm.so (include/bemfmm.h).

The shared library is built in-tree by ``bem_for_nonlinear_waves_amd.build``;
loading fails loudly when it is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libbemfmm.so")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)

OK = 0
E_INVALID, E_DEVICE, E_OOM, E_NONFINITE, E_NOT_CONVERGED, E_STATE = -1, -2, -3, -4, -5, -6
PRECISION_FP64, PRECISION_FP32_REFINE = 0, 1
PRECOND_NONE, PRECOND_JACOBI, PRECOND_SCHWARZ, PRECOND_ILU0, PRECOND_MILU0, PRECOND_ILUT = 0, 1, 2, 3, 4, 5


class Opts(ctypes.Structure):
    _fields_ = [("p", ctypes.c_int), ("theta", ctypes.c_double), ("leaf_max", ctypes.c_int),
                ("max_level", ctypes.c_int), ("q", ctypes.c_int), ("precision", ctypes.c_int),
                ("preconditioner", ctypes.c_int), ("refine_eta", ctypes.c_double), ("schwarz_max_core", ctypes.c_int),
                ("schwarz_max_block", ctypes.c_int), ("schwarz_overlap", ctypes.c_int), ("m2l_tile_pairs", ctypes.c_int),
                ("dof_reorder", ctypes.c_int), ("use_graphs", ctypes.c_int), ("m2l_kernel", ctypes.c_int),
                ("overlap_near", ctypes.c_int), ("refine_recycle", ctypes.c_int),
                ("ilu_neighborhood", ctypes.c_int), ("ilu_kring", ctypes.c_int), ("schwarz_variant", ctypes.c_int),
                ("schwarz_core_k", ctypes.c_int), ("milu_omega", ctypes.c_double), ("ilut_drop_tol", ctypes.c_double),
                ("ilut_max_entries", ctypes.c_int), ("ilut_pivot_min", ctypes.c_double),
                ("plane_wave_digits", ctypes.c_int), ("tree_on_gpu", ctypes.c_int), ("element_order", ctypes.c_int),
                ("near_rows_per_wave", ctypes.c_int), ("p2m_loads", ctypes.c_int), ("inner_p", ctypes.c_int)]


class Result(ctypes.Structure):
    _fields_ = [("iters", ctypes.c_int), ("n_matvec", ctypes.c_int), ("converged", ctypes.c_int),
                ("residual", ctypes.c_double), ("time_setup", ctypes.c_double), ("time_solve", ctypes.c_double),
                ("refinement_steps", ctypes.c_int)]


class Stats(ctypes.Structure):
    _fields_ = [("n_nodes", ctypes.c_int), ("n_faces", ctypes.c_int), ("n_dof", ctypes.c_int),
                ("n_levels", ctypes.c_int), ("n_boxes", ctypes.c_int), ("n_leaves", ctypes.c_int),
                ("n_m2l_pairs", ctypes.c_int64), ("near_nnz", ctypes.c_int64), ("n_m2l_ops", ctypes.c_int),
                ("L", ctypes.c_double), ("device_bytes", ctypes.c_double), ("t_tree", ctypes.c_double),
                ("t_lists", ctypes.c_double), ("t_near_struct", ctypes.c_double), ("t_p2m", ctypes.c_double),
                ("t_ops", ctypes.c_double), ("t_upload", ctypes.c_double), ("t_near_quad", ctypes.c_double),
                ("t_alpha", ctypes.c_double), ("n_precond_blocks", ctypes.c_int),
                ("precond_fallback_blocks", ctypes.c_int), ("precond_avg_block", ctypes.c_double),
                ("n_ops_uploaded", ctypes.c_int), ("m2l_tile_pairs", ctypes.c_int),
                ("m2l_rotation", ctypes.c_int), ("precond_nnz", ctypes.c_int64), ("precond_levels", ctypes.c_int),
                ("m2l_plane_wave", ctypes.c_int), ("tree_on_gpu", ctypes.c_int), ("t_tune", ctypes.c_double),
                ("t_precond", ctypes.c_double), ("t_fp32", ctypes.c_double),
                ("near_rows_per_wave", ctypes.c_int * 2), ("p2m_loads", ctypes.c_int * 2), ("inner_p", ctypes.c_int)]


# (name, restype, argtypes) of every entry point declared in include/bemfmm.h
SIGNATURES = [
    ("bemfmm_is_available", ctypes.c_int, []),
    ("bemfmm_device_name", ctypes.c_char_p, []),
    ("bemfmm_default_opts", None, [ctypes.POINTER(Opts)]),
    ("bemfmm_ctx_create", ctypes.c_void_p, [ctypes.c_int, ctypes.POINTER(Opts)]),
    ("bemfmm_ctx_destroy", None, [ctypes.c_void_p]),
    ("bemfmm_last_error", ctypes.c_char_p, [ctypes.c_void_p]),
    ("bemfmm_set_geometry", ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_int, _dp, ctypes.c_int, _ip, _ip, ctypes.c_int, _ip, _ip]),
    ("bemfmm_set_bc", ctypes.c_int, [ctypes.c_void_p, _dp, _dp]),
    ("bemfmm_set_robin", ctypes.c_int, [ctypes.c_void_p, _dp]),
    ("bemfmm_prepare", ctypes.c_int, [ctypes.c_void_p]),
    ("bemfmm_matvec", ctypes.c_int, [ctypes.c_void_p, _dp, _dp]),
    ("bemfmm_matvec_device", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    ("bemfmm_matvec_device_noalloc", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    ("bemfmm_matvec_fp32", ctypes.c_int, [ctypes.c_void_p, _dp, _dp]),
    ("bemfmm_matvec_fp32_device", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    ("bemfmm_sync", ctypes.c_int, [ctypes.c_void_p]),
    ("bemfmm_rhs", ctypes.c_int, [ctypes.c_void_p, _dp]),
    ("bemfmm_alpha", ctypes.c_int, [ctypes.c_void_p, _dp]),
    ("bemfmm_solve", ctypes.c_int,
     [ctypes.c_void_p, _dp, _dp, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.POINTER(Result)]),
    ("bemfmm_get_stats", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(Stats)]),
    ("bemfmm_matvec_profile", ctypes.c_int, [ctypes.c_void_p, _dp]),
    ("bemfmm_residual_history", ctypes.c_int, [ctypes.c_void_p, _dp, ctypes.c_int]),
    ("bemfmm_refinement_history", ctypes.c_int, [ctypes.c_void_p, _dp, ctypes.c_int]),
    ("bemfmm_precond_apply", ctypes.c_int, [ctypes.c_void_p, _dp, _dp]),
    # include/bemfmm_server.h (structures in serve.py; passed here as opaque pointers)
    ("bemfmm_server_default_opts", None, [ctypes.c_void_p]),
    ("bemfmm_server_create", ctypes.c_void_p, [ctypes.c_void_p]),
    ("bemfmm_server_destroy", None, [ctypes.c_void_p]),
    ("bemfmm_server_submit", ctypes.c_int64, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64]),
    ("bemfmm_server_wait", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, _dp, ctypes.c_void_p, _dp]),
    ("bemfmm_server_close_session", ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64]),
    ("bemfmm_server_get_stats", ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    ("bemfmm_server_last_error", ctypes.c_char_p, [ctypes.c_void_p]),
    ("bemfmm_dof_order", ctypes.c_int, [ctypes.c_void_p, _ip]),
    ("bemfmm_schwarz_blocks", ctypes.c_int,
     [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64), _ip, _ip]),
    ("bemfmm_tree_digest", ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint64)]),
]

_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m bem_for_nonlinear_waves_amd.build` "
                           "(there is no CPU fallback)")
    lib = ctypes.CDLL(LIB_PATH)
    for name, res, args in SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib
