"""ctypes loader for the MI355X WRAPD library (wrapd_amd/lib/libwrapd.so).

The library is the product: there is no CPU fallback.  Loading fails loudly when the
built library is missing.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(_HERE, "lib")
LIB_PATH = os.path.join(LIB_DIR, "libwrapd.so")
QUASI_PATH = os.path.join(LIB_DIR, "quasi")

D = C.POINTER(C.c_double)
I = C.POINTER(C.c_int)
F = C.POINTER(C.c_float)
P = C.c_void_p


class Material(C.Structure):
    _fields_ = [("model", C.c_int), ("mu", C.c_double), ("lam", C.c_double), ("beta_static", C.c_double),
                ("beta_min", C.c_double), ("beta_max", C.c_double)]


class LocalArgs(C.Structure):
    _fields_ = [("precision", C.c_int), ("prox_mode", C.c_int), ("prox_iters", C.c_int),
                ("update_cand", C.c_int), ("cand_use_hessian", C.c_int), ("newton_rtol", C.c_double),
                ("fallback_sigma", C.c_double)]


class LocalOut(C.Structure):
    _fields_ = [("delta_auglag", C.c_double), ("bef", C.c_double), ("aft", C.c_double), ("pe", C.c_double),
                ("max_wsq_ratio", C.c_double), ("prox_iters_total", C.c_double),
                ("fallback_count", C.c_double), ("neg_sigma_count", C.c_double),
                ("learned_accepted", C.c_double), ("evals_total", C.c_double), ("evals_max", C.c_double),
                ("learned_used", C.c_double)]


class LbfgsCfg(C.Structure):
    _fields_ = [("memory", C.c_int), ("max_iters", C.c_int), ("kappa", C.c_double), ("linesearch", C.c_int)]


class GlobalOut(C.Structure):
    _fields_ = [("iters", C.c_int), ("obj_bef", C.c_double), ("obj_aft", C.c_double), ("energy", C.c_double),
                ("flips", C.c_int), ("exit_reason", C.c_int)]


class SolverSettings(C.Structure):
    _fields_ = [("model", C.c_int), ("youngs", C.c_double), ("poisson", C.c_double),
                ("beta_static", C.c_double), ("beta_min", C.c_double), ("beta_max", C.c_double),
                ("rot_aware", C.c_int), ("reweighting", C.c_int), ("rw_delay", C.c_double),
                ("gamma", C.c_double), ("tol", C.c_double), ("max_admm_iters", C.c_int),
                ("global_max_iters", C.c_int), ("global_ls", C.c_int), ("kappa", C.c_double),
                ("lbfgs_m", C.c_int), ("prox_iters", C.c_int), ("precision", C.c_int),
                ("prox_mode", C.c_int), ("newton_rtol", C.c_double), ("fallback_sigma", C.c_double),
                ("fp64_switch", C.c_double), ("soln_provided", C.c_int), ("verbose", C.c_int), ("device", C.c_int)]


_PROTOS = {
    # wrapd_gpu.h
    "wrapd_create": (C.c_int, [C.POINTER(P), C.c_int]),
    "wrapd_destroy": (C.c_int, [P]),
    "wrapd_last_error": (C.c_char_p, [P]),
    "wrapd_upload_mesh": (C.c_int, [P, C.c_int, D, C.c_int, I, C.c_int, I, C.POINTER(Material), D, D]),
    "wrapd_sizes": (C.c_int, [P, I, I, I, I]),
    "wrapd_set_x_fix": (C.c_int, [P, D]),
    "wrapd_set_x_free": (C.c_int, [P, D]),
    "wrapd_get_x_free": (C.c_int, [P, D]),
    "wrapd_set_x_ref": (C.c_int, [P, D]),
    "wrapd_x_distance": (C.c_int, [P, D]),
    "wrapd_x_distance_launch": (C.c_int, [P]),
    "wrapd_x_distance_result": (C.c_int, [P, D]),
    "wrapd_set_rot_aware": (C.c_int, [P, C.c_int]),
    "wrapd_init_duals": (C.c_int, [P, C.c_int, D]),
    "wrapd_defo_energy": (C.c_int, [P, D, I]),
    "wrapd_candidate_weights": (C.c_int, [P, C.c_int, C.c_int]),
    "wrapd_max_wsq_ratio": (C.c_int, [P, D]),
    "wrapd_update_actual_weights": (C.c_int, [P]),
    "wrapd_local_step": (C.c_int, [P, C.POINTER(LocalArgs), C.c_double, C.POINTER(LocalOut)]),
    "wrapd_local_step_unaware": (C.c_int, [P, C.POINTER(LocalArgs)]),
    "wrapd_local_step_launch": (C.c_int, [P, C.POINTER(LocalArgs)]),
    "wrapd_local_step_result": (C.c_int, [P, C.c_double, C.POINTER(LocalOut)]),
    "wrapd_global_value_grad": (C.c_int, [P, D, D]),
    "wrapd_refactor": (C.c_int, [P]),
    "wrapd_solve3": (C.c_int, [P, D, D]),
    "wrapd_get_A": (C.c_int, [P, I, I, D]),
    "wrapd_chol_stats": (C.c_int, [P, D]),
    "wrapd_lbfgs_configure": (C.c_int, [P, C.POINTER(LbfgsCfg)]),
    "wrapd_lbfgs_reset": (C.c_int, [P]),
    "wrapd_global_step": (C.c_int, [P, C.c_double, C.POINTER(GlobalOut)]),
    "wrapd_global_step_begin": (C.c_int, [P]),
    "wrapd_global_step_unaware": (C.c_int, [P, C.POINTER(GlobalOut)]),
    "wrapd_get_elem": (C.c_int, [P, C.c_char_p, D]),
    "wrapd_state_save": (C.c_int, [P, P, C.c_size_t, C.POINTER(C.c_size_t)]),
    "wrapd_state_load": (C.c_int, [P, P, C.c_size_t]),
    "wrapd_set_elem": (C.c_int, [P, C.c_char_p, D]),
    "wrapd_profile_enable": (C.c_int, [P, C.c_int]),
    "wrapd_profile_only": (C.c_int, [P, C.c_char_p]),
    "wrapd_profile_get": (C.c_int, [P, C.c_char_p, D, D]),
    "wrapd_synchronize": (C.c_int, [P]),
    "wrapd_mlp_init": (C.c_int, [P, C.c_ulonglong]),
    "wrapd_mlp_set_params": (C.c_int, [P, F, F]),
    "wrapd_mlp_get_params": (C.c_int, [P, F, F]),
    "wrapd_mlp_infer": (C.c_int, [P, F, C.c_int, F]),
    "wrapd_mlp_grad": (C.c_int, [P, F, C.c_int, F, F]),
    "wrapd_mlp_grad_noalloc": (C.c_int, [P, F, C.c_int, F, F]),
    "wrapd_mlp_train_full": (C.c_int, [P, F, C.c_int, C.c_int, C.c_float, F]),
    "wrapd_svd3_batch": (C.c_int, [P, D, C.c_int, C.c_int, D, D, D]),
    "wrapd_linsolve_analyze": (C.c_int, [P, C.c_int, I, I, D]),
    "wrapd_linsolve_factorize": (C.c_int, [P, D]),
    "wrapd_add_to_b": (C.c_int, [P, D]),
    "wrapd_mlp_log_start": (C.c_int, [P, C.c_longlong]),
    "wrapd_mlp_log_stop": (C.c_int, [P]),
    "wrapd_mlp_log_count": (C.c_longlong, [P]),
    "wrapd_mlp_log_get": (C.c_int, [P, C.c_longlong, C.c_longlong, F]),
    "wrapd_mlp_log_set": (C.c_int, [P, F, C.c_longlong]),
    "wrapd_mlp_fit_norm": (C.c_int, [P]),
    "wrapd_mlp_augment": (C.c_int, [P, C.c_double, C.c_double, C.c_ulonglong]),
    "wrapd_mlp_train": (C.c_int, [P, C.c_int, C.c_int, C.c_float, C.c_ulonglong, F]),
    # wrapd_solver.h
    "wrapd_solver_default_settings": (None, [C.POINTER(SolverSettings)]),
    "wrapd_solver_create": (C.c_int, [C.POINTER(P), C.POINTER(SolverSettings), C.c_int, D, D, D, C.c_int, I,
                                      C.c_int, I, D, D]),
    "wrapd_solver_destroy": (C.c_int, [P]),
    "wrapd_solver_last_error": (C.c_char_p, [P]),
    "wrapd_solver_initialize": (C.c_int, [P]),
    "wrapd_solver_setup": (C.c_int, [P]),
    "wrapd_solver_iterate": (C.c_int, [P, I]),
    "wrapd_solver_finish": (C.c_int, [P]),
    "wrapd_solver_solve": (C.c_int, [P]),
    "wrapd_solver_num_iters": (C.c_int, [P]),
    "wrapd_solver_checkpoint": (C.c_int, [P, C.c_char_p]),
    "wrapd_solver_resume": (C.c_int, [P, C.c_char_p]),
    "wrapd_solver_get_log": (C.c_int, [P, C.c_char_p, D, C.c_int]),
    "wrapd_solver_get_x": (C.c_int, [P, D]),
    "wrapd_solver_get_x_free": (C.c_int, [P, D]),
    "wrapd_solver_get_runtime": (C.c_int, [P, D]),
    "wrapd_solver_ctx": (P, [P]),
}

_lib = None


def lib():
    """The loaded library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"WRAPD HIP library not built: {LIB_PATH} (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _PROTOS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def exported_symbols():
    return list(_PROTOS)


def dp(a):
    return a.ctypes.data_as(D)


def ip(a):
    return a.ctypes.data_as(I)


def fp(a):
    return None if a is None else a.ctypes.data_as(F)
