"""ctypes binding of the CPU FP64 oracle (oracle/_build/liboracle.so).

Test infrastructure: only tests/, __graft_entry__.smoke() and bench.py's CPU
baseline use it, always as the checker or the baseline, never as the product.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "_build", "liboracle.so")


class Settings(C.Structure):
    _fields_ = [
        ("model", C.c_int), ("mu", C.c_double), ("lam", C.c_double),
        ("beta_static", C.c_double), ("beta_min", C.c_double), ("beta_max", C.c_double),
        ("rot_aware", C.c_int), ("reweighting", C.c_int), ("rw_delay", C.c_double),
        ("gamma", C.c_double), ("tol", C.c_double), ("max_admm_iters", C.c_int),
        ("global_max_iters", C.c_int), ("global_ls", C.c_int), ("kappa", C.c_double),
        ("lbfgs_m", C.c_int), ("prox_iters", C.c_int), ("soln_provided", C.c_int),
    ]


_lib = None
_D = C.POINTER(C.c_double)
_I = C.POINTER(C.c_int)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        L.oracle_default_settings.argtypes = [C.POINTER(Settings)]
        L.oracle_svd3.argtypes = [_D, _D, _D, _D]
        L.oracle_nh_psi.argtypes = [_D, C.c_double, C.c_double]
        L.oracle_nh_psi.restype = C.c_double
        L.oracle_nh_hess.argtypes = [_D, C.c_double, C.c_double, _D]
        L.oracle_hess_lambda_max.argtypes = [_D, C.c_double, C.c_double]
        L.oracle_hess_lambda_max.restype = C.c_double
        L.oracle_sym_eig3.argtypes = [_D, _D]
        L.oracle_prox_lbfgs.argtypes = [_D, C.c_double, C.c_double, C.c_double, C.c_int, _D]
        L.oracle_prox_lbfgs.restype = C.c_int
        L.oracle_create.argtypes = [C.POINTER(Settings), C.c_int, _D, _D, _D, C.c_int, _I,
                                    C.c_int, _I, _D, _D]
        L.oracle_create.restype = C.c_void_p
        L.oracle_destroy.argtypes = [C.c_void_p]
        for fn in ("oracle_initialize", "oracle_setup_elements", "oracle_setup_solve", "oracle_iterate", "oracle_finish",
                   "oracle_solve", "oracle_num_iters", "oracle_n_free", "oracle_inversion_count"):
            getattr(L, fn).argtypes = [C.c_void_p]
            getattr(L, fn).restype = C.c_int
        L.oracle_get_log.argtypes = [C.c_void_p, C.c_char_p, _D, C.c_int]
        L.oracle_get_log.restype = C.c_int
        L.oracle_get_x.argtypes = [C.c_void_p, _D]
        L.oracle_get_x_free.argtypes = [C.c_void_p, _D]
        L.oracle_set_x_free.argtypes = [C.c_void_p, _D]
        L.oracle_get_elem.argtypes = [C.c_void_p, C.c_char_p, _D]
        L.oracle_set_elem.argtypes = [C.c_void_p, C.c_char_p, _D]
        L.oracle_get_scalar.argtypes = [C.c_void_p, C.c_char_p]
        L.oracle_get_scalar.restype = C.c_double
        L.oracle_set_scalar.argtypes = [C.c_void_p, C.c_char_p, C.c_double]
        L.oracle_local_step.argtypes = [C.c_void_p, C.c_int]
        L.oracle_local_step.restype = C.c_double
        L.oracle_update_defo_cache.argtypes = [C.c_void_p]
        L.oracle_global_value.argtypes = [C.c_void_p, _D]
        L.oracle_global_value.restype = C.c_double
        L.oracle_global_grad.argtypes = [C.c_void_p, _D, _D]
        L.oracle_global_grad.restype = C.c_double
        L.oracle_potential_energy.argtypes = [C.c_void_p]
        L.oracle_potential_energy.restype = C.c_double
        L.oracle_get_A.argtypes = [C.c_void_p, _I, _I, _D, C.c_int]
        L.oracle_get_A.restype = C.c_int
        L.oracle_solve3.argtypes = [C.c_void_p, _D, _D]
        _lib = L
    return _lib


def dp(a):
    return a.ctypes.data_as(_D)


def ip(a):
    return a.ctypes.data_as(_I)


def default_settings(**kw) -> Settings:
    s = Settings()
    lib().oracle_default_settings(C.byref(s))
    for k, v in kw.items():
        setattr(s, "lam" if k == "lambda" else k, v)
    return s


def settings_for(prob, **over) -> Settings:
    """Settings for a meshgen.Problem: its model/material plus its own settings dict."""
    mu, lam = prob.lame()
    kw = dict(model=1 if prob.model == "nh" else 0, mu=mu, lam=lam)
    kw.update(prob.settings)
    kw.update(over)
    return default_settings(**kw)


def svd3(F):
    F = np.ascontiguousarray(F, dtype=np.float64)
    s, U, V = np.zeros(3), np.zeros((3, 3)), np.zeros((3, 3))
    lib().oracle_svd3(dp(F), dp(s), dp(U), dp(V))
    return s, U, V


def prox_lbfgs(sig_hat, wov, mu, lam, x0, max_iters=20):
    sh = np.ascontiguousarray(sig_hat, dtype=np.float64)
    x = np.array(x0, dtype=np.float64)
    it = lib().oracle_prox_lbfgs(dp(sh), float(wov), float(mu), float(lam), int(max_iters), dp(x))
    return x, it


def hess_lambda_max(s, mu, lam):
    s = np.ascontiguousarray(s, dtype=np.float64)
    return lib().oracle_hess_lambda_max(dp(s), float(mu), float(lam))


def nh_psi(s, mu, lam):
    s = np.ascontiguousarray(s, dtype=np.float64)
    return lib().oracle_nh_psi(dp(s), float(mu), float(lam))


def nh_hess(s, mu, lam):
    s = np.ascontiguousarray(s, dtype=np.float64)
    H = np.zeros((3, 3))
    lib().oracle_nh_hess(dp(s), float(mu), float(lam), dp(H))
    return H


class Oracle:
    """One oracle solver instance (mirrors Solver/ADMM of the reference)."""

    def __init__(self, prob, settings: Settings | None = None, soln=None, **over):
        self.prob = prob
        self.st = settings if settings is not None else settings_for(prob, **over)
        self._keep = []
        rest = np.ascontiguousarray(prob.rest, dtype=np.float64)
        init = np.ascontiguousarray(prob.init, dtype=np.float64)
        tets = np.ascontiguousarray(prob.tets, dtype=np.int32)
        pins = np.ascontiguousarray(prob.pins, dtype=np.int32)
        sol = None if soln is None else np.ascontiguousarray(soln, dtype=np.float64)
        lp = prob.lame_per_elem()
        mu_e = lam_e = None
        if lp is not None:
            mu_e = np.ascontiguousarray(lp[0], dtype=np.float64)
            lam_e = np.ascontiguousarray(lp[1], dtype=np.float64)
        self._keep = [rest, init, tets, pins, sol, mu_e, lam_e]
        self.h = lib().oracle_create(C.byref(self.st), prob.n_verts, dp(rest), dp(init),
                                     None if sol is None else dp(sol), prob.n_elems, ip(tets),
                                     len(pins), ip(pins), None if mu_e is None else dp(mu_e),
                                     None if lam_e is None else dp(lam_e))
        if not self.h:
            raise RuntimeError("oracle_create failed (inverted rest element?)")
        self.n_v, self.n_e = prob.n_verts, prob.n_elems
        self.n_f = lib().oracle_n_free(self.h)

    def __del__(self):
        if getattr(self, "h", None):
            lib().oracle_destroy(self.h)
            self.h = None

    def initialize(self):
        r = lib().oracle_initialize(self.h)
        if r != 0:
            raise RuntimeError(f"oracle_initialize failed: {r}")

    def setup_elements(self):
        """Fixed positions, weights and x at X_init without the linear system (element-level checks)."""
        r = lib().oracle_setup_elements(self.h)
        if r != 0:
            raise RuntimeError(f"oracle_setup_elements failed: {r}")

    def setup_solve(self):
        r = lib().oracle_setup_solve(self.h)
        if r != 0:
            raise RuntimeError(f"oracle_setup_solve failed: {r}")

    def iterate(self) -> int:
        return lib().oracle_iterate(self.h)

    def solve(self) -> int:
        return lib().oracle_solve(self.h)

    def run(self, iters=None):
        """initialize + solve (max iterations from settings unless given)."""
        if iters is not None:
            self.st.max_admm_iters = iters
            # settings were copied at create; the loop bound is read from the copy
        self.initialize()
        self.setup_solve()
        n = 0
        status = 0
        limit = self.st.max_admm_iters if iters is None else iters
        for _ in range(limit):
            status = self.iterate()
            n += 1
            if status != 0:
                break
        lib().oracle_finish(self.h)
        return status

    def log(self, name):
        n = lib().oracle_get_log(self.h, name.encode(), None, 0)
        if n < 0:
            raise KeyError(name)
        out = np.zeros(max(n, 1))
        lib().oracle_get_log(self.h, name.encode(), dp(out), n)
        return out[:n]

    def x(self):
        X = np.zeros((self.n_v, 3))
        lib().oracle_get_x(self.h, dp(X))
        return X

    def x_free(self):
        x = np.zeros((self.n_f, 3))
        lib().oracle_get_x_free(self.h, dp(x))
        return x

    def set_x_free(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        lib().oracle_set_x_free(self.h, dp(x))

    _ELEM_SHAPES = {"U": (9,), "Z": (9,), "F": (9,), "symF": (9,), "wsq": (), "cand_wsq": (),
                    "sigma_last": (3,), "sigma": (3,), "vol": (), "Di": (12,), "mu": (),
                    "lambda": ()}

    def elem(self, name):
        shp = self._ELEM_SHAPES[name]
        out = np.zeros((self.n_e,) + shp)
        lib().oracle_get_elem(self.h, name.encode(), dp(out))
        if shp == (9,):
            return out.reshape(self.n_e, 3, 3)
        if shp == (12,):
            return out.reshape(self.n_e, 3, 4)
        return out

    def set_elem(self, name, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        lib().oracle_set_elem(self.h, name.encode(), dp(a))

    def scalar(self, name):
        return lib().oracle_get_scalar(self.h, name.encode())

    def set_scalar(self, name, v):
        lib().oracle_set_scalar(self.h, name.encode(), float(v))

    def local_step(self, update_cand=True):
        return lib().oracle_local_step(self.h, int(update_cand))

    def update_defo_cache(self):
        lib().oracle_update_defo_cache(self.h)

    def global_value(self, x_free):
        x = np.ascontiguousarray(x_free, dtype=np.float64)
        return lib().oracle_global_value(self.h, dp(x))

    def global_grad(self, x_free):
        x = np.ascontiguousarray(x_free, dtype=np.float64)
        g = np.zeros((self.n_f, 3))
        v = lib().oracle_global_grad(self.h, dp(x), dp(g))
        return v, g

    def potential_energy(self):
        return lib().oracle_potential_energy(self.h)

    def A_coo(self):
        nnz = lib().oracle_get_A(self.h, None, None, None, 0)
        r = np.zeros(nnz, np.int32)
        c = np.zeros(nnz, np.int32)
        v = np.zeros(nnz)
        lib().oracle_get_A(self.h, ip(r), ip(c), dp(v), nnz)
        return r, c, v

    def solve3(self, b):
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros_like(b)
        lib().oracle_solve3(self.h, dp(b), dp(x))
        return x
