"""Python front end of the MI355X WRAPD solver (thin ctypes layer over include/*.h).

`Solver` mirrors the reference's Solver/ADMM driving (samples/tests/quasi.cpp:63-97):
create from a mesh + settings, initialize, solve (or setup + iterate stepwise), read the
AlgorithmData logs and the final positions.  `Device` exposes the kernel-level C ABI
(include/wrapd_gpu.h) for element/vertex-level checks.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import dp, fp, ip, lib

MODEL = {"arap": 0, "nh": 1}
PRECISION = {"fp64": 0, "fp32": 1}
PROX = {"lbfgs": 0, "newton": 1, "learned": 2}


class WrapdError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"[status {status}] {msg}")
        self.status = status


def default_settings(**kw) -> _lib.SolverSettings:
    s = _lib.SolverSettings()
    lib().wrapd_solver_default_settings(C.byref(s))
    for k, v in kw.items():
        if k == "model" and isinstance(v, str):
            v = MODEL[v]
        elif k == "precision" and isinstance(v, str):
            v = PRECISION[v]
        elif k == "prox_mode" and isinstance(v, str):
            v = PROX[v]
        setattr(s, k, v)
    return s


def settings_for(prob, **over) -> _lib.SolverSettings:
    kw = dict(model=MODEL[prob.model], youngs=float(prob.youngs), poisson=float(prob.poisson))
    kw.update(prob.settings)
    kw.update(over)
    return default_settings(**kw)


class Solver:
    """One solve on the GPU (owns a device context)."""

    def __init__(self, prob, settings=None, soln=None, **over):
        self.prob = prob
        self.st = settings if settings is not None else settings_for(prob, **over)
        rest = np.ascontiguousarray(prob.rest, dtype=np.float64)
        init = np.ascontiguousarray(prob.init, dtype=np.float64)
        tets = np.ascontiguousarray(prob.tets, dtype=np.int32)
        pins = np.ascontiguousarray(prob.pins, dtype=np.int32)
        sol = None if soln is None else np.ascontiguousarray(soln, dtype=np.float64)
        if sol is not None:
            self.st.soln_provided = 1
        mu_e = lam_e = None
        lp = prob.lame_per_elem()
        if lp is not None:
            mu_e = np.ascontiguousarray(lp[0], dtype=np.float64)
            lam_e = np.ascontiguousarray(lp[1], dtype=np.float64)
        self._keep = [rest, init, tets, pins, sol, mu_e, lam_e]
        self.h = None
        h = C.c_void_p()
        st = lib().wrapd_solver_create(C.byref(h), C.byref(self.st), prob.n_verts, dp(rest), dp(init),
                                       None if sol is None else dp(sol), prob.n_elems, ip(tets), len(pins),
                                       ip(pins), None if mu_e is None else dp(mu_e),
                                       None if lam_e is None else dp(lam_e))
        self.h = h
        self._check(st, "create")
        self.n_v, self.n_e = prob.n_verts, prob.n_elems

    def _check(self, st, what):
        if st != 0:
            msg = lib().wrapd_solver_last_error(self.h).decode() if self.h else ""
            raise WrapdError(st, f"{what}: {msg}")

    def close(self):
        if getattr(self, "h", None):
            lib().wrapd_solver_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def initialize(self):
        self._check(lib().wrapd_solver_initialize(self.h), "initialize")

    def setup(self):
        self._check(lib().wrapd_solver_setup(self.h), "setup")

    def iterate(self) -> bool:
        conv = C.c_int(0)
        self._check(lib().wrapd_solver_iterate(self.h, C.byref(conv)), "iterate")
        return bool(conv.value)

    def finish(self):
        self._check(lib().wrapd_solver_finish(self.h), "finish")

    def checkpoint(self, path):
        """Write the whole ADMM state after the last iterate() to `path`."""
        self._check(lib().wrapd_solver_checkpoint(self.h, str(path).encode()), "checkpoint")

    def resume(self, path):
        """After setup(): load a checkpoint of a run with the same mesh and settings; the
        following iterate() calls continue that run bit for bit."""
        self._check(lib().wrapd_solver_resume(self.h, str(path).encode()), "resume")

    def solve(self):
        self._check(lib().wrapd_solver_solve(self.h), "solve")

    def run(self, iters=None, stop_on_converged=True):
        """initialize + setup + up to `iters` iterations (settings' -it by default) + finish."""
        self.initialize()
        self.setup()
        n = self.st.max_admm_iters if iters is None else iters
        done = 0
        for _ in range(n):
            c = self.iterate()
            done += 1
            if c and stop_on_converged:
                break
        self.finish()
        return done

    @property
    def num_iters(self):
        return lib().wrapd_solver_num_iters(self.h)

    def log(self, name):
        n = lib().wrapd_solver_get_log(self.h, name.encode(), None, 0)
        if n < 0:
            raise KeyError(name)
        out = np.zeros(max(n, 1))
        lib().wrapd_solver_get_log(self.h, name.encode(), dp(out), n)
        return out[:n]

    def x(self):
        X = np.zeros((self.n_v, 3))
        self._check(lib().wrapd_solver_get_x(self.h, dp(X)), "get_x")
        return X

    def x_free(self):
        n_f = self.n_v - len(self.prob.pins)
        x = np.zeros((n_f, 3))
        self._check(lib().wrapd_solver_get_x_free(self.h, dp(x)), "get_x_free")
        return x

    def runtime(self):
        o = np.zeros(11)
        lib().wrapd_solver_get_runtime(self.h, dp(o))
        keys = ["setup_ms", "local_ms", "global_ms", "refactor_ms", "inner_iters", "finish_ms", "prox_iters",
                "fallback_elems", "learned_accepted", "fp64_switch_iter", "learned_used"]
        return dict(zip(keys, o.tolist()))

    def device(self):
        return Device(lib().wrapd_solver_ctx(self.h), owner=self)


class Device:
    """Kernel-level access (include/wrapd_gpu.h) to one context."""

    _ELEM = {"U": 9, "P": 9, "Z": 9, "symF": 9, "wsq": 1, "cand_wsq": 1, "sigma_last": 3, "vol": 1, "Di": 12}

    def __init__(self, ctx=None, owner=None, device=0):
        self.owner = owner
        self._own = ctx is None
        self.ctx = None
        if ctx is None:
            h = C.c_void_p()
            st = lib().wrapd_create(C.byref(h), device)
            if st != 0:
                raise WrapdError(st, "wrapd_create failed (no HIP device?)")
            ctx = h.value
        self.ctx = C.c_void_p(ctx)
        self._keep = []

    def close(self):
        if getattr(self, "_own", False) and getattr(self, "ctx", None):
            lib().wrapd_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        self.close()

    def _check(self, st, what):
        if st != 0:
            raise WrapdError(st, f"{what}: {lib().wrapd_last_error(self.ctx).decode()}")

    def sizes(self):
        a, b, c, d = (C.c_int() for _ in range(4))
        lib().wrapd_sizes(self.ctx, C.byref(a), C.byref(b), C.byref(c), C.byref(d))
        return a.value, b.value, c.value, d.value

    def upload_mesh(self, prob, beta_static=1.0, beta_min=0.1, beta_max=10.0):
        mu, lam = prob.lame()
        m = _lib.Material(MODEL[prob.model], mu, lam, beta_static, beta_min, beta_max)
        rest = np.ascontiguousarray(prob.rest, dtype=np.float64)
        tets = np.ascontiguousarray(prob.tets, dtype=np.int32)
        pins = np.ascontiguousarray(prob.pins, dtype=np.int32)
        lp = prob.lame_per_elem()
        mu_e = lam_e = None
        if lp is not None:
            mu_e = np.ascontiguousarray(lp[0], dtype=np.float64)
            lam_e = np.ascontiguousarray(lp[1], dtype=np.float64)
        self._keep = [rest, tets, pins, mu_e, lam_e]
        self._check(lib().wrapd_upload_mesh(self.ctx, prob.n_verts, dp(rest), prob.n_elems, ip(tets), len(pins),
                                            ip(pins), C.byref(m), None if mu_e is None else dp(mu_e),
                                            None if lam_e is None else dp(lam_e)), "upload_mesh")

    def set_x_fix(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        self._check(lib().wrapd_set_x_fix(self.ctx, dp(x)), "set_x_fix")

    def set_x_free(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        self._check(lib().wrapd_set_x_free(self.ctx, dp(x)), "set_x_free")

    def x_free(self):
        n_f = self.sizes()[2]
        x = np.zeros((n_f, 3))
        self._check(lib().wrapd_get_x_free(self.ctx, dp(x)), "get_x_free")
        return x

    def set_rot_aware(self, rot_aware=True):
        self._check(lib().wrapd_set_rot_aware(self.ctx, int(rot_aware)), "set_rot_aware")

    def init_duals(self, rot_aware=True):
        pe = C.c_double()
        self._check(lib().wrapd_init_duals(self.ctx, int(rot_aware), C.byref(pe)), "init_duals")
        return pe.value

    def defo_energy(self):
        e, f = C.c_double(), C.c_int()
        self._check(lib().wrapd_defo_energy(self.ctx, C.byref(e), C.byref(f)), "defo_energy")
        return e.value, f.value

    def candidate_weights(self, use_hessian, direct=False):
        self._check(lib().wrapd_candidate_weights(self.ctx, int(use_hessian), int(direct)), "candidate_weights")

    def max_wsq_ratio(self):
        r = C.c_double()
        self._check(lib().wrapd_max_wsq_ratio(self.ctx, C.byref(r)), "max_wsq_ratio")
        return r.value

    def update_actual_weights(self):
        self._check(lib().wrapd_update_actual_weights(self.ctx), "update_actual_weights")

    def local_step(self, pe_prev=0.0, precision="fp64", prox="lbfgs", prox_iters=20, update_cand=True,
                   use_hessian=False, newton_rtol=1e-10, fallback_sigma=0.2):
        a = _lib.LocalArgs(PRECISION[precision], PROX[prox], prox_iters, int(update_cand), int(use_hessian),
                           newton_rtol, fallback_sigma)
        o = _lib.LocalOut()
        st = lib().wrapd_local_step(self.ctx, C.byref(a), pe_prev, C.byref(o))
        self._check(st, "local_step")
        return {k: getattr(o, k) for k, _ in _lib.LocalOut._fields_}

    def value_grad(self):
        n_f = self.sizes()[2]
        v = C.c_double()
        g = np.zeros((n_f, 3))
        self._check(lib().wrapd_global_value_grad(self.ctx, C.byref(v), dp(g)), "value_grad")
        return v.value, g

    def add_to_b(self):
        n_f = self.sizes()[2]
        b = np.zeros((n_f, 3))
        self._check(lib().wrapd_add_to_b(self.ctx, dp(b)), "add_to_b")
        return b

    def linsolve_analyze(self, row_ptr, cols, xyz=None):
        """Symbolic analysis of a caller-given SPD matrix's upper-triangle pattern (no mesh)."""
        rp = np.ascontiguousarray(row_ptr, dtype=np.int32)
        cl = np.ascontiguousarray(cols, dtype=np.int32)
        x = None if xyz is None else np.ascontiguousarray(xyz, dtype=np.float64)
        self._keep = [rp, cl, x]
        self._check(lib().wrapd_linsolve_analyze(self.ctx, len(rp) - 1, ip(rp), ip(cl), None if x is None else dp(x)),
                    "linsolve_analyze")

    def linsolve_factorize(self, vals):
        v = np.ascontiguousarray(vals, dtype=np.float64)
        self._check(lib().wrapd_linsolve_factorize(self.ctx, dp(v)), "linsolve_factorize")

    def refactor(self):
        self._check(lib().wrapd_refactor(self.ctx), "refactor")

    def solve3(self, b):
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros_like(b)
        self._check(lib().wrapd_solve3(self.ctx, dp(b), dp(x)), "solve3")
        return x

    def A_csr(self):
        n_f = self.sizes()[2]
        nnz = lib().wrapd_get_A(self.ctx, None, None, None)
        rp = np.zeros(n_f + 1, np.int32)
        cols = np.zeros(nnz, np.int32)
        vals = np.zeros(nnz)
        lib().wrapd_get_A(self.ctx, ip(rp), ip(cols), dp(vals))
        return rp, cols, vals

    def chol_stats(self):
        s = np.zeros(12)
        self._check(lib().wrapd_chol_stats(self.ctx, dp(s)), "chol_stats")
        return dict(zip(["nsuper", "nnz_L", "flops", "max_front", "leaf_tasks", "levels", "g_entries",
                         "factor_levels", "factor_launches", "factor_tile_flops", "pool_doubles", "flops_g"],
                        s.tolist()))

    def lbfgs_configure(self, memory=6, max_iters=100, kappa=1.0, linesearch=0):
        c = _lib.LbfgsCfg(memory, max_iters, kappa, linesearch)
        self._check(lib().wrapd_lbfgs_configure(self.ctx, C.byref(c)), "lbfgs_configure")

    def lbfgs_reset(self):
        self._check(lib().wrapd_lbfgs_reset(self.ctx), "lbfgs_reset")

    def global_step(self, tol):
        o = _lib.GlobalOut()
        self._check(lib().wrapd_global_step(self.ctx, float(tol), C.byref(o)), "global_step")
        return {k: getattr(o, k) for k, _ in _lib.GlobalOut._fields_}

    def state_save(self):
        """The device-resident ADMM state as bytes (wrapd_state_save)."""
        n = C.c_size_t(0)
        self._check(lib().wrapd_state_save(self.ctx, None, 0, C.byref(n)), "state_save")
        buf = (C.c_char * n.value)()
        self._check(lib().wrapd_state_save(self.ctx, buf, n.value, C.byref(n)), "state_save")
        return bytes(buf)

    def state_load(self, blob):
        buf = C.create_string_buffer(blob, len(blob))
        self._check(lib().wrapd_state_load(self.ctx, buf, len(blob)), "state_load")

    def elem(self, name):
        n_e = self.sizes()[1]
        k = self._ELEM[name]
        out = np.zeros((n_e, k))
        self._check(lib().wrapd_get_elem(self.ctx, name.encode(), dp(out)), f"get_elem {name}")
        if k == 9:
            return out.reshape(n_e, 3, 3)
        if k == 12:
            return out.reshape(n_e, 3, 4)
        return out if k > 1 else out[:, 0]

    def set_elem(self, name, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        self._check(lib().wrapd_set_elem(self.ctx, name.encode(), dp(a)), f"set_elem {name}")

    def svd3(self, F, precision="fp64"):
        """The device SVD of (n, 3, 3) matrices: sigma (n,3), U, V (n,3,3)."""
        F = np.ascontiguousarray(F, dtype=np.float64).reshape(-1, 9)
        n = len(F)
        s, U, V = np.zeros((n, 3)), np.zeros((n, 9)), np.zeros((n, 9))
        self._check(lib().wrapd_svd3_batch(self.ctx, dp(F), n, PRECISION[precision], dp(s), dp(U), dp(V)), "svd3")
        return s, U.reshape(n, 3, 3), V.reshape(n, 3, 3)

    def profile(self, on=True, only=None):
        """Times device work with HIP events; only: one scope name (None: all scopes)."""
        self._check(lib().wrapd_profile_only(self.ctx, only.encode() if only else None), "profile_only")
        self._check(lib().wrapd_profile_enable(self.ctx, int(on)), "profile_enable")

    def profile_get(self, kernel):
        t, n = C.c_double(), C.c_double()
        lib().wrapd_profile_get(self.ctx, kernel.encode(), C.byref(t), C.byref(n))
        return t.value, n.value

    def synchronize(self):
        self._check(lib().wrapd_synchronize(self.ctx), "synchronize")

    # ---- learned prox warm start (wd_mlp.hip) ----
    N_PARAMS, N_NORM = 4803, 18

    def mlp_init(self, seed=0):
        self._check(lib().wrapd_mlp_init(self.ctx, seed), "mlp_init")

    def mlp_set_params(self, params, norm):
        p = np.ascontiguousarray(params, dtype=np.float32)
        n = np.ascontiguousarray(norm, dtype=np.float32)
        assert p.size == self.N_PARAMS and n.size == self.N_NORM
        self._check(lib().wrapd_mlp_set_params(self.ctx, fp(p), fp(n)), "mlp_set_params")

    def mlp_get_params(self):
        p = np.zeros(self.N_PARAMS, np.float32)
        n = np.zeros(self.N_NORM, np.float32)
        self._check(lib().wrapd_mlp_get_params(self.ctx, fp(p), fp(n)), "mlp_get_params")
        return p, n

    def mlp_infer(self, inputs):
        x = np.ascontiguousarray(inputs, dtype=np.float32).reshape(-1, 6)
        out = np.zeros((len(x), 3), np.float32)
        self._check(lib().wrapd_mlp_infer(self.ctx, fp(x), len(x), fp(out)), "mlp_infer")
        return out

    def mlp_grad(self, tuples):
        t = np.ascontiguousarray(tuples, dtype=np.float32).reshape(-1, 9)
        g = np.zeros(self.N_PARAMS, np.float32)
        loss = np.zeros(1, np.float32)
        self._check(lib().wrapd_mlp_grad(self.ctx, fp(t), len(t), fp(g), fp(loss)), "mlp_grad")
        return float(loss[0]), g

    def mlp_grad_noalloc(self, tuples):
        """mlp_grad through the entry point that allocates no device memory (<= 65536 tuples)."""
        t = np.ascontiguousarray(tuples, dtype=np.float32).reshape(-1, 9)
        g = np.zeros(self.N_PARAMS, np.float32)
        loss = np.zeros(1, np.float32)
        self._check(lib().wrapd_mlp_grad_noalloc(self.ctx, fp(t), len(t), fp(g), fp(loss)), "mlp_grad_noalloc")
        return float(loss[0]), g

    def mlp_log_start(self, capacity):
        self._check(lib().wrapd_mlp_log_start(self.ctx, int(capacity)), "mlp_log_start")

    def mlp_log_stop(self):
        self._check(lib().wrapd_mlp_log_stop(self.ctx), "mlp_log_stop")

    def mlp_log_count(self):
        return int(lib().wrapd_mlp_log_count(self.ctx))

    def mlp_log_get(self, offset=0, n=None):
        if n is None:
            n = self.mlp_log_count() - offset
        out = np.zeros((n, 9), np.float32)
        self._check(lib().wrapd_mlp_log_get(self.ctx, int(offset), int(n), fp(out)), "mlp_log_get")
        return out

    def mlp_log_set(self, tuples):
        t = np.ascontiguousarray(tuples, dtype=np.float32).reshape(-1, 9)
        self._check(lib().wrapd_mlp_log_set(self.ctx, fp(t), len(t)), "mlp_log_set")

    def mlp_fit_norm(self):
        self._check(lib().wrapd_mlp_fit_norm(self.ctx), "mlp_fit_norm")

    def mlp_augment(self, w_lo=0.0, w_hi=0.0, seed=0):
        self._check(lib().wrapd_mlp_augment(self.ctx, float(w_lo), float(w_hi), seed), "mlp_augment")

    def mlp_train_full(self, tuples, steps, lr=1e-3):
        """Full-batch Adam steps on the given tuples; returns the loss of each step."""
        t = np.ascontiguousarray(tuples, dtype=np.float32).reshape(-1, 9)
        hist = np.zeros(max(steps, 1), np.float32)
        self._check(lib().wrapd_mlp_train_full(self.ctx, fp(t), len(t), int(steps), float(lr), fp(hist)),
                    "mlp_train_full")
        return hist[:steps]

    def mlp_train(self, steps, batch=8192, lr=1e-3, seed=0):
        hist = np.zeros(max(steps, 1), np.float32)
        self._check(lib().wrapd_mlp_train(self.ctx, int(steps), int(batch), float(lr), seed, fp(hist)), "mlp_train")
        return hist[:steps]
