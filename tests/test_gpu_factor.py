"""GPU numeric factorization (wrapd_amd/csrc/wd_factor.hip) through the C ABI.

The solve x = A^{-1} b is checked by its residual against A assembled on the device and
by the solution of an independent sparse LU (scipy's SuperLU), on the benchmark beam at
uniform and at strongly varying element weights, and an indefinite A must be reported as
such (the pivot check runs inside the kernel; every kernel still runs to its end).
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from wrapd_amd import meshgen
from wrapd_amd.solver import Device, WrapdError

pytestmark = pytest.mark.gpu


def _device(prob):
    d = Device()
    d.upload_mesh(prob)
    free = np.setdiff1d(np.arange(prob.n_verts), prob.pins)
    d.set_x_fix(prob.init[np.sort(prob.pins)])
    d.set_x_free(prob.init[free])
    d.candidate_weights(use_hessian=False, direct=True)
    d.update_actual_weights()
    return d


def _full_A(d):
    rp, cols, vals = d.A_csr()
    n = len(rp) - 1
    U = sp.csr_matrix((vals, cols, rp), shape=(n, n))
    return (U + sp.triu(U, 1).T).tocsc()


def _check_solve(d, seed):
    d.refactor()
    A = _full_A(d)
    b = np.random.default_rng(seed).standard_normal((A.shape[0], 3))
    x = d.solve3(b)
    r = np.linalg.norm(A @ x - b) / (spla.norm(A) * np.linalg.norm(x) + np.linalg.norm(b))
    x_lu = spla.splu(A).solve(b)
    print(f"n {A.shape[0]} scaled residual {r:.2e} vs LU {np.linalg.norm(x - x_lu) / np.linalg.norm(x_lu):.2e}")
    assert r < 1e-15 * 64
    assert np.linalg.norm(x - x_lu) <= 1e-9 * np.linalg.norm(x_lu)
    return d


def test_factor_benchmark_beam_uniform_weights():
    d = _check_solve(_device(meshgen.config_problem(2)), 0)
    st = d.chol_stats()
    assert st["factor_launches"] > 0 and st["pool_doubles"] > 0


def test_factor_varying_weights_and_refactor_reuse():
    prob = meshgen.config_problem(2)
    d = _device(prob)
    _check_solve(d, 1)
    w = d.elem("wsq")
    scale = np.exp(np.random.default_rng(2).uniform(np.log(0.1), np.log(10.0), len(w)))
    d.set_elem("wsq", w * scale)
    _check_solve(d, 3)   # same symbolic analysis and plan, new values


def test_factor_reports_indefinite_matrix():
    prob = meshgen.small_beam(12, 3, 3, seed=3)
    d = _device(prob)
    w = d.elem("wsq")
    d.set_elem("wsq", -w)
    with pytest.raises(WrapdError) as e:
        d.refactor()
    assert e.value.status == 4   # WRAPD_ERR_NOT_SPD
    d.set_elem("wsq", w)
    _check_solve(d, 4)


# the factorization plan's choices, each forced through the environment at context creation
PLAN_CHOICES = {
    "l21_launches_everywhere": {"WRAPD_L21_RIGHT_MAX": "0"},
    "l21_in_panel_steps": {"WRAPD_L21_RIGHT_MAX": "1e30"},
    "assembly_chunk_64": {"WRAPD_ASM_CHUNK": "64"},
    "assembly_chunk_1024": {"WRAPD_ASM_CHUNK": "1024"},
    "assembly_after_full_join": {"WRAPD_ASM_GROUPS": "0"},
    "diagonal_tiles_all_64x64": {"WRAPD_POTRF16": "0"},
}


def test_plan_choices_agree(monkeypatch):
    """Where L21 is formed (by the panel steps or by per-panel L21 launches), how many rows
    an assembly task builds, whether a level's fronts are assembled in groups overlapping the
    previous level's last FINAL launches, and whether narrow supernodes' diagonal tiles are
    factored by one wave are the plan's choices.  L21 launches apply the panel steps'
    products in their order, every assembly chunking and grouping adds A's entries and then
    the children's columns in the same order, and both diagonal-tile kernels run the same
    16x16 block step, so every choice gives the same factorization, bit for bit (and a race
    between an early assembly and the previous level's work would show as a difference)."""
    prob = meshgen.config_problem(3)   # fronts up to ~4000 rows, levels of both kinds
    b = np.random.default_rng(7).standard_normal((prob.n_verts - len(prob.pins), 3))
    env_names = sorted({k for v in PLAN_CHOICES.values() for k in v})

    def solve(env):
        for k in env_names:
            monkeypatch.delenv(k, raising=False)
        for k, v in env.items():
            monkeypatch.setenv(k, v)
        d = _device(prob)
        d.refactor()
        x = d.solve3(b)
        d.close()
        return x

    ref = solve({})
    for name, env in PLAN_CHOICES.items():
        assert np.array_equal(solve(env), ref), name
