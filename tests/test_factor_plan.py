"""The GPU factorization schedule (wrapd_amd/csrc/chol_plan.cpp) on the CPU.

tests/csrc/plan_emu.cpp executes the plan's launches task by task with the tiles the HIP
kernels use.  Checked: G_s = [L11^{-1}; -L21 L11^{-1}] for every supernode against a
dense LAPACK Cholesky of the permuted matrix, no two tasks of one launch touching the
same entry with one of them writing, front-pool reuse never handing out live memory
(the pool starts as NaN), and an indefinite matrix flagged.
"""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import oracle_lib as ol
from wrapd_amd import meshgen

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def emu(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("planemu") / "libplanemu.so")
    src = [os.path.join(ROOT, "tests", "csrc", "plan_emu.cpp"),
           os.path.join(ROOT, "wrapd_amd", "csrc", "chol_symbolic.cpp"),
           os.path.join(ROOT, "wrapd_amd", "csrc", "chol_plan.cpp")]
    subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-o", out] + src, check=True)
    lib = C.CDLL(out)
    lib.pe_run.restype = C.c_int
    lib.pe_set_options.argtypes = [C.c_double, C.c_int, C.c_int, C.c_int]
    return lib


def _upper_csr(prob):
    o = ol.Oracle(prob)
    o.initialize()
    r, c, v = o.A_coo()
    n = o.n_f
    m = r <= c
    r, c, v = r[m], c[m], v[m]
    order = np.lexsort((c, r))
    r, c, v = r[order], c[order], v[order]
    Ap = np.zeros(n + 1, np.int32)
    np.add.at(Ap, r + 1, 1)
    return n, np.cumsum(Ap).astype(np.int32), np.ascontiguousarray(c.astype(np.int32)), np.ascontiguousarray(v)


def _run(emu, prob, leaf_size, negate=False, coords=True, options=(256.0, 512, 1, 1)):
    n, Ap, Aj, Ax = _upper_csr(prob)
    emu.pe_set_options(*options)
    free = np.setdiff1d(np.arange(prob.n_verts), prob.pins)
    st = np.zeros(10)
    xyz = ol.dp(np.ascontiguousarray(prob.rest[free])) if coords else None
    emu.pe_analyze(n, ol.ip(Ap), ol.ip(Aj), xyz, leaf_size, ol.dp(st))
    nsup = int(st[0])
    perm = np.zeros(n, np.int32)
    col0 = np.zeros(nsup + 1, np.int32)
    fsize = np.zeros(nsup, np.int32)
    goff = np.zeros(nsup + 1, np.int64)
    rowptr = np.zeros(nsup + 1, np.int64)
    emu.pe_layout(ol.ip(perm), ol.ip(col0), ol.ip(fsize), goff.ctypes.data_as(C.c_void_p),
                  rowptr.ctypes.data_as(C.c_void_p), None)
    rows = np.zeros(int(rowptr[-1]), np.int32)
    emu.pe_layout(ol.ip(perm), ol.ip(col0), ol.ip(fsize), goff.ctypes.data_as(C.c_void_p),
                  rowptr.ctypes.data_as(C.c_void_p), ol.ip(rows))
    G = np.zeros(int(goff[-1]))
    conflicts = C.c_int64(0)
    fail = emu.pe_run(ol.dp(-Ax if negate else Ax), ol.dp(G), C.byref(conflicts))
    A = np.zeros((n, n))
    rr = np.repeat(np.arange(n), np.diff(Ap))
    A[rr, Aj] = Ax
    A = A + np.triu(A, 1).T
    return dict(st=st, perm=perm, col0=col0, fsize=fsize, goff=goff, rowptr=rowptr, rows=rows, G=G, fail=fail,
                conflicts=conflicts.value, A=A)


CASES = {
    "tiny": (lambda: meshgen.small_beam(6, 3, 3, seed=2), 48),
    "deep_tree": (lambda: meshgen.small_beam(24, 5, 4, seed=3), 8),
    "two_panel_separators": (lambda: meshgen.small_beam(16, 10, 10, seed=4), 48),
    "ragged": (lambda: meshgen.small_beam(9, 7, 5, seed=5, twist_deg=30.0), 16),
    # the ordering without vertex coordinates (breadth-first level structures of the graph),
    # as LinearSolver uses it for a matrix it is given
    "graph_order": (lambda: meshgen.small_beam(16, 10, 10, seed=4), 48),
    "graph_order_deep": (lambda: meshgen.small_beam(24, 5, 4, seed=3), 8),
}


# the plan's choices (L21 placement threshold, rows per assembly task): the defaults, L21
# launches on every level, L21 in the panel steps everywhere, 64-row assembly chunks
MODES = {"default": (256.0, 512, 1, 1), "left": (0.0, 512, 1, 1), "right": (1e30, 512, 1, 1),
         "chunk64": (256.0, 64, 1, 1), "joined": (256.0, 512, 0, 0)}


@pytest.mark.parametrize("mode", sorted(MODES))
@pytest.mark.parametrize("name", sorted(CASES))
def test_plan_matches_dense_cholesky(emu, name, mode):
    make, leaf = CASES[name]
    R = _run(emu, make(), leaf, coords=not name.startswith("graph_order"), options=MODES[mode])
    if mode == "left":
        assert R["st"][9] > 0   # L21 launches
    assert R["fail"] == 0
    assert R["conflicts"] == 0
    if name == "two_panel_separators":
        assert R["st"][3] > 64   # a supernode spans more than one 64-column panel
        assert R["st"][7] >= 1   # and a FINAL launch overlaps a wider supernode's panels
    if name.startswith("graph_order") and mode != "joined":
        assert R["st"][8] >= 1   # an assembly launch overlaps the previous level's later FINAL launches
    print(f"{name}: {R['st'][7]:.0f} launches on the second stream, {R['st'][8]:.0f} assembly launches "
          "waiting for part of it")
    p = R["perm"]
    L = np.linalg.cholesky(R["A"][np.ix_(p, p)])
    worst = 0.0
    for s in range(len(R["fsize"])):
        c0, c1, f = R["col0"][s], R["col0"][s + 1], R["fsize"][s]
        ns = c1 - c0
        rows = R["rows"][R["rowptr"][s]:R["rowptr"][s + 1]]
        assert np.array_equal(rows[:ns], np.arange(c0, c1))
        Linv = np.linalg.inv(L[c0:c1, c0:c1])
        expect = np.vstack([Linv, -L[np.ix_(rows[ns:], np.arange(c0, c1))] @ Linv])
        ld = (f + 15) // 16 * 16   # G's leading dimension: the front size rounded up to 16 rows
        got = R["G"][R["goff"][s]:R["goff"][s + 1]].reshape(ns, ld).T[:f]
        worst = max(worst, float(np.max(np.abs(got - expect)) / np.max(np.abs(expect))))
    assert worst < 1e-12, worst
    # every factor entry outside the supernodes' structure is zero
    Ls = np.zeros_like(L)
    for s in range(len(R["fsize"])):
        c0, c1 = R["col0"][s], R["col0"][s + 1]
        rows = R["rows"][R["rowptr"][s]:R["rowptr"][s + 1]]
        Ls[np.ix_(rows, np.arange(c0, c1))] = L[np.ix_(rows, np.arange(c0, c1))]
    assert np.allclose(np.tril(L), np.tril(Ls), atol=1e-12 * np.abs(L).max())


def test_plan_choices_same_bits(emu):
    """The left-looking L21 launches apply the right-looking panel steps' products in their
    order and assembly chunks add the same values in the same order, so G comes out bit for bit
    the same whichever the plan chooses."""
    make, leaf = CASES["two_panel_separators"]
    prob = make()
    Gs = [_run(emu, prob, leaf, options=MODES[m])["G"] for m in ("left", "right", "chunk64", "joined")]
    for g in Gs[1:]:
        assert np.array_equal(Gs[0], g, equal_nan=True)   # (G's row padding is never written)


def test_plan_pool_reuse_and_indefinite(emu):
    R = _run(emu, meshgen.small_beam(24, 5, 4, seed=3), 8, negate=True)
    assert R["fail"] == 1
    assert R["st"][5] < R["st"][6]   # reuse makes the pool smaller than the sum of fronts
