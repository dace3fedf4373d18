"""The reference's C++ interface over the MI355X solver, as a consumer uses it.

tests/csrc/integration_consumer.cpp (built with the library, linked against libwrapd.so,
including only the public headers) runs the reference's untwist sample through Solver,
then the Elements interface through a base-class pointer, System::get_A/get_b with
mcl::optlib::LinearSolver, wrapd::GlobalProblem as a SplitProblem, and
create_tets_from_mesh's vertex_offset.  The LinearSolver's C ABI is also checked on
matrices that come from no mesh.
"""
import os
import subprocess

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from test_oracle_golden import load, trace_bars, untwist_problem
from wrapd_amd import _lib, meshgen
from wrapd_amd.solver import Device, WrapdError

pytestmark = pytest.mark.gpu

CONSUMER = os.path.join(_lib.LIB_DIR, "integration_consumer")


def test_consumer_untwist(tmp_path):
    prob = untwist_problem()
    p = meshgen.write_problem(prob, str(tmp_path))
    out = tmp_path / "out"
    r = subprocess.run([CONSUMER, "-rest", p["rest"], "-init", p["init"], "-handles", p["handles"], "-it", "30",
                        "-tol", "0", "-p", "0", "-s", "1", "-out", str(out)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    f = r.stdout.split()
    kv = dict(zip(f[0::2], f[1::2]))
    print(r.stdout)
    # the solve: the reference's own trace of this sample
    z = load("trace_untwist.npz")
    obj = np.loadtxt(out / "objectives.txt")
    bar_e, _ = trace_bars(z)
    assert int(kv["iters"]) == 30 and int(kv["elements"]) == prob.n_elems
    assert np.max(np.abs(obj[:, 1] - z["objectives"])) <= bar_e * z["objectives"][0]
    # Elements through the base class: Z's energy is the last local step's, the penalty at
    # x is GlobalProblem's value at x
    assert float(kv["polar_pe"]) == pytest.approx(float(kv["last_local_pe"]), rel=1e-10)
    assert float(kv["penalty"]) == pytest.approx(float(kv["global_value"]), rel=1e-12)
    assert float(kv["primal_residual_sq"]) >= 0.0
    assert float(kv["grad_fd_rel_err"]) < 1e-5
    # System::get_A + LinearSolver (ordered without coordinates) on get_b
    assert int(kv["linsolve_info"]) == 0
    assert float(kv["linsolve_residual"]) < 1e-12
    assert int(kv["vertex_offset_ok"]) == 1


def _grid_spd(nx, ny, nz, seed, components=1):
    """A 7-point Laplacian on a grid plus a positive diagonal (an SPD matrix with no mesh);
    with components > 1 the grid is cut into disconnected slabs."""
    rng = np.random.default_rng(seed)
    n = nx * ny * nz
    idx = np.arange(n).reshape(nx, ny, nz)
    rows, cols = [], []
    for ax in range(3):
        a = np.take(idx, np.arange(idx.shape[ax] - 1), axis=ax).ravel()
        b = np.take(idx, np.arange(1, idx.shape[ax]), axis=ax).ravel()
        if components > 1 and ax == 0:
            keep = (np.take(np.arange(nx).reshape(-1, 1, 1) * np.ones((1, ny, nz), int), np.arange(nx - 1),
                            axis=0).ravel() + 1) % (nx // components) != 0
            a, b = a[keep], b[keep]
        rows.append(a)
        cols.append(b)
    r = np.concatenate(rows)
    c = np.concatenate(cols)
    w = rng.uniform(0.1, 10.0, len(r))
    A = sp.coo_matrix((-w, (r, c)), shape=(n, n))
    A = A + A.T
    d = -np.asarray(A.sum(axis=1)).ravel() + rng.uniform(1e-3, 1.0, n)
    A = (A + sp.diags(d)).tocsr()
    U = sp.triu(A).tocsr()
    U.sort_indices()
    xyz = np.stack(np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij"), -1).reshape(-1, 3) * 1.0
    return A, U, xyz


@pytest.mark.parametrize("coords,components", [(False, 1), (True, 1), (False, 3)])
def test_linear_solver_on_a_given_matrix(coords, components):
    A, U, xyz = _grid_spd(30, 24, 20, seed=3, components=components)
    d = Device()
    d.linsolve_analyze(U.indptr, U.indices, xyz if coords else None)
    d.linsolve_factorize(U.data)
    b = np.random.default_rng(0).standard_normal((A.shape[0], 3))
    x = d.solve3(b)
    r = np.linalg.norm(A @ x - b) / (spla.norm(A, 1) * np.linalg.norm(x) + np.linalg.norm(b))
    st = d.chol_stats()
    print(f"coords {coords}, components {components}: scaled residual {r:.2e}, nnz(L) {st['nnz_L']:.3e}")
    assert r < 64 * 2.2e-16
    assert np.linalg.norm(x - spla.spsolve(A.tocsc(), b)) <= 1e-10 * np.linalg.norm(x)
    # refactorization with new values on the same pattern
    d.linsolve_factorize(U.data * 2.0)
    assert np.allclose(d.solve3(b), 0.5 * x, rtol=1e-12, atol=1e-14 * np.abs(x).max())


def test_linear_solver_rejects_bad_input():
    A, U, _ = _grid_spd(6, 5, 4, seed=1)
    d = Device()
    with pytest.raises(WrapdError) as e:   # the lower triangle is not a valid pattern
        L = sp.tril(A).tocsr()
        d.linsolve_analyze(L.indptr, L.indices)
    assert e.value.status == 1
    d.linsolve_analyze(U.indptr, U.indices)
    with pytest.raises(WrapdError) as e:   # indefinite
        d.linsolve_factorize(-U.data)
    assert e.value.status == 4
    d.linsolve_factorize(U.data)           # and usable again after the failure
    b = np.ones((A.shape[0], 3))
    assert np.linalg.norm(A @ d.solve3(b) - b) < 1e-12 * np.linalg.norm(b)
