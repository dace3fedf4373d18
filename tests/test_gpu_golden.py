"""The GPU solver against the reference's own outputs (tests/golden/), through the C ABI,
and the `quasi` CLI on the reference's sample data.  Same bars as the oracle's golden tests:
equal inner-iteration counts and reweighting iterations, energies within
max(5e-9, 4 s_E) E_0, positions within max(1e-8, 8 s_x), s being the reference's measured
round-off spread of the same run (tests/test_oracle_golden.py).
"""
import os
import subprocess

import numpy as np
import pytest

from test_oracle_golden import TRACES, load, trace_bars, untwist_problem
from wrapd_amd import _lib, meshgen
from wrapd_amd.solver import Solver

pytestmark = pytest.mark.gpu

CASES = dict(TRACES)
CASES["untwist"] = lambda: (untwist_problem(), {})
CASES["c2_it20"] = lambda: (meshgen.config_problem(2), dict(lbfgs_m=6, max_admm_iters=5000))


@pytest.mark.parametrize("name", sorted(CASES))
def test_gpu_trace_matches_reference(name):
    prob, over = CASES[name]()
    z = load(f"trace_{name}.npz")
    iters = len(z["objectives"]) - 1
    s = Solver(prob, tol=0.0, **over)
    s.run(iters)
    e, e_ref = s.log("objectives"), z["objectives"]
    bar_e, bar_x = trace_bars(z)
    x = s.x()[z["x_rows"]]
    dx = np.linalg.norm(x - z["x"]) / np.linalg.norm(z["x"])
    print(f"{name}: max|dE|/E0 {np.max(np.abs(e - e_ref)) / abs(e_ref[0]):.2e} (bar {bar_e:.1e}), "
          f"|dx|/|x| {dx:.2e} (bar {bar_x:.1e})")
    assert len(e) == len(e_ref)
    assert np.array_equal(s.log("inner_iters").astype(int), z["inner_iters"])
    assert np.array_equal(s.log("reweighted").astype(int), z["reweighted"])
    assert np.max(np.abs(e - e_ref)) <= bar_e * abs(e_ref[0])
    assert dx <= bar_x


def test_quasi_cli_untwist(tmp_path):
    """quasi on the reference's untwist sample writes the reference's logs and meshes."""
    prob = untwist_problem()
    p = meshgen.write_problem(prob, str(tmp_path))
    out = tmp_path / "out"
    r = subprocess.run([_lib.QUASI_PATH, "-rest", p["rest"] + ".node", "-init", p["init"] + ".node", "-handles",
                        p["handles"], "-it", "30", "-tol", "0", "-p", "0", "-s", "1", "-out", str(out)],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    z = load("trace_untwist.npz")
    obj = np.loadtxt(out / "objectives.txt")
    assert obj.shape == (31, 2) and np.array_equal(obj[:, 0], np.arange(31))
    bar_e, bar_x = trace_bars(z)
    assert np.max(np.abs(obj[:, 1] - z["objectives"])) <= bar_e * z["objectives"][0]
    rw = np.loadtxt(out / "reweighted.txt")
    assert np.array_equal(rw[:, 1].astype(int), z["reweighted"])
    X, T = meshgen.load_elenode(str(out / "final_mesh"))
    assert np.array_equal(T, prob.tets)
    assert np.linalg.norm(X - z["x"]) <= bar_x * np.linalg.norm(z["x"])
    assert (out / "final_mesh.obj").exists()


def test_fp32_local_step_final_positions_within_1e6():
    """Mixed precision (FP32 local steps with a Newton prox, FP64 from residual 1e-8 on)
    reaches the FP64 fixed point: both runs stop on the reference's own criterion
    (|dE|/(|E|+1) < 1e-12, no flips) on a pushed and twisted beam (the oracle converges it
    in 272 iterations); final positions and energy within 1e-6 of the FP64 run."""
    prob = meshgen.small_beam(10, 3, 3, seed=9, twist_deg=90.0, axial=-0.2, youngs=1e6)
    ref = Solver(prob, tol=1e-12)
    n_ref = ref.run(3000)
    fast = Solver(prob, tol=1e-12, precision=1, prox_mode=1, fp64_switch=1e-8)
    n_fast = fast.run(3000)
    assert n_ref < 3000 and n_fast < 3000
    rt = fast.runtime()
    assert 0 < rt["fp64_switch_iter"] < n_fast
    xr, xf = ref.x(), fast.x()
    print(f"FP64 {n_ref} iterations, mixed {n_fast} (FP64 from {rt['fp64_switch_iter']:.0f}): "
          f"x rel err {np.linalg.norm(xf - xr) / np.linalg.norm(xr):.2e}")
    assert np.linalg.norm(xf - xr) <= 1e-6 * np.linalg.norm(xr)
    er, ef = ref.log("objectives"), fast.log("objectives")
    assert abs(ef[-1] - er[-1]) <= 1e-6 * abs(er[-1])


def test_newton_fp64_prox_same_fixed_point():
    prob = meshgen.small_beam(16, 3, 3, seed=4, twist_deg=120.0, axial=-0.3, youngs=1e6)
    ref = Solver(prob, tol=0.0)
    ref.run(200)
    nt = Solver(prob, tol=0.0, prox_mode=1)
    nt.run(200)
    assert np.linalg.norm(nt.x() - ref.x()) <= 1e-7 * np.linalg.norm(ref.x())


def test_gpu_uniform_per_element_material_is_the_homogeneous_run():
    """Per-element material arrays (the heterogeneous extension) at a uniform E follow the
    reference's homogeneous trace (E = 1e6) on the GPU, to the trace bars."""
    prob = meshgen.small_beam(20, 4, 4, seed=1, pin_far_end=False, youngs=1e6)
    prob.youngs_per_elem = np.full(prob.n_elems, 1e6)
    z = load("trace_c1geom_nh_E1e6.npz")
    s = Solver(prob, tol=0.0)
    s.run(len(z["objectives"]) - 1)
    bar_e, bar_x = trace_bars(z)
    e = s.log("objectives")
    assert np.array_equal(s.log("inner_iters").astype(int), z["inner_iters"])
    assert np.array_equal(s.log("reweighted").astype(int), z["reweighted"])
    assert np.max(np.abs(e - z["objectives"])) <= bar_e * abs(z["objectives"][0])
    assert np.linalg.norm(s.x()[z["x_rows"]] - z["x"]) <= bar_x * np.linalg.norm(z["x"])


def test_soln_x_errors_and_early_exit():
    """-soln (src/AlgorithmData.cpp:63,70, src/ADMM.cpp:164): the log of |x - x*| / |x0 - x*|
    over the free vertices, and the stop at the first iterate below 1e-4.  x* is the
    beam's converged solution; the log's first iterations are checked against the oracle."""
    import oracle_lib as ol
    prob = meshgen.small_beam(12, 3, 3, seed=5, youngs=1e6)
    ref = Solver(prob, tol=1e-12)
    assert ref.run(3000) < 3000
    xstar = ref.x()
    s = Solver(prob, soln=xstar, tol=0.0)
    n = s.run(3000)
    xe = s.log("x_errors")
    assert n < 3000 and len(xe) == n + 1
    assert xe[-1] < 1e-4 and np.all(xe[:-1] >= 1e-4)   # stopped at the first iterate below 1e-4
    o = ol.Oracle(prob, soln=xstar, tol=0.0)
    o.run(10)
    xo = o.log("x_errors")
    assert np.allclose(xe[:11], xo[:11], rtol=1e-7, atol=0.0)


@pytest.mark.parametrize("prox", ["newton", "lbfgs"])
def test_fp32_local_step_redoes_fallbacks_in_fp64(prox):
    """An FP32 local step defers the elements whose FP32 prox result is unusable (near
    inversion) to a second pass that redoes them in FP64, 64 per wave.  From one state (the
    Voronoi beam after 60 iterations, elements inverted), an FP32 step and an FP64 step with
    the fallback's prox settings: exactly the fallback elements come out bit for bit equal
    to the FP64 step's (U, P, the prox solution, the candidate weight).  With the fallback
    threshold above every singular value, every element is redone: all of them then equal
    the FP64 step's, and the step's sums (now all the second pass's) equal the FP64 step's
    up to summation order."""
    prob = meshgen.small_voronoi_beam(64, 8, 8, regions=16, twist_deg=360.0, axial=-0.2)
    s = Solver(prob, tol=0.0, precision=1, prox_mode=1, fp64_switch=0.0)
    s.initialize()
    s.setup()
    for _ in range(60):
        s.iterate()
    d = s.device()
    blob = d.state_save()
    names = ("U", "P", "sigma_last", "cand_wsq")
    kw = dict(pe_prev=1.0, prox=prox, newton_rtol=1e-12, update_cand=True, use_hessian=True)

    def step(**over):
        d.state_load(blob)
        o = d.local_step(**dict(kw, **over))
        return o, {k: d.elem(k).reshape(prob.n_elems, -1) for k in names}

    o64, a64 = step(precision="fp64")
    o32, a32 = step(precision="fp32")
    same = np.all([np.all(a32[k] == a64[k], axis=1) for k in names], axis=0)
    print(f"{prox}: {o32['fallback_count']:.0f} fallbacks, {same.sum()} elements equal to the FP64 step", flush=True)
    assert o32["fallback_count"] > 0 and o64["fallback_count"] == 0
    assert same.sum() == o32["fallback_count"]
    assert o32["bef"] == o64["bef"]
    assert o32["evals_total"] > o32["prox_iters_total"] > 0
    oall, aall = step(precision="fp32", fallback_sigma=1e300)
    assert oall["fallback_count"] == prob.n_elems
    assert all(np.array_equal(aall[k], a64[k]) for k in names)
    for k in ("aft", "pe", "prox_iters_total", "neg_sigma_count"):
        assert oall[k] == pytest.approx(o64[k], rel=1e-12, abs=0.0), k
    assert oall["max_wsq_ratio"] == o64["max_wsq_ratio"]
