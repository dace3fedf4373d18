"""The oracle against the reference's own outputs (tests/golden/, made by make_golden.py
from the reference compiled in oracle/_ref, single-threaded).

Bars: singular values and prox solutions to the precision their algorithms define;
ADMM traces at equal iteration counts: identical inner-iteration counts and reweighting
iterations; energies within max(5e-9, 4 s_E) E_0 and final positions within
max(1e-8, 8 s_x) relative, where s is the larger of two measured round-off spreads of
the reference's own trajectory: 1 vs 8 OpenMP threads, and initial positions perturbed
by 1e-15 relative.  Most traces have s_x below 1e-9; the kappa = 0.1 case amplifies
round-off to s_x = 5e-6 (any independent implementation lands a few s_x away).
"""
import os

import numpy as np
import pytest

import oracle_lib as ol
from wrapd_amd import meshgen

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(G, name), allow_pickle=False)


def test_svd_kat():
    z = load("kat_svd.npz")
    for F, s_ref in zip(z["F"], z["sigma"]):
        s, U, V = ol.svd3(F.reshape(3, 3))
        scale = max(1.0, abs(s_ref[0]))
        assert np.allclose(s, s_ref, atol=1e-13 * scale, rtol=0), (F, s, s_ref)
        assert np.allclose(U @ np.diag(s) @ V.T, F.reshape(3, 3), atol=1e-13 * scale)


def test_prox_kat():
    z = load("kat_prox.npz")
    worst = 0.0
    mism = 0
    for row, x_ref, it_ref in zip(z["inputs"], z["x"], z["iters"]):
        s, wov, mu, lam, last = row[:3], row[3], row[4], row[5], row[6:9]
        x, it = ol.prox_lbfgs(s, wov, mu, lam, last)
        worst = max(worst, float(np.max(np.abs(x - x_ref))))
        mism += int(it != it_ref)
    # the prox stops on |ds| < 1e-8; the same iterate sequence gives the same answer
    assert worst < 1e-9
    assert mism <= 0.01 * len(z["iters"])


def test_hessian_kat():
    z = load("kat_hess.npz")
    for row, k_ref in zip(z["inputs"], z["kmax"]):
        k = ol.hess_lambda_max(row[:3], row[3], row[4])
        assert abs(k - k_ref) <= 1e-11 * abs(k_ref) + 1e-9


TRACES = {
    "c1_arap": lambda: (meshgen.config_problem(1), dict(model=0, reweighting=0)),
    "c1geom_nh": lambda: (meshgen.small_beam(20, 4, 4, seed=1, pin_far_end=False), {}),
    "c1geom_nh_E1e6": lambda: (meshgen.small_beam(20, 4, 4, seed=1, pin_far_end=False, youngs=1e6), {}),
    "c1geom_nh_ls": lambda: (meshgen.small_beam(20, 4, 4, seed=1, pin_far_end=False), dict(global_ls=1)),
    "c1geom_nh_kappa": lambda: (meshgen.small_beam(20, 4, 4, seed=1, pin_far_end=False),
                                dict(kappa=0.1, gamma=1.2, rw_delay=10)),
    "nh_ra0": lambda: (meshgen.small_beam(12, 3, 3, seed=2), dict(rot_aware=0)),
    "arap_ra0": lambda: (meshgen.config_problem(1), dict(model=0, reweighting=0, rot_aware=0)),
    "push_twist_nh": lambda: (meshgen.small_beam(16, 3, 3, seed=4, twist_deg=120.0, axial=-0.3, youngs=1e6), {}),
}


def untwist_problem():
    z = load("trace_untwist.npz")
    return meshgen.Problem("untwist", z["rest"], z["init"], z["tets"], z["pins"])


def trace_bars(z):
    """Energy (relative to E_0) and position bars from the reference's round-off spreads."""
    s_e = max(float(z["spread_e"]), float(z["pert_e"]))
    s_x = max(float(z["spread_x"]), float(z["pert_x"]))
    return max(5e-9, 4.0 * s_e), max(1e-8, 8.0 * s_x)


def compare_trace(name, prob, over):
    z = load(f"trace_{name}.npz")
    iters = len(z["objectives"]) - 1
    o = ol.Oracle(prob, tol=0.0, **over)
    o.run(iters)
    e, e_ref = o.log("objectives"), z["objectives"]
    bar_e, bar_x = trace_bars(z)
    assert len(e) == len(e_ref)
    assert np.max(np.abs(e - e_ref)) <= bar_e * abs(e_ref[0]), np.max(np.abs(e - e_ref)) / abs(e_ref[0])
    assert np.array_equal(o.log("inner_iters").astype(int), z["inner_iters"])
    assert np.array_equal(o.log("reweighted").astype(int), z["reweighted"])
    x = o.x()[z["x_rows"]]
    assert np.linalg.norm(x - z["x"]) <= bar_x * np.linalg.norm(z["x"])


@pytest.mark.parametrize("name", sorted(TRACES))
def test_trace_matches_reference(name):
    prob, over = TRACES[name]()
    compare_trace(name, prob, over)


@pytest.mark.slow
def test_untwist_matches_reference():
    compare_trace("untwist", untwist_problem(), {})


def test_uniform_per_element_material_is_the_homogeneous_run():
    """The heterogeneous-material extension (per-element mu, lambda; no reference
    counterpart) with every element at the same E reproduces the reference's homogeneous
    trace (its -E path, E = 1e6)."""
    prob = meshgen.small_beam(20, 4, 4, seed=1, pin_far_end=False, youngs=1e6)
    prob.youngs_per_elem = np.full(prob.n_elems, 1e6)
    compare_trace("c1geom_nh_E1e6", prob, {})


def test_heterogeneous_oracle_fixture_reproduces():
    """The oracle regenerates the first 30 iterations of its heterogeneous-beam fixture
    (tests/golden/make_golden.py voronoi) within the fixture's own round-off envelope."""
    z = load("trace_voronoi_oracle.npz")
    prob = meshgen.small_voronoi_beam(64, 8, 8, regions=16, twist_deg=360.0, axial=-0.2)
    assert np.array_equal(prob.youngs_per_elem, z["youngs_per_elem"])
    o = ol.Oracle(prob, tol=0.0)
    o.run(30)
    e, eo = o.log("objectives"), z["objectives"][:31]
    env = np.maximum.accumulate(np.abs(z["objectives"] - z["objectives_perturbed"]) / abs(eo[0]))[:31]
    assert np.all(np.abs(e - eo) / abs(eo[0]) <= np.maximum(5e-9, 8.0 * env))
