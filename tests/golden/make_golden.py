"""Writes the golden vectors under tests/golden/ from the reference build (oracle/_ref).

Run in a container where the reference tree is present, after `make -C oracle ref`:
    OMP_NUM_THREADS=1 python tests/golden/make_golden.py [all|base|samples|samples_pert|c3|voronoi]
The reference runs single-threaded (its OpenMP reductions are order dependent).  Only
inputs and outputs are stored; the mesh inputs of the sample case come from the
reference's own data files (samples/data/*).
"""
from __future__ import annotations

import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
from wrapd_amd import meshgen  # noqa: E402

REF = os.path.join(ROOT, "oracle", "_ref", "ref_driver")
REF_TREE = "/root/reference"
ENV = dict(os.environ, OMP_NUM_THREADS="1")


def ref_lines(mode, rows):
    inp = "\n".join(" ".join(f"{v:.17g}" for v in r) for r in rows) + "\n"
    out = subprocess.run([REF, mode], input=inp, capture_output=True, text=True, env=ENV, check=True).stdout
    return np.array([[float(t) for t in line.split()] for line in out.strip().split("\n")])


def lame(E, nu):
    return E / (2 * (1 + nu)), E * nu / ((1 + nu) * (1 - 2 * nu))


def kat_svd(rng):
    F = []
    for i in range(300):
        kind = i % 6
        if kind == 0:
            A = rng.standard_normal((3, 3))
        elif kind == 1:   # near identity (typical deformation gradients)
            A = np.eye(3) + 0.2 * rng.standard_normal((3, 3))
        elif kind == 2:   # inverted
            A = np.eye(3) + 0.2 * rng.standard_normal((3, 3))
            A[:, 2] *= -1
        elif kind == 3:   # repeated singular values (rotation times scaled identity)
            Q, _ = np.linalg.qr(rng.standard_normal((3, 3)))
            A = Q * rng.uniform(0.5, 2.0)
        elif kind == 4:   # rank deficient
            A = rng.standard_normal((3, 2)) @ rng.standard_normal((2, 3))
        else:             # strongly anisotropic
            U, _ = np.linalg.qr(rng.standard_normal((3, 3)))
            V, _ = np.linalg.qr(rng.standard_normal((3, 3)))
            A = U @ np.diag([3.0, 0.7, 0.05]) @ V.T
        F.append(A.reshape(-1))
    F = np.array(F)
    out = ref_lines("svd", F)
    return dict(F=F, sigma=out[:, :3], U=out[:, 3:12], V=out[:, 12:21])


def kat_prox(rng):
    rows = []
    for i in range(600):
        E = [1e5, 1e6, 1e7][i % 3]
        nu = [0.3, 0.45, 0.49][(i // 3) % 3]
        mu, lam = lame(E, nu)
        k = 2.0066 * mu + 1.0122 * lam
        s = np.sort(rng.uniform(0.1, 2.5, 3))[::-1]
        if i % 10 == 0:
            s[2] = -rng.uniform(0.0, 0.3)
        wov = k * np.exp(rng.uniform(np.log(0.1), np.log(10.0)))
        last = np.ones(3) if i % 4 == 0 else np.abs(s) * rng.uniform(0.85, 1.15, 3) + 0.01
        rows.append([s[0], s[1], s[2], wov, mu, lam, last[0], last[1], last[2]])
    rows = np.array(rows)
    out = ref_lines("prox", rows)
    return dict(inputs=rows, x=out[:, :3], iters=out[:, 3].astype(int))


def kat_hess(rng):
    rows = []
    for i in range(300):
        E = [1e5, 1e6, 1e7][i % 3]
        mu, lam = lame(E, 0.45)
        s = rng.uniform(0.05, 2.0, 3)
        if i % 5 == 0:   # the J < J0 branch
            s = rng.uniform(0.05, 0.4, 3)
        rows.append([s[0], s[1], s[2], mu, lam])
    rows = np.array(rows)
    out = ref_lines("hess", rows)
    return dict(inputs=rows, kmax=out[:, 0])


def perturbed(prob, seed=7, rel=1e-15):
    """The same problem with the initial positions moved by ~1e-15 relative (round-off size)."""
    rng = np.random.default_rng(seed)
    init = prob.init * (1.0 + rel * rng.uniform(-1.0, 1.0, prob.init.shape))
    return meshgen.Problem(prob.name, prob.rest, init, prob.tets, prob.pins, prob.model, prob.youngs, prob.poisson,
                           prob.youngs_per_elem, dict(prob.settings))


def run_trace(prob, flags, youngs=None, threads=1):
    with tempfile.TemporaryDirectory() as d:
        p = meshgen.write_problem(prob, d)
        cmd = [REF, "solve", "-rest", p["rest"], "-init", p["init"], "-handles", p["handles"], "-p", "0", "-s", "0",
               "-out", d, "-tol", "0"] + flags
        if youngs is not None:
            cmd += ["-E", repr(youngs)]
        subprocess.run(cmd, env=dict(ENV, OMP_NUM_THREADS=str(threads)), check=True, capture_output=True)
        rd = lambda n: np.atleast_1d(np.loadtxt(os.path.join(d, n)))  # noqa: E731
        return dict(objectives=rd("objectives.txt"), reweighted=rd("reweighted.txt").astype(int),
                    inner_iters=rd("inner_iters.txt").astype(int), ls_dal=rd("ls_dal.txt"),
                    gs_dal=rd("gs_dal.txt"), x=np.loadtxt(os.path.join(d, "final_x.txt")))


def save_case(name, prob, flags, youngs=None, store_mesh=False, x_every=1):
    t = run_trace(prob, flags, youngs)
    # the reference's own run-to-run spread: the same run on 8 OpenMP threads
    t8 = run_trace(prob, flags, youngs, threads=8)
    spread_e = float(np.max(np.abs(t["objectives"] - t8["objectives"])) / abs(t["objectives"][0]))
    spread_x = float(np.linalg.norm(t["x"] - t8["x"]) / np.linalg.norm(t["x"]))
    # and its sensitivity to round-off: the run from initial positions perturbed by 1e-15
    tp = run_trace(perturbed(prob), flags, youngs)
    pert_e = float(np.max(np.abs(t["objectives"] - tp["objectives"])) / abs(t["objectives"][0]))
    pert_x = float(np.linalg.norm(t["x"] - tp["x"]) / np.linalg.norm(t["x"]))
    arrs = dict(flags=np.array(" ".join(flags)), x_rows=np.arange(0, prob.n_verts, x_every),
                spread_e=spread_e, spread_x=spread_x, pert_e=pert_e, pert_x=pert_x,
                spread_inner_equal=bool(np.array_equal(t["inner_iters"], t8["inner_iters"])),
                pert_inner_equal=bool(np.array_equal(t["inner_iters"], tp["inner_iters"])))
    arrs.update({k: v for k, v in t.items() if k != "x"})
    arrs["x"] = t["x"][::x_every]
    if store_mesh:
        arrs.update(rest=prob.rest, init=prob.init, tets=prob.tets.astype(np.int32), pins=prob.pins.astype(np.int32))
    np.savez_compressed(os.path.join(HERE, f"trace_{name}.npz"), **arrs)
    print(name, "iters", len(t["objectives"]) - 1, "E0", t["objectives"][0],
          "spread_e %.2e spread_x %.2e pert_e %.2e pert_x %.2e" % (spread_e, spread_x, pert_e, pert_x), flush=True)


SAMPLES = ("shear", "stretch", "dancer", "dino", "botijo", "homer", "cat", "centaur", "dinoskel", "rabbit")


def sample_inputs(case):
    """The reference's own sample data (samples/data/<case>): rest, init and the shipped
    neo-Hookean solution, tets, pins."""
    d = os.path.join(REF_TREE, "samples", "data", case)
    rest, tets = meshgen.load_elenode(os.path.join(d, f"{case}_rest"))
    init, _ = meshgen.load_elenode(os.path.join(d, f"{case}_init"))
    soln, _ = meshgen.load_elenode(os.path.join(d, f"{case}_nh_soln"))
    pins = np.sort(meshgen.load_handles(os.path.join(d, f"{case}_handles.txt")))
    return rest, init, soln, tets.astype(np.int32), pins.astype(np.int32)


def run_soln(case, threads, init=None):
    """The reference with its defaults and -soln: stops at the first x-error below 1e-4
    (src/ADMM.cpp:164, src/AlgorithmData.cpp:70).  With `init` given, those initial
    positions (on the case's tets) replace the case's own."""
    d = os.path.join(REF_TREE, "samples", "data", case)
    with tempfile.TemporaryDirectory() as out:
        init_path = os.path.join(d, f"{case}_init")
        if init is not None:
            _, tets = meshgen.load_elenode(init_path)
            init_path = os.path.join(out, "init_perturbed")
            meshgen.save_elenode(init_path, init, tets)
        cmd = [REF, "solve", "-rest", os.path.join(d, f"{case}_rest"), "-init", init_path,
               "-soln", os.path.join(d, f"{case}_nh_soln"), "-handles", os.path.join(d, f"{case}_handles.txt"),
               "-p", "0", "-s", "0", "-out", out]
        subprocess.run(cmd, env=dict(ENV, OMP_NUM_THREADS=str(threads)), check=True, capture_output=True)
        rd = lambda n: np.atleast_1d(np.loadtxt(os.path.join(out, n)))  # noqa: E731
        return dict(objectives=rd("objectives.txt"), reweighted=rd("reweighted.txt").astype(int),
                    inner_iters=rd("inner_iters.txt").astype(int), x_errors=rd("x_errors.txt"),
                    x=np.loadtxt(os.path.join(out, "final_x.txt")))


def samples():
    """One fixture per sample case: its input meshes and the reference's run to the -soln
    exit, single-threaded, with the 8-thread run's exit iteration and deviation."""
    for case in SAMPLES:
        if os.path.exists(os.path.join(HERE, f"sample_{case}.npz")) and "--redo" not in sys.argv:
            continue
        rest, init, soln, tets, pins = sample_inputs(case)
        t1 = run_soln(case, 1)
        t8 = run_soln(case, 8)
        n = min(len(t1["objectives"]), len(t8["objectives"]))
        spread_e = float(np.max(np.abs(t1["objectives"][:n] - t8["objectives"][:n])) / abs(t1["objectives"][0]))
        spread_x = float(np.linalg.norm(t1["x"] - t8["x"]) / np.linalg.norm(t1["x"]))
        np.savez_compressed(os.path.join(HERE, f"sample_{case}.npz"), rest=rest, init=init, soln=soln, tets=tets,
                            pins=pins, objectives=t1["objectives"], reweighted=t1["reweighted"],
                            inner_iters=t1["inner_iters"], x_errors=t1["x_errors"], x=t1["x"],
                            iters=len(t1["objectives"]) - 1, iters_8threads=len(t8["objectives"]) - 1,
                            spread_e=spread_e, spread_x=spread_x)
        print(case, "iters", len(t1["objectives"]) - 1, "(8 threads:", len(t8["objectives"]) - 1, ")",
              "spread_e %.2e spread_x %.2e" % (spread_e, spread_x), flush=True)


PERT_SEEDS = (7, 11, 13)


def samples_pert():
    """Adds to each sample fixture the reference's sensitivity to round-off: its runs from
    initial positions moved by ~1e-15 relative, one per seed of PERT_SEEDS (exit
    iterations; the largest energy, x-error and final-position deviations over the seeds)."""
    for case in SAMPLES:
        path = os.path.join(HERE, f"sample_{case}.npz")
        z = dict(np.load(path))
        done = [int(v) for v in np.atleast_1d(z.get("pert_seeds", [7] if "pert_iters" in z else []))]
        todo = [sd for sd in PERT_SEEDS if sd not in done or "--redo" in sys.argv]
        if not todo:
            continue
        iters = [int(v) for v in np.atleast_1d(z.get("pert_iters", []))] if "--redo" not in sys.argv else []
        pe, pxe, px = (float(z.get("pert_e", 0.0)), float(z.get("pert_xe", 0.0)), float(z.get("pert_x", 0.0))) \
            if "--redo" not in sys.argv else (0.0, 0.0, 0.0)
        if "--redo" in sys.argv:
            done = []
        for sd in todo:
            rng = np.random.default_rng(sd)
            init = z["init"] * (1.0 + 1e-15 * rng.uniform(-1.0, 1.0, z["init"].shape))
            tp = run_soln(case, 1, init=init)
            n = min(len(z["objectives"]), len(tp["objectives"]))
            iters.append(len(tp["objectives"]) - 1)
            pe = max(pe, float(np.max(np.abs(z["objectives"][:n] - tp["objectives"][:n])) / abs(z["objectives"][0])))
            pxe = max(pxe, float(np.max(np.abs(z["x_errors"][:n] - tp["x_errors"][:n]))))
            px = max(px, float(np.linalg.norm(z["x"] - tp["x"]) / np.linalg.norm(z["x"])))
            done.append(sd)
        z.update(pert_iters=np.array(iters), pert_seeds=np.array(done), pert_e=pe, pert_xe=pxe, pert_x=px)
        np.savez_compressed(path, **z)
        print(case, "perturbed: iters", iters, "(unperturbed", int(z["iters"]), ")",
              "pert_e %.2e pert_xe %.2e pert_x %.2e" % (pe, pxe, px), flush=True)


def c3():
    """The 32x32x240 beam (40% push + 180 degree twist, E = 1e6), 10 iterations at the
    reference's defaults (its L-BFGS memory 6)."""
    save_case("c3_it10", meshgen.config_problem(3), ["-it", "10"], youngs=1e6, x_every=101)


def voronoi_problem():
    """8x8x64 cells, 16 Voronoi regions of E in {1e5, 1e7}, 360 degree twist and 20% push:
    the 3.07M-tet benchmark beam's load on a 24,576-tet beam."""
    return meshgen.small_voronoi_beam(64, 8, 8, regions=16, twist_deg=360.0, axial=-0.2)


def voronoi():
    """The heterogeneous beam has no reference counterpart (the reference has one
    material): its fixture is the oracle's run, 100 iterations, with the oracle's own
    sensitivity to round-off (the run from initial positions moved by 1e-15 relative)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as ol
    prob = voronoi_problem()
    runs = []
    for pr in (prob, perturbed(prob)):
        o = ol.Oracle(pr, tol=0.0)
        o.run(100)
        runs.append(dict(objectives=o.log("objectives"), reweighted=o.log("reweighted").astype(int),
                         inner_iters=o.log("inner_iters").astype(int), flips=o.log("flips").astype(int), x=o.x()))
    t, tp = runs
    # the run is chaotic: the iterations over which the perturbed run stays within 1e-9 E0
    dev = np.abs(t["objectives"] - tp["objectives"]) / abs(t["objectives"][0])
    k_det = int(np.argmax(dev > 1e-9)) if np.any(dev > 1e-9) else len(dev)
    pert_e = float(np.max(np.abs(t["objectives"] - tp["objectives"])) / abs(t["objectives"][0]))
    pert_x = float(np.linalg.norm(t["x"] - tp["x"]) / np.linalg.norm(t["x"]))
    np.savez_compressed(os.path.join(HERE, "trace_voronoi_oracle.npz"), youngs_per_elem=prob.youngs_per_elem,
                        pert_e=pert_e, pert_x=pert_x, k_det=k_det, objectives_perturbed=tp["objectives"],
                        pert_inner_equal=bool(np.array_equal(t["inner_iters"], tp["inner_iters"])), **t)
    e = t["objectives"]
    print("voronoi E0 %.4e max %.4e end %.4e reweights %d pert_e %.2e pert_x %.2e, within 1e-9 E0 for %d iterations" % (
        e[0], e.max(), e[-1], t["reweighted"].sum(), pert_e, pert_x, k_det))


def main():
    if not os.path.exists(REF):
        sys.exit("build the reference first: make -C oracle ref")
    rng = np.random.default_rng(20211)
    np.savez_compressed(os.path.join(HERE, "kat_svd.npz"), **kat_svd(rng))
    np.savez_compressed(os.path.join(HERE, "kat_prox.npz"), **kat_prox(rng))
    np.savez_compressed(os.path.join(HERE, "kat_hess.npz"), **kat_hess(rng))
    c1 = meshgen.config_problem(1)
    save_case("c1_arap", c1, ["-model", "arap", "-rw", "0", "-it", "100"])
    nh = meshgen.small_beam(20, 4, 4, seed=1, pin_far_end=False)
    save_case("c1geom_nh", nh, ["-it", "100"])
    save_case("c1geom_nh_E1e6", nh, ["-it", "80"], youngs=1e6)
    save_case("c1geom_nh_ls", nh, ["-it", "60", "-global_ls", "1"])
    save_case("c1geom_nh_kappa", nh, ["-it", "60", "-kappa", "0.1", "-gamma", "1.2", "-rw_delay", "10"])
    sb = meshgen.small_beam(12, 3, 3, seed=2)
    save_case("nh_ra0", sb, ["-ra", "0", "-it", "40"])
    save_case("arap_ra0", c1, ["-model", "arap", "-rw", "0", "-ra", "0", "-it", "40"])
    push = meshgen.small_beam(16, 3, 3, seed=4, twist_deg=120.0, axial=-0.3)
    save_case("push_twist_nh", push, ["-it", "80"], youngs=1e6)
    # the reference's own sample: untwist (samples/data/untwist), 30 iterations
    ut = os.path.join(REF_TREE, "samples", "data", "untwist")
    if os.path.isdir(ut):
        rest, tets = meshgen.load_elenode(os.path.join(ut, "untwist_rest"))
        init, _ = meshgen.load_elenode(os.path.join(ut, "untwist_init"))
        soln, _ = meshgen.load_elenode(os.path.join(ut, "untwist_nh_soln"))
        pins = np.sort(meshgen.load_handles(os.path.join(ut, "untwist_handles.txt")))
        prob = meshgen.Problem("untwist", rest, init, tets, pins)
        save_case("untwist", prob, ["-it", "30"], store_mesh=True)
        np.savez_compressed(os.path.join(HERE, "untwist_soln.npz"), soln=soln)
    # the benchmark beam at the reference's defaults except E (-E 1e6), 20 iterations
    c2 = meshgen.config_problem(2)
    save_case("c2_it20", c2, ["-it", "20"], youngs=1e6, x_every=37)


if __name__ == "__main__":
    what = sys.argv[1] if len(sys.argv) > 1 else "all"
    if what in ("all", "base"):
        main()
    if what in ("all", "samples"):
        samples()
    if what in ("all", "samples", "samples_pert"):
        samples_pert()
    if what in ("all", "c3"):
        c3()
    if what in ("all", "voronoi"):
        voronoi()
