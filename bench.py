"""Benchmark: WRAPD ADMM on the twisted beam, one MI355X.

Default workload: the 16x16x160-cell beam (245,760 tets, 46,529 vertices), neo-Hookean
E = 1e6, nu = 0.45, both ends pinned, one twisted 360 degrees, dynamic reweighting,
L-BFGS memory 5 -- the plain-FP64 comparison beam and, with the learned warm start,
the learned-prox case on the same mesh and load.  `--config 5` runs the 3.07M-tet
heterogeneous beam instead.

One "step" is one ADMM iteration of the whole mesh (local step + L-BFGS global step with
the supernodal Cholesky preconditioner + reweighting, refactorization when it fires),
state resident in HBM.

Paths, each a fresh solve of `--steps` iterations from the same initial state:
  plain_fp64     the reference algorithm on the GPU (FP64 local step, L-BFGS prox)
  fp32_newton    FP32 local step (FP64 fallback) with the Newton prox, FP64 from
                 residual 1e-8 on
  learned_fp32   as fp32_newton, the Newton prox warm-started by the trained MLP
The MLP is trained first (untimed): tuples logged from a plain FP64 run of the same beam,
w^2/vol redrawn log-uniformly over the logged range and re-solved, Adam on the GPU.

Reported (one JSON line on stdout):
  value        seconds until the learned path's energy gap (E_k - E*)/|E*| <= 1e-6,
               E* = the lowest energy the plain FP64 run reaches (its converged value);
  paths        the same for every path, its ms/iteration and local-step elements/s;
  learned      training cost, how often the MLP start was used, Newton iterations saved;
  roofline     the dominant device phase: algorithmic bytes per launch / mean launch time
               (HIP events on the solver's stream), against 8 TB/s;
  cpu_baseline the reference (oracle/_ref/ref_driver) on this host's cores, sampled.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
BF16_PEAK_TFLOPS = 2500.0
SCOPES = ("local_step", "defo_grad", "gather_grad", "solve_fwd", "solve_bwd", "multidot", "axpy_q", "axpy_r", "xstep",
          "dot", "refactor", "assemble_A", "update_weights", "mlp_train")


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def algorithmic_bytes(prob, stats):
    """Unique bytes per launch of each profiled phase (each array read or written once)."""
    n_e, n_f = prob.n_elems, prob.n_verts - len(prob.pins)
    hetero = prob.youngs_per_elem is not None
    # local step per element: symF, U, P (3 x 72), w^2, vol, k (3 x 8), last s (24)
    # [+ mu, lambda], writes U, P (2 x 72), s (24), candidate w^2 (8)
    local = n_e * (3 * 72 + 3 * 8 + 24 + (16 if hetero else 0) + 2 * 72 + 24 + 8)
    # value + gradient pass per element: vertex indices and staging rows (2 x 16), D^-1 96,
    # P 72, w^2 8, x gather; each free corner's gradient (24) written to the vertex-major
    # staging, which the vertex gather reads back as contiguous ranges (plus its row pointers)
    # before writing the gradient
    free_corners = int(np.isin(np.asarray(prob.tets), np.asarray(prob.pins), invert=True).sum())
    defo = n_e * (2 * 16 + 96 + 72 + 8) + 24 * free_corners + 8 * 3 * n_f
    gather = 24 * free_corners + 4 * (n_f + 1) + 8 * 3 * n_f
    g = 8.0 * stats["g_entries"]
    vec = 8 * 3 * n_f
    m = int(prob.settings.get("lbfgs_m") or 0)
    return {
        "local_step": local,
        "defo_grad": defo,
        "gather_grad": gather,
        # G once, b (the gradient), y and the update rows, and the m history vectors y_j
        # from which the forward solve forms the two-loop's q = g - sum alpha_j y_j
        "solve_fwd": g + (3 + m) * vec,
        "solve_bwd": g + 3 * vec,      # G once, y, x
    }


def run_path(prob, iters, gap_ref=None, mlp=None, log_capacity=0, only=None, **over):
    """One solve of `iters` ADMM iterations, timed from the first iteration on.  The device
    work is timed with HIP events around every scope (only=None) or around the one scope
    named: each timed launch costs two event records in the stream, ~10% of an iteration
    when every scope is timed, so the wall-clock run times one scope and a second run of the
    same solve gives the per-kernel breakdown."""
    from wrapd_amd.solver import Solver
    s = Solver(prob, tol=0.0, max_admm_iters=iters, **over)
    dev = s.device()
    if mlp is not None:
        dev.mlp_set_params(*mlp)
    if log_capacity:
        dev.mlp_log_start(log_capacity)
    t0 = time.perf_counter()
    s.initialize()
    s.setup()
    t_setup = time.perf_counter() - t0
    dev.profile(True, only=only)
    dev.synchronize()
    walls = []
    t_start = time.perf_counter()
    for _ in range(iters):
        s.iterate()
        walls.append(time.perf_counter() - t_start)
    dev.synchronize()
    t_total = time.perf_counter() - t_start
    prof = {}
    for name in SCOPES:
        ms, n = dev.profile_get(name)
        if n > 0:
            prof[name] = dict(ms=ms, launches=int(n))
    dev.profile(False)
    res = dict(iters=iters, setup_s=t_setup, total_s=t_total, E=s.log("objectives"), wall=np.array(walls),
               runtime=s.runtime(), inner=s.log("inner_iters"), reweighted=s.log("reweighted"), profile=prof,
               chol=dev.chol_stats())
    s.finish()
    res["x"] = s.x()
    res["solver"] = s
    return res


def time_to_gap(res, e_star, target):
    gap = (res["E"][1:] - e_star) / abs(e_star)
    ok = np.nonzero(gap <= target)[0]
    if len(ok) == 0:
        return None, None, float(gap[-1])
    k = int(ok[0])
    return float(res["wall"][k]), k + 1, float(gap[k])


def train_mlp(prob, data_iters, train_steps, batch, lr):
    """Untimed: log tuples from a plain FP64 run, redraw w, fit, train (all on the GPU)."""
    t0 = time.perf_counter()
    res = run_path(prob, data_iters, log_capacity=prob.n_elems * data_iters)
    s = res["solver"]
    dev = s.device()
    n = dev.mlp_log_count()
    t_log = time.perf_counter() - t0
    t1 = time.perf_counter()
    dev.mlp_init(seed=5)
    dev.mlp_augment(seed=7)
    dev.mlp_fit_norm()
    hist = dev.mlp_train(train_steps, batch=batch, lr=lr, seed=9)
    dev.synchronize()
    t_train = time.perf_counter() - t1
    ms, nl = dev.profile_get("mlp_train")
    params, norm = dev.mlp_get_params()
    s.close()
    flops = 3 * 9344.0 * batch * train_steps   # forward + backward ~ 3x the forward
    return (params, norm), dict(tuples=n, data_run_s=t_log, train_s=t_train, train_steps=train_steps, batch=batch,
                                lr=lr, loss_first=float(hist[0]) if len(hist) else None,
                                loss_last=float(np.mean(hist[-20:])) if len(hist) else None,
                                train_tflops=flops / t_train / 1e12 if t_train > 0 else None)


def cpu_baseline(prob, n_iter_sample, iters_to_gap):
    """The reference (oracle/_ref/ref_driver) on this host, or the port (oracle) if absent."""
    ref = os.path.join(ROOT, "oracle", "_ref", "ref_driver")
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    from wrapd_amd import meshgen
    if os.path.exists(ref):
        with tempfile.TemporaryDirectory() as d:
            paths = meshgen.write_problem(prob, d)
            cmd = [ref, "solve", "-rest", paths["rest"], "-init", paths["init"], "-handles", paths["handles"],
                   "-p", "0", "-s", "0", "-out", d, "-it", str(n_iter_sample), "-tol", "0", "-E", repr(prob.youngs)]
            env = dict(os.environ, OMP_NUM_THREADS=str(threads))
            t0 = time.perf_counter()
            r = subprocess.run(cmd, env=env, capture_output=True, text=True, timeout=900)
            wall = time.perf_counter() - t0
            if r.returncode != 0:
                return dict(error=r.stderr[-300:])
            parts = r.stdout.split()
            acc = float(parts[1]) + float(parts[3]) + float(parts[5])
            per_iter = acc / n_iter_sample
        kind = "reference"
        sample = (f"reference ref_driver (Eigen SimplicialLDLT + AMD; PARDISO absent), {n_iter_sample} ADMM "
                  f"iterations of the same beam, OMP_NUM_THREADS={threads}: local+global+refactor {acc:.1f} s "
                  f"(wall {wall:.1f} s incl. mesh load and analysis; its L-BFGS memory is the compile-time 6, the GPU "
                  f"paths use 5); value = that per-iteration time x the plain path's iterations to the gap")
    else:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib
        o = oracle_lib.Oracle(prob, tol=0.0)
        o.initialize()
        o.setup_solve()
        t0 = time.perf_counter()
        for _ in range(n_iter_sample):
            o.iterate()
        per_iter = (time.perf_counter() - t0) / n_iter_sample
        kind = "port"
        sample = f"oracle port, {n_iter_sample} ADMM iterations; value = per-iteration time x iterations to gap"
    return dict(value=per_iter * iters_to_gap if iters_to_gap else None, unit="s", cores=threads, kind=kind,
                sample=sample, s_per_iter=per_iter)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300, help="ADMM iterations per path")
    ap.add_argument("--warmup", type=int, default=5, help="untimed iterations of a throw-away solve")
    ap.add_argument("--config", type=int, default=2)
    ap.add_argument("--gap", type=float, default=1e-6)
    ap.add_argument("--data-iters", type=int, default=200, help="logged plain iterations for the MLP's data")
    ap.add_argument("--train-steps", type=int, default=3000)
    ap.add_argument("--batch", type=int, default=16384)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--cpu-iters", type=int, default=60)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (--gpus must be 1).", file=sys.stderr)
        sys.exit(2)
    from wrapd_amd import meshgen
    prob = meshgen.config_problem(args.config)
    ne = prob.n_elems
    log(f"workload: {prob.name}: {ne} tets, {prob.n_verts} verts, {len(prob.pins)} pins")

    if args.warmup > 0:
        w = run_path(prob, args.warmup)
        w["solver"].close()
        log(f"warmup: {args.warmup} iterations")

    mlp, learned_info = train_mlp(prob, args.data_iters, args.train_steps, args.batch, args.lr)
    log(f"MLP: {learned_info['tuples']} tuples, {args.train_steps} Adam steps in {learned_info['train_s']:.2f} s, "
        f"loss {learned_info['loss_first']:.3e} -> {learned_info['loss_last']:.3e}")

    variants = (("plain_fp64", {}, None),
                ("fp32_newton", dict(precision=1, prox_mode=1, fp64_switch=1e-8), None),
                ("learned_fp32", dict(precision=1, prox_mode=2, fp64_switch=1e-8), mlp))
    # per-kernel breakdown: every scope timed (a run of its own: the events slow the loop)
    prof_runs = {}
    for name, over, net in variants:
        r = run_path(prob, args.steps, mlp=net, **over)
        r["solver"].close()
        prof_runs[name] = r["profile"]
    stats = r["chol"]
    abytes = algorithmic_bytes(prob, stats)
    # the roofline kernel: the phase with the most device time in the learned path
    pl = prof_runs["learned_fp32"]
    dom = max((k2 for k2 in abytes if k2 in pl), key=lambda k2: pl[k2]["ms"])
    # the measured runs: wall clock, with events around the roofline kernel only
    paths = {}
    for name, over, net in variants:
        r = run_path(prob, args.steps, mlp=net, only=dom, **over)
        r["solver"].close()
        paths[name] = r
        log(f"{name}: {args.steps} iterations in {r['total_s']:.3f} s, E {r['E'][0]:.6e} -> {r['E'][-1]:.9e}")
    plain = paths["plain_fp64"]
    e_star = float(np.min(plain["E"]))

    out_paths = {}
    for name, r in paths.items():
        t, k, g = time_to_gap(r, e_star, args.gap)
        p = prof_runs[name]
        lp = p.get("local_step")
        d = dict(time_to_gap_s=t, iters_to_gap=k, gap_at_end=float((r["E"][-1] - e_star) / abs(e_star)),
                 ms_per_iter=1e3 * r["total_s"] / args.steps, setup_s=r["setup_s"],
                 reweights=int(np.sum(r["reweighted"])), inner_iters=int(np.sum(r["inner"])),
                 local_step_elems_per_s=ne / (lp["ms"] / lp["launches"] * 1e-3) if lp else None,
                 prox_iters_per_elem_step=r["runtime"]["prox_iters"] / (ne * args.steps),
                 fallback_elems=r["runtime"]["fallback_elems"],
                 kernel_ms_per_iter={k2: v["ms"] / args.steps for k2, v in p.items()},
                 kernel_ms_per_iter_source="a second run of the same solve with every scope timed")
        if name != "plain_fp64":
            d["final_x_rel_err_vs_fp64"] = float(np.linalg.norm(r["x"] - plain["x"]) / np.linalg.norm(plain["x"]))
            d["final_energy_rel_err_vs_fp64"] = float(abs(r["E"][-1] - plain["E"][-1]) / abs(plain["E"][-1]))
            d["fp64_switch_iter"] = r["runtime"]["fp64_switch_iter"]
        out_paths[name] = d
    lr_ = paths["learned_fp32"]["runtime"]
    learned_info.update(
        used_per_elem_step=lr_["learned_used"] / (ne * args.steps),
        accepted_without_newton_per_elem_step=lr_["learned_accepted"] / (ne * args.steps),
        prox_iters_per_elem_step=out_paths["learned_fp32"]["prox_iters_per_elem_step"],
        prox_iters_per_elem_step_without=out_paths["fp32_newton"]["prox_iters_per_elem_step"])
    pi_l, pi_n = learned_info["prox_iters_per_elem_step"], learned_info["prox_iters_per_elem_step_without"]
    learned_info["newton_iters_saved_frac"] = (pi_n - pi_l) / pi_n if pi_n else None

    # every phase from the fully timed run; the roofline kernel from the measured run
    rows = {}
    for name, b in abytes.items():
        if name in pl:
            avg_s = pl[name]["ms"] / pl[name]["launches"] * 1e-3
            rows[name] = dict(achieved=b / avg_s / 1e9, bytes_per_launch=b, avg_launch_ms=avg_s * 1e3,
                              ms_per_iter=pl[name]["ms"] / args.steps)
    pd = paths["learned_fp32"]["profile"][dom]
    avg_dom = pd["ms"] / pd["launches"] * 1e-3
    if "refactor" in pl:
        rf = pl["refactor"]
        rows["refactor"] = dict(achieved_tflops=stats["factor_tile_flops"] / (rf["ms"] / rf["launches"] * 1e-3) / 1e12,
                                flops_per_launch=stats["factor_tile_flops"], avg_launch_ms=rf["ms"] / rf["launches"],
                                ms_per_iter=rf["ms"] / args.steps, peak_tflops=FP64_PEAK_TFLOPS)
    traffic = None
    tr_path = os.path.join(ROOT, "profiles", f"traffic_{dom}_c{args.config}.json")
    if os.path.exists(tr_path):
        with open(tr_path) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")
    ach = abytes[dom] / avg_dom / 1e9
    roofline = {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": ach / HBM_PEAK_GBS, "traffic": traffic, "kernel": dom,
                "bytes_per_launch": abytes[dom], "avg_launch_ms": avg_dom * 1e3, "launches": int(pd["launches"]),
                "timed": "HIP events on the solver's stream around every launch of this kernel in the measured run"}

    cpu = None
    if not args.no_cpu:
        try:
            cpu = cpu_baseline(prob, args.cpu_iters, out_paths["plain_fp64"]["iters_to_gap"])
        except Exception as ex:  # the baseline must never sink the line
            cpu = dict(error=str(ex)[:200])

    lp = out_paths["learned_fp32"]
    line = {
        "metric": "time-to-1e-6 rel. energy gap (s) + local-step elements/s, learned vs plain",
        "value": lp["time_to_gap_s"],
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": lp["ms_per_iter"],
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 (global step, Cholesky, solves, energies, FP64 tail); local step f32 with f64 fallback; "
                 "MLP bf16 x bf16 -> f32",
        "data": "synthetic twisted beam generated in-process (seeded SplitMix64 jitter); MLP trained on tuples "
                "logged from the plain run",
        "config": {"workload": prob.name, "tets": ne, "verts": prob.n_verts, "youngs": prob.youngs,
                   "poisson": prob.poisson, "lbfgs_m": prob.settings.get("lbfgs_m"), "reweighting": 1,
                   "gap_definition": "(E_k - E*)/|E*|, E* = lowest energy of the plain FP64 run",
                   "e_star": e_star},
        "speedup_vs_plain": (out_paths["plain_fp64"]["time_to_gap_s"] / lp["time_to_gap_s"]
                             if lp["time_to_gap_s"] and out_paths["plain_fp64"]["time_to_gap_s"] else None),
        "local_step_elems_per_s": {k: v["local_step_elems_per_s"] for k, v in out_paths.items()},
        "paths": out_paths,
        "learned": learned_info,
        "roofline": roofline,
        "roofline_all": rows,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
