"""Benchmark of the Barnes-Hut leapfrog workload on one MI355X.

Workload (the headline metric's configuration): N = 1,048,576 Plummer sphere
(a = 1, M = 1, G = 1, seed 2, untruncated radii), theta = 0.5, eps = 0.005,
quadrupole moments, leaf capacity 8, leapfrog dt = 1e-4, tree rebuilt every
step; FP32 force arithmetic, FP64 positions/velocities.

One "step" = one leapfrog kick-drift-kick step = one full Barnes-Hut force
evaluation (bbox, keys, radix sort, build, up-sweep, walk) + FP64 kick/drift.
W warm-up steps, then exactly K timed steps between device synchronisations;
the state is resident in HBM throughout.

Prints ONE JSON line.  Besides the contract fields it carries:
  roofline      the walk kernel (dominant): algorithmic FP32 FLOPs per launch
                (20 per particle-particle, 60 per particle-cell quadrupole
                interaction, counted by the kernel) / its mean duration from HIP
                events recorded on the calculator's stream during the timed steps
  stages        per-stage ms and achieved HBM GB/s on algorithmic bytes
  accuracy      median / p99 relative force error vs the FP64 direct sum over
                all N targets, and relative energy drift over the timed steps
  cpu_baseline  the compiled reference CPU Barnes-Hut (recursive tree, the only
                reference BH that runs this input) on a bounded sample
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "n-body-gravity-simulation_amd"))

N_DEFAULT = 1 << 20
THETA, EPS, DT, CAP, SEED = 0.5, 0.005, 1e-4, 8, 2
PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0
FLOP_PP, FLOP_PC_QUAD = 20, 60
# Published reference number for this metric: 373.34 ms per BH force evaluation
# at N = 1,000,000 (report/Report.pdf Table 4) -> 2.68e6 particle-steps/s.
REF_PUBLISHED = 1_000_000 / 0.37334


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def stage_bytes(n, nodes):
    """Algorithmic HBM bytes per force evaluation for each tree stage."""
    return {
        "bbox": 24 * n,                     # FP64 x, y, z read
        "keys": (24 + 8 + 4) * n,           # positions read, key + index written
        "sort": 8 * 32 * n,                 # 8 passes x (8 hist read + 12 read + 12 write)
        "gather": (4 + 32 + 16) * n,        # perm + FP64 x,y,z,m read, float4 written
        "build": 8 * n + 20 * nodes,        # sorted keys read, node topology written
        "upsweep": 16 * n + 2 * 40 * nodes,  # float4 particles + moment records in/out
    }


def cpu_baseline(pl, n_sample):
    """Compiled reference (oracle/_ref) recursive Barnes-Hut on a bounded
    sample: full tree build + walk of n_sample evenly spaced targets +
    destruction, extrapolated to one force evaluation over all N."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    n = len(pl["m"])
    if O.have_ref():
        _, ms = O.ref_bhrec_sample(pl["m"], pl["x"], pl["y"], pl["z"], n_sample, THETA, EPS)
        kind = "reference"
    else:
        return None
    t_eval = (ms["build_ms"] + ms["walk_ms"] * n / n_sample + ms["destroy_ms"]) * 1e-3
    return {"value": n / t_eval, "unit": "particle-steps/s", "cores": threads, "kind": kind,
            "sample": (f"reference recursive Barnes-Hut (CPU/src/force_calculator_barnes_hut.cpp, "
                       f"theta={THETA}, eps={EPS}) on the same N={n} Plummer input: tree build "
                       f"{ms['build_ms']:.0f} ms + walk of {n_sample} targets {ms['walk_ms']:.0f} ms "
                       f"(scaled x{n / n_sample:.0f}) + destroy {ms['destroy_ms']:.0f} ms; one force "
                       f"evaluation per leapfrog step"),
            "seconds": (ms["build_ms"] + ms["walk_ms"] + ms["destroy_ms"]) * 1e-3}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=N_DEFAULT)
    ap.add_argument("--leaf-cap", type=int, default=CAP)
    ap.add_argument("--cpu-sample", type=int, default=65536)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU: every box has one MI355X and the "
                                   "engine has no multi-GPU path; run with --gpus 1"}), flush=True)
        sys.exit(2)

    import nbody_amd as nb

    n, K, W = args.n, args.steps, args.warmup
    log(f"[bench] generating Plummer N={n}")
    pl = nb.ic_plummer(n, SEED)
    f = nb.BarnesHutForceCalculator(EPS, THETA, args.leaf_cap, "quadrupole", G=1.0)
    p = nb.Particles.from_arrays(pl["m"], pl["x"], pl["y"], pl["z"], pl["vx"], pl["vy"], pl["vz"])
    it = nb.LeapfrogIntegrator()
    if not args.no_accuracy:
        d0 = nb.diagnostics(p, 1.0, EPS)

    log(f"[bench] warmup {W} steps")
    it.run(f, p, DT, W)
    nb.synchronize()
    st_warm = f.stats()

    f.set_timing(True)
    f.reset_timing()
    log(f"[bench] timing {K} steps")
    nb.synchronize()
    t0 = time.perf_counter()
    it.run(f, p, DT, K)
    nb.synchronize()
    elapsed = time.perf_counter() - t0
    stages = f.stage_timing()
    f.set_timing(False)
    st = f.stats()
    calls = stages["calls"]
    value = n * K / elapsed

    # Walk roofline (FP32 VALU): interactions of the last evaluation are
    # counted by the kernel; the tree barely changes over 10 steps.
    walk_ms = stages["walk"] / calls
    flops = FLOP_PP * st["pp_interactions"] + FLOP_PC_QUAD * st["pc_interactions"]
    achieved = flops / (walk_ms * 1e-3) / 1e12
    roofline = {"kernel": "k_walk", "bound": "valu", "achieved": round(achieved, 2),
                "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP32_TFLOPS, 4),
                "traffic": None, "launch_ms": round(walk_ms, 4),
                "flops_per_launch": int(flops),
                "interactions_per_particle": {"pp": round(st["pp_interactions"] / n, 1),
                                              "pc": round(st["pc_interactions"] / n, 1)}}
    sb = stage_bytes(n, st["n_nodes"])
    stage_out = {}
    for k in ("bbox", "keys", "sort", "gather", "build", "upsweep", "walk"):
        ms = stages[k] / calls
        e = {"ms": round(ms, 4)}
        if k in sb:
            gbs = sb[k] / (ms * 1e-3) / 1e9
            e.update({"GB/s": round(gbs, 1), "hbm_frac": round(gbs / PEAK_HBM_GBS, 4),
                      "bytes": int(sb[k])})
        stage_out[k] = e

    out = {
        "metric": "interactions/s (direct) & particle-steps/s (BH, N=1M); serve p50/p99 latency",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": elapsed / K * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": value / REF_PUBLISHED,
        "dtype": "f32",
        "data": "synthetic (seeded Plummer sphere, Philox4x32-10 streams)",
        "config": {"workload": (f"Barnes-Hut leapfrog, N={n} Plummer (a=1, M=1, G=1, seed {SEED}, "
                                f"untruncated), theta={THETA}, eps={EPS}, quadrupole, "
                                f"leaf_capacity={args.leaf_cap}, dt={DT}, tree rebuilt every step, "
                                f"FP32 forces / FP64 state"),
                   "n": n, "theta": THETA, "eps": EPS, "dt": DT, "leaf_capacity": args.leaf_cap,
                   "baseline_published": "2.68e6 particle-steps/s: reference GPU BH, N=1,000,000, "
                                         "373.34 ms per force evaluation (report Table 4)"},
        "roofline": roofline,
        "stages": stage_out,
        "tree": {"nodes": st["n_nodes"], "leaves": st["n_leaves"], "groups": st["n_groups"],
                 "max_level": st["max_level"], "max_frontier": st["max_frontier"],
                 "frontier_overflow": st["frontier_overflow"], "oversize_cells": st["oversize_cells"],
                 "nodes_warmup": st_warm["n_nodes"]},
    }

    if not args.no_accuracy:
        log("[bench] accuracy vs FP64 direct sum (all targets)")
        d1 = nb.diagnostics(p, 1.0, EPS)
        E0 = d0["ke"] + d0["pe_soft"]
        E1 = d1["ke"] + d1["pe_soft"]
        f.computeAccelerations(p)
        a_bh = np.stack(p.acc(), 1)
        fd = nb.ForceCalculatorNaive(EPS, "f64", G=1.0)
        fd.set_timing(True)
        fd.computeAccelerations(p)
        a_ref = np.stack(p.acc(), 1)
        t_direct = fd.timing()["total_ms"]
        e = np.linalg.norm(a_bh - a_ref, axis=1) / np.linalg.norm(a_ref, axis=1)
        out["accuracy"] = {"median_rel": float(np.median(e)), "p99_rel": float(np.percentile(e, 99)),
                           "max_rel": float(e.max()),
                           "energy_drift_rel": float(abs(E1 - E0) / abs(E0)),
                           "steps_for_drift": W + K}
        out["speedup_vs_f64_direct"] = {"f64_direct_ms": t_direct,
                                        "bh_step_ms": elapsed / K * 1e3,
                                        "ratio": t_direct / (elapsed / K * 1e3)}

    if not args.no_cpu:
        log(f"[bench] CPU baseline (reference BH, sample {args.cpu_sample} targets)")
        cb = cpu_baseline(pl, args.cpu_sample)
        if cb is not None:
            out["cpu_baseline"] = cb
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
