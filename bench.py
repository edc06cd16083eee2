"""Benchmark of the N-body engine on one MI355X: one JSON line.

Headline (the metric's configuration): Barnes-Hut leapfrog on N = 1,048,576
Plummer sphere (a = 1, M = 1, G = 1, seed 2, untruncated radii), theta = 0.5,
eps = 0.005, quadrupole moments, leaf capacity 8, dt = 1e-4, tree rebuilt
every step; FP32 force arithmetic, FP64 positions/velocities.  One "step" =
one leapfrog kick-drift-kick step = one full Barnes-Hut force evaluation
(bbox, keys, radix sort, build, up-sweep, walk) + FP64 kick/drift.  W warm-up
steps, then exactly K timed steps between device synchronisations; the state
is resident in HBM throughout.

The default run (`python bench.py`) nests the other workloads in the same
line, each with its own rate, roofline, accuracy and CPU baseline:
  config1   direct FP64 (bit-exact arithmetic), N = 1000, leapfrog x100
  direct    direct tiled FP32, N = 262,144 Plummer + disk, RK4 x20
  ensemble  4096 systems in one persistent launch, 200 steps in a HIP graph
  serve     Poisson-arrival serving sweep, p50/p99 latency and throughput
  host      the CPU the baselines ran on
`--workload X` runs one of them alone and prints it as the line.

Fields beside the contract ones:
  roofline      the dominant kernel: algorithmic FLOPs per launch / its mean
                duration from HIP events on the stream it runs on
  stages        per-stage ms and achieved HBM GB/s on algorithmic bytes
  accuracy      relative force error against an FP64 direct sum on the GPU,
                and relative energy drift
  cpu_baseline  the compiled reference CPU path (oracle/_ref, built from the
                reference's own sources) timed on this host on a bounded
                sample; "port" = the CPU restatement, where the reference
                cannot run the input
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

METRIC = "interactions/s (direct) & particle-steps/s (BH, N=1M); serve p50/p99 latency"
N_DEFAULT = 1 << 20
THETA, EPS, DT, CAP, SEED = 0.5, 0.005, 1e-4, 8, 2
PEAK_FP32_TFLOPS = 157.3  # MI355X FP32 vector / dense FP32 MFMA
PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0
FLOP_PP, FLOP_PC_QUAD = 20, 60
STAGE_EVALS = 3  # untimed evaluations for the per-stage breakdown
# Published reference numbers (BASELINE): 373.34 ms per BH force evaluation at
# N = 1,000,000 (report/Report.pdf Table 4) -> 2.68e6 particle-steps/s; FP64
# GPU direct 4.12e9 interactions/s (CUDA/outputs/output1_1.txt:130).
REF_PUBLISHED_BH = 1_000_000 / 0.37334
REF_PUBLISHED_DIRECT_F64 = 4.12e9


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_threads():
    return int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)


def host_info():
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"cpu_model": model, "logical_cpus": os.cpu_count(), "omp_threads": cpu_threads()}


def checker():
    """The CPU checker package (oracle/): the compiled reference and its
    restatement.  Used only by the cpu_baseline legs below."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    return O


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC
    summary (profiles/traffic.json: FETCH_SIZE and WRITE_SIZE in KiB per
    dispatch, each scaled by the factor a 256 MiB device copy calibrates), or
    None when no measurement for this kernel is committed."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
        k = t["kernels"][kernel]
        c = t["calibration"]
        return {"bytes": int((k["fetch_kib"] * c["fetch_factor"] + k["write_kib"] * c["write_factor"])
                             * 1024),
                "source": "profiles/traffic.json (" + k["workload"] + ")"}
    except (OSError, KeyError, ValueError):
        return None


def traffic_fields(kernel):
    t = measured_traffic(kernel)
    return {"traffic": None if t is None else t["bytes"],
            "traffic_source": None if t is None else t["source"]}


def rel_err(a, r):
    return np.linalg.norm(a - r, axis=1) / np.maximum(np.linalg.norm(r, axis=1), 1e-300)


# --------------------------------------------------------------------- FP32 ceiling
# Instructions of the compiled terms besides their v_rsq_f32 (gfx950 ISA):
# the direct FP32 term with softening and the walk's particle term are 3 sub +
# 3 mul + 6 FMA (20 FLOP); the walk's quadrupole cell term is 3 add/sub + 9 mul
# + 18 FMA (60 FLOP).  A far cell (one lane builds its 24 local-expansion
# coefficients: 117 VALU + 2 rsqrt, 164 FLOP; the wave's reduce-scatter of a
# batch of 64 adds 82 VALU per lane, 30 of them adds) costs 199 other
# instructions and 194 FLOP; evaluating the expansion at a target, 105 VALU
# and 160 FLOP.
TERM_OTHER_INSTR = {"pp": 12, "pc": 30, "far": 199, "l2p": 105}
TERM_RSQ = {"pp": 1, "pc": 1, "far": 2, "l2p": 0}
TERM_FLOP = {"pp": 20, "pc": 60, "far": 194, "l2p": 160}


def fma_probe(nb, achieved, terms=None):
    """The FP32 ceiling without packed FP32 (the library is built without it):
    v_fma_f32 issue rate with VGPR operands, measured in this run outside the
    timed region, and the kernel's FLOP rate as a fraction of it.  With
    `terms` ({"pp": count, "pc": count, ...}), also the ceiling of that term
    mix: the FMA slot rate (probe / 2 FLOP) times FLOP per slot, a term taking
    its other instructions plus its rsqrts at the measured cost in slots."""
    r = nb.probe_fp32_fma()
    out = {"tflops": round(r["tflops"], 2), "frac": round(achieved / r["tflops"], 4),
           "what": "16 independent v_fma_f32 chains per lane, VGPR operands, 8 x 256-lane "
                   "blocks per CU (nb_probe_fp32_fma)"}
    if terms:
        c = nb.probe_fp32_rsq_cost()
        flop = sum(TERM_FLOP[k] * v for k, v in terms.items())
        slots = sum((TERM_OTHER_INSTR[k] + TERM_RSQ[k] * c) * v for k, v in terms.items())
        ceiling = r["tflops"] / 2.0 * flop / slots
        out["rsq_fma_slots"] = round(c, 2)
        out["mix_ceiling_tflops"] = round(ceiling, 2)
        out["mix_frac"] = round(achieved / ceiling, 4)
    return out


def walk_terms(st, n):
    """Term counts of one walk launch: particle and cell terms per target, far
    cells (one local-expansion build each) and, when there are far cells, one
    expansion evaluation per target."""
    t = {"pp": st["pp_interactions"], "pc": st["pc_interactions"]}
    if st.get("far_cells", 0) > 0:
        t["far"] = st["far_cells"]
        t["l2p"] = n
    return t


# --------------------------------------------------------------------- Barnes-Hut (headline)
def stage_bytes(n, nodes):
    """Algorithmic HBM bytes per force evaluation for each tree stage."""
    return {
        "bbox": 24 * n,                     # FP64 x, y, z read
        "keys": (32 + 8 + 4 + 16) * n,      # x, y, z, m read; key, index, float4 written (digit counts in registers)
        # 8 passes x (12 read + 12 write); the last pass also gathers the
        # float4 particles through the permutation (16 read + 16 written)
        "sort": 8 * 24 * n + 32 * n,
        "build": 8 * n + 20 * nodes,        # sorted keys read, node topology written
        "upsweep": 16 * n + 2 * 40 * nodes,  # float4 particles + moment records in/out
    }


def stage_report(stages, n, nodes):
    """Per-stage ms per evaluation, with achieved GB/s and HBM fraction on the
    algorithmic bytes of stage_bytes()."""
    sb = stage_bytes(n, nodes)
    out = {}
    for k in ("bbox", "keys", "sort", "gather", "build", "upsweep", "walk"):
        ms = stages[k] / stages["calls"]
        e = {"ms": round(ms, 4)}
        if k == "gather":
            e["note"] = "fused into the last radix pass (its bytes and time are in sort)"
        if k in sb:
            gbs = sb[k] / (ms * 1e-3) / 1e9
            e.update({"GB/s": round(gbs, 1), "hbm_frac": round(gbs / PEAK_HBM_GBS, 4),
                      "bytes": int(sb[k])})
        out[k] = e
    tree = sum(stages[k] for k in ("bbox", "keys", "sort", "gather", "build", "upsweep"))
    out["tree_total_ms"] = round(tree / stages["calls"], 4)
    return out


def stages_at_scale(nb, n, evals=3):
    """Tree stages of one force evaluation at a larger N (same Plummer model,
    theta, eps, leaf capacity): at N = 2^20 the tree stages are bound by
    kernel-launch latency, at 2^24 by memory traffic."""
    pl = nb.ic_plummer(n, SEED)
    f = nb.BarnesHutForceCalculator(EPS, THETA, CAP, "quadrupole", G=1.0)
    p = nb.Particles.from_arrays(pl["m"], pl["x"], pl["y"], pl["z"])
    del pl
    f.computeAccelerations(p)
    f.set_timing(True, stages=True)
    f.reset_timing()
    for _ in range(evals):
        f.computeAccelerations(p)
    stages = f.stage_timing()
    f.set_timing(False)
    nodes = f.stats()["n_nodes"]
    out = {"n": n, "nodes": nodes, "evals": evals, "stages": stage_report(stages, n, nodes)}
    p.free()
    f.close()
    return out


def cpu_baseline_bh(pl, n_sample):
    """Reference CPU Barnes-Hut on the same N = 2^20 input, on a bounded
    sample.  (i) the compiled reference recursive tree
    (CPU/src/force_calculator_barnes_hut.cpp): full build, walk of n_sample
    evenly spaced targets, destruction; (ii) the optimized pool tree
    (CPU/src/force_calculator_barnes_hut_optimized.cpp, leaf capacity 8),
    whose fixed 64-entry walk stack overflows on this input in the reference
    build, in the checker's restatement with a growable stack (a labelled
    deviation).  Both are extrapolated to one force evaluation over all N."""
    O = checker()
    n = len(pl["m"])
    out = {}
    if O.have_ref():
        _, ms = O.ref_bhrec_sample(pl["m"], pl["x"], pl["y"], pl["z"], n_sample, THETA, EPS)
        t_eval = (ms["build_ms"] + ms["walk_ms"] * n / n_sample + ms["destroy_ms"]) * 1e-3
        out["reference"] = {
            "value": n / t_eval, "unit": "particle-steps/s", "cores": cpu_threads(),
            "kind": "reference",
            "sample": (f"compiled reference recursive Barnes-Hut (CPU/src/force_calculator_barnes_hut"
                       f".cpp, theta={THETA}, eps={EPS}) on the same N={n} Plummer input: tree "
                       f"build {ms['build_ms']:.0f} ms + walk of {n_sample} targets "
                       f"{ms['walk_ms']:.0f} ms (scaled x{n / n_sample:.0f}) + destroy "
                       f"{ms['destroy_ms']:.0f} ms; one force evaluation per leapfrog step"),
            "seconds": (ms["build_ms"] + ms["walk_ms"] + ms["destroy_ms"]) * 1e-3}
    _, ms = O.bhopt_sample(pl["m"], pl["x"], pl["y"], pl["z"], n_sample, THETA, CAP, 1.0, EPS * EPS)
    t_eval = (ms["build_ms"] + ms["walk_ms"] * n / n_sample) * 1e-3
    out["optimized"] = {
        "value": n / t_eval, "unit": "particle-steps/s", "cores": cpu_threads(), "kind": "port",
        "sample": (f"optimized Barnes-Hut (leaf capacity {CAP}, monopole, "
                   f"CPU/src/force_calculator_barnes_hut_optimized.cpp) restated with a growable "
                   f"walk stack -- a deviation: the reference's int stack[64] (line 52) overflows "
                   f"on this input; serial build {ms['build_ms']:.0f} ms + walk of {n_sample} "
                   f"targets {ms['walk_ms']:.0f} ms (scaled x{n / n_sample:.0f})"),
        "seconds": (ms["build_ms"] + ms["walk_ms"]) * 1e-3}
    return out


def bench_bh(args):
    import nbody_amd as nb

    n, K, W = args.n, (args.steps if args.steps is not None else 10), args.warmup
    log(f"[bench] Barnes-Hut: generating Plummer N={n}")
    pl = nb.ic_plummer(n, SEED)
    f = nb.BarnesHutForceCalculator(EPS, THETA, args.leaf_cap, "quadrupole", G=1.0)
    p = nb.Particles.from_arrays(pl["m"], pl["x"], pl["y"], pl["z"], pl["vx"], pl["vy"], pl["vz"])
    it = nb.LeapfrogIntegrator()
    if not args.no_accuracy:
        d0 = nb.diagnostics(p, 1.0, EPS)

    log(f"[bench] warmup {W} steps")
    it.run(f, p, DT, W)
    nb.synchronize()
    st_warm = f.stats() if W > 0 else None

    # Timed steps record only three events per evaluation (start, walk start,
    # end): every event leaves the GPU idle for a few microseconds.
    f.set_timing(True, stages=False)
    f.reset_timing()
    log(f"[bench] timing {K} steps")
    nb.synchronize()
    t0 = time.perf_counter()
    it.run(f, p, DT, K)
    nb.synchronize()
    elapsed = time.perf_counter() - t0
    tm = f.stage_timing()
    tree_ms_timed = f.timing()["build_ms"] / tm["calls"]
    st = f.stats()
    value = n * K / elapsed
    walk_ms = tm["walk"] / tm["calls"]

    # Per-stage breakdown from separate, untimed evaluations of the final
    # state (the force is deterministic, so the state does not change).
    f.set_timing(True, stages=True)
    f.reset_timing()
    for _ in range(STAGE_EVALS):
        f.computeAccelerations(p)
    stages = f.stage_timing()
    f.set_timing(False)

    # Walk roofline (FP32 VALU): interactions of the last evaluation are
    # counted by the kernel; the tree barely changes over the timed steps.
    terms = walk_terms(st, n)
    flops = sum(TERM_FLOP[k] * v for k, v in terms.items())
    achieved = flops / (walk_ms * 1e-3) / 1e12
    # The same accepted cells evaluated at every target (NB_WALK_FAR=0) would
    # cost 60 FLOP per (target, cell) pair: the rate that work is done at.
    equiv = (FLOP_PP * st["pp_interactions"]
             + FLOP_PC_QUAD * (st["pc_interactions"] + st["far_target_terms"]))
    roofline = {"kernel": "k_walk", "bound": "valu", "achieved": round(achieved, 2),
                "peak_kind": "FP32 vector (157.3 TFLOP/s is also the dense FP32 MFMA figure)",
                "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP32_TFLOPS, 4),
                **traffic_fields("k_walk"),
                "launch_ms": round(walk_ms, 4),
                "flops_per_launch": int(flops),
                "interactions_per_particle": {"pp": round(st["pp_interactions"] / n, 1),
                                              "pc": round(st["pc_interactions"] / n, 1),
                                              "pc_in_far_expansions": round(st["far_target_terms"] / n, 1)},
                "far_cells_per_launch": st["far_cells"],
                "per_target_equivalent": {
                    "flops_per_launch": int(equiv),
                    "tflops": round(equiv / (walk_ms * 1e-3) / 1e12, 2),
                    "what": "FLOP of evaluating every accepted cell at every target (60 per pair) "
                            "and the particle terms (20), over the walk time"},
                "fma_probe": fma_probe(nb, achieved, terms)}
    out = {
        "metric": METRIC, "value": value, "unit": "particle-steps/s", "n_gpus": 1, "steps": K,
        "warmup": W, "ms_per_step": elapsed / K * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": value / REF_PUBLISHED_BH, "dtype": "f32",
        "data": "synthetic (seeded Plummer sphere, Philox4x32-10 streams)",
        "config": {"workload": (f"Barnes-Hut leapfrog, N={n} Plummer (a=1, M=1, G=1, seed {SEED}, "
                                f"untruncated), theta={THETA}, eps={EPS}, quadrupole, "
                                f"leaf_capacity={args.leaf_cap}, dt={DT}, tree rebuilt every step, "
                                f"FP32 forces / FP64 state"),
                   "n": n, "theta": THETA, "eps": EPS, "dt": DT, "leaf_capacity": args.leaf_cap,
                   "baseline_published": "2.68e6 particle-steps/s: reference GPU BH, N=1,000,000, "
                                         "373.34 ms per force evaluation (report Table 4)"},
        "roofline": roofline,
        "stages": stage_report(stages, n, st["n_nodes"]),
        # tree stages (bbox .. up-sweep) per step in the timed run, which
        # records no events between them
        "tree_ms_timed_steps": round(tree_ms_timed, 4),
        "tree": {"nodes": st["n_nodes"], "leaves": st["n_leaves"], "groups": st["n_groups"],
                 "max_level": st["max_level"], "max_frontier": st["max_frontier"],
                 "frontier_overflow": st["frontier_overflow"], "oversize_cells": st["oversize_cells"],
                 "nodes_warmup": st_warm["n_nodes"] if st_warm else None},
    }

    if args.scale_n > 0:
        log(f"[bench] tree stages at N={args.scale_n}")
        out["stages_at_scale"] = stages_at_scale(nb, args.scale_n)

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
        e = rel_err(a_bh, a_ref)
        out["accuracy"] = {"median_rel": float(np.median(e)), "p99_rel": float(np.percentile(e, 99)),
                           "max_rel": float(e.max()), "targets": n,
                           "reference": "FP64 direct sum on the GPU (rsqrt + Newton step)",
                           "energy_drift_rel": float(abs(E1 - E0) / abs(E0)),
                           "steps_for_drift": W + K}
        out["speedup_vs_f64_direct"] = {"f64_direct_ms": t_direct,
                                        "bh_step_ms": elapsed / K * 1e3,
                                        "ratio": t_direct / (elapsed / K * 1e3)}

    if not args.no_cpu:
        log(f"[bench] CPU baselines (reference BH, sample {args.cpu_sample} targets)")
        cb = cpu_baseline_bh(pl, args.cpu_sample)
        if "reference" in cb:
            out["cpu_baseline"] = cb["reference"]
        out["cpu_baseline_optimized"] = cb["optimized"]
    return out


# --------------------------------------------------------------------- config 1
def bench_config1(args):
    """Plain direct all-pairs FP64 with the reference CPU's exact arithmetic
    (k_direct_f64_exact), N = 1000 equal masses, positions U[-1,1]^3,
    velocities N(0, 0.1^2) per axis (seed 1), G = 1, eps = 0.01, leapfrog
    dt = 1e-3 for 100 steps (cached acceleration: 101 force evaluations)."""
    import nbody_amd as nb

    n, K, eps, dt = 1000, 100, 0.01, 1e-3
    d = nb.ic_uniform_gaussian(n, 1, 0.1)
    f = nb.ForceCalculatorNaive(eps, "f64_exact", G=1.0)
    it = nb.LeapfrogIntegrator()
    # one untimed run (module load, first launches), then the timed run from the same input
    for rep in range(2):
        p = nb.Particles.from_arrays(d["m"], d["x"], d["y"], d["z"], d["vx"], d["vy"], d["vz"])
        e0 = nb.diagnostics(p, 1.0, eps)
        nb.synchronize()
        t0 = time.perf_counter()
        it.reset()
        it.run(f, p, dt, K)
        nb.synchronize()
        elapsed = time.perf_counter() - t0
    e1 = nb.diagnostics(p, 1.0, eps)
    gpu = p.download()
    inter = (K + 1) * n * (n - 1)
    E0, E1 = e0["ke"] + e0["pe_soft"], e1["ke"] + e1["pe_soft"]
    out = {"value": inter / elapsed, "unit": "interactions/s", "steps": K,
           "ms_per_step": elapsed / K * 1e3, "dtype": "f64",
           "workload": (f"direct all-pairs FP64 (reference CPU arithmetic), N={n}, m=1/N, "
                        f"x~U[-1,1]^3, v~N(0,0.1^2) (seed 1), G=1, eps={eps}, leapfrog dt={dt} "
                        f"x {K} steps (cached acceleration: {K + 1} evaluations)"),
           "roofline": {"kernel": "k_direct_f64_exact_wave", "bound": "latency",
                        "note": "N=1000: one wave per target (1000 waves), the terms in "
                                "parallel and their in-order FP64 sums on three lanes: the "
                                "dependent chain of 999 adds per component and the launch "
                                "latency bound each evaluation"},
           "accuracy": {"energy_drift_rel": float(abs(E1 - E0) / abs(E0))}}
    if not args.no_cpu:
        O = checker()
        if O.have_ref():
            # The reference's own leapfrog (two force evaluations per step,
            # CPU/src/integrator_leapfrog.cpp:5,18) on the same input, timed.
            # The reference fixes G = 4 pi^2; the GPU path is run at that G on
            # the same input to check the bits.
            t0 = time.perf_counter()
            ref = O.ref_integrate_direct("leapfrog", d["m"], d["x"], d["y"], d["z"], d["vx"],
                                         d["vy"], d["vz"], dt, K, softening=eps)
            t_ref = time.perf_counter() - t0
            g = nb.ForceCalculatorNaive(eps, "f64_exact")  # G = 4 pi^2
            q = nb.Particles.from_arrays(d["m"], d["x"], d["y"], d["z"], d["vx"], d["vy"], d["vz"])
            nb.LeapfrogIntegrator().run(g, q, dt, K)
            s = q.download()
            same = all(np.array_equal(s[k], r) for k, r in
                       zip(("x", "y", "z", "vx", "vy", "vz", "ax", "ay", "az"), ref))
            out["accuracy"]["bit_exact_vs_reference"] = bool(same)
            out["accuracy"]["reference_check"] = ("final x, v, a after 100 steps, G=4pi^2, same "
                                                  "input, vs the compiled reference CPU build")
            out["cpu_baseline"] = {
                "value": 2 * K * n * (n - 1) / t_ref, "unit": "interactions/s",
                "cores": cpu_threads(), "kind": "reference",
                "sample": (f"the full run: compiled reference LeapfrogIntegrator + "
                           f"ForceCalculatorNaive, N={n}, {K} steps x 2 evaluations "
                           f"(CPU/src/integrator_leapfrog.cpp), {t_ref:.3f} s"),
                "seconds": t_ref, "steps_per_s": K / t_ref}
    out["gpu_steps_per_s"] = K / elapsed
    del gpu
    return out


# --------------------------------------------------------------------- config 3
def bench_direct(args):
    """Direct tiled FP32 all-pairs with FP64 RK4 state: N = 262,144 Plummer
    (90% mass) + cold exponential disk (10%), eps = 0.01, dt = 5e-4, 20 RK4
    steps (80 force evaluations); error against the FP64 kernel."""
    import nbody_amd as nb

    n = 262144
    K = 20
    log(f"[bench] direct: N={n}, RK4 x{K}")
    d = nb.ic_plummer_disk(n, 5)
    f = nb.ForceCalculatorNaive(0.01, "f32", G=1.0)
    p = nb.Particles.from_arrays(d["m"], d["x"], d["y"], d["z"], d["vx"], d["vy"], d["vz"])
    f.computeAccelerations(p)  # warm-up launch
    e0 = nb.diagnostics(p, 1.0, 0.01)
    it = nb.RK4Integrator()
    nb.synchronize()
    f.set_timing(True)
    f.reset_timing()
    t0 = time.perf_counter()
    it.run(f, p, 5e-4, K)
    nb.synchronize()
    elapsed = time.perf_counter() - t0
    tm = f.timing()
    e1 = nb.diagnostics(p, 1.0, 0.01)
    inter = 4 * K * n * (n - 1)
    kern_ms = tm["walk_ms"] / tm["calls"]
    achieved = FLOP_PP * n * (n - 1) / (kern_ms * 1e-3) / 1e12
    # accuracy of one evaluation against the FP64 kernels on the final state
    f.computeAccelerations(p)
    a32 = np.stack(p.acc(), 1)
    f64 = nb.ForceCalculatorNaive(0.01, "f64", G=1.0)
    f64.set_timing(True)
    f64.computeAccelerations(p)
    f64.computeAccelerations(p)
    a64 = np.stack(p.acc(), 1)
    t64 = f64.timing()["total_ms"] / 2
    fx = nb.ForceCalculatorNaive(0.01, "f64_exact", G=1.0)
    fx.set_timing(True)
    fx.computeAccelerations(p)
    ax = np.stack(p.acc(), 1)
    tx = fx.timing()["total_ms"]
    err = rel_err(a32, a64)
    E0, E1 = e0["ke"] + e0["pe_soft"], e1["ke"] + e1["pe_soft"]
    pairs = n * (n - 1)
    out = {
        "value": inter / elapsed, "unit": "interactions/s", "steps": K,
        "ms_per_step": elapsed / K * 1e3, "dtype": "f32",
        "vs_baseline": None,
        "workload": (f"direct tiled all-pairs FP32 forces, FP64 RK4 state, N={n} Plummer (90%) + "
                     f"cold exponential disk (10%, R_d=1, z_d=0.05), eps=0.01, dt=5e-4, {K} steps "
                     f"(4 evaluations per step)"),
        "roofline": {"kernel": "k_direct_f32", "bound": "valu", "achieved": round(achieved, 2),
                     "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / PEAK_FP32_TFLOPS, 4), **traffic_fields("k_direct_f32"),
                     "launch_ms": round(kern_ms, 4), "flops_per_launch": FLOP_PP * pairs,
                     "fma_probe": fma_probe(nb, achieved, {"pp": 1})},
        "accuracy": {"median_rel_vs_f64": float(np.median(err)),
                     "p99_rel_vs_f64": float(np.percentile(err, 99)),
                     "max_rel_vs_f64": float(err.max()),
                     "f64_fast_vs_f64_exact_max_rel": float(rel_err(a64, ax).max()),
                     "energy_drift_rel": float(abs(E1 - E0) / abs(E0)), "steps_for_drift": K},
        # Like-for-like with the reference's published FP64 GPU direct sum:
        "f64_direct": {
            "exact": {"kernel": "k_direct_f64_exact", "ms_per_eval": tx,
                      "interactions_per_s": pairs / (tx * 1e-3),
                      "vs_published_f64": pairs / (tx * 1e-3) / REF_PUBLISHED_DIRECT_F64,
                      "fp64_frac": round(FLOP_PP * pairs / (tx * 1e-3) / 1e12 / PEAK_FP64_TFLOPS, 4)},
            "fast": {"kernel": "k_direct_f64_fast", "ms_per_eval": t64,
                     "interactions_per_s": pairs / (t64 * 1e-3),
                     "vs_published_f64": pairs / (t64 * 1e-3) / REF_PUBLISHED_DIRECT_F64,
                     "fp64_frac": round(FLOP_PP * pairs / (t64 * 1e-3) / 1e12 / PEAK_FP64_TFLOPS, 4)},
            "baseline_published": "4.12e9 interactions/s: reference FP64 GPU direct, N=100,000 "
                                  "(CUDA/outputs/output1_1.txt:130)"},
    }
    if not args.no_cpu:
        O = checker()
        if O.have_ref():
            # The reference's ForceCalculatorNaive (FP64, OpenMP over targets)
            # on every 4th particle of the same input: one evaluation of a
            # 65,536-particle system.
            sub = {k: np.ascontiguousarray(d[k][::4]) for k in ("m", "x", "y", "z")}
            ns = len(sub["m"])
            t0 = time.perf_counter()
            O.ref_direct(sub["m"], sub["x"], sub["y"], sub["z"], 0.01)
            t_ref = time.perf_counter() - t0
            rate = ns * (ns - 1) / t_ref
            out["cpu_baseline"] = {
                "value": rate, "unit": "interactions/s", "cores": cpu_threads(),
                "kind": "reference",
                "sample": (f"compiled reference ForceCalculatorNaive (FP64) on every 4th particle "
                           f"of the same input (N={ns}, one evaluation, {t_ref:.2f} s); the "
                           f"{K}-step RK4 run ({4 * K} evaluations at N={n}) would take "
                           f"{inter / rate:.0f} s at this rate"),
                "seconds": t_ref}
    return out


# --------------------------------------------------------------------- config 4
def bench_ensemble(args):
    """4096 independent Plummer systems, N_k log-uniform in [64, 8192] and scale
    radius uniform in [0.5, 2] (seeded), direct FP32 forces with FP64 leapfrog,
    eps = 0.01, dt = 1e-3, 200 steps captured in one HIP graph; tile shapes per
    size bucket from the autotuner.  Compared with one-system-at-a-time
    launches of the same kernels (timed on a few steps)."""
    import nbody_amd as nb

    S, K = 4096, 200
    log(f"[bench] ensemble: {S} systems, {K} steps")
    rng = np.random.Generator(np.random.Philox(4))
    sizes = np.rint(np.exp(rng.uniform(np.log(64), np.log(8192), S))).astype(np.int64)
    radii = rng.uniform(0.5, 2.0, S)
    parts = [nb.ic_plummer(int(n), 1000 + k, a=float(a)) for k, (n, a) in enumerate(zip(sizes, radii))]
    cat = {k: np.concatenate([p[k] for p in parts]) for k in parts[0]}
    ntot = int(sizes.sum())
    offs = np.concatenate([[0], np.cumsum(sizes)])
    e = nb.Ensemble(sizes, G=1.0, softening=0.01, dt=1e-3)
    e.upload(cat["m"], cat["x"], cat["y"], cat["z"], cat["vx"], cat["vy"], cat["vz"])
    e.autotune()
    e.run(2, use_graph=True)  # warm-up: prime + capture a 2-step graph
    nb.synchronize()
    e.run(K, use_graph=True)  # captures the K-step graph (not timed)
    nb.synchronize()
    t0 = time.perf_counter()
    e.run(K, use_graph=True)
    nb.synchronize()
    t_b = time.perf_counter() - t0
    force_ms = e.time_force(5)
    pairs = float(np.sum(sizes * (sizes - 1)))
    achieved = FLOP_PP * pairs / (force_ms * 1e-3) / 1e12
    ku = 3
    t0 = time.perf_counter()
    e.run_unbatched(ku)
    nb.synchronize()
    t_u = (time.perf_counter() - t0) / ku
    value = ntot * K / t_b
    # Accuracy: the cached FP32 accelerations of the final state of 32 sampled
    # systems (every bucket, the largest included) against the FP64 direct sum.
    st = e.download()
    order = np.argsort(sizes)
    pick = np.unique(np.concatenate([order[np.linspace(0, S - 1, 31).astype(int)], [order[-1]]]))
    errs = []
    f64 = nb.ForceCalculatorNaive(0.01, "f64", G=1.0)
    for s in pick:
        a, b = offs[s], offs[s + 1]
        q = nb.Particles.from_arrays(cat["m"][a:b], st["x"][a:b], st["y"][a:b], st["z"][a:b])
        f64.computeAccelerations(q)
        r = np.stack(q.acc(), 1)
        g = np.stack([st["ax"][a:b], st["ay"][a:b], st["az"][a:b]], 1)
        errs.append(rel_err(g, r))
    err = np.concatenate(errs)
    out = {
        "value": value, "unit": "particle-steps/s", "steps": K, "ms_per_step": t_b / K * 1e3,
        "dtype": "f32",
        "workload": (f"ensemble of {S} Plummer systems, N log-uniform [64, 8192] (total {ntot}), "
                     f"a ~ U[0.5, 2], direct FP32 + FP64 leapfrog, eps=0.01, dt=1e-3, {K} steps "
                     f"in one HIP graph, autotuned shapes"),
        "shapes_per_bucket": e.tuning().tolist(),
        "interactions_per_s": pairs * K / t_b,
        "roofline": {"kernel": "k_ens_force", "bound": "valu", "achieved": round(achieved, 2),
                     "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / PEAK_FP32_TFLOPS, 4),
                     **traffic_fields("k_ens_force"), "launch_ms": round(force_ms, 4),
                     "flops_per_launch": int(FLOP_PP * pairs)},
        "accuracy": {"median_rel_vs_f64": float(np.median(err)),
                     "p99_rel_vs_f64": float(np.percentile(err, 99)),
                     "max_rel_vs_f64": float(err.max()),
                     "sample": f"final-state accelerations of {len(pick)} systems spanning every "
                               f"size bucket (N {int(sizes[pick].min())}..{int(sizes[pick].max())}) "
                               f"vs the FP64 direct sum on the GPU"},
        "unbatched": {"ms_per_step": t_u * 1e3, "particle_steps_per_s": ntot / t_u,
                      "speedup_batched": (ntot * K / t_b) / (ntot / t_u)},
    }
    if not args.no_cpu:
        O = checker()
        if O.have_ref():
            # The reference's leapfrog + naive force, one system at a time, on
            # every 32nd system by size (all buckets), 2 steps each,
            # extrapolated to all systems by sum N(N-1).
            samp = order[::32]
            t0 = time.perf_counter()
            for s in samp:
                a, b = offs[s], offs[s + 1]
                O.ref_integrate_direct("leapfrog", cat["m"][a:b], cat["x"][a:b], cat["y"][a:b],
                                       cat["z"][a:b], cat["vx"][a:b], cat["vy"][a:b],
                                       cat["vz"][a:b], 1e-3, 2, softening=0.01)
            t_s = time.perf_counter() - t0
            sp = float(np.sum(sizes[samp] * (sizes[samp] - 1)))
            t_step = t_s / 2 * pairs / sp
            out["cpu_baseline"] = {
                "value": ntot / t_step, "unit": "particle-steps/s", "cores": cpu_threads(),
                "kind": "reference",
                "sample": (f"compiled reference LeapfrogIntegrator + ForceCalculatorNaive, one "
                           f"system at a time, 2 steps on {len(samp)} of the {S} systems (every "
                           f"32nd by size), {t_s:.2f} s, extrapolated by sum N(N-1)"),
                "seconds": t_s}
    return out


# --------------------------------------------------------------------- config 5
def bench_serve(args):
    """Serving: one seeded trace of requests (Plummer systems with N
    log-uniform in [1e3, 1e6], steps uniform in [10, 1000], eps = 0.005,
    dt = 1e-4; direct method below N = 32,768, Barnes-Hut theta = 0.5 above;
    HBM budget 200 GB), replayed with Poisson arrivals at each rate of the
    sweep.  Requests are submitted at their arrival times while earlier ones
    run.  Per rate: p50/p99 completion latency (arrival -> result on the host),
    split into queueing (arrival -> admission) and service, and aggregate
    particle-steps/s.  The knee is the arrival rate the measured throughput
    sustains for this request mix."""
    import nbody_amd as nb

    rates = [int(r) for r in args.rates.split(",")]
    R = args.requests
    rng = np.random.Generator(np.random.Philox(1000))
    ns = np.rint(np.exp(rng.uniform(np.log(1e3), np.log(1e6), R))).astype(int)
    steps = rng.integers(10, 1001, R)
    log(f"[bench] serve: {R} requests, sum N x steps = {float(np.sum(ns * steps)):.3g}")
    systems = [nb.ic_plummer(int(n), 50000 + k) for k, n in enumerate(ns)]
    work = float(np.sum(ns * steps))
    rows, skipped = [], []
    for ri, rate in enumerate(rates):
        # Keep the whole run inside its time budget: a rate that would not
        # finish in time (judged by the previous rate's wall time) is skipped
        # and listed as such, never cut short.
        est = rows[-1]["wall_s"] if rows else work / 2.5e8
        if args.deadline and time.perf_counter() + 1.1 * est > args.deadline:
            skipped.append(rate)
            log(f"[serve] rate {rate}: skipped (time budget)")
            continue
        gaps = np.random.Generator(np.random.Philox(2000 + ri)).exponential(1.0 / rate, R)
        arrivals = np.cumsum(gaps) - gaps[0]
        srv = nb.Server(hbm_budget_bytes=200e9, direct_threshold=32768, bh_lanes=args.bh_lanes,
                        bh_steps_per_tick=args.bh_steps, tick_budget=args.tick_budget,
                        direct_steps_per_tick=args.direct_steps)
        t0 = time.perf_counter()
        sub = 0
        done, admitted, rid_of, failed = {}, {}, {}, 0
        while len(done) < R:
            now = time.perf_counter() - t0
            while sub < R and arrivals[sub] <= now:
                rid_of[srv.submit(systems[sub], int(steps[sub]), softening=0.005, dt=1e-4)] = sub
                sub += 1
            st = srv.status()
            if st["active"] == 0 and st["pending"] == 0:
                if sub < R:
                    time.sleep(max(0.0, arrivals[sub] - (time.perf_counter() - t0)))
                continue
            srv.tick()
            t = time.perf_counter() - t0
            for rid in srv.poll_admitted():
                admitted[rid_of[rid]] = t
            for rid in srv.poll():
                k = rid_of[rid]
                try:
                    srv.result(rid)
                except nb.NBError:
                    failed += 1
                done[k] = t - arrivals[k]
        wall = time.perf_counter() - t0
        lat = np.array([done[k] for k in range(R)])
        queue = np.array([admitted[k] - arrivals[k] for k in range(R)])
        service = lat - queue
        rows.append({"rate_per_s": rate, "requests": R, "p50_s": float(np.percentile(lat, 50)),
                     "p99_s": float(np.percentile(lat, 99)),
                     "queue_p50_s": float(np.percentile(queue, 50)),
                     "queue_p99_s": float(np.percentile(queue, 99)),
                     "service_p50_s": float(np.percentile(service, 50)),
                     "service_p99_s": float(np.percentile(service, 99)),
                     "particle_steps_per_s": work / wall, "wall_s": wall,
                     "arrival_window_s": float(arrivals[-1]), "failed": failed,
                     "ticks": srv.status()["ticks"]})
        log(f"[serve] rate {rate}: p50 {rows[-1]['p50_s']:.2f} s, p99 {rows[-1]['p99_s']:.2f} s "
            f"(queueing p50 {rows[-1]['queue_p50_s']:.2f} s), "
            f"{rows[-1]['particle_steps_per_s']:.3g} particle-steps/s, wall {wall:.1f} s")
    best = max(r["particle_steps_per_s"] for r in rows)
    mean_work = work / R
    # Same alone and in the mix: the smallest Barnes-Hut and direct requests
    # of the trace, run alone on a fresh server, against the last rate's run.
    check = {}
    for kind, mask in (("direct", ns < 32768), ("bh", ns >= 32768)):
        if not mask.any():
            continue
        k = int(np.argmin(np.where(mask, ns * steps, np.iinfo(np.int64).max)))
        check[kind] = {"n": int(ns[k]), "steps": int(steps[k])}
    for kind, c in check.items():
        k = [q for q in range(R) if ns[q] == c["n"] and steps[q] == c["steps"]][0]
        res = []
        for _ in range(2):  # alone, then with the whole trace submitted at once
            srv = nb.Server(hbm_budget_bytes=200e9, direct_threshold=32768,
                            bh_lanes=args.bh_lanes, bh_steps_per_tick=args.bh_steps,
                            tick_budget=args.tick_budget, direct_steps_per_tick=args.direct_steps)
            rid = srv.submit(systems[k], int(steps[k]), softening=0.005, dt=1e-4)
            if res:  # the mix: a few other requests sharing the ticks
                for q in np.argsort(ns * steps)[:8]:
                    if q != k:
                        srv.submit(systems[q], int(steps[q]), softening=0.005, dt=1e-4)
            out_r = None
            while out_r is None:
                srv.tick()
                for r in srv.poll():
                    if r == rid:
                        out_r = srv.result(r)
                    else:
                        srv.result(r)
            res.append(out_r)
        c["bitwise_alone_equals_mix"] = all(np.array_equal(res[0][f], res[1][f]) for f in res[0])
    return {
        "value": best, "unit": "particle-steps/s", "dtype": "f32",
        "workload": (f"serving: one trace of {R} requests replayed with Poisson arrivals at "
                     f"{rates} req/s; N log-uniform [1e3,1e6], steps U[10,1000], eps=0.005, "
                     f"dt=1e-4, direct below 32768 else BH theta=0.5, HBM budget 200 GB, "
                     f"{args.bh_lanes} BH lanes, "
                     + (f"shortest remaining work first, {args.tick_budget:.3g} BH particle-steps "
                        f"per tick" if args.tick_budget > 0 else
                        f"{args.bh_steps} BH step(s) per request per tick")
                     + f", {args.direct_steps} direct steps per tick"),
        "p50_s": {str(r["rate_per_s"]): round(r["p50_s"], 3) for r in rows},
        "p99_s": {str(r["rate_per_s"]): round(r["p99_s"], 3) for r in rows},
        "knee_rate_per_s": best / mean_work,
        "knee_note": (f"best throughput / mean request work ({mean_work:.3g} particle-steps): the "
                      f"arrival rate this mix saturates at; every swept rate is above it, so "
                      f"latency there is set by the backlog of the {R}-request trace"),
        "bh_requests": int(np.sum(ns >= 32768)),
        "rates_skipped_for_time": skipped,
        "sweep": rows,
        "accuracy": {"same_alone_and_in_mix": check,
                     "note": "request results are bit-identical alone and in the mix; a "
                             "Barnes-Hut request's forces meet the headline accuracy "
                             "(same kernels), a direct request's the direct FP32 one"},
    }


# --------------------------------------------------------------------- calibration
def copy_probe(nbytes=1 << 28):
    """A plain device-to-device copy of `nbytes` (HIP runtime copy kernel),
    used to calibrate the FETCH_SIZE / WRITE_SIZE counters."""
    import ctypes as C

    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(nbytes)) == 0
    for _ in range(3):
        assert hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"copy_probe_bytes": nbytes}), flush=True)


def as_line(obj, K=None, W=0):
    """A nested workload object as a stand-alone contract line."""
    line = {"metric": METRIC, "value": obj["value"], "unit": obj["unit"], "n_gpus": 1,
            "steps": obj.get("steps", K), "warmup": W, "ms_per_step": obj.get("ms_per_step"),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": obj.get("vs_baseline"),
            "dtype": obj.get("dtype", "f32"), "data": "synthetic (seeded)",
            "config": {"workload": obj["workload"]}}
    line.update({k: v for k, v in obj.items() if k not in line and k != "workload"})
    return line


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="Barnes-Hut timed steps (default 10)")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=N_DEFAULT)
    ap.add_argument("--leaf-cap", type=int, default=CAP)
    ap.add_argument("--cpu-sample", type=int, default=65536)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--workload", choices=("all", "bh", "config1", "direct", "ensemble", "serve"),
                    default="all")
    ap.add_argument("--requests", type=int, default=150, help="serve: requests in the trace")
    ap.add_argument("--bh-lanes", type=int, default=8,
                    help="serve: Barnes-Hut calculators (streams) stepping concurrently")
    ap.add_argument("--bh-steps", type=int, default=1,
                    help="serve: Barnes-Hut steps per request per tick (without a tick budget)")
    ap.add_argument("--tick-budget", type=float, default=4e7,
                    help="serve: Barnes-Hut particle-steps per tick, shortest remaining work first "
                         "(0: every request every tick)")
    ap.add_argument("--direct-steps", type=int, default=64, help="serve: direct steps per tick")
    ap.add_argument("--rates", default="10,25,50,100,250,500", help="serve: arrival rates, req/s")
    ap.add_argument("--scale-n", type=int, default=1 << 24,
                    help="also report the tree stages at this N (0: skip)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a 256 MiB device copy (counter calibration)")
    ap.add_argument("--time-budget", type=float, default=520.0,
                    help="seconds for the whole run: serving rates that would overrun are skipped")
    args = ap.parse_args()
    args.deadline = time.perf_counter() + args.time_budget if args.time_budget > 0 else None
    if args.copy_probe:
        copy_probe()
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU: every box has one MI355X and the "
                                   "engine has no multi-GPU path; run with --gpus 1"}), flush=True)
        sys.exit(2)
    nested = {"config1": bench_config1, "direct": bench_direct, "ensemble": bench_ensemble,
              "serve": bench_serve}
    if args.workload in nested:
        print(json.dumps(as_line(nested[args.workload](args))), flush=True)
        return
    out = bench_bh(args)
    if args.workload == "all":
        for name, fn in nested.items():
            t0 = time.perf_counter()
            out[name] = fn(args)
            out[name]["bench_seconds"] = round(time.perf_counter() - t0, 1)
        out["host"] = host_info()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
