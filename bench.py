"""Benchmark of the Barnes-Hut leapfrog workload on one MI355X.

Workload (the headline metric's configuration): N = 1,048,576 Plummer sphere
(a = 1, M = 1, G = 1, seed 2, untruncated radii), theta = 0.5, eps = 0.005,
quadrupole moments, leaf capacity 8, leapfrog dt = 1e-4, tree rebuilt every
step; FP32 force arithmetic, FP64 positions/velocities.

One "step" = one leapfrog kick-drift-kick step = one full Barnes-Hut force
evaluation (bbox, keys, radix sort, build, up-sweep, walk) + FP64 kick/drift.
W warm-up steps, then exactly K timed steps between device synchronisations;
the state is resident in HBM throughout.

--workload direct / ensemble / serve time the other workloads instead
(N = 262,144 Plummer + disk with RK4; 4096-system ensemble; Poisson-arrival
serving sweep with p50/p99 latency).

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
PEAK_FP32_TFLOPS = 157.3  # MI355X FP32 vector / dense FP32 MFMA
PEAK_HBM_GBS = 8000.0
FLOP_PP, FLOP_PC_QUAD = 20, 60
STAGE_EVALS = 3  # untimed evaluations for the per-stage breakdown
# Published reference number for this metric: 373.34 ms per BH force evaluation
# at N = 1,000,000 (report/Report.pdf Table 4) -> 2.68e6 particle-steps/s.
REF_PUBLISHED = 1_000_000 / 0.37334


def log(*a):
    print(*a, file=sys.stderr, flush=True)


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


def traffic_bytes(kernel):
    t = measured_traffic(kernel)
    return None if t is None else t["bytes"]


def traffic_source(kernel):
    t = measured_traffic(kernel)
    return None if t is None else t["source"]


def stage_bytes(n, nodes):
    """Algorithmic HBM bytes per force evaluation for each tree stage."""
    return {
        "bbox": 24 * n,                     # FP64 x, y, z read
        "keys": (32 + 8 + 4 + 16) * n,      # x, y, z, m read; key, index, float4 written
        "sort": (8 + 8 * 24) * n,           # keys read once for the digit counts; 8 passes x (12 read + 12 write)
        "gather": (4 + 16 + 16) * n,        # perm + float4 read through it, float4 written
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
        if k in sb:
            gbs = sb[k] / (ms * 1e-3) / 1e9
            e.update({"GB/s": round(gbs, 1), "hbm_frac": round(gbs / PEAK_HBM_GBS, 4),
                      "bytes": int(sb[k])})
        out[k] = e
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


def bench_direct(args):
    """Direct tiled FP32 all-pairs with FP64 RK4 state: N = 262,144 Plummer
    (90% mass) + cold exponential disk (10%), eps = 0.01, dt = 5e-4, 20 RK4
    steps (80 force evaluations); error and speed against the FP64 kernel."""
    import nbody_amd as nb

    n = args.n if args.n != N_DEFAULT else 262144
    K = args.steps if args.steps_set else 20
    W = args.warmup
    d = nb.ic_plummer_disk(n, 5)
    f = nb.ForceCalculatorNaive(0.01, "f32", G=1.0)
    p = nb.Particles.from_arrays(d["m"], d["x"], d["y"], d["z"], d["vx"], d["vy"], d["vz"])
    e0 = nb.diagnostics(p, 1.0, 0.01)
    it = nb.RK4Integrator()
    it.run(f, p, 5e-4, W)
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
    # accuracy of one evaluation against the FP64 kernel on the final state
    f.computeAccelerations(p)
    a32 = np.stack(p.acc(), 1)
    f64 = nb.ForceCalculatorNaive(0.01, "f64", G=1.0)
    f64.set_timing(True)
    f64.computeAccelerations(p)
    a64 = np.stack(p.acc(), 1)
    t64 = f64.timing()["total_ms"]
    err = np.linalg.norm(a32 - a64, axis=1) / np.linalg.norm(a64, axis=1)
    E0, E1 = e0["ke"] + e0["pe_soft"], e1["ke"] + e1["pe_soft"]
    return {
        "metric": "interactions/s (direct) & particle-steps/s (BH, N=1M); serve p50/p99 latency",
        "value": inter / elapsed, "unit": "interactions/s", "n_gpus": 1, "steps": K, "warmup": W,
        "ms_per_step": elapsed / K * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": (inter / elapsed) / 4.12e9, "dtype": "f32",
        "data": "synthetic (seeded Plummer + exponential disk, Philox4x32-10 streams)",
        "config": {"workload": f"direct tiled all-pairs FP32 forces, FP64 RK4 state, N={n} "
                               f"Plummer (90%) + cold exponential disk (10%, R_d=1, z_d=0.05), "
                               f"eps=0.01, dt=5e-4, {K} steps (4 evaluations per step)",
                   "baseline_published": "4.12e9 interactions/s: reference FP64 GPU direct, "
                                         "N=100,000 (CUDA/outputs/output1_1.txt:130)"},
        "roofline": {"kernel": "k_direct_f32", "bound": "valu", "achieved": round(achieved, 2),
                     "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / PEAK_FP32_TFLOPS, 4),
                     "traffic": traffic_bytes("k_direct_f32"),
                     "traffic_source": traffic_source("k_direct_f32"),
                     "launch_ms": round(kern_ms, 4), "flops_per_launch": FLOP_PP * n * (n - 1),
                     "fma_probe": fma_probe(nb, achieved, {"pp": 1})},
        "accuracy": {"median_rel_vs_f64": float(np.median(err)),
                     "p99_rel_vs_f64": float(np.percentile(err, 99)),
                     "energy_drift_rel": float(abs(E1 - E0) / abs(E0))},
        "f64_direct": {"ms_per_eval": t64, "interactions_per_s": n * (n - 1) / (t64 * 1e-3)},
    }


def bench_ensemble(args):
    """4096 independent Plummer systems, N_k log-uniform in [64, 8192] and scale
    radius uniform in [0.5, 2] (seeded), direct FP32 forces with FP64 leapfrog,
    eps = 0.01, dt = 1e-3, 200 steps captured in one HIP graph; tile shapes per
    size bucket from the autotuner.  Compared with one-system-at-a-time
    launches of the same kernels (timed on a few steps)."""
    import nbody_amd as nb

    S = 4096
    K = args.steps if args.steps_set else 200
    rng = np.random.Generator(np.random.Philox(4))
    sizes = np.rint(np.exp(rng.uniform(np.log(64), np.log(8192), S))).astype(np.int64)
    radii = rng.uniform(0.5, 2.0, S)
    parts = [nb.ic_plummer(int(n), 1000 + k, a=float(a)) for k, (n, a) in enumerate(zip(sizes, radii))]
    cat = {k: np.concatenate([p[k] for p in parts]) for k in parts[0]}
    ntot = int(sizes.sum())
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
    ku = 3
    t0 = time.perf_counter()
    e.run_unbatched(ku)
    nb.synchronize()
    t_u = (time.perf_counter() - t0) / ku
    pairs = float(np.sum(sizes * (sizes - 1)))
    value = ntot * K / t_b
    return {
        "metric": "interactions/s (direct) & particle-steps/s (BH, N=1M); serve p50/p99 latency",
        "value": value, "unit": "particle-steps/s", "n_gpus": 1, "steps": K, "warmup": 2,
        "ms_per_step": t_b / K * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded Plummer spheres)",
        "config": {"workload": f"ensemble of {S} Plummer systems, N log-uniform [64, 8192] "
                               f"(total {ntot}), a ~ U[0.5, 2], direct FP32 + FP64 leapfrog, "
                               f"eps=0.01, dt=1e-3, {K} steps in one HIP graph, autotuned shapes",
                   "shapes_per_bucket": e.tuning().tolist()},
        "interactions_per_s": pairs * K / t_b,
        "unbatched": {"ms_per_step": t_u * 1e3, "particle_steps_per_s": ntot / t_u,
                      "speedup_batched": (ntot * K / t_b) / (ntot / t_u)},
    }


def bench_serve(args):
    """Serving: Poisson arrivals swept over rates; each request is a seeded
    Plummer system with N log-uniform in [1e3, 1e6] and a step count uniform in
    [10, 1000], eps = 0.005, dt = 1e-4; direct method below N = 32,768,
    Barnes-Hut (theta = 0.5) above; HBM budget 200 GB.  Requests are submitted
    at their (replayed) arrival times while earlier ones run; reports p50/p99
    completion latency (arrival -> result on the host), split into queueing
    (arrival -> admission) and service (admission -> result), and aggregate
    particle-steps/s per rate."""
    import nbody_amd as nb

    rates = [10, 25, 50, 100, 250, 500]
    per_rate = args.requests
    rows = []
    for ri, rate in enumerate(rates):
        rng = np.random.Generator(np.random.Philox(1000 + ri))
        gaps = rng.exponential(1.0 / rate, per_rate)
        arrivals = np.cumsum(gaps) - gaps[0]
        ns = np.rint(np.exp(rng.uniform(np.log(1e3), np.log(1e6), per_rate))).astype(int)
        steps = rng.integers(10, 1001, per_rate)
        systems = [nb.ic_plummer(int(n), 50000 + 100 * ri + k) for k, n in enumerate(ns)]
        srv = nb.Server(hbm_budget_bytes=200e9, direct_threshold=32768, bh_lanes=args.bh_lanes)
        t0 = time.perf_counter()
        sub = 0
        done = {}
        admitted = {}
        rid_of = {}
        while len(done) < per_rate:
            now = time.perf_counter() - t0
            while sub < per_rate and arrivals[sub] <= now:
                rid_of[srv.submit(systems[sub], int(steps[sub]), softening=0.005, dt=1e-4)] = sub
                sub += 1
            st = srv.status()
            if st["active"] == 0 and st["pending"] == 0:
                if sub < per_rate:
                    time.sleep(max(0.0, arrivals[sub] - (time.perf_counter() - t0)))
                continue
            srv.tick()
            t = time.perf_counter() - t0
            for rid in srv.poll_admitted():
                admitted[rid_of[rid]] = t
            for rid in srv.poll():
                k = rid_of[rid]
                srv.result(rid)
                done[k] = t - arrivals[k]
        wall = time.perf_counter() - t0
        lat = np.array([done[k] for k in range(per_rate)])
        queue = np.array([admitted[k] - arrivals[k] for k in range(per_rate)])
        service = lat - queue
        work = float(np.sum(ns * steps))
        rows.append({"rate_per_s": rate, "requests": per_rate, "p50_s": float(np.percentile(lat, 50)),
                     "p99_s": float(np.percentile(lat, 99)),
                     "queue_p50_s": float(np.percentile(queue, 50)),
                     "queue_p99_s": float(np.percentile(queue, 99)),
                     "service_p50_s": float(np.percentile(service, 50)),
                     "service_p99_s": float(np.percentile(service, 99)),
                     "particle_steps_per_s": work / wall,
                     "wall_s": wall, "bh_requests": int(np.sum(ns >= 32768))})
        log(f"[serve] rate {rate}: p50 {rows[-1]['p50_s']:.2f} s, p99 {rows[-1]['p99_s']:.2f} s "
            f"(queueing p50 {rows[-1]['queue_p50_s']:.2f} s, service p50 {rows[-1]['service_p50_s']:.2f} s), "
            f"{rows[-1]['particle_steps_per_s']:.3g} particle-steps/s")
    best = max(r["particle_steps_per_s"] for r in rows)
    return {
        "metric": "interactions/s (direct) & particle-steps/s (BH, N=1M); serve p50/p99 latency",
        "value": best, "unit": "particle-steps/s", "n_gpus": 1, "steps": per_rate, "warmup": 0,
        "ms_per_step": None, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic (seeded Poisson arrival traces, Plummer systems)",
        "config": {"workload": f"serving: Poisson arrivals at {rates} req/s, {per_rate} requests "
                               f"per rate, N log-uniform [1e3,1e6], steps U[10,1000], eps=0.005, "
                               f"dt=1e-4, direct below 32768 else BH theta=0.5, HBM budget 200 GB"},
        "sweep": rows,
    }


# Instructions of the compiled interaction terms besides their one v_rsq_f32
# (gfx950 ISA of the inner loops): the direct FP32 term with softening and the
# walk's particle term are 3 sub + 3 mul + 6 FMA (20 FLOP); the walk's
# quadrupole cell term is 3 add/sub + 9 mul + 18 FMA (60 FLOP).
TERM_OTHER_INSTR = {"pp": 12, "pc": 30}
TERM_FLOP = {"pp": 20, "pc": 60}


def fma_probe(nb, achieved, terms=None):
    """The FP32 ceiling without packed FP32 (the library is built without it):
    v_fma_f32 issue rate with VGPR operands, measured in this run outside the
    timed region, and the kernel's FLOP rate as a fraction of it.  With
    `terms` ({"pp": count, "pc": count}), also the ceiling of that term mix:
    the FMA slot rate (probe / 2 FLOP) times FLOP per slot, a term taking its
    other instructions plus the measured rsqrt cost in slots."""
    r = nb.probe_fp32_fma()
    out = {"tflops": round(r["tflops"], 2), "frac": round(achieved / r["tflops"], 4),
           "what": "16 independent v_fma_f32 chains per lane, VGPR operands, 8 x 256-lane "
                   "blocks per CU (nb_probe_fp32_fma)"}
    if terms:
        c = nb.probe_fp32_rsq_cost()
        flop = sum(TERM_FLOP[k] * v for k, v in terms.items())
        slots = sum((TERM_OTHER_INSTR[k] + c) * v for k, v in terms.items())
        ceiling = r["tflops"] / 2.0 * flop / slots
        out["rsq_fma_slots"] = round(c, 2)
        out["mix_ceiling_tflops"] = round(ceiling, 2)
        out["mix_frac"] = round(achieved / ceiling, 4)
    return out


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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=N_DEFAULT)
    ap.add_argument("--leaf-cap", type=int, default=CAP)
    ap.add_argument("--cpu-sample", type=int, default=65536)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--workload", choices=("bh", "direct", "ensemble", "serve"), default="bh")
    ap.add_argument("--requests", type=int, default=16)
    ap.add_argument("--bh-lanes", type=int, default=8,
                    help="serve workload: Barnes-Hut calculators (streams) stepping concurrently")
    ap.add_argument("--scale-n", type=int, default=1 << 24,
                    help="also report the tree stages at this N (0: skip)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a 256 MiB device copy (counter calibration)")
    args = ap.parse_args()
    args.steps_set = args.steps is not None
    if args.copy_probe:
        copy_probe()
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU: every box has one MI355X and the "
                                   "engine has no multi-GPU path; run with --gpus 1"}), flush=True)
        sys.exit(2)
    if args.workload != "bh":
        out = {"direct": bench_direct, "ensemble": bench_ensemble,
               "serve": bench_serve}[args.workload](args)
        print(json.dumps(out), flush=True)
        return

    import nbody_amd as nb

    n, K, W = args.n, (args.steps if args.steps_set else 10), args.warmup
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
    calls = stages["calls"]

    # Walk roofline (FP32 VALU): interactions of the last evaluation are
    # counted by the kernel; the tree barely changes over 10 steps.
    flops = FLOP_PP * st["pp_interactions"] + FLOP_PC_QUAD * st["pc_interactions"]
    achieved = flops / (walk_ms * 1e-3) / 1e12
    roofline = {"kernel": "k_walk", "bound": "valu", "achieved": round(achieved, 2),
                "peak_kind": "FP32 vector (157.3 TFLOP/s is also the dense FP32 MFMA figure)",
                "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP32_TFLOPS, 4),
                "traffic": traffic_bytes("k_walk"), "traffic_source": traffic_source("k_walk"),
                "launch_ms": round(walk_ms, 4),
                "flops_per_launch": int(flops),
                "interactions_per_particle": {"pp": round(st["pp_interactions"] / n, 1),
                                              "pc": round(st["pc_interactions"] / n, 1)},
                "fma_probe": fma_probe(nb, achieved, {"pp": st["pp_interactions"],
                                                      "pc": st["pc_interactions"]})}
    stage_out = stage_report(stages, n, st["n_nodes"])

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
