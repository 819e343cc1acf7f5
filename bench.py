#!/usr/bin/env python3
"""Throughput of the force + integration hot path on one MI355X.

Headline (the `value`): LBVH Barnes-Hut at N = 2^20 bodies (half Plummer
sphere, half uniform cube [-4, 4]^3, FP32 positions/masses, eps = 0.005,
theta = 0.5). One step = one leapfrog step: box + Morton codes + radix sort +
Karras build + bottom-up moments + links + traversal + kick/drift, with inputs
resident in HBM. value = accepted body-node interactions per second.

Beside it, in the same JSON line: all-pairs FP32 (Kahan) at N = 65536 with its
own roofline, per-phase HIP-event timings, force errors against FP64
all-pairs, and the CPU path of the reference (restated in oracle/) timed on
this host's cores on a bounded sample.

python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "pair interactions/s (all-pairs N=64K, BH N=1M); adjoint/forward step time ratio"
FP32_PEAK_TFLOPS = 157.3  # MI355X published FP32 vector peak
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def rel_err(a, b):
    return np.linalg.norm(a - b, axis=1) / np.linalg.norm(b, axis=1)


def err_stats(e):
    return {"median": float(np.median(e)), "p99": float(np.quantile(e, 0.99)),
            "rms": float(np.sqrt(np.mean(e ** 2))), "max": float(e.max())}


def host_threads():
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def bench_bh(A, ctx, steps, warmup, method, n=1 << 20, theta=0.5, dt=5e-4):
    pos, vel, mass = A.ic_mixed(n, 20240)
    ctx.set_params(G=1.0, softening=0.005 ** 2, theta=theta, precision=A.FP32)
    ctx.set_state(pos, vel, mass)
    # accuracy of one evaluation against FP64 all-pairs (device path that is
    # bit-identical to the reference CPU path)
    log(f"[bench] BH n={n}: FP64 all-pairs reference")
    ref = ctx.forces(A.ALLPAIRS_FP64)
    acc = ctx.forces(method)
    ctx.bh_stats()
    errs = err_stats(rel_err(acc, ref))
    log(f"[bench] BH force error vs FP64 all-pairs: {errs}")
    # warmup, then K timed leapfrog steps
    ctx.step(method, A.LEAPFROG, dt, warmup)
    ctx.sync()
    ctx.bh_stats()  # reset counters
    ctx.profile(True)
    ctx.sync()
    t0 = time.perf_counter()
    ctx.mark(0)
    ctx.step(method, A.LEAPFROG, dt, steps)
    ctx.mark(1)
    ctx.sync()
    wall = time.perf_counter() - t0
    ev_ms = ctx.elapsed(0, 1)
    phases = ctx.phase_times()
    ctx.profile(False)
    inter, visits = ctx.bh_stats()
    log(f"[bench] BH {steps} steps: wall {wall * 1e3:.2f} ms, events {ev_ms:.2f} ms, phases {phases}")
    return dict(wall_s=wall, event_ms=ev_ms, phases=phases, interactions=inter, visits=visits, errors=errs,
                n=n, theta=theta, dt=dt)


def bench_allpairs(A, ctx, evals, n=65536):
    pos, vel, mass = A.ic_plummer(n, 3, mass_mode=1)
    ctx.set_params(G=1.0, softening=0.005 ** 2, theta=0.5)
    ctx.set_state(pos, vel, mass)
    ref = ctx.forces(A.ALLPAIRS_FP64)
    acc = ctx.forces(A.ALLPAIRS_FP32_KAHAN)
    errs = err_stats(rel_err(acc, ref))
    for _ in range(3):
        ctx.forces(A.ALLPAIRS_FP32_KAHAN, out=False)
    ctx.profile(True)
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(evals):
        ctx.forces(A.ALLPAIRS_FP32_KAHAN, out=False)
    ctx.sync()
    wall = time.perf_counter() - t0
    phases = ctx.phase_times()
    ctx.profile(False)
    ms_k, cnt = phases["allpairs"]
    ms_c, _ = phases.get("combine", (0.0, 1))
    kernel_ms = ms_k / cnt
    inter = float(n) * (n - 1)
    flops = 20.0 * inter
    achieved = flops / (kernel_ms * 1e-3) / 1e12
    log(f"[bench] all-pairs n={n}: kernel {kernel_ms:.3f} ms, combine {ms_c / cnt:.4f} ms, "
        f"{achieved:.1f} TFLOP/s, err {errs}")
    return {
        "n": n,
        "value": inter / ((ms_k + ms_c) / cnt * 1e-3),
        "unit": "interactions/s",
        "ms_per_eval": (ms_k + ms_c) / cnt,
        "wall_ms_per_eval": wall / evals * 1e3,
        "kernel_ms": kernel_ms,
        "force_error_vs_fp64": errs,
        "roofline": {"bound": "valu", "achieved": achieved, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP32_PEAK_TFLOPS, "traffic": None,
                     "kernel": "k_allpairs_f32<TILE_KAHAN>", "flop_per_interaction": 20},
    }, pos, vel, mass


def cpu_baseline_bh(orc, A, threads, n=1 << 20, theta=0.5, stride=16):
    pos, vel, mass = A.ic_mixed(n, 20240)
    box = orc.box_auto(pos)
    tb, tt, inter = orc.bh_sample(pos, vel, mass, box, theta, stride, soft=0.005 ** 2, threads=threads)
    value = inter / (tt + tb / stride)
    log(f"[bench] CPU BH sample: build {tb:.2f}s traverse {tt:.2f}s inter {inter} -> {value:.3e}/s")
    return {"value": value, "unit": "interactions/s", "cores": threads, "kind": "port",
            "sample": f"Aster CPU Barnes_Hut restated (FP64, reference MAC theta={theta}): make_tree over all "
                      f"{n} bodies ({tb:.2f} s) + traversal of every {stride}th body ({tt:.2f} s, {inter} "
                      f"interactions); build time amortised per sampled body",
            "cpu": cpu_model()}


def cpu_baseline_allpairs(orc, pos, vel, mass, threads, targets=16384):
    n = len(mass)
    _, sec = orc.allpairs_sample(pos, vel, mass, 0, targets, soft=0.005 ** 2, threads=threads)
    inter = float(targets) * (n - 1)
    log(f"[bench] CPU all-pairs sample: {targets} targets x {n} sources in {sec:.2f}s")
    return {"value": inter / sec, "unit": "interactions/s", "cores": threads, "kind": "port",
            "sample": f"Aster CPU Parallelized::compute_forces restated (FP64): {targets} targets x {n} sources",
            "cpu": cpu_model()}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--method", default="BH_GROUP")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (see DESIGN.md, "
              "'More than one GPU'); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    import aster_amd as A
    ctx = A.Context(0)
    method = getattr(A, args.method)
    bh = bench_bh(A, ctx, args.steps, args.warmup, method)
    ap_res, pos64k, vel64k, mass64k = bench_allpairs(A, ctx, max(args.steps, 10))

    # roofline of the dominant kernel of the headline step: the traversal
    ms_t, cnt_t = bh["phases"]["traverse"]
    trav_ms = ms_t / cnt_t
    inter_per_eval = bh["interactions"] / args.steps
    visits_per_eval = bh["visits"] / args.steps
    flops = 20.0 * inter_per_eval + 8.0 * visits_per_eval
    achieved = flops / (trav_ms * 1e-3) / 1e12
    ms_per_step = bh["wall_s"] / args.steps * 1e3
    value = inter_per_eval / (bh["wall_s"] / args.steps)

    cpu = None
    if not args.no_cpu:
        from oracle import oracle as orc
        th = host_threads()
        cpu = cpu_baseline_bh(orc, A, th)
        ap_res["cpu_baseline"] = cpu_baseline_allpairs(orc, pos64k, vel64k, mass64k, th)

    phases = {k: {"ms_per_launch": v[0] / v[1], "launches_per_step": v[1] / args.steps}
              for k, v in bh["phases"].items()}
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "interactions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 forces, f64 state",
        "data": "synthetic: seeded half Plummer sphere + half uniform cube [-4,4]^3, FP32 positions/masses",
        "config": {"workload": "LBVH Barnes-Hut N=1048576 theta=0.5 eps=0.005, one leapfrog step "
                               "(tree rebuild + traversal + kick/drift) per step, dt=5e-4",
                   "n": bh["n"], "theta": bh["theta"], "method": args.method, "dt": bh["dt"]},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP32_PEAK_TFLOPS, "traffic": None, "kernel": "traversal",
                     "kernel_ms": trav_ms, "flops": "20 per interaction + 8 per node visit"},
        "cpu_baseline": cpu,
        "interactions_per_step": inter_per_eval,
        "node_visits_per_step": visits_per_eval,
        "force_error_vs_fp64_allpairs": bh["errors"],
        "phases_ms": phases,
        "event_ms_per_step": bh["event_ms"] / args.steps,
        "allpairs": ap_res,
        "adjoint_forward_ratio": None,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
