#!/usr/bin/env python3
"""Throughput of the force + integration hot path on one MI355X.

Headline (the `value`): LBVH Barnes-Hut at N = 2^20 bodies (half Plummer
sphere, half uniform cube [-4, 4]^3, FP32 positions/masses, eps = 0.005,
theta = 0.5). One step = one leapfrog step: box + Morton codes + radix sort +
Karras build + bottom-up moments (which also write the traversal's child
records) + traversal + kick/drift, with inputs resident in HBM. value =
accepted body-node interactions per second.

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
FP32_PEAK_TFLOPS = 157.3  # MI355X published FP32 vector peak (packed FP32 FMAs)
# One wave64 FP32 FMA per SIMD quad-cycle (1024 SIMDs, 2.4 GHz): the ceiling
# without packed FP32 instructions, which this build does not emit.
FP32_NONPACKED_TFLOPS = 78.6
FP64_PEAK_TFLOPS = 78.6   # MI355X published FP64 vector peak
# FP32 operations (an FMA counts 2) per source-target pair of the fused
# adjoint kernel k_adj_fused: separation 3, r^2 6, rsqrt 1, inverse powers 2,
# m inv^3 1, acceleration 6, u = m_k w_j - m_j w_k 9, d.u 5, s = 3 (d.u) inv^2 2,
# J^T w = inv^3 (u - s d) 12, w.d 5, mass gradient 2.
ADJ_FLOP_PER_PAIR = 54
# FP64 operations per ordered pair and force evaluation of the ensemble
# kernel's fast mode with pairwise 1PN (what bench_batch runs): Newton 26
# (separation 3, r^2 6, rsqrt estimate + one Newton step 7, G m inv^3 4,
# accumulation 6) + 1PN 87 (unit vector 3, get_pn1 78, accumulation 6).
BATCH_FLOP_PER_PAIR = 113
HBM_PEAK_GBPS = 8000.0    # MI355X published HBM3E bandwidth

# Algorithmic HBM bytes per body of the bandwidth-bound phases of a BH step
# (what each phase must read and write once):
#  morton: Morton pass 24 B x + 8 B m read, 4 B code + 4 B index + 32 B
#          {x, m} record written (the box's maxima come from the drift that
#          wrote x; the drift reads x anyway);
#  sort: 4 passes x (4 B histogram read + 8 B key/value read + 8 B written);
#  build: 4 B code + 4 B permutation + 32 B record read, 48 B of leaf records
#         (FP64 {x, m}, FP32 sorted body) and 24 B of internal-node links
#         (left, right, parent, range, split) written, plus the group pass's
#         12 B (start flag, climb, root);
#  kickdrift: FP64 x, v, a read (72 B), x, v written (48 B).
PHASE_BYTES_PER_BODY = {"morton": 24 + 8 + 4 + 4 + 32, "sort": 4 * (4 + 8 + 8), "build": 4 + 4 + 32 + 48 + 24 + 12,
                        "kickdrift": 72 + 48}


def traffic_per_launch(kernel):
    """HBM bytes per launch of `kernel` (FETCH_SIZE + WRITE_SIZE, each scaled by
    the same counter's reading for a copy of known size), as measured with
    rocprofv3 --pmc and recorded in profiles/traffic.json; None if absent."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "traffic.json")
    try:
        with open(path) as f:
            rec = json.load(f)["kernels"][kernel]
        return rec["fetch_bytes"] + rec["write_bytes"]
    except (OSError, KeyError, ValueError):
        return None


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def rel_err(a, b):
    return np.linalg.norm(a - b, axis=1) / np.linalg.norm(b, axis=1)


def err_stats(e):
    return {"median": float(np.median(e)), "p99": float(np.quantile(e, 0.99)),
            "rms": float(np.sqrt(np.mean(e ** 2))), "max": float(e.max())}


def host_threads():
    """Host threads of the CPU baselines: the process's CPU share, i.e.
    OMP_NUM_THREADS when it is set (it is, on the GPU box), else every core
    of the affinity mask."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return max(1, n)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def params(A, ctx, **kw):
    """Set every context parameter: the library's defaults overridden by kw.
    Parameters persist on a context, so each section states all of them."""
    p = dict(G=6e-11, c=299792458.0, softening=1e-14, theta=0.2, box=(0.0, 0.0, 0.0), pn_flags=A.NEWTON,
             precision=A.FP32, lower=0, upper=-1, split=0)
    p.update(kw)
    return ctx.set_params(**p)


def bench_bh(A, ctx, steps, warmup, method, n=1 << 20, theta=0.5, dt=5e-4, reference=True):
    pos, vel, mass = A.ic_mixed(n, 20240)
    params(A, ctx, G=1.0, softening=0.005 ** 2, theta=theta, precision=A.FP32)
    ctx.set_state(pos, vel, mass)
    errs = None
    ref = None
    if reference:
        # accuracy of one evaluation against FP64 all-pairs (device path that
        # is bit-identical to the reference CPU path)
        log(f"[bench] BH n={n}: FP64 all-pairs reference")
        ref = ctx.forces(A.ALLPAIRS_FP64)
        acc = ctx.forces(method)
        acc_headline = acc
        ctx.bh_stats()
        errs = err_stats(rel_err(acc, ref))
        log(f"[bench] BH force error vs FP64 all-pairs: {errs}")
        # the reference's own MAC (s^2 < theta^2 d^2 with s = |box| 2^-depth) at the same theta
        ctx.forces(A.BH_PARITY, out=False)
        ctx.bh_stats()
        acc = ctx.forces(A.BH_PARITY)
        inter_ref, _ = ctx.bh_stats()
        errs["reference_mac"] = err_stats(rel_err(acc, ref))
        errs["reference_mac"]["interactions_per_body"] = inter_ref / n
        log(f"[bench] reference MAC at theta={theta}: {errs['reference_mac']}")
        # the workload's other opening angles: accuracy against the same
        # reference, then a short timed run from the same initial state
        sweep = {}
        for th in (0.3, 0.7):
            params(A, ctx, G=1.0, softening=0.005 ** 2, theta=th, precision=A.FP32)
            ctx.bh_stats()
            e = err_stats(rel_err(ctx.forces(method), ref))
            inter_th, _ = ctx.bh_stats()
            ctx.set_state(pos, vel, mass)
            ctx.step(method, A.LEAPFROG, dt, warmup)
            ctx.sync()
            ctx.bh_stats()
            t0 = time.perf_counter()
            ctx.step(method, A.LEAPFROG, dt, steps)
            ctx.sync()
            wall_th = time.perf_counter() - t0
            inter_steps, _ = ctx.bh_stats()
            e.update(interactions_per_body=inter_th / n, ms_per_step=wall_th / steps * 1e3,
                     interactions_per_s=inter_steps / wall_th)
            sweep[str(th)] = e
            log(f"[bench] theta={th}: {e}")
        errs["theta_sweep"] = sweep
        # the same traversal with the pairwise 1PN terms (added in FP64 over
        # the same lists): force evaluation time against the Newtonian one
        params(A, ctx, G=1.0, softening=0.005 ** 2, theta=theta, precision=A.FP32)
        ctx.set_state(pos, vel, mass)
        pn_ms = {}
        for name, flags in (("newton", A.NEWTON), ("newton_pn1", A.NEWTON | A.PN1)):
            params(A, ctx, G=1.0, softening=0.005 ** 2, theta=theta, precision=A.FP32, c=1e4, pn_flags=flags)
            ctx.forces(method, out=False)
            ctx.sync()
            ctx.profile(True)
            for _ in range(3):
                ctx.forces(method, out=False)
            ph = ctx.phase_times()
            ctx.profile(False)
            pn_ms[name] = ph["traverse"][0] / ph["traverse"][1]
        ctx.bh_stats()
        errs["pn1_traverse_ms"] = pn_ms
        log(f"[bench] traversal with the 1PN terms: {pn_ms}")
        params(A, ctx, G=1.0, softening=0.005 ** 2, theta=theta, precision=A.FP32)
    # warmup, then K timed leapfrog steps (no per-phase events inside the
    # timed region: they cost ~2.5% of a step)
    ctx.step(method, A.LEAPFROG, dt, warmup)
    ctx.sync()
    ctx.bh_stats()  # reset counters
    t0 = time.perf_counter()
    ctx.mark(0)
    ctx.step(method, A.LEAPFROG, dt, steps)
    ctx.mark(1)
    ctx.sync()
    wall = time.perf_counter() - t0
    ev_ms = ctx.elapsed(0, 1)
    inter, visits = ctx.bh_stats()
    # the same number of steps again with HIP events around each phase (the
    # per-kernel times the roofline uses)
    ctx.profile(True)
    ctx.step(method, A.LEAPFROG, dt, steps)
    ctx.sync()
    phases = ctx.phase_times()
    ctx.profile(False)
    ctx.bh_stats()
    log(f"[bench] BH {steps} steps: wall {wall * 1e3:.2f} ms, events {ev_ms:.2f} ms, phases {phases}")
    return dict(wall_s=wall, event_ms=ev_ms, phases=phases, interactions=inter, visits=visits, errors=errs,
                n=n, theta=theta, dt=dt, ref=ref, acc=acc_headline if reference else None, state=(pos, vel, mass))


def bench_scaling(A, ctx, method, sizes=(18, 20, 22, 24, 26)):
    """Barnes-Hut leapfrog step against N on the same workload family as the
    headline step (theta = 0.5); fewer timed steps at the larger sizes."""
    rows = []
    for lg in sizes:
        n = 1 << lg
        steps = 10 if lg <= 22 else (5 if lg <= 24 else 3)
        warmup = 20 if lg <= 22 else 3  # at least ~30 ms of steps before timing
        r = bench_bh(A, ctx, steps, warmup, method, n=n, reference=False)
        ms_t, cnt_t = r["phases"]["traverse"]
        step_ms = r["wall_s"] / steps * 1e3
        inter = r["interactions"] / steps
        rows.append({"n": n, "steps": steps, "ms_per_step": step_ms, "traverse_ms": ms_t / cnt_t,
                     "interactions_per_body": inter / n, "interactions_per_s": inter / (step_ms * 1e-3),
                     "bodies_per_s": n / (step_ms * 1e-3)})
        log(f"[bench] scaling {rows[-1]}")
    return rows


def bench_allpairs(A, ctx, evals, n=65536):
    """All-pairs at the mixed-precision configuration's shape (Plummer N=65536,
    log-uniform masses, eps=0.005): both FP32 methods, timed per kernel."""
    pos, vel, mass = A.ic_plummer(n, 3, mass_mode=1)
    params(A, ctx, G=1.0, softening=0.005 ** 2, theta=0.5)
    ctx.set_state(pos, vel, mass)
    ref = ctx.forces(A.ALLPAIRS_FP64)
    out = {}
    for name, method, vop in (("fp32_mixed_kahan", A.ALLPAIRS_FP32_KAHAN, 15), ("fp32", A.ALLPAIRS_FP32, 12)):
        errs = err_stats(rel_err(ctx.forces(method), ref))
        for _ in range(3):
            ctx.forces(method, out=False)
        ctx.profile(True)
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(evals):
            ctx.forces(method, out=False)
        ctx.sync()
        wall = time.perf_counter() - t0
        phases = ctx.phase_times()
        ctx.profile(False)
        eval_ms = sum(v[0] for v in phases.values()) / evals
        kernel_ms = phases["allpairs"][0] / phases["allpairs"][1]
        inter = float(n) * (n - 1)
        achieved = 20.0 * inter / (kernel_ms * 1e-3) / 1e12
        log(f"[bench] all-pairs {name} n={n}: kernel {kernel_ms:.3f} ms, eval {eval_ms:.3f} ms, "
            f"{achieved:.1f} TFLOP/s (20-flop convention), err {errs}")
        out[name] = {
            "value": inter / (eval_ms * 1e-3),
            "unit": "interactions/s",
            "ms_per_eval": eval_ms,
            "wall_ms_per_eval": wall / evals * 1e3,
            "kernel_ms": kernel_ms,
            "force_error_vs_fp64": errs,
            "roofline": {"bound": "valu", "achieved": achieved, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                         "frac": achieved / FP32_PEAK_TFLOPS,
                         "frac_of_nonpacked_ceiling": achieved / FP32_NONPACKED_TFLOPS,
                         "traffic": traffic_per_launch("k_allpairs_f32_mixed" if vop == 15 else "k_allpairs_f32_plain"),
                         "kernel": "k_allpairs_f32",
                         "flop_per_interaction": 20, "valu_instructions_per_interaction": vop + 1},
        }
    res = dict(out["fp32_mixed_kahan"])
    res["n"] = n
    res["method"] = "ALLPAIRS_FP32_KAHAN (FP64 separations, FP32 pair force, Kahan FP32 sums)"
    res["fp32_plain"] = out["fp32"]
    return res, pos, vel, mass


def bench_integrators(A, ctx, n=65536, dt=2.5e-4, steps=1000, steps64=20):
    """Mixed-precision configuration: Plummer N=65536 (masses log-uniform over
    two decades), eps=0.005, Yoshida-4 dt=2.5e-4 for 1000 steps with FP32 Kahan
    pair forces and FP64 state: relative drift of the Plummer energy and of
    |L|. The first `steps64` steps are repeated on the all-FP64 path (the
    bit-exact reference path) to compare drifts and positions. Then one step
    of each integrator of the family, timed with the mixed forces."""
    pos, vel, mass = A.ic_plummer(n, 3, mass_mode=1)
    soft = 0.005 ** 2
    params(A, ctx, G=1.0, softening=soft)

    def drift(method, nsteps):
        ctx.set_state(pos, vel, mass)
        _, e0, L0 = ctx.energy()
        t0 = time.perf_counter()
        ctx.step(method, A.YOSHIDA4, dt, nsteps)
        ctx.sync()
        wall = time.perf_counter() - t0
        _, e1, L1 = ctx.energy()
        p, _ = ctx.get_state()
        return dict(dE_rel=abs(e1 - e0) / abs(e0), dL_rel=float(np.linalg.norm(L1 - L0) / np.linalg.norm(L0)),
                    ms_per_step=wall / nsteps * 1e3), p

    mixed, _ = drift(A.ALLPAIRS_FP32_KAHAN, steps)
    mixed_short, p32 = drift(A.ALLPAIRS_FP32_KAHAN, steps64)
    fp64_short, p64 = drift(A.ALLPAIRS_FP64, steps64)
    dx = float(np.abs(p32 - p64).max())
    log(f"[bench] Yoshida-4 n={n}: mixed {steps} steps {mixed}; {steps64} steps mixed {mixed_short} "
        f"fp64 {fp64_short}, max |dx| {dx:.2e}")
    fam, fam_drift = {}, {}
    nfam = 40
    for name, integ, order in (("leapfrog", A.LEAPFROG, 0), ("yoshida4", A.YOSHIDA4, 0), ("saba4", A.SABA, 4),
                               ("saba10", A.SABA, 10)):
        ctx.set_state(pos, vel, mass)
        _, e0, L0 = ctx.energy()
        ctx.sync()
        t0 = time.perf_counter()
        ctx.integrate(A.ALLPAIRS_FP32_KAHAN, integ, order, dt, nfam)
        ctx.sync()
        fam[name] = (time.perf_counter() - t0) / nfam * 1e3
        _, e1, L1 = ctx.energy()
        fam_drift[name] = {"dE_rel": abs(e1 - e0) / abs(e0),
                           "dL_rel": float(np.linalg.norm(L1 - L0) / np.linalg.norm(L0))}
    log(f"[bench] integrator family ms/step (mixed forces, n={n}): {fam}; drift over {nfam} steps {fam_drift}")
    return {"n": n, "dt": dt, "integrator": "YOSHIDA4", "forces": "ALLPAIRS_FP32_KAHAN",
            f"mixed_{steps}_steps": mixed, f"mixed_{steps64}_steps": mixed_short, f"fp64_{steps64}_steps": fp64_short,
            f"max_abs_dx_mixed_vs_fp64_{steps64}_steps": dx, "ms_per_step_by_integrator": fam,
            f"drift_by_integrator_{nfam}_steps": fam_drift}


def bench_reference_case(A, ctx, cpu=True, n=4096, steps=100):
    """The reference's CPU-runnable case: Plummer N=4096 (equal masses, b = 1),
    eps = 0.01, FP64 all-pairs, leapfrog dt = (double)(float)1e-3, 100 steps.
    The device path is bit-identical to the reference's CPU path restated in
    the oracle; both are timed (the oracle on the host's cores) and the final
    states compared."""
    dt = float(np.float32(1e-3))
    pos, vel, mass = A.ic_plummer(n, 1)
    params(A, ctx, G=1.0, softening=1e-4)
    ctx.set_state(pos, vel, mass)
    e_ref0, e_pl0, L0 = ctx.energy()
    ctx.sync()
    t0 = time.perf_counter()
    ctx.step(A.ALLPAIRS_FP64, A.LEAPFROG, dt, steps)
    ctx.sync()
    wall = time.perf_counter() - t0
    p, v = ctx.get_state()
    e_ref1, e_pl1, L1 = ctx.energy()
    out = {"n": n, "steps": steps, "dt": dt, "gpu_ms_per_step": wall / steps * 1e3,
           "dE_rel_reference_definition": abs(e_ref1 - e_ref0) / abs(e_ref0),
           "dE_rel_plummer": abs(e_pl1 - e_pl0) / abs(e_pl0),
           "dL_rel": float(np.linalg.norm(L1 - L0) / np.linalg.norm(L0))}
    if cpu:
        from oracle import oracle as orc
        th = host_threads()
        t0 = time.perf_counter()
        rp, rv, _ = orc.integrate(pos, vel, mass, dt, steps, integrator=2, soft=1e-4, threads=th)
        cw = time.perf_counter() - t0
        out.update({"cpu_ms_per_step": cw / steps * 1e3, "cpu_threads": th, "cpu_kind": "port",
                    "bit_identical": bool(np.array_equal(p, rp) and np.array_equal(v, rv))})
    log(f"[bench] reference case (Plummer {n}, FP64, leapfrog, {steps} steps): {out}")
    return out


def bench_adjoint(A, ctx, n=65536, T=512, ckpt=64, dt=1e-3):
    """Differentiable-trajectory configuration: Plummer N=65536, FP32 forces /
    FP64 state, leapfrog dt=1e-3 over T=512 steps, 8 snapshots, checkpoints
    every 64 steps. Ratio = one backward step / one forward step."""
    pos, vel, mass = A.ic_plummer(n, 5)
    rng = np.random.default_rng(5)
    snaps = np.arange(1, 9) * (T // 8)
    params(A, ctx, G=1.0, softening=0.005 ** 2)
    ctx.set_state(pos + rng.normal(scale=1e-3, size=pos.shape), vel + rng.normal(scale=1e-3, size=vel.shape), mass)
    target = ctx.traj_forward(dt, T, ckpt, snaps)
    ctx.set_state(pos, vel, mass)
    ctx.sync()
    t0 = time.perf_counter()
    snap_pos = ctx.traj_forward(dt, T, ckpt, snaps)
    ctx.sync()
    t_fwd = time.perf_counter() - t0
    d = snap_pos - target
    loss = float((d ** 2).sum() / (len(snaps) * n))
    gsnap = 2.0 * d / (len(snaps) * n)
    t0 = time.perf_counter()
    gx, gv, gm = ctx.traj_backward(gsnap)
    t_bwd = time.perf_counter() - t0
    ratio = t_bwd / t_fwd
    log(f"[bench] adjoint n={n} T={T}: forward {t_fwd:.3f}s backward {t_bwd:.3f}s ratio {ratio:.3f}, loss {loss:.3e}")
    # the fused kernel's own time, from HIP events over a short backward pass
    ctx.set_state(pos, vel, mass)
    ctx.traj_forward(dt, ckpt, ckpt, np.array([ckpt]))
    ctx.profile(True)
    ctx.traj_backward(np.zeros((1, n, 3)))
    ph = ctx.phase_times()
    ctx.profile(False)
    k_ms = ph["adjoint"][0] / ph["adjoint"][1]
    achieved = ADJ_FLOP_PER_PAIR * float(n) * (n - 1) / (k_ms * 1e-3) / 1e12
    log(f"[bench] k_adj_fused: {k_ms:.3f} ms per launch, {achieved:.1f} TFLOP/s")
    return {"n": n, "steps": T, "checkpoint_every": ckpt, "forward_s": t_fwd, "backward_s": t_bwd,
            "forward_ms_per_step": t_fwd / T * 1e3, "backward_ms_per_step": t_bwd / T * 1e3,
            "ratio": ratio, "loss": loss, "grad_norms": [float(np.linalg.norm(g)) for g in (gx, gv, gm)],
            "roofline": {"bound": "valu", "kernel": "k_adj_fused", "kernel_ms": k_ms, "achieved": achieved,
                         "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": achieved / FP32_PEAK_TFLOPS,
                         "frac_of_nonpacked_ceiling": achieved / FP32_NONPACKED_TFLOPS,
                         "traffic": traffic_per_launch("k_adj_fused"),
                         "flop_per_pair": ADJ_FLOP_PER_PAIR}}


def bench_batch(A, ctx, nsys=8192, steps=10000, dt=1e-3):
    """Planetary ensembles: 8192 systems of N in [3, 64], 1PN (c=1e4) in FP64,
    Yoshida-4 dt=1e-3 for 10^4 steps in captured launches."""
    off, pos, vel, mass = A.ic_ensemble(nsys, 11, 3, 64)
    params(A, ctx, G=1.0, softening=0.0, c=1e4, pn_flags=A.NEWTON | A.PN1, precision=A.FP32)
    ctx.batch_set(off, pos, vel, mass)
    e0, L0 = ctx.batch_energy()
    wpb, spl = ctx.batch_autotune(A.YOSHIDA4, dt)
    ctx.batch_step(A.YOSHIDA4, dt, 20, steps_per_launch=spl)  # warm + capture of a short graph
    ctx.batch_set(off, pos, vel, mass)
    ctx.batch_step(A.YOSHIDA4, dt, 0)
    ctx.sync()
    t0 = time.perf_counter()
    ctx.batch_step(A.YOSHIDA4, dt, steps, steps_per_launch=spl, use_graph=True)
    ctx.sync()
    wall = time.perf_counter() - t0
    e1, L1 = ctx.batch_energy()
    de = np.abs(e1 - e0) / np.abs(e0)
    sizes = np.diff(off)
    pairs = float((sizes * (sizes - 1)).sum())
    ipairs, lane_slots, nwaves = ctx.batch_info()
    assert ipairs == pairs
    # two force evaluations per Yoshida substep with velocity-dependent forces
    inter = pairs * steps * 3 * 2
    achieved = BATCH_FLOP_PER_PAIR * inter / wall / 1e12
    log(f"[bench] batch {nsys} systems x {steps} steps: {wall:.3f}s, {inter / wall:.3e} pair interactions/s, "
        f"{achieved:.1f} FP64 TFLOP/s, lane use {pairs / lane_slots:.3f}, wpb {wpb} spl {spl}, "
        f"|dE/E| median {np.median(de):.2e}")
    # Newtonian comparison on the same ensemble: Wisdom-Holman (central mass =
    # body 0) against Yoshida-4 at the same coarse step (dt = 0.02, ~80 steps
    # per orbit of the innermost planets, a = 0.4), 500 steps each, where the
    # splitting's O(m/M dt^2) error is meant to pay off.
    wh = {}
    dtc, nc = 0.02, 500
    params(A, ctx, G=1.0, softening=0.0, pn_flags=A.NEWTON, precision=A.FP32)
    for name, integ in (("yoshida4", A.YOSHIDA4), ("wisdom_holman", A.WH_PLANETARY)):
        ctx.batch_set(off, pos, vel, mass)
        en0, _ = ctx.batch_energy()
        ctx.batch_step(integ, dtc, 0)
        ctx.sync()
        t0 = time.perf_counter()
        ctx.batch_step(integ, dtc, nc, steps_per_launch=spl, use_graph=True)
        ctx.sync()
        tw = time.perf_counter() - t0
        en1, _ = ctx.batch_energy()
        d = np.abs(en1 - en0) / np.abs(en0)
        wh[name] = {"ms_per_step": tw / nc * 1e3, "dE_rel_median": float(np.median(d)),
                    "dE_rel_p90": float(np.quantile(d, 0.9))}
    log(f"[bench] Newtonian ensemble, {nc} steps dt={dtc}: {wh}")
    # The same ensemble with the full N-body 1PN (EIH) law in exact FP64
    # mode, 1000 steps, against the pairwise 1PN law in exact mode: cost, and
    # how far the cross terms move the bodies at c = 1e4.
    ne = 1000
    eih = {}
    for name, flags in (("pairwise_1pn_exact", A.NEWTON | A.PN1), ("eih_exact", A.NEWTON | A.EIH)):
        params(A, ctx, G=1.0, softening=0.0, c=1e4, pn_flags=flags, precision=A.FP64)
        ctx.batch_set(off, pos, vel, mass)
        en0, _ = ctx.batch_energy()
        ctx.batch_step(A.YOSHIDA4, dt, 0)
        ctx.sync()
        t0 = time.perf_counter()
        ctx.batch_step(A.YOSHIDA4, dt, ne, steps_per_launch=spl, use_graph=True)
        ctx.sync()
        tw = time.perf_counter() - t0
        en1, _ = ctx.batch_energy()
        pe, _ = ctx.batch_get()
        d = np.abs(en1 - en0) / np.abs(en0)
        eih[name] = {"ms_per_step": tw / ne * 1e3, "pair_interactions_per_s": pairs * ne * 3 * 2 / tw,
                     "dE_rel_median": float(np.median(d)), "positions": pe}
    dx = np.abs(eih["eih_exact"].pop("positions") - eih["pairwise_1pn_exact"].pop("positions")).max(1)
    per_sys = np.maximum.reduceat(dx, off[:-1])  # largest |dx| of each system
    # an unsoftened close encounter (the maximum) amplifies any difference
    eih["per_system_max_abs_dx_eih_vs_pairwise"] = {"median": float(np.median(per_sys)),
                                                    "p99": float(np.quantile(per_sys, 0.99)),
                                                    "max": float(per_sys.max())}
    log(f"[bench] ensemble EIH vs pairwise 1PN, exact mode, {ne} steps: {eih}")
    return {"systems": nsys, "bodies": int(off[-1]), "steps": steps, "wall_s": wall,
            "system_steps_per_s": nsys * steps / wall, "pair_interactions_per_s": inter / wall,
            "roofline": {"bound": "valu", "kernel": "k_batch (fast FP64 mode, 1PN)", "achieved": achieved,
                         "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS,
                         "traffic": None, "flop_per_pair": BATCH_FLOP_PER_PAIR,
                         "note": "time = wall time of the graph replays (back to back on the device)"},
            "lane_use": {"pair_slots_used": pairs, "lane_pair_slots": lane_slots, "wavefronts": nwaves,
                         "frac": pairs / lane_slots},
            "waves_per_block": wpb, "steps_per_launch": spl,
            "energy_drift": {"median": float(np.median(de)), "p90": float(np.quantile(de, 0.9)),
                             "max": float(de.max())},
            "L_drift_max": float((np.abs(L1 - L0).max(1) / np.abs(L0).max(1)).max()),
            "newtonian_wh_vs_yoshida_dt0.02_500_steps": wh,
            "eih_vs_pairwise_1pn_exact_1000_steps": eih}


def bench_ic(A, n=1 << 24, host_n=1 << 20, reps=3):
    """Seeded Plummer sphere generated on the device by the cluster queue
    (fresh context per repetition: allocation + generation + recentring),
    against the host generator of the same law on a smaller sample."""
    times = []
    for r in range(reps + 1):
        with A.Context(0) as c:
            c.set_params(G=1.0)
            c.sync()
            t0 = time.perf_counter()
            c.cluster_add(A.CLUSTER_PLUMMER, n, 1000 + r, size=(1.0, 10.0, 0.0), mass=1.0)
            c.cluster_load()
            times.append(time.perf_counter() - t0)
    dev_s = min(times[1:])
    t0 = time.perf_counter()
    A.ic_plummer(host_n, 1000)
    host_s = time.perf_counter() - t0
    return {"n": n, "device_load_ms": dev_s * 1e3, "device_bodies_per_s": n / dev_s,
            "host_n": host_n, "host_ms": host_s * 1e3, "host_bodies_per_s": host_n / host_s,
            "note": "device: cluster_add + cluster_load of a Plummer sphere (allocation, per-body draws, mass "
                    "and barycentre sums, recentring); host: aster_ic_plummer, one thread"}


def gpu_parity_eval(A, ctx, theta, reps=3):
    """The device's FP64 reference-MAC traversal (ASTER_BH_PARITY, bit-identical
    to the reference's Barnes_Hut) on the loaded state: ms per evaluation
    (tree build + traversal, HIP events) and interactions per evaluation."""
    params(A, ctx, G=1.0, softening=0.005 ** 2, theta=theta, precision=A.FP64)
    acc = ctx.forces(A.BH_PARITY)
    ctx.sync()
    ctx.bh_stats()
    ctx.mark(2)
    for _ in range(reps):
        ctx.forces(A.BH_PARITY, out=False)
    ctx.mark(3)
    ms = ctx.elapsed(2, 3) / reps
    inter, _ = ctx.bh_stats()
    return acc, ms, inter / reps


def cpu_baselines_bh(orc, A, ctx, threads, bh, target_s=12.0):
    """CPU baselines of the headline Barnes-Hut evaluation, on its state
    (2^20 bodies, theta = 0.5), on this host's cores:
    1. equal work: the oracle's restatement of the group traversal (the same
       accepted lists as the device, FP64 sums) over every body, tree build
       included; interactions/s in the headline's unit;
    2. the reference's Barnes_Hut (make_tree + get_node_body, its MAC) at the
       same theta, restated, beside the device's bit-identical FP64 parity
       traversal on the same input (same algorithm, precision and count);
    3. the reference's Barnes_Hut at the theta where its median force error
       reaches the headline's (searched with the device's bit-identical FP64
       parity path), on a bounded sample of bodies: bodies/s at equal
       accuracy."""
    pos, vel, mass = bh["state"]
    n, theta, ref = bh["n"], bh["theta"], bh["ref"]
    box = orc.box_auto(pos)
    params(A, ctx, G=1.0, softening=0.005 ** 2, theta=theta, precision=A.FP64)
    ctx.set_state(pos, vel, mass)
    t0 = time.perf_counter()
    _, inter_g, _ = orc.bh_group(pos, mass, box, theta, soft=0.005 ** 2, threads=threads)
    t_group = time.perf_counter() - t0
    log(f"[bench] CPU group traversal (equal work): {t_group:.2f}s, {inter_g / t_group:.3e} interactions/s")
    out = {"value": inter_g / t_group, "unit": "interactions/s", "cores": threads, "kind": "port",
           "sample": f"equal work: the oracle's group traversal (the device's accepted lists, FP64 sums) over all "
                     f"{n} bodies at theta={theta}, tree build included ({t_group:.2f} s, {inter_g} interactions)",
           "bodies_per_s": n / t_group, "cpu": cpu_model()}
    # 2. the reference's algorithm at the same theta, CPU and device
    tb, tt, inter_r = orc.bh_sample(pos, vel, mass, box, theta, 1, soft=0.005 ** 2, threads=threads)
    acc_p, ms_p, inter_p = gpu_parity_eval(A, ctx, theta)
    e = rel_err(acc_p, ref)
    same = {"theta": theta, "interactions_per_body": inter_r / n, "median_error": float(np.median(e)),
            "cpu_s_per_eval": tb + tt, "cpu_build_s": tb, "cpu_interactions_per_s": inter_r / (tb + tt),
            "cpu_bodies_per_s": n / (tb + tt), "gpu_ms_per_eval": ms_p,
            "gpu_interactions_per_s": inter_p / (ms_p * 1e-3), "gpu_bodies_per_s": n / (ms_p * 1e-3),
            "same_interaction_count": bool(inter_p == inter_r)}
    log(f"[bench] reference Barnes_Hut theta={theta}: {same}")
    # 3. equal accuracy: the reference MAC's theta for the headline's median error
    target = float(np.median(rel_err(bh["acc"], ref)))
    th, found = 1e-4, None
    for _ in range(8):
        acc_t, ms_t, inter_t = gpu_parity_eval(A, ctx, th, reps=1)
        med = float(np.median(rel_err(acc_t, ref)))
        log(f"[bench] reference MAC theta={th:.3g}: median error {med:.3e}, {inter_t / n:.0f} interactions per body, "
            f"device {ms_t:.1f} ms")
        if med <= target:
            found = (th, med, ms_t, inter_t)
            break
        th *= 0.5
    eq = None
    if found:
        th, med, ms_t, inter_t = found
        probe = max(1, n // 2048)
        _, tp, _ = orc.bh_sample(pos, vel, mass, box, th, probe, soft=0.005 ** 2, threads=threads)
        per_body = tp / ((n + probe - 1) // probe)
        stride = max(1, int(np.ceil(per_body * n / target_s)))
        tb, tt, inter_s = orc.bh_sample(pos, vel, mass, box, th, stride, soft=0.005 ** 2, threads=threads)
        sampled = (n + stride - 1) // stride
        full_s = tb + tt * n / sampled
        eq = {"theta": th, "median_error": med, "target_median_error": target, "interactions_per_body": inter_t / n,
              "cpu_sampled_bodies": sampled, "cpu_build_s": tb, "cpu_traversal_s": tt,
              "cpu_s_per_eval_extrapolated": full_s, "cpu_bodies_per_s": n / full_s,
              "cpu_interactions_per_s": inter_s / tt, "gpu_ms_per_eval": ms_t, "gpu_bodies_per_s": n / (ms_t * 1e-3)}
        log(f"[bench] reference Barnes_Hut at equal accuracy: {eq}")
    out["reference_barnes_hut"] = {"same_theta": same, "equal_accuracy": eq}
    # restore the headline parameters
    params(A, ctx, G=1.0, softening=0.005 ** 2, theta=theta, precision=A.FP32)
    return out


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
    # the GPU takes ~30 ms of sustained work (about 20 steps at N = 2^20) to reach its steady clock
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--method", default="BH_GROUP")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint/forward ratio")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched ensemble")
    ap.add_argument("--no-integrators", action="store_true", help="skip the integrator section")
    ap.add_argument("--no-ic", action="store_true", help="skip the initial-condition generation section")
    ap.add_argument("--only", choices=["bh", "allpairs", "adjoint", "batch", "integrators", "ic", "reference",
                                       "scaling"],
                    help="run one section only (profiling); no contract line")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the library has no multi-GPU path; run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)
    import aster_amd as A
    ctx = A.Context(0)
    method = getattr(A, args.method)
    if args.only:
        # One section, for profiling runs; prints that section's result only.
        if args.only == "bh":
            r = bench_bh(A, ctx, args.steps, args.warmup, method, reference=False)
            r = {k: v for k, v in r.items() if k not in ("ref", "acc", "state")}
        elif args.only == "allpairs":
            r = bench_allpairs(A, ctx, max(args.steps, 10))[0]
        elif args.only == "adjoint":
            r = bench_adjoint(A, ctx)
        elif args.only == "integrators":
            r = bench_integrators(A, ctx)
        elif args.only == "ic":
            r = bench_ic(A)
        elif args.only == "reference":
            r = bench_reference_case(A, ctx, cpu=not args.no_cpu)
        elif args.only == "scaling":
            r = bench_scaling(A, ctx, method)
        else:
            r = bench_batch(A, ctx)
        print(json.dumps({"section": args.only, "result": r}, default=float), flush=True)
        return
    bh = bench_bh(A, ctx, args.steps, args.warmup, method)
    ap_res, pos64k, vel64k, mass64k = bench_allpairs(A, ctx, max(args.steps, 10))
    integ = None if args.no_integrators else bench_integrators(A, ctx)
    adj = None if args.no_adjoint else bench_adjoint(A, ctx)
    batch = None if args.no_batch else bench_batch(A, ctx)
    ic = None if args.no_ic else bench_ic(A)
    refcase = bench_reference_case(A, ctx, cpu=not args.no_cpu)

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
        cpu = cpu_baselines_bh(orc, A, ctx, th, bh)
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
        "bodies_per_s": bh["n"] / (ms_per_step * 1e-3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 forces, f64 state",
        "data": "synthetic: seeded half Plummer sphere + half uniform cube [-4,4]^3, FP32 positions/masses",
        "config": {"workload": "LBVH Barnes-Hut N=1048576 theta=0.5 eps=0.005, one leapfrog step "
                               "(tree rebuild + traversal + kick/drift) per step, dt=5e-4",
                   "n": bh["n"], "theta": bh["theta"], "method": args.method, "dt": bh["dt"]},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP32_PEAK_TFLOPS, "traffic": traffic_per_launch("k_bh_group"),
                     "frac_of_nonpacked_ceiling": achieved / FP32_NONPACKED_TFLOPS,
                     "kernel": "traversal",
                     "kernel_ms": trav_ms, "flops": "20 per interaction + 8 per node visit"},
        "cpu_baseline": cpu,
        "interactions_per_step": inter_per_eval,
        "node_visits_per_step": visits_per_eval,
        "force_error_vs_fp64_allpairs": bh["errors"],
        "phases_ms": phases,
        "hbm_phases": {k: {"bytes_per_body": b,
                           "achieved_GBps": b * bh["n"] / (phases[k]["ms_per_launch"] * 1e-3) / 1e9,
                           "frac_of_8TBps": b * bh["n"] / (phases[k]["ms_per_launch"] * 1e-3) / 1e9 / HBM_PEAK_GBPS}
                       for k, b in PHASE_BYTES_PER_BODY.items() if k in phases},
        "event_ms_per_step": bh["event_ms"] / args.steps,
        "allpairs": ap_res,
        "integrators": integ,
        "adjoint_forward_ratio": None if adj is None else adj["ratio"],
        "adjoint": adj,
        "batch": batch,
        "ic_generation": ic,
        "reference_case": refcase,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
