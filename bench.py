#!/usr/bin/env python3
"""Body-steps/s of the MI355X Barnes-Hut step on a 16M-body Plummer sphere.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n N] [--theta T]

Workload: Plummer sphere, N = 16,777,216 bodies, equal masses 1/N, G = 1,
scale radius 1, softening 0.01, theta 0.5 (the reference's half-longest-side
opening criterion), dt 1e-3, merging off, bounds recomputed every step; the
plain method (no learned correction). Bodies are generated on the host by
gravity_amd.initial_conditions.plummer and uploaded once; the timed region
is K calls of bh_step on resident device buffers, bracketed by stream
synchronisation.

Printed on one JSON line: value = alive bodies x K / wall time of the K
steps; the roofline entry for the dominant kernel (the walk, VALU-bound:
20 FLOP per body-node interaction / its mean duration from HIP events on the
library's stream); p50/p99 force error of the walk against an FP64 direct
sum for 1,024 sampled bodies; and the CPU baseline: the oracle's C port of
the same step timed on this host over a bounded sample (full-size O(N)
stages, the walk for 8 evenly spaced runs of sorted bodies, scaled to N).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "body-steps/sec @16M bodies at equal p99 force err; plain vs learned-corrected BH"
FP32_PEAK_TFLOPS = 157.3
HBM_PEAK_GBS = 8000.0
FLOP_PER_INTERACTION = 20


def measured_traffic(kernel: str):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC summary
    (2.0 x FETCH_SIZE + 1.0 x WRITE_SIZE, calibration in the file), or None."""
    path = os.path.join(ROOT, "profiles", "k_walk_plummer16M_step_pmc.md")
    try:
        with open(path) as fh:
            for line in fh:
                parts = line.split()
                if len(parts) == 3 and parts[0] == "traffic_bytes_per_launch" and parts[1] == kernel:
                    return float(parts[2])
    except OSError:
        pass
    return None


def hbm_copy_gbps(nbytes: int = 1 << 30, reps: int = 10):
    """Device-to-device copy bandwidth through the HIP runtime (read + write
    bytes per second): the attainable HBM rate next to the 8 TB/s vendor peak.
    None if the runtime calls fail."""
    import ctypes as C
    try:
        hip = C.CDLL("libamdhip64.so")
    except OSError:
        return None
    P = C.c_void_p
    a, b = P(), P()
    e0, e1 = P(), P()
    ms = C.c_float()
    try:
        if hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) or hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)):
            return None
        if hip.hipEventCreate(C.byref(e0)) or hip.hipEventCreate(C.byref(e1)):
            return None
        for k in range(reps + 2):
            if k == 2:
                hip.hipEventRecord(e0, None)
            if hip.hipMemcpyAsync(b, a, C.c_size_t(nbytes), 3, None):   # hipMemcpyDeviceToDevice
                return None
        hip.hipEventRecord(e1, None)
        hip.hipEventSynchronize(e1)
        if hip.hipEventElapsedTime(C.byref(ms), e0, e1) or ms.value <= 0:
            return None
        return 2.0 * nbytes * reps / (ms.value * 1e-3) / 1e9
    finally:
        for ptr in (a, b):
            if ptr.value:
                hip.hipFree(ptr)
        for ev in (e0, e1):
            if ev.value:
                hip.hipEventDestroy(ev)


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def force_error(bh, n: int, softening: float, n_sample: int = 1024, seed: int = 7) -> dict:
    """p50/p99 of |a - a_fp64| / |a_fp64| over sampled bodies."""
    from gravity_amd import barnes_hut as B
    rng = np.random.default_rng(seed)
    bodies = bh.bodies(B.BODIES_IN)
    alive = np.nonzero(bodies[:, 3] != 0)[0]
    sample = np.sort(rng.choice(alive, size=min(n_sample, alive.size), replace=False)).astype(np.uint32)
    acc = bh.accelerations()[sample].astype(np.float64)
    ref = bh.direct_sum_f64(sample, softening)
    err = np.linalg.norm(acc - ref, axis=1) / np.linalg.norm(ref, axis=1)
    return {"p50": float(np.percentile(err, 50)), "p99": float(np.percentile(err, 99)),
            "max": float(err.max()), "sample": int(sample.size), "exact_force": "fp64 direct sum (HIP kernel)"}


def variant_time(bh, prm, steps: int = 3) -> float:
    """Mean ms per step of `steps` steps with parameters prm (HIP-event stage sum)."""
    from gravity_amd import barnes_hut as B
    p = B.Params.from_buffer_copy(prm)
    p.flags |= B.FLAG_TIMING
    bh.step(p)                          # warm the variant
    bh.sync()
    bh.stage_times()
    t0 = time.perf_counter()
    for _ in range(steps):
        bh.step(p)
    bh.sync()
    return (time.perf_counter() - t0) * 1e3 / steps


def learned_leg(bh, n: int, soft: float, theta_plain: float, theta_wide: float, seed: int) -> dict:
    """Train the far-field network on 2M pairs from two 1M-body snapshots
    (Plummer, two-disk galaxy) with FP64 labels, then compare on the bench
    workload: plain walks over a sweep of angles vs the corrected walk at
    theta_wide and 0.7 (and the same box-aware walk with no network), each
    with its p50/p99 force error against the FP64 direct sum."""
    from gravity_amd import barnes_hut as B
    from gravity_amd import initial_conditions as ic
    from gravity_amd import learned as Lr
    t0 = time.perf_counter()
    feats, targs = [], []
    for k, snap in enumerate((ic.plummer(1 << 20, seed + 11), ic.two_galaxy(1 << 19, seed + 12))):
        sb = B.BarnesHut(snap, ic.tight_bounds(snap), B.make_params(theta=theta_wide, softening=soft))
        sb.init()
        sb.build_current_tree()
        f, t = Lr.collect_training_pairs(sb, theta_wide, theta_plain, 3, soft, max_node_bodies=1 << 20,
                                         cap=1 << 20)
        feats.append(f)
        targs.append(t)
        sb.close()
        log(f"learned: snapshot {k}: {f.shape[0]} pairs")
    x = np.concatenate(feats)
    t = np.concatenate(targs)
    t_data = time.perf_counter() - t0
    mlp = Lr.Mlp(params=Lr.init_params(seed))
    loss0, _ = mlp.loss_grad(x[:65536], t[:65536])
    t1 = time.perf_counter()
    mlp.train(x, t, epochs=20, batch=4096, lr=2e-3, seed=seed)
    t_train = time.perf_counter() - t1
    loss1, _ = mlp.loss_grad(x[:65536], t[:65536])
    log(f"learned: trained on {x.shape[0]} pairs in {t_train:.1f} s, loss {loss0:.4g} -> {loss1:.4g}")
    zero = Lr.Mlp(params=np.zeros(Lr.num_params(), np.float32))
    rows = []

    def run(method, th, net):
        flags = 0 if net is None else B.FLAG_LEARNED
        Lr.attach(bh, net, theta_plain)
        prm = B.make_params(theta=th, softening=soft, flags=flags)
        ms = variant_time(bh, prm)
        bh.params = prm
        err = force_error(bh, n, soft, n_sample=1024)
        row = {"method": method, "theta": th, "ms_per_step": ms, "body_steps_per_s": n * 1e3 / ms,
               "p50": err["p50"], "p99": err["p99"]}
        if net is not None:
            row["theta_plain"] = theta_plain
            row["network_evals_per_body"] = bh.counters()["corrected"] / n
        rows.append(row)
        log(f"learned leg: {method} theta {th}: {ms:.1f} ms, p50 {err['p50']:.3g} p99 {err['p99']:.3g}")

    for th in (0.3, 0.4, theta_plain, 0.6, 0.7, 0.8, theta_wide):
        run("plain", th, None)
    for th in (theta_wide, 0.7):
        run("box walk, no network", th, zero)
        run("learned", th, mlp)
    bh.params = B.make_params(theta=theta_plain, softening=soft)
    Lr.attach(bh, None, theta_plain)
    # equal-p99 comparison: the fastest plain angle whose p99 is at most the corrected walk's
    best = {}
    for r in rows:
        if r["method"] != "learned":
            continue
        ok = [q for q in rows if q["method"] == "plain" and q["p99"] <= r["p99"]]
        if ok:
            q = max(ok, key=lambda q: q["body_steps_per_s"])
            best[str(r["theta"])] = {"plain_theta_at_equal_p99": q["theta"],
                                     "speedup_vs_plain_at_equal_p99": r["body_steps_per_s"] / q["body_steps_per_s"]}
    mlp.close()
    zero.close()
    return {"train_pairs": int(x.shape[0]), "data_s": t_data, "train_s": t_train, "loss_before": loss0,
            "loss_after": loss1, "variants": rows, "equal_p99": best}


def ensemble_leg(n_sims: int, steps: int, soft: float, theta: float, dt: float, seed: int) -> dict:
    """Batched vs sequential: n_sims independent Plummer spheres (N per
    simulation log-uniform in [4096, 524288], virial ratio in [0.3, 0.7]) for
    `steps` steps, once as one ensemble (autotuned, hipGraph replay) and once
    one simulation after another through the single-simulation path. Both are
    aggregate body-steps/s over the timed steps with inputs resident."""
    from gravity_amd import barnes_hut as B
    from gravity_amd import ensemble as E
    from gravity_amd import initial_conditions as ic
    sims = ic.ensemble(n_sims, seed)
    total = int(sum(s.shape[0] for s in sims))
    prm = B.make_params(theta=theta, dt=dt, softening=soft)
    log(f"ensemble: {n_sims} simulations, {total} bodies")
    ens = E.Ensemble(sims, prm)
    ens.step(1)
    ens.sync()
    t0 = time.perf_counter()
    cfg, timings = ens.autotune()
    t_tune = time.perf_counter() - t0
    log(f"ensemble: autotune {t_tune:.1f} s chose {cfg.as_dict()}")
    ens.step(1)
    ens.sync()
    t0 = time.perf_counter()
    ens.step(steps)
    ens.sync()
    t_batched = time.perf_counter() - t0
    c = ens.counters()
    ens.close()
    log(f"ensemble: batched {steps} steps in {t_batched:.3f} s")
    t_seq = 0.0
    errs = []
    for k, b in enumerate(sims):
        bh = B.BarnesHut(b, ic.tight_bounds(b), prm)
        bh.init()
        bh.step()
        bh.sync()
        t0 = time.perf_counter()
        for _ in range(steps):
            bh.step()
        bh.sync()
        t_seq += time.perf_counter() - t0
        if k < 4:
            errs.append(force_error(bh, b.shape[0], soft, n_sample=256))
        bh.close()
    log(f"ensemble: sequential {steps} steps of each simulation in {t_seq:.3f} s")
    batched = total * steps / t_batched
    sequential = total * steps / t_seq
    return {"n_sims": n_sims, "n_bodies": total, "steps": steps,
            "sizes": {"min": int(min(s.shape[0] for s in sims)), "max": int(max(s.shape[0] for s in sims))},
            "batched_body_steps_per_s": batched, "batched_ms_per_step": t_batched * 1e3 / steps,
            "sequential_body_steps_per_s": sequential, "sequential_ms_per_step": t_seq * 1e3 / steps,
            "speedup_batched_vs_sequential": batched / sequential,
            "autotuned_config": cfg.as_dict(), "autotune_s": t_tune, "autotune_ms": timings,
            "interactions_per_body": c["interactions"] / total, "tree_height_max": int(c["tree_height"]),
            "error_flags": int(c["error_flags"]),
            "force_error_first_4_sims": [{"p50": e["p50"], "p99": e["p99"]} for e in errs]}


def cpu_baseline(bodies: np.ndarray, bounds: np.ndarray, theta: float, dt: float, softening: float,
                 threads: int, walk_sample: int) -> dict:
    """The oracle's C port of one step on this host: the O(N) stages at full
    size, the walk for `walk_sample` consecutive sorted bodies, scaled to N."""
    from oracle import bh_oracle as O
    from gravity_amd.initial_conditions import g_one_units
    n = bodies.shape[0]
    u = g_one_units()
    G, blc = O.units(u.mass, u.density, u.len, u.time)
    prm = O.params(theta=theta, dt=dt, softening=softening)
    src = bodies.copy()
    dst = np.zeros_like(src)
    idx = np.arange(n, dtype=np.uint32)
    t0 = time.perf_counter()
    out, died = O.dead_partition(src, dst, idx, n)
    num = n - died
    b = O.aabb(src, out, num, bounds)
    codes = O.morton_encode(src, out, num, b)
    keys, sidx = O.sort(codes, out[:num])
    out[:num] = sidx
    leaf, internal = O.build(keys, num, n)
    O.fill(src, out, num, n, leaf, internal, u.mass, u.density, max_passes=0, with_colors=False)
    t_stages = time.perf_counter() - t0
    chunks = 8
    per = max(1, walk_sample // chunks)
    walked = 0
    inter = 0
    t1 = time.perf_counter()
    for k in range(chunks):
        g0 = (num * (2 * k + 1)) // (2 * chunks) - per // 2
        g0 = max(0, min(g0, num - per))
        r = O.force_range(src, out, num, n, leaf, internal, prm, G, blc, b, g0, g0 + per, dst=dst, stack_cap=256,
                          threads=threads)
        walked += per
        inter += r["stats"].interactions
    t_walk = time.perf_counter() - t1
    t_step = t_stages + t_walk * (num / walked)
    return {"value": num / t_step, "unit": "body-steps/s", "cores": threads, "kind": "port",
            "sample": (f"one step of the C oracle (oracle/bh_oracle.c) at N={n}: dead partition, bounds, Morton, "
                       f"sort, build and fill at full size ({t_stages:.1f} s), walk + integrate for {walked} "
                       f"sorted bodies in {chunks} evenly spaced runs of {per} ({t_walk:.1f} s, "
                       f"{inter / walked:.0f} interactions/body), walk time scaled by N/{walked}"),
            "seconds_measured": t_stages + t_walk}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", choices=("plummer", "cube_core_merging", "two_galaxy"), default="plummer",
                    help="plummer: the headline 16M Plummer sphere (default); cube_core_merging: uniform cube "
                         "plus Plummer core, log-normal masses, merging on, dt 5e-4 (default N 4,194,304); "
                         "two_galaxy: two disk+bulge galaxies 10 apart on a parabolic orbit (default N 33,554,432)")
    ap.add_argument("--n", type=int, default=None)
    ap.add_argument("--theta", type=float, default=0.5)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-error", action="store_true")
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--ensemble-sims", type=int, default=256)
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--cpu-walk-sample", type=int, default=1048576)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(a split of one simulation over several GPUs is designed but not built)", file=sys.stderr)
        return 2

    from gravity_amd import barnes_hut as B
    from gravity_amd import initial_conditions as ic

    theta, soft = args.theta, 0.01
    if args.workload == "plummer":
        n, dt = args.n or 16_777_216, 1e-3
        log(f"generating Plummer N={n}")
        bodies = ic.plummer(n, args.seed)
        prm = B.make_params(theta=theta, dt=dt, softening=soft)
        workload = {"workload": f"plummer_{n}_plain_theta{theta}", "merging": "off"}
    elif args.workload == "cube_core_merging":
        # a softened distance (softening = eps^2) accepts every node smaller
        # than ~eps before its leaves are reached, and merges are only found
        # at leaves: this workload uses eps^2 = 1e-6 so that overlaps merge
        n, dt, soft = args.n or 4_194_304, 5e-4, 1e-6
        log(f"generating cube + Plummer core N={n}")
        bodies = ic.cube_plus_core(n, args.seed)
        prm = B.make_params(theta=theta, dt=dt, softening=soft, merging=B.MERGING)
        workload = {"workload": f"cube_plus_core_{n}_merging_theta{theta}", "merging": "on",
                    "masses": "log-normal(0, 0.5), total 1", "density": 64.0}
    else:
        n, dt = args.n or 33_554_432, 1e-3
        log(f"generating two galaxies N={n}")
        bodies = ic.two_galaxy(n // 2, args.seed)
        n = bodies.shape[0]
        prm = B.make_params(theta=theta, dt=dt, softening=soft)
        workload = {"workload": f"two_galaxy_{n}_plain_theta{theta}", "merging": "off"}
    bounds = ic.tight_bounds(bodies)
    headline = args.workload == "plummer"
    log("creating context and uploading")
    bh = B.BarnesHut(bodies, bounds, prm)
    bh.init()
    for k in range(args.warmup):
        bh.step()
        bh.sync()
        log(f"warmup step {k + 1}/{args.warmup}")
    timed = B.Params.from_buffer_copy(prm)
    timed.flags = B.FLAG_TIMING
    bh.stage_times()                      # clear the event ring
    bh.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        bh.step(timed)
    bh.sync()
    wall = time.perf_counter() - t0
    stage = bh.stage_times()
    c = bh.counters()
    log(f"timed {args.steps} steps in {wall:.3f} s; counters {c}")
    num = int(c["num_bodies"])
    value = num * args.steps / wall
    walk_ms = stage["force"] / args.steps
    achieved = c["interactions"] * FLOP_PER_INTERACTION / (walk_ms * 1e-3) / 1e12
    result = {
        "metric": METRIC,
        "value": value,
        "unit": "body-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": dict(workload, n_bodies=n, theta=theta, dt=dt, softening=soft, bounds="dynamic",
                       method="plain Barnes-Hut", mac="half longest AABB side x inverse softened distance < theta"),
        "roofline": {"bound": "valu", "kernel": "k_walk_q", "achieved": achieved, "peak": FP32_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / FP32_PEAK_TFLOPS,
                     "traffic": measured_traffic("k_walk_q") if headline and n == 16_777_216 else None,
                     "traffic_source": "profiles/k_walk_plummer16M_step_pmc.md (rocprofv3 FETCH_SIZE/WRITE_SIZE)",
                     "flop_per_interaction": FLOP_PER_INTERACTION,
                     "interactions_per_step": int(c["interactions"]), "walk_ms": walk_ms,
                     "interactions_per_s": c["interactions"] / (walk_ms * 1e-3)},
        "stage_ms": {k: v / args.steps for k, v in stage.items()},
        "counters": {"opened": int(c["opened"]), "interactions_per_body": c["interactions"] / max(num, 1),
                     "tree_height": int(c["tree_height"]), "max_stack": int(c["max_stack"]),
                     "error_flags": int(c["error_flags"]), "alive_bodies": num,
                     "merges_last_step": int(c["merges_applied"])},
    }
    # the HBM-bound stages at their algorithmic bytes per body (one sorted
    # pass over 24 B per key, 8 passes plus 8 histogram reads; the tree's 172
    # B; the encode's 24 B; census + bounds 16 B) against the measured copy
    # rate and the vendor peak
    copy = hbm_copy_gbps()
    sm = result["stage_ms"]
    alg = {"sort": 256, "tree": 172, "morton": 24, "dead": 16}
    result["hbm_stages"] = {
        "copy_GBps": copy, "peak_GBps": HBM_PEAK_GBS,
        **{k: {"ms": sm[k], "alg_bytes_per_body": b, "alg_GBps": b * num / (sm[k] * 1e-3) / 1e9,
               "frac_of_peak": b * num / (sm[k] * 1e-3) / 1e9 / HBM_PEAK_GBS}
           for k, b in alg.items() if sm.get(k, 0.0) > 0.0}}
    if not args.no_error:
        log("force error against the FP64 direct sum")
        result["force_error"] = force_error(bh, n, soft)
    if not args.no_learned and headline:
        log("learned-correction leg")
        result["learned"] = learned_leg(bh, n, soft, theta, 0.9, args.seed)
    bh.close()
    if not args.no_ensemble and headline:
        log("ensemble leg")
        result["ensemble"] = ensemble_leg(args.ensemble_sims, args.steps, soft, theta, dt, args.seed)
    if not args.no_cpu_baseline and headline:
        log("CPU baseline (oracle port)")
        result["cpu_baseline"] = cpu_baseline(bodies, bounds, theta, dt, soft, args.cpu_threads,
                                              args.cpu_walk_sample)
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
