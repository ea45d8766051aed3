#!/usr/bin/env python3
"""Body-steps/s of the MI355X Barnes-Hut step at 16M bodies, plain against
learned-corrected at equal p99 force error.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload W]

Headline workload: Plummer sphere, N = 16,777,216, equal masses 1/N, G = 1,
scale radius 1, softening 0.01, dt 1e-3, merging off, bounds recomputed every
step. Force error: |a - a_fp64| / |a_fp64| over 16,384 sampled bodies against
an FP64 direct sum (the GPU's FP64 kernel), p50 and p99.

1. The reference's configuration, the plain step at theta 0.5, is timed.
2. The correction's network is trained on the GPU on (body, node) pairs with
   FP64 labels from a 1M-body Plummer and a 1M-body two-galaxy snapshot.
3. On one frozen state, sweeps of the opening parameter: the plain walk's
   angle; the corrected walk's (exact quadrupole, no network) geometric
   angle; the relative test's tolerance for the exact quadrupole and for the
   learned-corrected walk (quadrupole + the network's residual); each with
   its p50/p99 and its step time; the sweeps stop at the metric's error
   level, p99 <= 1e-3.
4. The timed loop: K steps of the learned-corrected method at the largest
   tolerance whose p99 is <= 1e-3, bracketed by stream synchronisation,
   inputs resident in HBM. value = alive bodies x K / wall time of the K
   steps. equal_p99 compares it with the fastest plain angle at the same
   level.
5. The other workloads (the reference's CPU-runnable Plummer 65,536 x 10
   steps, cube + core merging 4,194,304 x 20 steps, two galaxies 33,554,432 x
   5 steps), the 256-simulation ensemble (plain at theta 0.5; the
   learned-corrected walk at theta 0.9 and with the relative test at the
   headline's tolerance; batched against sequential), and the CPU baseline
   (the oracle's C port of the plain step on this host over a bounded
   sample).

The roofline entry is the dominant kernel of the timed step, the corrected
walk (FP32 VALU: 20 FLOP per body-node interaction plus 33 per quadrupole
term, over its mean duration from HIP events on the library's stream; the
walk is bound by instruction issue and record latency, its HBM traffic is
taken from the committed rocprofv3 summary).
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
# the quadrupole term of an accepted internal node beyond its monopole: G T r (15
# with a fused multiply-add as 2), r.GT.r (5), 1/s^5 (1), the coefficient
# (7.5/s^2, one fma, the -3/s^5 factor: 4), its fma with the monopole's G M/s^3
# (1 more than a product) and the three fused updates of the sum (6)
FLOP_PER_QUAD_TERM = 33
P99_TARGET = 1e-3
N_SAMPLE = 16384


def measured_traffic(kernel: str, profile: str):
    """HBM bytes per launch of `kernel` from a committed rocprofv3 PMC summary
    (2.0 x FETCH_SIZE + 1.0 x WRITE_SIZE, calibration in the file), or None."""
    path = os.path.join(ROOT, "profiles", profile)
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


def sample_ids(bodies_or_n, n_sample: int = N_SAMPLE, seed: int = 7) -> np.ndarray:
    """Sorted random ids of alive bodies."""
    rng = np.random.default_rng(seed)
    if isinstance(bodies_or_n, np.ndarray):
        alive = np.nonzero(bodies_or_n[:, 3] != 0)[0]
    else:
        alive = np.arange(bodies_or_n)
    return np.sort(rng.choice(alive, size=min(n_sample, alive.size), replace=False)).astype(np.uint32)


def error_stats(acc: np.ndarray, exact: np.ndarray) -> dict:
    e = np.linalg.norm(acc.astype(np.float64) - exact, axis=1) / np.linalg.norm(exact, axis=1)
    return {"p50": float(np.percentile(e, 50)), "p99": float(np.percentile(e, 99)), "max": float(e.max()),
            "sample": int(e.size)}


def force_error(bh, softening: float, n_sample: int = N_SAMPLE, seed: int = 7, prm=None) -> dict:
    """p50/p99 of |a - a_fp64| / |a_fp64| over sampled alive bodies of bh's current state."""
    from gravity_amd import barnes_hut as B
    sample = sample_ids(bh.bodies(B.BODIES_IN), n_sample, seed)
    exact = bh.direct_sum_f64(sample, softening)
    out = error_stats(bh.accelerations(prm)[sample], exact)
    out["exact_force"] = "fp64 direct sum (HIP kernel)"
    return out


def variant_time(bh, prm, steps: int = 2) -> float:
    """Mean ms per step of `steps` steps with parameters prm (after one untimed step)."""
    from gravity_amd import barnes_hut as B
    p = B.Params.from_buffer_copy(prm)
    bh.step(p)
    bh.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        bh.step(p)
    bh.sync()
    return (time.perf_counter() - t0) * 1e3 / steps


def timed_loop(bh, prm, steps: int, warmup: int, stage_steps: int = 5) -> dict:
    """W untimed steps, then exactly K steps bracketed by stream
    synchronisation (the measurement), then `stage_steps` more with
    BH_FLAG_TIMING for the per-stage times (HIP events between the stages
    on the library's stream; they cost a little, so they stay out of the
    measured steps)."""
    from gravity_amd import barnes_hut as B
    for k in range(warmup):
        bh.step(prm)
        bh.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        bh.step(prm)
    bh.sync()
    wall = time.perf_counter() - t0
    c = bh.counters()
    num = int(c["num_bodies"])
    out = {"wall_s": wall, "ms_per_step": wall * 1e3 / steps, "body_steps_per_s": num * steps / wall,
           "counters": c, "num": num}
    if stage_steps:
        timed = B.Params.from_buffer_copy(prm)
        timed.flags = (timed.flags | B.FLAG_TIMING) & ~B.FLAG_GRAPH
        bh.stage_times()                      # clear the event ring
        bh.sync()
        t1 = time.perf_counter()
        for _ in range(stage_steps):
            bh.step(timed)
        bh.sync()
        wall_timed = time.perf_counter() - t1
        stage = bh.stage_times()
        out["stage_ms"] = {k: v / stage_steps for k, v in stage.items()}
        out["ms_per_step_with_stage_events"] = wall_timed * 1e3 / stage_steps
    return out


def train_network(soft: float, theta: float, tolerance: float, seed: int) -> tuple:
    """The correction's network trained on the GPU: (body, accepted node)
    pairs of the corrected walk's geometric test at `theta` on a 1M-body
    Plummer and a 1M-body two-galaxy snapshot, FP64 labels (the residual after
    monopole and exact quadrupole); its output scale calibrated with the
    relative test at `tolerance` on a third snapshot, a 16M-body Plummer of
    its own seed (the headline's workload, not its state)."""
    from gravity_amd import barnes_hut as B
    from gravity_amd import initial_conditions as ic
    from gravity_amd import learned as Lr
    t0 = time.perf_counter()
    feats, targs = [], []
    for k, snap in enumerate((ic.plummer(1 << 20, seed + 11), ic.two_galaxy(1 << 19, seed + 12))):
        sb = B.BarnesHut(snap, ic.tight_bounds(snap), B.make_params(theta=theta, softening=soft))
        sb.init()
        sb.build_current_tree()
        f, t = Lr.collect_training_pairs(sb, theta, 5, soft, max_node_bodies=1 << 20, cap=1 << 21)
        feats.append(f)
        targs.append(t)
        sb.close()
        log(f"learned: snapshot {k}: {f.shape[0]} pairs")
    x = np.concatenate(feats)
    t = np.concatenate(targs)
    t_data = time.perf_counter() - t0
    mlp = Lr.Mlp(params=Lr.init_params(seed))
    rows = np.random.default_rng(seed).permutation(x.shape[0])
    hold, train = rows[:65536], rows[65536:]
    loss0, _ = mlp.loss_grad(x[hold], t[hold])
    t1 = time.perf_counter()
    mlp.train(x[train], t[train], epochs=8, batch=4096, lr=1e-3, seed=seed)
    t_train = time.perf_counter() - t1
    loss1, _ = mlp.loss_grad(x[hold], t[hold])
    log(f"learned: trained on {train.size} pairs in {t_train:.1f} s, held-out loss {loss0:.4g} -> {loss1:.4g}")
    t2 = time.perf_counter()
    alpha, cal = Lr.calibrate(mlp, ic.plummer(1 << 24, seed + 13), 0.5, soft, n_sample=N_SAMPLE,
                              force_tolerance=tolerance)
    log(f"learned: output scale {alpha} on a validation snapshot ({time.perf_counter() - t2:.1f} s): {cal}")
    return mlp, {"train_pairs": int(train.size), "held_out_pairs": int(hold.size), "data_s": t_data,
                 "train_s": t_train, "held_out_loss_zero_network": loss0, "held_out_loss_trained": loss1,
                 "train_theta": theta, "calibration_force_tolerance": tolerance, "output_scale": alpha,
                 "calibration": [{"alpha": a, "p50": p50, "p99": p99} for a, p50, p99 in cal],
                 "calibration_snapshot": f"Plummer 16,777,216 (a seed of its own), {N_SAMPLE} sampled bodies"}


def sweep(bh, saved, sample, exact, soft: float, method: str, values, flags: int, relative: bool = False) -> list:
    """p50/p99 on the frozen state `saved` and step time of each value of the
    opening parameter (theta, or with `relative` the relative test's
    tolerance at theta 0.5 for its first walk), in order, until the first
    value whose p99 is at the target."""
    from gravity_amd import barnes_hut as B
    rows = []
    for v in values:
        if relative:
            prm = B.make_params(theta=0.5, softening=soft, flags=flags | B.FLAG_RELATIVE_MAC, force_tolerance=v)
        else:
            prm = B.make_params(theta=v, softening=soft, flags=flags)
        bh.upload(B.BODIES_IN, saved)
        err = error_stats(bh.accelerations(prm)[sample], exact)
        ms = variant_time(bh, prm)
        c = bh.counters()
        row = {"method": method, "mac": "relative" if relative else "geometric",
               ("force_tolerance" if relative else "theta"): v, "ms_per_step": ms,
               "body_steps_per_s": bh.n * 1e3 / ms, "p50": err["p50"], "p99": err["p99"],
               "interactions_per_body": c["interactions"] / bh.n}
        rows.append(row)
        log(f"sweep {method} {row['mac']} {v}: {ms:.1f} ms, p50 {err['p50']:.3g} p99 {err['p99']:.3g}, "
            f"{row['interactions_per_body']:.0f} interactions/body")
        if err["p99"] <= P99_TARGET:
            break
    return rows


def fastest_at_target(rows):
    ok = [r for r in rows if r["p99"] <= P99_TARGET]
    return max(ok, key=lambda r: r["body_steps_per_s"]) if ok else None


def other_workload(name: str, steps: int, seed: int) -> dict:
    """One of the other configurations: timed steps and the force error after them."""
    from gravity_amd import barnes_hut as B
    from gravity_amd import initial_conditions as ic
    soft = 0.01
    if name == "plummer_65536":
        bodies, dt = ic.plummer(65536, seed), 1e-3
        prm = B.make_params(theta=0.5, dt=dt, softening=soft)
        desc = {"n_bodies": 65536, "theta": 0.5, "dt": dt, "softening": soft, "merging": "off", "steps": steps}
    elif name == "cube_core_merging":
        # a softened distance (softening = eps^2) accepts every node smaller
        # than ~eps before its leaves are reached, and merges are only found
        # at leaves: this workload uses eps^2 = 1e-6 so that overlaps merge
        bodies, dt, soft = ic.cube_plus_core(4_194_304, seed), 5e-4, 1e-6
        prm = B.make_params(theta=0.5, dt=dt, softening=soft, merging=B.MERGING)
        desc = {"n_bodies": 4_194_304, "theta": 0.5, "dt": dt, "softening": soft, "merging": "on",
                "masses": "log-normal(0, 0.5), total 1", "density": 64.0, "steps": steps}
    else:
        bodies, dt = ic.two_galaxy(16_777_216, seed), 1e-3
        prm = B.make_params(theta=0.5, dt=dt, softening=soft)
        desc = {"n_bodies": int(bodies.shape[0]), "theta": 0.5, "dt": dt, "softening": soft, "merging": "off",
                "steps": steps}
    log(f"{name}: {bodies.shape[0]} bodies, {steps} steps")
    bounds = ic.tight_bounds(bodies)
    bh = B.BarnesHut(bodies, bounds, prm)
    bh.init()
    r = timed_loop(bh, prm, steps, 3)
    err = force_error(bh, soft)
    c = r["counters"]
    out = dict(desc, ms_per_step=r["ms_per_step"], body_steps_per_s=r["body_steps_per_s"],
               alive_bodies=r["num"], stage_ms=r["stage_ms"], p50=err["p50"], p99=err["p99"],
               error_sample=err["sample"], error_flags=int(c["error_flags"]), tree_height=int(c["tree_height"]),
               merges_last_step=int(c["merges_applied"]))
    bh.close()
    # the same steps from the same bodies, replayed as captured hipGraphs
    # (BH_FLAG_GRAPH: one launch per step instead of ~30)
    g = B.Params.from_buffer_copy(prm)
    g.flags |= B.FLAG_GRAPH
    bh = B.BarnesHut(bodies, bounds, prm)
    del bodies
    bh.init()
    rg = timed_loop(bh, g, steps, 3, stage_steps=0)   # (one direct step, then the two graphs captured)
    out["graph"] = {"ms_per_step": rg["ms_per_step"], "body_steps_per_s": rg["body_steps_per_s"],
                    "alive_bodies": rg["num"]}
    bh.close()
    log(f"{name}: {r['ms_per_step']:.2f} ms/step, {r['body_steps_per_s']:.3g} body-steps/s, "
        f"p50 {err['p50']:.3g} p99 {err['p99']:.3g}; as graphs {rg['ms_per_step']:.2f} ms/step")
    return out


def _ensemble_method(sims, prm, mlp, steps: int, soft: float) -> dict:
    """One method over the ensemble: batched (autotuned, hipGraph replay) and
    sequential (one single-simulation run after another), aggregate
    body-steps/s over `steps` timed steps each with inputs resident; the force
    error of the first 4 simulations from their single-simulation runs, whose
    walk takes the same nodes as the ensemble's."""
    from gravity_amd import barnes_hut as B
    from gravity_amd import ensemble as E
    from gravity_amd import initial_conditions as ic
    from gravity_amd import learned as Lr
    total = int(sum(s.shape[0] for s in sims))
    ens = E.Ensemble(sims, prm)
    if mlp is not None:
        ens.attach(mlp)
    ens.step(1)
    ens.sync()
    t0 = time.perf_counter()
    cfg, timings = ens.autotune()
    t_tune = time.perf_counter() - t0
    ens.step(1)
    ens.sync()
    t0 = time.perf_counter()
    ens.step(steps)
    ens.sync()
    t_batched = time.perf_counter() - t0
    c = ens.counters()
    ens.close()
    t_seq = 0.0
    errs = []
    for k, b in enumerate(sims):
        bh = B.BarnesHut(b, ic.tight_bounds(b), prm)
        bh.init()
        if mlp is not None:
            Lr.attach(bh, mlp)
        bh.step()
        bh.sync()
        t0 = time.perf_counter()
        for _ in range(steps):
            bh.step()
        bh.sync()
        t_seq += time.perf_counter() - t0
        if k < 4:
            errs.append(force_error(bh, soft, n_sample=1024, prm=prm))
        if mlp is not None:
            Lr.attach(bh, None)
        bh.close()
    batched = total * steps / t_batched
    sequential = total * steps / t_seq
    return {"batched_body_steps_per_s": batched, "batched_ms_per_step": t_batched * 1e3 / steps,
            "sequential_body_steps_per_s": sequential, "sequential_ms_per_step": t_seq * 1e3 / steps,
            "speedup_batched_vs_sequential": batched / sequential,
            "autotuned_config": cfg.as_dict(), "autotune_s": t_tune, "autotune_ms": timings,
            "interactions_per_body": c["interactions"] / total, "quad_terms_per_body": c["quad_terms"] / total,
            "network_evaluations_per_step": int(c["corrected"]), "tree_height_max": int(c["tree_height"]),
            "error_flags": int(c["error_flags"]),
            "force_error_first_4_sims": [{"p50": e["p50"], "p99": e["p99"]} for e in errs]}


def ensemble_leg(n_sims: int, steps: int, soft: float, theta: float, dt: float, seed: int, mlp=None,
                 tolerance: float | None = None) -> dict:
    """Batched vs sequential: n_sims independent Plummer spheres (N per
    simulation log-uniform in [4096, 524288], virial ratio in [0.3, 0.7]) for
    `steps` steps of the plain method at `theta`; with a network, also the
    learned-corrected method with the geometric test at theta 0.9 and with
    the relative test at `tolerance` (its first step geometric at 0.5)."""
    from gravity_amd import barnes_hut as B
    from gravity_amd import initial_conditions as ic
    sims = ic.ensemble(n_sims, seed)
    total = int(sum(s.shape[0] for s in sims))
    log(f"ensemble: {n_sims} simulations, {total} bodies")
    prm = B.make_params(theta=theta, dt=dt, softening=soft)
    out = {"n_sims": n_sims, "n_bodies": total, "steps": steps,
           "sizes": {"min": int(min(s.shape[0] for s in sims)), "max": int(max(s.shape[0] for s in sims))},
           "method": f"plain, theta {theta}"}
    r = _ensemble_method(sims, prm, None, steps, soft)
    log(f"ensemble plain: batched {r['batched_ms_per_step']:.2f} ms/step, sequential "
        f"{r['sequential_ms_per_step']:.2f} ms/step, chose {r['autotuned_config']}")
    out.update(r)
    if mlp is None:
        return out
    learned = []
    variants = [("learned-corrected, geometric test, theta 0.9",
                 B.make_params(theta=0.9, dt=dt, softening=soft, flags=B.FLAG_LEARNED))]
    if tolerance is not None:
        variants.append((f"learned-corrected, relative test, tolerance {tolerance}",
                         B.make_params(theta=0.5, dt=dt, softening=soft,
                                       flags=B.FLAG_LEARNED | B.FLAG_RELATIVE_MAC, force_tolerance=tolerance)))
    for name, p in variants:
        r = _ensemble_method(sims, p, mlp, steps, soft)
        e = r["force_error_first_4_sims"]
        log(f"ensemble {name}: batched {r['batched_ms_per_step']:.2f} ms/step, sequential "
            f"{r['sequential_ms_per_step']:.2f} ms/step, p99 of the first sims {[round(x['p99'], 5) for x in e]}")
        learned.append(dict(r, method=name))
    out["learned"] = learned
    return out


def render_feed_leg(bh, prm) -> dict:
    """Per-frame host feed for a renderer on the headline state: one step
    with colours, then the compacted feed (internal nodes of depth 1 .. 8,
    the node-glow default nodeGlowMaxDepth, and the alive bodies) against
    downloading the std430 InternalNodes and LeafNodes bindings and the
    bodies; best of 3 wall times, host buffers pageable."""
    from gravity_amd import barnes_hut as B
    p = B.Params.from_buffer_copy(prm)
    p.flags |= B.FLAG_COLORS
    bh.step(p)
    bh.sync()

    def best(fn):
        t = []
        for _ in range(3):
            t0 = time.perf_counter()
            r = fn()
            t.append(time.perf_counter() - t0)
        return min(t), r

    t_full, full = best(lambda: (bh.nodes(B.INTERNAL_NODES), bh.nodes(B.LEAF_NODES), bh.bodies(B.BODIES_IN)))
    full_bytes = sum(a.nbytes for a in full)
    del full
    t_nodes, nodes = best(lambda: bh.render_nodes(1, 8))
    t_bodies, bodies = best(lambda: bh.render_bodies())
    feed_bytes = nodes.nbytes + bodies.nbytes
    pn = B.PinnedBuffer((bh.n,), B.RENDER_NODE)
    pb = B.PinnedBuffer((bh.n,), B.RENDER_BODY)
    t_pn, _ = best(lambda: bh.render_nodes(1, 8, out=pn.array))
    t_pb, _ = best(lambda: bh.render_bodies(out=pb.array))
    pn.close()
    pb.close()
    out = {"std430_ms": t_full * 1e3, "std430_bytes": int(full_bytes),
           "feed_ms": (t_nodes + t_bodies) * 1e3, "feed_bytes": int(feed_bytes),
           "feed_pinned_ms": (t_pn + t_pb) * 1e3,
           "feed_nodes": int(nodes.shape[0]), "feed_bodies": int(bodies.shape[0]),
           "feed_nodes_ms": t_nodes * 1e3, "feed_bodies_ms": t_bodies * 1e3,
           "depth_range": [1, 8], "host_memory": "pageable (numpy); feed_pinned_ms: page-locked (bh_host_alloc)"}
    log(f"render feed: {out['feed_ms']:.1f} ms ({out['feed_pinned_ms']:.1f} ms into page-locked memory) for "
        f"{feed_bytes / 1e9:.2f} GB ({nodes.shape[0]} nodes, {bodies.shape[0]} bodies) vs std430 bindings "
        f"{out['std430_ms']:.1f} ms for {full_bytes / 1e9:.2f} GB")
    return out


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
    ap.add_argument("--workload", choices=("plummer", "plummer_65536", "cube_core_merging", "two_galaxy"),
                    default="plummer",
                    help="plummer: the headline 16M comparison (default, with every other leg); the others run "
                         "one of the other configurations alone")
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--no-configs", action="store_true")
    ap.add_argument("--ensemble-sims", type=int, default=256)
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--cpu-walk-sample", type=int, default=1048576)
    ap.add_argument("--profile-step", type=float, default=None, metavar="ALPHA",
                    help="only the headline's timed loop: the learned-corrected step with the relative opening "
                         "test at tolerance ALPHA and an untrained (zero-output) network, which runs the same "
                         "kernels; for rocprofv3 profiles")
    ap.add_argument("--plain-step", type=int, default=None, metavar="N",
                    help="only the plain theta 0.5 step's timed loop on an N-body Plummer sphere; for rocprofv3 "
                         "profiles")
    ap.add_argument("--ensemble-step", action="store_true",
                    help="only the ensemble's plain leg (batched and sequential); for rocprofv3 profiles")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(a split of one simulation over several GPUs is designed but not built)", file=sys.stderr)
        return 2

    from gravity_amd import barnes_hut as B
    from gravity_amd import initial_conditions as ic
    from gravity_amd import learned as Lr

    if args.workload != "plummer":
        steps = {"plummer_65536": 10, "cube_core_merging": 20, "two_galaxy": 5}[args.workload]
        print(json.dumps(other_workload(args.workload, steps, args.seed)), flush=True)
        return 0

    n, soft, dt = 16_777_216, 0.01, 1e-3
    if args.ensemble_step:
        print(json.dumps({"ensemble_step": ensemble_leg(args.ensemble_sims, args.steps, soft, 0.5, dt, args.seed)}),
              flush=True)
        return 0
    if args.plain_step is not None:
        n = args.plain_step
    log(f"generating Plummer N={n}")
    bodies = ic.plummer(n, args.seed)
    bounds = ic.tight_bounds(bodies)
    plain_ref = B.make_params(theta=0.5, dt=dt, softening=soft)
    bh = B.BarnesHut(bodies, bounds, plain_ref)
    bh.init()
    if args.profile_step is not None:
        mlp = Lr.Mlp(params=Lr.init_params(args.seed))      # zero output layer: no residual
        Lr.attach(bh, mlp)
        prm = B.make_params(theta=0.5, dt=dt, softening=soft, flags=B.FLAG_LEARNED | B.FLAG_RELATIVE_MAC,
                            force_tolerance=args.profile_step)
        run = timed_loop(bh, prm, args.steps, args.warmup)
        c = run["counters"]
        walk_ms = run["stage_ms"]["force"]
        flops = c["interactions"] * FLOP_PER_INTERACTION + c["quad_terms"] * FLOP_PER_QUAD_TERM
        print(json.dumps({"profile_step": args.profile_step, "ms_per_step": run["ms_per_step"],
                          "body_steps_per_s": run["body_steps_per_s"], "stage_ms": run["stage_ms"],
                          "walk_ms": walk_ms, "walk_tflops": flops / (walk_ms * 1e-3) / 1e12,
                          "interactions": int(c["interactions"]), "quad_terms": int(c["quad_terms"])}),
              flush=True)
        Lr.attach(bh, None)
        bh.close()
        mlp.close()
        return 0
    if args.plain_step is not None:
        run = timed_loop(bh, plain_ref, args.steps, args.warmup)
        c = run["counters"]
        walk_ms = run["stage_ms"]["force"]
        print(json.dumps({"plain_step": n, "ms_per_step": run["ms_per_step"],
                          "body_steps_per_s": run["body_steps_per_s"], "stage_ms": run["stage_ms"],
                          "walk_ms": walk_ms,
                          "walk_tflops": c["interactions"] * FLOP_PER_INTERACTION / (walk_ms * 1e-3) / 1e12,
                          "interactions_per_body": c["interactions"] / max(run["num"], 1),
                          "tree_height": int(c["tree_height"])}), flush=True)
        bh.close()
        return 0

    # 1. the reference's configuration: plain, theta 0.5
    ref = timed_loop(bh, plain_ref, args.steps, args.warmup)
    log(f"plain theta 0.5: {ref['ms_per_step']:.2f} ms/step, {ref['body_steps_per_s']:.4g} body-steps/s")

    # 2. the network
    mlp, train_info = train_network(soft, 0.6, 1.8e-3, args.seed)
    Lr.attach(bh, mlp)

    # 3. sweeps on one frozen state against one FP64 judge
    saved = np.ascontiguousarray(bh.bodies(B.BODIES_IN))
    sample = sample_ids(saved)
    exact = bh.direct_sum_f64(sample, soft)
    plain_rows = sweep(bh, saved, sample, exact, soft, "plain", (0.5, 0.3, 0.2, 0.15, 0.12, 0.1, 0.08, 0.06), 0)
    geo_rows = sweep(bh, saved, sample, exact, soft, "quadrupole", (0.8, 0.7, 0.6, 0.5, 0.45, 0.4, 0.35),
                     B.FLAG_QUADRUPOLE)
    tolerances = (0.004, 0.003, 0.0025, 0.002, 0.0018, 0.0016, 0.0015, 0.0012, 0.001, 0.0008, 0.0006)
    quad_rows = sweep(bh, saved, sample, exact, soft, "quadrupole", tolerances, B.FLAG_QUADRUPOLE, relative=True)
    learned_rows = sweep(bh, saved, sample, exact, soft, "learned", tolerances, B.FLAG_LEARNED, relative=True)
    plain_best = fastest_at_target(plain_rows)
    quad_best = fastest_at_target(quad_rows)
    learned_best = fastest_at_target(learned_rows)

    # 4. the timed loop: learned-corrected at its largest tolerance within the target
    pick = learned_best or learned_rows[-1]
    headline = B.make_params(theta=0.5, dt=dt, softening=soft, flags=B.FLAG_LEARNED | B.FLAG_RELATIVE_MAC,
                             force_tolerance=pick["force_tolerance"])
    bh.upload(B.BODIES_IN, saved)
    run = timed_loop(bh, headline, args.steps, args.warmup)
    c = run["counters"]
    num = run["num"]
    log(f"timed {args.steps} learned-corrected steps (tolerance {pick['force_tolerance']}): "
        f"{run['ms_per_step']:.2f} ms/step; "
        f"counters {c}")
    walk_ms = run["stage_ms"]["force"]
    flops = c["interactions"] * FLOP_PER_INTERACTION + c["quad_terms"] * FLOP_PER_QUAD_TERM
    achieved = flops / (walk_ms * 1e-3) / 1e12
    result = {
        "metric": METRIC,
        "value": run["body_steps_per_s"],
        "unit": "body-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": run["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"plummer_{n}_learned_corrected_at_p99_1e-3", "n_bodies": n,
                   "force_tolerance": pick["force_tolerance"], "dt": dt, "softening": soft, "merging": "off",
                   "bounds": "dynamic",
                   "method": "learned-corrected Barnes-Hut: monopole + softened quadrupole of the exact second "
                             "moment plus the network's residual, one network evaluation per node and step",
                   "mac": "relative: G M b^3 / s^5 < alpha |a_prev| and s > b (b: centre of mass to the farthest "
                          "box corner), decided for 64 bodies together (a node taken only if all accept it)",
                   "p50": pick["p50"], "p99": pick["p99"], "error_sample": int(sample.size),
                   "p99_target": P99_TARGET, "p99_met": bool(pick["p99"] <= P99_TARGET)},
        "roofline": {"bound": "valu", "kernel": "k_walk_q (corrected)", "achieved": achieved,
                     "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": achieved / FP32_PEAK_TFLOPS,
                     "traffic": measured_traffic("k_walk_q_corrected", "k_walk_q_corrected_plummer16M_pmc.md"),
                     "traffic_source": "profiles/k_walk_q_corrected_plummer16M_pmc.md (rocprofv3 FETCH_SIZE/WRITE_SIZE)",
                     "flop_per_interaction": FLOP_PER_INTERACTION, "flop_per_quad_term": FLOP_PER_QUAD_TERM,
                     "interactions_per_step": int(c["interactions"]), "quad_terms_per_step": int(c["quad_terms"]),
                     "walk_ms": walk_ms, "interactions_per_s": c["interactions"] / (walk_ms * 1e-3)},
        "stage_ms": run["stage_ms"],
        "ms_per_step_with_stage_events": run["ms_per_step_with_stage_events"],
        "counters": {"opened": int(c["opened"]), "interactions_per_body": c["interactions"] / max(num, 1),
                     "quad_terms_per_body": c["quad_terms"] / max(num, 1),
                     "network_evaluations": int(c["corrected"]), "tree_height": int(c["tree_height"]),
                     "max_stack": int(c["max_stack"]), "error_flags": int(c["error_flags"]), "alive_bodies": num},
        "equal_p99": {
            "p99_target": P99_TARGET,
            "plain": plain_best, "quadrupole": quad_best, "learned": learned_best,
            "speedup_learned_vs_plain": (learned_best["body_steps_per_s"] / plain_best["body_steps_per_s"]
                                         if learned_best and plain_best else None),
            "speedup_quadrupole_vs_plain": (quad_best["body_steps_per_s"] / plain_best["body_steps_per_s"]
                                            if quad_best and plain_best else None),
            "speedup_learned_vs_plain_theta_0_5": (learned_best["body_steps_per_s"] / ref["body_steps_per_s"]
                                                   if learned_best else None),
            "quadrupole_geometric": fastest_at_target(geo_rows),
            "sweeps": plain_rows + geo_rows + quad_rows + learned_rows},
        "learned_training": train_info,
        "plain_theta_0_5": {"ms_per_step": ref["ms_per_step"], "body_steps_per_s": ref["body_steps_per_s"],
                            "stage_ms": ref["stage_ms"], "p50": plain_rows[0]["p50"], "p99": plain_rows[0]["p99"],
                            "interactions": int(ref["counters"]["interactions"]),
                            "walk_tflops": ref["counters"]["interactions"] * FLOP_PER_INTERACTION
                            / (ref["stage_ms"]["force"] * 1e-3) / 1e12},
    }
    copy = hbm_copy_gbps()
    sm = ref["stage_ms"]
    # algorithmic bytes per body: one-sweep sort 8 + 8 x 24, tree 172, encode 24, census 16 + index copy 8
    alg = {"sort": 200, "tree": 172, "morton": 24, "dead": 24}
    result["hbm_stages_plain_theta_0_5"] = {
        "copy_GBps": copy, "peak_GBps": HBM_PEAK_GBS,
        **{k: {"ms": sm[k], "alg_bytes_per_body": b, "alg_GBps": b * n / (sm[k] * 1e-3) / 1e9,
               "frac_of_peak": b * n / (sm[k] * 1e-3) / 1e9 / HBM_PEAK_GBS}
           for k, b in alg.items() if sm.get(k, 0.0) > 0.0}}
    result["render_feed"] = render_feed_leg(bh, plain_ref)
    Lr.attach(bh, None)
    bh.close()
    if not args.no_configs:
        result["configs"] = {name: other_workload(name, steps, args.seed)
                             for name, steps in (("plummer_65536", 10), ("cube_core_merging", 20),
                                                 ("two_galaxy", 5))}
    if not args.no_ensemble:
        log("ensemble leg")
        result["ensemble"] = ensemble_leg(args.ensemble_sims, args.steps, soft, 0.5, dt, args.seed, mlp,
                                          pick["force_tolerance"])
    mlp.close()
    if not args.no_cpu_baseline:
        log("CPU baseline (oracle port)")
        result["cpu_baseline"] = cpu_baseline(bodies, bounds, 0.5, dt, soft, args.cpu_threads,
                                              args.cpu_walk_sample)
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
