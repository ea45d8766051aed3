"""Write the golden fixtures by running the reference's GLSL on Mesa llvmpipe.

Run in the development container only (it needs the reference checkout and
Mesa's swrast driver):

    python tests/golden/make_glsl_fixtures.py [--reference /root/reference]

It builds ``oracle/_ref/glsl_runner`` (oracle/llvmpipe/glsl_runner.c, which
replays simulation/BarnesHut.java:99-164 dispatch by dispatch), runs every
scenario below with ``LP_NUM_THREADS=0`` and packs the dumps into
``tests/golden/<scenario>.npz``.

Why LP_NUM_THREADS=0: the reference's tree fill resets the work-queue head
and tail from invocation 0 while other work groups already enqueue
(shaders/compute/tree/bh_tree.comp:144-149,188-193). llvmpipe with worker
threads runs work groups concurrently and loses queue entries (the root stays
unfilled and every body feels zero force); with no worker threads the work
groups run in dispatch order, which is the schedule the reference relies on.

Each fixture holds the run parameters, the input bodies, and per step ``s``:
``s{s}_<stage>_<field>`` arrays (see ``pack_step``). Inputs of every stage
are the reference's own outputs of the previous stage, so tests can check
each stage in isolation.
"""
from __future__ import annotations

import argparse
import hashlib
import json
import os
import struct
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from gravity_amd import initial_conditions as ic  # noqa: E402

MAGIC = 0x4E494842


def write_input(path, bodies, steps, theta, dt, soft, merging, dynamic, wrap, flags, dump_from, bounds,
                elasticity=0.5, units=None):
    u = ic.g_one_units() if units is None else units
    hdr = struct.pack("<8I9f6f", MAGIC, bodies.shape[0], steps, merging, dynamic, wrap, flags, dump_from,
                      theta, dt, soft, elasticity, u.mass, u.density, u.len, u.time, 1.0,
                      *[float(x) for x in bounds])
    with open(path, "wb") as f:
        f.write(hdr)
        f.write(np.ascontiguousarray(bodies, dtype=np.float32).tobytes())


def run(runner, shaders, inp, out, threads="0"):
    env = dict(os.environ, LP_NUM_THREADS=threads)
    r = subprocess.run([runner, shaders, inp, out], env=env, capture_output=True, text=True, check=True)
    return json.loads(r.stdout.strip().splitlines()[-1])


def ld(d, name, dtype=np.uint32):
    return np.fromfile(os.path.join(d, name), dtype=dtype)


def pack_step(d, s, n, full=True):
    """Arrays of step s. Internal nodes are split into topology (u32) and fill (f32)."""
    out = {}
    sim_dead = ld(d, f"s{s}_dead_sim.bin")
    num_before = int(sim_dead[0])
    died = int(sim_dead[2])
    num = num_before - died
    out[f"s{s}_num_before"] = np.uint32(num_before)
    out[f"s{s}_died"] = np.uint32(died)
    out[f"s{s}_G"] = sim_dead.view(np.float32)[21:22].copy()
    out[f"s{s}_blc"] = sim_dead.view(np.float32)[22:23].copy()
    out[f"s{s}_bounds_prev"] = None  # filled by caller
    sim_a = ld(d, f"s{s}_aabb_sim.bin").view(np.float32)
    out[f"s{s}_bounds"] = np.concatenate([sim_a[8:11], sim_a[12:15]]).astype(np.float32)
    sim_f = ld(d, f"s{s}_force_sim.bin")
    out[f"s{s}_opened"] = np.uint32(sim_f[23])
    out[f"s{s}_out_of_bounds"] = np.uint32(sim_f[4])
    mq = ld(d, f"s{s}_force_merge.bin")
    nt = int(mq[1])
    out[f"s{s}_tasks"] = mq[2:2 + 2 * nt].reshape(-1, 2).copy()
    if not full:
        return out, num
    out[f"s{s}_bodies_start"] = ld(d, f"s{s}_dead_bodies_in.bin", np.float32).reshape(n, 8)
    out[f"s{s}_index_dead"] = ld(d, f"s{s}_dead_index_in.bin")
    out[f"s{s}_morton_encoded"] = ld(d, f"s{s}_encode_morton_in.bin", np.uint64)[:num]
    out[f"s{s}_morton_sorted"] = ld(d, f"s{s}_sort_morton_in.bin", np.uint64)[:num]
    out[f"s{s}_index_sorted"] = ld(d, f"s{s}_sort_index_in.bin")[:num]
    ib = ld(d, f"s{s}_build_internal.bin").reshape(n - 1, 20)
    lb = ld(d, f"s{s}_build_leaf.bin").reshape(n, 20)
    out[f"s{s}_build_internal_u32"] = ib[: max(num - 1, 0), 10:16].copy()
    out[f"s{s}_build_leaf_parent"] = lb[:num, 15].copy()
    inf = ld(d, f"s{s}_fill_internal.bin").reshape(n - 1, 20)
    lf = ld(d, f"s{s}_fill_leaf.bin").reshape(n, 20)
    out[f"s{s}_fill_internal"] = inf[: max(num - 1, 0)].copy()
    out[f"s{s}_fill_leaf"] = lf[:num].copy()
    acc = ld(d, f"s{s}_accel_bodies_out.bin", np.float32).reshape(n, 8)[:, 4:7]
    out[f"s{s}_accel"] = acc.copy()
    out[f"s{s}_force_out"] = ld(d, f"s{s}_force_bodies_out.bin", np.float32).reshape(n, 8)
    out[f"s{s}_merge_out"] = ld(d, f"s{s}_merge_bodies_out.bin", np.float32).reshape(n, 8)
    return out, num


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def scenario_list():
    sc = []
    sc.append(dict(name="plummer_4096", bodies=ic.plummer(4096, 7), steps=2, theta=0.5, dt=1e-3, soft=0.01,
                   merging=0, dynamic=1, wrap=0))
    sc.append(dict(name="merge_4096", bodies=ic.cube_plus_core(4096, 11, density=1e3), steps=3, theta=0.5,
                   dt=5e-4, soft=1e-6, merging=1, dynamic=1, wrap=0))
    # COLLISION mode: impulses and (restitution 0) position corrections mid-walk
    sc.append(dict(name="collision_4096", bodies=ic.cube_plus_core(4096, 13, density=1e3), steps=2, theta=0.5,
                   dt=5e-4, soft=1e-6, merging=2, dynamic=1, wrap=0))
    # dead bodies, NaN/Inf and far outliers, N not a multiple of 256
    b = ic.plummer(1000, 21)
    rng = np.random.default_rng(5)
    dead = rng.choice(1000, 50, replace=False)
    b[dead, 3] = 0.0
    b[dead[:10], 0] = 7.0                      # dead bodies with stale, non-zero fields
    b[3, 0] = np.nan
    b[4, 1] = np.inf
    b[5, 2] = -np.inf
    b[6, 0:3] = [1e30, 0.0, 0.0]
    b[7, 0:3] = [0.0, -1e30, 5.0]
    sc.append(dict(name="edge_dead_nan_1000", bodies=b, steps=2, theta=0.5, dt=1e-3, soft=0.01, merging=0,
                   dynamic=1, wrap=0, bounds=ic.tight_bounds(np.delete(b, [3, 4, 5, 6, 7], axis=0))))
    # duplicated positions: equal Morton codes exercise the index tie-break
    base = ic.plummer(64, 3)
    b = np.repeat(base, 8, axis=0)
    b[:, 3] = 1.0 / b.shape[0]
    sc.append(dict(name="edge_duplicates_512", bodies=b, steps=2, theta=0.5, dt=1e-3, soft=0.01, merging=0,
                   dynamic=1, wrap=0))
    # static bounds smaller than the cloud: bodies outside are killed and counted
    b = ic.plummer(1024, 9)
    sc.append(dict(name="static_kill_1024", bodies=b, steps=2, theta=0.5, dt=1e-3, soft=0.01, merging=0,
                   dynamic=0, wrap=0, bounds=np.array([-1.5, -1.5, -1.5, 1.5, 1.5, 1.5], np.float32)))
    sc.append(dict(name="static_wrap_1024", bodies=ic.plummer(1024, 10), steps=2, theta=0.5, dt=0.05, soft=0.01,
                   merging=0, dynamic=0, wrap=1,
                   bounds=np.array([-1.5, -1.5, -1.5, 1.5, 1.5, 1.5], np.float32)))
    # theta = 0 opens every internal node: the walk is a direct sum in walk order
    sc.append(dict(name="theta0_512", bodies=ic.plummer(512, 4), steps=1, theta=0.0, dt=1e-3, soft=0.01,
                   merging=0, dynamic=1, wrap=0))
    # a tree higher than the 64 fill passes (GPU.java:41): 32 bodies at one
    # point (a subtree of height 5) under a chain of 63 bodies whose Morton
    # codes each hold one bit (height 63 above it); the top nodes stay
    # unfilled and the walk takes the empty root at once
    sc.append(dict(name="tall_tree_95", bodies=tall_tree_bodies(), steps=2, theta=0.5, dt=1e-3, soft=0.01,
                   merging=0, dynamic=0, wrap=0, bounds=np.array([0, 0, 0, 1, 1, 1], np.float32)))
    # the two-galaxy workload (disks with bulges, 10 apart, parabolic orbit)
    sc.append(dict(name="two_galaxy_4096", bodies=ic.two_galaxy(2048, 23), steps=2, theta=0.5, dt=1e-3,
                   soft=0.01, merging=0, dynamic=1, wrap=0))
    return sc


def tall_tree_bodies(dup=32):
    """Bodies in the box [0, 1]^3 whose Morton codes are 0 (dup bodies) and
    2^k for k = 0 .. 62: a quantised coordinate 2^j on one axis (x: bit 3j +
    2, y: 3j + 1, z: 3j) and 0 on the others, each at the middle of its cell
    (q = floor(p * 2097151))."""
    cells = [(0, 0, 0)] * dup
    for j in range(21):
        cells += [(0, 0, 1 << j), (0, 1 << j, 0), (1 << j, 0, 0)]
    q = np.array(cells, np.float64)
    b = np.zeros((len(cells), 8), np.float32)
    b[:, 0:3] = (q + 0.5) / 2097151.0
    b[:, 3] = 1.0 / len(cells)
    b[:, 7] = 1.0
    return b


def unit_sets():
    """(name, Units) pairs whose G and body-length constant the reference's
    UPDATE program derives in-shader (common/common.glsl:345-352): the G = 1
    units, the reference's named unit sets (simulation/UnitSet.java) and a few
    arbitrary ones, including ones whose G is an FP32 subnormal."""
    from gravity_amd import planets as pl
    out = [("g_one", ic.g_one_units())]
    for us in (pl.SOLAR_SYSTEM_SECOND, pl.SOLAR_SYSTEM_HOUR, pl.ASTRONOMICAL, pl.METRIC, pl.GALACTIC_MERGE):
        m, d, ln, t = us.uniforms()
        out.append((us.name, ic.Units(mass=m, density=d, len=ln, time=t)))
    for k, (m, d, ln, t) in enumerate([(1e10, 1e3, 1e3, 1e5), (3.0, 0.25, 0.5, 2.0), (1e20, 5e2, 1e6, 10.0),
                                       (7.0e29, 2.0e3, 3.0e10, 60.0), (1.0, 1.0, 1.0, 1.0)]):
        out.append((f"custom_{k}", ic.Units(mass=m, density=d, len=ln, time=t)))
    return out


def units_fixture(runner, shaders):
    """tests/golden/units.npz: per unit set the uniforms and the G and
    body-length constant bits the reference's UPDATE program leaves in
    SimulationValues (one step of two bodies per unit set)."""
    names, uni, G, blc = [], [], [], []
    bodies = ic.plummer(2, 1)
    for name, u in unit_sets():
        with tempfile.TemporaryDirectory() as td:
            inp = os.path.join(td, "in.bin")
            write_input(inp, bodies, 1, 0.5, 1e-3, 0.01, 0, 1, 0, 1, 0, ic.tight_bounds(bodies), units=u)
            run(runner, shaders, inp, td)
            sim = ld(td, "s0_dead_sim.bin").view(np.float32)
            names.append(name)
            uni.append([u.mass, u.density, u.len, u.time])
            G.append(sim[21])
            blc.append(sim[22])
    path = os.path.join(HERE, "units.npz")
    np.savez_compressed(path, names=np.array(names), uniforms=np.array(uni, np.float32),
                        G=np.array(G, np.float32), blc=np.array(blc, np.float32))
    print(f"units: {len(names)} unit sets -> {path}")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--reference", default="/root/reference")
    ap.add_argument("--only", default=None)
    ap.add_argument("--skip-config1", action="store_true")
    args = ap.parse_args()
    shaders = os.path.join(args.reference, "src/main/resources/shaders")
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"], check=True)
    runner = os.path.join(ROOT, "oracle", "_ref", "glsl_runner")

    for sc in scenario_list():
        if args.only and sc["name"] != args.only:
            continue
        bodies = sc["bodies"]
        n = bodies.shape[0]
        bounds = sc.get("bounds")
        if bounds is None:
            bounds = ic.tight_bounds(bodies)
        with tempfile.TemporaryDirectory() as td:
            inp = os.path.join(td, "in.bin")
            write_input(inp, bodies, sc["steps"], sc["theta"], sc["dt"], sc["soft"], sc["merging"], sc["dynamic"],
                        sc["wrap"], 1 | 2, 0, bounds)
            info = run(runner, shaders, inp, td)
            arrays = {"bodies0": bodies.astype(np.float32), "bounds0": np.asarray(bounds, np.float32),
                      "params": np.array([sc["theta"], sc["dt"], sc["soft"], 0.5], np.float32),
                      "modes": np.array([sc["merging"], sc["dynamic"], sc["wrap"], sc["steps"]], np.uint32)}
            prev = np.asarray(bounds, np.float32)
            for s in range(sc["steps"]):
                st, _ = pack_step(td, s, n)
                st[f"s{s}_bounds_prev"] = prev
                prev = st[f"s{s}_bounds"]
                arrays.update(st)
            arrays["final_bodies"] = ld(td, f"s{sc['steps'] - 1}_end_bodies_in.bin", np.float32).reshape(n, 8)
        path = os.path.join(HERE, sc["name"] + ".npz")
        np.savez_compressed(path, **arrays)
        print(f"{sc['name']}: N={n} steps={sc['steps']} -> {os.path.getsize(path) // 1024} KiB  ({info['total_ms']:.0f} ms)")

    if args.only in (None, "units"):
        units_fixture(runner, shaders)
    if args.only in (None, "plummer_65536") and not args.skip_config1:
        config1(runner, shaders)


def config1(runner, shaders):
    """Plummer N=65,536, theta 0.5, dt 1e-3, softening 0.01, 10 steps, merging off, dynamic bounds.

    Full arrays are too large to commit; the fixture keeps per-step hashes of
    the integer stages and of the bodies, the counters, the bounds and a
    1,024-body sample of accelerations and end-of-step bodies.
    """
    n, steps = 65536, 10
    bodies = ic.plummer(n, 1)
    bounds = ic.tight_bounds(bodies)
    sample = np.random.default_rng(2024).choice(n, 1024, replace=False).astype(np.uint32)
    arrays = {"params": np.array([0.5, 1e-3, 0.01, 0.5], np.float32),
              "modes": np.array([0, 1, 0, steps], np.uint32), "bounds0": bounds, "sample": sample,
              "seed": np.uint32(1)}
    hashes = {}
    with tempfile.TemporaryDirectory() as td:
        inp = os.path.join(td, "in.bin")
        write_input(inp, bodies, steps, 0.5, 1e-3, 0.01, 0, 1, 0, 1 | 2, 0, bounds)
        info = run(runner, shaders, inp, td)
        prev = bounds
        for s in range(steps):
            st, num = pack_step(td, s, n, full=True)
            arrays[f"s{s}_bounds_prev"] = prev
            prev = st[f"s{s}_bounds"]
            for k in ("G", "blc", "bounds", "opened", "died", "num_before"):
                arrays[f"s{s}_{k}"] = st[f"s{s}_{k}"]
            for k in ("morton_encoded", "morton_sorted", "index_sorted", "build_internal_u32", "build_leaf_parent",
                      "fill_internal", "fill_leaf", "accel", "force_out", "bodies_start"):
                hashes[f"s{s}_{k}"] = sha(st[f"s{s}_{k}"])
            arrays[f"s{s}_accel_sample"] = st[f"s{s}_accel"][sample]
            arrays[f"s{s}_force_out_sample"] = st[f"s{s}_force_out"][sample]
            arrays[f"s{s}_height"] = np.uint32(st[f"s{s}_fill_internal"][0, 12])
        arrays["final_sample"] = ld(td, f"s{steps - 1}_end_bodies_in.bin", np.float32).reshape(n, 8)[sample]
    arrays["hashes"] = np.array(json.dumps(hashes))
    path = os.path.join(HERE, "plummer_65536.npz")
    np.savez_compressed(path, **arrays)
    print(f"plummer_65536: N={n} steps={steps} -> {os.path.getsize(path) // 1024} KiB ({info['total_ms']:.0f} ms)")


if __name__ == "__main__":
    main()
