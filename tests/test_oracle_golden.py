"""The CPU oracle against the reference's GLSL run on Mesa llvmpipe.

Every stage is fed the reference's own inputs for that step (stage-isolated
parity), so a mismatch names the stage. Integer stages and the FP32 stages
(bounds, Morton quantisation, fill, accelerations, integration) must match
bit for bit.
"""
import collections
import json
import hashlib

import numpy as np
import pytest

from oracle import bh_oracle as O
from tests.golden_util import FULL, load, num_alive, params_of, stage_inputs, steps


def u32(a):
    return np.ascontiguousarray(a).view(np.uint32)


def assert_bits(actual, desired):
    """Bit-exact float comparison in which every NaN equals every NaN (the
    sign and payload of a NaN are not part of the contract)."""
    a = np.ascontiguousarray(actual, dtype=np.float32)
    d = np.ascontiguousarray(desired, dtype=np.float32)
    np.testing.assert_array_equal(np.isnan(a), np.isnan(d))
    keep = ~np.isnan(a)
    np.testing.assert_array_equal(a[keep].view(np.uint32), d[keep].view(np.uint32))


CASES = [(name, s) for name in FULL for s in range(steps(load(name)))]


@pytest.mark.parametrize("name,s", CASES)
def test_stage_parity(name, s):
    fx = load(name)
    n = fx["bodies0"].shape[0]
    prm = O.params(**params_of(fx))
    src, idx, num_before = stage_inputs(fx, s)
    dst = np.zeros_like(src)

    # dead partition (bh_dead.comp)
    out, died = O.dead_partition(src, dst, idx, num_before)
    assert died == int(fx[f"s{s}_died"])
    np.testing.assert_array_equal(out[:num_before], fx[f"s{s}_index_dead"][:num_before])
    assert_bits(src, fx[f"s{s}_bodies_start"])
    num = num_before - died
    src = fx[f"s{s}_bodies_start"].copy()
    idx = fx[f"s{s}_index_dead"].copy()

    # AABB (bh_morton.comp:114-231)
    bounds = fx[f"s{s}_bounds_prev"]
    if prm.dynamic == 1:
        bounds = O.aabb(src, idx, num, bounds)
    np.testing.assert_array_equal(u32(bounds), u32(fx[f"s{s}_bounds"]))

    # Morton encode
    codes = O.morton_encode(src, idx, num, fx[f"s{s}_bounds"])
    np.testing.assert_array_equal(codes, fx[f"s{s}_morton_encoded"])

    # radix sort
    keys, sidx = O.sort(fx[f"s{s}_morton_encoded"], idx[:num])
    np.testing.assert_array_equal(keys, fx[f"s{s}_morton_sorted"])
    np.testing.assert_array_equal(sidx, fx[f"s{s}_index_sorted"])

    # Karras build: childA, childB, nodeDepth(=min(i,j) before fill), count, ready, parent
    leaf, internal = O.build(fx[f"s{s}_morton_sorted"], num, n)
    np.testing.assert_array_equal(internal[: num - 1, 10:16], fx[f"s{s}_build_internal_u32"])
    np.testing.assert_array_equal(leaf[:num, 15], fx[f"s{s}_build_leaf_parent"])

    # fill, colours included: bit-exact
    sidx_full = idx.copy()
    sidx_full[:num] = fx[f"s{s}_index_sorted"]
    G = float(fx[f"s{s}_G"][0])
    blc = float(fx[f"s{s}_blc"][0])
    from gravity_amd.initial_conditions import g_one_units
    u = g_one_units()
    O.fill(src, sidx_full, num, n, leaf, internal, u.mass, u.density, max_passes=64)
    np.testing.assert_array_equal(internal[: num - 1], fx[f"s{s}_fill_internal"])
    np.testing.assert_array_equal(leaf[:num], fx[f"s{s}_fill_leaf"])

    # walk + integrate on the reference's tree
    leaf = np.zeros((n, 20), np.uint32)
    internal = np.zeros((n - 1, 20), np.uint32)
    leaf[:num] = fx[f"s{s}_fill_leaf"]
    internal[: num - 1] = fx[f"s{s}_fill_internal"]
    r = O.force(src, sidx_full, num, n, leaf, internal, prm, G, blc, fx[f"s{s}_bounds"], dst=dst)
    assert r["rc"] == 0
    assert r["stats"].opened == int(fx[f"s{s}_opened"])
    alive = sidx_full[:num]
    acc = np.zeros((n, 3), np.float32)
    acc[alive] = r["accel"][:num]
    killed = np.zeros(n, bool)
    if prm.dynamic == 0 and not prm.wrap_around:
        # the reference's acceleration dump comes from a FORCE pass that also
        # kills out-of-bounds bodies (bh_force.comp:121-132): their slot is empty
        b = fx[f"s{s}_bounds"]
        p = src[:, :3]
        killed = ((p < b[:3]) | (p > b[3:])).any(axis=1)
    keep = alive[~killed[alive]]
    assert_bits(acc[keep], fx[f"s{s}_accel"][keep])
    ref_out = fx[f"s{s}_force_out"]
    if prm.dynamic == 0 and not prm.wrap_around:
        # sim.outOfBounds accumulates over steps (bh_force.comp:128)
        before = int(fx[f"s{s - 1}_out_of_bounds"]) if s else 0
        assert r["stats"].out_of_bounds == int(fx[f"s{s}_out_of_bounds"]) - before
    assert_bits(dst[alive], ref_out[alive])

    # merge tasks: the same set (queue order differs between schedules)
    got = set(map(tuple, r["tasks"].tolist()))
    want = set(map(tuple, fx[f"s{s}_tasks"].tolist()))
    assert got == want

    # merge resolution: pairs whose bodies appear in no other task are
    # deterministic and must match bit for bit; contested bodies conserve mass
    mine = ref_out.copy()
    O.merge(mine, fx[f"s{s}_tasks"])
    ref_m = fx[f"s{s}_merge_out"]
    cnt = collections.Counter(fx[f"s{s}_tasks"].ravel().tolist())
    for a, b in fx[f"s{s}_tasks"].tolist():
        if cnt[a] == 1 and cnt[b] == 1:
            assert_bits(mine[[a, b]], ref_m[[a, b]])
    untouched = np.setdiff1d(np.arange(n), np.array(list(cnt.keys()), dtype=np.int64))
    assert_bits(mine[untouched], ref_m[untouched])
    np.testing.assert_allclose(mine[:, 3].astype(np.float64).sum(), ref_m[:, 3].astype(np.float64).sum(), rtol=1e-6)


@pytest.mark.parametrize("name", ["plummer_4096", "collision_4096", "edge_dead_nan_1000", "edge_duplicates_512",
                                  "two_galaxy_4096",
                                  "static_kill_1024", "static_wrap_1024", "theta0_512"])
def test_free_running_trajectory(name):
    """Without merges the whole multi-step trajectory is deterministic: the
    oracle's step loop must end on the reference's bodies bit for bit."""
    fx = load(name)
    prm = O.params(**params_of(fx))
    sim = O.Simulation(fx["bodies0"], fx["bounds0"], 1.0, 1.0, G=float(fx["s0_G"][0]),
                       body_len_const=float(fx["s0_blc"][0]))
    for s in range(steps(fx)):
        assert sim.step(prm) == 0
        assert sim.num_bodies == num_alive(fx, s)
    assert_bits(sim.bodies(), fx["final_bodies"])


def _sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def test_plummer_65536_ten_steps():
    """The reference's CPU-runnable case: Plummer N=65,536, 10 steps.

    Per step: hashes of the Morton codes, permutation, topology, fill,
    accelerations and integrated bodies; sampled accelerations; counters.
    """
    fx = load("plummer_65536")
    hashes = json.loads(str(fx["hashes"]))
    from gravity_amd import initial_conditions as ic
    n = 65536
    bodies = ic.plummer(n, int(fx["seed"]))
    np.testing.assert_array_equal(ic.tight_bounds(bodies), fx["bounds0"])
    prm = O.params(theta=0.5, dt=1e-3, softening=0.01, merging=0, dynamic=1)
    G = float(fx["s0_G"][0])
    src = bodies.copy()
    dst = np.zeros_like(src)
    idx = np.arange(n, dtype=np.uint32)
    bounds = fx["bounds0"]
    num = n
    sample = fx["sample"]
    for s in range(10):
        out, died = O.dead_partition(src, dst, idx, num)
        num -= died
        assert _sha(src) == hashes[f"s{s}_bodies_start"]
        bounds = O.aabb(src, out, num, bounds)
        np.testing.assert_array_equal(u32(bounds), u32(fx[f"s{s}_bounds"]))
        codes = O.morton_encode(src, out, num, bounds)
        assert _sha(codes) == hashes[f"s{s}_morton_encoded"]
        keys, sidx = O.sort(codes, out[:num])
        assert _sha(keys) == hashes[f"s{s}_morton_sorted"]
        assert _sha(sidx) == hashes[f"s{s}_index_sorted"]
        idx = out.copy()
        idx[:num] = sidx
        leaf, internal = O.build(keys, num, n)
        assert _sha(internal[: num - 1, 10:16]) == hashes[f"s{s}_build_internal_u32"]
        u = ic.g_one_units()
        h = O.fill(src, idx, num, n, leaf, internal, u.mass, u.density, max_passes=64, with_colors=True)
        assert h == int(fx[f"s{s}_height"])
        assert _sha(internal[: num - 1]) == hashes[f"s{s}_fill_internal"]
        assert _sha(leaf[:num]) == hashes[f"s{s}_fill_leaf"]
        r = O.force(src, idx, num, n, leaf, internal, prm, G, float(fx[f"s{s}_blc"][0]), bounds, dst=dst, threads=8)
        assert r["stats"].opened == int(fx[f"s{s}_opened"])
        acc = np.zeros((n, 3), np.float32)
        acc[idx[:num]] = r["accel"][:num]
        assert _sha(acc) == hashes[f"s{s}_accel"]
        np.testing.assert_array_equal(u32(acc[sample]), u32(fx[f"s{s}_accel_sample"]))
        np.testing.assert_array_equal(u32(dst[sample]), u32(fx[f"s{s}_force_out_sample"]))
        assert _sha(dst) == hashes[f"s{s}_force_out"]
        src, dst = dst, src
    np.testing.assert_array_equal(u32(src[sample]), u32(fx["final_sample"]))
