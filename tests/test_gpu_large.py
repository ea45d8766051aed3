"""HIP path against the CPU oracle at sizes beyond the llvmpipe fixtures.

At N = 1,048,576 (Plummer) every stage is compared bit for bit with the
oracle run on the same input: Morton codes, the permutation, the whole tree
(topology, counts, heights, centres of mass, boxes) and the accelerations.
At N = 4,194,304 the size-independent properties are checked: the keys are
sorted, the index is a permutation, the tree is a valid Karras tree (every
node's count is the sum of its children's, the root holds all the mass).
"""
import numpy as np
import pytest

from oracle import bh_oracle as O

pytestmark = pytest.mark.gpu


def _bits_equal(a, b):
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    return np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_gpu_matches_oracle_1m():
    from gravity_amd import barnes_hut as B
    from gravity_amd import initial_conditions as ic
    n = 1 << 20
    print(f"[gpu] large parity n={n}", flush=True)
    bodies = ic.plummer(n, 99)
    bounds = ic.tight_bounds(bodies)
    u = ic.g_one_units()
    G, blc = O.units(u.mass, u.density, u.len, u.time)
    prm = B.make_params(theta=0.5, dt=1e-3, softening=0.01, G=float(G), body_length_constant=float(blc))
    bh = B.BarnesHut(bodies, bounds, prm)
    bh.init()
    bh.reset_values()
    bh.partition_dead_bodies()
    bh.update_bounds()
    bh.generate_morton_codes()
    bh.radix_sort()
    bh.build_tree()
    keys = bh.morton()
    idx = bh.index()
    internal = bh.nodes(B.INTERNAL_NODES)
    gbounds = bh.bounds()

    src = bodies.copy()
    out, died = O.dead_partition(src, None, np.arange(n, dtype=np.uint32), n)
    assert died == 0
    ob = O.aabb(src, out, n, bounds)
    assert _bits_equal(ob, gbounds)
    codes = O.morton_encode(src, out, n, ob)
    okeys, oidx = O.sort(codes, out)
    np.testing.assert_array_equal(keys, okeys)
    np.testing.assert_array_equal(idx, oidx)
    leaf, ointernal = O.build(okeys, n, n)
    h = O.fill(src, oidx, n, n, leaf, ointernal, max_passes=0, with_colors=False)
    assert bh.counters()["tree_height"] == h
    np.testing.assert_array_equal(internal[:, 10:16], ointernal[:, 10:16])
    assert _bits_equal(internal[:, 0:10], ointernal[:, 0:10])

    acc = bh.accelerations()
    oprm = O.params(theta=0.5, dt=1.0, softening=0.01)
    zv = src.copy()
    zv[:, 4:7] = 0
    dst = np.zeros_like(zv)
    r = O.force(zv, oidx, n, n, leaf, ointernal, oprm, G, blc, ob, dst=dst, stack_cap=256, threads=16)
    assert r["rc"] == 0
    assert _bits_equal(acc[oidx], r["accel"])
    bh.close()


def test_gpu_tree_properties_4m():
    from gravity_amd import barnes_hut as B
    from gravity_amd import initial_conditions as ic
    n = 1 << 22
    print(f"[gpu] large properties n={n}", flush=True)
    bodies = ic.cube_plus_core(n, 5, density=64.0)
    bounds = ic.tight_bounds(bodies)
    bh = B.BarnesHut(bodies, bounds, B.make_params(theta=0.5, dt=5e-4, softening=1e-6))
    bh.init()
    bh.reset_values()
    bh.partition_dead_bodies()
    bh.update_bounds()
    bh.generate_morton_codes()
    bh.radix_sort()
    bh.build_tree()
    keys = bh.morton()
    idx = bh.index()
    assert np.all(keys[1:] >= keys[:-1])
    assert np.array_equal(np.sort(idx), np.arange(n, dtype=np.uint32))
    nodes = bh.nodes(B.INTERNAL_NODES)
    f = O.node_fields(nodes)
    count = f["bodiesContained"].astype(np.int64)
    a, b = f["childA"].astype(np.int64), f["childB"].astype(np.int64)
    ca = np.where(a < n, 1, count[np.clip(a - n, 0, n - 2)])
    cb = np.where(b < n, 1, count[np.clip(b - n, 0, n - 2)])
    np.testing.assert_array_equal(count, ca + cb)
    assert count[0] == n and f["parentId"][0] == 0xFFFFFFFF
    np.testing.assert_allclose(f["mass"][0], bodies[:, 3].astype(np.float64).sum(), rtol=1e-5)
    c = bh.counters()
    assert c["error_flags"] == 0
    bh.close()


def test_walk_inverse_sqrt_is_exact_on_its_domain():
    """The walk's fast 1/sqrt equals RN(1/RN(sqrt(x))) for every float in [2^-100, 2^100]."""
    from gravity_amd import barnes_hut as B
    from gravity_amd import initial_conditions as ic
    print("[gpu] math selftest", flush=True)
    b = ic.plummer(1024, 1)
    bh = B.BarnesHut(b, ic.tight_bounds(b))
    assert bh.selftest_math() == [0, 0, 0]
    bh.close()
