"""The one-simulation radix sort (bh_radix_sort) on keys uploaded through the
C ABI, against numpy's stable argsort: sizes around the 2,048-key tile of
sorts up to 262,144 keys and the 8,192-key tile above that, keys
with many ties, nearly ordered keys (the time-stepping case), all-equal keys,
a sort over fewer bodies than the context holds, and sorts in a row (each
sort's look-back words carry their own epoch).

Reference: the stable LSD sort of `shaders/compute/radix/bh_radix.comp:12-166`
driven by `BarnesHut.radixSort` (`simulation/BarnesHut.java:358-448`).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TILE = 8192
SMALL_TILE, SMALL_MAX = 2048, 1 << 18


def make_bh(n):
    from gravity_amd import barnes_hut as B
    from gravity_amd import initial_conditions as ic
    bodies = ic.plummer(n, 5)
    bh = B.BarnesHut(bodies, ic.tight_bounds(bodies), B.make_params(theta=0.5, dt=1e-3, softening=0.01, G=1.0))
    bh.init()
    return bh


def keys_of(kind, n, rng):
    if kind == "random":
        return rng.integers(0, 1 << 63, size=n, dtype=np.uint64)
    if kind == "ties":   # few distinct keys spread over all eight digits
        pool = rng.integers(0, 1 << 63, size=37, dtype=np.uint64)
        return pool[rng.integers(0, pool.size, size=n)]
    if kind == "nearly_ordered":   # sorted, then a tenth of the keys moved by a few positions
        k = np.sort(rng.integers(0, 1 << 63, size=n, dtype=np.uint64))
        i = rng.integers(0, n, size=n // 10)
        j = np.clip(i + rng.integers(-40, 41, size=i.size), 0, n - 1)
        k[i], k[j] = k[j].copy(), k[i].copy()
        return k
    if kind == "equal":
        return np.full(n, 0x5A5A5A5A5A5A5A5, np.uint64)
    raise ValueError(kind)


def check_sort(bh, keys, num):
    from gravity_amd import barnes_hut as B
    n = keys.size
    idx = np.random.default_rng(num).permutation(n).astype(np.uint32)
    bh.upload(B.MORTON_IN, keys)
    bh.upload(B.INDEX_IN, idx)
    bh.radix_sort()
    order = np.argsort(keys[:num], kind="stable")
    got_k, got_i = bh.morton(), bh.index()
    np.testing.assert_array_equal(got_k[:num], keys[:num][order])
    np.testing.assert_array_equal(got_i[:num], idx[:num][order])
    assert bh.counters()["error_flags"] == 0


@pytest.mark.parametrize("n", [2, 3, 1000, SMALL_TILE - 1, SMALL_TILE + 1, TILE - 1, TILE, TILE + 1, 3 * TILE + 17,
                               SMALL_MAX, SMALL_MAX + 1, (1 << 20) + 5])
@pytest.mark.parametrize("kind", ["random", "ties", "nearly_ordered", "equal"])
def test_gpu_sort_matches_stable_argsort(n, kind):
    print(f"[gpu] sort n={n} {kind}", flush=True)
    bh = make_bh(n)
    check_sort(bh, keys_of(kind, n, np.random.default_rng(n)), n)
    bh.close()


def test_gpu_sorts_in_a_row_and_fewer_bodies():
    """Three sorts in a row on one context (new epochs), the last over fewer
    bodies than the context holds (tiles past numBodies take no part)."""
    from gravity_amd import barnes_hut as B
    n = 5 * TILE + 123
    print(f"[gpu] sorts in a row n={n}", flush=True)
    bh = make_bh(n)
    rng = np.random.default_rng(11)
    check_sort(bh, keys_of("random", n, rng), n)
    check_sort(bh, keys_of("ties", n, rng), n)
    num = 2 * TILE + 5
    sim = bh.sim_values()
    sim[0] = num
    bh.upload(B.SIMULATION_VALUES, sim)
    check_sort(bh, keys_of("nearly_ordered", n, rng), num)
    bh.close()
