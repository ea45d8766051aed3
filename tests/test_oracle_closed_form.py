"""The CPU oracle against what the mathematics fixes, independent of the
reference: brute-force bit interleaving, textbook stable sort and partition,
Karras-2012 invariants with brute-force split positions, closed-form node
fill, theta = 0 direct summation and conservation laws of a merge.

Each check is chosen so that a plausible slip (a dropped term, a wrong
shift, a swapped operand, an unstable sort) fails it.
"""
import numpy as np
import pytest

from oracle import bh_oracle as O

EMPTY = 0xFFFFFFFF


def naive_interleave(x, y, z):
    code = np.zeros_like(x, dtype=np.uint64)
    for k in range(21):
        code |= ((x >> np.uint64(k)) & np.uint64(1)) << np.uint64(3 * k + 2)
        code |= ((y >> np.uint64(k)) & np.uint64(1)) << np.uint64(3 * k + 1)
        code |= ((z >> np.uint64(k)) & np.uint64(1)) << np.uint64(3 * k)
    return code


def test_expand_bits_all_21bit_inputs():
    v = np.arange(1 << 21, dtype=np.uint64)
    want = np.zeros_like(v)
    for k in range(21):
        want |= ((v >> np.uint64(k)) & np.uint64(1)) << np.uint64(3 * k)
    L = O.lib()
    got = np.array([L.bho_expand_bits21(int(x)) for x in v[::4099]], dtype=np.uint64)
    np.testing.assert_array_equal(got, want[::4099])
    # full range through the encoder: a body exactly on grid point q maps to q
    rng = np.random.default_rng(0)
    q = rng.integers(0, 1 << 21, size=(4096, 3)).astype(np.uint64)
    codes = np.array([L.bho_morton3d(int(a), int(b), int(c)) for a, b, c in q], dtype=np.uint64)
    np.testing.assert_array_equal(codes, naive_interleave(q[:, 0], q[:, 1], q[:, 2]))


def test_morton_quantisation_grid_points():
    """pNorm = k/2097151 exactly representable positions land on cell k."""
    n = 2048
    rng = np.random.default_rng(1)
    k = rng.integers(0, 2097151, size=(n, 3)).astype(np.float64)
    bodies = np.zeros((n, 8), np.float32)
    bodies[:, 0:3] = (k * 1.0).astype(np.float32)          # bounds [0, 2097151]: ext = 2097151
    bodies[:, 3] = 1.0
    bounds = np.array([0, 0, 0, 2097151, 2097151, 2097151], np.float32)
    codes = O.morton_encode(bodies, np.arange(n, dtype=np.uint32), n, bounds)
    # (k - 0)/2097151 * 2097151 may round one ulp below k, floor then gives k-1:
    # compute the expected cell with the same float32 arithmetic, but by numpy
    pn = (bodies[:, 0:3] - np.float32(0)) / np.float32(2097151)
    cell = np.clip(np.floor(pn.astype(np.float32) * np.float32(2097151.0)), 0, 2097151).astype(np.uint64)
    np.testing.assert_array_equal(codes, naive_interleave(cell[:, 0], cell[:, 1], cell[:, 2]))
    assert (np.abs(cell.astype(np.float64) - k) <= 1).all()


def test_aabb_is_exact_min_max_of_inliers():
    rng = np.random.default_rng(2)
    n = 5000
    b = np.zeros((n, 8), np.float32)
    b[:, 0:3] = rng.normal(size=(n, 3)).astype(np.float32)
    b[:, 3] = 1.0
    b[::7, 3] = 0.0                                  # dead bodies are ignored
    b[5, 0] = np.nan                                 # non-finite bodies are ignored
    b[6, 1] = 1e30                                   # outliers beyond 1.05x the window are ignored
    prev = np.array([-10, -10, -10, 10, 10, 10], np.float32)
    got = O.aabb(b, np.arange(n, dtype=np.uint32), n, prev)
    keep = (b[:, 3] != 0) & np.isfinite(b[:, :3]).all(1) & (np.abs(b[:, :3]) <= 10.5).all(1)
    lo = b[keep, :3].min(0)
    hi = b[keep, :3].max(0)
    # re-centred: centre +- half extent, computed in float32
    c = np.float32(0.5) * (lo + hi)
    h = np.float32(0.5) * (hi - lo)
    np.testing.assert_array_equal(got, np.concatenate([c - h, c + h]))
    assert np.all(np.abs(got - np.concatenate([lo, hi])) <= 2e-6)
    # growth clamp: the 1.05x inlier window keeps a box below 1.1x its previous
    # extent, except for a degenerate previous box whose window is floored at
    # +-1e-9: there the new extent is capped at 1.1 * 1e-9 around its centre
    prev = np.zeros(6, np.float32)
    b2 = np.zeros((2, 8), np.float32)
    b2[:, 3] = 1
    b2[0, :3] = -1e-9
    b2[1, :3] = 1e-9
    got = O.aabb(b2, np.arange(2, dtype=np.uint32), 2, prev)
    np.testing.assert_allclose(got[3:] - got[:3], 1.1e-9, rtol=1e-5)
    np.testing.assert_allclose(got[3:] + got[:3], 0.0, atol=1e-15)


def test_dead_partition_is_stable():
    rng = np.random.default_rng(3)
    n = 3000
    b = rng.normal(size=(n, 8)).astype(np.float32)
    b[rng.random(n) < 0.3, 3] = 0.0
    idx = rng.permutation(n).astype(np.uint32)
    src = b.copy()
    dst = np.ones_like(b)
    out, died = O.dead_partition(src, dst, idx, n)
    alive = b[idx, 3] != 0
    np.testing.assert_array_equal(out, np.concatenate([idx[alive], idx[~alive]]))
    assert died == (~alive).sum()
    assert (src[idx[~alive]] == 0).all() and (dst[idx[~alive]] == 0).all()
    assert (dst[idx[alive]] == 1).all()


def test_sort_is_stable_sort_by_key():
    rng = np.random.default_rng(4)
    for n in (1, 2, 17, 1000, 70000):
        keys = rng.integers(0, 1 << 63, size=n, dtype=np.uint64)
        keys[rng.random(n) < 0.3] = keys[0]          # many duplicates
        keys[: n // 4] &= np.uint64(0xFF)            # small keys: high digits zero
        idx = rng.permutation(n).astype(np.uint32)
        k, i = O.sort(keys, idx)
        order = np.argsort(keys, kind="stable")
        np.testing.assert_array_equal(k, keys[order])
        np.testing.assert_array_equal(i, idx[order])


def _karras_reference(keys):
    """Brute-force Karras-2012 tree over distinct augmented keys (code, index):
    node covering [lo, hi] splits at the highest differing bit of its range."""
    n = len(keys)
    aug = [(int(k) << 32) | i for i, k in enumerate(keys)]

    def split(lo, hi):
        if aug[lo] == aug[hi]:
            return (lo + hi) >> 1
        prefix = 96 - (aug[lo] ^ aug[hi]).bit_length()
        s = lo
        for j in range(lo, hi):
            if 96 - (aug[lo] ^ aug[j + 1]).bit_length() > prefix:
                s = j + 1
        return s

    ranges = {}
    stack = [(0, n - 1)]
    while stack:
        lo, hi = stack.pop()
        g = split(lo, hi)
        ranges[(lo, hi)] = g
        if g != lo:
            stack.append((lo, g))
        if g + 1 != hi:
            stack.append((g + 1, hi))
    return ranges


@pytest.mark.parametrize("n,dups", [(2, False), (3, False), (37, False), (64, True), (513, True)])
def test_karras_tree_invariants(n, dups):
    rng = np.random.default_rng(n)
    keys = np.sort(rng.integers(0, 1 << 63, size=n, dtype=np.uint64))
    if dups:
        keys = np.sort(np.repeat(keys[: n // 4 + 1], 4)[:n])
    leaf, internal = O.build(keys, n, n)
    f = O.node_fields(internal[: n - 1])
    lf = O.node_fields(leaf)
    # exactly one root: internal 0
    assert f["parentId"][0] == EMPTY
    # every internal node has two children and every node but the root one parent
    claimed = np.zeros(2 * n, int)
    ranges = {}
    for i in range(n - 1):
        for c in (f["childA"][i], f["childB"][i]):
            assert c != EMPTY and c < 2 * n - 1 + 1
            claimed[c] += 1
            parent = lf["parentId"][c] if c < n else f["parentId"][c - n]
            assert parent == i + n
        cnt = int(f["bodiesContained"][i])
        lo = int(f["nodeDepth"][i])               # min(i, j) before the fill
        ranges[(lo, lo + cnt - 1)] = i
    span = {i: r for r, i in ranges.items()}
    assert (claimed[:n] == 1).all()               # every leaf exactly once
    assert (claimed[n + 1: 2 * n - 1] == 1).all()  # every internal except the root once
    assert claimed[n] == 0
    # topology equals the brute-force split-at-highest-differing-bit tree
    want = _karras_reference(keys)
    assert set(want.keys()) == set(ranges.keys())
    for (lo, hi), g in want.items():
        i = ranges[(lo, hi)]
        a, b = int(f["childA"][i]), int(f["childB"][i])
        if g == lo:
            assert a == g
        else:
            assert a >= n and span[a - n] == (lo, g)
        if g + 1 == hi:
            assert b == hi
        else:
            assert b >= n and span[b - n] == (g + 1, hi)


def test_fill_closed_form():
    rng = np.random.default_rng(6)
    n = 777
    b = np.zeros((n, 8), np.float32)
    b[:, :3] = rng.normal(size=(n, 3))
    b[:, 3] = rng.uniform(0.1, 1.0, n)
    b[:, 7] = 1.0
    idx = np.arange(n, dtype=np.uint32)
    codes = O.morton_encode(b, idx, n, np.array([-5, -5, -5, 5, 5, 5], np.float32))
    keys, sidx = O.sort(codes, idx)
    leaf, internal = O.build(keys, n, n)
    h = O.fill(b, sidx, n, n, leaf, internal, 1.0, 1.0, max_passes=0)
    f = O.node_fields(internal[: n - 1])
    # recompute ranges from the topology and compare with float64 sums
    def leaves(node):
        out = []
        st = [node]
        while st:
            x = st.pop()
            if x < n:
                out.append(x)
            else:
                st += [int(f["childA"][x - n]), int(f["childB"][x - n])]
        return out

    def height(node):
        return 0 if node < n else 1 + max(height(int(f["childA"][node - n])), height(int(f["childB"][node - n])))

    assert h == height(n)
    for i in rng.choice(n - 1, 60, replace=False):
        ls = leaves(i + n)
        bb = b[sidx[ls]].astype(np.float64)
        m = bb[:, 3].sum()
        com = (bb[:, :3] * bb[:, 3:4]).sum(0) / m
        assert len(ls) == f["bodiesContained"][i]
        np.testing.assert_allclose(f["mass"][i], m, rtol=1e-5)
        np.testing.assert_allclose(f["com"][i], com, rtol=1e-4, atol=1e-5)
        np.testing.assert_array_equal(f["aabb"][i][:3], b[sidx[ls], :3].min(0))
        np.testing.assert_array_equal(f["aabb"][i][3:], b[sidx[ls], :3].max(0))
        assert f["nodeDepth"][i] == height(i + n)
        assert f["readyChildren"][i] == EMPTY


def test_theta_zero_is_direct_sum():
    rng = np.random.default_rng(7)
    n = 300
    b = np.zeros((n, 8), np.float32)
    b[:, :3] = rng.normal(size=(n, 3))
    b[:, 3] = rng.uniform(0.5, 1.5, n) / n
    b[:, 7] = 1
    idx = np.arange(n, dtype=np.uint32)
    bounds = np.concatenate([b[:, :3].min(0), b[:, :3].max(0)]).astype(np.float32)
    codes = O.morton_encode(b, idx, n, bounds)
    keys, sidx = O.sort(codes, idx)
    leaf, internal = O.build(keys, n, n)
    O.fill(b, sidx, n, n, leaf, internal)
    prm = O.params(theta=0.0, softening=0.01)
    r = O.force(b, sidx, n, n, leaf, internal, prm, 1.0, 1.0, bounds, dst=np.zeros_like(b))
    want = O.direct_sum_f64(b, sidx, 1.0, 0.01)
    err = np.linalg.norm(r["accel"] - want, axis=1) / np.linalg.norm(want, axis=1)
    assert err.max() < 1e-5
    assert r["stats"].interactions == n * (n - 1)


def test_theta_zero_opened_count():
    rng = np.random.default_rng(8)
    n = 64
    b = np.zeros((n, 8), np.float32)
    b[:, :3] = rng.normal(size=(n, 3))
    b[:, 3] = 1.0 / n
    idx = np.arange(n, dtype=np.uint32)
    bounds = np.concatenate([b[:, :3].min(0), b[:, :3].max(0)]).astype(np.float32)
    keys, sidx = O.sort(O.morton_encode(b, idx, n, bounds), idx)
    leaf, internal = O.build(keys, n, n)
    O.fill(b, sidx, n, n, leaf, internal)
    r = O.force(b, sidx, n, n, leaf, internal, O.params(theta=0.0), 1.0, 1.0, bounds)
    # theta = 0 opens every internal node for every body
    assert r["stats"].opened == n * (n - 1)


def test_direct_sum_two_bodies():
    b = np.zeros((2, 8), np.float32)
    b[0, :4] = [0, 0, 0, 2.0]
    b[1, :4] = [3, 4, 0, 5.0]
    a = O.direct_sum_f64(b, np.array([0, 1], np.uint32), G=1.5, softening=0.0)
    np.testing.assert_allclose(a[0], 1.5 * 5.0 * np.array([3, 4, 0]) / 125.0)
    np.testing.assert_allclose(a[1], -1.5 * 2.0 * np.array([3, 4, 0]) / 125.0)


def test_merge_pair_conserves_mass_momentum_volume():
    rng = np.random.default_rng(9)
    for _ in range(100):
        b1 = rng.uniform(0.1, 2.0, 8).astype(np.float32)
        b2 = rng.uniform(0.1, 2.0, 8).astype(np.float32)
        o = np.zeros(8, np.float32)
        O.lib().bho_merge_pair(O._p(b1), O._p(b2), O._p(o))
        m1, m2 = float(b1[3]), float(b2[3])
        np.testing.assert_allclose(o[3], m1 + m2, rtol=1e-6)
        np.testing.assert_allclose(o[4:7] * o[3], b1[4:7] * m1 + b2[4:7] * m2, rtol=1e-5)
        np.testing.assert_allclose(o[:3] * o[3], b1[:3] * m1 + b2[:3] * m2, rtol=1e-5)
        # volume m/rho is additive
        np.testing.assert_allclose(o[3] / o[7], m1 / b1[7] + m2 / b2[7], rtol=1e-5)


def test_merge_sequential_chain():
    d = np.zeros((4, 8), np.float32)
    d[:, 3] = [1, 2, 3, 4]
    d[:, 7] = 1
    d[:, 0] = [0, 1, 2, 3]
    tasks = np.array([[1, 2], [0, 1], [0, 3]], np.uint32)
    applied = O.merge(d, tasks)
    # ascending (a,b): (0,1) absorbs 1; (0,3) absorbs 3; (1,2): 1 is empty -> skipped
    assert applied == 2
    assert d[0, 3] == 7 and d[1, 3] == 0 and d[2, 3] == 3 and d[3, 3] == 0
