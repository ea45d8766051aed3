"""The oracle's box-aware walk (the learned-correction method's acceptance
test: a node is accepted only if half its longest side * inv < theta and the
target lies outside the node's AABB), pinned two ways:

* below theta = 1 / (2 sqrt 3) with no softening a node that contains the
  target always fails the angle test (|r| <= the box diagonal), so the walk
  equals the plain walk bit for bit;
* at theta = 0.9 an independent float32 numpy walk in the same order gives
  the same accelerations bit for bit, and differs from the plain walk.
"""
import numpy as np

from oracle import bh_oracle as O


def _tree(n, seed):
    rng = np.random.default_rng(seed)
    b = np.zeros((n, 8), np.float32)
    r = rng.standard_normal((n, 3)) * (1.0 / np.sqrt(3.0))
    b[:, :3] = r
    b[:, 3] = 1.0 / n
    idx = np.arange(n, dtype=np.uint32)
    bounds = np.concatenate([b[:, :3].min(0), b[:, :3].max(0)]).astype(np.float32)
    keys, sidx = O.sort(O.morton_encode(b, idx, n, bounds), idx)
    leaf, internal = O.build(keys, n, n)
    O.fill(b, sidx, n, n, leaf, internal, max_passes=0, with_colors=False)
    return b, sidx, leaf, internal, bounds


def _numpy_walk(b, sidx, leaf, internal, n, g, theta, soft):
    lf, fi = O.node_fields(leaf), O.node_fields(internal)
    f32 = np.float32
    x = b[sidx[g], :3]
    acc = np.zeros(3, np.float32)
    st = [n]
    while st:
        k = st.pop()
        com = lf["com"][k] if k < n else fi["com"][k - n]
        m = lf["mass"][k] if k < n else fi["mass"][k - n]
        r = (com - x).astype(np.float32)
        d2 = f32(r[0] * r[0] + f32(r[1] * r[1] + r[2] * r[2])) + f32(soft)
        inv = f32(1.0) / np.sqrt(f32(d2))
        if k >= n:
            bx = fi["aabb"][k - n]
            ext = bx[3:] - bx[:3]
            half = f32(max(ext[0], max(ext[1], ext[2])) / f32(2.0))
            inside = bool(np.all(x >= bx[:3]) and np.all(x <= bx[3:]))
            if half * inv < f32(theta) and not inside:
                acc = acc + f32(m) * r * inv * inv * inv
            else:
                st += [int(fi["childA"][k - n]), int(fi["childB"][k - n])]
        elif k != g:
            acc = acc + f32(m) * r * inv * inv * inv
    return acc


def test_boxmac_equals_plain_walk_below_the_containment_angle():
    n = 4096
    b, sidx, leaf, internal, bounds = _tree(n, 3)
    prm = O.params(theta=0.28, softening=0.0)
    plain = O.force(b, sidx, n, n, leaf, internal, prm, 1.0, 1.0, bounds, stack_cap=256)
    box = O.force_boxmac(b, sidx, n, n, leaf, internal, prm, 1.0, 1.0, bounds)
    assert plain["rc"] == 0 and box["rc"] == 0
    np.testing.assert_array_equal(box["accel"].view(np.uint32), plain["accel"].view(np.uint32))


def test_boxmac_matches_numpy_walk_at_wide_angle():
    n = 2048
    soft = 0.01
    b, sidx, leaf, internal, bounds = _tree(n, 4)
    prm = O.params(theta=0.9, softening=soft)
    box = O.force_boxmac(b, sidx, n, n, leaf, internal, prm, 1.0, 1.0, bounds)
    plain = O.force(b, sidx, n, n, leaf, internal, prm, 1.0, 1.0, bounds, stack_cap=256)
    assert box["rc"] == 0
    # the box test opens nodes the plain walk accepts
    assert box["stats"].opened > plain["stats"].opened
    differ = np.any(box["accel"] != plain["accel"], axis=1)
    assert differ.sum() > n // 10
    for g in [0, 1, 77, 500, 1023, 1500, n - 1]:
        want = _numpy_walk(b, sidx, leaf, internal, n, g, 0.9, soft)
        np.testing.assert_array_equal(box["accel"][g].view(np.uint32), want.view(np.uint32))
