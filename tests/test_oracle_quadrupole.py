"""The oracle's corrected walk (no reference counterpart): node second moments
and the quadrupole term, pinned by closed forms and by physics.

* Second moments: two and three bodies in closed form; every internal node's
  moment equals its children's moments moved to its centre of mass (the
  parallel-axis theorem).
* The quadrupole term: for a cluster of bodies seen from a distant target,
  the exact field minus the monopole is (G / s^4) B(u) M2 to leading order,
  and the corrected walk's error falls one order faster with distance than
  the plain walk's (a wrong coefficient, sign or trace term would leave an
  error of the monopole's order).
* The corrected walk's accepted set (opening test b < theta s, b the
  distance from the centre of mass to the farthest box corner) against an
  independent top-down restatement: same counts, same FP64 monopole sums.
"""
import numpy as np
import pytest

from oracle import bh_oracle as O


def _tree(b):
    n = b.shape[0]
    idx = np.arange(n, dtype=np.uint32)
    bounds = np.concatenate([b[:, :3].min(0), b[:, :3].max(0)]).astype(np.float32)
    keys, sidx = O.sort(O.morton_encode(b, idx, n, bounds), idx)
    leaf, internal = O.build(keys, n, n)
    O.fill(b, sidx, n, n, leaf, internal, max_passes=0, with_colors=False)
    return sidx, leaf, internal, bounds


def _records(m2, G=1.0):
    q = np.zeros((m2.shape[0], 8), np.float32)
    q[:, 0:6] = (G * m2).astype(np.float32)
    q[:, 6] = (-1.5 * G * (m2[:, 0] + m2[:, 3] + m2[:, 5])).astype(np.float32)
    return q


def test_moments_closed_forms():
    b = np.zeros((2, 8), np.float32)
    b[0, :4] = [-0.5, 0.0, 0.0, 2.0]
    b[1, :4] = [0.5, 0.0, 0.0, 2.0]
    _, leaf, internal, _ = _tree(b)
    m2 = O.node_moments(leaf, internal, 2, 2)
    np.testing.assert_allclose(m2[0], [1.0, 0, 0, 0, 0, 0], atol=1e-12)   # 2 m d^2 = 2 * 2 * 0.25
    # three bodies on a line, masses 1, 2, 1 at -1, 0, 1 along y: centre 0, M2_yy = 2
    b = np.zeros((3, 8), np.float32)
    b[:, 1] = [-1.0, 0.0, 1.0]
    b[:, 3] = [1.0, 2.0, 1.0]
    _, leaf, internal, _ = _tree(b)
    root = [k for k in range(2) if O.node_fields(internal)["bodiesContained"][k] == 3][0]
    m2 = O.node_moments(leaf, internal, 3, 3)
    np.testing.assert_allclose(m2[root], [0, 0, 0, 2.0, 0, 0], atol=1e-12)


def test_moments_parallel_axis():
    rng = np.random.default_rng(4)
    n = 2000
    b = np.zeros((n, 8), np.float32)
    b[:, :3] = rng.standard_normal((n, 3)) * [1.0, 0.5, 0.2]
    b[:, 3] = rng.uniform(0.5, 1.5, n) / n
    _, leaf, internal, _ = _tree(b)
    m2 = O.node_moments(leaf, internal, n, n)
    f = O.node_fields(internal)
    lf = O.node_fields(leaf)
    sym = lambda v: np.array([[v[0], v[1], v[2]], [v[1], v[3], v[4]], [v[2], v[4], v[5]]])   # noqa: E731
    for k in rng.choice(n - 1, 200, replace=False):
        c = f["com"][k].astype(np.float64)
        tot = np.zeros((3, 3))
        for ch in (int(f["childA"][k]), int(f["childB"][k])):
            if ch < n:
                pc, mc, mm = lf["com"][ch].astype(np.float64), float(lf["mass"][ch]), np.zeros((3, 3))
            else:
                pc, mc, mm = f["com"][ch - n].astype(np.float64), float(f["mass"][ch - n]), sym(m2[ch - n])
            d = pc - c
            tot += mm + mc * np.outer(d, d)
        # (the parent's centre of mass is rounded to FP32, which leaves a dipole of ~1e-7)
        np.testing.assert_allclose(sym(m2[k]), tot, rtol=1e-5, atol=1e-6 * np.abs(tot).max())


def _cluster(rng, k, size):
    p = rng.standard_normal((k, 3)) * [size, 0.6 * size, 0.3 * size]
    m = rng.uniform(0.5, 1.5, k)
    return p, m / m.sum()


def test_quadrupole_term_is_the_leading_residual():
    """exact - monopole = (G / s^4) B(u) M2 + O(s^-5)"""
    rng = np.random.default_rng(2)
    p, m = _cluster(rng, 40, 0.05)
    c = (m[:, None] * p).sum(0)
    d = p - c
    M2 = np.einsum("k,ki,kj->ij", m, d, d)
    m2v = np.array([M2[0, 0], M2[0, 1], M2[0, 2], M2[1, 1], M2[1, 2], M2[2, 2]])
    eps2 = 1e-4
    direction = np.array([0.6, -0.48, 0.64])
    rel = []
    for D in (1.0, 2.0, 4.0):
        x = c - D * direction                 # the target; r = c - x points along direction
        r = c - x
        s2 = r @ r + eps2
        rr = p - x
        exact = (m[:, None] * rr / ((rr * rr).sum(1) + eps2)[:, None] ** 1.5).sum(0)
        mono = r / s2 ** 1.5
        quad = O.mlp_quad_basis(r / np.sqrt(s2), m2v) / s2 ** 2
        rel.append((np.linalg.norm(exact - mono - quad), np.linalg.norm(exact - mono)))
    for (e_q, e_m) in rel:
        assert e_q < 0.2 * e_m
    # the residual after the quadrupole falls one order faster than after the monopole
    for (q0, m0), (q1, m1) in zip(rel, rel[1:]):
        assert 0.35 < (q1 / m1) / (q0 / m0) < 0.65


def test_corrected_walk_error_falls_one_order_faster():
    """A cluster and one distant target body: at theta 0.3 the target's walk
    opens the root (which holds the target itself) and accepts the cluster's
    node; its error with the quadrupole records falls as s^-5, the plain
    walk's as s^-4."""
    rng = np.random.default_rng(3)
    p, m = _cluster(rng, 63, 0.05)
    direction = np.array([0.36, 0.48, 0.8])
    errs = []
    for D in (2.0, 4.0):
        b = np.zeros((64, 8), np.float32)
        b[:63, :3] = p
        b[:63, 3] = m
        b[63, :3] = -D * direction
        b[63, 3] = 1e-6
        b = b.astype(np.float32)
        sidx, leaf, internal, bounds = _tree(b)
        q = _records(O.node_moments(leaf, internal, 64, 64))
        prm = O.params(theta=0.3, dt=1.0, softening=1e-6)
        t = int(np.nonzero(sidx == 63)[0][0])
        exact = O.direct_sum_f64(b, np.array([63], np.uint32), G=1.0, softening=1e-6)[0]
        plain = O.force(b, sidx, 64, 64, leaf, internal, prm, 1.0, 1.0, bounds, stack_cap=256)
        corr = O.force_corrected(leaf, internal, 64, 64, q, 0.3, 1e-6, group=1)
        assert plain["rc"] == 0 and corr["rc"] == 0
        assert plain["interactions"][t] == 1 and corr["interactions"][t] == 1   # one node: the whole cluster
        ep = np.linalg.norm(plain["accel"][t] - exact) / np.linalg.norm(exact)
        eq = np.linalg.norm(corr["accel"][t] - exact) / np.linalg.norm(exact)
        errs.append((ep, eq))
    (p1, q1), (p2, q2) = errs
    assert q1 < 0.25 * p1 and q2 < 0.15 * p2
    assert 0.15 < p2 / p1 < 0.35          # plain: relative error ~ (b / s)^2
    assert 0.07 < q2 / q1 < 0.18          # corrected: ~ (b / s)^3


def _accepted_sets(leaf, internal, n, theta, soft, group, amag=None, alpha=0.0):
    """The corrected walk's taken nodes per body, top-down over the tree for
    all bodies at once (an independent restatement of the traversal): the
    bodies of a group of `group` consecutive sorted positions reach the same
    nodes; they take an internal node if every one of them accepts it (the
    geometric test fl(s^2 theta^2) > b^2, or the relative s^2 > b^2 and
    fl(s^2 A) > K, in FP32 with fused multiply-adds emulated in FP64), else
    all reach both children; every leaf reached is taken but a body's own."""
    f32 = np.float32
    lf = O.node_fields(leaf)
    fi = O.node_fields(internal)
    p = lf["com"].astype(np.float32)
    th2 = f32(f32(theta) * f32(theta))
    gid = np.arange(n) // group
    ngroups = int(gid[-1]) + 1
    if amag is not None:
        A = np.power(float(alpha) * amag.astype(np.float64), 0.4).astype(np.float32)

    def fma(a, b, c):
        return (np.asarray(a, np.float64) * b + c).astype(np.float32)

    reach = {n: np.ones(n, bool)}
    taken = []          # (node id, mask of bodies)
    order = [n]
    while order:
        k = order.pop()
        m = reach.pop(k)
        if k < n:
            m = m.copy()
            m[k] = False
            taken.append((k, m))
            continue
        c = fi["com"][k - n].astype(np.float32)
        bb = fi["aabb"][k - n].astype(np.float32)
        d = np.maximum(c - bb[0:3], bb[3:6] - c).astype(np.float32)
        b2 = fma(d[0], d[0], fma(d[1], d[1], f32(d[2] * d[2])))
        r = (c[None, :] - p).astype(np.float32)
        s2 = fma(r[:, 0], r[:, 0], fma(r[:, 1], r[:, 1], fma(r[:, 2], r[:, 2], np.full(n, soft, np.float32))))
        if amag is None:
            ok = (s2 * th2).astype(np.float32) > b2
        else:
            gm = f32(f32(1.0) * f32(fi["mass"][k - n]))
            K = f32(np.power(float(gm), 0.4) * np.power(float(b2), 0.6))
            ok = (s2 > b2) & ((s2 * A).astype(np.float32) > K)
        fails = np.bincount(gid[m & ~ok], minlength=ngroups)
        all_ok = (fails == 0)[gid]
        taken.append((k, m & all_ok))
        rest = m & ~all_ok
        if rest.any():
            for ch in (int(fi["childA"][k - n]), int(fi["childB"][k - n])):
                reach[ch] = reach[ch] | rest if ch in reach else rest
                order.append(ch)
    return taken, lf, fi


@pytest.mark.parametrize("group", [1, 64])
@pytest.mark.parametrize("relative", [False, True])
def test_corrected_walk_accepts_what_the_traversal_restates(group, relative):
    """With zero records the oracle's corrected walk is the FP64 monopole sum
    over the nodes taken; the same set built independently top-down gives
    the same counts per body and the same sums; every body's nodes cover the
    other bodies once (or all of them, where a node holding the body itself
    was taken)."""
    rng = np.random.default_rng(5)
    n = 3000
    b = np.zeros((n, 8), np.float32)
    b[:, :3] = rng.standard_normal((n, 3))
    b[:, 3] = rng.uniform(0.5, 1.5, n) / n
    sidx, leaf, internal, bounds = _tree(b)
    theta, soft, alpha = 0.45, 0.01, 2e-3
    amag = rng.uniform(0.2, 2.0, n).astype(np.float32) if relative else None
    zero = O.force_corrected(leaf, internal, n, n, np.zeros((n - 1, 8), np.float32), theta, soft, amag=amag,
                             alpha=alpha, group=group)
    assert zero["rc"] == 0 and zero["quad_terms"] > 0
    taken, lf, fi = _accepted_sets(leaf, internal, n, theta, soft, group, amag, alpha)
    count = np.zeros(n, np.int64)
    covered = np.zeros(n, np.int64)
    acc = np.zeros((n, 3))
    p = lf["com"].astype(np.float64)
    for k, m in taken:
        if k < n:
            com, M, nb = p[k], float(lf["mass"][k]), 1
        else:
            com, M, nb = fi["com"][k - n].astype(np.float64), float(fi["mass"][k - n]), int(fi["bodiesContained"][k - n])
        count += m
        covered += m * nb
        r = (com.astype(np.float32)[None, :] - p.astype(np.float32)).astype(np.float64)
        s2 = (r * r).sum(1) + float(np.float32(soft))
        acc += m[:, None] * (float(np.float32(M)) * r / s2[:, None] ** 1.5)
    np.testing.assert_array_equal(count, zero["interactions"])
    assert ((covered == n - 1) | (covered == n)).all()
    np.testing.assert_allclose(zero["accel"], acc, rtol=1e-12, atol=1e-12 * np.abs(acc).max())
    if group == 64 and not relative:
        # a group takes no node that one of its bodies alone would open
        solo = O.force_corrected(leaf, internal, n, n, np.zeros((n - 1, 8), np.float32), theta, soft, group=1)
        assert (zero["interactions"] >= solo["interactions"]).all()
    if relative:
        return          # (|a| here is random: the gain below is the geometric test's)
    # and the records reduce the force error of the same walk against the FP64 direct sum
    q = _records(O.node_moments(leaf, internal, n, n))
    corr = O.force_corrected(leaf, internal, n, n, q, theta, soft, group=group)
    samp = np.arange(0, n, 7)
    exact = O.direct_sum_f64(b, sidx[samp].astype(np.uint32), G=1.0, softening=soft)
    ep = np.linalg.norm(zero["accel"][samp] - exact, axis=1) / np.linalg.norm(exact, axis=1)
    eq = np.linalg.norm(corr["accel"][samp] - exact, axis=1) / np.linalg.norm(exact, axis=1)
    assert np.percentile(eq, 50) < 0.33 * np.percentile(ep, 50)
    assert np.percentile(eq, 99) < 0.5 * np.percentile(ep, 99)


def test_corrected_walk_self_and_counts():
    """Two bodies: each walk opens the root and takes the other's leaf (never
    its own); with records, an accepted node's term is the closed form of the
    monopole plus quadrupole."""
    b = np.zeros((2, 8), np.float32)
    b[0, :4] = [0.0, 0.0, 0.0, 1.0]
    b[1, :4] = [1.0, 0.0, 0.0, 3.0]
    sidx, leaf, internal, _ = _tree(b)
    r = O.force_corrected(leaf, internal, 2, 2, np.zeros((1, 8), np.float32), 0.5, 0.0 + 1e-12, group=1)
    assert list(r["interactions"]) == [1, 1] and r["opened"] == 2 and r["quad_terms"] == 0
    a = r["accel"][np.argsort(sidx)]            # by body id
    s3 = (1.0 + 1e-12) ** 1.5
    np.testing.assert_allclose(a[0], [3.0 / s3, 0, 0], rtol=1e-6)
    np.testing.assert_allclose(a[1], [-1.0 / s3, 0, 0], rtol=1e-6)
