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
* With all-zero records the corrected walk accepts the plain walk's nodes and
  gives its accelerations within FP32 rounding.
"""
import numpy as np

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
        corr = O.force_corrected(leaf, internal, 64, 64, q, 0.3, 1e-6)
        assert plain["rc"] == 0 and corr["rc"] == 0
        assert plain["interactions"][t] == 1 and corr["interactions"][t] == 1   # one node: the whole cluster
        ep = np.linalg.norm(plain["accel"][t] - exact) / np.linalg.norm(exact)
        eq = np.linalg.norm(corr["accel"][t] - exact) / np.linalg.norm(exact)
        errs.append((ep, eq))
    (p1, q1), (p2, q2) = errs
    assert q1 < 0.25 * p1 and q2 < 0.15 * p2
    assert 0.15 < p2 / p1 < 0.35          # plain: relative error ~ (b / s)^2
    assert 0.07 < q2 / q1 < 0.18          # corrected: ~ (b / s)^3


def test_zero_records_give_the_plain_walk():
    """With zero records the corrected walk's opening test (theta^2 s^2 >
    half^2 in FP32) accepts the nodes of the reference's (half / s < theta)
    except at ties, and its FP64 sum is the plain walk's FP32 sum within the
    FP32 rounding of the terms."""
    rng = np.random.default_rng(5)
    n = 3000
    b = np.zeros((n, 8), np.float32)
    b[:, :3] = rng.standard_normal((n, 3))
    b[:, 3] = 1.0 / n
    sidx, leaf, internal, bounds = _tree(b)
    prm = O.params(theta=0.3, dt=1.0, softening=0.01)
    plain = O.force(b, sidx, n, n, leaf, internal, prm, 1.0, 1.0, bounds, stack_cap=256)
    zero = O.force_corrected(leaf, internal, n, n, np.zeros((n - 1, 8), np.float32), 0.3, 0.01)
    assert zero["rc"] == 0 and zero["quad_terms"] > 0
    same = plain["interactions"] == zero["interactions"]
    assert same.mean() >= 0.999
    err = np.linalg.norm(plain["accel"].astype(np.float64) - zero["accel"], axis=1)
    assert (err[same] <= 1e-5 * zero["absum"][same]).all()
    assert np.median(err / np.linalg.norm(zero["accel"], axis=1)) < 1e-6
    # and the records reduce the force error against the FP64 direct sum
    q = _records(O.node_moments(leaf, internal, n, n))
    corr = O.force_corrected(leaf, internal, n, n, q, 0.3, 0.01)
    samp = np.arange(0, n, 7)
    exact = O.direct_sum_f64(b, sidx[samp].astype(np.uint32), G=1.0, softening=0.01)
    ep = np.linalg.norm(plain["accel"][samp] - exact, axis=1) / np.linalg.norm(exact, axis=1)
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
    r = O.force_corrected(leaf, internal, 2, 2, np.zeros((1, 8), np.float32), 0.5, 0.0 + 1e-12)
    assert list(r["interactions"]) == [1, 1] and r["opened"] == 2 and r["quad_terms"] == 0
    a = r["accel"][np.argsort(sidx)]            # by body id
    s3 = (1.0 + 1e-12) ** 1.5
    np.testing.assert_allclose(a[0], [3.0 / s3, 0, 0], rtol=1e-6)
    np.testing.assert_allclose(a[1], [-1.0 / s3, 0, 0], rtol=1e-6)
