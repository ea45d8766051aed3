"""The learned far-field correction on the GPU: network kernels against the
FP64 oracle (forward, loss, gradient), BF16 MFMA inference, FP64 labels
against a CPU recomputation, the corrected walk with a zero network equal to
the plain walk bit for bit, and a trained network reducing the force error."""
import numpy as np
import pytest

from oracle import bh_oracle as O

pytestmark = pytest.mark.gpu


def _data(rng, n):
    x = rng.normal(size=(n, 12)).astype(np.float32)
    t = np.stack([np.sin(x[:, 0]) * x[:, 6], np.tanh(x[:, 1] - x[:, 2]), 0.3 * x[:, 3] * x[:, 4]], 1).astype(np.float32)
    return x, t


def test_mlp_forward_loss_grad_match_fp64_oracle():
    from gravity_amd import learned as Lr
    print("[gpu] mlp fp32 vs oracle", flush=True)
    rng = np.random.default_rng(0)
    p = (rng.normal(size=Lr.num_params()) * 0.2).astype(np.float32)
    x, t = _data(rng, 1000)
    m = Lr.Mlp(params=p)
    y = m.forward(x)
    y64 = O.mlp_forward(p, x)
    np.testing.assert_allclose(y, y64, rtol=1e-4, atol=1e-5)
    loss, g = m.loss_grad(x, t)
    loss64, g64 = O.mlp_loss_grad(p, x, t)
    np.testing.assert_allclose(loss, loss64, rtol=1e-5)
    assert np.linalg.norm(g - g64) <= 1e-4 * np.linalg.norm(g64)
    np.testing.assert_allclose(g, g64, rtol=1e-3, atol=1e-6 * np.abs(g64).max())
    m.close()


def test_mlp_mfma_forward_matches_fp64_within_bf16():
    from gravity_amd import learned as Lr
    print("[gpu] mlp mfma vs oracle", flush=True)
    rng = np.random.default_rng(1)
    p = (rng.normal(size=Lr.num_params()) * 0.2).astype(np.float32)
    x, _ = _data(rng, 777)      # not a multiple of 64: padding rows
    m = Lr.Mlp(params=p)
    y = m.forward(x, mfma=True)
    y64 = O.mlp_forward(p, x)
    err = np.abs(y - y64)
    scale = np.abs(y64).mean()
    assert err.max() <= 0.05 * scale + 1e-3, (err.max(), scale)
    assert np.sqrt((err ** 2).mean()) <= 0.01 * scale + 1e-4
    m.close()


def test_mlp_training_reduces_loss():
    from gravity_amd import learned as Lr
    print("[gpu] mlp training", flush=True)
    rng = np.random.default_rng(2)
    x, t = _data(rng, 20000)
    m = Lr.Mlp(params=Lr.init_params(3))
    l0, _ = m.loss_grad(x, t)
    m.train(x, t, epochs=30, batch=1024, lr=3e-3, seed=1)
    l1, _ = m.loss_grad(x, t)
    assert l1 < 0.1 * l0, (l0, l1)
    m.close()


def test_pair_labels_match_cpu_fp64():
    from gravity_amd import barnes_hut as B
    from gravity_amd import initial_conditions as ic
    from gravity_amd import learned as Lr
    print("[gpu] pair labels", flush=True)
    n = 1 << 14
    soft = 0.01
    bodies = ic.plummer(n, 4)
    bh = B.BarnesHut(bodies, ic.tight_bounds(bodies), B.make_params(theta=0.9, softening=soft, G=1.0))
    bh.init()
    bh.build_current_tree()
    f, t, pr = Lr.collect_training_pairs(bh, 0.9, 0.5, 7, soft, with_pairs=True)
    assert f.shape[0] > 1000
    assert np.abs(f[:, 6:9]).max() <= 1.0 + 1e-6
    assert (f[:, 11] > 0).all() and (f[:, 11] <= 1.0).all()
    lf = O.node_fields(bh.nodes(B.LEAF_NODES))
    fi = O.node_fields(bh.nodes(B.INTERNAL_NODES))
    pos = lf["com"].astype(np.float64)
    mass = lf["mass"].astype(np.float64)
    rng = np.random.default_rng(0)
    for q in rng.choice(f.shape[0], 40, replace=False):
        g, k = int(pr[q, 0]), int(pr[q, 1])
        assert g % 7 == 0
        leaves, st = [], [k + n]
        while st:
            x = st.pop()
            if x < n:
                leaves.append(x)
            else:
                st += [int(fi["childA"][x - n]), int(fi["childB"][x - n])]
        leaves = np.array(leaves)
        assert leaves.size == fi["bodiesContained"][k]
        others = leaves[leaves != g]
        r = pos[others] - pos[g]
        d2 = (r * r).sum(1) + soft
        a = (mass[others, None] * r / d2[:, None] ** 1.5).sum(0)
        c = fi["com"][k].astype(np.float64) - pos[g]
        ds2 = (c * c).sum() + soft
        M = float(fi["mass"][k])
        want = a * ds2 / M - c / np.sqrt(ds2)
        np.testing.assert_allclose(t[q], want, rtol=1e-4, atol=2e-5)
        # features: direction and softening share
        np.testing.assert_allclose(f[q, 6:9], c / np.sqrt(ds2), rtol=1e-5, atol=1e-6)
        np.testing.assert_allclose(f[q, 11], soft / ds2, rtol=1e-5)
    bh.close()


@pytest.mark.parametrize("general", [False, True])
def test_corrected_walk_with_zero_network_equals_box_aware_oracle_walk(general):
    """With a zero network the corrected walk is the box-aware walk: bit for
    bit the oracle's walk that opens every node containing its target, from
    the lean loop and from the general loop alike."""
    from gravity_amd import barnes_hut as B
    from gravity_amd import initial_conditions as ic
    from gravity_amd import learned as Lr
    print(f"[gpu] zero-network corrected walk, general loop {general}", flush=True)
    n = 1 << 16
    soft = 0.01
    bodies = ic.plummer(n, 5)
    bounds = ic.tight_bounds(bodies)
    prm = B.make_params(theta=0.9, softening=soft, G=1.0)
    bh = B.BarnesHut(bodies, bounds, prm)
    bh.init()
    m = Lr.Mlp(params=np.zeros(Lr.num_params(), np.float32))
    Lr.attach(bh, m, 0.5)
    lp = B.Params.from_buffer_copy(prm)
    lp.flags = B.FLAG_LEARNED | (B.FLAG_GENERAL_WALK if general else 0)
    corr = bh.accelerations(lp)
    c = bh.counters()
    assert c["corrected"] > 0
    idx = bh.index()
    leaf = bh.nodes(B.LEAF_NODES)
    internal = bh.nodes(B.INTERNAL_NODES)
    r = O.force_boxmac(bodies, idx, n, n, leaf, internal, O.params(theta=0.9, dt=1.0, softening=soft), 1.0, 1.0,
                       bh.bounds(), threads=16)
    assert r["rc"] == 0
    assert c["interactions"] == int(r["interactions"].astype(np.int64).sum())
    np.testing.assert_array_equal(corr[idx].view(np.uint32), r["accel"].view(np.uint32))
    bh.close()
    m.close()


def test_trained_network_reduces_force_error():
    """What the trained network gains: on a Plummer sphere not seen in
    training, the corrected walk at theta 0.9 has a lower p50 and p99 force
    error against the FP64 direct sum than the same walk without it."""
    from gravity_amd import barnes_hut as B
    from gravity_amd import initial_conditions as ic
    from gravity_amd import learned as Lr
    print("[gpu] trained network gain", flush=True)
    soft = 0.01
    xs, ts = [], []
    for seed in (21, 22):
        b = ic.plummer(1 << 17, seed)
        bh = B.BarnesHut(b, ic.tight_bounds(b), B.make_params(theta=0.9, softening=soft, G=1.0))
        bh.init()
        bh.build_current_tree()
        f, t = Lr.collect_training_pairs(bh, 0.9, 0.5, 3, soft)
        xs.append(f)
        ts.append(t)
        bh.close()
    x, t = np.concatenate(xs), np.concatenate(ts)
    m = Lr.Mlp(params=Lr.init_params(0))
    m.train(x, t, epochs=6, batch=4096, lr=2e-3, seed=0)
    n = 1 << 17
    b = ic.plummer(n, 23)
    bh = B.BarnesHut(b, ic.tight_bounds(b), B.make_params(theta=0.9, softening=soft, G=1.0))
    bh.init()
    sample = np.random.default_rng(0).choice(n, 512, replace=False).astype(np.uint32)
    exact = bh.direct_sum_f64(sample, soft)
    zero = Lr.Mlp(params=np.zeros(Lr.num_params(), np.float32))
    errs = {}
    for name, net in (("without", zero), ("with", m)):
        Lr.attach(bh, net, 0.5)
        lp = B.make_params(theta=0.9, softening=soft, G=1.0, flags=B.FLAG_LEARNED)
        a = bh.accelerations(lp)[sample].astype(np.float64)
        e = np.linalg.norm(a - exact, axis=1) / np.linalg.norm(exact, axis=1)
        errs[name] = (np.percentile(e, 50), np.percentile(e, 99))
    print("[gpu] force error p50/p99 without", errs["without"], "with", errs["with"], flush=True)
    assert errs["with"][0] < 0.7 * errs["without"][0]
    assert errs["with"][1] < errs["without"][1]
    bh.close()
    m.close()
    zero.close()
