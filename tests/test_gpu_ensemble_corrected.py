"""The corrected walk over an ensemble's trees on the GPU.

Each simulation of a corrected ensemble is walked from its own root with the
single-simulation corrected walk's opening tests and wave-wide rule, its 64
lanes taken from one simulation only. Every simulation's accelerations are
compared with the oracle's corrected walk (FP64 sums over the nodes that the
GPU's FP32 tests take, in groups of 64 sorted positions of the simulation)
on the tree the oracle rebuilds from the ensemble's sorted order: the counts
summed over the simulations exactly, each body's error at most 2e-5 of the
sum of its terms' magnitudes and the median relative error below 2e-6.
Covered: the geometric test with the exact quadrupole, the relative test
(the first step geometric, the second reading the first step's |a|), and a
network whose last layer is only a bias b (records M2 + M h^2 b). The
autotuner must leave a relative ensemble's state unchanged.
"""
import numpy as np
import pytest

from oracle import bh_oracle as O

pytestmark = pytest.mark.gpu

SOFT = 0.01
THETA = 0.5
DT = 1.0 / 1024     # a power of two: a = v / dt exactly while v starts at zero
SIZES = [5000, 20000, 2, 3333, 65, 4096]


def _sims(seed):
    from gravity_amd import initial_conditions as ic
    out = []
    for k, n in enumerate(SIZES):
        b = ic.plummer(n, seed * 100 + k)
        b[:, 4:7] = 0.0
        out.append(b)
    return out


def _params(flags, alpha=0.0):
    from gravity_amd import barnes_hut as B
    return B.make_params(theta=THETA, dt=DT, softening=SOFT, G=1.0, flags=flags, force_tolerance=alpha)


def _f32_norm(a):
    """|a| as the walk stores it: sqrt(fma(x, x, fma(y, y, z z))) in FP32."""
    a = a.astype(np.float32)
    zz = (a[:, 2] * a[:, 2]).astype(np.float32)
    yz = (a[:, 1].astype(np.float64) * a[:, 1] + zz).astype(np.float32)
    s = (a[:, 0].astype(np.float64) * a[:, 0] + yz).astype(np.float32)
    return np.sqrt(s)


def _oracle_tree(bodies, keys, idx, ens_internal):
    """The oracle's tree over the ensemble's sorted order of one simulation;
    its internal nodes' geometry and topology must equal the ensemble's."""
    n = bodies.shape[0]
    leaf, internal = O.build(keys, n, n)
    O.fill(np.ascontiguousarray(bodies, np.float32), np.ascontiguousarray(idx, np.uint32), n, n, leaf, internal)
    np.testing.assert_array_equal(internal[: n - 1, 0:12], ens_internal[:, 0:12])
    return leaf, internal


def _records(leaf, internal, n, bias=None):
    """The oracle's node records: G T (xx xy xz yy yz zz), -1.5 G tr T with
    T = M2, or M2 + M h^2 b for a bias-only network (G = 1)."""
    m2 = O.node_moments(leaf, internal, n, n, threads=16)
    if bias is not None:
        f = O.node_fields(internal[: n - 1])
        ext = (f["aabb"][:, 3:6] - f["aabb"][:, 0:3]).astype(np.float32)
        h = (np.max(ext, axis=1) / np.float32(2)).astype(np.float64)
        m2 = m2 + (f["mass"].astype(np.float64) * h * h)[:, None] * bias[None, :].astype(np.float64)
    q = np.zeros((n - 1, 8), np.float32)
    q[:, 0:6] = m2
    q[:, 6] = -1.5 * (m2[:, 0] + m2[:, 3] + m2[:, 5])
    return q


def _check(ens, before, after, amag_by_body=None, alpha=0.0, bias=None):
    """The last step of `ens` against the oracle: `before` the bodies the
    step started from, `after` those it left (accelerations = dv / dt)."""
    srt = ens.sorted()
    nodes = ens.internal_nodes()
    c = ens.counters()
    assert c["error_flags"] == 0, c
    inter = quads = opened = 0
    worst, meds = 0.0, []
    for s, n in enumerate(SIZES):
        keys, idx = srt[s]
        leaf, internal = _oracle_tree(before[s], keys, idx, nodes[s])
        q = _records(leaf, internal, n, bias)
        amag = None if amag_by_body is None else amag_by_body[s][idx]
        r = O.force_corrected(leaf, internal, n, n, q, THETA, SOFT, amag=amag, alpha=alpha, group=64, threads=16)
        assert r["rc"] == 0, s
        inter += int(r["interactions"].astype(np.int64).sum())
        quads += r["quad_terms"]
        opened += r["opened"]
        dv = after[s][:, 4:7].astype(np.float64) - before[s][:, 4:7].astype(np.float64)
        a = dv[idx] / DT
        err = np.linalg.norm(a - r["accel"], axis=1)
        worst = max(worst, float((err / r["absum"]).max()))
        meds.append(err / np.linalg.norm(r["accel"], axis=1))
    med = float(np.median(np.concatenate(meds)))
    print(f"[gpu] ensemble corrected walk error: max {worst:.3g} of the term sum, median relative {med:.3g}",
          flush=True)
    assert c["interactions"] == inter and c["quad_terms"] == quads and c["opened"] == opened
    assert worst <= 2e-5 and med <= 2e-6


def test_ensemble_corrected_walk_geometric():
    from gravity_amd import barnes_hut as B
    from gravity_amd import ensemble as E
    print("[gpu] ensemble corrected walk, geometric test", flush=True)
    sims = _sims(1)
    ens = E.Ensemble(sims, _params(B.FLAG_QUADRUPOLE), config=E.Config(8, 4, 192, 0))
    ens.step(1)
    _check(ens, sims, ens.bodies())
    assert ens.counters()["corrected"] == 0
    ens.close()


@pytest.mark.parametrize("use_graph", [0, 1])
def test_ensemble_corrected_walk_relative(use_graph):
    from gravity_amd import barnes_hut as B
    from gravity_amd import ensemble as E
    print(f"[gpu] ensemble corrected walk, relative test, graph {use_graph}", flush=True)
    alpha = 1e-3
    sims = _sims(2)
    ens = E.Ensemble(sims, _params(B.FLAG_QUADRUPOLE | B.FLAG_RELATIVE_MAC, alpha),
                     config=E.Config(8, 4, 192, use_graph))
    ens.step(1)          # the first step: the geometric test at theta
    first = ens.bodies()
    _check(ens, sims, first)
    amag = [_f32_norm(b[:, 4:7].astype(np.float64) / DT) for b in first]
    ens.step(1)          # the relative test on the first step's |a|
    _check(ens, first, ens.bodies(), amag_by_body=amag, alpha=alpha)
    ens.close()


def test_ensemble_learned_records_with_bias_network():
    from gravity_amd import barnes_hut as B
    from gravity_amd import ensemble as E
    from gravity_amd import learned as Lr
    print("[gpu] ensemble learned walk, bias-only network", flush=True)
    p = (np.random.default_rng(0).normal(size=Lr.num_params()) * 0.2).astype(np.float32)
    p[-6 - 6 * 64:-6] = 0.0
    b = np.array([0.5, -0.25, 0.125, 1.0, 0.0, -0.5], np.float32)
    p[-6:] = b
    m = Lr.Mlp(params=p)
    sims = _sims(3)
    ens = E.Ensemble(sims, _params(B.FLAG_LEARNED), config=E.Config(8, 4, 192, 1))
    with pytest.raises(RuntimeError):
        ens.step(1)      # no network attached
    ens.attach(m)
    ens.step(1)
    _check(ens, sims, ens.bodies(), bias=b)
    assert ens.counters()["corrected"] == sum(SIZES) - len(SIZES)
    ens.close()
    m.close()


def test_ensemble_relative_autotune_leaves_the_state_unchanged():
    from gravity_amd import barnes_hut as B
    from gravity_amd import ensemble as E
    print("[gpu] ensemble relative autotune", flush=True)
    sims = _sims(4)
    prm = _params(B.FLAG_QUADRUPOLE | B.FLAG_RELATIVE_MAC, 1e-3)
    ref = E.Ensemble(sims, prm, config=E.Config(8, 4, 192, 1))
    tuned = E.Ensemble(sims, prm, config=E.Config(8, 4, 192, 1))
    ref.step(2)
    tuned.step(2)
    tuned.autotune()
    ref.set_config(tuned.config())
    ref.step(2)
    tuned.step(2)
    a, b = ref.bodies(), tuned.bodies()
    for s in range(len(sims)):
        assert np.array_equal(a[s].view(np.uint32), b[s].view(np.uint32)), s
    ref.close()
    tuned.close()
