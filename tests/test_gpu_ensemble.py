"""Ensembles of independent simulations on the GPU.

Each simulation of an ensemble must evolve bit for bit as it does stepped
alone: bounds, Morton codes, permutation, tree (every field of every internal
node) and bodies are compared with single-simulation runs of the HIP path
(itself pinned to the oracle) and, for two simulations, with the CPU oracle
directly. Graph replay, every sort digit width and every walk shape must give
the same bits as eager launches with the defaults, and the autotuner must
leave the state unchanged.
"""
import numpy as np
import pytest

from oracle import bh_oracle as O

pytestmark = pytest.mark.gpu

SOFT = 0.01


def _sims(sizes, seed):
    from gravity_amd import initial_conditions as ic
    return [ic.plummer(int(n), seed * 100 + k) for k, n in enumerate(sizes)]


def _params(theta=0.5):
    from gravity_amd import barnes_hut as B
    return B.make_params(theta=theta, dt=1e-3, softening=SOFT, G=1.0)


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


SIZES = [4096, 5000, 2, 777, 12345, 2048, 2049, 30000]


def test_ensemble_step_matches_single_simulations():
    from gravity_amd import barnes_hut as B
    from gravity_amd import ensemble as E
    from gravity_amd import initial_conditions as ic
    print("[gpu] ensemble vs single simulations", flush=True)
    sims = _sims(SIZES, 1)
    prm = _params()
    ens = E.Ensemble(sims, prm, config=E.Config(8, 4, 192, 0))
    singles = []
    for b in sims:
        bh = B.BarnesHut(b, ic.tight_bounds(b), prm)
        bh.init()
        singles.append(bh)
    for step in range(3):
        ens.step(1)
        for bh in singles:
            bh.step(prm)
        ens_bounds = ens.bounds()
        ens_sorted = ens.sorted()
        ens_nodes = ens.internal_nodes()
        ens_bodies = ens.bodies()
        for s, bh in enumerate(singles):
            n = SIZES[s]
            assert np.array_equal(_bits(ens_bounds[s]), _bits(bh.bounds())), (step, s)
            keys, idx = ens_sorted[s]
            np.testing.assert_array_equal(keys, bh.morton(), err_msg=f"step {step} sim {s}")
            np.testing.assert_array_equal(idx, bh.index()[:n], err_msg=f"step {step} sim {s}")
            np.testing.assert_array_equal(ens_nodes[s], bh.nodes(B.INTERNAL_NODES)[: n - 1],
                                          err_msg=f"step {step} sim {s}")
            assert np.array_equal(_bits(ens_bodies[s]), _bits(bh.bodies(B.BODIES_IN))), (step, s)
    c = ens.counters()
    assert c["error_flags"] == 0 and c["interactions"] > 0
    for bh in singles:
        bh.close()
    ens.close()


def test_ensemble_matches_cpu_oracle():
    from gravity_amd import ensemble as E
    from gravity_amd import initial_conditions as ic
    print("[gpu] ensemble vs oracle", flush=True)
    sizes = [3000, 9000, 4096]
    sims = _sims(sizes, 2)
    ens = E.Ensemble(sims, _params(), config=E.Config(8, 4, 192, 1))
    ens.step(2)
    got = ens.bodies()
    oprm = O.params(theta=0.5, dt=1e-3, softening=SOFT)
    for s in (0, 1):
        sim = O.Simulation(sims[s], ic.tight_bounds(sims[s]), 1.0, 1.0, G=1.0, stack_cap=256, threads=8)
        for _ in range(2):
            assert sim.step(oprm) == 0
        assert np.array_equal(_bits(got[s]), _bits(sim.bodies())), s
    ens.close()


def test_graph_replay_matches_eager_launches():
    from gravity_amd import ensemble as E
    print("[gpu] ensemble graph replay", flush=True)
    sims = _sims([5000, 20000, 2, 4096, 65536], 3)
    eager = E.Ensemble(sims, _params(), config=E.Config(8, 4, 192, 0))
    graph = E.Ensemble(sims, _params(), config=E.Config(8, 4, 192, 1))
    for k in range(5):
        eager.step(1)
        graph.step(1)
        a, b = eager.bodies(), graph.bodies()      # synchronising reads between the replays
        for s in range(len(sims)):
            assert np.array_equal(_bits(a[s]), _bits(b[s])), (k, s)
    eager.step(3)
    graph.step(3)
    a, b = eager.bodies(), graph.bodies()
    for s in range(len(sims)):
        assert np.array_equal(_bits(a[s]), _bits(b[s])), s
    eager.close()
    graph.close()


def test_general_walk_loop_gives_identical_results():
    """BH_FLAG_GENERAL_WALK runs the walk's LDS-stack loop for every task: the
    same bits as the shorter loop."""
    from gravity_amd import barnes_hut as B
    from gravity_amd import ensemble as E
    print("[gpu] ensemble general walk loop", flush=True)
    sims = _sims([5000, 20000, 2, 3333], 5)
    general = B.make_params(theta=0.5, dt=1e-3, softening=SOFT, G=1.0, flags=B.FLAG_GENERAL_WALK)
    ref = E.Ensemble(sims, _params(), config=E.Config(8, 4, 192, 0))
    alt = E.Ensemble(sims, general, config=E.Config(8, 4, 192, 0))
    ref.step(3)
    alt.step(3)
    a, b = ref.bodies(), alt.bodies()
    for s in range(len(sims)):
        assert np.array_equal(_bits(a[s]), _bits(b[s])), s
    ref.close()
    alt.close()


@pytest.mark.parametrize("bits,waves,depth", [(5, 1, 64), (6, 2, 128), (7, 8, 192), (7, 4, 64), (8, 8, 64)])
def test_digit_widths_and_walk_shapes_give_identical_results(bits, waves, depth):
    from gravity_amd import ensemble as E
    print(f"[gpu] ensemble shape radix {bits} walk {waves}x{depth}", flush=True)
    sims = _sims([5000, 20000, 2, 3333], 4)
    ref = E.Ensemble(sims, _params(), config=E.Config(8, 4, 192, 0))
    alt = E.Ensemble(sims, _params(), config=E.Config(bits, waves, depth, 1))
    ref.step(3)
    alt.step(3)
    a, b = ref.bodies(), alt.bodies()
    for s in range(len(sims)):
        assert np.array_equal(_bits(a[s]), _bits(b[s])), s
    ka, kb = ref.sorted(), alt.sorted()
    for s in range(len(sims)):
        np.testing.assert_array_equal(ka[s][1], kb[s][1])
    ref.close()
    alt.close()


def test_autotune_leaves_the_state_unchanged():
    from gravity_amd import ensemble as E
    from gravity_amd import initial_conditions as ic
    print("[gpu] ensemble autotune", flush=True)
    sims = [ic.ensemble(12, 5)[k] for k in range(12)]
    ref = E.Ensemble(sims, _params(), config=E.Config(8, 4, 192, 1))
    tuned = E.Ensemble(sims, _params(), config=E.Config(8, 4, 192, 1))
    ref.step(2)
    tuned.step(2)
    cfg, timings = tuned.autotune()
    print("[gpu] autotune chose", cfg.as_dict(), flush=True)
    assert 5 <= cfg.radix_bits <= 8 and cfg.walk_waves in (1, 2, 4, 8) and cfg.stack_depth in (64, 128, 192)
    assert all(timings[f"radix_bits_{b}"] > 0 for b in range(5, 9))
    assert cfg.use_graph == 1
    ref.step(2)
    tuned.step(2)
    a, b = ref.bodies(), tuned.bodies()
    for s in range(len(sims)):
        assert np.array_equal(_bits(a[s]), _bits(b[s])), s
    ref.close()
    tuned.close()
