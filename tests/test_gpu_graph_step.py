"""bh_step replayed as a hipGraph (BH_FLAG_GRAPH) against the direct step.

Two contexts start from the same bodies; one steps directly, the other with
the graph flag (its first step runs directly: no previous positions yet; the
next ones capture a graph for each state of the buffer roles and replay it).
After every step both download their bodies (work between replays) and the
bodies must be equal bit for bit: the plain step, the merging step and the
corrected walk with the relative opening test.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _pair(bodies, prm):
    from gravity_amd import barnes_hut as B
    from gravity_amd import initial_conditions as ic
    out = []
    for _ in range(2):
        bh = B.BarnesHut(bodies, ic.tight_bounds(bodies), prm)
        bh.init()
        out.append(bh)
    return out


def _run(direct, graph, prm, steps):
    from gravity_amd import barnes_hut as B
    g = B.Params.from_buffer_copy(prm)
    g.flags |= B.FLAG_GRAPH
    for s in range(steps):
        direct.step(prm)
        graph.step(g)
        a, b = direct.bodies(), graph.bodies()
        np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32), err_msg=f"step {s}")
    ca, cb = direct.counters(), graph.counters()
    for k in ("num_bodies", "interactions", "opened", "merges_applied", "quad_terms", "error_flags"):
        assert ca[k] == cb[k], k


def test_graph_step_plain():
    from gravity_amd import barnes_hut as B
    from gravity_amd import initial_conditions as ic
    print("[gpu] graph step, plain, 65536", flush=True)
    bodies = ic.plummer(1 << 16, 31)
    prm = B.make_params(theta=0.5, dt=1e-3, softening=0.01, G=1.0)
    d, g = _pair(bodies, prm)
    _run(d, g, prm, 7)
    d.close()
    g.close()


def test_graph_step_merging():
    from gravity_amd import barnes_hut as B
    from gravity_amd import initial_conditions as ic
    print("[gpu] graph step, merging, 32768", flush=True)
    bodies = ic.cube_plus_core(1 << 15, 32, density=64.0)
    prm = B.make_params(theta=0.5, dt=5e-4, softening=1e-6, merging=B.MERGING, G=1.0,
                        body_length_constant=0.6203504909)
    d, g = _pair(bodies, prm)
    _run(d, g, prm, 6)
    assert d.counters()["num_bodies"] <= bodies.shape[0]
    d.close()
    g.close()


def test_graph_step_corrected_relative():
    from gravity_amd import barnes_hut as B
    from gravity_amd import initial_conditions as ic
    print("[gpu] graph step, corrected walk with the relative test, 65536", flush=True)
    bodies = ic.plummer(1 << 16, 33)
    prm = B.make_params(theta=0.5, dt=1e-3, softening=0.01, G=1.0,
                        flags=B.FLAG_QUADRUPOLE | B.FLAG_RELATIVE_MAC, force_tolerance=1.5e-3)
    d, g = _pair(bodies, prm)
    _run(d, g, prm, 7)
    d.close()
    g.close()
