"""The render feed on the GPU against the std430 bindings it compacts.

bh_render_nodes must hold exactly the internal nodes of the downloaded
InternalNodes binding that the node-glow and region shaders do not cull
statically (non-empty box, mass > 0, bodies > 0, depth in range), in node
order, with their centre of mass, mass, box and depth bit for bit and their
colour quantised from avgColor; bh_render_bodies exactly the alive bodies of
BodiesIn in id order, with the leaf's star colour (the same function of the
same body) quantised, and the radius bodyLengthConstant (m / rho)^(1/3)
within 2 ulp of a CPU power. The cap truncates the copy, not the count.
"""
import numpy as np
import pytest

from tests.golden_util import load

pytestmark = pytest.mark.gpu


def _rgba(c):
    q = np.rint(np.clip(c.astype(np.float32), 0.0, 1.0) * np.float32(255.0)).astype(np.uint32)
    return q[:, 0] | (q[:, 1] << 8) | (q[:, 2] << 16) | np.uint32(255 << 24)


def _stepped(bodies, bounds, steps=2, **kw):
    from gravity_amd import barnes_hut as B
    prm = B.make_params(theta=0.5, dt=1e-3, softening=0.01, flags=B.FLAG_COLORS, **kw)
    bh = B.BarnesHut(bodies, bounds, prm)
    bh.init()
    for _ in range(steps):
        bh.step()
    # the tree of the current bodies (the last step built it from the bodies before)
    bh.build_current_tree(prm)
    return bh


def _expected_nodes(bh, dmin, dmax):
    from gravity_amd import barnes_hut as B
    num = int(bh.sim_values()[0])
    raw = bh.nodes(B.INTERNAL_NODES)[: max(num - 1, 0)]
    f = raw.view(np.float32)
    depth = raw[:, 12]
    keep = ~(f[:, 4] > f[:, 7]) & (f[:, 3] > 0) & (raw[:, 13] > 0) & (depth >= dmin) & (depth <= dmax)
    return raw[keep], f[keep]


@pytest.mark.parametrize("dmin,dmax", [(0, 0xFFFFFFFF), (1, 8), (5, 12)])
def test_render_nodes_match_the_binding(dmin, dmax):
    from gravity_amd import initial_conditions as ic
    print(f"[gpu] render nodes depth {dmin}..{dmax}", flush=True)
    b = ic.plummer(50000, 31)
    bh = _stepped(b, ic.tight_bounds(b))
    got = bh.render_nodes(dmin, dmax)
    raw, f = _expected_nodes(bh, dmin, dmax)
    assert bh.render_total == raw.shape[0] == got.shape[0] and got.shape[0] > 0
    np.testing.assert_array_equal(got["com"].view(np.uint32), f[:, 0:3].view(np.uint32))
    np.testing.assert_array_equal(got["mass"].view(np.uint32), f[:, 3].view(np.uint32))
    np.testing.assert_array_equal(got["aabb_min"].view(np.uint32), f[:, 4:7].view(np.uint32))
    np.testing.assert_array_equal(got["aabb_max"].view(np.uint32), f[:, 7:10].view(np.uint32))
    np.testing.assert_array_equal(got["depth"], raw[:, 12])
    np.testing.assert_array_equal(got["rgba"], _rgba(f[:, 16:19]))
    assert (got["rgba"] & 0xFFFFFF).any()
    # a cap below the count: the first records, the full count
    part = bh.render_nodes(dmin, dmax, cap=got.shape[0] // 3)
    assert bh.render_total == got.shape[0]
    np.testing.assert_array_equal(part.view(np.uint8), got[: part.shape[0]].view(np.uint8))
    bh.close()


def test_render_bodies_match_the_binding():
    """Dead bodies (mass 0), NaN and infinite positions: the fixture with them."""
    from gravity_amd import barnes_hut as B
    from gravity_amd import initial_conditions as ic
    print("[gpu] render bodies", flush=True)
    fx = load("edge_dead_nan_1000")
    bh = _stepped(fx["bodies0"], fx["bounds0"], steps=1)
    got = bh.render_bodies()
    body = bh.bodies(B.BODIES_IN)
    alive = np.nonzero(body[:, 3] != 0)[0]
    assert bh.render_total == alive.size == got.shape[0] and alive.size < body.shape[0]
    np.testing.assert_array_equal(got["pos"].view(np.uint32), body[alive, 0:3].view(np.uint32))
    blc = np.float32(bh.sim_values().view(np.float32)[22])
    want_r = blc * np.power(body[alive, 3] / body[alive, 7], np.float32(1.0 / 3.0), dtype=np.float32)
    fin = np.isfinite(want_r)
    assert (np.isfinite(got["radius"]) == fin).all()
    np.testing.assert_array_max_ulp(got["radius"][fin], want_r[fin], maxulp=2)
    # the leaves' colours are the same function of the same bodies
    leaf = bh.nodes(B.LEAF_NODES)
    idx = bh.index()
    num = int(bh.sim_values()[0])
    color = np.zeros((body.shape[0], 3), np.float32)
    color[idx[:num]] = leaf[:num].view(np.float32)[:, 16:19]
    np.testing.assert_array_equal(got["rgba"], _rgba(color[alive]))
    bh.close()


def test_render_feed_without_colours_has_none():
    from gravity_amd import barnes_hut as B
    from gravity_amd import initial_conditions as ic
    print("[gpu] render nodes without colours", flush=True)
    b = ic.plummer(4096, 3)
    prm = B.make_params(theta=0.5, dt=1e-3, softening=0.01)
    bh = B.BarnesHut(b, ic.tight_bounds(b), prm)
    bh.init()
    bh.step()
    got = bh.render_nodes()
    assert got.shape[0] == bh.render_total > 0 and not got["rgba"].any()
    bh.close()


def test_render_feed_into_page_locked_memory():
    from gravity_amd import barnes_hut as B
    from gravity_amd import initial_conditions as ic
    print("[gpu] render feed into page-locked memory", flush=True)
    b = ic.plummer(20000, 5)
    bh = _stepped(b, ic.tight_bounds(b), steps=1)
    want_n, want_b = bh.render_nodes(1, 8), bh.render_bodies()
    pn = B.PinnedBuffer((bh.n,), B.RENDER_NODE)
    pb = B.PinnedBuffer((bh.n,), B.RENDER_BODY)
    got_n = bh.render_nodes(1, 8, out=pn.array)
    got_b = bh.render_bodies(out=pb.array)
    np.testing.assert_array_equal(got_n.view(np.uint8), want_n.view(np.uint8))
    np.testing.assert_array_equal(got_b.view(np.uint8), want_b.view(np.uint8))
    del got_n, got_b
    pn.close()
    pb.close()
    bh.close()
