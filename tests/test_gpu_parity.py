"""HIP path against the reference's llvmpipe fixtures and the CPU oracle.

Stage-isolated: every stage of the GPU pipeline is fed the reference's own
inputs for that step through the C ABI (bh_upload by binding id) and its
outputs are downloaded and compared. Integer stages and the FP32 stages are
bit-exact (NaN equals NaN); merge resolution is compared with the oracle's
deterministic order and, where the reference's racy resolution is
unambiguous (bodies in exactly one task), with the reference itself.
"""
import collections

import numpy as np
import pytest

from oracle import bh_oracle as O
from tests.golden_util import FULL, load, num_alive, params_of, stage_inputs, steps

pytestmark = pytest.mark.gpu

EMPTY = 0xFFFFFFFF


def assert_bits(actual, desired, what=""):
    a = np.ascontiguousarray(actual, dtype=np.float32)
    d = np.ascontiguousarray(desired, dtype=np.float32)
    np.testing.assert_array_equal(np.isnan(a), np.isnan(d), err_msg=what)
    keep = ~np.isnan(a)
    np.testing.assert_array_equal(a[keep].view(np.uint32), d[keep].view(np.uint32), err_msg=what)


def sim_array(n, num, bounds, G, blc, units, out_of_bounds=0):
    s = np.zeros(224, np.uint32)
    f = s.view(np.float32)
    s[0] = num
    s[1] = n
    s[4] = out_of_bounds
    f[8:11] = bounds[:3]
    f[12:15] = bounds[3:]
    f[16:20] = [units.mass, units.density, units.len, units.time]
    f[21] = G
    f[22] = blc
    return s


def make_bh(fx, s=0, flags=0):
    from gravity_amd import barnes_hut as B
    from gravity_amd.initial_conditions import g_one_units
    p = params_of(fx)
    G = float(fx[f"s{s}_G"][0])
    blc = float(fx[f"s{s}_blc"][0])
    prm = B.make_params(theta=p["theta"], dt=p["dt"], softening=p["softening"], elasticity=p["elasticity"],
                        merging=p["merging"], dynamic=p["dynamic"], wrap_around=p["wrap_around"],
                        G=G, body_length_constant=blc, flags=flags)
    bh = B.BarnesHut(fx["bodies0"], fx["bounds0"], prm)
    return bh, prm, g_one_units(), G, blc


CASES = [(name, s) for name in FULL for s in range(steps(load(name)))]
# fixtures whose tree is higher than the reference's 64 fill passes: the
# reference's behaviour is BH_FLAG_FILL_CAP
CAPPED = {"tall_tree_95"}


@pytest.mark.parametrize("name,s", CASES)
def test_gpu_stage_parity(name, s):
    from gravity_amd import barnes_hut as B
    fx = load(name)
    n = fx["bodies0"].shape[0]
    print(f"[gpu] stage parity {name} step {s} n={n}", flush=True)
    capped = name in CAPPED
    bh, prm, units, G, blc = make_bh(fx, s, flags=B.FLAG_COLORS | (B.FLAG_FILL_CAP if capped else 0))
    src, idx, num_before = stage_inputs(fx, s)
    oob_before = int(fx[f"s{s - 1}_out_of_bounds"]) if s else 0
    bh.upload(B.SIMULATION_VALUES, sim_array(n, num_before, fx[f"s{s}_bounds_prev"], G, blc, units, oob_before))
    bh.upload(B.BODIES_IN, src)
    bh.upload(B.BODIES_OUT, np.zeros_like(src))
    bh.upload(B.INDEX_IN, idx)

    bh.reset_values()
    bh.partition_dead_bodies()
    num = num_alive(fx, s)
    sim = bh.sim_values()
    assert sim[0] == num and sim[2] == 0
    assert bh.counters()["just_died"] == int(fx[f"s{s}_died"])
    np.testing.assert_array_equal(bh.index()[:num_before], fx[f"s{s}_index_dead"][:num_before])
    assert_bits(bh.bodies(), fx[f"s{s}_bodies_start"], "bodies after dead partition")

    if prm.static_or_dynamic == B.DYNAMIC:
        bh.update_bounds()
    assert_bits(bh.bounds(), fx[f"s{s}_bounds"], "bounds")

    bh.generate_morton_codes()
    np.testing.assert_array_equal(bh.morton()[:num], fx[f"s{s}_morton_encoded"])

    bh.radix_sort()
    np.testing.assert_array_equal(bh.morton()[:num], fx[f"s{s}_morton_sorted"])
    np.testing.assert_array_equal(bh.index()[:num], fx[f"s{s}_index_sorted"])

    bh.build_tree()
    internal = bh.nodes(B.INTERNAL_NODES)[: num - 1]
    leaf = bh.nodes(B.LEAF_NODES)[:num]
    ref_i = fx[f"s{s}_fill_internal"]
    ref_l = fx[f"s{s}_fill_leaf"]
    np.testing.assert_array_equal(internal[:, 10:16], ref_i[:, 10:16])            # topology + height + count
    np.testing.assert_array_equal(leaf[:, 10:16], ref_l[:, 10:16])
    assert_bits(internal[:, 0:10].view(np.float32), ref_i[:, 0:10].view(np.float32), "internal fill")
    assert_bits(leaf[:, 0:10].view(np.float32), ref_l[:, 0:10].view(np.float32), "leaf fill")
    # star colours (common.glsl:235-276: pow and log of the HIP math library
    # against the reference's) and their mass-weighted averages
    for what, got_c, ref_c in (("leaf", leaf, ref_l), ("internal", internal, ref_i)):
        d = np.abs(got_c[:, 16:20].view(np.float32) - ref_c[:, 16:20].view(np.float32))
        exact = float((got_c[:, 16:20] == ref_c[:, 16:20]).all(axis=1).mean()) if d.size else 1.0
        print(f"[gpu] {what} colours: max difference {float(d.max()) if d.size else 0.0:.3g}, "
              f"exact fraction {exact:.4f}", flush=True)
        assert_bits(got_c[:, 16:20].view(np.float32), ref_c[:, 16:20].view(np.float32), f"{what} colours")
    if capped:   # (the reference's root keeps the build's nodeDepth; the counter has the height)
        assert bh.counters()["tree_height"] > 64 and ref_i[0, 3] == 0
    else:
        assert bh.counters()["tree_height"] == int(ref_i[0, 12])

    # accelerations: the reference's own extraction (velocities 0, dt 1 -> vel' = a)
    start = fx[f"s{s}_bodies_start"]
    zv = start.copy()
    zv[:, 4:7] = 0
    bh.upload(B.BODIES_IN, zv)
    p1 = B.Params.from_buffer_copy(prm)
    p1.dt = 1.0
    bh.reset_values(p1)
    bh.compute_force(p1)
    acc = bh.bodies(B.BODIES_OUT)[:, 4:7]
    alive = fx[f"s{s}_index_sorted"]
    killed = np.zeros(n, bool)
    if prm.static_or_dynamic == B.STATIC and not prm.wrap_around:
        b = fx[f"s{s}_bounds"]
        killed = ((start[:, :3] < b[:3]) | (start[:, :3] > b[3:])).any(axis=1)
    keep = alive[~killed[alive]]
    assert_bits(acc[keep], fx[f"s{s}_accel"][keep], "accelerations")

    # the real force pass
    bh.upload(B.BODIES_IN, start)
    bh.upload(B.BODIES_OUT, np.zeros_like(start))
    bh.reset_values()
    sim = sim_array(n, num, fx[f"s{s}_bounds"], G, blc, units, oob_before)
    bh.upload(B.SIMULATION_VALUES, sim)
    bh.compute_force()
    c = bh.counters()
    assert c["opened"] == int(fx[f"s{s}_opened"])
    assert c["error_flags"] == 0
    assert bh.sim_values()[23] == int(fx[f"s{s}_opened"]) & 0xFFFFFFFF
    assert bh.sim_values()[4] == int(fx[f"s{s}_out_of_bounds"])
    assert_bits(bh.bodies(B.BODIES_OUT)[alive], fx[f"s{s}_force_out"][alive], "integrated bodies")
    got = set(map(tuple, bh.merge_tasks().tolist()))
    want = set(map(tuple, fx[f"s{s}_tasks"].tolist()))
    assert got == want

    # merge resolution on the reference's tasks and bodies
    if prm.merging_collision_or_neither == B.MERGING:
        tasks = fx[f"s{s}_tasks"]
        blob = np.zeros(2 + 2 * tasks.shape[0], np.uint32)
        blob[1] = tasks.shape[0]
        blob[2:] = tasks.ravel()
        bh.upload(B.MERGE_TASKS, blob)
        bh.upload(B.BODIES_OUT, fx[f"s{s}_force_out"])
        bh.merge_bodies()
        gpu = bh.bodies(B.BODIES_OUT)
        cpu = fx[f"s{s}_force_out"].copy()
        applied = O.merge(cpu, tasks)
        assert_bits(gpu, cpu, "merge vs oracle order")
        assert bh.counters()["merges_applied"] == applied
        ref = fx[f"s{s}_merge_out"]
        cnt = collections.Counter(tasks.ravel().tolist())
        for a, b in tasks.tolist():
            if cnt[a] == 1 and cnt[b] == 1:
                assert_bits(gpu[[a, b]], ref[[a, b]], "isolated merge pair")
    bh.close()


@pytest.mark.parametrize("name", ["plummer_4096", "collision_4096", "edge_dead_nan_1000", "edge_duplicates_512",
                                  "two_galaxy_4096",
                                  "static_kill_1024", "static_wrap_1024", "theta0_512"])
def test_gpu_free_running_steps(name):
    """bh_step from the initial bodies reproduces the reference trajectory bit for bit."""
    from gravity_amd import barnes_hut as B
    fx = load(name)
    print(f"[gpu] free-running {name}", flush=True)
    bh, prm, _, _, _ = make_bh(fx)
    bh.init()
    for s in range(steps(fx)):
        bh.step()
        c = bh.counters()
        assert c["num_bodies"] == num_alive(fx, s)
        assert c["opened"] == int(fx[f"s{s}_opened"])
        assert c["error_flags"] == 0
    assert_bits(bh.bodies(), fx["final_bodies"], "final bodies")
    bh.close()


@pytest.mark.parametrize("name", ["plummer_4096", "collision_4096", "static_kill_1024", "theta0_512"])
def test_gpu_general_walk_loop(name):
    """BH_FLAG_GENERAL_WALK (the walk's LDS-stack loop instead of the shorter
    one) reproduces the reference trajectory bit for bit as well."""
    from gravity_amd import barnes_hut as B
    fx = load(name)
    print(f"[gpu] general walk loop {name}", flush=True)
    bh, prm, _, _, _ = make_bh(fx, flags=B.FLAG_GENERAL_WALK)
    bh.init()
    for s in range(steps(fx)):
        bh.step()
        c = bh.counters()
        assert c["opened"] == int(fx[f"s{s}_opened"])
        assert c["error_flags"] == 0
    assert_bits(bh.bodies(), fx["final_bodies"], "final bodies")
    bh.close()


@pytest.mark.parametrize("general", [False, True])
def test_gpu_merge_steps_match_oracle(general):
    """Merging on: bh_step and the oracle's step agree on every step (same
    deterministic task order), and the reference agrees on step 0's tasks;
    with either walk loop."""
    from gravity_amd import barnes_hut as B
    fx = load("merge_4096")
    print(f"[gpu] merge steps (general loop {general})", flush=True)
    bh, prm, _, G, blc = make_bh(fx, flags=B.FLAG_GENERAL_WALK if general else 0)
    bh.init()
    sim = O.Simulation(fx["bodies0"], fx["bounds0"], 1.0, 1.0, G=G, body_len_const=blc)
    oprm = O.params(**params_of(fx))
    for s in range(steps(fx)):
        bh.step()
        assert sim.step(oprm) == 0
        c = bh.counters()
        assert c["num_bodies"] == sim.num_bodies
        assert c["merge_tasks"] == sim.st.last.n_tasks
        assert_bits(bh.bodies(), sim.bodies(), f"bodies after step {s}")
    bh.close()


def test_gpu_steps_with_uploads_between_match_oracle():
    """bh_step carries the walk's positions by sorted position into the next
    step's encode and tree; an upload of new bodies between steps (and a
    stage entry point called on its own) must drop them: with bodies
    perturbed on both sides before step 3 the trajectory still equals the
    oracle's bit for bit, and so does a trajectory with deaths (merging)."""
    import ctypes
    from gravity_amd import barnes_hut as B
    for name in ("plummer_4096", "merge_4096"):
        fx = load(name)
        print(f"[gpu] steps with uploads between: {name}", flush=True)
        bh, prm, _, G, blc = make_bh(fx)
        bh.init()
        sim = O.Simulation(fx["bodies0"], fx["bounds0"], 1.0, 1.0, G=G, body_len_const=blc)
        oprm = O.params(**params_of(fx))
        rng = np.random.default_rng(7)
        for s in range(6):
            if s == 3:
                b = np.ascontiguousarray(bh.bodies(), np.float32)
                alive = b[:, 3] != 0.0
                b[alive, 0:3] += rng.normal(0.0, 1e-3, size=(int(alive.sum()), 3)).astype(np.float32)
                bh.upload(B.BODIES_IN, b)
                ctypes.memmove(sim.st.bodies[sim.st.cur], b.ctypes.data, b.nbytes)
            if s == 4:
                bh.generate_morton_codes()
            bh.step()
            assert sim.step(oprm) == 0
            c = bh.counters()
            assert c["num_bodies"] == sim.num_bodies
            assert c["error_flags"] == 0
            assert_bits(bh.bodies(), sim.bodies(), f"{name} bodies after step {s}")
            np.testing.assert_array_equal(bh.index()[:sim.num_bodies], sim.index()[:sim.num_bodies])
        bh.close()


@pytest.mark.parametrize("n", [2, 3, 63, 64, 65, 1023, 1025, 4095, 4097, 8191, 8193, 16385])
def test_gpu_step_sizes_match_oracle(n):
    """bh_step at sizes around the wave (64), tree block (1,024) and sort
    tile (8,192) boundaries equals the oracle's step bit for bit, with the
    positions carried from step to step."""
    from gravity_amd import barnes_hut as B
    from gravity_amd import initial_conditions as ic
    print(f"[gpu] step sizes n={n}", flush=True)
    bodies = ic.plummer(n, 100 + n)
    bounds = ic.tight_bounds(bodies)
    prm = B.make_params(theta=0.5, dt=1e-3, softening=0.01, G=1.0)
    bh = B.BarnesHut(bodies, bounds, prm)
    bh.init()
    sim = O.Simulation(bodies, bounds, 1.0, 1.0, G=1.0, body_len_const=1.0)
    oprm = O.params(theta=0.5, dt=1e-3, softening=0.01)
    for s in range(3):
        bh.step()
        assert sim.step(oprm) == 0
        c = bh.counters()
        assert c["error_flags"] == 0
        assert_bits(bh.bodies(), sim.bodies(), f"n={n} bodies after step {s}")
        np.testing.assert_array_equal(bh.index(), sim.index())
    bh.close()


@pytest.mark.parametrize("name", ["plummer_4096", "merge_4096"])
def test_gpu_step_with_colors_equals_stage_calls(name):
    """bh_step (positions sorted, bodies read through them, new index written
    by the tree) and the stage entry points called one by one (ids sorted,
    bodies gathered by id) give the same bodies, index and std430 nodes,
    colours included, step after step."""
    from gravity_amd import barnes_hut as B
    fx = load(name)
    print(f"[gpu] bh_step with colours vs stage calls: {name}", flush=True)
    a, prm, _, _, _ = make_bh(fx, flags=B.FLAG_COLORS)
    b, _, _, _, _ = make_bh(fx, flags=B.FLAG_COLORS)
    a.init()
    b.init()
    for s in range(4):
        a.step()
        b.reset_values()
        b.partition_dead_bodies()
        if prm.static_or_dynamic == B.DYNAMIC:
            b.update_bounds()
        b.generate_morton_codes()
        b.radix_sort()
        b.build_tree()
        b.compute_force()
        if prm.merging_collision_or_neither == B.MERGING:
            b.merge_bodies()
        np.testing.assert_array_equal(a.nodes(B.LEAF_NODES), b.nodes(B.LEAF_NODES), f"{name} leaves step {s}")
        np.testing.assert_array_equal(a.nodes(B.INTERNAL_NODES), b.nodes(B.INTERNAL_NODES),
                                      f"{name} internal nodes step {s}")
        b.swap_body_buffers()
        assert_bits(a.bodies(), b.bodies(), f"{name} bodies after step {s}")
        np.testing.assert_array_equal(a.index(), b.index())
    a.close()
    b.close()


def test_gpu_two_bodies_merge_to_one():
    """Two overlapping bodies merge in the first step; from the next step on
    one body is left and the walk runs on the reference's one-body tree
    (bh_tree.comp:46-62). Every step equals the oracle's bit for bit."""
    from gravity_amd import barnes_hut as B
    print("[gpu] two bodies merge to one", flush=True)
    bodies = np.zeros((2, 8), np.float32)
    bodies[0, :4] = [0.0, 0.0, 0.0, 0.5]
    bodies[1, :4] = [0.01, 0.0, 0.0, 0.5]
    bodies[:, 4:8] = [[0.0, 0.1, 0.0, 1.0], [0.0, -0.1, 0.0, 1.0]]
    bounds = np.array([-1, -1, -1, 1, 1, 1], np.float32)
    prm = B.make_params(theta=0.5, dt=1e-3, softening=1e-6, merging=B.MERGING, G=1.0, body_length_constant=1.0)
    bh = B.BarnesHut(bodies, bounds, prm)
    bh.init()
    sim = O.Simulation(bodies, bounds, 1.0, 1.0, G=1.0, body_len_const=1.0)
    oprm = O.params(theta=0.5, dt=1e-3, softening=1e-6, merging=1)
    alive = []
    for s in range(4):
        bh.step()
        assert sim.step(oprm) == 0
        c = bh.counters()
        assert c["error_flags"] == 0
        assert c["num_bodies"] == sim.num_bodies
        alive.append(c["num_bodies"])
        assert_bits(bh.bodies(), sim.bodies(), f"bodies after step {s}")
    assert alive[0] == 2 and alive[-1] == 1
    bh.close()


def test_gpu_plummer_65536_ten_steps():
    """The reference's CPU-runnable case end to end: 10 steps of bh_step."""
    from gravity_amd import barnes_hut as B
    from gravity_amd import initial_conditions as ic
    fx = load("plummer_65536")
    print("[gpu] plummer 65536 x 10", flush=True)
    bodies = ic.plummer(65536, int(fx["seed"]))
    prm = B.make_params(theta=0.5, dt=1e-3, softening=0.01, G=float(fx["s0_G"][0]),
                        body_length_constant=float(fx["s0_blc"][0]))
    bh = B.BarnesHut(bodies, fx["bounds0"], prm)
    bh.init()
    for s in range(10):
        bh.step()
        c = bh.counters()
        assert c["opened"] == int(fx[f"s{s}_opened"])
        assert c["tree_height"] == int(fx[f"s{s}_height"])
        assert_bits(bh.bounds(), fx[f"s{s}_bounds"], f"bounds step {s}")
    assert_bits(bh.bodies()[fx["sample"]], fx["final_sample"], "final sample")
    bh.close()


def test_gpu_tall_tree_filled_without_the_cap():
    """Without BH_FLAG_FILL_CAP every level of a tree higher than 64 is
    filled: nodes and accelerations equal the oracle's with no fill cap (its
    walk stack raised to the tree's height)."""
    from gravity_amd import barnes_hut as B
    fx = load("tall_tree_95")
    n = fx["bodies0"].shape[0]
    print(f"[gpu] tall tree without the fill cap n={n}", flush=True)
    bh, prm, units, G, blc = make_bh(fx)
    bodies = fx["bodies0"].copy()
    assert not bodies[:, 4:7].any()          # at rest: one step at dt 1 gives the accelerations as velocities
    p1 = B.Params.from_buffer_copy(prm)
    p1.dt = 1.0
    bh.init()
    bh.step(p1)
    c = bh.counters()
    assert c["tree_height"] == 68 and c["error_flags"] == 0
    sim = O.Simulation(bodies, fx["bounds0"], units.mass, units.density, G=G, body_len_const=blc, fill_cap=0,
                       stack_cap=256)
    assert sim.step(O.params(theta=float(fx["params"][0]), dt=1.0, softening=float(fx["params"][2]),
                             dynamic=0)) == 0
    assert c["opened"] == sim.st.last.opened and c["opened"] > 0
    got = bh.bodies()
    assert_bits(got, sim.bodies(), "bodies after one uncapped step")
    assert np.abs(got[:, 4:7]).max() > 0
    bh.close()
