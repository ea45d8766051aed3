"""The merging, collision and two-galaxy workloads on the GPU, and the FP64
force judge, against the CPU oracle.

* Two galaxies (2 x 524,288 bodies, exponential disks with Hernquist bulges,
  10 apart on a parabolic orbit): every stage bit for bit against the oracle
  on the same input (Morton codes, permutation, whole tree, accelerations),
  through the persistent walk (from 2^20 bodies).
* Cube + Plummer core at 2^20 bodies with merging (and with collisions):
  two bh_step calls against the oracle's step, bodies bit for bit, task and
  merge counts equal; the persistent walk emits the merge tasks.
* Merge resolution with more chained tasks than its one-block list holds
  after the grid rounds (the path that scans the whole task array).
* The FP64 direct sum (the judge of every force-error figure) against the
  oracle's, with an explicit G, dead bodies and a sample.
* Two galaxies at the full 33,554,432 bodies for 5 steps: size-independent
  properties (sorted keys, a permutation, the Karras tree's counts and mass,
  no error flags, finite bodies) and the force error against the FP64 judge
  on 16,384 sampled bodies.
"""
import numpy as np
import pytest

from oracle import bh_oracle as O

pytestmark = pytest.mark.gpu


def _bits_equal(a, b):
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    return a.shape == b.shape and np.array_equal(a.view(np.uint32), b.view(np.uint32))


def _g_one():
    from gravity_amd import initial_conditions as ic
    u = ic.g_one_units()
    G, blc = O.units(u.mass, u.density, u.len, u.time)
    return float(G), float(blc)


def test_gpu_two_galaxy_matches_oracle_1m():
    from gravity_amd import barnes_hut as B
    from gravity_amd import initial_conditions as ic
    bodies = ic.two_galaxy(1 << 19, 21)
    n = bodies.shape[0]
    print(f"[gpu] two galaxies stage parity n={n}", flush=True)
    bounds = ic.tight_bounds(bodies)
    G, blc = _g_one()
    prm = B.make_params(theta=0.5, dt=1e-3, softening=0.01, G=G, body_length_constant=blc)
    bh = B.BarnesHut(bodies, bounds, prm)
    bh.init()
    bh.reset_values()
    bh.partition_dead_bodies()
    bh.update_bounds()
    bh.generate_morton_codes()
    bh.radix_sort()
    bh.build_tree()
    keys, idx = bh.morton(), bh.index()
    internal = bh.nodes(B.INTERNAL_NODES)
    gbounds = bh.bounds()

    src = bodies.copy()
    out, died = O.dead_partition(src, None, np.arange(n, dtype=np.uint32), n)
    assert died == 0
    ob = O.aabb(src, out, n, bounds)
    assert _bits_equal(ob, gbounds)
    okeys, oidx = O.sort(O.morton_encode(src, out, n, ob), out)
    np.testing.assert_array_equal(keys, okeys)
    np.testing.assert_array_equal(idx, oidx)
    leaf, ointernal = O.build(okeys, n, n)
    h = O.fill(src, oidx, n, n, leaf, ointernal, max_passes=0, with_colors=False)
    assert bh.counters()["tree_height"] == h
    np.testing.assert_array_equal(internal[:, 10:16], ointernal[:, 10:16])
    assert _bits_equal(internal[:, 0:10], ointernal[:, 0:10])

    acc = bh.accelerations()
    c = bh.counters()
    zv = src.copy()
    zv[:, 4:7] = 0
    r = O.force(zv, oidx, n, n, leaf, ointernal, O.params(theta=0.5, dt=1.0, softening=0.01), G, blc, ob,
                dst=np.zeros_like(zv), stack_cap=256)
    assert r["rc"] == 0
    assert _bits_equal(acc[oidx], r["accel"])
    assert c["opened"] == r["stats"].opened
    assert c["error_flags"] == 0
    bh.close()


@pytest.mark.parametrize("mode", ["merging", "collision"])
def test_gpu_cube_core_steps_match_oracle_1m(mode):
    """Two steps of the cube + core workload at 2^20 bodies (persistent walk)
    with merging (tasks emitted by the walk, resolved in ascending order) or
    collisions (bodies changed in the walk), bit for bit against the oracle."""
    from gravity_amd import barnes_hut as B
    from gravity_amd import initial_conditions as ic
    n = 1 << 20
    print(f"[gpu] cube + core {mode} steps n={n}", flush=True)
    bodies = ic.cube_plus_core(n, 17, density=64.0)
    bounds = ic.tight_bounds(bodies)
    G, blc = _g_one()
    sel = B.MERGING if mode == "merging" else B.COLLISION
    prm = B.make_params(theta=0.5, dt=5e-4, softening=1e-6, merging=sel, G=G, body_length_constant=blc)
    bh = B.BarnesHut(bodies, bounds, prm)
    bh.init()
    u = ic.g_one_units()
    sim = O.Simulation(bodies, bounds, u.mass, u.density, G=G, body_len_const=blc, fill_cap=0, stack_cap=256)
    oprm = O.params(theta=0.5, dt=5e-4, softening=1e-6, merging=sel)
    for s in range(2):
        bh.step()
        assert sim.step(oprm) == 0
        c = bh.counters()
        assert c["error_flags"] == 0
        assert c["num_bodies"] == sim.num_bodies
        if mode == "merging":
            assert c["merge_tasks"] == sim.st.last.n_tasks
            assert c["merge_tasks"] > 0, "the workload must emit merge tasks"
        assert _bits_equal(bh.bodies(), sim.bodies()), f"{mode}: bodies after step {s}"
    bh.close()


def test_gpu_merge_long_chains_match_oracle():
    """Chains of tasks (i, i+1) commit one task per round, so after the grid
    rounds 40,000 tasks are still pending, more than the finishing block's
    list holds: that block then scans the whole task array every round. The
    result equals the oracle's ascending-order resolution bit for bit."""
    from gravity_amd import barnes_hut as B
    from gravity_amd import initial_conditions as ic
    chains, length = 20000, 10
    n = chains * (length + 1)
    print(f"[gpu] merge long chains n={n}", flush=True)
    bodies = ic.plummer(n, 4)
    bh = B.BarnesHut(bodies, ic.tight_bounds(bodies))
    bh.init()
    a = (np.arange(chains)[:, None] * (length + 1) + np.arange(length)[None, :]).ravel().astype(np.uint32)
    tasks = np.stack([a, a + 1], axis=1)
    rng = np.random.default_rng(3)
    tasks = tasks[rng.permutation(tasks.shape[0])]        # emission order does not matter
    header = np.array([0, tasks.shape[0]], np.uint32)
    bh.upload(B.BODIES_OUT, bodies)
    bh.upload(B.MERGE_TASKS, np.concatenate([header, tasks.ravel()]))
    bh.merge_bodies()
    got = bh.bodies(B.BODIES_OUT)
    want = bodies.copy()
    applied = O.merge(want, tasks)
    assert bh.counters()["merges_applied"] == applied
    assert applied == chains * (length // 2)
    assert _bits_equal(got, want)
    bh.close()


def test_gpu_direct_sum_matches_oracle():
    """The GPU's FP64 direct sum against the oracle's on the same sample:
    G from the context (explicit 0.5 here), dead bodies skipped, the sampled
    body itself skipped. Both sum in double precision in different orders;
    the smallest single term of these sums is above 1e-9 of |a|, so a
    relative tolerance of 1e-11 on |a| would catch one dropped body."""
    from gravity_amd import barnes_hut as B
    from gravity_amd import initial_conditions as ic
    for n, m in ((4096, 4096), (1 << 20, 64)):
        print(f"[gpu] direct sum vs oracle n={n} sample={m}", flush=True)
        bodies = ic.plummer(n, 12)
        bodies[::61, 3] = 0.0
        prm = B.make_params(G=0.5)
        bh = B.BarnesHut(bodies, ic.tight_bounds(bodies), prm)
        rng = np.random.default_rng(n)
        sample = np.sort(rng.choice(n, size=m, replace=False)).astype(np.uint32)
        got = bh.direct_sum_f64(sample, 0.01)
        want = O.direct_sum_f64(bodies, sample, G=0.5, softening=0.01)
        err = np.linalg.norm(got - want, axis=1) / np.linalg.norm(want, axis=1)
        assert err.max() < 1e-11, err.max()
        # the G is the context's: a G of 1 gives twice the sums
        bh2 = B.BarnesHut(bodies, ic.tight_bounds(bodies), B.make_params(G=1.0))
        np.testing.assert_allclose(bh2.direct_sum_f64(sample[:8], 0.01), 2.0 * got[:8], rtol=1e-15)
        bh.close()
        bh2.close()


def test_gpu_two_galaxy_full_size_properties():
    from gravity_amd import barnes_hut as B
    from gravity_amd import initial_conditions as ic
    bodies = ic.two_galaxy(1 << 24, 2024)
    n = bodies.shape[0]
    print(f"[gpu] two galaxies full size n={n}, 5 steps", flush=True)
    bounds = ic.tight_bounds(bodies)
    bh = B.BarnesHut(bodies, bounds, B.make_params(theta=0.5, dt=1e-3, softening=0.01))
    del bodies
    bh.init()
    for s in range(5):
        bh.step()
    c = bh.counters()
    assert c["error_flags"] == 0 and c["num_bodies"] == n
    after = bh.bodies()
    assert np.isfinite(after).all()
    # the tree of the last step's bodies
    bh.build_current_tree()
    keys, idx = bh.morton(), bh.index()
    assert np.all(keys[1:] >= keys[:-1])
    assert np.array_equal(np.sort(idx), np.arange(n, dtype=np.uint32))
    del keys, idx
    f = O.node_fields(bh.nodes(B.INTERNAL_NODES))
    count = f["bodiesContained"].astype(np.int64)
    a, b = f["childA"].astype(np.int64), f["childB"].astype(np.int64)
    ca = np.where(a < n, 1, count[np.clip(a - n, 0, n - 2)])
    cb = np.where(b < n, 1, count[np.clip(b - n, 0, n - 2)])
    np.testing.assert_array_equal(count, ca + cb)
    assert count[0] == n and f["parentId"][0] == 0xFFFFFFFF
    np.testing.assert_allclose(f["mass"][0], after[:, 3].astype(np.float64).sum(), rtol=1e-5)
    del f, count, a, b, ca, cb
    # force error of the walk against the FP64 judge, 16,384 sampled bodies
    rng = np.random.default_rng(9)
    sample = np.sort(rng.choice(n, size=16384, replace=False)).astype(np.uint32)
    acc = bh.accelerations()[sample].astype(np.float64)
    ref = bh.direct_sum_f64(sample, 0.01)
    err = np.linalg.norm(acc - ref, axis=1) / np.linalg.norm(ref, axis=1)
    p50, p99 = np.percentile(err, [50, 99])
    print(f"[gpu] two galaxies full size: p50 {p50:.3g} p99 {p99:.3g}", flush=True)
    assert p50 < 0.05 and p99 < 0.2
    bh.close()


def test_gpu_unit_derivation_matches_oracle_and_takes_explicit_bits():
    """The library derives G and the body-length constant from the unit
    uniforms (at bh_create and at every bh_reset_values that sets units) with
    the oracle's FP32 expression, bit for bit; explicit values (here the
    reference shader's own bits from tests/golden/units.npz) are taken as
    given, which is the bit-exact contract across GL drivers."""
    import os
    from gravity_amd import barnes_hut as B
    from gravity_amd import initial_conditions as ic
    d = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "units.npz"))
    bodies = ic.plummer(64, 1)
    bounds = ic.tight_bounds(bodies)
    f32 = lambda x: np.float32(x).view(np.uint32)   # noqa: E731
    other = ic.Units(mass=3.0, density=0.25, len=0.5, time=2.0)
    for name, u, g_ref, b_ref in zip(d["names"], d["uniforms"], d["G"], d["blc"]):
        print(f"[gpu] units {name}", flush=True)
        units = ic.Units(*[float(x) for x in u])
        g, b = O.units(*[float(x) for x in u])
        bh = B.BarnesHut(bodies, bounds, B.make_params(units=other))
        s = bh.sim_values().view(np.float32)
        g0, b0 = O.units(other.mass, other.density, other.len, other.time)
        assert f32(s[21]) == f32(g0) and f32(s[22]) == f32(b0), "bh_create"
        bh.reset_values(B.make_params(units=units))
        s = bh.sim_values().view(np.float32)
        assert f32(s[21]) == f32(g) and f32(s[22]) == f32(b), f"{name}: bh_reset_values"
        assert tuple(s[16:20]) == tuple(np.float32(x) for x in u)
        bh.close()
        if g_ref != 0.0:
            bh = B.BarnesHut(bodies, bounds, B.make_params(units=units, G=float(g_ref),
                                                           body_length_constant=float(b_ref)))
            s = bh.sim_values().view(np.float32)
            assert f32(s[21]) == f32(g_ref) and f32(s[22]) == f32(b_ref), f"{name}: explicit bits"
            bh.close()
