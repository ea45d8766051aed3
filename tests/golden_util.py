"""Loading helpers for the llvmpipe golden fixtures (tests/golden/*.npz)."""
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FULL = ["plummer_4096", "merge_4096", "collision_4096", "edge_dead_nan_1000", "edge_duplicates_512",
        "static_kill_1024", "static_wrap_1024", "theta0_512", "two_galaxy_4096", "tall_tree_95"]


def load(name):
    return dict(np.load(os.path.join(GOLDEN, name + ".npz")))


def steps(fx):
    return int(fx["modes"][3])


def stage_inputs(fx, s):
    """Bodies and index the reference's step s started from (before the dead partition)."""
    n = fx["bodies0"].shape[0]
    if s == 0:
        return fx["bodies0"].copy(), np.arange(n, dtype=np.uint32), n
    prev = fx[f"s{s - 1}_merge_out"].copy()
    idx = np.zeros(n, np.uint32)
    srt = fx[f"s{s - 1}_index_sorted"]
    idx[: srt.shape[0]] = srt
    return prev, idx, int(fx[f"s{s}_num_before"])


def num_alive(fx, s):
    return int(fx[f"s{s}_num_before"]) - int(fx[f"s{s}_died"])


def params_of(fx):
    theta, dt, soft, elasticity = [float(x) for x in fx["params"]]
    merging, dynamic, wrap, _ = [int(x) for x in fx["modes"]]
    return dict(theta=theta, dt=dt, softening=soft, elasticity=elasticity, merging=merging, dynamic=dynamic,
                wrap_around=wrap)
