"""The tree's upper levels where a group of first-launch blocks leaves more
subtrees than one window of the level kernel takes (1,024): sorted keys
built so that every 512-position block leaves about 49 maximal subtrees
(Plummer spheres leave about 10), so each block of the middle level runs
several windows and hands what they leave to the last level. The build from
uploaded keys is compared bit for bit with the oracle's build and fill
(every level filled).

Reference: treeBuildBinaryRadixTreeKernel / treeInitLeafNodesKernel /
treePropagateNodesKernel (`shaders/compute/tree/bh_tree.comp:43-259`).
"""
import numpy as np
import pytest

from oracle import bh_oracle as O

pytestmark = pytest.mark.gpu

BLOCK = 512        # first-launch block of the plain tree above 131,072 bodies
WINDOW = 1024      # entries per window of the level kernel
GROUP = 64         # first-launch blocks per block of the level kernel


def caterpillar_keys(n, k=24):
    """Sorted 64-bit keys whose tree around every block boundary p is a
    two-sided caterpillar: the split between p-1 and p is the deepest and the
    splits outward alternate sides (flips of bit 4 at p, of the odd bits 5, 7,
    .. to its right and the even bits 6, 8, .. to its left), so each of the
    leaves on either side is a subtree of its own whose parent crosses p. The
    rest of each block is one run of equal keys; a block counter in bits 54
    and up separates the blocks."""
    keys = np.zeros(n, np.uint64)
    key = 0
    for j in range(n - 1):                # split j: between positions j and j + 1
        r = (j + 1) % BLOCK
        if r == 0:
            bit = 4                       # a block boundary: the deepest split
        elif r < k:
            bit = 5 + 2 * (r - 1)         # right of it: 5, 7, ..
        elif r == BLOCK - k:
            bit = 54                      # towards the next block: the counter
        elif r > BLOCK - k:
            bit = 4 + 2 * (BLOCK - r)     # left of the next boundary: .., 8, 6
        else:
            bit = None                    # inside the run of equal keys
        if bit == 54:
            key = ((key >> 54) + 1) << 54
        elif bit is not None:
            assert not (key >> bit) & 1
            key = ((key >> bit) | 1) << bit
        keys[j + 1] = key
    return keys


def maximal_subtrees_per_block(keys):
    """Per block, the maximal subtrees of the sorted keys' radix tree that lie
    inside it (what the first launch leaves), from the split lengths."""
    n = keys.size
    x = keys[:-1] ^ keys[1:]
    d = np.empty(n - 1, np.int64)
    for s, v in enumerate(x.tolist()):
        d[s] = 64 - int(v).bit_length() if v else 64 + 32 - int(s ^ (s + 1)).bit_length()
    # node of split s covers (previous smaller split, next smaller split]
    S = n - 1
    pse = np.full(S, -1, np.int64)
    nse = np.full(S, S, np.int64)
    st = []
    for s in range(S):
        while st and d[st[-1]] > d[s]:
            nse[st.pop()] = s
        pse[s] = st[-1] if st else -1
        st.append(s)
    lo, hi = pse + 1, nse
    dv = lambda a: np.where((a < 0) | (a >= S), -1, d[np.clip(a, 0, S - 1)])
    par = np.where(dv(pse) > dv(nse), pse, nse)
    ok = (par >= 0) & (par < S)
    plo = np.where(ok, lo[np.clip(par, 0, S - 1)], 0)
    phi = np.where(ok, hi[np.clip(par, 0, S - 1)], n - 1)
    inside = lo // BLOCK == hi // BLOCK
    maximal = inside & ((plo // BLOCK != phi // BLOCK) | ~ok)
    pos = np.arange(n)
    ps = np.clip(np.where(dv(pos - 1) > dv(pos), pos - 1, pos), 0, S - 1)
    leaves = lo[ps] // BLOCK != hi[ps] // BLOCK
    counts = np.bincount(lo[maximal] // BLOCK, minlength=n // BLOCK)
    counts += np.bincount(pos[leaves] // BLOCK, minlength=n // BLOCK)
    return counts


def test_gpu_tree_levels_with_several_windows():
    from gravity_amd import barnes_hut as B
    from gravity_amd import initial_conditions as ic
    n = 512 * BLOCK                          # 512 first-launch blocks, 8 level blocks
    keys = caterpillar_keys(n)
    per_block = maximal_subtrees_per_block(keys)
    assert per_block[1:-1].min() > WINDOW // GROUP, per_block[:8]   # more than a window per group
    print(f"[gpu] tree levels n={n}: {per_block.mean():.1f} subtrees per block", flush=True)
    bodies = ic.plummer(n, 7)
    idx = np.random.default_rng(3).permutation(n).astype(np.uint32)
    bh = B.BarnesHut(bodies, ic.tight_bounds(bodies), B.make_params(theta=0.5, dt=1e-3, softening=0.01, G=1.0))
    bh.init()
    bh.upload(B.MORTON_IN, keys)
    bh.upload(B.INDEX_IN, idx)
    bh.radix_sort()                          # (stable: keys and ids as uploaded)
    np.testing.assert_array_equal(bh.morton(), keys)
    bh.build_tree()
    c = bh.counters()
    assert c["error_flags"] == 0
    got_i = bh.nodes(B.INTERNAL_NODES)[: n - 1]
    got_l = bh.nodes(B.LEAF_NODES)[:n]
    leaf, internal = O.build(keys, n, n)
    O.fill(bodies, bh.index().copy(), n, n, leaf, internal, max_passes=256, with_colors=False)
    np.testing.assert_array_equal(got_i[:, 10:16], internal[:, 10:16])   # topology, height, count
    np.testing.assert_array_equal(got_l[:, 10:16], leaf[:, 10:16])
    np.testing.assert_array_equal(got_i[:, 0:10], internal[:, 0:10])     # centre of mass, mass, box (bits)
    np.testing.assert_array_equal(got_l[:, 0:10], leaf[:, 0:10])
    assert c["tree_height"] == int(internal[0, 12])
    bh.close()
