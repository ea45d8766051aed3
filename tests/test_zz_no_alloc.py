"""bh_direct_sum_f64_noalloc, the entry point that allocates no device memory,
against bh_direct_sum_f64 through the C ABI: two calls in a row (the second
with a smaller sample, so the static scratch still holds the first call's
larger contents), each bit-identical to the allocating entry point, plus the
size limit."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _call(L, ctx, fn, sample, soft):
    s = np.ascontiguousarray(sample, dtype=np.uint32)
    out = np.full((s.shape[0], 3), np.nan, np.float64)
    rc = getattr(L, fn)(ctx, s.ctypes.data_as(C.c_void_p), s.shape[0], soft, out.ctypes.data_as(C.c_void_p))
    return rc, out


def test_noalloc_direct_sum_matches_allocating_entry_point():
    from gravity_amd import barnes_hut as B
    from gravity_amd import initial_conditions as ic
    n = 16384
    print(f"[gpu] no-allocation direct sum n={n}", flush=True)
    bodies = ic.plummer(n, 31)
    bodies[::97, 3] = 0.0                      # some dead bodies: skipped by both
    bh = B.BarnesHut(bodies, ic.tight_bounds(bodies))
    bh.init()
    L, ctx = bh.L, bh.ctx
    rng = np.random.default_rng(5)
    for k, m in enumerate((4096, 333)):        # the second call is smaller than the first
        sample = rng.choice(n, size=m, replace=False).astype(np.uint32)
        rc0, want = _call(L, ctx, "bh_direct_sum_f64", sample, 0.01)
        rc1, got = _call(L, ctx, "bh_direct_sum_f64_noalloc", sample, 0.01)
        assert rc0 == 0 and rc1 == 0, (rc0, rc1, L.bh_last_error(ctx))
        assert np.isfinite(got).all()
        assert np.array_equal(got.view(np.uint64), want.view(np.uint64)), f"call {k}: results differ"
    # the size limit is an error with a message, not a fault
    big = np.zeros(B.NOALLOC_MAX_SAMPLE + 1, np.uint32)
    rc, _ = _call(L, ctx, "bh_direct_sum_f64_noalloc", big, 0.01)
    assert rc == -5
    assert b"static scratch" in L.bh_last_error(ctx)
    # an id out of range is rejected on the host
    rc, _ = _call(L, ctx, "bh_direct_sum_f64_noalloc", np.array([n], np.uint32), 0.01)
    assert rc == -1
    bh.close()
