// bh_kernels.hip — the Barnes-Hut step as hand-written HIP kernels for gfx950.
//
// Stages (reference program in brackets; shaders/compute/...):
//   k_init                 [COMPUTE_INIT, init/bh_init.comp:10-28]
//   k_reset                [COMPUTE_UPDATE flag=1, update/bh_update.comp:10-37]
//   k_dead_count/scan/scatter
//                          [COMPUTE_DEAD_*, dead/bh_dead.comp:13-112 + UPDATE flag=0]
//   k_bounds               [COMPUTE_MORTON_AABB_REPOPULATE + COLLAPSE, morton/bh_morton.comp:50-231]
//   k_encode               [COMPUTE_MORTON_ENCODE, morton/bh_morton.comp:6-34,238-262]
//   k_sort_digits/bases/onesweep
//                          [COMPUTE_RADIX_*, radix/bh_radix.comp:12-166]
//   k_tree/k_tree_mid      [COMPUTE_TREE_BUILD + INIT_LEAVES + PROPAGATE, tree/bh_tree.comp:6-259]
//   (walk: bh_walk.hip)    [COMPUTE_FORCE_COMPUTE]
//   k_merge                [COMPUTE_MERGE_BODIES, merge/bh_merge.comp:5-71]
//   k_node_to_std430 / k_std430_to_node: the 80-byte Node layout at the boundary.
#include "bh_device.hpp"
#include "bh_kernels.hpp"

namespace bh {

// ------------------------------------------------------------------ init/reset
__global__ void k_init(uint32_t* __restrict__ index_in, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) index_in[i] = i;
}

__global__ void k_reset(SimValues* __restrict__ sim, DevCounters* __restrict__ ctr, int set_units, float mass,
                        float density, float len, float time, float G, float blc) {
    if (threadIdx.x != 0) return;
    ctr->opened = 0;
    ctr->interactions = 0;
    ctr->nTasks = 0;
    ctr->mergesApplied = 0;
    ctr->maxStack = 0;
    ctr->errorFlags = 0;
    ctr->corrected = 0;
    ctr->quadTerms = 0;
    if (set_units) {
        sim->unitMass = mass;
        sim->unitDensity = density;
        sim->unitLen = len;
        sim->unitTime = time;
        sim->gravitationalConstant = G;
        sim->bodyLengthConstant = blc;
    }
}

// ------------------------------------------------------------------ dead partition
// The reference examines sorted positions [0, numBodies) through the index
// and moves the bodies of zero mass behind the others, stably, setting them
// EMPTY in both body buffers. Bodies at positions >= numBodies are always
// EMPTY (a body leaves [0, numBodies) only through this partition), so the
// census reads the bodies in id order (coalesced, no gather through the
// index): it sets every zero-mass body EMPTY (idempotent for those already
// dead) and counts them. If that count equals n - numBodies nobody died and
// the partition is the identity: the scatter copies the index. Otherwise the
// stable partition runs in index order.

__device__ __forceinline__ bool nobody_died(const SimValues* sim, const DevCounters* ctr) {
    return ctr->zeroMassIds == sim->initialNumBodies - sim->numBodies;
}

// Tile of 1024 sorted positions per block of 256 threads: alive bodies per tile.
__global__ __launch_bounds__(256) void k_dead_count(const SimValues* __restrict__ sim,
                                                    const DevCounters* __restrict__ ctr,
                                                    const float4* __restrict__ src,
                                                    const uint32_t* __restrict__ index_in,
                                                    uint32_t* __restrict__ block_alive) {
    __shared__ uint32_t s_w[4];
    const uint32_t num = sim->numBodies;
    const uint32_t base = blockIdx.x * kPartTile;
    if (base >= num || nobody_died(sim, ctr)) return;
    uint32_t alive = 0;
    for (int r = 0; r < kPartTile / 256; r++) {
        uint32_t e = base + r * 256 + threadIdx.x;
        if (e < num && src[2 * index_in[e]].w != 0.0f) alive++;
    }
    for (int off = 32; off > 0; off >>= 1) alive += __shfl_xor(alive, off, 64);
    if (lane_id() == 0) s_w[threadIdx.x >> 6] = alive;
    __syncthreads();
    if (threadIdx.x == 0) block_alive[blockIdx.x] = s_w[0] + s_w[1] + s_w[2] + s_w[3];
}

// One block of 1024 threads: exclusive scan of the per-tile alive counts;
// then numBodies -= justDied, justDied = 0, uintDebug[0] = 0 (bh_update.comp:31-34).
__global__ __launch_bounds__(1024) void k_dead_scan(SimValues* __restrict__ sim, DevCounters* __restrict__ ctr,
                                                    uint32_t* __restrict__ block_alive) {
    __shared__ uint32_t s_w[16];
    __shared__ uint32_t s_total;
    const uint32_t num = sim->numBodies;
    if (nobody_died(sim, ctr)) {
        __syncthreads();   // every thread has read the census before it is cleared
        if (threadIdx.x == 0) {
            ctr->numBefore = num;
            ctr->died = 0;
            ctr->zeroMassIds = 0;
            sim->justDied = 0;
            sim->uintDebug[0] = 0;
        }
        return;
    }
    const uint32_t nb = (num + kPartTile - 1) / kPartTile;
    const uint32_t per = (nb + 1023) / 1024;
    const uint32_t t = threadIdx.x;
    uint32_t start = t * per, end = min(start + per, nb);
    uint32_t sum = 0;
    for (uint32_t k = start; k < end; k++) sum += block_alive[k];
    // block exclusive scan of sum
    uint32_t incl = sum;
    for (int off = 1; off < 64; off <<= 1) {
        uint32_t v = __shfl_up(incl, off, 64);
        if (lane_id() >= (uint32_t)off) incl += v;
    }
    if (lane_id() == 63) s_w[t >> 6] = incl;
    __syncthreads();
    if (t < 16) {
        uint32_t v = s_w[t];
        for (int off = 1; off < 16; off <<= 1) {
            uint32_t u = __shfl_up(v, off, 16);
            if (t >= (uint32_t)off) v += u;
        }
        s_w[t] = v;  // inclusive over waves
        if (t == 15) s_total = v;
    }
    __syncthreads();
    uint32_t run = incl - sum + ((t >> 6) ? s_w[(t >> 6) - 1] : 0);
    for (uint32_t k = start; k < end; k++) {
        uint32_t v = block_alive[k];
        block_alive[k] = run;
        run += v;
    }
    if (t == 0) {
        uint32_t alive = s_total;
        ctr->numBefore = num;
        ctr->died = num - alive;
        ctr->zeroMassIds = 0;
        sim->numBodies = alive;
        sim->justDied = 0;
        sim->uintDebug[0] = 0;
    }
}

// Stable scatter: alive first in order, dead after them in order (bh_dead.comp:141-181).
// When nobody died it is a copy of positions [0, numBodies).
__global__ __launch_bounds__(256) void k_dead_scatter(const SimValues* __restrict__ sim,
                                                      const DevCounters* __restrict__ ctr,
                                                      const float4* __restrict__ src,
                                                      const uint32_t* __restrict__ index_in,
                                                      uint32_t* __restrict__ index_out,
                                                      const uint32_t* __restrict__ block_alive_off) {
    __shared__ uint32_t s_wa[4], s_wd[4];
    const uint32_t num = ctr->numBefore;
    const uint32_t alive_total = sim->numBodies;
    const uint32_t base = blockIdx.x * kPartTile;
    if (base >= num) return;
    if (ctr->died == 0) {
        for (int r = 0; r < kPartTile / 256; r++) {
            const uint32_t e = base + r * 256 + threadIdx.x;
            if (e < num) index_out[e] = index_in[e];
        }
        return;
    }
    const uint32_t aoff = block_alive_off[blockIdx.x];
    const uint32_t doff = base - aoff;  // dead bodies in earlier tiles (all earlier tiles are full)
    const uint32_t w = threadIdx.x >> 6;
    uint32_t run_a = 0, run_d = 0;
    for (int r = 0; r < kPartTile / 256; r++) {
        uint32_t e = base + r * 256 + threadIdx.x;
        bool in = e < num;
        uint32_t b = in ? index_in[e] : 0;
        bool alive = in && src[2 * b].w != 0.0f;
        bool dead = in && !alive;
        uint64_t ma = ballot(alive), md = ballot(dead);
        uint32_t ra = __popcll(ma & lanemask_lt()), rd = __popcll(md & lanemask_lt());
        if (lane_id() == 0) {
            s_wa[w] = __popcll(ma);
            s_wd[w] = __popcll(md);
        }
        __syncthreads();
        uint32_t pa = 0, pd = 0, ta = 0, td = 0;
        for (uint32_t k = 0; k < 4; k++) {
            if (k < w) { pa += s_wa[k]; pd += s_wd[k]; }
            ta += s_wa[k];
            td += s_wd[k];
        }
        if (alive) index_out[aoff + run_a + pa + ra] = b;
        if (dead) index_out[alive_total + doff + run_d + pd + rd] = b;
        run_a += ta;
        run_d += td;
        __syncthreads();
    }
}

// ------------------------------------------------------------------ AABB
__device__ __forceinline__ void window_axis(float pmin, float pmax, float& lo, float& hi) {
    // mortonInlierWindow, bh_morton.comp:50-56
    float center = fmul(0.5f, fadd(pmin, pmax));
    float half = gmax(fmul(fmul(0.5f, fsub(pmax, pmin)), 1.05f), 1e-9f);
    lo = fsub(center, half);
    hi = fadd(center, half);
}

// Inlier window of the previous bounds (bh_morton.comp:50-63): lo/hi per axis.
struct BoundsWindow {
    float lo[3], hi[3];
    bool prev_empty;
};

__device__ __forceinline__ BoundsWindow bounds_window(const float* pmin, const float* pmax) {
    BoundsWindow W;
    W.prev_empty = pmin[0] > pmax[0];
    for (int c = 0; c < 3; c++) window_axis(pmin[c], pmax[c], W.lo[c], W.hi[c]);
    return W;
}

// fold one body into the running min (r[0..2]) / max (r[3..5])
__device__ __forceinline__ void bounds_accum(const BoundsWindow& W, float4 pm, float* r) {
    if (pm.w == 0.0f) return;                          // isEmpty
    if (!finite3(pm.x, pm.y, pm.z)) return;            // mortonContributorAABB :61-63
    if (!W.prev_empty && (pm.x < W.lo[0] || pm.y < W.lo[1] || pm.z < W.lo[2] || pm.x > W.hi[0] ||
                          pm.y > W.hi[1] || pm.z > W.hi[2]))
        return;
    r[0] = gmin(r[0], pm.x); r[1] = gmin(r[1], pm.y); r[2] = gmin(r[2], pm.z);
    r[3] = gmax(r[3], pm.x); r[4] = gmax(r[4], pm.y); r[5] = gmax(r[5], pm.z);
}

// min/max over the 256 threads of a block; the result is valid in thread 0
__device__ __forceinline__ void bounds_block_reduce(float* r, float (*s_r)[4]) {
    for (int off = 32; off > 0; off >>= 1) {
        for (int k = 0; k < 3; k++) r[k] = gmin(r[k], __shfl_xor(r[k], off, 64));
        for (int k = 3; k < 6; k++) r[k] = gmax(r[k], __shfl_xor(r[k], off, 64));
    }
    const uint32_t w = threadIdx.x >> 6;
    if (lane_id() == 0)
        for (int k = 0; k < 6; k++) s_r[k][w] = r[k];
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 0; k < 3; k++) r[k] = gmin(gmin(s_r[k][0], s_r[k][1]), gmin(s_r[k][2], s_r[k][3]));
        for (int k = 3; k < 6; k++) r[k] = gmax(gmax(s_r[k][0], s_r[k][1]), gmax(s_r[k][2], s_r[k][3]));
    }
}

// New bounds from the reduced inliers res and the previous bounds
// (bh_morton.comp:77-88, 205-231).
__device__ __forceinline__ void bounds_collapse(const float* prev, const float* res, float* mn, float* mx) {
    const bool prev_empty = prev[0] > prev[3];
    if (res[0] > res[3]) {                               // no finite inliers (:221-226)
        if (prev_empty)
            for (int c = 0; c < 3; c++) { mn[c] = res[c]; mx[c] = res[3 + c]; }
        return;
    }
    if (prev_empty) {
        for (int c = 0; c < 3; c++) { mn[c] = res[c]; mx[c] = res[3 + c]; }
        return;
    }
    for (int c = 0; c < 3; c++) {                        // clampAABBExtentGrowth :77-88
        float prevExt = gmax(fsub(prev[3 + c], prev[c]), 1e-9f);
        float maxExt = fmul(prevExt, 1.1f);
        float ctr_c = fmul(0.5f, fadd(res[c], res[3 + c]));
        float h = fmul(0.5f, fsub(res[3 + c], res[c]));
        float mh = fmul(0.5f, maxExt);
        h = gmin(h, mh);
        mn[c] = fsub(ctr_c, h);
        mx[c] = fadd(ctr_c, h);
    }
}

// The block's min/max (r) to partials; the last block to finish reduces
// every block's partial (min/max: order-free) and collapses the bounds. All
// 256 threads call it.
__device__ void bounds_finish(SimValues* __restrict__ sim, DevCounters* __restrict__ ctr,
                              float* __restrict__ partials, const float* prev, float* r) {
    __shared__ float s_r[6][4];
    __shared__ bool s_last;
    bounds_block_reduce(r, s_r);
    if (threadIdx.x == 0) {
        for (int k = 0; k < 6; k++)
            __hip_atomic_store(&partials[6 * blockIdx.x + k], r[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        uint32_t ticket = __hip_atomic_fetch_add(&ctr->aabbTicket, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        s_last = ticket == gridDim.x - 1;
    }
    __syncthreads();
    if (!s_last) return;
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    float res[6] = {1e38f, 1e38f, 1e38f, -1e38f, -1e38f, -1e38f};
    for (uint32_t k = threadIdx.x; k < gridDim.x; k += blockDim.x) {
        for (int c = 0; c < 3; c++)
            res[c] = gmin(res[c], __hip_atomic_load(&partials[6 * k + c], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
        for (int c = 3; c < 6; c++)
            res[c] = gmax(res[c], __hip_atomic_load(&partials[6 * k + c], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    }
    bounds_block_reduce(res, s_r);
    if (threadIdx.x != 0) return;
    ctr->aabbTicket = 0;
    bounds_collapse(prev, res, sim->minCorner, sim->maxCorner);
}

// The census of the dead partition (see above), here because it can also
// make the bounds.
// With kBounds the same pass also makes the bounds of k_bounds (which skips
// the zero-mass bodies the census counts) and collapses them.
//
// step_pos: the previous walk's positions by sorted position, when they are
// the bodies' (no merge since): the census then reads them, 16 bytes a body
// in order, instead of the 32-byte bodies by id. A body of zero mass there is
// one that died (the walk killed it); it is found by its id and set EMPTY in
// both buffers, and the bodies behind numBodies, EMPTY already, are counted
// without reading them.
template <bool kBounds>
__global__ __launch_bounds__(256) void k_dead_census(SimValues* __restrict__ sim, DevCounters* __restrict__ ctr,
                                                     float4* __restrict__ src, float4* __restrict__ dst,
                                                     float* __restrict__ partials,
                                                     const float4* __restrict__ step_pos,
                                                     const uint32_t* __restrict__ index_in) {
    __shared__ uint32_t s_w[4];
    const uint32_t n = sim->initialNumBodies;
    float prev[6], r[6] = {1e38f, 1e38f, 1e38f, -1e38f, -1e38f, -1e38f};
    BoundsWindow W;
    if constexpr (kBounds) {
        for (int k = 0; k < 3; k++) {
            prev[k] = sim->minCorner[k];
            prev[3 + k] = sim->maxCorner[k];
        }
        W = bounds_window(prev, prev + 3);
    }
    uint32_t zero = 0;
    // bh_dead.comp:99-102: a dead body is EMPTY in both buffers
    auto empty = [&](uint32_t b) {
        src[2 * b] = make_float4(0.f, 0.f, 0.f, 0.f);
        src[2 * b + 1] = make_float4(0.f, 0.f, 0.f, 0.f);
        dst[2 * b] = make_float4(0.f, 0.f, 0.f, 0.f);
        dst[2 * b + 1] = make_float4(0.f, 0.f, 0.f, 0.f);
    };
    if (step_pos) {
        const uint32_t num = sim->numBodies;
        if (blockIdx.x == 0 && threadIdx.x == 0) zero = n - num;
        for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < num; g += gridDim.x * blockDim.x) {
            const float4 pm = step_pos[g];
            if constexpr (kBounds) bounds_accum(W, pm, r);
            if (pm.w == 0.0f) {
                empty(index_in[g]);
                zero++;
            }
        }
    } else {
        for (uint32_t b = blockIdx.x * blockDim.x + threadIdx.x; b < n; b += gridDim.x * blockDim.x) {
            const float4 pm = src[2 * b];
            if constexpr (kBounds) bounds_accum(W, pm, r);
            if (pm.w == 0.0f) {
                empty(b);
                zero++;
            }
        }
    }
    for (int off = 32; off > 0; off >>= 1) zero += __shfl_xor(zero, off, 64);
    if (lane_id() == 0) s_w[threadIdx.x >> 6] = zero;
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t t = s_w[0] + s_w[1] + s_w[2] + s_w[3];
        if (t) atomicAdd(&ctr->zeroMassIds, t);
    }
    if constexpr (kBounds) bounds_finish(sim, ctr, partials, prev, r);
}

__global__ __launch_bounds__(256) void k_bounds(SimValues* __restrict__ sim, DevCounters* __restrict__ ctr,
                                                const float4* __restrict__ src,
                                                const uint32_t* __restrict__ index_in,
                                                float* __restrict__ partials) {
    const uint32_t num = sim->numBodies;
    const float prev[6] = {sim->minCorner[0], sim->minCorner[1], sim->minCorner[2],
                           sim->maxCorner[0], sim->maxCorner[1], sim->maxCorner[2]};
    const BoundsWindow W = bounds_window(prev, prev + 3);
    float r[6] = {1e38f, 1e38f, 1e38f, -1e38f, -1e38f, -1e38f};
    // The alive bodies are those at positions [0, numBodies); every body at a
    // later position is EMPTY, which bounds_accum skips, so reading all ids in
    // order (coalesced) covers the same bodies as the gather through the index.
    (void)index_in;
    (void)num;
    const uint32_t n = sim->initialNumBodies;
    for (uint32_t b = blockIdx.x * blockDim.x + threadIdx.x; b < n; b += gridDim.x * blockDim.x)
        bounds_accum(W, src[2 * b], r);
    bounds_finish(sim, ctr, partials, prev, r);
}

// ------------------------------------------------------------------ Morton
__device__ __forceinline__ uint64_t morton_of(float4 pm, const float* bmin, const float* bmax, bool empty) {
    const float MAXV = 2097151.0f;
    float p[3] = {pm.x, pm.y, pm.z};
    if (!finite3(pm.x, pm.y, pm.z) || empty) {
        for (int k = 0; k < 3; k++) p[k] = fmul(0.5f, fadd(bmin[k], bmax[k]));
    } else {
        for (int k = 0; k < 3; k++) p[k] = gmin(gmax(p[k], bmin[k]), bmax[k]);
    }
    uint32_t q[3];
    for (int k = 0; k < 3; k++) {
        float ext = gmax(fsub(bmax[k], bmin[k]), 1e-9f);
        float pn = fdiv(fsub(p[k], bmin[k]), ext);
        float f = gmin(gmax(floorf(fmul(pn, MAXV)), 0.0f), MAXV);
        q[k] = (uint32_t)floorf(f);
    }
    return (expand_bits21(q[0]) << 2) | (expand_bits21(q[1]) << 1) | expand_bits21(q[2]);
}

// The body at sorted position g is src[index_in[g]]; step_pos, when given
// and the partition before this encode removed nobody (so that the order of
// positions is the one the last walk wrote them in), holds the same values
// by position and is read in order instead of gathered.
__global__ __launch_bounds__(256) void k_encode(const SimValues* __restrict__ sim,
                                                const DevCounters* __restrict__ ctr, const float4* __restrict__ src,
                                                const uint32_t* __restrict__ index_in,
                                                const float4* __restrict__ step_pos,
                                                uint64_t* __restrict__ morton_in) {
    const uint32_t num = sim->numBodies;
    uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= num) return;
    float bmin[3] = {sim->minCorner[0], sim->minCorner[1], sim->minCorner[2]};
    float bmax[3] = {sim->maxCorner[0], sim->maxCorner[1], sim->maxCorner[2]};
    const float4 pm = (step_pos && ctr->died == 0) ? step_pos[g] : src[2 * index_in[g]];
    morton_in[g] = morton_of(pm, bmin, bmax, bmin[0] > bmax[0]);
}

// ------------------------------------------------------------------ radix sort
// Stable LSD sort of (key, index) pairs over positions [0, numBodies), eight
// passes of eight bits. One simulation: one histogram launch reads every key
// once and counts all eight digits, a small launch sums the blocks' counts,
// then each pass is one launch: tiles take their ids in launch order, scan the
// pass's digit totals, rank their keys in-wave by ballots, publish their digit counts and
// find the counts of all earlier tiles by a decoupled look-back, then write
// their keys out in coalesced digit runs. The ensembles keep tile histograms
// and a scan per simulation (k_ens_sort_*).
// Digit histogram of one tile of up to kSortTile keys, per wave without
// atomics: keys of a tile share their high digits (the input is nearly in
// order), so per-key LDS atomics would serialize on one bin. Each row of 64
// keys finds the lanes holding its digit (one ballot per digit bit) and the
// lowest of them adds the group's size to the wave's own bins (distinct
// digits: distinct addresses). Called by all 256 threads; returns thread t's
// digit count for the tile.
template <int BITS, int TILE = kSortTile, int W = 4>
__device__ __forceinline__ uint32_t tile_histogram(uint32_t (&s_h)[W][256], const uint64_t* __restrict__ keys,
                                                   uint32_t start, uint32_t count, uint32_t shift) {
    constexpr uint32_t R = 1u << BITS;
    constexpr int kPer = TILE / W, kRows = kPer / 64;
    const uint32_t t = threadIdx.x, w = t >> 6, lane = lane_id();
    const uint64_t lt = lanemask_lt();
    for (int k = 0; k < 4; k++) s_h[w][lane * 4 + k] = 0;
    uint64_t key[kRows];
    for (int r = 0; r < kRows; r++) {
        const uint32_t e = w * kPer + r * 64 + lane;
        key[r] = e < count ? keys[start + e] : 0ull;
    }
    for (int r = 0; r < kRows; r++) {
        const uint32_t e = w * kPer + r * 64 + lane;
        const bool valid = e < count;
        const uint32_t d = (uint32_t)(key[r] >> shift) & (R - 1u);
        uint64_t peers = ballot(valid);
        for (int bit = 0; bit < BITS; bit++) {
            const uint64_t bb = ballot(valid && ((d >> bit) & 1u));
            peers &= ((d >> bit) & 1u) ? bb : ~bb;
        }
        if (valid && __popcll(peers & lt) == 0) s_h[w][d] += (uint32_t)__popcll(peers);
    }
    __syncthreads();
    uint32_t c = 0;
    if (t < 256)
        for (int k = 0; k < W; k++) c += s_h[k][t];
    return c;
}

// Stable scatter of one tile of up to TILE keys. Wave w owns the tile's
// elements [w TILE/4, (w + 1) TILE/4) and ranks them by digit on its own (in-wave
// ballot matching, a per-wave running count per digit in LDS, no block
// barrier); the block then turns the four waves' digit counts into per-wave
// offsets and the tile's digit starts, reorders the tile by digit in LDS and
// writes it out so that consecutive threads write consecutive positions of
// each digit's run. The order within a digit is the element order: stable.
// L.base[d]: the global position of the tile's first key with digit d.
template <int TILE, int W = 4>
struct ScatterLds {
    uint64_t key[TILE];
    uint32_t idx[TILE];
    uint32_t base[256], start[256];
    uint32_t cnt[W][256];   // per-wave digit counts, then per-wave offsets
    uint32_t w_tot[4];      // the digit scan over 256 digits (4 waves)
};

// Positions travel through the sort multiplied by an odd constant (a
// bijection of 32-bit words): on MI355X the scatter passes run ~7 % slower
// when their values are the positions themselves (small, nearly ordered
// words) than with these scrambled ones, measured over whole sorts.
constexpr uint32_t kPosScramble = 0x9E3779B1u, kPosUnscramble = 0x0E8B2F51u;
static_assert(kPosScramble * kPosUnscramble == 1u, "inverse mod 2^32");

// The tile's digit counts before they turn into global positions: a scatter
// whose bases are known up front (the ensembles) passes NoLookback; the
// one-simulation sort passes a TileLookback, which publishes the counts and
// adds those of all earlier tiles.
struct NoLookback {
    static constexpr bool kActive = false;
    __device__ void publish(uint32_t, uint32_t) const {}
    __device__ uint32_t resolve(uint32_t, uint32_t) const { return 0; }
};

// idx_in nullptr: the values start as the scrambled positions start + e.
template <int BITS, int TILE, int W = 4, class Lookback = NoLookback>
__device__ __forceinline__ void scatter_tile(ScatterLds<TILE, W>& L, const uint64_t* __restrict__ keys_in,
                                             const uint32_t* __restrict__ idx_in, uint32_t start, uint32_t count,
                                             uint64_t* __restrict__ keys_out, uint32_t* __restrict__ idx_out,
                                             uint32_t shift, const Lookback& lb = Lookback()) {
    constexpr uint32_t R = 1u << BITS;
    constexpr int kPer = TILE / W;         // elements per wave
    constexpr int kRows = kPer / 64;       // rows of 64 per wave
    const uint32_t t = threadIdx.x, w = t >> 6, lane = lane_id();
    const uint64_t lt = lanemask_lt();
    for (int k = 0; k < 4; k++) L.cnt[w][lane * 4 + k] = 0;   // this wave's counts
    uint64_t key[kRows];
    uint32_t idx[kRows], rank[kRows];
    for (int r = 0; r < kRows; r++) {
        const uint32_t e = w * kPer + r * 64 + lane;
        const bool valid = e < count;
        key[r] = valid ? keys_in[start + e] : 0ull;
        idx[r] = valid ? (idx_in ? idx_in[start + e] : (start + e) * kPosScramble) : 0u;
    }
    for (int r = 0; r < kRows; r++) {
        const uint32_t e = w * kPer + r * 64 + lane;
        const bool valid = e < count;
        const uint32_t d = (uint32_t)(key[r] >> shift) & (R - 1u);
        uint64_t peers = ballot(valid);
        for (int bit = 0; bit < BITS; bit++) {
            const uint64_t bb = ballot(valid && ((d >> bit) & 1u));
            peers &= ((d >> bit) & 1u) ? bb : ~bb;
        }
        const uint32_t before = valid ? L.cnt[w][d] : 0u;   // same digit: same value in every peer
        const uint32_t in_row = __popcll(peers & lt);
        rank[r] = before + in_row;
        if (valid && in_row == 0) L.cnt[w][d] = before + __popcll(peers);
    }
    __syncthreads();
    uint32_t tot = 0;   // thread t < 256: the tile's count of digit t
    {   // thread t < 256 = digit t: per-wave offsets and the tile's digit starts
        uint32_t c[W], incl = 0;
        if (t < 256) {
            for (int k = 0; k < W; k++) {
                c[k] = L.cnt[k][t];
                tot += c[k];
            }
            lb.publish(t, tot);
            incl = tot;
            for (int off = 1; off < 64; off <<= 1) {
                const uint32_t v = __shfl_up(incl, off, 64);
                if (lane >= (uint32_t)off) incl += v;
            }
            if (lane == 63) L.w_tot[w] = incl;
        }
        __syncthreads();
        if (t < 256) {
            uint32_t before = 0;
            for (uint32_t k = 0; k < w; k++) before += L.w_tot[k];
            L.start[t] = before + incl - tot;
            uint32_t run = 0;
            for (int k = 0; k < W; k++) {
                L.cnt[k][t] = run;
                run += c[k];
            }
        }
    }
    __syncthreads();
    for (int r = 0; r < kRows; r++) {
        const uint32_t e = w * kPer + r * 64 + lane;
        if (e < count) {
            const uint32_t d = (uint32_t)(key[r] >> shift) & (R - 1u);
            const uint32_t pos = L.start[d] + L.cnt[w][d] + rank[r];
            L.key[pos] = key[r];
            L.idx[pos] = idx[r];
        }
    }
    // the earlier tiles' counts arrive while the tile sits reordered in LDS
    if (Lookback::kActive && t < 256) L.base[t] += lb.resolve(t, tot);
    __syncthreads();
    for (uint32_t j = t; j < count; j += W * 64) {
        const uint64_t k = L.key[j];
        const uint32_t d = (uint32_t)(k >> shift) & (R - 1u);
        const uint32_t dest = L.base[d] + (j - L.start[d]);
        keys_out[dest] = k;
        idx_out[dest] = L.idx[j];
    }
}

// ---- one simulation: all digit counts in one read, one launch per pass.
// Counts of all eight digits of the keys at [0, numBodies): each block counts
// a contiguous chunk in LDS (a row of 64 keys whose digit is the same in
// every lane, the common case for the high digits of nearly ordered keys,
// adds with one atomic; other rows with one per key) and writes its 8 x 256
// counts; block 0 also starts a new sort: the next epoch, the passes' tile
// ids back to zero.
__global__ __launch_bounds__(512) void k_sort_digits(const SimValues* __restrict__ sim,
                                                     const uint64_t* __restrict__ keys,
                                                     uint32_t* __restrict__ partials, SortState* __restrict__ st) {
    __shared__ uint32_t h[kSortPasses * 256];
    const uint32_t t = threadIdx.x, lane = lane_id(), w = t >> 6;
    for (uint32_t i = t; i < kSortPasses * 256; i += 512) h[i] = 0;
    __syncthreads();
    const uint32_t num = sim->numBodies;
    const uint32_t chunk = ((num + gridDim.x - 1) / gridDim.x + 63u) & ~63u;
    const uint32_t begin = min(blockIdx.x * chunk, num), end = min(begin + chunk, num);
    constexpr uint32_t kRows = 4;                    // rows of 64 keys a wave has in flight
    for (uint32_t r0 = begin + w * 64 * kRows; r0 < end; r0 += 512 * kRows) {
        uint64_t key[kRows];
        for (uint32_t r = 0; r < kRows; r++) {
            const uint32_t e = r0 + r * 64 + lane;
            key[r] = e < end ? keys[e] : 0ull;
        }
        for (uint32_t r = 0; r < kRows; r++) {
            if (r0 + r * 64 >= end) break;           // (wave-uniform)
            const bool valid = r0 + r * 64 + lane < end;   // a prefix of the row's lanes
            const uint32_t nvalid = (uint32_t)__popcll(ballot(valid));
            for (uint32_t p = 0; p < kSortPasses; p++) {
                const uint32_t d = (uint32_t)(key[r] >> (p * kSortBits)) & 255u;
                const uint32_t d0 = uniform(d);      // lane 0 is valid
                if (ballot(valid && d != d0) == 0ull) {
                    if (lane == 0) atomicAdd(&h[p * 256 + d0], nvalid);
                } else if (valid) {
                    atomicAdd(&h[p * 256 + d], 1u);
                }
            }
        }
    }
    __syncthreads();
    for (uint32_t i = t; i < kSortPasses * 256; i += 512) partials[(size_t)blockIdx.x * kSortPasses * 256 + i] = h[i];
    if (blockIdx.x == 0 && t == 0) {
        const uint32_t e = st->epoch + 1u;
        st->epoch = e ? e : 1u;                      // 0 marks a status word never written
        for (uint32_t p = 0; p < kSortPasses; p++) st->tile[p] = 0;
    }
}

// Each digit's total over the partial counts: block (pass, group of 32
// digits), 32 threads per digit over every 32nd count set, then summed in LDS.
constexpr uint32_t kTotalsDigits = 32;
__global__ __launch_bounds__(1024) void k_sort_totals(const uint32_t* __restrict__ partials, uint32_t nparts,
                                                      uint32_t* __restrict__ totals) {
    __shared__ uint32_t s_part[32][kTotalsDigits + 1];
    const uint32_t p = blockIdx.x / (256 / kTotalsDigits);
    const uint32_t d0 = (blockIdx.x % (256 / kTotalsDigits)) * kTotalsDigits;
    const uint32_t t = threadIdx.x, dl = t % kTotalsDigits, q = t / kTotalsDigits;
    constexpr uint32_t kU = 4;
    uint32_t part = 0;
    for (uint32_t b0 = q; b0 < nparts; b0 += 32 * kU) {
        uint32_t v[kU];
        for (uint32_t j = 0; j < kU; j++)
            v[j] = b0 + 32 * j < nparts ? partials[(size_t)(b0 + 32 * j) * kSortPasses * 256 + p * 256 + d0 + dl] : 0u;
        for (uint32_t j = 0; j < kU; j++) part += v[j];
    }
    s_part[q][dl] = part;
    __syncthreads();
    if (t < kTotalsDigits) {
        uint32_t tot = 0;
        for (uint32_t k = 0; k < 32; k++) tot += s_part[k][t];
        totals[p * 256 + d0 + t] = tot;
    }
}

// Decoupled look-back over the tiles of one pass. Status word of (tile,
// digit): epoch << 32 | inclusive << 31 | count, where count is the tile's
// own count of the digit (an aggregate) or, with the inclusive bit, the
// count over this and every earlier tile. A word of another epoch is a
// previous sort's: not yet written. Tile ids follow the order in which the
// blocks started, so every tile a block waits for is running or done; the
// wait is bounded all the same (kLookbackSpins polls, then error bit 3).
constexpr uint32_t kLookbackInclusive = 0x80000000u;
constexpr uint32_t kLookbackSpins = 1u << 20;
// earlier tiles' words read at once: 4 gave 1.33 ms for a 16M sort of 4,096-key
// tiles, 1 gave 1.42 ms and 8 1.36 ms
constexpr int kLookbackWindow = 4;
struct TileLookback {
    static constexpr bool kActive = true;
    unsigned long long* status;   // [tiles][256] of this pass
    uint32_t tile, epoch;
    DevCounters* ctr;
    __device__ void publish(uint32_t d, uint32_t cnt) const {
        const unsigned long long v = ((unsigned long long)epoch << 32) | (tile ? 0u : kLookbackInclusive) | cnt;
        __hip_atomic_store(&status[(size_t)tile * 256 + d], v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    // the words of kLookbackWindow earlier tiles are read at once; counts are
    // taken in order up to the first inclusive word or the first word not yet
    // written, which is read again (after a short sleep) with the next window
    __device__ uint32_t resolve(uint32_t d, uint32_t cnt) const {
        if (tile == 0) return 0;
        uint32_t excl = 0, spins = 0;
        int j = (int)tile - 1;
        for (;;) {
            unsigned long long s[kLookbackWindow];
            for (int q = 0; q < kLookbackWindow; q++)
                s[q] = j - q >= 0 ? __hip_atomic_load(&status[(size_t)(j - q) * 256 + d], __ATOMIC_RELAXED,
                                                      __HIP_MEMORY_SCOPE_AGENT)
                                  : 0ull;
            int q = 0;
            bool done = false;
            for (; q < kLookbackWindow; q++) {
                if ((uint32_t)(s[q] >> 32) != epoch) break;
                excl += (uint32_t)s[q] & ~kLookbackInclusive;
                if (((uint32_t)s[q] & kLookbackInclusive) || j - q == 0) {
                    done = true;
                    break;
                }
            }
            if (done) break;
            j -= q;
            if (q < kLookbackWindow) {
                if (++spins > kLookbackSpins) {
                    atomicOr(&ctr->errorFlags, 8u);
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
        }
        const unsigned long long v = ((unsigned long long)epoch << 32) | kLookbackInclusive | (excl + cnt);
        __hip_atomic_store(&status[(size_t)tile * 256 + d], v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return excl;
    }
};

template <int kTile, int kWaves>
__global__ __launch_bounds__(kWaves * 64) void k_sort_onesweep(const SimValues* __restrict__ sim,
                                                                      DevCounters* __restrict__ ctr,
                                                                      const uint64_t* __restrict__ keys_in,
                                                                      const uint32_t* __restrict__ idx_in,
                                                                      uint64_t* __restrict__ keys_out,
                                                                      uint32_t* __restrict__ idx_out, uint32_t pass,
                                                                      const uint32_t* __restrict__ totals,
                                                                      unsigned long long* __restrict__ status,
                                                                      SortState* __restrict__ st) {
    __shared__ ScatterLds<kTile, kWaves> L;
    __shared__ uint32_t s_tile;
    const uint32_t t = threadIdx.x;
    if (t == 0) s_tile = atomicAdd(&st->tile[pass], 1u);
    __syncthreads();
    const uint32_t tile = s_tile;
    const uint32_t num = sim->numBodies;
    const uint32_t base = tile * (uint32_t)kTile;
    if (base >= num) return;                          // ids are dense: the tiles below num are all taken
    {   // where each digit's run starts: the exclusive scan of the pass's digit totals
        __shared__ uint32_t s_w[4];
        const uint32_t lane = lane_id();
        const uint32_t tot = t < 256 ? totals[pass * 256 + t] : 0u;
        uint32_t incl = tot;
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t v = __shfl_up(incl, off, 64);
            if (lane >= (uint32_t)off) incl += v;
        }
        if (t < 256 && lane == 63) s_w[t >> 6] = incl;
        __syncthreads();
        if (t < 256) {
            uint32_t before = 0;
            for (uint32_t k = 0; k < (t >> 6); k++) before += s_w[k];
            L.base[t] = before + incl - tot;
        }
    }
    const TileLookback lb = {status, tile, st->epoch, ctr};
    scatter_tile<8, kTile, kWaves>(L, keys_in, idx_in, base, min((uint32_t)kTile, num - base), keys_out, idx_out,
                                   pass * kSortBits, lb);
}


// ------------------------------------------------------------------ tree (build + fill)
// star colour, common/common.glsl:235-276
__device__ float4 star_color(float mass, float density) {
    float temp = fdiv(fmul(fmul(5.95589e-19f, powf(mass, 0.875f - 1.0f / 6.0f)), powf(density, 1.0f / 6.0f)), 100.0f);
    float r, g, b;
    if (temp <= 66.0f) r = 255.0f;
    else r = fmul(329.698727446f, powf(fsub(temp, 60.0f), -0.1332047592f));
    if (temp <= 66.0f) g = fsub(fmul(99.4708025861f, logf(gmax(temp, 1e-3f))), 161.1195681661f);
    else g = fmul(288.1221695283f, powf(temp, -0.0755148492f));
    if (temp >= 66.0f) b = 255.0f;
    else b = fsub(fmul(138.5177312231f, logf(gmax(fsub(temp, 10.0f), 1e-3f))), 305.0447927307f);
    return make_float4(fdiv(gmin(gmax(r, 0.f), 255.f), 255.f), fdiv(gmin(gmax(g, 0.f), 255.f), 255.f),
                       fdiv(gmin(gmax(b, 0.f), 255.f), 255.f), 0.0f);
}

__device__ __forceinline__ int delta_at(const uint64_t* __restrict__ keys, int num, int k) {
    return safe_lcp(keys, num, k, k + 1);
}

// Cross-XCD exchange of a child's summary between the two threads that meet
// at a parent. MI355X L2 caches are per XCD and not coherent with each other;
// stores and loads with the sc1 coherence bit (what agent-scope atomic
// accesses compile to) are, so the record goes through them and no agent-scope
// fence (an L2 write-back plus invalidate on this chip) is needed. The record
// stores are waited for before the arrival exchange that publishes them.
// 16-byte store with the coherence bit the compiler uses for agent-scope
// atomic stores (sc1): the record is visible to every XCD once the store has
// completed. The hardware reads the data VGPRs of a store wider than 8 bytes
// after it issues, and a VALU that rewrites them in the next cycles corrupts
// the stored data; the compiler pads its own stores for this but cannot see
// into inline assembly, so the padding is part of the asm block.
__device__ __forceinline__ void store_coherent(float4* p, float4 v) {
    typedef float f32x4_t __attribute__((ext_vector_type(4)));
    f32x4_t w = {v.x, v.y, v.z, v.w};
    asm volatile("global_store_dwordx4 %0, %1, off sc1\n\ts_nop 2" ::"v"(p), "v"(w) : "memory");
}
__device__ __forceinline__ float4 load_coherent(const float4* p) {
    const unsigned long long* q = (const unsigned long long*)p;
    unsigned long long a = __hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    unsigned long long b = __hip_atomic_load(q + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return make_float4(__uint_as_float((uint32_t)a), __uint_as_float((uint32_t)(a >> 32)),
                       __uint_as_float((uint32_t)b), __uint_as_float((uint32_t)(b >> 32)));
}

// Wait until this wave's stores are acknowledged, so that the exchange that
// publishes them cannot overtake them (a workgroup-scope release fence emits
// no wait on this target; an agent-scope one would also write back the L2).
__device__ __forceinline__ void stores_done() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }

struct NodeState {
    float4 com;          // centre of mass, mass
    float4 mn, mx;       // AABB
    float4 col;          // mass-weighted colour
    uint32_t height, id; // height, global node id
    // (trees with second moments) about com: sum m (x - c)(x - c)^T as xx, xy,
    // xz, yy, yz, zz, and the first moment sum m (x - c), which the rounding of
    // the FP32 centre of mass leaves slightly off zero
    float m2[6];
    float dm[3];
};

// The moments of the parent p of A and B about p's centre of mass, from the
// children's about theirs: with d = c_child - c_p, a child's bodies add
// sum m (x - c_p)(x - c_p)^T = M2 + D d^T + d D^T + m d d^T and
// sum m (x - c_p) = D + m d. Carrying D keeps this exact up to FP32 rounding
// of terms of the node's own size, however far the FP32 centre lies from the
// true one (the parallel-axis form without D is off by D d^T, of relative size
// ulp(|x|) / h for a node of half side h).
__device__ __forceinline__ void merge_moments(const NodeState& A, const NodeState& B, NodeState& p) {
    float m2[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f}, dm[3] = {0.f, 0.f, 0.f};
    const NodeState* ch[2] = {&A, &B};
    for (int c = 0; c < 2; c++) {
        const NodeState& C = *ch[c];
        const float d[3] = {C.com.x - p.com.x, C.com.y - p.com.y, C.com.z - p.com.z};
        const float hm = 0.5f * C.com.w;
        float u[3];   // D + m d / 2: then D d^T + d D^T + m d d^T = u d^T + d u^T
        for (int i = 0; i < 3; i++) {
            u[i] = __builtin_fmaf(hm, d[i], C.dm[i]);
            dm[i] += __builtin_fmaf(C.com.w, d[i], C.dm[i]);
        }
        int q = 0;
        for (int i = 0; i < 3; i++)
            for (int j = i; j < 3; j++, q++) m2[q] += __builtin_fmaf(u[i], d[j], __builtin_fmaf(d[i], u[j], C.m2[q]));
    }
    for (int q = 0; q < 6; q++) p.m2[q] = m2[q];
    for (int i = 0; i < 3; i++) p.dm[i] = dm[i];
}

// the three 16-byte words of a node's moments in the exchange and frontier records
__device__ __forceinline__ void moment_words(const NodeState& s, float4* w) {
    w[0] = make_float4(s.m2[0], s.m2[1], s.m2[2], s.m2[3]);
    w[1] = make_float4(s.m2[4], s.m2[5], s.dm[0], s.dm[1]);
    w[2] = make_float4(s.dm[2], 0.f, 0.f, 0.f);
}
__device__ __forceinline__ void moments_from_words(const float4* w, NodeState& s) {
    s.m2[0] = w[0].x; s.m2[1] = w[0].y; s.m2[2] = w[0].z; s.m2[3] = w[0].w;
    s.m2[4] = w[1].x; s.m2[5] = w[1].y; s.dm[0] = w[1].z; s.dm[1] = w[1].w;
    s.dm[2] = w[2].x;
}

// record of one side of a split: 3 (4 with colours) 16-byte words
__device__ __forceinline__ void put_state(float4* x, const NodeState& s, int colors) {
    store_coherent(x + 0, s.com);
    store_coherent(x + 1, make_float4(s.mn.x, s.mn.y, s.mn.z, s.mx.x));
    store_coherent(x + 2, make_float4(s.mx.y, s.mx.z, __uint_as_float(s.height), __uint_as_float(s.id)));
    if (colors) store_coherent(x + 3, s.col);
}

__device__ __forceinline__ void put_moments(float4* xm, const NodeState& s) {
    float4 w[3];
    moment_words(s, w);
    for (int k = 0; k < 3; k++) store_coherent(xm + k, w[k]);
}
__device__ __forceinline__ void get_moments(const float4* xm, NodeState& s) {
    float4 w[3];
    for (int k = 0; k < 3; k++) w[k] = load_coherent(xm + k);
    moments_from_words(w, s);
}

__device__ __forceinline__ void get_state(const float4* x, NodeState& s, int colors) {
    s.com = load_coherent(x + 0);
    const float4 a = load_coherent(x + 1), b = load_coherent(x + 2);
    s.mn = make_float4(a.x, a.y, a.z, 0.f);
    s.mx = make_float4(a.w, b.x, b.y, 0.f);
    s.height = __float_as_uint(b.z);
    s.id = __float_as_uint(b.w);
    s.col = colors ? load_coherent(x + 3) : make_float4(0.f, 0.f, 0.f, 0.f);
}

// Bottom-up build of the Karras-2012 radix tree (Apetrei 2014 agglomeration:
// one thread per leaf climbs while it is the second child to arrive), fused
// with the centre-of-mass/AABB/height fill of bh_tree.comp:183-259. The node
// numbering is Karras's: an internal node covering [lo, hi] is numbered hi
// when it is its parent's left child and lo when it is the right child or the
// root, which is exactly the node index treeBuildBinaryRadixTreeKernel assigns.
//
// tree_climb builds the tree of one segment of the sorted order: positions
// [off, off + num) (local positions 0..num-1, keys given from the segment's
// first key), internal nodes stored at ibase + local index, leaf ids
// off + local, internal ids n_ids + ibase + local. A single simulation is the
// segment off = 0, ibase = 0, n_ids = n; an ensemble has one per simulation.
struct TreeSeg {
    uint32_t off, num, ibase, n_ids;
    float G;                // the walk records carry G m
    uint32_t fill_cap = 0;  // nonzero: nodes higher than this are left unfilled
};

// The reference fills the tree in PROPAGATE_NODES_ITERATIONS = 64 passes of
// one height each (gpu/GPU.java:41, tree/bh_tree.comp:183-259): a node
// higher than that keeps what the build gave it (tree/bh_tree.comp:114-117):
// no mass, the empty box (1e38, -1e38), no colour. Its walk record has the
// half side of that box, -1e38, which every body accepts at once (no force),
// and a half2 of -1 (never opened by the wide-margin test).
__device__ __forceinline__ void unfilled_node(uint32_t k, uint32_t height, uint32_t idA, uint32_t idB,
                                              const TreeArrays& T, int colors) {
    WalkNode nd;
    nd.comMass = make_float4(0.f, 0.f, 0.f, 0.f);
    nd.half = fdiv(fsub(-1e38f, 1e38f), 2.0f);
    nd.childA = idA;
    nd.childB = idB;
    nd.half2 = -1.0f;
    T.nodes[k] = nd;
    T.node_aabb[2 * k] = make_float4(1e38f, 1e38f, 1e38f, __uint_as_float(height));
    T.node_aabb[2 * k + 1] = make_float4(-1e38f, -1e38f, -1e38f, 0.0f);
    if (colors) T.node_color[k] = make_float4(0.f, 0.f, 0.f, 0.f);
}

// The summary of leaf (sorted position) j as its own thread starts it.
__device__ __forceinline__ NodeState leaf_state(float4 pm, uint32_t id) {
    NodeState me;
    me.com = pm;
    me.mn = me.mx = make_float4(pm.x, pm.y, pm.z, 0.f);
    me.height = 0;
    me.id = id;
    me.col = make_float4(0.f, 0.f, 0.f, 0.f);
    for (int q = 0; q < 6; q++) me.m2[q] = 0.f;
    for (int i = 0; i < 3; i++) me.dm[i] = 0.f;
    return me;
}

// The walk's record of a leaf: its position and G m, with a half side of 0
// and a half2 of -1 (so that the test for nodes every lane opens never passes).
__device__ __forceinline__ WalkNode leaf_record(float4 pm, float G) {
    WalkNode w;
    w.comMass = make_float4(pm.x, pm.y, pm.z, fmul(G, pm.w));
    w.half = 0.0f;
    w.childA = 0;
    w.childB = 0;
    w.half2 = -1.0f;
    return w;
}

struct LeafSource {
    const float4* src;          // bodies (written by earlier kernels: coherent)
    const uint32_t* index_in;   // sorted position -> body id
    float unit_mass, unit_density;
    // With from set, the sort has sorted positions, not ids: from[j] is the
    // (scrambled) position before the sort of the body now at j (close to j
    // for most bodies), index_in is still in that order, and the first launch writes
    // the ids in the new order to ids_out. With step_pos set too, the body is
    // step_pos[from[j]] (the previous walk's output by its positions: read
    // nearly in order, where the body buffer by id is in no useful order).
    const float4* step_pos = nullptr;
    const uint32_t* from = nullptr;
    uint32_t* ids_out = nullptr;
};

// the body at sorted position j and its id (the first launch)
__device__ __forceinline__ float4 leaf_body(const LeafSource& L, uint32_t j, uint32_t& id) {
    if (!L.from) {
        id = L.index_in[j];
        return L.src[2 * id];
    }
    const uint32_t f = L.from[j] * kPosUnscramble;
    id = L.index_in[f];
    L.ids_out[j] = id;
    return L.step_pos ? L.step_pos[f] : L.src[2 * id];
}

// a leaf met during the climb: the first launch has stored every leaf's body in leaf_pos
__device__ __forceinline__ NodeState leaf_at(const LeafSource& L, const TreeArrays& T, uint32_t j, int colors) {
    const float4 pm = T.leaf_pos[j];
    NodeState me = leaf_state(pm, j);
    if (colors) {
        const float4 vd = L.src[2 * L.index_in[j] + 1];
        if (pm.w > 0.0f && vd.w > 0.0f) me.col = star_color(fmul(pm.w, L.unit_mass), fmul(vd.w, L.unit_density));
    }
    return me;
}

// Whether a child is a leaf is known from the keys alone: the sibling of a
// node ending (starting) at a split is the single leaf next to that split iff
// that leaf's own parent lies towards the split. So no exchange is needed at a
// parent with a leaf child: the internal child's thread (or, for two leaves,
// the left leaf's thread) builds the parent at once, rebuilding the leaf's
// summary from the body arrays, which earlier kernels wrote (so every XCD
// sees them), and a leaf whose sibling is internal simply stops. Only parents
// of two internal children use the arrival exchange and the coherent
// records; that is about a third of the nodes of a radix tree.
//
// The build runs in two launches. The first (tree_block) builds, per block of
// tree_block_size sorted positions, every node inside the block with the exchange
// in LDS, and lists the roots of the subtrees it leaves. The second runs
// tree_climb from each of them to the root, with the exchange in HBM.
struct FrontierEntry {
    float4 com, a, b;   // NodeState as put_state packs it
    uint4 range;        // lo, hi (local positions), segment (ensemble simulation), unused
};

// The summary of the parent of A and B over local positions [plo, phi]
// with Karras index pid, stored in the node arrays (bh_tree.comp:183-259);
// a_left: A is the left child (childA). Every summary value is symmetric in
// the two children (IEEE addition and multiplication commute, and the
// moments add the two children's terms to zero in either order), so only
// the child ids depend on which one is left: one inlined copy serves a wave
// whose lanes arrive from both sides.
__device__ __forceinline__ NodeState tree_parent(const NodeState& A, const NodeState& B, bool a_left, int plo,
                                                 int phi, uint32_t pid, const TreeSeg& S, const TreeArrays& T,
                                                 uint32_t* __restrict__ height_out, bool height_max, int colors,
                                                 bool moments) {
    const uint32_t idA = a_left ? A.id : B.id, idB = a_left ? B.id : A.id;
    NodeState p;
    const float M = fadd(A.com.w, B.com.w);
    if (M > 0.0f) {
        p.com.x = fdiv(fadd(fmul(A.com.w, A.com.x), fmul(B.com.w, B.com.x)), M);
        p.com.y = fdiv(fadd(fmul(A.com.w, A.com.y), fmul(B.com.w, B.com.y)), M);
        p.com.z = fdiv(fadd(fmul(A.com.w, A.com.z), fmul(B.com.w, B.com.z)), M);
        p.col.x = fdiv(fadd(fmul(A.com.w, A.col.x), fmul(B.com.w, B.col.x)), M);
        p.col.y = fdiv(fadd(fmul(A.com.w, A.col.y), fmul(B.com.w, B.col.y)), M);
        p.col.z = fdiv(fadd(fmul(A.com.w, A.col.z), fmul(B.com.w, B.col.z)), M);
    } else {
        p.com.x = fmul(fadd(A.com.x, B.com.x), 0.5f);
        p.com.y = fmul(fadd(A.com.y, B.com.y), 0.5f);
        p.com.z = fmul(fadd(A.com.z, B.com.z), 0.5f);
        p.col.x = fmul(fadd(A.col.x, B.col.x), 0.5f);
        p.col.y = fmul(fadd(A.col.y, B.col.y), 0.5f);
        p.col.z = fmul(fadd(A.col.z, B.col.z), 0.5f);
    }
    p.com.w = M;
    p.col.w = 0.0f;
    p.mn = make_float4(gmin(A.mn.x, B.mn.x), gmin(A.mn.y, B.mn.y), gmin(A.mn.z, B.mn.z), 0.f);
    p.mx = make_float4(gmax(A.mx.x, B.mx.x), gmax(A.mx.y, B.mx.y), gmax(A.mx.z, B.mx.z), 0.f);
    p.height = 1u + (A.height > B.height ? A.height : B.height);
    const uint32_t k = S.ibase + pid;                 // storage index of the parent
    p.id = S.n_ids + k;
    WalkNode nd;
    nd.comMass = make_float4(p.com.x, p.com.y, p.com.z, fmul(S.G, M));
    const float ext = gmax(fsub(p.mx.x, p.mn.x), gmax(fsub(p.mx.y, p.mn.y), fsub(p.mx.z, p.mn.z)));
    nd.half = fdiv(ext, 2.0f);
    nd.childA = idA;
    nd.childB = idB;
    nd.half2 = fmul(nd.half, nd.half);
    // these arrays are read only by later kernels
    if (S.fill_cap && p.height > S.fill_cap) {
        unfilled_node(k, p.height, idA, idB, T, colors);   // (its ancestors are higher still)
    } else {
        T.nodes[k] = nd;
        T.node_aabb[2 * k] = make_float4(p.mn.x, p.mn.y, p.mn.z, __uint_as_float(p.height));
        T.node_aabb[2 * k + 1] = make_float4(p.mx.x, p.mx.y, p.mx.z, M);
        if (colors) T.node_color[k] = p.col;
    }
    if (moments) {
        merge_moments(A, B, p);
        T.node_mom[2 * k] = make_float4(p.m2[0], p.m2[1], p.m2[2], p.m2[3]);
        T.node_mom[2 * k + 1] = make_float4(p.m2[4], p.m2[5], 0.f, 0.f);
    }
    T.node_range[k] = make_uint2(S.off + (uint32_t)plo, (uint32_t)(phi - plo + 1));
    if (A.id < S.n_ids) T.leaf_parent[A.id] = p.id; else T.node_parent[A.id - S.n_ids] = p.id;
    if (B.id < S.n_ids) T.leaf_parent[B.id] = p.id; else T.node_parent[B.id - S.n_ids] = p.id;
    if (pid == 0) {
        T.node_parent[k] = kEmpty;
        if (height_max) atomicMax(height_out, p.height);
        else *height_out = p.height;
    }
    return p;
}

__device__ __forceinline__ void tree_climb(int lo, int hi, const TreeSeg& S, NodeState me,
                                           const uint64_t* __restrict__ keys, const LeafSource& L,
                                           const TreeArrays& T, uint32_t* __restrict__ height_out, bool height_max,
                                           int colors, bool moments) {
    const int num = (int)S.num;
    for (;;) {
        if (lo == 0 && hi == num - 1) break;
        const bool left = (lo == 0) || (hi != num - 1 && delta_at(keys, num, hi) > delta_at(keys, num, lo - 1));
        const int split = left ? hi : lo - 1;
        // each side of a split has its own record; the sibling's is the other one
        // the sibling covers [split + 1, ...] (I am the left child) or
        // [..., split]; it is the leaf split + 1 (split) iff that leaf's parent
        // is towards the split
        const bool sib_leaf = left ? (split + 1 == num - 1 || delta_at(keys, num, split + 1) <= delta_at(keys, num, split))
                                   : (split == 0 || delta_at(keys, num, split) > delta_at(keys, num, split - 1));
        uint32_t other;
        NodeState sib;
        if (sib_leaf) {
            if (lo == hi && !left) return;                // two leaves: the left one builds the parent
            other = (uint32_t)(left ? split + 1 : split);
            sib = leaf_at(L, T, S.off + other, colors);
        } else {
            if (lo == hi) return;                         // the internal sibling's thread builds the parent
            float4* x = T.xchg + 8 * ((size_t)S.ibase + split);
            float4* xm = moments ? T.xchg_mom + 6 * ((size_t)S.ibase + split) : nullptr;
            put_state(x + (left ? 0 : 4), me, colors);
            if (moments) put_moments(xm + (left ? 0 : 3), me);
            stores_done();
            other = __hip_atomic_exchange(&T.meet[S.ibase + split], (uint32_t)(left ? lo : hi), __ATOMIC_RELAXED,
                                          __HIP_MEMORY_SCOPE_AGENT);
            if (other == kEmpty) return;                  // first to arrive: the sibling continues
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
            get_state(x + (left ? 4 : 0), sib, colors);
            if (moments) get_moments(xm + (left ? 3 : 0), sib);
            T.meet[S.ibase + split] = kEmpty;             // only this thread touches the slot now
        }
        const int plo = left ? lo : (int)other;
        const int phi = left ? (int)other : hi;
        uint32_t pid;
        if (plo == 0 && phi == num - 1) {
            pid = 0;
        } else {
            const bool pl = (plo == 0) || (phi != num - 1 && delta_at(keys, num, phi) > delta_at(keys, num, plo - 1));
            pid = (uint32_t)(pl ? phi : plo);
        }
        me = tree_parent(me, sib, left, plo, phi, pid, S, T, height_out, height_max, colors, moments);
        lo = plo;
        hi = phi;
    }
}

// ---- first launch: block-local build.
// A block of tree_block_size() threads takes as many consecutive sorted positions
// [blo, bhi] of one segment and builds every internal node whose range lies
// inside them, with the same agglomeration as tree_climb but with the
// arrival slots, the records and the split lengths in LDS. A climb stops at a
// node whose parent split is outside [blo, bhi - 1] (an edge node: at most one
// per side), and a first arrival whose sibling never comes (its sibling
// crosses the block's edge) is left in its slot. Both kinds go to the
// frontier list for the second launch. Each side of a block is crossed by a
// chain of nested nodes, one per level, so a block leaves at most
// 2 * 98 + 2 entries (a tree over 96-bit keys is at most 98 levels high).
//
// Records are indexed by the child's own Karras index (hi for a left child,
// lo for a right one), which is unique among internal nodes; the arrival slot
// of split j holds the first arrival's far end, with kSideRight set for a
// right child.
constexpr uint32_t kSideRight = 0x80000000u;

// A record of the first launch's exchange: com, box, height and id, colour
// (3 or 4 words); with moments the id is left out (the reader knows the
// node's Karras index) and the 9 moment floats fill the words to 5 (6 with
// colours), so that three blocks of 512 positions fit a CU's LDS.
template <bool kColors, bool kMoments, uint32_t kBlock = tree_block_size(kMoments)>
struct TreeBlockLds {
    static constexpr int kRec = kMoments ? (kColors ? 6 : 5) : (kColors ? 4 : 3);   // float4 words of one record
    static constexpr uint32_t kB = kBlock;         // positions per block
    int d[kB + 1];                                 // delta(blo - 1 + j)
    uint32_t meet[kB];                             // arrival slot of split blo + j
    float4 pm[kB];                                 // posMass of leaf blo + j
    float4 col[kColors ? kB : 1];                  // colour of leaf blo + j
    float4 rec[kRec * kB];                         // summary of internal node blo + j
    float4 edge[2][kRec];                          // edge nodes: left child at split bhi, right child at blo - 1
    int2 edge_range[2];
    uint32_t edge_id[2];
    uint32_t nf, base;
};

// The arrival slot of a split whose parent has a leaf child is not used for
// an exchange: the thread that builds such a parent marks it (sorted output),
// and a leaf that left its parent to its internal sibling finds out there
// whether the parent was built in the block.
constexpr uint32_t kLeafParentBuilt = 0xFFFFFFFEu;

template <bool kColors, bool kMoments>
__device__ __forceinline__ void lds_put(float4* x, const NodeState& s) {
    x[0] = s.com;
    x[1] = make_float4(s.mn.x, s.mn.y, s.mn.z, s.mx.x);
    if (kMoments) {
        x[2] = make_float4(s.mx.y, s.mx.z, __uint_as_float(s.height), s.m2[0]);
        x[3] = make_float4(s.m2[1], s.m2[2], s.m2[3], s.m2[4]);
        x[4] = make_float4(s.m2[5], s.dm[0], s.dm[1], s.dm[2]);
        if (kColors) x[5] = s.col;
    } else {
        x[2] = make_float4(s.mx.y, s.mx.z, __uint_as_float(s.height), __uint_as_float(s.id));
        if (kColors) x[3] = s.col;
    }
}

// id: the node's global id (a record with moments does not hold it)
template <bool kColors, bool kMoments>
__device__ __forceinline__ NodeState lds_get(const float4* x, uint32_t id) {
    NodeState s;
    s.com = x[0];
    const float4 a = x[1], b = x[2];
    s.mn = make_float4(a.x, a.y, a.z, 0.f);
    s.mx = make_float4(a.w, b.x, b.y, 0.f);
    s.height = __float_as_uint(b.z);
    if (kMoments) {
        const float4 c = x[3], d = x[4];
        s.id = id;
        s.m2[0] = b.w; s.m2[1] = c.x; s.m2[2] = c.y; s.m2[3] = c.z; s.m2[4] = c.w; s.m2[5] = d.x;
        s.dm[0] = d.y; s.dm[1] = d.z; s.dm[2] = d.w;
        s.col = kColors ? x[5] : make_float4(0.f, 0.f, 0.f, 0.f);
    } else {
        s.id = __float_as_uint(b.w);
        s.col = kColors ? x[3] : make_float4(0.f, 0.f, 0.f, 0.f);
    }
    return s;
}

__device__ __forceinline__ FrontierEntry frontier_entry(const NodeState& me, int lo, int hi, uint32_t segment) {
    FrontierEntry e;
    e.com = me.com;
    e.a = make_float4(me.mn.x, me.mn.y, me.mn.z, me.mx.x);
    e.b = make_float4(me.mx.y, me.mx.z, __uint_as_float(me.height), __uint_as_float(me.id));
    e.range = make_uint4((uint32_t)lo, (uint32_t)hi, segment, 0u);
    return e;
}

// Rank of key among the first nf of list (distinct keys): the sorted order
// of a block's few frontier entries.
__device__ __forceinline__ uint32_t rank_in(const int* list, uint32_t nf, int key) {
    uint32_t r = 0;
    for (uint32_t k = 0; k < nf; k++) r += list[k] < key;
    return r;
}

// Positions [blo, blo + cnt) of segment S (local positions; keys given from
// the segment's first key); thread t holds leaf blo + t. Called by every
// thread of the block (it has barriers). With T.frontier_cnt set, the
// entries go, sorted by their first position, to slots
// [out_block kTreeBlockFrontier, ...), and a leaf whose parent is not built
// here is an entry too (a mid level needs every position covered).
template <bool kColors, bool kMoments, uint32_t kBlock>
__device__ __forceinline__ void tree_block(TreeBlockLds<kColors, kMoments, kBlock>& s, const TreeSeg& S, int blo, int cnt,
                                           const uint64_t* __restrict__ keys, const LeafSource& L,
                                           const TreeArrays& T, uint32_t* __restrict__ height_out, bool height_max,
                                           uint32_t segment, uint32_t out_block = 0,
                                           uint32_t* __restrict__ error_flags = nullptr) {
    constexpr int R = TreeBlockLds<kColors, kMoments>::kRec;
    const int num = (int)S.num;
    const int t = threadIdx.x;
    const int bhi = blo + cnt - 1;
    if (t < cnt) {
        const uint32_t g = S.off + (uint32_t)(blo + t);
        uint32_t b;
        const float4 pm = leaf_body(L, g, b);
        T.leaf_pos[g] = pm;
        const WalkNode lr = leaf_record(pm, S.G);
        T.walk[g] = lr;
        if (kMoments) {   // the corrected walk's leaf record: position and G m, b^2 = K = -1
            T.corr_leaf[4 * (size_t)g] = lr.comMass;
            T.corr_leaf[4 * (size_t)g + 1] = make_float4(-1.0f, __uint_as_float(0u), __uint_as_float(0u), lr.half2);
        }
        s.pm[t] = pm;
        if (kColors) {
            const float4 vd = L.src[2 * b + 1];
            float4 c = make_float4(0.f, 0.f, 0.f, 0.f);
            if (pm.w > 0.0f && vd.w > 0.0f) c = star_color(fmul(pm.w, L.unit_mass), fmul(vd.w, L.unit_density));
            T.leaf_color[g] = c;
            s.col[t] = c;
        }
    }
    s.d[t + 1] = delta_at(keys, num, blo + t);
    if (t == 0) {
        s.d[0] = delta_at(keys, num, blo - 1);
        s.nf = 0;
        s.edge_range[0] = s.edge_range[1] = make_int2(-1, -1);
    }
    s.meet[t] = kEmpty;
    const bool sorted_out = T.frontier_cnt != nullptr;
    int leaf_split = -1;   // this thread's leaf left its parent (at this split) to its internal sibling
    __syncthreads();
    auto D = [&](int x) { return s.d[x - blo + 1]; };   // x in [blo - 1, bhi]
    if (t < cnt) {
        int lo = blo + t, hi = lo;
        NodeState me = leaf_state(s.pm[t], S.off + (uint32_t)lo);
        if (kColors) me.col = s.col[t];
        for (;;) {
            if (lo == 0 && hi == num - 1) break;
            const bool left = (lo == 0) || (hi != num - 1 && D(hi) > D(lo - 1));
            const int split = left ? hi : lo - 1;
            if (split < blo || split >= bhi) {             // the parent crosses the block's edge
                const int side = left ? 0 : 1;
                lds_put<kColors, kMoments>(s.edge[side], me);
                s.edge_range[side] = make_int2(lo, hi);
                s.edge_id[side] = me.id;
                break;
            }
            const bool sib_leaf = left ? (split + 1 == num - 1 || D(split + 1) <= D(split))
                                       : (split == 0 || D(split) > D(split - 1));
            int other;
            NodeState sib;
            if (sib_leaf) {
                if (lo == hi && !left) break;              // two leaves: the left one builds the parent
                other = left ? split + 1 : split;
                sib = leaf_state(s.pm[other - blo], S.off + (uint32_t)other);
                if (kColors) sib.col = s.col[other - blo];
                if (sorted_out) s.meet[split - blo] = kLeafParentBuilt;
            } else {
                if (lo == hi) {                            // the internal sibling's thread builds the parent
                    leaf_split = split;
                    break;
                }
                // the exchange orders LDS only: the fences are for the local
                // address space, so they do not wait for this wave's node stores
                lds_put<kColors, kMoments>(&s.rec[R * ((left ? hi : lo) - blo)], me);
                const uint32_t v = left ? (uint32_t)lo : ((uint32_t)hi | kSideRight);
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
                const uint32_t old =
                    __hip_atomic_exchange(&s.meet[split - blo], v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                if (old == kEmpty) break;                  // first to arrive
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
                s.meet[split - blo] = kEmpty;              // only this thread touches the slot now
                other = (int)(old & ~kSideRight);
                const int sk = left ? split + 1 : split;   // the sibling's Karras index
                sib = lds_get<kColors, kMoments>(&s.rec[R * (sk - blo)], S.n_ids + S.ibase + (uint32_t)sk);
            }
            const int plo = left ? lo : other;
            const int phi = left ? other : hi;
            uint32_t pid;
            if (plo == 0 && phi == num - 1) {
                pid = 0;
            } else {
                const bool pl = (plo == 0) || (phi != num - 1 && D(phi) > D(plo - 1));
                pid = (uint32_t)(pl ? phi : plo);
            }
            me = tree_parent(me, sib, left, plo, phi, pid, S, T, height_out, height_max, kColors, kMoments);
            lo = plo;
            hi = phi;
        }
    }
    __syncthreads();
    // first arrivals left waiting, the edge nodes and (sorted output) the
    // leaves no parent here took
    FrontierEntry e0, e1, e2;
    NodeState n0, n1;   // (their moments)
    int slot0 = -1, slot1 = -1, slot2 = -1;
    if (t < cnt - 1) {
        const uint32_t v = s.meet[t];
        if (v != kEmpty && v != kLeafParentBuilt) {
            const int split = blo + t, far = (int)(v & ~kSideRight);
            const bool right = (v & kSideRight) != 0;
            const int lo = right ? split + 1 : far, hi = right ? far : split;
            const int k0 = right ? lo : hi;   // its Karras index
            n0 = lds_get<kColors, kMoments>(&s.rec[R * (k0 - blo)], S.n_ids + S.ibase + (uint32_t)k0);
            e0 = frontier_entry(n0, lo, hi, segment);
            slot0 = (int)atomicAdd(&s.nf, 1u);
        }
    }
    if (t < 2 && s.edge_range[t].x >= 0) {
        n1 = lds_get<kColors, kMoments>(s.edge[t], s.edge_id[t]);
        e1 = frontier_entry(n1, s.edge_range[t].x, s.edge_range[t].y, segment);
        slot1 = (int)atomicAdd(&s.nf, 1u);
    }
    if (sorted_out && leaf_split >= 0 && s.meet[leaf_split - blo] != kLeafParentBuilt) {
        NodeState lf = leaf_state(s.pm[t], S.off + (uint32_t)(blo + t));
        if (kColors) lf.col = s.col[t];
        e2 = frontier_entry(lf, blo + t, blo + t, segment);
        slot2 = (int)atomicAdd(&s.nf, 1u);
    }
    __syncthreads();
    if (sorted_out) {
        // the entries in order of their first position (s.d is free now)
        const uint32_t nf = s.nf;
        if (slot0 >= 0) s.d[slot0] = (int)e0.range.x;
        if (slot1 >= 0) s.d[slot1] = (int)e1.range.x;
        if (slot2 >= 0) s.d[slot2] = (int)e2.range.x;
        __syncthreads();
        if (slot0 >= 0) slot0 = (int)rank_in(s.d, nf, (int)e0.range.x);
        if (slot1 >= 0) slot1 = (int)rank_in(s.d, nf, (int)e1.range.x);
        if (slot2 >= 0) slot2 = (int)rank_in(s.d, nf, (int)e2.range.x);
        if (nf > kTreeBlockFrontier) {   // (the bound above: not reached)
            if (t == 0 && error_flags) atomicOr(error_flags, 16u);
            slot0 = slot0 < (int)kTreeBlockFrontier ? slot0 : -1;
            slot1 = slot1 < (int)kTreeBlockFrontier ? slot1 : -1;
            slot2 = slot2 < (int)kTreeBlockFrontier ? slot2 : -1;
        }
        if (t == 0) T.frontier_cnt[out_block] = nf < kTreeBlockFrontier ? nf : kTreeBlockFrontier;
        if (t == 0) s.base = out_block * kTreeBlockFrontier;
        __syncthreads();
    } else {
        if (t == 0 && s.nf) s.base = atomicAdd(T.frontier_count, s.nf);
        __syncthreads();
    }
    FrontierEntry* F = reinterpret_cast<FrontierEntry*>(T.frontier);
    if (slot0 >= 0) F[s.base + slot0] = e0;
    if (slot1 >= 0) F[s.base + slot1] = e1;
    if (slot2 >= 0) F[s.base + slot2] = e2;
    if (kMoments) {
        if (slot0 >= 0) moment_words(n0, T.frontier_mom + 3 * (size_t)(s.base + slot0));
        if (slot1 >= 0) moment_words(n1, T.frontier_mom + 3 * (size_t)(s.base + slot1));
        if (slot2 >= 0) {   // a leaf's moments about itself are zero
            float4* w = T.frontier_mom + 3 * (size_t)(s.base + slot2);
            w[0] = w[1] = w[2] = make_float4(0.f, 0.f, 0.f, 0.f);
        }
    }
}

template <bool kColors, bool kMoments, uint32_t kB>
__global__ __launch_bounds__(kB) void k_tree(const SimValues* __restrict__ sim,
                                                     DevCounters* __restrict__ ctr, const float4* __restrict__ src,
                                                     const uint32_t* __restrict__ index_in,
                                                     const uint64_t* __restrict__ keys, TreeArrays T,
                                                     const float4* __restrict__ step_pos,
                                                     const uint32_t* __restrict__ from,
                                                     uint32_t* __restrict__ ids_out, uint32_t fill_cap) {
    __shared__ TreeBlockLds<kColors, kMoments, kB> s;
    const uint32_t n = sim->initialNumBodies;
    const int num = (int)sim->numBodies;
    const int blo = (int)(blockIdx.x * kB);
    if (from) {   // ids past numBodies (the dead) keep their positions
        const uint32_t lo = (uint32_t)blo > (uint32_t)num ? (uint32_t)blo : (uint32_t)num;
        const uint32_t hi = (uint32_t)blo + kB < n ? (uint32_t)blo + kB : n;
        for (uint32_t j = lo + threadIdx.x; j < hi; j += kB) ids_out[j] = index_in[j];
    }
    if (blo >= num || num == 1) {   // (no entries: the levels above read block 0's count)
        if (T.frontier_cnt && blockIdx.x == 0 && threadIdx.x == 0) T.frontier_cnt[0] = 0u;
        if (blo >= num) return;
    }
    if (num == 1) {
        if (threadIdx.x != 0) return;
        // bh_tree.comp:46-62: a one-body tree is an empty root over leaf 0
        if (from) ids_out[0] = index_in[0];
        const float4 pm = src[2 * index_in[0]];
        T.leaf_pos[0] = pm;
        T.walk[0] = leaf_record(pm, sim->gravitationalConstant);
        if (kColors) {
            const float4 vd = src[2 * index_in[0] + 1];
            float4 c = make_float4(0.f, 0.f, 0.f, 0.f);
            if (pm.w > 0.0f && vd.w > 0.0f) c = star_color(fmul(pm.w, sim->unitMass), fmul(vd.w, sim->unitDensity));
            T.leaf_color[0] = c;
        }
        WalkNode r;
        r.comMass = make_float4(0.f, 0.f, 0.f, 0.f);
        r.half = -1e38f;
        r.childA = 0;
        r.childB = kEmpty;
        r.half2 = 0.0f;
        T.nodes[0] = r;
        T.node_aabb[0] = make_float4(1e38f, 1e38f, 1e38f, 0.f);
        T.node_aabb[1] = make_float4(-1e38f, -1e38f, -1e38f, 0.f);
        T.node_parent[0] = kEmpty;
        T.node_range[0] = make_uint2(0, 0);
        T.leaf_parent[0] = n;
        ctr->treeHeight = 0;
        return;
    }
    const TreeSeg S = {0u, (uint32_t)num, 0u, n, sim->gravitationalConstant, fill_cap};
    // the previous walk's positions hold this step's bodies unless the
    // partition before the sort removed some (then their order is another)
    const bool by_pos = from && step_pos && ctr->died == 0;
    const LeafSource L = {src, index_in, sim->unitMass, sim->unitDensity, by_pos ? step_pos : nullptr, from,
                          ids_out};
    const int cnt = num - blo < (int)kB ? num - blo : (int)kB;
    tree_block<kColors, kMoments>(s, S, blo, cnt, keys, L, T, &ctr->treeHeight, false, 0u, blockIdx.x,
                                  &ctr->errorFlags);
}

// (the colour of a frontier node, when colours are on, is in node_color)
__device__ __forceinline__ NodeState frontier_state(const FrontierEntry& e) {
    NodeState me;
    me.com = e.com;
    me.mn = make_float4(e.a.x, e.a.y, e.a.z, 0.f);
    me.mx = make_float4(e.a.w, e.b.x, e.b.y, 0.f);
    me.height = __float_as_uint(e.b.z);
    me.id = __float_as_uint(e.b.w);
    me.col = make_float4(0.f, 0.f, 0.f, 0.f);
    for (int q = 0; q < 6; q++) me.m2[q] = 0.f;
    for (int i = 0; i < 3; i++) me.dm[i] = 0.f;
    return me;
}

// entry t's moments, written by the first launch
__device__ __forceinline__ void frontier_moments(const TreeArrays& T, uint32_t t, NodeState& me) {
    float4 w[3];
    for (int k = 0; k < 3; k++) w[k] = T.frontier_mom[3 * (size_t)t + k];
    moments_from_words(w, me);
}


// ---- the levels between: the entries of kTreeMidGroup blocks as leaves.
// Entry u of a window covers positions [lo_u, hi_u], hi_u + 1 = lo_{u+1}, so
// the split between entries u and u + 1 is the split at position hi_u and a
// node over entries [cl, ch] is the node over positions [lo_cl, hi_ch]: the
// same agglomeration as tree_block with d[u + 1] = delta(hi_u) and
// d[0] = delta(lo_0 - 1), every entry meeting its sibling through the arrival
// slots (an entry may be a leaf or a subtree; either one waits for its sibling
// here). A node's id follows from its position range: the Karras index of a
// left child is its last position, of a right child its first, and a node
// over one position is that leaf.
struct TreeMidArgs {
    const float4* in;         // the level below: 4 words per entry, in_cap slots per block
    const float4* in_mom;     // 3 moment words per entry (moments only)
    const uint32_t* in_cnt;   // entries per block of the level below
    float4* out;
    float4* out_mom;
    uint32_t* out_cnt;
    uint32_t in_cap, out_cap;
    uint64_t in_span;         // sorted positions per block of the level below
    // The last level (one block): what a window leaves climbs to the root at
    // once, through the arrival slots in HBM (tree_climb), meeting leaves by
    // their ids in index_in; nothing is written out.
    int last;
    const float4* src;
    const uint32_t* index_in;
};

template <bool kColors, bool kMoments>
struct TreeMidLds {
    static constexpr int kRec = TreeBlockLds<kColors, kMoments>::kRec;
    int d[kTreeMidUnits + 1];               // d[u + 1] = delta(hi_u); after the climb the entries' first positions
    uint32_t meet[kTreeMidUnits];           // arrival slot of the split after entry u
    int2 ur[kTreeMidUnits];                 // positions [lo_u, hi_u] of entry u
    float4 rec[kRec * kTreeMidUnits];       // entry u's summary, then the node with compressed Karras index u
    float4 edge[2][kRec];
    int2 edge_range[2];
    uint32_t edge_id[2];
    uint32_t pre[kTreeMidGroup + 1];        // exclusive prefix of the group's entry counts
    uint32_t nf;
};

template <bool kColors, bool kMoments>
__global__ __launch_bounds__(kTreeMidUnits) void k_tree_mid(const SimValues* __restrict__ sim,
                                                            DevCounters* __restrict__ ctr,
                                                            const uint64_t* __restrict__ keys, TreeArrays T,
                                                            TreeMidArgs A, uint32_t fill_cap) {
    __shared__ TreeMidLds<kColors, kMoments> s;
    constexpr int R = TreeMidLds<kColors, kMoments>::kRec;
    const uint32_t n = sim->initialNumBodies;
    const int num = (int)sim->numBodies;
    const int t = threadIdx.x;
    const uint32_t nb_in = num > 1 ? (uint32_t)(((uint64_t)num + A.in_span - 1) / A.in_span) : 0u;
    const uint32_t g0 = blockIdx.x * kTreeMidGroup;
    if (g0 >= nb_in) {
        if (blockIdx.x == 0 && t == 0) A.out_cnt[0] = 0u;
        return;
    }
    const uint32_t ng = nb_in - g0 < kTreeMidGroup ? nb_in - g0 : kTreeMidGroup;
    if (t < 64) {   // one wave scans the group's counts
        const uint32_t c = (uint32_t)t < ng ? A.in_cnt[g0 + t] : 0u;
        uint32_t x = c;
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t y = __shfl_up(x, o, 64);
            if (t >= o) x += y;
        }
        s.pre[t + 1] = x;
        if (t == 0) s.pre[0] = 0u;
    }
    __syncthreads();
    const uint32_t U = s.pre[ng];
    const TreeSeg S = {0u, (uint32_t)num, 0u, n, sim->gravitationalConstant, fill_cap};
    uint32_t out_n = 0;   // entries this block has written (the same in every thread)
    const FrontierEntry* Fin = reinterpret_cast<const FrontierEntry*>(A.in);
    FrontierEntry* Fout = reinterpret_cast<FrontierEntry*>(A.out) + (size_t)blockIdx.x * A.out_cap;
    for (uint32_t w0 = 0; w0 < U; w0 += kTreeMidUnits) {
        const int W = (int)(U - w0 < kTreeMidUnits ? U - w0 : kTreeMidUnits);
        NodeState me;
        if (t < W) {
            const uint32_t u = w0 + (uint32_t)t;
            int b = 0, bh = (int)ng - 1;   // the block holding entry u: the last b with pre[b] <= u
            while (b < bh) {
                const int mid = (b + bh + 1) >> 1;
                if (s.pre[mid] <= u) b = mid; else bh = mid - 1;
            }
            const size_t slot = (size_t)(g0 + (uint32_t)b) * A.in_cap + (u - s.pre[b]);
            const FrontierEntry e = Fin[slot];
            me = frontier_state(e);
            if (kColors) me.col = me.id < n ? T.leaf_color[me.id] : T.node_color[me.id - n];
            if (kMoments) {
                float4 w[3];
                for (int k = 0; k < 3; k++) w[k] = A.in_mom[3 * slot + k];
                moments_from_words(w, me);
            }
            s.ur[t] = make_int2((int)e.range.x, (int)e.range.y);
            s.d[t + 1] = delta_at(keys, num, (int)e.range.y);
            if (t == 0) s.d[0] = delta_at(keys, num, (int)e.range.x - 1);
        }
        s.meet[t] = kEmpty;
        if (t < 2) s.edge_range[t] = make_int2(-1, -1);
        if (t == 0) s.nf = 0u;
        __syncthreads();
        if (t < W) {
            int cl = t, ch = t;
            int lo = s.ur[t].x, hi = s.ur[t].y;
            for (;;) {
                if (lo == 0 && hi == num - 1) break;
                const bool left = (lo == 0) || (hi != num - 1 && s.d[ch + 1] > s.d[cl]);
                const int split = left ? ch : cl - 1;
                if (split < 0 || split >= W - 1) {   // the parent takes entries of another window
                    const int side = left ? 0 : 1;
                    lds_put<kColors, kMoments>(s.edge[side], me);
                    s.edge_range[side] = make_int2(lo, hi);
                    s.edge_id[side] = me.id;
                    break;
                }
                lds_put<kColors, kMoments>(&s.rec[R * (left ? ch : cl)], me);
                const uint32_t v = left ? (uint32_t)cl : ((uint32_t)ch | kSideRight);
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
                const uint32_t old =
                    __hip_atomic_exchange(&s.meet[split], v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                if (old == kEmpty) break;            // first to arrive
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
                s.meet[split] = kEmpty;              // only this thread touches the slot now
                const int other = (int)(old & ~kSideRight);
                const int slo = left ? hi + 1 : s.ur[other].x;   // the sibling's positions
                const int shi = left ? s.ur[other].y : lo - 1;
                const uint32_t sid = slo == shi ? S.off + (uint32_t)slo
                                                : S.n_ids + S.ibase + (uint32_t)(left ? slo : shi);
                const NodeState sib = lds_get<kColors, kMoments>(&s.rec[R * (left ? split + 1 : split)], sid);
                const int pcl = left ? cl : other, pch = left ? other : ch;
                const int plo = left ? lo : slo, phi = left ? shi : hi;
                uint32_t pid;
                if (plo == 0 && phi == num - 1) {
                    pid = 0;
                } else {
                    const bool pl = (plo == 0) || (phi != num - 1 && s.d[pch + 1] > s.d[pcl]);
                    pid = (uint32_t)(pl ? phi : plo);
                }
                me = tree_parent(me, sib, left, plo, phi, pid, S, T, &ctr->treeHeight, false, kColors, kMoments);
                cl = pcl;
                ch = pch;
                lo = plo;
                hi = phi;
            }
        }
        __syncthreads();
        // what this window leaves: first arrivals still waiting, and the edges
        FrontierEntry e0, e1;
        NodeState n0, n1;
        int slot0 = -1, slot1 = -1;
        if (t < W - 1) {
            const uint32_t v = s.meet[t];
            if (v != kEmpty) {
                const bool right = (v & kSideRight) != 0;
                const int far = (int)(v & ~kSideRight);
                const int cl = right ? t + 1 : far, ch = right ? far : t;
                const int lo = s.ur[cl].x, hi = s.ur[ch].y;
                const uint32_t id = lo == hi ? S.off + (uint32_t)lo : S.n_ids + S.ibase + (uint32_t)(right ? lo : hi);
                n0 = lds_get<kColors, kMoments>(&s.rec[R * (right ? cl : ch)], id);
                e0 = frontier_entry(n0, lo, hi, 0u);
                slot0 = (int)atomicAdd(&s.nf, 1u);
            }
        }
        if (t < 2 && s.edge_range[t].x >= 0) {
            n1 = lds_get<kColors, kMoments>(s.edge[t], s.edge_id[t]);
            e1 = frontier_entry(n1, s.edge_range[t].x, s.edge_range[t].y, 0u);
            slot1 = (int)atomicAdd(&s.nf, 1u);
        }
        __syncthreads();
        const uint32_t nf = s.nf;
        if (A.last) {
            const LeafSource L = {A.src, A.index_in, sim->unitMass, sim->unitDensity};
            if (slot0 >= 0)
                tree_climb((int)e0.range.x, (int)e0.range.y, S, n0, keys, L, T, &ctr->treeHeight, false, kColors,
                           kMoments);
            if (slot1 >= 0)
                tree_climb((int)e1.range.x, (int)e1.range.y, S, n1, keys, L, T, &ctr->treeHeight, false, kColors,
                           kMoments);
            __syncthreads();   // (the next window reuses every array)
            continue;
        }
        if (slot0 >= 0) s.d[slot0] = (int)e0.range.x;
        if (slot1 >= 0) s.d[slot1] = (int)e1.range.x;
        __syncthreads();
        if (out_n + nf > A.out_cap) {   // (a window leaves at most kTreeBlockFrontier: not reached)
            if (t == 0) atomicOr(&ctr->errorFlags, 16u);
            slot0 = slot1 = -1;
        }
        if (slot0 >= 0) {
            const size_t k = out_n + rank_in(s.d, nf, (int)e0.range.x);
            Fout[k] = e0;
            if (kMoments) moment_words(n0, A.out_mom + 3 * ((size_t)blockIdx.x * A.out_cap + k));
        }
        if (slot1 >= 0) {
            const size_t k = out_n + rank_in(s.d, nf, (int)e1.range.x);
            Fout[k] = e1;
            if (kMoments) moment_words(n1, A.out_mom + 3 * ((size_t)blockIdx.x * A.out_cap + k));
        }
        out_n += nf;
        __syncthreads();   // (the next window reuses every array)
    }
    if (t == 0) A.out_cnt[blockIdx.x] = out_n < A.out_cap ? out_n : A.out_cap;
}

// ------------------------------------------------------------------ merge
// bh_merge.comp:5-23
__device__ __forceinline__ void merge_pair(float4 p1, float4 v1, float4 p2, float4 v2, float4& po, float4& vo) {
    float m = fadd(p1.w, p2.w);
    if (m == 0.0f) { po = vo = make_float4(0.f, 0.f, 0.f, 0.f); return; }
    if (p1.w == 0.0f) { po = p2; vo = v2; return; }
    if (p2.w == 0.0f) { po = p1; vo = v1; return; }
    po.x = fdiv(fadd(fmul(p1.x, p1.w), fmul(p2.x, p2.w)), m);
    po.y = fdiv(fadd(fmul(p1.y, p1.w), fmul(p2.y, p2.w)), m);
    po.z = fdiv(fadd(fmul(p1.z, p1.w), fmul(p2.z, p2.w)), m);
    po.w = m;
    vo.x = fdiv(fadd(fmul(v1.x, p1.w), fmul(v2.x, p2.w)), m);
    vo.y = fdiv(fadd(fmul(v1.y, p1.w), fmul(v2.y, p2.w)), m);
    vo.z = fdiv(fadd(fmul(v1.z, p1.w), fmul(v2.z, p2.w)), m);
    vo.w = fdiv(fadd(p1.w, p2.w), fadd(fdiv(p1.w, v1.w), fdiv(p2.w, v2.w)));
}

// Deterministic resolution in ascending (a, b) order (deterministic
// reservations: a task commits once it holds the smallest pending key on both
// of its bodies, so every earlier task touching them has already committed;
// the result equals processing the tasks one by one in ascending order).
// The reservation words are all ~0 between calls.
//
// A round is reserve -> commit -> release. The first kMergeGridRounds rounds
// run over the whole grid, one launch per phase (most tasks commit in the
// first rounds); k_merge, one block, then finishes the remaining chains with
// block barriers between its phases. Eight grid rounds measured fastest on
// the 4M cube + core merging step (0.25 ms for ~27K tasks, against 0.70 ms
// with four and 0.28 ms with sixteen).
constexpr int kMergeGridRounds = 8;
constexpr uint32_t kMergeGridBlocks = 1024;

__device__ __forceinline__ unsigned long long task_key(uint2 ab) {
    return ((unsigned long long)ab.x << 32) | ab.y;
}

// commit one task if it holds both reservations; returns whether it is done
__device__ __forceinline__ bool merge_commit(float4* __restrict__ dst, uint2 ab,
                                             const unsigned long long* __restrict__ reserve, uint32_t& applied) {
    const unsigned long long key = task_key(ab);
    const unsigned long long ra = __hip_atomic_load(&reserve[ab.x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned long long rb = __hip_atomic_load(&reserve[ab.y], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (ra != key || rb != key) return false;
    const float4 pa = dst[2 * ab.x], va = dst[2 * ab.x + 1];
    const float4 pb = dst[2 * ab.y], vb = dst[2 * ab.y + 1];
    if (pa.w != 0.0f && pb.w != 0.0f) {
        float4 po, vo;
        merge_pair(pa, va, pb, vb, po, vo);
        dst[2 * ab.x] = po;
        dst[2 * ab.x + 1] = vo;
        dst[2 * ab.y] = make_float4(0.f, 0.f, 0.f, 0.f);
        dst[2 * ab.y + 1] = make_float4(0.f, 0.f, 0.f, 0.f);
        applied++;
    }
    return true;
}

__global__ __launch_bounds__(256) void k_merge_round(float4* __restrict__ dst, const uint2* __restrict__ tasks,
                                                     uint32_t cap, DevCounters* __restrict__ ctr,
                                                     unsigned long long* __restrict__ reserve,
                                                     uint32_t* __restrict__ done, int phase) {
    const uint32_t nt = min(ctr->nTasks, cap);
    const uint32_t stride = gridDim.x * blockDim.x;
    uint32_t applied = 0;
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < nt; t += stride) {
        const uint2 ab = tasks[t];
        if (phase == 0) {          // start: nothing done yet; a self-pair is done at once (bh_merge.comp:40)
            done[t] = ab.x == ab.y ? 1u : 0u;
        } else if (phase == 1) {   // reserve
            if (done[t]) continue;
            atomicMin(&reserve[ab.x], task_key(ab));
            atomicMin(&reserve[ab.y], task_key(ab));
        } else if (phase == 2) {   // commit
            if (done[t]) continue;
            if (merge_commit(dst, ab, reserve, applied)) done[t] = 2u;   // 2: done in this round, release pending
        } else {                   // release every reservation this round made
            if (done[t] == 1u) continue;
            __hip_atomic_store(&reserve[ab.x], ~0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&reserve[ab.y], ~0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (done[t] == 2u) done[t] = 1u;
        }
    }
    if (phase == 0 && blockIdx.x == 0 && threadIdx.x == 0) ctr->mergesApplied = 0;
    for (int off = 32; off > 0; off >>= 1) applied += __shfl_xor(applied, off, 64);
    if (applied && lane_id() == 0) atomicAdd(&ctr->mergesApplied, applied);
}

// The remaining chains, one block of 1024 threads. The pending tasks are
// gathered into a list in LDS once, each round works on the list only and
// compacts it; more pending tasks than the list holds go the long way: every
// round scans the whole task array.
constexpr uint32_t kMergeList = 16384;
static_assert(kMergeList / 1024 <= 32, "one keep bit per list entry of a thread");

// block-wide exclusive prefix of one flag per thread (1024 threads); returns
// the thread's offset, *total the block's sum
__device__ __forceinline__ uint32_t block_prefix_1024(bool flag, uint32_t* s_wsum, uint32_t* total) {
    const uint32_t lane = lane_id(), w = threadIdx.x >> 6;
    const unsigned long long b = ballot(flag);
    if (lane == 0) s_wsum[w] = (uint32_t)__popcll(b);
    __syncthreads();
    uint32_t before = 0, sum = 0;
    for (uint32_t k = 0; k < 16; k++) {
        const uint32_t v = s_wsum[k];
        before += k < w ? v : 0u;
        sum += v;
    }
    __syncthreads();
    *total = sum;
    return before + (uint32_t)__popcll(b & lanemask_lt());
}

__global__ __launch_bounds__(1024) void k_merge(float4* __restrict__ dst, const uint2* __restrict__ tasks,
                                                uint32_t cap, DevCounters* __restrict__ ctr,
                                                unsigned long long* __restrict__ reserve,
                                                uint32_t* __restrict__ done) {
    __shared__ uint32_t s_list[kMergeList];
    __shared__ uint32_t s_wsum[16];
    __shared__ int s_pending;
    const uint32_t nt = min(ctr->nTasks, cap);
    const uint32_t tid = threadIdx.x;
    uint32_t applied = 0;
    // the pending tasks, in index order
    uint32_t L = 0;
    for (uint32_t base = 0; base < nt; base += 1024) {
        const uint32_t t = base + tid;
        const bool pend = t < nt && done[t] == 0u;
        uint32_t tot;
        const uint32_t off = block_prefix_1024(pend, s_wsum, &tot);
        if (pend && L + off < kMergeList) s_list[L + off] = t;
        L += tot;
    }
    __syncthreads();
    if (L <= kMergeList) {
        while (L) {
            // a round: reserve, commit, release (each a barrier apart), then
            // keep the tasks that did not commit
            for (uint32_t i = tid; i < L; i += 1024) {
                const uint2 ab = tasks[s_list[i]];
                atomicMin(&reserve[ab.x], task_key(ab));
                atomicMin(&reserve[ab.y], task_key(ab));
            }
            __syncthreads();
            uint32_t keep = 0;   // bit k: entry tid + 1024 k stays pending
            for (uint32_t i = tid, k = 0; i < L; i += 1024, k++)
                if (!merge_commit(dst, tasks[s_list[i]], reserve, applied)) keep |= 1u << k;
            __syncthreads();
            for (uint32_t i = tid; i < L; i += 1024) {
                const uint2 ab = tasks[s_list[i]];
                __hip_atomic_store(&reserve[ab.x], ~0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_store(&reserve[ab.y], ~0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            __syncthreads();
            // in place, chunk by chunk: every entry of a chunk is read before
            // the prefix's barrier and written at or below its own place
            uint32_t kept = 0;
            for (uint32_t base = 0, k = 0; base < L; base += 1024, k++) {
                const uint32_t i = base + tid;
                const uint32_t t = i < L ? s_list[i] : 0u;
                const bool stay = i < L && ((keep >> k) & 1u);
                uint32_t tot;
                const uint32_t off = block_prefix_1024(stay, s_wsum, &tot);
                if (stay) s_list[kept + off] = t;
                kept += tot;
            }
            __syncthreads();
            L = kept;
        }
    } else {
        for (;;) {
            if (tid == 0) s_pending = 0;
            __syncthreads();
            for (uint32_t t = tid; t < nt; t += blockDim.x) {
                if (done[t]) continue;
                const uint2 ab = tasks[t];
                atomicMin(&reserve[ab.x], task_key(ab));
                atomicMin(&reserve[ab.y], task_key(ab));
                s_pending = 1;
            }
            __syncthreads();
            if (!s_pending) break;
            for (uint32_t t = tid; t < nt; t += blockDim.x) {
                if (done[t]) continue;
                if (merge_commit(dst, tasks[t], reserve, applied)) done[t] = 2u;
            }
            __syncthreads();
            for (uint32_t t = tid; t < nt; t += blockDim.x) {
                if (done[t] == 1u) continue;
                const uint2 ab = tasks[t];
                __hip_atomic_store(&reserve[ab.x], ~0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_store(&reserve[ab.y], ~0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (done[t] == 2u) done[t] = 1u;
            }
            __syncthreads();
        }
    }
    for (int off = 32; off > 0; off >>= 1) applied += __shfl_xor(applied, off, 64);
    if (applied && lane_id() == 0) atomicAdd(&ctr->mergesApplied, applied);
}

// ------------------------------------------------------------------ render feed
// Stable compaction in three launches: per-block counts, one block turning
// them into offsets (and the total), then each block writes its records at
// its offset in item order.
__device__ __forceinline__ uint32_t pack_rgba8(float4 c) {
    auto q = [](float x) -> uint32_t { return __float2uint_rn(gmin(gmax(x, 0.0f), 1.0f) * 255.0f); };
    return q(c.x) | (q(c.y) << 8) | (q(c.z) << 16) | (255u << 24);
}

struct FeedNodes {
    const SimValues* sim;
    TreeArrays T;
    int colors;
    uint32_t dmin, dmax;
    __device__ uint32_t items() const { const uint32_t m = sim->numBodies; return m > 1 ? m - 1 : 0; }
    // the static culls of the node-glow and region shaders (emptyAABB: min.x > max.x)
    __device__ bool keep(uint32_t k) const {
        const float4 mn = T.node_aabb[2 * k], mx = T.node_aabb[2 * k + 1];
        const uint32_t d = __float_as_uint(mn.w);
        return !(mn.x > mx.x) && mx.w > 0.0f && T.node_range[k].y > 0 && d >= dmin && d <= dmax;
    }
    __device__ void write(uint32_t k, bh_render_node* out) const {
        const float4 mn = T.node_aabb[2 * k], mx = T.node_aabb[2 * k + 1];
        const float4 com = T.nodes[k].comMass;
        bh_render_node r;
        r.com[0] = com.x; r.com[1] = com.y; r.com[2] = com.z; r.mass = mx.w;
        r.aabb_min[0] = mn.x; r.aabb_min[1] = mn.y; r.aabb_min[2] = mn.z; r.depth = __float_as_uint(mn.w);
        r.aabb_max[0] = mx.x; r.aabb_max[1] = mx.y; r.aabb_max[2] = mx.z;
        r.rgba = colors ? pack_rgba8(T.node_color[k]) : 0u;
        *out = r;
    }
};

struct FeedBodies {
    const SimValues* sim;
    const float4* bodies;
    __device__ uint32_t items() const { return sim->initialNumBodies; }
    __device__ bool keep(uint32_t i) const { return bodies[2 * i].w != 0.0f; }   // alive (common.glsl:203-205)
    __device__ void write(uint32_t i, bh_render_body* out) const {
        const float4 pm = bodies[2 * i], vd = bodies[2 * i + 1];
        bh_render_body r;
        r.pos[0] = pm.x; r.pos[1] = pm.y; r.pos[2] = pm.z;
        // common/common.glsl:228-230 and :235-276, as the leaves' colours
        r.radius = fmul(sim->bodyLengthConstant, powf(fdiv(pm.w, vd.w), 1.0f / 3.0f));
        float4 c = make_float4(0.f, 0.f, 0.f, 0.f);
        if (pm.w > 0.0f && vd.w > 0.0f) c = star_color(fmul(pm.w, sim->unitMass), fmul(vd.w, sim->unitDensity));
        r.rgba = pack_rgba8(c);
        *out = r;
    }
};

template <class F>
__global__ __launch_bounds__(kFeedBlock) void k_feed_count(const F f, uint32_t* __restrict__ counts) {
    __shared__ uint32_t s_w[kFeedBlock / 64];
    const uint32_t i = blockIdx.x * kFeedBlock + threadIdx.x;
    const uint64_t b = ballot(i < f.items() && f.keep(i));
    if (lane_id() == 0) s_w[threadIdx.x >> 6] = (uint32_t)__popcll(b);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (uint32_t w = 0; w < kFeedBlock / 64; w++) t += s_w[w];
        counts[blockIdx.x] = t;
    }
}

// one block of 1024: counts[0 .. nb) -> exclusive offsets, counts[nb] = total
__global__ __launch_bounds__(1024) void k_feed_scan(uint32_t* __restrict__ counts, uint32_t nb) {
    __shared__ uint32_t s[1024];
    const uint32_t per = (nb + 1023) / 1024, t0 = threadIdx.x * per, t1 = min(nb, t0 + per);
    uint32_t sum = 0;
    for (uint32_t b = t0; b < t1; b++) sum += counts[b];
    s[threadIdx.x] = sum;
    __syncthreads();
    for (uint32_t o = 1; o < 1024; o <<= 1) {   // inclusive scan (Hillis-Steele)
        const uint32_t v = threadIdx.x >= o ? s[threadIdx.x - o] : 0u;
        __syncthreads();
        s[threadIdx.x] += v;
        __syncthreads();
    }
    uint32_t run = s[threadIdx.x] - sum;
    for (uint32_t b = t0; b < t1; b++) {
        const uint32_t c = counts[b];
        counts[b] = run;
        run += c;
    }
    if (threadIdx.x == 1023) counts[nb] = s[1023];
}

template <class F, class R>
__global__ __launch_bounds__(kFeedBlock) void k_feed_write(const F f, const uint32_t* __restrict__ offsets,
                                                           R* __restrict__ out, uint32_t cap) {
    __shared__ uint32_t s_w[kFeedBlock / 64];
    const uint32_t i = blockIdx.x * kFeedBlock + threadIdx.x, w = threadIdx.x >> 6;
    const bool k = i < f.items() && f.keep(i);
    const uint64_t b = ballot(k);
    if (lane_id() == 0) s_w[w] = (uint32_t)__popcll(b);
    __syncthreads();
    uint32_t pos = offsets[blockIdx.x];
    for (uint32_t v = 0; v < w; v++) pos += s_w[v];
    pos += (uint32_t)__popcll(b & ((1ull << lane_id()) - 1ull));
    if (k && pos < cap) f.write(i, out + pos);
}

// ------------------------------------------------------------------ boundary layouts
// fill_cap: the cap of the tree's build (0: none); a node higher than it is
// unfilled and shows the build's nodeDepth, the start of its range
// (tree/bh_tree.comp:127), and in readyChildren the number of its children
// that were filled (leaves count as filled).
__global__ void k_nodes_to_std430(const SimValues* __restrict__ sim, TreeArrays T, int colors, int which,
                                  uint32_t fill_cap, NodeStd430* __restrict__ out) {
    const uint32_t n = sim->initialNumBodies, num = sim->numBodies;
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    NodeStd430 o = {};
    if (which == 0) {
        if (k >= n) return;
        if (k < num) {
            float4 p = T.leaf_pos[k];
            o.comMass[0] = p.x; o.comMass[1] = p.y; o.comMass[2] = p.z; o.comMass[3] = p.w;
            o.aabb[0] = o.aabb[3] = p.x; o.aabb[1] = o.aabb[4] = p.y; o.aabb[2] = o.aabb[5] = p.z;
            o.childA = o.childB = kEmpty;
            o.nodeDepth = 0;
            o.readyChildren = kEmpty;
            o.parentId = T.leaf_parent[k];
            if (colors) { float4 c = T.leaf_color[k]; o.avgColor[0] = c.x; o.avgColor[1] = c.y; o.avgColor[2] = c.z; }
        }
        out[k] = o;
    } else {
        if (k + 1 >= n) return;
        if (k + 1 < num || (num == 1 && k == 0)) {
            WalkNode w = T.nodes[k];
            float4 mn = T.node_aabb[2 * k], mx = T.node_aabb[2 * k + 1];
            o.comMass[0] = w.comMass.x; o.comMass[1] = w.comMass.y; o.comMass[2] = w.comMass.z; o.comMass[3] = mx.w;
            o.aabb[0] = mn.x; o.aabb[1] = mn.y; o.aabb[2] = mn.z; o.aabb[3] = mx.x; o.aabb[4] = mx.y; o.aabb[5] = mx.z;
            o.childA = w.childA;
            o.childB = w.childB;
            o.nodeDepth = __float_as_uint(mn.w);
            o.bodiesContained = T.node_range[k].y;
            o.readyChildren = kEmpty;
            if (fill_cap && o.nodeDepth > fill_cap) {
                auto filled = [&](uint32_t c) -> uint32_t {
                    return c < n || __float_as_uint(T.node_aabb[2 * (c - n)].w) <= fill_cap ? 1u : 0u;
                };
                o.nodeDepth = T.node_range[k].x;
                o.readyChildren = filled(w.childA) + filled(w.childB);
            }
            o.parentId = T.node_parent[k];
            if (colors) { float4 c = T.node_color[k]; o.avgColor[0] = c.x; o.avgColor[1] = c.y; o.avgColor[2] = c.z; }
        }
        out[k] = o;
    }
}

__global__ void k_std430_to_nodes(const SimValues* __restrict__ sim, TreeArrays T,
                                  const NodeStd430* __restrict__ leaf_in, const NodeStd430* __restrict__ int_in,
                                  int which) {
    const uint32_t n = sim->initialNumBodies;
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (which == 0 && k < n) {
        const NodeStd430& o = leaf_in[k];
        T.leaf_pos[k] = make_float4(o.comMass[0], o.comMass[1], o.comMass[2], o.comMass[3]);
        T.walk[k] = leaf_record(T.leaf_pos[k], sim->gravitationalConstant);
        T.leaf_parent[k] = o.parentId;
        T.leaf_color[k] = make_float4(o.avgColor[0], o.avgColor[1], o.avgColor[2], o.avgColor[3]);
    }
    if (which == 1 && k + 1 < n) {
        const NodeStd430& o = int_in[k];
        WalkNode w;
        w.comMass = make_float4(o.comMass[0], o.comMass[1], o.comMass[2], fmul(sim->gravitationalConstant, o.comMass[3]));
        float ext = gmax(fsub(o.aabb[3], o.aabb[0]), gmax(fsub(o.aabb[4], o.aabb[1]), fsub(o.aabb[5], o.aabb[2])));
        w.half = fdiv(ext, 2.0f);
        w.childA = o.childA;
        w.childB = o.childB;
        // (an unfilled node's box is empty: a negative half side, never
        // opened by the wide-margin test)
        w.half2 = w.half < 0.0f ? -1.0f : fmul(w.half, w.half);
        T.nodes[k] = w;
        T.node_aabb[2 * k] = make_float4(o.aabb[0], o.aabb[1], o.aabb[2], __uint_as_float(o.nodeDepth));
        T.node_aabb[2 * k + 1] = make_float4(o.aabb[3], o.aabb[4], o.aabb[5], o.comMass[3]);
        T.node_parent[k] = o.parentId;
        T.node_range[k] = make_uint2(0, o.bodiesContained);
        T.node_color[k] = make_float4(o.avgColor[0], o.avgColor[1], o.avgColor[2], o.avgColor[3]);
    }
}

// ids of the positions past numBodies (the dead bodies, which keep their
// places) into the new index
// fills a u32 array (graph-safe replacement for a memset)
__global__ void k_fill_u32(uint32_t* __restrict__ p, uint32_t v, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ------------------------------------------------------------------ launchers
static inline uint32_t cdiv(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

void launch_init(hipStream_t s, uint32_t* index_in, uint32_t n) {
    k_init<<<cdiv(n, 256), 256, 0, s>>>(index_in, n);
}
void launch_reset(hipStream_t s, SimValues* sim, DevCounters* ctr, int set_units, float mass, float density,
                  float len, float time, float G, float blc) {
    k_reset<<<1, 64, 0, s>>>(sim, ctr, set_units, mass, density, len, time, G, blc);
}
// census blocks: at most 1,024, so that the last block's reduction of the
// bounds partials stays short (16M bodies: 61 us with 1,024 blocks, 65 with
// 2,048, 93 with 512; 84 with 4,096)
constexpr uint32_t kCensusBlocks = 1024;
void launch_dead_partition(hipStream_t s, uint32_t n, SimValues* sim, DevCounters* ctr, float4* src, float4* dst,
                           const uint32_t* index_in, uint32_t* index_out, uint32_t* block_scratch,
                           float* bounds_partials, uint32_t partial_cap, const float4* step_pos) {
    uint32_t nb = cdiv(n, kPartTile);
    uint32_t nc = cdiv(n, 256 * 8);
    if (nc > kCensusBlocks) nc = kCensusBlocks;
    if (nc < 1) nc = 1;
    if (bounds_partials) {
        if (nc > partial_cap) nc = partial_cap;
        k_dead_census<true><<<nc, 256, 0, s>>>(sim, ctr, src, dst, bounds_partials, step_pos, index_in);
    } else {
        k_dead_census<false><<<nc, 256, 0, s>>>(sim, ctr, src, dst, nullptr, step_pos, index_in);
    }
    k_dead_count<<<nb, 256, 0, s>>>(sim, ctr, src, index_in, block_scratch);
    k_dead_scan<<<1, 1024, 0, s>>>(sim, ctr, block_scratch);
    k_dead_scatter<<<nb, 256, 0, s>>>(sim, ctr, src, index_in, index_out, block_scratch);
}
void launch_bounds(hipStream_t s, uint32_t n, SimValues* sim, DevCounters* ctr, const float4* src,
                   const uint32_t* index_in, float* partials, uint32_t partial_cap) {
    uint32_t nb = cdiv(n, 256 * 8);
    if (nb > partial_cap) nb = partial_cap;
    if (nb < 1) nb = 1;
    k_bounds<<<nb, 256, 0, s>>>(sim, ctr, src, index_in, partials);
}
void launch_encode(hipStream_t s, uint32_t n, const SimValues* sim, const DevCounters* ctr, const float4* src,
                   const uint32_t* index_in, const float4* step_pos, uint64_t* morton_in) {
    k_encode<<<cdiv(n, 256), 256, 0, s>>>(sim, ctr, src, index_in, step_pos, morton_in);
}
// 16,384 keys per block, at most 1,024 blocks, and (from 65,536 keys) at
// least 64 blocks: a few blocks over a small sort leave the chip idle
uint32_t sort_digit_blocks(uint32_t n) {
    const uint32_t lo = cdiv(n, 1024) < 64 ? cdiv(n, 1024) : 64;
    uint32_t b = cdiv(n, 16384);
    if (b < lo) b = lo;
    return b < 1 ? 1 : (b > 1024 ? 1024 : b);
}
void launch_sort(hipStream_t s, uint32_t n, const SimValues* sim, DevCounters* ctr, uint64_t* keys[2],
                 uint32_t* vals[2], int in, bool positions, uint32_t* partials, uint32_t* totals,
                 unsigned long long* status, uint32_t tiles_cap, SortState* st) {
    const uint32_t nb = sort_digit_blocks(n), tile = sort_tile(n), nt = cdiv(n, tile);
    k_sort_digits<<<nb, 512, 0, s>>>(sim, keys[in], partials, st);
    k_sort_totals<<<kSortPasses * (256 / kTotalsDigits), 1024, 0, s>>>(partials, nb, totals);
    for (uint32_t pass = 0; pass < kSortPasses; pass++) {
        const int i = (in + (int)pass) & 1;
        const uint32_t* vin = (positions && pass == 0) ? nullptr : vals[i];
        unsigned long long* sp = status + (size_t)pass * tiles_cap * 256;
        if (tile == kSortTileSmall)
            k_sort_onesweep<kSortTileSmall, kSortWavesSmall><<<nt, kSortWavesSmall * 64, 0, s>>>(
                sim, ctr, keys[i], vin, keys[1 - i], vals[1 - i], pass, totals, sp, st);
        else
            k_sort_onesweep<kSortTileOne, kSortWavesOne><<<nt, kSortWavesOne * 64, 0, s>>>(
                sim, ctr, keys[i], vin, keys[1 - i], vals[1 - i], pass, totals, sp, st);
    }
}
void launch_tree(hipStream_t s, uint32_t n, const SimValues* sim, DevCounters* ctr, const float4* src,
                 const uint32_t* index_in, const uint64_t* keys, const TreeArrays& T, int colors,
                 const float4* step_pos, const uint32_t* from, uint32_t* ids_out, uint32_t fill_cap) {
    const bool moments = T.node_mom != nullptr;
    const uint32_t B = sim_tree_block(n, moments);
    const TreeLevels Lv = tree_levels(n, B);
    const dim3 grid(cdiv(n, B));
#define BH_TREE(C, M, K) k_tree<C, M, K><<<grid, K, 0, s>>>(sim, ctr, src, index_in, keys, T, step_pos, from, ids_out, fill_cap)
    if (moments) {
        if (colors) BH_TREE(true, true, kTreeBlockMoments);
        else BH_TREE(false, true, kTreeBlockMoments);
    } else if (B == kTreeBlock) {
        if (colors) BH_TREE(true, false, kTreeBlock);
        else BH_TREE(false, false, kTreeBlock);
    } else {
        if (colors) BH_TREE(true, false, kTreeBlockMoments);
        else BH_TREE(false, false, kTreeBlockMoments);
    }
#undef BH_TREE
    // (a single first-launch block builds the whole tree: no levels above)
    auto level_entries = [&](int l) { return T.frontier + 4 * Lv.entry_off[l]; };
    auto level_mom = [&](int l) { return moments ? T.frontier_mom + 3 * Lv.entry_off[l] : nullptr; };
    for (int l = 1; l < Lv.count; l++) {
        // the last level meets leaves by their ids in the new order
        const TreeMidArgs A = {level_entries(l - 1), level_mom(l - 1), T.frontier_cnt + Lv.count_off[l - 1],
                               level_entries(l), level_mom(l), T.frontier_cnt + Lv.count_off[l],
                               Lv.cap[l - 1], Lv.cap[l], Lv.span[l - 1], l == Lv.count - 1 ? 1 : 0, src,
                               from ? ids_out : index_in};
        const dim3 g(Lv.blocks[l]);
        if (moments) {
            if (colors) k_tree_mid<true, true><<<g, kTreeMidUnits, 0, s>>>(sim, ctr, keys, T, A, fill_cap);
            else k_tree_mid<false, true><<<g, kTreeMidUnits, 0, s>>>(sim, ctr, keys, T, A, fill_cap);
        } else if (colors) {
            k_tree_mid<true, false><<<g, kTreeMidUnits, 0, s>>>(sim, ctr, keys, T, A, fill_cap);
        } else {
            k_tree_mid<false, false><<<g, kTreeMidUnits, 0, s>>>(sim, ctr, keys, T, A, fill_cap);
        }
    }
}
void launch_merge(hipStream_t s, float4* dst, const uint2* tasks, uint32_t cap, DevCounters* ctr,
                  unsigned long long* reserve, uint32_t* done) {
    k_merge_round<<<kMergeGridBlocks, 256, 0, s>>>(dst, tasks, cap, ctr, reserve, done, 0);
    for (int r = 0; r < kMergeGridRounds; r++)
        for (int phase = 1; phase <= 3; phase++)
            k_merge_round<<<kMergeGridBlocks, 256, 0, s>>>(dst, tasks, cap, ctr, reserve, done, phase);
    k_merge<<<1, 1024, 0, s>>>(dst, tasks, cap, ctr, reserve, done);
}
void launch_nodes_to_std430(hipStream_t s, uint32_t n, const SimValues* sim, const TreeArrays& T, int colors,
                            int which, uint32_t fill_cap, NodeStd430* out) {
    k_nodes_to_std430<<<cdiv(n, 256), 256, 0, s>>>(sim, T, colors, which, fill_cap, out);
}
void launch_render_nodes(hipStream_t s, uint32_t n, const SimValues* sim, const TreeArrays& T, int colors,
                         uint32_t depth_min, uint32_t depth_max, uint32_t* counts, bh_render_node* out, uint32_t cap) {
    const FeedNodes f = {sim, T, colors, depth_min, depth_max};
    const uint32_t nb = render_feed_blocks(n);   // (at most n - 1 internal nodes)
    k_feed_count<<<nb, kFeedBlock, 0, s>>>(f, counts);
    k_feed_scan<<<1, 1024, 0, s>>>(counts, nb);
    k_feed_write<<<nb, kFeedBlock, 0, s>>>(f, counts, out, cap);
}
void launch_render_bodies(hipStream_t s, uint32_t n, const SimValues* sim, const float4* bodies, uint32_t* counts,
                          bh_render_body* out, uint32_t cap) {
    const FeedBodies f = {sim, bodies};
    const uint32_t nb = render_feed_blocks(n);
    k_feed_count<<<nb, kFeedBlock, 0, s>>>(f, counts);
    k_feed_scan<<<1, 1024, 0, s>>>(counts, nb);
    k_feed_write<<<nb, kFeedBlock, 0, s>>>(f, counts, out, cap);
}
void launch_std430_to_nodes(hipStream_t s, uint32_t n, const SimValues* sim, const TreeArrays& T,
                            const NodeStd430* leaf_in, const NodeStd430* int_in, int which) {
    k_std430_to_nodes<<<cdiv(n, 256), 256, 0, s>>>(sim, T, leaf_in, int_in, which);
}
void launch_fill_u32(hipStream_t s, uint32_t* p, uint32_t v, size_t n) {
    if (n) k_fill_u32<<<cdiv(n, 256), 256, 0, s>>>(p, v, n);
}

}  // namespace bh

namespace bh {

// ------------------------------------------------------------ FP64 direct sum
// a_i = sum_{j != i, m_j != 0} G m_j r_ij (|r_ij|^2 + softening)^-3/2 in FP64 for
// sampled bodies i; one block per sampled body, fixed-order reduction.
__global__ __launch_bounds__(256) void k_direct_sum_f64(const float4* __restrict__ bodies, uint32_t n,
                                                        const uint32_t* __restrict__ sample, double G,
                                                        double soft, double* __restrict__ out) {
    __shared__ double s_acc[3][4];
    const uint32_t i = sample[blockIdx.x];
    const float4 pi = bodies[2 * i];
    const double xi = pi.x, yi = pi.y, zi = pi.z;
    double ax = 0.0, ay = 0.0, az = 0.0;
    for (uint32_t j = threadIdx.x; j < n; j += blockDim.x) {
        const float4 pj = bodies[2 * j];
        if (j == i || pj.w == 0.0f) continue;
        const double rx = (double)pj.x - xi, ry = (double)pj.y - yi, rz = (double)pj.z - zi;
        const double d2 = rx * rx + ry * ry + rz * rz + soft;
        const double inv = 1.0 / sqrt(d2);
        const double f = G * (double)pj.w * inv * inv * inv;
        ax += f * rx;
        ay += f * ry;
        az += f * rz;
    }
    for (int off = 32; off > 0; off >>= 1) {
        ax += __shfl_xor(ax, off, 64);
        ay += __shfl_xor(ay, off, 64);
        az += __shfl_xor(az, off, 64);
    }
    const uint32_t w = threadIdx.x >> 6;
    if (lane_id() == 0) { s_acc[0][w] = ax; s_acc[1][w] = ay; s_acc[2][w] = az; }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 0; k < 3; k++) out[3 * blockIdx.x + k] = (s_acc[k][0] + s_acc[k][1]) + (s_acc[k][2] + s_acc[k][3]);
    }
}

void launch_direct_sum_f64(hipStream_t s, const float4* bodies, uint32_t n, const uint32_t* sample,
                           uint32_t n_sample, double G, double soft, double* out) {
    if (n_sample) k_direct_sum_f64<<<n_sample, 256, 0, s>>>(bodies, n, sample, G, soft, out);
}

}  // namespace bh

namespace bh {

// Self-test of the walk's fast exact inverse square root against the general
// IEEE sequences, over every float bit pattern of its fast domain.
__global__ void k_selftest_math(uint32_t* __restrict__ bad) {
    const uint32_t lo = 0x0D800000u;   // 2^-100
    const uint32_t hi = 0x71800000u;   // 2^100
    const uint32_t rlo = 0x26800000u;  // 2^-50
    const uint32_t rhi = 0x58800000u;  // 2^50
    uint32_t b0 = 0, b1 = 0, b2 = 0;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t u = lo + blockIdx.x * blockDim.x + threadIdx.x; u <= hi; u += stride) {
        const float x = __uint_as_float(u);
        const float s = sqrt_cr_normal(x);
        if (__float_as_uint(s) != __float_as_uint(fsqrt(x))) b0++;
        if (__float_as_uint(inv_sqrt_exact(x)) != __float_as_uint(fdiv(1.0f, fsqrt(x)))) b2++;
        if (u >= rlo && u <= rhi && __float_as_uint(rcp_cr_normal(x)) != __float_as_uint(fdiv(1.0f, x))) b1++;
    }
    if (b0) atomicAdd(&bad[0], b0);
    if (b1) atomicAdd(&bad[1], b1);
    if (b2) atomicAdd(&bad[2], b2);
}

void launch_selftest_math(hipStream_t s, uint32_t* bad) { k_selftest_math<<<4096, 256, 0, s>>>(bad); }

}  // namespace bh

// ====================================================================== ensembles
// Independent simulations concatenated in one set of arrays (no reference
// counterpart: the reference runs one simulation). Every stage runs over all
// simulations in one launch; simulation s touches only its own segment of
// positions, body ids and internal nodes, so each simulation's bounds, codes,
// permutation, tree and accelerations are those of a run of it alone.
namespace bh {

__global__ void k_ens_reset(DevCounters* __restrict__ ctr, EnsQueue* __restrict__ queues) {
    if (threadIdx.x < kEnsQueues) queues[threadIdx.x].head = 0;
    if (threadIdx.x != 0) return;
    ctr->opened = 0;
    ctr->interactions = 0;
    ctr->nTasks = 0;
    ctr->mergesApplied = 0;
    ctr->maxStack = 0;
    ctr->errorFlags = 0;
    ctr->corrected = 0;
    ctr->quadTerms = 0;
    ctr->treeHeight = 0;
}

// One block per tile; the last tile of a simulation to finish (ticket)
// reduces that simulation's tile partials and updates its bounds.
__global__ __launch_bounds__(256) void k_ens_bounds(EnsSim* __restrict__ sims, const EnsTile* __restrict__ tiles,
                                                    const float4* __restrict__ src, float* __restrict__ partials) {
    __shared__ float s_r[6][4];
    __shared__ bool s_last;
    const EnsTile tl = tiles[blockIdx.x];
    EnsSim* S = sims + tl.sim;
    const float prev[6] = {S->bmin[0], S->bmin[1], S->bmin[2], S->bmax[0], S->bmax[1], S->bmax[2]};
    const BoundsWindow W = bounds_window(prev, prev + 3);
    float r[6] = {1e38f, 1e38f, 1e38f, -1e38f, -1e38f, -1e38f};
    // the bodies of a simulation are its ids [off, off + num): read them in id order
    for (uint32_t e = threadIdx.x; e < tl.count; e += blockDim.x) bounds_accum(W, src[2 * (tl.start + e)], r);
    bounds_block_reduce(r, s_r);
    if (threadIdx.x == 0) {
        for (int k = 0; k < 6; k++)
            __hip_atomic_store(&partials[6 * blockIdx.x + k], r[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        uint32_t ticket = __hip_atomic_fetch_add(&S->ticket, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        s_last = ticket == S->ntiles - 1;
    }
    __syncthreads();
    if (!s_last) return;
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    float res[6] = {1e38f, 1e38f, 1e38f, -1e38f, -1e38f, -1e38f};
    for (uint32_t k = S->tile0 + threadIdx.x; k < S->tile0 + S->ntiles; k += blockDim.x) {
        for (int c = 0; c < 3; c++)
            res[c] = gmin(res[c], __hip_atomic_load(&partials[6 * k + c], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
        for (int c = 3; c < 6; c++)
            res[c] = gmax(res[c], __hip_atomic_load(&partials[6 * k + c], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    }
    bounds_block_reduce(res, s_r);
    if (threadIdx.x != 0) return;
    S->ticket = 0;
    bounds_collapse(prev, res, S->bmin, S->bmax);
}

__global__ __launch_bounds__(256) void k_ens_encode(const EnsSim* __restrict__ sims,
                                                    const EnsTile* __restrict__ tiles,
                                                    const float4* __restrict__ src,
                                                    const uint32_t* __restrict__ index_in,
                                                    uint64_t* __restrict__ keys) {
    const EnsTile tl = tiles[blockIdx.x];
    const EnsSim* S = sims + tl.sim;
    const float bmin[3] = {S->bmin[0], S->bmin[1], S->bmin[2]};
    const float bmax[3] = {S->bmax[0], S->bmax[1], S->bmax[2]};
    const bool empty = bmin[0] > bmax[0];
    for (uint32_t e = threadIdx.x; e < tl.count; e += blockDim.x) {
        const uint32_t g = tl.start + e;
        keys[g] = morton_of(src[2 * index_in[g]], bmin, bmax, empty);
    }
}

// Segmented LSD radix sort: digit histograms per tile, then per simulation an
// exclusive scan over (digit, tile) that starts at the simulation's first
// position, then a stable scatter with in-wave ballot ranking. A simulation's
// keys never leave its segment, so the result is the stable sort by
// (simulation, Morton code) of the whole ensemble.
template <int BITS>
__global__ __launch_bounds__(256) void k_ens_sort_hist(const EnsTile* __restrict__ tiles,
                                                       const uint64_t* __restrict__ keys, uint32_t shift,
                                                       uint32_t* __restrict__ hist) {
    constexpr uint32_t R = 1u << BITS;
    __shared__ uint32_t s_h[4][256];
    const EnsTile tl = tiles[blockIdx.x];
    const uint32_t c = tile_histogram<BITS>(s_h, keys, tl.start, tl.count, shift);
    if (threadIdx.x < R) hist[(size_t)blockIdx.x * 256 + threadIdx.x] = c;
}

template <int BITS>
__global__ __launch_bounds__(256) void k_ens_sort_scan(const EnsSim* __restrict__ sims, uint32_t* __restrict__ hist) {
    constexpr uint32_t R = 1u << BITS;
    __shared__ uint32_t s_w[4];
    const EnsSim S = sims[blockIdx.x];
    uint32_t* h = hist + (size_t)S.tile0 * 256;
    const uint32_t d = threadIdx.x, lane = lane_id();
    // this digit's total over the simulation's tiles (8 loads in flight)
    constexpr uint32_t kU = 8;
    uint32_t total = 0;
    if (d < R)
        for (uint32_t k0 = 0; k0 < S.ntiles; k0 += kU) {
            uint32_t v[kU];
            for (uint32_t j = 0; j < kU; j++) v[j] = k0 + j < S.ntiles ? h[(size_t)(k0 + j) * 256 + d] : 0u;
            for (uint32_t j = 0; j < kU; j++) total += v[j];
        }
    uint32_t incl = total;
    for (int off = 1; off < 64; off <<= 1) {
        uint32_t v = __shfl_up(incl, off, 64);
        if (lane >= (uint32_t)off) incl += v;
    }
    if (lane == 63) s_w[d >> 6] = incl;
    __syncthreads();
    uint32_t before = S.off;
    for (uint32_t k = 0; k < (d >> 6); k++) before += s_w[k];
    // each tile's start of this digit: the digit's base plus the counts of
    // the simulation's earlier tiles
    uint32_t run = before + incl - total;
    if (d < R)
        for (uint32_t k0 = 0; k0 < S.ntiles; k0 += kU) {
            uint32_t v[kU];
            for (uint32_t j = 0; j < kU; j++) v[j] = k0 + j < S.ntiles ? h[(size_t)(k0 + j) * 256 + d] : 0u;
            for (uint32_t j = 0; j < kU; j++)
                if (k0 + j < S.ntiles) {
                    h[(size_t)(k0 + j) * 256 + d] = run;
                    run += v[j];
                }
        }
}

template <int BITS>
__global__ __launch_bounds__(256) void k_ens_sort_scatter(const EnsTile* __restrict__ tiles,
                                                          const uint64_t* __restrict__ keys_in,
                                                          const uint32_t* __restrict__ idx_in,
                                                          uint64_t* __restrict__ keys_out,
                                                          uint32_t* __restrict__ idx_out, uint32_t shift,
                                                          const uint32_t* __restrict__ hist) {
    constexpr uint32_t R = 1u << BITS;
    __shared__ ScatterLds<kEnsTile> L;
    const EnsTile tl = tiles[blockIdx.x];
    const uint32_t t = threadIdx.x;
    L.base[t] = t < R ? hist[(size_t)blockIdx.x * 256 + t] : 0u;
    scatter_tile<BITS, kEnsTile>(L, keys_in, idx_in, tl.start, tl.count, keys_out, idx_out, shift);
}

// Tree of every simulation: the block-local first launch over two blocks per
// tile (tiles never span simulations).
template <bool kMoments>
__global__ __launch_bounds__(tree_block_size(kMoments)) void k_ens_tree(const EnsSim* __restrict__ sims,
                                                         const EnsTile* __restrict__ tiles, uint32_t N,
                                                         DevCounters* __restrict__ ctr,
                                                         const float4* __restrict__ src,
                                                         const uint32_t* __restrict__ index_in,
                                                         const uint64_t* __restrict__ keys, TreeArrays T,
                                                         float G) {
    __shared__ TreeBlockLds<false, kMoments> s;
    constexpr uint32_t kB = tree_block_size(kMoments), kPerTile = kEnsTile / kB;
    const EnsTile tl = tiles[blockIdx.x / kPerTile];
    const uint32_t first = (blockIdx.x % kPerTile) * kB;
    if (first >= tl.count) return;
    const EnsSim& S = sims[tl.sim];
    const int cnt = tl.count - first < kB ? (int)(tl.count - first) : (int)kB;
    const TreeSeg seg = {S.off, S.num, S.ibase, N, G};
    const LeafSource L = {src, index_in, 0.0f, 0.0f};
    tree_block<false, kMoments>(s, seg, (int)(tl.start + first - S.off), cnt, keys + S.off, L, T, &ctr->treeHeight,
                                true, tl.sim);
}

__global__ __launch_bounds__(256) void k_ens_tree_upper(const EnsSim* __restrict__ sims, uint32_t N,
                                                        DevCounters* __restrict__ ctr, const float4* __restrict__ src,
                                                        const uint32_t* __restrict__ index_in,
                                                        const uint64_t* __restrict__ keys, TreeArrays T, float G) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= *T.frontier_count) return;
    const FrontierEntry e = reinterpret_cast<const FrontierEntry*>(T.frontier)[t];
    const EnsSim& S = sims[e.range.z];
    const TreeSeg seg = {S.off, S.num, S.ibase, N, G};
    const LeafSource L = {src, index_in, 0.0f, 0.0f};
    NodeState me = frontier_state(e);
    const bool moments = T.node_mom != nullptr;
    if (moments) frontier_moments(T, t, me);
    tree_climb((int)e.range.x, (int)e.range.y, seg, me, keys + S.off, L, T, &ctr->treeHeight, true, 0, moments);
}

__global__ void k_ens_nodes_std430(TreeArrays T, uint32_t n_internal, NodeStd430* __restrict__ out) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n_internal) return;
    NodeStd430 o = {};
    const WalkNode w = T.nodes[k];
    const float4 mn = T.node_aabb[2 * k], mx = T.node_aabb[2 * k + 1];
    o.comMass[0] = w.comMass.x; o.comMass[1] = w.comMass.y; o.comMass[2] = w.comMass.z; o.comMass[3] = mx.w;
    o.aabb[0] = mn.x; o.aabb[1] = mn.y; o.aabb[2] = mn.z; o.aabb[3] = mx.x; o.aabb[4] = mx.y; o.aabb[5] = mx.z;
    o.childA = w.childA;
    o.childB = w.childB;
    o.nodeDepth = __float_as_uint(mn.w);
    o.bodiesContained = T.node_range[k].y;
    o.readyChildren = kEmpty;
    o.parentId = T.node_parent[k];
    out[k] = o;
}

void launch_ens_reset(hipStream_t s, DevCounters* ctr, EnsQueue* queues) {
    k_ens_reset<<<1, 64, 0, s>>>(ctr, queues);
}

void launch_ens_bounds(hipStream_t s, EnsSim* sims, const EnsTile* tiles, uint32_t ntiles, const float4* src,
                       float* partials) {
    k_ens_bounds<<<ntiles, 256, 0, s>>>(sims, tiles, src, partials);
}

void launch_ens_encode(hipStream_t s, const EnsSim* sims, const EnsTile* tiles, uint32_t ntiles, const float4* src,
                       const uint32_t* index_in, uint64_t* keys) {
    k_ens_encode<<<ntiles, 256, 0, s>>>(sims, tiles, src, index_in, keys);
}

template <int BITS>
static void ens_sort_pass(hipStream_t s, const EnsSim* sims, uint32_t nsims, const EnsTile* tiles, uint32_t ntiles,
                          const uint64_t* keys_in, const uint32_t* idx_in, uint64_t* keys_out, uint32_t* idx_out,
                          uint32_t shift, uint32_t* hist) {
    k_ens_sort_hist<BITS><<<ntiles, 256, 0, s>>>(tiles, keys_in, shift, hist);
    k_ens_sort_scan<BITS><<<nsims, 256, 0, s>>>(sims, hist);
    k_ens_sort_scatter<BITS><<<ntiles, 256, 0, s>>>(tiles, keys_in, idx_in, keys_out, idx_out, shift, hist);
}

int launch_ens_sort_pass(hipStream_t s, int bits, const EnsSim* sims, uint32_t nsims, const EnsTile* tiles,
                         uint32_t ntiles, const uint64_t* keys_in, const uint32_t* idx_in, uint64_t* keys_out,
                         uint32_t* idx_out, uint32_t shift, uint32_t* hist) {
    switch (bits) {
        case 5: ens_sort_pass<5>(s, sims, nsims, tiles, ntiles, keys_in, idx_in, keys_out, idx_out, shift, hist); break;
        case 6: ens_sort_pass<6>(s, sims, nsims, tiles, ntiles, keys_in, idx_in, keys_out, idx_out, shift, hist); break;
        case 7: ens_sort_pass<7>(s, sims, nsims, tiles, ntiles, keys_in, idx_in, keys_out, idx_out, shift, hist); break;
        case 8: ens_sort_pass<8>(s, sims, nsims, tiles, ntiles, keys_in, idx_in, keys_out, idx_out, shift, hist); break;
        default: return 1;
    }
    return 0;
}

void launch_ens_tree(hipStream_t s, const EnsSim* sims, const EnsTile* tiles, uint32_t ntiles, uint32_t N,
                     DevCounters* ctr, const float4* src, const uint32_t* index_in, const uint64_t* keys,
                     const TreeArrays& T, float G) {
    k_fill_u32<<<1, 64, 0, s>>>(T.frontier_count, 0u, 1);
    const uint32_t B = tree_block_size(T.node_mom != nullptr), blocks = ens_tree_blocks(ntiles, T.node_mom != nullptr);
    if (T.node_mom)
        k_ens_tree<true><<<blocks, B, 0, s>>>(sims, tiles, N, ctr, src, index_in, keys, T, G);
    else
        k_ens_tree<false><<<blocks, B, 0, s>>>(sims, tiles, N, ctr, src, index_in, keys, T, G);
    k_ens_tree_upper<<<cdiv(frontier_cap(N, blocks), 256), 256, 0, s>>>(sims, N, ctr, src, index_in, keys, T, G);
}

void launch_ens_nodes_std430(hipStream_t s, const TreeArrays& T, uint32_t n_internal, NodeStd430* out) {
    if (n_internal) k_ens_nodes_std430<<<cdiv(n_internal, 256), 256, 0, s>>>(T, n_internal, out);
}

}  // namespace bh
