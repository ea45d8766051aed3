// bh_kernels.hpp — launcher declarations and the tree arrays shared by the
// kernels and the host side of the C ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "bh_device.hpp"

#include "../../include/bh_abi.h"

namespace bh {

constexpr uint32_t kPartTile = 1024;   // dead partition tile (sorted positions per block)
constexpr uint32_t kSortTile = 2048;   // radix sort tile of the ensembles (keys per block)
// radix sort tile of one simulation: 8,192 keys per block of 16 waves (8 rows
// of 64 keys each), one block per CU (114 KB of LDS). Larger tiles shorten
// the look-back and lengthen the digit runs written out: at 16M keys the
// sort takes 1.24 ms with these tiles, 1.33 ms with 4,096 and 1.78 ms with
// 2,048 (MI355X, same box)
constexpr uint32_t kSortTileOne = 8192;
constexpr int kSortWavesOne = kSortTileOne / 512;
// Up to kSortSmallMax keys a pass has too few 8,192-key tiles to fill the
// chip (8 at 65,536 keys), and each takes as long as at 16M: smaller tiles
// of 8 waves then.
constexpr uint32_t kSortTileSmall = 2048;
constexpr int kSortWavesSmall = 8;
constexpr uint32_t kSortSmallMax = 1u << 18;
inline uint32_t sort_tile(uint32_t n) { return n <= kSortSmallMax ? kSortTileSmall : kSortTileOne; }
constexpr uint32_t kSortBits = 8;
constexpr uint32_t kSortPasses = 8;    // 64 key bits; even, so the result lands back in the In buffers
constexpr int kWalkWaves = 16;         // waves per walk block (adjacent body groups share a CU's scalar cache)
// persistent walk: one queue of 64-body groups per CU slot (8 XCDs x 4 shader
// engines x 9 CU ids; 32 of the 36 slots of an XCD hold a CU), each head in
// its own 128-byte line
constexpr uint32_t kWalkSlots = 36, kWalkQueues = 8 * kWalkSlots;
constexpr bool kWalkPersistent = true;
// below this many bodies a queue holds too few groups to repay its claims:
// the walk runs as one block of kWalkSmallWaves waves per 4 groups instead
// (MI355X, Plummer: that is faster at 524,288 bodies, slower at 1,048,576),
// small blocks spreading a small walk over all CUs
constexpr uint32_t kWalkPersistentMin = 1u << 20;
constexpr int kWalkSmallWaves = 4;     // (2 and 8 measured no faster at 65,536-524,288 bodies)
constexpr int kWalkSmallBodies = 64;   // bodies per wave of that walk
// up to this many bodies a 64-body wave per SIMD leaves the walk waiting on
// its own latency: the plain walk (whose sums do not depend on how bodies are
// grouped) runs waves of 32 bodies (two per SIMD) instead
constexpr uint32_t kWalkHalfMax = 1u << 16;
constexpr int kStackDepth = 192;       // per-wave walk stack; a tree over 96-bit keys is at most 96 high

// Tree in HBM. Leaves are sorted positions 0..n-1, internal nodes 0..n-2
// (global id n + k). The walk reads `nodes` (32 B) and `leaf_pos` (16 B);
// the rest serves the fill and the 80-byte Node layout at the boundary.
struct TreeArrays {
    float4* leaf_pos;        // [n] posMass in sorted order
    uint32_t* leaf_parent;   // [n]
    float4* leaf_color;      // [n]
    WalkNode* walk;          // [2n-1] the walk's records (centre of mass and G m): leaf j at j (half 0,
                             // half2 -1), then the internal nodes; node id i is walk[i]
    WalkNode* nodes;         // = walk + n: internal node k
    float4* node_aabb;       // [2(n-1)] min, max; min.w holds the height (nodeDepth) as uint bits, max.w the mass
    uint32_t* node_parent;   // [n-1]
    uint2* node_range;       // [n-1] (first sorted position, bodies contained)
    float4* node_color;      // [n-1]
    uint32_t* meet;          // [n-1] bottom-up arrival slots, ~0 between builds
    float4* xchg;            // [8(n-1)] summaries of the two children meeting at each split (4 words each)
    float4* frontier;        // roots of the subtrees the first build launch left, 4 words each: the
                             // ensembles' list (frontier_cap entries), one simulation's level slots
                             // (frontier_slots entries)
    uint32_t* frontier_count = nullptr;
    // With frontier_cnt set, the first launch writes block b's entries in
    // sorted order to slots [b kTreeBlockFrontier, ...) of frontier and their
    // number to frontier_cnt[b] (the levels of tree_levels() follow); unset,
    // it appends them in any order at frontier_count (the ensemble build).
    uint32_t* frontier_cnt = nullptr;
    // Second moments, built with the tree for the corrected walk (all null
    // otherwise): node_mom holds each internal node's second moment about its
    // (FP32) centre of mass, the records below carry it and the node's first
    // moment about that centre through the exchanges.
    float4* node_mom = nullptr;       // [2(n-1)] (xx, xy, xz, yy), (yz, zz, 0, 0)
    float4* xchg_mom = nullptr;       // [6(n-1)] the moment words of xchg's two sides (3 words each)
    float4* frontier_mom = nullptr;   // the moment words of the frontier entries (3 per entry)
    float4* corr_leaf = nullptr;      // the corrected walk's records: the first launch writes the first
                                      // half of every leaf's (leaf j at 4 j, 4 j + 1)
};

// The first tree launch builds each block of kTreeBlock sorted positions on
// its own. A block leaves at most one frontier entry per level of the two
// chains of nodes crossing its edges, plus the two edge nodes (a tree over
// 96-bit keys is at most 98 levels high); the entries are disjoint subtrees,
// so there are also at most as many as positions.
constexpr uint32_t kTreeBlock = 1024;
// The build with second moments carries three more words per record in LDS,
// so a block of 1,024 positions leaves room for one block per CU: blocks of
// 512 positions measured faster (16M Plummer: first launch 1.85 against 2.25
// ms, the second 0.43 against 0.30 ms). The ensembles' plain build keeps
// 1,024 (its second launch climbs through HBM, slower from twice as many
// entries).
constexpr uint32_t kTreeBlockMoments = 512;
__host__ __device__ constexpr uint32_t tree_block_size(bool moments) {
    return moments ? kTreeBlockMoments : kTreeBlock;
}
// The one-simulation build, whose upper levels are cheap (k_tree_mid), takes
// blocks of 512 positions for the plain tree too above kTreeWideMax bodies
// (four blocks per CU: MI355X, Plummer, plain tree stage 1.31 against 1.69 ms
// at 16M and 0.149 against 0.178 ms at 1M), and 1,024 up to it (0.074 against
// 0.085 ms at 65,536).
constexpr uint32_t kTreeWideMax = 1u << 17;
inline uint32_t sim_tree_block(uint32_t n, bool moments) {
    return moments || n > kTreeWideMax ? kTreeBlockMoments : kTreeBlock;
}
constexpr uint32_t kTreeBlockFrontier = 2 * 98 + 2;
inline uint32_t frontier_cap(uint32_t n, uint32_t blocks) {
    const uint64_t b = (uint64_t)blocks * kTreeBlockFrontier;
    return (uint32_t)(b < n ? b : n) + 1;
}

// The levels of the one-simulation build. Level 0 is the first launch (one
// block per tree_block_size positions). Each later level is one launch of
// k_tree_mid: a block takes the entries that kTreeMidGroup consecutive
// blocks of the level below left, in sorted order, as the leaves of a
// smaller tree (a node covers whole entries, and the split between two
// neighbours is the split between their position ranges), builds in LDS
// every node whose entries are all its own, and leaves the rest, sorted, to
// the next level. It takes kTreeMidUnits entries at a time, so a block that
// gets more runs several windows, each leaving at most kTreeBlockFrontier
// entries. The last level has one block, which climbs from what its windows
// leave (nothing unless there are several) to the root through the arrival
// slots in HBM. A single first-launch block builds the whole tree.
// Plummer spheres leave about 10 entries per first-launch block, so one
// window of 1,024 takes a whole group of 64 blocks.
constexpr uint32_t kTreeMidGroup = 64;
constexpr uint32_t kTreeMidUnits = 1024;
constexpr int kTreeMaxLevels = 8;
struct TreeLevels {
    int count = 0;
    uint32_t blocks[kTreeMaxLevels] = {}, cap[kTreeMaxLevels] = {};   // blocks, entry slots per block
    uint64_t span[kTreeMaxLevels] = {};                                // sorted positions per block
    size_t entry_off[kTreeMaxLevels] = {}, count_off[kTreeMaxLevels] = {};
    size_t entries = 0, counts = 0;                                    // totals over the levels
};
inline TreeLevels tree_levels(uint32_t n, uint32_t B) {
    TreeLevels L;
    uint64_t blocks = n > B ? ((uint64_t)n + B - 1) / B : 1;
    uint64_t cap = kTreeBlockFrontier < B ? kTreeBlockFrontier : B;
    uint64_t span = B;
    for (;;) {
        const int l = L.count++;
        L.blocks[l] = (uint32_t)blocks;
        L.cap[l] = (uint32_t)cap;
        L.span[l] = span;
        L.entry_off[l] = L.entries;
        L.count_off[l] = L.counts;
        L.entries += blocks * cap;
        L.counts += blocks;
        if (blocks == 1) break;
        const uint64_t units = (blocks < kTreeMidGroup ? blocks : kTreeMidGroup) * cap;
        const uint64_t windows = (units + kTreeMidUnits - 1) / kTreeMidUnits;
        cap = units < windows * kTreeBlockFrontier ? units : windows * kTreeBlockFrontier;
        blocks = (blocks + kTreeMidGroup - 1) / kTreeMidGroup;
        span *= kTreeMidGroup;
    }
    return L;
}
// entry slots and counts for either plan
inline size_t frontier_slots(uint32_t n) {
    const size_t a = tree_levels(n, kTreeBlock).entries, b = tree_levels(n, kTreeBlockMoments).entries;
    return a > b ? a : b;
}
inline size_t frontier_counts(uint32_t n) {
    const size_t a = tree_levels(n, kTreeBlock).counts, b = tree_levels(n, kTreeBlockMoments).counts;
    return a > b ? a : b;
}

struct WalkArgs {
    const SimValues* sim;
    DevCounters* ctr;
    const float4* src;
    float4* dst;
    const uint32_t* index_in;
    const float4* leaf_pos;
    const WalkNode* nodes;
    uint2* tasks;
    uint32_t task_cap;
    float theta, dt, softening;
    uint32_t merging, dynamic, wrap;
    float elasticity, restitution;   // COLLISION mode
    float* leaf_radius;              // [n] scratch of the merging and collision modes
    uint32_t general;                // 1: always the general walk loop
    float4* pos_out = nullptr;       // [n] new position-mass by sorted position as well, or nullptr
    uint32_t* queues = nullptr;      // [kWalkQueues * 32] per-CU work queues (persistent walk), or nullptr
    // the corrected walk: its records [4 corr_records(n)], or nullptr
    const float4* qrec = nullptr;
    // the corrected walk's |a| by body id [n]: read by the relative opening
    // test (rel), written by every corrected walk
    float* acc_mag = nullptr;
    uint32_t rel = 0;
    float alpha = 0.0f;
};

// The corrected walk's records, 64 bytes per node id (four float4):
//   (centre of mass, G M), (K of corr_rel_k, children word, childB, b^2 of
//   corr_size2), -3 G (T_xx, T_xy, T_xz, T_yy), (-3 G T_yz, -3 G T_zz,
//   -1.5 G tr T, 0).
// Leaf j is record j (its position and G m, K -1, b^2 -1, no children,
// zero quadrupole terms); internal node k is record n + kCorrGap + k, and the
// child ids are in the same numbering, so every internal id is at least 64
// above every leaf id. The kCorrGap records in between are never read.
constexpr uint32_t kCorrGap = 64;
__host__ __device__ inline uint32_t corr_internal_id(uint32_t n, uint32_t k) { return n + kCorrGap + k; }
__host__ __device__ inline size_t corr_records(uint32_t n) { return 2 * (size_t)n - 1 + kCorrGap; }
// The children word names both children in 32 bits: childA's id, bit 30 set
// if childA is internal, bit 31 if childB is. The tree splits a node's sorted
// range [lo, hi] at one position g: childA is leaf g (g = lo) or the internal
// node of index g, childB leaf g + 1 (g + 1 = hi) or the internal node of
// index g + 1 (bh_tree.comp:96-113). So childB's id is childA's plus
// 1, minus ioff if childA is internal, plus ioff if childB is: ioff is the
// difference between the id of internal node g and that of leaf g (n +
// kCorrGap for one simulation; an ensemble's simulation s, whose leaves
// start at sorted position off_s and internal nodes at index off_s - s, has
// N + kCorrGap - s).
__host__ __device__ inline uint32_t corr_children_word(uint32_t a_id, bool a_internal, bool b_internal) {
    return a_id | (a_internal ? 1u << 30 : 0u) | (b_internal ? 1u << 31 : 0u);
}
__host__ __device__ inline uint32_t corr_child_a(uint32_t w) { return w & 0x3FFFFFFFu; }
__host__ __device__ inline uint32_t corr_child_b(uint32_t w, uint32_t ioff) {
    return (w & 0x3FFFFFFFu) + 1u + ((w >> 31) ? ioff : 0u) - (((w >> 30) & 1u) ? ioff : 0u);
}
// every record's byte offset fits 32 bits
__host__ __device__ inline bool corr_offsets_32bit(uint32_t n) { return corr_records(n) <= (1ull << 26); }
// The corrected walk's range: softening at least 1e-14 and every s^2 at most
// 16 h_root^2 + softening <= 1e15, so that 1/s^5 stays a normal FP32 number
// between 1e-37.5 and 1e35 (the walk forms the monopole as G M s^2 / s^5).
__host__ __device__ inline bool corr_range_ok(float root_half, float soft) {
    return soft >= 1e-14f && 16.0f * root_half * root_half + soft <= 1e15f;
}

// The size in the corrected walk's opening test (record word 7): the squared
// distance b^2 from the node's centre of mass to the farthest corner of its
// box, which bounds the distance to every body of the node; the walk accepts
// the node when fl(fl(s^2) theta^2) > b^2, s^2 = fma(x, x, fma(y, y, fma(z, z,
// eps^2))). A node whose centre of mass sits off the box centre counts as
// larger than its half side, in step with the error of its expansion.
__device__ __forceinline__ float corr_size2(float4 com, float4 mn, float4 mx) {
    const float dx = fmaxf(com.x - mn.x, mx.x - com.x);
    const float dy = fmaxf(com.y - mn.y, mx.y - com.y);
    const float dz = fmaxf(com.z - mn.z, mx.z - com.z);
    return __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
}
// The relative opening test: a node's next-order (octupole) error is about
// G M b^3 / s^5; the walk accepts it where that is below alpha |a| (|a| the
// body's acceleration in the previous corrected walk), raised to the power
// 2/5: fl(s^2) A > K with K = (G M b^3)^(2/5) per node and A = (alpha |a|)^(2/5)
// per body, both in FP64 rounded to FP32, together with s^2 > b^2 (the body
// outside the sphere around the centre of mass that holds the node's bodies).
__device__ __forceinline__ float corr_rel_k(float gm, float b2) {
    return (float)(pow((double)gm, 0.4) * pow((double)b2, 0.6));
}
__device__ __forceinline__ float corr_rel_a(float alpha, float amag) {
    return (float)pow((double)alpha * (double)amag, 0.4);
}
__device__ __forceinline__ float corr_dist2(float rx, float ry, float rz, float soft) {
    return __builtin_fmaf(rx, rx, __builtin_fmaf(ry, ry, __builtin_fmaf(rz, rz, soft)));
}


// ---- ensembles: independent simulations concatenated in one set of arrays.
// Simulation s owns sorted positions and body ids [off, off + num) and
// internal-node storage [ibase, ibase + num - 1) with ibase = off - s; node
// ids are global (leaves < N, internal N + storage index).
struct EnsSim {
    uint32_t off, num, ibase, tile0;   // first position, bodies, first internal node, first sort tile
    uint32_t ntiles, ticket, height, pad0;
    float bmin[4], bmax[4];            // bounds (xyz; w unused)
};
static_assert(sizeof(EnsSim) == 64, "EnsSim layout");
struct EnsTile {                       // sort tile: positions [start, start + count) of simulation sim
    uint32_t sim, start, count, pad;
};
struct EnsTask {                       // walk task: up to 64 positions of one simulation
    uint32_t sim, start, count, pad;
};
// Walk work queues, one per XCD: tasks [begin, begin + count) of the task
// list, handed out through head. One 128-byte line each.
constexpr int kEnsQueues = 8;
struct EnsQueue {
    uint32_t begin, count, head, pad[29];
};
static_assert(sizeof(EnsQueue) == 128, "EnsQueue layout");
constexpr uint32_t kEnsTile = 2048;    // positions per sort/bounds/encode tile
static_assert(kEnsTile % kTreeBlock == 0 && kEnsTile % kTreeBlockMoments == 0, "tiles are whole tree blocks");
// first tree launch blocks over ntiles ensemble tiles
inline uint32_t ens_tree_blocks(uint32_t ntiles, bool moments) {
    return ntiles * (kEnsTile / tree_block_size(moments));
}
static_assert(kEnsTile == kSortTile, "ensemble tiles are histogrammed as sort tiles");

void launch_ens_reset(hipStream_t s, DevCounters* ctr, EnsQueue* queues);
void launch_ens_bounds(hipStream_t s, EnsSim* sims, const EnsTile* tiles, uint32_t ntiles, const float4* src,
                       float* partials);
void launch_ens_encode(hipStream_t s, const EnsSim* sims, const EnsTile* tiles, uint32_t ntiles, const float4* src,
                       const uint32_t* index_in, uint64_t* keys);
// one LSD pass of `bits`-bit digits over every simulation's segment
int launch_ens_sort_pass(hipStream_t s, int bits, const EnsSim* sims, uint32_t nsims, const EnsTile* tiles,
                         uint32_t ntiles, const uint64_t* keys_in, const uint32_t* idx_in, uint64_t* keys_out,
                         uint32_t* idx_out, uint32_t shift, uint32_t* hist);
// tree: whole blocks of tree_block_size positions per tile
void launch_ens_tree(hipStream_t s, const EnsSim* sims, const EnsTile* tiles, uint32_t ntiles, uint32_t N,
                     DevCounters* ctr, const float4* src, const uint32_t* index_in, const uint64_t* keys,
                     const TreeArrays& T, float G);
// persistent walk over the task queue; returns nonzero for an unsupported shape
int launch_ens_walk(hipStream_t s, int waves, int depth, uint32_t blocks, const EnsSim* sims, const EnsTask* tasks,
                    EnsQueue* queues, DevCounters* ctr, const float4* src, float4* dst,
                    const uint32_t* index_in, const float4* leaf_pos, const WalkNode* nodes, uint32_t N, float G,
                    float theta, float dt, float soft, uint32_t general, const float4* qrec = nullptr,
                    float* acc_mag = nullptr, uint32_t rel = 0, float alpha = 0.0f);
int ens_walk_occupancy(int waves, int depth, int* blocks_per_cu);
void launch_ens_nodes_std430(hipStream_t s, const TreeArrays& T, uint32_t n_internal, NodeStd430* out);

void launch_init(hipStream_t s, uint32_t* index_in, uint32_t n);
void launch_reset(hipStream_t s, SimValues* sim, DevCounters* ctr, int set_units, float mass, float density,
                  float len, float time, float G, float blc);
// bounds_partials: the census pass also updates the bounds (as launch_bounds)
// step_pos: the census reads the previous walk's positions by sorted
// position instead of the bodies (only when no body changed since that walk)
void launch_dead_partition(hipStream_t s, uint32_t n, SimValues* sim, DevCounters* ctr, float4* src, float4* dst,
                           const uint32_t* index_in, uint32_t* index_out, uint32_t* block_scratch,
                           float* bounds_partials = nullptr, uint32_t partial_cap = 0,
                           const float4* step_pos = nullptr);
void launch_bounds(hipStream_t s, uint32_t n, SimValues* sim, DevCounters* ctr, const float4* src,
                   const uint32_t* index_in, float* partials, uint32_t partial_cap);
// step_pos: the bodies by sorted position as the previous step's walk wrote
// them (read when the partition removed nobody), or nullptr: gathered from src
void launch_encode(hipStream_t s, uint32_t n, const SimValues* sim, const DevCounters* ctr, const float4* src,
                   const uint32_t* index_in, const float4* step_pos, uint64_t* morton_in);
// One LSD pass over (key, value) at positions [0, numBodies); idx_in nullptr:
// the values start as the positions (scrambled: times an odd constant).
// launch order of the one-simulation sort's tiles and the epoch that tells
// this sort's look-back words from an earlier sort's
struct SortState {
    uint32_t epoch;
    uint32_t tile[kSortPasses];
};
uint32_t sort_digit_blocks(uint32_t n);   // histogram blocks (partial count sets of 8 x 256)
// Sorts the keys of keys[in] with the values of vals[in] (positions: the
// values start as the positions 0..num-1 instead) in kSortPasses passes; the
// result lands back in keys[in], vals[in].
void launch_sort(hipStream_t s, uint32_t n, const SimValues* sim, DevCounters* ctr, uint64_t* keys[2],
                 uint32_t* vals[2], int in, bool positions, uint32_t* partials, uint32_t* totals,
                 unsigned long long* status, uint32_t tiles_cap, SortState* st);
// from: when the sort sorted positions instead of ids, the (scrambled)
// position before the sort of the body at each sorted position; index_in is then in the
// order before the sort and the tree writes the ids in the new order to
// ids_out. step_pos (with from): the previous walk's positions by its sorted
// positions, read instead of the bodies when the partition removed nobody.
// fill_cap: 0, or the reference's pass count (kReferenceFillPasses): nodes
// higher than it are left unfilled.
void launch_tree(hipStream_t s, uint32_t n, const SimValues* sim, DevCounters* ctr, const float4* src,
                 const uint32_t* index_in, const uint64_t* keys, const TreeArrays& T, int colors,
                 const float4* step_pos = nullptr, const uint32_t* from = nullptr, uint32_t* ids_out = nullptr,
                 uint32_t fill_cap = 0);
// the reference's tree-fill passes, one height each (gpu/GPU.java:41)
constexpr uint32_t kReferenceFillPasses = 64;
void launch_walk_counters_reset(hipStream_t s, const SimValues* sim, DevCounters* ctr);
void launch_walk(hipStream_t s, uint32_t n, const WalkArgs& a);
// The corrected walk's node records on the current tree, which was built
// with its second moments (T.node_mom): per internal node the second moment
// about its centre of mass plus the network's residual (w, bias: the packed
// network, or nullptr for none), into
// the records qrec [4 corr_records(n)] (the leaves' quadrupole halves must
// hold zeros: they are never written); ctr->corrected counts the network
// evaluations.
void launch_node_quadrupoles(hipStream_t s, uint32_t n, const SimValues* sim, DevCounters* ctr, const TreeArrays& T,
                             const __bf16* w, const float* bias, float4* qrec);
// the packed BF16 weights and biases of a network (include/bh_abi.h's bh_mlp), or nullptr
void mlp_device_weights(const bh_mlp* m, const __bf16** w, const float** bias);
// The same for an ensemble's trees (N bodies, n_internal = N - S internal
// nodes, node ids global).
void launch_ens_node_quadrupoles(hipStream_t s, uint32_t N, uint32_t n_internal, float G, DevCounters* ctr,
                                 const TreeArrays& T, const __bf16* w, const float* bias, float4* qrec);
// count [2]: pairs found, and 1 where a walk ran out of stack (its pairs are incomplete)
void launch_collect_pairs(hipStream_t s, uint32_t n, const SimValues* sim, const float4* leaf_pos,
                          const WalkNode* nodes, const float4* node_aabb, float theta, float soft, uint32_t stride,
                          uint2* pairs, uint32_t cap, uint32_t* count);
void launch_pair_labels(hipStream_t s, const uint2* pairs, uint32_t npairs, const float4* leaf_pos,
                        const WalkNode* nodes, const float4* node_aabb, const uint2* node_range, double G,
                        float soft, uint32_t max_bodies, float* feat, float* targ);
void launch_gather_rows(hipStream_t s, const float* src, const uint32_t* perm, uint32_t offset, uint32_t rows,
                        int cols, float* dst);
// ws (ws_floats floats, or nullptr): workspace for splitting a long K
// (at most kGemmSlices slices of M x N partial sums)
constexpr int kGemmSlices = 64, kGemmSliceK = 256;
void launch_gemm(hipStream_t s, const float* A, int sam, int sak, const float* B, int sbk, int sbn, float* C, int M,
                 int N, int K, int accumulate, float* ws = nullptr, uint32_t ws_floats = 0);
void launch_bias_act(hipStream_t s, float* Z, float* H, const float* bias, int M, int N);
void launch_loss_grad(hipStream_t s, const float* Y, const float* T, float* dY, int M, float* partial,
                      uint32_t* nblocks);
void launch_act_grad(hipStream_t s, const float* Z, float* dH, size_t n);
void launch_colsum(hipStream_t s, const float* dZ, float* db, int M, int N);
void launch_adam(hipStream_t s, float* p, const float* g, float* m1, float* m2, int n, float lr, int t);
void launch_pack_weights(hipStream_t s, const float* p, __bf16* w, float* bias);
void launch_mlp_forward_mfma(hipStream_t s, const float* x, uint32_t n, const __bf16* w, const float* bias, float* y);
void launch_merge(hipStream_t s, float4* dst, const uint2* tasks, uint32_t cap, DevCounters* ctr,
                  unsigned long long* reserve, uint32_t* done);
void launch_nodes_to_std430(hipStream_t s, uint32_t n, const SimValues* sim, const TreeArrays& T, int colors,
                            int which, uint32_t fill_cap, NodeStd430* out);
// Render feed: compaction of the nodes / bodies that pass (counts: the
// per-block counts and offsets, render_feed_blocks(items) + 1 entries; the
// last receives the total). out: device records, at most cap written.
constexpr uint32_t kFeedBlock = 256;
inline uint32_t render_feed_blocks(uint32_t items) { return (items + kFeedBlock - 1) / kFeedBlock; }
void launch_render_nodes(hipStream_t s, uint32_t n, const SimValues* sim, const TreeArrays& T, int colors,
                         uint32_t depth_min, uint32_t depth_max, uint32_t* counts, bh_render_node* out, uint32_t cap);
void launch_render_bodies(hipStream_t s, uint32_t n, const SimValues* sim, const float4* bodies, uint32_t* counts,
                          bh_render_body* out, uint32_t cap);
void launch_std430_to_nodes(hipStream_t s, uint32_t n, const SimValues* sim, const TreeArrays& T,
                            const NodeStd430* leaf_in, const NodeStd430* int_in, int which);
void launch_fill_u32(hipStream_t s, uint32_t* p, uint32_t v, size_t n);
void launch_selftest_math(hipStream_t s, uint32_t* bad);
void launch_direct_sum_f64(hipStream_t s, const float4* bodies, uint32_t n, const uint32_t* sample,
                           uint32_t n_sample, double G, double soft, double* out);

}  // namespace bh
