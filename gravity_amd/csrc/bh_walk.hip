// bh_walk.hip — tree walk, force and integration (shaders/compute/force/bh_force.comp:6-142).
//
// One wavefront walks the tree for 64 consecutive sorted bodies with one
// shared LIFO stack. Each stack entry carries the 64-bit mask of the lanes
// that still need that node; a node is loaded once per wave (its index is
// wave-uniform, so the 32-byte record comes in through a scalar load), every
// lane in the mask evaluates its own acceptance test, and the children are
// pushed with the mask of the lanes that opened it. Restricted to one lane,
// this visits exactly the nodes of that body's own per-thread walk in exactly
// the reference's order (push childA, push childB, pop childB first: here
// childA is pushed and childB visited directly), so the accepted set, the
// opened-node count and the FP32 sums are the reference's. Two loops run it:
// walk_lean (stack in lanes of three registers, for the common case) and
// walk_chunk (stack in LDS, any tree, any distance).
#include "bh_device.hpp"
#include "bh_kernels.hpp"

namespace bh {

__device__ __forceinline__ float body_radius(float m, float rho, float blc) {
    // common/common.glsl:228-230
    return fmul(blc, powf(fdiv(m, rho), 1.0f / 3.0f));
}

// Radius of the body at every sorted position, for the overlap tests of the
// merging and collision modes: the walk reads it per leaf instead of
// evaluating pow for every (lane, leaf) visit. Same function, same inputs,
// same bits as the per-visit evaluation.
__global__ __launch_bounds__(256) void k_leaf_radius(const SimValues* __restrict__ sim,
                                                     const float4* __restrict__ src,
                                                     const uint32_t* __restrict__ index_in,
                                                     const float4* __restrict__ leaf_pos,
                                                     float* __restrict__ leaf_radius) {
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= sim->numBodies) return;
    leaf_radius[g] = body_radius(leaf_pos[g].w, src[2 * index_in[g] + 1].w, sim->bodyLengthConstant);
}

// Accumulated results of one wave's walk.
struct WalkAcc {
    float ax, ay, az;
    unsigned long long opened, inter;
    uint32_t maxsp, err;
    unsigned long long quads;   // quadrupole terms evaluated (corrected walk)
};

// a += G m r inv^3 in the reference's operation order (bh_force.comp:22-25,74,82)
__device__ __forceinline__ void add_force(float& ax, float& ay, float& az, float gm, float rx, float ry, float rz,
                                          float inv) {
    ax = fadd(ax, fmul(fmul(fmul(fmul(gm, rx), inv), inv), inv));
    ay = fadd(ay, fmul(fmul(fmul(fmul(gm, ry), inv), inv), inv));
    az = fadd(az, fmul(fmul(fmul(fmul(gm, rz), inv), inv), inv));
}

// The overlap test of a leaf visit in the merging (1) and collision (2)
// modes (bh_force.comp:85-112); use: this lane interacts with the leaf.
template <uint32_t mode>
__device__ __forceinline__ void leaf_contact(bool use, uint32_t node, float rx, float ry, float rz,
                                             float4& pm, float4& vd, float Ri, const float* __restrict__ leaf_radius,
                                             uint32_t bodyIdx, float elasticity, float restitution,
                                             const uint32_t* __restrict__ index_in, const float4* __restrict__ src,
                                             DevCounters* __restrict__ ctr, uint2* __restrict__ tasks,
                                             uint32_t task_cap, uint32_t& err) {
    const float Rj = leaf_radius[node];
    const float dist = fsqrt(dot3(rx, ry, rz));
    const float pen = fsub(fadd(Ri, Rj), dist);
    const bool hit = use && pen > 0.0f;
    if (ballot(hit)) {
        const uint32_t other = uniform(index_in[node]);
        if constexpr (mode == 1u) {
            // merge task from the lower-index body
            if (hit && bodyIdx < other) {
                uint32_t slot = atomicAdd(&ctr->nTasks, 1u);
                if (slot < task_cap) tasks[slot] = make_uint2(bodyIdx, other);
                else err |= 2u;
            }
        } else if (hit) {
            // collision (bh_force.comp:93-107): normalize(r) = r * inversesqrt(r.r)
            const float4 ov = src[2 * other + 1];
            const float mj = src[2 * other].w;                    // the leaf's mass
            const float il = fdiv(1.0f, fsqrt(dot3(rx, ry, rz)));
            const float nx = fmul(rx, il), ny = fmul(ry, il), nz = fmul(rz, il);
            const float vimp = dot3(fsub(ov.x, vd.x), fsub(ov.y, vd.y), fsub(ov.z, vd.z), nx, ny, nz);
            if (vimp < 0.0f) {
                const float meff = fdiv(1.0f, fadd(fdiv(1.0f, pm.w), fdiv(1.0f, mj)));
                const float imp = fmul(fmul(fadd(1.0f, elasticity), meff), vimp);
                vd.x = fadd(vd.x, fdiv(fmul(nx, imp), pm.w));
                vd.y = fadd(vd.y, fdiv(fmul(ny, imp), pm.w));
                vd.z = fadd(vd.z, fdiv(fmul(nz, imp), pm.w));
            }
            const float cs = fmul(fdiv(pen, fadd(pm.w, mj)), restitution);
            pm.x = fsub(pm.x, fmul(cs, nx));
            pm.y = fsub(pm.y, fmul(cs, ny));
            pm.z = fsub(pm.z, fmul(cs, nz));
        }
    }
}

// The masked-stack walk of one wave over the tree rooted at `root`: lanes in
// `mask` walk for the bodies at sorted positions base + lane (position g,
// position-mass pm, velocity-density vd); leaf ids are below n_ids, internal
// node id k >= n_ids is stored at nodes[k - n_ids]. s_node/s_mask are this
// wave's LDS stack of kDepth entries. mode (mergingCollisionOrNeither):
// 1 = leaf visits emit merge tasks, 2 = an overlapping leaf changes this
// lane's body at once, velocity impulse and position correction, and the rest
// of the walk and the integration use the changed body (bh_force.comp:85-112).
template <int kDepth, uint32_t mode, bool kFast>
__device__ __forceinline__ void walk_chunk(uint32_t* __restrict__ s_node, unsigned long long* __restrict__ s_mask,
                                           uint32_t root, uint32_t n_ids, uint32_t base, uint32_t g,
                                           unsigned long long mask, float4& pm, float4& vd,
                                           const float4* __restrict__ leaf_pos, const WalkNode* __restrict__ nodes,
                                           float G, float theta, float soft, float Ri,
                                           const float* __restrict__ leaf_radius, uint32_t bodyIdx,
                                           float elasticity, float restitution,
                                           const uint32_t* __restrict__ index_in, const float4* __restrict__ src,
                                           DevCounters* __restrict__ ctr, uint2* __restrict__ tasks,
                                           uint32_t task_cap, WalkAcc& o) {
    const uint32_t lane = lane_id();
    // the sums and counters live in locals for the loop (a struct carried
    // around the loop costs register copies on every iteration)
    float ax = o.ax, ay = o.ay, az = o.az;
    unsigned long long opened = o.opened, inter = o.inter;
    uint32_t maxsp = o.maxsp, err = o.err;
    // current node in scalar registers; the stack holds the pending childA
    // entries (after opening a node its childB, the top of the reference's
    // stack, is visited next without a round trip through LDS)
    uint32_t node = root;
    int sp = 0;
    const float theta_hi = theta * 1.0000153f;   // theta (1 + 2^-16), rounded up
    const uint32_t wbase = __builtin_amdgcn_readfirstlane(base);
    // Lane predicates are taken from 64-bit wave masks (inverse ballot: a
    // lane's bit of a scalar mask) so that they cost no vector instructions.
    for (;;) {
        // one body-node interaction per lane, for an internal node (its centre
        // of mass, if the lane accepts it) or a leaf (its body, unless it is
        // the lane's own); r, d2 and the sum are shared by both kinds
        const bool internal = node >= n_ids;
        float4 sp4;                 // position and mass of the source
        float half = 0.0f;
        uint32_t childA = 0, childB = 0;
        if (internal) {
            const WalkNode nd = nodes[node - n_ids];
            sp4 = nd.comMass;
            half = nd.half;
            childA = nd.childA;
            childB = nd.childB;
        } else {
            sp4 = leaf_pos[node];
        }
        const float rx = fsub(sp4.x, pm.x);
        const float ry = fsub(sp4.y, pm.y);
        const float rz = fsub(sp4.z, pm.z);
        const float d2 = fadd(dot3(rx, ry, rz), soft);
        unsigned long long open = 0ull, take;
        // bh_force.comp:77-79 (self skip): the lane whose own leaf this is
        const uint32_t rel = node - wbase;
        const unsigned long long use_m = mask & ~(rel < (uint32_t)kWave ? (1ull << rel) : 0ull);
        // The hardware 1/sqrt (within 2 ulp) decides a node that every lane of
        // the mask opens by a wide margin without the correctly rounded
        // inverse square root: half * rsq >= theta (1 + 2^-16) implies the
        // exact ratio is >= theta. Only the decision is taken from it; every
        // value that is accumulated uses the exact one.
        float inv;
        if (internal) {
            if ((ballot(!(fmul(half, __builtin_amdgcn_rsqf(d2)) >= theta_hi)) & mask) == 0ull) {
                open = mask;
                take = 0ull;
            } else {
                inv = kFast ? inv_sqrt_fast(d2) : inv_sqrt_exact(d2);
                take = ballot(fmul(half, inv) < theta) & mask;
                open = mask & ~take;
            }
        } else {
            inv = kFast ? inv_sqrt_fast(d2) : inv_sqrt_exact(d2);
            take = use_m;
        }
        // an internal record holds G m; a leaf's position-mass holds m
        if (__builtin_amdgcn_inverse_ballot_w64(take))
            add_force(ax, ay, az, internal ? sp4.w : fmul(G, sp4.w), rx, ry, rz, inv);
        inter += (unsigned long long)__popcll(take);
        if constexpr (mode != 0u)
            if (!internal && use_m)
                leaf_contact<mode>(__builtin_amdgcn_inverse_ballot_w64(use_m), node, rx, ry, rz, pm, vd, Ri,
                                   leaf_radius, bodyIdx, elasticity, restitution, index_in, src, ctr, tasks, task_cap,
                                   err);
        bool descend = false;
        if (open) {
            if (sp + 1 > kDepth) {
                err |= 1u;
                break;
            }
            if (lane == 0) {
                s_node[sp] = childA;
                s_mask[sp] = open;
            }
            sp += 1;
            opened += (unsigned long long)__popcll(open);
            if ((uint32_t)sp + 1 > maxsp) maxsp = (uint32_t)sp + 1;
            node = uniform(childB);
            mask = open;
            descend = true;
        }
        if (!descend) {
            if (sp == 0) break;
            sp--;
            node = uniform(s_node[sp]);
            const unsigned long long mraw = s_mask[sp];
            mask = ((unsigned long long)uniform((uint32_t)(mraw >> 32)) << 32) | uniform((uint32_t)mraw);
        }
    }
    o.ax = ax;
    o.ay = ay;
    o.az = az;
    o.opened = opened;
    o.inter = inter;
    o.maxsp = maxsp;
    o.err = err;
}

// The same walk for the common case, with less scalar bookkeeping per visit
// (the walk issues about as many scalar as vector instructions, and each
// SIMD issues at most one of each per four cycles): every d2 is in the fast
// domain of the inverse square root, the tree is at most kLeanHeight high so
// that the stack (never deeper than the tree) fits one entry per lane of
// three registers (node, mask low, mask high: v_writelane / v_readlane
// instead of LDS), nothing can overflow it, and the per-wave counts fit
// 32 bits (at most 2^26 bodies). Same visits, same order, same sums.
constexpr uint32_t kLeanHeight = 62;
typedef uint32_t u32x8 __attribute__((ext_vector_type(8)));
// a corrected-walk record (64 bytes, 64-byte aligned)
typedef uint32_t u32x16 __attribute__((ext_vector_type(16)));
// v_writelane_b32 (the LLVM intrinsic; this compiler has no builtin for it)
__device__ int lane_write(int val, int lane, int old) __asm("llvm.amdgcn.writelane.i32");
constexpr uint32_t kLeanBodies = 1u << 26;

__device__ __forceinline__ bool walk_lean_ok(float root_half, float soft, uint32_t height, uint32_t num) {
    return inv_sqrt_fast_domain(root_half, soft) && height <= kLeanHeight && num <= kLeanBodies;
}

template <uint32_t mode>
__device__ __forceinline__ void walk_lean(uint32_t root, uint32_t n_ids, uint32_t g, unsigned long long mask,
                                          float4& pm, float4& vd, const WalkNode* __restrict__ nodes, float G,
                                          float theta, float soft, float Ri, const float* __restrict__ leaf_radius,
                                          uint32_t bodyIdx, float elasticity, float restitution,
                                          const uint32_t* __restrict__ index_in, const float4* __restrict__ src,
                                          DevCounters* __restrict__ ctr, uint2* __restrict__ tasks,
                                          uint32_t task_cap, WalkAcc& o) {
    float ax = o.ax, ay = o.ay, az = o.az;
    uint32_t opened = 0, inter = 0, maxsp = o.maxsp, err = o.err;
    // One record per node id, leaves included (the tree arrays keep the leaf
    // records right before the internal ones), addressed as a base plus a
    // 32-bit byte offset (ids < 2^27).
    const char* __restrict__ wbase = reinterpret_cast<const char*>(nodes - n_ids);
    int st_node = 0, st_lo = 0, st_hi = 0;       // stack entry i in lane i
    uint32_t node = root;
    int sp = 0;
    // Every lane of the mask opens a node by a wide margin if d2 theta^2
    // (1 + 2^-10) <= half^2: then half / sqrt(d2) >= theta (1 + 2^-11), and
    // the correctly rounded half * inv is >= theta (a few ulps of error).
    // A leaf's half2 of -1 never passes.
    const float open_k = fmul(fmul(theta, theta), 1.0009766f);
    for (;;) {
        // one 32-byte scalar load (a vector type keeps the compiler from
        // splitting it and sinking the half side into a later dependent load)
        const u32x8 rec = *reinterpret_cast<const u32x8*>(wbase + (uint32_t)(node << 5));
        WalkNode nd;
        nd.comMass = make_float4(__uint_as_float(rec[0]), __uint_as_float(rec[1]), __uint_as_float(rec[2]),
                                 __uint_as_float(rec[3]));
        nd.half = __uint_as_float(rec[4]);
        nd.childA = rec[5];
        nd.childB = rec[6];
        nd.half2 = __uint_as_float(rec[7]);
        const float rx = fsub(nd.comMass.x, pm.x);
        const float ry = fsub(nd.comMass.y, pm.y);
        const float rz = fsub(nd.comMass.z, pm.z);
        const float d2 = fadd(dot3(rx, ry, rz), soft);
        unsigned long long open = mask;
        if (ballot(!(fmul(d2, open_k) <= nd.half2)) & mask) {
            // a node some lane may accept, or a leaf (every lane takes it but
            // the one whose own body it is: bh_force.comp:77-79)
            const bool internal = node >= n_ids;
            const float inv = inv_sqrt_fast(d2);
            const unsigned long long accm = internal ? ballot(fmul(nd.half, inv) < theta) : ~0ull;
            const unsigned long long take = accm & mask & ballot(node != g);
            open = mask & ~accm;
            if (__builtin_amdgcn_inverse_ballot_w64(take))
                add_force(ax, ay, az, nd.comMass.w, rx, ry, rz, inv);   // the record holds G m
            inter += (uint32_t)__popcll(take);
            if constexpr (mode != 0u)
                if (!internal && take)
                    leaf_contact<mode>(__builtin_amdgcn_inverse_ballot_w64(take), node, rx, ry, rz, pm,
                                       vd, Ri, leaf_radius, bodyIdx, elasticity, restitution, index_in, src, ctr,
                                       tasks, task_cap, err);
        }
        // Push (childA, open) and go to childB if any lane opened the node,
        // else pop; both without a branch. The top entry is read first (lane
        // -1 reads as lane 63 when the stack is empty: not used), and the new
        // entry goes to lane sp either way (above the top, it is never read).
        const int top = sp - 1;
        const uint32_t pop_node = (uint32_t)__builtin_amdgcn_readlane(st_node, top);
        const unsigned long long pop_mask =
            ((unsigned long long)(uint32_t)__builtin_amdgcn_readlane(st_hi, top) << 32) |
            (uint32_t)__builtin_amdgcn_readlane(st_lo, top);
        st_node = lane_write((int)nd.childA, sp, st_node);
        st_lo = lane_write((int)(uint32_t)open, sp, st_lo);
        st_hi = lane_write((int)(uint32_t)(open >> 32), sp, st_hi);
        opened += (uint32_t)__popcll(open);
        const bool descend = open != 0ull;
        node = descend ? nd.childB : pop_node;
        mask = descend ? open : pop_mask;
        sp = descend ? sp + 1 : top;
        maxsp = (uint32_t)sp + 1 > maxsp ? (uint32_t)sp + 1 : maxsp;   // a pop never raises it
        if (sp < 0) break;
    }
    o.ax = ax;
    o.ay = ay;
    o.az = az;
    o.opened += opened;
    o.inter += inter;
    o.maxsp = maxsp;
    o.err = err;
}

// The corrected walk (no reference counterpart). Every node a body takes adds
// its softened monopole and quadrupole,
//     a += G M r / s^3 + (7.5 (r.GT.r) / s^2 - 1.5 G tr T) r / s^5 - 3 GT r / s^5,
// s^2 = |r|^2 + eps^2 (a leaf: its monopole), from one 64-byte record per node
// id (corr_records in bh_kernels.hpp: leaves have zero quadrupole terms and
// b^2 = K = -1). It needs no bit parity with the reference, so it takes the
// shortest FP32 sequence: s^2 by fused multiply-adds, opening tests without a
// root (exactly reproducible on the host), the hardware 1/sqrt (within 1 ulp)
// and one fused update of both terms.
//
// Decisions are wave-wide: the 64 bodies of a wave take a node together when
// every lane's test accepts it (the geometric fl(s^2) theta^2 > b^2, or the
// relative s^2 > b^2 and fl(s^2) A > K), and open it together otherwise. The
// lane that needs the finer nodes decides for its neighbours, which then add
// more, smaller terms (never less accurate than their own test asks); the
// mask never changes, so the stack holds node ids only and every internal
// node taken adds a term in all lanes. The lane whose own leaf a record is
// skips it (r = 0 would add exactly 0; it does not count as an interaction).
// kLds: stack in LDS (any tree height, overflow checked, 64-bit record
// offsets) instead of two lane registers (trees at most kLeanHeight high,
// offsets below 2^32).
// Needs eps^2 >= 1e-14 and 16 h_root^2 + eps^2 <= 1e15 (corr_range_ok).
__device__ __forceinline__ bool walk_corrected_ok(float root_half, float soft) {
    return corr_range_ok(root_half, soft);
}

// The traversal evaluates the two children of an opened node together: the
// stack holds sibling pairs (one children word each, corr_children_word),
// an iteration pops one pair, loads both records at once (one latency for two
// nodes) and pushes the child pair of each of the two that is opened (b's
// last: its subtree comes next). Which nodes are taken does not depend on the
// order. Counting: the root, then two nodes per pair; every opened node
// pushes one pair.
// root_k: the root's internal storage index (0 for one simulation; an
// ensemble's simulations have their own); ioff: the children word's offset
// (corr_child_b); kCap: entries of the LDS stack.
template <bool kLds, bool kRel, int kCap = kStackDepth>
__device__ __forceinline__ void walk_corrected_pairs(uint32_t n_ids, uint32_t root_k, uint32_t ioff, uint32_t base,
                                                     unsigned long long mask, const float4& pm,
                                                     const float4* __restrict__ crec, float afac, float soft,
                                                     unsigned long long* __restrict__ s_pair, WalkAcc& o) {
    typedef typename std::conditional<kLds, unsigned long long, uint32_t>::type Count;
    float ax = o.ax, ay = o.ay, az = o.az;
    Count pairs = 0, leafvis = 0, leafint = 0;
    uint32_t err = o.err;
    int maxsp = 0;
    const char* __restrict__ cbase = reinterpret_cast<const char*>(crec);
    auto rec = [&](uint32_t id) {
        return kLds ? *reinterpret_cast<const u32x16*>(cbase + ((size_t)id << 6))
                    : *reinterpret_cast<const u32x16*>(cbase + (uint32_t)(id << 6));
    };
    const uint32_t wbase = __builtin_amdgcn_readfirstlane(base);
    // The node's terms (record r, offset r, s^2 d2) added to this lane's sum,
    // from the record's S = -3 G T:
    //     a += (W r + S r) / s^5,  W = G M s^2 - 2.5 (r.S.r) / s^2 - 1.5 G tr T
    // (25 operations; within the range of walk_corrected_ok 1/s^5 is normal).
    auto terms = [&](const u32x16& r, float rx, float ry, float rz, float d2) {
        const float inv = __builtin_amdgcn_rsqf(d2);
        const float i2 = inv * inv;
        const float i5 = (i2 * i2) * inv;
        const float h = -2.5f * i2;
        const float q0 = __uint_as_float(r[8]), q1 = __uint_as_float(r[9]), q2 = __uint_as_float(r[10]);
        const float q3 = __uint_as_float(r[11]), q4 = __uint_as_float(r[12]), q5 = __uint_as_float(r[13]);
        const float tx = __builtin_fmaf(q0, rx, __builtin_fmaf(q1, ry, q2 * rz));
        const float ty = __builtin_fmaf(q1, rx, __builtin_fmaf(q3, ry, q4 * rz));
        const float tz = __builtin_fmaf(q2, rx, __builtin_fmaf(q4, ry, q5 * rz));
        const float rsr = __builtin_fmaf(rx, tx, __builtin_fmaf(ry, ty, rz * tz));
        const float W = __builtin_fmaf(__uint_as_float(r[3]), d2, __builtin_fmaf(rsr, h, __uint_as_float(r[14])));
        ax = __builtin_fmaf(i5, __builtin_fmaf(W, rx, tx), ax);
        ay = __builtin_fmaf(i5, __builtin_fmaf(W, ry, ty), ay);
        az = __builtin_fmaf(i5, __builtin_fmaf(W, rz, tz), az);
    };
    // the lanes of the mask that fail the opening test for record r
    auto fails = [&](const u32x16& r, float d2) -> unsigned long long {
        const float b2 = __uint_as_float(r[7]);
        const unsigned long long f =
            kRel ? ballot(!(d2 > b2)) | ballot(!(d2 * afac > __uint_as_float(r[4]))) : ballot(!(d2 * afac > b2));
        return f & mask;
    };
    // One node for the whole wave: 1 if some lane of the mask fails the
    // opening test (the wave opens the node), else 0 and the node's terms are
    // added in every lane of the wave, under a scalar branch (a lane outside
    // the mask holds no body: its sum is never read). A leaf (b^2 = K = -1)
    // passes every test; the lane whose own leaf it is adds exactly zero
    // (r = 0, no quadrupole terms, s^2 = eps^2 > 0). The decision is kept an
    // integer (a count of 0 .. 64 bits is nonzero when (count + 63) >> 6 is
    // 1): the compiler keeps a condition carried on as a lane mask and turns
    // it back into an integer with vector instructions.
    auto visit = [&](const u32x16& r) -> uint32_t {
        const float rx = __uint_as_float(r[0]) - pm.x;
        const float ry = __uint_as_float(r[1]) - pm.y;
        const float rz = __uint_as_float(r[2]) - pm.z;
        const float d2 = corr_dist2(rx, ry, rz, soft);
        const uint32_t opn = ((uint32_t)__popcll(fails(r, d2)) + 63u) >> 6;
        if (opn == 0u) terms(r, rx, ry, rz, d2);
        return opn;
    };
    // The interactions of leaf id x with the mask's lanes: all but the lane
    // whose own body it is (bh_force.comp:77-79).
    const uint32_t mask_lanes = (uint32_t)__popcll(mask);
    auto leaf_interactions = [&](uint32_t x) -> uint32_t {
        const uint32_t rel = x - wbase;
        const unsigned long long own = rel < (uint32_t)kWave ? (1ull << rel) & mask : 0ull;
        return mask_lanes - (uint32_t)__popcll(own);
    };
    // the root; then pairs until the stack is empty
    const uint32_t root = corr_internal_id(n_ids, root_k);
    const u32x16 rr = rec(root);
    if (visit(rr)) {
        // lane stack: pair i in lane i of st; every value it ever holds is a
        // children word, so a read above the top names records in the array
        int st = 0;
        int sp;
        if constexpr (kLds) {
            if (lane_id() == 0) s_pair[0] = rr[5];
        } else {
            st = lane_write((int)rr[5], 0, st);
        }
        sp = 1;
        for (;;) {
            uint32_t w;
            if constexpr (kLds)
                w = uniform((uint32_t)s_pair[sp - 1]);
            else
                w = (uint32_t)__builtin_amdgcn_readlane(st, sp - 1);
            const uint32_t a = corr_child_a(w), b = corr_child_b(w, ioff);
            sp -= 1;
            u32x16 ra = rec(a);
            u32x16 rb = rec(b);
            // both records before either is used: the two loads are issued
            // back to back and waited for once (every use of either record
            // comes after this statement, which needs both)
            asm volatile("" : "+s"(ra), "+s"(rb));
            // a leaf is never opened (the internal bit of its side is 0)
            const uint32_t open_a = visit(ra) & ((w >> 30) & 1u);
            const uint32_t open_b = visit(rb) & (w >> 31);
            if (w < 0xC0000000u) {   // a leaf in the pair (rare): its counts
                if (!((w >> 30) & 1u)) {
                    leafvis += 1;
                    leafint += (Count)leaf_interactions(a);
                }
                if (!(w >> 31)) {
                    leafvis += 1;
                    leafint += (Count)leaf_interactions(b);
                }
            }
            pairs += 1;
            // push a's children if a was opened, then b's
            if constexpr (kLds) {
                if (sp + (int)(open_a + open_b) > kCap) {
                    err |= 1u;
                    break;
                }
                if (lane_id() == 0) {
                    if (open_a) s_pair[sp] = ra[5];
                    if (open_b) s_pair[sp + open_a] = rb[5];
                }
            } else {
                // without a branch: a's pair goes to sp, b's to sp or sp + 1
                // (a write above the new top is never read)
                st = lane_write((int)ra[5], sp, st);
                st = lane_write((int)rb[5], sp + (int)open_a, st);
            }
            sp += (int)(open_a + open_b);
            maxsp = sp > maxsp ? sp : maxsp;
            if (sp == 0) break;
        }
    }
    const Count lanes = (Count)__popcll(mask);
    const Count evaluated = 1 + 2 * pairs;
    const Count opens = pairs;                 // (each opened node pushed one pair)
    const Count takes = evaluated - opens - leafvis;
    o.ax = ax;
    o.ay = ay;
    o.az = az;
    o.quads += (unsigned long long)takes * lanes;
    o.inter += (unsigned long long)takes * lanes + leafint;
    o.opened += (unsigned long long)opens * lanes;
    o.maxsp = (uint32_t)maxsp + 1 > o.maxsp ? (uint32_t)maxsp + 1 : o.maxsp;
    o.err = err;
}

// The corrected walk's inputs for the walk kernels besides its records (a
// __restrict__ parameter of their own, so that they come in through scalar
// loads): the |a| of the previous corrected walk by body id (read by the
// relative test, rewritten by every corrected walk) and the relative test's
// alpha.
struct CorrArgs {
    float* acc_mag;
    float alpha;
};

// Totals a wave accumulates over its 64-body groups, added to the counters once.
struct WalkTotals {
    unsigned long long opened, inter;
    uint32_t maxsp, err, oob;
    unsigned long long quads;
};

// Walk and integrate the kBodies sorted bodies [base, base + kBodies) of one
// wave (lanes kBodies.. idle).
template <uint32_t kMode, int kCorr, int kBodies = kWave>
__device__ __forceinline__ void walk_group(uint32_t base, uint32_t* s_node, unsigned long long* s_mask,
                                           SimValues* __restrict__ sim, DevCounters* __restrict__ ctr,
                                           const float4* __restrict__ src, float4* __restrict__ dst,
                                           const uint32_t* __restrict__ index_in,
                                           const float4* __restrict__ leaf_pos,
                                           const WalkNode* __restrict__ nodes, const float4* __restrict__ qrec,
                                           const CorrArgs corr, uint2* __restrict__ tasks,
                                           uint32_t task_cap, float theta_, float dt_, float softening,
                                           uint32_t dynamic, uint32_t wrap, float elasticity, float restitution,
                                           const float* __restrict__ leaf_radius, uint32_t general,
                                           float4* __restrict__ pos_out, WalkTotals& tot) {
    const uint32_t n = sim->initialNumBodies;
    const uint32_t num = sim->numBodies;
    const uint32_t lane = lane_id();
    const uint32_t g = base + lane;
    const bool active = (kBodies == kWave || lane < (uint32_t)kBodies) && g < num;
    const uint32_t bodyIdx = active ? index_in[g] : 0;
    float4 pm = active ? leaf_pos[g] : make_float4(0.f, 0.f, 0.f, 0.f);
    float4 vd = active ? src[2 * bodyIdx + 1] : make_float4(0.f, 0.f, 0.f, 1.f);
    const float G = sim->gravitationalConstant;
    const float Ri = (kMode != 0u && active) ? leaf_radius[g] : 0.0f;

    WalkAcc o = {0.0f, 0.0f, 0.0f, 0ull, 0ull, 1u, 0u, 0ull};
    if (num > 1) {
        if constexpr (kCorr != 0) {
            // kCorr 1: geometric test (theta^2), 2: relative (A of the body)
            const float afac =
                kCorr == 2 ? corr_rel_a(corr.alpha, active ? corr.acc_mag[bodyIdx] : 0.0f) : theta_ * theta_;
            if (!walk_corrected_ok(nodes[0].half, softening))
                o.err |= 4u;   // outside the corrected walk's range: no forces
            else if (!general && ctr->treeHeight <= kLeanHeight && corr_offsets_32bit(n))
                walk_corrected_pairs<false, kCorr == 2>(n, 0u, n + kCorrGap, base, ballot(active), pm, qrec, afac, softening,
                                                        s_mask, o);
            else
                walk_corrected_pairs<true, kCorr == 2>(n, 0u, n + kCorrGap, base, ballot(active), pm, qrec, afac, softening,
                                                       s_mask, o);
        } else if (!general && walk_lean_ok(nodes[0].half, softening, ctr->treeHeight, num)) {
            walk_lean<kMode>(n, n, g, ballot(active), pm, vd, nodes, G, theta_, softening, Ri, leaf_radius, bodyIdx,
                             elasticity, restitution, index_in, src, ctr, tasks, task_cap, o);
        } else if (inv_sqrt_fast_domain(nodes[0].half, softening)) {
            walk_chunk<kStackDepth, kMode, true>(s_node, s_mask, n, n, base, g, ballot(active), pm, vd, leaf_pos, nodes,
                                                 G, theta_, softening, Ri, leaf_radius, bodyIdx, elasticity,
                                                 restitution, index_in, src, ctr, tasks, task_cap, o);
        } else {
            walk_chunk<kStackDepth, kMode, false>(s_node, s_mask, n, n, base, g, ballot(active), pm, vd, leaf_pos,
                                                  nodes, G, theta_, softening, Ri, leaf_radius, bodyIdx, elasticity,
                                                  restitution, index_in, src, ctr, tasks, task_cap, o);
        }
    }
    const float ax = o.ax, ay = o.ay, az = o.az;

    // kick-drift into the output buffer (bh_force.comp:118-140)
    uint32_t oob = 0;
    if (active) {
        const float dt = dt_;
        const float nvx = fadd(vd.x, fmul(ax, dt));
        const float nvy = fadd(vd.y, fmul(ay, dt));
        const float nvz = fadd(vd.z, fmul(az, dt));
        float npx = fadd(pm.x, fmul(nvx, dt));
        float npy = fadd(pm.y, fmul(nvy, dt));
        float npz = fadd(pm.z, fmul(nvz, dt));
        bool kill = false;
        if (dynamic == 0u) {
            const float mn0 = sim->minCorner[0], mn1 = sim->minCorner[1], mn2 = sim->minCorner[2];
            const float mx0 = sim->maxCorner[0], mx1 = sim->maxCorner[1], mx2 = sim->maxCorner[2];
            if (wrap) {
                float x, y;
                x = fsub(npx, mn0); y = fsub(mx0, mn0); npx = fadd(fsub(x, fmul(y, floorf(fdiv(x, y)))), mn0);
                x = fsub(npy, mn1); y = fsub(mx1, mn1); npy = fadd(fsub(x, fmul(y, floorf(fdiv(x, y)))), mn1);
                x = fsub(npz, mn2); y = fsub(mx2, mn2); npz = fadd(fsub(x, fmul(y, floorf(fdiv(x, y)))), mn2);
            } else if (pm.x < mn0 || pm.x > mx0 || pm.y < mn1 || pm.y > mx1 || pm.z < mn2 || pm.z > mx2) {
                kill = true;
            }
        }
        const float4 np = kill ? make_float4(0.f, 0.f, 0.f, 0.f) : make_float4(npx, npy, npz, pm.w);
        const float4 nv = kill ? make_float4(0.f, 0.f, 0.f, 0.f) : make_float4(nvx, nvy, nvz, vd.w);
        dst[2 * bodyIdx] = np;
        dst[2 * bodyIdx + 1] = nv;
        if (pos_out) pos_out[g] = np;   // the same value by sorted position
        if constexpr (kCorr != 0)       // |a| for the next relative test (correctly rounded root)
            corr.acc_mag[bodyIdx] = fsqrt(__builtin_fmaf(ax, ax, __builtin_fmaf(ay, ay, az * az)));
        oob = kill ? 1u : 0u;
    }
    tot.opened += o.opened;
    tot.inter += o.inter;
    tot.quads += o.quads;
    tot.maxsp = o.maxsp > tot.maxsp ? o.maxsp : tot.maxsp;
    tot.err |= o.err;
    tot.oob += __popcll(ballot(oob != 0));
}

// one set of atomics per wave
__device__ __forceinline__ void walk_totals_flush(SimValues* sim, DevCounters* ctr, const WalkTotals& t) {
    const uint32_t oob_w = t.oob;   // wave-uniform
    const uint32_t err_w =
        (ballot(t.err & 1u) ? 1u : 0u) | (ballot(t.err & 2u) ? 2u : 0u) | (ballot(t.err & 4u) ? 4u : 0u);
    if (lane_id() == 0) {
        atomicAdd(&ctr->opened, t.opened);
        atomicAdd(&ctr->interactions, t.inter);
        if (t.quads) atomicAdd(&ctr->quadTerms, t.quads);
        atomicMax(&ctr->maxStack, t.maxsp);
        if (oob_w) atomicAdd(&sim->outOfBounds, oob_w);
        if (err_w) atomicOr(&ctr->errorFlags, err_w);
        // sim.uintDebug[0] counts opened nodes as a 32-bit word (bh_force.comp:72)
        atomicAdd(&sim->uintDebug[0], (uint32_t)t.opened);
    }
}

// Pointers are separate __restrict__ parameters so that the compiler can
// prove the tree is not written during the kernel and fetch node records with
// scalar loads.
template <uint32_t kMode, int kW, int kCorr, int kBodies>
__global__ __launch_bounds__(kW * 64) void k_walk(SimValues* __restrict__ sim, DevCounters* __restrict__ ctr,
                                              const float4* __restrict__ src, float4* __restrict__ dst,
                                              const uint32_t* __restrict__ index_in,
                                              const float4* __restrict__ leaf_pos,
                                              const WalkNode* __restrict__ nodes, const float4* __restrict__ qrec,
                                              const CorrArgs corr, uint2* __restrict__ tasks,
                                              uint32_t task_cap, float theta_, float dt_, float softening,
                                              uint32_t dynamic, uint32_t wrap, float elasticity, float restitution,
                                              const float* __restrict__ leaf_radius, uint32_t general,
                                              float4* __restrict__ pos_out) {
    __shared__ uint32_t s_node[kW][kStackDepth];
    __shared__ unsigned long long s_mask[kW][kStackDepth];
    const uint32_t w = threadIdx.x >> 6;
    const uint32_t lb = xcd_block(blockIdx.x, gridDim.x);
    const uint32_t base = (lb * kW + w) * kBodies;
    if (base >= sim->numBodies) return;
    WalkTotals t = {0ull, 0ull, 1u, 0u, 0u, 0ull};
    walk_group<kMode, kCorr, kBodies>(base, s_node[w], s_mask[w], sim, ctr, src, dst, index_in, leaf_pos, nodes, qrec, corr, tasks,
                             task_cap, theta_, dt_, softening, dynamic, wrap, elasticity, restitution, leaf_radius,
                             general, pos_out, t);
    walk_totals_flush(sim, ctr, t);
}

// The CU slot (XCD x 36 + shader engine x 9 + CU id) the calling wave runs on,
// read from the HW_ID and XCC_ID hardware registers.
__device__ __forceinline__ uint32_t walk_home_queue() {
    const uint32_t hw = __builtin_amdgcn_s_getreg((31 << 11) | 4);          // HW_ID, 32 bits
    const uint32_t xcc = __builtin_amdgcn_s_getreg((15 << 11) | 20) & 7u;   // XCC_ID
    const uint32_t f = (hw >> 8) & 0xFFu;   // CU_ID [3:0], SH_ID [4], SE_ID [7:5]
    const uint32_t slot = (((f >> 5) & 7u) * 9u + (f & 15u)) % kWalkSlots;
    return xcc * kWalkSlots + slot;
}

// 64-body groups [begin, end) of queue q: the sorted bodies split into
// kWalkQueues consecutive ranges, XCD x owning queues 36x .. 36x + 35.
__device__ __forceinline__ uint32_t walk_queue_begin(uint32_t groups, uint32_t q) {
    return (uint32_t)(((unsigned long long)groups * q) / kWalkQueues);
}

__global__ void k_walk_queue_reset(uint32_t* __restrict__ queues) {
    if (threadIdx.x < kWalkQueues) queues[threadIdx.x * 32] = 0;
}

// Persistent walk: each wave takes the next 64-body group of its own CU's
// queue, so the waves resident on one CU walk neighbouring bodies and share
// the records in that CU's scalar cache; a wave whose queue is drained moves
// to the nearest queue (by slot, then XCD) that still holds groups. A queue
// that a wave found drained (by its atomic or by a load) is never tried again
// by that wave, so each wave ends after at most kWalkQueues failed takes.
template <uint32_t kMode, int kCorr>
__global__ __launch_bounds__(kWalkWaves * 64) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_walk_q(SimValues* __restrict__ sim, DevCounters* __restrict__ ctr,
                                              const float4* __restrict__ src, float4* __restrict__ dst,
                                              const uint32_t* __restrict__ index_in,
                                              const float4* __restrict__ leaf_pos,
                                              const WalkNode* __restrict__ nodes, const float4* __restrict__ qrec,
                                              const CorrArgs corr, uint2* __restrict__ tasks,
                                              uint32_t task_cap, float theta_, float dt_, float softening,
                                              uint32_t dynamic, uint32_t wrap, float elasticity, float restitution,
                                              const float* __restrict__ leaf_radius, uint32_t general,
                                              float4* __restrict__ pos_out, uint32_t* __restrict__ queues) {
    __shared__ uint32_t s_node[kWalkWaves][kStackDepth];
    __shared__ unsigned long long s_mask[kWalkWaves][kStackDepth];
    const uint32_t w = threadIdx.x >> 6, lane = lane_id();
    const uint32_t groups = (sim->numBodies + kWave - 1) / kWave;
    const uint32_t home = uniform(walk_home_queue());
    uint32_t q = home;
    uint32_t drained = 0;   // bit k: queue lane + 64 k is known to be drained
    WalkTotals t = {0ull, 0ull, 1u, 0u, 0u, 0ull};
    for (;;) {
        uint32_t k = 0;
        if (lane == 0) k = atomicAdd(&queues[q * 32], 1u);
        k = uniform(__shfl(k, 0, 64));
        const uint32_t qb = walk_queue_begin(groups, q);
        if (k < walk_queue_begin(groups, q + 1) - qb) {
            walk_group<kMode, kCorr>((qb + k) * kWave, s_node[w], s_mask[w], sim, ctr, src, dst, index_in, leaf_pos,
                                     nodes, qrec, corr, tasks, task_cap, theta_, dt_, softening, dynamic, wrap, elasticity,
                                     restitution, leaf_radius, general, pos_out, t);
            continue;
        }
        if (lane == (q & 63u)) drained |= 1u << (q >> 6);
        // the nearest queue after home that may still hold groups
        uint32_t best = 0xFFFFFFFFu;
        for (uint32_t j = 0; j < (kWalkQueues + 63) / 64; j++) {
            const uint32_t qq = lane + 64u * j;
            if (qq >= kWalkQueues || ((drained >> j) & 1u)) continue;
            const uint32_t h = __hip_atomic_load(&queues[qq * 32], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (h >= walk_queue_begin(groups, qq + 1) - walk_queue_begin(groups, qq)) {
                drained |= 1u << j;
                continue;
            }
            const uint32_t d = (qq + kWalkQueues - home) % kWalkQueues;
            best = d < best ? d : best;
        }
        for (uint32_t o = 32; o > 0; o >>= 1) {
            const uint32_t v = __shfl_xor(best, o, 64);
            best = v < best ? v : best;
        }
        best = uniform(best);
        if (best == 0xFFFFFFFFu) break;
        q = (home + best) % kWalkQueues;
    }
    walk_totals_flush(sim, ctr, t);
}

// the walk's totals back to their state before a walk (between the two walks
// of a first relative-test call)
__global__ void k_walk_counters_reset(SimValues* __restrict__ sim, DevCounters* __restrict__ ctr) {
    if (threadIdx.x != 0) return;
    ctr->opened = 0;
    ctr->interactions = 0;
    ctr->quadTerms = 0;
    ctr->maxStack = 0;
    sim->uintDebug[0] = 0;
}

void launch_walk_counters_reset(hipStream_t s, const SimValues* sim, DevCounters* ctr) {
    k_walk_counters_reset<<<1, 64, 0, s>>>(const_cast<SimValues*>(sim), ctr);
}

void launch_walk(hipStream_t s, uint32_t n, const WalkArgs& a) {
    const CorrArgs corr = {a.acc_mag, a.alpha};
    const uint32_t per_block = kWalkWaves * kWave;
    const uint32_t nb = (n + per_block - 1) / per_block;
    // (the corrected walk decides for whole 64-body groups: its results depend on the group)
    const bool half = n <= kWalkHalfMax && !a.qrec;
    const uint32_t per_wave = half ? kWalkSmallBodies / 2 : kWalkSmallBodies;
    const uint32_t nb_small = (n + kWalkSmallWaves * per_wave - 1) / (kWalkSmallWaves * per_wave);
#define BH_WALK_B(M, Q, B)                                                                                       \
    k_walk<M, kWalkSmallWaves, Q, B><<<nb_small, kWalkSmallWaves * kWave, 0, s>>>(const_cast<SimValues*>(a.sim), a.ctr, a.src, a.dst, a.index_in, \
                                                a.leaf_pos, a.nodes, a.qrec, corr, a.tasks, a.task_cap, a.theta, a.dt, \
                                                a.softening, a.dynamic, a.wrap, a.elasticity, a.restitution,     \
                                                a.leaf_radius, a.general, a.pos_out)
#define BH_WALK(M, Q)                                                                                            \
    do {                                                                                                         \
        if (half) BH_WALK_B(M, Q, kWalkSmallBodies / 2);                                                          \
        else BH_WALK_B(M, Q, kWalkSmallBodies);                                                                   \
    } while (0)
    if (a.merging != 0u)
        k_leaf_radius<<<(n + 255) / 256, 256, 0, s>>>(a.sim, a.src, a.index_in, a.leaf_pos, a.leaf_radius);
#define BH_WALK_Q(M, Q)                                                                                          \
    k_walk_q<M, Q><<<nq, kWalkWaves * kWave, 0, s>>>(const_cast<SimValues*>(a.sim), a.ctr, a.src, a.dst, a.index_in, \
                                                  a.leaf_pos, a.nodes, a.qrec, corr, a.tasks, a.task_cap, a.theta, a.dt, \
                                                  a.softening, a.dynamic, a.wrap, a.elasticity, a.restitution,     \
                                                  a.leaf_radius, a.general, a.pos_out, a.queues)
    // the corrected walk (quadrupole records) runs without merging or collisions
    if (kWalkPersistent && a.queues && n >= kWalkPersistentMin) {
        // two blocks of 16 waves fill a CU (8 waves per SIMD); blocks beyond
        // what is resident find the queues drained and end
        const uint32_t nq = std::min(nb, 2u * 256u);
        k_walk_queue_reset<<<1, kWalkQueues, 0, s>>>(a.queues);
        if (a.qrec && a.rel) BH_WALK_Q(0u, 2);
        else if (a.qrec) BH_WALK_Q(0u, 1);
        else if (a.merging == 1u) BH_WALK_Q(1u, 0);
        else if (a.merging == 2u) BH_WALK_Q(2u, 0);
        else BH_WALK_Q(0u, 0);
    } else if (a.qrec && a.rel) BH_WALK(0u, 2);
    else if (a.qrec) BH_WALK(0u, 1);
    else if (a.merging == 1u) BH_WALK(1u, 0);
    else if (a.merging == 2u) BH_WALK(2u, 0);
    else BH_WALK(0u, 0);
#undef BH_WALK
#undef BH_WALK_B
#undef BH_WALK_Q
}

// Ensemble walk (no reference counterpart): a persistent grid whose waves
// pull 64-body tasks from work queues until every queue is drained; each
// task walks its simulation's tree (root = that simulation's first internal
// node) with the same masked-stack walk and integrates. There is one queue
// per XCD holding a contiguous, equal-cost share of the task list (tasks in
// simulation order), so an XCD's waves walk few trees and keep them in its own
// L2; a wave whose queue is empty takes tasks from the next queues.
template <int kWaves, int kDepth, int kCorr>
__global__ __launch_bounds__(kWaves * 64) void k_ens_walk(const EnsSim* __restrict__ sims,
                                                          const EnsTask* __restrict__ tasks,
                                                          EnsQueue* __restrict__ queues, DevCounters* __restrict__ ctr,
                                                          const float4* __restrict__ src, float4* __restrict__ dst,
                                                          const uint32_t* __restrict__ index_in,
                                                          const float4* __restrict__ leaf_pos,
                                                          const WalkNode* __restrict__ nodes, uint32_t N, float G,
                                                          float theta, float dt, float soft, uint32_t general,
                                                          const float4* __restrict__ qrec, const CorrArgs corr) {
    __shared__ uint32_t s_node[kWaves][kDepth];
    __shared__ unsigned long long s_mask[kWaves][kDepth];
    const uint32_t w = threadIdx.x >> 6, lane = lane_id();
    // workgroups are dealt to the 8 XCDs round robin: this block's XCD
    const uint32_t home = blockIdx.x % kEnsQueues;
    const uint32_t height = ctr->treeHeight;   // the highest tree of the ensemble
    unsigned long long opened = 0, inter = 0, quads = 0;
    uint32_t maxsp = 1, err = 0;
    for (uint32_t qi = 0; qi < kEnsQueues; qi++) {
        EnsQueue* q = queues + (home + qi) % kEnsQueues;
        const uint32_t qbegin = q->begin, qcount = q->count;
        for (;;) {
            uint32_t t = 0;
            if (lane == 0) t = atomicAdd(&q->head, 1u);
            t = uniform(__shfl(t, 0, 64));
            if (t >= qcount) break;   // this queue is drained; every wave ends after the last queue
            const EnsTask tk = tasks[qbegin + t];
            const uint32_t root = N + sims[tk.sim].ibase;
            const uint32_t base = tk.start, g = base + lane;
            const bool active = lane < tk.count;
            const uint32_t bodyIdx = active ? index_in[g] : 0;
            float4 pm = active ? leaf_pos[g] : make_float4(0.f, 0.f, 0.f, 0.f);
            float4 vd = active ? src[2 * bodyIdx + 1] : make_float4(0.f, 0.f, 0.f, 1.f);
            WalkAcc o = {0.0f, 0.0f, 0.0f, 0ull, 0ull, 1u, 0u, 0ull};
            if constexpr (kCorr != 0) {
                // the corrected walk (as walk_group's), from this simulation's root
                const float afac =
                    kCorr == 2 ? corr_rel_a(corr.alpha, active ? corr.acc_mag[bodyIdx] : 0.0f) : theta * theta;
                // (every simulation of an ensemble holds at least two bodies)
                if (!walk_corrected_ok(nodes[root - N].half, soft)) {
                    o.err |= 4u;
                } else if (!general && height <= kLeanHeight && corr_offsets_32bit(N)) {
                    walk_corrected_pairs<false, kCorr == 2, kDepth>(N, root - N, N + kCorrGap - tk.sim, base, ballot(active), pm, qrec, afac,
                                                                    soft, s_mask[w], o);
                } else {
                    walk_corrected_pairs<true, kCorr == 2, kDepth>(N, root - N, N + kCorrGap - tk.sim, base, ballot(active), pm, qrec, afac,
                                                                   soft, s_mask[w], o);
                }
            } else if (!general && walk_lean_ok(nodes[root - N].half, soft, height, N))
                walk_lean<0u>(root, N, g, ballot(active), pm, vd, nodes, G, theta, soft, 0.0f, nullptr,
                              bodyIdx, 0.0f, 0.0f, index_in, src, ctr, nullptr, 0u, o);
            else if (inv_sqrt_fast_domain(nodes[root - N].half, soft))
                walk_chunk<kDepth, 0u, true>(s_node[w], s_mask[w], root, N, base, g, ballot(active), pm, vd, leaf_pos,
                                             nodes, G, theta, soft, 0.0f, nullptr, bodyIdx, 0.0f, 0.0f,
                                             index_in, src, ctr, nullptr, 0u, o);
            else
                walk_chunk<kDepth, 0u, false>(s_node[w], s_mask[w], root, N, base, g, ballot(active), pm, vd, leaf_pos,
                                              nodes, G, theta, soft, 0.0f, nullptr, bodyIdx, 0.0f, 0.0f,
                                              index_in, src, ctr, nullptr, 0u, o);
            if (active) {   // kick-drift (bh_force.comp:118-140), dynamic bounds
                const float nvx = fadd(vd.x, fmul(o.ax, dt));
                const float nvy = fadd(vd.y, fmul(o.ay, dt));
                const float nvz = fadd(vd.z, fmul(o.az, dt));
                dst[2 * bodyIdx] = make_float4(fadd(pm.x, fmul(nvx, dt)), fadd(pm.y, fmul(nvy, dt)),
                                               fadd(pm.z, fmul(nvz, dt)), pm.w);
                dst[2 * bodyIdx + 1] = make_float4(nvx, nvy, nvz, vd.w);
                if constexpr (kCorr != 0)
                    corr.acc_mag[bodyIdx] = fsqrt(__builtin_fmaf(o.ax, o.ax, __builtin_fmaf(o.ay, o.ay, o.az * o.az)));
            }
            opened += o.opened;
            inter += o.inter;
            quads += o.quads;
            maxsp = o.maxsp > maxsp ? o.maxsp : maxsp;
            err |= o.err;
        }
    }
    if (lane == 0) {
        atomicAdd(&ctr->opened, opened);
        atomicAdd(&ctr->interactions, inter);
        if (quads) atomicAdd(&ctr->quadTerms, quads);
        atomicMax(&ctr->maxStack, maxsp);
        if (err) atomicOr(&ctr->errorFlags, err);
    }
}

#define BH_ENS_WALK_SHAPES(X) \
    X(1, 64) X(1, 128) X(1, 192) X(2, 64) X(2, 128) X(2, 192) X(4, 64) X(4, 128) X(4, 192) X(8, 64) X(8, 128) X(8, 192)

int launch_ens_walk(hipStream_t s, int waves, int depth, uint32_t blocks, const EnsSim* sims, const EnsTask* tasks,
                    EnsQueue* queues, DevCounters* ctr, const float4* src, float4* dst,
                    const uint32_t* index_in, const float4* leaf_pos, const WalkNode* nodes, uint32_t N, float G,
                    float theta, float dt, float soft, uint32_t general, const float4* qrec, float* acc_mag,
                    uint32_t rel, float alpha) {
    const CorrArgs corr = {acc_mag, alpha};
#define BH_LAUNCH(W, D)                                                                                            \
    if (waves == W && depth == D) {                                                                                \
        if (qrec && rel)                                                                                           \
            k_ens_walk<W, D, 2><<<blocks, W * 64, 0, s>>>(sims, tasks, queues, ctr, src, dst, index_in, leaf_pos,  \
                                                          nodes, N, G, theta, dt, soft, general, qrec, corr);      \
        else if (qrec)                                                                                             \
            k_ens_walk<W, D, 1><<<blocks, W * 64, 0, s>>>(sims, tasks, queues, ctr, src, dst, index_in, leaf_pos,  \
                                                          nodes, N, G, theta, dt, soft, general, qrec, corr);      \
        else                                                                                                       \
            k_ens_walk<W, D, 0><<<blocks, W * 64, 0, s>>>(sims, tasks, queues, ctr, src, dst, index_in, leaf_pos,  \
                                                          nodes, N, G, theta, dt, soft, general, qrec, corr);      \
        return 0;                                                                                                  \
    }
    BH_ENS_WALK_SHAPES(BH_LAUNCH)
#undef BH_LAUNCH
    return 1;
}

int ens_walk_occupancy(int waves, int depth, int* blocks_per_cu) {
#define BH_OCC(W, D)                                                                                               \
    if (waves == W && depth == D)                                                                                  \
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, k_ens_walk<W, D, 0>, W * 64, 0) ==      \
                       hipSuccess                                                                                  \
                   ? 0                                                                                             \
                   : 2;
    BH_ENS_WALK_SHAPES(BH_OCC)
#undef BH_OCC
    return 1;
}

}  // namespace bh
