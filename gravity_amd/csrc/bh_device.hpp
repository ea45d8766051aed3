// bh_device.hpp — device-side data layout and exact FP32 helpers shared by
// the Barnes-Hut kernels (gfx950, wave64).
//
// Arithmetic contract: the kernels reproduce the reference shaders' FP32
// results bit for bit. Every expression keeps the GLSL operation order with
// IEEE-rounded add/mul/div/sqrt and no multiply-add contraction (the file is
// compiled with -ffp-contract=off and the helpers below use the _rn
// intrinsics); GLSL min/max are IEEE minNum/maxNum (fminf/fmaxf).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bh {

constexpr uint32_t kEmpty = 0xFFFFFFFFu;
constexpr int kWave = 64;

// SimulationValues, std430 as GLSL lays it out (common/common.glsl:86-88):
// the 28-byte UnitSet puts uintDebug at byte 92 and floatDebug at 492.
struct SimValues {
    uint32_t numBodies, initialNumBodies, justDied, merged, outOfBounds, relativeTo, pad1, pad2;
    float minCorner[3];
    float padA;
    float maxCorner[3];
    float padB;
    float unitMass, unitDensity, unitLen, unitTime, cameraScale, gravitationalConstant, bodyLengthConstant;
    uint32_t uintDebug[100];
    float floatDebug[100];
    uint32_t tailPad;
};
static_assert(sizeof(SimValues) == 896, "SimulationValues is 896 bytes");

// std430 Node, 80 bytes (common/common.glsl:40-58, gpu/Node.java:31-41)
struct NodeStd430 {
    float comMass[4];
    float aabb[6];
    uint32_t childA, childB, nodeDepth, bodiesContained, readyChildren, parentId;
    float avgColor[4];
};
static_assert(sizeof(NodeStd430) == 80, "Node is 80 bytes");

// What the walk reads per internal node: one 32-byte record, one scalar load.
struct alignas(32) WalkNode {
    float4 comMass;     // centre of mass, G * mass (the walk's factor, rounded once as each interaction would)
    float half;         // half of the longest AABB side (bh_force.comp:14-20)
    uint32_t childA;    // global node ids: < n leaf, >= n internal (n = initialNumBodies)
    uint32_t childB;
    float half2;        // half * half (the walk's cheap test for nodes every lane opens)
};
static_assert(sizeof(WalkNode) == 32, "WalkNode is 32 bytes");

// Counters the host reads back; reset by the reset stage of every step.
struct DevCounters {
    unsigned long long opened;
    unsigned long long interactions;
    uint32_t nTasks;
    uint32_t mergesApplied;
    uint32_t treeHeight;
    uint32_t maxStack;
    uint32_t errorFlags;       // bit0 stack overflow, bit1 task overflow
    uint32_t numBefore;        // numBodies before the last dead partition
    uint32_t died;             // bodies removed by the last dead partition
    uint32_t aabbTicket;       // last-block-done ticket of the bounds reduction
    uint32_t mergeRound;
    uint32_t zeroMassIds;      // body ids with zero mass found by the dead-body census
    unsigned long long corrected;   // network evaluations of the corrected walk's tree stage
    unsigned long long quadTerms;   // quadrupole terms evaluated by the corrected walk
    uint32_t pad[2];
};

// ---- exact FP32 operations (no contraction, IEEE rounding)
__device__ __forceinline__ float fadd(float a, float b) { return __fadd_rn(a, b); }
__device__ __forceinline__ float fsub(float a, float b) { return __fsub_rn(a, b); }
__device__ __forceinline__ float fmul(float a, float b) { return __fmul_rn(a, b); }
__device__ __forceinline__ float fdiv(float a, float b) { return __fdiv_rn(a, b); }
// __builtin_sqrtf is lowered to the correctly rounded sequence (hardware
// v_sqrt_f32 plus a one-ulp residual correction); __fsqrt_rn is not on gfx950.
__device__ __forceinline__ float fsqrt(float a) { return __builtin_sqrtf(a); }

// 1 / sqrt(d2) rounded exactly as RN(1 / RN(sqrt(d2))), the reference's
// inversesqrt, for normal finite d2 (the walk's d2 >= softening > 0).
// The square root starts from the hardware 1/sqrt y (a few ulps): s = x y is
// then within a few ulps of sqrt(x), and one residual step with h = y / 2
// (r = x - s s exactly, by fma; s + r h) rounds it correctly; the reciprocal
// is the hardware v_rcp_f32 (within one ulp) plus one Newton step. For every
// float of [2^-100, 2^100] both are the correctly rounded results, as
// bh_selftest_math checks against the general IEEE sequences on the device.
__device__ __forceinline__ float sqrt_cr_normal(float x) {
    const float y = __builtin_amdgcn_rsqf(x);
    const float s = __fmul_rn(x, y);
    const float h = __fmul_rn(0.5f, y);
    const float r = __builtin_fmaf(-s, s, x);
    return __builtin_fmaf(r, h, s);
}
__device__ __forceinline__ float rcp_cr_normal(float y) {
    float r = __builtin_amdgcn_rcpf(y);
    float e = __builtin_fmaf(-y, r, 1.0f);
    return __builtin_fmaf(e, r, r);
}
// Whether every d2 = |r|^2 + soft of a walk lies in the fast domain: r joins
// two points of the root's box (bodies and centres of mass), so |r|^2 is at
// most 12 half^2 for the root's half longest side; a NaN or infinite box
// fails the test.
__device__ __forceinline__ bool inv_sqrt_fast_domain(float root_half, float soft) {
    return soft >= 1e-30f && fadd(fmul(fmul(root_half, root_half), 16.0f), soft) <= 1e29f;
}
__device__ __forceinline__ float inv_sqrt_fast(float d2) { return rcp_cr_normal(sqrt_cr_normal(d2)); }
__device__ __forceinline__ float inv_sqrt_exact(float d2) {
    // fast sequence on [2^-100, 2^100], the general IEEE one elsewhere
    if (d2 >= 7.88860905e-31f && d2 <= 1.26765060e30f) return rcp_cr_normal(sqrt_cr_normal(d2));
    return fdiv(1.0f, fsqrt(d2));
}
__device__ __forceinline__ float gmin(float a, float b) { return fminf(a, b); }
__device__ __forceinline__ float gmax(float a, float b) { return fmaxf(a, b); }

// GLSL dot(r, r) as the reference executes it: x*x + (y*y + z*z)
__device__ __forceinline__ float dot3(float x, float y, float z) {
    return fadd(fmul(x, x), fadd(fmul(y, y), fmul(z, z)));
}
__device__ __forceinline__ float dot3(float ax, float ay, float az, float bx, float by, float bz) {
    return fadd(fmul(ax, bx), fadd(fmul(ay, by), fmul(az, bz)));
}

__device__ __forceinline__ bool finite3(float x, float y, float z) {
    return !(__builtin_isnan(x) || __builtin_isnan(y) || __builtin_isnan(z) || __builtin_isinf(x) ||
             __builtin_isinf(y) || __builtin_isinf(z));
}

// ---- wave helpers
__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }
__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }
__device__ __forceinline__ uint64_t lanemask_lt() {
    uint32_t l = lane_id();
    return l == 0 ? 0ull : (~0ull >> (64 - l));
}
__device__ __forceinline__ uint32_t uniform(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }

// Morton interleave of a 21-bit integer (shaders/compute/morton/bh_morton.comp:6-15)
__device__ __forceinline__ uint64_t expand_bits21(uint32_t v) {
    uint64_t x = (uint64_t)v & 0x1FFFFFull;
    x = (x | (x << 32)) & 0x1F00000000FFFFull;
    x = (x | (x << 16)) & 0x1F0000FF0000FFull;
    x = (x | (x << 8)) & 0x100F00F00F00F00Full;
    x = (x | (x << 4)) & 0x10C30C30C30C30C3ull;
    x = (x | (x << 2)) & 0x1249249249249249ull;
    return x;
}

// Karras prefix length of sorted positions i, j (shaders/compute/tree/bh_tree.comp:6-37):
// equal codes tie-break on the index, i.e. the 96-bit key (code, index).
__device__ __forceinline__ int safe_lcp(const uint64_t* __restrict__ keys, int num, int i, int j) {
    if (i < 0 || j < 0 || i >= num || j >= num) return -1;
    uint64_t a = keys[i], b = keys[j];
    if (a == b) return i == j ? 64 : 64 + __clz((uint32_t)i ^ (uint32_t)j);
    return __clzll(a ^ b);
}

// Block-id remap so that each XCD (blocks are dealt round-robin over 8 XCDs)
// works on one contiguous range of sorted bodies and keeps its share of the
// tree in its own L2.
__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t nblocks) {
    const uint32_t kXcd = 8;
    uint32_t per = nblocks / kXcd;
    uint32_t rem = nblocks % kXcd;
    uint32_t x = b % kXcd, k = b / kXcd;
    // XCD x owns per (+1 for x < rem) consecutive logical blocks
    uint32_t start = x * per + (x < rem ? x : rem);
    return start + k;
}

}  // namespace bh
