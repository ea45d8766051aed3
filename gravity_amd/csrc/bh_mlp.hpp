// bh_mlp.hpp — the far-field correction: node features, the quadrupole basis,
// the network layout and the wave-level BF16 MFMA forward pass shared by
// training, the per-node evaluation and the tests.
//
// Correction. An internal node accepted by the walk contributes its monopole
// plus the softened quadrupole of an effective tensor T = M2 + M h^2 o, where
// M2 is the node's exact second moment about its centre of mass, h half its
// longest side, and o the network's output for that node (o = 0: the exact
// quadrupole). For a target at r = com - x, s^2 = |r|^2 + eps^2, the
// quadrupole term of the Plummer-softened kernel is
//     a_q = G [ -3 T r / s^5 - 1.5 tr(T) r / s^5 + 7.5 (r.T.r) r / s^7 ],
// and in units of the monopole scale G M / s^2 the network's share is
//     a_q(M h^2 o) / (G M / s^2) = (h / s)^2 B(u) o,   u = r / s,
//     B(u) o = -3 o u - 1.5 tr(o) u + 7.5 (u.o.u) u.
// Tensors are stored as 6 values: xx, xy, xz, yy, yz, zz.
//
// Network: 12 inputs -> 64 (SiLU) -> 64 (SiLU) -> 6. Parameters, FP32 master
// copy, one flat array: W1[64][12], b1[64], W2[64][64], b2[64], W3[6][64],
// b3[6]. Inference runs on v_mfma_f32_16x16x32_bf16 with the weights as the
// A operand (rows = output neurons) and 16 samples as the B operand (columns),
// so every activation stays in registers: the D fragment of one layer (lane
// holds neurons 4*(lane/16) + t of sample lane%16 in each 16-neuron block) is
// the B fragment of the next once the hidden dimension is read in the order
// hid(s, grp, j) = 16*(2s + j/4) + 4*grp + j%4 (K slice s, lane group grp,
// element j). The device keeps the BF16 weights in that fragment order:
// 14 fragments (W1 by output block, W2 by output block and K slice, W3 by K
// slice), fragment f of lane L at elements (64 f + L) * 8 .. +7, so each lane
// loads its 14 fragments with coalesced 16-byte loads and holds them in
// registers. Biases (FP32, b1[64] b2[64] b3[16]) seed the accumulators.
//
// The packed copy carries log2(e) in the hidden pre-activations: W1, b1 and
// b2 are scaled by log2(e) and W3 by ln(2), so a hidden unit holds
// a' = log2(e) a and SiLU becomes a' / (1 + 2^-a') = log2(e) silu(a), whose
// factor the next layer's unscaled W2 (or the ln(2) in W3) takes back.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "bh_device.hpp"

namespace bh {

constexpr int kMlpIn = 12, kMlpHidden = 64, kMlpOut = 6;
constexpr int kMlpInPad = 32, kMlpOutPad = 16;
constexpr int kMlpTarget = 7;   // a training row's target: u (3), the scaled residual (3), its weight
constexpr int kMlpParams = kMlpHidden * kMlpIn + kMlpHidden + kMlpHidden * kMlpHidden + kMlpHidden +
                           kMlpOut * kMlpHidden + kMlpOut;  // 5,382
// offsets into the flat FP32 parameter array
constexpr int kOffW1 = 0, kOffB1 = kOffW1 + kMlpHidden * kMlpIn, kOffW2 = kOffB1 + kMlpHidden;
constexpr int kOffB2 = kOffW2 + kMlpHidden * kMlpHidden, kOffW3 = kOffB2 + kMlpHidden;
constexpr int kOffB3 = kOffW3 + kMlpOut * kMlpHidden;
// BF16 weight fragments: W1 4, W2 8, W3 2; 64 lanes x 8 elements each
constexpr int kFragW1 = 0, kFragW2 = 4, kFragW3 = 12, kFrags = 14;
constexpr int kBfTotal = kFrags * 64 * 8;   // 7,168 bf16 = 14 KiB
constexpr int kBiasTotal = 2 * kMlpHidden + kMlpOutPad;   // b1, b2, b3 padded: 144 floats
// Features per sample staged in LDS: 12 used, padded to 16.
constexpr int kMlpXStride = 16;
// Outputs per sample staged in LDS: 6 used, padded to 8.
constexpr int kMlpYStride = 8;

// hidden neuron read as element j of K slice s by lane group grp
__host__ __device__ constexpr int mlp_hid(int s, int grp, int j) { return 16 * (2 * s + (j >> 2)) + 4 * grp + (j & 3); }

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// The 12 dimensionless features of an internal node, all of order one:
//   0-2  AABB extents / (2 h)                (the longest is 1)
//   3-5  (centre of mass - box centre) / h
//   6-11 M2 / (M h^2): xx, xy, xz, yy, yz, zz
// h = half the longest side, M2 the second moment about the centre of mass.
__device__ __forceinline__ void node_features(float4 mn, float4 mx, float cx, float cy, float cz, float h,
                                              const double* m2_over_mh2, float* f) {
    const float ih = 1.0f / h;
    f[0] = 0.5f * (mx.x - mn.x) * ih;
    f[1] = 0.5f * (mx.y - mn.y) * ih;
    f[2] = 0.5f * (mx.z - mn.z) * ih;
    f[3] = (cx - 0.5f * (mn.x + mx.x)) * ih;
    f[4] = (cy - 0.5f * (mn.y + mx.y)) * ih;
    f[5] = (cz - 0.5f * (mn.z + mx.z)) * ih;
    for (int c = 0; c < 6; c++) f[6 + c] = (float)m2_over_mh2[c];
}

// B(u) o (see above), and the gradient g = B(u)^T e
__host__ __device__ __forceinline__ void quad_basis(const float* u, const float* o, float* p) {
    const float ou0 = o[0] * u[0] + o[1] * u[1] + o[2] * u[2];
    const float ou1 = o[1] * u[0] + o[3] * u[1] + o[4] * u[2];
    const float ou2 = o[2] * u[0] + o[4] * u[1] + o[5] * u[2];
    const float uou = u[0] * ou0 + u[1] * ou1 + u[2] * ou2;
    const float c = 7.5f * uou - 1.5f * (o[0] + o[3] + o[5]);
    p[0] = c * u[0] - 3.0f * ou0;
    p[1] = c * u[1] - 3.0f * ou1;
    p[2] = c * u[2] - 3.0f * ou2;
}
__host__ __device__ __forceinline__ void quad_basis_t(const float* u, const float* e, float* g) {
    const float eu = e[0] * u[0] + e[1] * u[1] + e[2] * u[2];
    g[0] = -3.0f * e[0] * u[0] - 1.5f * eu + 7.5f * u[0] * u[0] * eu;
    g[1] = -3.0f * (e[0] * u[1] + e[1] * u[0]) + 15.0f * u[0] * u[1] * eu;
    g[2] = -3.0f * (e[0] * u[2] + e[2] * u[0]) + 15.0f * u[0] * u[2] * eu;
    g[3] = -3.0f * e[1] * u[1] - 1.5f * eu + 7.5f * u[1] * u[1] * eu;
    g[4] = -3.0f * (e[1] * u[2] + e[2] * u[1]) + 15.0f * u[1] * u[2] * eu;
    g[5] = -3.0f * e[2] * u[2] - 1.5f * eu + 7.5f * u[2] * u[2] * eu;
}

// SiLU of a pre-activation held as x = log2(e) a, returning log2(e) silu(a),
// with the hardware exp2 and reciprocal (the result is rounded to BF16 next,
// far coarser than their error)
__device__ __forceinline__ float silu_fast(float x) {
    return x * __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(-x));
}

__device__ __forceinline__ bf16x8 pack_silu(f32x4 lo, f32x4 hi) {
    bf16x8 r;
    for (int t = 0; t < 4; t++) {
        r[t] = (__bf16)silu_fast(lo[t]);
        r[4 + t] = (__bf16)silu_fast(hi[t]);
    }
    return r;
}

// The network held in one lane's registers (see the layout above).
struct MlpRegs {
    bf16x8 w1[4], w2[4][2], w3[2];
    f32x4 b1[4], b2[4], b3;
};

__device__ __forceinline__ void mlp_load(MlpRegs& R, const __bf16* __restrict__ wfrag,
                                         const float* __restrict__ bias) {
    const int lane = (int)lane_id(), grp = lane >> 4;
    const bf16x8* f = (const bf16x8*)wfrag;
    for (int mo = 0; mo < 4; mo++) R.w1[mo] = f[(kFragW1 + mo) * 64 + lane];
    for (int mo = 0; mo < 4; mo++)
        for (int s = 0; s < 2; s++) R.w2[mo][s] = f[(kFragW2 + 2 * mo + s) * 64 + lane];
    for (int s = 0; s < 2; s++) R.w3[s] = f[(kFragW3 + s) * 64 + lane];
    const f32x4* b = (const f32x4*)bias;
    for (int mo = 0; mo < 4; mo++) {
        R.b1[mo] = b[4 * mo + grp];
        R.b2[mo] = b[16 + 4 * mo + grp];
    }
    R.b3 = b[32 + grp];
}

// One wave: for each 16-sample block ns set in `blocks`, y = MLP(x) of the
// samples 16 ns .. 16 ns + 15. xs: LDS [64][16] bf16 features (12 used, the
// rest zero) written by the caller; ys: LDS [64][8] FP32 outputs (6 used).
// Rows of blocks not evaluated are left as they were.
__device__ __forceinline__ void mlp_wave_forward(const MlpRegs& R, const __bf16* xs, float* ys, uint32_t blocks) {
    const int lane = (int)lane_id();
    const int r16 = lane & 15, grp = lane >> 4;
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    const bf16x8 zero = {};
    for (int ns = 0; ns < 4; ns++) {
        if (!((blocks >> ns) & 1u)) continue;
        const bf16x8 xb = grp < 2 ? *(const bf16x8*)(xs + (16 * ns + r16) * kMlpXStride + 8 * grp) : zero;
        f32x4 a[4];
        for (int mo = 0; mo < 4; mo++) a[mo] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(R.w1[mo], xb, R.b1[mo], 0, 0, 0);
        bf16x8 h0 = pack_silu(a[0], a[1]), h1 = pack_silu(a[2], a[3]);
        for (int mo = 0; mo < 4; mo++) {
            a[mo] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(R.w2[mo][0], h0, R.b2[mo], 0, 0, 0);
            a[mo] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(R.w2[mo][1], h1, a[mo], 0, 0, 0);
        }
        h0 = pack_silu(a[0], a[1]);
        h1 = pack_silu(a[2], a[3]);
        f32x4 o = __builtin_amdgcn_mfma_f32_16x16x32_bf16(R.w3[0], h0, R.b3, 0, 0, 0);
        o = __builtin_amdgcn_mfma_f32_16x16x32_bf16(R.w3[1], h1, o, 0, 0, 0);
        // outputs 0..3 of sample 16 ns + r16 sit in lane group 0, outputs 4..7
        // (4 and 5 used) in lane group 1
        if (grp < 2) *(f32x4*)(ys + (16 * ns + r16) * kMlpYStride + 4 * grp) = o;
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
}

// A lane's features as two 16-byte LDS stores into row `row` of xs.
__device__ __forceinline__ void mlp_stage_row(__bf16* xs, int row, const float* f) {
    bf16x8 lo, hi;
    for (int c = 0; c < 8; c++) lo[c] = (__bf16)f[c];
    for (int c = 0; c < 4; c++) hi[c] = (__bf16)f[8 + c];
    for (int c = 4; c < 8; c++) hi[c] = (__bf16)0.0f;
    *(bf16x8*)(xs + row * kMlpXStride) = lo;
    *(bf16x8*)(xs + row * kMlpXStride + 8) = hi;
}
__device__ __forceinline__ void mlp_stage_features(__bf16* xs, const float* f) {
    mlp_stage_row(xs, (int)lane_id(), f);
}

}  // namespace bh
