// bh_mlp.hpp — the far-field correction network: features, layout and the
// wave-level BF16 MFMA forward pass shared by training and the walk.
//
// Network: 12 inputs -> 64 (SiLU) -> 64 (SiLU) -> 3. Parameters, FP32 master
// copy, one flat array: W1[64][12], b1[64], W2[64][64], b2[64], W3[3][64], b3[3].
//
// Inference runs on v_mfma_f32_16x16x32_bf16 with the weights as the A
// operand (rows = output neurons) and 16 samples as the B operand (columns),
// so every activation stays in registers: the D fragment of one layer (lane
// holds neurons 4*(lane/16) + t of sample lane%16 in each 16-neuron block) is
// the B fragment of the next once the hidden dimension is read in the order
// hid(s, grp, j) = 16*(2s + j/4) + 4*grp + j%4 (K slice s, lane group grp,
// element j). The device keeps the BF16 weights in that fragment order:
// 14 fragments (W1 by output block, W2 by output block and K slice, W3 by K
// slice), fragment f of lane L at elements (64 f + L) * 8 .. +7, so each lane
// loads its 14 fragments with coalesced 16-byte loads and holds them in
// registers for the whole kernel. Biases (FP32, b1[64] b2[64] b3[16]) seed
// the accumulators.
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

constexpr int kMlpIn = 12, kMlpHidden = 64, kMlpOut = 3;
constexpr int kMlpInPad = 32, kMlpOutPad = 16;
constexpr int kMlpParams = kMlpHidden * kMlpIn + kMlpHidden + kMlpHidden * kMlpHidden + kMlpHidden +
                           kMlpOut * kMlpHidden + kMlpOut;  // 5,187
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

// hidden neuron read as element j of K slice s by lane group grp
__host__ __device__ constexpr int mlp_hid(int s, int grp, int j) { return 16 * (2 * s + (j >> 2)) + 4 * grp + (j & 3); }

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// 12 dimensionless features of a (target, node) pair. r = com - x, the
// softened distance d = sqrt(|r|^2 + eps^2):
//   0-2  AABB extents / d            3-5  (com - box centre) / d
//   6-8  r / d (direction)           9    ln(mass) / 16
//   10   1 / sqrt(bodies in node)    11   eps^2 / d^2 (softening share)
// Feature 10 is the node's discreteness scale: it tends to 0 for the large
// nodes whose residual depends only on the smooth mass profile, so a network
// trained at one N does not extrapolate in it at a larger N.
__device__ __forceinline__ void mlp_features(float rx, float ry, float rz, float inv, float soft, float mass,
                                             uint32_t count, float4 mn, float4 mx, float4 com, float* f) {
    f[0] = (mx.x - mn.x) * inv;
    f[1] = (mx.y - mn.y) * inv;
    f[2] = (mx.z - mn.z) * inv;
    f[3] = (com.x - 0.5f * (mn.x + mx.x)) * inv;
    f[4] = (com.y - 0.5f * (mn.y + mx.y)) * inv;
    f[5] = (com.z - 0.5f * (mn.z + mx.z)) * inv;
    f[6] = rx * inv;
    f[7] = ry * inv;
    f[8] = rz * inv;
    f[9] = __logf(fmaxf(mass, 1e-30f)) * 0.0625f;
    f[10] = __builtin_amdgcn_rsqf((float)count);
    f[11] = soft * inv * inv;
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
// rest zero) written by the caller; ys: LDS [64][4] FP32 outputs (3 used).
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
        // outputs 0..2 of sample 16 ns + r16 sit in lane group 0, elements 0..2
        if (grp == 0) *(f32x4*)(ys + (16 * ns + r16) * 4) = o;
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
}

// mlp_wave_forward with the fragments read from LDS as each MFMA needs them
// (wfrag, bias as for mlp_load_lds): a few fragments live at a time instead
// of all 14, which keeps the calling kernel's register count low.
__device__ __forceinline__ void mlp_wave_forward_lds(const bf16x8* wfrag, const f32x4* bias, const __bf16* xs,
                                                     float* ys, uint32_t blocks) {
    const int lane = (int)lane_id();
    const int r16 = lane & 15, grp = lane >> 4;
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    const bf16x8 zero = {};
    for (int ns = 0; ns < 4; ns++) {
        if (!((blocks >> ns) & 1u)) continue;
        const bf16x8 xb = grp < 2 ? *(const bf16x8*)(xs + (16 * ns + r16) * kMlpXStride + 8 * grp) : zero;
        f32x4 a[4];
        for (int mo = 0; mo < 4; mo++)
            a[mo] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wfrag[(kFragW1 + mo) * 64 + lane], xb,
                                                            bias[4 * mo + grp], 0, 0, 0);
        bf16x8 h0 = pack_silu(a[0], a[1]), h1 = pack_silu(a[2], a[3]);
        for (int mo = 0; mo < 4; mo++) {
            a[mo] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wfrag[(kFragW2 + 2 * mo) * 64 + lane], h0,
                                                            bias[16 + 4 * mo + grp], 0, 0, 0);
            a[mo] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wfrag[(kFragW2 + 2 * mo + 1) * 64 + lane], h1, a[mo],
                                                            0, 0, 0);
        }
        h0 = pack_silu(a[0], a[1]);
        h1 = pack_silu(a[2], a[3]);
        f32x4 o = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wfrag[kFragW3 * 64 + lane], h0, bias[32 + grp], 0, 0, 0);
        o = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wfrag[(kFragW3 + 1) * 64 + lane], h1, o, 0, 0, 0);
        if (grp == 0) *(f32x4*)(ys + (16 * ns + r16) * 4) = o;
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
