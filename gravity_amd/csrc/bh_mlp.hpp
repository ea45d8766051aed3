// bh_mlp.hpp — the far-field correction network: features, layout and the
// wave-level BF16 MFMA forward pass shared by training and the walk.
//
// Network: 12 inputs -> 64 (SiLU) -> 64 (SiLU) -> 3. Parameters, FP32 master
// copy, one flat array: W1[64][12], b1[64], W2[64][64], b2[64], W3[3][64], b3[3].
// On the device the weights are also kept as BF16, row-major [out][in] with
// the input dimension padded to a multiple of 32 (W1: [64][32], W2: [64][64],
// W3: [16][64], padding zero), which is the B-operand layout of
// v_mfma_f32_16x16x32_bf16 (each lane reads 8 consecutive inputs of one
// output). A wave evaluates 64 samples at once: rows of the A operand are
// samples, staged in LDS.
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
// BF16 weight block (elements): W1 [64][32], W2 [64][64], W3 [16][64]
constexpr int kBfW1 = 0, kBfW2 = kBfW1 + kMlpHidden * kMlpInPad, kBfW3 = kBfW2 + kMlpHidden * kMlpHidden;
constexpr int kBfTotal = kBfW3 + kMlpOutPad * kMlpHidden;   // 7,168 bf16 = 14 KiB

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// 12 dimensionless features of a (target, node) pair. r = com - x, the
// softened distance d = sqrt(|r|^2 + eps^2):
//   0-2  AABB extents / d            3-5  (com - box centre) / d
//   6-8  r / d (direction)           9    ln(mass) / 16
//   10   ln(bodies in node) / 16     11   eps^2 / d^2 (softening share)
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
    f[10] = __logf((float)count) * 0.0625f;
    f[11] = soft * inv * inv;
}

__device__ __forceinline__ float silu(float x) { return x / (1.0f + __expf(-x)); }

__device__ __forceinline__ bf16x8 lds_frag8(const __bf16* p) { return *(const bf16x8*)p; }

// One wave: y[64][3] = MLP(x[64][32]) with x (bf16, row-major, padded) in
// LDS `xs`, hidden scratch `hs` ([64][64] bf16) in LDS, weights `w` (bf16
// block, LDS) and FP32 biases `bias` (LDS: b1[64], b2[64], b3[16]).
// Returns the 3 outputs of the sample in row `lane`... of this lane's row.
__device__ __forceinline__ void mlp_wave_forward(const __bf16* xs, __bf16* hs, const __bf16* w,
                                                 const float* bias, float* ys /* [64][4] LDS */) {
    const int lane = (int)lane_id();
    const int r16 = lane & 15, grp = lane >> 4;
    // layer 1: [64 x 32] x [32 x 64]
    f32x4 acc[4][4];
    for (int m = 0; m < 4; m++)
        for (int nn = 0; nn < 4; nn++) acc[m][nn] = (f32x4){0.f, 0.f, 0.f, 0.f};
    for (int m = 0; m < 4; m++) {
        bf16x8 a = lds_frag8(xs + (16 * m + r16) * kMlpInPad + grp * 8);
        for (int nn = 0; nn < 4; nn++) {
            bf16x8 b = lds_frag8(w + kBfW1 + (16 * nn + r16) * kMlpInPad + grp * 8);
            acc[m][nn] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[m][nn], 0, 0, 0);
        }
    }
    // bias + SiLU -> hs (D layout: row 16m + 4 grp + t, column 16 nn + r16)
    for (int m = 0; m < 4; m++)
        for (int nn = 0; nn < 4; nn++)
            for (int t = 0; t < 4; t++) {
                int i = 16 * m + 4 * grp + t, j = 16 * nn + r16;
                hs[i * kMlpHidden + j] = (__bf16)silu(acc[m][nn][t] + bias[j]);
            }
    __builtin_amdgcn_wave_barrier();
    // layer 2: [64 x 64] x [64 x 64], result kept in registers until all reads are done
    for (int m = 0; m < 4; m++)
        for (int nn = 0; nn < 4; nn++) acc[m][nn] = (f32x4){0.f, 0.f, 0.f, 0.f};
    for (int s = 0; s < 2; s++)
        for (int m = 0; m < 4; m++) {
            bf16x8 a = lds_frag8(hs + (16 * m + r16) * kMlpHidden + 32 * s + grp * 8);
            for (int nn = 0; nn < 4; nn++) {
                bf16x8 b = lds_frag8(w + kBfW2 + (16 * nn + r16) * kMlpHidden + 32 * s + grp * 8);
                acc[m][nn] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[m][nn], 0, 0, 0);
            }
        }
    __builtin_amdgcn_wave_barrier();
    for (int m = 0; m < 4; m++)
        for (int nn = 0; nn < 4; nn++)
            for (int t = 0; t < 4; t++) {
                int i = 16 * m + 4 * grp + t, j = 16 * nn + r16;
                hs[i * kMlpHidden + j] = (__bf16)silu(acc[m][nn][t] + bias[64 + j]);
            }
    __builtin_amdgcn_wave_barrier();
    // layer 3: [64 x 64] x [64 x 16]
    f32x4 out[4];
    for (int m = 0; m < 4; m++) out[m] = (f32x4){0.f, 0.f, 0.f, 0.f};
    for (int s = 0; s < 2; s++)
        for (int m = 0; m < 4; m++) {
            bf16x8 a = lds_frag8(hs + (16 * m + r16) * kMlpHidden + 32 * s + grp * 8);
            bf16x8 b = lds_frag8(w + kBfW3 + r16 * kMlpHidden + 32 * s + grp * 8);
            out[m] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, out[m], 0, 0, 0);
        }
    // column r16 < 3 holds output r16 of rows 16m + 4 grp + t
    if (r16 < kMlpOut)
        for (int m = 0; m < 4; m++)
            for (int t = 0; t < 4; t++) ys[(16 * m + 4 * grp + t) * 4 + r16] = out[m][t] + bias[128 + r16];
    __builtin_amdgcn_wave_barrier();
}

}  // namespace bh
