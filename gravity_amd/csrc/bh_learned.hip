// bh_learned.hip — the quadrupole-corrected walk's node records and the
// learned residual (no reference counterpart).
//
//  * k_qnodes: per internal node, the second moment about its centre of mass
//    (built bottom-up with the tree, T.node_mom), the network's residual (BF16
//    MFMA, one evaluation per node), and the 64-byte record the walk reads.
//  * k_collect_pairs / k_pair_labels: (target, accepted node) pairs of the walk
//    and their FP64 labels: the exact field of the node's bodies minus its
//    monopole and exact quadrupole, in units of (h / s)^2 G M / s^2.
//  * training: FP32 GEMM / activation / Adam kernels (forward and backward of
//    the 12-64-64-6 SiLU network, the loss taken through the quadrupole
//    basis), checked against the FP64 oracle.
#include "bh_device.hpp"
#include "bh_kernels.hpp"
#include "bh_mlp.hpp"

namespace bh {

// ------------------------------------------------------------------ node records
// The counts the record kernels work on: one simulation's from its device
// values (bodies alive, internal nodes = alive - 1), or an ensemble's given
// on the host (N bodies, all alive; N - S internal nodes).
struct RecScope {
    const SimValues* sim;
    uint32_t n, num, ninternal;
    float G;
    __device__ __forceinline__ uint32_t n_ids() const { return sim ? sim->initialNumBodies : n; }
    __device__ __forceinline__ uint32_t alive() const { return sim ? sim->numBodies : num; }
    __device__ __forceinline__ uint32_t internal() const {
        if (!sim) return ninternal;
        const uint32_t m = sim->numBodies;
        return m > 1 ? m - 1 : 0;
    }
    __device__ __forceinline__ float grav() const { return sim ? sim->gravitationalConstant : G; }
};

// The corrected walk's records (layout with corr_records; the tree's first
// launch wrote the first half of every leaf's): the whole record of every
// internal node, T = M2 + M h^2 o with o the network's output for the node's
// features (o = 0 without a network), children renumbered. Each wave takes 64
// consecutive nodes per turn; the grid strides over the nodes so that every
// wave loads the network's fragments once.
__global__ __launch_bounds__(256) void k_qnodes(const RecScope sc, DevCounters* __restrict__ ctr,
                                                const WalkNode* __restrict__ walk,
                                                const float4* __restrict__ node_aabb,
                                                const float4* __restrict__ node_mom, const __bf16* __restrict__ wg,
                                                const float* __restrict__ bg, float4* __restrict__ qrec) {
    __shared__ __attribute__((aligned(16))) __bf16 s_x[4][64 * kMlpXStride];
    __shared__ __attribute__((aligned(16))) float s_y[4][64 * kMlpYStride];
    const uint32_t w = threadIdx.x >> 6, lane = lane_id();
    const uint32_t n = sc.n_ids();
    const uint32_t ninternal = sc.internal();
    const float G = sc.grav();
    const WalkNode* __restrict__ nodes = walk + n;
    const uint32_t waves = gridDim.x * 4;
    MlpRegs R;
    if (wg) mlp_load(R, wg, bg);
    uint32_t evals = 0;
    for (uint32_t base = (blockIdx.x * 4 + w) * 64; base < ninternal; base += waves * 64) {
        const uint32_t k = base + lane;
        const bool active = k < ninternal;
        double m2[6] = {0, 0, 0, 0, 0, 0};
        float M = 0.0f, h = 0.0f;
        float4 mn = make_float4(0.f, 0.f, 0.f, 0.f), mx = mn, com = mn;
        WalkNode nd;
        if (active) {
            nd = nodes[k];
            com = nd.comMass;
            h = nd.half;
            mn = node_aabb[2 * k];
            mx = node_aabb[2 * k + 1];
            M = mx.w;
            const float4 a = node_mom[2 * k], b = node_mom[2 * k + 1];
            m2[0] = a.x; m2[1] = a.y; m2[2] = a.z; m2[3] = a.w; m2[4] = b.x; m2[5] = b.y;
        }
        const bool ok = active && M > 0.0f && h > 0.0f;
        double o[6] = {0, 0, 0, 0, 0, 0};
        if (wg && ballot(ok)) {
            // every lane takes part in the MFMAs; lanes without a node stage zeros
            float f[kMlpIn];
            double nm[6];
            const double mh2 = ok ? (double)M * (double)h * (double)h : 1.0;
            for (int c = 0; c < 6; c++) nm[c] = ok ? m2[c] / mh2 : 0.0;
            if (ok) node_features(mn, mx, com.x, com.y, com.z, h, nm, f);
            else
                for (int c = 0; c < kMlpIn; c++) f[c] = 0.0f;
            mlp_stage_features(s_x[w], f);
            mlp_wave_forward(R, s_x[w], s_y[w], 0xFu);
            if (ok)
                for (int c = 0; c < 6; c++) o[c] = (double)s_y[w][lane * kMlpYStride + c] * mh2;
            evals += (uint32_t)__popcll(ballot(ok));
        }
        if (active) {
            float T[6];
            for (int c = 0; c < 6; c++) T[c] = (float)(m2[c] + o[c]);
            const float tr = T[0] + T[3] + T[5];
            const uint32_t a = nd.childA >= n ? nd.childA + kCorrGap : nd.childA;
            const uint32_t b = nd.childB >= n ? nd.childB + kCorrGap : nd.childB;
            float4* rec = qrec + 4 * (size_t)corr_internal_id(n, k);
            rec[0] = nd.comMass;
            const float b2 = corr_size2(com, mn, mx);
            rec[1] = make_float4(corr_rel_k(nd.comMass.w, b2),
                                 __uint_as_float(corr_children_word(a, nd.childA >= n, nd.childB >= n)),
                                 __uint_as_float(b), b2);
            const float g3 = -3.0f * G;
            rec[2] = make_float4(g3 * T[0], g3 * T[1], g3 * T[2], g3 * T[3]);
            rec[3] = make_float4(g3 * T[4], g3 * T[5], -1.5f * (G * tr), 0.0f);
        }
    }
    if (evals && lane == 0) atomicAdd(&ctr->corrected, (unsigned long long)evals);
}

static void node_quadrupoles(hipStream_t s, uint32_t n, const RecScope& sc, DevCounters* ctr, const TreeArrays& T,
                             const __bf16* w, const float* bias, float4* qrec) {
    const uint32_t groups = (n + 63) / 64;
    const uint32_t blocks = std::min<uint32_t>((groups + 3) / 4, 2048u);
    k_qnodes<<<blocks, 256, 0, s>>>(sc, ctr, T.walk, T.node_aabb, T.node_mom, w, bias, qrec);
}

void launch_node_quadrupoles(hipStream_t s, uint32_t n, const SimValues* sim, DevCounters* ctr, const TreeArrays& T,
                             const __bf16* w, const float* bias, float4* qrec) {
    const RecScope sc = {sim, 0u, 0u, 0u, 0.0f};
    node_quadrupoles(s, n, sc, ctr, T, w, bias, qrec);
}

void launch_ens_node_quadrupoles(hipStream_t s, uint32_t N, uint32_t n_internal, float G, DevCounters* ctr,
                                 const TreeArrays& T, const __bf16* w, const float* bias, float4* qrec) {
    const RecScope sc = {nullptr, N, N, n_internal, G};
    node_quadrupoles(s, N, sc, ctr, T, w, bias, qrec);
}

// ------------------------------------------------------------------ pairs
// The pairs of the corrected walk: every stride-th sorted body walks with the
// corrected walk's opening test at theta (corr_size2) and records each
// internal node it accepts.
__global__ __launch_bounds__(256) void k_collect_pairs(const SimValues* __restrict__ sim,
                                                       const float4* __restrict__ leaf_pos,
                                                       const WalkNode* __restrict__ nodes,
                                                       const float4* __restrict__ node_aabb, float theta, float soft,
                                                       uint32_t stride, uint2* __restrict__ pairs, uint32_t cap,
                                                       uint32_t* __restrict__ count) {
    const uint32_t n = sim->initialNumBodies, num = sim->numBodies;
    const uint32_t g = (blockIdx.x * blockDim.x + threadIdx.x) * stride;
    if (g >= num || num < 2) return;
    const float4 pm = leaf_pos[g];
    const float th2 = theta * theta;
    uint32_t stack[kStackDepth];
    int sp = 0;
    stack[sp++] = n;
    while (sp > 0) {
        const uint32_t node = stack[--sp];
        if (node < n) continue;
        const WalkNode nd = nodes[node - n];
        const float rx = nd.comMass.x - pm.x, ry = nd.comMass.y - pm.y, rz = nd.comMass.z - pm.z;
        const float b2 = corr_size2(nd.comMass, node_aabb[2 * (node - n)], node_aabb[2 * (node - n) + 1]);
        if (corr_dist2(rx, ry, rz, soft) * th2 > b2) {
            uint32_t slot = atomicAdd(count, 1u);
            if (slot < cap) pairs[slot] = make_uint2(g, node - n);
        } else if (sp + 2 <= kStackDepth) {
            stack[sp++] = nd.childA;
            stack[sp++] = nd.childB;
        } else {
            atomicOr(count + 1, 1u);   // reported: the pairs of this walk are incomplete
            break;
        }
    }
}

// One wave per pair (grid-stride over the pairs): in FP64 over the node's
// bodies (sorted positions [lo, lo + count)), the exact field on the target
// (itself excluded) and the node's second moment about its centre of mass;
// the label is the exact field minus the monopole and the quadrupole of that
// moment, in units of unit = (h / s)^2 G M / s^2, with u = r / s:
// targ = (u, label, unit). (The host turns the unit into the row's weight.)
__global__ __launch_bounds__(256) void k_pair_labels(const uint2* __restrict__ pairs, uint32_t npairs,
                                                     const float4* __restrict__ leaf_pos,
                                                     const WalkNode* __restrict__ nodes,
                                                     const float4* __restrict__ node_aabb,
                                                     const uint2* __restrict__ node_range, double G, float soft,
                                                     uint32_t max_bodies, float* __restrict__ feat,
                                                     float* __restrict__ targ) {
    const uint32_t lane = lane_id();
    const uint32_t waves = gridDim.x * (blockDim.x >> 6);
    for (uint32_t p = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); p < npairs; p += waves) {
        const uint32_t g = pairs[p].x, k = pairs[p].y;
        const float4 x = leaf_pos[g];
        const WalkNode nd = nodes[k];
        const uint2 rg = node_range[k];
        // Too large to label exactly, or a node that holds the target itself
        // (its monopole then counts the target's own mass, a residual of the
        // pair, not of the node's far field): marked, dropped by the host.
        if (rg.y > max_bodies || (g >= rg.x && g < rg.x + rg.y)) {
            if (lane == 0)
                for (int c = 0; c < kMlpTarget; c++) targ[(size_t)kMlpTarget * p + c] = __builtin_nanf("");
            continue;
        }
        const double cx = nd.comMass.x, cy = nd.comMass.y, cz = nd.comMass.z;
        double a[3] = {0, 0, 0}, m2[6] = {0, 0, 0, 0, 0, 0};
        for (uint32_t j = rg.x + lane; j < rg.x + rg.y; j += kWave) {
            const float4 b = leaf_pos[j];
            const double dx = (double)b.x - cx, dy = (double)b.y - cy, dz = (double)b.z - cz, m = b.w;
            m2[0] += m * dx * dx;
            m2[1] += m * dx * dy;
            m2[2] += m * dx * dz;
            m2[3] += m * dy * dy;
            m2[4] += m * dy * dz;
            m2[5] += m * dz * dz;
            if (j == g) continue;
            const double rx = (double)b.x - x.x, ry = (double)b.y - x.y, rz = (double)b.z - x.z;
            const double d2 = rx * rx + ry * ry + rz * rz + (double)soft;
            const double inv = 1.0 / sqrt(d2);
            const double f = G * m * inv * inv * inv;
            a[0] += f * rx;
            a[1] += f * ry;
            a[2] += f * rz;
        }
        for (int off = 32; off > 0; off >>= 1) {
            for (int c = 0; c < 3; c++) a[c] += __shfl_xor(a[c], off, 64);
            for (int c = 0; c < 6; c++) m2[c] += __shfl_xor(m2[c], off, 64);
        }
        if (lane != 0) continue;
        const double r[3] = {cx - x.x, cy - x.y, cz - x.z};
        const double s2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2] + (double)soft;
        const double s = sqrt(s2);
        const double M = node_aabb[2 * k + 1].w;                  // (the record holds G M)
        const double h = nd.half;
        // monopole and exact quadrupole, double precision
        const double i2 = 1.0 / s2, i3 = i2 / s, i5 = i3 * i2;
        const double tr[3] = {m2[0] * r[0] + m2[1] * r[1] + m2[2] * r[2], m2[1] * r[0] + m2[3] * r[1] + m2[4] * r[2],
                              m2[2] * r[0] + m2[4] * r[1] + m2[5] * r[2]};
        const double rtr = r[0] * tr[0] + r[1] * tr[1] + r[2] * tr[2];
        const double c = (7.5 * rtr * i2 - 1.5 * (m2[0] + m2[3] + m2[5])) * i5;
        const double unit = G * M / s2 * (h * h / s2);            // (h / s)^2 G M / s^2
        float* t = targ + (size_t)kMlpTarget * p;
        for (int q = 0; q < 3; q++) {
            const double mono = G * M * r[q] * i3;
            const double quad = G * (c * r[q] - 3.0 * tr[q] * i5);
            t[q] = (float)(r[q] / s);
            t[3 + q] = (float)((a[q] - mono - quad) / unit);
        }
        t[6] = (float)unit;
        double nm[6];
        const double mh2 = M * h * h;
        for (int q = 0; q < 6; q++) nm[q] = m2[q] / mh2;
        node_features(node_aabb[2 * k], node_aabb[2 * k + 1], nd.comMass.x, nd.comMass.y, nd.comMass.z, nd.half, nm,
                      feat + (size_t)kMlpIn * p);
    }
}

// ------------------------------------------------------------------ FP32 training kernels
// C[M][N] = sum_k A(m,k) B(k,n) (+ C if accumulate); A(m,k) = A[m*sam + k*sak],
// B(k,n) = B[k*sbk + n*sbn] (strides express the transposes). 16x16 tiles.
__global__ __launch_bounds__(256) void k_gemm(const float* __restrict__ A, int sam, int sak,
                                              const float* __restrict__ B, int sbk, int sbn,
                                              float* __restrict__ C, int M, int N, int K, int accumulate) {
    __shared__ float sA[16][17], sB[16][17];
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    const int m = blockIdx.y * 16 + ty, nn = blockIdx.x * 16 + tx;
    float acc = 0.0f;
    for (int k0 = 0; k0 < K; k0 += 16) {
        int ka = k0 + tx, kb = k0 + ty;
        sA[ty][tx] = (m < M && ka < K) ? A[(size_t)m * sam + (size_t)ka * sak] : 0.0f;
        int nb = blockIdx.x * 16 + tx;
        sB[ty][tx] = (kb < K && nb < N) ? B[(size_t)kb * sbk + (size_t)nb * sbn] : 0.0f;
        __syncthreads();
        for (int kk = 0; kk < 16; kk++) acc += sA[ty][kk] * sB[kk][tx];
        __syncthreads();
    }
    if (m < M && nn < N) {
        float* c = C + (size_t)m * N + nn;
        *c = accumulate ? *c + acc : acc;
    }
}

// The same product with the K range cut into gridDim.z slices: slice z
// writes its partial sums to ws[z][M][N] (the weight gradients have K = the
// batch and only M x N <= 64 x 64 outputs, too few blocks for the plain one)
__global__ __launch_bounds__(256) void k_gemm_splitk(const float* __restrict__ A, int sam, int sak,
                                                     const float* __restrict__ B, int sbk, int sbn,
                                                     float* __restrict__ ws, int M, int N, int K, int chunk) {
    __shared__ float sA[16][17], sB[16][17];
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    const int m = blockIdx.y * 16 + ty, nn = blockIdx.x * 16 + tx;
    const int k_begin = blockIdx.z * chunk, k_end = min(K, k_begin + chunk);
    float acc = 0.0f;
    for (int k0 = k_begin; k0 < k_end; k0 += 16) {
        int ka = k0 + tx, kb = k0 + ty;
        sA[ty][tx] = (m < M && ka < k_end) ? A[(size_t)m * sam + (size_t)ka * sak] : 0.0f;
        int nb = blockIdx.x * 16 + tx;
        sB[ty][tx] = (kb < k_end && nb < N) ? B[(size_t)kb * sbk + (size_t)nb * sbn] : 0.0f;
        __syncthreads();
        for (int kk = 0; kk < 16; kk++) acc += sA[ty][kk] * sB[kk][tx];
        __syncthreads();
    }
    if (m < M && nn < N) ws[((size_t)blockIdx.z * M + m) * N + nn] = acc;
}

// C = (accumulate ? C : 0) + the slices' partial sums, added in slice order
__global__ void k_splitk_reduce(const float* __restrict__ ws, float* __restrict__ C, int MN, int slices,
                                int accumulate) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= MN) return;
    float acc = accumulate ? C[i] : 0.0f;
    for (int z = 0; z < slices; z++) acc += ws[(size_t)z * MN + i];
    C[i] = acc;
}

// Z = Z + bias (row broadcast); H = SiLU(Z) (if H)
__global__ void k_bias_act(float* __restrict__ Z, float* __restrict__ H, const float* __restrict__ bias, int M,
                           int N) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)M * N) return;
    float z = Z[i] + bias[i % N];
    Z[i] = z;
    if (H) H[i] = z / (1.0f + expf(-z));
}

// Loss through the quadrupole basis: row m has network outputs Y[m][6] and
// target T[m] = (u (3), label (3), weight w); the prediction is B(u) Y[m],
// the loss w times the squared error summed over its 3 components.
// dY = 2 w B(u)^T (B(u) Y - label) / batch; per-block partial sums of the loss.
__global__ void k_loss_grad(const float* __restrict__ Y, const float* __restrict__ T, float* __restrict__ dY,
                            int M, float* __restrict__ loss_partial) {
    __shared__ float s[256];
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    float e2 = 0.0f;
    if (m < M) {
        const float* t = T + (size_t)m * kMlpTarget;
        float o[6], pr[3], e[3], g[6];
        for (int c = 0; c < 6; c++) o[c] = Y[(size_t)m * kMlpOut + c];
        quad_basis(t, o, pr);
        const float wgt = t[6];
        for (int c = 0; c < 3; c++) {
            e[c] = pr[c] - t[3 + c];
            e2 += wgt * e[c] * e[c];
        }
        quad_basis_t(t, e, g);
        for (int c = 0; c < 6; c++) dY[(size_t)m * kMlpOut + c] = 2.0f * wgt * g[c] / (float)M;
    }
    s[threadIdx.x] = e2;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) loss_partial[blockIdx.x] = s[0];
}

// dZ = dH * silu'(Z)
__global__ void k_act_grad(const float* __restrict__ Z, float* __restrict__ dH, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float z = Z[i];
    float sg = 1.0f / (1.0f + expf(-z));
    dH[i] = dH[i] * sg * (1.0f + z * (1.0f - sg));
}

// db[n] = sum over rows of dZ[m][n]
__global__ void k_colsum(const float* __restrict__ dZ, float* __restrict__ db, int M, int N) {
    int nn = blockIdx.x;
    float s = 0.0f;
    for (int m = threadIdx.x; m < M; m += blockDim.x) s += dZ[(size_t)m * N + nn];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    __shared__ float w[4];
    if (lane_id() == 0) w[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) db[nn] = (w[0] + w[1]) + (w[2] + w[3]);
}

__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m1,
                       float* __restrict__ m2, int n, float lr, float b1, float b2, float eps, float c1, float c2) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float gi = g[i];
    float a = b1 * m1[i] + (1.0f - b1) * gi;
    float b = b2 * m2[i] + (1.0f - b2) * gi * gi;
    m1[i] = a;
    m2[i] = b;
    p[i] -= lr * (a / c1) / (sqrtf(b / c2) + eps);
}

// FP32 master parameters -> BF16 MFMA weights (padded, zero elsewhere) + FP32 bias block
// FP32 master parameters -> BF16 weight fragments (layout in bh_mlp.hpp) and
// the padded FP32 biases.
__global__ void k_pack_weights(const float* __restrict__ p, __bf16* __restrict__ w, float* __restrict__ bias) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < kBfTotal) {
        const int f = i >> 9, lane = (i >> 3) & 63, j = i & 7;
        const int r16 = lane & 15, grp = lane >> 4;
        float v = 0.0f;
        if (f < kFragW2) {
            const int k = 8 * grp + j;
            if (k < kMlpIn) v = p[kOffW1 + (16 * f + r16) * kMlpIn + k];
        } else if (f < kFragW3) {
            const int mo = (f - kFragW2) >> 1, sl = (f - kFragW2) & 1;
            v = p[kOffW2 + (16 * mo + r16) * kMlpHidden + mlp_hid(sl, grp, j)];
        } else {
            const int sl = f - kFragW3;
            if (r16 < kMlpOut) v = p[kOffW3 + r16 * kMlpHidden + mlp_hid(sl, grp, j)];
        }
        // log2(e) into the hidden pre-activations, ln(2) out of the last layer's inputs
        const float scale = f < kFragW2 ? 1.44269504f : (f < kFragW3 ? 1.0f : 0.693147181f);
        w[i] = (__bf16)(v * scale);
    }
    if (i < kBiasTotal) {
        float v = 0.0f;
        if (i < 64) v = p[kOffB1 + i] * 1.44269504f;
        else if (i < 128) v = p[kOffB2 + i - 64] * 1.44269504f;
        else if (i - 128 < kMlpOut) v = p[kOffB3 + i - 128];
        bias[i] = v;
    }
}

// ------------------------------------------------------------------ MFMA inference
// 64 samples per wave, 4 waves per block; x is [n][12] FP32, y is [n][6].
__global__ __launch_bounds__(256) void k_mlp_forward_mfma(const float* __restrict__ x, uint32_t n,
                                                          const __bf16* __restrict__ wg,
                                                          const float* __restrict__ bg, float* __restrict__ y) {
    __shared__ __attribute__((aligned(16))) __bf16 s_x[4][64 * kMlpXStride];
    __shared__ __attribute__((aligned(16))) float s_y[4][64 * kMlpYStride];
    const int w = threadIdx.x >> 6, lane = (int)lane_id();
    MlpRegs R;
    mlp_load(R, wg, bg);
    const uint32_t base = (blockIdx.x * 4 + w) * 64;
    if (base >= n) return;
    const uint32_t row = base + lane;
    float f[kMlpIn];
    for (int k = 0; k < kMlpIn; k++) f[k] = row < n ? x[(size_t)row * kMlpIn + k] : 0.0f;
    mlp_stage_features(s_x[w], f);
    const uint32_t live = n - base;   // rows of this wave in range
    const uint32_t blocks = live >= 64 ? 0xFu : (1u << ((live + 15) / 16)) - 1u;
    mlp_wave_forward(R, s_x[w], s_y[w], blocks);
    if (row < n)
        for (int j = 0; j < kMlpOut; j++) y[(size_t)row * kMlpOut + j] = s_y[w][lane * kMlpYStride + j];
}

// ------------------------------------------------------------------ launchers
static inline uint32_t cdiv2(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

void launch_collect_pairs(hipStream_t s, uint32_t n, const SimValues* sim, const float4* leaf_pos,
                          const WalkNode* nodes, const float4* node_aabb, float theta, float soft, uint32_t stride,
                          uint2* pairs, uint32_t cap, uint32_t* count) {
    uint32_t threads = (n + stride - 1) / stride;
    k_collect_pairs<<<cdiv2(threads, 256), 256, 0, s>>>(sim, leaf_pos, nodes, node_aabb, theta, soft, stride, pairs,
                                                        cap, count);
}
void launch_pair_labels(hipStream_t s, const uint2* pairs, uint32_t npairs, const float4* leaf_pos,
                        const WalkNode* nodes, const float4* node_aabb, const uint2* node_range, double G,
                        float soft, uint32_t max_bodies, float* feat, float* targ) {
    // 4 waves per block, at most 2^16 blocks: the waves stride over the pairs
    const uint32_t blocks = npairs < (1u << 18) ? cdiv2(npairs, 4) : (1u << 16);
    if (npairs) k_pair_labels<<<blocks, 256, 0, s>>>(pairs, npairs, leaf_pos, nodes, node_aabb, node_range, G, soft,
                                                     max_bodies, feat, targ);
}
// rows of x (n x cols) gathered through perm[offset .. offset + rows)
__global__ void k_gather_rows(const float* __restrict__ src, const uint32_t* __restrict__ perm, uint32_t offset,
                              uint32_t rows, int cols, float* __restrict__ dst) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)rows * cols) return;
    uint32_t r = (uint32_t)(i / cols);
    int c = (int)(i % cols);
    dst[i] = src[(size_t)perm[offset + r] * cols + c];
}
void launch_gather_rows(hipStream_t s, const float* src, const uint32_t* perm, uint32_t offset, uint32_t rows,
                        int cols, float* dst) {
    if (rows) k_gather_rows<<<cdiv2((size_t)rows * cols, 256), 256, 0, s>>>(src, perm, offset, rows, cols, dst);
}
void launch_gemm(hipStream_t s, const float* A, int sam, int sak, const float* B, int sbk, int sbn, float* C, int M,
                 int N, int K, int accumulate, float* ws, uint32_t ws_floats) {
    dim3 grid(cdiv2(N, 16), cdiv2(M, 16));
    // few output tiles and a long K: split K over up to kGemmSlices slices
    int slices = (K + kGemmSliceK - 1) / kGemmSliceK;
    if (slices > kGemmSlices) slices = kGemmSlices;
    if (ws && grid.x * grid.y < 64 && slices > 1 && (size_t)slices * M * N <= ws_floats) {
        const int chunk = ((K + slices - 1) / slices + 15) / 16 * 16;
        slices = (K + chunk - 1) / chunk;
        grid.z = slices;
        k_gemm_splitk<<<grid, 256, 0, s>>>(A, sam, sak, B, sbk, sbn, ws, M, N, K, chunk);
        k_splitk_reduce<<<cdiv2((size_t)M * N, 256), 256, 0, s>>>(ws, C, M * N, slices, accumulate);
        return;
    }
    k_gemm<<<grid, 256, 0, s>>>(A, sam, sak, B, sbk, sbn, C, M, N, K, accumulate);
}
void launch_bias_act(hipStream_t s, float* Z, float* H, const float* bias, int M, int N) {
    k_bias_act<<<cdiv2((size_t)M * N, 256), 256, 0, s>>>(Z, H, bias, M, N);
}
void launch_loss_grad(hipStream_t s, const float* Y, const float* T, float* dY, int M, float* partial,
                      uint32_t* nblocks) {
    *nblocks = cdiv2((size_t)M, 256);
    k_loss_grad<<<*nblocks, 256, 0, s>>>(Y, T, dY, M, partial);
}
void launch_act_grad(hipStream_t s, const float* Z, float* dH, size_t n) {
    k_act_grad<<<cdiv2(n, 256), 256, 0, s>>>(Z, dH, n);
}
void launch_colsum(hipStream_t s, const float* dZ, float* db, int M, int N) {
    k_colsum<<<N, 256, 0, s>>>(dZ, db, M, N);
}
void launch_adam(hipStream_t s, float* p, const float* g, float* m1, float* m2, int n, float lr, int t) {
    const float b1 = 0.9f, b2 = 0.999f;
    float c1 = 1.0f - powf(b1, (float)t), c2 = 1.0f - powf(b2, (float)t);
    k_adam<<<cdiv2(n, 256), 256, 0, s>>>(p, g, m1, m2, n, lr, b1, b2, 1e-8f, c1, c2);
}
void launch_pack_weights(hipStream_t s, const float* p, __bf16* w, float* bias) {
    k_pack_weights<<<cdiv2(kBfTotal, 256), 256, 0, s>>>(p, w, bias);
}
void launch_mlp_forward_mfma(hipStream_t s, const float* x, uint32_t n, const __bf16* w, const float* bias, float* y) {
    if (n) k_mlp_forward_mfma<<<cdiv2(n, 256), 256, 0, s>>>(x, n, w, bias, y);
}

}  // namespace bh
