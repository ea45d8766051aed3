// bh_walk.hip — tree walk, force and integration (shaders/compute/force/bh_force.comp:6-142).
//
// One wavefront walks the tree for 64 consecutive sorted bodies with one
// shared LIFO stack in LDS. Each stack entry carries the 64-bit mask of the
// lanes that still need that node; a node is loaded once per wave (its index
// is wave-uniform, so the 32-byte record comes in through scalar loads), every
// lane in the mask evaluates its own acceptance test, and the children are
// pushed with the mask of the lanes that opened it. Restricted to one lane,
// this visits exactly the nodes of that body's own per-thread walk in exactly
// the reference's order (push childA, push childB, pop childB first: here
// childA is pushed and childB visited directly), so the
// accepted set, the opened-node count and the FP32 sums are the reference's.
#include "bh_device.hpp"
#include "bh_kernels.hpp"

namespace bh {

__device__ __forceinline__ float body_radius(float m, float rho, float blc) {
    // common/common.glsl:228-230
    return fmul(blc, powf(fdiv(m, rho), 1.0f / 3.0f));
}

// Pointers are separate __restrict__ parameters so that the compiler can
// prove the tree is not written during the kernel and fetch node records with
// scalar loads.
__global__ __launch_bounds__(256) void k_walk(SimValues* __restrict__ sim, DevCounters* __restrict__ ctr,
                                              const float4* __restrict__ src, float4* __restrict__ dst,
                                              const uint32_t* __restrict__ index_in,
                                              const float4* __restrict__ leaf_pos,
                                              const WalkNode* __restrict__ nodes, uint2* __restrict__ tasks,
                                              uint32_t task_cap, float theta_, float dt_, float softening,
                                              uint32_t merging_mode, uint32_t dynamic, uint32_t wrap) {
    __shared__ uint32_t s_node[kWalkWaves][kStackDepth];
    __shared__ unsigned long long s_mask[kWalkWaves][kStackDepth];

    const uint32_t n = sim->initialNumBodies;
    const uint32_t num = sim->numBodies;
    const uint32_t w = threadIdx.x >> 6, lane = lane_id();
    const uint32_t lb = xcd_block(blockIdx.x, gridDim.x);
    const uint32_t base = (lb * kWalkWaves + w) * kWave;
    if (base >= num) return;
    const uint32_t g = base + lane;
    const bool active = g < num;
    const uint32_t bodyIdx = active ? index_in[g] : 0;
    const float4 pm = active ? leaf_pos[g] : make_float4(0.f, 0.f, 0.f, 0.f);
    const float4 vd = active ? src[2 * bodyIdx + 1] : make_float4(0.f, 0.f, 0.f, 1.f);
    const float G = sim->gravitationalConstant;
    const float blc = sim->bodyLengthConstant;
    const float theta = theta_, soft = softening;
    const bool merging = merging_mode == 1u;
    const float Ri = (merging && active) ? body_radius(pm.w, vd.w, blc) : 0.0f;

    float ax = 0.0f, ay = 0.0f, az = 0.0f;
    unsigned long long opened = 0, inter = 0;
    uint32_t maxsp = 1, err = 0;

    if (num > 1) {
        // current node in scalar registers; the stack holds the pending
        // childA entries (after opening a node its childB, the top of the
        // reference's stack, is visited next without a round trip through LDS)
        uint32_t node = n;
        unsigned long long mask = ballot(active);
        int sp = 0;
        for (;;) {
            const bool mine = (mask >> lane) & 1ull;
            bool descend = false;
            if (node >= n) {
                const WalkNode nd = nodes[node - n];
                const float rx = fsub(nd.comMass.x, pm.x);
                const float ry = fsub(nd.comMass.y, pm.y);
                const float rz = fsub(nd.comMass.z, pm.z);
                const float d2 = fadd(dot3(rx, ry, rz), soft);
                const float inv = inv_sqrt_exact(d2);
                const bool accept = fmul(nd.half, inv) < theta;
                const unsigned long long open = ballot(mine && !accept);
                if (mine && accept) {
                    const float gm = fmul(G, nd.comMass.w);
                    ax = fadd(ax, fmul(fmul(fmul(fmul(gm, rx), inv), inv), inv));
                    ay = fadd(ay, fmul(fmul(fmul(fmul(gm, ry), inv), inv), inv));
                    az = fadd(az, fmul(fmul(fmul(fmul(gm, rz), inv), inv), inv));
                }
                inter += (unsigned long long)__popcll(mask & ~open);
                if (open) {
                    if (sp + 1 > kStackDepth) {
                        err |= 1u;
                        break;
                    }
                    if (lane == 0) {
                        s_node[w][sp] = nd.childA;
                        s_mask[w][sp] = open;
                    }
                    sp += 1;
                    opened += (unsigned long long)__popcll(open);
                    if ((uint32_t)sp + 1 > maxsp) maxsp = (uint32_t)sp + 1;
                    node = uniform(nd.childB);
                    mask = open;
                    descend = true;
                }
            } else {
                const float4 lp = leaf_pos[node];
                const bool use = mine && node != g;          // bh_force.comp:77-79 (self skip)
                const float rx = fsub(lp.x, pm.x);
                const float ry = fsub(lp.y, pm.y);
                const float rz = fsub(lp.z, pm.z);
                const float d2 = fadd(dot3(rx, ry, rz), soft);
                const float inv = inv_sqrt_exact(d2);
                if (use) {
                    const float gm = fmul(G, lp.w);
                    ax = fadd(ax, fmul(fmul(fmul(fmul(gm, rx), inv), inv), inv));
                    ay = fadd(ay, fmul(fmul(fmul(fmul(gm, ry), inv), inv), inv));
                    az = fadd(az, fmul(fmul(fmul(fmul(gm, rz), inv), inv), inv));
                }
                const unsigned long long selfbit =
                    (node >= base && node < base + kWave) ? (1ull << (node - base)) : 0ull;
                inter += (unsigned long long)__popcll(mask & ~selfbit);
                if (merging && (mask & ~selfbit)) {
                    // bh_force.comp:85-112: overlap test against the leaf's body,
                    // merge task from the lower-index body
                    const uint32_t other = uniform(index_in[node]);
                    const float rho = src[2 * other + 1].w;
                    const float Rj = body_radius(lp.w, rho, blc);
                    const float dist = fsqrt(dot3(rx, ry, rz));
                    const float pen = fsub(fadd(Ri, Rj), dist);
                    if (use && pen > 0.0f && bodyIdx < other) {
                        uint32_t slot = atomicAdd(&ctr->nTasks, 1u);
                        if (slot < task_cap) tasks[slot] = make_uint2(bodyIdx, other);
                        else err |= 2u;
                    }
                }
            }
            if (!descend) {
                if (sp == 0) break;
                sp--;
                node = uniform(s_node[w][sp]);
                const unsigned long long mraw = s_mask[w][sp];
                mask = ((unsigned long long)uniform((uint32_t)(mraw >> 32)) << 32) | uniform((uint32_t)mraw);
            }
        }
    }

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
        if (kill) {
            dst[2 * bodyIdx] = make_float4(0.f, 0.f, 0.f, 0.f);
            dst[2 * bodyIdx + 1] = make_float4(0.f, 0.f, 0.f, 0.f);
            oob = 1;
        } else {
            dst[2 * bodyIdx] = make_float4(npx, npy, npz, pm.w);
            dst[2 * bodyIdx + 1] = make_float4(nvx, nvy, nvz, vd.w);
        }
    }
    // one set of atomics per wave
    const uint32_t oob_w = __popcll(ballot(oob != 0));
    const uint32_t err_w = (ballot(err & 1u) ? 1u : 0u) | (ballot(err & 2u) ? 2u : 0u);
    if (lane == 0) {
        atomicAdd(&ctr->opened, opened);
        atomicAdd(&ctr->interactions, inter);
        atomicMax(&ctr->maxStack, maxsp);
        if (oob_w) atomicAdd(&sim->outOfBounds, oob_w);
        if (err_w) atomicOr(&ctr->errorFlags, err_w);
        // sim.uintDebug[0] counts opened nodes as a 32-bit word (bh_force.comp:72)
        atomicAdd(&sim->uintDebug[0], (uint32_t)opened);
    }
}

void launch_walk(hipStream_t s, uint32_t n, const WalkArgs& a) {
    const uint32_t per_block = kWalkWaves * kWave;
    const uint32_t nb = (n + per_block - 1) / per_block;
    k_walk<<<nb, kWalkWaves * kWave, 0, s>>>(const_cast<SimValues*>(a.sim), a.ctr, a.src, a.dst, a.index_in,
                                              a.leaf_pos, a.nodes, a.tasks, a.task_cap, a.theta, a.dt,
                                              a.softening, a.merging, a.dynamic, a.wrap);
}

}  // namespace bh

// ---------------------------------------------------------------------------
// The corrected walk (no reference counterpart): the same masked-stack walk
// with a wide opening angle `theta`, where a node is accepted only if the
// target also lies outside the node's box (a node that contains its target
// has a residual the network cannot know from the node summary); an internal
// node accepted at `theta` that the plain angle `theta_plain` would have
// opened gets, besides its monopole, the network's residual:
// a += G M / d^2 * MLP(features). The wave stages the 64 lanes' features in
// LDS and evaluates them together with BF16 MFMA (FP32 accumulation).
#include "bh_mlp.hpp"

namespace bh {

__global__ __launch_bounds__(256) void k_walk_learned(SimValues* __restrict__ sim, DevCounters* __restrict__ ctr,
                                                      const float4* __restrict__ src, float4* __restrict__ dst,
                                                      const uint32_t* __restrict__ index_in,
                                                      const float4* __restrict__ leaf_pos,
                                                      const WalkNode* __restrict__ nodes,
                                                      const float4* __restrict__ node_aabb,
                                                      const uint2* __restrict__ node_range,
                                                      const __bf16* __restrict__ mlp_w,
                                                      const float* __restrict__ mlp_b, float theta_, float theta_plain,
                                                      float dt_, float softening) {
    __shared__ uint32_t s_node[kWalkWaves][kStackDepth];
    __shared__ unsigned long long s_mask[kWalkWaves][kStackDepth];
    __shared__ __attribute__((aligned(16))) __bf16 s_x[kWalkWaves][64 * kMlpXStride];
    __shared__ __attribute__((aligned(16))) float s_y[kWalkWaves][64 * 4];

    MlpRegs R;
    mlp_load(R, mlp_w, mlp_b);

    const uint32_t n = sim->initialNumBodies;
    const uint32_t num = sim->numBodies;
    const uint32_t w = threadIdx.x >> 6, lane = lane_id();
    const uint32_t lb = xcd_block(blockIdx.x, gridDim.x);
    const uint32_t base = (lb * kWalkWaves + w) * kWave;
    if (base >= num) return;
    const uint32_t g = base + lane;
    const bool active = g < num;
    const uint32_t bodyIdx = active ? index_in[g] : 0;
    const float4 pm = active ? leaf_pos[g] : make_float4(0.f, 0.f, 0.f, 0.f);
    const float4 vd = active ? src[2 * bodyIdx + 1] : make_float4(0.f, 0.f, 0.f, 1.f);
    const float G = sim->gravitationalConstant;
    const float theta = theta_, soft = softening;

    float ax = 0.0f, ay = 0.0f, az = 0.0f;
    unsigned long long opened = 0, inter = 0, corrected = 0;
    uint32_t maxsp = 1, err = 0;

    if (num > 1) {
        uint32_t node = n;
        unsigned long long mask = ballot(active);
        int sp = 0;
        for (;;) {
            const bool mine = (mask >> lane) & 1ull;
            bool descend = false;
            if (node >= n) {
                const WalkNode nd = nodes[node - n];
                const float rx = fsub(nd.comMass.x, pm.x);
                const float ry = fsub(nd.comMass.y, pm.y);
                const float rz = fsub(nd.comMass.z, pm.z);
                const float d2 = fadd(dot3(rx, ry, rz), soft);
                const float inv = inv_sqrt_exact(d2);
                const float ratio = fmul(nd.half, inv);
                const uint32_t k = node - n;
                const float4 mn = node_aabb[2 * k], mx = node_aabb[2 * k + 1];
                const bool inside = pm.x >= mn.x && pm.x <= mx.x && pm.y >= mn.y && pm.y <= mx.y && pm.z >= mn.z &&
                                    pm.z <= mx.z;
                const bool accept = ratio < theta && !inside;
                const bool correct = accept && ratio >= theta_plain;
                const unsigned long long open = ballot(mine && !accept);
                const unsigned long long corr = ballot(mine && correct);
                float cx = 0.0f, cy = 0.0f, cz = 0.0f;
                if (corr) {
                    const uint32_t cnt = node_range[k].y;
                    float f[kMlpIn];
                    mlp_features(rx, ry, rz, inv, soft, nd.comMass.w, cnt, mn, mx, nd.comMass, f);
                    mlp_stage_features(s_x[w], f);
                    // 16-sample blocks holding at least one lane to correct
                    uint32_t blocks = 0;
                    for (int b = 0; b < 4; b++) blocks |= ((corr >> (16 * b)) & 0xFFFFull) ? (1u << b) : 0u;
                    mlp_wave_forward(R, s_x[w], s_y[w], blocks);
                    cx = s_y[w][lane * 4 + 0];
                    cy = s_y[w][lane * 4 + 1];
                    cz = s_y[w][lane * 4 + 2];
                    corrected += (unsigned long long)__popcll(corr);
                }
                if (mine && accept) {
                    const float gm = fmul(G, nd.comMass.w);
                    ax = fadd(ax, fmul(fmul(fmul(fmul(gm, rx), inv), inv), inv));
                    ay = fadd(ay, fmul(fmul(fmul(fmul(gm, ry), inv), inv), inv));
                    az = fadd(az, fmul(fmul(fmul(fmul(gm, rz), inv), inv), inv));
                    if (correct) {
                        const float s2 = gm * inv * inv;
                        ax += s2 * cx;
                        ay += s2 * cy;
                        az += s2 * cz;
                    }
                }
                inter += (unsigned long long)__popcll(mask & ~open);
                if (open) {
                    if (sp + 1 > kStackDepth) {
                        err |= 1u;
                        break;
                    }
                    if (lane == 0) {
                        s_node[w][sp] = nd.childA;
                        s_mask[w][sp] = open;
                    }
                    sp += 1;
                    opened += (unsigned long long)__popcll(open);
                    if ((uint32_t)sp + 1 > maxsp) maxsp = (uint32_t)sp + 1;
                    node = uniform(nd.childB);
                    mask = open;
                    descend = true;
                }
            } else {
                const float4 lp = leaf_pos[node];
                const bool use = mine && node != g;
                const float rx = fsub(lp.x, pm.x);
                const float ry = fsub(lp.y, pm.y);
                const float rz = fsub(lp.z, pm.z);
                const float d2 = fadd(dot3(rx, ry, rz), soft);
                const float inv = inv_sqrt_exact(d2);
                if (use) {
                    const float gm = fmul(G, lp.w);
                    ax = fadd(ax, fmul(fmul(fmul(fmul(gm, rx), inv), inv), inv));
                    ay = fadd(ay, fmul(fmul(fmul(fmul(gm, ry), inv), inv), inv));
                    az = fadd(az, fmul(fmul(fmul(fmul(gm, rz), inv), inv), inv));
                }
                const unsigned long long selfbit =
                    (node >= base && node < base + kWave) ? (1ull << (node - base)) : 0ull;
                inter += (unsigned long long)__popcll(mask & ~selfbit);
            }
            if (!descend) {
                if (sp == 0) break;
                sp--;
                node = uniform(s_node[w][sp]);
                const unsigned long long mraw = s_mask[w][sp];
                mask = ((unsigned long long)uniform((uint32_t)(mraw >> 32)) << 32) | uniform((uint32_t)mraw);
            }
        }
    }
    if (active) {
        const float dt = dt_;
        const float nvx = fadd(vd.x, fmul(ax, dt));
        const float nvy = fadd(vd.y, fmul(ay, dt));
        const float nvz = fadd(vd.z, fmul(az, dt));
        dst[2 * bodyIdx] = make_float4(fadd(pm.x, fmul(nvx, dt)), fadd(pm.y, fmul(nvy, dt)), fadd(pm.z, fmul(nvz, dt)),
                                       pm.w);
        dst[2 * bodyIdx + 1] = make_float4(nvx, nvy, nvz, vd.w);
    }
    const uint32_t err_w = (ballot(err & 1u) ? 1u : 0u);
    if (lane == 0) {
        atomicAdd(&ctr->opened, opened);
        atomicAdd(&ctr->interactions, inter);
        atomicAdd(&ctr->corrected, corrected);
        atomicMax(&ctr->maxStack, maxsp);
        if (err_w) atomicOr(&ctr->errorFlags, err_w);
        atomicAdd(&sim->uintDebug[0], (uint32_t)opened);
    }
}

void launch_walk_learned(hipStream_t s, uint32_t n, const WalkArgs& a, const float4* node_aabb,
                         const uint2* node_range, const __bf16* w, const float* bias, float theta_plain) {
    const uint32_t per_block = kWalkWaves * kWave;
    const uint32_t nb = (n + per_block - 1) / per_block;
    k_walk_learned<<<nb, kWalkWaves * kWave, 0, s>>>(const_cast<SimValues*>(a.sim), a.ctr, a.src, a.dst, a.index_in,
                                                      a.leaf_pos, a.nodes, node_aabb, node_range, w, bias, a.theta,
                                                      theta_plain, a.dt, a.softening);
}

}  // namespace bh
