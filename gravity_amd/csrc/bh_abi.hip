// bh_abi.hip — host side of the C ABI declared in include/bh_abi.h.
//
// A context owns the device memory behind the reference's 16 bindings
// (gpu/GPU.java:199-341) plus the tree and scratch arrays of this
// implementation, and one HIP stream. Swaps exchange pointers, as the
// reference exchanges GL buffer names (gpu/GPU.java:1060-1086).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <utility>
#include <vector>

#include "../../include/bh_abi.h"
#include "bh_device.hpp"
#include "bh_kernels.hpp"
#include "bh_mlp.hpp"

using namespace bh;

struct bh_mlp {
    int device = 0;
    hipStream_t stream = nullptr;
    float* params = nullptr;   // FP32 master [kMlpParams]
    float* grad = nullptr;
    float* m1 = nullptr;
    float* m2 = nullptr;
    __bf16* w = nullptr;       // BF16 MFMA weights [kBfTotal]
    float* bias = nullptr;     // [kBiasTotal]
    int adam_t = 0;
    // training workspace for `cap` rows
    uint32_t cap = 0;
    float *x = nullptr, *t = nullptr, *z1 = nullptr, *h1 = nullptr, *z2 = nullptr, *h2 = nullptr, *y = nullptr,
          *dy = nullptr, *partial = nullptr;
    float* gemm_ws = nullptr;   // split-K partial sums of the weight gradients
    char err[512] = {0};
};

struct bh_context {
    int device = 0;
    hipStream_t stream = nullptr;
    uint32_t n = 0;

    // reference bindings
    float4* bodies[2] = {nullptr, nullptr};   // 3 BodiesIn = bodies[body_in], 4 BodiesOut
    uint64_t* morton[2] = {nullptr, nullptr}; // 5 MortonIn = morton[mi_in], 6 MortonOut
    uint32_t* index[2] = {nullptr, nullptr};  // 7 IndexIn = index[mi_in], 8 IndexOut
    SimValues* sim = nullptr;                 // 2
    uint2* tasks = nullptr;                   // 14 (pairs)
    uint32_t* walk_queues = nullptr;          // heads of the persistent walk's per-CU queues
    int body_in = 0, mi_in = 0;

    TreeArrays tree{};
    DevCounters* ctr = nullptr;
    uint32_t* part_scratch = nullptr;         // per-tile alive counts / offsets
    float* bounds_partials = nullptr;
    uint32_t bounds_partial_cap = 0;
    uint32_t* sort_partials = nullptr;        // [sort_digit_blocks(n)][8][256] digit counts
    uint32_t tiles_cap = 0;                   // sort tiles of n keys
    uint32_t* digit_totals = nullptr;         // [8][256] keys per digit value of each pass
    unsigned long long* sort_status = nullptr;   // [8][tiles_cap][256] look-back words
    SortState* sort_state = nullptr;
    unsigned long long* reserve = nullptr;
    uint32_t* task_done = nullptr;
    NodeStd430* staging_nodes = nullptr;      // boundary conversion of bindings 0 and 1 (and the render feed)
    uint32_t* feed_counts = nullptr;          // [render_feed_blocks(n) + 1] render feed offsets
    float* leaf_radius = nullptr;             // body radius per sorted position (merging/collision walks)
    // Positions the last bh_step's walk wrote, by sorted position (the same
    // values it wrote to the body buffer by id). While valid, the next step's
    // Morton encode reads them in order instead of gathering bodies through
    // the index; any other call that can change bodies, index or order ends
    // the validity, and a partition that removes bodies skips them on device.
    float4* step_pos = nullptr;
    bool step_pos_valid = false;
    bool step_merged = false;   // the step that wrote step_pos merged bodies after its walk
    // bh_step sorts positions instead of ids: sort_pos[1] ends up holding,
    // per sorted position, the position the body had before the sort; the
    // tree reads the bodies through it (from step_pos, nearly in order) and
    // writes the ids in the new order to sort_pos[0], which then becomes
    // IndexIn (the buffers trade places).
    uint32_t* sort_pos[2] = {nullptr, nullptr};
    bool colors = false;
    uint32_t fill_cap = 0;   // the last tree's fill cap (BH_FLAG_FILL_CAP), for its std430 download
    const bh_mlp* correction = nullptr;       // attached network of the corrected walk
    // corrected walk (allocated on first use): the 64-byte records of every
    // node id (layout with corr_records) and the FP64 moment prefix sums they
    // are made from. quad_valid: the records belong to the current tree (made
    // by its tree stage under a corrected-walk flag). acc_mag: |a| by body id
    // from the last corrected walk, for the relative opening test (accmag_valid:
    // a corrected walk has run since the bodies were last set from the host).
    float4* qrec = nullptr;                   // [4 corr_records(n)]
    bool quad_valid = false;
    float* acc_mag = nullptr;                 // [n]
    bool accmag_valid = false;

    // bh_step under BH_FLAG_GRAPH: one instantiated graph per (host-side
    // buffer roles before the step, parameters), with the roles after it
    struct Roles {
        int body_in, mi_in;
        uint32_t *index0, *index1, *sort0, *sort1;
        bool step_pos_valid, step_merged, quad_valid, accmag_valid, colors;
        uint32_t fill_cap;
        bool operator==(const Roles& o) const {
            return body_in == o.body_in && mi_in == o.mi_in && index0 == o.index0 && index1 == o.index1 &&
                   sort0 == o.sort0 && sort1 == o.sort1 && step_pos_valid == o.step_pos_valid &&
                   step_merged == o.step_merged &&
                   quad_valid == o.quad_valid && accmag_valid == o.accmag_valid && colors == o.colors &&
                   fill_cap == o.fill_cap;
        }
    };
    struct StepGraph {
        Roles pre, post;
        bh_params prm;
        hipGraphExec_t exec;
    };
    std::vector<StepGraph> graphs;
    // the tree's second moments for the corrected walk (TreeArrays' node_mom,
    // xchg_mom, frontier_mom), set in the tree's arrays only for its builds
    float4* node_mom = nullptr;
    float4* xchg_mom = nullptr;
    float4* frontier_mom = nullptr;

    // per-stage HIP events of up to kTimedSteps steps, read without
    // synchronising inside the stepping loop
    static constexpr int kTimedSteps = 512;
    hipEvent_t* ev = nullptr;            // [kTimedSteps][BH_STAGE_COUNT + 1]
    int timed = 0;                       // steps recorded since the last read
    bool timing = false;
    char err[512] = {0};
};

namespace {

int fail(bh_context* c, int code, const char* fmt, ...) {
    if (c) {
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(c->err, sizeof c->err, fmt, ap);
        va_end(ap);
    }
    return code;
}

#define HIPCHECK(c, expr)                                                                              \
    do {                                                                                               \
        hipError_t e_ = (expr);                                                                        \
        if (e_ != hipSuccess) return fail((c), BH_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_));    \
    } while (0)

template <typename T>
int dalloc(bh_context* c, T** p, size_t count) {
    size_t bytes = count * sizeof(T);
    if (bytes == 0) bytes = sizeof(T);
    hipError_t e = hipMalloc((void**)p, bytes);
    if (e != hipSuccess) return fail(c, BH_ERR_OUT_OF_MEMORY, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
    return BH_OK;
}

int check_launch(bh_context* c, const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(c, BH_ERR_HIP, "%s launch: %s", what, hipGetErrorString(e));
    return BH_OK;
}

void derive_units(const bh_params* p, float* G, float* blc) {
    // common/common.glsl:346-352, evaluated in FP32 as the shader does
    const float GC = 6.67430e-11f;
    *G = GC * powf(p->unit_len, -3.0f) * powf(p->unit_time, 2.0f) * powf(p->unit_mass, 1.0f);
    *blc = 0.6203504909f * powf(p->unit_mass / p->unit_density, 1.0f / 3.0f) / p->unit_len;
    if (p->gravitational_constant != 0.0f) *G = p->gravitational_constant;
    if (p->body_length_constant != 0.0f) *blc = p->body_length_constant;
}

int mfail(bh_mlp* m, int code, const char* fmt, ...) {
    if (m) {
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(m->err, sizeof m->err, fmt, ap);
        va_end(ap);
    }
    return code;
}

#define MHIP(m, expr)                                                                                  \
    do {                                                                                               \
        hipError_t e_ = (expr);                                                                        \
        if (e_ != hipSuccess) return mfail((m), BH_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_));   \
    } while (0)

constexpr uint32_t kGemmWsFloats = (uint32_t)kGemmSlices * kMlpHidden * kMlpHidden;

int mlp_workspace(bh_mlp* m, uint32_t rows) {
    if (rows <= m->cap) return BH_OK;
    float** bufs[] = {&m->x, &m->t, &m->z1, &m->h1, &m->z2, &m->h2, &m->y, &m->dy, &m->partial};
    for (float** b : bufs) {
        if (*b) (void)hipFree(*b);
        *b = nullptr;
    }
    size_t r = rows;
    size_t sizes[] = {r * kMlpIn, r * kMlpTarget, r * kMlpHidden, r * kMlpHidden, r * kMlpHidden, r * kMlpHidden,
                      r * kMlpOut, r * kMlpOut, (r + 255) / 256 + 1};
    for (int k = 0; k < 9; k++) MHIP(m, hipMalloc((void**)bufs[k], sizes[k] * sizeof(float)));
    if (!m->gemm_ws) MHIP(m, hipMalloc((void**)&m->gemm_ws, kGemmWsFloats * sizeof(float)));
    m->cap = rows;
    return BH_OK;
}

// forward over rows already in m->x; leaves Z1,H1,Z2,H2,Y in the workspace
void mlp_forward_fp32(bh_mlp* m, uint32_t rows) {
    hipStream_t s = m->stream;
    const int B = (int)rows;
    launch_gemm(s, m->x, kMlpIn, 1, m->params + kOffW1, 1, kMlpIn, m->z1, B, kMlpHidden, kMlpIn, 0);
    launch_bias_act(s, m->z1, m->h1, m->params + kOffB1, B, kMlpHidden);
    launch_gemm(s, m->h1, kMlpHidden, 1, m->params + kOffW2, 1, kMlpHidden, m->z2, B, kMlpHidden, kMlpHidden, 0);
    launch_bias_act(s, m->z2, m->h2, m->params + kOffB2, B, kMlpHidden);
    launch_gemm(s, m->h2, kMlpHidden, 1, m->params + kOffW3, 1, kMlpHidden, m->y, B, kMlpOut, kMlpHidden, 0);
    launch_bias_act(s, m->y, nullptr, m->params + kOffB3, B, kMlpOut);
}

// backward for targets in m->t: gradient of the mean squared error into m->grad
uint32_t mlp_backward_fp32(bh_mlp* m, uint32_t rows) {
    hipStream_t s = m->stream;
    const int B = (int)rows;
    float* g = m->grad;
    uint32_t nparts = 0;
    launch_loss_grad(s, m->y, m->t, m->dy, B, m->partial, &nparts);
    // layer 3
    launch_gemm(s, m->dy, 1, kMlpOut, m->h2, kMlpHidden, 1, g + kOffW3, kMlpOut, kMlpHidden, B, 0, m->gemm_ws,
                kGemmWsFloats);
    launch_colsum(s, m->dy, g + kOffB3, B, kMlpOut);
    launch_gemm(s, m->dy, kMlpOut, 1, m->params + kOffW3, kMlpHidden, 1, m->h2, B, kMlpHidden, kMlpOut, 0);
    launch_act_grad(s, m->z2, m->h2, (size_t)B * kMlpHidden);          // h2 now holds dZ2
    // layer 2
    launch_gemm(s, m->h2, 1, kMlpHidden, m->h1, kMlpHidden, 1, g + kOffW2, kMlpHidden, kMlpHidden, B, 0,
                m->gemm_ws, kGemmWsFloats);
    launch_colsum(s, m->h2, g + kOffB2, B, kMlpHidden);
    launch_gemm(s, m->h2, kMlpHidden, 1, m->params + kOffW2, kMlpHidden, 1, m->h1, B, kMlpHidden, kMlpHidden, 0);
    launch_act_grad(s, m->z1, m->h1, (size_t)B * kMlpHidden);          // h1 now holds dZ1
    // layer 1
    launch_gemm(s, m->h1, 1, kMlpHidden, m->x, kMlpIn, 1, g + kOffW1, kMlpHidden, kMlpIn, B, 0, m->gemm_ws,
                kGemmWsFloats);
    launch_colsum(s, m->h1, g + kOffB1, B, kMlpHidden);
    return nparts;
}

// Static device scratch of bh_direct_sum_f64_noalloc: the sampled ids
// (BH_NOALLOC_MAX_SAMPLE x 4 bytes) followed by the FP64 sums (x 24 bytes),
// 1.75 MiB. Every GPU holds its own copy at its own address. In host code the
// name `scratch` is only a stand-in for the symbol: kernels, copies and
// memsets get the device address that hipGetSymbolAddress returned on that
// device (scratch_ptr), never the name itself.
constexpr size_t kNoAllocIdsBytes = (size_t)BH_NOALLOC_MAX_SAMPLE * 4;
constexpr size_t kNoAllocBytes = kNoAllocIdsBytes + (size_t)BH_NOALLOC_MAX_SAMPLE * 24;
static_assert(kNoAllocIdsBytes % 256 == 0, "the sums start 256-byte aligned");
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

constexpr int kMaxDevices = 64;
std::mutex scratch_mu;
void* scratch_ptr[kMaxDevices] = {};   // device address of `scratch` per device, set once

// With `device` current: resolve that device's address of the scratch array
// (once per device, at context creation).
int scratch_setup(bh_context* c, int device) {
    if (device < 0 || device >= kMaxDevices)
        return fail(c, BH_ERR_INVALID_ARGUMENT, "device %d outside 0..%d", device, kMaxDevices - 1);
    std::lock_guard<std::mutex> lock(scratch_mu);
    if (scratch_ptr[device]) return BH_OK;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return fail(c, BH_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", device, hipGetErrorString(e));
    scratch_ptr[device] = p;
    return BH_OK;
}

unsigned char* scratch_of(int device) {
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(scratch_mu);
    return static_cast<unsigned char*>(scratch_ptr[device]);
}

// stage events are created on the first timed step
int ensure_events(bh_context* c) {
    if (c->ev) return BH_OK;
    const int total = bh_context::kTimedSteps * (BH_STAGE_COUNT + 1);
    c->ev = new (std::nothrow) hipEvent_t[total]();
    if (!c->ev) return fail(c, BH_ERR_OUT_OF_MEMORY, "events");
    for (int k = 0; k < total; k++)
        HIPCHECK(c, hipEventCreate(&c->ev[k]));
    return BH_OK;
}

void stage_begin(bh_context* c, int k) {
    if (c->timing && c->timed < bh_context::kTimedSteps)
        (void)hipEventRecord(c->ev[c->timed * (BH_STAGE_COUNT + 1) + k], c->stream);
}

}  // namespace

extern "C" {

const char* bh_last_error(const bh_context* c) { return c ? c->err : "null context"; }

int bh_create(int device, uint32_t n, const float* bounds6, const bh_params* units, bh_context** out) {
    if (!out) return BH_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    if (n < 2 || !bounds6) return BH_ERR_INVALID_ARGUMENT;
    // 32-bit positions and node ids: n + (n - 1) must fit below 0xFFFFFFFF
    if (n > (1u << 30)) return BH_ERR_TOO_LARGE;
    bh_context* c = new (std::nothrow) bh_context();
    if (!c) return BH_ERR_OUT_OF_MEMORY;
    c->device = device;
    c->n = n;
    int rc;
#define TRY(x)                     \
    do {                           \
        rc = (x);                  \
        if (rc != BH_OK) goto bad; \
    } while (0)
    {
        hipError_t e = hipSetDevice(device);
        if (e != hipSuccess) { rc = fail(c, BH_ERR_HIP, "hipSetDevice(%d): %s", device, hipGetErrorString(e)); goto bad; }
        TRY(scratch_setup(c, device));
        e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
        if (e != hipSuccess) { rc = fail(c, BH_ERR_HIP, "hipStreamCreate: %s", hipGetErrorString(e)); goto bad; }
        for (int k = 0; k < 2; k++) {
            TRY(dalloc(c, &c->bodies[k], 2 * (size_t)n));
            TRY(dalloc(c, &c->morton[k], (size_t)n));
            TRY(dalloc(c, &c->index[k], (size_t)n));
            TRY(dalloc(c, &c->sort_pos[k], (size_t)n));
        }
        TRY(dalloc(c, &c->sim, 1));
        TRY(dalloc(c, &c->tasks, (size_t)n));
        TRY(dalloc(c, &c->walk_queues, (size_t)kWalkQueues * 32));
        TRY(dalloc(c, &c->ctr, 1));
        TreeArrays& T = c->tree;
        TRY(dalloc(c, &T.leaf_pos, (size_t)n));
        TRY(dalloc(c, &T.leaf_parent, (size_t)n));
        TRY(dalloc(c, &T.leaf_color, (size_t)n));
        TRY(dalloc(c, &T.walk, 2 * (size_t)n - 1));
        T.nodes = T.walk + n;
        TRY(dalloc(c, &T.node_aabb, 2 * ((size_t)n - 1)));
        TRY(dalloc(c, &T.node_parent, (size_t)n - 1));
        TRY(dalloc(c, &T.node_range, (size_t)n - 1));
        TRY(dalloc(c, &T.node_color, (size_t)n - 1));
        TRY(dalloc(c, &T.meet, (size_t)n - 1));
        TRY(dalloc(c, &T.xchg, 8 * ((size_t)n - 1)));
        // the entry slots and counts of every build level, for the larger of
        // the two plans (the build with second moments has smaller blocks)
        TRY(dalloc(c, &T.frontier, 4 * frontier_slots(n)));
        TRY(dalloc(c, &T.frontier_cnt, frontier_counts(n)));
        TRY(dalloc(c, &c->part_scratch, (size_t)(n + kPartTile - 1) / kPartTile));
        c->bounds_partial_cap = 4096;
        TRY(dalloc(c, &c->bounds_partials, 6 * (size_t)c->bounds_partial_cap));
        c->tiles_cap = (n + sort_tile(n) - 1) / sort_tile(n);
        TRY(dalloc(c, &c->sort_partials, (size_t)sort_digit_blocks(n) * kSortPasses * 256));
        TRY(dalloc(c, &c->digit_totals, kSortPasses * 256));
        TRY(dalloc(c, &c->sort_status, (size_t)kSortPasses * c->tiles_cap * 256));
        TRY(dalloc(c, &c->sort_state, 1));
        TRY(dalloc(c, &c->reserve, (size_t)n));
        TRY(dalloc(c, &c->task_done, (size_t)n));
        TRY(dalloc(c, &c->staging_nodes, (size_t)n));
        TRY(dalloc(c, &c->feed_counts, (size_t)render_feed_blocks(n) + 1));
        TRY(dalloc(c, &c->leaf_radius, (size_t)n));
        TRY(dalloc(c, &c->step_pos, (size_t)n));
        // zero everything the reference zero-initialises, then pack sim
        // (gpu/GPU.java:1096-1143)
        for (int k = 0; k < 2; k++) {
            (void)hipMemsetAsync(c->bodies[k], 0, 32 * (size_t)n, c->stream);
            (void)hipMemsetAsync(c->morton[k], 0, 8 * (size_t)n, c->stream);
            (void)hipMemsetAsync(c->index[k], 0, 4 * (size_t)n, c->stream);
        }
        (void)hipMemsetAsync(c->ctr, 0, sizeof(DevCounters), c->stream);
        // epoch 0 marks look-back words no sort has written
        (void)hipMemsetAsync(c->sort_status, 0, 8 * (size_t)kSortPasses * c->tiles_cap * 256, c->stream);
        (void)hipMemsetAsync(c->sort_state, 0, sizeof(SortState), c->stream);
        (void)hipMemsetAsync(T.meet, 0xFF, 4 * ((size_t)n - 1), c->stream);
        (void)hipMemsetAsync(c->reserve, 0xFF, 8 * (size_t)n, c->stream);
        (void)hipMemsetAsync(T.leaf_color, 0, 16 * (size_t)n, c->stream);
        (void)hipMemsetAsync(T.node_color, 0, 16 * ((size_t)n - 1), c->stream);
        SimValues h;
        memset(&h, 0, sizeof h);
        h.numBodies = n;
        h.initialNumBodies = n;
        memcpy(h.minCorner, bounds6, 12);
        memcpy(h.maxCorner, bounds6 + 3, 12);
        if (units) {
            h.unitMass = units->unit_mass;
            h.unitDensity = units->unit_density;
            h.unitLen = units->unit_len;
            h.unitTime = units->unit_time;
            if (units->unit_mass != 0.0f) derive_units(units, &h.gravitationalConstant, &h.bodyLengthConstant);
        }
        e = hipMemcpyAsync(c->sim, &h, sizeof h, hipMemcpyHostToDevice, c->stream);
        if (e != hipSuccess) { rc = fail(c, BH_ERR_HIP, "upload sim: %s", hipGetErrorString(e)); goto bad; }
        e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) { rc = fail(c, BH_ERR_HIP, "create sync: %s", hipGetErrorString(e)); goto bad; }
    }
#undef TRY
    *out = c;
    return BH_OK;
bad:
    fprintf(stderr, "bh_create: %s\n", c->err);
    bh_destroy(c);
    return rc;
}

void bh_destroy(bh_context* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (auto& g : c->graphs) (void)hipGraphExecDestroy(g.exec);
    c->graphs.clear();
    void* ptrs[] = {c->bodies[0], c->bodies[1], c->morton[0], c->morton[1], c->index[0], c->index[1], c->sim,
                    c->tasks, c->ctr, c->tree.leaf_pos, c->tree.leaf_parent, c->tree.leaf_color, c->tree.walk,
                    c->tree.node_aabb, c->tree.node_parent, c->tree.node_range, c->tree.node_color, c->tree.meet,
                    c->tree.xchg, c->tree.frontier, c->tree.frontier_cnt, c->part_scratch, c->bounds_partials, c->sort_partials, c->digit_totals, c->sort_status, c->sort_state,
                    c->reserve, c->task_done, c->staging_nodes, c->leaf_radius, c->step_pos, c->sort_pos[0],
                    c->sort_pos[1], c->walk_queues, c->qrec, c->node_mom, c->xchg_mom, c->frontier_mom, c->acc_mag,
                    c->feed_counts};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    if (c->ev) {
        for (int k = 0; k < bh_context::kTimedSteps * (BH_STAGE_COUNT + 1); k++)
            if (c->ev[k]) (void)hipEventDestroy(c->ev[k]);
        delete[] c->ev;
    }
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

size_t bh_binding_size(const bh_context* c, int b) {
    if (!c) return 0;
    const size_t n = c->n;
    switch (b) {
        case BH_BINDING_LEAF_NODES: return n * 80;
        case BH_BINDING_INTERNAL_NODES: return (n - 1) * 80;
        case BH_BINDING_SIMULATION_VALUES: return sizeof(SimValues);
        case BH_BINDING_BODIES_IN:
        case BH_BINDING_BODIES_OUT: return n * 32;
        case BH_BINDING_MORTON_IN:
        case BH_BINDING_MORTON_OUT: return n * 8;
        case BH_BINDING_INDEX_IN:
        case BH_BINDING_INDEX_OUT: return n * 4;
        case BH_BINDING_MERGE_TASKS: return 8 + n * 8;
        case BH_BINDING_NODE_QUADRUPOLES: return (n - 1) * 32;
        default: return 0;
    }
}

static void* binding_ptr(bh_context* c, int b) {
    switch (b) {
        case BH_BINDING_SIMULATION_VALUES: return c->sim;
        case BH_BINDING_BODIES_IN: return c->bodies[c->body_in];
        case BH_BINDING_BODIES_OUT: return c->bodies[1 - c->body_in];
        case BH_BINDING_MORTON_IN: return c->morton[c->mi_in];
        case BH_BINDING_MORTON_OUT: return c->morton[1 - c->mi_in];
        case BH_BINDING_INDEX_IN: return c->index[c->mi_in];
        case BH_BINDING_INDEX_OUT: return c->index[1 - c->mi_in];
        default: return nullptr;
    }
}

static int check_range(bh_context* c, int b, size_t off, size_t bytes) {
    size_t sz = bh_binding_size(c, b);
    if (sz == 0) return fail(c, BH_ERR_UNSUPPORTED, "binding %d is reference scratch with no counterpart here", b);
    if (off > sz || bytes > sz - off) return fail(c, BH_ERR_INVALID_ARGUMENT, "binding %d: range %zu+%zu > %zu", b, off, bytes, sz);
    return BH_OK;
}

// the corrected walk's arrays, allocated on first use (about 260 bytes per body)
static int ensure_quad(bh_context* c) {
    if (c->qrec) return BH_OK;
    int rc;
    if ((rc = dalloc(c, &c->node_mom, 2 * ((size_t)c->n - 1)))) return rc;
    if ((rc = dalloc(c, &c->xchg_mom, 6 * ((size_t)c->n - 1)))) return rc;
    if ((rc = dalloc(c, &c->frontier_mom, 3 * tree_levels(c->n, kTreeBlockMoments).entries))) return rc;
    if ((rc = dalloc(c, &c->qrec, 4 * corr_records(c->n)))) return rc;
    if ((rc = dalloc(c, &c->acc_mag, c->n))) return rc;
    // the quadrupole halves of the leaf records stay zero
    HIPCHECK(c, hipMemsetAsync(c->qrec, 0, 64 * corr_records(c->n), c->stream));
    return BH_OK;
}

// Binding 16 is the quadrupole half (bytes 32..63) of every internal node's
// record: copied whole, with a 64-byte pitch on the device. The records hold
// S = -3 G T; the binding holds G T (converted in FP32 on the host).
static int quad_binding_copy(bh_context* c, void* host, size_t off, size_t bytes, bool upload) {
    if (off != 0 || bytes != bh_binding_size(c, BH_BINDING_NODE_QUADRUPOLES))
        return fail(c, BH_ERR_INVALID_ARGUMENT, "the quadrupole records are copied whole");
    char* dev = reinterpret_cast<char*>(c->qrec + 4 * (size_t)corr_internal_id(c->n, 0)) + 32;
    const size_t count = (size_t)c->n - 1;
    if (upload) {
        std::vector<float> s(8 * count);
        const float* in = static_cast<const float*>(host);
        for (size_t k = 0; k < count; k++)
            for (int q = 0; q < 8; q++) s[8 * k + q] = q < 6 ? -3.0f * in[8 * k + q] : in[8 * k + q];
        HIPCHECK(c, hipMemcpy2DAsync(dev, 64, s.data(), 32, 32, count, hipMemcpyHostToDevice, c->stream));
        HIPCHECK(c, hipStreamSynchronize(c->stream));
    } else {
        HIPCHECK(c, hipMemcpy2DAsync(host, 32, dev, 64, 32, count, hipMemcpyDeviceToHost, c->stream));
        HIPCHECK(c, hipStreamSynchronize(c->stream));
        float* out = static_cast<float*>(host);
        for (size_t k = 0; k < count; k++)
            for (int q = 0; q < 6; q++) out[8 * k + q] = out[8 * k + q] / -3.0f;
    }
    return BH_OK;
}

int bh_upload(bh_context* c, int b, const void* host, size_t off, size_t bytes) {
    if (!c || (!host && bytes)) return BH_ERR_INVALID_ARGUMENT;
    c->step_pos_valid = false;
    if (b == BH_BINDING_BODIES_IN || b == BH_BINDING_BODIES_OUT) c->accmag_valid = false;
    int rc = check_range(c, b, off, bytes);
    if (rc) return rc;
    HIPCHECK(c, hipSetDevice(c->device));
    if (b == BH_BINDING_NODE_QUADRUPOLES) {
        if ((rc = ensure_quad(c))) return rc;
        return quad_binding_copy(c, const_cast<void*>(host), off, bytes, true);
    }
    if (b == BH_BINDING_LEAF_NODES || b == BH_BINDING_INTERNAL_NODES) {
        c->quad_valid = false;   // the tree changes under the fused records
        if (off != 0 || bytes != bh_binding_size(c, b))
            return fail(c, BH_ERR_INVALID_ARGUMENT, "node bindings are uploaded whole");
        HIPCHECK(c, hipMemcpyAsync(c->staging_nodes, host, bytes, hipMemcpyHostToDevice, c->stream));
        int which = b == BH_BINDING_LEAF_NODES ? 0 : 1;
        launch_std430_to_nodes(c->stream, c->n, c->sim, c->tree, c->staging_nodes, c->staging_nodes, which);
        return check_launch(c, "std430_to_nodes");
    }
    if (b == BH_BINDING_MERGE_TASKS) {
        // layout of bh_merge.comp / common.glsl:117: head, tail, uvec2 tasks[]
        const uint8_t* h = (const uint8_t*)host;
        if (off != 0 || bytes < 8) return fail(c, BH_ERR_INVALID_ARGUMENT, "merge tasks upload starts at 0 with head/tail");
        uint32_t tail;
        memcpy(&tail, h + 4, 4);
        size_t pairs = (bytes - 8) / 8;
        if (tail > pairs) tail = (uint32_t)pairs;
        HIPCHECK(c, hipMemcpyAsync(c->tasks, h + 8, (size_t)tail * 8, hipMemcpyHostToDevice, c->stream));
        HIPCHECK(c, hipMemcpyAsync(&c->ctr->nTasks, &tail, 4, hipMemcpyHostToDevice, c->stream));
        return hipStreamSynchronize(c->stream) == hipSuccess ? BH_OK : fail(c, BH_ERR_HIP, "sync");
    }
    HIPCHECK(c, hipMemcpyAsync((uint8_t*)binding_ptr(c, b) + off, host, bytes, hipMemcpyHostToDevice, c->stream));
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    return BH_OK;
}

extern "C++" {
namespace {
// the render feed's records, staged in the node conversion buffer (80 bytes
// per body: more than either kind needs), then copied in one piece
template <class R>
int feed_copy(bh_context* c, R* out, uint32_t cap, uint32_t* count, const char* what) {
    int rc = check_launch(c, what);
    if (rc) return rc;
    uint32_t total = 0;
    HIPCHECK(c, hipMemcpyAsync(&total, c->feed_counts + render_feed_blocks(c->n), 4, hipMemcpyDeviceToHost,
                               c->stream));
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    const uint32_t m = total < cap ? total : cap;
    if (m) {
        HIPCHECK(c, hipMemcpyAsync(out, c->staging_nodes, (size_t)m * sizeof(R), hipMemcpyDeviceToHost, c->stream));
        HIPCHECK(c, hipStreamSynchronize(c->stream));
    }
    *count = total;
    return BH_OK;
}
}  // namespace
}  // extern "C++"

int bh_host_alloc(size_t bytes, void** out) {
    if (!out) return BH_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    if (hipHostMalloc(out, bytes ? bytes : 1, hipHostMallocDefault) != hipSuccess) {
        *out = nullptr;
        return BH_ERR_OUT_OF_MEMORY;
    }
    return BH_OK;
}

void bh_host_free(void* p) {
    if (p) (void)hipHostFree(p);
}

int bh_render_nodes(bh_context* c, uint32_t depth_min, uint32_t depth_max, bh_render_node* out, uint32_t cap,
                    uint32_t* count) {
    if (!c || !count || (!out && cap)) return BH_ERR_INVALID_ARGUMENT;
    HIPCHECK(c, hipSetDevice(c->device));
    static_assert(sizeof(bh_render_node) <= sizeof(NodeStd430), "staging");
    launch_render_nodes(c->stream, c->n, c->sim, c->tree, c->colors ? 1 : 0, depth_min, depth_max, c->feed_counts,
                        reinterpret_cast<bh_render_node*>(c->staging_nodes), cap < c->n ? cap : c->n);
    return feed_copy(c, out, cap, count, "render nodes");
}

int bh_render_bodies(bh_context* c, bh_render_body* out, uint32_t cap, uint32_t* count) {
    if (!c || !count || (!out && cap)) return BH_ERR_INVALID_ARGUMENT;
    HIPCHECK(c, hipSetDevice(c->device));
    static_assert(sizeof(bh_render_body) <= sizeof(NodeStd430), "staging");
    launch_render_bodies(c->stream, c->n, c->sim, c->bodies[c->body_in], c->feed_counts,
                         reinterpret_cast<bh_render_body*>(c->staging_nodes), cap < c->n ? cap : c->n);
    return feed_copy(c, out, cap, count, "render bodies");
}

int bh_download(bh_context* c, int b, void* host, size_t off, size_t bytes) {
    if (!c || (!host && bytes)) return BH_ERR_INVALID_ARGUMENT;
    int rc = check_range(c, b, off, bytes);
    if (rc) return rc;
    HIPCHECK(c, hipSetDevice(c->device));
    if (b == BH_BINDING_NODE_QUADRUPOLES) {
        if ((rc = ensure_quad(c))) return rc;
        return quad_binding_copy(c, host, off, bytes, false);
    }
    if (b == BH_BINDING_LEAF_NODES || b == BH_BINDING_INTERNAL_NODES) {
        launch_nodes_to_std430(c->stream, c->n, c->sim, c->tree, c->colors, b == BH_BINDING_LEAF_NODES ? 0 : 1,
                               c->fill_cap, c->staging_nodes);
        rc = check_launch(c, "nodes_to_std430");
        if (rc) return rc;
        HIPCHECK(c, hipMemcpyAsync(host, (uint8_t*)c->staging_nodes + off, bytes, hipMemcpyDeviceToHost, c->stream));
        HIPCHECK(c, hipStreamSynchronize(c->stream));
        return BH_OK;
    }
    if (b == BH_BINDING_MERGE_TASKS) {
        uint32_t nt = 0;
        HIPCHECK(c, hipMemcpyAsync(&nt, &c->ctr->nTasks, 4, hipMemcpyDeviceToHost, c->stream));
        HIPCHECK(c, hipStreamSynchronize(c->stream));
        if (nt > c->n) nt = c->n;
        uint8_t* tmp = new (std::nothrow) uint8_t[8 + (size_t)nt * 8];
        if (!tmp) return BH_ERR_OUT_OF_MEMORY;
        uint32_t ht[2] = {0, nt};
        memcpy(tmp, ht, 8);
        hipError_t e = hipMemcpy(tmp + 8, c->tasks, (size_t)nt * 8, hipMemcpyDeviceToHost);
        if (e == hipSuccess) {
            size_t avail = 8 + (size_t)nt * 8;
            memset(host, 0, bytes);
            if (off < avail) memcpy(host, tmp + off, bytes < avail - off ? bytes : avail - off);
        }
        delete[] tmp;
        return e == hipSuccess ? BH_OK : fail(c, BH_ERR_HIP, "download tasks: %s", hipGetErrorString(e));
    }
    HIPCHECK(c, hipMemcpyAsync(host, (uint8_t*)binding_ptr(c, b) + off, bytes, hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    return BH_OK;
}

int bh_swap_body_buffers(bh_context* c) {
    if (!c) return BH_ERR_INVALID_ARGUMENT;
    c->step_pos_valid = false;
    c->body_in = 1 - c->body_in;
    return BH_OK;
}

int bh_swap_morton_and_index_buffers(bh_context* c) {
    if (!c) return BH_ERR_INVALID_ARGUMENT;
    c->step_pos_valid = false;
    c->mi_in = 1 - c->mi_in;
    return BH_OK;
}

int bh_init(bh_context* c) {
    if (!c) return BH_ERR_INVALID_ARGUMENT;
    c->step_pos_valid = false;
    c->accmag_valid = false;
    HIPCHECK(c, hipSetDevice(c->device));
    launch_init(c->stream, c->index[c->mi_in], c->n);
    return check_launch(c, "init");
}

int bh_reset_values(bh_context* c, const bh_params* p) {
    if (!c || !p) return BH_ERR_INVALID_ARGUMENT;
    float G = 0.0f, blc = 0.0f;
    int set = p->unit_mass != 0.0f;
    if (set) derive_units(p, &G, &blc);
    launch_reset(c->stream, c->sim, c->ctr, set, p->unit_mass, p->unit_density, p->unit_len, p->unit_time, G, blc);
    return check_launch(c, "reset");
}

namespace {
// with_bounds: the partition's census pass also updates the bounds (what
// bh_update_bounds does; the bounds skip the bodies the partition removes)
// step_pos: the previous walk's positions, when they are the bodies' (the census reads them)
int partition_dead_bodies(bh_context* c, bool with_bounds, const float4* step_pos = nullptr) {
    launch_dead_partition(c->stream, c->n, c->sim, c->ctr, c->bodies[c->body_in], c->bodies[1 - c->body_in],
                          c->index[c->mi_in], c->index[1 - c->mi_in], c->part_scratch,
                          with_bounds ? c->bounds_partials : nullptr, c->bounds_partial_cap, step_pos);
    int rc = check_launch(c, "dead partition");
    if (rc) return rc;
    return bh_swap_morton_and_index_buffers(c);   // BarnesHut.java:117
}
}  // namespace

int bh_partition_dead_bodies(bh_context* c) {
    if (!c) return BH_ERR_INVALID_ARGUMENT;
    c->step_pos_valid = false;
    return partition_dead_bodies(c, false);
}

int bh_update_bounds(bh_context* c) {
    if (!c) return BH_ERR_INVALID_ARGUMENT;
    launch_bounds(c->stream, c->n, c->sim, c->ctr, c->bodies[c->body_in], c->index[c->mi_in], c->bounds_partials,
                  c->bounds_partial_cap);
    return check_launch(c, "bounds");
}

namespace {
int generate_morton_codes(bh_context* c, const float4* step_pos) {
    launch_encode(c->stream, c->n, c->sim, c->ctr, c->bodies[c->body_in], c->index[c->mi_in], step_pos,
                  c->morton[c->mi_in]);
    return check_launch(c, "encode");
}
}  // namespace

int bh_generate_morton_codes(bh_context* c) {
    if (!c) return BH_ERR_INVALID_ARGUMENT;
    c->step_pos_valid = false;
    return generate_morton_codes(c, nullptr);
}

namespace {
// Stable LSD sort of the keys in MortonIn, landing back in MortonIn (an even
// number of passes), with the ids in IndexIn (positions: with the positions
// 0..num-1 instead, into sort_pos[1], IndexIn untouched).
static_assert(kSortPasses % 2 == 0, "the sort ends in the In buffers");
int radix_sort(bh_context* c, bool positions) {
    uint32_t* vals[2] = {positions ? c->sort_pos[0] : c->index[0], positions ? c->sort_pos[1] : c->index[1]};
    launch_sort(c->stream, c->n, c->sim, c->ctr, c->morton, vals, c->mi_in, positions, c->sort_partials,
                c->digit_totals, c->sort_status, c->tiles_cap, c->sort_state);
    return check_launch(c, "sort");
}

// positions: the sort before sorted positions (radix_sort(c, true)); the
// tree then puts the ids in the new order into IndexIn.
int build_tree(bh_context* c, const bh_params* p, bool positions, const float4* step_pos) {
    c->colors = p && (p->flags & BH_FLAG_COLORS);
    const bool quad = p && (p->flags & (BH_FLAG_QUADRUPOLE | BH_FLAG_LEARNED));
    const bool learned = p && (p->flags & BH_FLAG_LEARNED);
    if (learned && !c->correction) return fail(c, BH_ERR_INVALID_ARGUMENT, "BH_FLAG_LEARNED without an attached network");
    const uint32_t cap = p && (p->flags & BH_FLAG_FILL_CAP) ? kReferenceFillPasses : 0u;
    if (quad && cap) return fail(c, BH_ERR_UNSUPPORTED, "BH_FLAG_FILL_CAP is for the plain method");
    int rc;
    if (quad && (rc = ensure_quad(c))) return rc;
    c->fill_cap = cap;
    TreeArrays T = c->tree;
    if (quad) {   // the tree also builds every node's second moment
        T.node_mom = c->node_mom;
        T.xchg_mom = c->xchg_mom;
        T.frontier_mom = c->frontier_mom;
        T.corr_leaf = c->qrec;
    }
    launch_tree(c->stream, c->n, c->sim, c->ctr, c->bodies[c->body_in], c->index[c->mi_in], c->morton[c->mi_in],
                T, c->colors ? 1 : 0, positions ? step_pos : nullptr,
                positions ? c->sort_pos[c->mi_in] : nullptr, positions ? c->sort_pos[1 - c->mi_in] : nullptr, cap);
    rc = check_launch(c, "tree");
    if (rc == BH_OK && positions) std::swap(c->index[c->mi_in], c->sort_pos[1 - c->mi_in]);
    c->quad_valid = false;
    if (rc == BH_OK && quad) {
        launch_node_quadrupoles(c->stream, c->n, c->sim, c->ctr, T,
                                learned ? c->correction->w : nullptr, learned ? c->correction->bias : nullptr,
                                c->qrec);
        rc = check_launch(c, "node quadrupoles");
        c->quad_valid = rc == BH_OK;
    }
    return rc;
}
}  // namespace

int bh_radix_sort(bh_context* c) {
    if (!c) return BH_ERR_INVALID_ARGUMENT;
    c->step_pos_valid = false;
    return radix_sort(c, false);
}

int bh_build_tree(bh_context* c, const bh_params* p) {
    if (!c) return BH_ERR_INVALID_ARGUMENT;
    c->step_pos_valid = false;
    return build_tree(c, p, false, nullptr);
}

namespace {
// step_pos: where the plain walk also writes the new positions by sorted
// position (nullptr: not written); *wrote tells whether it did.
int compute_force(bh_context* c, const bh_params* p, float4* step_pos, bool* wrote) {
    *wrote = false;
    if (p->merging_collision_or_neither > BH_COLLISION || p->static_or_dynamic > 1)
        return fail(c, BH_ERR_INVALID_ARGUMENT, "bad selector");
    WalkArgs a;
    a.sim = c->sim;
    a.ctr = c->ctr;
    a.src = c->bodies[c->body_in];
    a.dst = c->bodies[1 - c->body_in];
    a.index_in = c->index[c->mi_in];
    a.leaf_pos = c->tree.leaf_pos;
    a.nodes = c->tree.nodes;
    a.tasks = c->tasks;
    a.task_cap = c->n;
    a.theta = p->theta;
    a.dt = p->dt;
    a.softening = p->softening;
    a.merging = p->merging_collision_or_neither;
    a.dynamic = p->static_or_dynamic;
    a.wrap = p->wrap_around;
    a.elasticity = p->elasticity;
    // the reference never uploads restitution to the force program
    // (gpu/GPU.java:775-783 vs :442-444), so the shader's correction term
    // multiplies by 0; bh_params.restitution is kept for the interface
    a.restitution = 0.0f;
    a.leaf_radius = c->leaf_radius;
    a.general = (p->flags & BH_FLAG_GENERAL_WALK) ? 1u : 0u;
    a.pos_out = step_pos;
    a.queues = c->walk_queues;
    if (p->flags & (BH_FLAG_QUADRUPOLE | BH_FLAG_LEARNED)) {
        if (p->merging_collision_or_neither != BH_NEITHER)
            return fail(c, BH_ERR_UNSUPPORTED, "the corrected walk runs without merging or collisions");
        if (!(p->softening >= 1e-14f))
            return fail(c, BH_ERR_UNSUPPORTED, "the corrected walk needs softening >= 1e-14");
        if (!c->qrec || !c->quad_valid)
            return fail(c, BH_ERR_INVALID_ARGUMENT,
                        "no records of the current tree: build it with BH_FLAG_QUADRUPOLE or BH_FLAG_LEARNED first");
        a.qrec = c->qrec;
        a.acc_mag = c->acc_mag;
        if (p->flags & BH_FLAG_RELATIVE_MAC) {
            if (!(p->force_tolerance > 0.0f))
                return fail(c, BH_ERR_INVALID_ARGUMENT, "BH_FLAG_RELATIVE_MAC needs force_tolerance > 0");
            if (!c->accmag_valid) {
                // no |a| yet: a walk with the geometric test at theta first,
                // whose counts are then cleared (its output is overwritten)
                launch_walk(c->stream, c->n, a);
                launch_walk_counters_reset(c->stream, c->sim, c->ctr);
            }
            a.rel = 1;
            a.alpha = p->force_tolerance;
        }
    }
    launch_walk(c->stream, c->n, a);
    const int rc = check_launch(c, "walk");
    if (rc == BH_OK && a.qrec) c->accmag_valid = true;
    *wrote = rc == BH_OK && step_pos != nullptr;
    return rc;
}
}  // namespace

int bh_compute_force(bh_context* c, const bh_params* p) {
    if (!c || !p) return BH_ERR_INVALID_ARGUMENT;
    c->step_pos_valid = false;
    bool wrote;
    return compute_force(c, p, nullptr, &wrote);
}

int bh_merge_bodies(bh_context* c) {
    if (!c) return BH_ERR_INVALID_ARGUMENT;
    c->step_pos_valid = false;
    launch_merge(c->stream, c->bodies[1 - c->body_in], c->tasks, c->n, c->ctr, c->reserve, c->task_done);
    return check_launch(c, "merge");
}

namespace {
// The stages of one step on c->stream (no host synchronisation); advances
// the host's record of the buffer roles.
int step_enqueue(bh_context* c, const bh_params* p) {
    int rc;
    // the entry points below end the validity; this step's encode may use
    // what the previous step's walk left, if nothing else ran in between
    const bool use_pos = c->step_pos_valid;
    stage_begin(c, 0);
    if ((rc = bh_reset_values(c, p))) return rc;
    stage_begin(c, 1);
    // dynamic bounds are made by the partition's census pass (one read of
    // the bodies for both); the bounds stage then records no work
    // the walk's positions are the bodies' unless a merge followed it
    if ((rc = partition_dead_bodies(c, p->static_or_dynamic == BH_DYNAMIC,
                                    use_pos && !c->step_merged ? c->step_pos : nullptr)))
        return rc;
    stage_begin(c, 2);
    stage_begin(c, 3);
    if ((rc = generate_morton_codes(c, use_pos ? c->step_pos : nullptr))) return rc;
    stage_begin(c, 4);
    if ((rc = radix_sort(c, true))) return rc;
    stage_begin(c, 5);
    if ((rc = build_tree(c, p, true, use_pos ? c->step_pos : nullptr))) return rc;
    stage_begin(c, 6);
    bool wrote = false;
    c->step_pos_valid = false;
    if ((rc = compute_force(c, p, c->step_pos, &wrote))) return rc;
    stage_begin(c, 7);
    if (p->merging_collision_or_neither == BH_MERGING && (rc = bh_merge_bodies(c))) return rc;
    stage_begin(c, 8);
    if (c->timing && c->timed < bh_context::kTimedSteps) c->timed++;
    rc = bh_swap_body_buffers(c);
    c->step_pos_valid = rc == BH_OK && wrote;
    c->step_merged = p->merging_collision_or_neither == BH_MERGING;
    return rc;
}

bh_context::Roles roles_of(const bh_context* c) {
    return {c->body_in,  c->mi_in,          c->index[0],   c->index[1],      c->sort_pos[0],
            c->sort_pos[1], c->step_pos_valid, c->step_merged, c->quad_valid, c->accmag_valid, c->colors,
            c->fill_cap};
}

void set_roles(bh_context* c, const bh_context::Roles& r) {
    c->body_in = r.body_in;
    c->mi_in = r.mi_in;
    c->index[0] = r.index0;
    c->index[1] = r.index1;
    c->sort_pos[0] = r.sort0;
    c->sort_pos[1] = r.sort1;
    c->step_pos_valid = r.step_pos_valid;
    c->step_merged = r.step_merged;
    c->quad_valid = r.quad_valid;
    c->accmag_valid = r.accmag_valid;
    c->colors = r.colors;
    c->fill_cap = r.fill_cap;
}

// The step as an instantiated graph: the one captured for the current roles
// and parameters, captured now if there is none (capturing runs nothing; the
// roles are restored, the graph launched, and the roles it leaves applied).
int graph_step(bh_context* c, const bh_params* p) {
    const bh_context::Roles cur = roles_of(c);
    for (const auto& g : c->graphs)
        if (g.pre == cur && memcmp(&g.prm, p, sizeof *p) == 0) {
            HIPCHECK(c, hipGraphLaunch(g.exec, c->stream));
            set_roles(c, g.post);
            return BH_OK;
        }
    if (c->graphs.size() >= 8) {
        for (auto& g : c->graphs) (void)hipGraphExecDestroy(g.exec);
        c->graphs.clear();
    }
    HIPCHECK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    int rc = step_enqueue(c, p);
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(c->stream, &graph);
    const bh_context::Roles post = roles_of(c);
    set_roles(c, cur);
    if (rc || e != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc ? rc : fail(c, BH_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
    }
    bh_context::StepGraph g;
    g.pre = cur;
    g.post = post;
    g.prm = *p;
    const hipError_t ei = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ei != hipSuccess) return fail(c, BH_ERR_HIP, "hipGraphInstantiate: %s", hipGetErrorString(ei));
    c->graphs.push_back(g);
    HIPCHECK(c, hipGraphLaunch(g.exec, c->stream));
    set_roles(c, post);
    return BH_OK;
}
}  // namespace

int bh_step(bh_context* c, const bh_params* p) {
    if (!c || !p) return BH_ERR_INVALID_ARGUMENT;
    HIPCHECK(c, hipSetDevice(c->device));
    c->timing = (p->flags & BH_FLAG_TIMING) != 0;
    int rc;
    if (c->timing && (rc = ensure_events(c))) return rc;
    // a graph only in the steady state: the previous step's positions are
    // valid, the records' arrays exist, and (relative test) a previous |a|
    const bool quad = (p->flags & (BH_FLAG_QUADRUPOLE | BH_FLAG_LEARNED)) != 0;
    if ((p->flags & BH_FLAG_GRAPH) && !c->timing && c->step_pos_valid && (!quad || c->qrec) &&
        (!(p->flags & BH_FLAG_RELATIVE_MAC) || c->accmag_valid))
        return graph_step(c, p);
    return step_enqueue(c, p);
}

int bh_sync(bh_context* c) {
    if (!c) return BH_ERR_INVALID_ARGUMENT;
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    return BH_OK;
}

int bh_get_counters(bh_context* c, bh_counters* out) {
    if (!c || !out) return BH_ERR_INVALID_ARGUMENT;
    SimValues h;
    DevCounters d;
    HIPCHECK(c, hipMemcpyAsync(&h, c->sim, sizeof h, hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(c, hipMemcpyAsync(&d, c->ctr, sizeof d, hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    memset(out, 0, sizeof *out);
    out->num_bodies = h.numBodies;
    out->initial_num_bodies = h.initialNumBodies;
    out->just_died = d.died;
    out->out_of_bounds = h.outOfBounds;
    out->opened = d.opened;
    out->interactions = d.interactions;
    out->merge_tasks = d.nTasks;
    out->merges_applied = d.mergesApplied;
    out->tree_height = d.treeHeight;
    out->max_stack = d.maxStack;
    out->error_flags = d.errorFlags | (d.nTasks > c->n ? 2u : 0u);
    out->corrected = d.corrected;
    out->quad_terms = d.quadTerms;
    return BH_OK;
}

int bh_get_stage_times(bh_context* c, float* ms, int n) {
    if (!c || !ms || n < 0) return BH_ERR_INVALID_ARGUMENT;
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    for (int k = 0; k < n; k++) ms[k] = 0.0f;
    for (int st = 0; st < c->timed; st++) {
        for (int k = 0; k < n && k < BH_STAGE_COUNT; k++) {
            float t = 0.0f;
            hipEvent_t* e = c->ev + st * (BH_STAGE_COUNT + 1);
            if (hipEventElapsedTime(&t, e[k], e[k + 1]) == hipSuccess) ms[k] += t;
        }
    }
    if (n > BH_STAGE_COUNT) ms[BH_STAGE_COUNT] = (float)c->timed;
    c->timed = 0;
    return BH_OK;
}

// ------------------------------------------------------------------ learned correction
int bh_mlp_create(int device, bh_mlp** out) {
    if (!out) return BH_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    bh_mlp* m = new (std::nothrow) bh_mlp();
    if (!m) return BH_ERR_OUT_OF_MEMORY;
    m->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess) {
        delete m;
        return BH_ERR_HIP;
    }
    float** bufs[] = {&m->params, &m->grad, &m->m1, &m->m2};
    for (float** b : bufs)
        if (hipMalloc((void**)b, kMlpParams * sizeof(float)) != hipSuccess) { bh_mlp_destroy(m); return BH_ERR_OUT_OF_MEMORY; }
    if (hipMalloc((void**)&m->w, kBfTotal * sizeof(__bf16)) != hipSuccess ||
        hipMalloc((void**)&m->bias, kBiasTotal * sizeof(float)) != hipSuccess) {
        bh_mlp_destroy(m);
        return BH_ERR_OUT_OF_MEMORY;
    }
    for (float** b : bufs) (void)hipMemsetAsync(*b, 0, kMlpParams * sizeof(float), m->stream);
    launch_pack_weights(m->stream, m->params, m->w, m->bias);
    if (hipStreamSynchronize(m->stream) != hipSuccess) { bh_mlp_destroy(m); return BH_ERR_HIP; }
    *out = m;
    return BH_OK;
}

void bh_mlp_destroy(bh_mlp* m) {
    if (!m) return;
    (void)hipSetDevice(m->device);
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    void* ptrs[] = {m->params, m->grad, m->m1, m->m2, m->w, m->bias, m->x, m->t, m->z1, m->h1, m->z2, m->h2,
                    m->y, m->dy, m->partial, m->gemm_ws};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    if (m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
}

const char* bh_mlp_last_error(const bh_mlp* m) { return m ? m->err : "null network"; }

uint32_t bh_mlp_num_params(void) { return (uint32_t)kMlpParams; }

int bh_mlp_set_params(bh_mlp* m, const float* params, uint32_t n) {
    if (!m || !params || n != (uint32_t)kMlpParams) return BH_ERR_INVALID_ARGUMENT;
    MHIP(m, hipSetDevice(m->device));
    MHIP(m, hipMemcpyAsync(m->params, params, n * sizeof(float), hipMemcpyHostToDevice, m->stream));
    MHIP(m, hipMemsetAsync(m->m1, 0, n * sizeof(float), m->stream));
    MHIP(m, hipMemsetAsync(m->m2, 0, n * sizeof(float), m->stream));
    m->adam_t = 0;
    launch_pack_weights(m->stream, m->params, m->w, m->bias);
    MHIP(m, hipStreamSynchronize(m->stream));
    return BH_OK;
}

int bh_mlp_get_params(bh_mlp* m, float* params, uint32_t n) {
    if (!m || !params || n != (uint32_t)kMlpParams) return BH_ERR_INVALID_ARGUMENT;
    MHIP(m, hipMemcpyAsync(params, m->params, n * sizeof(float), hipMemcpyDeviceToHost, m->stream));
    MHIP(m, hipStreamSynchronize(m->stream));
    return BH_OK;
}

int bh_mlp_forward(bh_mlp* m, const float* x, uint32_t rows, float* y, int use_mfma) {
    if (!m || (!x && rows) || (!y && rows)) return BH_ERR_INVALID_ARGUMENT;
    if (rows == 0) return BH_OK;
    MHIP(m, hipSetDevice(m->device));
    int rc = mlp_workspace(m, rows);
    if (rc) return rc;
    MHIP(m, hipMemcpyAsync(m->x, x, (size_t)rows * kMlpIn * 4, hipMemcpyHostToDevice, m->stream));
    if (use_mfma) launch_mlp_forward_mfma(m->stream, m->x, rows, m->w, m->bias, m->y);
    else mlp_forward_fp32(m, rows);
    MHIP(m, hipGetLastError());
    MHIP(m, hipMemcpyAsync(y, m->y, (size_t)rows * kMlpOut * 4, hipMemcpyDeviceToHost, m->stream));
    MHIP(m, hipStreamSynchronize(m->stream));
    return BH_OK;
}

static int mlp_loss_value(bh_mlp* m, uint32_t nparts, uint32_t rows, float* loss) {
    float* part = new (std::nothrow) float[nparts];
    if (!part) return BH_ERR_OUT_OF_MEMORY;
    hipError_t e = hipMemcpyAsync(part, m->partial, nparts * 4, hipMemcpyDeviceToHost, m->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
    double acc = 0.0;
    for (uint32_t k = 0; k < nparts; k++) acc += part[k];
    delete[] part;
    *loss = (float)(acc / rows);
    return e == hipSuccess ? BH_OK : mfail(m, BH_ERR_HIP, "loss readback: %s", hipGetErrorString(e));
}

int bh_mlp_loss_grad(bh_mlp* m, const float* x, const float* t, uint32_t rows, float* grad, float* loss) {
    if (!m || !x || !t || !grad || !loss || rows == 0) return BH_ERR_INVALID_ARGUMENT;
    MHIP(m, hipSetDevice(m->device));
    int rc = mlp_workspace(m, rows);
    if (rc) return rc;
    MHIP(m, hipMemcpyAsync(m->x, x, (size_t)rows * kMlpIn * 4, hipMemcpyHostToDevice, m->stream));
    MHIP(m, hipMemcpyAsync(m->t, t, (size_t)rows * kMlpTarget * 4, hipMemcpyHostToDevice, m->stream));
    mlp_forward_fp32(m, rows);
    uint32_t nparts = mlp_backward_fp32(m, rows);
    MHIP(m, hipGetLastError());
    MHIP(m, hipMemcpyAsync(grad, m->grad, kMlpParams * 4, hipMemcpyDeviceToHost, m->stream));
    return mlp_loss_value(m, nparts, rows, loss);
}

int bh_mlp_train(bh_mlp* m, const float* x, const float* t, uint32_t rows, uint32_t epochs, uint32_t batch,
                 float lr, uint64_t seed, float* final_loss) {
    if (!m || !x || !t || rows == 0 || batch == 0 || !final_loss) return BH_ERR_INVALID_ARGUMENT;
    MHIP(m, hipSetDevice(m->device));
    if (batch > rows) batch = rows;
    int rc = mlp_workspace(m, batch);
    if (rc) return rc;
    float *dx = nullptr, *dt = nullptr;
    uint32_t* dperm = nullptr;
    MHIP(m, hipMalloc((void**)&dx, (size_t)rows * kMlpIn * 4));
    MHIP(m, hipMalloc((void**)&dt, (size_t)rows * kMlpTarget * 4));
    MHIP(m, hipMalloc((void**)&dperm, (size_t)rows * 4));
    uint32_t* perm = new (std::nothrow) uint32_t[rows];
    hipError_t e = perm ? hipSuccess : hipErrorOutOfMemory;
    if (e == hipSuccess) e = hipMemcpyAsync(dx, x, (size_t)rows * kMlpIn * 4, hipMemcpyHostToDevice, m->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(dt, t, (size_t)rows * kMlpTarget * 4, hipMemcpyHostToDevice, m->stream);
    uint64_t st = seed * 6364136223846793005ull + 1442695040888963407ull;
    uint32_t nparts = 0, last = batch;
    for (uint32_t ep = 0; ep < epochs && e == hipSuccess; ep++) {
        for (uint32_t i = 0; i < rows; i++) perm[i] = i;
        for (uint32_t i = rows - 1; i > 0; i--) {   // Fisher-Yates with a 64-bit LCG
            st = st * 6364136223846793005ull + 1442695040888963407ull;
            uint32_t j = (uint32_t)((st >> 33) % (i + 1));
            uint32_t tmp = perm[i]; perm[i] = perm[j]; perm[j] = tmp;
        }
        e = hipMemcpyAsync(dperm, perm, (size_t)rows * 4, hipMemcpyHostToDevice, m->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(m->stream);   // perm is reused next epoch
        for (uint32_t off = 0; off + batch <= rows && e == hipSuccess; off += batch) {
            launch_gather_rows(m->stream, dx, dperm, off, batch, kMlpIn, m->x);
            launch_gather_rows(m->stream, dt, dperm, off, batch, kMlpTarget, m->t);
            mlp_forward_fp32(m, batch);
            nparts = mlp_backward_fp32(m, batch);
            m->adam_t++;
            launch_adam(m->stream, m->params, m->grad, m->m1, m->m2, kMlpParams, lr, m->adam_t);
            last = batch;
            e = hipGetLastError();
        }
    }
    if (e == hipSuccess) launch_pack_weights(m->stream, m->params, m->w, m->bias);
    if (e == hipSuccess) e = hipGetLastError();
    if (e == hipSuccess) rc = mlp_loss_value(m, nparts, last, final_loss);
    else rc = mfail(m, BH_ERR_HIP, "train: %s", hipGetErrorString(e));
    (void)hipStreamSynchronize(m->stream);
    (void)hipFree(dx);
    (void)hipFree(dt);
    (void)hipFree(dperm);
    delete[] perm;
    return rc;
}

int bh_collect_training_pairs(bh_context* c, float theta, uint32_t stride, float softening,
                              uint32_t max_node_bodies, uint32_t cap, float* features, float* targets,
                              uint32_t* pairs_out, uint32_t* n_out) {
    if (!c || !features || !targets || !n_out || stride == 0 || cap == 0) return BH_ERR_INVALID_ARGUMENT;
    HIPCHECK(c, hipSetDevice(c->device));
    uint2* pairs = nullptr;
    uint32_t* count = nullptr;
    float *f = nullptr, *t = nullptr;
    SimValues h;
    HIPCHECK(c, hipMemcpyAsync(&h, c->sim, sizeof h, hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(c, hipMalloc((void**)&pairs, (size_t)cap * 8));
    HIPCHECK(c, hipMalloc((void**)&count, 8));   // pairs found, stack overflow flag
    hipError_t e = hipMemsetAsync(count, 0, 8, c->stream);
    if (e == hipSuccess) {
        launch_collect_pairs(c->stream, c->n, c->sim, c->tree.leaf_pos, c->tree.nodes, c->tree.node_aabb, theta,
                             softening, stride, pairs, cap, count);
        e = hipGetLastError();
    }
    uint32_t counts[2] = {0, 0};
    if (e == hipSuccess) e = hipMemcpyAsync(counts, count, 8, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    const uint32_t found = counts[0];
    if (e == hipSuccess && counts[1]) {
        (void)hipFree(pairs);
        (void)hipFree(count);
        return fail(c, BH_ERR_STACK_OVERFLOW, "collect pairs: a walk needed more than %d stack entries",
                    (int)kStackDepth);
    }
    const uint32_t np = found > cap ? cap : found;
    if (e == hipSuccess) e = hipMalloc((void**)&f, (size_t)(np ? np : 1) * kMlpIn * 4);
    if (e == hipSuccess) e = hipMalloc((void**)&t, (size_t)(np ? np : 1) * kMlpTarget * 4);
    if (e == hipSuccess) {
        launch_pair_labels(c->stream, pairs, np, c->tree.leaf_pos, c->tree.nodes, c->tree.node_aabb,
                           c->tree.node_range, (double)h.gravitationalConstant, softening, max_node_bodies, f, t);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(features, f, (size_t)np * kMlpIn * 4, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(targets, t, (size_t)np * kMlpTarget * 4, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess && pairs_out)
        e = hipMemcpyAsync(pairs_out, pairs, (size_t)np * 8, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(pairs);
    (void)hipFree(count);
    if (f) (void)hipFree(f);
    if (t) (void)hipFree(t);
    *n_out = found;
    return e == hipSuccess ? BH_OK : fail(c, BH_ERR_HIP, "collect pairs: %s", hipGetErrorString(e));
}

extern "C++" {
namespace bh {
void mlp_device_weights(const bh_mlp* m, const __bf16** w, const float** bias) {
    *w = m ? m->w : nullptr;
    *bias = m ? m->bias : nullptr;
}
}  // namespace bh
}

int bh_attach_correction(bh_context* c, const bh_mlp* m) {
    if (!c) return BH_ERR_INVALID_ARGUMENT;
    if (c->correction != m) {
        // captured steps hold the previous network's weight pointers
        for (auto& g : c->graphs) (void)hipGraphExecDestroy(g.exec);
        c->graphs.clear();
    }
    c->correction = m;
    return BH_OK;
}

int bh_selftest_math(bh_context* c, uint32_t* mismatches3) {
    if (!c || !mismatches3) return BH_ERR_INVALID_ARGUMENT;
    HIPCHECK(c, hipSetDevice(c->device));
    uint32_t* d = nullptr;
    HIPCHECK(c, hipMalloc((void**)&d, 12));
    hipError_t e = hipMemsetAsync(d, 0, 12, c->stream);
    if (e == hipSuccess) {
        launch_selftest_math(c->stream, d);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(mismatches3, d, 12, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(d);
    return e == hipSuccess ? BH_OK : fail(c, BH_ERR_HIP, "selftest: %s", hipGetErrorString(e));
}

int bh_direct_sum_f64(bh_context* c, const uint32_t* sample, uint32_t n_sample, double softening, double* out) {
    if (!c || (!sample && n_sample) || (!out && n_sample)) return BH_ERR_INVALID_ARGUMENT;
    if (n_sample == 0) return BH_OK;
    // one 256-thread block per sample: the grid must stay below 2^32 work-items
    if (n_sample > (1u << 22)) return fail(c, BH_ERR_INVALID_ARGUMENT, "n_sample %u > 2^22", n_sample);
    HIPCHECK(c, hipSetDevice(c->device));
    for (uint32_t k = 0; k < n_sample; k++)
        if (sample[k] >= c->n) return fail(c, BH_ERR_INVALID_ARGUMENT, "sample id %u >= n", sample[k]);
    uint32_t* d_sample = nullptr;
    double* d_out = nullptr;
    SimValues h;
    HIPCHECK(c, hipMemcpyAsync(&h, c->sim, sizeof h, hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    HIPCHECK(c, hipMalloc((void**)&d_sample, 4 * (size_t)n_sample));
    hipError_t e = hipMalloc((void**)&d_out, 24 * (size_t)n_sample);
    if (e == hipSuccess) e = hipMemcpyAsync(d_sample, sample, 4 * (size_t)n_sample, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) {
        launch_direct_sum_f64(c->stream, c->bodies[c->body_in], c->n, d_sample, n_sample,
                              (double)h.gravitationalConstant, softening, d_out);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(out, d_out, 24 * (size_t)n_sample, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(d_sample);
    if (d_out) (void)hipFree(d_out);
    return e == hipSuccess ? BH_OK : fail(c, BH_ERR_HIP, "direct sum: %s", hipGetErrorString(e));
}

int bh_direct_sum_f64_noalloc(bh_context* c, const uint32_t* sample, uint32_t n_sample, double softening,
                              double* out) {
    if (!c || (!sample && n_sample) || (!out && n_sample)) return BH_ERR_INVALID_ARGUMENT;
    if (n_sample > BH_NOALLOC_MAX_SAMPLE)
        return fail(c, BH_ERR_TOO_LARGE, "n_sample %u > %u, the static scratch of the no-allocation direct sum",
                    n_sample, BH_NOALLOC_MAX_SAMPLE);
    if (n_sample == 0) return BH_OK;
    HIPCHECK(c, hipSetDevice(c->device));
    for (uint32_t k = 0; k < n_sample; k++)
        if (sample[k] >= c->n) return fail(c, BH_ERR_INVALID_ARGUMENT, "sample id %u >= n", sample[k]);
    unsigned char* base = scratch_of(c->device);
    if (!base) return fail(c, BH_ERR_HIP, "no scratch address for device %d", c->device);
    uint32_t* d_sample = reinterpret_cast<uint32_t*>(base);
    double* d_out = reinterpret_cast<double*>(base + kNoAllocIdsBytes);
    SimValues h;
    HIPCHECK(c, hipMemcpyAsync(&h, c->sim, sizeof h, hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(c, hipMemcpyAsync(d_sample, sample, 4 * (size_t)n_sample, hipMemcpyHostToDevice, c->stream));
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    // every output word is written by the kernel: nothing in the shared
    // array has to start at zero
    launch_direct_sum_f64(c->stream, c->bodies[c->body_in], c->n, d_sample, n_sample,
                          (double)h.gravitationalConstant, softening, d_out);
    int rc = check_launch(c, "direct sum (no allocation)");
    if (rc) return rc;
    HIPCHECK(c, hipMemcpyAsync(out, d_out, 24 * (size_t)n_sample, hipMemcpyDeviceToHost, c->stream));
    HIPCHECK(c, hipStreamSynchronize(c->stream));
    return BH_OK;
}

}  // extern "C"
