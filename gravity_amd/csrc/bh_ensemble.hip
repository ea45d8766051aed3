// bh_ensemble.hip — host side of include/bh_ensemble.h: ensembles of
// independent simulations stepped with one launch per stage, the step
// captured as a hipGraph, and the autotuner.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <new>
#include <vector>

#include "../../include/bh_ensemble.h"
#include "bh_device.hpp"
#include "bh_kernels.hpp"

using namespace bh;

struct bh_ensemble {
    int device = 0;
    hipStream_t stream = nullptr;
    uint32_t S = 0, N = 0;
    std::vector<uint32_t> off, num;

    EnsSim* sims = nullptr;
    EnsTile* tiles = nullptr;
    uint32_t ntiles = 0;
    EnsTask* tasks = nullptr;
    uint32_t ntasks = 0;

    float4* bodies[2] = {nullptr, nullptr};
    uint64_t* keys[2] = {nullptr, nullptr};
    uint32_t* index[2] = {nullptr, nullptr};
    int body_in = 0, mi_in = 0;

    TreeArrays tree{};
    DevCounters* ctr = nullptr;
    EnsQueue* queues = nullptr;
    float* partials = nullptr;
    uint32_t* hist = nullptr;

    float theta = 0.5f, dt = 1e-3f, soft = 0.01f, G = 1.0f;
    uint32_t general = 0;      // BH_FLAG_GENERAL_WALK: always the walk's general loop

    // the corrected walk (BH_FLAG_QUADRUPOLE / BH_FLAG_LEARNED, optionally
    // BH_FLAG_RELATIVE_MAC): records over all simulations' trees, moment
    // prefix sums over all sorted positions, |a| by body id
    bool corrected = false, learned = false, relative = false;
    float alpha = 0.0f;
    const bh_mlp* correction = nullptr;
    float4* qrec = nullptr;

    float* acc_mag = nullptr;
    bool accmag_valid = false;   // a corrected walk has run on these bodies
    bh_ensemble_config cfg{8, 4, 192, 1};
    uint32_t walk_blocks = 0;
    int cus = 0;

    // two captured steps: the pointer roles repeat with period 2
    hipGraph_t graph[2] = {nullptr, nullptr};
    hipGraphExec_t exec[2] = {nullptr, nullptr};
    int start_body[2] = {0, 0}, start_mi[2] = {0, 0};
    bool graphs_valid = false;

    char err[512] = {0};
};

namespace {

int efail(bh_ensemble* e, int code, const char* fmt, ...) {
    if (e) {
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(e->err, sizeof e->err, fmt, ap);
        va_end(ap);
    }
    return code;
}

#define EHIP(e, expr)                                                                                  \
    do {                                                                                               \
        hipError_t e_ = (expr);                                                                        \
        if (e_ != hipSuccess) return efail((e), BH_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_));   \
    } while (0)

template <typename T>
int ealloc(bh_ensemble* e, T** p, size_t count) {
    size_t bytes = (count ? count : 1) * sizeof(T);
    hipError_t r = hipMalloc((void**)p, bytes);
    if (r != hipSuccess) return efail(e, BH_ERR_OUT_OF_MEMORY, "hipMalloc(%zu): %s", bytes, hipGetErrorString(r));
    return BH_OK;
}

uint32_t sort_passes(uint32_t bits) { return (63u + bits - 1u) / bits; }

bool valid_config(const bh_ensemble_config& c) {
    const bool bits = c.radix_bits >= 5 && c.radix_bits <= 8;
    const bool waves = c.walk_waves == 1 || c.walk_waves == 2 || c.walk_waves == 4 || c.walk_waves == 8;
    const bool depth = c.stack_depth == 64 || c.stack_depth == 128 || c.stack_depth == 192;
    return bits && waves && depth && c.use_graph <= 1;
}

int walk_grid(bh_ensemble* e, int waves, int depth, uint32_t* blocks) {
    int per_cu = 0;
    if (ens_walk_occupancy(waves, depth, &per_cu) != 0)
        return efail(e, BH_ERR_HIP, "occupancy query for the walk (%d waves, depth %d) failed", waves, depth);
    if (per_cu < 1) per_cu = 1;
    *blocks = (uint32_t)per_cu * (uint32_t)e->cus;
    return BH_OK;
}

void drop_graphs(bh_ensemble* e) {
    for (int k = 0; k < 2; k++) {
        if (e->exec[k]) (void)hipGraphExecDestroy(e->exec[k]);
        if (e->graph[k]) (void)hipGraphDestroy(e->graph[k]);
        e->exec[k] = nullptr;
        e->graph[k] = nullptr;
    }
    e->graphs_valid = false;
}

// The stages of one step, enqueued on e->stream; advances the host's record
// of which buffers are current.
int enqueue_sort(bh_ensemble* e, uint32_t bits) {
    const uint32_t passes = sort_passes(bits);
    for (uint32_t p = 0; p < passes; p++) {
        const int mi = e->mi_in;
        if (launch_ens_sort_pass(e->stream, (int)bits, e->sims, e->S, e->tiles, e->ntiles, e->keys[mi], e->index[mi],
                                 e->keys[1 - mi], e->index[1 - mi], p * bits, e->hist))
            return efail(e, BH_ERR_INVALID_ARGUMENT, "radix bits %u", bits);
        e->mi_in = 1 - mi;
    }
    return BH_OK;
}

int enqueue_walk(bh_ensemble* e, int waves, int depth, uint32_t blocks) {
    const int b = e->body_in;
    // the relative test reads the |a| of a previous corrected walk: the first
    // walk uses the geometric test at theta
    const uint32_t rel = e->relative && e->accmag_valid ? 1u : 0u;
    if (launch_ens_walk(e->stream, waves, depth, blocks, e->sims, e->tasks, e->queues, e->ctr,
                        e->bodies[b], e->bodies[1 - b], e->index[e->mi_in], e->tree.leaf_pos, e->tree.nodes, e->N,
                        e->G, e->theta, e->dt, e->soft, e->general, e->corrected ? e->qrec : nullptr, e->acc_mag,
                        rel, e->alpha))
        return efail(e, BH_ERR_INVALID_ARGUMENT, "walk shape %d x %d", waves, depth);
    return BH_OK;
}

int enqueue_step(bh_ensemble* e) {
    const int b = e->body_in;
    launch_ens_reset(e->stream, e->ctr, e->queues);
    launch_ens_bounds(e->stream, e->sims, e->tiles, e->ntiles, e->bodies[b], e->partials);
    launch_ens_encode(e->stream, e->sims, e->tiles, e->ntiles, e->bodies[b], e->index[e->mi_in], e->keys[e->mi_in]);
    int rc = enqueue_sort(e, e->cfg.radix_bits);
    if (rc) return rc;
    launch_ens_tree(e->stream, e->sims, e->tiles, e->ntiles, e->N, e->ctr, e->bodies[b], e->index[e->mi_in],
                    e->keys[e->mi_in], e->tree, e->G);
    if (e->corrected) {
        if (e->learned && !e->correction)
            return efail(e, BH_ERR_INVALID_ARGUMENT, "BH_FLAG_LEARNED without an attached network");
        const __bf16* w = nullptr;
        const float* bias = nullptr;
        if (e->learned) mlp_device_weights(e->correction, &w, &bias);
        launch_ens_node_quadrupoles(e->stream, e->N, e->N - e->S, e->G, e->ctr, e->tree, w, bias, e->qrec);
    }
    if ((rc = enqueue_walk(e, (int)e->cfg.walk_waves, (int)e->cfg.stack_depth, e->walk_blocks))) return rc;
    e->body_in = 1 - b;
    return BH_OK;
}

int capture_graphs(bh_ensemble* e) {
    drop_graphs(e);
    const int b0 = e->body_in, m0 = e->mi_in;
    // enqueueing a step advances the host's buffer roles although nothing
    // runs: every return restores them
    auto bail = [&](int rc) {
        e->body_in = b0;
        e->mi_in = m0;
        drop_graphs(e);
        return rc;
    };
    for (int k = 0; k < 2; k++) {
        e->start_body[k] = e->body_in;
        e->start_mi[k] = e->mi_in;
        hipError_t r = hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal);
        if (r != hipSuccess) return bail(efail(e, BH_ERR_HIP, "hipStreamBeginCapture: %s", hipGetErrorString(r)));
        int rc = enqueue_step(e);
        hipGraph_t g = nullptr;
        r = hipStreamEndCapture(e->stream, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return bail(rc);
        }
        if (r != hipSuccess) return bail(efail(e, BH_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(r)));
        e->graph[k] = g;
        r = hipGraphInstantiate(&e->exec[k], g, nullptr, nullptr, 0);
        if (r != hipSuccess) return bail(efail(e, BH_ERR_HIP, "hipGraphInstantiate: %s", hipGetErrorString(r)));
    }
    // nothing ran: the current buffers are those before the two captured steps
    e->body_in = b0;
    e->mi_in = m0;
    if (e->body_in != e->start_body[0] || e->mi_in != e->start_mi[0])
        return efail(e, BH_ERR_HIP, "graph capture: inconsistent buffer roles");
    e->graphs_valid = true;
    return BH_OK;
}

int graph_step(bh_ensemble* e) {
    if (!e->graphs_valid) {
        int rc = capture_graphs(e);
        if (rc) return rc;
    }
    int k = -1;
    for (int j = 0; j < 2; j++)
        if (e->start_body[j] == e->body_in && e->start_mi[j] == e->mi_in) k = j;
    if (k < 0) {
        int rc = capture_graphs(e);
        if (rc) return rc;
        k = 0;
    }
    EHIP(e, hipGraphLaunch(e->exec[k], e->stream));
    // graph k leaves the roles graph 1 - k starts from (the roles repeat
    // every two steps: bodies swap once, the sort flips its buffers an even
    // number of times over two steps)
    e->body_in = e->start_body[1 - k];
    e->mi_in = e->start_mi[1 - k];
    return BH_OK;
}

}  // namespace

extern "C" {

const char* bh_ensemble_last_error(const bh_ensemble* e) { return e ? e->err : "null ensemble"; }

void bh_ensemble_destroy(bh_ensemble* e) {
    if (!e) return;
    (void)hipSetDevice(e->device);
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    drop_graphs(e);
    void* ptrs[] = {e->sims, e->tiles, e->tasks, e->bodies[0], e->bodies[1], e->keys[0], e->keys[1], e->index[0],
                    e->index[1], e->tree.leaf_pos, e->tree.leaf_parent, e->tree.walk, e->tree.node_aabb,
                    e->tree.node_parent, e->tree.node_range, e->tree.meet, e->tree.xchg, e->tree.frontier,
                    e->tree.frontier_count, e->ctr, e->queues,
                    e->partials, e->hist, e->qrec, e->tree.node_mom, e->tree.xchg_mom, e->tree.frontier_mom,
                    e->acc_mag};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
}

int bh_ensemble_create(int device, uint32_t n_sims, const uint32_t* counts, const float* bodies,
                       const float* bounds, const bh_params* params, const bh_ensemble_config* cfg,
                       bh_ensemble** out) {
    if (!out || !counts || !bodies || !bounds || !params || n_sims == 0) return BH_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    if (cfg && !valid_config(*cfg)) return BH_ERR_INVALID_ARGUMENT;
    if (params->merging_collision_or_neither != BH_NEITHER || params->static_or_dynamic != BH_DYNAMIC ||
        (params->flags & BH_FLAG_FILL_CAP))
        return BH_ERR_UNSUPPORTED;
    const bool corrected = (params->flags & (BH_FLAG_QUADRUPOLE | BH_FLAG_LEARNED)) != 0;
    const bool relative = corrected && (params->flags & BH_FLAG_RELATIVE_MAC);
    if (corrected && !(params->softening >= 1e-14f)) return BH_ERR_UNSUPPORTED;
    if (relative && !(params->force_tolerance > 0.0f)) return BH_ERR_INVALID_ARGUMENT;
    uint64_t total = 0;
    for (uint32_t s = 0; s < n_sims; s++) {
        if (counts[s] < 2) return BH_ERR_INVALID_ARGUMENT;
        total += counts[s];
    }
    if (total > (1u << 30)) return BH_ERR_TOO_LARGE;
    for (uint64_t i = 0; i < total; i++)
        if (!(bodies[8 * i + 3] != 0.0f)) return BH_ERR_UNSUPPORTED;   // every body alive
    bh_ensemble* e = new (std::nothrow) bh_ensemble();
    if (!e) return BH_ERR_OUT_OF_MEMORY;
    e->device = device;
    e->S = n_sims;
    e->N = (uint32_t)total;
    if (cfg) e->cfg = *cfg;
    e->theta = params->theta;
    e->dt = params->dt;
    e->soft = params->softening;
    e->general = (params->flags & BH_FLAG_GENERAL_WALK) ? 1u : 0u;
    e->corrected = corrected;
    e->learned = (params->flags & BH_FLAG_LEARNED) != 0;
    e->relative = relative;
    e->alpha = params->force_tolerance;
    float G = params->gravitational_constant;
    if (G == 0.0f && params->unit_mass != 0.0f) {
        // common/common.glsl:346-352, in FP32 as the shader does
        G = 6.67430e-11f * powf(params->unit_len, -3.0f) * powf(params->unit_time, 2.0f) *
            powf(params->unit_mass, 1.0f);
    }
    e->G = G;
    int rc = BH_OK;
    const uint32_t N = e->N, S = e->S;
    std::vector<EnsSim> hs(S);
    std::vector<EnsTile> ht;
    std::vector<EnsTask> hk;
    e->off.resize(S);
    e->num.assign(counts, counts + S);
    uint32_t o = 0;
    for (uint32_t s = 0; s < S; s++) {
        EnsSim& x = hs[s];
        memset(&x, 0, sizeof x);
        x.off = o;
        x.num = counts[s];
        x.ibase = o - s;
        x.tile0 = (uint32_t)ht.size();
        for (uint32_t t = 0; t < counts[s]; t += kEnsTile)
            ht.push_back({s, o + t, std::min(kEnsTile, counts[s] - t), 0u});
        x.ntiles = (uint32_t)ht.size() - x.tile0;
        for (int c = 0; c < 3; c++) {
            x.bmin[c] = bounds[6 * s + c];
            x.bmax[c] = bounds[6 * s + 3 + c];
        }
        e->off[s] = o;
        o += counts[s];
    }
    // walk tasks in simulation order, split into one contiguous share per XCD
    // queue of about equal cost (64-body tasks cost ~ log2 of their
    // simulation's size: the depth of its walks)
    std::vector<double> cost_prefix(1, 0.0);
    for (uint32_t s = 0; s < S; s++)
        for (uint32_t t = 0; t < counts[s]; t += kWave) {
            const uint32_t c = std::min(64u, counts[s] - t);
            hk.push_back({s, hs[s].off + t, c, 0u});
            cost_prefix.push_back(cost_prefix.back() + c * std::log2((double)counts[s] + 1.0));
        }
    std::vector<EnsQueue> hq(kEnsQueues);
    {
        uint32_t begin = 0;
        for (int q = 0; q < kEnsQueues; q++) {
            const double goal = cost_prefix.back() * (q + 1) / kEnsQueues;
            uint32_t end = begin;
            while (end < hk.size() && (q == kEnsQueues - 1 || cost_prefix[end + 1] <= goal)) end++;
            memset(&hq[q], 0, sizeof(EnsQueue));
            hq[q].begin = begin;
            hq[q].count = end - begin;
            begin = end;
        }
    }
    e->ntiles = (uint32_t)ht.size();
    e->ntasks = (uint32_t)hk.size();
    std::vector<uint32_t> iota(N);
    for (uint32_t i = 0; i < N; i++) iota[i] = i;
#define TRY(x)                     \
    do {                           \
        rc = (x);                  \
        if (rc != BH_OK) goto bad; \
    } while (0)
    {
        hipError_t r = hipSetDevice(device);
        if (r != hipSuccess) { rc = efail(e, BH_ERR_HIP, "hipSetDevice: %s", hipGetErrorString(r)); goto bad; }
        r = hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking);
        if (r != hipSuccess) { rc = efail(e, BH_ERR_HIP, "hipStreamCreate: %s", hipGetErrorString(r)); goto bad; }
        r = hipDeviceGetAttribute(&e->cus, hipDeviceAttributeMultiprocessorCount, device);
        if (r != hipSuccess || e->cus < 1) { rc = efail(e, BH_ERR_HIP, "CU count"); goto bad; }
        TRY(ealloc(e, &e->sims, S));
        TRY(ealloc(e, &e->tiles, e->ntiles));
        TRY(ealloc(e, &e->tasks, e->ntasks));
        for (int k = 0; k < 2; k++) {
            TRY(ealloc(e, &e->bodies[k], 2 * (size_t)N));
            TRY(ealloc(e, &e->keys[k], (size_t)N));
            TRY(ealloc(e, &e->index[k], (size_t)N));
        }
        const size_t ni = (size_t)N - S;
        TreeArrays& T = e->tree;
        TRY(ealloc(e, &T.leaf_pos, (size_t)N));
        TRY(ealloc(e, &T.leaf_parent, (size_t)N));
        TRY(ealloc(e, &T.walk, (size_t)N + ni));
        T.nodes = T.walk + N;
        TRY(ealloc(e, &T.node_aabb, 2 * ni));
        TRY(ealloc(e, &T.node_parent, ni));
        TRY(ealloc(e, &T.node_range, ni));
        TRY(ealloc(e, &T.meet, ni));
        TRY(ealloc(e, &T.xchg, 8 * ni));
        TRY(ealloc(e, &T.frontier, 4 * (size_t)frontier_cap(N, ens_tree_blocks(e->ntiles, corrected))));
        TRY(ealloc(e, &T.frontier_count, 1));
        TRY(ealloc(e, &e->ctr, 1));
        TRY(ealloc(e, &e->queues, kEnsQueues));
        TRY(ealloc(e, &e->partials, 6 * (size_t)e->ntiles));
        TRY(ealloc(e, &e->hist, 256 * (size_t)e->ntiles));
        if (corrected) {
            TRY(ealloc(e, &e->qrec, 4 * corr_records(N)));
            // the trees also build every node's second moment
            TRY(ealloc(e, &T.node_mom, 2 * ni));
            TRY(ealloc(e, &T.xchg_mom, 6 * ni));
            TRY(ealloc(e, &T.frontier_mom, 3 * (size_t)frontier_cap(N, ens_tree_blocks(e->ntiles, true))));
            T.corr_leaf = e->qrec;
            TRY(ealloc(e, &e->acc_mag, (size_t)N));
            // the quadrupole halves of the leaf records stay zero
            r = hipMemsetAsync(e->qrec, 0, 64 * corr_records(N), e->stream);
            if (r != hipSuccess) { rc = efail(e, BH_ERR_HIP, "records: %s", hipGetErrorString(r)); goto bad; }
        }
        TRY(walk_grid(e, (int)e->cfg.walk_waves, (int)e->cfg.stack_depth, &e->walk_blocks));
        const hipMemcpyKind H2D = hipMemcpyHostToDevice;
        if (hipMemcpyAsync(e->sims, hs.data(), S * sizeof(EnsSim), H2D, e->stream) != hipSuccess ||
            hipMemcpyAsync(e->tiles, ht.data(), ht.size() * sizeof(EnsTile), H2D, e->stream) != hipSuccess ||
            hipMemcpyAsync(e->tasks, hk.data(), hk.size() * sizeof(EnsTask), H2D, e->stream) != hipSuccess ||
            hipMemcpyAsync(e->queues, hq.data(), kEnsQueues * sizeof(EnsQueue), H2D, e->stream) != hipSuccess ||
            hipMemcpyAsync(e->bodies[0], bodies, 32 * (size_t)N, H2D, e->stream) != hipSuccess ||
            hipMemcpyAsync(e->index[0], iota.data(), 4 * (size_t)N, H2D, e->stream) != hipSuccess ||
            hipMemsetAsync(e->bodies[1], 0, 32 * (size_t)N, e->stream) != hipSuccess ||
            hipMemsetAsync(e->keys[0], 0, 8 * (size_t)N, e->stream) != hipSuccess ||
            hipMemsetAsync(e->ctr, 0, sizeof(DevCounters), e->stream) != hipSuccess ||
            hipMemsetAsync(T.meet, 0xFF, 4 * ni, e->stream) != hipSuccess) {
            rc = efail(e, BH_ERR_HIP, "upload");
            goto bad;
        }
        r = hipStreamSynchronize(e->stream);
        if (r != hipSuccess) { rc = efail(e, BH_ERR_HIP, "create sync: %s", hipGetErrorString(r)); goto bad; }
    }
#undef TRY
    *out = e;
    return BH_OK;
bad:
    fprintf(stderr, "bh_ensemble_create: %s\n", e->err);
    bh_ensemble_destroy(e);
    return rc;
}

int bh_ensemble_get_config(const bh_ensemble* e, bh_ensemble_config* cfg) {
    if (!e || !cfg) return BH_ERR_INVALID_ARGUMENT;
    *cfg = e->cfg;
    return BH_OK;
}

int bh_ensemble_set_config(bh_ensemble* e, const bh_ensemble_config* cfg) {
    if (!e || !cfg || !valid_config(*cfg)) return BH_ERR_INVALID_ARGUMENT;
    EHIP(e, hipSetDevice(e->device));
    uint32_t blocks = 0;
    int rc = walk_grid(e, (int)cfg->walk_waves, (int)cfg->stack_depth, &blocks);
    if (rc) return rc;
    EHIP(e, hipStreamSynchronize(e->stream));
    drop_graphs(e);
    e->cfg = *cfg;
    e->walk_blocks = blocks;
    return BH_OK;
}

int bh_ensemble_step(bh_ensemble* e, uint32_t steps) {
    if (!e) return BH_ERR_INVALID_ARGUMENT;
    EHIP(e, hipSetDevice(e->device));
    for (uint32_t k = 0; k < steps; k++) {
        if (e->relative && !e->accmag_valid) {
            // the first corrected walk uses the geometric test (no |a| yet);
            // run directly, and capture graphs from the next step on
            int rc = enqueue_step(e);
            if (rc) return rc;
            hipError_t r = hipGetLastError();
            if (r != hipSuccess) return efail(e, BH_ERR_HIP, "step launch: %s", hipGetErrorString(r));
            e->accmag_valid = true;
            drop_graphs(e);
            continue;
        }
        int rc = e->cfg.use_graph ? graph_step(e) : enqueue_step(e);
        if (rc) return rc;
        if (!e->cfg.use_graph) {
            hipError_t r = hipGetLastError();
            if (r != hipSuccess) return efail(e, BH_ERR_HIP, "step launch: %s", hipGetErrorString(r));
        }
    }
    return BH_OK;
}

int bh_ensemble_sync(bh_ensemble* e) {
    if (!e) return BH_ERR_INVALID_ARGUMENT;
    EHIP(e, hipStreamSynchronize(e->stream));
    return BH_OK;
}

int bh_ensemble_download_bodies(bh_ensemble* e, float* bodies) {
    if (!e || !bodies) return BH_ERR_INVALID_ARGUMENT;
    EHIP(e, hipMemcpyAsync(bodies, e->bodies[e->body_in], 32 * (size_t)e->N, hipMemcpyDeviceToHost, e->stream));
    EHIP(e, hipStreamSynchronize(e->stream));
    return BH_OK;
}

int bh_ensemble_download_bounds(bh_ensemble* e, float* bounds) {
    if (!e || !bounds) return BH_ERR_INVALID_ARGUMENT;
    std::vector<EnsSim> hs(e->S);
    EHIP(e, hipMemcpyAsync(hs.data(), e->sims, e->S * sizeof(EnsSim), hipMemcpyDeviceToHost, e->stream));
    EHIP(e, hipStreamSynchronize(e->stream));
    for (uint32_t s = 0; s < e->S; s++)
        for (int c = 0; c < 3; c++) {
            bounds[6 * s + c] = hs[s].bmin[c];
            bounds[6 * s + 3 + c] = hs[s].bmax[c];
        }
    return BH_OK;
}

int bh_ensemble_download_sorted(bh_ensemble* e, uint64_t* keys, uint32_t* index) {
    if (!e || (!keys && !index)) return BH_ERR_INVALID_ARGUMENT;
    if (keys) EHIP(e, hipMemcpyAsync(keys, e->keys[e->mi_in], 8 * (size_t)e->N, hipMemcpyDeviceToHost, e->stream));
    if (index) EHIP(e, hipMemcpyAsync(index, e->index[e->mi_in], 4 * (size_t)e->N, hipMemcpyDeviceToHost, e->stream));
    EHIP(e, hipStreamSynchronize(e->stream));
    return BH_OK;
}

int bh_ensemble_download_internal_nodes(bh_ensemble* e, void* nodes) {
    if (!e || !nodes) return BH_ERR_INVALID_ARGUMENT;
    const uint32_t ni = e->N - e->S;
    NodeStd430* tmp = nullptr;
    EHIP(e, hipMalloc((void**)&tmp, (size_t)ni * sizeof(NodeStd430)));
    launch_ens_nodes_std430(e->stream, e->tree, ni, tmp);
    hipError_t r = hipMemcpyAsync(nodes, tmp, (size_t)ni * sizeof(NodeStd430), hipMemcpyDeviceToHost, e->stream);
    if (r == hipSuccess) r = hipStreamSynchronize(e->stream);
    (void)hipFree(tmp);
    if (r != hipSuccess) return efail(e, BH_ERR_HIP, "download nodes: %s", hipGetErrorString(r));
    return BH_OK;
}

int bh_ensemble_get_counters(bh_ensemble* e, bh_counters* out) {
    if (!e || !out) return BH_ERR_INVALID_ARGUMENT;
    DevCounters d;
    EHIP(e, hipMemcpyAsync(&d, e->ctr, sizeof d, hipMemcpyDeviceToHost, e->stream));
    EHIP(e, hipStreamSynchronize(e->stream));
    memset(out, 0, sizeof *out);
    out->num_bodies = e->N;
    out->initial_num_bodies = e->N;
    out->opened = d.opened;
    out->interactions = d.interactions;
    out->tree_height = d.treeHeight;
    out->max_stack = d.maxStack;
    out->error_flags = d.errorFlags;
    out->corrected = d.corrected;
    out->quad_terms = d.quadTerms;
    return BH_OK;
}

int bh_ensemble_attach_correction(bh_ensemble* e, const bh_mlp* m) {
    if (!e) return BH_ERR_INVALID_ARGUMENT;
    if (!e->learned) return efail(e, BH_ERR_INVALID_ARGUMENT, "the ensemble was created without BH_FLAG_LEARNED");
    EHIP(e, hipSetDevice(e->device));
    EHIP(e, hipStreamSynchronize(e->stream));
    drop_graphs(e);   // captured steps hold the network's weight pointers
    e->correction = m;
    return BH_OK;
}

int bh_ensemble_autotune(bh_ensemble* e, bh_ensemble_config* chosen, float* timings) {
    if (!e) return BH_ERR_INVALID_ARGUMENT;
    EHIP(e, hipSetDevice(e->device));
    EHIP(e, hipStreamSynchronize(e->stream));
    float local[16];
    float* tm = timings ? timings : local;
    for (int k = 0; k < 16; k++) tm[k] = -1.0f;
    // the state every trial starts from: bounds and the previous sorted order
    EnsSim* saved_sims = nullptr;
    uint32_t* saved_index = nullptr;
    float* saved_amag = nullptr;   // the relative test's |a| (every walk trial rewrites it)
    const bool keep_amag = e->relative && e->accmag_valid;
    hipEvent_t ev[2] = {nullptr, nullptr};
    const int b0 = e->body_in, m0 = e->mi_in;
    int rc = BH_OK;
    bh_ensemble_config best = e->cfg;
    auto elapsed = [&](float* ms) -> int {
        EHIP(e, hipEventSynchronize(ev[1]));
        EHIP(e, hipEventElapsedTime(ms, ev[0], ev[1]));
        return BH_OK;
    };
    auto prepare = [&]() -> int {
        e->mi_in = m0;
        EHIP(e, hipMemcpyAsync(e->sims, saved_sims, e->S * sizeof(EnsSim), hipMemcpyDeviceToDevice, e->stream));
        EHIP(e, hipMemcpyAsync(e->index[m0], saved_index, 4 * (size_t)e->N, hipMemcpyDeviceToDevice, e->stream));
        launch_ens_reset(e->stream, e->ctr, e->queues);
        launch_ens_bounds(e->stream, e->sims, e->tiles, e->ntiles, e->bodies[b0], e->partials);
        launch_ens_encode(e->stream, e->sims, e->tiles, e->ntiles, e->bodies[b0], e->index[m0], e->keys[m0]);
        return BH_OK;
    };
    do {
        if ((rc = ealloc(e, &saved_sims, e->S)) || (rc = ealloc(e, &saved_index, (size_t)e->N))) break;
        if (keep_amag) {
            if ((rc = ealloc(e, &saved_amag, (size_t)e->N))) break;
            if (hipMemcpyAsync(saved_amag, e->acc_mag, 4 * (size_t)e->N, hipMemcpyDeviceToDevice, e->stream) !=
                hipSuccess) {
                rc = efail(e, BH_ERR_HIP, "autotune save");
                break;
            }
        }
        if (hipEventCreate(&ev[0]) != hipSuccess || hipEventCreate(&ev[1]) != hipSuccess) {
            rc = efail(e, BH_ERR_HIP, "events");
            break;
        }
        if (hipMemcpyAsync(saved_sims, e->sims, e->S * sizeof(EnsSim), hipMemcpyDeviceToDevice, e->stream) !=
                hipSuccess ||
            hipMemcpyAsync(saved_index, e->index[m0], 4 * (size_t)e->N, hipMemcpyDeviceToDevice, e->stream) !=
                hipSuccess) {
            rc = efail(e, BH_ERR_HIP, "autotune save");
            break;
        }
        // digit width of the segmented sort
        float best_ms = 1e30f;
        for (uint32_t bits = 5; bits <= 8 && !rc; bits++) {
            float t_min = 1e30f;
            for (int rep = 0; rep < 3 && !rc; rep++) {
                if ((rc = prepare())) break;
                (void)hipEventRecord(ev[0], e->stream);
                if ((rc = enqueue_sort(e, bits))) break;
                (void)hipEventRecord(ev[1], e->stream);
                float ms = 0.0f;
                if ((rc = elapsed(&ms))) break;
                t_min = std::min(t_min, ms);
            }
            tm[bits - 5] = t_min;
            if (t_min < best_ms) {
                best_ms = t_min;
                best.radix_bits = bits;
            }
        }
        if (rc) break;
        // walk shape on the tree of the current state
        if ((rc = prepare()) || (rc = enqueue_sort(e, best.radix_bits))) break;
        launch_ens_tree(e->stream, e->sims, e->tiles, e->ntiles, e->N, e->ctr, e->bodies[b0], e->index[e->mi_in],
                        e->keys[e->mi_in], e->tree, e->G);
        if (e->corrected) {   // the node records of this tree
            const __bf16* w = nullptr;
            const float* bias = nullptr;
            if (e->learned) {
                if (!e->correction) {
                    rc = efail(e, BH_ERR_INVALID_ARGUMENT, "BH_FLAG_LEARNED without an attached network");
                    break;
                }
                mlp_device_weights(e->correction, &w, &bias);
            }
            launch_ens_node_quadrupoles(e->stream, e->N, e->N - e->S, e->G, e->ctr, e->tree, w, bias, e->qrec);
        }
        DevCounters d;
        if (hipMemcpyAsync(&d, e->ctr, sizeof d, hipMemcpyDeviceToHost, e->stream) != hipSuccess ||
            hipStreamSynchronize(e->stream) != hipSuccess) {
            rc = efail(e, BH_ERR_HIP, "autotune tree");
            break;
        }
        const uint32_t need = d.treeHeight + 2;
        const int waves_c[4] = {1, 2, 4, 8}, depth_c[3] = {64, 128, 192};
        best_ms = 1e30f;
        for (int wi = 0; wi < 4 && !rc; wi++)
            for (int di = 0; di < 3 && !rc; di++) {
                if ((uint32_t)depth_c[di] < need && depth_c[di] != 192) continue;
                uint32_t blocks = 0;
                if ((rc = walk_grid(e, waves_c[wi], depth_c[di], &blocks))) break;
                float t_min = 1e30f;
                for (int rep = 0; rep < 3 && !rc; rep++) {
                    if (saved_amag &&
                        hipMemcpyAsync(e->acc_mag, saved_amag, 4 * (size_t)e->N, hipMemcpyDeviceToDevice,
                                       e->stream) != hipSuccess) {
                        rc = efail(e, BH_ERR_HIP, "autotune restore");
                        break;
                    }
                    launch_ens_reset(e->stream, e->ctr, e->queues);
                    (void)hipEventRecord(ev[0], e->stream);
                    if ((rc = enqueue_walk(e, waves_c[wi], depth_c[di], blocks))) break;
                    (void)hipEventRecord(ev[1], e->stream);
                    float ms = 0.0f;
                    if ((rc = elapsed(&ms))) break;
                    t_min = std::min(t_min, ms);
                }
                tm[4 + 3 * wi + di] = t_min;
                if (t_min < best_ms) {
                    best_ms = t_min;
                    best.walk_waves = (uint32_t)waves_c[wi];
                    best.stack_depth = (uint32_t)depth_c[di];
                }
            }
    } while (0);
    // restore the state the trials started from
    e->mi_in = m0;
    e->body_in = b0;
    if (saved_sims && saved_index) {
        (void)hipMemcpyAsync(e->sims, saved_sims, e->S * sizeof(EnsSim), hipMemcpyDeviceToDevice, e->stream);
        (void)hipMemcpyAsync(e->index[m0], saved_index, 4 * (size_t)e->N, hipMemcpyDeviceToDevice, e->stream);
    }
    if (saved_amag)
        (void)hipMemcpyAsync(e->acc_mag, saved_amag, 4 * (size_t)e->N, hipMemcpyDeviceToDevice, e->stream);
    (void)hipStreamSynchronize(e->stream);
    if (saved_sims) (void)hipFree(saved_sims);
    if (saved_index) (void)hipFree(saved_index);
    if (saved_amag) (void)hipFree(saved_amag);
    for (hipEvent_t x : ev)
        if (x) (void)hipEventDestroy(x);
    if (rc) return rc;
    rc = bh_ensemble_set_config(e, &best);
    if (rc) return rc;
    if (chosen) *chosen = best;
    return BH_OK;
}

}  // extern "C"
