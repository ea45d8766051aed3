/*
 * bh_abi.h — C ABI of the MI355X Barnes-Hut step (libgravity_amd.so).
 *
 * This is the drop-in boundary for the reference's compute-shader pipeline.
 * The Java host of the reference drives that pipeline through
 * ComputeProgram.run() (gpu/ComputeProgram.java:51-55, gpu/GLSLProgram.java:72-92)
 * on 16 SSBO bindings (shaders/common/common.glsl:80-120) from
 * BarnesHut.step() (simulation/BarnesHut.java:99-164). Every entry point below
 * replaces one of those calls; buffers keep the reference's binding ids and
 * std430 layouts so the Java mirrors Body (gpu/Body.java:22-25), Node
 * (gpu/Node.java:31-41) and SimulationValues (gpu/GPU.java:1096-1143) stay valid.
 *
 * Conventions: plain pointers and sizes, no C++ or framework types; every
 * function returns BH_OK (0) or a negative BH_ERR_* code and never lets an
 * exception or abort cross the boundary; bh_last_error() gives the text.
 * A context is single-threaded and stream-ordered: calls enqueue work on the
 * context's HIP stream and return without waiting, except bh_download,
 * bh_sync, bh_get_counters and bh_get_stage_times, which wait for the stream.
 */
#ifndef BH_ABI_H
#define BH_ABI_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (the reference prints GL errors and continues,
 *      simulation/GPUSimulation.java:300-305; here every failure is a status) */
#define BH_OK 0
#define BH_ERR_INVALID_ARGUMENT (-1)
#define BH_ERR_OUT_OF_MEMORY (-2)
#define BH_ERR_HIP (-3)
#define BH_ERR_UNSUPPORTED (-4)
#define BH_ERR_TOO_LARGE (-5)        /* replaces gpu/GPU.java:201-209 and the int overflow of GPU.java:241,246 */
#define BH_ERR_STACK_OVERFLOW (-6)   /* a walk needed more stack than BH_WALK_STACK_DEPTH */
#define BH_ERR_TASK_OVERFLOW (-7)    /* more merge tasks than the queue holds (N pairs, gpu/GPU.java:323-330) */

/* ---- binding ids: shaders/common/common.glsl:80-120 */
#define BH_BINDING_LEAF_NODES 0
#define BH_BINDING_INTERNAL_NODES 1
#define BH_BINDING_SIMULATION_VALUES 2
#define BH_BINDING_BODIES_IN 3
#define BH_BINDING_BODIES_OUT 4
#define BH_BINDING_MORTON_IN 5
#define BH_BINDING_MORTON_OUT 6
#define BH_BINDING_INDEX_IN 7
#define BH_BINDING_INDEX_OUT 8
#define BH_BINDING_WORK_QUEUE_IN 9
#define BH_BINDING_WORK_QUEUE_OUT 10
#define BH_BINDING_RADIX_WG_HIST 11
#define BH_BINDING_RADIX_WG_SCANNED 12
#define BH_BINDING_RADIX_BUCKET_TOTALS 13
#define BH_BINDING_MERGE_TASKS 14
#define BH_BINDING_MERGE_BODY_LOCKS 15
/* no reference counterpart: the corrected walk's quadrupole records of the
 * internal nodes, 32 bytes each (float G T_xx, T_xy, T_xz, T_yy, T_yz, T_zz,
 * -1.5 G tr T, 0), written by bh_build_tree under BH_FLAG_QUADRUPOLE or
 * BH_FLAG_LEARNED and read by bh_compute_force under the same flags */
#define BH_BINDING_NODE_QUADRUPOLES 16

/* mergingCollisionOrNeither selector (common/common.glsl:136-138,
 * settings/defaultProperties.json:64-70) and staticOrDynamic (:71-77) */
#define BH_NEITHER 0u
#define BH_MERGING 1u
#define BH_COLLISION 2u
#define BH_STATIC 0u
#define BH_DYNAMIC 1u

/* bh_params.flags */
#define BH_FLAG_COLORS 1u      /* compute leaf star colours and node avgColor (render interop) */
#define BH_FLAG_TIMING 2u      /* record per-stage HIP-event timings */
#define BH_FLAG_LEARNED 4u     /* corrected walk with the attached network: every internal node taken adds
                                  its softened monopole and the softened quadrupole of T = M2 + M h^2
                                  net(node), M2 its exact second moment about its centre of mass, h half
                                  its longest side, net the network's residual (no reference counterpart;
                                  see "corrected walk" below) */
#define BH_FLAG_GENERAL_WALK 8u /* run the walk's general loop (LDS stack, overflow check, IEEE inverse
                                   square root outside its fast domain) even where the shorter loop
                                   applies; same results (for tests of both loops) */
#define BH_FLAG_QUADRUPOLE 16u /* corrected walk with the exact quadrupole (T = M2, no network) (no
                                  reference counterpart) */
#define BH_FLAG_RELATIVE_MAC 32u /* with BH_FLAG_QUADRUPOLE or BH_FLAG_LEARNED: the corrected walk's
                                    relative opening test, tolerance bh_params.force_tolerance (no
                                    reference counterpart) */
#define BH_FLAG_GRAPH 64u      /* bh_step replays the step as a hipGraph, captured on first use for each
                                  state of the buffer roles (they cycle every few steps) and each
                                  parameter set (at most 8 kept). Same results as the direct step. A
                                  step after anything but bh_step (an upload, a stage call, bh_init)
                                  or with BH_FLAG_TIMING runs directly. Attaching another network
                                  drops the graphs. */
#define BH_FLAG_FILL_CAP 128u  /* fill the tree as the reference does, in 64 passes of one height each
                                  (gpu/GPU.java:41, tree/bh_tree.comp:183-259): a node higher than 64
                                  keeps the build's state (no mass, the empty box, nodeDepth the start
                                  of its range, readyChildren its filled children), so in a taller tree
                                  the root is empty and every body feels no force. Without it every
                                  level is filled. Plain method only. */

/* The corrected walk (BH_FLAG_QUADRUPOLE / BH_FLAG_LEARNED; no reference
 * counterpart). bh_build_tree under either flag also makes one 64-byte record
 * per node (the quadrupole terms from second moments the tree builds with
 * its fill, plus the network's residual under BH_FLAG_LEARNED);
 * bh_compute_force under the same
 * flag then walks with them, without merging or collisions, softening >=
 * 1e-14 and 16 h_root^2 + softening <= 1e15 (h_root: half the root box's
 * longest side; beyond it the walk adds no forces and sets error bit 2).
 * Every node taken adds G M r / s^3 + (7.5 (r.GT.r) / s^2 - 1.5 G tr
 * T) r / s^5 - 3 GT r / s^5, s^2 = |r|^2 + softening (a leaf: its monopole),
 * in FP32 with the hardware inverse square root. Opening tests, with b the
 * distance from a node's centre of mass to the farthest corner of its box:
 *   geometric: b < theta s;
 *   relative (BH_FLAG_RELATIVE_MAC): G M b^3 / s^5 < force_tolerance |a|,
 *     |a| the body's acceleration in the previous corrected walk on this
 *     context, and b < s. A call with no previous corrected walk since
 *     bh_init or a body upload first walks with the geometric test at theta.
 * The 64 bodies of consecutive sorted positions [64k, 64k + 64) decide
 * together: they take a node when every one of them accepts it, and open it
 * for all of them otherwise (each body then adds more, smaller terms than
 * its own test asks for). */

#define BH_WALK_STACK_DEPTH 192

/* The uniforms of the step (gpu/GPU.java:402-498) as one explicit struct. */
typedef struct {
    float theta;            /* opening angle of the half-longest-side MAC (bh_force.comp:6-20); the
                               corrected walk's geometric test (see above) */
    float dt;
    float softening;        /* added to |r|^2, i.e. epsilon^2 (bh_force.comp:27-32) */
    float elasticity;
    float restitution;      /* never uploaded to the reference's FORCE program: 0 there */
    uint32_t merging_collision_or_neither;
    uint32_t static_or_dynamic;
    uint32_t wrap_around;
    /* unit uniforms of COMPUTE_UPDATE (bh_update.comp:10-37); unit_mass == 0
     * keeps the current units, as in the reference */
    float unit_mass, unit_density, unit_len, unit_time;
    /* optional exact values of the derived units; 0 derives them from the
     * unit uniforms with the formulas of common/common.glsl:346-352 */
    float gravitational_constant;
    float body_length_constant;
    uint32_t flags;
    /* no reference counterpart: alpha of the corrected walk's relative opening
     * test (BH_FLAG_RELATIVE_MAC), > 0; unused otherwise */
    float force_tolerance;
} bh_params;

/* Integer counters the reference keeps in SimulationValues, plus walk totals. */
typedef struct {
    uint32_t num_bodies;          /* sim.numBodies */
    uint32_t initial_num_bodies;  /* sim.initialNumBodies */
    uint32_t just_died;           /* bodies removed by the last dead partition */
    uint32_t out_of_bounds;       /* sim.outOfBounds (accumulates) */
    uint64_t opened;              /* internal nodes opened by the last walk (sim.uintDebug[0], unwrapped) */
    uint64_t interactions;        /* accepted internal nodes + non-self leaves in the last walk */
    uint32_t merge_tasks;         /* tasks emitted by the last walk */
    uint32_t merges_applied;      /* merges applied by the last merge resolution */
    uint32_t tree_height;         /* height of the last tree (the reference fills at most 64 levels) */
    uint32_t max_stack;           /* deepest walk stack of the last walk */
    uint32_t error_flags;         /* bit0 stack overflow, bit1 task overflow, bit2 the corrected walk's box
                                     beyond its range (16 h_root^2 + softening > 1e15: no forces), bit3 a
                                     sort tile waited past its bound for an earlier tile's counts (the
                                     order is not valid), bit4 a tree build level left more subtrees
                                     than its slots hold (the tree is not valid) */
    uint32_t reserved;
    uint64_t corrected;           /* network evaluations (one per internal node) of the last tree build */
    uint64_t quad_terms;          /* quadrupole terms added by the last corrected walk (accepted internal nodes) */
} bh_counters;

typedef struct bh_context bh_context;

/* Context create/destroy: device memory for n_initial bodies, a HIP stream,
 * sim.numBodies = sim.initialNumBodies = n_initial, sim.bounds = bounds
 * (min xyz, max xyz), sim.units from the unit uniforms. n_initial must be
 * at least 2 (the internal-node buffers hold n_initial - 1 nodes; a run whose
 * bodies merge down to one takes the reference's one-body tree) and at most
 * 2^30 (BH_ERR_TOO_LARGE).
 * Replaces GPU.initGPU (gpu/GPU.java:172-188) + BarnesHut.init (BarnesHut.java:170-173). */
int bh_create(int device, uint32_t n_initial, const float *bounds6, const bh_params *units,
              bh_context **out);
void bh_destroy(bh_context *ctx);
const char *bh_last_error(const bh_context *ctx);

/* Buffers keyed by binding id (SSBO.bind, gpu/SSBO.java:204-218; upload
 * gpu/GPU.java:1166-1196; readback gpu/SSBO.java:274-284). Node bindings are
 * converted to and from the 80-byte std430 Node on the device. Bindings 9-13
 * and 15 are the reference's scratch, with no counterpart here:
 * BH_ERR_UNSUPPORTED. */
size_t bh_binding_size(const bh_context *ctx, int binding);
int bh_upload(bh_context *ctx, int binding, const void *host, size_t offset, size_t bytes);
int bh_download(bh_context *ctx, int binding, void *host, size_t offset, size_t bytes);

/* Buffer-name swaps: GPU.swapBodyBuffers / swapMortonAndIndexBuffers (gpu/GPU.java:1060-1086) */
int bh_swap_body_buffers(bh_context *ctx);
int bh_swap_morton_and_index_buffers(bh_context *ctx);

/* BarnesHut.init(): COMPUTE_INIT, indexIn[i] = i (bh_init.comp:10-28) */
int bh_init(bh_context *ctx);

/* One entry per stage of BarnesHut.step (simulation/BarnesHut.java:99-164) */
int bh_reset_values(bh_context *ctx, const bh_params *p);          /* COMPUTE_UPDATE flag=1, :193-212 */
int bh_partition_dead_bodies(bh_context *ctx);                     /* DEAD_COUNT/SCAN/SCATTER + swap + UPDATE flag=0, :117,:217-288 */
int bh_update_bounds(bh_context *ctx);                             /* AABB_REPOPULATE + COLLAPSE, :293-329 */
int bh_generate_morton_codes(bh_context *ctx);                     /* MORTON_ENCODE, :334-353 */
int bh_radix_sort(bh_context *ctx);                                /* 16 x {HIST,BUCKET_SCAN,GLOBAL_SCAN,SCATTER}, :358-448 */
int bh_build_tree(bh_context *ctx, const bh_params *p);            /* TREE_BUILD + INIT_LEAVES + 64 x PROPAGATE, :453-535 */
int bh_compute_force(bh_context *ctx, const bh_params *p);         /* FORCE_COMPUTE, :540-558 */
int bh_merge_bodies(bh_context *ctx);                              /* MERGE_BODIES, :563-579 */

/* The whole step: all stages above in order, then the body-buffer swap. */
int bh_step(bh_context *ctx, const bh_params *p);

/* glFinish equivalent; counters and per-stage times of the last step */
int bh_sync(bh_context *ctx);
int bh_get_counters(bh_context *ctx, bh_counters *out);
#define BH_STAGE_COUNT 8
/* Per-stage times summed over the steps run with BH_FLAG_TIMING since the
 * last call (at most 512 steps), in milliseconds, order: reset, dead, bounds,
 * morton, sort, tree, force, merge; ms[8], if n > 8, receives the number of
 * steps summed. Waits for the stream. */
int bh_get_stage_times(bh_context *ctx, float *ms, int n);

/* Render feed (no reference counterpart): what the reference's renderer
 * reads each frame, compacted on the device and copied in one piece, instead
 * of the 80-byte std430 nodes of every binding. The node-glow and region
 * shaders draw InternalNodes' comMass, aabb, nodeDepth and avgColor and cull
 * empty boxes, massless nodes and depths outside their range
 * (render/nodeglow/nodeglow.vert:55-123, render/regions/regions.vert:20-42);
 * the impostor shader draws each body with its radius and star colour
 * (render/impostor/impostor.vert:29-33, common/common.glsl:228-276).
 *
 * bh_render_nodes: the internal nodes of the last tree with a non-empty box,
 * mass > 0 and depth_min <= nodeDepth <= depth_max, in node order, 48 bytes
 * each; colours are those of the last tree built with BH_FLAG_COLORS (else
 * 0). bh_render_bodies: the alive bodies of BodiesIn (mass != 0) in id
 * order, 20 bytes each: position, radius bodyLengthConstant (m / rho)^(1/3),
 * star colour. Colours are packed R | G << 8 | B << 16 | 255 << 24 with each
 * channel round(clamp(c, 0, 1) 255). *count receives the number of records
 * that pass; at most cap are written. Wait for the stream. */
typedef struct {
    float com[3], mass;
    float aabb_min[3];
    uint32_t depth;   /* nodeDepth: the node's height */
    float aabb_max[3];
    uint32_t rgba;
} bh_render_node;
typedef struct {
    float pos[3], radius;
    uint32_t rgba;
} bh_render_body;
int bh_render_nodes(bh_context *ctx, uint32_t depth_min, uint32_t depth_max, bh_render_node *out, uint32_t cap,
                    uint32_t *count);
int bh_render_bodies(bh_context *ctx, bh_render_body *out, uint32_t cap, uint32_t *count);
/* Page-locked host memory (hipHostMalloc; no reference counterpart): copies
 * into it, a per-frame render feed or a binding download, run at the link's
 * rate instead of through the runtime's staging of pageable memory. */
int bh_host_alloc(size_t bytes, void **out);
void bh_host_free(void *p);

/* FP64 direct summation over the bodies of BodiesIn for n_sample sampled
 * body ids: accel_out[3k..3k+2] = sum_{j != i, m_j != 0} G m_j r_ij
 * (|r_ij|^2 + softening)^-3/2 with i = sample[k], in double precision, G and
 * softening taken from sim and from `softening`. This exact force is what the
 * walk's force error is measured against (no reference counterpart; the reference publishes no
 * error figure). Host arrays in and out; waits for the stream. */
int bh_direct_sum_f64(bh_context *ctx, const uint32_t *sample, uint32_t n_sample, double softening,
                      double *accel_out);

/* bh_direct_sum_f64 without any device allocation. No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, serving or monitoring loops that forbid hipMalloc).
 * Same kernel, same arguments, same results bit for bit; the sampled ids and
 * the FP64 sums are staged in one static device array of the library
 * (BH_NOALLOC_MAX_SAMPLE samples, 28 bytes each, held by every GPU the library
 * runs on until the process ends) instead of per-call hipMalloc buffers.
 * n_sample > BH_NOALLOC_MAX_SAMPLE returns BH_ERR_TOO_LARGE. Every call on one
 * GPU shares that GPU's array: calls that overlap on one GPU (other threads,
 * other contexts of the same device) race and must be serialised by the
 * caller. Waits for the stream. */
#define BH_NOALLOC_MAX_SAMPLE 65536u
int bh_direct_sum_f64_noalloc(bh_context *ctx, const uint32_t *sample, uint32_t n_sample, double softening,
                              double *accel_out);

/* ---- learned far-field correction (no reference counterpart) ----
 * A 12-64-64-6 SiLU network predicts, once per internal node and step, a
 * residual o (6 values: xx, xy, xz, yy, yz, zz) to the node's quadrupole
 * tensor: T = M2 + M h^2 o. Its 12 inputs describe the node: AABB extents /
 * (2h) (3), (centre of mass - box centre) / h (3), M2 / (M h^2) (6). The
 * target of a training row is (u, y, w): u = r / s (3) of a (target body,
 * node) pair with r = com - x, s^2 = |r|^2 + softening, y (3) the residual of
 * the node's exact field after its monopole and exact quadrupole in units of
 * (h / s)^2 G M / s^2, and w the row's weight; the network's prediction for
 * the row is B(u) o = -3 o u - 1.5 tr(o) u + 7.5 (u.o.u) u, and the loss
 * w |B(u) o - y|^2 averaged over rows.
 * Parameters: FP32 master copy (5,382 values, order W1[64][12], b1,
 * W2[64][64], b2, W3[6][64], b3); training uses FP32 GEMM kernels and Adam;
 * the tree stage evaluates it with BF16 MFMA. */
typedef struct bh_mlp bh_mlp;
int bh_mlp_create(int device, bh_mlp **out);
void bh_mlp_destroy(bh_mlp *m);
const char *bh_mlp_last_error(const bh_mlp *m);
uint32_t bh_mlp_num_params(void);
int bh_mlp_set_params(bh_mlp *m, const float *params, uint32_t n);
int bh_mlp_get_params(bh_mlp *m, float *params, uint32_t n);
/* y[rows][6] = network(x[rows][12]); use_mfma = 0: FP32 kernels, 1: BF16 MFMA (the tree stage's routine) */
int bh_mlp_forward(bh_mlp *m, const float *x, uint32_t rows, float *y, int use_mfma);
/* the loss above for targets t[rows][7] = (u, y, w), and its gradient (FP32) */
int bh_mlp_loss_grad(bh_mlp *m, const float *x, const float *t, uint32_t rows, float *grad, float *loss);
/* Adam (beta 0.9/0.999) over shuffled minibatches; loss of the last minibatch */
int bh_mlp_train(bh_mlp *m, const float *x, const float *t, uint32_t rows, uint32_t epochs, uint32_t batch,
                 float lr, uint64_t seed, float *final_loss);
/* On the context's current tree (after bh_build_tree): every stride-th sorted
 * body walks on its own with the corrected walk's geometric test at theta;
 * every internal node it accepts makes a pair with FP32 node features [12] and an FP64-labelled
 * target [7] = (u, y, unit) as above, unit = (h / s)^2 G M / s^2 (a caller
 * turns it into the row's weight; nodes holding more than max_node_bodies bodies,
 * and nodes holding the target body itself, get NaN targets). n_out = pairs found; when it exceeds cap only the first
 * cap (in no particular order) are written, and a caller that wants an
 * unbiased sample raises stride and calls again. BH_ERR_STACK_OVERFLOW if a
 * walk needs more than BH_WALK_STACK_DEPTH stack entries. */
int bh_collect_training_pairs(bh_context *ctx, float theta, uint32_t stride, float softening,
                              uint32_t max_node_bodies, uint32_t cap, float *features, float *targets,
                              uint32_t *pairs_out /* optional: (sorted body, internal node) */, uint32_t *n_out);
/* Attach (or with m = NULL detach) the network used under BH_FLAG_LEARNED;
 * it must outlive the steps that use it. */
int bh_attach_correction(bh_context *ctx, const bh_mlp *m);

/* Self-test of the walk's arithmetic: counts, over every float of the walk's
 * fast domain, the results of its correctly rounded sqrt [0], reciprocal [1]
 * and inverse square root [2] that differ from the general IEEE sequences.
 * All three must be 0. (No reference counterpart.) */
int bh_selftest_math(bh_context *ctx, uint32_t *mismatches3);

#ifdef __cplusplus
}
#endif
#endif
