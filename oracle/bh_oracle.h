/*
 * bh_oracle — CPU restatement of the reference's Barnes-Hut step.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker for the HIP path: only
 * the tests, the smoke check and bench.py's CPU-baseline leg may load it. The
 * product path (gravity_amd) never links, loads or calls it.
 *
 * Each function restates one program of the reference's GLSL pipeline
 * (src/main/resources/shaders/compute/) with the same FP32 operation order;
 * division is IEEE division and no multiply-add is contracted, which is how
 * Mesa llvmpipe executes the reference shaders (established bit-exactly
 * against the llvmpipe fixtures under tests/golden/). Pinning of every
 * function is listed in bh_oracle.c's header.
 */
#ifndef BH_ORACLE_H
#define BH_ORACLE_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* std430 Node, 80 bytes (common/common.glsl:40-58, gpu/Node.java:31-41) */
typedef struct {
    float comMass[4];
    float aabb[6];
    uint32_t childA, childB, nodeDepth, bodiesContained, readyChildren, parentId;
    float avgColor[4];
} bho_node;

/* std430 SimulationValues as GLSL lays it out (common/common.glsl:86-88):
 * the UnitSet struct is 28 bytes, so uintDebug starts at byte 92. */
typedef struct {
    uint32_t numBodies, initialNumBodies, justDied, merged, outOfBounds, relativeTo, pad1, pad2;
    float minCorner[3]; float padA;
    float maxCorner[3]; float padB;
    float unitMass, unitDensity, unitLen, unitTime, cameraScale, gravitationalConstant, bodyLengthConstant;
    uint32_t uintDebug[100];
    float floatDebug[100];
    uint32_t tail_pad;
} bho_sim;

typedef struct {
    float theta, dt, softening, elasticity, restitution;
    uint32_t merging;        /* mergingCollisionOrNeither: 0 neither, 1 merge, 2 collision */
    uint32_t dynamic;        /* staticOrDynamic: 0 static, 1 dynamic */
    uint32_t wrap_around;
} bho_params;

/* outputs of the walk that the reference keeps only implicitly */
typedef struct {
    uint64_t opened;         /* uintDebug[0] increments: internal nodes opened */
    uint64_t interactions;   /* accepted internal nodes + non-self leaves */
    uint64_t mac_tests;      /* internal nodes tested */
    uint32_t out_of_bounds;  /* bodies killed by the static-mode bounds test */
    uint32_t max_stack;      /* deepest stack occupancy seen */
    uint32_t n_tasks;        /* merge tasks emitted */
    uint32_t task_overflow;  /* tasks dropped because the queue was full */
} bho_walk_stats;

/* units: common/common.glsl:346-360 (powf of the host C library) */
void bho_units(float mass, float density, float len, float time, float *G, float *body_len_const);

/* dead partition: shaders/compute/dead/bh_dead.comp:13-112. Stable partition of
 * index_in[0..num) into alive-then-dead; dead bodies are set to EMPTY in both
 * body buffers. Returns justDied. */
uint32_t bho_dead_partition(float *src_bodies, float *dst_bodies, const uint32_t *index_in,
                            uint32_t num, uint32_t *index_out);

/* AABB repopulate + collapse: shaders/compute/morton/bh_morton.comp:50-231.
 * bounds_inout = {min.xyz, max.xyz}; on entry the previous bounds. */
void bho_aabb(const float *bodies, const uint32_t *index, uint32_t num, float *bounds_inout);

/* Morton encode: bh_morton.comp:6-34,238-262 */
uint64_t bho_expand_bits21(uint32_t v);
uint64_t bho_morton3d(uint32_t x, uint32_t y, uint32_t z);
void bho_morton_encode(const float *bodies, const uint32_t *index, uint32_t num,
                       const float *bounds, uint64_t *codes);

/* radix sort result: shaders/compute/radix/bh_radix.comp:12-166 is a stable
 * LSD sort of (key, index) pairs; this is the same stable order. */
void bho_sort(uint64_t *keys, uint32_t *index, uint32_t num);

/* Karras build: shaders/compute/tree/bh_tree.comp:6-133 */
int bho_safe_lcp(const uint64_t *keys, uint32_t num, int i, int j);
void bho_build(const uint64_t *keys, uint32_t num, uint32_t n_init, bho_node *leaf, bho_node *internal);

/* leaf init + level-synchronous propagate: bh_tree.comp:140-259.
 * max_passes = 64 reproduces the reference cap (nodes taller than max_passes stay
 * unfilled); max_passes = 0 fills every node. Returns the tree height. */
uint32_t bho_fill(const float *bodies, const uint32_t *index, uint32_t num, uint32_t n_init,
                  float unit_mass, float unit_density, int max_passes, int with_colors,
                  bho_node *leaf, bho_node *internal);

/* star colour: common/common.glsl:235-276 */
void bho_star_color(float mass, float density, float rgb[3]);

/* walk + integrate: shaders/compute/force/bh_force.comp:6-142. Writes dst for
 * each walked body; accel (3 floats per sorted position) and per-body
 * interaction counts are optional outputs. tasks receives (a, b) pairs in
 * sorted-position order of the emitting body, up to task_cap pairs. */
int bho_force(const float *src, float *dst, const uint32_t *index, uint32_t num, uint32_t n_init,
              const bho_node *leaf, const bho_node *internal, const bho_params *p,
              float G, float body_len_const, const float *bounds, int stack_cap,
              float *accel, uint32_t *interactions, uint32_t *tasks, uint32_t task_cap,
              bho_walk_stats *stats, int threads);

/* the same walk for the sorted positions [g_begin, g_end) only */
int bho_force_range(const float *src, float *dst, const uint32_t *index, uint32_t num, uint32_t n_init,
                    const bho_node *leaf, const bho_node *internal, const bho_params *p,
                    float G, float body_len_const, const float *bounds, int stack_cap,
                    float *accel, uint32_t *interactions, uint32_t *tasks, uint32_t task_cap,
                    bho_walk_stats *stats, int threads, uint32_t g_begin, uint32_t g_end);
/* the corrected walk (no reference counterpart) in FP64 over the nodes the
 * GPU's FP32 tests take, in groups of `group` (1..64; the GPU: 64) sorted
 * positions that take a node only together: geometric test at
 * theta (amag NULL) or relative test at alpha with amag [num] (|a| by sorted
 * position); quad[8k .. 8k+6] = G T xx xy xz yy yz zz, -1.5 G tr T of internal
 * node k. accel [num][3], absum [num] (sum of the taken terms' magnitudes),
 * inter [num] by sorted position; the stats' mac_tests field counts the
 * quadrupole terms. Returns -1 on stack overflow. */
int bho_force_corrected(const bho_node *leaf, const bho_node *internal, uint32_t num, uint32_t n_init,
                        const float *quad, float theta, float softening, float G, const float *amag, float alpha,
                        uint32_t group, int stack_cap, double *accel, double *absum, uint32_t *inter,
                        bho_walk_stats *st, int threads);
/* FP64 second moment of each internal node's bodies about its centre of mass
 * (m2 [num-1][6]: xx xy xz yy yz zz), by direct summation */
void bho_node_moments(const bho_node *leaf, const bho_node *internal, uint32_t num, uint32_t n_init, double *m2,
                      int threads);

/* merge resolution: shaders/compute/merge/bh_merge.comp:5-71, resolved in
 * ascending (a, b) task order. Returns the number of merges applied. */
uint32_t bho_merge(float *dst, uint32_t *tasks, uint32_t n_tasks);
void bho_merge_pair(const float *b1, const float *b2, float *out);

/* the whole step of simulation/BarnesHut.java:99-164 on host buffers.
 * bodies_in/out are ping-pong buffers (swapped on return), index/morton hold
 * the persistent sorted order. */
typedef struct {
    uint32_t n_init;
    float *bodies[2];
    uint32_t *index[2];
    uint64_t *morton;
    bho_node *leaf, *internal;
    uint32_t *tasks; uint32_t task_cap;
    bho_sim sim;
    int cur;                 /* which of bodies[] is BodiesIn */
    int fill_cap;            /* 64 = reference cap, 0 = none */
    int stack_cap;           /* 64 = reference stack */
    int threads;
    bho_walk_stats last;
    uint32_t last_height;
} bho_state;

int bho_state_init(bho_state *s, uint32_t n, const float *bodies, const float *bounds,
                   float unit_mass, float unit_density, float unit_len, float unit_time);
void bho_state_free(bho_state *s);
int bho_step(bho_state *s, const bho_params *p);

/* FP64 direct sum for the sampled bodies: a_i = sum_j G m_j r_ij (|r_ij|^2 + soft)^-3/2, j != i */
void bho_direct_sum_f64(const float *bodies, uint32_t n, const uint32_t *sample, uint32_t n_sample,
                        double G, double softening, double *accel_out, int threads);

#ifdef __cplusplus
}
#endif
#endif
