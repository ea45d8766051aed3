/* bh_ensemble.h — ensembles of independent simulations stepped together on
 * one GPU.
 *
 * No counterpart in the reference, which runs one simulation per process
 * (BarnesHut.java:99-164 is the step this batches). The callers are sweeps
 * that run many small simulations (parameter studies, initial-condition
 * ensembles), for which one simulation per launch leaves the GPU idle.
 *
 * The bodies of all simulations live in one set of device arrays,
 * simulation s owning body ids [off_s, off_s + n_s). One step runs every
 * stage over all simulations with one launch per stage: per-simulation
 * bounds, Morton codes against each simulation's own bounds, a segmented
 * radix sort keyed on (simulation, Morton code), one bottom-up tree build over
 * all segments, and a persistent walk whose waves pull 64-body tasks from
 * global work queues (one per XCD, each holding an equal-cost contiguous share
 * of the tasks; a wave whose queue is empty takes from the others). Each
 * simulation's bounds, codes,
 * permutation, tree and integrated bodies are bit for bit those of the same
 * simulation stepped alone through include/bh_abi.h.
 *
 * The whole step can be captured once as a hipGraph and replayed. Digit width
 * of the sort, walk workgroup size and walk stack depth are chosen by
 * bh_ensemble_autotune from measurements on the ensemble itself.
 *
 * Scope: dynamic bounds, merging off, every level of each tree filled (no
 * BH_FLAG_FILL_CAP) and every body alive (non-zero mass); each simulation
 * has at least 2 bodies. The walk is the plain method, or
 * the corrected walk of include/bh_abi.h (params->flags BH_FLAG_QUADRUPOLE
 * or BH_FLAG_LEARNED, optionally BH_FLAG_RELATIVE_MAC with
 * params->force_tolerance): node records over all simulations' trees,
 * evaluated by the attached network once per node under BH_FLAG_LEARNED,
 * each simulation walked from its own root with the same opening tests and
 * wave-wide decisions (a wave's 64 bodies belong to one simulation). With the
 * relative test the first step runs directly with the geometric test at
 * theta; the graph of the step is captured from the second on.
 */
#ifndef BH_ENSEMBLE_H
#define BH_ENSEMBLE_H

#include <stdint.h>

#include "bh_abi.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct bh_ensemble bh_ensemble;

typedef struct {
    uint32_t radix_bits;   /* sort digit width: 5, 6, 7 or 8 (ceil(63 / bits) passes) */
    uint32_t walk_waves;   /* walk workgroup: 1, 2, 4 or 8 waves of 64 */
    uint32_t stack_depth;  /* walk stack entries per wave: 64, 128 or 192 */
    uint32_t use_graph;    /* 1: bh_ensemble_step replays a captured hipGraph of the step */
} bh_ensemble_config;

/* n_sims simulations of counts[s] bodies; bodies = the simulations' Body
 * records (8 floats each, include/bh_abi.h) one after another; bounds =
 * n_sims x {min xyz, max xyz} initial bounds; params: theta, dt, softening
 * and the units (or gravitational_constant) of every simulation;
 * cfg = NULL selects radix 8, 4 waves, depth 192, graph on. */
int bh_ensemble_create(int device, uint32_t n_sims, const uint32_t *counts, const float *bodies,
                       const float *bounds, const bh_params *params, const bh_ensemble_config *cfg,
                       bh_ensemble **out);
void bh_ensemble_destroy(bh_ensemble *e);
const char *bh_ensemble_last_error(const bh_ensemble *e);

int bh_ensemble_get_config(const bh_ensemble *e, bh_ensemble_config *cfg);
int bh_ensemble_set_config(bh_ensemble *e, const bh_ensemble_config *cfg);

/* enqueue `steps` steps (asynchronous; bh_ensemble_sync waits) */
int bh_ensemble_step(bh_ensemble *e, uint32_t steps);
int bh_ensemble_sync(bh_ensemble *e);

/* current bodies (all simulations, [total][8] floats) */
int bh_ensemble_download_bodies(bh_ensemble *e, float *bodies);
/* current bounds, [n_sims][6] */
int bh_ensemble_download_bounds(bh_ensemble *e, float *bounds);
/* Morton codes and body ids in sorted order as the last step left them */
int bh_ensemble_download_sorted(bh_ensemble *e, uint64_t *keys, uint32_t *index);
/* internal nodes of the last step's trees in the reference's 80-byte Node
 * layout: simulation s's nodes at [off_s - s, off_s - s + n_s - 1); node ids
 * are global (leaves < total, internal total + storage index) */
int bh_ensemble_download_internal_nodes(bh_ensemble *e, void *nodes);
/* totals over all simulations for the last step (tree_height: the highest) */
int bh_ensemble_get_counters(bh_ensemble *e, bh_counters *out);

/* the network of an ensemble created with BH_FLAG_LEARNED (no reference
 * counterpart; include/bh_abi.h's bh_mlp); it must outlive the steps that
 * use it. Drops the captured graphs. */
int bh_ensemble_attach_correction(bh_ensemble *e, const bh_mlp *m);

/* Time the candidate digit widths (sort) and walk shapes (workgroup x stack
 * depth, depths below the current tree height + 2 excluded) on the current
 * state, without advancing it; keep the fastest and return it. timings
 * (optional, 16 floats, ms; -1 = not run): [0..3] radix 5..8, then the 12
 * walk shapes waves {1,2,4,8} x depth {64,128,192}. */
int bh_ensemble_autotune(bh_ensemble *e, bh_ensemble_config *chosen, float *timings);

#ifdef __cplusplus
}
#endif

#endif
