/*
 * inscy.h — C ABI of the MI355X INSCY engine (libinscy_hip.so).
 *
 * The reference exposes INSCY through a pybind/torch extension whose ops take
 * (X, neighborhood_size, F, num_obj, min_size, r, number_of_cells,
 * rectangular, entropy_order) and return [[subspaces], [clusterings]]
 * (reference: inscy_map.cpp:230-270 run_cpu3_weak, :979-1067 run_gpu_4,
 * :1070-1158 run_gpu_5, :1161-1249 run_gpu_memory, :1344-1432 run_gpu_star,
 * module table :1844-1872; Python wrappers python/inscy.py:121-125,189-214).
 * This header is the boundary those ops bind to instead: plain pointers and
 * sizes, no torch types, errors reported by return code.
 *
 * Results are exact: the subspace list and, per subspace, the partition of the
 * points into clusters equal those of the reference's CPU INSCY
 * (INSCYCPU3Weak, src/algorithms/inscy/InscyNodeCpu.cu:159-235).  Labels are
 * canonical: -1 for noise, clusters numbered 0.. in order of their smallest
 * point id (the reference's raw labels differ only by renaming, and its -2
 * border marks are noise).
 *
 * Threading: calls on different inscy_result objects may run concurrently on
 * different devices; one device is driven by one call at a time.
 */
#ifndef INSCY_H
#define INSCY_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Return codes. */
#define INSCY_OK 0
#define INSCY_ERR_ARGUMENT 1    /* invalid argument (sizes, eps <= 0, cells < 1, ...) */
#define INSCY_ERR_DEVICE 2      /* a HIP runtime call failed */
#define INSCY_ERR_MEMORY 3      /* device or host allocation failed */
#define INSCY_ERR_UNSUPPORTED 4 /* a parameter combination this build does not implement */

/* Engine modes. */
#define INSCY_MODE_EXACT 1   /* one restricted tree at a time (reference GPU / GPU4 role) */
#define INSCY_MODE_BATCHED 2 /* level-batched exact mode (reference GPU5 role); same results */
#define INSCY_MODE_LEARNED 3 /* level-batched with the learned pruner (inscy_set_pruner): approximate */
#define INSCY_MODE_COLLECT 4 /* level-batched exact, and records the pruner's training pairs (inscy_result_pairs) */

/* Parameters of one run; field meaning as in the reference call
 * INSCY.CPU(X, neighborhood_size, F, num_obj, min_size, r, number_of_cells,
 * rectangular, entropy_order) (python/inscy.py:121-125). */
typedef struct inscy_params {
    float neighborhood_size; /* eps */
    float F;                 /* density factor */
    int32_t num_obj;         /* mu: minimum neighbours */
    int32_t min_size;        /* minimum cluster / region size (absolute count) */
    float r;                 /* redundancy factor */
    int32_t number_of_cells; /* c: grid cells per dimension */
    int32_t rectangular;     /* 1: neighbour-count density, 0: Epanechnikov density */
    int32_t entropy_order;   /* 0: identity level order (the only order this build runs) */
} inscy_params;

/* Counters of one run (the reference prints `calls` only, InscyNodeCpu.cu:166). */
typedef struct inscy_stats {
    int64_t calls;          /* INSCY recursion invocations */
    int64_t trees;          /* restricted (subspace, region) trees evaluated by the prune */
    int64_t queries;        /* epsilon-neighbourhood counts computed on the device */
    int64_t candidates;     /* candidate pairs examined by the neighbourhood kernels */
    int64_t subspaces;      /* subspaces in the result */
    int64_t clusters;       /* clusters in the result */
    int64_t components;     /* subspaces whose dense components were built */
    double seconds;         /* wall time of the run inside the library */
    /* filled when profiling is on (inscy_profile): */
    double count_kernel_ms; /* device time of the neighbour-count kernel (HIP events on the run's stream) */
    int64_t count_launches; /* launches of that kernel */
    int64_t count_bytes;    /* its algorithmic bytes: per counted point 4k+4, per neighbour found 4+4k */
    /* wall time of the phases of the run (seconds): growing the restricted
     * trees (restrict, prunes, counts, dense components), the replay of
     * pruneRedundancy and join (level-synchronous mode; inside the first
     * phase in exact mode), and reading the result back */
    double expand_seconds;
    double replay_seconds;
    double emit_seconds;
    /* learned and collect modes: (tree, dimension) pairs scored by the
     * pruner's features, and pairs skipped (learned mode) */
    int64_t pairs_scored;
    int64_t pairs_skipped;
    /* level-synchronous driver: grid indexes whose sort was taken from an
     * earlier batch of the run (same index dimensions) instead of sorted */
    int64_t sorts_kept;
    /* level-synchronous driver: bytes of the candidates the neighbour-count
     * kernel examined (packed cells and coordinates), most of them cache hits */
    int64_t count_examined_bytes;
} inscy_stats;

typedef struct inscy_result inscy_result;

/* Runs INSCY on a host array X (row-major float32 [n, d], values in [0, 1]).
 * Replaces run_cpu3_weak / run_gpu_4 / run_gpu_5 (inscy_map.cpp:230,979,1070):
 * X is borrowed, the result is owned by the caller (inscy_result_free). */
int inscy_run(const float *X, int64_t n, int32_t d, const inscy_params *params, int32_t mode,
              inscy_result **out);

/* Same with X already in device memory (row-major float32 [n, d]) on the
 * current HIP device; `stream` is a hipStream_t or NULL for the library's own
 * stream.  This is the form whose time excludes the host-to-device copy. */
int inscy_run_device(const float *d_X, int64_t n, int32_t d, const inscy_params *params, int32_t mode,
                     void *stream, inscy_result **out);

/* Runs INSCY on `count` independent data sets (config 3 of the measured
 * workloads): X[i] is a row-major float32 [n[i], d[i]] host array, params[i]
 * its parameters, out[i] receives its result.  Worker threads with one stream
 * each (4, or the INSCY_WORKERS environment variable) take the sets in order
 * of estimated cost n^2 d, largest first.  On an error no result is returned
 * and the message names the data set.  The reference has no counterpart: its
 * harness calls one op per data set in a loop (experiments/run.py:122-175). */
int inscy_run_many(const float *const *X, const int64_t *n, const int32_t *d, const inscy_params *params,
                   int32_t count, int32_t mode, inscy_result **out);

/* Same with every X[i] already in device memory on the current device. */
int inscy_run_many_device(const float *const *d_X, const int64_t *n, const int32_t *d, const inscy_params *params,
                          int32_t count, int32_t mode, inscy_result **out);

/* Number of subspaces in the result, in the reference's output order
 * (ascending dimension bitmask, util.cu:633-653 vec_cmp). */
int32_t inscy_result_num_subspaces(const inscy_result *res);

/* Dimensions of subspace i (ascending) into dims_out (capacity d); returns their count. */
int32_t inscy_result_subspace(const inscy_result *res, int32_t i, int32_t *dims_out);

/* Canonical labels of subspace i: n int32 owned by the result. */
const int32_t *inscy_result_labels(const inscy_result *res, int32_t i);

/* Labels of every subspace as one contiguous row-major [num_subspaces][n]
 * array owned by the result (row i is inscy_result_labels(res, i)).  No
 * counterpart in the reference, which returns one tensor per subspace; it lets
 * a caller take the whole result without copying. */
const int32_t *inscy_result_labels_all(const inscy_result *res);

int64_t inscy_result_num_points(const inscy_result *res);

int inscy_result_stats(const inscy_result *res, inscy_stats *out);

void inscy_result_free(inscy_result *res);

/* Profiling switch for later runs of this process: brackets every launch of
 * the neighbour-count kernel with HIP events (no counterpart in the
 * reference; used by the benchmark's roofline figures). */
int inscy_profile(int on);

/* Tuning of later calls of this process (no counterpart in the reference;
 * gpu_inscy_amd/autotune.py measures and applies it): max_batch child
 * subspaces per level-synchronous batch (default 1024, still bounded by
 * memory), workers streams of inscy_run_many (default 8); 0 restores a
 * default.  INSCY_ERR_ARGUMENT outside [0, 65536] and [0, 64].  Results do
 * not depend on either. */
int32_t inscy_set_tuning(int32_t max_batch, int32_t workers);

/* Device and pinned host blocks of finished runs are cached for reuse by
 * later runs (freeing them would synchronise the whole device); this returns
 * them to the runtime.  Call it only while no run is in progress.  No
 * counterpart in the reference, which frees everything inside each call. */
int inscy_release_memory(void);

/* ---- learned pruner (no counterpart in the reference) ----
 *
 * In INSCY_MODE_LEARNED the level-batched driver scores every (restricted
 * tree T, dimension m) pair before the neighbour counts of T's points in
 * S u {m}: a small network (64 -> 128 -> 128 -> 1, ReLU, bf16 weights, fp32
 * accumulation) maps the pair's features to the probability that T's subtree
 * along m reaches a cluster of the exact result.  Pairs scored below the
 * threshold are skipped with everything below them; all other work, and all
 * integer work, is exact, so a threshold of 0 gives the exact result.
 *
 * Features of a pair (64 float32): T's size and depth, m, the value and cell
 * histograms of T's points in dimension m, T's points' counts in S, and the
 * dimension bitmask of S u {m} (bit k at feature 32 + k % 32). */
#define INSCY_MLP_INPUTS 64
#define INSCY_MLP_HIDDEN 128
#define INSCY_MLP_PARAMS (128 * 64 + 128 + 128 * 128 + 128 + 128 + 1)

/* Network of later INSCY_MODE_LEARNED runs of this process: `weights` holds
 * INSCY_MLP_PARAMS float32 values, W1[128][64] b1[128] W2[128][128] b2[128]
 * w3[128] b3 (rounded to bf16 here).  Pairs scoring below `threshold` are
 * skipped, except those of trees with fewer than `min_depth` restricted
 * dimensions.  weights == NULL unloads the network. */
int inscy_set_pruner(const float *weights, float threshold, int32_t min_depth);

/* Training pairs recorded by an INSCY_MODE_COLLECT run: returns their count
 * and sets (each optional) features [count][INSCY_MLP_INPUTS] float32,
 * labels [count] (1: the exact result holds a cluster found in the pair's
 * subtree), info [count][4] int32 (restricted dimensions of T, m, points of
 * T, T's index), cost [count][2] int64 (points counted for the pair itself,
 * and for the pair and every pair below it).  Arrays are owned by the result. */
int64_t inscy_result_pairs(const inscy_result *res, const float **features, const uint8_t **labels,
                           const int32_t **info, const int64_t **cost);

/* The network's kernels on device arrays, for training (stream: hipStream_t or
 * NULL).  params: INSCY_MLP_PARAMS bf16 values (uint16 bit patterns).
 * forward: prob[i] = sigmoid(network(features[i])).
 * gradient: grad[0..INSCY_MLP_PARAMS) = gradient, grad[INSCY_MLP_PARAMS] =
 *   value, of inv_weight_sum * sum_i weights[i] * BCE(logit_i, labels[i]);
 *   partials: nblocks * (INSCY_MLP_PARAMS + 1) float32 scratch.
 * adam: one Adam step (beta 0.9, 0.999, eps 1e-8, bias-corrected at `step`
 *   >= 1) on the float32 master weights, then params = bf16(master).
 * round: params = bf16(master). */
int inscy_mlp_forward(const uint16_t *d_params, const float *d_features, int64_t count, float *d_prob, void *stream);
int inscy_mlp_gradient(const uint16_t *d_params, const float *d_features, const float *d_labels,
                       const float *d_weights, int64_t count, float inv_weight_sum, float *d_partials,
                       int32_t nblocks, float *d_grad, void *stream);
int inscy_mlp_adam(float *d_master, const float *d_grad, float *d_m, float *d_v, uint16_t *d_params, int32_t step,
                   float lr, void *stream);
int inscy_mlp_round(const float *d_master, uint16_t *d_params, void *stream);

/* Message of the last failing call on this thread ("" if none). */
const char *inscy_last_error(void);

/* ---- kernel-level entry points (device pointers; used by parity tests) ---- */

/* Discretisation of X into the c-grid and the ceil(1/eps)-grid plus the
 * s-connection marks (reference: ScyTreeNode.cu:100-102 get_cell_no and
 * :34-49 construct_s_connection).  Outputs, all dimension-major [d][n]:
 * cell_c (uint8), cell_e (uint16), near_mask[n][(d + 63) / 64] (bit k % 64 of
 * word k / 64 set when point p lies within eps of the upper edge of its c-grid
 * cell in dimension k; one word per point for d <= 64). */
int inscy_discretise(const float *d_X, int64_t n, int32_t d, int32_t cells, float eps, uint8_t *d_cell_c,
                     uint16_t *d_cell_e, uint64_t *d_near_mask);

/* epsilon-neighbour counts |N_S(p)| for query points (reference:
 * ClusteringCpu.cu:52-81 neighborhood with the candidate rule of
 * ScyTreeNode.cu:896-946): q != p, |cell_e difference| <= 1 in every dimension
 * of S, and the float distance in S (S order, no fused multiply-add) <= eps.
 * dims: k ascending dimension ids; queries: nq point ids; out: nq counts. */
int inscy_neighbour_counts(const float *d_X, int64_t n, int32_t d, float eps, const int32_t *dims, int32_t k,
                           const int32_t *d_queries, int64_t nq, int32_t *d_out);

/* inscy_neighbour_counts without any device allocation.  No counterpart in
 * the reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The same counts, for k ascending dimension ids; out[i] = -1 for
 * a query id outside [0, n).  Every buffer of the pass lives in one 64 MiB
 * device array per device, held by the library until the process ends; inputs
 * whose pass needs more (about (12 k + 16) n bytes plus 8 bytes per grid cell)
 * are refused with INSCY_ERR_ARGUMENT before any device call.  The work is
 * queued on `stream` (a hipStream_t, NULL for the null stream) and the
 * function returns without waiting: d_out holds the counts once the stream's
 * work is done.  X values outside [0, 1] are clamped, not rejected.  Calls on
 * one device share the array: calls that overlap on one device (different
 * streams or threads) race and must be serialised by the caller. */
int inscy_neighbour_counts_noalloc(const float *d_X, int64_t n, int32_t d, float eps, const int32_t *dims, int32_t k,
                                   const int32_t *d_queries, int64_t nq, int32_t *d_out, void *stream);

/* Dense components of subspace S over all points (reference:
 * ClusteringCpu.cu:317-389 BFS expansion): comp[p] = smallest point id of p's
 * component in the dense-dense epsilon graph, or -1 when p is not dense with
 * the |S|-dimensional threshold. */
int inscy_dense_components(const float *d_X, int64_t n, int32_t d, const inscy_params *params,
                           const int32_t *dims, int32_t k, int32_t *d_comp);

#ifdef __cplusplus
}
#endif

#endif /* INSCY_H */
