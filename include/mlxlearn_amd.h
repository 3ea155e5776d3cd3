/*
 * mlxlearn_amd -- C ABI of the MI355X (gfx950) exact neighbour-search library.
 *
 * Every entry point returns an int status: 0 = ok, > 0 = invalid input (the
 * Python layer raises ValueError), < 0 = device / allocation failure
 * (BackendExecutionError / MemoryError). No C++ exception crosses the ABI and
 * no partial output is promised on error. mlk_last_error() returns the
 * calling thread's last message.
 *
 * Pointers named *_dev are device (HBM) pointers; outputs are caller-owned.
 * `stream` is a hipStream_t (NULL = the default stream). One index may be used
 * from several host threads: calls on one index serialise on its lock.
 *
 * Exactness contract (all routes, all precisions): k-NN rows are ordered by the
 * exact real squared distance S(q, x) of the float32 coordinates, lower
 * training index first on exact ties; radius membership is exact S <= R2 with
 * R2 = fl64(r*r); k-means labels are the exact nearest centre, lowest index on
 * ties. Reported distances are sqrt of the canonical float64 sum
 * sum_k (double(q_k) - double(x_k))**2 evaluated left to right.
 */
#ifndef MLXLEARN_AMD_H
#define MLXLEARN_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mlk_index mlk_index;
typedef struct mlk_radius_result mlk_radius_result;

/* status codes */
#define MLK_OK 0
#define MLK_EINVAL 1
#define MLK_ECAPACITY 2
#define MLK_ENOMEM (-1)
#define MLK_EDEVICE (-2)

/* query flags */
#define MLK_EXCLUDE_SELF 0x1u     /* query row i is training row q_offset + i; drop it   */
#define MLK_PREC_FP32 0x00u       /* screen precision (bits 4-5)                          */
#define MLK_PREC_BF16 0x10u
#define MLK_PREC_FP16 0x20u
#define MLK_PREC_MASK 0x30u
#define MLK_SORT_RESULTS 0x100u   /* radius: order rows by (distance, index)              */
#define MLK_NEED_DISTANCE 0x200u  /* radius: keep distances                               */
#define MLK_SQUARED_DIST 0x400u   /* k-NN: dist_dev receives squared distances s           */
#define MLK_GRAPH_REPLAY 0x800u   /* k-NN: capture the call's device chain in a HIP graph and
                                     replay it on later calls with the same arguments and
                                     buffers (one launch, one read-back per call; no device
                                     times in mlk_get_stats)                               */

/* weights for mlk_knn_votes */
#define MLK_WEIGHTS_UNIFORM 0
#define MLK_WEIGHTS_DISTANCE 1

/* test hooks (mlk_set_test_hooks) */
#define MLK_HOOK_FORCE_FALLBACK 0x1u   /* every query takes the exact fallback route        */
#define MLK_HOOK_CAPACITY 0x2u         /* candidate capacity per (query, split) = param     */
#define MLK_HOOK_FAIL 0x4u             /* the next library call fails with MLK_EDEVICE      */
#define MLK_HOOK_FORCE_UNCERTIFIED 0x8u /* rows listed by mlk_set_forced_rows lose certificates */
#define MLK_HOOK_SORT_LIMIT 0x10u      /* radius: counts sorted or selected in one pass must be <= param
                                          (default INT32_MAX, the limit of the library sorts)       */
#define MLK_HOOK_FORCE_ORDER 0x20u     /* regroup the queries by locality at any call size (k-NN
                                          calls without self-exclusion, radius calls)               */
#define MLK_HOOK_NO_ORDER 0x40u        /* keep the caller's query order at any call size            */

/* Per-call statistics of the last query on this thread. */
typedef struct {
    int64_t queries;
    int64_t candidates;        /* screen candidates appended over all rounds       */
    int64_t reranked;          /* candidates that reached the float64 exact rank    */
    int64_t fallback_queries;  /* queries resolved by the exact fallback scan       */
    int64_t rounds;            /* screen rounds launched                            */
    int64_t splits;            /* database splits per round                         */
    int64_t capacity;          /* candidate slots per (query, split)                */
    int64_t screen_precision;  /* MLK_PREC_* actually used                          */
    double screen_ms;          /* device time of the screen kernels (events)        */
    double total_ms;           /* device time of the whole query (events)           */
} mlk_stats;

const char* mlk_last_error(void);
int mlk_version(void);
int mlk_get_stats(mlk_stats* out);
int mlk_set_test_hooks(uint32_t hooks, int64_t param);
/* Rows (query positions within the next call) whose screen certificate is withdrawn. */
int mlk_set_forced_rows(const int64_t* rows_host, int64_t count);

/* Tuning overrides (0 = automatic): database splits, candidate slots per
 * sub-list, query tiles of 32 per screen wave (1-2), stages of the screen's
 * LDS tile ring (2-4) and the cumulative round fractions of each split.
 * Results never depend on them; a shape that is not built falls back to the
 * exact scan. */
int mlk_set_tuning(int32_t splits, int32_t capacity, int32_t query_tiles, int32_t stages,
                   int32_t n_rounds, const double* round_fractions);

/*
 * Autotuner. Replaces the tile planner resolve_blocks
 * (src/mlxlearn_device/_neighbors_plan.py:74-126) and the per-GPU-family
 * selection tables (src/mlxlearn_device/_neighbors_general.py:80-170).
 * mlk_autotune times screen plans (shape, database splits, round schedule,
 * list capacity) on the caller's queries Q_dev (m x d) with the index and
 * keeps the fastest in a process-wide cache keyed by (k-NN or radius,
 * screen precision, padded width, k, n and m in half-octave buckets); later
 * calls of that shape use it, including mlk_kmeans_assign (k = 1 over a
 * centres index of the same shape). With MLK_TUNE_RADIUS it tunes
 * mlk_radius_query at R2 (k ignored). max_trials bounds the plans timed
 * (each twice). Results never depend on the plan. mlk_get_plan reports the
 * plan a call of that shape would use; mlk_clear_plans empties the cache.
 */
#define MLK_TUNE_RADIUS 0x1000u
#define MLK_MAX_ROUNDS 8
typedef struct {
    int32_t splits;       /* database splits per screen round                       */
    int32_t capacity;     /* candidate slots per (query, split, lane half), k-NN    */
    int32_t query_tiles;  /* 32-query tiles per screen wave                         */
    int32_t stages;       /* stages of the screen's LDS tile ring                   */
    int32_t n_rounds;     /* screen rounds (k-NN; 0 for radius)                     */
    double rounds[MLK_MAX_ROUNDS]; /* cumulative fractions of each split, last = 1   */
    int32_t tuned;        /* 1 if the plan came from the autotuner                  */
    int32_t trials;       /* plans timed by the autotuner                           */
    double tuned_ms;      /* device time of one call with this plan                 */
    double default_ms;    /* device time of one call with the untuned plan          */
} mlk_plan;
int mlk_autotune(mlk_index* index, const float* Q_dev, int64_t m, int32_t k, double R2,
                 uint32_t flags, int32_t max_trials, mlk_plan* out, void* stream);
int mlk_get_plan(const mlk_index* index, int64_t m, int32_t k, uint32_t flags, mlk_plan* out);
int mlk_clear_plans(void);
/* The plan cache as text (one line per tuned plan), so a plan tuned once can
 * be reused by later processes: mlk_export_plans writes *len bytes into buf
 * (buf NULL and cap 0 only report *len); mlk_import_plans adds the plans of a
 * text made by mlk_export_plans, rejecting the whole text (MLK_EINVAL) if a
 * line is not a valid plan for a built screen shape. */
int mlk_export_plans(char* buf, int64_t cap, int64_t* len);
int mlk_import_plans(const char* text, int64_t len);

/*
 * Exact brute-force index over X (n x d float32, row-major, device memory).
 * The index owns a copy of X plus its centre, float64 row norms and the
 * screen mirrors. Replaces DeviceNeighborIndex / fit_neighbors
 * (src/mlxlearn_device/_neighbors.py:243-270, 447-466) and
 * build_index (src/mlxlearn/_kernels/distance.py:366-392).
 */
int mlk_index_create(const float* X_dev, int64_t n, int32_t d, uint32_t flags, void* stream,
                     mlk_index** out);
int mlk_index_destroy(mlk_index* index);
int mlk_index_info(const mlk_index* index, int64_t* n, int32_t* d, int32_t* integral);

/*
 * Exact k nearest neighbours of Q (m x d float32, device). Writes idx_dev
 * (m x k int64) and, when dist_dev is not NULL, dist_dev (m x k float64,
 * Euclidean). Replaces DeviceNeighborIndex.query
 * (src/mlxlearn_device/_neighbors.py:271-443) behind NeighborIndex.query /
 * query_device (src/mlxlearn/_kernels/distance.py:211-270).
 */
int mlk_knn_query(mlk_index* index, const float* Q_dev, int64_t m, int32_t k, int64_t q_offset,
                  uint32_t flags, double* dist_dev, int64_t* idx_dev, void* stream);

/*
 * mlk_knn_query without any device allocation. No counterpart in the
 * reference: it is for callers that cannot allocate (serving loops with a
 * fixed memory budget, harnesses that forbid hipMalloc). Temporaries live in
 * a fixed 128 MiB file-scope device array of the library (held on each GPU
 * the library runs on until the process ends) instead of the index's
 * workspace; queries are processed in chunks that fit it. Same arguments,
 * results and errors as mlk_knn_query, and: the index must already hold the
 * screen mirror of the requested precision (one mlk_knn_query call of that
 * precision prepares it) -> else MLK_EINVAL; an index of more than about 3.3
 * million rows (its exact fallback would not fit the array) -> MLK_ECAPACITY.
 * Calls that overlap on one GPU race on the array; no call records events, so
 * mlk_get_stats reports no device times for it.
 */
int mlk_knn_query_noalloc(mlk_index* index, const float* Q_dev, int64_t m, int32_t k, int64_t q_offset,
                          uint32_t flags, double* dist_dev, int64_t* idx_dev, void* stream);

/*
 * Several query batches (concatenated rows; batch b is rows
 * [offsets[b], offsets[b+1])) answered by one screen launch per round.
 * Results are identical to one mlk_knn_query per batch.
 */
int mlk_knn_query_batched(mlk_index* index, const float* Q_dev, const int64_t* offsets_host,
                          int32_t n_batches, int32_t k, uint32_t flags, double* dist_dev,
                          int64_t* idx_dev, void* stream);

/*
 * Exact radius neighbourhoods: S(q, x) <= R2. Two phases: mlk_radius_query
 * computes membership and fills indptr_dev (m+1 int64, CSR row pointers) and
 * *nnz; mlk_radius_fill copies indices (int64) and optional distances
 * (float64, sqrt(min(s, R2))) into caller buffers of nnz entries; then
 * mlk_radius_free. Rows are in ascending index order, or (distance, index)
 * with MLK_SORT_RESULTS. Replaces NeighborIndex.query_radius_csr
 * (src/mlxlearn/_kernels/distance.py:285-336) and the Metal route's device
 * screen + host finisher (radius_metal_v1.py:257-611; _radius_cpu_native.cpp:129-200).
 */
int mlk_radius_query(mlk_index* index, const float* Q_dev, int64_t m, double R2, int64_t q_offset,
                     uint32_t flags, int64_t* indptr_dev, int64_t* nnz, mlk_radius_result** out,
                     void* stream);
int mlk_radius_fill(mlk_radius_result* result, int64_t* indices_dev, double* dist_dev, void* stream);
int mlk_radius_free(mlk_radius_result* result);

/*
 * Reductions of a radius result (before or instead of mlk_radius_fill), for
 * RadiusNeighborsClassifier / RadiusNeighborsRegressor. Each query's entries
 * are summed in CSR order (ascending training index), in float64.
 * mlk_radius_votes: votes_dev (m x n_classes) = per-class sums of weights, the
 * class of training row i being codes_dev[i]; counts_dev (m, optional) =
 * neighbours per query. mlk_radius_mean: pred_dev (m) = weighted mean of
 * targets_dev over the neighbours, NaN for an empty row. Weights: uniform, or
 * distance (1/d; a row holding a zero distance uses the indicator of zero
 * distances), which needs a result made with MLK_NEED_DISTANCE. Replaces
 * reduce_votes_csr / reduce_mean_csr / distance_weights
 * (src/mlxlearn/_kernels/radius_reduce.py:19-73) behind
 * RadiusNeighborsClassifier / RadiusNeighborsRegressor
 * (src/mlxlearn/neighbors/_radius.py:162-338).
 */
int mlk_radius_votes(const mlk_radius_result* result, const int64_t* indptr_dev, const int32_t* codes_dev,
                     int32_t n_classes, int32_t weights, double* votes_dev, int64_t* counts_dev, void* stream);
int mlk_radius_mean(const mlk_radius_result* result, const int64_t* indptr_dev, const double* targets_dev,
                    int32_t weights, double* pred_dev, int64_t* counts_dev, void* stream);

/*
 * k-means assignment: exact nearest centre of every row of X (n x d) among
 * C (kc x d), lowest index on ties; residual_dev gets the canonical float64
 * squared distance. Replaces KMeans._assignment / e_means.assign
 * (src/mlxlearn/cluster/_k_means.py:326-349; src/mlxlearn/_kernels/e_means.py:245-345).
 */
int mlk_kmeans_assign(const float* X_dev, int64_t n, int32_t d, const float* C_dev, int64_t kc,
                      uint32_t flags, int64_t* labels_dev, double* residual_dev, void* stream);

/*
 * k-means assignment that reuses the previous labels prev_dev (n int64) where
 * a certificate proves them exact: a row keeps centre j when the distance from
 * j to its nearest other centre exceeds twice the row's distance to j (float64
 * bounds, so j is then the unique nearest centre); every other row gets the
 * full exact assignment. Outputs are bit-identical to mlk_kmeans_assign's;
 * labels_dev may be prev_dev. *reused (host, optional) receives the number of
 * rows that kept their label without a search. Replaces reusable_labels /
 * separation_lower and the pruned assignment tiles of the Lloyd loop
 * (src/mlxlearn/_kernels/lloyd_bounds.py:24-76;
 * src/mlxlearn/cluster/_k_means.py:159-175, 309-360).
 */
int mlk_kmeans_assign_reuse(const float* X_dev, int64_t n, int32_t d, const float* C_dev, int64_t kc,
                            const int64_t* prev_dev, uint32_t flags, int64_t* labels_dev,
                            double* residual_dev, int64_t* reused, void* stream);

/*
 * k-means update: float64 per-cluster coordinate sums (kc x d), counts (kc)
 * and inertia (sum of residuals), deterministic (fixed summation order).
 * Replaces KMeans._statistics / e_means.batch_stats
 * (src/mlxlearn/cluster/_k_means.py:351-361; src/mlxlearn/_kernels/e_means.py:371-382).
 */
int mlk_kmeans_update(const float* X_dev, int64_t n, int32_t d, const int64_t* labels_dev,
                      int64_t kc, const double* residual_dev, double* sums_dev, int64_t* counts_dev,
                      double* inertia_dev, void* stream);

/*
 * Class votes of k neighbours: pred_dev (m) = class code with the largest
 * vote, smallest code on ties; proba_dev (m x n_classes, optional) = votes
 * normalised per row. Distance weights follow sklearn (1/d; rows with a zero
 * distance use the indicator of zero distances). Replaces
 * class_predictions / class_probabilities (src/mlxlearn_device/_neighbors_votes.py:188-244)
 * and the CPU votes of KNeighborsClassifier (src/mlxlearn/neighbors/_classification.py:141-154).
 */
int mlk_knn_votes(const int64_t* idx_dev, const double* dist_dev, int64_t m, int32_t k,
                  const int32_t* codes_dev, int32_t n_classes, int32_t weights, int64_t* pred_dev,
                  double* proba_dev, void* stream);

/*
 * Diagnostics, no counterpart in the reference: runs `blocks` blocks of a
 * register-only bf16 MFMA loop (4 waves x 4 independent 32x32x16 MFMA chains,
 * `iterations` steps) twice and writes the second run's rate to *tflops. It
 * measures the attainable dense MFMA rate at the device's current clock and
 * is the reference point for MFMA busy-cycle counters.
 */
int mlk_mfma_probe(int32_t blocks, int32_t iterations, double* tflops, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* MLXLEARN_AMD_H */
