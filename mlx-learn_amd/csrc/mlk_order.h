// Query locality order (mlk_order.hip): regroup a call's queries by their
// nearest order centre so that queries with shared neighbourhoods share the
// screen's tiles.
#pragma once

#include <hip/hip_runtime.h>

#include "mlk_internal.h"

namespace mlk {

constexpr int ORDER_CENTRES = 1024;          // database rows sampled as order centres (at most)
constexpr int ORDER_MAX_D = 128;             // widest rows the order kernels take
constexpr int64_t ORDER_MIN_QUERIES = 4096;  // smaller calls keep their order
// larger calls keep their order too: the regrouping's cost grows with the
// queries while its gain is spread over fewer hits per query (10^7 k-means
// rows over 65,536 centroids: 20 ms of regrouping for a 10 ms shorter screen)
constexpr int64_t ORDER_MAX_QUERIES = int64_t(1) << 20;
constexpr int64_t ORDER_MIN_ROWS = 65536;

// Whether a call of m queries on this index is regrouped (test hooks
// MLK_HOOK_FORCE_ORDER / MLK_HOOK_NO_ORDER override the size rule).
bool order_wanted(const mlk_index* index, int64_t m);

// Device bytes of query_order's scratch for m queries.
size_t order_scratch_bytes(int64_t m);

// The screen order of m queries: qperm[p] = original index of the query at
// position p (m_pad entries; padding positions p >= m map to p), Qp = the
// queries in that order (m x d float32). Stream-ordered; allocates only the
// index's order centres, on its first use.
int query_order(mlk_index* index, const float* Q, int64_t m, int64_t m_pad, void* scratch, int* qperm, float* Qp,
                hipStream_t stream);

// The database rows grouped by nearest order centre (index->order_rows,
// order_start), built on first use: the start-threshold samples of regrouped
// k-NN queries come from their own neighbourhoods.
int prepare_order_rows(mlk_index* index, hipStream_t stream);

// Start-threshold sample rows of regrouped queries: block b of the start
// threshold kernel (queries at positions b * qpb .. + qpb - 1) samples up to
// `samples` distinct rows, srows[b * samples + j] (-1: no sample): rows of the
// centres of its queries, and rows at a stride. `cell` is query_order's
// per-query centre (the scratch's first m ints).
int tau_sample_rows(const mlk_index* index, const int* cell, const int* qperm, int64_t m, int qpb, int64_t blocks,
                    int samples, int* srows, hipStream_t stream);

}  // namespace mlk
