// Query locality order: a call's queries regrouped by their nearest order
// centre (one of up to ORDER_CENTRES database rows sampled at a fixed stride)
// before the screen. Queries that share a neighbourhood then share the MFMA
// tiles of the screen, so one database row that lies within several of their
// thresholds is appended for all of them by one pass of the hit path instead
// of one pass each. The order changes no result: every query is screened,
// thresholded and certified on its own, and the candidates carry the query's
// original index back (the map position -> query).
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>

#include "mlk_device.h"
#include "mlk_internal.h"
#include "mlk_order.h"

namespace mlk {

bool order_wanted(const mlk_index* index, int64_t m) {
    const TestHooks hk = hooks_snapshot();
    if (hk.hooks & MLK_HOOK_NO_ORDER) return false;
    if (index->d > ORDER_MAX_D || index->n < 2) return false;
    if (hk.hooks & MLK_HOOK_FORCE_ORDER) return true;
    return m >= ORDER_MIN_QUERIES && m <= ORDER_MAX_QUERIES && index->n >= ORDER_MIN_ROWS;
}

// padded sizes of the bf16 centre copy: centres to a multiple of 32 (the
// MFMA tile), features to a multiple of 16 (one k-step)
static int centres_pad(int K) { return (K + 31) / 32 * 32; }
static int features_pad(int d) { return (d + 15) / 16 * 16; }

// Centre j < K is database row j * (n / K): its bf16 copy (zero-padded to dp
// features) and its float32 squared norm; padding centres get +inf norms, so
// no row picks them.
__global__ void k_order_centres(const float* __restrict__ X, int64_t n, int d, int K, int dp,
                                __bf16* __restrict__ Cb, float* __restrict__ cn) {
    const int j = blockIdx.x;
    const int64_t row = (int64_t)j * (n / (K > 0 ? K : 1));
    float s = 0.0f;
    for (int e = threadIdx.x; e < dp; e += 64) {
        const float v = (j < K && e < d) ? X[row * d + e] : 0.0f;
        Cb[(size_t)j * dp + e] = (__bf16)v;
        s = fmaf(v, v, s);
    }
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if (threadIdx.x == 0) cn[j] = j < K ? s : INFINITY;
}

int prepare_order(mlk_index* index, hipStream_t stream) {
    if (index->order_k) return MLK_OK;
    const int K = (int)std::min<int64_t>(ORDER_CENTRES, index->n);
    const int kp = centres_pad(K), dp = features_pad(index->d);
    MLK_TRY(index->order_cb.reserve((size_t)kp * dp * sizeof(__bf16)));
    MLK_TRY(index->order_cn.reserve((size_t)kp * sizeof(float)));
    hipLaunchKernelGGL(k_order_centres, dim3((unsigned)kp), dim3(64), 0, stream, index->X.as<float>(), index->n,
                       index->d, K, dp, index->order_cb.as<__bf16>(), index->order_cn.as<float>());
    MLK_CHECK_HIP(hipGetLastError());
    index->order_k = K;
    return MLK_OK;
}

__global__ void k_order_zero(int* __restrict__ v, int count) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) v[i] = 0;
}

// Nearest centre of each row by |c|^2 - 2 v.c, the products on bf16 MFMA
// (32x32x16). A block takes 32 rows, held by each of its waves as B fragments
// (row = lane % 32; lane half h supplies features 16 j + 8 h .. + 8 of k-step
// j); wave w walks centre tiles w, w + 4, ... as A fragments (the next tile's
// loaded while the current one multiplies), so a lane's 16 accumulators are 16
// centres of its row. Each lane keeps its row's best (ties to the lower
// centre); the lane halves and then the waves combine. Rounding may pick a
// near-nearest centre: any choice is a valid order. NB k-steps (16 NB <= 128
// features).
template <int NB>
__global__ __launch_bounds__(256) void k_order_assign(const float* __restrict__ V, int64_t rows, int d,
                                                      const __bf16* __restrict__ Cb, const float* __restrict__ cn,
                                                      int kp, int* __restrict__ cell, int* __restrict__ count) {
    constexpr int DP = 16 * NB;
    __shared__ float sbest[4][32];
    __shared__ int sidx[4][32];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, h = lane >> 5, col = lane & 31;
    const int64_t r = (int64_t)blockIdx.x * 32 + col;
    bf16x8 bfr[NB];
#pragma unroll
    for (int j = 0; j < NB; ++j)
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int f = 16 * j + 8 * h + e;
            bfr[j][e] = (__bf16)((r < rows && f < d) ? V[r * d + f] : 0.0f);
        }
    float best = INFINITY;
    int bi = 0;
    const int tiles = kp / 32;
    auto load_tile = [&](int t, bf16x8(&a)[NB], f32x4(&cv)[4]) {
        const __bf16* cr = Cb + (size_t)(t * 32 + col) * DP + 8 * h;
#pragma unroll
        for (int j = 0; j < NB; ++j) a[j] = *reinterpret_cast<const bf16x8*>(cr + 16 * j);
        // norms of the lane's 16 centres: four runs of four
#pragma unroll
        for (int q = 0; q < 4; ++q) cv[q] = *reinterpret_cast<const f32x4*>(cn + t * 32 + 8 * q + 4 * h);
    };
    bf16x8 a[NB];
    f32x4 cv[4];
    if (w < tiles) load_tile(w, a, cv);
    for (int t = w; t < tiles; t += 4) {
        bf16x8 an[NB];
        f32x4 cvn[4];
        if (t + 4 < tiles) load_tile(t + 4, an, cvn);
        f32x16 acc = {};
#pragma unroll
        for (int j = 0; j < NB; ++j) acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[j], bfr[j], acc, 0, 0, 0);
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const float s = cv[i >> 2][i & 3] - 2.0f * acc[i];
            const int c = t * 32 + mfma32_row(i, lane);
            if (s < best) {
                best = s;
                bi = c;
            }
        }
        if (t + 4 < tiles) {
#pragma unroll
            for (int j = 0; j < NB; ++j) a[j] = an[j];
#pragma unroll
            for (int q = 0; q < 4; ++q) cv[q] = cvn[q];
        }
    }
    const float ob = __shfl_xor(best, 32);
    const int oi = __shfl_xor(bi, 32);
    if (ob < best || (ob == best && oi < bi)) {
        best = ob;
        bi = oi;
    }
    if (h == 0) {
        sbest[w][col] = best;
        sidx[w][col] = bi;
    }
    __syncthreads();
    if (threadIdx.x < 32 && (int64_t)blockIdx.x * 32 + threadIdx.x < rows) {
        float b0 = sbest[0][threadIdx.x];
        int i0 = sidx[0][threadIdx.x];
        for (int v = 1; v < 4; ++v) {
            const float bv = sbest[v][threadIdx.x];
            const int iv = sidx[v][threadIdx.x];
            if (bv < b0 || (bv == b0 && iv < i0)) {
                b0 = bv;
                i0 = iv;
            }
        }
        cell[(int64_t)blockIdx.x * 32 + threadIdx.x] = i0;
        atomicAdd(&count[i0], 1);
    }
}

// Nearest centre of every row of V (rows x d): cell[r], and the number of
// rows per centre in count (ORDER_CENTRES entries, cleared here).
static int assign_centres(const mlk_index* index, const float* V, int64_t rows, int* cell, int* count,
                          hipStream_t stream) {
    const int kp = centres_pad(index->order_k), nb = features_pad(index->d) / 16;
    const __bf16* Cb = index->order_cb.as<__bf16>();
    const float* cn = index->order_cn.as<float>();
    hipLaunchKernelGGL(k_order_zero, dim3((ORDER_CENTRES + 255) / 256), dim3(256), 0, stream, count, ORDER_CENTRES);
    const dim3 grid((unsigned)((rows + 31) / 32));
    switch (nb) {
#define MLK_ASSIGN_CASE(N) \
        case N: hipLaunchKernelGGL(k_order_assign<N>, grid, dim3(256), 0, stream, V, rows, index->d, Cb, cn, kp, cell, count); break;
        MLK_ASSIGN_CASE(1) MLK_ASSIGN_CASE(2) MLK_ASSIGN_CASE(3) MLK_ASSIGN_CASE(4)
        MLK_ASSIGN_CASE(5) MLK_ASSIGN_CASE(6) MLK_ASSIGN_CASE(7) MLK_ASSIGN_CASE(8)
#undef MLK_ASSIGN_CASE
        default:
            set_error("query order: row width above the order kernels' limit");
            return MLK_EINVAL;
    }
    MLK_CHECK_HIP(hipGetLastError());
    return MLK_OK;
}

// Exclusive scan of the K <= 1024 cell counts (one block of 1024 threads).
__global__ __launch_bounds__(1024) void k_order_scan(const int* __restrict__ count, int K, int* __restrict__ start) {
    __shared__ int wsum[16];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int v = t < K ? count[t] : 0;
    int inc = v;
    for (int o = 1; o < 64; o <<= 1) {
        const int u = __shfl_up(inc, o);
        if (lane >= o) inc += u;
    }
    if (lane == 63) wsum[w] = inc;
    __syncthreads();
    int before = 0;
    for (int i = 0; i < w; ++i) before += wsum[i];
    if (t < K) start[t] = before + inc - v;
}

// Each query takes the next position of its cell (the order within a cell
// follows the atomics and does not matter). Positions m .. m_pad - 1 (padding
// queries) map to themselves.
__global__ void k_order_scatter(const int* __restrict__ cell, int64_t m, int64_t m_pad, int* __restrict__ start,
                                int* __restrict__ qperm) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= m_pad) return;
    if (q >= m) {
        qperm[q] = (int)q;
        return;
    }
    const int pos = atomicAdd(&start[cell[q]], 1);
    qperm[pos] = (int)q;
}

// Qp row p = Q row qperm[p].
__global__ void k_order_gather(const float* __restrict__ Q, const int* __restrict__ qperm, int64_t m, int d,
                               float* __restrict__ Qp) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m * d) return;
    const int64_t p = i / d, e = i - p * d;
    Qp[i] = Q[(int64_t)qperm[p] * d + e];
}

// scratch: cell (m ints), count and start (ORDER_CENTRES ints each)
size_t order_scratch_bytes(int64_t m) { return (size_t)m * sizeof(int) + 2 * ORDER_CENTRES * sizeof(int) + 256; }

int query_order(mlk_index* index, const float* Q, int64_t m, int64_t m_pad, void* scratch, int* qperm, float* Qp,
                hipStream_t stream) {
    MLK_TRY(prepare_order(index, stream));
    const int K = index->order_k, d = index->d;
    int* cell = static_cast<int*>(scratch);
    int* count = cell + m;
    int* start = count + ORDER_CENTRES;
    MLK_TRY(assign_centres(index, Q, m, cell, count, stream));
    hipLaunchKernelGGL(k_order_scan, dim3(1), dim3(1024), 0, stream, count, K, start);
    hipLaunchKernelGGL(k_order_scatter, dim3((unsigned)((m_pad + 255) / 256)), dim3(256), 0, stream, cell, m, m_pad,
                       start, qperm);
    hipLaunchKernelGGL(k_order_gather, dim3((unsigned)((m * d + 255) / 256)), dim3(256), 0, stream, Q, qperm, m, d,
                       Qp);
    MLK_CHECK_HIP(hipGetLastError());
    return MLK_OK;
}

__global__ void k_order_iota(int* __restrict__ v, int64_t count) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) v[i] = (int)i;
}

__global__ void k_order_last(int* __restrict__ start, int K, int64_t n) {
    if (threadIdx.x == 0) start[K] = (int)n;
}

int prepare_order_rows(mlk_index* index, hipStream_t stream) {
    if (index->order_rows_ready) return MLK_OK;
    MLK_TRY(prepare_order(index, stream));
    const int64_t n = index->n;
    const int K = index->order_k;
    if (n > INT32_MAX) {
        set_error("query order: more rows than the order's 32-bit row lists hold");
        return MLK_ECAPACITY;
    }
    MLK_TRY(index->order_rows.reserve((size_t)n * sizeof(int)));
    MLK_TRY(index->order_start.reserve((size_t)(K + 1) * sizeof(int)));
    MLK_TRY(index->order_cell.reserve((size_t)n * sizeof(int)));
    // temporaries: the row ids, the sorted centres, the counts and the sort's scratch
    DevBuf ids, cell_out, count, tmp;
    MLK_TRY(ids.reserve((size_t)n * sizeof(int)));
    MLK_TRY(cell_out.reserve((size_t)n * sizeof(int)));
    MLK_TRY(count.reserve((size_t)ORDER_CENTRES * sizeof(int)));
    int* cell = index->order_cell.as<int>();
    MLK_TRY(assign_centres(index, index->X.as<float>(), n, cell, count.as<int>(), stream));
    hipLaunchKernelGGL(k_order_scan, dim3(1), dim3(1024), 0, stream, count.as<int>(), K, index->order_start.as<int>());
    hipLaunchKernelGGL(k_order_last, dim3(1), dim3(64), 0, stream, index->order_start.as<int>(), K, n);
    hipLaunchKernelGGL(k_order_iota, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, ids.as<int>(), n);
    // a stable sort by centre: each centre's rows in ascending order
    int bits = 1;
    while ((1 << bits) < K) ++bits;
    size_t tmp_bytes = 0;
    MLK_CHECK_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, cell, cell_out.as<int>(), ids.as<int>(),
                                                     index->order_rows.as<int>(), (int)n, 0, bits, stream));
    MLK_TRY(tmp.reserve(tmp_bytes + 16));
    MLK_CHECK_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.ptr, tmp_bytes, cell, cell_out.as<int>(), ids.as<int>(),
                                                     index->order_rows.as<int>(), (int)n, 0, bits, stream));
    MLK_CHECK_HIP(hipGetLastError());
    MLK_CHECK_HIP(hipStreamSynchronize(stream));  // the temporaries are freed on return
    index->order_rows_ready = true;
    return MLK_OK;
}

constexpr int TAU_CELLS = 8;  // centres sampled per start-threshold block (at most)

// Block b's start-threshold samples: three quarters from the centres of its
// queries (the first TAU_CELLS distinct ones in position order, an equal share
// each, ascending rows of the centre), the rest database rows at a stride;
// a stride row of a sampled centre is dropped (-1), so every sample is a
// distinct row. The stride rows bound the threshold of a query whose own
// centre is not sampled as an unregrouped query's would be.
__global__ void k_tau_rows(const int* __restrict__ cell, const int* __restrict__ qperm, int64_t m, int qpb,
                           int samples, int64_t n, const int* __restrict__ rows, const int* __restrict__ start,
                           const int* __restrict__ row_cell, int* __restrict__ srows) {
    const int64_t b = blockIdx.x;
    const int j = threadIdx.x;
    if (j >= samples) return;
    int64_t p0 = b * qpb, p1 = p0 + qpb - 1;
    if (p1 >= m) p1 = m - 1;
    if (p0 > p1) p0 = p1;  // a block of padding queries
    int cells[TAU_CELLS];
    int nc = 0;
    for (int64_t p = p0; p <= p1 && nc < TAU_CELLS; ++p) {
        const int c = cell[qperm[p]];
        bool seen = false;
        for (int i = 0; i < nc; ++i) seen |= cells[i] == c;
        if (!seen) cells[nc++] = c;
    }
    const int share = samples * 3 / 4 / nc;
    const int from_cells = share * nc;
    int row = -1;
    if (j < from_cells) {
        const int c = cells[j / share], t = j % share;
        if (t < start[c + 1] - start[c]) row = rows[start[c] + t];
    } else {
        const int64_t stride = n / (samples - from_cells);
        const int64_t r = (int64_t)(j - from_cells) * stride;
        const int c = row_cell[r];
        bool taken = false;
        for (int i = 0; i < nc; ++i) taken |= cells[i] == c;
        row = taken ? -1 : (int)r;
    }
    srows[b * samples + j] = row;
}

int tau_sample_rows(const mlk_index* index, const int* cell, const int* qperm, int64_t m, int qpb, int64_t blocks,
                    int samples, int* srows, hipStream_t stream) {
    if (samples > 256 || samples > index->n || samples < 1) {
        set_error("start-threshold samples outside the row builder's range");
        return MLK_EINVAL;
    }
    hipLaunchKernelGGL(k_tau_rows, dim3((unsigned)blocks), dim3(256), 0, stream, cell, qperm, m, qpb, samples,
                       index->n, index->order_rows.as<int>(), index->order_start.as<int>(),
                       index->order_cell.as<int>(), srows);
    MLK_CHECK_HIP(hipGetLastError());
    return MLK_OK;
}

}  // namespace mlk
