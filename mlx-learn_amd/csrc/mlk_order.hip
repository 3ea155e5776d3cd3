// Query locality order: a call's queries regrouped by their nearest order
// centre (one of up to ORDER_CENTRES database rows sampled at a fixed stride)
// before the screen. Queries that share a neighbourhood then share the MFMA
// tiles of the screen, so one database row that lies within several of their
// thresholds is appended for all of them by one pass of the hit path instead
// of one pass each. The order changes no result: every query is screened,
// thresholded and certified on its own, and the candidates carry the query's
// original index back (the map position -> query).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "mlk_internal.h"
#include "mlk_order.h"

namespace mlk {

// rows of d features (d <= ORDER_MAX_D) padded to DW = 32, 64, 96 or 128 floats
static int order_width(int d) { return (d + 31) / 32 * 32; }

bool order_wanted(const mlk_index* index, int64_t m) {
    const TestHooks hk = hooks_snapshot();
    if (hk.hooks & MLK_HOOK_NO_ORDER) return false;
    if (index->d > ORDER_MAX_D || index->n < 2) return false;
    if (hk.hooks & MLK_HOOK_FORCE_ORDER) return true;
    return m >= ORDER_MIN_QUERIES && index->n >= ORDER_MIN_ROWS;
}

// Centre j is database row j * (n / K), zero-padded to DW floats; cn[j] its
// squared norm (float32).
__global__ void k_order_centres(const float* __restrict__ X, int64_t n, int d, int K, int dw,
                                float* __restrict__ C, float* __restrict__ cn) {
    const int j = blockIdx.x;
    const int64_t row = (int64_t)j * (n / K);
    float s = 0.0f;
    for (int e = threadIdx.x; e < dw; e += 64) {
        const float v = e < d ? X[row * d + e] : 0.0f;
        C[(size_t)j * dw + e] = v;
        s = fmaf(v, v, s);
    }
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if (threadIdx.x == 0) cn[j] = s;
}

int prepare_order(mlk_index* index, hipStream_t stream) {
    if (index->order_k) return MLK_OK;
    const int K = (int)std::min<int64_t>(ORDER_CENTRES, index->n);
    const int dw = order_width(index->d);
    MLK_TRY(index->order_c.reserve((size_t)K * (dw + 1) * sizeof(float)));
    float* C = index->order_c.as<float>();
    hipLaunchKernelGGL(k_order_centres, dim3((unsigned)K), dim3(64), 0, stream, index->X.as<float>(), index->n,
                       index->d, K, dw, C, C + (size_t)K * dw);
    MLK_CHECK_HIP(hipGetLastError());
    index->order_k = K;
    return MLK_OK;
}

__global__ void k_order_zero(int* __restrict__ v, int count) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) v[i] = 0;
}

// One thread per query: the nearest centre by |c|^2 - 2 q.c in float32 (ties
// and rounding may pick a near-nearest centre: any choice is a valid order).
// Centres stream through LDS in 32 KiB chunks; the query stays in registers.
template <int DW>
__global__ __launch_bounds__(256) void k_order_assign(const float* __restrict__ Q, int64_t m, int d,
                                                      const float* __restrict__ C, const float* __restrict__ cn,
                                                      int K, int* __restrict__ cell, int* __restrict__ count) {
    constexpr int CH = 8192 / DW;
    __shared__ __attribute__((aligned(16))) float sc[CH * DW];
    __shared__ float scn[CH];
    const int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x;
    float qv[DW];
#pragma unroll
    for (int j = 0; j < DW; ++j) qv[j] = (q < m && j < d) ? Q[q * d + j] : 0.0f;
    float best = INFINITY;
    int bi = 0;
    for (int c0 = 0; c0 < K; c0 += CH) {
        const int nc = K - c0 < CH ? K - c0 : CH;
        __syncthreads();
        const float4* src = reinterpret_cast<const float4*>(C + (size_t)c0 * DW);
        for (int i = threadIdx.x; i < nc * (DW / 4); i += 256) reinterpret_cast<float4*>(sc)[i] = src[i];
        for (int i = threadIdx.x; i < nc; i += 256) scn[i] = cn[c0 + i];
        __syncthreads();
        for (int c = 0; c < nc; ++c) {
            const float4* cv = reinterpret_cast<const float4*>(sc + c * DW);
            float dot = 0.0f;
#pragma unroll
            for (int j = 0; j < DW / 4; ++j) {
                const float4 v = cv[j];
                dot = fmaf(qv[4 * j], v.x, dot);
                dot = fmaf(qv[4 * j + 1], v.y, dot);
                dot = fmaf(qv[4 * j + 2], v.z, dot);
                dot = fmaf(qv[4 * j + 3], v.w, dot);
            }
            const float s = scn[c] - 2.0f * dot;
            if (s < best) {
                best = s;
                bi = c0 + c;
            }
        }
    }
    if (q < m) {
        cell[q] = bi;
        atomicAdd(&count[bi], 1);
    }
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
// follows the atomics and does not matter).
// Positions m .. m_pad - 1 (padding queries) map to themselves.
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

size_t order_scratch_bytes(int64_t m) { return (size_t)m * sizeof(int) + 2 * ORDER_CENTRES * sizeof(int) + 256; }

int query_order(mlk_index* index, const float* Q, int64_t m, int64_t m_pad, void* scratch, int* qperm, float* Qp,
                hipStream_t stream) {
    MLK_TRY(prepare_order(index, stream));
    const int K = index->order_k, d = index->d, dw = order_width(d);
    const float* C = index->order_c.as<float>();
    int* cell = static_cast<int*>(scratch);
    int* count = cell + m;
    int* start = count + ORDER_CENTRES;
    hipLaunchKernelGGL(k_order_zero, dim3((K + 255) / 256), dim3(256), 0, stream, count, K);
    const dim3 grid((unsigned)((m + 255) / 256));
    switch (dw) {
        case 32: hipLaunchKernelGGL(k_order_assign<32>, grid, dim3(256), 0, stream, Q, m, d, C, C + (size_t)K * dw, K, cell, count); break;
        case 64: hipLaunchKernelGGL(k_order_assign<64>, grid, dim3(256), 0, stream, Q, m, d, C, C + (size_t)K * dw, K, cell, count); break;
        case 96: hipLaunchKernelGGL(k_order_assign<96>, grid, dim3(256), 0, stream, Q, m, d, C, C + (size_t)K * dw, K, cell, count); break;
        case 128: hipLaunchKernelGGL(k_order_assign<128>, grid, dim3(256), 0, stream, Q, m, d, C, C + (size_t)K * dw, K, cell, count); break;
        default:
            set_error("query order: row width above the order kernels' limit");
            return MLK_EINVAL;
    }
    hipLaunchKernelGGL(k_order_scan, dim3(1), dim3(1024), 0, stream, count, K, start);
    hipLaunchKernelGGL(k_order_scatter, dim3((unsigned)((m_pad + 255) / 256)), dim3(256), 0, stream, cell, m, m_pad,
                       start, qperm);
    hipLaunchKernelGGL(k_order_gather, dim3((unsigned)((m * d + 255) / 256)), dim3(256), 0, stream, Q, qperm, m, d,
                       Qp);
    MLK_CHECK_HIP(hipGetLastError());
    return MLK_OK;
}

}  // namespace mlk
