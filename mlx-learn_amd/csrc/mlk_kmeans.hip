// k-means assignment / update and k-NN class votes.
//
// Assignment is the k-NN search at k = 1 with the centroids as the database:
// exact nearest centre, lowest index on ties (the tie rule of the reference's
// argmin kernel, src/mlxlearn/_kernels/e_means.py:50-69, which itself is not
// certified). The update replaces the reference's float32 atomic scatter
// (e_means.py:86-98, 371-382) by float64 sums in a fixed order: rows are
// stably sorted by label and each cluster is summed in row order, so results
// are bitwise reproducible.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <cmath>
#include <vector>

#include "mlk_device.h"
#include "mlk_internal.h"

namespace mlk {

int create_index(const float* X, int64_t n, int d, hipStream_t stream, mlk_index** out);

static inline unsigned nblk(int64_t n, int t = 256) { return (unsigned)((n + t - 1) / t); }

// Labels outside [0, kc) are counted in counts[kc] and reported as an input
// error by the host; their rows go to the sentinel key kc and are not summed.
__global__ void k_labels_u32(const int64_t* __restrict__ labels, int64_t n, int64_t kc, unsigned* __restrict__ keys,
                             int* __restrict__ rows, unsigned long long* __restrict__ counts) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t raw = labels[i];
    const unsigned l = (raw >= 0 && raw < kc) ? (unsigned)raw : (unsigned)kc;
    keys[i] = l;
    rows[i] = (int)i;
    atomicAdd(&counts[l], 1ull);
}

// One wave per cluster: float64 sums of its rows in ascending row order.
__global__ void k_cluster_sums(const float* __restrict__ X, int d, const int* __restrict__ order,
                               const unsigned long long* __restrict__ offsets, int64_t kc, double* __restrict__ sums,
                               int64_t* __restrict__ counts_out) {
    const int lane = threadIdx.x & 63;
    const int64_t c = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (c >= kc) return;
    const unsigned long long b = offsets[c], e = offsets[c + 1];
    for (int j0 = 0; j0 < d; j0 += 64) {
        const int j = j0 + lane;
        double s = 0.0;
        if (j < d)
            for (unsigned long long r = b; r < e; ++r) s += (double)X[(int64_t)order[r] * d + j];
        if (j < d) sums[c * d + j] = s;
    }
    if (lane == 0) counts_out[c] = (int64_t)(e - b);
}

// Fixed-order float64 reduction: block partials then one block.
__global__ void k_sum_partials(const double* __restrict__ v, int64_t n, double* __restrict__ partial) {
    __shared__ double sh[256];
    double s = 0.0;
    const int64_t per = (n + gridDim.x - 1) / gridDim.x;
    const int64_t b = (int64_t)blockIdx.x * per;
    const int64_t e = b + per < n ? b + per : n;
    for (int64_t i = b + threadIdx.x; i < e; i += blockDim.x) s += v[i];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

__global__ void k_sum_final(const double* __restrict__ partial, int np, double* __restrict__ out) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        double s = 0.0;
        for (int i = 0; i < np; ++i) s += partial[i];
        *out = s;
    }
}

// One thread per query row: votes of its k neighbours in neighbour order.
__global__ void k_votes(const int64_t* __restrict__ idx, const double* __restrict__ dist, int64_t m, int k,
                        const int32_t* __restrict__ codes, int n_classes, int weights, double* __restrict__ votes,
                        int64_t* __restrict__ pred) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= m) return;
    double* v = votes + q * n_classes;
    for (int c = 0; c < n_classes; ++c) v[c] = 0.0;
    bool coincident = false;
    if (weights == MLK_WEIGHTS_DISTANCE)
        for (int j = 0; j < k; ++j) coincident |= dist[q * k + j] == 0.0;
    for (int j = 0; j < k; ++j) {
        const int c = codes[idx[q * k + j]];
        double w = 1.0;
        if (weights == MLK_WEIGHTS_DISTANCE) {
            const double dj = dist[q * k + j];
            w = coincident ? (dj == 0.0 ? 1.0 : 0.0) : 1.0 / dj;
        }
        v[c] += w;
    }
    int best = -1;
    double bv = 0.0, total = 0.0;
    for (int c = 0; c < n_classes; ++c) {
        total += v[c];
        if (v[c] > bv) { bv = v[c]; best = c; }  // strict: the smallest code wins ties
    }
    pred[q] = best;  // -1 when every weight is zero
    for (int c = 0; c < n_classes; ++c) v[c] = total > 0.0 ? v[c] / total : NAN;
}

}  // namespace mlk

using namespace mlk;

extern "C" int mlk_kmeans_assign(const float* X_dev, int64_t n, int32_t d, const float* C_dev, int64_t kc,
                                 uint32_t flags, int64_t* labels_dev, double* residual_dev, void* stream) {
    if (n < 0 || d < 1 || kc < 1 || (n && (!X_dev || !labels_dev)) || !C_dev) {
        set_error("kmeans_assign: invalid shapes or pointers");
        return MLK_EINVAL;
    }
    if (kc >= (int64_t(1) << 31)) {
        set_error("too many centres for 32-bit labels");
        return MLK_EINVAL;
    }
    if (n == 0) return MLK_OK;
    mlk_index* centres = nullptr;
    MLK_TRY(create_index(C_dev, kc, d, (hipStream_t)stream, &centres));
    struct G {
        mlk_index* p;
        ~G() { (void)hipDeviceSynchronize(); delete p; }
    } g{centres};
    const uint32_t qflags = (flags & MLK_PREC_MASK) | MLK_SQUARED_DIST;
    std::lock_guard<std::mutex> lk(centres->lock);
    return knn_query_impl(centres, X_dev, n, 1, 0, qflags, residual_dev, labels_dev, (hipStream_t)stream, nullptr, 0);
}

extern "C" int mlk_kmeans_update(const float* X_dev, int64_t n, int32_t d, const int64_t* labels_dev, int64_t kc,
                                 const double* residual_dev, double* sums_dev, int64_t* counts_dev, double* inertia_dev,
                                 void* stream_) {
    hipStream_t stream = (hipStream_t)stream_;
    if (n < 0 || d < 1 || kc < 1 || !sums_dev || !counts_dev || (n && (!X_dev || !labels_dev))) {
        set_error("kmeans_update: invalid shapes or pointers");
        return MLK_EINVAL;
    }
    if (n >= (int64_t(1) << 31)) {
        set_error("kmeans_update supports fewer than 2^31 rows per call");
        return MLK_EINVAL;
    }
    DevBuf keys, keys2, rows, rows2, cnt, off, tmp;
    MLK_TRY(keys.reserve((size_t)std::max<int64_t>(n, 1) * 4));
    MLK_TRY(keys2.reserve((size_t)std::max<int64_t>(n, 1) * 4));
    MLK_TRY(rows.reserve((size_t)std::max<int64_t>(n, 1) * 4));
    MLK_TRY(rows2.reserve((size_t)std::max<int64_t>(n, 1) * 4));
    MLK_TRY(cnt.reserve((size_t)(kc + 1) * 8));
    MLK_TRY(off.reserve((size_t)(kc + 1) * 8));
    MLK_CHECK_HIP(hipMemsetAsync(cnt.ptr, 0, (size_t)(kc + 1) * 8, stream));
    if (n) {
        hipLaunchKernelGGL(k_labels_u32, dim3(nblk(n)), dim3(256), 0, stream, labels_dev, n, kc, keys.as<unsigned>(),
                           rows.as<int>(), cnt.as<unsigned long long>());
        unsigned long long bad = 0;
        MLK_CHECK_HIP(hipMemcpyAsync(&bad, cnt.as<unsigned long long>() + kc, sizeof(bad), hipMemcpyDeviceToHost, stream));
        MLK_CHECK_HIP(hipStreamSynchronize(stream));
        if (bad) {
            set_error("kmeans_update: " + std::to_string(bad) + " labels outside [0, n_clusters)");
            return MLK_EINVAL;
        }
        int bits = 1;
        while ((1ll << bits) <= kc) ++bits;
        size_t tb = 0;
        MLK_CHECK_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, keys.as<unsigned>(), keys2.as<unsigned>(),
                                                         rows.as<int>(), rows2.as<int>(), (int)n, 0, bits, stream));
        MLK_TRY(tmp.reserve(tb + 16));
        MLK_CHECK_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.ptr, tb, keys.as<unsigned>(), keys2.as<unsigned>(),
                                                         rows.as<int>(), rows2.as<int>(), (int)n, 0, bits, stream));
    }
    {
        size_t tb = 0;
        MLK_CHECK_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, cnt.as<unsigned long long>(),
                                                       off.as<unsigned long long>(), (int)(kc + 1), stream));
        MLK_TRY(tmp.reserve(tb + 16));
        MLK_CHECK_HIP(hipcub::DeviceScan::ExclusiveSum(tmp.ptr, tb, cnt.as<unsigned long long>(),
                                                       off.as<unsigned long long>(), (int)(kc + 1), stream));
    }
    hipLaunchKernelGGL(k_cluster_sums, dim3(nblk(kc, 4)), dim3(256), 0, stream, X_dev, d, rows2.as<int>(),
                       off.as<unsigned long long>(), kc, sums_dev, counts_dev);
    MLK_CHECK_HIP(hipGetLastError());
    if (inertia_dev) {
        if (!residual_dev && n) {
            set_error("inertia requires residuals");
            return MLK_EINVAL;
        }
        const int parts = 1024;
        DevBuf partial;
        MLK_TRY(partial.reserve(parts * sizeof(double)));
        if (n) {
            hipLaunchKernelGGL(k_sum_partials, dim3(parts), dim3(256), 0, stream, residual_dev, n, partial.as<double>());
            hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(64), 0, stream, partial.as<double>(), parts, inertia_dev);
        } else {
            MLK_CHECK_HIP(hipMemsetAsync(inertia_dev, 0, sizeof(double), stream));
        }
        MLK_CHECK_HIP(hipGetLastError());
    }
    MLK_CHECK_HIP(hipStreamSynchronize(stream));
    return MLK_OK;
}

extern "C" int mlk_knn_votes(const int64_t* idx_dev, const double* dist_dev, int64_t m, int32_t k,
                             const int32_t* codes_dev, int32_t n_classes, int32_t weights, int64_t* pred_dev,
                             double* proba_dev, void* stream_) {
    hipStream_t stream = (hipStream_t)stream_;
    if (m < 0 || k < 1 || n_classes < 1 || (m && (!idx_dev || !codes_dev || !pred_dev))) {
        set_error("knn_votes: invalid shapes or pointers");
        return MLK_EINVAL;
    }
    if (weights != MLK_WEIGHTS_UNIFORM && weights != MLK_WEIGHTS_DISTANCE) {
        set_error("weights must be uniform or distance");
        return MLK_EINVAL;
    }
    if (weights == MLK_WEIGHTS_DISTANCE && m && !dist_dev) {
        set_error("weights='distance' requires neighbor distances.");
        return MLK_EINVAL;
    }
    if (m == 0) return MLK_OK;
    DevBuf scratch;
    double* votes = proba_dev;
    if (!votes) {
        MLK_TRY(scratch.reserve((size_t)m * n_classes * sizeof(double)));
        votes = scratch.as<double>();
    }
    hipLaunchKernelGGL(k_votes, dim3(nblk(m)), dim3(256), 0, stream, idx_dev, dist_dev, m, k, codes_dev, n_classes,
                       weights, votes, pred_dev);
    MLK_CHECK_HIP(hipGetLastError());
    MLK_CHECK_HIP(hipStreamSynchronize(stream));
    return MLK_OK;
}
