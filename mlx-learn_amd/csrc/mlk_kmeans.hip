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

// The centres index of one assignment: it borrows the device's pooled
// workspace and hands it back (after the device has finished with it).
struct TransientIndex {
    mlk_index* p;
    explicit TransientIndex(mlk_index* index) : p(index) { workspace_adopt(p); }
    ~TransientIndex() {
        (void)hipDeviceSynchronize();
        workspace_return(p);
        delete p;
    }
};

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

// One wave per cluster: float64 sums of its rows in ascending row order. Lane
// l owns features l + 64 t (t < NJ) of a 64 NJ-feature block; the wave reads
// 64 row numbers at a time (one coalesced load) and keeps eight rows' values
// in flight, while each feature's additions stay one row at a time in order.
template <int NJ>
__global__ void k_cluster_sums(const float* __restrict__ X, int d, const int* __restrict__ order,
                               const unsigned long long* __restrict__ offsets, int64_t kc, double* __restrict__ sums,
                               int64_t* __restrict__ counts_out) {
    constexpr int U = 8;  // rows in flight
    const int lane = threadIdx.x & 63;
    const int64_t c = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (c >= kc) return;
    const unsigned long long b = offsets[c], e = offsets[c + 1];
    for (int j0 = 0; j0 < d; j0 += 64 * NJ) {
        double s[NJ];
#pragma unroll
        for (int t = 0; t < NJ; ++t) s[t] = 0.0;
        for (unsigned long long r0 = b; r0 < e; r0 += 64) {
            const int cnt = e - r0 < 64 ? (int)(e - r0) : 64;
            const int my_row = lane < cnt ? order[r0 + lane] : 0;
            for (int u0 = 0; u0 < cnt; u0 += U) {
                float v[U][NJ];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const int row = __shfl(my_row, u0 + u < cnt ? u0 + u : 0);
#pragma unroll
                    for (int t = 0; t < NJ; ++t) {
                        const int j = j0 + lane + 64 * t;
                        v[u][t] = (u0 + u < cnt && j < d) ? X[(int64_t)row * d + j] : 0.0f;
                    }
                }
#pragma unroll
                for (int u = 0; u < U; ++u)
                    if (u0 + u < cnt)
#pragma unroll
                        for (int t = 0; t < NJ; ++t) s[t] += (double)v[u][t];
            }
        }
#pragma unroll
        for (int t = 0; t < NJ; ++t) {
            const int j = j0 + lane + 64 * t;
            if (j < d) sums[c * d + j] = s[t];
        }
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

// ------------------------------------------------------- label reuse (Lloyd)
// Certified reuse of the previous labels (reference lloyd_bounds.py:24-76,
// which certifies against float32 scores with rounding allowances; here the
// certificate is exact). With sep_j the distance from centre j to its nearest
// other centre and R the distance from row x to centre j: if sep_j > 2R, then
// for every other centre i, |x - c_i| >= |c_i - c_j| - |x - c_j| >= sep_j - R > R,
// so j is the unique nearest centre and the row keeps it. sep_j is bounded
// below and R above from their canonical float64 squares (relative error at
// most f64_interval(d)), with a 2^-50 margin for the square roots.

// Lower bound of each centre's separation from its exact nearest-other-centre
// squared distance (float64, +inf when there is no other centre).
__global__ void k_sep_lower(const double* __restrict__ sep2, int64_t kc, int d, double* __restrict__ sep_lo) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= kc) return;
    const double s = sep2 ? sep2[j] : INFINITY;
    sep_lo[j] = s < INFINITY ? sqrt(fmax(s, 0.0) * (1.0 - f64_interval(d))) * (1.0 - 0x1p-50) : INFINITY;
}

// One thread per row: canonical residual to the previous centre and the
// certificate; uncertain rows (and labels outside [0, kc), counted in
// words[1]) are appended to `uncertain`, one atomic per wave.
template <bool VEC>
__global__ void k_reuse(const float* __restrict__ X, int64_t n, int d, const float* __restrict__ C, int64_t kc,
                        const int64_t* __restrict__ prev, const double* __restrict__ sep_lo,
                        int64_t* __restrict__ labels, double* __restrict__ residual, int* __restrict__ uncertain,
                        unsigned* __restrict__ words) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63;
    bool keep = false, bad = false;
    if (i < n) {
        const int64_t j = prev[i];
        bad = j < 0 || j >= kc;
        if (!bad) {
            const double r2 = VEC ? canonical_sq4(X + i * d, C + j * d, d) : canonical_sq(X + i * d, C + j * d, d);
            const double r_hi = sqrt(r2 * (1.0 + f64_interval(d))) * (1.0 + 0x1p-50);
            keep = sep_lo[j] > 2.0 * r_hi;
            if (keep) {
                labels[i] = j;
                residual[i] = r2;
            }
        }
    }
    const bool send = i < n && !keep;
    const unsigned long long mask = ballot(send);
    const unsigned long long badmask = ballot(bad);
    unsigned base = 0;
    if (lane == 0 && mask) base = atomicAdd(&words[0], (unsigned)__popcll(mask));
    if (lane == 0 && badmask) atomicAdd(&words[1], (unsigned)__popcll(badmask));
    base = __shfl(base, 0);
    if (send) {
        const unsigned below = __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0));
        uncertain[base + below] = (int)i;
    }
}

// Labels outside [0, kc) counted in words[1].
__global__ void k_prev_check(const int64_t* __restrict__ prev, int64_t n, int64_t kc, unsigned* __restrict__ words) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool bad = i < n && (prev[i] < 0 || prev[i] >= kc);
    const unsigned long long mask = ballot(bad);
    if ((threadIdx.x & 63) == 0 && mask) atomicAdd(&words[1], (unsigned)__popcll(mask));
}

// Certified rows among `samples` evenly spaced rows (words[2]); the gate that
// decides whether a reuse pass can pay for itself.
__global__ void k_reuse_sample(const float* __restrict__ X, int64_t n, int d, const float* __restrict__ C,
                               int64_t kc, const int64_t* __restrict__ prev, const double* __restrict__ sep_lo,
                               int samples, unsigned* __restrict__ words) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    bool keep = false;
    if (t < samples) {
        const int64_t i = (int64_t)t * (n / samples);
        const int64_t j = prev[i];
        if (j >= 0 && j < kc) {
            const double r2 = canonical_sq(X + i * d, C + j * d, d);
            keep = sep_lo[j] > 2.0 * sqrt(r2 * (1.0 + f64_interval(d))) * (1.0 + 0x1p-50);
        }
    }
    const unsigned long long mask = ballot(keep);
    if ((threadIdx.x & 63) == 0 && mask) atomicAdd(&words[2], (unsigned)__popcll(mask));
}

// Gather the uncertain rows into a compact matrix (one block per row).
__global__ void k_gather_rows(const float* __restrict__ X, int d, const int* __restrict__ rows, int64_t count,
                              float* __restrict__ out) {
    const int64_t r = blockIdx.x;
    if (r >= count) return;
    const float* src = X + (int64_t)rows[r] * d;
    for (int j = threadIdx.x; j < d; j += blockDim.x) out[r * d + j] = src[j];
}

__global__ void k_scatter_labels(const int* __restrict__ rows, int64_t count, const int64_t* __restrict__ lab,
                                 const double* __restrict__ res, int64_t* __restrict__ labels,
                                 double* __restrict__ residual) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= count) return;
    labels[rows[r]] = lab[r];
    residual[rows[r]] = res[r];
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
    TransientIndex g{centres};
    const uint32_t qflags = (flags & MLK_PREC_MASK) | MLK_SQUARED_DIST;
    std::lock_guard<std::mutex> lk(centres->lock);
    return knn_query_impl(centres, X_dev, n, 1, 0, qflags, residual_dev, labels_dev, (hipStream_t)stream);
}

extern "C" int mlk_kmeans_assign_reuse(const float* X_dev, int64_t n, int32_t d, const float* C_dev, int64_t kc,
                                       const int64_t* prev_dev, uint32_t flags, int64_t* labels_dev,
                                       double* residual_dev, int64_t* reused, void* stream_) {
    hipStream_t stream = (hipStream_t)stream_;
    if (reused) *reused = 0;
    if (n < 0 || d < 1 || kc < 1 || (n && (!X_dev || !labels_dev || !residual_dev || !prev_dev)) || !C_dev) {
        set_error("kmeans_assign_reuse: invalid shapes or pointers");
        return MLK_EINVAL;
    }
    if (kc >= (int64_t(1) << 31) || n >= (int64_t(1) << 31)) {
        set_error("kmeans_assign_reuse: rows and centres must fit 32-bit indices");
        return MLK_EINVAL;
    }
    if (flags & ~MLK_PREC_MASK) {
        set_error("kmeans_assign_reuse: only precision flags are accepted");
        return MLK_EINVAL;
    }
    if (n == 0) return MLK_OK;
    mlk_index* centres = nullptr;
    MLK_TRY(create_index(C_dev, kc, d, stream, &centres));
    TransientIndex g{centres};
    std::lock_guard<std::mutex> lk(centres->lock);
    const uint32_t qflags = (flags & MLK_PREC_MASK) | MLK_SQUARED_DIST;
    DevBuf sep2, sep_lo, near, unc, words, Xu, lab_u, res_u;
    MLK_TRY(sep_lo.reserve((size_t)kc * sizeof(double)));
    if (kc > 1) {
        // exact nearest other centre of every centre: the same exact search
        MLK_TRY(sep2.reserve((size_t)kc * sizeof(double)));
        MLK_TRY(near.reserve((size_t)kc * sizeof(int64_t)));
        MLK_TRY(knn_query_impl(centres, C_dev, kc, 1, 0, qflags | MLK_EXCLUDE_SELF, sep2.as<double>(),
                               near.as<int64_t>(), stream));
    }
    hipLaunchKernelGGL(k_sep_lower, dim3(nblk(kc)), dim3(256), 0, stream, kc > 1 ? sep2.as<double>() : nullptr, kc, d,
                       sep_lo.as<double>());
    MLK_TRY(words.reserve(64));
    MLK_CHECK_HIP(hipMemsetAsync(words.ptr, 0, 16, stream));
    // Gate (as the reference's sampled test, _k_means.py:309-324): with dense
    // centres the separation test rarely holds; then the plain assignment is
    // cheaper than a reuse pass that certifies few rows.
    constexpr int SAMPLES = 1024;
    if (n >= 4 * SAMPLES) {
        hipLaunchKernelGGL(k_reuse_sample, dim3(SAMPLES / 256), dim3(256), 0, stream, X_dev, n, d, C_dev, kc, prev_dev,
                           sep_lo.as<double>(), SAMPLES, words.as<unsigned>());
        MLK_CHECK_HIP(hipGetLastError());
        unsigned w3[3] = {0, 0, 0};
        MLK_CHECK_HIP(hipMemcpyAsync(w3, words.ptr, sizeof(w3), hipMemcpyDeviceToHost, stream));
        MLK_CHECK_HIP(hipStreamSynchronize(stream));
        if (w3[2] < SAMPLES * 3 / 4) {
            // the labels are still validated: the caller's contract does not
            // depend on the gate
            hipLaunchKernelGGL(k_prev_check, dim3(nblk(n)), dim3(256), 0, stream, prev_dev, n, kc, words.as<unsigned>());
            MLK_CHECK_HIP(hipGetLastError());
            unsigned w[2] = {0, 0};
            MLK_CHECK_HIP(hipMemcpyAsync(w, words.ptr, sizeof(w), hipMemcpyDeviceToHost, stream));
            MLK_CHECK_HIP(hipStreamSynchronize(stream));
            if (w[1]) {
                set_error("kmeans_assign_reuse: " + std::to_string(w[1]) + " previous labels outside [0, n_clusters)");
                return MLK_EINVAL;
            }
            MLK_TRY(knn_query_impl(centres, X_dev, n, 1, 0, qflags, residual_dev, labels_dev, stream));
            MLK_CHECK_HIP(hipStreamSynchronize(stream));
            return MLK_OK;
        }
    }
    MLK_TRY(unc.reserve((size_t)n * sizeof(int)));
    const bool vec = d % 4 == 0 && ((uintptr_t)X_dev % 16) == 0 && ((uintptr_t)C_dev % 16) == 0;
    hipLaunchKernelGGL(vec ? k_reuse<true> : k_reuse<false>, dim3(nblk(n)), dim3(256), 0, stream, X_dev, n, d, C_dev,
                       kc, prev_dev,
                       sep_lo.as<double>(), labels_dev, residual_dev, unc.as<int>(), words.as<unsigned>());
    MLK_CHECK_HIP(hipGetLastError());
    unsigned w[2] = {0, 0};
    MLK_CHECK_HIP(hipMemcpyAsync(w, words.ptr, sizeof(w), hipMemcpyDeviceToHost, stream));
    MLK_CHECK_HIP(hipStreamSynchronize(stream));
    if (w[1]) {
        set_error("kmeans_assign_reuse: " + std::to_string(w[1]) + " previous labels outside [0, n_clusters)");
        return MLK_EINVAL;
    }
    const int64_t nu = w[0];
    if (reused) *reused = n - nu;
    if (nu) {
        // the rows without a certificate: the full exact assignment
        MLK_TRY(Xu.reserve((size_t)nu * d * sizeof(float)));
        MLK_TRY(lab_u.reserve((size_t)nu * sizeof(int64_t)));
        MLK_TRY(res_u.reserve((size_t)nu * sizeof(double)));
        hipLaunchKernelGGL(k_gather_rows, dim3((unsigned)nu), dim3(d < 256 ? 64 : 256), 0, stream, X_dev, d,
                           unc.as<int>(), nu, Xu.as<float>());
        MLK_CHECK_HIP(hipGetLastError());
        MLK_TRY(knn_query_impl(centres, Xu.as<float>(), nu, 1, 0, qflags, res_u.as<double>(), lab_u.as<int64_t>(),
                               stream));
        hipLaunchKernelGGL(k_scatter_labels, dim3(nblk(nu)), dim3(256), 0, stream, unc.as<int>(), nu,
                           lab_u.as<int64_t>(), res_u.as<double>(), labels_dev, residual_dev);
        MLK_CHECK_HIP(hipGetLastError());
    }
    MLK_CHECK_HIP(hipStreamSynchronize(stream));
    return MLK_OK;
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
    hipLaunchKernelGGL(d <= 64 ? k_cluster_sums<1> : d <= 128 ? k_cluster_sums<2> : k_cluster_sums<4>,
                       dim3(nblk(kc, 4)), dim3(256), 0, stream, X_dev, d, rows2.as<int>(), off.as<unsigned long long>(),
                       kc, sums_dev, counts_dev);
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
