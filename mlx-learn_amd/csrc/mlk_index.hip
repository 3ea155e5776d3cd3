// Index construction: owned float32 copy, centre, per-precision screen
// mirrors with float64 norms, safe-range flags and Gram biases.
//
// Reference counterpart: DeviceNeighborIndex preparation
// (src/mlxlearn_device/_neighbors.py:232-239 _prepare_gram_arrays: anchored
// mean centre, centred train, float32 norms). Here the centre is rounded to an
// integer for integer data (so centred integer data stay exact in the screen
// types), norms are float64, and each row carries a safe-range flag instead of
// the reference's "unsafe -> lower bound 0" rule (_neighbors_general.py:188-223).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "mlk_device.h"
#include "mlk_internal.h"
#include "mlk_screen.h"

namespace mlk {

int DevBuf::reserve(size_t need) {
    if (need <= bytes && ptr) return MLK_OK;
    // a buffer that has to grow gets an eighth more than asked, so calls of
    // slightly different sizes do not free and allocate it again each time
    const bool regrow = ptr != nullptr;
    release();
    size_t want = need < 256 ? 256 : need;
    if (regrow) want += want / 8;
    hipError_t e = hipMalloc(&ptr, want);
    if (e == hipErrorOutOfMemory && workspace_release_pool()) {
        // the pooled workspace of transient indexes held the memory
        (void)hipGetLastError();
        e = hipMalloc(&ptr, want);
    }
    if (e != hipSuccess) {
        ptr = nullptr;
        bytes = 0;
        (void)hipGetLastError();
        set_error(std::string("device allocation of ") + std::to_string(want) + " bytes failed: " +
                  hipGetErrorString(e));
        return MLK_ENOMEM;
    }
    bytes = want;
    return MLK_OK;
}

void DevBuf::release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
}

// ------------------------------------------------ pooled workspace
constexpr int POOL_DEVICES = 64;
constexpr int WS_SLOTS = sizeof(mlk_index::ws) / sizeof(DevBuf);
static std::mutex g_pool_lock;
// never destroyed: freeing device memory from a static destructor could run
// after the HIP runtime has shut down
static DevBuf (&g_pool)[POOL_DEVICES][WS_SLOTS] = *new DevBuf[1][POOL_DEVICES][WS_SLOTS];

static void swap_buf(DevBuf& a, DevBuf& b) {
    std::swap(a.ptr, b.ptr);
    std::swap(a.bytes, b.bytes);
}

void workspace_adopt(mlk_index* index) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= POOL_DEVICES) return;
    std::lock_guard<std::mutex> g(g_pool_lock);
    for (int i = 0; i < WS_SLOTS; ++i)
        if (!index->ws[i].ptr) swap_buf(index->ws[i], g_pool[dev][i]);
}

// The caller has synchronised the device: no queued work uses the buffers.
void workspace_return(mlk_index* index) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= POOL_DEVICES) return;
    std::lock_guard<std::mutex> g(g_pool_lock);
    for (int i = 0; i < WS_SLOTS; ++i) {
        // keep the larger of the two buffers of each slot
        if (index->ws[i].bytes > g_pool[dev][i].bytes) swap_buf(index->ws[i], g_pool[dev][i]);
    }
}

bool workspace_release_pool() {
    std::lock_guard<std::mutex> g(g_pool_lock);
    bool any = false;
    for (auto& dev : g_pool)
        for (DevBuf& b : dev) {
            any |= b.ptr != nullptr;
            b.release();
        }
    return any;
}

// Per-block float64 column sums (rows [b*rows_per_block, ...)) and an
// "is every value an integer" flag.
__global__ void k_colsum(const float* __restrict__ X, int64_t n, int d, int64_t rows_per_block,
                         double* __restrict__ partial, unsigned* __restrict__ nonint) {
    const int64_t r0 = (int64_t)blockIdx.x * rows_per_block;
    const int64_t r1 = r0 + rows_per_block < n ? r0 + rows_per_block : n;
    bool frac = false;
    for (int c = threadIdx.x; c < d; c += blockDim.x) {
        double s = 0.0;
#pragma unroll 8
        for (int64_t r = r0; r < r1; ++r) {
            float v = X[r * d + c];
            s += (double)v;
            frac |= rintf(v) != v;
        }
        partial[(size_t)blockIdx.x * d + c] = s;
    }
    if (frac) atomicOr(nonint, 1u);
}

// max |x - c| over the matrix (non-negative float bits order like integers).
__global__ void k_maxabs(const float* __restrict__ X, int64_t total, int d, const float* __restrict__ c,
                         unsigned* __restrict__ out) {
    float m = 0.0f;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * blockDim.x) {
        double v = fabs((double)X[i] - (double)c[i % d]);
        m = fmaxf(m, (float)v);
    }
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
    if ((threadIdx.x & 63) == 0) atomicMax(out, __float_as_uint(m));
}

template <int PREC> struct ScreenType;
template <> struct ScreenType<PREC_FP32> { typedef float T; };
template <> struct ScreenType<PREC_BF16> { typedef __bf16 T; };
template <> struct ScreenType<PREC_FP16> { typedef _Float16 T; };

// Safe range of the scaled centred value for each screen type: zero, or a
// magnitude whose rounding stays normal and whose products with any other safe
// value are normal float32 numbers, up to `hi` (safe_limit).
template <int PREC>
__device__ __forceinline__ bool safe_value(double v, double hi) {
    double a = fabs(v);
    if (a == 0.0) return true;
    if (PREC == PREC_FP16) return a >= 0x1p-14 && a <= 0x1p15;
    return a >= 0x1p-60 && a <= hi;
}

// Largest safe magnitude for rows of d features: 2^56 (products of two safe
// values stay below 2^112), lowered for very wide rows so that d hi^2 <= 2^124:
// a safe row's dot products and norms then stay below 2^124 and its screen
// accumulator below 2^125, under the screen's lowest threshold (THR_FLOOR).
static double safe_limit(int prec, int d) {
    if (prec == PREC_FP16) return 0x1p15;
    int e = 0;
    while ((int64_t(1) << e) < d) ++e;  // 2^e >= d
    return std::ldexp(1.0, std::min(56, (124 - e) / 2));
}

template <int PREC>
__device__ __forceinline__ typename ScreenType<PREC>::T to_screen(double v) {
    // float64 -> float32 (RN) -> screen type (RN); the double rounding is
    // covered by the screen's error constant.
    return (typename ScreenType<PREC>::T)(float)v;
}

// One wave per row: screen mirror, float64 norm, safe flag, exactness flag.
// Used for database rows (flags_out per row) and for query rows.
template <int PREC, int R, int T>
__global__ __launch_bounds__(256) void k_mirror(const float* __restrict__ X, int64_t n, int64_t n_pad, int d, int dp,
                                                const float* __restrict__ c, double scale, double safe_hi,
                                                typename ScreenType<PREC>::T* __restrict__ xs,
                                                double* __restrict__ norms, uint8_t* __restrict__ flags,
                                                unsigned* __restrict__ inexact, double* __restrict__ nt_out,
                                                double* __restrict__ d2_out) {
    // R lanes per row (64 / R rows per wave); lane r holds features r + R t (t < T,
    // T R >= dp) as float64 between the two passes, so every row is read once
    typedef typename ScreenType<PREC>::T T_;
    const int lane = threadIdx.x & 63, r = lane % R;
    const int64_t row = ((int64_t)blockIdx.x * 256 + threadIdx.x) / R;
    if (row >= n_pad) return;  // whole row groups leave together
    const bool real = row < n;
    // pass 1: norm (float64, any order), safe range, exactness, and the squared
    // norms of the rounded row (st) and of its rounding residual (sd; each
    // difference x - x~ is exact); pass 2: the mirror values (zero for rows
    // outside the safe range, whose screen products must stay finite)
    double s = 0.0, st = 0.0, sd = 0.0;
    bool safe = true, exact = true;
    if constexpr (T == 0) {  // any width: two reads of the row, 64 lanes
        static_assert(R == 64, "looping mirror uses whole waves");
        for (int j = r; j < d && real; j += R) {
            const double x = ((double)X[row * d + j] - (double)c[j]) * scale;
            s += x * x;
            safe &= safe_value<PREC>(x, safe_hi);
            const double xt = (double)(float)to_screen<PREC>(x);
            exact &= xt == x;
            st += xt * xt;
            sd += (x - xt) * (x - xt);
        }
        for (int o = R / 2; o > 0; o >>= 1) {
            s += __shfl_xor(s, o, R);
            st += __shfl_xor(st, o, R);
            sd += __shfl_xor(sd, o, R);
            safe = __shfl_xor((int)safe, o, R) & safe;
            exact = __shfl_xor((int)exact, o, R) & exact;
        }
        for (int j = r; j < dp; j += R) {
            T_ out = (T_)0.0f;
            if (real && j < d && safe) out = to_screen<PREC>(((double)X[row * d + j] - (double)c[j]) * scale);
            xs[row * dp + j] = out;
        }
    } else {
    double v[T];
#pragma unroll
    for (int t = 0; t < T; ++t) {
        const int j = r + R * t;
        v[t] = 0.0;
        if (real && j < d) {
            v[t] = ((double)X[row * d + j] - (double)c[j]) * scale;
            s += v[t] * v[t];
            safe &= safe_value<PREC>(v[t], safe_hi);
            const double xt = (double)(float)to_screen<PREC>(v[t]);
            exact &= xt == v[t];
            st += xt * xt;
            sd += (v[t] - xt) * (v[t] - xt);
        }
    }
#pragma unroll
    for (int o = R / 2; o > 0; o >>= 1) {
        s += __shfl_xor(s, o, R);
        st += __shfl_xor(st, o, R);
        sd += __shfl_xor(sd, o, R);
        safe = __shfl_xor((int)safe, o, R) & safe;
        exact = __shfl_xor((int)exact, o, R) & exact;
    }
#pragma unroll
    for (int t = 0; t < T; ++t) {
        const int j = r + R * t;
        if (j < dp) xs[row * dp + j] = (real && j < d && safe) ? to_screen<PREC>(v[t]) : (T_)0.0f;
    }
    }
    if (r == 0) {
        norms[row] = real ? s : 0.0;
        if (nt_out) nt_out[row] = real ? st : 0.0;
        if (d2_out) d2_out[row] = real ? sd : 0.0;
        flags[row] = (real && safe) ? 1 : 0;
    }
    // one flag for the whole mirror (null: not wanted). Every check of it is an
    // access to one address, serialised in the L2, so one lane per wave checks
    // for the wave's rows, and the atomic is skipped once the flag is set.
    if (inexact) {
        const unsigned long long any = ballot(r == 0 && real && safe && !exact);
        if (any && lane == __builtin_ctzll(any) &&
            __hip_atomic_load(inexact, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0u)
            atomicOr(inexact, 1u);
    }
}

// Negated bias -a_x, a_x = screen_row_bound(...) rounded down (the screen's
// accumulator start value); unsafe rows get a start value no threshold reaches.
__global__ void k_bias(const double* __restrict__ nx, const double* __restrict__ nt, const double* __restrict__ d2,
                       const uint8_t* __restrict__ flags, int64_t n, Consts c, float* __restrict__ bias) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    bias[i] = flags[i] ? -__double2float_rd(screen_row_bound(nx[i], nt[i], d2[i], c)) : -3.0e38f;
}

__global__ void k_to_f32(const double* __restrict__ v, int64_t n, float* __restrict__ out) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (float)v[i];
}

__global__ void k_unsafe_list(const uint8_t* __restrict__ flags, int64_t n, int* __restrict__ list,
                              unsigned* __restrict__ count) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && !flags[i]) {
        unsigned at = atomicAdd(count, 1u);
        list[at] = (int)i;
    }
}

template <int PREC>
static void launch_mirror(const float* X, int64_t n, int64_t n_pad, int d, int dp, const float* c,
                          double scale, void* xs, double* norms, uint8_t* flags, unsigned* inexact, double* nt,
                          double* d2, hipStream_t stream) {
    typedef typename ScreenType<PREC>::T T_;
    const double hi = safe_limit(PREC, d);
    auto go = [&](auto kern, int lanes) {
        const int64_t rows_per_block = 256 / lanes;
        dim3 grid((unsigned)((n_pad + rows_per_block - 1) / rows_per_block));
        hipLaunchKernelGGL(kern, grid, dim3(256), 0, stream, X, n, n_pad, d, dp, c, scale, hi, (T_*)xs, norms, flags,
                           inexact, nt, d2);
    };
    // lanes per row: about four features per lane, at least 16 lanes
    if (dp <= 64) go(k_mirror<PREC, 16, 4>, 16);
    else if (dp <= 128) go(k_mirror<PREC, 32, 4>, 32);
    else if (dp <= 256) go(k_mirror<PREC, 64, 4>, 64);
    else if (dp <= 832) go(k_mirror<PREC, 64, 13>, 64);
    else go(k_mirror<PREC, 64, 0>, 64);  // wider: the row is read twice
}

int mirror_rows(int prec, const float* X, int64_t n, int64_t n_pad, int d, int dp, const float* c,
                double scale, void* xs, double* norms, uint8_t* flags, unsigned* inexact, double* nt, double* d2,
                hipStream_t stream) {
    if (n_pad == 0) return MLK_OK;
    switch (prec) {
        case PREC_FP32:
            launch_mirror<PREC_FP32>(X, n, n_pad, d, dp, c, scale, xs, norms, flags, inexact, nt, d2, stream);
            break;
        case PREC_BF16:
            launch_mirror<PREC_BF16>(X, n, n_pad, d, dp, c, scale, xs, norms, flags, inexact, nt, d2, stream);
            break;
        default: launch_mirror<PREC_FP16>(X, n, n_pad, d, dp, c, scale, xs, norms, flags, inexact, nt, d2, stream); break;
    }
    MLK_CHECK_HIP(hipGetLastError());
    return MLK_OK;
}



int prepare_mirror(mlk_index* index, int prec, hipStream_t stream) {
    Mirror& mr = index->mirror[prec];
    if (mr.ready) return MLK_OK;
    const int64_t n = index->n;
    const int d = index->d;
    mr.dp = screen_dp(prec, d);
    double scale = 1.0;
    DevBuf tmp;
    MLK_TRY(tmp.reserve(16));
    unsigned* dflag = tmp.as<unsigned>();
    if (prec == PREC_FP16) {
        MLK_CHECK_HIP(hipMemsetAsync(dflag, 0, 4, stream));
        hipLaunchKernelGGL(k_maxabs, dim3(1024), dim3(256), 0, stream, index->X.as<float>(), n * (int64_t)d, d,
                           index->centre.as<float>(), dflag);
        MLK_CHECK_HIP(hipGetLastError());
        unsigned bits = 0;
        MLK_CHECK_HIP(hipMemcpyAsync(&bits, dflag, 4, hipMemcpyDeviceToHost, stream));
        MLK_CHECK_HIP(hipStreamSynchronize(stream));
        float mx;
        std::memcpy(&mx, &bits, 4);
        if (mx > 0.0f && std::isfinite(mx)) {
            int e;
            std::frexp((double)mx, &e);  // mx < 2^e
            scale = std::ldexp(1.0, 14 - e);  // scaled max < 2^14
        }
    }
    mr.scale = scale;
    MLK_TRY(mr.xs.reserve((size_t)n * mr.dp * elem_bytes(prec)));
    MLK_TRY(mr.nx.reserve((size_t)n * sizeof(double)));
    MLK_TRY(mr.nt.reserve((size_t)n * sizeof(double)));
    MLK_TRY(mr.d2.reserve((size_t)n * sizeof(double)));
    MLK_TRY(mr.flags.reserve((size_t)n));
    MLK_CHECK_HIP(hipMemsetAsync(dflag, 0, 8, stream));
    MLK_TRY(mirror_rows(prec, index->X.as<float>(), n, n, d, mr.dp, index->centre.as<float>(), scale,
                        mr.xs.ptr, mr.nx.as<double>(), mr.flags.as<uint8_t>(), dflag, mr.nt.as<double>(),
                        mr.d2.as<double>(), stream));
    MLK_TRY(mr.nx32.reserve((size_t)n * sizeof(float)));
    hipLaunchKernelGGL(k_to_f32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, mr.nx.as<double>(), n,
                       mr.nx32.as<float>());
    MLK_TRY(mr.unsafe.reserve((size_t)n * sizeof(int)));
    hipLaunchKernelGGL(k_unsafe_list, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream,
                       mr.flags.as<uint8_t>(), n, mr.unsafe.as<int>(), dflag + 1);
    MLK_CHECK_HIP(hipGetLastError());
    unsigned host[2] = {0, 0};
    MLK_CHECK_HIP(hipMemcpyAsync(host, dflag, 8, hipMemcpyDeviceToHost, stream));
    MLK_CHECK_HIP(hipStreamSynchronize(stream));
    mr.exact = host[0] == 0;
    mr.n_unsafe = host[1];
    if (mr.n_unsafe) {
        // Deterministic order of the unsafe list (atomics appended in any order).
        std::vector<int> rows(mr.n_unsafe);
        MLK_CHECK_HIP(hipMemcpyAsync(rows.data(), mr.unsafe.ptr, rows.size() * sizeof(int), hipMemcpyDeviceToHost,
                                     stream));
        MLK_CHECK_HIP(hipStreamSynchronize(stream));
        std::sort(rows.begin(), rows.end());
        MLK_CHECK_HIP(hipMemcpyAsync(mr.unsafe.ptr, rows.data(), rows.size() * sizeof(int), hipMemcpyHostToDevice,
                                     stream));
        MLK_CHECK_HIP(hipStreamSynchronize(stream));  // `rows` leaves scope
    }
    mr.bias_t1 = mr.bias_t2 = -1.0;
    mr.ready = true;
    return MLK_OK;
}

int CallMem::get(int slot, size_t bytes, void** out) {
    if (!arena) {
        MLK_TRY(ws[slot].reserve(bytes));
        *out = ws[slot].ptr;
        return MLK_OK;
    }
    const size_t at = (used + 255) / 256 * 256;
    if (at > cap || bytes > cap - at) {
        set_error("the call needs " + std::to_string(at + bytes) + " bytes of its fixed scratch arena of " +
                  std::to_string(cap) + " bytes");
        return MLK_ECAPACITY;
    }
    *out = arena + at;
    used = at + bytes;
    return MLK_OK;
}

int prepare_bias(mlk_index* index, int prec, const Consts& c, hipStream_t stream, bool may_allocate) {
    Mirror& mr = index->mirror[prec];
    if (mr.bias_t1 == c.t1 && mr.bias_t2 == c.t2 && mr.bias_gamma == c.gamma && mr.bias_eps == c.eps) return MLK_OK;
    if (!may_allocate && mr.bias.bytes < (size_t)index->n * sizeof(float)) {
        set_error("the index has no screen bias for this precision yet: one mlk_knn_query call of the same "
                  "precision prepares it");
        return MLK_EINVAL;
    }
    MLK_TRY(mr.bias.reserve((size_t)index->n * sizeof(float)));
    hipLaunchKernelGGL(k_bias, dim3((unsigned)((index->n + 255) / 256)), dim3(256), 0, stream,
                       mr.nx.as<double>(), mr.nt.as<double>(), mr.d2.as<double>(), mr.flags.as<uint8_t>(), index->n, c,
                       mr.bias.as<float>());
    MLK_CHECK_HIP(hipGetLastError());
    mr.bias_t1 = c.t1;
    mr.bias_t2 = c.t2;
    mr.bias_gamma = c.gamma;
    mr.bias_eps = c.eps;
    return MLK_OK;
}

int create_index(const float* X, int64_t n, int d, hipStream_t stream, mlk_index** out) {
    mlk_index* index = new (std::nothrow) mlk_index();
    if (!index) {
        set_error("host allocation failed");
        return MLK_ENOMEM;
    }
    index->n = n;
    index->d = d;
    int rc = MLK_OK;
    do {
        if ((rc = index->X.reserve((size_t)n * d * sizeof(float)))) break;
        if (hipMemcpyAsync(index->X.ptr, X, (size_t)n * d * sizeof(float), hipMemcpyDeviceToDevice, stream) !=
            hipSuccess) {
            set_error("copy of the training matrix failed");
            rc = MLK_EDEVICE;
            break;
        }
        // Centre: float64 column means; integral data get an integer centre.
        // About 1,024 blocks of row sums (a k-means centroid index of 65,536
        // rows is built on every assignment: 16 blocks of 4,096 rows took 1.1 ms)
        const int64_t rows_per_block = std::max<int64_t>(64, (n + 1023) / 1024);
        const int64_t blocks = (n + rows_per_block - 1) / rows_per_block;
        DevBuf partial;
        if ((rc = partial.reserve((size_t)blocks * d * sizeof(double) + 16))) break;
        unsigned* nonint = (unsigned*)((char*)partial.ptr + (size_t)blocks * d * sizeof(double));
        if (hipMemsetAsync(nonint, 0, 4, stream) != hipSuccess) { rc = MLK_EDEVICE; break; }
        hipLaunchKernelGGL(k_colsum, dim3((unsigned)blocks), dim3(256), 0, stream, index->X.as<float>(), n, d,
                           rows_per_block, partial.as<double>(), nonint);
        if (hipGetLastError() != hipSuccess) { set_error("column sum launch failed"); rc = MLK_EDEVICE; break; }
        std::vector<double> host((size_t)blocks * d);
        unsigned host_nonint = 0;
        if (hipMemcpyAsync(host.data(), partial.ptr, host.size() * sizeof(double), hipMemcpyDeviceToHost, stream) !=
                hipSuccess ||
            hipMemcpyAsync(&host_nonint, nonint, 4, hipMemcpyDeviceToHost, stream) != hipSuccess ||
            hipStreamSynchronize(stream) != hipSuccess) {
            set_error("reading column sums failed");
            rc = MLK_EDEVICE;
            break;
        }
        index->integral = host_nonint == 0;
        std::vector<float> centre(d);
        for (int c = 0; c < d; ++c) {
            double s = 0.0;
            for (int64_t b = 0; b < blocks; ++b) s += host[(size_t)b * d + c];
            double mean = s / (double)n;
            centre[c] = index->integral ? (float)std::nearbyint(mean) : (float)mean;
            if (!std::isfinite(centre[c])) centre[c] = 0.0f;
        }
        if ((rc = index->centre.reserve((size_t)d * sizeof(float)))) break;
        if (hipMemcpy(index->centre.ptr, centre.data(), d * sizeof(float), hipMemcpyHostToDevice) != hipSuccess) {
            set_error("uploading the centre failed");
            rc = MLK_EDEVICE;
            break;
        }
    } while (0);
    if (rc) {
        delete index;
        return rc;
    }
    *out = index;
    return MLK_OK;
}

}  // namespace mlk
