// Fused MFMA screen shared by the k-NN and radius paths: certified Gram
// bounds, the tile kernel, and the per-query threshold helpers.
//
// Reference counterpart: gram_lower + select_topk / compact_gram
// (src/mlxlearn_device/_neighbors_general.py:188-223, _neighbors_fused.py:421-506,
// _radius.py:128-185), which write the rows x N Gram tile to memory and
// re-read it for selection; here the tile stays in MFMA accumulators and only
// threshold hits leave the kernel.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "mlk_device.h"
#include "mlk_internal.h"

namespace mlk {

struct Consts {
    double H;       // relative error constant of the screen bound
    double G;       // bias constant (>= H + gamma)
    double gamma;   // accumulation constant
    double f0;      // absolute floor (flushed subnormal partial sums)
    double scale2;  // sigma^2 of the screen mirror
};

__device__ __forceinline__ float thr_from_tau(double tau, double nq, const Consts& c) {
    if (!(tau < INFINITY)) return -INFINITY;
    const double H = c.H + 0x1p-46;
    return __double2float_rd(((1.0 - H) * nq - tau * (1.0 + 0x1p-46)) * 0.5 - c.f0);
}

__device__ __forceinline__ double upper_bound(float acc, float bias, double nq, double nx, const Consts& c) {
    const double H = c.H + 0x1p-46;
    const double p = (double)acc + (double)bias;
    return (nq + nx) * (1.0 + H) - 2.0 * p + c.gamma * nx + 2.0 * c.f0 + 0x1p-46 * fabs(p);
}

__device__ __forceinline__ double lower_bound(float acc, float bias, double nq, double nx, const Consts& c) {
    const double H = c.H + 0x1p-46;
    const double p = (double)acc + (double)bias;
    return (nq + nx) * (1.0 - H) - 2.0 * p - c.gamma * nx - 2.0 * c.f0 - 0x1p-46 * fabs(p);
}

// k-th smallest (1-based) of non-negative float values spread over a wave:
// binary search on the bit pattern. Returns +inf bits if fewer than k values.
template <class GetFn>
__device__ uint32_t wave_kth_bits(int count, int k, GetFn get) {
    const int lane = threadIdx.x & 63;
    uint32_t ans = 0;
    for (int b = 31; b >= 0; --b) {
        const uint32_t trial = ans | (1u << b);
        int below = 0;
        for (int j = lane; j < count; j += 64) below += get(j) < trial;
        for (int o = 32; o > 0; o >>= 1) below += __shfl_xor(below, o);
        if (below < k) ans = trial;
    }
    return ans;  // largest pattern with fewer than k values strictly below it
}

// ---------------------------------------------------------------- screen
struct ScreenArgs {
    const void* xs;
    const float* bias;
    const void* qs;
    const float* thr;
    int* cnt;
    int* cidx;
    float* cacc;
    int64_t n;
    int64_t split_rows;
    int64_t off_lo, off_hi;
    int S, cap, QB;
    int64_t self_off;
    int exclude;
    int xcd_map;
    // radius mode: hits append (query, row) pairs to one global buffer
    int2* pairs;
    unsigned long long* pair_count;
    unsigned long long pair_cap;
};

constexpr int SCREEN_LISTS = 0;  // k-NN: per-(query, split) candidate lists
constexpr int SCREEN_PAIRS = 1;  // radius: global (query, row) pair buffer
constexpr int PAIR_STAGE = 512;  // per-wave LDS staging of radius pairs

template <int PREC> struct PT;
template <> struct PT<PREC_FP32> { typedef float T; };
template <> struct PT<PREC_BF16> { typedef __bf16 T; typedef bf16x8 V; };
template <> struct PT<PREC_FP16> { typedef _Float16 T; typedef f16x8 V; };

template <int PREC>
__device__ __forceinline__ f32x16 mfma16(const i32x4& a, const i32x4& b, const f32x16& c) {
    if constexpr (PREC == PREC_BF16)
        return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c, 0, 0, 0);
    else
        return __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(f16x8, a), __builtin_bit_cast(f16x8, b), c, 0, 0, 0);
}

// Work item (query block, split) of a block; splits sharing database rows are
// kept on one XCD (blocks are dispatched round-robin over the 8 XCDs).
__device__ __forceinline__ void block_item(const ScreenArgs& a, int& qb, int& split) {
    const int b = blockIdx.x;
    if (a.xcd_map) {
        const int xcd = b & 7, slot = b >> 3;
        qb = slot % a.QB;
        split = (slot / a.QB) * 8 + xcd;
    } else {
        qb = b % a.QB;
        split = b / a.QB;
    }
}

// Narrow screen (padded D <= 128): each wave keeps QT x 32 queries' operand
// fragments in registers for the whole database pass; 32*RT-row database
// tiles stream through LDS (double buffered) and are shared by the 4 waves.
template <int PREC, int DP, int QT, int RT, int MODE>
__global__ __launch_bounds__(256) void k_screen(const ScreenArgs a) {
    typedef typename PT<PREC>::T T;
    constexpr int EB = sizeof(T);
    constexpr int BN = RT * 32;
    constexpr int RS = DP * EB + 16;          // padded LDS row stride (bytes)
    constexpr int CPR = DP * EB / 16;         // 16-byte chunks per row
    constexpr int NCH = BN * CPR / 256;       // chunks per thread per tile
    constexpr int BQ = 4 * QT * 32;
    static_assert(NCH * 256 == BN * CPR, "tile must split evenly over the block");
    __shared__ __attribute__((aligned(16))) char sA[2][BN * RS];
    __shared__ float sBias[2][BN];
    __shared__ int2 sPair[MODE == SCREEN_PAIRS ? 4 : 1][MODE == SCREEN_PAIRS ? PAIR_STAGE : 1];
    __shared__ int sPairN[4];

    int qb, split;
    block_item(a, qb, split);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, h = lane >> 5, col = lane & 31;
    const int64_t s_lo = (int64_t)split * a.split_rows;
    const int64_t r_begin = s_lo + a.off_lo;
    int64_t r_end = s_lo + (a.off_hi < a.split_rows ? a.off_hi : a.split_rows);
    if (r_end > a.n) r_end = a.n;
    if (r_begin >= r_end) return;

    // k-NN mode: each lane owns the candidate sub-list (query col, split, lane
    // half h), so appends need no atomics; counts live in registers.
    if constexpr (MODE == SCREEN_PAIRS) {
        if (threadIdx.x < 4) sPairN[threadIdx.x] = 0;
    }
    // Radius mode: copy this wave's staged pairs to the global buffer.
    auto flush_pairs = [&]() {
        const int nbuf = sPairN[wave];
        unsigned long long base = 0;
        if (lane == 0 && nbuf) base = atomicAdd(a.pair_count, (unsigned long long)nbuf);
        base = __shfl(base, 0);
        for (int j = lane; j < nbuf; j += 64)
            if (base + j < a.pair_cap) a.pairs[base + j] = sPair[wave][j];
        __builtin_amdgcn_wave_barrier();
        if (lane == 0) sPairN[wave] = 0;
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
    };

    // Query operand fragments: lane (col, h) holds features h*DP/2 + [8j, 8j+8)
    // (float32: h*DP/2 + j) of query col of each tile.
    const int64_t qrow0 = (int64_t)qb * BQ + wave * QT * 32;
    float thr[QT];
#pragma unroll
    for (int qt = 0; qt < QT; ++qt) thr[qt] = a.thr[qrow0 + qt * 32 + col];
    int cnt_r[QT];
    if constexpr (MODE == SCREEN_LISTS) {
#pragma unroll
        for (int qt = 0; qt < QT; ++qt)
            cnt_r[qt] = a.cnt[(((size_t)(qrow0 + qt * 32 + col)) * a.S + split) * 2 + h];
    }
    bool live = false;
#pragma unroll
    for (int qt = 0; qt < QT; ++qt) live |= thr[qt] < INFINITY;
    const bool wave_live = ballot(live) != 0;

    constexpr int NB = PREC == PREC_FP32 ? DP / 2 : DP / 16;
    typedef typename std::conditional<PREC == PREC_FP32, float, i32x4>::type BFrag;
    BFrag bq[QT][NB];
#pragma unroll
    for (int qt = 0; qt < QT; ++qt) {
        const T* qp = (const T*)a.qs + (size_t)(qrow0 + qt * 32 + col) * DP + h * (DP / 2);
        if constexpr (PREC == PREC_FP32) {
#pragma unroll
            for (int j = 0; j < NB; j += 4) {
                f32x4 v = *(const f32x4*)(qp + j);
                bq[qt][j] = v[0]; bq[qt][j + 1] = v[1]; bq[qt][j + 2] = v[2]; bq[qt][j + 3] = v[3];
            }
        } else {
#pragma unroll
            for (int j = 0; j < NB; ++j) bq[qt][j] = *(const i32x4*)(qp + 8 * j);
        }
    }

    const int64_t ntiles = (r_end - r_begin + BN - 1) / BN;
    i32x4 pf[NCH];
    float pb = 0.0f;
    auto fetch = [&](int64_t t) {
        const int64_t r0 = r_begin + t * BN;
#pragma unroll
        for (int c = 0; c < NCH; ++c) {
            const int ch = threadIdx.x + c * 256;
            const int64_t row = r0 + ch / CPR;
            if (row < r_end)
                pf[c] = *(const i32x4*)((const char*)a.xs + (size_t)row * DP * EB + (ch % CPR) * 16);
            else
                pf[c] = i32x4{0, 0, 0, 0};
        }
        if (threadIdx.x < BN) {
            const int64_t row = r0 + threadIdx.x;
            pb = row < r_end ? a.bias[row] : 3.0e38f;
        }
    };
    auto stash = [&](int buf) {
#pragma unroll
        for (int c = 0; c < NCH; ++c) {
            const int ch = threadIdx.x + c * 256;
            *(i32x4*)(sA[buf] + (ch / CPR) * RS + (ch % CPR) * 16) = pf[c];
        }
        if (threadIdx.x < BN) sBias[buf][threadIdx.x] = pb;
    };

    fetch(0);
    stash(0);
    __syncthreads();
    for (int64_t t = 0; t < ntiles; ++t) {
        const int buf = (int)(t & 1);
        if (t + 1 < ntiles) fetch(t + 1);
        if (wave_live) {
            const int64_t r0 = r_begin + t * BN;
#pragma unroll
            for (int rt = 0; rt < RT; ++rt) {
                f32x16 init;
#pragma unroll
                for (int i = 0; i < 16; ++i) init[i] = -sBias[buf][rt * 32 + mfma32_row(i, lane)];
                f32x16 acc[QT];
#pragma unroll
                for (int qt = 0; qt < QT; ++qt) acc[qt] = init;
                const char* arow = sA[buf] + (rt * 32 + col) * RS + h * (DP / 2) * EB;
                if constexpr (PREC == PREC_FP32) {
#pragma unroll
                    for (int j = 0; j < NB; j += 4) {
                        f32x4 av = *(const f32x4*)(arow + j * 4);
#pragma unroll
                        for (int e = 0; e < 4; ++e)
#pragma unroll
                            for (int qt = 0; qt < QT; ++qt)
                                acc[qt] = __builtin_amdgcn_mfma_f32_32x32x2f32(av[e], bq[qt][j + e], acc[qt], 0, 0, 0);
                    }
                } else {
#pragma unroll
                    for (int j = 0; j < NB; ++j) {
                        i32x4 av = *(const i32x4*)(arow + j * 16);
#pragma unroll
                        for (int qt = 0; qt < QT; ++qt) acc[qt] = mfma16<PREC>(av, bq[qt][j], acc[qt]);
                    }
                }
                // Epilogue: a hit is acc >= thr (i.e. L(x) <= tau); hits are rare.
#pragma unroll
                for (int qt = 0; qt < QT; ++qt) {
                    bool any = false;
#pragma unroll
                    for (int i = 0; i < 16; ++i) any |= acc[qt][i] >= thr[qt];
                    if (ballot(any)) {
                        const int64_t qg = qrow0 + qt * 32 + col;
                        const int64_t self = a.exclude ? a.self_off + qg : -1;
#pragma unroll
                        for (int i = 0; i < 16; ++i) {
                            const int64_t row = r0 + rt * 32 + mfma32_row(i, lane);
                            const bool hit = acc[qt][i] >= thr[qt] && row < r_end && row != self;
                            if constexpr (MODE == SCREEN_LISTS) {
                                if (hit) {
                                    const int slot = cnt_r[qt]++;
                                    if (slot < a.cap) {
                                        const size_t o = (((size_t)qg * a.S + split) * 2 + h) * a.cap + slot;
                                        a.cidx[o] = (int)row;
                                        a.cacc[o] = acc[qt][i];
                                    }
                                }
                            } else {
                                if (ballot(hit)) {
                                    if (sPairN[wave] + 64 > PAIR_STAGE) flush_pairs();
                                    if (hit) {
                                        const int slot = atomicAdd(&sPairN[wave], 1);
                                        sPair[wave][slot] = int2{(int)qg, (int)row};
                                    }
                                    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
                                    __builtin_amdgcn_wave_barrier();
                                }
                            }
                        }
                    }
                }
            }
        }
        if (t + 1 < ntiles) stash(buf ^ 1);
        __syncthreads();
    }
    if constexpr (MODE == SCREEN_LISTS) {
#pragma unroll
        for (int qt = 0; qt < QT; ++qt)
            a.cnt[(((size_t)(qrow0 + qt * 32 + col)) * a.S + split) * 2 + h] = cnt_r[qt];
    } else {
        flush_pairs();
    }
}

}  // namespace mlk
