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

// Certified bounds of S (scaled units) from a stored hit: p = acc + a_x and the
// row norm, both rounded to float32 (their rounding is inside the extra
// 2^-22 terms), and the float64 query norm nq.
__device__ __forceinline__ void hit_bounds(float p32, float nx32, double nq, const Consts& c, double& lo, double& hi) {
    const double H = c.H + 0x1p-22;
    const double p = (double)p32;
    const double nx = (double)nx32;
    const double sum = nq + nx;
    const double pad = c.gamma * nx + 2.0 * c.f0 + 0x1p-22 * fabs(p);
    lo = sum * (1.0 - H) - 2.0 * p - pad;
    hi = sum * (1.0 + H) - 2.0 * p + pad;
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
    const void* xs;      // database screen mirror, n x DP
    const float* bias;   // per row: -a_x (accumulator start value)
    const float* nx32;   // per row: squared norm rounded to float32
    const void* qs;      // query screen mirror, m_pad x DP
    const float* thr;    // per query: hit threshold on the accumulator
    int* cnt;            // per (query, split, lane half): appended count
    int* cidx;           // candidate rows
    float* cP;           // acc + a_x (the screen's dot product estimate), float32
    float* cN;           // the row's squared norm, float32
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

typedef __attribute__((address_space(3))) char lds_char;

// Raw buffer resource over [base, base + bytes): loads past `bytes` return zero
// (buffer range check), so tile loads need no clamping.
#if defined(__HIP_DEVICE_COMPILE__)
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buffer_rsrc(const void* base, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)bytes, 0x00020000);
}
#endif

__device__ __forceinline__ float max16(const f32x16& v) {
    float a = fmaxf(fmaxf(v[0], v[1]), v[2]);
    float b = fmaxf(fmaxf(v[3], v[4]), v[5]);
    float c = fmaxf(fmaxf(v[6], v[7]), v[8]);
    float d = fmaxf(fmaxf(v[9], v[10]), v[11]);
    float e = fmaxf(fmaxf(v[12], v[13]), v[14]);
    return fmaxf(fmaxf(fmaxf(a, b), fmaxf(c, d)), fmaxf(e, v[15]));
}

// Narrow screen (padded D <= 128). A block of 4*RG waves walks 64-row tiles
// of its split for BQ = 4*QT*32 queries: wave (row group rg, query group
// qg) keeps QT x 32 queries' operand fragments in registers and covers 64/RG
// tile rows, so every A-operand byte read from LDS feeds QT MFMAs. 64-row
// database tiles stream from HBM straight into LDS (global_load_lds, double
// buffered, no staging registers; rows padded by one 16-byte chunk so the 32
// rows of an operand read hit distinct banks). Per tile a wave issues all its MFMAs, then
// one branch-free max-tree test of every accumulator against its query's
// threshold; only a hit enters the append path, which stores the row, the
// dot-product estimate and the row norm (float32) for the certified bounds.
template <int PREC, int DP, int QT, int RG, int MODE>
__global__ __launch_bounds__(256 * RG) void k_screen(const ScreenArgs a) {
    // the body uses device-only types (buffer resources): compile it in the
    // device pass only, so the host pass still emits the kernel's launch stub
#if defined(__HIP_DEVICE_COMPILE__)
    typedef typename PT<PREC>::T T;
    constexpr int NW = 4 * RG;                // waves per block
    constexpr int RTW = 2 / RG;               // 32-row tiles per wave
    constexpr int EB = sizeof(T);
    constexpr int BN = 64;                    // rows per tile
    constexpr int RB = DP * EB;               // bytes per database row
    constexpr int CPR = RB / 16;              // 16-byte chunks per row
    constexpr int RS = RB + 16;               // LDS row stride: one pad chunk per row
    constexpr int NLOAD = CPR + 1;            // 1 KiB DMA instructions per 64-row tile
    constexpr int LPW = (NLOAD + NW - 1) / NW;
    constexpr int BQ = 4 * QT * 32;
    static_assert(RG == 1 || RG == 2, "row groups");
    // two distinct buffers (not one 2-D array) so the compiler can tell the
    // tile being read from the tile being loaded and keep the DMA in flight
    __shared__ __attribute__((aligned(1024))) char sA0[BN * RS];
    __shared__ __attribute__((aligned(1024))) char sA1[BN * RS];
    __shared__ __attribute__((aligned(256))) float sB0[2 * BN];  // [0, BN): -a_x, [BN, 2BN): nx32
    __shared__ __attribute__((aligned(256))) float sB1[2 * BN];
    __shared__ int2 sPair[MODE == SCREEN_PAIRS ? NW : 1][MODE == SCREEN_PAIRS ? PAIR_STAGE : 1];
    __shared__ int sPairN[NW];

    int qb, split;
    block_item(a, qb, split);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, h = lane >> 5, col = lane & 31;
    const int qg_i = wave & 3, rg = wave >> 2;
    const int64_t s_lo = (int64_t)split * a.split_rows;
    const int64_t r_begin = s_lo + a.off_lo;
    int64_t r_end = s_lo + (a.off_hi < a.split_rows ? a.off_hi : a.split_rows);
    if (r_end > a.n) r_end = a.n;
    if (r_begin >= r_end) return;

    if constexpr (MODE == SCREEN_PAIRS) {
        if (threadIdx.x < NW) sPairN[threadIdx.x] = 0;
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
    const int64_t qrow0 = (int64_t)qb * BQ + qg_i * QT * 32;
    float thr[QT];
#pragma unroll
    for (int qt = 0; qt < QT; ++qt) thr[qt] = a.thr[qrow0 + qt * 32 + col];
    // k-NN mode: each lane owns the candidate sub-list (query col, split, lane
    // half h, row group rg), so appends need no atomics; counts live in registers.
    const int SL = 2 * RG;  // sub-lists per (query, split)
    int cnt_r[QT];
    size_t list_base[QT];
    if constexpr (MODE == SCREEN_LISTS) {
#pragma unroll
        for (int qt = 0; qt < QT; ++qt) {
            const size_t sub = (((size_t)(qrow0 + qt * 32 + col)) * a.S + split) * SL + rg * 2 + h;
            cnt_r[qt] = a.cnt[sub];
            list_base[qt] = sub * a.cap;
        }
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
    // DMA plan: LDS chunk L (of 64 (CPR + 1) per tile) is row L / (CPR + 1),
    // position L % (CPR + 1); the last position of each row is padding (loaded
    // out of range, i.e. zero). Source offsets within a tile are lane constants;
    // the tile base goes into the buffer resource (SGPRs).
    uint32_t voff[LPW];
#pragma unroll
    for (int i = 0; i < LPW; ++i) {
        const int L = (wave + i * NW) * 64 + lane;
        const int row = L / (CPR + 1), pos = L % (CPR + 1);
        voff[i] = pos < CPR ? (uint32_t)(row * RB + pos * 16) : 0x7ffffff0u;
    }
    auto load_tile = [&](int64_t t, char* dstA, float* dstB) {
        const int64_t r0 = r_begin + t * BN;
        const int64_t rows = r_end - r0 < BN ? r_end - r0 : BN;
        const __amdgpu_buffer_rsrc_t ra = buffer_rsrc((const char*)a.xs + (size_t)r0 * RB, (uint32_t)(rows * RB));
#pragma unroll
        for (int i = 0; i < LPW; ++i) {
            const int g = wave + i * NW;
            if (g < NLOAD)
                __builtin_amdgcn_raw_ptr_buffer_load_lds(ra, (lds_char*)(dstA + g * 1024), 16, voff[i], 0, 0, 0);
        }
        if (wave < 2) {
            const float* src = wave == 0 ? a.bias + r0 : a.nx32 + r0;
            const __amdgpu_buffer_rsrc_t rb = buffer_rsrc(src, (uint32_t)(rows * 4));
            __builtin_amdgcn_raw_ptr_buffer_load_lds(rb, (lds_char*)(dstB + wave * BN), 4, lane * 4, 0, 0, 0);
        }
    };

    // one tile: prefetch tile t+1 into (nA, nB) while computing on (cA, cB)
    auto step = [&](int64_t t, const char* cA, const float* cB, char* nA, float* nB) {
        if (t + 1 < ntiles) load_tile(t + 1, nA, nB);
        if (wave_live) {
            const int64_t r0 = r_begin + t * BN;
            f32x16 acc[RTW][QT];
#pragma unroll
            for (int rt = 0; rt < RTW; ++rt) {
                // accumulator start = -a_x of this lane's 16 rows (4 runs of 4 rows)
                const float* nb = cB + (rg * RTW + rt) * 32 + (h << 2);
                const f32x4 b0 = *(const f32x4*)(nb), b1 = *(const f32x4*)(nb + 8);
                const f32x4 b2 = *(const f32x4*)(nb + 16), b3 = *(const f32x4*)(nb + 24);
                const f32x16 init = {b0[0], b0[1], b0[2], b0[3], b1[0], b1[1], b1[2], b1[3],
                                     b2[0], b2[1], b2[2], b2[3], b3[0], b3[1], b3[2], b3[3]};
#pragma unroll
                for (int qt = 0; qt < QT; ++qt) acc[rt][qt] = init;
            }
#pragma unroll
            for (int rt = 0; rt < RTW; ++rt) {
                const int r = (rg * RTW + rt) * 32 + col;
                const char* arow = cA + r * RS + h * (CPR / 2) * 16;
                if constexpr (PREC == PREC_FP32) {
#pragma unroll
                    for (int j = 0; j < NB; j += 4) {
                        const f32x4 av = *(const f32x4*)(arow + (j / 4) * 16);
#pragma unroll
                        for (int e = 0; e < 4; ++e)
#pragma unroll
                            for (int qt = 0; qt < QT; ++qt)
                                acc[rt][qt] = __builtin_amdgcn_mfma_f32_32x32x2f32(av[e], bq[qt][j + e], acc[rt][qt], 0, 0, 0);
                    }
                } else {
#pragma unroll
                    for (int j = 0; j < NB; ++j) {
                        const i32x4 av = *(const i32x4*)(arow + j * 16);
#pragma unroll
                        for (int qt = 0; qt < QT; ++qt) acc[rt][qt] = mfma16<PREC>(av, bq[qt][j], acc[rt][qt]);
                    }
                }
            }
            // Hit test: acc >= thr  <=>  L(x) <= tau. Branch-free over the tile.
            bool hit_rq[RTW][QT];
            bool any = false;
#pragma unroll
            for (int rt = 0; rt < RTW; ++rt)
#pragma unroll
                for (int qt = 0; qt < QT; ++qt) {
                    hit_rq[rt][qt] = max16(acc[rt][qt]) >= thr[qt];
                    any |= hit_rq[rt][qt];
                }
            if (ballot(any)) {
#pragma unroll
                for (int rt = 0; rt < RTW; ++rt)
#pragma unroll
                for (int qt = 0; qt < QT; ++qt) {
                    if (!ballot(hit_rq[rt][qt])) continue;
                    const int64_t qg = qrow0 + qt * 32 + col;
                    const int64_t self = a.exclude ? a.self_off + qg : -1;
                    const int rbase = (rg * RTW + rt) * 32;
                    // rolled loop: the register index i is uniform (v_movrels)
#pragma unroll 1
                    for (int i = 0; i < 16; ++i) {
                        const float v = acc[rt][qt][i];
                        const int rl = rbase + mfma32_row(i, lane);
                        const int64_t row = r0 + rl;
                        const bool hit = v >= thr[qt] && row < r_end && row != self;
                        if constexpr (MODE == SCREEN_LISTS) {
                            if (hit) {
                                const int slot = cnt_r[qt]++;
                                if (slot < a.cap) {
                                    const size_t o = list_base[qt] + slot;
                                    a.cidx[o] = (int)row;
                                    a.cP[o] = v - cB[rl];
                                    a.cN[o] = cB[BN + rl];
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
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
    };

    load_tile(0, sA0, sB0);
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    for (int64_t t = 0; t < ntiles; t += 2) {
        step(t, sA0, sB0, sA1, sB1);
        if (t + 1 < ntiles) step(t + 1, sA1, sB1, sA0, sB0);
    }
    if constexpr (MODE == SCREEN_LISTS) {
#pragma unroll
        for (int qt = 0; qt < QT; ++qt)
            a.cnt[(((size_t)(qrow0 + qt * 32 + col)) * a.S + split) * SL + rg * 2 + h] = cnt_r[qt];
    } else {
        flush_pairs();
    }
#else
    (void)a;
#endif
}

}  // namespace mlk
