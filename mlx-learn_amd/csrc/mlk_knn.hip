// Exact k-nearest-neighbour search on gfx950.
//
// Pipeline (one stream, rounds launched back to back):
//   qprep   queries -> screen mirror (bf16 / fp16 / fp32), float64 norms, flags
//   tau0    per query: k-th smallest certified upper bound over 64 sampled rows
//   screen  fused MFMA Gram tile + threshold filter; hits append (row, acc) to a
//           per-(query, split) candidate list; the N x M matrix never leaves
//           registers (replaces gram_lower + select_topk, reference
//           src/mlxlearn_device/_neighbors_general.py:188-329 and
//           _neighbors_fused.py:421-506)
//   merge   per query: k-th smallest certified upper bound over all
//           candidates -> tighter threshold for the next round
//   rank    per query: lower-bound filter, float64 distances, exact (S, index)
//           order with interval tests and exact integer fall back (replaces the
//           refine / candidate-order / certificate kernels,
//           _neighbors_fused.py:545-834 and _neighbors_precision.py)
//   fallback  queries whose candidate lists overflowed (or whose coordinates
//           leave the screen's safe range): float64 scan of every row, radix
//           select of the k-th value, collection of the interval band, exact
//           rank (replaces _neighbors_fallback.py partitioned_neighbors).
//
// Certificate. For a screen value acc (MFMA accumulation of the rounded
// centred coordinates, started from the bias -a_x), every row x satisfies
//   L(x) <= S(q, x) <= U(x),  L/U = (nq + nx)(1 -/+ H) - 2(acc + a_x) -/+ (g nx + 2 f0)
// with H covering operand rounding (2 eta + eta^2), accumulation in any order
// with any rounding (gamma_{K+1}, unit 2^-23), float64 norm errors and flushed
// subnormal partial sums (f0). A query keeps every row with L(x) <= tau, where
// tau is an upper bound of its k-th smallest S (the k-th smallest U seen so
// far), so the kept set contains every true neighbour; the rank stage orders
// the kept set exactly. Overflow of a candidate list withdraws the
// certificate and sends the query to the exact fallback.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "mlk_device.h"
#include "mlk_internal.h"
#include "mlk_order.h"
#include "mlk_screen.h"

namespace mlk {

int mirror_rows(int prec, const float* X, int64_t n, int64_t n_pad, int d, int dp, const float* c,
                double scale, void* xs, double* norms, uint8_t* flags, unsigned* inexact, double* nt, double* d2,
                hipStream_t stream);

// status bits per query
constexpr int ST_UNSAFE = 1;     // query coordinates outside the screen's safe range
constexpr int ST_OVERFLOW = 2;   // a candidate list overflowed
constexpr int ST_FORCED = 4;     // certificate withdrawn by a test hook
constexpr int ST_FALLBACK = 8;   // resolved by the exact fallback
constexpr int ST_TIE = 16;       // k = 1 group rank: a comparison needs the exact comparator (resolved by the lane rank)
constexpr int ST_BIG = 32;       // more than LCAP rows survive the bound filter (ranked again with LCAP_BIG)

constexpr int LCAP = 512;        // candidates ranked in LDS per query
constexpr int LCAP_BIG = 4096;   // the same for queries with more survivors (one wave per block)
constexpr int MCAP = 512;        // upper bounds held in LDS per query by the merge

// ---------------------------------------------------------------- tau0
constexpr int TAU0_SAMPLES = 256;

// float64 squared distance in four interleaved partial sums (any evaluation
// order is covered by the relative bound f64_interval(d)).
__device__ __forceinline__ double fast_sq(const float* __restrict__ q, const float* __restrict__ x, int d) {
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    int i = 0;
    for (; i + 4 <= d; i += 4) {
        const double t0 = (double)q[i] - (double)x[i], t1 = (double)q[i + 1] - (double)x[i + 1];
        const double t2 = (double)q[i + 2] - (double)x[i + 2], t3 = (double)q[i + 3] - (double)x[i + 3];
        s0 += t0 * t0; s1 += t1 * t1; s2 += t2 * t2; s3 += t3 * t3;
    }
    for (; i < d; ++i) {
        const double t = (double)q[i] - (double)x[i];
        s0 += t * t;
    }
    return (s0 + s1) + (s2 + s3);
}

// Start threshold: for each query, the k-th smallest certified upper bound of
// S over up to 256 evenly spaced rows is an upper bound of its k-th smallest
// S. The block stages the sample rows in LDS, 32 features at a time, with
// coalesced loads (32 lanes per 128-byte row segment), and each wave stages
// its QW queries' features beside them; features past d, samples past the
// sample count and padding queries are staged as zeros (they add exact zeros
// to the sums), so the arithmetic needs no guards. The global loads of the
// next 32 features are issued into registers before the current ones are
// used, so their latency overlaps the arithmetic. Lane j of each wave owns
// samples j + 64u (u < U) and takes their features into registers eight at a
// time while it walks the wave's QW queries (broadcast LDS reads), so each
// staged value serves 4 * QW queries. float32 differences and fused
// squares-and-adds, turned into a certified upper bound of S by the relative
// and absolute error terms below. Each wave also initialises its queries'
// candidate sub-list counters and k-best bounds (screened calls only).
template <int U, int QW>
__global__ __launch_bounds__(256) void k_tau0(const float* __restrict__ X, int64_t n, int d,
                                              const float* __restrict__ Q, int64_t m, int64_t m_pad, int k,
                                              int max_samples, int64_t self_off, int exclude,
                                              const double* __restrict__ qa, const uint8_t* __restrict__ qflags,
                                              Consts c, float* __restrict__ tau, float* __restrict__ thr,
                                              int* __restrict__ status, int S2, int* __restrict__ cnt,
                                              int* __restrict__ seen, float* __restrict__ kbest,
                                              int* __restrict__ kbest_row, const int* __restrict__ srows) {
    constexpr int FL = 32;      // features staged per pass
    constexpr int FC = 8;       // features per register chunk
    constexpr int NS = 64 * U;  // sample slots of the block
    constexpr int PER = NS * FL / 256;           // staged sample values per thread per pass
    constexpr int QS = (QW * FL + 63) / 64;      // staged query values per lane per pass
    __shared__ float sx[NS][FL + 1];
    __shared__ float sq[4][QW][FL];
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int64_t q0 = ((int64_t)blockIdx.x * 4 + w) * QW;
    // queries of this wave: padding included (nqw), real ones (live)
    const int nqw = q0 >= m_pad ? 0 : (m_pad - q0 < QW ? (int)(m_pad - q0) : QW);
    const int live = q0 >= m ? 0 : (m - q0 < nqw ? (int)(m - q0) : nqw);
    if (cnt) {
        for (int qi = 0; qi < nqw; ++qi) {
            const int64_t q = q0 + qi;
            for (int j = lane; j < S2; j += 64) {
                cnt[q * S2 + j] = 0;
                seen[q * S2 + j] = 0;
            }
            for (int j = lane; j < k; j += 64) {
                kbest[q * k + j] = INFINITY;
                if (kbest_row) kbest_row[q * k + j] = -1;
            }
        }
    }
    const int samples = n < max_samples ? (int)n : max_samples;  // <= NS
    // sample j is row j * stride, or srows[j] of this block's list (distinct rows)
    const int64_t stride = n / samples;
    const int* srow = srows ? srows + (int64_t)blockIdx.x * samples : nullptr;
    // staging registers: value i of this thread is sample j0 + 8 i, feature e of the
    // pass (j0 = threadIdx.x / 32, e = threadIdx.x % 32); query value i of this lane is
    // query qw0 + 2 i of the wave, feature e (qw0 = lane / 32)
    static_assert(256 % FL == 0 && 64 % FL == 0, "staging layout");
    float stg[PER], qstg[QS];
    const int j0 = (int)threadIdx.x / FL, e_st = (int)threadIdx.x % FL, qw0 = lane / FL;
    const int64_t row_step = (int64_t)(256 / FL) * stride * d;  // eight sample rows
    const int64_t x_first = (int64_t)j0 * stride * d + e_st;
    const int64_t q_first = (q0 + qw0) * d + e_st;
    // listed samples: the block's rows, read from memory once and kept in LDS
    // for every pass (-1: no sample)
    __shared__ int srr[NS];
    if (srow)
        for (int j = threadIdx.x; j < NS; j += 256) srr[j] = j < samples ? srow[j] : -1;
    if (srow) __syncthreads();
    auto fetch = [&](int f0) {
        // one running offset per thread, opaque to the compiler at each pass, so the
        // per-value addresses are formed at their loads instead of being kept live
        int64_t xo = x_first + f0, qo = q_first + f0;
        asm volatile("" : "+v"(xo), "+v"(qo));
        const bool fe = f0 + e_st < d;
        if (srow) {
#pragma unroll
            for (int i = 0; i < PER; ++i) {
                const int64_t r = srr[j0 + (256 / FL) * i];
                stg[i] = X[fe && r >= 0 ? r * d + f0 + e_st : 0];
            }
        } else {
#pragma unroll
            for (int i = 0; i < PER; ++i) {
                const bool ok = fe && j0 + (256 / FL) * i < samples;
                stg[i] = X[ok ? xo + i * row_step : 0];  // row 0 stands in for padding (zeroed at the store)
            }
        }
#pragma unroll
        for (int i = 0; i < QS; ++i) {
            const int qi = qw0 + (64 / FL) * i;
            const bool ok = fe && qi < live;
            qstg[i] = Q[ok ? qo + (int64_t)(64 / FL) * i * d : 0];
        }
    };
    auto store = [&](int f0) {
        const bool fe = f0 + e_st < d;
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            const int j = j0 + (256 / FL) * i;
            sx[j][e_st] = (fe && j < samples) ? stg[i] : 0.0f;
        }
#pragma unroll
        for (int i = 0; i < QS; ++i) {
            const int qi = qw0 + (64 / FL) * i;
            if (qi < QW) sq[w][qi][e_st] = (fe && qi < live) ? qstg[i] : 0.0f;
        }
    };
    float acc[QW][U];
#pragma unroll
    for (int qi = 0; qi < QW; ++qi)
#pragma unroll
        for (int u = 0; u < U; ++u) acc[qi][u] = 0.0f;
    fetch(0);
    for (int f0 = 0; f0 < d; f0 += FL) {
        __syncthreads();  // every read of the previous pass is done
        store(f0);
        __syncthreads();
        if (f0 + FL < d) fetch(f0 + FL);  // in flight during this pass
#pragma unroll 1
        for (int c0 = 0; c0 < FL; c0 += FC) {
            float sv[U][FC];
#pragma unroll
            for (int u = 0; u < U; ++u)
#pragma unroll
                for (int e = 0; e < FC; ++e) sv[u][e] = sx[lane + 64 * u][c0 + e];
#pragma unroll
            for (int qi = 0; qi < QW; ++qi) {
#pragma unroll
                for (int e = 0; e < FC; ++e) {
                    const float dq = sq[w][qi][c0 + e];
#pragma unroll
                    for (int u = 0; u < U; ++u) {
                        const float t = dq - sv[u][e];
                        acc[qi][u] = fmaf(t, t, acc[qi][u]);
                    }
                }
            }
        }
    }
    // float32 rounding: t = fl(q - x) = (q - x)(1 + delta) and a_i = fma(t, t, a_(i-1))
    // = (t^2 + a_(i-1))(1 + eps) with |delta|, |eps| <= 2^-24, and at most 2^-126
    // lost to a subnormal or flushed result per step, give
    // a_d >= (1 - 2^-24)^(d+2) S - d 2^-125, hence
    // S <= (a_d + (d + 1) 2^-125)(1 + (d + 4) 2^-23), evaluated in float64 (its own
    // rounding is inside the spare 2^-22) and rounded up; an overflow gives +inf
    const double pad = (double)(d + 1) * 0x1p-125;
    const double rel = 1.0 + (double)(d + 4) * 0x1p-23;
#pragma unroll
    for (int qi = 0; qi < QW; ++qi) {
        const int64_t q = q0 + qi;
        if (qi >= live) {  // padding query (or past m_pad)
            if (lane == 0 && qi < nqw) { tau[q] = -1.0f; thr[q] = INFINITY; status[q] = 0; }
            continue;
        }
        uint32_t bits[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int j = lane + 64 * u;
            const int64_t row = j < samples ? (srow ? (int64_t)srow[j] : (int64_t)j * stride) : -1;
            bits[u] = 0x7f800000u;
            if (row >= 0 && !(exclude && row == self_off + q))
                bits[u] = __float_as_uint(__double2float_ru(((double)acc[qi][u] + pad) * rel * c.scale2));
        }
        uint32_t kb;
        if (k == 1) {  // the minimum (non-negative float bit patterns order as integers)
            kb = bits[0];
#pragma unroll
            for (int u = 1; u < U; ++u) kb = kb < bits[u] ? kb : bits[u];
            for (int o = 32; o > 0; o >>= 1) {
                const uint32_t other = (uint32_t)__shfl_xor((int)kb, o);
                kb = kb < other ? kb : other;
            }
        } else {
            kb = wave_kth_bits(64 * U, k, [&](int j) { return bits[j >> 6]; });
        }
        if (lane == 0) {
            const float t0 = __uint_as_float(kb);
            const int st = qflags[q] ? 0 : ST_UNSAFE;
            status[q] = st;
            tau[q] = t0;
            thr[q] = st ? INFINITY : thr_from_tau((double)t0, qa[q], c);
        }
    }
}

// ---------------------------------------------------------------- merge
// Incremental merge after a round. Per query, kbest holds the k smallest
// certified upper bounds over every candidate merged so far; the entries
// appended since the previous merge (seen .. min(count, cap) of each sub-list)
// are folded in, the k-th smallest value of the union becomes the new tau
// (an upper bound of the query's k-th smallest S), and kbest is rewritten with
// the k smallest values of the union.
//
// With rows (kbest_row != nullptr, k <= 32) the merge also keeps the row of
// every entry: it takes the min(2k, 32) smallest bounds, replaces each by the
// smaller of it and its row's float64 distance raised by the order-free bound
// g (rounded up), and keeps the k smallest results as kbest. These are upper
// bounds of k distinct rows, so the k-th is a valid tau, now close to the
// exact k-th distance instead of a screen band above it, which shrinks the
// next rounds' candidate sets.
__global__ void k_merge(int64_t m, int k, const int* __restrict__ cnt, int* __restrict__ seen,
                        const int2* __restrict__ cand, const float* __restrict__ nx32, int S2, int cap,
                        float* __restrict__ kbest, const double* __restrict__ nq, const double* __restrict__ qa,
                        Consts c, float* __restrict__ tau, float* __restrict__ thr, int* __restrict__ status,
                        int* __restrict__ kbest_row,
                        const float* __restrict__ X, const float* __restrict__ Q, int d) {
    __shared__ uint32_t sU[4][MCAP];
    __shared__ int sR[4][MCAP];
    __shared__ float sKV[4][32];
    __shared__ int sKR[4][32];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int64_t q = (int64_t)blockIdx.x * 4 + w;
    if (q >= m) return;
    if (status[q]) return;
    // sub-list bounds [seen, min(count, cap)), two sub-lists per lane (s0 + lane,
    // s0 + lane + 64); the first 128 stay in registers for the append pass
    int fresh = 0;
    bool over = false;
    int lo0[2] = {0, 0}, hi0[2] = {0, 0};
    for (int s0 = 0; s0 < S2; s0 += 128) {
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            const int s = s0 + lane + 64 * t;
            if (s < S2) {
                const int cs = cnt[q * S2 + s];
                const int sn = seen[q * S2 + s];
                over |= cs > cap;
                const int hi = cs < cap ? cs : cap;
                fresh += hi - sn;
                if (s0 == 0) {
                    lo0[t] = sn;
                    hi0[t] = hi;
                }
            }
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        fresh += __shfl_xor(fresh, o);
        over = __shfl_xor((int)over, o) | over;
    }
    if (over) {
        if (lane == 0) { status[q] = ST_OVERFLOW; thr[q] = INFINITY; }
        return;
    }
    if (fresh == 0 || k + fresh > MCAP) return;  // tau stays valid (just not tighter)
    uint32_t* buf = sU[w];
    int* rbuf = sR[w];
    const bool rows = kbest_row != nullptr && k <= 32;
    for (int j = lane; j < k; j += 64) {
        buf[j] = __float_as_uint(kbest[q * k + j]);
        if (rows) rbuf[j] = kbest_row[q * k + j];
    }
    int base = k;
    const double nqq = nq[q];
    for (int s0 = 0; s0 < S2; s0 += 128) {
        int lo[2], hi[2], at[2];
        size_t off[2];
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            const int s = s0 + lane + 64 * t;
            lo[t] = hi[t] = 0;
            if (s0 == 0) {
                lo[t] = lo0[t];
                hi[t] = hi0[t];
            } else if (s < S2) {
                const int cs = cnt[q * S2 + s];
                hi[t] = cs < cap ? cs : cap;
                lo[t] = seen[q * S2 + s];
            }
            if (s < S2) seen[q * S2 + s] = hi[t];
            off[t] = (size_t)q * cap * S2 + s;  // entry j at off + j * S2
        }
        // positions in sub-list order: s0 .. s0 + 63 (t = 0), then s0 + 64 .. (t = 1)
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            const int mine = hi[t] - lo[t];
            int incl = mine;
            for (int o = 1; o < 64; o <<= 1) {
                const int v = __shfl_up(incl, o);
                if (lane >= o) incl += v;
            }
            at[t] = base + incl - mine;
            base += __shfl(incl, 63);
        }
        const int nmax = hi[0] - lo[0] > hi[1] - lo[1] ? hi[0] - lo[0] : hi[1] - lo[1];
        for (int j0 = 0; j0 < nmax; j0 += 4) {
            // four entries of each sub-list per batch: loads and norm gathers
            // issued together, from valid slots (clamped to the sub-list's last
            // new entry; a sub-list without new entries borrows the other's)
            int2 e[2][4];
            float nxv[2][4];
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                const int tt = hi[t] > lo[t] ? t : 1 - t;
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int j = lo[tt] + j0 + u < hi[tt] ? lo[tt] + j0 + u : hi[tt] - 1;
                    e[t][u] = cand[off[tt] + (size_t)j * S2];
                }
            }
#pragma unroll
            for (int t = 0; t < 2; ++t)
#pragma unroll
                for (int u = 0; u < 4; ++u) nxv[t][u] = nx32[e[t][u].x];
#pragma unroll
            for (int t = 0; t < 2; ++t)
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    if (lo[t] + j0 + u >= hi[t]) continue;
                    double L, U;
                    hit_bounds(__int_as_float(e[t][u].y), nxv[t][u], nqq, c, L, U);
                    if (rows) rbuf[at[t]] = e[t][u].x;
                    buf[at[t]++] = __float_as_uint(fmaxf(__double2float_ru(U), 0.0f));
                }
        }
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    const int total = base;
    if (!rows) {
        const uint32_t kb = wave_kth_bits(total, k, [&](int j) { return buf[j]; });
        // kbest <- every value below the k-th, then copies of the k-th
        int put = 0;
        for (int j0 = 0; j0 < total; j0 += 64) {
            const int j = j0 + lane;
            const bool below = j < total && buf[j] < kb;
            const unsigned long long mask = ballot(below);
            const int pos = put + __popcll(mask & ((1ull << lane) - 1));
            if (below) kbest[q * k + pos] = __uint_as_float(buf[j]);
            put += __popcll(mask);
        }
        for (int j = put + lane; j < k; j += 64) kbest[q * k + j] = __uint_as_float(kb);
        if (lane == 0) {
            const float t_new = fminf(tau[q], __uint_as_float(kb));
            tau[q] = t_new;
            thr[q] = thr_from_tau((double)t_new, qa[q], c);
        }
        return;
    }
    // the K2 = min(2k, 32) smallest bounds with their rows (every value below the
    // K2-th, then entries equal to it, then empty slots); their float64 distances
    // replace the bounds where smaller, and the k smallest results become kbest.
    // Entries carried over from kbest hold such a distance already (an earlier
    // merge computed it): they are kept as -2 - row and not computed again.
    const int K2 = 2 * k < 32 ? 2 * k : 32;
    const uint32_t kb2 = wave_kth_bits(total, K2, [&](int j) { return buf[j]; });
    float* kv = sKV[w];
    int* kr = sKR[w];
    int put = 0;
    for (int pass = 0; pass < 2; ++pass) {
        for (int j0 = 0; j0 < total && put < K2; j0 += 64) {
            const int j = j0 + lane;
            const bool sel = j < total && (pass == 0 ? buf[j] < kb2 : buf[j] == kb2);
            const unsigned long long mask = ballot(sel);
            const int pos = put + __popcll(mask & ((1ull << lane) - 1));
            if (sel && pos < K2) {
                kv[pos] = __uint_as_float(buf[j]);
                kr[pos] = (j < k && rbuf[j] >= 0) ? -2 - rbuf[j] : rbuf[j];
            }
            put += __popcll(mask);
        }
    }
    for (int j = put + lane; j < K2; j += 64) {
        kv[j] = INFINITY;
        kr[j] = -1;
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    // float64 distances of the new rows, 8 lanes per row (any summation
    // order); sixteen rows per pass, the loads of a 64-feature slice of both
    // groups of eight issued before either sum
    const double g = f64_interval(d);
    const int gl = lane & 7, grp = lane >> 3;
    const float* qv = Q + q * d;
    for (int j0 = 0; j0 < K2; j0 += 16) {
        const int ja = j0 + grp, jb = j0 + 8 + grp;
        const int ra = ja < K2 ? kr[ja] : -1, rb = jb < K2 ? kr[jb] : -1;
        if (!ballot(ra >= 0 || rb >= 0)) continue;
        const float* xa = X + (size_t)(ra >= 0 ? ra : 0) * d;
        const float* xb = X + (size_t)(rb >= 0 ? rb : 0) * d;
        double pa = 0.0, pb = 0.0;
        for (int e0 = 0; e0 < d; e0 += 64) {
            // unconditional loads (features clamped to d - 1); padded terms zeroed
            float qq[8], va[8], vb[8];
#pragma unroll
            for (int t = 0; t < 8; ++t) {
                const int e = e0 + gl + 8 * t;
                const int ec = e < d ? e : d - 1;
                qq[t] = qv[ec];
                va[t] = xa[ec];
                vb[t] = xb[ec];
            }
#pragma unroll
            for (int t = 0; t < 8; ++t) {
                const bool in = e0 + gl + 8 * t < d;
                const double ta = in ? (double)qq[t] - (double)va[t] : 0.0;
                const double tb = in ? (double)qq[t] - (double)vb[t] : 0.0;
                pa += ta * ta;
                pb += tb * tb;
            }
        }
        for (int o = 4; o > 0; o >>= 1) {
            pa += __shfl_xor(pa, o, 8);
            pb += __shfl_xor(pb, o, 8);
        }
        if (gl == 0) {
            if (ra >= 0) {
                const float up = __double2float_ru(pa * (1.0 + g) * c.scale2);
                if (up < kv[ja]) kv[ja] = up;
            }
            if (rb >= 0) {
                const float up = __double2float_ru(pb * (1.0 + g) * c.scale2);
                if (up < kv[jb]) kv[jb] = up;
            }
        }
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    // rank of each entry by (value, slot); the k lowest become kbest, the k-th is tau
    float kth = 0.0f;
    if (lane < K2) {
        const float v = kv[lane];
        int rank = 0;
        for (int i = 0; i < K2; ++i) {
            const float u = kv[i];
            rank += (u < v) || (u == v && i < lane);
        }
        if (rank < k) {
            const int r = kr[lane];
            kbest[q * k + rank] = v;
            kbest_row[q * k + rank] = r < -1 ? -2 - r : r;
        }
        if (rank == k - 1) kth = v;
    }
    for (int o = 32; o > 0; o >>= 1) kth = fmaxf(kth, __shfl_xor(kth, o));
    if (lane == 0) {
        const float t_new = fminf(tau[q], kth);
        tau[q] = t_new;
        thr[q] = thr_from_tau((double)t_new, qa[q], c);
    }
}

// Merge with one lane per query, for few sub-lists per query and k <= KM
// (k-means assignment, small k): the k smallest upper bounds stay sorted in
// registers and every new entry goes through a compare-exchange chain. Same
// result as k_merge (kbest = the k smallest bounds of the union, tau = min of
// the old tau and the k-th).
template <int KM>
__global__ void k_merge_lane(int64_t m, int k, const int* __restrict__ cnt, int* __restrict__ seen,
                             const int2* __restrict__ cand, const float* __restrict__ nx32, int S2, int cap,
                             float* __restrict__ kbest, const double* __restrict__ nq, const double* __restrict__ qa,
                             Consts c, float* __restrict__ tau, float* __restrict__ thr, int* __restrict__ status) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= m || status[q]) return;
    float best[KM];
#pragma unroll
    for (int i = 0; i < KM; ++i) best[i] = INFINITY;
    auto insert = [&](float v) {
#pragma unroll
        for (int i = 0; i < KM; ++i) {
            const float lo = fminf(v, best[i]);
            v = fmaxf(v, best[i]);
            best[i] = lo;
        }
    };
    for (int i = 0; i < k; ++i) insert(kbest[q * k + i]);
    const double nqq = nq[q];
    bool over = false;
    int fresh = 0;
    for (int s = 0; s < S2; ++s) {
        const int cs = cnt[q * S2 + s];
        over |= cs > cap;
        const int hi = cs < cap ? cs : cap;
        const int lo = seen[q * S2 + s];
        seen[q * S2 + s] = hi;
        const size_t off = (size_t)q * cap * S2 + s;  // entry j at off + j * S2
        // four entries per step: their loads, then their norm gathers, issued together
        for (int j0 = lo; j0 < hi; j0 += 4) {
            int2 e[4];
            float nxv[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) e[u] = cand[off + (size_t)(j0 + u < hi ? j0 + u : hi - 1) * S2];
#pragma unroll
            for (int u = 0; u < 4; ++u) nxv[u] = nx32[e[u].x];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                if (j0 + u >= hi) continue;
                double L, U;
                hit_bounds(__int_as_float(e[u].y), nxv[u], nqq, c, L, U);
                insert(fmaxf(__double2float_ru(U), 0.0f));
                ++fresh;
            }
        }
    }
    if (over) {
        status[q] = ST_OVERFLOW;
        thr[q] = INFINITY;
        return;
    }
    if (fresh == 0) return;
#pragma unroll
    for (int i = 0; i < KM; ++i)
        if (i < k) kbest[q * k + i] = best[i];
    float kth = best[0];
#pragma unroll
    for (int i = 0; i < KM; ++i)
        if (i == k - 1) kth = best[i];
    const float t_new = fminf(tau[q], kth);
    tau[q] = t_new;
    thr[q] = thr_from_tau((double)t_new, qa[q], c);
}

// ---------------------------------------------------------------- rank
struct RankArgs {
    const float* X;
    int64_t n;
    int d;
    const float* Q;
    int64_t m;
    int k;
    // screen candidate lists
    const int* cnt;
    const int2* cand;
    const float* nx32;
    const double* nq;
    int S, cap;
    const float* tau;
    Consts c;
    const int* unsafe;
    int n_unsafe;
    // fallback plain lists: candidates of fallback slot f are plist/ps[f*pstride + j], j < pcount[f]
    const int* fq;
    const int* plist;
    const double* ps;
    const int* pcount;
    int64_t pstride;
    int mode;  // 0 = screen lists, 1 = fallback lists
    int64_t self_off;
    int exclude;
    int* status;
    int64_t* out_idx;
    double* out_dist;
    int squared;
    int* qstats;  // per query: list entries, after the bound filter, ranked
};

struct Cand {
    double s;
    long long i;
};

__device__ __forceinline__ Cand shfl_cand(Cand v, int o) {
    Cand r;
    r.s = __shfl_xor(v.s, o);
    r.i = __shfl_xor(v.i, o);
    return r;
}

// One wave ranks one query (item `item` of the call; query q): exact (S, index)
// order of its candidate set, held in LDS (li / ls: LC rows; *sN: the wave's
// counter; orow: 64 result rows). In mode 0, a query whose bound filter keeps
// more than LC rows is marked ST_BIG when LC < LCAP_BIG (ranked again with
// LCAP_BIG rows), else ST_OVERFLOW (exact fallback).
template <int LC>
__device__ void rank_query(const RankArgs& a, int64_t item, int64_t q, int* __restrict__ li, double* __restrict__ ls,
                           int* sN, int* orow) {
    const int lane = threadIdx.x & 63;
    const int d = a.d, k = a.k;
    const double g = f64_interval(d);
    const float* qv = a.Q + q * d;
    int count = 0;
    if (a.mode == 0) {
        const float tau = a.tau[q];
        const double nqq = a.nq[q];
        const int64_t self = a.exclude ? a.self_off + q : -1;
        // Rows whose certified lower bound does not exceed tau, appended in one
        // pass through an LDS counter (the rank below does not depend on the
        // order of the candidates).
        if (lane == 0) *sN = 0;
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        int listed = 0;
        bool over = false;
        // two sub-lists per lane (s0 + lane and s0 + lane + 64): their counts,
        // then their entries, then the entries' norms are loaded together
        for (int s0 = 0; s0 < a.S; s0 += 128) {
            int cs[2];
            size_t off[2];
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                const int s = s0 + lane + 64 * t;
                cs[t] = 0;
                if (s < a.S) {
                    cs[t] = a.cnt[q * a.S + s];
                    over |= cs[t] > a.cap;
                    cs[t] = cs[t] < a.cap ? cs[t] : a.cap;
                }
                listed += cs[t];
                off[t] = (size_t)q * a.cap * a.S + s;  // entry j at off + j * S
            }
            const int cm = cs[0] > cs[1] ? cs[0] : cs[1];
            for (int j0 = 0; j0 < cm; j0 += 8) {
                // unconditional loads from valid slots (a sub-list's slot j is
                // clamped to its last entry; an empty sub-list borrows the other
                // one's), so the compiler issues them together; slots past the
                // counts are ignored
                int2 e[2][8];
                float nxv[2][8];
#pragma unroll
                for (int t = 0; t < 2; ++t) {
                    const int tt = cs[t] > 0 ? t : 1 - t;
#pragma unroll
                    for (int u = 0; u < 8; ++u) {
                        const int jc = j0 + u < cs[tt] ? j0 + u : cs[tt] - 1;
                        e[t][u] = a.cand[off[tt] + (size_t)jc * a.S];
                    }
                }
#pragma unroll
                for (int t = 0; t < 2; ++t)
#pragma unroll
                    for (int u = 0; u < 8; ++u) nxv[t][u] = a.nx32[e[t][u].x];
#pragma unroll
                for (int t = 0; t < 2; ++t)
#pragma unroll
                    for (int u = 0; u < 8; ++u) {
                        if (j0 + u >= cs[t]) continue;
                        double L, U;
                        hit_bounds(__int_as_float(e[t][u].y), nxv[t][u], nqq, a.c, L, U);
                        if (L <= (double)tau) {
                            const int at = atomicAdd(sN, 1);
                            if (at < LC) li[at] = e[t][u].x;
                        }
                    }
            }
        }
        for (int o = 32; o > 0; o >>= 1) over = __shfl_xor((int)over, o) | over;
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        const int total = *sN;
        int extra = 0;
        for (int u = 0; u < a.n_unsafe; ++u) extra += a.unsafe[u] != self;
        if (over || total + extra > LC) {
            if (lane == 0) {
                a.status[q] = (!over && LC < LCAP_BIG) ? ST_BIG : ST_OVERFLOW;
                if (a.qstats) a.qstats[3 * q] = a.qstats[3 * q + 1] = a.qstats[3 * q + 2] = 0;
            }
            return;
        }
        if (lane == 0) {
            int u2 = total;
            for (int u = 0; u < a.n_unsafe; ++u)
                if (a.unsafe[u] != self) li[u2++] = a.unsafe[u];
        }
        count = total + extra;
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        // float64 distances, 8 lanes per candidate (any summation order: the
        // interval tests use the order-free bound g)
        const double cut = (double)tau / a.c.scale2;
        int kept = 0;
        const int gl = lane & 7;
        for (int b = 0; b < count; b += 16) {
            // 8 lanes per candidate, two candidates per lane group; all loads
            // of a 128-feature slice are issued before the arithmetic
            const int j0 = b + (lane >> 3), j1 = j0 + 8;
            const float* x0 = a.X + (size_t)(j0 < count ? li[j0] : 0) * d;
            const float* x1 = a.X + (size_t)(j1 < count ? li[j1] : 0) * d;
            double s0 = 0.0, s1 = 0.0;
            for (int e0 = 0; e0 < d; e0 += 128) {
                // unconditional loads (features clamped to d - 1, rows to row 0
                // past the list); padded terms are zeroed after the loads
                float qq[16], xa[16], xb[16];
#pragma unroll
                for (int t = 0; t < 16; ++t) {
                    const int e = e0 + gl + 8 * t;
                    const int ec = e < d ? e : d - 1;
                    qq[t] = qv[ec];
                    xa[t] = x0[ec];
                    xb[t] = x1[ec];
                }
#pragma unroll
                for (int t = 0; t < 16; ++t) {
                    const bool in = e0 + gl + 8 * t < d;
                    const double ta = in ? (double)qq[t] - (double)xa[t] : 0.0;
                    const double tb = in ? (double)qq[t] - (double)xb[t] : 0.0;
                    s0 += ta * ta;
                    s1 += tb * tb;
                }
            }
            for (int o = 4; o > 0; o >>= 1) {
                s0 += __shfl_xor(s0, o, 8);
                s1 += __shfl_xor(s1, o, 8);
            }
            if (gl == 0) {
                if (j0 < count) {
                    const bool ok = s0 * (1.0 - g) <= cut;
                    ls[j0] = ok ? s0 : INFINITY;
                    kept += ok;
                }
                if (j1 < count) {
                    const bool ok = s1 * (1.0 - g) <= cut;
                    ls[j1] = ok ? s1 : INFINITY;
                    kept += ok;
                }
            }
        }
        for (int o = 32; o > 0; o >>= 1) {
            kept += __shfl_xor(kept, o);
            listed += __shfl_xor(listed, o);
        }
        if (lane == 0 && a.qstats) {
            a.qstats[3 * q] = listed;
            a.qstats[3 * q + 1] = count;
            a.qstats[3 * q + 2] = kept;
        }
        if (kept < k) {
            if (lane == 0) a.status[q] = ST_OVERFLOW;
            return;
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    } else {
        count = a.pcount[item];
        if (count <= LC) {
            for (int j = lane; j < count; j += 64) {
                li[j] = a.plist[item * a.pstride + j];
                ls[j] = a.ps[item * a.pstride + j];
            }
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        } else {
            // Large tie sets: k rounds of exact successor extraction from global memory.
            Cand prev{-1.0, -1};
            for (int r = 0; r < k; ++r) {
                Cand best{INFINITY, -1};
                for (int j = lane; j < count; j += 64) {
                    Cand cj{a.ps[item * a.pstride + j], (long long)a.plist[item * a.pstride + j]};
                    if (prev.i >= 0 && !exact_precedes(prev.s, prev.i, cj.s, cj.i, g, qv, a.X, d)) continue;
                    if (best.i < 0 || exact_precedes(cj.s, cj.i, best.s, best.i, g, qv, a.X, d)) best = cj;
                }
                for (int o = 32; o > 0; o >>= 1) {
                    Cand other = shfl_cand(best, o);
                    if (other.i >= 0 &&
                        (best.i < 0 || exact_precedes(other.s, other.i, best.s, best.i, g, qv, a.X, d)))
                        best = other;
                }
                if (lane == 0) {
                    a.out_idx[q * k + r] = best.i;
                    if (a.out_dist) {
                        const double sc = canonical_sq(qv, a.X + (size_t)best.i * d, d);
                        a.out_dist[q * k + r] = a.squared ? sc : sqrt(sc);
                    }
                }
                prev = best;
            }
            return;
        }
    }
    // rank counting: position of each kept candidate in the exact (S, index) order
    if (lane < 64) orow[lane] = 0;  // every slot holds a valid row even if a rank stays unset
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    for (int j = lane; j < count; j += 64) {
        const double sj = ls[j];
        if (!(sj < INFINITY)) continue;
        const long long ij = li[j];
        int rank = 0;
        for (int i = 0; i < count && rank < k; ++i) {
            const double si = ls[i];
            if (!(si < INFINITY) || i == j) continue;
            rank += exact_precedes(si, li[i], sj, ij, g, qv, a.X, d);
        }
        if (rank < k) {
            a.out_idx[q * k + rank] = ij;
            if (k <= 64) {
                orow[rank] = (int)ij;
            } else if (a.out_dist) {
                // fallback route with k > 64: no staging, the lane reads its row
                const double sc = canonical_sq(qv, a.X + (size_t)ij * d, d);
                a.out_dist[q * k + rank] = a.squared ? sc : sqrt(sc);
            }
        }
    }
    if (!a.out_dist || k > 64) return;
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    // Canonical distances of the k results: the query and a chunk of result
    // rows are copied into LDS (the candidate distance array is free now) with
    // coalesced loads; then one lane per row forms the left-to-right sum.
    float* stage = (float*)ls;
    const int rs = d + 1;  // padded row stride
    const int ch = d < LC * 2 ? (LC * 2 - d) / rs : 0;
    if (ch == 0) {
        for (int r = lane; r < k; r += 64) {
            const double sc = canonical_sq(qv, a.X + (size_t)orow[r] * d, d);
            a.out_dist[q * k + r] = a.squared ? sc : sqrt(sc);
        }
        return;
    }
    float* sq_ = stage + ch * rs;
    for (int e = lane; e < d; e += 64) sq_[e] = qv[e];
    for (int r0 = 0; r0 < k; r0 += ch) {
        const int nr = k - r0 < ch ? k - r0 : ch;
        // 8 rows x 128 features per batch, all loads issued before the stores
        // (rows clamped to the chunk's last, features to d - 1)
        for (int e0 = 0; e0 < d; e0 += 128) {
            for (int rr0 = 0; rr0 < nr; rr0 += 8) {
                float v[8][2];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const int rr = rr0 + u < nr ? rr0 + u : nr - 1;
                    const float* xv = a.X + (size_t)orow[r0 + rr] * d;
#pragma unroll
                    for (int h2 = 0; h2 < 2; ++h2) {
                        const int e = e0 + lane + 64 * h2;
                        v[u][h2] = xv[e < d ? e : d - 1];
                    }
                }
#pragma unroll
                for (int u = 0; u < 8; ++u)
#pragma unroll
                    for (int h2 = 0; h2 < 2; ++h2) {
                        const int e = e0 + lane + 64 * h2;
                        if (rr0 + u < nr && e < d) stage[(rr0 + u) * rs + e] = v[u][h2];
                    }
            }
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        if (lane < nr) {
            const float* xr = stage + lane * rs;
            double sc = 0.0;
            for (int e = 0; e < d; ++e) {
                const double t = (double)sq_[e] - (double)xr[e];
                sc = add_rn(sc, mul_rn(t, t));
            }
            a.out_dist[q * k + r0 + lane] = a.squared ? sc : sqrt(sc);
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    }
}

// One wave per query, four per block.
__global__ __launch_bounds__(256) void k_rank(const RankArgs a, int64_t n_items) {
    __shared__ int sI[4][LCAP];
    __shared__ double sS[4][LCAP];
    __shared__ int sN[4];
    __shared__ int sO[4][64];  // result rows by rank
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int64_t item = (int64_t)blockIdx.x * 4 + w;
    if (item >= n_items) return;
    int64_t q;
    if (a.mode == 0) {
        q = item;
        if (a.status[q]) {  // handled by the fallback
            if (lane == 0 && a.qstats) a.qstats[3 * q] = a.qstats[3 * q + 1] = a.qstats[3 * q + 2] = 0;
            return;
        }
    } else {
        q = a.fq[item];
    }
    rank_query<LCAP>(a, item, q, sI[w], sS[w], &sN[w], sO[w]);
}

// Queries of the listed set (fq, n_items) marked ST_BIG by k_rank: ranked again
// with room for LCAP_BIG rows, one wave per block. A query that still does not
// fit keeps ST_OVERFLOW (exact fallback); the others are resolved (status 0) and
// their list statistics are added to sums[0..3).
__global__ __launch_bounds__(64) void k_rank_big(const RankArgs a, const int* __restrict__ fq, int64_t n_items,
                                                 unsigned long long* __restrict__ sums) {
    __shared__ int sI[LCAP_BIG];
    __shared__ double sS[LCAP_BIG];
    __shared__ int sN;
    __shared__ int sO[64];
    const int64_t item = blockIdx.x;
    if (item >= n_items) return;
    const int64_t q = fq[item];
    if (a.status[q] != ST_BIG) return;
    __builtin_amdgcn_wave_barrier();
    if (threadIdx.x == 0) a.status[q] = 0;
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    rank_query<LCAP_BIG>(a, q, q, sI, sS, &sN, sO);
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    if (threadIdx.x == 0 && a.qstats && a.status[q] == 0) {
        for (int c = 0; c < 3; ++c) atomicAdd(&sums[c], (unsigned long long)(unsigned)a.qstats[3 * q + c]);
    }
}

// fast_sq read four features at a time (d % 4 == 0, 16-byte aligned rows).
__device__ __forceinline__ double fast_sq4(const float* __restrict__ q, const float* __restrict__ x, int d) {
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    for (int i = 0; i < d; i += 4) {
        const float4 a = *reinterpret_cast<const float4*>(q + i);
        const float4 b = *reinterpret_cast<const float4*>(x + i);
        const double t0 = (double)a.x - (double)b.x, t1 = (double)a.y - (double)b.y;
        const double t2 = (double)a.z - (double)b.z, t3 = (double)a.w - (double)b.w;
        s0 += t0 * t0; s1 += t1 * t1; s2 += t2 * t2; s3 += t3 * t3;
    }
    return (s0 + s1) + (s2 + s3);
}

// Rank for k = 1, one lane per query (many queries with short candidate
// lists, as k-means assignment): the lane streams its query's sub-lists, keeps
// candidates whose certified lower bound does not exceed tau, computes their
// float64 distances (order-free bound g) and keeps the minimum under the exact
// (S, index) order; the result's distance is the canonical sum. Statuses and
// statistics are those of k_rank (no database row is unsafe on this route).
template <bool VEC>
__global__ __launch_bounds__(256) void k_rank1_lane(const RankArgs a, int64_t m, int ties_only) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= m) return;
    if (ties_only) {  // after the group rank: only its undecided queries
        if (a.status[q] != ST_TIE) return;
        a.status[q] = 0;
    } else if (a.status[q]) {  // handled by the fallback
        if (a.qstats) a.qstats[3 * q] = a.qstats[3 * q + 1] = a.qstats[3 * q + 2] = 0;
        return;
    }
    const int d = a.d;
    const double g = f64_interval(d);
    const float* qv = a.Q + q * d;
    const float tau = a.tau[q];
    const double nqq = a.nq[q];
    const double cut = (double)tau / a.c.scale2;
    int listed = 0, count = 0, kept = 0;
    bool over = false;
    double best_s = INFINITY;
    long long best_i = -1;
    for (int sl = 0; sl < a.S; ++sl) {
        int cs = a.cnt[q * a.S + sl];
        over |= cs > a.cap;
        cs = cs < a.cap ? cs : a.cap;
        listed += cs;
        const int2* lst = a.cand + (size_t)q * a.cap * a.S + sl;  // entry j at lst[j * S]
        for (int j = 0; j < cs; ++j) {
            const int2 e = lst[(size_t)j * a.S];
            double L, U;
            hit_bounds(__int_as_float(e.y), a.nx32[e.x], nqq, a.c, L, U);
            if (!(L <= (double)tau)) continue;
            ++count;
            const double sj = VEC ? fast_sq4(qv, a.X + (size_t)e.x * d, d) : fast_sq(qv, a.X + (size_t)e.x * d, d);
            if (!(sj * (1.0 - g) <= cut)) continue;
            ++kept;
            if (best_i < 0 || exact_precedes(sj, e.x, best_s, best_i, g, qv, a.X, d)) {
                best_s = sj;
                best_i = e.x;
            }
        }
    }
    if (over || kept < 1) {
        a.status[q] = ST_OVERFLOW;
        if (a.qstats) a.qstats[3 * q] = a.qstats[3 * q + 1] = a.qstats[3 * q + 2] = 0;
        return;
    }
    if (a.qstats) {
        a.qstats[3 * q] = listed;
        a.qstats[3 * q + 1] = count;
        a.qstats[3 * q + 2] = kept;
    }
    a.out_idx[q] = best_i;
    if (a.out_dist) {
        const float* xb = a.X + (size_t)best_i * d;
        const double sc = VEC ? canonical_sq4(qv, xb, d) : canonical_sq(qv, xb, d);
        a.out_dist[q] = a.squared ? sc : sqrt(sc);
    }
}

// Rank for k = 1 with a group of eight lanes per query (rows of at most
// 32 * NF4 features, d % 4 == 0, 16-byte aligned rows). The group holds its
// query row in registers (lane gl owns the 16-byte chunks gl + 8 t) and takes
// the list entries round-robin for the certified-bound filter. The rows that
// pass are compacted, in list order, into the group's LDS slots; their float64
// distances are then formed SB rows at a time from coalesced 16-byte segments
// whose loads are issued together (one memory latency per SB rows), reduced
// over the group by a butterfly (every lane gets the same sum; any order is
// covered by the order-free bound g). Loop bounds and cross-lane steps are
// uniform over the wave: groups with shorter lists, or no query, take part
// with empty batches, so no shuffle runs under a partial mask. Candidates are
// compared by their float64 intervals only; a query with an undecided
// comparison (a near or exact tie) is marked ST_TIE and ranked again by
// k_rank1_lane, whose exact comparator resolves it. The output distance is the
// canonical sum: for d <= 128 the lanes that own the features form its
// products and the group's first lane adds them in feature order from LDS.
// Results and statistics are those of k_rank1_lane.
template <int NF4>
__global__ __launch_bounds__(256) void k_rank1_group(const RankArgs a, int64_t m) {
    constexpr int G = 8;
    constexpr int SB = NF4 <= 4 ? 4 : 2;  // surviving rows whose loads are issued together
    constexpr bool LDS_SUM = NF4 <= 4;    // canonical products staged in LDS
    __shared__ int srows[256 / G][4 * G];
    __shared__ double sprod[LDS_SUM ? 256 / G : 1][LDS_SUM ? 32 * NF4 : 1];
    const int lane = threadIdx.x & 63, gl = lane % G, gbase = lane - gl, grp = threadIdx.x / G;
    const int64_t q = ((int64_t)blockIdx.x * 256 + threadIdx.x) / G;
    bool live = q < m;
    if (live && a.status[q]) {  // handled by the fallback
        if (gl == 0 && a.qstats) a.qstats[3 * q] = a.qstats[3 * q + 1] = a.qstats[3 * q + 2] = 0;
        live = false;
    }
    if (ballot(live) == 0) return;  // the whole wave leaves together
    const int d = a.d;
    const double g = f64_interval(d);
    const int64_t qs = live ? q : 0;
    f32x4 qr[NF4];
#pragma unroll
    for (int t = 0; t < NF4; ++t) {
        const int f = (t * G + gl) * 4;
        qr[t] = live && f < d ? *(const f32x4*)(a.Q + qs * d + f) : f32x4{0.f, 0.f, 0.f, 0.f};
    }
    const float tau = live ? a.tau[q] : 0.f;
    const double nqq = live ? a.nq[q] : 0.0;
    const double cut = (double)tau / a.c.scale2;
    int listed = 0, count = 0, kept = 0;
    bool over = false, tie = false;
    double best_s = INFINITY;
    long long best_i = -1;
    for (int sl = 0; sl < a.S; ++sl) {
        int cs = live ? a.cnt[q * a.S + sl] : 0;
        over |= cs > a.cap;
        cs = cs < a.cap ? cs : a.cap;
        listed += cs;
        int cw = cs;  // longest list of the wave (cs is uniform in the group)
#pragma unroll
        for (int o = 32; o >= G; o >>= 1) cw = max(cw, __shfl_xor(cw, o));
        const int2* lst = a.cand + (size_t)qs * a.cap * a.S + sl;  // entry j at lst[j * S]
        for (int j0 = 0; j0 < cw; j0 += 4 * G) {
            int2 e[4];
            float nxv[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int j = j0 + gl + G * u;
                e[u] = j < cs ? lst[(size_t)j * a.S] : make_int2(0, 0);
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) nxv[u] = a.nx32[e[u].x];
            int ns = 0;  // rows of this chunk passing the bound filter (uniform in the group)
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                bool pass = false;
                if (j0 + gl + G * u < cs) {
                    double L, U;
                    hit_bounds(__int_as_float(e[u].y), nxv[u], nqq, a.c, L, U);
                    pass = L <= (double)tau;
                }
                const uint32_t mu = (uint32_t)(ballot(pass) >> gbase) & ((1u << G) - 1u);
                if (pass) srows[grp][ns + __popc(mu & ((1u << gl) - 1u))] = e[u].x;
                ns += __popc(mu);
            }
            count += ns;
            int nw = ns;
#pragma unroll
            for (int o = 32; o >= G; o >>= 1) nw = max(nw, __shfl_xor(nw, o));
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            for (int b0 = 0; b0 < nw; b0 += SB) {
                int rows[SB];
#pragma unroll
                for (int t = 0; t < SB; ++t) rows[t] = b0 + t < ns ? srows[grp][b0 + t] : -1;
                f32x4 xr[SB][NF4];
#pragma unroll
                for (int t = 0; t < SB; ++t) {
#pragma unroll
                    for (int tt = 0; tt < NF4; ++tt) {
                        const int f = (tt * G + gl) * 4;
                        xr[t][tt] = rows[t] >= 0 && f < d ? *(const f32x4*)(a.X + (size_t)rows[t] * d + f)
                                                          : f32x4{0.f, 0.f, 0.f, 0.f};
                    }
                }
                double part[SB];
#pragma unroll
                for (int t = 0; t < SB; ++t) {
                    part[t] = 0.0;
#pragma unroll
                    for (int tt = 0; tt < NF4; ++tt) {
#pragma unroll
                        for (int c = 0; c < 4; ++c) {
                            const double dt = (double)qr[tt][c] - (double)xr[t][tt][c];
                            part[t] = fma(dt, dt, part[t]);
                        }
                    }
                }
#pragma unroll
                for (int t = 0; t < SB; ++t) {
#pragma unroll
                    for (int o = G / 2; o > 0; o >>= 1) part[t] += __shfl_xor(part[t], o, G);
                }
#pragma unroll
                for (int t = 0; t < SB; ++t) {
                    if (rows[t] < 0) continue;
                    const double sj = part[t];
                    if (!(sj * (1.0 - g) <= cut)) continue;
                    ++kept;
                    // the interval decides almost every comparison; an undecided one (a
                    // near or exact tie) leaves the query to the lane rank's exact comparator
                    if (best_i < 0 || sj * (1.0 + g) < best_s * (1.0 - g)) {
                        best_s = sj;
                        best_i = rows[t];
                    } else if (!(best_s * (1.0 + g) < sj * (1.0 - g))) {
                        tie = true;
                    }
                }
            }
        }
    }
    if (!live) return;
    if (over || kept < 1) {
        if (gl == 0) {
            a.status[q] = ST_OVERFLOW;
            if (a.qstats) a.qstats[3 * q] = a.qstats[3 * q + 1] = a.qstats[3 * q + 2] = 0;
        }
        return;
    }
    if (tie) {
        if (gl == 0) a.status[q] = ST_TIE;
        return;
    }
    if (gl == 0) {
        if (a.qstats) {
            a.qstats[3 * q] = listed;
            a.qstats[3 * q + 1] = count;
            a.qstats[3 * q + 2] = kept;
        }
        a.out_idx[q] = best_i;
    }
    if (!a.out_dist) return;
    if constexpr (LDS_SUM) {
        // canonical products by the lanes owning the features (the same roundings
        // as canonical_sq4), summed left to right by the group's first lane
        const float* xb = a.X + (size_t)best_i * d;
#pragma unroll
        for (int t = 0; t < NF4; ++t) {
            const int f = (t * G + gl) * 4;
            if (f < d) {
                const f32x4 x4 = *(const f32x4*)(xb + f);
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    const double dt = (double)qr[t][c] - (double)x4[c];
                    sprod[grp][f + c] = mul_rn(dt, dt);
                }
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (gl == 0) {
            double sc = 0.0;
            for (int i = 0; i < d; ++i) sc = add_rn(sc, sprod[grp][i]);
            a.out_dist[q] = a.squared ? sc : sqrt(sc);
        }
    } else if (gl == 0) {
        const double sc = canonical_sq4(a.Q + q * d, a.X + (size_t)best_i * d, d);
        a.out_dist[q] = a.squared ? sc : sqrt(sc);
    }
}

// ---------------------------------------------------------------- fallback
__global__ void k_fb_dist(const float* __restrict__ X, int64_t n, int d, const float* __restrict__ Q,
                          const int* __restrict__ fq, int64_t self_off, int exclude, double* __restrict__ s_out,
                          int k, unsigned* __restrict__ hist, unsigned long long* __restrict__ prefix,
                          unsigned* __restrict__ rank) {
    // 16 lanes per row, 16 rows per block pass; float64 partial sums in any order.
    const int f = blockIdx.y;
    if (blockIdx.x == 0) {  // radix-select state of this query
        hist[f * 256 + threadIdx.x] = 0;
        if (threadIdx.x == 0) {
            prefix[f] = 0;
            rank[f] = (unsigned)k;
        }
    }
    const int64_t q = fq[f];
    const float* qv = Q + q * d;
    const int g = threadIdx.x & 15;
    const int64_t row = (int64_t)blockIdx.x * 16 + (threadIdx.x >> 4);
    double s = 0.0;
    if (row < n) {
        const float* xv = X + row * d;
        for (int j = g; j < d; j += 16) {
            const double t = (double)qv[j] - (double)xv[j];
            s += t * t;
        }
    }
    for (int o = 8; o > 0; o >>= 1) s += __shfl_xor(s, o, 16);
    if (row < n && g == 0) s_out[(size_t)f * n + row] = (exclude && row == self_off + q) ? INFINITY : s;
}

// k-th smallest float64 of each fallback query by 8-bit radix select over the
// bit pattern (non-negative doubles order like their bits): eight passes, each
// a histogram of the values that share the digits chosen so far, spread over
// many blocks per query (k_fb_hist), then one block per query that picks the
// digit holding the k-th value (k_fb_pick). k_fb_dist sets prefix 0, rank k and
// clears the histograms.
__global__ void k_fb_hist(const double* __restrict__ s_in, int64_t n, int shift,
                          const unsigned long long* __restrict__ prefix, unsigned* __restrict__ hist) {
    __shared__ unsigned h[256];
    const int f = blockIdx.y;
    h[threadIdx.x] = 0;
    __syncthreads();
    const double* s = s_in + (size_t)f * n;
    const unsigned long long pre = prefix[f];
    const unsigned long long mask = shift >= 56 ? 0ull : ~0ull << (shift + 8);
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const unsigned long long v = (unsigned long long)__double_as_longlong(s[i]);
        if ((v & mask) == pre) atomicAdd(&h[(v >> shift) & 255], 1u);
    }
    __syncthreads();
    if (h[threadIdx.x]) atomicAdd(&hist[f * 256 + threadIdx.x], h[threadIdx.x]);
}

__global__ void k_fb_pick(int shift, int d, unsigned* __restrict__ hist, unsigned long long* __restrict__ prefix,
                          unsigned* __restrict__ rank, double* __restrict__ cut) {
    __shared__ unsigned h[256];
    const int f = blockIdx.x;
    h[threadIdx.x] = hist[f * 256 + threadIdx.x];
    hist[f * 256 + threadIdx.x] = 0;  // ready for the next pass
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned want = rank[f];  // 1-based rank within the current prefix class
        unsigned acc = 0;
        int b = 0;
        for (; b < 256; ++b) {
            if (acc + h[b] >= want) break;
            acc += h[b];
        }
        if (b == 256) b = 255;
        const unsigned long long p = prefix[f] | ((unsigned long long)b << shift);
        prefix[f] = p;
        rank[f] = want - acc;
        if (shift == 0) {
            const double tk = __longlong_as_double((long long)p);
            const double g = f64_interval(d);
            cut[f] = tk < INFINITY ? tk * (1.0 + g) / (1.0 - g) * (1.0 + 0x1p-50) : INFINITY;
        }
    }
}

__global__ void k_fb_collect(const double* __restrict__ s_in, int64_t n, const double* __restrict__ cut,
                             int* __restrict__ pcount, int* __restrict__ plist, double* __restrict__ ps,
                             int64_t pstride) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int f = blockIdx.y;
    if (i >= n) return;
    const double s = s_in[(size_t)f * n + i];
    if (s <= cut[f]) {
        const int at = atomicAdd(&pcount[f], 1);
        plist[(size_t)f * pstride + at] = (int)i;
        ps[(size_t)f * pstride + at] = s;
    }
}

// After the rank: queries without a certificate are listed for the exact
// fallback, and the rank's statistics (three counts per query, when qs is
// given) are summed into out (zeroed by the caller; block partials, then one
// atomic per count; integer sums, any order).
__global__ void k_finish_status(int64_t m, const int* __restrict__ status, int* __restrict__ fq,
                                unsigned* __restrict__ nfq, const int* __restrict__ qs,
                                unsigned long long* __restrict__ out) {
    __shared__ unsigned long long sh[3][256];
    unsigned long long a0 = 0, a1 = 0, a2 = 0;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < m; q += (int64_t)gridDim.x * blockDim.x) {
        if (status[q]) {
            const unsigned at = atomicAdd(nfq, 1u);
            fq[at] = (int)q;
        }
        if (qs) {
            a0 += (unsigned)qs[3 * q];
            a1 += (unsigned)qs[3 * q + 1];
            a2 += (unsigned)qs[3 * q + 2];
        }
    }
    if (!qs) return;
    sh[0][threadIdx.x] = a0; sh[1][threadIdx.x] = a1; sh[2][threadIdx.x] = a2;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o)
            for (int c = 0; c < 3; ++c) sh[c][threadIdx.x] += sh[c][threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x < 3 && sh[threadIdx.x][0]) atomicAdd(&out[threadIdx.x], sh[threadIdx.x][0]);
}

// Query part of every screen threshold (screen_qa), once per call.
__global__ void k_qa(int64_t m_pad, const double* __restrict__ nq, const double* __restrict__ nt,
                     const double* __restrict__ d2, Consts c, double* __restrict__ qa) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < m_pad) qa[q] = screen_qa(nq[q], nt[q], d2[q], c);
}

__global__ void k_fill_int(int* __restrict__ out, int64_t count, int value) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) out[i] = value;
}

__global__ void k_force(int64_t m, const int64_t* __restrict__ rows, int64_t count, int* __restrict__ status,
                        float* __restrict__ thr) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count && rows[i] >= 0 && rows[i] < m) {
        status[rows[i]] |= ST_FORCED;
        thr[rows[i]] = INFINITY;
    }
}

// ---------------------------------------------------------------- host
static double unit_roundoff(int prec) {
    switch (prec) {
        case PREC_FP32: return 0x1p-24;
        case PREC_BF16: return 0x1p-8 + 0x1p-23;
        default: return 0x1p-11 + 0x1p-23;
    }
}

Consts make_consts(int prec, int dp, int d, bool rows_exact, bool queries_exact, double scale) {
    Consts c;
    const double eta = rows_exact && queries_exact ? 0.0 : unit_roundoff(prec);
    const double Kp = dp + 1;
    c.gamma = Kp * 0x1p-23 / (1.0 - Kp * 0x1p-23);
    c.H = 2 * eta + eta * eta + c.gamma * (1 + eta) * (1 + eta) + (d + 8) * 0x1p-50;
    c.G = c.H + c.gamma + 0x1p-40;
    c.f0 = (dp + 2) * 0x1p-126;
    c.scale2 = scale * scale;
    // splitting parameter of the screen bound: half the unit roundoff sits near
    // the measured ratio |v - v~| / |v~| of rounded rows (about 0.42 of it for
    // bf16 and fp16 on Gaussian data); 0 where that side's residuals vanish
    const double t = 0.5 * unit_roundoff(prec);
    c.t1 = rows_exact ? 0.0 : t;
    c.t2 = queries_exact ? 0.0 : t;
    c.eps = (d + 8) * 0x1p-50;
    return c;
}

// The buffers of one k-NN call: the plan and every temporary's size. Shared by
// the call itself and by knn_arena_bytes, so a fixed arena is sized exactly.
struct KnnLayout {
    Plan plan;
    bool all_fallback = false;
    bool lane_merge = false;
    bool kbest_rows = false;
    int64_t QB = 0, m_pad = 0;
    int S = 0, S2 = 0, cap = 0;
    size_t bytes_main = 0;  // everything but the fallback's per-query buffers
};

static KnnLayout knn_layout(const mlk_index* index, int prec, int64_t m, int k, const TestHooks& hk,
                            size_t list_budget) {
    KnnLayout L;
    const Mirror& mr = index->mirror[prec];
    const int dp = mr.dp;
    L.plan = resolve_plan(PLAN_KNN, prec, dp, index->n, m, k);
    const ScreenFn screen = pick_screen(SCREEN_LISTS, prec, dp, L.plan.shape);
    L.all_fallback = !screen || k > 64 || mr.n_unsafe > 64 || (hk.hooks & MLK_HOOK_FORCE_FALLBACK);
    if (hk.hooks & MLK_HOOK_CAPACITY) L.plan.cap = (int)std::max<int64_t>(1, hk.param);
    const int BQ = 4 * L.plan.shape.qt * 32;
    L.QB = (m + BQ - 1) / BQ;
    L.m_pad = L.QB * BQ;
    // a bounded list budget: fewer database splits (whole XCD groups while
    // there are at least eight), the capacity following the split count
    while (list_budget && L.plan.splits > 1 &&
           (size_t)L.m_pad * 2 * L.plan.splits * (size_t)L.plan.cap * sizeof(int2) > list_budget) {
        const int s = L.plan.splits / 2;
        L.plan.splits = s >= 8 ? s / 8 * 8 : std::max(1, s);
        if (!(hk.hooks & MLK_HOOK_CAPACITY)) L.plan.cap = knn_capacity(prec, L.plan.splits, k);
    }
    L.S = L.plan.splits;
    L.S2 = 2 * L.S;
    L.cap = L.plan.cap;
    // one lane per query merges when a query has few sub-lists and k is small
    L.lane_merge = L.S2 <= 16 && k <= 8;
    // rows of the kbest entries (sized whenever the call may use them; the call
    // decides once the screen's error constant is known)
    L.kbest_rows = !L.lane_merge && k <= 32;
    const size_t mp = (size_t)L.m_pad, al = 256;
    auto piece = [&](size_t b) { return (b + al - 1) / al * al; };
    size_t t = 0;
    t += piece(mp * dp * elem_bytes(prec)) + 4 * piece(mp * 8) + piece(mp + 64) + 3 * piece(mp * 4) + piece(64);
    if (!L.all_fallback) {
        t += 2 * piece(mp * L.S2 * 4) + piece(mp * L.S2 * (size_t)L.cap * 8) + piece(mp * k * 4);
        if (L.kbest_rows) t += piece(mp * k * 4);
        t += piece((size_t)m * 12);
    }
    if (!hk.forced_rows.empty()) t += piece(hk.forced_rows.size() * 8);
    t += piece((size_t)m * 4) + piece(64 * 5);
    L.bytes_main = t;
    return L;
}

// Bytes of the candidate lists of a call (0 if it screens nothing).
static size_t knn_list_bytes(const mlk_index* index, int prec, int64_t m, int k) {
    const KnnLayout L = knn_layout(index, prec, m, k, hooks_snapshot(), 0);
    return L.all_fallback ? 0 : (size_t)L.m_pad * L.S2 * (size_t)L.cap * sizeof(int2);
}
// candidate-list bytes of one chunk of queries: 10^7 k-means rows over 65,536
// centres (15.4 GB of lists) run in one piece (four chunks of a 4 GiB budget
// cost 4 % more screen time in launch tails), far less than HBM holds
constexpr size_t KNN_LIST_BUDGET = size_t(16) << 30;
constexpr int64_t KNN_CHUNK_MIN = 65536;              // smallest chunk of a split call

size_t knn_fallback_bytes_per_query(int64_t n) {
    // float64 distances, candidate rows, their float64 values, and the radix state
    return (size_t)n * (sizeof(double) * 2 + sizeof(int)) + sizeof(int) + sizeof(double) +
           sizeof(unsigned long long) + sizeof(unsigned) + 1024;
}

size_t knn_arena_bytes(const mlk_index* index, int prec, int64_t m, int k, uint32_t flags, size_t list_budget) {
    (void)flags;
    if (!index->mirror[prec].ready) return 0;
    return knn_layout(index, prec, m, k, hooks_snapshot(), list_budget).bytes_main;
}

// One k-NN call between its enqueued device chain and its completion.
struct KnnCall {
    int64_t m = 0, n = 0, q_offset = 0;
    int d = 0, k = 0, prec = 0;
    uint32_t flags = 0;
    bool exclude = false;
    const float* Q = nullptr;
    const float* X = nullptr;
    double* dist = nullptr;
    int64_t* idx = nullptr;
    int* status = nullptr;
    int* fq = nullptr;
    unsigned* dflag = nullptr;
    bool big_rank = false;  // k_rank ran: queries marked ST_BIG are ranked again
    // regrouped queries: the chain's results (idx, dist) are by screen
    // position; knn_complete moves row p to the caller's row qperm[p]
    const int* qperm = nullptr;
    int64_t* user_idx = nullptr;
    double* user_dist = nullptr;
    RankArgs ra{};
    Consts c{};
    int64_t rounds = 0, splits = 0, capacity = 0;
    // timed calls: the start of the chain's first kernel and the end of its
    // status kernel, bound to those launches (no marker packets in the stream)
    hipEvent_t ev_first = nullptr, ev_status = nullptr;
};

// Enqueue a call's device chain on `stream`, up to the per-query status
// (k_finish_status). `capture`: the stream is being captured into a graph, so
// nothing may wait for the device (the queries' exactness is not read back:
// the screen then assumes inexact queries, which only loosens its bound), and
// no event is recorded. `events` (optional) times the call and its screens.
static int knn_enqueue(mlk_index* index, const float* Q, int64_t m, int k, int64_t q_offset, uint32_t flags,
                       double* dist, int64_t* idx, hipStream_t stream, CallMem& mem, bool capture, KnnCall& call,
                       Events* events, std::vector<hipEvent_t>* round_events) {
    mlk_stats& st = stats();
    const bool fixed = mem.fixed();  // no allocation
    const bool timed = events != nullptr;
    const int64_t n = index->n;
    const int d = index->d;
    const int prec = effective_prec((int)((flags & MLK_PREC_MASK) >> 4), d);
    const bool exclude = (flags & MLK_EXCLUDE_SELF) != 0;
    st.screen_precision = (uint32_t)prec << 4;  // the precision actually screened
    if (timed) MLK_CHECK_HIP(events->make(&call.ev_first));

    if (fixed && !index->mirror[prec].ready) {
        set_error("the index has no screen mirror for this precision yet: one mlk_knn_query call of the same "
                  "precision prepares it");
        return MLK_EINVAL;
    }
    MLK_TRY(prepare_mirror(index, prec, stream));
    Mirror& mr = index->mirror[prec];
    const int dp = mr.dp;
    const TestHooks hk = hooks_snapshot();
    const KnnLayout lay = knn_layout(index, prec, m, k, hk, mem.list_budget);
    const Plan& plan = lay.plan;
    const ScreenFn screen = pick_screen(SCREEN_LISTS, prec, dp, plan.shape);
    const bool all_fallback = lay.all_fallback;
    const int64_t QB = lay.QB;
    const int64_t m_pad = lay.m_pad;
    const int S = lay.S;
    const int cap = lay.cap;
    st.splits = S;
    st.capacity = cap;

    // temporaries
    const size_t eb = elem_bytes(prec);
    void* qmirror;
    double* nq;
    uint8_t* qflags;
    MLK_TRY(mem.get(0, (size_t)m_pad * dp * eb, &qmirror));     // query mirror
    MLK_TRY(mem.get(1, (size_t)m_pad * sizeof(double), &nq));   // nq
    MLK_TRY(mem.get(2, (size_t)m_pad + 64, &qflags));           // qflags
    double* qnt;
    double* qd2;
    double* qa;
    MLK_TRY(mem.get(22, (size_t)m_pad * sizeof(double), &qnt));  // |q~|^2
    MLK_TRY(mem.get(23, (size_t)m_pad * sizeof(double), &qd2));  // |q - q~|^2
    MLK_TRY(mem.get(24, (size_t)m_pad * sizeof(double), &qa));   // threshold part
    float* tau;
    float* thr;
    int* status;
    MLK_TRY(mem.get(3, (size_t)m_pad * sizeof(float), &tau));
    MLK_TRY(mem.get(4, (size_t)m_pad * sizeof(float), &thr));
    MLK_TRY(mem.get(5, (size_t)m_pad * sizeof(int), &status));
    unsigned* dflag;
    MLK_TRY(mem.get(6, 64, &dflag));
    // flag words: [0] query rounding inexact, [2] fallback count, [4..10) three
    // 64-bit statistics sums (cleared by a kernel: graphs replay no memset nodes)
    if (timed)
        hipExtLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, stream, call.ev_first, nullptr, 0, (int*)dflag,
                              (int64_t)10, 0);
    else
        hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, stream, (int*)dflag, (int64_t)10, 0);
    // Query locality order (calls without self-exclusion that may allocate):
    // the chain below runs on the regrouped queries, the start threshold of
    // each block of them samples rows of their own neighbourhood, and the
    // results are written by position for knn_complete to move to the caller's
    // rows. A captured call regroups only if an eager call prepared the rows.
    int64_t* const idx_caller = idx;
    double* const dist_caller = dist;
    const int* qperm = nullptr;
    const int* qcell = nullptr;
    if (!exclude && !fixed && order_wanted(index, m) && (!capture || index->order_rows_ready)) {
        void* oscr;
        int* qp;
        float* Qp;
        MLK_TRY(mem.get(25, order_scratch_bytes(m), &oscr));
        MLK_TRY(mem.get(26, (size_t)m_pad * sizeof(int), &qp));
        MLK_TRY(mem.get(27, (size_t)m * d * sizeof(float) + 16, &Qp));
        MLK_TRY(prepare_order_rows(index, stream));
        MLK_TRY(query_order(index, Q, m, m_pad, oscr, qp, Qp, stream));
        int64_t* idx_pos;
        MLK_TRY(mem.get(28, (size_t)m * k * sizeof(int64_t), &idx_pos));
        double* dist_pos = nullptr;
        if (dist) MLK_TRY(mem.get(29, (size_t)m * k * sizeof(double), &dist_pos));
        Q = Qp;
        idx = idx_pos;
        dist = dist_pos;
        qperm = qp;
        qcell = static_cast<const int*>(oscr);
    }
    // the queries' rounding only matters when every database value is exact
    // in the screen type (integer data): only then is it recorded and read
    // back (otherwise no flag traffic and no wait here)
    MLK_TRY(mirror_rows(prec, Q, m, m_pad, d, dp, index->centre.as<float>(), mr.scale, qmirror, nq, qflags,
                        mr.exact ? dflag : nullptr, qnt, qd2, stream));
    unsigned q_inexact = 1;
    if (mr.exact && !capture) {
        MLK_CHECK_HIP(hipMemcpyAsync(&q_inexact, dflag, 4, hipMemcpyDeviceToHost, stream));
        MLK_CHECK_HIP(hipStreamSynchronize(stream));
    }
    const Consts c = make_consts(prec, dp, d, mr.exact, mr.exact && q_inexact == 0, mr.scale);
    MLK_TRY(prepare_bias(index, prec, c, stream, !fixed));
    hipLaunchKernelGGL(k_qa, dim3((unsigned)((m_pad + 255) / 256)), dim3(256), 0, stream, m_pad, nq, qnt, qd2, c, qa);

    const float* X = index->X.as<float>();
    // sampled start threshold: 32 rows per neighbour asked for, 64 to 256 rows
    const int tau_samples = std::min(TAU0_SAMPLES, std::max(64, 32 * k));
    const int S2 = lay.S2;  // candidate sub-lists per query (split x lane half)
    const bool lane_merge = lay.lane_merge;
    int* kbest_row = nullptr;
    int* cnt = nullptr;
    int2* cand = nullptr;
    int* seen = nullptr;
    float* kbest = nullptr;
    if (!all_fallback) {
        MLK_TRY(mem.get(8, (size_t)m_pad * S2 * sizeof(int), &cnt));                 // counts
        MLK_TRY(mem.get(9, (size_t)m_pad * S2 * cap * sizeof(int2), &cand));         // (row, dot-product estimate)
        MLK_TRY(mem.get(17, (size_t)m_pad * S2 * sizeof(int), &seen));               // merged counts
        MLK_TRY(mem.get(18, (size_t)m_pad * k * sizeof(float), &kbest));             // k smallest upper bounds
        // rows of the kbest entries: the wave merge tightens tau with their float64
        // distances, which pays where the screen's band is wide (relative error
        // constant >= 2^-11: bf16, and fp16 on inexact inputs); float32 screens and
        // inputs that are exact in the screen type have a narrow band already
        if (lay.kbest_rows && c.H >= 0x1p-11) MLK_TRY(mem.get(21, (size_t)m_pad * k * sizeof(int), &kbest_row));
    }
    {
        // U sample rows per lane (64 U samples), QW queries per wave: U * QW = 32
        // float64 accumulators per lane
        const int samples = (int)std::min<int64_t>(n, tau_samples);
        int launch_rc = MLK_OK;
        auto launch = [&](auto kern, int qw) {
            const int64_t waves = (m_pad + qw - 1) / qw;
            const int64_t blocks = (waves + 3) / 4;
            int* srows = nullptr;  // regrouped queries: each block's sample rows
            if (qperm) {
                launch_rc = mem.get(30, (size_t)blocks * samples * sizeof(int), &srows);
                if (launch_rc == MLK_OK)
                    launch_rc = tau_sample_rows(index, qcell, qperm, m, 4 * qw, blocks, samples, srows, stream);
                if (launch_rc != MLK_OK) return;
            }
            hipLaunchKernelGGL(kern, dim3((unsigned)blocks), dim3(256), 0, stream, X, n, d, Q, m, m_pad,
                               k, tau_samples, q_offset, (int)exclude, qa, qflags, c, tau, thr, status, S2,
                               all_fallback ? nullptr : cnt, seen, kbest, kbest_row, srows);
        };
        // the most queries per wave (register reuse of the samples) that still
        // leaves about 4096 waves (4 per SIMD) to fill the device
        auto qw_for = [&](int most) {
            int qw = most;
            while (qw > 1 && (m_pad + qw - 1) / qw < 4096) qw >>= 1;
            return qw;
        };
        if (samples <= 64) {
            const int qw = qw_for(32);
            if (qw == 32) launch(k_tau0<1, 32>, 32);
            else if (qw >= 8) launch(k_tau0<1, 8>, 8);
            else if (qw >= 2) launch(k_tau0<1, 2>, 2);
            else launch(k_tau0<1, 1>, 1);
        } else if (samples <= 128) {
            const int qw = qw_for(16);
            if (qw == 16) launch(k_tau0<2, 16>, 16);
            else if (qw >= 4) launch(k_tau0<2, 4>, 4);
            else if (qw >= 2) launch(k_tau0<2, 2>, 2);
            else launch(k_tau0<2, 1>, 1);
        } else {
            const int qw = qw_for(8);
            if (qw == 8) launch(k_tau0<4, 8>, 8);
            else if (qw >= 4) launch(k_tau0<4, 4>, 4);
            else if (qw >= 2) launch(k_tau0<4, 2>, 2);
            else launch(k_tau0<4, 1>, 1);
        }
        MLK_TRY(launch_rc);
    }
    MLK_CHECK_HIP(hipGetLastError());

    if ((hk.hooks & MLK_HOOK_FORCE_UNCERTIFIED) && !hk.forced_rows.empty()) {
        const auto& rows = hk.forced_rows;
        int64_t* drows;
        MLK_TRY(mem.get(7, rows.size() * sizeof(int64_t), &drows));
        MLK_CHECK_HIP(hipMemcpyAsync(drows, rows.data(), rows.size() * sizeof(int64_t), hipMemcpyHostToDevice,
                                     stream));
        hipLaunchKernelGGL(k_force, dim3((unsigned)((rows.size() + 255) / 256)), dim3(256), 0, stream, m, drows,
                           (int64_t)rows.size(), status, thr);
        MLK_CHECK_HIP(hipStreamSynchronize(stream));  // the host rows are a snapshot of this call
        MLK_CHECK_HIP(hipGetLastError());
    }

    const int* rank_stats = nullptr;  // per-query counts written by the rank (screened calls)
    bool big_rank = false;
    RankArgs ra{};
    if (!all_fallback) {
        ScreenArgs sa{};
        sa.xs = mr.xs.ptr;
        sa.bias = mr.bias.as<float>();
        sa.nx32 = mr.nx32.as<float>();
        sa.qs = qmirror;
        sa.thr = thr;
        sa.cnt = cnt;
        sa.cand = cand;
        sa.n = n;
        sa.split_rows = (n + S - 1) / S;
        sa.S = S;
        sa.cap = cap;
        sa.QB = (int)QB;
        sa.self_off = q_offset;
        sa.exclude = exclude;
        sa.nchunk = screen_nchunk(SCREEN_LISTS, prec, dp, plan.shape);
        sa.qgroup = screen_qgroup(prec, dp);
        const unsigned blocks = (unsigned)screen_grid(sa);
        auto merge = [&]() -> int {
            if (lane_merge)
                hipLaunchKernelGGL(k == 1 ? k_merge_lane<1> : k_merge_lane<8>, dim3((unsigned)((m + 255) / 256)),
                                   dim3(256), 0, stream, m, k, sa.cnt, seen, sa.cand, sa.nx32, S2, cap, kbest, nq, qa,
                                   c, tau, thr, status);
            else
                hipLaunchKernelGGL(k_merge, dim3((unsigned)((m + 3) / 4)), dim3(256), 0, stream, m, k, sa.cnt,
                                   seen, sa.cand, sa.nx32, S2, cap, kbest, nq, qa, c, tau, thr, status, kbest_row, X,
                                   Q, d);
            MLK_CHECK_HIP(hipGetLastError());
            return MLK_OK;
        };
        double prev = 0.0;
        for (size_t r = 0; r < plan.rounds.size(); ++r) {
            const double hi = r + 1 == plan.rounds.size() ? 1.0 : plan.rounds[r];
            sa.off_lo = (int64_t)std::llround(prev * sa.split_rows);
            sa.off_hi = r + 1 == plan.rounds.size() ? sa.split_rows : (int64_t)std::llround(hi * sa.split_rows);
            prev = hi;
            if (sa.off_hi <= sa.off_lo) continue;
            hipEvent_t ea = nullptr, eb = nullptr;
            if (timed) {
                MLK_CHECK_HIP(events->make(&ea));
                MLK_CHECK_HIP(events->make(&eb));
                if (round_events) {
                    round_events->push_back(ea);
                    round_events->push_back(eb);
                }
                // the launch's own start and end times
                hipExtLaunchKernelGGL(screen, dim3(blocks), dim3(256), 0, stream, ea, eb, 0, sa);
            } else {
                hipLaunchKernelGGL(screen, dim3(blocks), dim3(256), 0, stream, sa);
            }
            MLK_CHECK_HIP(hipGetLastError());
            ++st.rounds;
            if (r + 1 < plan.rounds.size()) MLK_TRY(merge());
        }
        // final overflow check (the last round has no merge)
        MLK_TRY(merge());

        ra.X = X; ra.n = n; ra.d = d; ra.Q = Q; ra.m = m; ra.k = k;
        ra.cnt = sa.cnt; ra.cand = sa.cand; ra.nx32 = sa.nx32; ra.nq = nq;
        ra.S = S2; ra.cap = cap;
        ra.tau = tau; ra.c = c;
        ra.unsafe = mr.unsafe.as<int>(); ra.n_unsafe = (int)mr.n_unsafe;
        ra.mode = 0; ra.self_off = q_offset; ra.exclude = exclude; ra.status = status;
        ra.out_idx = idx; ra.out_dist = dist; ra.squared = (flags & MLK_SQUARED_DIST) != 0;
        int* qstats;
        MLK_TRY(mem.get(19, (size_t)m * 3 * sizeof(int), &qstats));
        ra.qstats = qstats;  // written for every query by k_rank
        const bool vec = d % 4 == 0 && ((uintptr_t)Q % 16) == 0 && ((uintptr_t)X % 16) == 0;
        if (k == 1 && mr.n_unsafe == 0 && vec && d <= 256) {
            hipLaunchKernelGGL(d <= 64    ? k_rank1_group<2>
                               : d <= 96  ? k_rank1_group<3>
                               : d <= 128 ? k_rank1_group<4>
                                          : k_rank1_group<8>,
                               dim3((unsigned)((m + 31) / 32)), dim3(256), 0, stream, ra, m);
            hipLaunchKernelGGL(k_rank1_lane<true>, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, stream, ra, m, 1);
        } else if (k == 1 && mr.n_unsafe == 0)
            hipLaunchKernelGGL(vec ? k_rank1_lane<true> : k_rank1_lane<false>, dim3((unsigned)((m + 255) / 256)),
                               dim3(256), 0, stream, ra, m, 0);
        else {
            hipLaunchKernelGGL(k_rank, dim3((unsigned)((m + 3) / 4)), dim3(256), 0, stream, ra, m);
            big_rank = true;
        }
        MLK_CHECK_HIP(hipGetLastError());
        rank_stats = ra.qstats;
    } else {
        hipLaunchKernelGGL(k_fill_int, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, stream, status, m,
                           ST_FALLBACK);
        MLK_CHECK_HIP(hipGetLastError());
    }

    // queries without a certificate -> exact fallback
    int* fq;
    MLK_TRY(mem.get(11, (size_t)m * sizeof(int), &fq));
    const dim3 gs((unsigned)std::min<int64_t>(1024, (m + 255) / 256));
    // the end of the timed chain, unless regrouped results are moved after it
    // (then the move's launch carries the end event instead)
    if (timed && !qperm) {
        MLK_CHECK_HIP(events->make(&call.ev_status));
        hipExtLaunchKernelGGL(k_finish_status, gs, dim3(256), 0, stream, nullptr, call.ev_status, 0, m, status, fq,
                              dflag + 2, rank_stats, (unsigned long long*)(dflag + 4));
    } else {
        hipLaunchKernelGGL(k_finish_status, gs, dim3(256), 0, stream, m, status, fq, dflag + 2, rank_stats,
                           (unsigned long long*)(dflag + 4));
    }
    MLK_CHECK_HIP(hipGetLastError());
    call.m = m; call.n = n; call.q_offset = q_offset; call.d = d; call.k = k; call.prec = prec;
    call.flags = flags; call.exclude = exclude; call.Q = Q; call.X = X; call.dist = dist; call.idx = idx;
    call.status = status; call.fq = fq; call.dflag = dflag; call.big_rank = big_rank; call.ra = ra; call.c = c;
    call.rounds = st.rounds; call.splits = S; call.capacity = cap;
    call.qperm = qperm; call.user_idx = idx_caller; call.user_dist = dist_caller;
    return MLK_OK;
}

// Regrouped queries: result row p (by screen position) to the caller's row qperm[p].
__global__ void k_unpermute_results(const int* __restrict__ qperm, int64_t m, int k, const int64_t* __restrict__ idx,
                                    const double* __restrict__ dist, int64_t* __restrict__ idx_out,
                                    double* __restrict__ dist_out) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= m * k) return;
    const int64_t p = t / k, j = t - p * k;
    const int64_t o = (int64_t)qperm[p] * k + j;
    idx_out[o] = idx[t];
    if (dist) dist_out[o] = dist[t];
}

// Completion of an enqueued call: one read-back of the fallback count and the
// statistics, then the exact fallback of the queries without a certificate.
static int knn_complete(mlk_index* index, const KnnCall& call, hipStream_t stream, CallMem& mem, Events* events,
                        const std::vector<hipEvent_t>& round_events, hipEvent_t ev0) {
    mlk_stats& st = stats();
    const bool fixed = mem.fixed();
    const int64_t m = call.m, n = call.n, q_offset = call.q_offset;
    const int d = call.d, k = call.k;
    const uint32_t flags = call.flags;
    const bool exclude = call.exclude;
    const float* Q = call.Q;
    const float* X = call.X;
    double* dist = call.dist;
    int64_t* idx = call.idx;
    int* status = call.status;
    int* fq = call.fq;
    unsigned* dflag = call.dflag;
    const bool big_rank = call.big_rank;
    const RankArgs& ra = call.ra;
    st.screen_precision = (uint32_t)call.prec << 4;
    st.rounds = call.rounds;
    st.splits = call.splits;
    st.capacity = call.capacity;
    (void)index;
    // regrouped queries: results to the caller's rows, enqueued before the
    // read-back so that the common call (no fallback) ends with it; a fallback
    // rewrites some rows by position and moves them again below
    // the end of the chain before the read-back (timed calls)
    hipEvent_t ev_end = call.ev_status;
    auto unpermute = [&](hipEvent_t stop) -> int {
        if (!call.qperm) return MLK_OK;
        const dim3 g((unsigned)((m * k + 255) / 256));
        if (stop)
            hipExtLaunchKernelGGL(k_unpermute_results, g, dim3(256), 0, stream, nullptr, stop, 0, call.qperm, m, k,
                                  idx, dist, call.user_idx, call.user_dist);
        else
            hipLaunchKernelGGL(k_unpermute_results, g, dim3(256), 0, stream, call.qperm, m, k, idx, dist,
                               call.user_idx, call.user_dist);
        MLK_CHECK_HIP(hipGetLastError());
        return MLK_OK;
    };
    if (events && call.qperm) MLK_CHECK_HIP(events->make(&ev_end));
    MLK_TRY(unpermute(events ? ev_end : nullptr));
    // one read-back: the fallback count and the statistics sums
    struct {
        unsigned nf, pad;
        unsigned long long cnts[3];
    } words{};
    static_assert(sizeof(words) == 32, "flag words 2..10");
    MLK_CHECK_HIP(hipMemcpyAsync(&words, dflag + 2, sizeof(words), hipMemcpyDeviceToHost, stream));
    MLK_CHECK_HIP(hipStreamSynchronize(stream));
    unsigned nf = words.nf;
    const bool rewritten = nf != 0;  // some rows get ranked again below (by position)
    std::vector<int> fq_host;
    if (nf) {
        fq_host.resize(nf);
        MLK_CHECK_HIP(hipMemcpyAsync(fq_host.data(), fq, nf * sizeof(int), hipMemcpyDeviceToHost, stream));
        MLK_CHECK_HIP(hipStreamSynchronize(stream));
    }
    if (nf && big_rank) {
        // queries with more survivors than one wave's LDS list: rank them again
        // with the larger list, then keep only what still needs the exact scan
        hipLaunchKernelGGL(k_rank_big, dim3(nf), dim3(64), 0, stream, ra, fq, (int64_t)nf,
                           (unsigned long long*)(dflag + 4));
        MLK_CHECK_HIP(hipGetLastError());
        std::vector<int> st_host(m);
        MLK_CHECK_HIP(hipMemcpyAsync(st_host.data(), status, m * sizeof(int), hipMemcpyDeviceToHost, stream));
        MLK_CHECK_HIP(hipMemcpyAsync(&words, dflag + 2, sizeof(words), hipMemcpyDeviceToHost, stream));
        MLK_CHECK_HIP(hipStreamSynchronize(stream));
        fq_host.erase(std::remove_if(fq_host.begin(), fq_host.end(), [&](int q) { return st_host[q] == 0; }),
                      fq_host.end());
        nf = (unsigned)fq_host.size();
    }
    st.fallback_queries = nf;
    st.candidates = (int64_t)words.cnts[0];
    st.reranked = (int64_t)words.cnts[2];
    if (trial_abandon_fallback(nf, m)) return MLK_OK;  // autotuner trial: rejected, outputs unused
    if (nf) {
        std::sort(fq_host.begin(), fq_host.end());
        MLK_CHECK_HIP(hipMemcpyAsync(fq, fq_host.data(), nf * sizeof(int), hipMemcpyHostToDevice, stream));
        // batches bounded by ~1 GiB of float64 distances + candidate lists (in
        // fixed mode: by what is left of the arena)
        const int64_t per_q = (int64_t)knn_fallback_bytes_per_query(n) + 8 * 256;
        const int64_t budget = fixed ? (int64_t)mem.left() - 16 * 256 : (int64_t(1) << 30);
        int64_t fb = std::min<int64_t>(nf, budget / std::max<int64_t>(per_q, 1));
        if (fb < 1) {
            if (fixed) {
                set_error("the exact fallback needs " + std::to_string(per_q) + " bytes per query; the fixed "
                          "scratch arena has " + std::to_string(mem.left()) + " left");
                return MLK_ECAPACITY;
            }
            fb = 1;
        }
        double* fb_dist;
        int* fb_plist;
        double* fb_ps;
        MLK_TRY(mem.get(12, (size_t)fb * n * sizeof(double), &fb_dist));  // distances
        MLK_TRY(mem.get(13, (size_t)fb * n * sizeof(int), &fb_plist));    // plist
        MLK_TRY(mem.get(14, (size_t)fb * n * sizeof(double), &fb_ps));    // ps
        // per fallback query: count, cut, radix prefix and rank, 256-bin histogram
        const size_t per = sizeof(int) + sizeof(double) + sizeof(unsigned long long) + sizeof(unsigned) + 1024;
        char* wp;
        MLK_TRY(mem.get(15, (size_t)fb * per + 5 * 64, &wp));  // five pieces, each 64-byte aligned
        auto carve = [&](size_t bytes) {
            char* at = wp;
            wp += (bytes + 63) / 64 * 64;
            return at;
        };
        int* pcount = (int*)carve(fb * sizeof(int));
        double* cut = (double*)carve(fb * sizeof(double));
        unsigned long long* rprefix = (unsigned long long*)carve(fb * sizeof(unsigned long long));
        unsigned* rrank = (unsigned*)carve(fb * sizeof(unsigned));
        unsigned* hist = (unsigned*)carve(fb * 1024);
        for (int64_t f0 = 0; f0 < (int64_t)nf; f0 += fb) {
            const int64_t nb = std::min<int64_t>(fb, nf - f0);
            dim3 g2((unsigned)((n + 255) / 256), (unsigned)nb);
            dim3 g16((unsigned)((n + 15) / 16), (unsigned)nb);
            hipLaunchKernelGGL(k_fb_dist, g16, dim3(256), 0, stream, X, n, d, Q, fq + f0, q_offset, (int)exclude,
                               fb_dist, k, hist, rprefix, rrank);
            // about 4096 values per histogram block
            const dim3 gh((unsigned)std::max<int64_t>(1, std::min<int64_t>(256, (n + 4095) / 4096)), (unsigned)nb);
            for (int shift = 56; shift >= 0; shift -= 8) {
                hipLaunchKernelGGL(k_fb_hist, gh, dim3(256), 0, stream, fb_dist, n, shift, rprefix, hist);
                hipLaunchKernelGGL(k_fb_pick, dim3((unsigned)nb), dim3(256), 0, stream, shift, d, hist, rprefix,
                                   rrank, cut);
            }
            MLK_CHECK_HIP(hipMemsetAsync(pcount, 0, nb * sizeof(int), stream));
            hipLaunchKernelGGL(k_fb_collect, g2, dim3(256), 0, stream, fb_dist, n, cut, pcount,
                               fb_plist, fb_ps, n);
            RankArgs fa{};
            fa.X = X; fa.n = n; fa.d = d; fa.Q = Q; fa.m = m; fa.k = k;
            fa.fq = fq + f0; fa.plist = fb_plist; fa.ps = fb_ps; fa.pcount = pcount;
            fa.pstride = n; fa.mode = 1; fa.self_off = q_offset; fa.exclude = exclude; fa.status = status;
            fa.out_idx = idx; fa.out_dist = dist; fa.squared = (flags & MLK_SQUARED_DIST) != 0;
            hipLaunchKernelGGL(k_rank, dim3((unsigned)((nb + 3) / 4)), dim3(256), 0, stream, fa, nb);
            MLK_CHECK_HIP(hipGetLastError());
        }
    }
    if (rewritten) MLK_TRY(unpermute(nullptr));
    if (!events) {
        // no events: the outputs are ordered on the stream; the fallback's host
        // list must outlive its copy
        if (nf) MLK_CHECK_HIP(hipStreamSynchronize(stream));
        return MLK_OK;
    }
    if (nf) {
        // the fallback's kernels end the call: one more event and wait
        MLK_CHECK_HIP(events->make(&ev_end));
        MLK_CHECK_HIP(hipEventRecord(ev_end, stream));
        MLK_CHECK_HIP(hipEventSynchronize(ev_end));
    }  // else every event was reached before the read-back's wait ended
    if (!ev0) ev0 = call.ev_first;
    float ms = 0.0f;
    // screen kernels only: sum of the per-round launch intervals
    st.screen_ms = 0.0;
    for (size_t i = 0; i + 1 < round_events.size(); i += 2)
        if (hipEventElapsedTime(&ms, round_events[i], round_events[i + 1]) == hipSuccess) st.screen_ms += ms;
    if (hipEventElapsedTime(&ms, ev0, ev_end) == hipSuccess) st.total_ms = ms;
    return MLK_OK;
}

// ---------------------------------------------------------------- graph replay
// With MLK_GRAPH_REPLAY a call's device chain (query mirror, start threshold,
// screen rounds and merges, rank, status) is captured once into a HIP graph;
// a later call with the same arguments, plan and buffers replays it with one
// launch and no host round trip before its single read-back. A graph holds
// device addresses, so an entry is used only while every buffer it captured
// is still the one in place.
struct GraphEntry {
    // what the captured chain depends on
    const float* Q = nullptr;
    int64_t m = 0, q_offset = 0;
    int k = 0, device = -1;
    uint32_t flags = 0;
    double* dist = nullptr;
    int64_t* idx = nullptr;
    Plan plan;
    std::vector<void*> buffers;  // every workspace slot and the mirror's buffers at capture time
    hipGraphExec_t exec = nullptr;
    KnnCall call;
    uint64_t last_use = 0;
};
struct GraphCache {
    std::vector<GraphEntry> entries;
    hipStream_t capture_stream = nullptr;
    uint64_t clock = 0;
};
constexpr size_t GRAPH_CACHE_MAX = 8;

static std::vector<void*> index_buffers(const mlk_index* index, int prec) {
    std::vector<void*> v;
    for (const DevBuf& b : index->ws) v.push_back(b.ptr);
    const Mirror& mr = index->mirror[prec];
    for (const DevBuf* b : {&mr.xs, &mr.nx, &mr.nx32, &mr.flags, &mr.unsafe, &mr.bias, &index->X, &index->centre})
        v.push_back(b->ptr);
    return v;
}

static bool same_plan_fields(const Plan& a, const Plan& b) {
    return a.shape.qt == b.shape.qt && a.shape.ns == b.shape.ns && a.splits == b.splits && a.cap == b.cap &&
           a.rounds == b.rounds;
}

void destroy_graphs(mlk_index* index) {
    GraphCache* gc = static_cast<GraphCache*>(index->graphs);
    if (!gc) return;
    for (GraphEntry& e : gc->entries)
        if (e.exec) (void)hipGraphExecDestroy(e.exec);
    if (gc->capture_stream) (void)hipStreamDestroy(gc->capture_stream);
    delete gc;
    index->graphs = nullptr;
}

static int knn_query_graph(mlk_index* index, const float* Q, int64_t m, int k, int64_t q_offset, uint32_t flags,
                           double* dist, int64_t* idx, hipStream_t stream) {
    const int prec = effective_prec((int)((flags & MLK_PREC_MASK) >> 4), index->d);
    Mirror& mr = index->mirror[prec];
    int device = 0;
    MLK_CHECK_HIP(hipGetDevice(&device));
    if (!index->graphs) index->graphs = new GraphCache();
    GraphCache& gc = *static_cast<GraphCache*>(index->graphs);
    const Plan plan = resolve_plan(PLAN_KNN, prec, mr.dp, index->n, m, k);
    CallMem mem;
    mem.ws = index->ws;
    GraphEntry* hit = nullptr;
    for (GraphEntry& e : gc.entries)
        if (e.Q == Q && e.m == m && e.k == k && e.q_offset == q_offset && e.flags == flags && e.dist == dist &&
            e.idx == idx && e.device == device && same_plan_fields(e.plan, plan) &&
            e.buffers == index_buffers(index, prec)) {
            hit = &e;
            break;
        }
    if (!hit) {
        // First call of this shape: an ordinary call answers it and sizes the
        // workspace (nothing is allocated while capturing); then the chain is
        // captured, with the bias its thresholds assume (queries taken as
        // inexact) built beforehand.
        {
            mlk_stats& st0 = stats();
            st0 = mlk_stats{};
            st0.queries = m;
            KnnCall probe;
            MLK_TRY(knn_enqueue(index, Q, m, k, q_offset, flags, dist, idx, stream, mem, false, probe, nullptr,
                                nullptr));
            MLK_TRY(knn_complete(index, probe, stream, mem, nullptr, {}, nullptr));
        }
        const mlk_stats answered = stats();
        const Consts c = make_consts(prec, mr.dp, index->d, mr.exact, false, mr.scale);
        MLK_TRY(prepare_bias(index, prec, c, stream));
        if (!gc.capture_stream) MLK_CHECK_HIP(hipStreamCreateWithFlags(&gc.capture_stream, hipStreamNonBlocking));
        MLK_CHECK_HIP(hipStreamSynchronize(stream));
        GraphEntry e;
        e.Q = Q; e.m = m; e.k = k; e.q_offset = q_offset; e.flags = flags; e.dist = dist; e.idx = idx;
        e.device = device; e.plan = plan;
        MLK_CHECK_HIP(hipStreamBeginCapture(gc.capture_stream, hipStreamCaptureModeRelaxed));
        const int rc = knn_enqueue(index, Q, m, k, q_offset, flags, dist, idx, gc.capture_stream, mem, true, e.call,
                                   nullptr, nullptr);
        hipGraph_t graph = nullptr;
        const hipError_t ec = hipStreamEndCapture(gc.capture_stream, &graph);
        if (rc != MLK_OK || ec != hipSuccess) {
            if (graph) (void)hipGraphDestroy(graph);
            if (rc != MLK_OK) return rc;
            MLK_CHECK_HIP(ec);
        }
        const hipError_t ei = hipGraphInstantiate(&e.exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        MLK_CHECK_HIP(ei);
        e.buffers = index_buffers(index, prec);
        if (gc.entries.size() >= GRAPH_CACHE_MAX) {
            auto oldest = std::min_element(gc.entries.begin(), gc.entries.end(),
                                           [](const GraphEntry& a, const GraphEntry& b) { return a.last_use < b.last_use; });
            (void)hipGraphExecDestroy(oldest->exec);
            gc.entries.erase(oldest);
        }
        e.last_use = ++gc.clock;
        gc.entries.push_back(std::move(e));
        stats() = answered;
        return MLK_OK;
    }
    hit->last_use = ++gc.clock;
    // the index's bias must be the one the graph's thresholds assume
    MLK_TRY(prepare_bias(index, prec, hit->call.c, stream));
    mlk_stats& st = stats();
    st = mlk_stats{};
    st.queries = m;
    MLK_CHECK_HIP(hipGraphLaunch(hit->exec, stream));
    return knn_complete(index, hit->call, stream, mem, nullptr, {}, nullptr);
}

int knn_query_impl(mlk_index* index, const float* Q, int64_t m, int k, int64_t q_offset, uint32_t flags,
                   double* dist, int64_t* idx, hipStream_t stream, CallMem* mem_in) {
    mlk_stats& st = stats();
    st = mlk_stats{};
    st.queries = m;
    if (m == 0) return MLK_OK;
    const int prec = effective_prec((int)((flags & MLK_PREC_MASK) >> 4), index->d);
    // graph replay: allocating calls without test hooks on a prepared index
    if ((flags & MLK_GRAPH_REPLAY) && !mem_in && index->mirror[prec].ready && hooks_snapshot().hooks == 0)
        return knn_query_graph(index, Q, m, k, q_offset, flags, dist, idx, stream);
    CallMem own;
    own.ws = index->ws;
    CallMem& mem = mem_in ? *mem_in : own;
    KnnCall call;
    if (!mem.fixed() && m > KNN_CHUNK_MIN) {
        // Candidate lists beyond KNN_LIST_BUDGET bytes: the queries go in chunks
        // (each with its own plan), which bounds the workspace instead of growing
        // it to the whole call. The plan that sizes the lists needs the screen
        // mirror (its width and unsafe rows), so a first call prepares it here.
        MLK_TRY(prepare_mirror(index, prec, stream));
        const size_t lists = knn_list_bytes(index, prec, m, k);
        if (lists > KNN_LIST_BUDGET) {
            const int64_t parts = (int64_t)((lists + KNN_LIST_BUDGET - 1) / KNN_LIST_BUDGET);
            const int64_t chunk = std::max<int64_t>(KNN_CHUNK_MIN, (m + parts - 1) / parts / 512 * 512);
            mlk_stats total{};
            for (int64_t off = 0; off < m; off += chunk) {
                const int64_t mc = std::min<int64_t>(chunk, m - off);
                MLK_TRY(knn_query_impl(index, Q + off * index->d, mc, k, q_offset + off, flags & ~MLK_GRAPH_REPLAY,
                                       dist ? dist + off * k : nullptr, idx + off * k, stream, nullptr));
                const mlk_stats& sc = stats();
                total.queries += sc.queries;
                total.candidates += sc.candidates;
                total.reranked += sc.reranked;
                total.fallback_queries += sc.fallback_queries;
                total.rounds += sc.rounds;
                total.splits = sc.splits;
                total.capacity = sc.capacity;
                total.screen_precision = sc.screen_precision;
                total.screen_ms += sc.screen_ms;
                total.total_ms += sc.total_ms;
            }
            stats() = total;
            return MLK_OK;
        }
    }
    if (mem.fixed()) {
        MLK_TRY(knn_enqueue(index, Q, m, k, q_offset, flags, dist, idx, stream, mem, false, call, nullptr, nullptr));
        return knn_complete(index, call, stream, mem, nullptr, {}, nullptr);
    }
    Events events;  // every event of this call
    std::vector<hipEvent_t> round_events;
    MLK_TRY(knn_enqueue(index, Q, m, k, q_offset, flags, dist, idx, stream, mem, false, call, &events, &round_events));
    return knn_complete(index, call, stream, mem, &events, round_events, nullptr);
}

}  // namespace mlk
