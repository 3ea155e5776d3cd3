// Exact radius neighbourhoods on gfx950.
//
// Membership contract (reference radius_contract.py:1-25, 129-217;
// _radius_common.h:1-18, 337-357): row x belongs to query q iff the exact
// S(q, x) <= R2 = fl64(r*r). The screen keeps every row whose certified lower
// bound L(x) <= R2 (MFMA tile + threshold, hits staged per wave in LDS and
// appended to one pair buffer, so heavy-tailed neighbourhoods need no
// per-query capacity). Candidate pairs are sorted by (query, row); each pair
// gets the canonical float64 s, is classified against the outward-rounded
// band [f64_in, f64_out] and the band is resolved exactly against
// floor(R2 * 2^298). Members are compacted into CSR rows in ascending index
// order, or stably re-sorted by distance for sort_results. Distances are
// sqrt(min(s, R2)). This replaces the reference's per-tile device screen with
// host finishing (radius_metal_v1.py:257-611, _radius_cpu_native.cpp:129-200)
// by a device-only pipeline.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

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
Consts make_consts(int prec, int dp, int d, bool rows_exact, bool queries_exact, double scale);

static void band(int d, double R2, double* f_in, double* f_out) {
    // restates thresholds() of radius_contract.py:129-153
    if (std::isinf(R2)) { *f_in = *f_out = INFINITY; return; }
    if (R2 == 0.0) { *f_in = *f_out = 0.0; return; }
    const double u = 0x1p-53;
    const double g = (d + 2) * u / (1.0 - (d + 2) * u);
    auto dn = [](double v) { return std::nextafter(v, -INFINITY); };
    auto up = [](double v) { return std::nextafter(v, INFINITY); };
    *f_in = std::max(dn(dn(R2 * dn(1.0 - 2.0 * g))), 0.0);
    *f_out = up(up(R2 * up(1.0 + 2.0 * g)));
}

// floor(R2 * 2^298) in 20 little-endian 32-bit chunks; false when R2 >= 2^342.
static bool r2_key(double R2, uint32_t key[SACC_LIMBS]) {
    std::memset(key, 0, sizeof(uint32_t) * SACC_LIMBS);
    if (!(R2 < 0x1p342)) return false;
    if (!(R2 > 0)) return true;
    int e;
    const double fr = std::frexp(R2, &e);
    const uint64_t mant = (uint64_t)std::ldexp(fr, 53);  // R2 = mant * 2^(e-53)
    const int shift = e - 53 + 298;
    if (shift >= 0) {
        const unsigned limb = (unsigned)shift / 32, bit = (unsigned)shift % 32;
        unsigned __int128 wide = (unsigned __int128)mant << bit;
        for (unsigned i = limb; i < (unsigned)SACC_LIMBS && wide; ++i) {
            key[i] = (uint32_t)wide;
            wide >>= 32;
        }
    } else if (-shift < 64) {
        const uint64_t v = mant >> (unsigned)(-shift);
        key[0] = (uint32_t)v;
        key[1] = (uint32_t)(v >> 32);
    }
    return true;
}

struct RadiusConsts {
    double R2, f_in, f_out;
    uint32_t key[SACC_LIMBS];
};

__global__ void k_radius_thr(int64_t m, int64_t m_pad, const double* __restrict__ nq, const double* __restrict__ nt,
                             const double* __restrict__ d2, const uint8_t* __restrict__ qflags, double tau, Consts c,
                             float* __restrict__ thr) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= m_pad) return;
    if (q >= m) { thr[q] = INFINITY; return; }
    // Unsafe query (screen values zero): every safe row is a candidate, and
    // THR_FLOOR passes exactly the safe rows; k_unsafe_pairs adds the unsafe ones once.
    thr[q] = qflags[q] ? thr_from_tau(tau, screen_qa(nq[q], nt[q], d2[q], c), c) : THR_FLOOR;
}

// Unsafe database rows never hit the screen; they are candidates of every query.
// Pairs name queries by screen position; qmap (nullable) maps a position to the
// query's index.
__global__ void k_unsafe_pairs(int64_t m, const int* __restrict__ unsafe, int n_unsafe, int64_t self_off, int exclude,
                               const int* __restrict__ qmap, int2* __restrict__ pairs,
                               unsigned long long* __restrict__ count, unsigned long long cap) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= m * n_unsafe) return;
    const int64_t q = t / n_unsafe;
    const int row = unsafe[t % n_unsafe];
    if (exclude && row == self_off + (qmap ? (int64_t)qmap[q] : q)) return;
    const unsigned long long at = atomicAdd(count, 1ull);
    if (at < cap) pairs[at] = int2{(int)q, row};
}

// Pair keys (query << rb) | row, rb = bits of the row count, with the query's
// index (qmap of its screen position when the queries were regrouped): sorting
// them orders pairs by (query, row) with as few key bits as the shape needs.
__global__ void k_pair_keys(const int2* __restrict__ pairs, int64_t np, int rb, const int* __restrict__ qmap,
                            unsigned long long* __restrict__ keys) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    const int2 p = pairs[i];
    const unsigned q = qmap ? (unsigned)qmap[p.x] : (unsigned)p.x;
    keys[i] = ((unsigned long long)q << rb) | (unsigned)p.y;
}

// Membership and distance of every sorted candidate pair.
// VEC: d % 4 == 0 and 16-byte aligned rows (four features per load).
template <bool VEC>
__global__ void k_classify(const unsigned long long* __restrict__ keys, int64_t np, int rb, const float* __restrict__ X,
                           const float* __restrict__ Q, int d, RadiusConsts rc, uint8_t* __restrict__ member,
                           double* __restrict__ pdist) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    const unsigned long long key = keys[i];
    const int64_t q = (int64_t)(key >> rb);
    const int64_t row = (int64_t)(key & ((1ull << rb) - 1));
    const float* qv = Q + q * d;
    const float* xv = X + row * d;
    const double s = VEC ? canonical_sq4(qv, xv, d) : canonical_sq(qv, xv, d);
    bool in;
    if (s <= rc.f_in) in = true;
    else if (!(s <= rc.f_out)) in = false;
    else in = exact_cmp_key(qv, xv, d, rc.key) <= 0;
    member[i] = in;
    pdist[i] = sqrt(s < rc.R2 ? s : rc.R2);
}

// k_classify for d % 4 == 0 and 16-byte aligned rows, with the database rows
// of the block's 256 pairs staged through LDS in 32-feature chunks: eight
// lanes load one row chunk (128 contiguous bytes) instead of one lane reading
// a whole row in eight scattered pieces, a quarter of the cache-line
// requests. Each thread then adds its pair's chunk from LDS in feature order,
// so the sum is bitwise canonical_sq's. The query rows (shared by consecutive
// pairs) are read from memory.
constexpr int CLS_CHUNK = 32;

__global__ __launch_bounds__(256) void k_classify_staged(const unsigned long long* __restrict__ keys, int64_t np,
                                                         int rb, const float* __restrict__ X,
                                                         const float* __restrict__ Q, int d, RadiusConsts rc,
                                                         uint8_t* __restrict__ member, double* __restrict__ pdist) {
    // rows padded to 36 floats: eight lanes reading 16 bytes each cover 32 banks once
    __shared__ __attribute__((aligned(16))) float sx[256][CLS_CHUNK + 4];
    __shared__ int64_t srow[256];
    const int t = threadIdx.x;
    const int64_t i0 = (int64_t)blockIdx.x * 256, i = i0 + t;
    const bool live = i < np;
    const unsigned long long key = live ? keys[i] : 0ull;
    const int64_t q = (int64_t)(key >> rb);
    const int64_t row = (int64_t)(key & ((1ull << rb) - 1));
    srow[t] = row;
    const float* qv = Q + q * d;
    double s = 0.0;
    for (int c0 = 0; c0 < d; c0 += CLS_CHUNK) {
        __syncthreads();  // srow written / the previous chunk consumed
#pragma unroll
        for (int k = 0; k < CLS_CHUNK / 4; ++k) {
            const int idx = t + 256 * k;
            const int r = idx / (CLS_CHUNK / 4), part = idx % (CLS_CHUNK / 4);
            const int f = c0 + 4 * part;
            float4 v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (i0 + r < np && f < d) v = *reinterpret_cast<const float4*>(X + srow[r] * d + f);
            *reinterpret_cast<float4*>(&sx[r][4 * part]) = v;
        }
        __syncthreads();
        const int nf = d - c0 < CLS_CHUNK ? d - c0 : CLS_CHUNK;
        for (int e = 0; e < nf; e += 4) {
            const float4 a = *reinterpret_cast<const float4*>(qv + c0 + e);
            const float4 b = *reinterpret_cast<const float4*>(&sx[t][e]);
            double u = (double)a.x - (double)b.x;
            s = add_rn(s, mul_rn(u, u));
            u = (double)a.y - (double)b.y;
            s = add_rn(s, mul_rn(u, u));
            u = (double)a.z - (double)b.z;
            s = add_rn(s, mul_rn(u, u));
            u = (double)a.w - (double)b.w;
            s = add_rn(s, mul_rn(u, u));
        }
    }
    if (!live) return;
    const float* xv = X + row * d;
    bool in;
    if (s <= rc.f_in) in = true;
    else if (!(s <= rc.f_out)) in = false;
    else in = exact_cmp_key(qv, xv, d, rc.key) <= 0;
    member[i] = in;
    pdist[i] = sqrt(s < rc.R2 ? s : rc.R2);
}

// Stable compaction of the members (pairs in (query, row) order): blocks of
// MEMBER_TILE pairs. Pass 1 counts each block's members (16 consecutive flags
// per thread); a scan of the counts gives each block's first output position;
// pass 2 walks the block's pairs in rounds of 256 and writes every member's
// row, distance and query at its position.
constexpr int MEMBER_PER_THREAD = 16;
constexpr int MEMBER_TILE = 256 * MEMBER_PER_THREAD;

__device__ __forceinline__ int member_count16(const uint8_t* __restrict__ member, int64_t np, int64_t at) {
    int c = 0;
    if (at + MEMBER_PER_THREAD <= np && (at & 15) == 0) {
        const uint4 v = *reinterpret_cast<const uint4*>(member + at);
        // flags are 0 or 1: the byte sum of each word is its popcount
        for (uint32_t w : {v.x, v.y, v.z, v.w}) c += __popc(w);
    } else {
        for (int j = 0; j < MEMBER_PER_THREAD; ++j)
            if (at + j < np) c += member[at + j];
    }
    return c;
}

__global__ __launch_bounds__(256) void k_member_count(const uint8_t* __restrict__ member, int64_t np,
                                                      unsigned long long* __restrict__ bcount) {
    __shared__ int part[4];
    const int64_t at = (int64_t)blockIdx.x * MEMBER_TILE + (int64_t)threadIdx.x * MEMBER_PER_THREAD;
    int c = member_count16(member, np, at);
    for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) bcount[blockIdx.x] = (unsigned long long)(part[0] + part[1] + part[2] + part[3]);
}

__global__ __launch_bounds__(256) void k_member_scatter(const uint8_t* __restrict__ member,
                                                        const unsigned long long* __restrict__ keys,
                                                        const double* __restrict__ pdist, int64_t np, int rb,
                                                        const unsigned long long* __restrict__ bend,
                                                        const unsigned long long* __restrict__ bcount,
                                                        int64_t* __restrict__ idx, double* __restrict__ dist,
                                                        int* __restrict__ mq) {
    // MEMBER_PER_THREAD rounds of 256 consecutive pairs: in each, a member's
    // position is the block's running count plus the members before it in the
    // round (its wave's earlier lanes, then the earlier waves), so consecutive
    // members go to consecutive positions (coalesced reads and writes)
    __shared__ int wcount[MEMBER_PER_THREAD][4];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t base = (int64_t)blockIdx.x * MEMBER_TILE;
    bool flag[MEMBER_PER_THREAD];
    int rank[MEMBER_PER_THREAD];
#pragma unroll
    for (int j = 0; j < MEMBER_PER_THREAD; ++j) {
        const int64_t i = base + (int64_t)j * 256 + threadIdx.x;
        flag[j] = i < np && member[i];
        const unsigned long long mask = ballot(flag[j]);
        rank[j] = __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0));
        if (lane == 0) wcount[j][wave] = __popcll(mask);
    }
    __syncthreads();
    int64_t run = (int64_t)(bend[blockIdx.x] - bcount[blockIdx.x]);
    const uint64_t row_mask = (1ull << rb) - 1;
#pragma unroll
    for (int j = 0; j < MEMBER_PER_THREAD; ++j) {
        int before = 0, total = 0;
#pragma unroll
        for (int w = 0; w < 4; ++w) {
            const int c = wcount[j][w];
            before += w < wave ? c : 0;
            total += c;
        }
        if (flag[j]) {
            const int64_t i = base + (int64_t)j * 256 + threadIdx.x;
            const int64_t pos = run + before + rank[j];
            const unsigned long long key = keys[i];
            idx[pos] = (int64_t)(key & row_mask);
            if (dist) dist[pos] = pdist[i];
            mq[pos] = (int)(key >> rb);
        }
        run += total;
    }
}

// CSR row pointers from the members' queries (non-decreasing): indptr[q] is
// the first member whose query is >= q (binary search; indptr[m] = nnz).
__global__ void k_indptr_search(const int* __restrict__ mq, int64_t nnz, int64_t m, int64_t* __restrict__ indptr) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q > m) return;
    int64_t lo = 0, hi = nnz;
    while (lo < hi) {
        const int64_t mid = lo + ((hi - lo) >> 1);
        if ((int64_t)mq[mid] < q) lo = mid + 1; else hi = mid;
    }
    indptr[q] = lo;
}

// All rows (except self) are members: R2 >= 2^342 or +inf.
__global__ void k_everything(const float* __restrict__ X, int64_t n, int d, const float* __restrict__ Q, int64_t m,
                             int64_t self_off, int exclude, double R2, int64_t* __restrict__ idx,
                             double* __restrict__ dist) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t per = n - (exclude ? 1 : 0);
    if (t >= m * per) return;
    const int64_t q = t / per;
    int64_t row = t % per;
    if (exclude && row >= self_off + q) ++row;
    idx[t] = row;
    if (dist) {
        const double s = canonical_sq(Q + q * d, X + row * d, d);
        dist[t] = sqrt(s < R2 ? s : R2);
    }
}

__global__ void k_everything_indptr(int64_t m, int64_t per, int64_t* __restrict__ indptr) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q <= m) indptr[q] = q * per;
}

static inline unsigned blocks_for(int64_t n, int t = 256) { return (unsigned)((n + t - 1) / t); }

// ------------------------------------------------------ CSR reductions
// Radius votes and means (reference radius_reduce.py:19-73): one thread per
// query row sums its entries in CSR order (ascending index, as the
// reference's np.bincount accumulates them), so the float64 results are
// bitwise those of the reference's reduction. Distance weights are 1/d,
// except on rows holding a zero distance, which use the indicator of zero
// distances (sklearn's _get_weights).
__device__ __forceinline__ bool row_coincident(const double* __restrict__ dist, int64_t b, int64_t e) {
    for (int64_t j = b; j < e; ++j)
        if (dist[j] == 0.0) return true;
    return false;
}

__device__ __forceinline__ double entry_weight(const double* __restrict__ dist, int64_t j, bool coincident) {
    return coincident ? (dist[j] == 0.0 ? 1.0 : 0.0) : 1.0 / dist[j];
}

__global__ void k_radius_votes(const int64_t* __restrict__ indptr, int64_t m, const int64_t* __restrict__ idx,
                               const double* __restrict__ dist, const int32_t* __restrict__ codes, int n_classes,
                               int weighted, double* __restrict__ votes, int64_t* __restrict__ counts) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= m) return;
    const int64_t b = indptr[q], e = indptr[q + 1];
    if (counts) counts[q] = e - b;
    double* v = votes + q * n_classes;
    for (int c = 0; c < n_classes; ++c) v[c] = 0.0;
    const bool coincident = weighted && row_coincident(dist, b, e);
    for (int64_t j = b; j < e; ++j) v[codes[idx[j]]] += weighted ? entry_weight(dist, j, coincident) : 1.0;
}

__global__ void k_radius_mean(const int64_t* __restrict__ indptr, int64_t m, const int64_t* __restrict__ idx,
                              const double* __restrict__ dist, const double* __restrict__ targets, int weighted,
                              double* __restrict__ pred, int64_t* __restrict__ counts) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= m) return;
    const int64_t b = indptr[q], e = indptr[q + 1];
    if (counts) counts[q] = e - b;
    const bool coincident = weighted && row_coincident(dist, b, e);
    double num = 0.0, den = 0.0;
    for (int64_t j = b; j < e; ++j) {
        if (weighted) {
            // product rounded before the sum (no contraction), as the reference
            // forms values * weights before adding
            const double w = entry_weight(dist, j, coincident);
            num = add_rn(num, mul_rn(targets[idx[j]], w));
            den = add_rn(den, w);
        } else {
            num = add_rn(num, targets[idx[j]]);
        }
    }
    if (!weighted) den = (double)(e - b);
    pred[q] = e == b ? NAN : num / den;
}

// Sort keys of the CSR entries by distance: the float32 rounding of each
// (non-negative) distance, a monotone map that never inverts two distances,
// and the entry's position.
__global__ void k_row_sort_keys(const double* __restrict__ dist, int64_t nnz, uint32_t* __restrict__ keys,
                                uint32_t* __restrict__ pos) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nnz) return;
    keys[i] = __float_as_uint(__double2float_rn(dist[i]));
    pos[i] = (uint32_t)i;
}

// First entry of every non-empty row (runs never cross a row boundary).
__global__ void k_row_starts(const int64_t* __restrict__ indptr, int64_t m, uint8_t* __restrict__ start) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < m && indptr[q] < indptr[q + 1]) start[indptr[q]] = 1;
}

__global__ void k_row_gather(const uint32_t* __restrict__ pos, int64_t nnz, const double* __restrict__ dist,
                             const int64_t* __restrict__ idx, double* __restrict__ dist_out,
                             int64_t* __restrict__ idx_out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nnz) return;
    const uint32_t p = pos[i];
    dist_out[i] = dist[p];
    idx_out[i] = idx[p];
}

// Rows of (lo, hi] entries, one block per row (grid: every row; other rows
// leave at once): the row's float32 distance keys and local positions are
// sorted in LDS by a block radix sort, then the block writes the row's
// (distance, index) entries in key order, coalesced, with their keys (for the
// exact repair of equal-key runs). Padding slots carry the largest key and
// sort last.
template <int IPT>
__global__ __launch_bounds__(256) void k_row_sort_block(const int64_t* __restrict__ indptr, int lo, int hi,
                                                        const double* __restrict__ dist,
                                                        const int64_t* __restrict__ idx, double* __restrict__ dist_out,
                                                        int64_t* __restrict__ idx_out, uint32_t* __restrict__ keys_out) {
    using Sort = hipcub::BlockRadixSort<uint32_t, 256, IPT, uint32_t>;
    __shared__ typename Sort::TempStorage tmp;
    const int64_t q = blockIdx.x;
    const int64_t b = indptr[q];
    const int64_t L = indptr[q + 1] - b;
    if (L <= lo || L > hi) return;  // uniform over the block
    uint32_t key[IPT], pos[IPT];
#pragma unroll
    for (int i = 0; i < IPT; ++i) {
        const int j = i * 256 + (int)threadIdx.x;
        key[i] = j < L ? __float_as_uint(__double2float_rn(dist[b + j])) : 0xffffffffu;
        pos[i] = (uint32_t)j;
    }
    Sort(tmp).SortBlockedToStriped(key, pos);
#pragma unroll
    for (int i = 0; i < IPT; ++i) {
        const int r = i * 256 + (int)threadIdx.x;  // rank of this slot
        if (r < L) {
            const int64_t p = b + pos[i];
            dist_out[b + r] = dist[p];
            idx_out[b + r] = idx[p];
            keys_out[b + r] = key[i];
        }
    }
}

// Rows of more than `above` entries, listed for the device-wide segmented sort.
__global__ void k_row_list_big(const int64_t* __restrict__ indptr, int64_t m, int64_t above, int* __restrict__ rows,
                               unsigned* __restrict__ count, int64_t* __restrict__ seg_begin,
                               int64_t* __restrict__ seg_end) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= m) return;
    const int64_t b = indptr[q], e = indptr[q + 1];
    if (e - b <= above) return;
    const unsigned at = atomicAdd(count, 1u);
    rows[at] = (int)q;
    seg_begin[at] = b;
    seg_end[at] = e;
}

// Entries of the listed rows in the order of their sorted positions.
__global__ __launch_bounds__(256) void k_row_gather_rows(const int* __restrict__ rows, const int64_t* __restrict__ indptr,
                                                         const uint32_t* __restrict__ pos,
                                                         const double* __restrict__ dist,
                                                         const int64_t* __restrict__ idx, double* __restrict__ dist_out,
                                                         int64_t* __restrict__ idx_out) {
    const int64_t q = rows[blockIdx.x];
    for (int64_t i = indptr[q] + threadIdx.x; i < indptr[q + 1]; i += 256) {
        const uint32_t p = pos[i];
        dist_out[i] = dist[p];
        idx_out[i] = idx[p];
    }
}

__device__ __forceinline__ bool dist_idx_less(double da, int64_t ia, double db, int64_t ib) {
    return da < db || (da == db && ia < ib);
}

// Entries of a row whose float32 keys are equal stay in their input order
// after the stable key sort; the thread at the start of each such run orders
// it by (distance, index): insertion for short runs, an in-place heap sort for
// long ones (many distinct distances inside one float32 step).
__global__ void k_row_fix_runs(const uint32_t* __restrict__ keys, const uint8_t* __restrict__ start, int64_t nnz,
                               double* __restrict__ dist, int64_t* __restrict__ idx) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nnz) return;
    if (!(i == 0 || start[i] || keys[i] != keys[i - 1])) return;
    int64_t e = i + 1;
    while (e < nnz && !start[e] && keys[e] == keys[i]) ++e;
    const int64_t L = e - i;
    if (L < 2) return;
    double* D = dist + i;
    int64_t* I = idx + i;
    if (L <= 32) {
        for (int64_t a = 1; a < L; ++a) {
            const double dv = D[a];
            const int64_t iv = I[a];
            int64_t b = a - 1;
            while (b >= 0 && dist_idx_less(dv, iv, D[b], I[b])) {
                D[b + 1] = D[b];
                I[b + 1] = I[b];
                --b;
            }
            D[b + 1] = dv;
            I[b + 1] = iv;
        }
        return;
    }
    auto sift = [&](int64_t root, int64_t end) {
        while (2 * root + 1 < end) {
            int64_t child = 2 * root + 1;
            if (child + 1 < end && dist_idx_less(D[child], I[child], D[child + 1], I[child + 1])) ++child;
            if (!dist_idx_less(D[root], I[root], D[child], I[child])) return;
            const double td = D[root]; D[root] = D[child]; D[child] = td;
            const int64_t ti = I[root]; I[root] = I[child]; I[child] = ti;
            root = child;
        }
    };
    for (int64_t r = L / 2 - 1; r >= 0; --r) sift(r, L);
    for (int64_t end2 = L - 1; end2 > 0; --end2) {
        const double td = D[0]; D[0] = D[end2]; D[end2] = td;
        const int64_t ti = I[0]; I[0] = I[end2]; I[end2] = ti;
        sift(0, end2);
    }
}

// The library sorts and selects take int counts (and the row sort 32-bit
// positions): a count above INT32_MAX (or the test hook's lower limit) is
// refused with MLK_ECAPACITY instead of sorting a wrapped count.
static int check_sort_count(int64_t count, const char* what) {
    int64_t limit = INT32_MAX;
    const TestHooks hk = hooks_snapshot();
    if (hk.hooks & MLK_HOOK_SORT_LIMIT) limit = std::min<int64_t>(limit, std::max<int64_t>(hk.param, 0));
    if (count > limit) {
        set_error(std::string("radius search: ") + what + " (" + std::to_string(count) +
                  ") exceeds the one-pass sort limit (" + std::to_string(limit) +
                  "); split the queries into smaller calls");
        return MLK_ECAPACITY;
    }
    return MLK_OK;
}

// Every CSR row ordered by (distance, index). Rows are first ordered by 32-bit
// keys (the float32 rounding of the distance), up to runs of equal keys, which
// are then ordered exactly (so neither the rows' input order nor the key
// sort's stability matters). Rows of up to ROW_SORT_BLOCK entries are sorted
// in LDS, one block each (two block sizes); longer rows (few: a query inside a
// dense clump) go through the device-wide segmented radix sort on (key,
// position) and a gather. Temporaries live in the index workspace (ws[14] sort
// scratch, ws[10..11] keys, ws[21..22] positions, ws[23] row starts, ws[19..20]
// sorted distances and indices, ws[24] the long rows); the sorted arrays are
// swapped into the result.
constexpr int ROW_SORT_SMALL = 256 * 4;   // rows sorted by the 4-items-per-thread block
constexpr int ROW_SORT_BLOCK = 256 * 32;  // longest row sorted by one block
static int sort_rows_by_distance(mlk_radius_result* res, const int64_t* indptr, int64_t m, int64_t nnz, DevBuf* ws,
                                 hipStream_t stream) {
    if (!nnz) return MLK_OK;
    MLK_TRY(check_sort_count(nnz, "result entries"));
    MLK_TRY(check_sort_count(m, "query rows"));
    DevBuf& tmp = ws[14];
    DevBuf& dist_out = ws[19];
    DevBuf& idx_out = ws[20];
    MLK_TRY(ws[11].reserve((size_t)nnz * 4));
    MLK_TRY(ws[23].reserve((size_t)nnz));
    MLK_TRY(dist_out.reserve((size_t)nnz * 8));
    MLK_TRY(idx_out.reserve((size_t)nnz * 8));
    MLK_TRY(ws[24].reserve((size_t)m * (sizeof(int) + 2 * sizeof(int64_t)) + 64));
    uint32_t* keys_out = ws[11].as<uint32_t>();
    uint8_t* start = ws[23].as<uint8_t>();
    const double* dist = res->dist.as<double>();
    const int64_t* idx = res->indices.as<int64_t>();
    double* dout = dist_out.as<double>();
    int64_t* iout = idx_out.as<int64_t>();
    MLK_CHECK_HIP(hipMemsetAsync(start, 0, (size_t)nnz, stream));
    hipLaunchKernelGGL(k_row_starts, dim3(blocks_for(m)), dim3(256), 0, stream, indptr, m, start);
    hipLaunchKernelGGL(k_row_sort_block<4>, dim3((unsigned)m), dim3(256), 0, stream, indptr, 0, ROW_SORT_SMALL, dist,
                       idx, dout, iout, keys_out);
    hipLaunchKernelGGL(k_row_sort_block<32>, dim3((unsigned)m), dim3(256), 0, stream, indptr, ROW_SORT_SMALL,
                       ROW_SORT_BLOCK, dist, idx, dout, iout, keys_out);
    // longer rows: listed on the device, counted on the host
    char* lw = (char*)ws[24].ptr;
    unsigned* nbig_d = (unsigned*)lw;
    int64_t* seg_begin = (int64_t*)(lw + 64);
    int64_t* seg_end = seg_begin + m;
    int* big_rows = (int*)(seg_end + m);
    MLK_CHECK_HIP(hipMemsetAsync(nbig_d, 0, sizeof(unsigned), stream));
    hipLaunchKernelGGL(k_row_list_big, dim3(blocks_for(m)), dim3(256), 0, stream, indptr, m, (int64_t)ROW_SORT_BLOCK,
                       big_rows, nbig_d, seg_begin, seg_end);
    unsigned nbig = 0;
    MLK_CHECK_HIP(hipMemcpyAsync(&nbig, nbig_d, sizeof(unsigned), hipMemcpyDeviceToHost, stream));
    MLK_CHECK_HIP(hipStreamSynchronize(stream));
    if (nbig) {
        MLK_TRY(ws[10].reserve((size_t)nnz * 4));
        MLK_TRY(ws[21].reserve((size_t)nnz * 4));
        MLK_TRY(ws[22].reserve((size_t)nnz * 4));
        uint32_t* keys = ws[10].as<uint32_t>();
        uint32_t* pos = ws[21].as<uint32_t>();
        uint32_t* pos_out = ws[22].as<uint32_t>();
        hipLaunchKernelGGL(k_row_sort_keys, dim3(blocks_for(nnz)), dim3(256), 0, stream, dist, nnz, keys, pos);
        size_t tmp_bytes = 0;
        MLK_CHECK_HIP(hipcub::DeviceSegmentedRadixSort::SortPairs(nullptr, tmp_bytes, keys, keys_out, pos, pos_out,
                                                                  (int)nnz, (int)nbig, seg_begin, seg_end, 0, 31,
                                                                  stream));
        MLK_TRY(tmp.reserve(tmp_bytes + 16));
        MLK_CHECK_HIP(hipcub::DeviceSegmentedRadixSort::SortPairs(tmp.ptr, tmp_bytes, keys, keys_out, pos, pos_out,
                                                                  (int)nnz, (int)nbig, seg_begin, seg_end, 0, 31,
                                                                  stream));
        hipLaunchKernelGGL(k_row_gather_rows, dim3(nbig), dim3(256), 0, stream, big_rows, indptr, pos_out, dist, idx,
                           dout, iout);
    }
    hipLaunchKernelGGL(k_row_fix_runs, dim3(blocks_for(nnz)), dim3(256), 0, stream, keys_out, start, nnz, dout, iout);
    MLK_CHECK_HIP(hipGetLastError());
    std::swap(res->dist.ptr, dist_out.ptr);
    std::swap(res->dist.bytes, dist_out.bytes);
    // the sorted indices become the result's; its unsorted buffer becomes workspace
    std::swap(res->indices.ptr, idx_out.ptr);
    std::swap(res->indices.bytes, idx_out.bytes);
    return MLK_OK;
}

int radius_query_impl(mlk_index* index, const float* Q, int64_t m, double R2, int64_t q_offset, uint32_t flags,
                      int64_t* indptr, int64_t* nnz_out, mlk_radius_result** out, hipStream_t stream) {
    mlk_stats& st = stats();
    st = mlk_stats{};
    st.queries = m;
    mlk_radius_result* res = new (std::nothrow) mlk_radius_result();
    if (!res) { set_error("host allocation failed"); return MLK_ENOMEM; }
    struct Guard {
        mlk_radius_result*& r;
        bool keep = false;
        ~Guard() { if (!keep) delete r; }
    } guard{res};
    const bool sort_rows = (flags & MLK_SORT_RESULTS) != 0;
    res->m = m;
    res->has_dist = (flags & (MLK_NEED_DISTANCE | MLK_SORT_RESULTS)) != 0;
    const int64_t n = index->n;
    const int d = index->d;
    const bool exclude = (flags & MLK_EXCLUDE_SELF) != 0;
    const float* X = index->X.as<float>();

    if (m == 0) {
        MLK_CHECK_HIP(hipMemsetAsync(indptr, 0, sizeof(int64_t), stream));
        MLK_CHECK_HIP(hipStreamSynchronize(stream));
        *nnz_out = 0;
        guard.keep = true;
        *out = res;
        return MLK_OK;
    }
    RadiusConsts rc;
    rc.R2 = R2;
    band(d, R2, &rc.f_in, &rc.f_out);
    if (!r2_key(R2, rc.key)) {
        const int64_t per = n - (exclude ? 1 : 0);
        const int64_t total = m * per;
        if (sort_rows) MLK_TRY(check_sort_count(total, "result entries"));
        MLK_TRY(res->indices.reserve((size_t)std::max<int64_t>(total, 1) * sizeof(int64_t)));
        if (res->has_dist) MLK_TRY(res->dist.reserve((size_t)std::max<int64_t>(total, 1) * sizeof(double)));
        hipLaunchKernelGGL(k_everything, dim3(blocks_for(total)), dim3(256), 0, stream, X, n, d, Q, m, q_offset,
                           (int)exclude, R2, res->indices.as<int64_t>(), res->has_dist ? res->dist.as<double>() : nullptr);
        hipLaunchKernelGGL(k_everything_indptr, dim3(blocks_for(m + 1)), dim3(256), 0, stream, m, per, indptr);
        MLK_CHECK_HIP(hipGetLastError());
        if (sort_rows) MLK_TRY(sort_rows_by_distance(res, indptr, m, total, index->ws, stream));
        MLK_CHECK_HIP(hipStreamSynchronize(stream));
        res->nnz = total;
        *nnz_out = total;
        guard.keep = true;
        *out = res;
        return MLK_OK;
    }

    const int prec = radius_prec((int)((flags & MLK_PREC_MASK) >> 4), d);
    st.screen_precision = (uint32_t)prec << 4;  // the precision actually screened
    MLK_TRY(prepare_mirror(index, prec, stream));
    Mirror& mr = index->mirror[prec];
    const int dp = mr.dp;
    const Plan plan = resolve_plan(PLAN_RADIUS, prec, dp, n, m, 0);
    ScreenFn screen = pick_screen(SCREEN_PAIRS, prec, dp, plan.shape);
    const int qt = plan.shape.qt;
    const int BQ = 4 * qt * 32;
    const int64_t QB = (m + BQ - 1) / BQ;
    const int64_t m_pad = QB * BQ;
    const int S = plan.splits;
    st.splits = S;

    DevBuf* ws = index->ws;
    const size_t eb = elem_bytes(prec);
    MLK_TRY(ws[0].reserve((size_t)m_pad * std::max(dp, 32) * eb));
    MLK_TRY(ws[1].reserve((size_t)m_pad * sizeof(double)));
    MLK_TRY(ws[2].reserve((size_t)m_pad + 64));
    MLK_TRY(ws[4].reserve((size_t)m_pad * sizeof(float)));
    MLK_TRY(ws[6].reserve(64));
    MLK_TRY(ws[25].reserve((size_t)m_pad * sizeof(double)));  // |q~|^2
    MLK_TRY(ws[26].reserve((size_t)m_pad * sizeof(double)));  // |q - q~|^2
    unsigned* dflag = ws[6].as<unsigned>();
    MLK_CHECK_HIP(hipMemsetAsync(dflag, 0, 16, stream));
    // Query locality order: regrouped queries share the screen's tiles, and
    // the pairs carry positions that qmap turns back into query indices.
    const float* Qs = Q;  // the queries in screen order
    const int* qmap = nullptr;
    if (order_wanted(index, m)) {
        MLK_TRY(ws[27].reserve(order_scratch_bytes(m)));
        MLK_TRY(ws[28].reserve((size_t)m_pad * sizeof(int)));
        MLK_TRY(ws[29].reserve((size_t)m * d * sizeof(float) + 16));
        MLK_TRY(query_order(index, Q, m, m_pad, ws[27].ptr, ws[28].as<int>(), ws[29].as<float>(), stream));
        Qs = ws[29].as<float>();
        qmap = ws[28].as<int>();
    }
    // the queries' rounding is recorded (and read back) only for an exact
    // database mirror (integer data); otherwise the queries count as inexact
    MLK_TRY(mirror_rows(prec, Qs, m, m_pad, d, dp, index->centre.as<float>(), mr.scale, ws[0].ptr, ws[1].as<double>(),
                        ws[2].as<uint8_t>(), mr.exact ? dflag : nullptr, ws[25].as<double>(), ws[26].as<double>(),
                        stream));
    unsigned q_inexact = 1;
    if (mr.exact) {
        MLK_CHECK_HIP(hipMemcpyAsync(&q_inexact, dflag, 4, hipMemcpyDeviceToHost, stream));
        MLK_CHECK_HIP(hipStreamSynchronize(stream));
    }
    const Consts c = make_consts(prec, dp, d, mr.exact, q_inexact == 0, mr.scale);
    MLK_TRY(prepare_bias(index, prec, c, stream));
    const double tau = R2 * c.scale2 * (1.0 + 0x1p-50);
    float* thr = ws[4].as<float>();
    hipLaunchKernelGGL(k_radius_thr, dim3(blocks_for(m_pad)), dim3(256), 0, stream, m, m_pad, ws[1].as<double>(),
                       ws[25].as<double>(), ws[26].as<double>(), ws[2].as<uint8_t>(), tau, c, thr);
    MLK_CHECK_HIP(hipGetLastError());

    // candidate pairs; the buffer grows and the screen reruns on overflow
    unsigned long long cap = std::max<unsigned long long>((unsigned long long)m * 64, 1ull << 20);
    if (ws[9].bytes / sizeof(int2) > cap) cap = ws[9].bytes / sizeof(int2);
    unsigned long long count = 0;
    unsigned long long* dcount = (unsigned long long*)(dflag + 2);
    Events events;
    hipEvent_t e0, e1;
    MLK_CHECK_HIP(events.make(&e0));
    MLK_CHECK_HIP(events.make(&e1));
    for (int attempt = 0; attempt < 3; ++attempt) {
        MLK_TRY(ws[9].reserve((size_t)cap * sizeof(int2)));
        cap = ws[9].bytes / sizeof(int2);
        MLK_CHECK_HIP(hipMemsetAsync(dcount, 0, sizeof(unsigned long long), stream));
        if (screen) {
            ScreenArgs sa{};
            sa.xs = mr.xs.ptr;
            sa.bias = mr.bias.as<float>();
            sa.nx32 = mr.nx32.as<float>();
            sa.qs = ws[0].ptr;
            sa.thr = thr;
            sa.n = n;
            sa.split_rows = (n + S - 1) / S;
            sa.off_lo = 0;
            sa.off_hi = sa.split_rows;
            sa.S = S;
            sa.QB = (int)QB;
            sa.self_off = q_offset;
            sa.exclude = exclude;
            sa.nchunk = screen_nchunk(SCREEN_PAIRS, prec, dp, plan.shape);
            sa.qgroup = screen_qgroup(prec, dp);
            sa.pairs = ws[9].as<int2>();
            sa.qmap = qmap;
            sa.pair_count = dcount;
            sa.pair_cap = cap;
            // the events take the launch's own start and end times
            hipExtLaunchKernelGGL(screen, dim3((unsigned)screen_grid(sa)), dim3(256), 0, stream, e0, e1, 0, sa);
            MLK_CHECK_HIP(hipGetLastError());
        } else {
            set_error("no radius screen kernel for this row width");
            return MLK_EINVAL;
        }
        if (mr.n_unsafe) {
            hipLaunchKernelGGL(k_unsafe_pairs, dim3(blocks_for(m * mr.n_unsafe)), dim3(256), 0, stream, m,
                               mr.unsafe.as<int>(), (int)mr.n_unsafe, q_offset, (int)exclude, qmap, ws[9].as<int2>(),
                               dcount, cap);
            MLK_CHECK_HIP(hipGetLastError());
        }
        MLK_CHECK_HIP(hipMemcpyAsync(&count, dcount, sizeof(count), hipMemcpyDeviceToHost, stream));
        MLK_CHECK_HIP(hipStreamSynchronize(stream));
        if (count <= cap) break;
        cap = count + count / 8 + 1024;
        if (attempt == 2) {
            set_error("radius candidate buffer kept overflowing");
            return MLK_ECAPACITY;
        }
    }
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, e0, e1) == hipSuccess) st.screen_ms = ms;
    const int64_t np = (int64_t)count;
    st.candidates = np;
    MLK_TRY(check_sort_count(np, "candidate pairs"));

    // sort pairs by (query, row): keys (query << rb) | row, rb = bits of n.
    // Rows that are re-sorted by distance afterwards need only the query
    // order (the distance sort orders every row by (distance, index) whatever
    // its input order), so their sort skips the row bits (c4: 3 radix passes
    // instead of 5).
    int rb = 1;
    while ((1ll << rb) < n) ++rb;
    MLK_TRY(ws[10].reserve((size_t)std::max<int64_t>(np, 1) * 8));   // keys in
    MLK_TRY(ws[11].reserve((size_t)std::max<int64_t>(np, 1) * 8));   // keys sorted
    MLK_TRY(ws[12].reserve((size_t)std::max<int64_t>(np, 1) * sizeof(double)));  // distances of the pairs
    MLK_TRY(ws[13].reserve((size_t)std::max<int64_t>(np, 1) + 16));  // member flags
    unsigned long long* keys = ws[10].as<unsigned long long>();
    unsigned long long* skeys = ws[11].as<unsigned long long>();
    if (np) {
        hipLaunchKernelGGL(k_pair_keys, dim3(blocks_for(np)), dim3(256), 0, stream, ws[9].as<int2>(), np, rb, qmap,
                           keys);
        int qbits = 1;
        while ((1ll << qbits) < m) ++qbits;
        const int begin_bit = sort_rows ? rb : 0;
        size_t tmp_bytes = 0;
        MLK_CHECK_HIP(hipcub::DeviceRadixSort::SortKeys(nullptr, tmp_bytes, keys, skeys, (int)np, begin_bit, rb + qbits,
                                                        stream));
        MLK_TRY(ws[14].reserve(tmp_bytes + 16));
        MLK_CHECK_HIP(hipcub::DeviceRadixSort::SortKeys(ws[14].ptr, tmp_bytes, keys, skeys, (int)np, begin_bit,
                                                        rb + qbits, stream));
        const bool vec = d % 4 == 0 && ((uintptr_t)Q % 16) == 0 && ((uintptr_t)X % 16) == 0;
        if (vec)
            hipLaunchKernelGGL(k_classify_staged, dim3(blocks_for(np)), dim3(256), 0, stream, skeys, np, rb, X, Q, d,
                               rc, ws[13].as<uint8_t>(), ws[12].as<double>());
        else
            hipLaunchKernelGGL(k_classify<false>, dim3(blocks_for(np)), dim3(256), 0, stream, skeys, np, rb, X, Q, d,
                               rc, ws[13].as<uint8_t>(), ws[12].as<double>());
        MLK_CHECK_HIP(hipGetLastError());
    }
    // compact the members (order preserved): block counts, their inclusive
    // scan (the last entry is nnz), then each block's members at its offset
    const int64_t nblk = (np + MEMBER_TILE - 1) / MEMBER_TILE;
    MLK_TRY(ws[15].reserve((size_t)std::max<int64_t>(nblk, 1) * 16 + 64));
    unsigned long long* bcount = ws[15].as<unsigned long long>();
    unsigned long long* bend = bcount + std::max<int64_t>(nblk, 1);
    int64_t nnz = 0;
    if (np) {
        hipLaunchKernelGGL(k_member_count, dim3((unsigned)nblk), dim3(256), 0, stream, ws[13].as<uint8_t>(), np,
                           bcount);
        size_t tmp_bytes = 0;
        MLK_CHECK_HIP(hipcub::DeviceScan::InclusiveSum(nullptr, tmp_bytes, bcount, bend, (int)nblk, stream));
        MLK_TRY(ws[14].reserve(tmp_bytes + 16));
        MLK_CHECK_HIP(hipcub::DeviceScan::InclusiveSum(ws[14].ptr, tmp_bytes, bcount, bend, (int)nblk, stream));
        unsigned long long total = 0;
        MLK_CHECK_HIP(hipMemcpyAsync(&total, bend + nblk - 1, sizeof(total), hipMemcpyDeviceToHost, stream));
        MLK_CHECK_HIP(hipStreamSynchronize(stream));
        nnz = (int64_t)total;
    }
    MLK_TRY(res->indices.reserve((size_t)std::max<int64_t>(nnz, 1) * sizeof(int64_t)));
    if (res->has_dist) MLK_TRY(res->dist.reserve((size_t)std::max<int64_t>(nnz, 1) * sizeof(double)));
    MLK_TRY(ws[16].reserve((size_t)std::max<int64_t>(nnz, 1) * sizeof(int)));  // each member's query
    int* mq = ws[16].as<int>();
    if (nnz)
        hipLaunchKernelGGL(k_member_scatter, dim3((unsigned)nblk), dim3(256), 0, stream, ws[13].as<uint8_t>(), skeys,
                           ws[12].as<double>(), np, rb, bend, bcount, res->indices.as<int64_t>(),
                           res->has_dist ? res->dist.as<double>() : nullptr, mq);
    // CSR row pointers
    hipLaunchKernelGGL(k_indptr_search, dim3(blocks_for(m + 1)), dim3(256), 0, stream, mq, nnz, m, indptr);
    MLK_CHECK_HIP(hipGetLastError());
    if (sort_rows) MLK_TRY(sort_rows_by_distance(res, indptr, m, nnz, ws, stream));
    MLK_CHECK_HIP(hipGetLastError());
    MLK_CHECK_HIP(hipStreamSynchronize(stream));
    res->nnz = nnz;
    *nnz_out = nnz;
    st.reranked = np;
    guard.keep = true;
    *out = res;
    return MLK_OK;
}

}  // namespace mlk

using namespace mlk;

extern "C" int mlk_radius_votes(const mlk_radius_result* r, const int64_t* indptr_dev, const int32_t* codes_dev,
                                int32_t n_classes, int32_t weights, double* votes_dev, int64_t* counts_dev,
                                void* stream_) {
    if (!r || (r->m && (!indptr_dev || !votes_dev)) || (r->nnz && !codes_dev) || n_classes < 1) {
        set_error("radius_votes: result, row pointers, codes and votes are required");
        return MLK_EINVAL;
    }
    if (weights != MLK_WEIGHTS_UNIFORM && weights != MLK_WEIGHTS_DISTANCE) {
        set_error("weights must be uniform or distance");
        return MLK_EINVAL;
    }
    if (weights == MLK_WEIGHTS_DISTANCE && !r->has_dist) {
        set_error("distance weights require a radius result with distances");
        return MLK_EINVAL;
    }
    hipStream_t stream = (hipStream_t)stream_;
    if (r->m) {
        hipLaunchKernelGGL(k_radius_votes, dim3(blocks_for(r->m)), dim3(256), 0, stream, indptr_dev, r->m,
                           r->indices.as<int64_t>(), r->has_dist ? r->dist.as<double>() : nullptr, codes_dev,
                           n_classes, (int)(weights == MLK_WEIGHTS_DISTANCE), votes_dev, counts_dev);
        MLK_CHECK_HIP(hipGetLastError());
    }
    MLK_CHECK_HIP(hipStreamSynchronize(stream));
    return MLK_OK;
}

extern "C" int mlk_radius_mean(const mlk_radius_result* r, const int64_t* indptr_dev, const double* targets_dev,
                               int32_t weights, double* pred_dev, int64_t* counts_dev, void* stream_) {
    if (!r || (r->m && (!indptr_dev || !pred_dev)) || (r->nnz && !targets_dev)) {
        set_error("radius_mean: result, row pointers, targets and output are required");
        return MLK_EINVAL;
    }
    if (weights != MLK_WEIGHTS_UNIFORM && weights != MLK_WEIGHTS_DISTANCE) {
        set_error("weights must be uniform or distance");
        return MLK_EINVAL;
    }
    if (weights == MLK_WEIGHTS_DISTANCE && !r->has_dist) {
        set_error("distance weights require a radius result with distances");
        return MLK_EINVAL;
    }
    hipStream_t stream = (hipStream_t)stream_;
    if (r->m) {
        hipLaunchKernelGGL(k_radius_mean, dim3(blocks_for(r->m)), dim3(256), 0, stream, indptr_dev, r->m,
                           r->indices.as<int64_t>(), r->has_dist ? r->dist.as<double>() : nullptr, targets_dev,
                           (int)(weights == MLK_WEIGHTS_DISTANCE), pred_dev, counts_dev);
        MLK_CHECK_HIP(hipGetLastError());
    }
    MLK_CHECK_HIP(hipStreamSynchronize(stream));
    return MLK_OK;
}
