// The fused MFMA screen kernel (k-NN candidate lists and radius pairs).
//
// Reference counterpart: gram_lower + select_topk / compact_gram
// (src/mlxlearn_device/_neighbors_general.py:188-223, _neighbors_fused.py:421-506,
// _radius.py:128-185), which write the rows x N Gram tile to memory and
// re-read it for selection; here the tile stays in MFMA accumulators and only
// threshold hits leave the kernel.
//
// This translation unit is compiled with -fno-honor-nans: every screen value
// is finite (rows outside the safe range are zero in the mirror), so the
// max-tree hit test needs no NaN canonicalisation.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "mlk_device.h"
#include "mlk_internal.h"
#include "mlk_screen.h"

namespace mlk {

template <int PREC> struct PT;
template <> struct PT<PREC_FP32> { typedef float T; };
template <> struct PT<PREC_BF16> { typedef __bf16 T; };
template <> struct PT<PREC_FP16> { typedef _Float16 T; };

template <int PREC>
__device__ __forceinline__ f32x16 mfma16(const i32x4& a, const i32x4& b, const f32x16& c) {
    if constexpr (PREC == PREC_BF16)
        return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c, 0, 0, 0);
    else
        return __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(f16x8, a), __builtin_bit_cast(f16x8, b), c, 0, 0, 0);
}

// Work item (query block, split) of a block; the QB blocks of one split are
// consecutive slots of one XCD (blocks are dispatched round-robin over the 8
// XCDs), so they stream the same rows through that XCD's L2 together. A block
// past the last item gets split >= S (no rows: it exits).
__device__ __forceinline__ void block_item(const ScreenArgs& a, int& qb, int& split) {
    const int b = blockIdx.x;
    if (a.xcd_map && a.qgroup > 1) {
        // an XCD's consecutive blocks cover qgroup query blocks x its splits, so
        // the blocks resident on it at one time share both their query rows
        // and their database rows in its L2 (deep rows reload query fragments)
        const int xcd = b & 7, slot = b >> 3, sg = a.S >> 3;
        const int g = slot / (a.qgroup * sg);
        const int gsize = min(a.qgroup, a.QB - g * a.qgroup);
        const int local = slot - g * a.qgroup * sg;
        qb = g * a.qgroup + local % gsize;
        split = (local / gsize) * 8 + xcd;
    } else if (a.xcd_map) {
        // XCD x takes items [x * per, (x + 1) * per) of the split-major order
        const int xcd = b & 7, slot = b >> 3;
        const int64_t per = ((int64_t)a.S * a.QB + 7) / 8;
        const int64_t item = xcd * per + slot;
        qb = (int)(item % a.QB);
        split = (int)(item / a.QB);
    } else {
        qb = b % a.QB;
        split = b / a.QB;
    }
}

typedef __attribute__((address_space(3))) char lds_char;

#if defined(__HIP_DEVICE_COMPILE__)
// Raw buffer resource over [base, base + bytes): loads past `bytes` return zero
// (buffer range check), so tile loads need no clamping.
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

// s_waitcnt with vmcnt <= N and lgkmcnt = 0 (expcnt unconstrained); gfx9
// encoding: vmcnt[3:0] in bits 3:0, vmcnt[5:4] in bits 15:14.
template <int N>
__device__ __forceinline__ void wait_vm_lgkm0() {
    static_assert(N >= 0 && N < 64, "vmcnt range");
    __builtin_amdgcn_s_waitcnt((N & 15) | (7 << 4) | (0 << 8) | ((N >> 4) << 14));
}

// Workgroup barrier that the compiler may not move memory accesses across.
// The waits that make the tile data valid are explicit (wait_vm_lgkm0).
__device__ __forceinline__ void tile_barrier() { asm volatile("s_barrier" ::: "memory"); }

// The screen. A block of 4 waves walks BN-row tiles (BN = 32 * RTW) of its
// split for BQ = 4 * QT * 32 queries: each wave keeps the operand fragments of
// its QT x 32 queries in registers and multiplies them with every row of the
// tile, so every A-operand byte read from LDS feeds QT MFMAs. Tiles stream
// from memory straight into an NS-stage LDS ring (buffer loads with the LDS
// flag, no staging registers; rows padded by one 16-byte chunk so the 32 rows
// of an operand read hit distinct banks): while a wave computes tile t, tiles
// t+1 .. t+NS-1 are in flight, and at the end of each step a wave waits only
// for tile t+1 (vmcnt counts its outstanding tile loads in issue order).
// Per tile a wave issues its MFMAs (accumulators start at the rows' biases
// -a_x), then one branch-free max-tree test of every accumulator against its
// query's threshold; only a hit enters the append path. With several query
// tiles per wave (narrow rows) a step holds two 32-row tiles that are
// multiplied and tested one after the other, so one accumulator set is live
// and the step's barrier and tile load serve twice the MFMA work; their A
// fragments are read three k-steps ahead of the MFMAs that use them.
// Chunked rows (CH): step u multiplies feature chunk u % nc of tile u / nc and
// the accumulators carry over a tile's chunks. Deep rows reload the query
// fragments of each chunk; with RC = NS resident chunks (wide rows in two
// halves) the fragments of every chunk stay in registers and LDS stage c
// always holds chunk c, so the chunk index is a compile-time constant.
template <int PREC, int DP, int QT, int RTW, int NS, int MODE, bool CH = false, int RC = 1>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(RC > 1 ? 2 : 1))) void k_screen(
    const ScreenArgs a) {
    // the body uses device-only types (buffer resources): compile it in the
    // device pass only, so the host pass still emits the kernel's launch stub
#if defined(__HIP_DEVICE_COMPILE__)
    typedef typename PT<PREC>::T T;
    constexpr int NW = 4;                     // waves per block
    constexpr int BN = 32 * RTW;              // rows per tile (RTW 32-row MFMA tiles per step)
    constexpr int EB = sizeof(T);
    constexpr int RB = DP * EB;               // bytes per database row
    constexpr int CPR = RB / 16;              // 16-byte chunks per row
    constexpr int RS = RB + 16;               // LDS row stride: one pad chunk per row
    constexpr int NCH = BN * (CPR + 1);       // LDS chunks per tile
    constexpr int NLOAD = (NCH + 63) / 64;    // 1 KiB load instructions per tile
    constexpr int ABYTES = NLOAD * 1024;
    constexpr int LPW = (NLOAD + NW - 1) / NW;
    constexpr int BQ = NW * QT * 32;
    static_assert(NS >= 2 && NS <= 4, "stages");
    static_assert(QT == 1 || QT == 2 || QT == 4, "query tiles");
    static_assert(RTW == 1 || RTW == 2, "row tiles per step");
    static_assert(!CH || NS == 2, "deep rows: two stages (each step waits for its query reload)");
    static_assert(RC == 1 || (CH && RC == NS), "resident chunks: one LDS stage per chunk");
    static_assert(NS * (ABYTES + 512) + (MODE == SCREEN_PAIRS ? NW * PAIR_STAGE * 8 : 0) <= 160 * 1024,
                  "LDS of one block");
    // distinct buffers (not one array) so the compiler can tell the tile
    // being read from the tiles being loaded and keeps those loads in flight
    __shared__ __attribute__((aligned(1024))) char sA0[ABYTES];
    __shared__ __attribute__((aligned(1024))) char sA1[ABYTES];
    __shared__ __attribute__((aligned(1024))) char sA2[NS > 2 ? ABYTES : 16];
    __shared__ __attribute__((aligned(1024))) char sA3[NS > 3 ? ABYTES : 16];
    // [0, 64): -a_x of the tile rows (float32)
    __shared__ __attribute__((aligned(256))) float sB0[64];
    __shared__ __attribute__((aligned(256))) float sB1[64];
    __shared__ __attribute__((aligned(256))) float sB2[NS > 2 ? 64 : 4];
    __shared__ __attribute__((aligned(256))) float sB3[NS > 3 ? 64 : 4];
    __shared__ int2 sPair[MODE == SCREEN_PAIRS ? NW : 1][MODE == SCREEN_PAIRS ? PAIR_STAGE : 1];
    __shared__ int sPairN[NW];

    int qb, split;
    block_item(a, qb, split);
    const int lane = threadIdx.x & 63, h = lane >> 5, col = lane & 31;
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int64_t s_lo = (int64_t)split * a.split_rows;
    const int64_t r_begin = s_lo + a.off_lo;
    int64_t r_end = s_lo + (a.off_hi < a.split_rows ? a.off_hi : a.split_rows);
    if (r_end > a.n) r_end = a.n;
    if (r_begin >= r_end) return;
    // Deep rows (CH): a row is nc chunks of DP features; step u of the walk is
    // chunk u % nc of tile u / nc, and the accumulators carry over the chunks
    // of a tile. Otherwise nc = 1 and a step is a tile.
    const int nc = CH ? a.nchunk : 1;
    const uint32_t rbg = (uint32_t)RB * (uint32_t)nc;  // bytes per database row in memory
    const size_t qstride = (size_t)DP * nc;          // elements per query row

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
    const int64_t qrow0 = (int64_t)qb * BQ + wave * QT * 32;
    float thr[QT];
#pragma unroll
    for (int qt = 0; qt < QT; ++qt) thr[qt] = a.thr[qrow0 + qt * 32 + col];
    // k-NN mode: each lane owns the candidate sub-list (query col, split, lane
    // half h), so appends need no atomics; counts live in registers.
    int cnt_r[QT];
    if constexpr (MODE == SCREEN_LISTS) {
#pragma unroll
        for (int qt = 0; qt < QT; ++qt) {
            const size_t q = (size_t)(qrow0 + qt * 32 + col);
            const size_t sub = (q * a.S + split) * 2 + h;
            cnt_r[qt] = a.cnt[sub];
        }
    }
    // entry j of sub-list s of query q lives at (q * cap + j) * 2S + s: the j-th
    // entries of a query's sub-lists are adjacent, so the list walks that visit
    // every sub-list of a query read them with coalesced loads (formed on the
    // hit path only, so no registers hold it across the walk)
    auto list_base = [&](int qt) -> size_t {
        const size_t q = (size_t)(qrow0 + qt * 32 + col);
        return q * a.cap * (size_t)(2 * a.S) + (size_t)split * 2 + h;
    };
    bool live = false;
#pragma unroll
    for (int qt = 0; qt < QT; ++qt) live |= thr[qt] < INFINITY;
    const bool wave_live = ballot(live) != 0;

    constexpr int NB = PREC == PREC_FP32 ? DP / 2 : DP / 16;
    typedef typename std::conditional<PREC == PREC_FP32, float, i32x4>::type BFrag;
    // query fragments: one chunk (deep rows reload them per chunk), or all RC
    // chunks of the row (wide rows split in RC chunks: stage c holds chunk c)
    BFrag bq[QT][NB * RC];
    // deep rows: fragments of feature chunk cc into dst
    auto load_query = [&](BFrag(&dst)[QT][NB * RC], int cc) {
#pragma unroll
        for (int qt = 0; qt < QT; ++qt) {
            const T* qp = (const T*)a.qs + (size_t)(qrow0 + qt * 32 + col) * qstride + (size_t)cc * DP + h * (DP / 2);
            if constexpr (PREC == PREC_FP32) {
#pragma unroll
                for (int j = 0; j < NB; j += 4) {
                    f32x4 v = *(const f32x4*)(qp + j);
                    dst[qt][j] = v[0]; dst[qt][j + 1] = v[1]; dst[qt][j + 2] = v[2]; dst[qt][j + 3] = v[3];
                }
            } else {
#pragma unroll
                for (int j = 0; j < NB; ++j) dst[qt][j] = *(const i32x4*)(qp + 8 * j);
            }
        }
    };
    // the first chunk (the whole row unless CH), written out rather than through
    // the lambda: the narrow kernels then keep their register budget (128 at
    // two query tiles of 96 features, occupancy 4)
#pragma unroll
    for (int qt = 0; qt < QT; ++qt) {
        const T* qp = (const T*)a.qs + (size_t)(qrow0 + qt * 32 + col) * qstride + h * (DP / 2);
        if constexpr (PREC == PREC_FP32) {
#pragma unroll
            for (int j = 0; j < NB; j += 4) {
                f32x4 v = *(const f32x4*)(qp + j);
                bq[qt][j] = v[0]; bq[qt][j + 1] = v[1]; bq[qt][j + 2] = v[2]; bq[qt][j + 3] = v[3];
            }
        } else {
#pragma unroll
            for (int c = 0; c < RC; ++c)
#pragma unroll
                for (int j = 0; j < NB; ++j) bq[qt][c * NB + j] = *(const i32x4*)(qp + c * DP + 8 * j);
        }
    }

    const int64_t ntiles = (r_end - r_begin + BN - 1) / BN;
    const int64_t nsteps = ntiles * nc;
    // Load plan: LDS chunk L (of NLOAD * 64 per tile) is row L / (CPR + 1),
    // position L % (CPR + 1); the last position of each row and every chunk
    // past the tile are padding (loaded out of range, i.e. zero). Source
    // offsets within a tile are lane constants; the tile base goes into the
    // buffer resource (SGPRs). Wave 1 also loads the rows' biases.
    auto chunk_off = [&](int i) -> uint32_t {
        const int L = (wave + i * NW) * 64 + lane;
        const int row = L / (CPR + 1), pos = L % (CPR + 1);
        return (pos < CPR && row < BN) ? (uint32_t)row * rbg + (uint32_t)(pos * 16) : 0x7ffffff0u;
    };
    // narrow rows: the few offsets stay in registers; wide rows recompute them
    constexpr bool KEEP_OFF = LPW <= 4;
    uint32_t voff[KEEP_OFF ? LPW : 1];
    if constexpr (KEEP_OFF) {
#pragma unroll
        for (int i = 0; i < LPW; ++i) voff[i] = chunk_off(i);
    }
    // tile loads this wave issues (uniform): its share of the NLOAD chunks, plus one
    const int my_loads = (NLOAD > wave ? (NLOAD - wave + NW - 1) / NW : 0) + (wave == 1);

    auto bufA = [&](auto I) -> char* {
        constexpr int i = decltype(I)::value;
        if constexpr (i == 0) return sA0; else if constexpr (i == 1) return sA1;
        else if constexpr (i == 2) return sA2; else return sA3;
    };
    auto bufB = [&](auto I) -> float* {
        constexpr int i = decltype(I)::value;
        if constexpr (i == 0) return sB0; else if constexpr (i == 1) return sB1;
        else if constexpr (i == 2) return sB2; else return sB3;
    };
    auto load_tile = [&](int64_t u, char* dstA, float* dstB) {
        const int64_t t = CH ? u / nc : u;
        const uint32_t coff = CH ? (uint32_t)(u - t * nc) * RB : 0u;  // byte offset of the chunk in a row
        const int64_t r0 = r_begin + t * BN;
        const int64_t rows = r_end - r0 < BN ? r_end - r0 : BN;
        const __amdgpu_buffer_rsrc_t ra = buffer_rsrc((const char*)a.xs + (size_t)r0 * rbg, (uint32_t)rows * rbg);
#pragma unroll
        for (int i = 0; i < LPW; ++i) {
            const int g = wave + i * NW;
            if (g < NLOAD) {
                uint32_t off;
                if constexpr (KEEP_OFF) off = voff[i]; else off = chunk_off(i);
                if constexpr (CH) off = off == 0x7ffffff0u ? off : off + coff;
                __builtin_amdgcn_raw_ptr_buffer_load_lds(ra, (lds_char*)(dstA + g * 1024), 16, off, 0, 0, 0);
            }
        }
        if (wave == 1) {
            const __amdgpu_buffer_rsrc_t rb = buffer_rsrc(a.bias + r0, (uint32_t)(rows * 4));
            __builtin_amdgcn_raw_ptr_buffer_load_lds(rb, (lds_char*)dstB, 4, lane * 4, 0, 0, 0);
        }
    };
    // wait until at most `stages` tiles of this wave's loads are outstanding
    // (and every LDS read has returned), then meet the other waves
    auto sync_tiles = [&](int stages) {
        const int allowed = stages * my_loads;
        switch (allowed) {
#define MLK_WAIT_CASE(n) case n: wait_vm_lgkm0<n>(); break;
            MLK_WAIT_CASE(1) MLK_WAIT_CASE(2) MLK_WAIT_CASE(3) MLK_WAIT_CASE(4) MLK_WAIT_CASE(5)
            MLK_WAIT_CASE(6) MLK_WAIT_CASE(7) MLK_WAIT_CASE(8) MLK_WAIT_CASE(9) MLK_WAIT_CASE(10)
            MLK_WAIT_CASE(11) MLK_WAIT_CASE(12) MLK_WAIT_CASE(13) MLK_WAIT_CASE(14) MLK_WAIT_CASE(15)
            MLK_WAIT_CASE(16) MLK_WAIT_CASE(17) MLK_WAIT_CASE(18) MLK_WAIT_CASE(19) MLK_WAIT_CASE(20)
            MLK_WAIT_CASE(21) MLK_WAIT_CASE(22) MLK_WAIT_CASE(23) MLK_WAIT_CASE(24) MLK_WAIT_CASE(25)
            MLK_WAIT_CASE(26) MLK_WAIT_CASE(27) MLK_WAIT_CASE(28) MLK_WAIT_CASE(29) MLK_WAIT_CASE(30)
            MLK_WAIT_CASE(31) MLK_WAIT_CASE(32)
#undef MLK_WAIT_CASE
            default: wait_vm_lgkm0<0>(); break;
        }
        tile_barrier();
    };

    f32x16 acc_keep[RTW][QT];  // deep rows: sums carried over the chunks of a tile
    // one step: prefetch step u + NS - 1 into the buffer freed by step u - 1,
    // compute step u from buffer CUR, wait for step u + 1
    auto step = [&](auto CUR, int64_t u) {
        constexpr int c = decltype(CUR)::value;
        constexpr int nxt = (c + NS - 1) % NS;
        constexpr int cb = RC > 1 ? c * NB : 0;  // resident chunks: stage c multiplies chunk c
        const bool full = u + NS - 1 < nsteps;
        const int64_t t = CH ? u / nc : u;
        const int cc = CH ? (int)(u - t * nc) : 0;  // feature chunk of this step
        if (full) load_tile(u + NS - 1, bufA(std::integral_constant<int, nxt>{}), bufB(std::integral_constant<int, nxt>{}));
        const char* cA = bufA(CUR);
        const float* cB = bufB(CUR);
        bool stored = false;
        if (wave_live) {
            const int64_t r0 = r_begin + t * BN;
            f32x16 acc_step[RTW][QT];
            f32x16(&acc)[RTW][QT] = *(CH ? &acc_keep : &acc_step);
            // MFMAs of row tile rt of the step
            auto mfma_rt = [&](int rt) {
                // accumulator start = -a_x of this lane's 16 rows (4 runs of 4 rows)
                const float* nb = cB + rt * 32 + (h << 2);
                const f32x4 b0 = *(const f32x4*)(nb), b1 = *(const f32x4*)(nb + 8);
                const f32x4 b2 = *(const f32x4*)(nb + 16), b3 = *(const f32x4*)(nb + 24);
                const f32x16 init = {b0[0], b0[1], b0[2], b0[3], b1[0], b1[1], b1[2], b1[3],
                                     b2[0], b2[1], b2[2], b2[3], b3[0], b3[1], b3[2], b3[3]};
                const char* arow = cA + (rt * 32 + col) * RS + h * (CPR / 2) * 16;
                if constexpr (PREC == PREC_FP32) {
                    if (cc == 0) {
#pragma unroll
                        for (int qt = 0; qt < QT; ++qt) acc[rt][qt] = init;
                    }
#pragma unroll
                    for (int j = 0; j < NB; j += 4) {
                        const f32x4 av = *(const f32x4*)(arow + (j / 4) * 16);
#pragma unroll
                        for (int e = 0; e < 4; ++e)
#pragma unroll
                            for (int qt = 0; qt < QT; ++qt)
                                acc[rt][qt] = __builtin_amdgcn_mfma_f32_32x32x2f32(av[e], bq[qt][j + e], acc[rt][qt], 0, 0, 0);
                    }
                } else if constexpr (!CH && QT >= 2 && NB >= 4) {
                    // A fragments read PF k-steps ahead of their MFMAs: the bias
                    // and the first PF fragments, then one read per k-step
                    constexpr int PF = 3;
                    i32x4 av[NB];
#pragma unroll
                    for (int j = 0; j < NB; ++j) av[j] = *(const i32x4*)(arow + j * 16);
#pragma unroll
                    for (int qt = 0; qt < QT; ++qt) acc[rt][qt] = mfma16<PREC>(av[0], bq[qt][cb], init);
#pragma unroll
                    for (int j = 1; j < NB; ++j)
#pragma unroll
                        for (int qt = 0; qt < QT; ++qt) acc[rt][qt] = mfma16<PREC>(av[j], bq[qt][cb + j], acc[rt][qt]);
                    __builtin_amdgcn_sched_group_barrier(0x100, 4 + PF, 0);
#pragma unroll
                    for (int j = 0; j < NB; ++j) {
                        __builtin_amdgcn_sched_group_barrier(0x008, QT, 0);
                        if (j + PF < NB) __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
                    }
                } else if constexpr (RC > 1) {
                    // resident chunks: the biases start the accumulators in the
                    // first chunk only (no separate bias registers stay live)
                    if (cc == 0) {
#pragma unroll
                        for (int qt = 0; qt < QT; ++qt) acc[rt][qt] = init;
                    }
#pragma unroll
                    for (int j = 0; j < NB; ++j) {
                        const i32x4 av = *(const i32x4*)(arow + j * 16);
#pragma unroll
                        for (int qt = 0; qt < QT; ++qt) acc[rt][qt] = mfma16<PREC>(av, bq[qt][cb + j], acc[rt][qt]);
                    }
                } else {
                    // the first k-step reads the bias registers as its C operand
                    // (deep rows: only in the first chunk of the tile)
                    const i32x4 a0 = *(const i32x4*)(arow);
#pragma unroll
                    for (int qt = 0; qt < QT; ++qt)
                        acc[rt][qt] = mfma16<PREC>(a0, bq[qt][cb], cc == 0 ? init : acc[rt][qt]);
#pragma unroll
                    for (int j = 1; j < NB; ++j) {
                        const i32x4 av = *(const i32x4*)(arow + j * 16);
#pragma unroll
                        for (int qt = 0; qt < QT; ++qt) acc[rt][qt] = mfma16<PREC>(av, bq[qt][cb + j], acc[rt][qt]);
                    }
                }
            };
            // Hit test of row tiles [rt0, rt1): acc >= thr  <=>  L(x) <= tau.
            // Branch-free over the tiles; only a hit enters the append path.
            auto test_rt = [&](int rt0, int rt1) {
                bool hit_rq[RTW][QT];
                bool any = false;
#pragma unroll
                for (int rt = 0; rt < RTW; ++rt)
#pragma unroll
                    for (int qt = 0; qt < QT; ++qt) {
                        hit_rq[rt][qt] = rt >= rt0 && rt < rt1 && max16(acc[rt][qt]) >= thr[qt];
                        any |= hit_rq[rt][qt];
                    }
                if (!ballot(any)) return;
                stored = true;
#pragma unroll
                for (int rt = 0; rt < RTW; ++rt)
#pragma unroll
                for (int qt = 0; qt < QT; ++qt) {
                    if (rt < rt0 || rt >= rt1 || !ballot(hit_rq[rt][qt])) continue;
                    const int64_t qg = qrow0 + qt * 32 + col;
                    int64_t self = -1;
                    if (a.exclude) self = a.self_off + (MODE == SCREEN_PAIRS && a.qmap ? (int64_t)a.qmap[qg] : qg);
                    if constexpr (MODE == SCREEN_LISTS) {
                        // wave-uniform mask of the accumulator registers with a hit
                        // in some lane; only those registers are visited (uniform
                        // register index, v_movrels)
                        uint32_t regs = 0;
#pragma unroll
                        for (int i = 0; i < 16; ++i) regs |= (ballot(acc[rt][qt][i] >= thr[qt]) != 0 ? 1u : 0u) << i;
#pragma unroll 1
                        while (regs) {
                            const int i = __builtin_ctz(regs);
                            regs &= regs - 1;
                            const float v = acc[rt][qt][i];
                            const int rl = rt * 32 + mfma32_row(i, lane);
                            const int64_t row = r0 + rl;
                            if (v >= thr[qt] && row < r_end && row != self) {
                                const int slot = cnt_r[qt]++;
                                if (slot < a.cap)
                                    a.cand[list_base(qt) + (size_t)slot * (2 * a.S)] =
                                        int2{(int)row, __float_as_int(v - cB[rl])};
                            }
                        }
                    } else {
                        // registers with a hit in some lane, visited in order;
                        // each one's hits are compacted into the wave's LDS
                        // staging with one ballot (no per-lane atomics)
                        uint32_t regs = 0;
#pragma unroll
                        for (int i = 0; i < 16; ++i) regs |= (ballot(acc[rt][qt][i] >= thr[qt]) != 0 ? 1u : 0u) << i;
#pragma unroll 1
                        while (regs) {
                            const int i = __builtin_ctz(regs);
                            regs &= regs - 1;
                            const float v = acc[rt][qt][i];
                            const int rl = rt * 32 + mfma32_row(i, lane);
                            const int64_t row = r0 + rl;
                            const bool hit = v >= thr[qt] && row < r_end && row != self;
                            const unsigned long long mask = ballot(hit);
                            if (!mask) continue;
                            if (sPairN[wave] + 64 > PAIR_STAGE) flush_pairs();
                            const int base = sPairN[wave];
                            const int below = __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                                                        __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0));
                            if (hit) sPair[wave][base + below] = int2{(int)qg, (int)row};
                            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
                            __builtin_amdgcn_wave_barrier();
                            if (lane == 0) sPairN[wave] = base + __popcll(mask);
                            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
                            __builtin_amdgcn_wave_barrier();
                        }
                    }
                }
            };
            // several query tiles per wave: each row tile is tested right after
            // its MFMAs, so the accumulators of one row tile are live at a time
            constexpr bool SEQ = !CH && QT >= 2 && RTW > 1;
            if constexpr (SEQ) {
#pragma unroll
                for (int rt = 0; rt < RTW; ++rt) {
                    mfma_rt(rt);
                    test_rt(rt, rt + 1);
                }
            } else {
#pragma unroll
                for (int rt = 0; rt < RTW; ++rt) mfma_rt(rt);
                // deep rows: the next step's fragments replace this step's once its
                // MFMAs are issued (two stages: the step's final wait covers them)
                if constexpr (CH && RC == 1) {
                    if (u + 1 < nsteps) load_query(bq, cc + 1 == nc ? 0 : cc + 1);
                }
                // deep rows: the test only after the tile's last chunk
                if (!CH || cc == nc - 1) test_rt(0, RTW);
            }
        }
        // Stores of the hit path also count in vmcnt: after one, wait for
        // everything rather than rely on the order of loads and stores.
        sync_tiles(full && !stored ? NS - 2 : 0);
    };

    // prologue: steps 0 .. NS-2 in flight, wait for step 0
    load_tile(0, sA0, sB0);
    if constexpr (NS > 2) if (1 < nsteps) load_tile(1, sA1, sB1);
    if constexpr (NS > 3) if (2 < nsteps) load_tile(2, sA2, sB2);
    sync_tiles(nsteps >= NS - 1 ? NS - 2 : 0);
    for (int64_t u = 0; u < nsteps; u += NS) {
        step(std::integral_constant<int, 0>{}, u);
        if (u + 1 < nsteps) step(std::integral_constant<int, 1>{}, u + 1);
        if constexpr (NS > 2) if (u + 2 < nsteps) step(std::integral_constant<int, 2>{}, u + 2);
        if constexpr (NS > 3) if (u + 3 < nsteps) step(std::integral_constant<int, 3>{}, u + 3);
    }
    if constexpr (MODE == SCREEN_LISTS) {
#pragma unroll
        for (int qt = 0; qt < QT; ++qt)
            a.cnt[(((size_t)(qrow0 + qt * 32 + col)) * a.S + split) * 2 + h] = cnt_r[qt];
    } else {
        flush_pairs();
    }
#else
    (void)a;
#endif
}

// Software-pipelined screen for k-NN lists on rows of at most 128 features
// (bf16 / fp16): 32-row tiles, QT query tiles per wave (4 * QT * 32 queries
// per block), a three-stage LDS tile ring. Step u issues the MFMAs of tile u
// into one accumulator set and, in the same instruction stream, the
// threshold test (max-tree and compare) of tile u - 1 held in the other set,
// so the test's VALU work issues in the shadow of the running MFMAs instead
// of after them; the A fragments of each tile are read from LDS four ahead of
// their MFMAs. Only a wave with a hit leaves the straight-line path (the
// append, as in k_screen, with the row's bias read from memory because its
// LDS slot may already hold a later tile).
template <int PREC, int DP, int QT>
__global__ __launch_bounds__(256) void k_screen_pp(const ScreenArgs a) {
#if defined(__HIP_DEVICE_COMPILE__)
    static_assert(PREC != PREC_FP32, "bf16 / fp16 operands");
    static_assert(DP % 32 == 0 && DP <= 128, "narrow rows");
    constexpr int NW = 4, NS = 3, BN = 32;
    constexpr int RB = DP * 2;                // bytes per database row
    constexpr int CPR = RB / 16;              // 16-byte chunks per row
    constexpr int RS = RB + 16;               // LDS row stride: one pad chunk per row
    constexpr int NCH = BN * (CPR + 1);
    constexpr int NLOAD = (NCH + 63) / 64;    // 1 KiB load instructions per tile
    constexpr int ABYTES = NLOAD * 1024;
    constexpr int LPW = (NLOAD + NW - 1) / NW;
    constexpr int BQ = NW * QT * 32;
    constexpr int NB = DP / 16;               // k-steps (A fragments) per tile
    constexpr int PF = NB < 4 ? NB : 4;       // A fragments read ahead
    static_assert(NS * (ABYTES + 256) <= 160 * 1024, "LDS of one block");
    // distinct buffers with compile-time selection, so the compiler can tell the
    // tile being read from the tiles in flight and leaves those loads running
    __shared__ __attribute__((aligned(1024))) char sA0[ABYTES];
    __shared__ __attribute__((aligned(1024))) char sA1[ABYTES];
    __shared__ __attribute__((aligned(1024))) char sA2[ABYTES];
    __shared__ __attribute__((aligned(256))) float sB0[64];  // [0, 32): -a_x of the tile rows
    __shared__ __attribute__((aligned(256))) float sB1[64];
    __shared__ __attribute__((aligned(256))) float sB2[64];
    auto bufA = [&](auto I) -> char* {
        constexpr int i = decltype(I)::value;
        if constexpr (i == 0) return sA0; else if constexpr (i == 1) return sA1; else return sA2;
    };
    auto bufB = [&](auto I) -> float* {
        constexpr int i = decltype(I)::value;
        if constexpr (i == 0) return sB0; else if constexpr (i == 1) return sB1; else return sB2;
    };

    int qb, split;
    block_item(a, qb, split);
    const int lane = threadIdx.x & 63, h = lane >> 5, col = lane & 31;
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int64_t s_lo = (int64_t)split * a.split_rows;
    const int64_t r_begin = s_lo + a.off_lo;
    int64_t r_end = s_lo + (a.off_hi < a.split_rows ? a.off_hi : a.split_rows);
    if (r_end > a.n) r_end = a.n;
    if (r_begin >= r_end) return;

    const int64_t qrow0 = (int64_t)qb * BQ + wave * QT * 32;
    float thr[QT];
    int cnt_r[QT];
#pragma unroll
    for (int qt = 0; qt < QT; ++qt) {
        const size_t q = (size_t)(qrow0 + qt * 32 + col);
        thr[qt] = a.thr[q];
        cnt_r[qt] = a.cnt[(q * a.S + split) * 2 + h];
    }
    bool live = false;
#pragma unroll
    for (int qt = 0; qt < QT; ++qt) live |= thr[qt] < INFINITY;
    const bool wave_live = ballot(live) != 0;

    // query fragments: lane (col, h) holds features h * DP/2 + [8j, 8j + 8) of query col
    i32x4 bq[QT][NB];
#pragma unroll
    for (int qt = 0; qt < QT; ++qt) {
        const __bf16* qp = (const __bf16*)a.qs + (size_t)(qrow0 + qt * 32 + col) * DP + h * (DP / 2);
#pragma unroll
        for (int j = 0; j < NB; ++j) bq[qt][j] = *(const i32x4*)(qp + 8 * j);
    }

    const int64_t ntiles = (r_end - r_begin + BN - 1) / BN;
    auto chunk_off = [&](int i) -> uint32_t {
        const int L = (wave + i * NW) * 64 + lane;
        const int row = L / (CPR + 1), pos = L % (CPR + 1);
        return (pos < CPR && row < BN) ? (uint32_t)row * RB + (uint32_t)(pos * 16) : 0x7ffffff0u;
    };
    uint32_t voff[LPW];
#pragma unroll
    for (int i = 0; i < LPW; ++i) voff[i] = chunk_off(i);
    const int my_loads = (NLOAD > wave ? (NLOAD - wave + NW - 1) / NW : 0) + (wave == 1);
    auto load_tile = [&](int64_t t, char* dstA, float* dstB) {
        const int64_t r0 = r_begin + t * BN;
        const int64_t rows = r_end - r0 < BN ? r_end - r0 : BN;
        const __amdgpu_buffer_rsrc_t ra = buffer_rsrc((const char*)a.xs + (size_t)r0 * RB, (uint32_t)rows * RB);
#pragma unroll
        for (int i = 0; i < LPW; ++i)
            if (wave + i * NW < NLOAD)
                __builtin_amdgcn_raw_ptr_buffer_load_lds(ra, (lds_char*)(dstA + (wave + i * NW) * 1024), 16, voff[i],
                                                         0, 0, 0);
        if (wave == 1) {
            const __amdgpu_buffer_rsrc_t rb = buffer_rsrc(a.bias + r0, (uint32_t)(rows * 4));
            __builtin_amdgcn_raw_ptr_buffer_load_lds(rb, (lds_char*)dstB, 4, lane * 4, 0, 0, 0);
        }
    };
    auto sync_tiles = [&](int stages) {
        switch (stages * my_loads) {
            case 0: wait_vm_lgkm0<0>(); break;
#define MLK_WAIT_CASE(n) case n: wait_vm_lgkm0<n>(); break;
            MLK_WAIT_CASE(1) MLK_WAIT_CASE(2) MLK_WAIT_CASE(3) MLK_WAIT_CASE(4) MLK_WAIT_CASE(5)
            MLK_WAIT_CASE(6) MLK_WAIT_CASE(7) MLK_WAIT_CASE(8)
#undef MLK_WAIT_CASE
            default: wait_vm_lgkm0<0>(); break;
        }
        tile_barrier();
    };

    f32x16 accA[QT], accB[QT];
    // MFMAs of tile t (LDS buffer buf) into acc
    auto mfma_tile = [&](f32x16(&acc)[QT], const char* tA, const float* tB) {
        const char* arow = tA + col * RS + h * (CPR / 2) * 16;
        const float* nb = tB + (h << 2);
        const f32x4 b0 = *(const f32x4*)(nb), b1 = *(const f32x4*)(nb + 8);
        const f32x4 b2 = *(const f32x4*)(nb + 16), b3 = *(const f32x4*)(nb + 24);
        i32x4 av[NB];
#pragma unroll
        for (int j = 0; j < PF; ++j) av[j] = *(const i32x4*)(arow + j * 16);
        const f32x16 init = {b0[0], b0[1], b0[2], b0[3], b1[0], b1[1], b1[2], b1[3],
                             b2[0], b2[1], b2[2], b2[3], b3[0], b3[1], b3[2], b3[3]};
#pragma unroll
        for (int j = 0; j < NB; ++j) {
            if (j + PF < NB) av[j + PF] = *(const i32x4*)(arow + (j + PF) * 16);
#pragma unroll
            for (int qt = 0; qt < QT; ++qt) acc[qt] = mfma16<PREC>(av[j], bq[qt][j], j == 0 ? init : acc[qt]);
        }
    };
    // hit flags of a finished tile (the branch-free part of the test)
    auto test_tile = [&](const f32x16(&acc)[QT], bool(&hit)[QT]) {
        bool any = false;
#pragma unroll
        for (int qt = 0; qt < QT; ++qt) {
            hit[qt] = max16(acc[qt]) >= thr[qt];
            any |= hit[qt];
        }
        return any;
    };
    // appends of a tile whose test found a hit in some lane
    auto append_tile = [&](const f32x16(&acc)[QT], const bool(&hit)[QT], int64_t r0) {
#pragma unroll
        for (int qt = 0; qt < QT; ++qt) {
            if (!ballot(hit[qt])) continue;
            const int64_t qg = qrow0 + qt * 32 + col;
            const int64_t self = a.exclude ? a.self_off + qg : -1;
            const size_t base = (size_t)qg * a.cap * (size_t)(2 * a.S) + (size_t)split * 2 + h;
            uint32_t regs = 0;
#pragma unroll
            for (int i = 0; i < 16; ++i) regs |= (ballot(acc[qt][i] >= thr[qt]) != 0 ? 1u : 0u) << i;
#pragma unroll 1
            while (regs) {
                const int i = __builtin_ctz(regs);
                regs &= regs - 1;
                const float v = acc[qt][i];
                const int64_t row = r0 + mfma32_row(i, lane);
                if (v >= thr[qt] && row < r_end && row != self) {
                    const int slot = cnt_r[qt]++;
                    if (slot < a.cap) a.cand[base + (size_t)slot * (2 * a.S)] = int2{(int)row, __float_as_int(v - a.bias[row])};
                }
            }
        }
    };

    // prologue: tiles 0 and 1 in flight, wait for tile 0; the "previous tile"
    // of step 0 is an all -inf accumulator set, which no threshold reaches, so
    // every step runs the same straight-line code
    load_tile(0, sA0, sB0);
    if (ntiles > 1) load_tile(1, sA1, sB1);
    sync_tiles(ntiles > 1 ? 1 : 0);
#pragma unroll
    for (int qt = 0; qt < QT; ++qt)
#pragma unroll
        for (int i = 0; i < 16; ++i) accB[qt][i] = -INFINITY;
    // one step: prefetch tile u + 2 into the buffer of tile u - 1, MFMAs of tile
    // u into CUR, test of tile u - 1 in PREV
    auto step = [&](auto BUF, f32x16(&cur)[QT], f32x16(&prev)[QT], int64_t u) {
        constexpr int b = decltype(BUF)::value;
        const bool full = u + 2 < ntiles;
        if (full) load_tile(u + 2, bufA(std::integral_constant<int, (b + 2) % NS>{}),
                            bufB(std::integral_constant<int, (b + 2) % NS>{}));
        bool stored = false;
        if (wave_live) {
            mfma_tile(cur, bufA(BUF), bufB(BUF));
            bool hit[QT];
            if (ballot(test_tile(prev, hit))) {
                append_tile(prev, hit, r_begin + (u - 1) * BN);
                stored = true;
            }
        }
        // stores also count in vmcnt: after one, wait for everything
        sync_tiles(full && !stored ? 1 : 0);
    };
    using I0 = std::integral_constant<int, 0>;
    using I1 = std::integral_constant<int, 1>;
    using I2 = std::integral_constant<int, 2>;
    // buffer u % 3, accumulator set u % 2: six steps per pass of the loop
    for (int64_t u = 0; u < ntiles; u += 6) {
        step(I0{}, accA, accB, u);
        if (u + 1 < ntiles) step(I1{}, accB, accA, u + 1);
        if (u + 2 < ntiles) step(I2{}, accA, accB, u + 2);
        if (u + 3 < ntiles) step(I0{}, accB, accA, u + 3);
        if (u + 4 < ntiles) step(I1{}, accA, accB, u + 4);
        if (u + 5 < ntiles) step(I2{}, accB, accA, u + 5);
    }
    // the last tile's test
    if (wave_live) {
        const f32x16(&last)[QT] = ((ntiles - 1) & 1) ? accB : accA;
        bool hit[QT];
        if (ballot(test_tile(last, hit))) append_tile(last, hit, r_begin + (ntiles - 1) * BN);
    }
#pragma unroll
    for (int qt = 0; qt < QT; ++qt)
        a.cnt[(((size_t)(qrow0 + qt * 32 + col)) * a.S + split) * 2 + h] = cnt_r[qt];
#else
    (void)a;
#endif
}

// Narrow rows: 64-row steps for one and two query tiles per wave (measured on
// the 1M x 128 mixture: two query tiles with 64-row steps, 2.07 ms per call,
// against 2.2 ms with 32-row steps), 32-row steps for four.
template <int PREC, int QT, int NS, int MODE>
static ScreenFn pick_dp(int dp) {
    constexpr int RTW = QT <= 2 ? 2 : 1;
    switch (dp) {
        case 32: return k_screen<PREC, 32, QT, RTW, NS, MODE>;
        case 64: return k_screen<PREC, 64, QT, RTW, NS, MODE>;
        case 96: return k_screen<PREC, 96, QT, RTW, NS, MODE>;
        case 128: return k_screen<PREC, 128, QT, RTW, NS, MODE>;
        default: return nullptr;
    }
}

template <int PREC, int QT>
static ScreenFn pick_pp(int dp) {
    switch (dp) {
        case 32: return k_screen_pp<PREC, 32, QT>;
        case 64: return k_screen_pp<PREC, 64, QT>;
        case 96: return k_screen_pp<PREC, 96, QT>;
        case 128: return k_screen_pp<PREC, 128, QT>;
        default: return nullptr;
    }
}

// Wide rows (bf16 / fp16, padded width in {256, 384, 512, 640, 800}): one
// query tile per wave, its fragments for the whole width in registers, 32-row
// tiles. Shape ns = 3: whole rows per LDS stage (three stages for k-NN lists,
// two beside the pair staging; 51 KB stages at 800 features, so one block per
// CU). Shape ns = 2: rows split in two chunks, stage c holding chunk c of a
// tile (26 KB stages at 800 features), the accumulators carried over the two
// chunks, at most 256 registers: two blocks per CU (k-NN lists only: beside
// the pair staging the 800-feature kernel would spill).
// Deeper rows (multiples of DEEP_CHUNK features): the deep kernel walks each
// 64-row tile in DEEP_CHUNK-feature chunks, reloading the query fragments of
// each chunk (they do not fit in registers for the whole row).
template <int PREC, int MODE>
static ScreenFn pick_wide(int dp, int ns) {
    constexpr int NS = MODE == SCREEN_PAIRS ? 2 : 3;
    if constexpr (MODE == SCREEN_LISTS) {
        if (ns == 2) switch (dp) {
            case 256: return k_screen<PREC, 128, 1, 1, 2, MODE, true, 2>;
            case 384: return k_screen<PREC, 192, 1, 1, 2, MODE, true, 2>;
            case 512: return k_screen<PREC, 256, 1, 1, 2, MODE, true, 2>;
            case 640: return k_screen<PREC, 320, 1, 1, 2, MODE, true, 2>;
            case 800: return k_screen<PREC, 400, 1, 1, 2, MODE, true, 2>;
            default: break;
        }
    }
    switch (dp) {
        case 256: return k_screen<PREC, 256, 1, 1, NS, MODE>;
        case 384: return k_screen<PREC, 384, 1, 1, NS, MODE>;
        case 512: return k_screen<PREC, 512, 1, 1, NS, MODE>;
        case 640: return k_screen<PREC, 640, 1, 1, NS, MODE>;
        case 800: return k_screen<PREC, 800, 1, 1, NS, MODE>;
        default:
            // deeper rows: DEEP_CHUNK-feature chunks, 64-row tiles, two stages
            if (dp > WIDE_MAX && dp % DEEP_CHUNK == 0) return k_screen<PREC, DEEP_CHUNK, 1, 2, 2, MODE, true>;
            return nullptr;
    }
}

template <int PREC, int MODE>
static ScreenFn pick_shape(int dp, int qt, int ns) {
    if (dp > 128) {
        // float32 rows wider than 128 features: 128-feature chunks (512 bytes a
        // row, the bf16 deep chunk's size), 64-row tiles, two stages
        if constexpr (PREC == PREC_FP32)
            return dp % FP32_CHUNK == 0 ? k_screen<PREC, FP32_CHUNK, 1, 2, 2, MODE, true> : nullptr;
        else return pick_wide<PREC, MODE>(dp, ns);
    }
    if constexpr (PREC == PREC_FP32) {
        // float32 operands: two query tiles would need 128 registers of fragments
        if (qt == 1 && ns == 2) return pick_dp<PREC, 1, 2, MODE>(dp);
        if constexpr (MODE == SCREEN_LISTS)
            if (qt == 1 && ns == 3) return pick_dp<PREC, 1, 3, MODE>(dp);
        return nullptr;
    } else {
        if (qt == 1 && ns == 2) return pick_dp<PREC, 1, 2, MODE>(dp);
        // two query tiles: every LDS fragment and bias read feeds two MFMAs
        // (radius pairs at 32 features: two MFMAs per tile and query tile)
        if constexpr (MODE == SCREEN_PAIRS)
            if (qt == 2 && ns == 2) return pick_dp<PREC, 2, 2, MODE>(dp);
        if constexpr (MODE == SCREEN_LISTS) {
            // four query tiles per wave: every LDS fragment feeds four MFMAs and a
            // block's 512 queries share each database tile (two waves per SIMD)
            if (qt == 4 && ns == 3) return pick_dp<PREC, 4, 3, MODE>(dp);
            if (qt == 4 && ns == 2) return pick_dp<PREC, 4, 2, MODE>(dp);
            if (qt == 2 && ns == 4) return pick_dp<PREC, 2, 4, MODE>(dp);
            // three stages, two query tiles: the software-pipelined kernel
            if (qt == 2 && ns == 3) return pick_pp<PREC, 2>(dp);
            if (qt == 2 && ns == 2) return pick_dp<PREC, 2, 2, MODE>(dp);
        }
        return nullptr;
    }
}

int screen_dp(int prec, int d) {
    const int p32 = (d + 31) / 32 * 32;
    if (p32 <= 128) return p32;
    if (prec == PREC_FP32) return (d + FP32_CHUNK - 1) / FP32_CHUNK * FP32_CHUNK;  // float32 deep kernel
    for (int w : {256, 384, 512, 640, WIDE_MAX})
        if (d <= w) return w;
    return (d + DEEP_CHUNK - 1) / DEEP_CHUNK * DEEP_CHUNK;  // deep kernels
}

// Measured: for bf16 / fp16 rows of at most 128 features, two query tiles per
// wave with 64-row steps: k-NN lists on the 1M x 128 mixture 2.07 ms per call
// against 2.28 ms with one; radius pairs on the 2M x 32 clumped workload 37.8 ms
// per call against 41.6 ms (at 32 features a tile's bias and fragment reads and
// its barrier then serve four MFMAs instead of two). float32 operands keep one
// query tile (two would need 128 registers of query fragments).
ScreenShape default_shape(int mode, int prec, int dp) {
    // wide k-NN rows: two chunks per row, two blocks per CU (60k x 784 pixels, 10,000
    // queries: 1.36 ms per call against 2.10 ms with whole-row stages)
    if (prec != PREC_FP32 && dp > 128 && dp <= WIDE_MAX) return ScreenShape{1, mode == SCREEN_LISTS ? 2 : 3};
    if (prec != PREC_FP32 && dp <= 128) return ScreenShape{2, 2};
    return ScreenShape{1, 2};
}

ScreenFn pick_screen(int mode, int prec, int dp, ScreenShape shape) {
    if (mode == SCREEN_LISTS) {
        if (prec == PREC_FP32) return pick_shape<PREC_FP32, SCREEN_LISTS>(dp, shape.qt, shape.ns);
        if (prec == PREC_BF16) return pick_shape<PREC_BF16, SCREEN_LISTS>(dp, shape.qt, shape.ns);
        return pick_shape<PREC_FP16, SCREEN_LISTS>(dp, shape.qt, shape.ns);
    }
    if (prec == PREC_FP32) return pick_shape<PREC_FP32, SCREEN_PAIRS>(dp, shape.qt, shape.ns);
    if (prec == PREC_BF16) return pick_shape<PREC_BF16, SCREEN_PAIRS>(dp, shape.qt, shape.ns);
    return pick_shape<PREC_FP16, SCREEN_PAIRS>(dp, shape.qt, shape.ns);
}

std::vector<ScreenShape> screen_shapes(int mode, int prec, int dp) {
    std::vector<ScreenShape> out;
    if (dp > 128) {  // wide rows: whole-row stages and two-chunk stages; deeper rows: one shape
        const ScreenShape first = default_shape(mode, prec, dp);
        if (pick_screen(mode, prec, dp, first)) out.push_back(first);
        if (mode == SCREEN_LISTS && prec != PREC_FP32 && dp <= WIDE_MAX) {
            const ScreenShape other{1, first.ns == 2 ? 3 : 2};
            if (pick_screen(mode, prec, dp, other)) out.push_back(other);
        }
        return out;
    }
    for (int qt : {1, 2, 4})
        for (int ns = 2; ns <= 4; ++ns)
            if (pick_screen(mode, prec, dp, ScreenShape{qt, ns})) out.push_back(ScreenShape{qt, ns});
    return out;
}

}  // namespace mlk
