// Fused MFMA screen shared by the k-NN and radius paths: certified Gram
// bounds, the per-query threshold helpers, and the screen launch interface
// (the kernel itself lives in mlk_screen.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "mlk_device.h"
#include "mlk_internal.h"

namespace mlk {

struct Consts {
    double H;       // relative error constant of the stored-hit bounds (hit_bounds)
    double G;       // H + gamma + 2^-40
    double gamma;   // accumulation constant
    double f0;      // absolute floor (flushed subnormal partial sums)
    double scale2;  // sigma^2 of the screen mirror
    // screen test (separable rounding bound): splitting parameters of the
    // query-norm x row-residual and query-residual x row-norm products (0 when
    // the side whose residual they weigh is exact), and the relative error of
    // the float64 norms
    double t1, t2, eps;
};

// The screen's separable bound. For a query q and a row x with screen values
// q~, x~ (exact), residuals dq = q - q~, dx = x - x~ and accumulator
// acc = fl32(-a_x + sum q~_i x~_i) (any order, any rounding: gamma, plus f0
// for flushed subnormals),
//   |q.x - q~.x~| <= |q~||dx| + |dq||x~| + |dq||dx|   (Cauchy-Schwarz)
//   |acc - (q~.x~ - a_x)| <= gamma (|a_x| + |q~||x~|) + f0,
// and 2ab <= t a^2 + b^2 / t splits every product into a query part alpha and a
// row part beta:
//   gamma |q~||x~| + |q~||dx| + |dq||x~| + |dq||dx| <= alpha(q) + beta(x),
//   alpha = (gamma + t1) |q~|^2 / 2 + (1 / t2 + 1) |dq|^2 / 2,
//   beta  = (gamma + t2) |x~|^2 / 2 + (1 / t1 + 1) |dx|^2 / 2,
// with the |dq|^2 term dropped when every query residual is zero (t2 = 0) and
// the |dx|^2 term when every row residual is zero (t1 = 0). The squared norms
// nt = |v~|^2 and d2 = |v - v~|^2 are float64 sums (relative error <= eps).
__host__ __device__ __forceinline__ double screen_alpha(double nt, double d2, const Consts& c) {
    const double T2 = nt * (1.0 + c.eps), D2 = d2 * (1.0 + c.eps);
    double a = (c.gamma + c.t1) * T2 * 0.5;
    if (D2 > 0.0) a += c.t2 > 0.0 ? (1.0 / c.t2 + 1.0) * D2 * 0.5 : INFINITY;
    return a * (1.0 + 0x1p-50);
}
__host__ __device__ __forceinline__ double screen_beta(double nt, double d2, const Consts& c) {
    const double T2 = nt * (1.0 + c.eps), D2 = d2 * (1.0 + c.eps);
    double b = (c.gamma + c.t2) * T2 * 0.5;
    if (D2 > 0.0) b += c.t1 > 0.0 ? (1.0 / c.t1 + 1.0) * D2 * 0.5 : INFINITY;
    return b * (1.0 + 0x1p-50);
}
// Query part of the screen threshold: qa <= nq* (1 - eps) - 2 alpha, rounded
// down (nq = float64 squared norm of the scaled centred query).
__host__ __device__ __forceinline__ double screen_qa(double nq, double nt, double d2, const Consts& c) {
    const double al = screen_alpha(nt, d2, c);
    const double v = nq * (1.0 - c.eps) - 2.0 * al;
    return v - 0x1p-50 * (fabs(nq) + 2.0 * al);
}
// The row's accumulator start -a_x: a_x(1 + gamma) <= (nx* (1 - eps) - 2 beta) / 2
// (a_x(1 - gamma) when negative), rounded down to float32, so that
//   sigma^2 S >= 2 (qa / 2 - f0 - acc)   for every row,
// and acc >= thr = rd((qa - tau) / 2 - f0) holds for every row with sigma^2 S <= tau.
__host__ __device__ __forceinline__ double screen_row_bound(double nx, double nt, double d2, const Consts& c) {
    const double be = screen_beta(nt, d2, c);
    const double v = nx * (1.0 - c.eps) - 2.0 * be;
    const double lo = v - 0x1p-50 * (fabs(nx) + 2.0 * be);
    return lo >= 0.0 ? lo * 0.5 / (1.0 + c.gamma) * (1.0 - 0x1p-50) : lo * 0.5 / (1.0 - c.gamma) * (1.0 + 0x1p-50);
}

// Lowest screen threshold. A safe row's accumulator has magnitude below 2^125
// (the mirror's safe range bounds d |v|^2 by 2^124), an unsafe row's is -3e38
// (its bias; its mirror values are zero): a threshold of THR_FLOOR passes every
// safe row and no unsafe one, so unsafe rows reach a query only through the
// unsafe-row lists, never twice.
constexpr float THR_FLOOR = -1.0e38f;

// Screen threshold of a query from its tau (an upper bound of its k-th smallest
// sigma^2 S) and its qa: acc >= thr <=> the row may lie within tau. Thresholds
// below THR_FLOOR (an infinite tau, or one beyond float32) are raised to it.
__device__ __forceinline__ float thr_from_tau(double tau, double qa, const Consts& c) {
    if (!(tau < INFINITY)) return THR_FLOOR;
    const float t = __double2float_rd((qa - tau * (1.0 + 0x1p-46)) * 0.5 - c.f0);
    return t > THR_FLOOR ? t : THR_FLOOR;
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
// binary search on the bit pattern, counting with ballots (no shuffles).
// Up to 256 values stay in registers for the whole search. Returns the
// largest pattern with fewer than k values strictly below it (+inf bits if
// there are fewer than k values).
template <class GetFn>
__device__ uint32_t wave_kth_bits(int count, int k, GetFn get) {
    const int lane = threadIdx.x & 63;
    uint32_t ans = 0;
    if (count <= 256) {
        uint32_t v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int j = lane + 64 * u;
            v[u] = j < count ? get(j) : 0xffffffffu;
        }
        for (int b = 31; b >= 0; --b) {
            const uint32_t trial = ans | (1u << b);
            int below = 0;
#pragma unroll
            for (int u = 0; u < 4; ++u) below += __popcll(ballot(v[u] < trial));
            if (below < k) ans = trial;
        }
        return ans;
    }
    for (int b = 31; b >= 0; --b) {
        const uint32_t trial = ans | (1u << b);
        int below = 0;
        for (int j0 = 0; j0 < count; j0 += 64) {
            const int j = j0 + lane;
            below += __popcll(ballot(j < count && get(j) < trial));
        }
        if (below < k) ans = trial;
    }
    return ans;
}

// ---------------------------------------------------------------- screen
struct ScreenArgs {
    const void* xs;      // database screen mirror, n x DP
    const float* bias;   // per row: -a_x (accumulator start value)
    const float* nx32;   // per row: squared norm rounded to float32
    const void* qs;      // query screen mirror, m_pad x DP
    const float* thr;    // per query: hit threshold on the accumulator
    int* cnt;            // per (query, split, lane half): appended count
    int2* cand;          // candidates: (row, float32 bits of p = acc + a_x, the screen's dot product)
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
    // rows wider than the widest register-held query (deep kernels): the row
    // is nchunk chunks of the kernel's chunk width; 1 for every other kernel
    int nchunk;
    // with xcd_map: query blocks per group of an XCD's consecutive blocks
    // (0 or 1: every query block of a split before the next split)
    int qgroup;
    // radius mode: query position -> original query index (nullptr: the
    // identity), for the self-exclusion test of regrouped queries
    const int* qmap;
};

// Block order of a screen launch (sets xcd_map; S, QB and qgroup must be set):
// with at least eight splits, the launch's work items (query block, split),
// split-major, are dealt to the eight XCDs in contiguous shares, so the
// blocks an XCD runs together stream the same database rows through its L2
// (deep rows group query blocks too, and need a multiple of eight splits).
// Returns the grid size; blocks past the last item exit at once.
inline int64_t screen_grid(ScreenArgs& a) {
    a.xcd_map = a.S >= 8 && (a.qgroup <= 1 || a.S % 8 == 0);
    const int64_t items = (int64_t)a.S * a.QB;
    return a.xcd_map ? (items + 7) / 8 * 8 : items;
}

constexpr int WIDE_MAX = 800;    // widest row whose query fragments stay in registers (bf16 / fp16)
constexpr int DEEP_CHUNK = 256;  // features per chunk of the bf16 / fp16 deep kernels
constexpr int FP32_CHUNK = 128;  // features per chunk of the float32 deep kernel (512 bytes, as DEEP_CHUNK halves)
constexpr int DEEP_QGROUP = 8;   // query blocks per XCD block group of the deep kernels

constexpr int SCREEN_LISTS = 0;  // k-NN: per-(query, split) candidate lists
constexpr int SCREEN_PAIRS = 1;  // radius: global (query, row) pair buffer
constexpr int PAIR_STAGE = 512;  // per-wave LDS staging of radius pairs

typedef void (*ScreenFn)(const ScreenArgs);

// Screen shape: QT query tiles of 32 per wave (1: 64-row tiles, 2: 32-row
// tiles) and an NS-stage LDS tile ring.
struct ScreenShape {
    int qt, ns;
};

// Measured default shape for a screen mode (SCREEN_LISTS / SCREEN_PAIRS), a
// precision and a padded width.
ScreenShape default_shape(int mode, int prec, int dp);

// Padded width of the screen mirror: a multiple of 32 up to 128; above that,
// bf16 / fp16 use the wide-kernel widths 256, 384, 512, 640, 800 and then
// multiples of DEEP_CHUNK, float32 multiples of FP32_CHUNK.
int screen_dp(int prec, int d);

// Features per chunk of a padded width (the deep kernels walk a row in chunks
// whose query fragments fit in registers): dp itself for the narrow and wide
// kernels.
inline int screen_chunk(int prec, int dp) {
    if (prec == PREC_FP32) return dp > 128 ? FP32_CHUNK : dp;
    return dp > WIDE_MAX ? DEEP_CHUNK : dp;
}
inline int screen_chunks(int prec, int dp) { return dp / screen_chunk(prec, dp); }

// Chunks per row of the kernel that a plan's shape selects: the deep kernels
// walk rows in chunks (query fragments reloaded per chunk); wide bf16 / fp16
// rows screened for k-NN lists with two stages are split in two chunks whose
// query fragments all stay in registers (stage c multiplies chunk c).
inline int screen_nchunk(int mode, int prec, int dp, ScreenShape shape) {
    if (prec != PREC_FP32 && dp > 128 && dp <= WIDE_MAX) return mode == SCREEN_LISTS && shape.ns == 2 ? 2 : 1;
    return screen_chunks(prec, dp);
}
// Query blocks per XCD block group: the deep kernels, which reload query
// fragments per chunk, group their blocks so resident ones share query rows.
inline int screen_qgroup(int prec, int dp) { return screen_chunks(prec, dp) > 1 ? DEEP_QGROUP : 1; }

// The screen precision a request runs: every request runs the precision it
// names, at every width, for k-NN and radius search alike.
inline int effective_prec(int prec, int d) {
    (void)d;
    return prec;
}
inline int radius_prec(int prec, int d) { return effective_prec(prec, d); }

// Launch function of the screen for a padded width dp in {32, 64, 96, 128};
// nullptr when the shape is not built for that mode and precision. Blocks
// have 256 threads and cover 4 * qt * 32 queries.
ScreenFn pick_screen(int mode, int prec, int dp, ScreenShape shape);

// Every shape built for a mode, precision and padded width.
std::vector<ScreenShape> screen_shapes(int mode, int prec, int dp);

// ---------------------------------------------------------------- plans
// How a k-NN or radius call tiles its screen: the shape, the database splits
// of each round, the candidate slots per (query, split, lane half) sub-list
// and the cumulative round fractions (k-NN only). Results never depend on
// the plan; only the time does.
struct Plan {
    ScreenShape shape{1, 2};
    int splits = 1;
    int cap = 0;
    std::vector<double> rounds;
};
enum PlanKind { PLAN_KNN = 0, PLAN_RADIUS = 1 };

// The measured default plan for a call's shape. `shape` (qt > 0) replaces the
// default screen shape; the splits follow from the query block it implies.
Plan heuristic_plan(int kind, int prec, int dp, int64_t n, int64_t m, int k, ScreenShape shape = {0, 0});

// Candidate slots per sub-list: at least 6k per query over its 2 * splits
// sub-lists, with a per-precision floor.
int knn_capacity(int prec, int splits, int k);

// The plan a call uses: the autotuner's trial plan of the calling thread if
// one is set, else the tuned plan cached for the call's shape (kind,
// precision, padded width, k, n and m in half-octave buckets), else the
// heuristic; mlk_set_tuning's non-zero knobs override fields of the last two.
// *tuned reports whether a cached plan was used.
Plan resolve_plan(int kind, int prec, int dp, int64_t n, int64_t m, int k, bool* tuned = nullptr);

// An autotuner trial call on this thread whose exact fallback would take nf of
// its m queries (more than m / 256 + 16, a cost no plan recovers from) is cut
// short: the call skips the fallback (its outputs are the tuner's scratch) and
// the tuner rejects the plan. Returns whether the call is to stop there.
bool trial_abandon_fallback(int64_t nf, int64_t m);

}  // namespace mlk
