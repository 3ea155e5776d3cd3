// Screen plans and the autotuner.
//
// Reference counterpart: the tile planner resolve_blocks
// (src/mlxlearn_device/_neighbors_plan.py:74-126) and the per-GPU-family
// selection tables (_neighbors_general.py:80-170), which fix tile sizes and
// selection strategies from a device model. Here a measured heuristic gives
// the default plan, and mlk_autotune times alternatives on the caller's own
// queries and caches the fastest per call shape. Only time depends on the
// plan: every plan returns the same bits.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <limits>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <mutex>
#include <sstream>
#include <string>

#include "mlk_internal.h"
#include "mlk_order.h"
#include "mlk_screen.h"

namespace mlk {

int knn_capacity(int prec, int splits, int k) {
    const int lists = 2 * splits;
    // few lists per query (small databases, k-means centroids): more slots each
    const int base = (prec == PREC_FP32 ? 24 : 48) * (lists <= 16 ? 2 : 1);
    const int need = (6 * k + lists - 1) / lists + 4;
    return std::max(base, need);
}

// Geometric rounds: each threshold comes from four to eight times fewer rows
// than the round it filters. Measured on the 1M x 128 mixture: five rounds
// beat seven (each launch and merge has a fixed cost that outweighs the extra
// candidates of the coarser steps).
static std::vector<double> default_rounds() { return {1.0 / 256, 1.0 / 32, 1.0 / 8, 1.0 / 2, 1.0}; }

static int round_splits(int64_t want, int64_t n) {
    want = std::min<int64_t>(want, std::max<int64_t>(1, n / 128));  // at least 128 rows per split
    want = std::max<int64_t>(want, 1);
    if (want >= 8) want = (want + 7) / 8 * 8;  // whole XCD groups
    return (int)std::min<int64_t>(want, 4096);
}

Plan heuristic_plan(int kind, int prec, int dp, int64_t n, int64_t m, int k, ScreenShape shape) {
    Plan p;
    p.shape = shape.qt > 0 ? shape : default_shape(kind == PLAN_KNN ? SCREEN_LISTS : SCREEN_PAIRS, prec, dp);
    const int64_t BQ = 4 * p.shape.qt * 32;
    const int64_t qb = std::max<int64_t>(1, (m + BQ - 1) / BQ);
    if (kind == PLAN_KNN) {
        // about 10 blocks per CU in every round: fewer splits mean fewer
        // sub-lists to merge and rank (tuned 1M x 128, 10,000 queries: 32 splits)
        p.splits = round_splits((2560 + qb - 1) / qb, n);
        p.cap = knn_capacity(prec, p.splits, k);
        // calls large enough to be regrouped start from a threshold sampled in
        // the queries' own neighbourhoods: four rounds (tuned 1M x 128, 10,000
        // queries: 1/64, 1/8, 1/2, 1)
        const bool regrouped = m >= ORDER_MIN_QUERIES && m <= ORDER_MAX_QUERIES && n >= ORDER_MIN_ROWS &&
                               dp <= ORDER_MAX_D;
        p.rounds = regrouped ? std::vector<double>{1.0 / 64, 1.0 / 8, 1.0 / 2, 1.0} : default_rounds();
    } else {
        // one pass; about 128 blocks per CU so that the query blocks with many
        // neighbours (dense clumps, grouped together by the query order) do
        // not leave a long tail (2M x 32, 132,759 regrouped queries: 32 splits,
        // a 21.5 ms screen against 24 ms at 8)
        p.splits = round_splits((32768 + qb - 1) / qb, n);
    }
    return p;
}

// ------------------------------------------------------------ plan cache
typedef std::array<int64_t, 6> PlanKey;

static int half_octave(int64_t x) { return x <= 1 ? 0 : (int)std::lround(2.0 * std::log2((double)x)); }

static PlanKey plan_key(int kind, int prec, int dp, int64_t n, int64_t m, int k) {
    return PlanKey{kind, prec, dp, kind == PLAN_KNN ? k : 0, half_octave(n), half_octave(m)};
}

struct TunedEntry {
    Plan plan;
    double ms = 0.0, default_ms = 0.0;
    int trials = 0;
};

static std::mutex g_plan_lock;
static std::map<PlanKey, TunedEntry> g_plans;
static thread_local const Plan* t_trial = nullptr;
static thread_local bool t_trial_abandoned = false;
static thread_local bool t_trial_may_abandon = false;  // alternatives only, never the starting plan

bool trial_abandon_fallback(int64_t nf, int64_t m) {
    if (!t_trial || !t_trial_may_abandon || nf <= m / 256 + 16) return false;
    t_trial_abandoned = true;
    return true;
}

static bool lookup(const PlanKey& key, TunedEntry* out) {
    std::lock_guard<std::mutex> g(g_plan_lock);
    auto it = g_plans.find(key);
    if (it == g_plans.end()) return false;
    if (out) *out = it->second;
    return true;
}

Plan resolve_plan(int kind, int prec, int dp, int64_t n, int64_t m, int k, bool* tuned) {
    if (tuned) *tuned = false;
    if (t_trial) return *t_trial;
    const Tuning t = tuning_snapshot();
    ScreenShape knob{0, 0};
    if (dp <= 128 && (t.query_tiles || t.stages)) {
        knob = default_shape(kind == PLAN_KNN ? SCREEN_LISTS : SCREEN_PAIRS, prec, dp);
        if (t.query_tiles) knob.qt = t.query_tiles;
        if (t.stages) knob.ns = t.stages;
    } else if (kind == PLAN_KNN && prec != PREC_FP32 && dp <= WIDE_MAX && t.stages) {
        // wide rows: one query tile; two stages = rows in two chunks, three = whole rows
        knob = ScreenShape{1, t.stages};
    }
    Plan p;
    TunedEntry e;
    if (lookup(plan_key(kind, prec, dp, n, m, k), &e)) {
        p = e.plan;
        if (tuned) *tuned = true;
        if (knob.qt) p.shape = knob;
    } else {
        p = heuristic_plan(kind, prec, dp, n, m, k, knob);
    }
    if (t.splits) {
        p.splits = t.splits;
        if (kind == PLAN_KNN) p.cap = knn_capacity(prec, p.splits, k);
    }
    if (kind == PLAN_KNN) {
        if (t.capacity) p.cap = t.capacity;
        if (!t.rounds.empty()) p.rounds = t.rounds;
    }
    return p;
}

static void fill_plan(const Plan& p, mlk_plan* out) {
    out->splits = p.splits;
    out->capacity = p.cap;
    out->query_tiles = p.shape.qt;
    out->stages = p.shape.ns;
    out->n_rounds = (int32_t)std::min<size_t>(p.rounds.size(), MLK_MAX_ROUNDS);
    for (int i = 0; i < MLK_MAX_ROUNDS; ++i) out->rounds[i] = i < out->n_rounds ? p.rounds[i] : 0.0;
}

static bool same_plan(const Plan& a, const Plan& b) {
    return a.shape.qt == b.shape.qt && a.shape.ns == b.shape.ns && a.splits == b.splits && a.cap == b.cap &&
           a.rounds == b.rounds;
}

// device bytes of a k-NN plan's candidate lists for m queries
static double list_bytes(const Plan& p, int64_t m) {
    const int64_t BQ = 4 * p.shape.qt * 32;
    const int64_t m_pad = (m + BQ - 1) / BQ * BQ;
    return (double)m_pad * 2.0 * p.splits * p.cap * 8.0;
}

int plan_info_impl(const mlk_index* index, int64_t m, int k, uint32_t flags, mlk_plan* out) {
    const int kind = (flags & MLK_TUNE_RADIUS) ? PLAN_RADIUS : PLAN_KNN;
    const int req = (int)((flags & MLK_PREC_MASK) >> 4);
    const int prec = kind == PLAN_RADIUS ? radius_prec(req, index->d) : effective_prec(req, index->d);
    const int dp = screen_dp(prec, index->d);
    *out = mlk_plan{};
    bool tuned = false;
    const Plan p = resolve_plan(kind, prec, dp, index->n, m, k, &tuned);
    fill_plan(p, out);
    out->tuned = tuned ? 1 : 0;
    TunedEntry e;
    if (tuned && lookup(plan_key(kind, prec, dp, index->n, m, k), &e)) {
        out->tuned_ms = e.ms;
        out->default_ms = e.default_ms;
        out->trials = e.trials;
    }
    return MLK_OK;
}

void clear_plans_impl() {
    std::lock_guard<std::mutex> g(g_plan_lock);
    g_plans.clear();
}

// Text form of the plan cache, one plan per line:
//   plan <kind> <prec> <dp> <k> <n bucket> <m bucket> <qt> <ns> <splits> <cap>
//        <rounds> <fraction>... <tuned ms> <default ms> <trials>
// (fractions and times as hexadecimal floating point, so they round-trip).
std::string export_plans_impl() {
    std::lock_guard<std::mutex> g(g_plan_lock);
    std::string out;
    char buf[128];
    for (const auto& kv : g_plans) {
        const PlanKey& key = kv.first;
        const TunedEntry& e = kv.second;
        out += "plan";
        for (int64_t v : key) {
            snprintf(buf, sizeof buf, " %lld", (long long)v);
            out += buf;
        }
        snprintf(buf, sizeof buf, " %d %d %d %d %zu", e.plan.shape.qt, e.plan.shape.ns, e.plan.splits, e.plan.cap,
                 e.plan.rounds.size());
        out += buf;
        for (double f : e.plan.rounds) {
            snprintf(buf, sizeof buf, " %a", f);
            out += buf;
        }
        snprintf(buf, sizeof buf, " %a %a %d\n", e.ms, e.default_ms, e.trials);
        out += buf;
    }
    return out;
}

int import_plans_impl(const char* text, size_t len) {
    std::map<PlanKey, TunedEntry> add;
    std::string all(text, len);
    size_t pos = 0;
    int line_no = 0;
    while (pos < all.size()) {
        size_t end = all.find('\n', pos);
        if (end == std::string::npos) end = all.size();
        const std::string line = all.substr(pos, end - pos);
        pos = end + 1;
        ++line_no;
        if (line.empty()) continue;
        std::istringstream in(line);
        std::string tag;
        PlanKey key{};
        TunedEntry e;
        size_t nr = 0;
        in >> tag;
        for (auto& v : key) in >> v;
        in >> e.plan.shape.qt >> e.plan.shape.ns >> e.plan.splits >> e.plan.cap >> nr;
        bool ok = tag == "plan" && !in.fail() && nr <= MLK_MAX_ROUNDS;
        std::string tok;
        for (size_t i = 0; ok && i < nr; ++i) {
            ok = static_cast<bool>(in >> tok);
            if (ok) e.plan.rounds.push_back(std::strtod(tok.c_str(), nullptr));
        }
        std::string t1, t2;
        ok = ok && static_cast<bool>(in >> t1 >> t2 >> e.trials);
        if (ok) {
            e.ms = std::strtod(t1.c_str(), nullptr);
            e.default_ms = std::strtod(t2.c_str(), nullptr);
        }
        // the plan must name a built screen shape and a well-formed schedule
        const int kind = (int)key[0], prec = (int)key[1], dp = (int)key[2];
        ok = ok && (kind == PLAN_KNN || kind == PLAN_RADIUS) && prec >= PREC_FP32 && prec <= PREC_FP16 &&
             e.plan.splits >= 1 && e.plan.splits <= 4096 && (kind == PLAN_RADIUS || e.plan.cap >= 1) &&
             pick_screen(kind == PLAN_KNN ? SCREEN_LISTS : SCREEN_PAIRS, prec, dp, e.plan.shape) != nullptr;
        double prev = 0.0;
        for (double f : e.plan.rounds) {
            ok = ok && f > prev && f <= 1.0;
            prev = f;
        }
        ok = ok && (kind == PLAN_RADIUS ? e.plan.rounds.empty() : (!e.plan.rounds.empty() && prev == 1.0));
        if (!ok) {
            set_error("plan text, line " + std::to_string(line_no) + ": not a valid plan");
            return MLK_EINVAL;
        }
        add[key] = e;
    }
    std::lock_guard<std::mutex> g(g_plan_lock);
    for (const auto& kv : add) g_plans[kv.first] = kv.second;
    return MLK_OK;
}

// Coordinate descent over the plan: screen shape, then database splits (a
// line search by factors of two each way, then counts that fill whole waves
// of resident blocks), then the round schedule, then (if queries fell back)
// list capacity. A plan replaces the best one only when it
// is at least 2% faster; each plan is timed as the best of two calls, after
// three untimed calls (the first builds the mirrors and workspace; the others
// let the device's clocks settle, which otherwise inflates the default plan's
// time, the first one taken). A winner other than the default is timed again
// beside the default before it is kept.
int autotune_impl(mlk_index* index, const float* Q, int64_t m, int k, double R2, uint32_t flags, int max_trials,
                  mlk_plan* out, hipStream_t stream) {
    const int kind = (flags & MLK_TUNE_RADIUS) ? PLAN_RADIUS : PLAN_KNN;
    const uint32_t qflags = flags & (MLK_PREC_MASK | MLK_SORT_RESULTS | MLK_NEED_DISTANCE);
    const int req = (int)((flags & MLK_PREC_MASK) >> 4);
    const int prec = kind == PLAN_RADIUS ? radius_prec(req, index->d) : effective_prec(req, index->d);
    const int dp = screen_dp(prec, index->d);
    const int64_t n = index->n;
    const int mode = kind == PLAN_KNN ? SCREEN_LISTS : SCREEN_PAIRS;
    const std::vector<ScreenShape> shapes = screen_shapes(mode, prec, dp);
    *out = mlk_plan{};
    const Plan base = heuristic_plan(kind, prec, dp, n, m, k);
    if (shapes.empty() || (kind == PLAN_KNN && k > 64)) {
        // no screen for this call (exact scan only): nothing to tune
        fill_plan(base, out);
        return MLK_OK;
    }

    DevBuf idx, dist, indptr;
    if (kind == PLAN_KNN) {
        MLK_TRY(idx.reserve((size_t)m * k * sizeof(int64_t)));
        MLK_TRY(dist.reserve((size_t)m * k * sizeof(double)));
    } else {
        MLK_TRY(indptr.reserve((size_t)(m + 1) * sizeof(int64_t)));
    }
    Events events;
    hipEvent_t e0, e1;
    MLK_CHECK_HIP(events.make(&e0));
    MLK_CHECK_HIP(events.make(&e1));
    struct TrialGuard {
        explicit TrialGuard(const Plan* p) { t_trial = p; }
        ~TrialGuard() { t_trial = nullptr; }
    };

    auto call = [&](const Plan& p) -> int {
        TrialGuard tg(&p);
        if (kind == PLAN_KNN)
            return knn_query_impl(index, Q, m, k, 0, qflags, dist.as<double>(), idx.as<int64_t>(), stream);
        int64_t nnz = 0;
        mlk_radius_result* res = nullptr;
        const int rc = radius_query_impl(index, Q, m, R2, 0, qflags, indptr.as<int64_t>(), &nnz, &res, stream);
        delete res;
        return rc;
    };
    int trials = 0;
    auto time_plan = [&](const Plan& p, double& ms, int64_t& fallback, bool may_abandon) -> int {
        ms = std::numeric_limits<double>::infinity();
        t_trial_abandoned = false;
        t_trial_may_abandon = may_abandon;
        for (int rep = 0; rep < 2; ++rep) {
            MLK_CHECK_HIP(hipEventRecord(e0, stream));
            MLK_TRY(call(p));
            MLK_CHECK_HIP(hipEventRecord(e1, stream));
            MLK_CHECK_HIP(hipEventSynchronize(e1));
            if (t_trial_abandoned) {  // too many fallback queries: the plan loses
                ms = std::numeric_limits<double>::infinity();
                break;
            }
            float t = 0.0f;
            MLK_CHECK_HIP(hipEventElapsedTime(&t, e0, e1));
            ms = std::min(ms, (double)t);
        }
        t_trial_may_abandon = false;
        fallback = stats().fallback_queries;
        ++trials;
        return MLK_OK;
    };

    for (int warm = 0; warm < 3; ++warm) MLK_TRY(call(base));  // untimed: mirrors, biases, workspace, clocks
    Plan best = base;
    double best_ms = 0.0;
    int64_t best_fb = 0;
    MLK_TRY(time_plan(best, best_ms, best_fb, false));
    const double default_ms = best_ms;
    const double mem_limit = std::max(4.0 * list_bytes(base, m), 1024.0 * 1024.0 * 1024.0);
    auto consider = [&](const Plan& p) -> int {
        if (trials >= max_trials || same_plan(p, best)) return 0;
        if (!pick_screen(mode, prec, dp, p.shape)) return 0;
        if (kind == PLAN_KNN && list_bytes(p, m) > mem_limit) return 0;
        double ms = 0.0;
        int64_t fb = 0;
        MLK_TRY(time_plan(p, ms, fb, true));
        if (ms < 0.98 * best_ms) {
            best = p;
            best_ms = ms;
            best_fb = fb;
            return 1;
        }
        return 0;
    };
    auto check = [](int r) { return r < 0 ? r : MLK_OK; };

    // 1. screen shape (splits re-derived for its query block)
    for (const ScreenShape& s : shapes) {
        const int r = consider(heuristic_plan(kind, prec, dp, n, m, k, s));
        MLK_TRY(check(r));
    }
    // 2. splits
    const int max_splits = (int)std::min<int64_t>(4096, std::max<int64_t>(1, n / 32));
    for (int dir : {2, -2}) {
        for (int step = 0; step < 4; ++step) {
            Plan p = best;
            int s = dir > 0 ? p.splits * 2 : p.splits / 2;
            if (s >= 8) s = (s + 7) / 8 * 8;
            s = std::max(1, std::min(s, max_splits));
            if (s == best.splits) break;
            p.splits = s;
            if (kind == PLAN_KNN) p.cap = knn_capacity(prec, s, k);
            const int r = consider(p);
            MLK_TRY(check(r));
            if (r != 1) break;
        }
    }
    // 2b. splits that fill whole waves of resident blocks: the last wave of a
    // launch otherwise runs part-empty (its tail)
    {
        int dev = 0, cus = 0, per_cu = 0;
        MLK_CHECK_HIP(hipGetDevice(&dev));
        MLK_CHECK_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        ScreenFn fn = pick_screen(mode, prec, dp, best.shape);
        if (fn && hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(fn), 256, 0) ==
                      hipSuccess && per_cu > 0 && cus > 0) {
            const int64_t resident = (int64_t)cus * per_cu;
            const int64_t BQ = 4 * best.shape.qt * 32;
            const int64_t qb = std::max<int64_t>(1, (m + BQ - 1) / BQ);
            const double waves = (double)qb * best.splits / (double)resident;
            const Plan around = best;
            for (int64_t j : {(int64_t)std::floor(waves), (int64_t)std::ceil(waves), (int64_t)std::ceil(waves) + 1}) {
                if (j < 1) continue;
                const int64_t fit = j * resident / qb;  // blocks just within j waves
                for (int64_t sp : {fit, fit / 8 * 8}) {
                    if (sp < 1 || sp > max_splits || sp == around.splits) continue;
                    Plan p = around;
                    p.splits = (int)sp;
                    if (kind == PLAN_KNN) p.cap = knn_capacity(prec, p.splits, k);
                    MLK_TRY(check(consider(p)));
                }
            }
        }
    }
    if (kind == PLAN_KNN) {
        // 3. round schedule
        const std::vector<std::vector<double>> schedules = {
            {1.0 / 64, 1.0 / 8, 1.0 / 2, 1.0},
            {1.0 / 1024, 1.0 / 128, 1.0 / 16, 1.0 / 4, 1.0},
            {1.0 / 256, 1.0 / 64, 1.0 / 16, 1.0 / 4, 1.0},
            {1.0 / 16, 1.0 / 2, 1.0},
        };
        for (const auto& rs : schedules) {
            Plan p = best;
            p.rounds = rs;
            MLK_TRY(check(consider(p)));
        }
        // one or two rounds: they win where the start threshold is already
        // close (regrouped queries sample their own neighbourhood) and where
        // calls of well under a millisecond are bound by launches
        for (const auto& rs : std::vector<std::vector<double>>{{1.0 / 8, 1.0}, {1.0 / 4, 1.0}, {1.0}}) {
            Plan p = best;
            p.rounds = rs;
            MLK_TRY(check(consider(p)));
        }
        // 4. capacity, when candidate lists overflowed into the exact fallback
        for (int step = 0; step < 2 && best_fb > 0; ++step) {
            Plan p = best;
            p.cap *= 2;
            const int r = consider(p);
            MLK_TRY(check(r));
            if (r != 1) break;
        }
    }

    double base_ms = default_ms;
    if (!same_plan(best, base)) {
        // confirmation: both timed once more, side by side; each keeps its
        // best time, and the winner must still be 2% faster
        // (not counted as trials: max_trials bounds the search)
        const int searched = trials;
        double again = 0.0;
        int64_t fb = 0;
        MLK_TRY(time_plan(base, again, fb, false));
        base_ms = std::min(base_ms, again);
        MLK_TRY(time_plan(best, again, fb, false));
        best_ms = std::min(best_ms, again);
        trials = searched;
        if (!(best_ms < 0.98 * base_ms)) {
            best = base;
            best_ms = base_ms;
        }
    }
    TunedEntry e;
    e.plan = best;
    e.ms = best_ms;
    e.default_ms = base_ms;
    e.trials = trials;
    {
        std::lock_guard<std::mutex> g(g_plan_lock);
        g_plans[plan_key(kind, prec, dp, n, m, k)] = e;
    }
    fill_plan(best, out);
    out->tuned = 1;
    out->tuned_ms = best_ms;
    out->default_ms = base_ms;
    out->trials = trials;
    return MLK_OK;
}

}  // namespace mlk
