// extern "C" entry points: argument validation, error strings, test hooks.
// Reference seam: DeviceNeighborIndex / fit_neighbors (src/mlxlearn_device/_neighbors.py:243-466)
// and the native module conventions of _neighbors_cpu_native.cpp:28-65, 818-828
// (caller-owned outputs, validated shapes, ValueError-style status on bad input).
#include <hip/hip_runtime.h>

#include <cstring>
#include <string>

#include "mlk_internal.h"

namespace mlk {

static thread_local std::string g_error;
static thread_local mlk_stats g_stats;
static Tuning g_tuning;
static TestHooks g_hooks;
static std::mutex g_config_lock;

// Event pool: (device, event) pairs, at most EVENT_POOL_MAX kept.
static std::mutex g_event_lock;
static std::vector<std::pair<int, hipEvent_t>> g_event_pool;
constexpr size_t EVENT_POOL_MAX = 256;

hipError_t event_acquire(int device, hipEvent_t* out) {
    {
        std::lock_guard<std::mutex> g(g_event_lock);
        for (size_t i = g_event_pool.size(); i-- > 0;) {
            if (g_event_pool[i].first == device) {
                *out = g_event_pool[i].second;
                g_event_pool.erase(g_event_pool.begin() + (std::ptrdiff_t)i);
                return hipSuccess;
            }
        }
    }
    // timing only: no system-scope fence (cache write-back and invalidate)
    // when the event is reached, so a recorded event costs the stream next to
    // nothing and leaves the caches of the work after it warm
    return hipEventCreateWithFlags(out, hipEventDisableSystemFence);
}

void event_release(int device, hipEvent_t e) {
    {
        std::lock_guard<std::mutex> g(g_event_lock);
        if (g_event_pool.size() < EVENT_POOL_MAX) {
            g_event_pool.emplace_back(device, e);
            return;
        }
    }
    (void)hipEventDestroy(e);
}

void set_error(const std::string& msg) { g_error = msg; }
Tuning tuning_snapshot() {
    std::lock_guard<std::mutex> g(g_config_lock);
    return g_tuning;
}
TestHooks hooks_snapshot() {
    std::lock_guard<std::mutex> g(g_config_lock);
    return g_hooks;
}
mlk_stats& stats() { return g_stats; }

int create_index(const float* X, int64_t n, int d, hipStream_t stream, mlk_index** out);

int check_fail_hook() {
    std::lock_guard<std::mutex> g(g_config_lock);
    if (g_hooks.hooks & MLK_HOOK_FAIL) {
        g_hooks.hooks &= ~MLK_HOOK_FAIL;
        set_error("injected failure (test hook)");
        return MLK_EDEVICE;
    }
    return MLK_OK;
}

static int invalid(const std::string& msg) {
    set_error(msg);
    return MLK_EINVAL;
}

int validate_knn(mlk_index* index, const float* Q, int64_t m, int32_t k, int64_t q_offset, uint32_t flags,
                        const int64_t* idx) {
    if (!index) return invalid("index is NULL");
    if (m < 0) return invalid("query row count must be non-negative");
    if (m && (!Q || !idx)) return invalid("queries and index output must be device pointers");
    if (k < 1) return invalid("Expected n_neighbors > 0. Got " + std::to_string(k));
    const bool exclude = (flags & MLK_EXCLUDE_SELF) != 0;
    const int64_t limit = index->n - (exclude ? 1 : 0);
    if (k > limit)
        return invalid("n_neighbors must be <= " + std::to_string(limit) + ", got " + std::to_string(k));
    if ((flags & MLK_PREC_MASK) == MLK_PREC_MASK) return invalid("unknown screen precision");
    if (exclude && (q_offset < 0 || q_offset + m > index->n))
        return invalid("exclude_self requires query rows to be training rows");
    return MLK_OK;
}

}  // namespace mlk

using namespace mlk;

extern "C" {

const char* mlk_last_error(void) { return g_error.c_str(); }

int mlk_version(void) { return 1; }

int mlk_get_stats(mlk_stats* out) {
    if (!out) return invalid("stats output is NULL");
    *out = g_stats;
    return MLK_OK;
}

int mlk_set_test_hooks(uint32_t h, int64_t param) {
    std::lock_guard<std::mutex> g(g_config_lock);
    g_hooks.hooks = h;
    g_hooks.param = param;
    if (!(h & MLK_HOOK_FORCE_UNCERTIFIED)) g_hooks.forced_rows.clear();
    return MLK_OK;
}

int mlk_set_forced_rows(const int64_t* rows, int64_t count) {
    std::lock_guard<std::mutex> g(g_config_lock);
    if (count < 0 || (count && !rows)) return invalid("forced rows must be a host array");
    g_hooks.forced_rows.assign(rows, rows + count);
    return MLK_OK;
}

int mlk_set_tuning(int32_t splits, int32_t capacity, int32_t query_tiles, int32_t stages, int32_t n_rounds,
                   const double* round_fractions) {
    std::lock_guard<std::mutex> g(g_config_lock);
    if (splits < 0 || capacity < 0 || query_tiles < 0 || query_tiles == 3 || query_tiles > 4 || stages < 0 ||
        stages == 1 || stages > 4 || n_rounds < 0)
        return invalid("tuning values must be non-negative (query_tiles 1, 2 or 4, stages 2-4)");
    Tuning t;
    t.splits = splits;
    t.capacity = capacity;
    t.query_tiles = query_tiles;
    t.stages = stages;
    if (n_rounds) {
        if (!round_fractions) return invalid("round fractions missing");
        double prev = 0.0;
        for (int i = 0; i < n_rounds; ++i) {
            double f = round_fractions[i];
            if (!(f > prev && f <= 1.0)) return invalid("round fractions must increase within (0, 1]");
            t.rounds.push_back(f);
            prev = f;
        }
        if (t.rounds.back() != 1.0) t.rounds.push_back(1.0);
    }
    g_tuning = t;
    return MLK_OK;
}

int mlk_index_create(const float* X_dev, int64_t n, int32_t d, uint32_t flags, void* stream, mlk_index** out) {
    (void)flags;
    if (!out) return invalid("output handle pointer is NULL");
    *out = nullptr;
    MLK_TRY(check_fail_hook());
    if (!X_dev) return invalid("X must be a device pointer");
    if (n < 1 || d < 1) return invalid("X must be a nonempty two-dimensional matrix");
    if (n >= (int64_t(1) << 31)) return invalid("training row count must fit int32 indices");
    if (d > 65536) return invalid("at most 65536 features are supported");
    return create_index(X_dev, n, d, (hipStream_t)stream, out);
}

int mlk_index_destroy(mlk_index* index) {
    if (!index) return MLK_OK;
    (void)hipDeviceSynchronize();
    destroy_graphs(index);
    delete index;
    return MLK_OK;
}

int mlk_index_info(const mlk_index* index, int64_t* n, int32_t* d, int32_t* integral) {
    if (!index) return invalid("index is NULL");
    if (n) *n = index->n;
    if (d) *d = index->d;
    if (integral) *integral = index->integral ? 1 : 0;
    return MLK_OK;
}

int mlk_knn_query(mlk_index* index, const float* Q_dev, int64_t m, int32_t k, int64_t q_offset, uint32_t flags,
                  double* dist_dev, int64_t* idx_dev, void* stream) {
    MLK_TRY(check_fail_hook());
    MLK_TRY(validate_knn(index, Q_dev, m, k, q_offset, flags, idx_dev));
    std::lock_guard<std::mutex> g(index->lock);
    return knn_query_impl(index, Q_dev, m, k, q_offset, flags, dist_dev, idx_dev, (hipStream_t)stream);
}

int mlk_knn_query_batched(mlk_index* index, const float* Q_dev, const int64_t* offsets, int32_t n_batches, int32_t k,
                          uint32_t flags, double* dist_dev, int64_t* idx_dev, void* stream) {
    MLK_TRY(check_fail_hook());
    if (n_batches < 0 || (n_batches && !offsets)) return invalid("batch offsets must be a host array");
    if (flags & MLK_EXCLUDE_SELF) return invalid("exclude_self is not supported for batched queries");
    int64_t m = 0;
    if (n_batches) {
        if (offsets[0] != 0) return invalid("batch offsets must start at 0");
        for (int b = 0; b < n_batches; ++b)
            if (offsets[b + 1] < offsets[b]) return invalid("batch offsets must be non-decreasing");
        m = offsets[n_batches];
    }
    MLK_TRY(validate_knn(index, Q_dev, m, k, 0, flags, idx_dev));
    std::lock_guard<std::mutex> g(index->lock);
    // Every query is screened independently, so the batches share one
    // screen launch per round over their concatenated rows.
    return knn_query_impl(index, Q_dev, m, k, 0, flags, dist_dev, idx_dev, (hipStream_t)stream);
}

int mlk_radius_query(mlk_index* index, const float* Q_dev, int64_t m, double R2, int64_t q_offset, uint32_t flags,
                     int64_t* indptr_dev, int64_t* nnz, mlk_radius_result** out, void* stream) {
    MLK_TRY(check_fail_hook());
    if (!index || !out || !nnz) return invalid("index, nnz and result pointers are required");
    *out = nullptr;
    if (m < 0 || (m && !Q_dev) || !indptr_dev) return invalid("queries and indptr must be device pointers");
    if (!(R2 >= 0)) return invalid("R2 must be a non-negative number");
    if ((flags & MLK_EXCLUDE_SELF) && (q_offset < 0 || q_offset + m > index->n))
        return invalid("exclude_self requires query rows to be training rows");
    std::lock_guard<std::mutex> g(index->lock);
    return radius_query_impl(index, Q_dev, m, R2, q_offset, flags, indptr_dev, nnz, out, (hipStream_t)stream);
}

int mlk_autotune(mlk_index* index, const float* Q_dev, int64_t m, int32_t k, double R2, uint32_t flags,
                 int32_t max_trials, mlk_plan* out, void* stream) {
    MLK_TRY(check_fail_hook());
    if (!out) return invalid("plan output is NULL");
    if (!index) return invalid("index is NULL");
    if (m < 1 || !Q_dev) return invalid("autotuning needs at least one query row on the device");
    if (max_trials < 0) return invalid("max_trials must be non-negative");
    if (flags & ~(MLK_PREC_MASK | MLK_TUNE_RADIUS | MLK_SORT_RESULTS | MLK_NEED_DISTANCE))
        return invalid("unsupported autotune flags");
    if ((flags & MLK_PREC_MASK) == MLK_PREC_MASK) return invalid("unknown screen precision");
    if (flags & MLK_TUNE_RADIUS) {
        if (!(R2 >= 0)) return invalid("R2 must be a non-negative number");
    } else {
        if (k < 1) return invalid("Expected n_neighbors > 0. Got " + std::to_string(k));
        if (k > index->n)
            return invalid("n_neighbors must be <= " + std::to_string(index->n) + ", got " + std::to_string(k));
    }
    std::lock_guard<std::mutex> g(index->lock);
    return autotune_impl(index, Q_dev, m, k, R2, flags, max_trials ? max_trials : 24, out, (hipStream_t)stream);
}

int mlk_get_plan(const mlk_index* index, int64_t m, int32_t k, uint32_t flags, mlk_plan* out) {
    if (!index || !out) return invalid("index and plan output are required");
    if (m < 0) return invalid("query row count must be non-negative");
    if (!(flags & MLK_TUNE_RADIUS) && k < 1) return invalid("Expected n_neighbors > 0. Got " + std::to_string(k));
    if ((flags & MLK_PREC_MASK) == MLK_PREC_MASK) return invalid("unknown screen precision");
    return plan_info_impl(index, m, k, flags, out);
}

int mlk_clear_plans(void) {
    clear_plans_impl();
    return MLK_OK;
}

int mlk_export_plans(char* buf, int64_t cap, int64_t* len) {
    if (!len || cap < 0 || (cap && !buf)) return invalid("export_plans: length output and buffer are required");
    const std::string text = export_plans_impl();
    *len = (int64_t)text.size();
    if ((int64_t)text.size() > cap) {
        if (buf) return invalid("export_plans: buffer too small (" + std::to_string(text.size()) + " bytes needed)");
        return MLK_OK;  // size query
    }
    std::memcpy(buf, text.data(), text.size());
    return MLK_OK;
}

int mlk_import_plans(const char* text, int64_t len) {
    if (len < 0 || (len && !text)) return invalid("import_plans: text is required");
    return import_plans_impl(text, (size_t)len);
}

int mlk_radius_fill(mlk_radius_result* r, int64_t* indices_dev, double* dist_dev, void* stream) {
    if (!r) return invalid("result is NULL");
    if (r->nnz && !indices_dev) return invalid("indices output is NULL");
    if (dist_dev && !r->has_dist) return invalid("distances were not computed for this result");
    hipStream_t s = (hipStream_t)stream;
    if (r->nnz) {
        MLK_CHECK_HIP(hipMemcpyAsync(indices_dev, r->indices.ptr, r->nnz * sizeof(int64_t), hipMemcpyDeviceToDevice, s));
        if (dist_dev)
            MLK_CHECK_HIP(hipMemcpyAsync(dist_dev, r->dist.ptr, r->nnz * sizeof(double), hipMemcpyDeviceToDevice, s));
    }
    MLK_CHECK_HIP(hipStreamSynchronize(s));
    return MLK_OK;
}

int mlk_radius_free(mlk_radius_result* r) {
    delete r;
    return MLK_OK;
}

}  // extern "C"
