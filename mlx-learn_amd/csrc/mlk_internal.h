// Host-side internals shared by the library's translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/mlxlearn_amd.h"

namespace mlk {

// ------------------------------------------------------------------ errors
void set_error(const std::string& msg);
struct Status {
    int code = MLK_OK;
    static Status ok() { return {}; }
};

#define MLK_CHECK_HIP(expr)                                                                   \
    do {                                                                                      \
        hipError_t _e = (expr);                                                               \
        if (_e != hipSuccess) {                                                               \
            ::mlk::set_error(std::string(#expr " failed: ") + hipGetErrorString(_e));          \
            return _e == hipErrorOutOfMemory ? MLK_ENOMEM : MLK_EDEVICE;                      \
        }                                                                                     \
    } while (0)

#define MLK_TRY(expr)                                                                         \
    do {                                                                                      \
        int _rc = (expr);                                                                     \
        if (_rc != MLK_OK) return _rc;                                                        \
    } while (0)

// HIP events owned by one call: every event made here is destroyed when the
// owner goes out of scope, on every return path.
// Events come from a per-device pool (creating and destroying a dozen events
// per call costs more than a small call's kernels); released events are
// re-recorded by later calls. They are timing events: waiting on one orders
// the host after the work before it, but device memory is read back through
// stream-ordered copies, never through an event.
hipError_t event_acquire(int device, hipEvent_t* out);
void event_release(int device, hipEvent_t e);

struct Events {
    std::vector<hipEvent_t> ev;
    int device = -1;
    Events() = default;
    Events(const Events&) = delete;
    Events& operator=(const Events&) = delete;
    ~Events() {
        for (hipEvent_t e : ev) event_release(device, e);
    }
    hipError_t make(hipEvent_t* out) {
        if (device < 0) {
            const hipError_t g = hipGetDevice(&device);
            if (g != hipSuccess) {
                device = -1;
                return g;
            }
        }
        const hipError_t r = event_acquire(device, out);
        if (r == hipSuccess) ev.push_back(*out);
        return r;
    }
};

// Grow-only device buffer.
struct DevBuf {
    void* ptr = nullptr;
    size_t bytes = 0;
    int reserve(size_t need);
    void release();
    ~DevBuf() { release(); }
    template <class T> T* as() const { return static_cast<T*>(ptr); }
};

// Workspace pool of transient indexes (the centres index of one k-means
// assignment): such an index borrows the device's pooled per-call workspace
// when it is made and hands its own back when it is destroyed, so repeated
// assignments of one shape reuse their candidate lists instead of allocating
// and freeing gigabytes per call. One workspace is kept per device, until the
// process ends or a device allocation fails (it is then freed and the
// allocation tried again).
void workspace_adopt(mlk_index* index);
void workspace_return(mlk_index* index);
bool workspace_release_pool();

// Device memory of one call's temporaries. Allocating mode (arena == nullptr):
// slot i of the index's grow-only workspace, reserved with hipMalloc when it
// is too small. Fixed mode: consecutive 256-byte aligned pieces of a
// preallocated arena, never grown; a piece that does not fit is an
// MLK_ECAPACITY error. get() returns the piece in *out.
struct CallMem {
    DevBuf* ws = nullptr;
    char* arena = nullptr;
    size_t cap = 0, used = 0;
    // > 0: k-NN calls shrink their plan's database splits until the candidate
    // lists take at most this many bytes (fixed arenas)
    size_t list_budget = 0;
    bool fixed() const { return arena != nullptr; }
    size_t left() const { return cap > used ? cap - used : 0; }
    int get(int slot, size_t bytes, void** out);
    template <class T> int get(int slot, size_t bytes, T** out) {
        void* p = nullptr;
        const int rc = get(slot, bytes, &p);
        *out = static_cast<T*>(p);
        return rc;
    }
};

// Screen precisions.
enum Prec { PREC_FP32 = 0, PREC_BF16 = 1, PREC_FP16 = 2 };

inline int elem_bytes(int prec) { return prec == PREC_FP32 ? 4 : 2; }

// Per-precision screen mirror of the database.
struct Mirror {
    bool ready = false;
    int dp = 0;             // padded feature count (multiple of 32)
    double scale = 1.0;     // power of two applied before rounding (float16 range)
    bool exact = false;     // every scaled centred coordinate is exact in the screen type
    int64_t n_unsafe = 0;   // rows whose coordinates leave the screen's safe range
    DevBuf xs;              // n x dp screen values
    DevBuf nx;              // n float64 squared norms of scale * (x - c)
    DevBuf nx32;            // the same norms rounded to float32 (screen hit path)
    DevBuf flags;           // n uint8: bit0 = safe
    DevBuf unsafe;          // n_unsafe int32 row ids
    DevBuf nt;              // n float64 squared norms of the rounded rows (screen values)
    DevBuf d2;              // n float64 squared norms of the rounding residuals
    DevBuf bias;            // n float: -a_x for the constants below
    double bias_t1 = -1.0, bias_t2 = -1.0, bias_gamma = -1.0, bias_eps = -1.0;
};

// Tuning knobs (0 = automatic).
struct Tuning {
    int splits = 0;
    int capacity = 0;
    int query_tiles = 0;
    int stages = 0;
    std::vector<double> rounds;
};
Tuning tuning_snapshot();  // a copy taken under the configuration lock

struct TestHooks {
    uint32_t hooks = 0;
    int64_t param = 0;
    std::vector<int64_t> forced_rows;
};
TestHooks hooks_snapshot();  // a copy taken under the configuration lock
mlk_stats& stats();

}  // namespace mlk

struct mlk_index {
    int64_t n = 0;
    int32_t d = 0;
    bool integral = false;   // every coordinate is an integer
    mlk::DevBuf X;           // n x d float32 copy (authoritative coordinates)
    mlk::DevBuf centre;      // d float32
    mlk::Mirror mirror[3];
    std::mutex lock;
    // per-call workspace
    mlk::DevBuf ws[32];
    // captured k-NN call graphs (MLK_GRAPH_REPLAY), owned by mlk_knn.hip
    void* graphs = nullptr;
    // query order centres (mlk_order.hip): order_k database rows as bf16
    // (padded to a multiple of 32 centres and of 16 features) and their
    // float32 squared norms (+inf for the padding); order_k = 0 until first use
    int order_k = 0;
    mlk::DevBuf order_cb, order_cn;
    // database rows grouped by nearest order centre (ascending within a
    // centre), each centre's first position (order_k + 1 entries) and each
    // row's centre
    mlk::DevBuf order_rows, order_start, order_cell;
    bool order_rows_ready = false;
};

struct mlk_radius_result {
    int64_t m = 0;
    int64_t nnz = 0;
    bool has_dist = false;
    mlk::DevBuf indices;  // nnz int64
    mlk::DevBuf dist;     // nnz double
};

namespace mlk {
// implemented in mlk_capi.hip: the injected-failure hook and the k-NN argument checks
int check_fail_hook();
int validate_knn(mlk_index* index, const float* Q, int64_t m, int32_t k, int64_t q_offset, uint32_t flags,
                 const int64_t* idx);
// implemented in mlk_index.hip
int prepare_mirror(mlk_index* index, int prec, hipStream_t stream);
// may_allocate false: the bias buffer must exist already (a bias for other
// constants is rebuilt in place); otherwise MLK_EINVAL.
struct Consts;
int prepare_bias(mlk_index* index, int prec, const Consts& c, hipStream_t stream, bool may_allocate = true);
// implemented in mlk_knn.hip
// mem: where the call's temporaries live (nullptr: the index's workspace).
// With a fixed arena the call allocates nothing, records no events and needs
// the screen mirror and bias of its precision to exist already.
int knn_query_impl(mlk_index* index, const float* Q, int64_t m, int k, int64_t q_offset, uint32_t flags,
                   double* dist, int64_t* idx, hipStream_t stream, CallMem* mem = nullptr);
// Device bytes a fixed arena needs for one knn_query_impl call of m queries
// with the plan such a call would use (0 if the index is not prepared for the
// precision), not counting the exact fallback's per-query buffers.
size_t knn_arena_bytes(const mlk_index* index, int prec, int64_t m, int k, uint32_t flags, size_t list_budget);
// Device bytes of the exact fallback's buffers for one query against n rows.
size_t knn_fallback_bytes_per_query(int64_t n);
void destroy_graphs(mlk_index* index);
// implemented in mlk_tune.hip
int autotune_impl(mlk_index* index, const float* Q, int64_t m, int k, double R2, uint32_t flags, int max_trials,
                  mlk_plan* out, hipStream_t stream);
int plan_info_impl(const mlk_index* index, int64_t m, int k, uint32_t flags, mlk_plan* out);
void clear_plans_impl();
std::string export_plans_impl();
int import_plans_impl(const char* text, size_t len);
// implemented in mlk_radius.hip
int radius_query_impl(mlk_index* index, const float* Q, int64_t m, double R2, int64_t q_offset,
                      uint32_t flags, int64_t* indptr, int64_t* nnz, mlk_radius_result** out,
                      hipStream_t stream);
}  // namespace mlk
