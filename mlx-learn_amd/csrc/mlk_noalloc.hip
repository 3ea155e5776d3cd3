// The k-NN entry point that allocates nothing: its temporaries live in a
// fixed, file-scope device array instead of the index's grow-only workspace.
//
// No counterpart in the reference. It serves callers that cannot call
// hipMalloc (serving loops with a fixed memory budget, harnesses that forbid
// allocation): the array is part of the library's code object, so each GPU
// the library runs on holds it from the library's first use on that GPU until
// the process ends.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "mlk_internal.h"
#include "mlk_screen.h"

namespace mlk {

// The scratch arena: query mirrors, thresholds, candidate lists, rank state
// and the exact fallback's buffers of one call (or one chunk of its queries).
constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(128) << 20;
// Half of it is kept for the exact fallback, whose buffers grow with the
// index: one fallback query needs about 20 bytes per database row.
constexpr size_t NOALLOC_FALLBACK_BYTES = NOALLOC_SCRATCH_BYTES / 2;

static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

// Device address of `scratch` on each device. In host code the array's name is
// a host-side stand-in, not a device address: only these looked-up pointers
// are passed to kernels and copies.
constexpr int NOALLOC_MAX_DEVICES = 64;
static std::mutex g_scratch_lock;
static void* g_scratch_ptr[NOALLOC_MAX_DEVICES];

static int scratch_for_current_device(char** out) {
    int dev = -1;
    MLK_CHECK_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) {
        set_error("device ordinal out of range for the scratch table");
        return MLK_EDEVICE;
    }
    std::lock_guard<std::mutex> g(g_scratch_lock);
    if (!g_scratch_ptr[dev]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) {
            (void)hipGetLastError();
            set_error(std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
            return MLK_EDEVICE;
        }
        g_scratch_ptr[dev] = ptr;
    }
    *out = static_cast<char*>(g_scratch_ptr[dev]);
    return MLK_OK;
}

}  // namespace mlk

using namespace mlk;

// Same arguments, results and errors as mlk_knn_query. Additionally: the index
// must already hold the screen mirror and bias of the requested precision (one
// mlk_knn_query call of that precision prepares them), else MLK_EINVAL; an
// index whose exact fallback would not fit half the scratch (more than about
// 3.3 million rows) is refused with MLK_ECAPACITY. Queries are processed in
// chunks whose temporaries fit the scratch. Calls that overlap on one GPU (on
// different streams or threads, any index) race on its scratch.
extern "C" int mlk_knn_query_noalloc(mlk_index* index, const float* Q_dev, int64_t m, int32_t k, int64_t q_offset,
                                     uint32_t flags, double* dist_dev, int64_t* idx_dev, void* stream) {
    MLK_TRY(check_fail_hook());
    MLK_TRY(validate_knn(index, Q_dev, m, k, q_offset, flags, idx_dev));
    std::lock_guard<std::mutex> g(index->lock);
    if (m == 0) {
        stats() = mlk_stats{};
        return MLK_OK;
    }
    const int prec = effective_prec((int)((flags & MLK_PREC_MASK) >> 4), index->d);
    if (!index->mirror[prec].ready) {
        set_error("the index has no screen mirror for this precision yet: one mlk_knn_query call of the same "
                  "precision prepares it");
        return MLK_EINVAL;
    }
    const size_t per_fallback = knn_fallback_bytes_per_query(index->n) + 8 * 256;
    if (per_fallback > NOALLOC_FALLBACK_BYTES) {
        set_error("mlk_knn_query_noalloc serves indexes of at most " +
                  std::to_string(NOALLOC_FALLBACK_BYTES / (per_fallback / index->n)) + " rows (this one has " +
                  std::to_string(index->n) + "): the exact fallback would not fit its fixed scratch");
        return MLK_ECAPACITY;
    }
    char* arena = nullptr;
    MLK_TRY(scratch_for_current_device(&arena));
    // the largest chunk of queries whose temporaries leave the fallback's half
    // free, the candidate lists taking at most three quarters of the rest (the
    // plan's database splits shrink to fit)
    const size_t budget = NOALLOC_SCRATCH_BYTES - NOALLOC_FALLBACK_BYTES;
    const size_t lists = budget / 4 * 3;
    int64_t chunk = m;
    while (chunk > 1 && knn_arena_bytes(index, prec, chunk, k, flags, lists) > budget) chunk = (chunk + 1) / 2;
    if (knn_arena_bytes(index, prec, chunk, k, flags, lists) > budget) {
        set_error("one query's temporaries do not fit the fixed scratch of mlk_knn_query_noalloc");
        return MLK_ECAPACITY;
    }
    mlk_stats total{};
    for (int64_t off = 0; off < m; off += chunk) {
        const int64_t mc = std::min<int64_t>(chunk, m - off);
        CallMem mem;
        mem.arena = arena;
        mem.cap = NOALLOC_SCRATCH_BYTES;
        mem.list_budget = lists;
        MLK_TRY(knn_query_impl(index, Q_dev + off * index->d, mc, k, q_offset + off, flags,
                               dist_dev ? dist_dev + off * k : nullptr, idx_dev + off * k, (hipStream_t)stream, &mem));
        const mlk_stats& st = stats();
        total.queries += st.queries;
        total.candidates += st.candidates;
        total.reranked += st.reranked;
        total.fallback_queries += st.fallback_queries;
        total.rounds += st.rounds;
        total.splits = st.splits;
        total.capacity = st.capacity;
        total.screen_precision = st.screen_precision;
    }
    stats() = total;
    return MLK_OK;
}
