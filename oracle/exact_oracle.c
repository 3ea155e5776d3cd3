/*
 * Exact neighbour oracle -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference's exact ordering contract, used by
 * tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py as the
 * checker. Nothing in the product library links or calls this file.
 *
 * Contract restated (reference):
 *  - k-NN rows are ordered by the exact real squared distance S(q, x) of the
 *    float32 coordinates; only mathematically equal S falls back to the lower
 *    training index (src/mlxlearn/_kernels/neighbors_cpu.py:174-181, 213-262;
 *    docs/knn-precision-ordering.md).
 *  - float32 values are integers in units of 2**-149
 *    (neighbors_cpu.py:39-47, _integer_row), so S * 2**298 is an integer.
 *  - Radius membership is exact S <= R2 with R2 = fl64(r*r); candidates are
 *    classified by the canonical float64 sum against an outward-rounded band
 *    and the band is resolved by S * 2**298 <= floor(R2 * 2**298)
 *    (radius_contract.py:129-217; _radius_common.h:1-18, 337-357).
 *  - Reported distances are sqrt of the canonical float64 sum
 *    s = sum_k (double(q_k) - double(x_k))**2 evaluated strictly left to right
 *    (_radius_common.h:349-357); radius distances are sqrt(min(s, R2)).
 *
 * Method (deliberately different from the product's device arithmetic): every
 * coordinate becomes a signed 288-bit integer I(v) = v * 2**149; each
 * difference I(q) - I(x) is formed in big-integer arithmetic and squared by
 * schoolbook multiplication into a 640-bit key.
 *
 * Build: oracle/build_oracle.py (gcc -O2 -ffp-contract=off -shared -fPIC).
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define MAG_LIMBS 9   /* |I(v)| < 2**277 */
#define KEY_LIMBS 20  /* S * 2**298 < 2**640 for d < 2**60 */

typedef struct { uint32_t w[KEY_LIMBS]; } okey;

/* |v| * 2**149 as a little-endian 288-bit magnitude; returns the sign. */
static int f32_to_mag(float v, uint32_t mag[MAG_LIMBS]) {
    uint32_t bits;
    memcpy(&bits, &v, 4);
    uint32_t exponent = (bits >> 23) & 255u;
    uint64_t significand = (bits & 0x7fffffu) | (exponent ? 0x800000u : 0u);
    unsigned shift = exponent ? exponent - 1 : 0;  /* value = significand * 2**(shift-149) */
    memset(mag, 0, sizeof(uint32_t) * MAG_LIMBS);
    unsigned limb = shift / 32, bit = shift % 32;
    uint64_t wide = significand << bit;  /* at most 24 + 31 = 55 bits */
    mag[limb] = (uint32_t)wide;
    if (limb + 1 < MAG_LIMBS) mag[limb + 1] = (uint32_t)(wide >> 32);
    return (int)(bits >> 31);
}

static int mag_cmp(const uint32_t* a, const uint32_t* b, int n) {
    for (int i = n - 1; i >= 0; --i) {
        if (a[i] != b[i]) return a[i] < b[i] ? -1 : 1;
    }
    return 0;
}

/* out = |I(a) - I(b)| */
static void abs_diff(float a, float b, uint32_t out[MAG_LIMBS + 1]) {
    uint32_t ma[MAG_LIMBS], mb[MAG_LIMBS];
    int sa = f32_to_mag(a, ma), sb = f32_to_mag(b, mb);
    memset(out, 0, sizeof(uint32_t) * (MAG_LIMBS + 1));
    if (sa != sb) { /* magnitudes add */
        uint64_t carry = 0;
        for (int i = 0; i < MAG_LIMBS; ++i) {
            uint64_t t = (uint64_t)ma[i] + mb[i] + carry;
            out[i] = (uint32_t)t;
            carry = t >> 32;
        }
        out[MAG_LIMBS] = (uint32_t)carry;
        return;
    }
    const uint32_t *hi = ma, *lo = mb;
    if (mag_cmp(ma, mb, MAG_LIMBS) < 0) { hi = mb; lo = ma; }
    int64_t borrow = 0;
    for (int i = 0; i < MAG_LIMBS; ++i) {
        int64_t t = (int64_t)hi[i] - lo[i] - borrow;
        borrow = t < 0;
        out[i] = (uint32_t)(t + (borrow ? ((int64_t)1 << 32) : 0));
    }
}

/* key += m * m (schoolbook square of a (MAG_LIMBS+1)-limb magnitude). */
static void add_square(okey* key, const uint32_t m[MAG_LIMBS + 1]) {
    uint64_t prod[2 * (MAG_LIMBS + 1)];
    memset(prod, 0, sizeof(prod));
    for (int i = 0; i <= MAG_LIMBS; ++i) {
        if (!m[i]) continue;
        uint64_t carry = 0;
        for (int j = 0; j <= MAG_LIMBS; ++j) {
            uint64_t t = (uint64_t)m[i] * m[j] + prod[i + j] + carry;
            prod[i + j] = t & 0xffffffffu;
            carry = t >> 32;
        }
        int p = i + MAG_LIMBS + 1;
        while (carry) {
            uint64_t t = prod[p] + carry;
            prod[p] = t & 0xffffffffu;
            carry = t >> 32;
            ++p;
        }
    }
    uint64_t carry = 0;
    for (int i = 0; i < KEY_LIMBS; ++i) {
        uint64_t t = (uint64_t)key->w[i] + (i < 2 * (MAG_LIMBS + 1) ? prod[i] : 0) + carry;
        key->w[i] = (uint32_t)t;
        carry = t >> 32;
    }
}

/* S(q, x) * 2**298 exactly. */
static void exact_key(const float* q, const float* x, int d, okey* key) {
    memset(key, 0, sizeof(*key));
    uint32_t diff[MAG_LIMBS + 1];
    for (int i = 0; i < d; ++i) {
        abs_diff(q[i], x[i], diff);
        add_square(key, diff);
    }
}

static int key_cmp(const okey* a, const okey* b) { return mag_cmp(a->w, b->w, KEY_LIMBS); }

/* floor(R2 * 2**298) for finite R2 >= 0; returns 0 when R2 >= 2**342 (everything). */
static int r2_key(double R2, okey* key) {
    memset(key, 0, sizeof(*key));
    if (!(R2 < 0x1p342)) return 0;
    if (!(R2 > 0)) return 1;
    int e;
    double m = frexp(R2, &e);                 /* R2 = m * 2**e, 0.5 <= m < 1 */
    uint64_t mant = (uint64_t)ldexp(m, 53);   /* R2 = mant * 2**(e-53) */
    int shift = e - 53 + 298;
    if (shift >= 0) {
        unsigned limb = (unsigned)shift / 32, bit = (unsigned)shift % 32;
        unsigned __int128 wide = (unsigned __int128)mant << bit;
        for (unsigned i = limb; i < KEY_LIMBS && wide; ++i) {
            key->w[i] = (uint32_t)wide;
            wide >>= 32;
        }
    } else if (-shift < 64) {
        uint64_t v = mant >> (unsigned)(-shift);
        key->w[0] = (uint32_t)v;
        key->w[1] = (uint32_t)(v >> 32);
    }
    return 1;
}

/* Canonical float64 squared distance, strictly sequential (built with -ffp-contract=off). */
static double canonical_sq(const float* q, const float* x, int d) {
    double s = 0.0;
    for (int i = 0; i < d; ++i) {
        double t = (double)q[i] - (double)x[i];
        s = s + t * t;
    }
    return s;
}

/* ------------------------------------------------------------------ k-NN */

typedef struct {
    const float* X; int64_t n; int d;
    const float* Q; int64_t m; int k; int exclude_self; int64_t q_offset;
    int64_t* out_idx; double* out_dist;
    int64_t next; pthread_mutex_t lock;
    int status;
} knn_job;

typedef struct { double s; int64_t idx; okey key; } cand;

static int cand_cmp(const void* pa, const void* pb) {
    const cand* a = (const cand*)pa;
    const cand* b = (const cand*)pb;
    int c = key_cmp(&a->key, &b->key);
    if (c) return c;
    return a->idx < b->idx ? -1 : (a->idx > b->idx);
}

/* k-th smallest (0-based) of a[0..n) by quickselect; a is permuted. */
static double select_kth(double* a, int64_t n, int64_t kth) {
    int64_t lo = 0, hi = n - 1;
    while (lo < hi) {
        double pivot = a[lo + (hi - lo) / 2];
        int64_t i = lo, j = hi;
        while (i <= j) {
            while (a[i] < pivot) ++i;
            while (a[j] > pivot) --j;
            if (i <= j) { double t = a[i]; a[i] = a[j]; a[j] = t; ++i; --j; }
        }
        if (kth <= j) hi = j;
        else if (kth >= i) lo = i;
        else return a[kth];
    }
    return a[kth];
}

static int knn_one(knn_job* job, int64_t row, double* s, double* scratch, cand** cbuf, int64_t* ccap) {
    const int d = job->d, k = job->k;
    const float* q = job->Q + row * (int64_t)d;
    int64_t self = job->exclude_self ? job->q_offset + row : -1;
    int64_t valid = 0;
    for (int64_t i = 0; i < job->n; ++i) {
        s[i] = (i == self) ? INFINITY : canonical_sq(q, job->X + i * (int64_t)d, d);
        if (i != self) scratch[valid++] = s[i];
    }
    if (valid < k) return 1;
    double tk = select_kth(scratch, valid, k - 1);
    /* |s - S| <= (d+2) u S with u = 2**-53; a generous (d+4) 2**-52 margin. */
    double g = (d + 4) * 0x1p-52;
    double cut = tk * (1 + g) / (1 - g);
    cut = nextafter(nextafter(cut, INFINITY), INFINITY);
    int64_t count = 0;
    for (int64_t i = 0; i < job->n; ++i)
        if (i != self && s[i] <= cut) ++count;
    if (count > *ccap) {
        free(*cbuf);
        *ccap = count;
        *cbuf = (cand*)malloc(sizeof(cand) * (size_t)count);
        if (!*cbuf) return 2;
    }
    cand* c = *cbuf;
    int64_t at = 0;
    for (int64_t i = 0; i < job->n; ++i) {
        if (i == self || !(s[i] <= cut)) continue;
        c[at].s = s[i];
        c[at].idx = i;
        exact_key(q, job->X + i * (int64_t)d, d, &c[at].key);
        ++at;
    }
    qsort(c, (size_t)at, sizeof(cand), cand_cmp);
    for (int j = 0; j < k; ++j) {
        job->out_idx[row * k + j] = c[j].idx;
        if (job->out_dist) job->out_dist[row * k + j] = sqrt(c[j].s);
    }
    return 0;
}

static void* knn_worker(void* arg) {
    knn_job* job = (knn_job*)arg;
    double* s = (double*)malloc(sizeof(double) * (size_t)(job->n ? job->n : 1));
    double* scratch = (double*)malloc(sizeof(double) * (size_t)(job->n ? job->n : 1));
    cand* cbuf = NULL;
    int64_t ccap = 0;
    if (!s || !scratch) { job->status = 2; free(s); free(scratch); return NULL; }
    for (;;) {
        pthread_mutex_lock(&job->lock);
        int64_t row = job->next++;
        pthread_mutex_unlock(&job->lock);
        if (row >= job->m) break;
        int rc = knn_one(job, row, s, scratch, &cbuf, &ccap);
        if (rc) job->status = rc;
    }
    free(s); free(scratch); free(cbuf);
    return NULL;
}

static int run_threads(void* (*fn)(void*), void* job, int threads) {
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    pthread_t tid[256];
    int started = 0;
    for (int t = 0; t < threads; ++t) {
        if (pthread_create(&tid[t], NULL, fn, job) == 0) ++started;
    }
    if (!started) fn(job);
    for (int t = 0; t < started; ++t) pthread_join(tid[t], NULL);
    return 0;
}

/* Exact k-NN: indices ordered by (exact S, index); distances sqrt(canonical s). */
int oracle_knn(const float* X, int64_t n, int d, const float* Q, int64_t m, int k,
               int exclude_self, int64_t q_offset, int64_t* out_idx, double* out_dist,
               int threads) {
    if (k < 1 || d < 1 || n < 1) return 1;
    knn_job job = {X, n, d, Q, m, k, exclude_self, q_offset, out_idx, out_dist, 0,
                   PTHREAD_MUTEX_INITIALIZER, 0};
    run_threads(knn_worker, &job, threads);
    return job.status;
}

/* Order two training rows for one query: -1, 0, +1 by exact S (no index tie-break). */
int oracle_exact_compare(const float* q, const float* a, const float* b, int d) {
    okey ka, kb;
    exact_key(q, a, d, &ka);
    exact_key(q, b, d, &kb);
    return key_cmp(&ka, &kb);
}

/* S * 2**298 as 20 little-endian uint32 limbs. */
void oracle_exact_key(const float* q, const float* x, int d, uint32_t* out20) {
    okey key;
    exact_key(q, x, d, &key);
    memcpy(out20, key.w, sizeof(key.w));
}

double oracle_canonical_sq(const float* q, const float* x, int d) { return canonical_sq(q, x, d); }

/* ---------------------------------------------------------------- radius */

static double down(double v) { return nextafter(v, -INFINITY); }
static double up(double v) { return nextafter(v, INFINITY); }

/* Restates thresholds() (radius_contract.py:129-153): the float64 band. */
void oracle_radius_band(int d, double R2, double* f64_in, double* f64_out) {
    if (isinf(R2)) { *f64_in = *f64_out = INFINITY; return; }
    if (R2 == 0.0) { *f64_in = *f64_out = 0.0; return; }
    double u = 0x1p-53;
    double g = (d + 2) * u / (1.0 - (d + 2) * u);
    double lo = down(down(R2 * down(1.0 - 2.0 * g)));
    *f64_in = lo > 0.0 ? lo : 0.0;
    *f64_out = up(up(R2 * up(1.0 + 2.0 * g)));
}

typedef struct {
    const float* X; int64_t n; int d;
    const float* Q; int64_t m; double R2; int exclude_self; int64_t q_offset;
    double f64_in, f64_out; okey rkey; int everything;
    int64_t* counts;            /* pass 1 */
    const int64_t* indptr;      /* pass 2 */
    int64_t* indices; double* distances;
    int pass;
    int64_t next; pthread_mutex_t lock;
} rad_job;

static int rad_member(rad_job* job, const float* q, const float* x, double s) {
    if (job->everything) return 1;
    if (s <= job->f64_in) return 1;
    if (!(s <= job->f64_out)) return 0;
    okey key;
    exact_key(q, x, job->d, &key);
    return key_cmp(&key, &job->rkey) <= 0;
}

static void* rad_worker(void* arg) {
    rad_job* job = (rad_job*)arg;
    for (;;) {
        pthread_mutex_lock(&job->lock);
        int64_t row = job->next++;
        pthread_mutex_unlock(&job->lock);
        if (row >= job->m) break;
        const float* q = job->Q + row * (int64_t)job->d;
        int64_t self = job->exclude_self ? job->q_offset + row : -1;
        int64_t count = 0, at = job->pass == 2 ? job->indptr[row] : 0;
        for (int64_t i = 0; i < job->n; ++i) {
            if (i == self) continue;
            const float* x = job->X + i * (int64_t)job->d;
            double s = canonical_sq(q, x, job->d);
            if (!rad_member(job, q, x, s)) continue;
            if (job->pass == 1) { ++count; continue; }
            job->indices[at] = i;
            if (job->distances) job->distances[at] = sqrt(s < job->R2 ? s : job->R2);
            ++at;
        }
        if (job->pass == 1) job->counts[row] = count;
    }
    return NULL;
}

/* Two-pass exact radius CSR in ascending index order. pass=1 fills counts[m];
 * pass=2 fills indices/distances at indptr. */
int oracle_radius(const float* X, int64_t n, int d, const float* Q, int64_t m, double R2,
                  int exclude_self, int64_t q_offset, int pass, int64_t* counts,
                  const int64_t* indptr, int64_t* indices, double* distances, int threads) {
    if (d < 1 || !(R2 >= 0)) return 1;
    rad_job job;
    memset(&job, 0, sizeof(job));
    job.X = X; job.n = n; job.d = d; job.Q = Q; job.m = m; job.R2 = R2;
    job.exclude_self = exclude_self; job.q_offset = q_offset;
    oracle_radius_band(d, R2, &job.f64_in, &job.f64_out);
    job.everything = !r2_key(R2, &job.rkey);
    job.counts = counts; job.indptr = indptr; job.indices = indices; job.distances = distances;
    job.pass = pass;
    pthread_mutex_init(&job.lock, NULL);
    run_threads(rad_worker, &job, threads);
    return 0;
}

/* Exact membership S(q, x) <= R2 for one pair. */
int oracle_exact_le(const float* q, const float* x, int d, double R2) {
    okey key, rk;
    if (!r2_key(R2, &rk)) return 1;
    exact_key(q, x, d, &key);
    return key_cmp(&key, &rk) <= 0;
}

/* ---------------------------------------------------------------- k-means */

/* Nearest centre by exact S, lowest index on ties (e_means.py:50-69 tie rule);
 * residual = canonical float64 s to the chosen centre. */
int oracle_kmeans_assign(const float* X, int64_t n, int d, const float* C, int64_t kc,
                         int64_t* labels, double* residual, int threads) {
    knn_job job = {C, kc, d, X, n, 1, 0, 0, labels, residual, 0, PTHREAD_MUTEX_INITIALIZER, 0};
    run_threads(knn_worker, &job, threads);
    /* knn_one wrote sqrt(s); the residual is the canonical s itself. */
    if (residual && !job.status)
        for (int64_t i = 0; i < n; ++i)
            residual[i] = canonical_sq(X + i * (int64_t)d, C + labels[i] * (int64_t)d, d);
    return job.status;
}
