// Device-side arithmetic shared by the kernels: exact squared-distance
// comparison, the canonical float64 sum, MFMA fragment types and layouts.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mlk {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(8 * sizeof(_Float16)))) _Float16 f16x8;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float f32x16;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;
typedef __attribute__((__vector_size__(4 * sizeof(int)))) int i32x4;

constexpr int WAVE = 64;

// Row of the 32x32 MFMA output held by accumulator register `i` of `lane`
// (column = lane % 32): rows come in groups of four, interleaved by lane half.
__device__ __forceinline__ int mfma32_row(int i, int lane) {
    return ((i >> 2) << 3) + ((lane >> 5) << 2) + (i & 3);
}

// ---------------------------------------------------------------------------
// float64 product and sum, each rounded on its own. HIP compiles with
// contraction on, and __dmul_rn / __dadd_rn are plain operators there, so
// __dadd_rn(s, __dmul_rn(t, t)) still becomes one fused multiply-add; these
// helpers carry the no-contraction pragma, so the compiler keeps the two
// roundings.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double mul_rn(double a, double b) {
#pragma clang fp contract(off)
    return a * b;
}
__device__ __forceinline__ double add_rn(double a, double b) {
#pragma clang fp contract(off)
    return a + b;
}

// ---------------------------------------------------------------------------
// canonical float64 squared distance: s = sum_k (double(q_k) - double(x_k))^2,
// strictly left to right, no contraction (matches the reference's canonical
// sum, _radius_common.h:349-357). |s - S| <= (d+2) 2^-53 S.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double canonical_sq(const float* __restrict__ q, const float* __restrict__ x, int d) {
    double s = 0.0;
    for (int i = 0; i < d; ++i) {
        const double t = (double)q[i] - (double)x[i];
        s = add_rn(s, mul_rn(t, t));
    }
    return s;
}

// The same canonical sum read four features at a time: d % 4 == 0 and both
// rows 16-byte aligned. The arithmetic and its order are unchanged, so the
// result is bitwise that of canonical_sq.
__device__ __forceinline__ double canonical_sq4(const float* __restrict__ q, const float* __restrict__ x, int d) {
    double s = 0.0;
    for (int i = 0; i < d; i += 4) {
        const float4 a = *reinterpret_cast<const float4*>(q + i);
        const float4 b = *reinterpret_cast<const float4*>(x + i);
        double t = (double)a.x - (double)b.x;
        s = add_rn(s, mul_rn(t, t));
        t = (double)a.y - (double)b.y;
        s = add_rn(s, mul_rn(t, t));
        t = (double)a.z - (double)b.z;
        s = add_rn(s, mul_rn(t, t));
        t = (double)a.w - (double)b.w;
        s = add_rn(s, mul_rn(t, t));
    }
    return s;
}

// Relative error bound of any float64 evaluation of S used for interval tests
// (covers the canonical order and every other summation order).
__host__ __device__ __forceinline__ double f64_interval(int d) { return (double)(d + 4) * 0x1p-52; }

// ---------------------------------------------------------------------------
// Exact comparison of S(q, a) and S(q, b).
//
// Each difference q_k - x_k of float32 values is split exactly into a float64
// pair (hi + lo) by TwoSum; its square is six float64 values from FMA-based
// TwoProd. Every such value is an integer multiple of 2^-298 and below 2^260,
// so all of them add exactly into a fixed-point accumulator of 32-bit chunks
// held in int64 limbs (lsb 2^-298). Signs of a's terms are +, b's are -.
// ---------------------------------------------------------------------------
constexpr int SACC_LIMBS = 20;

__device__ __forceinline__ void sacc_add(long long* w, double v) {
    if (v == 0.0) return;
    unsigned long long bits = (unsigned long long)__double_as_longlong(v);
    int e = (int)((bits >> 52) & 0x7ff);
    unsigned long long m = bits & ((1ull << 52) - 1);
    if (e) m |= 1ull << 52; else e = 1;
    int pos = e - 1075 + 298;  // |v| = m * 2^(e-1075) = m * 2^(pos-298)
    if (pos < 0) { m >>= -pos; pos = 0; }  // exact: v is a multiple of 2^-298
    const bool neg = (bits >> 63) != 0;
    const int limb = pos >> 5, sh = pos & 31;
    unsigned long long lo = (m & 0xffffffffull) << sh;  // < 2^63
    unsigned long long hi = (m >> 32) << sh;            // < 2^52
    long long c0 = (long long)(lo & 0xffffffffull);
    long long c1 = (long long)(lo >> 32) + (long long)(hi & 0xffffffffull);
    long long c2 = (long long)(hi >> 32);
    if (neg) { c0 = -c0; c1 = -c1; c2 = -c2; }
    w[limb] += c0;
    w[limb + 1] += c1;
    w[limb + 2] += c2;
}

__device__ __forceinline__ void sacc_add_square(long long* w, float qf, float xf, bool negate) {
    const double A = (double)qf, B = -(double)xf;
    const double hi = A + B;  // TwoSum(A, B): hi + lo == A + B exactly
    const double bb = hi - A;
    const double lo = (A - (hi - bb)) + (B - bb);
    const double p1 = hi * hi, e1 = __fma_rn(hi, hi, -p1);
    const double p2 = hi * lo, e2 = __fma_rn(hi, lo, -p2);
    const double p3 = lo * lo, e3 = __fma_rn(lo, lo, -p3);
    const double s = negate ? -1.0 : 1.0;
    sacc_add(w, s * p1);
    sacc_add(w, s * e1);
    sacc_add(w, s * 2.0 * p2);
    sacc_add(w, s * 2.0 * e2);
    sacc_add(w, s * p3);
    sacc_add(w, s * e3);
}

// Sign of the accumulated value after carry normalisation.
__device__ __forceinline__ int sacc_sign(long long* w) {
    for (int j = 0; j < SACC_LIMBS - 1; ++j) {
        long long carry = w[j] >> 32;  // arithmetic shift
        w[j] -= carry << 32;
        w[j + 1] += carry;
    }
    if (w[SACC_LIMBS - 1] != 0) return w[SACC_LIMBS - 1] > 0 ? 1 : -1;
    for (int j = SACC_LIMBS - 2; j >= 0; --j)
        if (w[j] != 0) return 1;
    return 0;
}

// Float64 sum that reports whether every operation was exact. Features are
// read sixteen at a time so their loads are in flight together.
__device__ __forceinline__ double exact_flag_sum(const float* __restrict__ q, const float* __restrict__ x, int d,
                                                 bool& exact) {
    double s = 0.0;
    bool ok = true;
    for (int i0 = 0; i0 < d; i0 += 16) {
        float qa[16], xa[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) {  // unconditional loads (clamped index)
            const int i = i0 + u < d ? i0 + u : d - 1;
            qa[u] = q[i];
            xa[u] = x[i];
        }
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            if (i0 + u >= d) break;
            const double A = (double)qa[u], B = -(double)xa[u];
            const double t = A + B;
            const double bb = t - A;
            ok &= ((A - (t - bb)) + (B - bb)) == 0.0;
            const double p = t * t;
            ok &= __fma_rn(t, t, -p) == 0.0;
            const double s2 = s + p;
            const double b2 = s2 - s;
            ok &= ((s - (s2 - b2)) + (p - b2)) == 0.0;
            s = s2;
        }
    }
    exact = ok;
    return s;
}

// Bitwise equality of two rows, sixteen features per batch of loads.
__device__ __forceinline__ bool rows_identical(const float* __restrict__ a, const float* __restrict__ b, int d) {
    for (int i0 = 0; i0 < d; i0 += 16) {
        uint32_t va[16], vb[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) {  // unconditional loads (clamped index)
            const int i = i0 + u < d ? i0 + u : d - 1;
            va[u] = __float_as_uint(a[i]);
            vb[u] = __float_as_uint(b[i]);
        }
        bool same = true;
#pragma unroll
        for (int u = 0; u < 16; ++u) same &= va[u] == vb[u];
        if (!same) return false;
    }
    return true;
}

// sign(S(q, a) - S(q, b)), exactly.
__device__ __noinline__ int exact_cmp(const float* __restrict__ q, const float* __restrict__ a,
                                      const float* __restrict__ b, int d) {
    if (rows_identical(a, b, d)) return 0;
    bool ea, eb;
    const double sa = exact_flag_sum(q, a, d, ea);
    const double sb = exact_flag_sum(q, b, d, eb);
    if (ea && eb) return sa < sb ? -1 : (sa > sb ? 1 : 0);
    long long w[SACC_LIMBS];
#pragma unroll
    for (int j = 0; j < SACC_LIMBS; ++j) w[j] = 0;
    for (int i = 0; i < d; ++i) {
        sacc_add_square(w, q[i], a[i], false);
        sacc_add_square(w, q[i], b[i], true);
    }
    return sacc_sign(w);
}

// sign(S(q, x) - R2key) where R2key = floor(R2 * 2^298) in 32-bit little-endian chunks.
__device__ __noinline__ int exact_cmp_key(const float* __restrict__ q, const float* __restrict__ x, int d,
                                          const uint32_t* __restrict__ r2key) {
    long long w[SACC_LIMBS];
#pragma unroll
    for (int j = 0; j < SACC_LIMBS; ++j) w[j] = -(long long)r2key[j];
    for (int i = 0; i < d; ++i) sacc_add_square(w, q[i], x[i], false);
    return sacc_sign(w);
}

// (s_a, ia) precedes (s_b, ib) in the exact (S, index) order. s_* are float64
// evaluations of S with relative error at most g.
__device__ __forceinline__ bool exact_precedes(double sa, long long ia, double sb, long long ib, double g,
                                               const float* q, const float* X, int d) {
    if (ia == ib) return false;
    if (sa * (1.0 + g) < sb * (1.0 - g)) return true;
    if (sb * (1.0 + g) < sa * (1.0 - g)) return false;
    int c = exact_cmp(q, X + (size_t)ia * d, X + (size_t)ib * d, d);
    if (c != 0) return c < 0;
    return ia < ib;
}

__device__ __forceinline__ unsigned long long ballot(bool p) { return __builtin_amdgcn_ballot_w64(p); }

}  // namespace mlk
