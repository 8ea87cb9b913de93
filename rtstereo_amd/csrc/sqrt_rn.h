/* Correctly rounded single-precision square root without double precision.
 *
 * rts_sqrt_fix(x, s) returns RN(sqrt(x)) given an approximation s within a
 * few ulp of it (the hardware's v_sqrt_f32 is within one): it moves s by
 * single ulps until sqrt(x) lies between the midpoints below and above it,
 * testing x against each squared midpoint exactly in 64-bit integers.  A
 * midpoint between a float a = Ma * 2^q and its successor is
 * (2Ma + 1) * 2^(q-1); its square has an odd 49-50 bit mantissa, which no
 * float equals, so sqrt(x) is never a tie.  Inputs below 2^-100 (denormals
 * included) are scaled by 2^64 first; the root then scales back by 2^-32,
 * exactly, since it is a normal float.
 *
 * Plain C and HIP: the device code and the CPU check (tests/sqrt_rn_check.c)
 * compile this same text.
 */
#ifndef RTS_SQRT_RN_H
#define RTS_SQRT_RN_H

#include <stdint.h>

#if defined(__HIPCC__)
#define RTS_HD __host__ __device__
#else
#define RTS_HD
#endif

static inline RTS_HD uint32_t rts_f2u(float f) {
    uint32_t u;
    __builtin_memcpy(&u, &f, 4);
    return u;
}

static inline RTS_HD float rts_u2f(uint32_t u) {
    float f;
    __builtin_memcpy(&f, &u, 4);
    return f;
}

static inline RTS_HD int rts_bitlen64(uint64_t v) { return v ? 64 - __builtin_clzll(v) : 0; }

/* value = M * 2^q for a finite float u >= 0 */
static inline RTS_HD void rts_float_parts(uint32_t u, uint64_t *M, int *q) {
    const int e = (int)(u >> 23);
    const uint32_t m = u & 0x7FFFFFu;
    *M = e ? (uint64_t)(m | 0x800000u) : (uint64_t)m;
    *q = e ? e - 150 : -149;
}

/* x >= m^2 for m the midpoint between a > 0 and the next float above a */
static inline RTS_HD int rts_sqrt_above_mid(uint32_t ux, uint32_t ua) {
    uint64_t X, Ma;
    int qx, qa;
    rts_float_parts(ux, &X, &qx);
    rts_float_parts(ua, &Ma, &qa);
    const uint64_t m2 = (2 * Ma + 1) * (2 * Ma + 1);   /* m^2 = m2 * 2^(2qa-2), m2 < 2^50 */
    const int k = qx - (2 * qa - 2);                   /* compare X * 2^k with m2 */
    const int lx = rts_bitlen64(X) + k, lm = rts_bitlen64(m2);
    if (lx != lm) return lx > lm;
    return k >= 0 ? (X << k) >= m2 : X >= (m2 << -k);
}

/* RN(sqrt(x)) for a positive normal x from an approximation s */
static inline RTS_HD float rts_sqrt_fix(float x, float s) {
    const uint32_t ux = rts_f2u(x);
    uint32_t us = rts_f2u(s);
    for (int it = 0; it < 8; it++) {
        if (rts_sqrt_above_mid(ux, us))
            us++;                                              /* sqrt(x) above the midpoint above s */
        else if (us > 0x00800000u && !rts_sqrt_above_mid(ux, us - 1))
            us--;                                              /* sqrt(x) below the midpoint below s */
        else
            break;
    }
    return rts_u2f(us);
}

/* RN(sqrt(x)) for every float x, with approx(v) an approximate square root
 * (within a few ulp for positive normal v, exact on +-0, +inf and NaN). */
static inline RTS_HD float rts_sqrt_rn_with(float x, float (*approx)(float)) {
    if (!(x > 0.0f) || !(x <= 3.40282347e38f)) return approx(x);   /* +-0, negative, NaN, +inf */
    const int tiny = x < 7.88860905e-31f;                          /* 2^-100: denormals included */
    const float v = tiny ? x * 18446744073709551616.0f : x;        /* x * 2^64, exact */
    const float r = rts_sqrt_fix(v, approx(v));
    return tiny ? r * 2.3283064365386963e-10f : r;                 /* root * 2^-32, exact */
}

#endif /* RTS_SQRT_RN_H */
