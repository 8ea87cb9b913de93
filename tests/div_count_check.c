/* Exhaustive-over-counts check of the count quotient used by the cross
 * sweep: q0 = a*y, r0 = fma(-q0, c, a), q1 = fma(r0, y, q0) with y within two
 * ulp of 1/c, falling back to a/c when q1 is NaN, infinite, denormal or zero, must
 * equal the IEEE quotient a/c bit for bit.  The hardware reciprocal is
 * replaced by every float within two ulp of RN(1/c), so the check covers any
 * reciprocal of that accuracy.
 *
 * usage: div_count_check CMAX SAMPLES SEED   (exit status 0 when all agree)
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static uint32_t bits(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static float from_bits(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }

static uint64_t rng_state;
static uint32_t rnd(void) {
    rng_state ^= rng_state << 13;
    rng_state ^= rng_state >> 7;
    rng_state ^= rng_state << 17;
    return (uint32_t)(rng_state >> 11);
}

static float quotient(float a, float c, float y) {
    const float q0 = a * y;
    const float r0 = fmaf(-q0, c, a);
    const float q1 = fmaf(r0, y, q0);
    const int cls = fpclassify(q1);
    if (cls == FP_NAN || cls == FP_INFINITE || cls == FP_SUBNORMAL || cls == FP_ZERO) return a / c;
    return q1;
}

static long check(float a, uint32_t c) {
    const float cf = (float)c;
    const float want = a / cf;
    const float y0 = 1.0f / cf;
    long bad = 0;
    for (int k = -2; k <= 2; k++) {
        const float y = from_bits(bits(y0) + (uint32_t)k);
        const float got = quotient(a, cf, y);
        if (bits(got) != bits(want) && !(isnan(got) && isnan(want))) {
            if (bad == 0)
                fprintf(stderr, "mismatch a=%a c=%u y=%a: %a vs %a\n", (double)a, c, (double)y, (double)got,
                        (double)want);
            bad++;
        }
    }
    return bad;
}

int main(int argc, char **argv) {
    const uint32_t cmax = argc > 1 ? (uint32_t)atol(argv[1]) : 1225;
    const long samples = argc > 2 ? atol(argv[2]) : 20000;
    rng_state = argc > 3 ? (uint64_t)atoll(argv[3]) * 2654435761u + 1 : 88172645463325252ull;
    long bad = 0, n = 0;
    const float specials[] = {0.0f, -0.0f, 1.0f, -1.0f, 0x1p-126f, 0x1.fffffep-127f, 0x1p-149f, 0x1p-140f, 0x1p-115f,
                              0x1.fffffep127f, -0x1.fffffep127f, INFINITY, -INFINITY, NAN, 3.0f, 1e-30f, 12345.678f};
    for (uint32_t c = 1; c <= cmax; c++) {
        for (size_t i = 0; i < sizeof specials / sizeof specials[0]; i++, n++) bad += check(specials[i], c);
        /* multiples of c: exact quotients, and their neighbours */
        for (int m = 1; m <= 64; m++) {
            const float a = (float)c * (float)m;
            bad += check(a, c) + check(nextafterf(a, 0.0f), c) + check(nextafterf(a, INFINITY), c);
            n += 3;
        }
        for (long s = 0; s < samples; s++, n++) {
            /* random sign and mantissa, exponents across the whole range
               (normal and denormal dividends included) */
            uint32_t u = rnd();
            const uint32_t e = rnd() % 256;
            u = (u & 0x807FFFFFu) | (e << 23);
            if (e == 255) u &= 0x807FFFFFu;
            bad += check(from_bits(u), c);
        }
    }
    printf("%ld cases x 5 reciprocals, %ld mismatches\n", n, bad);
    return bad != 0;
}
