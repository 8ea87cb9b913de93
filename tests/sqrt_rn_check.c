/* The correctly rounded square root of rtstereo_amd/csrc/sqrt_rn.h against
 * the C library's (IEEE, correctly rounded) sqrtf: every positive float of
 * [1, 4) (every mantissa of both exponent parities), every denormal, every
 * integer 0..2^18 (the weight and proximity distances' inputs), specials,
 * and a sample across all exponents, each from approximations up to two ulp
 * away on either side.
 *
 * usage: sqrt_rn_check SAMPLES SEED   (exit status 0 when all agree)
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "sqrt_rn.h"

static int g_off;   /* ulp offset of the approximation */
static float approx(float v) {
    const float t = sqrtf(v);
    if (!(v > 0.0f) || isinf(v)) return t;
    uint32_t u = rts_f2u(t);
    u = (uint32_t)((int64_t)u + g_off);
    return rts_u2f(u);
}

static long bad = 0, n = 0;
static void check(float x) {
    const float want = sqrtf(x);
    for (g_off = -2; g_off <= 2; g_off++, n++) {
        const float got = rts_sqrt_rn_with(x, approx);
        if (rts_f2u(got) != rts_f2u(want) && !(isnan(got) && isnan(want))) {
            if (bad < 5) fprintf(stderr, "x=%a off=%d: %a vs %a\n", (double)x, g_off, (double)got, (double)want);
            bad++;
        }
    }
}

int main(int argc, char **argv) {
    const long samples = argc > 1 ? atol(argv[1]) : 1000000;
    uint64_t st = argc > 2 ? (uint64_t)atoll(argv[2]) * 2654435761u + 7 : 88172645463325252ull;
    for (uint32_t u = rts_f2u(1.0f); u < rts_f2u(4.0f); u++) check(rts_u2f(u));
    for (uint32_t u = 1; u < 0x00800000u; u++) check(rts_u2f(u));
    for (uint32_t i = 0; i <= (1u << 18); i++) check((float)i);
    const float sp[] = {0.0f, -0.0f, INFINITY, -INFINITY, NAN, -1.0f, 0x1p-126f, 0x1p-149f, 0x1.fffffep127f,
                        0x1p-100f, 0x1.fffffep-101f, 0x1p127f, 3.0f};
    for (size_t i = 0; i < sizeof sp / sizeof sp[0]; i++) check(sp[i]);
    for (long s = 0; s < samples; s++) {
        st ^= st << 13;
        st ^= st >> 7;
        st ^= st << 17;
        const uint32_t u = (uint32_t)(st >> 16) & 0x7F7FFFFFu;   /* positive, finite */
        check(rts_u2f(u));
    }
    printf("%ld cases, %ld mismatches\n", n, bad);
    return bad != 0;
}
