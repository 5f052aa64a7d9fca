/* Host check of csrc/div_rn.hpp's division by a constant (restated here in C
 * with libm's correctly rounded fmaf) against double-precision division, whose
 * rounding to FP32 is the correctly rounded quotient (double rounding from 53
 * to 24 bits is innocuous for division: 53 >= 2 * 24 + 2).  As the fused sweep
 * uses it: b = alpha within [2^-60, 2^36], y = RN(1/b), a = |p| or |q| >= 0.
 *   1. Random numerators (exponents -90 .. 63) and divisors.
 *   2. Numerators placed at rounding midpoints and at exact quotients.
 *   3. a = 0 gives +0; numerators below 2^-90 give a quotient on the same side
 *      of 1 as a / b.
 * usage: div_rn_host_check N   (N cases of each kind)
 * prints the counts; exit status 0 when every case matches. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static float fb(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static uint32_t bf(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static uint64_t st = 0x9e3779b97f4a7c15ull;
static uint64_t rnd(void) { st ^= st << 13; st ^= st >> 7; st ^= st << 17; return st; }

static float div_rn(float a, float b, float y) {
    float q0 = a * y;
    float r = fmaf(b, q0, -a);
    return fmaf(-r, y, q0);
}
static float rand_divisor(void) {   /* [2^-60, 2^36) */
    return ldexpf(fb(0x3f800000u | (uint32_t)(rnd() & 0x7fffffu)), -60 + (int)(rnd() % 96));
}

int main(int argc, char** argv) {
    long n = argc > 1 ? atol(argv[1]) : 10000000L;
    long q_cases = 0, q_bad = 0, z_bad = 0, t_bad = 0;
    for (long i = 0; i < n; ++i) {
        float b = rand_divisor(), y = (float)(1.0 / (double)b);
        /* 1 */
        float a = ldexpf(fb(0x3f800000u | (uint32_t)(rnd() & 0x7fffffu)), -90 + (int)(rnd() % 154));
        ++q_cases;
        if (bf(div_rn(a, b, y)) != bf((float)((double)a / (double)b))) ++q_bad;
        /* 2: quotients at midpoints and exact, kept within range */
        float q = ldexpf(fb(0x3f800000u | (uint32_t)(rnd() & 0x7fffffu)), -120 + (int)(rnd() % 240));
        double ulpq = (double)fb(bf(q) + 1) - (double)q;
        for (int k = 0; k < 3; ++k) {
            double mq = k == 0 ? (double)q + 0.5 * ulpq : k == 1 ? (double)q : (double)q - 0.5 * ulpq;
            float a2 = (float)(mq * (double)b);
            if (!(a2 >= 0x1p-90f) || !isfinite(a2)) continue;
            ++q_cases;
            if (bf(div_rn(a2, b, y)) != bf((float)((double)a2 / (double)b))) ++q_bad;
        }
        /* 3 */
        if (bf(div_rn(0.0f, b, y)) != 0u) ++z_bad;
        float at = fb((uint32_t)(rnd() % 0x12800000u));   /* [0, 2^-90) */
        if ((div_rn(at, b, y) > 1.0f) != ((double)at / (double)b > 1.0)) ++t_bad;
    }
    printf("quotients %ld mismatches %ld; zero mismatches %ld; tiny comparisons mismatches %ld\n",
           q_cases, q_bad, z_bad, t_bad);
    return q_bad == 0 && z_bad == 0 && t_bad == 0 ? 0 : 1;
}
