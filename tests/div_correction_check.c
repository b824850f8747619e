/* Checks the quotient correction of the integrate kernel (div_by_rcp in
 * cu-reconstruction_amd/csrc/recon_common.hpp) against IEEE float division:
 * y = RN(1/b), q = RN(a*y), r = fma(-q, b, a), result fma(r, y, q).
 *   1. every float32 mantissa of a in [1, 2) against every divisor 1..256
 *      (the weight update's table; a power-of-two scale of a scales every
 *      step exactly in the normal range);
 *   2. random pixel-axis quotients (pa*f)/pz over camera ranges, and random
 *      a, b over a wide exponent range with normal quotients.
 * Prints the number of mismatches of each part; exit status 1 if any. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

static float from_bits(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static int differ(float x, float y) { return memcmp(&x, &y, 4) != 0; }
static float corrected(float a, float b, float y) {
    const float q = a * y;
    return fmaf(fmaf(-q, b, a), y, q);
}
static uint64_t state = 0x9E3779B97F4A7C15ull;
static uint64_t next(void) {
    state ^= state << 13; state ^= state >> 7; state ^= state << 17;
    return state;
}
static float uniform(float lo, float hi) {
    return lo + (hi - lo) * (float)((next() >> 11) * (1.0 / 9007199254740992.0));
}

int main(void) {
    volatile float one = 1.0f;
    long bad_table = 0, bad_pixel = 0, bad_wide = 0;
    for (int k = 1; k <= 256; ++k) {
        const float b = (float)k, y = one / b;
        for (uint32_t m = 0; m < (1u << 23); ++m) {
            const float a = from_bits(0x3F800000u | m);
            bad_table += differ(corrected(a, b, y), a / b);
        }
    }
    for (long i = 0; i < 20000000; ++i) {
        const float a = uniform(-6.f, 6.f) * uniform(200.f, 700.f), pz = uniform(0.05f, 12.f);
        bad_pixel += differ(corrected(a, pz, one / pz), a / pz);
        const uint32_t ua = ((uint32_t)next() & 0x807FFFFFu) | ((uint32_t)(64 + next() % 128) << 23);
        const uint32_t ub = ((uint32_t)next() & 0x007FFFFFu) | ((uint32_t)(64 + next() % 128) << 23);
        const float a2 = from_bits(ua), b2 = from_bits(ub), q = a2 / b2;
        if (fabsf(q) > 1e-30f && fabsf(q) < 1e30f) bad_wide += differ(corrected(a2, b2, one / b2), q);
    }
    printf("table %ld pixel %ld wide %ld\n", bad_table, bad_pixel, bad_wide);
    return (bad_table || bad_pixel || bad_wide) ? 1 : 0;
}
