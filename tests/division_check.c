/* Host check of the device's division from a stored reciprocal
 * (tt_div_by_rcp in master_thesis_amd/csrc/tt_device.h): q0 = a y,
 * q1 = q0 + (a - b q0) y, q2 = q1 + (a - b q1) y with fused multiply-adds and
 * y = RN(1 / b), against IEEE a / b.  Argument: "full" runs the sequence,
 * "product" the bare product a y (which must fail: the check has teeth).
 * Prints the count of cases and of mismatches. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

static float u2f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static uint32_t f2u(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static int full = 1;

static float div_rcp(float a, float b, float y)
{
    if (!full) return a * y;
    const float q0 = a * y;
    const float r0 = fmaf(-q0, b, a);
    const float q1 = fmaf(r0, y, q0);
    const float r1 = fmaf(-q1, b, a);
    return fmaf(r1, y, q1);
}

static uint64_t mix(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

int main(int argc, char **argv)
{
    full = !(argc > 1 && strcmp(argv[1], "product") == 0);
    long long n = 0, bad = 0;
    /* every divisor significand, against 32 numerator significands spread
     * over the binade, numerator in the divisor's binade and the next */
    for (uint32_t mb = 0; mb < (1u << 23); ++mb) {
        const float b = u2f(0x3F800000u | mb), y = 1.0f / b;
        for (uint32_t k = 0; k < 32; ++k) {
            const uint32_t ma = (uint32_t)(mix(((uint64_t)mb << 5) | k) & 0x7FFFFF);
            const float a1 = u2f(0x3F800000u | ma), a2 = u2f(0x40000000u | ma);
            n += 2;
            bad += f2u(a1 / b) != f2u(div_rcp(a1, b, y));
            bad += f2u(a2 / b) != f2u(div_rcp(a2, b, y));
        }
    }
    /* random significands and exponents in [-40, 39] */
    for (uint64_t t = 0; t < 50000000ull; ++t) {
        const uint64_t r = mix(t * 0x2545F4914F6CDD1Dull), e = mix(r);
        const float a = u2f((uint32_t)((int)(e & 0x7F) % 80 + 87) << 23 | (uint32_t)(r & 0x7FFFFF));
        const float b = u2f((uint32_t)((int)((e >> 8) & 0x7F) % 80 + 87) << 23 | (uint32_t)((r >> 23) & 0x7FFFFF));
        n += 1;
        bad += f2u(a / b) != f2u(div_rcp(a, b, 1.0f / b));
    }
    printf("%lld %lld\n", n, bad);
    return 0;
}
