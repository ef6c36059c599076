// Host check of two exact rewrites in the device rain noise
// (master_thesis_amd/csrc/tt_noise.h, tt_erosion.h), compiled from the
// product header itself:
//   "zterm": tt_lattice_z_term(z), the integer-only form of (float)z * 149812.0f,
//            against the float product for every z in [0, 28669);
//   "fma":   fmaf(f, 2^-31, -1) against (f / 2^32) * 2 - 1 for every float f
//            that (float)hash can produce (0 and every float in [1, 2^32]);
//   "trunc": the z term with its rounding step left out (truncation), which
//            must fail: the comparison detects a wrong rounding.
// Prints the count of cases and of mismatches.
#include <math.h>
#include <stdio.h>
#include <string.h>

#include "../master_thesis_amd/csrc/tt_noise.h"

static float trunc_term(uint32_t z)
{
    const uint32_t p = z * 149812u;
    if (p == 0u) return 0.0f;
    const int nb = 32 - __builtin_clz(p);
    uint32_t q;
    int e;
    if (nb <= 24) {
        q = p << (24 - nb);
        e = nb - 24;
    } else {
        q = p >> (nb - 24);
        e = nb - 24;
    }
    return tt_u2f(((uint32_t)(150 + e) << 23) | (q & 0x007fffffu));
}

int main(int argc, char **argv)
{
    const char *mode = argc > 1 ? argv[1] : "zterm";
    long long n = 0, bad = 0;
    if (!strcmp(mode, "zterm") || !strcmp(mode, "trunc")) {
        const bool tr = !strcmp(mode, "trunc");
        for (uint32_t z = 0; z < 28669u; ++z) {
            volatile float fz = (float)z;
            const float ref = fz * 149812.0f;
            const float got = tr ? trunc_term(z) : tt_lattice_z_term(z);
            ++n;
            bad += tt_f2u(ref) != tt_f2u(got);
        }
    } else {
        // 0 and every float in [1, 2^32]
        for (uint32_t u = 0x3F800000u; u <= 0x4F800000u; ++u) {
            const float f = tt_u2f(u);
            volatile float a = f / 4294967296.0f;
            volatile float b = a * 2.0f;
            const float ref = b - 1.0f;
            const float got = fmaf(f, 0x1p-31f, -1.0f);
            ++n;
            bad += tt_f2u(ref) != tt_f2u(got);
        }
        ++n;
        bad += tt_f2u(fmaf(0.0f, 0x1p-31f, -1.0f)) != tt_f2u(((0.0f / 4294967296.0f) * 2.0f) - 1.0f);
    }
    printf("%lld %lld\n", n, bad);
    return 0;
}
