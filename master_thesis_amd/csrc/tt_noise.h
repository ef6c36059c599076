// tt_noise.h -- the integer hash and the two noise bases the terrain step uses,
// restated from the published cudaNoise formulas that the reference vendors in
// src/cuda_compute/cuda_noise.cuh.  Shared by the gfx950 kernels and the CPU
// restatement; all arithmetic is IEEE basic operations in a fixed order (no
// contraction), and signed-integer wraparound is written with uint32 so it is
// defined on the host as well.
#pragma once

#include "tt_math.h"

// Bob Jenkins' 32-bit integer mix (cuda_noise.cuh:44-54).
TT_HD uint32_t tt_hash(uint32_t seed)
{
    seed = (seed + 0x7ed55d16u) + (seed << 12);
    seed = (seed ^ 0xc761c23cu) ^ (seed >> 19);
    seed = (seed + 0x165667b1u) + (seed << 5);
    seed = (seed + 0xd3a2646cu) ^ (seed << 9);
    seed = (seed + 0xfd7046c5u) + (seed << 3);
    seed = (seed ^ 0xb55a4f09u) ^ (seed >> 16);
    return seed;
}

// hash / 0xffffffff as float, in [0, 1] (cuda_noise.cuh:66-71).
TT_HD float tt_random_float(uint32_t seed)
{
    const uint32_t v = tt_hash(seed);
    return (float)v / 4294967296.0f;  // (float)0xffffffff rounds to 2^32
}

// random value for an integer lattice point, in [-1, 1] (cuda_noise.cuh:112-115):
// the float expression x*1723 + y*93241 + z*149812 + 3824 + seed is evaluated
// left to right in float and converted to unsigned int.
TT_HD float tt_random_grid(int x, int y, int z, int seed)
{
    const float a = (float)x * 1723.0f;
    const float b = (float)y * 93241.0f;
    const float c = (float)z * 149812.0f;
    const float s = (((a + b) + c) + 3824.0f) + (float)seed;
    return tt_random_float(tt_f2u32_sat(s)) * 2.0f - 1.0f;
}

// (float)z * 149812.0f, the lattice term of tt_random_grid's third coordinate,
// for 0 <= z < 28669 (the exact product z * 149812 stays below 2^32), from
// integer operations only: the product rounded to 24 significant bits, ties
// to even, and packed as a float.  A wave-uniform z then runs on the scalar
// unit.
TT_HD float tt_lattice_z_term(uint32_t z)
{
    const uint32_t p = z * 149812u;
    if (p == 0u) return 0.0f;
    const int nb = 32 - __builtin_clz(p);  // significant bits of p
    uint32_t q;
    int e;  // p ~ q * 2^e with 2^23 <= q < 2^24
    if (nb <= 24) {
        q = p << (24 - nb);
        e = nb - 24;
    } else {
        const int sh = nb - 24;
        q = p >> sh;
        const uint32_t rem = p & ((1u << sh) - 1u), half = 1u << (sh - 1);
        if (rem > half || (rem == half && (q & 1u))) ++q;
        e = sh;
        if (q == (1u << 24)) {
            q >>= 1;
            ++e;
        }
    }
    return tt_u2f(((uint32_t)(150 + e) << 23) | (q & 0x007fffffu));
}

// nearest-lattice noise (cuda_noise.cuh:470-477).
TT_HD float tt_discrete_noise(float px, float py, float pz, float scale, int seed)
{
    const int ix = (int)(px * scale);
    const int iy = (int)(py * scale);
    const int iz = (int)(pz * scale);
    return tt_random_grid(ix, iy, iz, seed);
}

// gradient table of cuda_noise.cuh:158-160 (12 used entries).
TT_HD float tt_grad_dot(uint32_t gi, float x, float y, float z)
{
    float g0, g1, g2;
    switch (gi) {
        case 0: g0 = 1.0f; g1 = 1.0f; g2 = 0.0f; break;
        case 1: g0 = -1.0f; g1 = 1.0f; g2 = 0.0f; break;
        case 2: g0 = 1.0f; g1 = -1.0f; g2 = 0.0f; break;
        case 3: g0 = -1.0f; g1 = -1.0f; g2 = 0.0f; break;
        case 4: g0 = 1.0f; g1 = 0.0f; g2 = 1.0f; break;
        case 5: g0 = -1.0f; g1 = 0.0f; g2 = 1.0f; break;
        case 6: g0 = 1.0f; g1 = 0.0f; g2 = -1.0f; break;
        case 7: g0 = -1.0f; g1 = 0.0f; g2 = -1.0f; break;
        case 8: g0 = 0.0f; g1 = 1.0f; g2 = 1.0f; break;
        case 9: g0 = 0.0f; g1 = -1.0f; g2 = 1.0f; break;
        case 10: g0 = 0.0f; g1 = 1.0f; g2 = -1.0f; break;
        default: g0 = 0.0f; g1 = -1.0f; g2 = -1.0f; break;
    }
    return ((g0 * x) + (g1 * y)) + (g2 * z);
}

TT_HD uint32_t tt_perm12(uint32_t p) { return tt_hash(p) % 12u; }

TT_HD float tt_simplex_corner(float x, float y, float z, uint32_t gi)
{
    float t = ((0.6f - x * x) - y * y) - z * z;
    if (t < 0.0f) return 0.0f;
    t = t * t;
    return (t * t) * tt_grad_dot(gi, x, y, z);
}

// 3-D simplex noise of Gustavson / Eastman as adapted in cuda_noise.cuh:228-319.
TT_HD float tt_simplex(float px, float py, float pz, float scale, int seed)
{
    const float xin = px * scale;
    const float yin = py * scale;
    const float zin = pz * scale;
    const float F3 = 1.0f / 3.0f;
    const float G3 = 1.0f / 6.0f;
    const float s = ((xin + yin) + zin) * F3;
    const int i = (int)tt_floorf(xin + s);
    const int j = (int)tt_floorf(yin + s);
    const int k = (int)tt_floorf(zin + s);
    const float t = (float)(i + j + k) * G3;
    const float X0 = (float)i - t;
    const float Y0 = (float)j - t;
    const float Z0 = (float)k - t;
    const float x0 = xin - X0;
    const float y0 = yin - Y0;
    const float z0 = zin - Z0;
    int i1, j1, k1, i2, j2, k2;
    if (x0 >= y0) {
        if (y0 >= z0) { i1 = 1; j1 = 0; k1 = 0; i2 = 1; j2 = 1; k2 = 0; }
        else if (x0 >= z0) { i1 = 1; j1 = 0; k1 = 0; i2 = 1; j2 = 0; k2 = 1; }
        else { i1 = 0; j1 = 0; k1 = 1; i2 = 1; j2 = 0; k2 = 1; }
    } else {
        if (y0 < z0) { i1 = 0; j1 = 0; k1 = 1; i2 = 0; j2 = 1; k2 = 1; }
        else if (x0 < z0) { i1 = 0; j1 = 1; k1 = 0; i2 = 0; j2 = 1; k2 = 1; }
        else { i1 = 0; j1 = 1; k1 = 0; i2 = 1; j2 = 1; k2 = 0; }
    }
    const float x1 = (x0 - (float)i1) + G3;
    const float y1 = (y0 - (float)j1) + G3;
    const float z1 = (z0 - (float)k1) + G3;
    const float x2 = (x0 - (float)i2) + 2.0f * G3;
    const float y2 = (y0 - (float)j2) + 2.0f * G3;
    const float z2 = (z0 - (float)k2) + 2.0f * G3;
    const float x3 = (x0 - 1.0f) + 3.0f * G3;
    const float y3 = (y0 - 1.0f) + 3.0f * G3;
    const float z3 = (z0 - 1.0f) + 3.0f * G3;
    const uint32_t us = (uint32_t)seed, ui = (uint32_t)i, uj = (uint32_t)j, uk = (uint32_t)k;
    const uint32_t A = 607495u, B = 359609u, C = 654846u;
    const uint32_t gi0 = tt_perm12(us + ui * A + uj * B + uk * C);
    const uint32_t gi1 = tt_perm12(us + (ui + (uint32_t)i1) * A + (uj + (uint32_t)j1) * B + (uk + (uint32_t)k1) * C);
    const uint32_t gi2 = tt_perm12(us + (ui + (uint32_t)i2) * A + (uj + (uint32_t)j2) * B + (uk + (uint32_t)k2) * C);
    const uint32_t gi3 = tt_perm12(us + (ui + 1u) * A + (uj + 1u) * B + (uk + 1u) * C);
    const float n0 = tt_simplex_corner(x0, y0, z0, gi0);
    const float n1 = tt_simplex_corner(x1, y1, z1, gi1);
    const float n2 = tt_simplex_corner(x2, y2, z2, gi2);
    const float n3 = tt_simplex_corner(x3, y3, z3, gi3);
    return 32.0f * (((n0 + n1) + n2) + n3);
}
