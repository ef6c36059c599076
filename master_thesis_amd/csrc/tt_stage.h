// tt_stage.h -- device helpers shared by several kernels of one step.
#pragma once

#include "tt_device.h"

// continentalPropagation (plate_tectonics_kernel.cu:975-1027): voxel march
// against the plate direction while the (pre-step) height stays at or above
// continentalPropThreshold.  Coordinates are returned unwrapped.
__device__ __forceinline__ void tt_cont_prop(const tt_dev &d, const float *height, int x, int y, float dirx,
                                             float diry, float pcx, float pcy, float startH, int &ox, int &oy)
{
    ox = x;
    oy = y;
    const float thr = d.ks.continentalPropThreshold;
    if (startH < thr) return;
    const float px = (float)x + pcx, py = (float)y + pcy;
    const float ndx = -dirx, ndy = -diry;
    const int stepX = ndx > 0.0f ? 1 : -1;
    const int stepY = ndy > 0.0f ? 1 : -1;
    const float tDeltaX = tt_fabsf(1.0f / ndx);
    const float tDeltaY = tt_fabsf(1.0f / ndy);
    const int nextX = stepX > 0 ? x + 1 : x;
    const int nextY = stepY > 0 ? y + 1 : y;
    float tMaxX = ((float)nextX - px) / ndx;
    float tMaxY = ((float)nextY - py) / ndy;
    int cx = x, cy = y, qx = x, qy = y;
    for (int i = 0; i < d.ks.continentalPropMaxIters; ++i) {
        if (tMaxX < tMaxY) { cx += stepX; tMaxX = tMaxX + tDeltaX; }
        else { cy += stepY; tMaxY = tMaxY + tDeltaY; }
        const float ch = height[(size_t)tt_wrapg(cy, d.H) * d.W + tt_wrapg(cx, d.W)];
        if (ch < thr) { ox = cx; oy = cy; return; }
        qx = cx;
        qy = cy;
    }
    ox = qx;
    oy = qy;
}

// atomicInsertPlatePair (plate_tectonics_kernel.cu:1029-1069) on a word owned
// by the caller (canonical order makes the CAS loop unnecessary).
__device__ __forceinline__ bool tt_insert_pair(uint32_t &inv, int a, int b)
{
    uint32_t i0 = (uint32_t)(TT_MAXP - a), i1 = (uint32_t)(TT_MAXP - b);
    const uint32_t old = inv;
    uint32_t nw = old;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const uint32_t pl = (old >> (8 * i)) & 0xFFu;
        if (pl == 0 && i0 != 0) { nw |= i0 << (8 * i); i0 = 0; continue; }
        if (pl == i0) i0 = 0;
        if (pl == 0 && i1 != 0) { nw |= i1 << (8 * i); i1 = 0; continue; }
        if (pl == i1) i1 = 0;
    }
    if (i0 != 0 || i1 != 0) return false;
    inv = nw;
    return true;
}

// plate-id bit set of up to 256 plates held in four 64-bit registers
struct tt_pset {
    unsigned long long m0, m1, m2, m3;
    __device__ __forceinline__ void clear() { m0 = m1 = m2 = m3 = 0ull; }
    __device__ __forceinline__ void add(int p)
    {
        const unsigned long long b = 1ull << (p & 63);
        switch (p >> 6) {
            case 0: m0 |= b; break;
            case 1: m1 |= b; break;
            case 2: m2 |= b; break;
            default: m3 |= b; break;
        }
    }
    // remove and return the smallest member (256 if empty)
    __device__ __forceinline__ int pop_min()
    {
        if (m0) { const int b = __ffsll((long long)m0) - 1; m0 &= m0 - 1; return b; }
        if (m1) { const int b = __ffsll((long long)m1) - 1; m1 &= m1 - 1; return 64 + b; }
        if (m2) { const int b = __ffsll((long long)m2) - 1; m2 &= m2 - 1; return 128 + b; }
        if (m3) { const int b = __ffsll((long long)m3) - 1; m3 &= m3 - 1; return 192 + b; }
        return 256;
    }
};
