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
// The four smallest of a set of distinct plate ids, ascending (256 for
// missing entries): an insertion through a min/max chain.
struct tt_min4 {
    unsigned s0, s1, s2, s3;
    __device__ __forceinline__ void clear() { s0 = s1 = s2 = s3 = 256u; }
    __device__ __forceinline__ void add(unsigned p)
    {
        unsigned t = p, a;
        a = min(s0, t); t = max(s0, t); s0 = a;
        a = min(s1, t); t = max(s1, t); s1 = a;
        a = min(s2, t); t = max(s2, t); s2 = a;
        s3 = min(s3, t);
    }
};
