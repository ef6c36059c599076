// k_batch.hip -- erosion of many worlds in one launch per pass.
//
// A batch keeps, in device memory, one descriptor per member world (sizes,
// settings, both buffers of every ping-pong field and the buffer indices at
// the start of the step's erosion) and a tile table per tile height: entry
// (world << 20) | (tile row << 8) | tile column, so a workgroup finds its tile
// without dividing by the member's tiles per row.  One launch of k_thermal_batch runs a pair of thermal
// passes of every member (after one launch of the single first pass of the
// members with an odd pass count), one launch of k_hydraulic_batch<TH> runs hydraulic
// sub-iteration k of every member whose tiles have TH rows.  A member whose
// pass count is smaller skips the later launches.  The tile bodies are the
// one-world kernels' (tt_erosion.h), so a world's bits do not depend on the
// batch it is in.
#include "tt_kernels.h"
#include "tt_erosion.h"
#include "tt_batch.h"

// TWO: launch j runs passes of a pair (launch index j + T % 2 of the member,
// members with fewer pairs return); otherwise the launch runs the single first
// pass of the members with an odd pass count (the others return).
template <bool TWO>
__global__ __launch_bounds__(256) void k_thermal_batch(const tt_bworld *__restrict__ worlds,
                                                       const uint32_t *__restrict__ tiles, int j)
{
    __shared__ std::conditional_t<TWO, tt_therm2_lds<TT_TILE_H>, tt_therm_lds<TT_TILE_H>> S;
    const uint32_t e = tiles[tt_xcd_tile(blockIdx.x, gridDim.x)];
    const tt_bworld &B = worlds[e >> 20];
    const int odd = B.T & 1;
    if (TWO ? j >= B.T / 2 : !odd) return;
    const int l = TWO ? j + odd : 0;
    const int tx = (int)(e & 0xFFu), ty = (int)((e >> 8) & 0xFFFu);
    tt_therm_args A;
    A.W = B.W;
    A.H = B.H;
    A.talus = B.talus;
    A.amp = B.amp;
    A.str = B.str;
    A.hin = B.h[B.hc0 ^ (l & 1)];
    A.hout = B.h[B.hc0 ^ ((l + 1) & 1)];
    A.ch = nullptr;
    A.frozen = B.frozen;
    A.chg = B.chg;
    A.pass = l;
    A.bandMagic = 0;
    A.nbxMagic = 0;
    if constexpr (TWO) tt_thermal_tile2<TT_TILE_H>(A, S, tx * TT_TILE_W, ty * TT_TILE_H);
    else tt_thermal_tile<TT_TILE_H>(A, S, tx * TT_TILE_W, ty * TT_TILE_H);
}

template <int TH>
__global__ __launch_bounds__(64 * tt_hyd_waves<TH>()) void k_hydraulic_batch(const tt_bworld *__restrict__ worlds,
                                                         const uint32_t *__restrict__ tiles, int k)
{
    constexpr int NW = tt_hyd_waves<TH>();
    __shared__ tt_hyd_lds<TH> S;
    const uint32_t e = tiles[tt_xcd_tile(blockIdx.x, gridDim.x)];
    const tt_bworld &B = worlds[e >> 20];
    if (k >= B.K) return;
    const int tx = (int)(e & 0xFFu), ty = (int)((e >> 8) & 0xFFFu);
    tt_hyd_args A;
    A.W = B.W;
    A.H = B.H;
    A.rseed = B.rseed;
    A.rain = B.rain;
    A.evap = B.evap;
    A.cap = B.cap;
    A.ke = B.ke;
    A.kd = B.kd;
    A.fk = B.fk;
    A.cpxMagic = 0;
    A.g = B.g;
    A.L = B.L;
    A.rL = B.rL;
    A.divL = B.divL;
    A.sqrtOk = B.sqrtOk;
    A.cvtOk = B.cvtOk;
    A.sin003 = B.sin003;
    const int hc = B.hc1 ^ (k & 1), fc = B.fc ^ (k & 1);
    A.hin = B.h[hc];
    A.hout = B.h[hc ^ 1];
    tt_hyd_water_bufs(B.water[0], B.water[1], B.wn, B.wc, k, B.K, &A.win, &A.wout);
    A.fin = B.flux[fc];
    A.fout = B.flux[fc ^ 1];
    A.vel = B.vel;
    A.ch = nullptr;
    A.csprev = nullptr;
    A.csout = nullptr;
    if (k == 0) {
        A.sprev = B.sed[B.sc];
        A.sout = B.sed[B.sc];
    } else {
        A.sprev = B.sed[B.sc ^ ((k - 1) & 1)];
        A.sout = B.sed[B.sc ^ (k & 1)];
    }
    const int x0 = tx * TT_TILE_W, y0 = ty * TH;
    const bool first = k == 0, last = k == B.K - 1;
    // interior tiles (two-cell halo inside the grid) take the wrap-free body
    const bool in = x0 >= 2 && x0 + TT_TILE_W + 2 <= B.W && y0 >= 2 && y0 + TH + 2 <= B.H;
    if (in) {
        if (first) {
            if (last) tt_hydraulic_tile<TH, true, true, false, true, NW>(A, S, x0, y0);
            else tt_hydraulic_tile<TH, true, false, false, true, NW>(A, S, x0, y0);
        } else {
            if (last) tt_hydraulic_tile<TH, false, true, false, true, NW>(A, S, x0, y0);
            else tt_hydraulic_tile<TH, false, false, false, true, NW>(A, S, x0, y0);
        }
    } else {
        if (first) {
            if (last) tt_hydraulic_tile<TH, true, true, false, false, NW>(A, S, x0, y0);
            else tt_hydraulic_tile<TH, true, false, false, false, NW>(A, S, x0, y0);
        } else {
            if (last) tt_hydraulic_tile<TH, false, true, false, false, NW>(A, S, x0, y0);
            else tt_hydraulic_tile<TH, false, false, false, false, NW>(A, S, x0, y0);
        }
    }
}

hipError_t tt_launch_thermal_batch(hipStream_t s, const tt_bworld *worlds, const uint32_t *tiles, int ntiles, int j,
                                   bool two)
{
    if (ntiles <= 0) return hipSuccess;
    if (two) hipLaunchKernelGGL(k_thermal_batch<true>, dim3((unsigned)ntiles), dim3(256), 0, s, worlds, tiles, j);
    else hipLaunchKernelGGL(k_thermal_batch<false>, dim3((unsigned)ntiles), dim3(256), 0, s, worlds, tiles, j);
    return hipGetLastError();
}

hipError_t tt_launch_hydraulic_batch(hipStream_t s, int rows, const tt_bworld *worlds, const uint32_t *tiles,
                                     int ntiles, int k)
{
    if (ntiles <= 0) return hipSuccess;
    if (rows == 8) hipLaunchKernelGGL(k_hydraulic_batch<8>, dim3((unsigned)ntiles), dim3(64 * tt_hyd_waves<8>()), 0, s, worlds, tiles, k);
    else if (rows == 32) hipLaunchKernelGGL(k_hydraulic_batch<32>, dim3((unsigned)ntiles), dim3(64 * tt_hyd_waves<32>()), 0, s, worlds, tiles, k);
    else hipLaunchKernelGGL(k_hydraulic_batch<16>, dim3((unsigned)ntiles), dim3(64 * tt_hyd_waves<16>()), 0, s, worlds, tiles, k);
    return hipGetLastError();
}
