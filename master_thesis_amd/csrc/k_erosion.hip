// k_erosion.hip -- one-world kernels of stage D (thermal erosion) and stage E
// (hydraulic erosion); the tile bodies are in tt_erosion.h.  Tiles of 64
// columns by TH rows (8, 16 or 32: `tile_rows`, chosen per world size by the
// tile autotuner), XCD-aware tile order.
#include "tt_kernels.h"
#include "tt_erosion.h"

// Tile (column, row) of a thermal workgroup: bands of 4 tile rows dealt
// round-robin to the XCDs, so that neighbouring tiles share an L2 and the rough
// parts of the terrain (the tiles that do not return early) spread over all
// XCDs.  Workgroup j of XCD x takes band x + 8 (j / band) and tile j % band of
// it (band = 4 nbx tiles); the divisions are multiply-highs by host constants,
// so the map runs on the scalar unit.
__device__ __forceinline__ void tt_therm_tile_xy(const tt_therm_args &A, int nbx, int nby, int &tx, int &ty)
{
    const int blk = (int)blockIdx.x, band = 4 * nbx;
    if ((nby & 3) == 0 && ((nby >> 2) & 7) == 0) {
        const int x = blk & 7, j = blk >> 3;
        const int q = (int)__umulhi((uint32_t)j, A.bandMagic);
        const int m = j - q * band;
        const int r = nbx == 1 ? m : (int)__umulhi((uint32_t)m, A.nbxMagic);
        tx = m - r * nbx;
        ty = (x + 8 * q) * 4 + r;
    } else {
        const int b = tt_xcd_tile(blk, nbx * nby);
        tx = b % nbx;
        ty = b / nbx;
    }
}

template <int TH, bool COMP>
__global__ __launch_bounds__(256) void k_thermal(tt_therm_args A)
{
    __shared__ tt_therm_lds<TH> S;
    int tx, ty;
    tt_therm_tile_xy(A, A.W / TT_TILE_W, A.H / TH, tx, ty);
    tt_thermal_tile<TH, COMP>(A, S, tx * TT_TILE_W, ty * TH);
}

// two passes per launch (plain precision, 64x16 tiles)
__global__ __launch_bounds__(256) void k_thermal2(tt_therm_args A)
{
    __shared__ tt_therm2_lds<TT_TILE_H> S;
    int tx, ty;
    tt_therm_tile_xy(A, A.W / TT_TILE_W, A.H / TT_TILE_H, tx, ty);
    tt_thermal_tile2<TT_TILE_H>(A, S, tx * TT_TILE_W, ty * TT_TILE_H);
}

// Tile (column, row) of a hydraulic workgroup.  The dispatcher deals
// consecutive workgroups to the 8 XCDs round-robin.  A grid of at least 4096
// tiles passes through the chip in several waves of resident blocks: there each
// XCD owns a block of nbx / 8 tile columns and all XCDs advance through the
// rows together, so a wave of blocks covers one band of rows (fewer DRAM pages
// open) and a tile's neighbours above, below and beside it share its L2 except
// at the edges of the column block.  A smaller grid is resident almost at once,
// and the XCD map of contiguous tile ranges keeps neighbouring rows in one L2.
// Every map is a bijection on the tiles, so the results do not depend on it.
// The column-block map's row division is a multiply-high by a host constant,
// so the map runs on the scalar unit.
__device__ __forceinline__ void tt_hyd_tile_xy(int nbx, int nby, uint32_t cpxMagic, int &tx, int &ty)
{
    const int nb = nbx * nby, blk = (int)blockIdx.x;
    if (nb >= 4096 && (nbx & 7) == 0) {
        const int cpx = nbx >> 3, x = blk & 7, k = blk >> 3;
        const int r = cpx == 1 ? k : (int)__umulhi((uint32_t)k, cpxMagic);
        tx = x * cpx + (k - r * cpx);
        ty = r;
        return;
    }
    const int b = nb >= 4096 ? blk : tt_xcd_tile(blk, nb);
    tx = b % nbx;
    ty = b / nbx;
}

template <int TH, bool FIRST, bool LAST, bool COMP>
__global__ __launch_bounds__(64 * tt_hyd_waves<TH>()) void k_hydraulic(tt_hyd_args A)
{
    constexpr int NW = tt_hyd_waves<TH>();
    __shared__ tt_hyd_lds<TH> S;
    int tx, ty;
    tt_hyd_tile_xy(A.W / TT_TILE_W, A.H / TH, A.cpxMagic, tx, ty);
    const int x0 = tx * TT_TILE_W, y0 = ty * TH;
    if (x0 >= 2 && x0 + TT_TILE_W + 2 <= A.W && y0 >= 2 && y0 + TH + 2 <= A.H)
        tt_hydraulic_tile<TH, FIRST, LAST, COMP, true, NW>(A, S, x0, y0);
    else
        tt_hydraulic_tile<TH, FIRST, LAST, COMP, false, NW>(A, S, x0, y0);
}

int tt_tile_rows_valid(int rows, int H) { return (rows == 8 || rows == 16 || rows == 32) && (H % rows) == 0; }

// Per 64x16 tile: minimum and maximum height (a NaN yields the other operand).
__global__ __launch_bounds__(256) void k_tile_minmax(tt_dev d, const float *__restrict__ h, float2 *__restrict__ mm)
{
    __shared__ float rmin[4], rmax[4];
    const int nbx = d.W / TT_TILE_W;
    const int x0 = (int)(blockIdx.x % (unsigned)nbx) * TT_TILE_W, y0 = (int)(blockIdx.x / (unsigned)nbx) * TT_TILE_H;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    float mn = tt_u2f(0x7f800000u), mx = tt_u2f(0xff800000u);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const float v = h[(y0 + wv * 4 + r) * d.W + x0 + lane];
        mn = tt_fminf(mn, v);
        mx = tt_fmaxf(mx, v);
    }
    for (int o = 32; o > 0; o >>= 1) {
        mn = tt_fminf(mn, __shfl_xor(mn, o, 64));
        mx = tt_fmaxf(mx, __shfl_xor(mx, o, 64));
    }
    if (lane == 0) { rmin[wv] = mn; rmax[wv] = mx; }
    __syncthreads();
    if (threadIdx.x == 0)
        mm[blockIdx.x] = make_float2(tt_fminf(tt_fminf(rmin[0], rmin[1]), tt_fminf(rmin[2], rmin[3])),
                                     tt_fmaxf(tt_fmaxf(rmax[0], rmax[1]), tt_fmaxf(rmax[2], rmax[3])));
}

// A tile is frozen for T passes when the tile grown by T cells on every side
// spans at most the talus: a cell changes in a pass only through an adjacent
// pair differing by more than the talus, so after pass k the cells within
// T - k of the tile are unchanged and the region stays within the talus.
// The grown region is covered by the neighbouring tiles within ceil(T / 16)
// tile rows and ceil(T / 64) tile columns (toroidal).
__global__ __launch_bounds__(256) void k_tile_frozen(tt_dev d, const float2 *__restrict__ mm, int T,
                                                     uint8_t *__restrict__ frozen)
{
    const int nbx = d.W / TT_TILE_W, nby = d.H / TT_TILE_H;
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= nbx * nby) return;
    const int tx = t % nbx, ty = t / nbx;
    const int ry = (T + TT_TILE_H - 1) / TT_TILE_H, rx = (T + TT_TILE_W - 1) / TT_TILE_W;
    float mn = tt_u2f(0x7f800000u), mx = tt_u2f(0xff800000u);
    for (int dy = -ry; dy <= ry; ++dy)
        for (int dx = -rx; dx <= rx; ++dx) {
            const float2 v = mm[tt_wrapg(ty + dy, nby) * nbx + tt_wrapg(tx + dx, nbx)];
            mn = tt_fminf(mn, v.x);
            mx = tt_fmaxf(mx, v.y);
        }
    const float talus = d.ks.thermalThresholdAngle * d.ks.thermalCellSize;
    frozen[t] = (mx - mn <= talus) ? 1 : 0;
}

hipError_t tt_launch_thermal_frozen(hipStream_t s, const tt_dev &d, int T, const float *h, float2 *mm, uint8_t *frozen)
{
    const unsigned nt = (unsigned)((d.W / TT_TILE_W) * (d.H / TT_TILE_H));
    hipLaunchKernelGGL(k_tile_minmax, dim3(nt), dim3(256), 0, s, d, h, mm);
    hipLaunchKernelGGL(k_tile_frozen, dim3((nt + 255) / 256), dim3(256), 0, s, d, mm, T, frozen);
    return hipGetLastError();
}

hipError_t tt_launch_thermal(hipStream_t s, const tt_dev &d, int rows, const float *hin, float *hout, float *ch,
                             const uint8_t *frozen, uint8_t *chg, int pass, bool two)
{
    tt_therm_args A = tt_therm_make(d, hin, hout);
    A.ch = ch;
    A.frozen = rows == TT_TILE_H ? frozen : nullptr;
    A.chg = rows == TT_TILE_H ? chg : nullptr;
    A.pass = pass;
    const unsigned blocks = (unsigned)((d.W / TT_TILE_W) * (d.H / rows));
    if (two) {
        if (ch || rows != TT_TILE_H) return hipErrorInvalidValue;
        hipLaunchKernelGGL(k_thermal2, dim3(blocks), dim3(256), 0, s, A);
    } else if (ch) {
        if (rows == 8) hipLaunchKernelGGL((k_thermal<8, true>), dim3(blocks), dim3(256), 0, s, A);
        else if (rows == 32) hipLaunchKernelGGL((k_thermal<32, true>), dim3(blocks), dim3(256), 0, s, A);
        else hipLaunchKernelGGL((k_thermal<16, true>), dim3(blocks), dim3(256), 0, s, A);
    } else {
        if (rows == 8) hipLaunchKernelGGL((k_thermal<8, false>), dim3(blocks), dim3(256), 0, s, A);
        else if (rows == 32) hipLaunchKernelGGL((k_thermal<32, false>), dim3(blocks), dim3(256), 0, s, A);
        else hipLaunchKernelGGL((k_thermal<16, false>), dim3(blocks), dim3(256), 0, s, A);
    }
    return hipGetLastError();
}

template <int TH, bool COMP>
static void launch_hyd(hipStream_t s, unsigned blocks, bool first, bool last, const tt_hyd_args &A)
{
    const dim3 block(64 * tt_hyd_waves<TH>());
    if (first && last) hipLaunchKernelGGL((k_hydraulic<TH, true, true, COMP>), dim3(blocks), block, 0, s, A);
    else if (first) hipLaunchKernelGGL((k_hydraulic<TH, true, false, COMP>), dim3(blocks), block, 0, s, A);
    else if (last) hipLaunchKernelGGL((k_hydraulic<TH, false, true, COMP>), dim3(blocks), block, 0, s, A);
    else hipLaunchKernelGGL((k_hydraulic<TH, false, false, COMP>), dim3(blocks), block, 0, s, A);
}

hipError_t tt_launch_hydraulic(hipStream_t s, const tt_dev &d, int rows, int first, int last, int rseed,
                               const float *hin, float *hout, const float *win, float *wout, const float4 *fin,
                               float4 *fout, float2 *vel, const float *sprev, float *sout, float *ch,
                               const float *csprev, float *csout)
{
    tt_hyd_args A;
    tt_hyd_settings(d, rseed, A);
    A.hin = hin;
    A.hout = hout;
    A.win = win;
    A.wout = wout;
    A.fin = fin;
    A.fout = fout;
    A.vel = vel;
    A.sprev = sprev;
    A.sout = sout;
    A.ch = ch;
    A.csprev = csprev;
    A.csout = csout;
    const unsigned blocks = (unsigned)((d.W / TT_TILE_W) * (d.H / rows));
    if (ch) {
        if (rows == 8) launch_hyd<8, true>(s, blocks, first, last, A);
        else if (rows == 32) launch_hyd<32, true>(s, blocks, first, last, A);
        else launch_hyd<16, true>(s, blocks, first, last, A);
    } else {
        if (rows == 8) launch_hyd<8, false>(s, blocks, first, last, A);
        else if (rows == 32) launch_hyd<32, false>(s, blocks, first, last, A);
        else launch_hyd<16, false>(s, blocks, first, last, A);
    }
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_transport(tt_dev d, const float *sprev, const float2 *vel, float *sout,
                                                   const float *csprev, float *csout)
{
    const uint32_t c = blockIdx.x * 256u + threadIdx.x;
    if (c >= (uint32_t)d.N) return;
    int x, y;
    tt_cell_xy(c, d, x, y);
    tt_transport_cell(d.W, d.H, c, x, y, sprev, vel, sout);
    if (csprev) tt_transport_cell(d.W, d.H, c, x, y, csprev, vel, csout);
}

hipError_t tt_launch_transport(hipStream_t s, const tt_dev &d, const float *sprev, const float2 *vel, float *sout,
                               const float *csprev, float *csout)
{
    const unsigned blocks = (unsigned)(((long long)d.N + 255) / 256);
    hipLaunchKernelGGL(k_transport, dim3(blocks), dim3(256), 0, s, d, sprev, vel, sout, csprev, csout);
    return hipGetLastError();
}
