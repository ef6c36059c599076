// k_erosion.hip -- one-world kernels of stage D (thermal erosion) and stage E
// (hydraulic erosion); the tile bodies are in tt_erosion.h.  Tiles of 64
// columns by TH rows (8, 16 or 32: `tile_rows`, chosen per world size by the
// tile autotuner), XCD-aware tile order.
#include "tt_kernels.h"
#include "tt_erosion.h"

template <int TH, bool COMP>
__global__ __launch_bounds__(256) void k_thermal(tt_therm_args A)
{
    __shared__ tt_therm_lds<TH> S;
    const int nbx = A.W / TT_TILE_W, nby = A.H / TH;
    const int b = tt_xcd_tile(blockIdx.x, nbx * nby);
    tt_thermal_tile<TH, COMP>(A, S, (b % nbx) * TT_TILE_W, (b / nbx) * TH);
}

template <int TH, bool FIRST, bool LAST, bool COMP>
__global__ __launch_bounds__(256) void k_hydraulic(tt_hyd_args A)
{
    __shared__ tt_hyd_lds<TH> S;
    const int nbx = A.W / TT_TILE_W, nby = A.H / TH;
    const int b = tt_xcd_tile(blockIdx.x, nbx * nby);
    tt_hydraulic_tile<TH, FIRST, LAST, COMP>(A, S, (b % nbx) * TT_TILE_W, (b / nbx) * TH);
}

int tt_tile_rows_valid(int rows, int H) { return (rows == 8 || rows == 16 || rows == 32) && (H % rows) == 0; }

hipError_t tt_launch_thermal(hipStream_t s, const tt_dev &d, int rows, const float *hin, float *hout, float *ch)
{
    tt_therm_args A = tt_therm_make(d, hin, hout);
    A.ch = ch;
    const unsigned blocks = (unsigned)((d.W / TT_TILE_W) * (d.H / rows));
    if (ch) {
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
    if (first && last) hipLaunchKernelGGL((k_hydraulic<TH, true, true, COMP>), dim3(blocks), dim3(256), 0, s, A);
    else if (first) hipLaunchKernelGGL((k_hydraulic<TH, true, false, COMP>), dim3(blocks), dim3(256), 0, s, A);
    else if (last) hipLaunchKernelGGL((k_hydraulic<TH, false, true, COMP>), dim3(blocks), dim3(256), 0, s, A);
    else hipLaunchKernelGGL((k_hydraulic<TH, false, false, COMP>), dim3(blocks), dim3(256), 0, s, A);
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
