// tt_erosion.h -- tile bodies of stage D (thermal erosion) and stage E
// (hydraulic erosion), shared by the one-world kernels (k_erosion.hip) and the
// batched kernels over many worlds (k_batch.hip).
//
// Tiles are 64 columns by TH rows (TH = 8, 16 or 32); lane l of a wavefront
// owns column l of its rows.  A thermal block has 4 wavefronts (TH/4 rows
// each), a hydraulic block TH/2 (two rows each, tt_hyd_waves).  The kernel owns
// the LDS and passes it in, so a kernel that can run several tile variants
// still allocates it once.
//
// Thermal: thermalErosionKernel (plate_tectonics_kernel.cu:2735-2786) in the
// canonical Jacobi form -- each pass reads the previous heights; a cell's new
// height is its old height minus its own outflows (directions 0..7) plus the
// inflows it receives (directions 0..7).  Two-cell halo.
//
// Hydraulic: one pass per sub-iteration fuses rain (:2430-2437), flux
// (:2439-2485), flow (:2487-2533), evaporate (:2587-2593) and the
// erosion/deposition exchange (:2535-2570 with CudaTexture::Slope), and begins
// with the semi-Lagrangian transport (:2572-2585, cuda_helper.cuh:49-65) of the
// previous sub-iteration, gathered from its sediment and velocity.  The order
// of float operations is the reference's; dt = 1 (plate_tectonic_sim.cu:338-363).
#pragma once

#include <type_traits>

#include "tt_device.h"

#define TT_EW (TT_TILE_W + 4)   // tile width with a two-cell halo
#define TT_FW (TT_TILE_W + 2)   // tile width with a one-cell halo

__device__ __constant__ static const int c_toff[8][2] = {{-1, 0}, {1, 0}, {0, -1}, {0, 1},
                                                         {-1, -1}, {-1, 1}, {1, -1}, {1, 1}};

// ---------------------------------------------------------------------------
// thermal
// ---------------------------------------------------------------------------
struct tt_therm_args {
    int W, H;
    float talus, amp, str;
    const float *hin;
    float *hout;
    float *ch;              // precision mode: Kahan compensation of the height (in place)
    const uint8_t *frozen;  // per 64x16 tile: unchanged by every pass of this step (null: unknown)
    uint8_t *chg;           // per 64x16 tile and pass parity: the pass changed a cell (null: not tracked)
    int pass;               // pass index within the step
    uint32_t bandMagic;     // ceil(2^32 / (4 nbx)) and ceil(2^32 / nbx) (0 for nbx = 1): the
    uint32_t nbxMagic;      // tile map's divisions as multiply-highs (nbx = W / 64)
};

// Kahan-compensated sum += d: c carries the low-order part lost so far
__device__ __forceinline__ float tt_kahan_add(float sum, float &c, float d)
{
    const float y = d - c;
    const float t = sum + y;
    c = (t - sum) - y;
    return t;
}

template <int TH>
struct tt_therm_lds {
    float hs[TH + 4][TT_EW];
    float tot[TH + 2][TT_FW];
    float rcp[TH + 2][TT_FW];   // RN(1 / tot) where tot > 0
    float rmin[4], rmax[4];
};

__host__ __device__ inline tt_therm_args tt_therm_make(const tt_dev &d, const float *hin, float *hout)
{
    tt_therm_args A;
    A.W = d.W;
    A.H = d.H;
    A.talus = d.ks.thermalThresholdAngle * d.ks.thermalCellSize;
    A.amp = d.ks.thermalErosionAmplitude;
    A.str = d.ks.thermalErosionStrength;
    A.hin = hin;
    A.hout = hout;
    A.ch = nullptr;
    A.frozen = nullptr;
    A.chg = nullptr;
    A.pass = 0;
    // exact for every block index below 2^16 and every nbx up to 256 (checked
    // exhaustively on the host)
    const uint32_t nbx = (uint32_t)(d.W / TT_TILE_W);
    A.bandMagic = 0xFFFFFFFFu / (4u * nbx) + 1u;
    A.nbxMagic = nbx >= 2 ? 0xFFFFFFFFu / nbx + 1u : 0u;
    return A;
}

template <int TH, bool COMP = false>
__device__ __forceinline__ void tt_thermal_tile(const tt_therm_args &A, tt_therm_lds<TH> &S, int x0, int y0)
{
    constexpr int RPW = TH / 4;
    const int tid = threadIdx.x;
    const int W = A.W, H = A.H;
    const int lane = tid & 63, wv = tid >> 6;
    const int nbx = W / TT_TILE_W, nby = H / TT_TILE_H;
    const int tile = (y0 / TT_TILE_H) * nbx + x0 / TT_TILE_W;
    uint8_t *chgOut = (A.chg && TH == TT_TILE_H) ? A.chg + (size_t)(A.pass & 1) * (nbx * nby) + tile : nullptr;
    // A frozen tile keeps its heights through every pass of the step: the
    // first pass copies it to the other buffer, later passes find both
    // buffers equal and do nothing.
    // A cell's update reads the heights within two cells (and its own
    // compensation).  If no cell of this tile or of its eight neighbouring
    // tiles changed in the previous pass, this pass computes what the previous
    // one did: nothing changes, and the output buffer (the previous pass's
    // input) already holds the heights.
    // Both flags are loaded together (one memory round trip), and a tile that
    // is not frozen issues its height loads before it learns whether it is
    // unchanged.
    const bool track = chgOut && A.pass > 0;
    const bool fz = A.frozen && TH == TT_TILE_H && A.frozen[tile];
    int any = 0;
    if (track && tid < 9) {
        const uint8_t *prev = A.chg + (size_t)((A.pass - 1) & 1) * (nbx * nby);
        const int tx = x0 / TT_TILE_W, ty = y0 / TT_TILE_H;
        const int dy = tid / 3 - 1, dx = tid % 3 - 1;
        any = prev[tt_wrap1(ty + dy, nby) * nbx + tt_wrap1(tx + dx, nbx)];
    }
    if (fz) {
        if (A.pass == 0)
#pragma unroll
            for (int r = 0; r < RPW; ++r) {
                const int c = (y0 + wv * RPW + r) * W + x0 + lane;
                A.hout[c] = A.hin[c];
            }
        if (chgOut && tid == 0) *chgOut = 0;
        return;
    }
    constexpr int NHR = (TH + 4 + 3) / 4;   // halo rows per wavefront
    float ha[NHR], hb[NHR];
    {
        const int gxa = tt_wrap1(x0 + lane - 2, W), gxb = tt_wrap1(x0 + 64 + lane - 2, W);
#pragma unroll
        for (int j = 0; j < NHR; ++j) {
            const int ly = wv + 4 * j;
            if (ly < TH + 4) {
                const int row = tt_wrap1(y0 + ly - 2, H) * W;
                ha[j] = A.hin[row + gxa];
                if (lane < TT_EW - 64) hb[j] = A.hin[row + gxb];
            }
        }
    }
    if (track && !__syncthreads_or(any)) {
        if (tid == 0) *chgOut = 0;
        return;
    }
    float mn = tt_u2f(0x7f800000u), mx = tt_u2f(0xff800000u);
#pragma unroll
    for (int j = 0; j < NHR; ++j) {
        const int ly = wv + 4 * j;
        if (ly < TH + 4) {
            S.hs[ly][lane] = ha[j];
            mn = tt_fminf(mn, ha[j]);
            mx = tt_fmaxf(mx, ha[j]);
            if (lane < TT_EW - 64) {
                S.hs[ly][64 + lane] = hb[j];
                mn = tt_fminf(mn, hb[j]);
                mx = tt_fmaxf(mx, hb[j]);
            }
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        mn = tt_fminf(mn, __shfl_xor(mn, o, 64));
        mx = tt_fmaxf(mx, __shfl_xor(mx, o, 64));
    }
    if (lane == 0) { S.rmin[wv] = mn; S.rmax[wv] = mx; }
    __syncthreads();
    const float talus = A.talus, amp = A.amp, str = A.str;
    // No transfer anywhere in the tile when the tile plus its two-cell halo
    // spans at most the talus: every difference h_a - h_b rounds to at most
    // max - min (rounding is monotone), so no "diff > talus" test can pass and
    // every cell keeps its height exactly.
    {
        const float tmn = tt_fminf(tt_fminf(S.rmin[0], S.rmin[1]), tt_fminf(S.rmin[2], S.rmin[3]));
        const float tmx = tt_fmaxf(tt_fmaxf(S.rmax[0], S.rmax[1]), tt_fmaxf(S.rmax[2], S.rmax[3]));
        if (tmx - tmn <= talus) {
#pragma unroll
            for (int r = 0; r < RPW; ++r) {
                const int ry = wv * RPW + r;
                A.hout[(y0 + ry) * W + x0 + lane] = S.hs[ry + 2][lane + 2];
            }
            if (chgOut && tid == 0) *chgOut = 0;
            return;
        }
    }
    // sum of the downhill differences above the talus, on the tile + 1 halo,
    // and its reciprocal: a cell's total divides its 8 outflows here and is
    // the divisor of its neighbours' inflows from it
    int wide = 0;
    for (int i = tid; i < (TH + 2) * TT_FW; i += 256) {
        const int ty = i / TT_FW, tx = i - ty * TT_FW;
        const int ly = ty + 1, lx = tx + 1;
        const float hc = S.hs[ly][lx];
        float nb[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) nb[k] = S.hs[ly + c_toff[k][1]][lx + c_toff[k][0]];
        float t = 0.0f;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const float diff = hc - nb[k];
            t = (diff > talus) ? t + diff : t;
        }
        S.tot[ty][tx] = t;
        S.rcp[ty][tx] = t > 0.0f ? 1.0f / t : 0.0f;
        wide |= t > 0.0f && !(t <= 0x1p40f);
    }
    // Every quotient is diff / total with talus < diff <= total (the total
    // sums the diff with other positive terms).  With the talus at least
    // 2^-40 and every total at most 2^40 both lie in tt_div_by_rcp's range,
    // which then gives the IEEE quotient from the stored reciprocal.
    const bool byRcp = !__syncthreads_or(wide) && A.talus >= 0x1p-40f;
    int changed = 0;
#pragma unroll
    for (int r = 0; r < RPW; ++r) {
        const int ry = wv * RPW + r;
        const int ly = ry + 2, lx = lane + 2;
        const float hc = S.hs[ly][lx];
        const float tc = S.tot[ry + 1][lane + 1];
        float dOut[8], dIn[8], tIn[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            dOut[k] = hc - S.hs[ly + c_toff[k][1]][lx + c_toff[k][0]];
            const int sy = ly - c_toff[k][1], sx = lx - c_toff[k][0];
            dIn[k] = S.hs[sy][sx] - hc;
            tIn[k] = S.tot[sy - 1][sx - 1];
        }
        // own outflows (directions 0..7), then inflows (directions 0..7); the
        // divisions sit in branches that a wavefront skips when no lane needs them
        float v = hc;
        const int cidx = (y0 + ry) * W + x0 + lane;
        float cc = 0.0f;
        if (COMP) cc = A.ch[cidx];
        const float cc0 = cc;
        if (byRcp) {
            if (tc > 0.0f) {
                const float rc = S.rcp[ry + 1][lane + 1];
#pragma unroll
                for (int k = 0; k < 8; ++k)
                    if (dOut[k] > talus) {
                        const float q = (amp * tt_div_by_rcp(dOut[k], tc, rc)) * str;
                        v = COMP ? tt_kahan_add(v, cc, -q) : v - q;
                    }
            }
#pragma unroll
            for (int k = 0; k < 8; ++k)
                if (dIn[k] > talus && tIn[k] > 0.0f) {
                    const int sy = ly - c_toff[k][1], sx = lx - c_toff[k][0];
                    const float q = (amp * tt_div_by_rcp(dIn[k], tIn[k], S.rcp[sy - 1][sx - 1])) * str;
                    v = COMP ? tt_kahan_add(v, cc, q) : v + q;
                }
        } else {
            if (tc > 0.0f) {
#pragma unroll
                for (int k = 0; k < 8; ++k)
                    if (dOut[k] > talus) {
                        const float q = (amp * (dOut[k] / tc)) * str;
                        v = COMP ? tt_kahan_add(v, cc, -q) : v - q;
                    }
            }
#pragma unroll
            for (int k = 0; k < 8; ++k)
                if (dIn[k] > talus && tIn[k] > 0.0f) {
                    const float q = (amp * (dIn[k] / tIn[k])) * str;
                    v = COMP ? tt_kahan_add(v, cc, q) : v + q;
                }
        }
        if (COMP) {
            changed |= __float_as_uint(cc) != __float_as_uint(cc0);
            A.ch[cidx] = cc;
        }
        changed |= __float_as_uint(v) != __float_as_uint(hc);
        A.hout[cidx] = v;
    }
    if (chgOut) {
        changed = __syncthreads_or(changed);
        if (tid == 0) *chgOut = changed ? 1 : 0;
    }
}

// Two passes in one launch.  The tile is loaded with a four-cell halo; the
// first pass computes the tile plus two cells (what the second pass reads) in
// LDS, the second pass the tile.  Every cell's update is the one-pass update of
// the same heights, so the result is bit for bit that of two single passes,
// with half the loads, stores and launches.
template <int TH>
struct tt_therm2_lds {
    float h0[TH + 8][TT_TILE_W + 8];    // input heights, four-cell halo
    float h1[TH + 4][TT_EW];            // the first pass's heights, two-cell halo
    float tot[TH + 6][TT_TILE_W + 6];   // totals of the pass being computed
    float rcp[TH + 6][TT_TILE_W + 6];   // RN(1 / tot) where tot > 0
    float rmin[4], rmax[4];
};

// Totals of the downhill differences above the talus, and their reciprocals,
// on an R x C region whose cell (ty, tx) sits at hs[ty + 1][tx + 1]; the sum
// runs over directions 0..7 as in tt_thermal_tile.  Returns whether some total
// lies beyond 2^40 (outside tt_div_by_rcp's range).
template <int R, int C, int HP, int TP>
__device__ __forceinline__ int tt_therm_totals(const float (*hs)[HP], float (*tot)[TP], float (*rcp)[TP], float talus)
{
    int wide = 0;
    for (int i = threadIdx.x; i < R * C; i += 256) {
        const int ty = i / C, tx = i - ty * C;
        const int ly = ty + 1, lx = tx + 1;
        const float hc = hs[ly][lx];
        float nb[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) nb[k] = hs[ly + c_toff[k][1]][lx + c_toff[k][0]];
        float t = 0.0f;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const float diff = hc - nb[k];
            t = (diff > talus) ? t + diff : t;
        }
        tot[ty][tx] = t;
        rcp[ty][tx] = t > 0.0f ? 1.0f / t : 0.0f;
        wide |= t > 0.0f && !(t <= 0x1p40f);
    }
    return wide;
}

// New height of the cell at hs[ly][lx] from the heights hs and the totals of
// the same pass (a cell's total at tot[ly - 1][lx - 1]): own outflows
// (directions 0..7), then inflows (directions 0..7), as tt_thermal_tile.
template <int HP, int TP>
__device__ __forceinline__ float tt_therm_cell(const float (*hs)[HP], const float (*tot)[TP], const float (*rcp)[TP],
                                               int ly, int lx, float talus, float amp, float str, bool byRcp)
{
    const float hc = hs[ly][lx];
    const float tc = tot[ly - 1][lx - 1];
    float dOut[8], dIn[8], tIn[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        dOut[k] = hc - hs[ly + c_toff[k][1]][lx + c_toff[k][0]];
        const int sy = ly - c_toff[k][1], sx = lx - c_toff[k][0];
        dIn[k] = hs[sy][sx] - hc;
        tIn[k] = tot[sy - 1][sx - 1];
    }
    float v = hc;
    if (byRcp) {
        if (tc > 0.0f) {
            const float rc = rcp[ly - 1][lx - 1];
#pragma unroll
            for (int k = 0; k < 8; ++k)
                if (dOut[k] > talus) v = v - (amp * tt_div_by_rcp(dOut[k], tc, rc)) * str;
        }
#pragma unroll
        for (int k = 0; k < 8; ++k)
            if (dIn[k] > talus && tIn[k] > 0.0f) {
                const int sy = ly - c_toff[k][1], sx = lx - c_toff[k][0];
                v = v + (amp * tt_div_by_rcp(dIn[k], tIn[k], rcp[sy - 1][sx - 1])) * str;
            }
    } else {
        if (tc > 0.0f) {
#pragma unroll
            for (int k = 0; k < 8; ++k)
                if (dOut[k] > talus) v = v - (amp * (dOut[k] / tc)) * str;
        }
#pragma unroll
        for (int k = 0; k < 8; ++k)
            if (dIn[k] > talus && tIn[k] > 0.0f) v = v + (amp * (dIn[k] / tIn[k])) * str;
    }
    return v;
}

// Passes p and p + 1 of the step on one 64 x TH tile (256 threads; not for
// the compensated mode, whose passes run one at a time).  The frozen and
// change flags keep their meaning with a launch in place of a pass: a launch
// whose input equals the previous launch's input on the tile and its eight
// neighbouring tiles (the two passes read heights within four cells) writes
// what the previous launch wrote, which its output buffer already holds.
template <int TH>
__device__ __forceinline__ void tt_thermal_tile2(const tt_therm_args &A, tt_therm2_lds<TH> &S, int x0, int y0)
{
    static_assert(TH == TT_TILE_H, "frozen and change flags are kept per 64x16 tile");
    constexpr int RPW = TH / 4;
    constexpr int NHR = (TH + 8) / 4;   // halo rows per wavefront
    constexpr int XC = 8;               // halo columns beyond the first 64 (60..67)
    const int tid = threadIdx.x;
    const int W = A.W, H = A.H;
    const int lane = tid & 63, wv = tid >> 6;
    const int nbx = W / TT_TILE_W, nby = H / TT_TILE_H;
    const int tile = (y0 / TT_TILE_H) * nbx + x0 / TT_TILE_W;
    uint8_t *chgOut = A.chg ? A.chg + (size_t)(A.pass & 1) * (nbx * nby) + tile : nullptr;
    const bool track = chgOut && A.pass > 0;
    const bool fz = A.frozen && A.frozen[tile];
    int any = 0;
    if (track && tid < 9) {
        const uint8_t *prev = A.chg + (size_t)((A.pass - 1) & 1) * (nbx * nby);
        const int tx = x0 / TT_TILE_W, ty = y0 / TT_TILE_H;
        const int dy = tid / 3 - 1, dx = tid % 3 - 1;
        any = prev[tt_wrap1(ty + dy, nby) * nbx + tt_wrap1(tx + dx, nbx)];
    }
    if (fz) {
        if (A.pass == 0)
#pragma unroll
            for (int r = 0; r < RPW; ++r) {
                const int c = (y0 + wv * RPW + r) * W + x0 + lane;
                A.hout[c] = A.hin[c];
            }
        if (chgOut && tid == 0) *chgOut = 0;
        return;
    }
    float ha[NHR], hb[NHR];
    {
        const int gxa = tt_wrap1(x0 + lane - 4, W), gxb = tt_wrap1(x0 + 60 + lane, W);
#pragma unroll
        for (int j = 0; j < NHR; ++j) {
            const int row = tt_wrap1(y0 + wv + 4 * j - 4, H) * W;
            ha[j] = A.hin[row + gxa];
            if (lane < XC) hb[j] = A.hin[row + gxb];
        }
    }
    if (track && !__syncthreads_or(any)) {
        if (tid == 0) *chgOut = 0;
        return;
    }
    float mn = tt_u2f(0x7f800000u), mx = tt_u2f(0xff800000u);
#pragma unroll
    for (int j = 0; j < NHR; ++j) {
        const int ly = wv + 4 * j;
        S.h0[ly][lane] = ha[j];
        mn = tt_fminf(mn, ha[j]);
        mx = tt_fmaxf(mx, ha[j]);
        if (lane < XC) {
            S.h0[ly][64 + lane] = hb[j];
            mn = tt_fminf(mn, hb[j]);
            mx = tt_fmaxf(mx, hb[j]);
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        mn = tt_fminf(mn, __shfl_xor(mn, o, 64));
        mx = tt_fmaxf(mx, __shfl_xor(mx, o, 64));
    }
    if (lane == 0) { S.rmin[wv] = mn; S.rmax[wv] = mx; }
    __syncthreads();
    const float talus = A.talus, amp = A.amp, str = A.str;
    // a tile plus four cells within the talus: neither pass moves anything
    // there (tt_thermal_tile's argument, applied to each pass)
    {
        const float tmn = tt_fminf(tt_fminf(S.rmin[0], S.rmin[1]), tt_fminf(S.rmin[2], S.rmin[3]));
        const float tmx = tt_fmaxf(tt_fmaxf(S.rmax[0], S.rmax[1]), tt_fmaxf(S.rmax[2], S.rmax[3]));
        if (tmx - tmn <= talus) {
#pragma unroll
            for (int r = 0; r < RPW; ++r) {
                const int ry = wv * RPW + r;
                A.hout[(y0 + ry) * W + x0 + lane] = S.h0[ry + 4][lane + 4];
            }
            if (chgOut && tid == 0) *chgOut = 0;
            return;
        }
    }
    const bool rcpOk = talus >= 0x1p-40f;
    // first pass: totals on the tile plus three cells, new heights on the tile
    // plus two
    const bool byA = !__syncthreads_or(tt_therm_totals<TH + 6, TT_TILE_W + 6>(S.h0, S.tot, S.rcp, talus)) && rcpOk;
    for (int i = tid; i < (TH + 4) * TT_EW; i += 256) {
        const int ry = i / TT_EW, rx = i - ry * TT_EW;
        S.h1[ry][rx] = tt_therm_cell(S.h0, S.tot, S.rcp, ry + 2, rx + 2, talus, amp, str, byA);
    }
    __syncthreads();
    // second pass: totals on the tile plus one cell, new heights on the tile
    const bool byB = !__syncthreads_or(tt_therm_totals<TH + 2, TT_FW>(S.h1, S.tot, S.rcp, talus)) && rcpOk;
    int changed = 0;
#pragma unroll
    for (int r = 0; r < RPW; ++r) {
        const int ry = wv * RPW + r;
        const float v = tt_therm_cell(S.h1, S.tot, S.rcp, ry + 2, lane + 2, talus, amp, str, byB);
        changed |= __float_as_uint(v) != __float_as_uint(S.h0[ry + 4][lane + 4]);
        A.hout[(y0 + ry) * W + x0 + lane] = v;
    }
    if (chgOut) {
        changed = __syncthreads_or(changed);
        if (tid == 0) *chgOut = changed ? 1 : 0;
    }
}

// ---------------------------------------------------------------------------
// hydraulic
// ---------------------------------------------------------------------------
struct tt_hyd_args {
    int W, H, rseed;
    float rain, evap, cap, ke, kd, fk, g, L;
    float rL;             // RN(1 / L)
    int divL;             // 1: the division by L may take tt_div_const
    int sqrtOk;           // 1: square roots of normal arguments may take tt_sqrt_normal
    int cvtOk;            // 1: the rain's saturating conversion may take tt_f2u32_hw
    float sin003;         // tt_sinf(0.03f): the capacity's sine below the slope threshold
    uint32_t cpxMagic;    // ceil(2^32 / (W / 512)) for the column-block tile map (0: W < 1024)
    const float *hin;
    float *hout;
    const float *win;     // first sub-iteration: the state water; later: the previous one's flow water
    float *wout;          // last sub-iteration: the state water after evaporation; earlier: the flow water
    const float4 *fin;
    float4 *fout;
    float2 *vel;          // written by the last sub-iteration only
    const float *sprev;
    float *sout;
    float *ch;            // precision mode: height compensation (in place)
    const float *csprev;  // precision mode: sediment compensation, gathered like sprev
    float *csout;
};

template <int TH>
struct tt_hyd_lds {
    float M[TH + 4][TT_EW];
    float Wt[TH + 4][TT_EW];
    float4 F[TH + 2][TT_FW];
};

__host__ __device__ inline void tt_hyd_settings(const tt_dev &d, int rseed, tt_hyd_args &A)
{
    A.W = d.W;
    A.H = d.H;
    A.rseed = rseed;
    A.rain = d.ks.hydrationRainfall;
    A.evap = d.ks.hydrationEvaporation;
    A.cap = d.ks.sedimentCapacity;
    A.ke = d.ks.sedimentErosionRate;
    A.kd = d.ks.sedimentDepositionRate;
    A.fk = 1.0f * d.ks.hydrationPipeCrossSection;
    A.g = d.ks.gravity;
    A.L = d.ks.hydrationPipeLength;
    A.rL = d.rL;
    A.divL = d.divL;
    A.sqrtOk = d.sqrtOk;
    A.cvtOk = d.cvtOk;
    A.sin003 = tt_sinf(0.03f);
    // k / cpx = umulhi(k, ceil(2^32 / cpx)) for every block index k < 2^17 and
    // cpx in 2..32 (checked exhaustively on the host)
    const uint32_t cpx = (uint32_t)(d.W / TT_TILE_W) >> 3;
    A.cpxMagic = cpx >= 2 ? 0xFFFFFFFFu / cpx + 1u : 0u;
}

// Squared slopes up to this bound have a square root below 0.03f (the root is
// monotone and 0.00089 < 0.03f^2), so max(slope, 0.03f) is 0.03f there.
#define TT_SLOPE2_SMALL 0.00089f

// IEEE square root; a wavefront whose arguments all lie in [2^-96, FLT_MAX]
// takes tt_sqrt_normal when it has been checked (ok)
__device__ __forceinline__ float tt_sqrt_wave(float x, int ok)
{
    if (__builtin_amdgcn_readfirstlane(ok) && __builtin_amdgcn_ballot_w64(!tt_sqrt_in_range(x)) == 0)
        return tt_sqrt_normal(x);
    return sqrtf(x);
}

// bilinear sample with toroidal wrap (interpolate, cuda_helper.cuh:49-65).
// |v| <= 50 keeps the sample column within [-51, W + 50], where a single
// conditional wrap is exact (tt_wrap_near); other columns (non-finite
// velocities) take the general modulo.  IEEE floor is exact, so the hardware
// floor equals the oracle's bit-level floor.

// the four taps and the weights of a sample, split from the mix so that a
// caller can issue the tap loads early
struct tt_taps {
    int a, b, c, d;  // (x0, y0), (x0 + 1, y0), (x0, y0 + 1), (x0 + 1, y0 + 1)
    float wx, wy;
};

__device__ __forceinline__ tt_taps tt_bilinear_taps(float tx, float ty, int W, int H)
{
    const float fx = __builtin_floorf(tx), fy = __builtin_floorf(ty);
    const int x0 = (int)fx, y0 = (int)fy;
    tt_taps t;
    t.wx = tx - (float)x0;
    t.wy = ty - (float)y0;
    const int xa = tt_wrap_near(x0, W), xb = tt_wrap_near(x0 + 1, W);
    const int ra = tt_wrap_near(y0, H) * W, rb = tt_wrap_near(y0 + 1, H) * W;
    t.a = ra + xa;
    t.b = ra + xb;
    t.c = rb + xa;
    t.d = rb + xb;
    return t;
}

// As tt_bilinear_taps; when no sample of the wavefront has its 2x2 footprint
// across the grid edge (the common case), the indices are formed without the
// wrap, which gives the same indices there.
__device__ __forceinline__ tt_taps tt_bilinear_taps_nowrap(float tx, float ty, int W, int H)
{
    const float fx = __builtin_floorf(tx), fy = __builtin_floorf(ty);
    const int x0 = (int)fx, y0 = (int)fy;
    const bool inside = (unsigned)x0 < (unsigned)(W - 1) && (unsigned)y0 < (unsigned)(H - 1);
    if (__builtin_amdgcn_ballot_w64(!inside) == 0) {
        tt_taps t;
        t.wx = tx - (float)x0;
        t.wy = ty - (float)y0;
        t.a = (int)(__umul24((uint32_t)y0, (uint32_t)W) + (uint32_t)x0);   // 0 <= y0, x0 < 2^24
        t.b = t.a + 1;
        t.c = t.a + W;
        t.d = t.c + 1;
        return t;
    }
    return tt_bilinear_taps(tx, ty, W, H);
}

__device__ __forceinline__ float tt_bilinear_mix(float s00, float s10, float s01, float s11, float wx, float wy)
{
    const float I0 = (1.0f - wx) * s00 + wx * s10;
    const float I1 = (1.0f - wx) * s01 + wx * s11;
    return (1.0f - wy) * I0 + wy * I1;
}

__device__ __forceinline__ float tt_bilinear(const float *__restrict__ s, float tx, float ty, int W, int H)
{
    const tt_taps t = tt_bilinear_taps(tx, ty, W, H);
    return tt_bilinear_mix(s[t.a], s[t.b], s[t.c], s[t.d], t.wx, t.wy);
}

// tt_sinf for x >= 0.03 (finite or not): when every lane of the wavefront has
// x < 0.78, the reduction's quadrant count rint(x * 2/pi) is 0 for every lane
// (0.78 * 2/pi < 0.497, far from the rounding boundary 0.5), so the reduced
// argument is x itself (x - 0 * pi/2 pieces) and tt_sinf is the sine
// polynomial of x; the other wavefronts (NaN included) take tt_sinf.  Same
// bits either way; the float compare replaces three fp64 operations per cell.
__device__ __forceinline__ float tt_sinf_wave(float x)
{
    if (__builtin_amdgcn_ballot_w64(!(x < 0.78f)) == 0) return (float)tt_sin_poly((double)x);
    return tt_sinf(x);
}

// evaporate (plate_tectonics_kernel.cu:2587-2593) with dt = 1
__device__ __forceinline__ float tt_evaporate(float w, float evap)
{
    return tt_fmaxf(0.0f, w - (tt_fminf(w, 20.0f) * evap) * 1.0f);
}

// rain amount of a cell: max(0, discreteNoise((x, 0, y), 1, seed + iter)) * dt * rain.
// For a cell (0 <= x, y < 2^24) the lattice point is the cell itself, and the
// lattice sum ((x*1723 + 0*93241) + y*149812) + 3824 + seed of tt_random_grid
// loses its middle term: adding +0 to the non-negative x*1723 changes nothing.
// tt_rain_from takes the sum's first two terms, a = (float)x * 1723 +
// (float)y * 149812, and rseedf = (float)(seed + iter).  The noise value
// tt_random_float(u) * 2 - 1 is formed as one fused multiply-add by 2^-31:
// the scaling by 2^-32 and the doubling are exact, so both round the same
// value once.  HW: the saturating conversion is the hardware's (checked).
template <bool HW>
__device__ __forceinline__ float tt_rain_from(float a, float rseedf, float rain)
{
    const float s = (a + 3824.0f) + rseedf;
    const uint32_t u = HW ? tt_f2u32_hw(s) : tt_f2u32_sat(s);
    const float n = __builtin_fmaf((float)tt_hash(u), 0x1p-31f, -1.0f);
    return (tt_fmaxf(0.0f, n) * 1.0f) * rain;
}

// flow velocity of a cell from the fluxes of the cell and its four
// neighbours and its water after the flow (flow, plate_tectonics_kernel.cu:
// 2487-2533): netFlow / max(w, 0.01), clamped to length 50
__device__ __forceinline__ void tt_flow_velocity(const float4 &Sf, const float4 &Lf, const float4 &Uf, const float4 &Rf,
                                                 const float4 &Df, float wn, float &vx, float &vy, float &vmag,
                                                 int sqrtOk)
{
    const float nx = (((Lf.z - Sf.x) + Sf.z) - Rf.x) * 0.5f;
    const float ny = (((Uf.w - Sf.y) + Sf.y) - Df.w) * 0.5f;
    const float hh = tt_fmaxf(wn, 0.01f);
    vx = nx / hh;
    vy = ny / hh;
    const float vm = tt_sqrt_wave(vx * vx + vy * vy, sqrtOk);
    vmag = vm;
    if (vm > 50.0f) {
        const float sc = 50.0f / vm;
        vx = vx * sc;
        vy = vy * sc;
        vmag = sqrtf(vx * vx + vy * vy);
    }
}

// The velocity alone, as tt_flow_velocity computes it: a squared magnitude of
// at most 2500 has a root of at most 50 (the root is correctly rounded and
// monotone), so no clamp applies and the root is not needed; a wavefront with
// a larger (or NaN) one computes it.
__device__ __forceinline__ void tt_flow_velocity_dir(const float4 &Sf, const float4 &Lf, const float4 &Uf,
                                                     const float4 &Rf, const float4 &Df, float wn, float &vx,
                                                     float &vy)
{
    const float nx = (((Lf.z - Sf.x) + Sf.z) - Rf.x) * 0.5f;
    const float ny = (((Uf.w - Sf.y) + Sf.y) - Df.w) * 0.5f;
    const float hh = tt_fmaxf(wn, 0.01f);
    vx = nx / hh;
    vy = ny / hh;
    const float s2 = vx * vx + vy * vy;
    if (__builtin_amdgcn_ballot_w64(!(s2 <= 2500.0f)) == 0) return;
    const float vm = sqrtf(s2);
    if (vm > 50.0f) {
        const float sc = 50.0f / vm;
        vx = vx * sc;
        vy = vy * sc;
    }
}

// Wavefronts per hydraulic block: two own rows per wavefront at every tile
// height.  A thread then needs at most 64 VGPRs, so a SIMD holds eight
// wavefronts (a 16-row block has 29 KB of LDS: four 512-thread blocks fit a
// CU); with four own rows per wavefront the 16-row block needed 85 VGPRs and a
// SIMD held five.
template <int TH> constexpr int tt_hyd_waves() { return TH / 2; }

// One sub-iteration on one tile.  FIRST: the water input is the step's
// state water and the sediment is the step's state; otherwise the water input
// is the previous sub-iteration's water after its flow (before evaporation),
// which this one evaporates while loading, and the sediment is the previous
// sub-iteration's output transported along its velocity.  Every sub-iteration
// adds its rain while loading.  LAST: the water output is the state water
// after evaporation; otherwise it is the water after the flow.
//
// Between sub-iterations only that flow water is stored (4 bytes a cell; the
// evaporated water and the velocity are functions of it): the next
// sub-iteration evaporates it, and recomputes the velocity, bit for bit, from
// it and the fluxes it reads anyway (its incoming flux ring).  Only the last
// sub-iteration writes the velocity field, for the closing transport and the
// state.  Loads are issued together before the first barrier (heights and
// water with their halo, the flux ring, the own flow water); the sediment taps
// follow the velocity and are in flight while the new fluxes are computed.
//
// The kernel sits at the balance of its vector ALU work and its HBM traffic,
// so addressing is kept off the vector units where it can be: the wavefront
// index is read into a scalar register (row arithmetic runs on the scalar
// unit), and IN (a tile whose two-cell halo does not cross the grid edge)
// drops the wrap of every halo and flux-ring index.
template <int TH, bool FIRST, bool LAST, bool COMP = false, bool IN = false, int NW = 4>
__device__ __forceinline__ void tt_hydraulic_tile(const tt_hyd_args &A, tt_hyd_lds<TH> &S, int x0, int y0)
{
    static_assert(TH % NW == 0, "each wavefront owns whole rows of the tile");
    constexpr int NT = 64 * NW;                  // threads of the block
    constexpr int RPW = TH / NW;                 // own rows per wavefront
    constexpr int NFI = (TH + 2) * TT_FW;        // flux items (tile plus a one-cell halo)
    constexpr int NFR = (NFI + NT - 1) / NT;     // flux items per thread
    constexpr int NHR = (TH + 4 + NW - 1) / NW;  // halo rows per wavefront (columns -2..61)
    constexpr int NXE = 4 * (TH + 4);            // the halo's last 4 columns (62..65), packed
    static_assert(NXE <= NT, "one extra-column element per thread");
    const int tid = threadIdx.x;
    const int W = A.W, H = A.H;
    const int lane = tid & 63, wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    const float rain = A.rain;
    const int rseed = A.rseed;
    const int gx = x0 + lane;

    // 1. the previous sub-iteration's flow water of the own cells
    float wnp[RPW];
    if (!FIRST) {
#pragma unroll
        for (int r = 0; r < RPW; ++r) wnp[r] = tt_at(A.win, (uint32_t)((y0 + wv * RPW + r) * W + gx));
    }
    // 2. heights and water on the tile plus a two-cell halo.  Columns
    // -2..61 of the halo, a row per wavefront pass (row addresses stay
    // scalar); the 4 columns 62..65 of all rows are packed over the first
    // lanes of the block, so that they cost two wavefront passes instead of a
    // partial pass per row.
    const int gxa = IN ? x0 + lane - 2 : tt_wrap1(x0 + lane - 2, W);
    float hm[NHR], hw[NHR];
#pragma unroll
    for (int j = 0; j < NHR; ++j) {
        const int ly = wv + NW * j;
        if (ly < TH + 4) {
            const int row = (IN ? y0 + ly - 2 : tt_wrap1(y0 + ly - 2, H)) * W;
            hm[j] = tt_at(A.hin, (uint32_t)(row + gxa));
            hw[j] = tt_at(A.win, (uint32_t)(row + gxa));
        }
    }
    const int xey = tid >> 2, xex = TT_EW - 4 + (tid & 3);
    const int xgy = IN ? y0 + xey - 2 : tt_wrap1(y0 + xey - 2, H);
    const int xgx = IN ? x0 + xex - 2 : tt_wrap1(x0 + xex - 2, W);
    float xm = 0.0f, xw = 0.0f;
    if (tid < NXE) {
        const uint32_t o = __umul24((uint32_t)xgy, (uint32_t)W) + (uint32_t)xgx;
        xm = tt_at(A.hin, o);
        xw = tt_at(A.win, o);
    }
    // 3. incoming fluxes of the flux ring.  Item i = tid + NT k sits at row
    // i / TT_FW, column i % TT_FW of the ring; NT = (NT / TT_FW) TT_FW +
    // NT % TT_FW, so the coordinates of item k follow from those of item
    // k - 1 by additions.
    int fyk[NFR], fxk[NFR];
    {
        int fy = tid / TT_FW, fx = tid - (tid / TT_FW) * TT_FW;
#pragma unroll
        for (int k = 0; k < NFR; ++k) {
            fyk[k] = fy;
            fxk[k] = fx;
            fx += NT % TT_FW;
            fy += NT / TT_FW;
            if (fx >= TT_FW) { fx -= TT_FW; fy += 1; }
        }
    }
    float4 cf[NFR];
#pragma unroll
    for (int k = 0; k < NFR; ++k) {
        const int i = tid + NT * k;
        if (i < NFI) {
            const int fy = fyk[k], fx = fxk[k];
            // rows and columns below 2^24: 24-bit products (full rate)
            cf[k] = tt_at(A.fin, IN ? __umul24((uint32_t)(y0 + fy - 1), (uint32_t)W) + (uint32_t)(x0 + fx - 1)
                                    : __umul24((uint32_t)tt_wrap1(y0 + fy - 1, H), (uint32_t)W) +
                                          (uint32_t)tt_wrap1(x0 + fx - 1, W));
        }
    }
    const float evap = A.evap;
    // sediment: the state itself on the first sub-iteration
    float s4[RPW][4], c4[RPW][4];
    float wxs[RPW], wys[RPW];
    if (FIRST) {
#pragma unroll
        for (int r = 0; r < RPW; ++r) {
            const int c = (y0 + wv * RPW + r) * W + gx;
            s4[r][0] = tt_at(A.sprev, (uint32_t)c);
            if (COMP) c4[r][0] = A.csprev[c];
        }
    }
    // the water this sub-iteration starts from: the state water plus rain
    // (first), or the previous flow water evaporated plus rain.  The row's
    // lattice term of the rain noise is wave-uniform and comes from the scalar
    // unit; the column's term is the lane's own, formed once.
    const float rseedf = (float)rseed, xterm = (float)gxa * 1723.0f;
    auto store_halo = [&](auto hw_cvt) {
        constexpr bool HW = decltype(hw_cvt)::value;
#pragma unroll
        for (int j = 0; j < NHR; ++j) {
            const int ly = wv + NW * j;
            if (ly < TH + 4) {
                const int gy = IN ? y0 + ly - 2 : tt_wrap1(y0 + ly - 2, H);
                float w = hw[j];
                if (!FIRST) w = tt_evaporate(w, evap);
                w = w + tt_rain_from<HW>(xterm + tt_lattice_z_term((uint32_t)gy), rseedf, rain);
                S.M[ly][lane] = hm[j];
                S.Wt[ly][lane] = w;
            }
        }
        if (tid < NXE) {
            float w = xw;
            if (!FIRST) w = tt_evaporate(w, evap);
            w = w + tt_rain_from<HW>((float)xgx * 1723.0f + (float)xgy * 149812.0f, rseedf, rain);
            S.M[xey][xex] = xm;
            S.Wt[xey][xex] = w;
        }
    };
    if (A.cvtOk) store_halo(std::true_type{});
    else store_halo(std::false_type{});
    if (!FIRST) {
        // the incoming fluxes, for the previous sub-iteration's velocities
#pragma unroll
        for (int k = 0; k < NFR; ++k) {
            const int i = tid + NT * k;
            if (i < NFI) S.F[fyk[k]][fxk[k]] = cf[k];
        }
    }
    __syncthreads();
    if (!FIRST) {
        // 4. the previous sub-iteration's velocity of the own cells (as it
        // computed it) and the sediment taps it addresses
#pragma unroll
        for (int r = 0; r < RPW; ++r) {
            const int ry = wv * RPW + r;
            const int gy = y0 + ry;
            const int fy = ry + 1, fx = lane + 1;
            float vx, vy;
            tt_flow_velocity_dir(S.F[fy][fx], S.F[fy][fx - 1], S.F[fy + 1][fx], S.F[fy][fx + 1], S.F[fy - 1][fx],
                                 wnp[r], vx, vy);
            const tt_taps t = tt_bilinear_taps_nowrap((float)gx - vx * 1.0f, (float)gy - vy * 1.0f, W, H);
            wxs[r] = t.wx;
            wys[r] = t.wy;
            s4[r][0] = tt_at(A.sprev, (uint32_t)t.a);
            s4[r][1] = tt_at(A.sprev, (uint32_t)t.b);
            s4[r][2] = tt_at(A.sprev, (uint32_t)t.c);
            s4[r][3] = tt_at(A.sprev, (uint32_t)t.d);
            if (COMP) {
                c4[r][0] = A.csprev[t.a];
                c4[r][1] = A.csprev[t.b];
                c4[r][2] = A.csprev[t.c];
                c4[r][3] = A.csprev[t.d];
            }
        }
        __syncthreads();
    }
    // 5. outflow flux on the tile plus a one-cell halo
    const float fk = A.fk, g = A.g, L = A.L;
#pragma unroll
    for (int k = 0; k < NFR; ++k) {
        const int i = tid + NT * k;
        if (i >= NFI) break;
        const int fy = fyk[k], fx = fxk[k];
        const int ly = fy + 1, lx = fx + 1;
        const float mc = S.M[ly][lx], wc = S.Wt[ly][lx];
        const float sc = mc + wc;
        const float n0 = g * ((sc - S.M[ly][lx - 1]) - S.Wt[ly][lx - 1]);
        const float n1 = g * ((sc - S.M[ly + 1][lx]) - S.Wt[ly + 1][lx]);
        const float n2 = g * ((sc - S.M[ly][lx + 1]) - S.Wt[ly][lx + 1]);
        const float n3 = g * ((sc - S.M[ly - 1][lx]) - S.Wt[ly - 1][lx]);
        float q0, q1, q2, q3;
        const int safe = tt_div_safe(n0) & tt_div_safe(n1) & tt_div_safe(n2) & tt_div_safe(n3);
        if (A.divL && __builtin_amdgcn_ballot_w64(!safe) == 0) {
            q0 = tt_div_const(n0, L, A.rL);
            q1 = tt_div_const(n1, L, A.rL);
            q2 = tt_div_const(n2, L, A.rL);
            q3 = tt_div_const(n3, L, A.rL);
        } else {
            q0 = n0 / L;
            q1 = n1 / L;
            q2 = n2 / L;
            q3 = n3 / L;
        }
        float4 o;
        o.x = tt_fmaxf(0.0f, cf[k].x + fk * q0);
        o.y = tt_fmaxf(0.0f, cf[k].y + fk * q1);
        o.z = tt_fmaxf(0.0f, cf[k].z + fk * q2);
        o.w = tt_fmaxf(0.0f, cf[k].w + fk * q3);
        const float tot = (((o.x + o.y) + o.z) + o.w) + 1e-6f;
        const float den = tot * 1.0f + 1e-6f;
        // K = max(min(1, wc / den), 0.01): wc >= den gives a quotient of at
        // least 1 (the division is correctly rounded), so K = 1 and the
        // scaling is the identity; a wavefront where that holds in every lane
        // skips the division and the products
        if (__builtin_amdgcn_ballot_w64(!(wc >= den)) != 0) {
            float K = tt_fminf(1.0f, wc / den);
            K = tt_fmaxf(K, 0.01f);
            o.x = o.x * K;
            o.y = o.y * K;
            o.z = o.z * K;
            o.w = o.w * K;
        }
        S.F[fy][fx] = o;
    }
    __syncthreads();
    const float cap = A.cap, ke = A.ke, kd = A.kd;
#pragma unroll
    for (int r = 0; r < RPW; ++r) {
        const int ry = wv * RPW + r;
        const int gy = y0 + ry;
        const int c = gy * W + gx;
        const int fy = ry + 1, fx = lane + 1;
        const float4 Sf = S.F[fy][fx], Lf = S.F[fy][fx - 1], Uf = S.F[fy + 1][fx], Rf = S.F[fy][fx + 1],
                     Df = S.F[fy - 1][fx];
        float in = 0.0f;
        in = in + Lf.z;
        in = in + Uf.w;
        in = in + Rf.x;
        in = in + Df.y;
        const float out = ((Sf.x + Sf.y) + Sf.z) + Sf.w;
        const float dv = 1.0f * (in - out);
        const int ly = ry + 2, lx = lane + 2;
        const float wn = tt_fmaxf(0.0f, S.Wt[ly][lx] + dv);
        float vx, vy, vmag;
        tt_flow_velocity(Sf, Lf, Uf, Rf, Df, wn, vx, vy, vmag, A.sqrtOk);
        __builtin_nontemporal_store(LAST ? tt_evaporate(wn, evap) : wn, tt_ptr(A.wout, (uint32_t)c));
        float4 *fo = tt_ptr(A.fout, (uint32_t)c);
        __builtin_nontemporal_store(Sf.x, &fo->x);
        __builtin_nontemporal_store(Sf.y, &fo->y);
        __builtin_nontemporal_store(Sf.z, &fo->z);
        __builtin_nontemporal_store(Sf.w, &fo->w);
        float sed, cs = 0.0f;
        if (FIRST) {
            sed = s4[r][0];
            if (COMP) cs = c4[r][0];
        } else {
            sed = tt_bilinear_mix(s4[r][0], s4[r][1], s4[r][2], s4[r][3], wxs[r], wys[r]);
            if (COMP) cs = tt_bilinear_mix(c4[r][0], c4[r][1], c4[r][2], c4[r][3], wxs[r], wys[r]);
        }
        if (LAST) {
            float2 *vo = tt_ptr(A.vel, (uint32_t)c);
            __builtin_nontemporal_store(vx, &vo->x);
            __builtin_nontemporal_store(vy, &vo->y);
        }
        const float mc = S.M[ly][lx];
        // Slope: (m[+1] - m[-1]) / 2.0 in the reference; halving is exact in float
        const float dzdx = (S.M[ly][lx + 1] - S.M[ly][lx - 1]) * 0.5f;
        const float dzdy = (S.M[ly + 1][lx] - S.M[ly - 1][lx]) * 0.5f;
        // sin(max(slope, 0.03)): a wavefront whose slopes all lie below 0.03
        // takes the constant; the others compute the root and the sine
        const float s2 = dzdx * dzdx + dzdy * dzdy;
        float sn = A.sin003;
        if (__builtin_amdgcn_ballot_w64(!(s2 <= TT_SLOPE2_SMALL)) != 0)
            sn = tt_sinf_wave(tt_fmaxf(tt_sqrt_wave(s2, A.sqrtOk), 0.03f));
        float C = (cap * sn) * vmag;
        C = tt_fminf(C, 5.0f);
        float hnew;
        float chv = 0.0f;
        if (COMP) chv = A.ch[c];
        if (C > sed) {
            float s = (ke * (C - sed)) * 1.0f;
            s = tt_fminf(s, mc);
            if (COMP) {
                hnew = tt_kahan_add(mc, chv, -s);
                sed = tt_kahan_add(sed, cs, s);
            } else {
                hnew = mc - s;
                sed = sed + s;
            }
        } else {
            float s = (kd * (sed - C)) * 1.0f;
            s = tt_fminf(s, sed);
            if (COMP) {
                sed = tt_kahan_add(sed, cs, -s);
                hnew = tt_kahan_add(mc, chv, s);
            } else {
                sed = sed - s;
                hnew = mc + s;
            }
        }
        if (COMP) {
            A.ch[c] = chv;
            A.csout[c] = cs;
        }
        __builtin_nontemporal_store(hnew, tt_ptr(A.hout, (uint32_t)c));
        __builtin_nontemporal_store(sed, tt_ptr(A.sout, (uint32_t)c));
    }
}

// closing transport of the step (the last sub-iteration's sediment along its velocity)
__device__ __forceinline__ void tt_transport_cell(int W, int H, uint32_t c, int x, int y, const float *sprev,
                                                  const float2 *vel, float *sout)
{
    const float2 v = vel[c];
    sout[c] = tt_bilinear(sprev, (float)x - v.x * 1.0f, (float)y - v.y * 1.0f, W, H);
}
