// k_erosion.hip -- stage D (thermal erosion) and stage E (hydraulic erosion)
// on gfx950, LDS-tiled stencils over 64x16 tiles (one wavefront per row).
//
// Thermal: thermalErosionKernel (plate_tectonics_kernel.cu:2735-2786) in the
// canonical Jacobi form -- each pass reads the previous heights; a cell's new
// height is its old height minus its own outflows (directions 0..7) plus the
// inflows it receives (directions 0..7).  The tile holds a two-cell halo.
//
// Hydraulic: one kernel per sub-iteration fuses rain (:2430-2437), flux
// (:2439-2485), flow (:2487-2533), evaporate (:2587-2593) and the
// erosion/deposition exchange (:2535-2570 with CudaTexture::Slope), and begins
// with the semi-Lagrangian transport (:2572-2585, cuda_helper.cuh:49-65) of the
// previous sub-iteration, gathered from its sediment and velocity.  The order of
// float operations is the reference's; dt = 1 (plate_tectonic_sim.cu:338-363).
#include "tt_kernels.h"

static_assert(TT_TILE_W == 64 && TT_TILE_H == 16, "tiles are 64 x 16");

#define TH_W (TT_TILE_W + 4)
#define TH_H (TT_TILE_H + 4)
#define TT_W (TT_TILE_W + 2)
#define TT_H (TT_TILE_H + 2)

__constant__ int c_toff[8][2] = {{-1, 0}, {1, 0}, {0, -1}, {0, 1}, {-1, -1}, {-1, 1}, {1, -1}, {1, 1}};

__global__ __launch_bounds__(256) void k_thermal(tt_dev d, const float *__restrict__ hin, float *__restrict__ hout)
{
    __shared__ float hs[TH_H][TH_W];
    __shared__ float tot[TT_H][TT_W];
    const int tid = threadIdx.x;
    const int W = d.W, H = d.H;
    const int nbx = W / TT_TILE_W, nby = H / TT_TILE_H;
    const int b = tt_xcd_tile(blockIdx.x, nbx * nby);
    const int x0 = (b % nbx) * TT_TILE_W, y0 = (b / nbx) * TT_TILE_H;
    const int lane = tid & 63, wv = tid >> 6;
    __shared__ float rmin[4], rmax[4];
    float mn = tt_u2f(0x7f800000u), mx = tt_u2f(0xff800000u);
    {
        const int gxa = tt_wrap1(x0 + lane - 2, W), gxb = tt_wrap1(x0 + 64 + lane - 2, W);
        for (int ly = wv; ly < TH_H; ly += 4) {
            const int row = tt_wrap1(y0 + ly - 2, H) * W;
            const float a = hin[row + gxa];
            hs[ly][lane] = a;
            mn = tt_fminf(mn, a);
            mx = tt_fmaxf(mx, a);
            if (lane < TH_W - 64) {
                const float b = hin[row + gxb];
                hs[ly][64 + lane] = b;
                mn = tt_fminf(mn, b);
                mx = tt_fmaxf(mx, b);
            }
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        mn = tt_fminf(mn, __shfl_xor(mn, o, 64));
        mx = tt_fmaxf(mx, __shfl_xor(mx, o, 64));
    }
    if (lane == 0) { rmin[wv] = mn; rmax[wv] = mx; }
    __syncthreads();
    const float talus = d.ks.thermalThresholdAngle * d.ks.thermalCellSize;
    const float amp = d.ks.thermalErosionAmplitude, str = d.ks.thermalErosionStrength;
    // No transfer anywhere in the tile when the tile plus its two-cell halo
    // spans at most the talus: every difference h_a - h_b rounds to at most
    // max - min (rounding is monotone), so no "diff > talus" test can pass and
    // every cell keeps its height exactly.
    {
        const float tmn = tt_fminf(tt_fminf(rmin[0], rmin[1]), tt_fminf(rmin[2], rmin[3]));
        const float tmx = tt_fmaxf(tt_fmaxf(rmax[0], rmax[1]), tt_fmaxf(rmax[2], rmax[3]));
        if (tmx - tmn <= talus) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int ry = wv * 4 + r;
                hout[(y0 + ry) * W + x0 + lane] = hs[ry + 2][lane + 2];
            }
            return;
        }
    }
    // sum of the downhill differences above the talus, on the tile + 1 halo
    for (int i = tid; i < TT_H * TT_W; i += 256) {
        const int ty = i / TT_W, tx = i - ty * TT_W;
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
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int ry = wv * 4 + r;
        const int ly = ry + 2, lx = lane + 2;
        const float hc = hs[ly][lx];
        const float tc = tot[ry + 1][lane + 1];
        float dOut[8], dIn[8], tIn[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            dOut[k] = hc - hs[ly + c_toff[k][1]][lx + c_toff[k][0]];
            const int sy = ly - c_toff[k][1], sx = lx - c_toff[k][0];
            dIn[k] = hs[sy][sx] - hc;
            tIn[k] = tot[sy - 1][sx - 1];
        }
        // own outflows (directions 0..7), then inflows (directions 0..7); the
        // divisions sit in branches that a wavefront skips when no lane needs them
        float v = hc;
        if (tc > 0.0f) {
#pragma unroll
            for (int k = 0; k < 8; ++k)
                if (dOut[k] > talus) v = v - (amp * (dOut[k] / tc)) * str;
        }
#pragma unroll
        for (int k = 0; k < 8; ++k)
            if (dIn[k] > talus && tIn[k] > 0.0f) v = v + (amp * (dIn[k] / tIn[k])) * str;
        hout[(y0 + ry) * W + x0 + lane] = v;
    }
}

hipError_t tt_launch_thermal(hipStream_t s, const tt_dev &d, const float *hin, float *hout)
{
    const unsigned blocks = (unsigned)((d.W / TT_TILE_W) * (d.H / TT_TILE_H));
    hipLaunchKernelGGL(k_thermal, dim3(blocks), dim3(256), 0, s, d, hin, hout);
    return hipGetLastError();
}

// bilinear sample with toroidal wrap (interpolate, cuda_helper.cuh:49-65).
// |v| <= 50 keeps the sample column within [-51, W + 50], where a single
// conditional wrap is exact; other columns (non-finite velocities) take the
// general modulo.  IEEE floor is exact, so the hardware floor equals the
// oracle's bit-level floor.
__device__ __forceinline__ int tt_wrap_near(int v, int n)
{
    return ((unsigned)(v + n) < (unsigned)(3 * n)) ? tt_wrap1(v, n) : tt_wrapg(v, n);
}

__device__ __forceinline__ float tt_bilinear(const float *__restrict__ s, float tx, float ty, int W, int H)
{
    const float fx = __builtin_floorf(tx), fy = __builtin_floorf(ty);
    const int x0 = (int)fx, y0 = (int)fy;
    const float wx = tx - (float)x0, wy = ty - (float)y0;
    const int xa = tt_wrap_near(x0, W), xb = tt_wrap_near(x0 + 1, W);
    const int ra = tt_wrap_near(y0, H) * W, rb = tt_wrap_near(y0 + 1, H) * W;
    const float I0 = (1.0f - wx) * s[ra + xa] + wx * s[ra + xb];
    const float I1 = (1.0f - wx) * s[rb + xa] + wx * s[rb + xb];
    return (1.0f - wy) * I0 + wy * I1;
}

// rain amount of a cell: max(0, discreteNoise((x, 0, y), 1, seed + iter)) * dt * rain
__device__ __forceinline__ float tt_rain_at(int gx, int gy, int rseed, float rain)
{
    const float n = tt_discrete_noise((float)gx, 0.0f, (float)gy, 1.0f, rseed);
    return (tt_fmaxf(0.0f, n) * 1.0f) * rain;
}

// One sub-iteration.  FIRST: the water input has no rain yet (rain is added
// while loading) and the sediment is the step's state; otherwise the water
// input already holds this sub-iteration's rain (added by the previous one on
// its final write) and the sediment is the previous sub-iteration's output
// transported along its velocity.  LAST: the water output gets no rain.
template <bool FIRST, bool LAST>
__global__ __launch_bounds__(256) void k_hydraulic(tt_dev d, int rseed, const float *__restrict__ hin,
                                                   float *__restrict__ hout, const float *__restrict__ win,
                                                   float *__restrict__ wout, const float4 *__restrict__ fin,
                                                   float4 *__restrict__ fout, float2 *__restrict__ vel,
                                                   const float *__restrict__ sprev, float *__restrict__ sout)
{
    __shared__ float M[TH_H][TH_W];
    __shared__ float Wt[TH_H][TH_W];
    __shared__ float4 F[TT_H][TT_W];
    const int tid = threadIdx.x;
    const int W = d.W, H = d.H;
    const int nbx = W / TT_TILE_W, nby = H / TT_TILE_H;
    const int b = tt_xcd_tile(blockIdx.x, nbx * nby);
    const int x0 = (b % nbx) * TT_TILE_W, y0 = (b / nbx) * TT_TILE_H;
    const float rain = d.ks.hydrationRainfall;
    const int lane = tid & 63, wv = tid >> 6;
    // heights and water on the tile plus a two-cell halo: rows wv, wv+4, ...;
    // lanes 0..63 take columns 0..63, lanes 0..3 also columns 64..67
    {
        const int gxa = tt_wrap1(x0 + lane - 2, W);
        const int gxb = tt_wrap1(x0 + 64 + lane - 2, W);
        for (int ly = wv; ly < TH_H; ly += 4) {
            const int gy = tt_wrap1(y0 + ly - 2, H);
            const int row = gy * W;
            float m = hin[row + gxa], w = win[row + gxa];
            if (FIRST) w = w + tt_rain_at(gxa, gy, rseed, rain);
            M[ly][lane] = m;
            Wt[ly][lane] = w;
            if (lane < TH_W - 64) {
                float m2 = hin[row + gxb], w2 = win[row + gxb];
                if (FIRST) w2 = w2 + tt_rain_at(gxb, gy, rseed, rain);
                M[ly][64 + lane] = m2;
                Wt[ly][64 + lane] = w2;
            }
        }
    }
    __syncthreads();
    // outflow flux on the tile plus a one-cell halo
    const float fk = 1.0f * d.ks.hydrationPipeCrossSection;
    const float g = d.ks.gravity, L = d.ks.hydrationPipeLength;
    for (int i = tid; i < TT_H * TT_W; i += 256) {
        const int fy = i / TT_W, fx = i - fy * TT_W;
        const int ly = fy + 1, lx = fx + 1;
        const int gx = tt_wrap1(x0 + fx - 1, W), gy = tt_wrap1(y0 + fy - 1, H);
        const float4 cf = fin[gy * W + gx];
        const float mc = M[ly][lx], wc = Wt[ly][lx];
        const float sc = mc + wc;
        float4 o;
        o.x = tt_fmaxf(0.0f, cf.x + fk * ((g * ((sc - M[ly][lx - 1]) - Wt[ly][lx - 1])) / L));
        o.y = tt_fmaxf(0.0f, cf.y + fk * ((g * ((sc - M[ly + 1][lx]) - Wt[ly + 1][lx])) / L));
        o.z = tt_fmaxf(0.0f, cf.z + fk * ((g * ((sc - M[ly][lx + 1]) - Wt[ly][lx + 1])) / L));
        o.w = tt_fmaxf(0.0f, cf.w + fk * ((g * ((sc - M[ly - 1][lx]) - Wt[ly - 1][lx])) / L));
        const float tot = (((o.x + o.y) + o.z) + o.w) + 1e-6f;
        float K = tt_fminf(1.0f, wc / (tot * 1.0f + 1e-6f));
        K = tt_fmaxf(K, 0.01f);
        o.x = o.x * K;
        o.y = o.y * K;
        o.z = o.z * K;
        o.w = o.w * K;
        F[fy][fx] = o;
    }
    __syncthreads();
    const float evap = d.ks.hydrationEvaporation;
    const float cap = d.ks.sedimentCapacity, ke = d.ks.sedimentErosionRate, kd = d.ks.sedimentDepositionRate;
    const int gx = x0 + lane;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int ry = wv * 4 + r;
        const int gy = y0 + ry;
        const int c = gy * W + gx;
        const int fy = ry + 1, fx = lane + 1;
        const float4 S = F[fy][fx], Lf = F[fy][fx - 1], Uf = F[fy + 1][fx], Rf = F[fy][fx + 1], Df = F[fy - 1][fx];
        float in = 0.0f;
        in = in + Lf.z;
        in = in + Uf.w;
        in = in + Rf.x;
        in = in + Df.y;
        const float out = ((S.x + S.y) + S.z) + S.w;
        const float dv = 1.0f * (in - out);
        const int ly = ry + 2, lx = lane + 2;
        const float wn = tt_fmaxf(0.0f, Wt[ly][lx] + dv);
        const float nx = (((Lf.z - S.x) + S.z) - Rf.x) * 0.5f;
        const float ny = (((Uf.w - S.y) + S.y) - Df.w) * 0.5f;
        const float hh = tt_fmaxf(wn, 0.01f);
        float vx = nx / hh, vy = ny / hh;
        const float vm = sqrtf(vx * vx + vy * vy);
        float vmag = vm;
        if (vm > 50.0f) {
            const float sc = 50.0f / vm;
            vx = vx * sc;
            vy = vy * sc;
            vmag = sqrtf(vx * vx + vy * vy);
        }
        float wo = tt_fmaxf(0.0f, wn - (tt_fminf(wn, 20.0f) * evap) * 1.0f);
        if (!LAST) wo = wo + tt_rain_at(gx, gy, rseed, rain);
        wout[c] = wo;
        fout[c] = S;
        float sed;
        if (FIRST) {
            sed = sprev[c];
        } else {
            const float2 pv = vel[c];
            sed = tt_bilinear(sprev, (float)gx - pv.x * 1.0f, (float)gy - pv.y * 1.0f, W, H);
        }
        vel[c] = make_float2(vx, vy);
        const float mc = M[ly][lx];
        // Slope: (m[+1] - m[-1]) / 2.0 in the reference; halving is exact in float
        const float dzdx = (M[ly][lx + 1] - M[ly][lx - 1]) * 0.5f;
        const float dzdy = (M[ly + 1][lx] - M[ly - 1][lx]) * 0.5f;
        const float slope = sqrtf(dzdx * dzdx + dzdy * dzdy);
        float C = (cap * tt_sinf(tt_fmaxf(slope, 0.03f))) * vmag;
        C = tt_fminf(C, 5.0f);
        float hnew;
        if (C > sed) {
            float s = (ke * (C - sed)) * 1.0f;
            s = tt_fminf(s, mc);
            hnew = mc - s;
            sed = sed + s;
        } else {
            float s = (kd * (sed - C)) * 1.0f;
            s = tt_fminf(s, sed);
            sed = sed - s;
            hnew = mc + s;
        }
        hout[c] = hnew;
        sout[c] = sed;
    }
}

hipError_t tt_launch_hydraulic(hipStream_t s, const tt_dev &d, int first, int last, int rseed, const float *hin,
                               float *hout, const float *win, float *wout, const float4 *fin, float4 *fout,
                               float2 *vel, const float *sprev, float *sout)
{
    const unsigned blocks = (unsigned)((d.W / TT_TILE_W) * (d.H / TT_TILE_H));
#define TT_HYD(F_, L_) \
    hipLaunchKernelGGL((k_hydraulic<F_, L_>), dim3(blocks), dim3(256), 0, s, d, rseed, hin, hout, win, wout, fin, \
                       fout, vel, sprev, sout)
    if (first && last) TT_HYD(true, true);
    else if (first) TT_HYD(true, false);
    else if (last) TT_HYD(false, true);
    else TT_HYD(false, false);
#undef TT_HYD
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_transport(tt_dev d, const float *sprev, const float2 *vel, float *sout)
{
    const long long c = (long long)blockIdx.x * 256 + threadIdx.x;
    if (c >= d.N) return;
    int x, y;
    tt_cell_xy((uint32_t)c, d, x, y);
    const float2 v = vel[c];
    sout[c] = tt_bilinear(sprev, (float)x - v.x * 1.0f, (float)y - v.y * 1.0f, d.W, d.H);
}

hipError_t tt_launch_transport(hipStream_t s, const tt_dev &d, const float *sprev, const float2 *vel, float *sout)
{
    const unsigned blocks = (unsigned)(((long long)d.N + 255) / 256);
    hipLaunchKernelGGL(k_transport, dim3(blocks), dim3(256), 0, s, d, sprev, vel, sout);
    return hipGetLastError();
}
