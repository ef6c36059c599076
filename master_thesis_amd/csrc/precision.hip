// precision.hip -- error of the float erosion against an fp64 run
// (tt_precision_study in tectonic.h).
//
// Three copies of one world take the same erosion-only steps (thermal and
// hydraulic stages, so every variant makes the same discrete decisions):
//   * the reference float arithmetic (TT_PRECISION_FP32),
//   * the compensated float arithmetic (TT_PRECISION_COMPENSATED),
//   * an fp64 twin: the same formulas in double, pass by pass in the
//     reference's order (rain, flux, flow, sediment, transport, evaporate;
//     plate_tectonic_sim.cu:331-373), Jacobi thermal passes before them.
// The report gives, per field (height, water, sediment), the maximum and RMS
// difference of each float variant from the fp64 twin, and the drift of the
// total of height + sediment over the run for all three.
#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "tt_kernels.h"
#include "tt_world.h"

void tt_set_error(const std::string &msg);

namespace {

struct d4 { double x, y, z, w; };
struct d2 { double x, y; };

struct P64 {
    int W, H, N;
    double talus, amp, str, rain, evap, cap, ke, kd, fk, g, L;
    double *h[2], *water, *sed[2], *tot;
    d4 *flux[2];
    d2 *vel;
};

__device__ __forceinline__ int wrapd(int v, int n)
{
    int r = v % n;
    return r + (n & (r >> 31));
}

__device__ __forceinline__ uint32_t cell(const P64 &P, int x, int y) { return (uint32_t)wrapd(y, P.H) * P.W + wrapd(x, P.W); }

#define GRID_LOOP(P)                                                              \
    for (uint32_t c = blockIdx.x * 256u + threadIdx.x; c < (uint32_t)(P).N;       \
         c += gridDim.x * 256u)

__constant__ int k64_off[8][2] = {{-1, 0}, {1, 0}, {0, -1}, {0, 1}, {-1, -1}, {-1, 1}, {1, -1}, {1, 1}};

__global__ void k64_from32(P64 P, const float *h, const float *w, const float4 *f, const float2 *v, const float *s)
{
    GRID_LOOP(P)
    {
        P.h[0][c] = h[c];
        P.water[c] = w[c];
        P.flux[0][c] = d4{f[c].x, f[c].y, f[c].z, f[c].w};
        P.vel[c] = d2{v[c].x, v[c].y};
        P.sed[0][c] = s[c];
    }
}

__global__ void k64_thermal_tot(P64 P, int cur)
{
    GRID_LOOP(P)
    {
        const int x = (int)(c % (uint32_t)P.W), y = (int)(c / (uint32_t)P.W);
        const double hc = P.h[cur][c];
        double t = 0.0;
        for (int k = 0; k < 8; ++k) {
            const double diff = hc - P.h[cur][cell(P, x + k64_off[k][0], y + k64_off[k][1])];
            if (diff > P.talus) t += diff;
        }
        P.tot[c] = t;
    }
}

__global__ void k64_thermal(P64 P, int cur)
{
    GRID_LOOP(P)
    {
        const int x = (int)(c % (uint32_t)P.W), y = (int)(c / (uint32_t)P.W);
        const double hc = P.h[cur][c], tc = P.tot[c];
        double v = hc;
        if (tc > 0.0)
            for (int k = 0; k < 8; ++k) {
                const double diff = hc - P.h[cur][cell(P, x + k64_off[k][0], y + k64_off[k][1])];
                if (diff > P.talus) v -= (P.amp * (diff / tc)) * P.str;
            }
        for (int k = 0; k < 8; ++k) {
            const uint32_t s = cell(P, x - k64_off[k][0], y - k64_off[k][1]);
            const double diff = P.h[cur][s] - hc;
            if (diff > P.talus && P.tot[s] > 0.0) v += (P.amp * (diff / P.tot[s])) * P.str;
        }
        P.h[cur ^ 1][c] = v;
    }
}

__global__ void k64_rain(P64 P, int rseed)
{
    GRID_LOOP(P)
    {
        const int x = (int)(c % (uint32_t)P.W), y = (int)(c / (uint32_t)P.W);
        const float n = tt_discrete_noise((float)x, 0.0f, (float)y, 1.0f, rseed);
        P.water[c] += (double)tt_fmaxf(0.0f, n) * P.rain;
    }
}

__global__ void k64_flux(P64 P, int hc, int fc)
{
    GRID_LOOP(P)
    {
        const int x = (int)(c % (uint32_t)P.W), y = (int)(c / (uint32_t)P.W);
        const double sc = P.h[hc][c] + P.water[c];
        auto sub = [&](double f, int nx, int ny) {
            const uint32_t n = cell(P, nx, ny);
            return std::max(0.0, f + P.fk * ((P.g * ((sc - P.h[hc][n]) - P.water[n])) / P.L));
        };
        const d4 cf = P.flux[fc][c];
        d4 o{sub(cf.x, x - 1, y), sub(cf.y, x, y + 1), sub(cf.z, x + 1, y), sub(cf.w, x, y - 1)};
        const double tot = (((o.x + o.y) + o.z) + o.w) + 1e-6;
        double K = std::min(1.0, P.water[c] / (tot + 1e-6));
        K = std::max(K, 0.01);
        P.flux[fc ^ 1][c] = d4{o.x * K, o.y * K, o.z * K, o.w * K};
    }
}

__global__ void k64_flow(P64 P, int fc)
{
    GRID_LOOP(P)
    {
        const int x = (int)(c % (uint32_t)P.W), y = (int)(c / (uint32_t)P.W);
        const d4 *F = P.flux[fc];
        const d4 S = F[c], Lf = F[cell(P, x - 1, y)], Uf = F[cell(P, x, y + 1)], Rf = F[cell(P, x + 1, y)],
                 Df = F[cell(P, x, y - 1)];
        const double in = ((Lf.z + Uf.w) + Rf.x) + Df.y;
        const double out = ((S.x + S.y) + S.z) + S.w;
        const double wn = std::max(0.0, P.water[c] + (in - out));
        P.water[c] = wn;
        const double nx = (((Lf.z - S.x) + S.z) - Rf.x) * 0.5;
        const double ny = (((Uf.w - S.y) + S.y) - Df.w) * 0.5;
        const double hh = std::max(wn, 0.01);
        double vx = nx / hh, vy = ny / hh;
        const double vm = sqrt(vx * vx + vy * vy);
        if (vm > 50.0) {
            vx *= 50.0 / vm;
            vy *= 50.0 / vm;
        }
        P.vel[c] = d2{vx, vy};
    }
}

__global__ void k64_sediment(P64 P, int hc, int sc)
{
    GRID_LOOP(P)
    {
        const int x = (int)(c % (uint32_t)P.W), y = (int)(c / (uint32_t)P.W);
        const double *M = P.h[hc];
        const double dzdx = (M[cell(P, x + 1, y)] - M[cell(P, x - 1, y)]) * 0.5;
        const double dzdy = (M[cell(P, x, y + 1)] - M[cell(P, x, y - 1)]) * 0.5;
        const double slope = sqrt(dzdx * dzdx + dzdy * dzdy);
        const d2 v = P.vel[c];
        double C = (P.cap * tt_sin_d(std::max(slope, 0.03))) * sqrt(v.x * v.x + v.y * v.y);
        C = std::min(C, 5.0);
        double sed = P.sed[sc][c], h = M[c];
        if (C > sed) {
            const double s = std::min(P.ke * (C - sed), h);
            h -= s;
            sed += s;
        } else {
            const double s = std::min(P.kd * (sed - C), sed);
            sed -= s;
            h += s;
        }
        P.h[hc ^ 1][c] = h;
        P.sed[sc][c] = sed;
    }
}

__global__ void k64_transport(P64 P, int sc)
{
    GRID_LOOP(P)
    {
        const int x = (int)(c % (uint32_t)P.W), y = (int)(c / (uint32_t)P.W);
        const d2 v = P.vel[c];
        const double tx = (double)x - v.x, ty = (double)y - v.y;
        const double fx = floor(tx), fy = floor(ty);
        const int x0 = (int)fx, y0 = (int)fy;
        const double wx = tx - fx, wy = ty - fy;
        const double *S = P.sed[sc];
        const double I0 = (1.0 - wx) * S[cell(P, x0, y0)] + wx * S[cell(P, x0 + 1, y0)];
        const double I1 = (1.0 - wx) * S[cell(P, x0, y0 + 1)] + wx * S[cell(P, x0 + 1, y0 + 1)];
        P.sed[sc ^ 1][c] = (1.0 - wy) * I0 + wy * I1;
    }
}

__global__ void k64_evaporate(P64 P)
{
    GRID_LOOP(P)
    {
        const double w = P.water[c];
        P.water[c] = std::max(0.0, w - std::min(w, 20.0) * P.evap);
    }
}

// per block: max |a - r|, sum (a - r)^2, sum r^2
__global__ void k64_compare(int n, const float *a, const double *r, double *out)
{
    __shared__ double sm[3][256];
    double mx = 0.0, s2 = 0.0, r2 = 0.0;
    for (int c = blockIdx.x * 256 + threadIdx.x; c < n; c += gridDim.x * 256) {
        const double d = (double)a[c] - r[c];
        mx = std::max(mx, fabs(d));
        s2 += d * d;
        r2 += r[c] * r[c];
    }
    sm[0][threadIdx.x] = mx;
    sm[1][threadIdx.x] = s2;
    sm[2][threadIdx.x] = r2;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            sm[0][threadIdx.x] = std::max(sm[0][threadIdx.x], sm[0][threadIdx.x + o]);
            sm[1][threadIdx.x] += sm[1][threadIdx.x + o];
            sm[2][threadIdx.x] += sm[2][threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[3 * blockIdx.x] = sm[0][0];
        out[3 * blockIdx.x + 1] = sm[1][0];
        out[3 * blockIdx.x + 2] = sm[2][0];
    }
}

// per block: sum of a + b (float or double inputs)
template <typename T>
__global__ void k64_total(int n, const T *a, const T *b, double *out)
{
    __shared__ double sm[256];
    double s = 0.0;
    for (int c = blockIdx.x * 256 + threadIdx.x; c < n; c += gridDim.x * 256) s += (double)a[c] + (double)b[c];
    sm[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sm[threadIdx.x] += sm[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = sm[0];
}

constexpr int kBlocks = 1024;

}  // namespace

// fp64 erosion step of the twin (T thermal passes, K hydraulic sub-iterations)
static hipError_t p64_step(hipStream_t s, P64 &P, int &hc, int &fc, int &sc, int T, int K, int rseed)
{
    for (int t = 0; t < T; ++t) {
        hipLaunchKernelGGL(k64_thermal_tot, dim3(kBlocks), dim3(256), 0, s, P, hc);
        hipLaunchKernelGGL(k64_thermal, dim3(kBlocks), dim3(256), 0, s, P, hc);
        hc ^= 1;
    }
    for (int k = 0; k < K; ++k) {
        hipLaunchKernelGGL(k64_rain, dim3(kBlocks), dim3(256), 0, s, P, rseed);
        hipLaunchKernelGGL(k64_flux, dim3(kBlocks), dim3(256), 0, s, P, hc, fc);
        fc ^= 1;
        hipLaunchKernelGGL(k64_flow, dim3(kBlocks), dim3(256), 0, s, P, fc);
        hipLaunchKernelGGL(k64_sediment, dim3(kBlocks), dim3(256), 0, s, P, hc, sc);
        hc ^= 1;
        hipLaunchKernelGGL(k64_transport, dim3(kBlocks), dim3(256), 0, s, P, sc);
        sc ^= 1;
        hipLaunchKernelGGL(k64_evaporate, dim3(kBlocks), dim3(256), 0, s, P);
    }
    return hipGetLastError();
}

#define PCK(x)                                                                              \
    do {                                                                                    \
        hipError_t e_ = (x);                                                                \
        if (e_ != hipSuccess) {                                                             \
            tt_set_error(std::string(#x) + ": " + hipGetErrorString(e_));                   \
            st = TT_ERR_DEVICE;                                                             \
            goto done;                                                                      \
        }                                                                                   \
    } while (0)

extern "C" int tt_precision_study(const tt_sim_config *cfg_in, const tt_kernel_settings *ks, int32_t steps,
                                  tt_precision_report *out)
{
    if (!cfg_in || !ks || !out || steps < 0) return TT_ERR_INVALID;
    tt_sim_config cfg = *cfg_in;
    cfg.stage_mask = TT_STAGE_THERMAL | TT_STAGE_HYDRAULIC;
    tt_sim_config cfgA = cfg, cfgB = cfg;
    cfgA.precision = TT_PRECISION_FP32;
    cfgB.precision = TT_PRECISION_COMPENSATED;
    tt_world *A = nullptr, *B = nullptr;
    int st = tt_world_create(&cfgA, ks, &A);
    if (st == TT_OK) st = tt_world_init(A);
    if (st == TT_OK) st = tt_world_create(&cfgB, ks, &B);
    if (st == TT_OK) st = tt_world_init(B);
    if (st != TT_OK) {
        if (A) tt_world_destroy(A);
        if (B) tt_world_destroy(B);
        return st;
    }
    const size_t N = (size_t)A->N;
    const int T = cfg.thermal_iterations, K = cfg.hydraulic_iterations;
    P64 P;
    memset(&P, 0, sizeof P);
    P.W = A->W;
    P.H = A->H;
    P.N = A->N;
    P.talus = (double)(ks->thermalThresholdAngle * ks->thermalCellSize);
    P.amp = ks->thermalErosionAmplitude;
    P.str = ks->thermalErosionStrength;
    P.rain = ks->hydrationRainfall;
    P.evap = ks->hydrationEvaporation;
    P.cap = ks->sedimentCapacity;
    P.ke = ks->sedimentErosionRate;
    P.kd = ks->sedimentDepositionRate;
    P.fk = ks->hydrationPipeCrossSection;
    P.g = ks->gravity;
    P.L = ks->hydrationPipeLength;
    hipStream_t s = A->stream;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    double *part = nullptr;
    std::vector<double> hp((size_t)kBlocks * 3);
    int hc = 0, fc = 0, sc = 0;
    double msA = 0, msB = 0, msR = 0;
    tt_precision_report R;
    memset(&R, 0, sizeof R);
    auto totals = [&](double *res3) -> hipError_t {
        hipLaunchKernelGGL(k64_total<double>, dim3(kBlocks), dim3(256), 0, s, P.N, P.h[hc], P.sed[sc], part);
        hipLaunchKernelGGL(k64_total<float>, dim3(kBlocks), dim3(256), 0, s, P.N, A->h[A->hc], A->sed[A->sc],
                           part + kBlocks);
        hipLaunchKernelGGL(k64_total<float>, dim3(kBlocks), dim3(256), 0, s, P.N, B->h[B->hc], B->sed[B->sc],
                           part + 2 * kBlocks);
        hipError_t e = hipMemcpyAsync(hp.data(), part, hp.size() * 8, hipMemcpyDeviceToHost, s);
        if (e != hipSuccess) return e;
        e = hipStreamSynchronize(s);
        for (int v = 0; v < 3; ++v) {
            res3[v] = 0.0;
            for (int i = 0; i < kBlocks; ++i) res3[v] += hp[(size_t)v * kBlocks + i];
        }
        return e;
    };
    auto compare = [&](const float *a, const double *r, double *mx, double *rms, double *rref) -> hipError_t {
        hipLaunchKernelGGL(k64_compare, dim3(kBlocks), dim3(256), 0, s, P.N, a, r, part);
        hipError_t e = hipMemcpyAsync(hp.data(), part, hp.size() * 8, hipMemcpyDeviceToHost, s);
        if (e != hipSuccess) return e;
        e = hipStreamSynchronize(s);
        double m = 0, s2 = 0, r2 = 0;
        for (int i = 0; i < kBlocks; ++i) {
            m = std::max(m, hp[3 * i]);
            s2 += hp[3 * i + 1];
            r2 += hp[3 * i + 2];
        }
        *mx = m;
        *rms = sqrt(s2 / (double)P.N);
        if (rref) *rref = sqrt(r2 / (double)P.N);
        return e;
    };
    double t0[3], t1[3];
    {
        PCK(hipMalloc((void **)&P.h[0], N * 8));
        PCK(hipMalloc((void **)&P.h[1], N * 8));
        PCK(hipMalloc((void **)&P.water, N * 8));
        PCK(hipMalloc((void **)&P.sed[0], N * 8));
        PCK(hipMalloc((void **)&P.sed[1], N * 8));
        PCK(hipMalloc((void **)&P.tot, N * 8));
        PCK(hipMalloc((void **)&P.flux[0], N * 32));
        PCK(hipMalloc((void **)&P.flux[1], N * 32));
        PCK(hipMalloc((void **)&P.vel, N * 16));
        PCK(hipMalloc((void **)&part, hp.size() * 8));
        PCK(hipEventCreate(&e0));
        PCK(hipEventCreate(&e1));
        hipLaunchKernelGGL(k64_from32, dim3(kBlocks), dim3(256), 0, s, P, A->h[A->hc], A->water[A->wc],
                           A->flux[A->fc], A->vel, A->sed[A->sc]);
        PCK(hipGetLastError());
        PCK(totals(t0));
        for (int k = 0; k < steps; ++k) {
            const int rseed = (int)(cfg.seed + A->iter);
            float ms = 0.0f;
            PCK(hipEventRecord(e0, s));
            PCK(p64_step(s, P, hc, fc, sc, T, K, rseed));
            PCK(hipEventRecord(e1, s));
            PCK(hipEventSynchronize(e1));
            PCK(hipEventElapsedTime(&ms, e0, e1));
            msR += ms;
            PCK(hipEventRecord(e0, A->stream));
            st = tt_world_step(A, 1);
            if (st != TT_OK) goto done;
            PCK(hipEventRecord(e1, A->stream));
            PCK(hipEventSynchronize(e1));
            PCK(hipEventElapsedTime(&ms, e0, e1));
            msA += ms;
            PCK(hipEventRecord(e0, B->stream));
            st = tt_world_step(B, 1);
            if (st != TT_OK) goto done;
            PCK(hipEventRecord(e1, B->stream));
            PCK(hipEventSynchronize(e1));
            PCK(hipEventElapsedTime(&ms, e0, e1));
            msB += ms;
        }
        PCK(compare(A->h[A->hc], P.h[hc], &R.height.max_abs_fp32, &R.height.rms_fp32, &R.height.rms_ref));
        PCK(compare(B->h[B->hc], P.h[hc], &R.height.max_abs_comp, &R.height.rms_comp, nullptr));
        PCK(compare(A->water[A->wc], P.water, &R.water.max_abs_fp32, &R.water.rms_fp32, &R.water.rms_ref));
        PCK(compare(B->water[B->wc], P.water, &R.water.max_abs_comp, &R.water.rms_comp, nullptr));
        PCK(compare(A->sed[A->sc], P.sed[sc], &R.sediment.max_abs_fp32, &R.sediment.rms_fp32, &R.sediment.rms_ref));
        PCK(compare(B->sed[B->sc], P.sed[sc], &R.sediment.max_abs_comp, &R.sediment.rms_comp, nullptr));
        PCK(totals(t1));
        R.steps = steps;
        R.width = P.W;
        R.height_cells = P.H;
        R.mass_start = t0[0];
        R.mass_ref = t1[0];
        R.mass_fp32 = t1[1];
        R.mass_comp = t1[2];
        R.drift_ref = (t1[0] - t0[0]) / t0[0];
        R.drift_fp32 = (t1[1] - t0[1]) / t0[1];
        R.drift_comp = (t1[2] - t0[2]) / t0[2];
        R.ms_fp32 = steps ? msA / steps : 0.0;
        R.ms_comp = steps ? msB / steps : 0.0;
        R.ms_ref = steps ? msR / steps : 0.0;
        *out = R;
        st = TT_OK;
    }
done:
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    void *bufs[] = {P.h[0], P.h[1], P.water, P.sed[0], P.sed[1], P.tot, P.flux[0], P.flux[1], P.vel, part};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    tt_world_destroy(A);
    tt_world_destroy(B);
    return st;
}
