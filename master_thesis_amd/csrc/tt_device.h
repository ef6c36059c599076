// tt_device.h -- definitions shared by the gfx950 kernels of libtectonic:
// device-side layouts, the per-step plate table, index helpers and the launch
// wrappers' declarations.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/tectonic.h"
#include "tt_math.h"
#include "tt_noise.h"

#define TT_MAXP 255
#define TT_NUM_TRI (TT_MAXP * (TT_MAXP - 1) / 2)
#define TT_NUM_TRI_WORDS ((TT_NUM_TRI + 31) / 32)

// fixed-point scales of the exact per-plate sums
#define TT_MASS_BITS 20
#define TT_ASTH_BITS 32
#define TT_FRIC_BITS 48
#define TT_ANG_BITS 40

// 2-D tiles of the stencil kernels: 64 columns (one wavefront per row) by 16
// rows, 256 threads per block (4 wavefronts, each owning 4 rows).
#define TT_TILE_W 64
#define TT_TILE_H 16
#define TT_BLOCK 256

// rounds of a tiled fixed-point kernel enqueued per host check, and the ring of
// per-round change counters (two batches)
#define TT_ROUND_BATCH 32
#define TT_RCHG (2 * TT_ROUND_BATCH)

struct tt_uplift { float cum; uint8_t types; uint8_t pad[3]; };
struct tt_prop { float cum; uint32_t inv; uint8_t types; uint8_t pad[3]; };
struct tt_dist { int32_t dist; uint32_t orig; };

// Per-plate values derived once per step from the plate table.
struct tt_plate_step {
    int ox, oy;       // floor(pixelCenter + direction * velocity)
    float vel, mass;
    float dirx, diry;
    float pcx, pcy;
};

// Per-step control block in device memory (one per world).
struct tt_step_ctl {
    // advection
    int n_distinct;
    int max_r;
    int zero_distinct;   // some plate (0..254) has the offset (0, 0)
    int dox[256], doy[256];
    // outputs of the step read back by the host
    int overflow_cells;
    int complex_sources;
    int n_roots;
    int n_unassigned;
    int merges;
    int split_plate;
    int split_new;
    int changed;         // relaxation / flood-fill change counter
    int dead_end;
    int n_plate_cells;
    int largest_value, largest_id, heaviest_id, n_plates;
    float heaviest_value;
    int split_seed;      // cell where the split's new id starts, -1 if none
    int split_negative;  // some cell of the plate to split has a height below 0 (or NaN)
    long long ledger;
};

// Water buffers of hydraulic sub-iteration k of K.  Between sub-iterations
// only the water after the flow (before evaporation) is stored: the next
// sub-iteration evaporates it and adds its rain while loading, and recomputes
// the previous velocity from it.  The first sub-iteration reads the state
// water water[wc]; sub-iteration k < K - 1 writes its flow water to X
// (k even) or Y (k odd), with X = water[wc ^ 1] and Y = flow; the last writes
// the state water after evaporation into water[wc] (free once the first has
// read it), or into water[wc ^ 1] when it is also the first.  The state water
// index after the step is tt_hyd_water_index_after(wc, K).
__host__ __device__ inline void tt_hyd_water_bufs(float *w0, float *w1, float *flow, int wc, int k, int K,
                                                  const float **win, float **wout)
{
    // selects, not an indexed array: a runtime index into a local array
    // puts it in scratch memory in device code
    float *cur = wc ? w1 : w0, *other = wc ? w0 : w1;
    float *X = other, *Y = flow;
    *win = k == 0 ? cur : (((k - 1) & 1) ? Y : X);
    if (k == K - 1) *wout = K == 1 ? other : cur;
    else *wout = (k & 1) ? Y : X;
}

__host__ __device__ inline int tt_hyd_water_index_after(int wc, int K) { return K == 1 ? wc ^ 1 : wc; }

// Element i of a buffer addressed as base + 32-bit byte offset: with a
// wavefront-uniform base the compiler uses the scalar-base addressing mode
// (SGPR base + VGPR offset) instead of per-lane 64-bit address arithmetic.
// Valid while i * sizeof(T) < 2^32 (sides of at most 16384 cells).
template <typename T>
__device__ __forceinline__ const T &tt_at(const T *base, uint32_t i)
{
    return *(const T *)((const char *)base + i * (uint32_t)sizeof(T));
}
template <typename T>
__device__ __forceinline__ T *tt_ptr(T *base, uint32_t i)
{
    return (T *)((char *)base + i * (uint32_t)sizeof(T));
}

// uint8 plate id arithmetic of the reference (MAX_PLATE_COUNT - byte)
__device__ __forceinline__ int tt_slot_plate(uint32_t word, int slot)
{
    return (int)(uint8_t)(TT_MAXP - ((word >> (8 * slot)) & 0xFFu));
}

__device__ __forceinline__ unsigned tt_tri(unsigned i, unsigned j) { return i * (2u * TT_MAXP - i - 1u) / 2u + (j - i - 1u); }

// wrap for |v| < 2n (conditional add / subtract)
__device__ __forceinline__ int tt_wrap1(int v, int n)
{
    v = v < 0 ? v + n : v;
    return v >= n ? v - n : v;
}
// general wrap (CudaTexture::wrapCoordinate, cuda_texture.cuh:67-76)
__device__ __forceinline__ int tt_wrapg(int v, int n)
{
    int r = v % n;
    return r + (n & (r >> 31));
}
// the general wrap, with a conditional add / subtract when -n <= v < 2n
__device__ __forceinline__ int tt_wrap_near(int v, int n)
{
    return ((unsigned)(v + n) < (unsigned)(3 * n)) ? tt_wrap1(v, n) : tt_wrapg(v, n);
}

// XCD-aware block order: the dispatcher hands consecutive block ids to the 8
// XCDs round-robin; remapping so that each XCD receives a contiguous range of
// tiles keeps the halo rows of neighbouring tiles in the same L2.
__device__ __forceinline__ int tt_xcd_tile(int b, int nb)
{
    const int per = nb >> 3;
    if (per == 0 || (nb & 7)) return b;
    return (b & 7) * per + (b >> 3);
}

// asthenosphere vector at cell (x, y): computeCurl of the 256x256 noise table
// (plate_tectonics_kernel.cu:2788-2822), evaluated on the fly.
__device__ __forceinline__ float tt_noise_at(const uint8_t *noise, int x, int y)
{
    return (float)noise[tt_wrapg(y, 256) * 256 + tt_wrapg(x, 256)] / 255.0f;
}
// the vector of noise texel (sx, sy); cell (x, y) samples texel
// ((int)(x * 256 / W), (int)(y * 256 / H)) in double
__device__ __forceinline__ void tt_asth_texel(const uint8_t *noise, int sx, int sy, float &ax, float &ay)
{
    const float left = tt_noise_at(noise, sx - 1, sy), right = tt_noise_at(noise, sx + 1, sy);
    const float bottom = tt_noise_at(noise, sx, sy - 1), top = tt_noise_at(noise, sx, sy + 1);
    ay = (float)((double)(right - left) * 0.5);
    ax = (float)(-((double)(top - bottom) * 0.5));
}
__device__ __forceinline__ void tt_asth_at(const uint8_t *noise, int x, int y, double xs, double ys, float &ax, float &ay)
{
    tt_asth_texel(noise, (int)((double)x * xs), (int)((double)y * ys), ax, ay);
}

// ---------------------------------------------------------------------------
// device world view passed to kernels by value
// ---------------------------------------------------------------------------
struct tt_dev {
    int W, H, N;
    uint32_t iter;
    int seed;
    uint32_t wmul;          // ceil(2^(31 + l) / W), l = ceil(log2 W)
    int wsh;                // l - 1
    double axs, ays;        // 256 / W, 256 / H (noise scale)
    tt_kernel_settings ks;
    float rL;               // RN(1 / hydrationPipeLength)
    int divL;               // 1: tt_div_const is exact for the pipe length (checked)
    int sqrtOk;             // 1: tt_sqrt_normal is exact over its range (checked)
    int cvtOk;              // 1: tt_f2u32_hw equals tt_f2u32_sat for every float (checked)
};

// float -> uint32 by the hardware conversion alone.  v_cvt_u32_f32 clamps:
// negative values and NaN give 0, values of 2^32 and above give 0xffffffff,
// which is tt_f2u32_sat's definition; whether the device agrees for all 2^32
// bit patterns is checked once per process (k_check_sqrt) before a kernel uses
// it.  Written as one instruction because the C conversion is undefined
// outside [0, 2^32).
__device__ __forceinline__ uint32_t tt_f2u32_hw(float f)
{
    uint32_t r;
    asm("v_cvt_u32_f32 %0, %1" : "=v"(r) : "v"(f));
    return r;
}

// x / L as RN(x * rL) corrected by one fused multiply-add with the negated
// residual L * q0 - x, rL = RN(1/L).
// For numerators accepted by tt_div_safe the correction is exact and the
// result is the correctly rounded quotient (Markstein's theorem); whether it
// equals IEEE x / L for every such x is checked per divisor over all 2^32
// bit patterns (k_check_divisor) before it is used.
__device__ __forceinline__ float tt_div_const(float x, float L, float rL)
{
    const float q0 = x * rL;
    const float e = __builtin_fmaf(L, q0, -x);
    return __builtin_fmaf(-e, rL, q0);
}

// a / b from y = RN(1 / b), correctly rounded, for a and b in [2^-40, 2^40]
// (the caller checks the range).  q0 = RN(a y) is within two ulps of a / b;
// one correction q1 = RN(q0 + (a - b q0) y) leaves a relative error of about
// 2^-46 before its rounding, so q1 is a faithful quotient (one of the two
// floats around a / b); then a - b q1 is exact in one fused multiply-add and
// Markstein's correction RN(q1 + (a - b q1) y) is the correctly rounded
// quotient, given y = RN(1 / b) (P. Markstein, IBM J. Res. Dev. 34(1), 1990;
// Cornea, Golliver, Markstein, ARITH 1999, Theorem 1).  The range keeps the
// quotient, the residuals and their products with y normal.  Five
// operations, against about ten for an IEEE division; worth it where one
// reciprocal serves several quotients.
__device__ __forceinline__ float tt_div_by_rcp(float a, float b, float y)
{
    const float q0 = a * y;
    const float r0 = __builtin_fmaf(-q0, b, a);
    const float q1 = __builtin_fmaf(r0, y, q0);
    const float r1 = __builtin_fmaf(-q1, b, a);
    return __builtin_fmaf(r1, y, q1);
}

__device__ __forceinline__ int tt_div_rcp_in_range(float x)
{
    return (int)(x >= 0x1p-40f) & (int)(x <= 0x1p40f);
}

// Square root of 2^-96 <= x <= FLT_MAX: the hardware root (within 1 ulp) moved
// to a neighbour when the fused residual x - s' * s shows that the neighbour
// s' is the better root.  This is the correction that the IEEE square root of
// the compiler applies; below 2^-96 it also scales the argument, and it treats
// zeros and infinity apart, which this range leaves out.  Whether it equals
// the IEEE square root over the whole range is checked once per process
// (k_check_sqrt) before it is used.
__device__ __forceinline__ float tt_sqrt_normal(float x)
{
    const float s = __builtin_amdgcn_sqrtf(x);
    const float dn = __uint_as_float(__float_as_uint(s) - 1u);
    const float up = __uint_as_float(__float_as_uint(s) + 1u);
    float t = __builtin_fmaf(-dn, s, x) <= 0.0f ? dn : s;
    t = __builtin_fmaf(-up, s, x) > 0.0f ? up : t;
    return t;
}

__device__ __forceinline__ int tt_sqrt_in_range(float x)
{
    return (int)(x >= 0x1p-96f) & (int)(x <= 0x1.fffffep127f);
}

// 2^-100 <= |x| <= 2^100: the quotient and the residual stay normal (two
// compares; zeros, NaN and infinities are left to the IEEE division)
__device__ __forceinline__ int tt_div_safe(float x)
{
    const float a = __builtin_fabsf(x);
    return (int)(a >= 0x1p-100f) & (int)(a <= 0x1p100f);
}

// Row and column of a cell index c < 2^31 without a division: with
// m = ceil(2^(31+l) / W) and 2^(l-1) < W <= 2^l, m W - 2^(31+l) < W <= 2^l,
// so floor(c / W) = floor(c m / 2^(31+l)) (Granlund-Montgomery) for all
// c < 2^31.
__device__ __forceinline__ void tt_cell_xy(uint32_t c, const tt_dev &d, int &x, int &y)
{
    const uint32_t q = __umulhi(c, d.wmul) >> d.wsh;
    y = (int)q;
    x = (int)(c - q * (uint32_t)d.W);
}

// host side: the constants of tt_cell_xy
static inline void tt_dev_set_divisor(tt_dev &d)
{
    int l = 0;
    while ((1 << l) < d.W) ++l;
    if (l == 0) l = 1;
    const unsigned long long two = 1ull << (31 + l);
    d.wmul = (uint32_t)((two + (unsigned long long)d.W - 1ull) / (unsigned long long)d.W);
    d.wsh = l - 1;
}
