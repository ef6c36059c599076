// k_uplift.hip -- stage C: continental propagation (moveContinentalCollisions,
// plate_tectonics_kernel.cu:1087-1185), combinePropagatedCollisions
// (:1187-1218), jump flood (JFAInitialize / JFAPass / JFAComputeDistance,
// :1498-1600, host loop plate_tectonic_sim.cu:289-311) and applyUplift
// (:1662-1738, formulas :1646-1660).
//
// Canonical order of the propagation (the reference races through CAS inserts
// and float atomics): sources are applied one after another in ascending cell
// order.  Sources whose requests all target their own cell and whose cell no
// other source targets are independent and run in parallel; the rest (remote
// voxel-march targets, or cells that receive remote requests) are applied by
// one serial pass in cell order.
#include "tt_kernels.h"
#include "tt_stage.h"

__device__ __forceinline__ tt_prop tt_combine(tt_prop P, const tt_uplift &U, uint32_t word)
{
    if (P.inv != 0) {
        P.types = 1;
    } else if ((U.types & 1u) == 0) {
        P.types = U.types;
        P.cum = U.cum;
        P.inv = word;
    } else {
        P.inv = 1;
    }
    return P;
}

__device__ __forceinline__ int tt_seed_of(const tt_prop &P, int x, int y)
{
    return (P.inv >> 8) != 0 ? ((y << 16) | x) : -1;
}

__global__ __launch_bounds__(256) void k_propagate_local(tt_dev d, const uint32_t *word, const tt_uplift *uplift,
                                                         const uint8_t *srcFlag, const uint8_t *contested, tt_prop *prop,
                                                         int *jfa, uint8_t *serialFlag)
{
    const long long c = (long long)blockIdx.x * 256 + threadIdx.x;
    if (c >= d.N) return;
    int x, y;
    tt_cell_xy((uint32_t)c, d, x, y);
    const uint8_t sf = srcFlag[c];
    const uint8_t ct = contested[c];
    tt_prop P;
    P.cum = 0.0f;
    P.inv = 0u;
    P.types = 0;
    P.pad[0] = P.pad[1] = P.pad[2] = 0;
    serialFlag[c] = (ct && sf) ? 1 : 0;
    if (ct) {
        // applied by the serial pass and combined afterwards
        prop[c] = P;
        return;
    }
    const uint32_t cv = word[c];
    const tt_uplift U = uplift[c];
    if (sf == 1) {
        const int prim = tt_slot_plate(cv, (int)((U.types >> 1) & 3u));
        int total = 0;
        uint32_t inv = 0;
#pragma unroll
        for (int ci = 0; ci < 4; ++ci) {
            const int sec = tt_slot_plate(cv, ci);
            if (sec == TT_MAXP || sec == prim) continue;
            if (tt_insert_pair(inv, prim, sec)) total += 2;
            total += 2;  // the unconditional counts of both positions (same cell)
        }
        P.inv = inv;
        if (total > 0) P.cum = 0.0f + (U.cum / (float)total) * (float)total;
    }
    P = tt_combine(P, U, cv);
    prop[c] = P;
    jfa[c] = tt_seed_of(P, x, y);
}

hipError_t tt_launch_propagate_local(hipStream_t s, const tt_dev &d, const uint32_t *word, const tt_uplift *uplift,
                                     const uint8_t *srcFlag, const uint8_t *contested, tt_prop *prop, int *jfa,
                                     uint8_t *serialFlag)
{
    const unsigned blocks = (unsigned)(((long long)d.N + 255) / 256);
    hipLaunchKernelGGL(k_propagate_local, dim3(blocks), dim3(256), 0, s, d, word, uplift, srcFlag, contested, prop,
                       jfa, serialFlag);
    return hipGetLastError();
}

// One thread applies the listed sources in ascending cell order, exactly as the
// reference program would run them one after another: the walks and every
// read-modify-write of a target's record go to global memory.  Used when the
// sources or their distinct targets exceed what k_prop_apply holds in LDS.
__device__ void tt_propagate_serial_global(const tt_dev &d, const uint32_t *list, uint32_t n, const uint32_t *word,
                                           const tt_uplift *uplift, const float *height, const tt_plate_step *ps,
                                           tt_prop *prop)
{
    for (uint32_t li = 0; li < n; ++li) {
        const uint32_t c = list[li];
        int x, y;
        tt_cell_xy((uint32_t)c, d, x, y);
        const uint32_t cv = word[c];
        const tt_uplift U = uplift[c];
        const int prim = tt_slot_plate(cv, (int)((U.types >> 1) & 3u));
        const float hc = height[c];
        int ppx, ppy;
        tt_cont_prop(d, height, x, y, ps[prim].dirx, ps[prim].diry, ps[prim].pcx, ps[prim].pcy, hc, ppx, ppy);
        int posx[6] = {0, 0, 0, 0, 0, 0}, posy[6] = {0, 0, 0, 0, 0, 0}, cnt[6] = {0, 0, 0, 0, 0, 0};
        int total = 0;
        auto count_at = [&](int qx, int qy, int inc) {
            bool done = false;
#pragma unroll
            for (int i = 0; i < 6; ++i)
                if (!done && (cnt[i] == 0 || (posx[i] == qx && posy[i] == qy))) {
                    posx[i] = qx; posy[i] = qy; cnt[i] += inc; total += inc;
                    done = true;
                }
        };
        for (int ci = 0; ci < 4; ++ci) {
            const int sec = tt_slot_plate(cv, ci);
            if (sec == TT_MAXP || sec == prim) continue;
            int spx, spy;
            tt_cont_prop(d, height, x, y, ps[sec].dirx, ps[sec].diry, ps[sec].pcx, ps[sec].pcy, hc, spx, spy);
            tt_prop *tp = &prop[(size_t)tt_wrapg(ppy, d.H) * d.W + tt_wrapg(ppx, d.W)];
            if (ppx == spx && ppy == spy) {
                uint32_t inv = tp->inv;
                if (tt_insert_pair(inv, prim, sec)) { tp->inv = inv; count_at(ppx, ppy, 2); }
            } else {
                uint32_t inv = tp->inv;
                if (tt_insert_pair(inv, prim, sec)) { tp->inv = inv; count_at(ppx, ppy, 1); }
                tt_prop *ts = &prop[(size_t)tt_wrapg(spy, d.H) * d.W + tt_wrapg(spx, d.W)];
                uint32_t inv2 = ts->inv;
                if (tt_insert_pair(inv2, prim, sec)) { ts->inv = inv2; count_at(spx, spy, 1); }
            }
            count_at(ppx, ppy, 1);
            count_at(spx, spy, 1);
        }
        bool end = false;
#pragma unroll
        for (int i = 0; i < 6; ++i) {
            end = end || cnt[i] == 0;
            if (end) continue;
            tt_prop *t = &prop[(size_t)tt_wrapg(posy[i], d.H) * d.W + tt_wrapg(posx[i], d.W)];
            t->cum = t->cum + (U.cum / (float)total) * (float)cnt[i];
        }
    }
}

// The canonical serial pass in three kernels' worth of work, with the same
// result.  What makes the one-thread pass slow is memory latency: every step
// of a voxel-march walk and every update of a target is a dependent global
// access.  The walks depend only on the heights, so they run in parallel
// (k_prop_targets: one thread per walk).  Then one block gives every distinct
// target cell a slot in LDS and one thread applies the sources in ascending
// cell order to those slots; only the order-dependent part stays serial, at
// LDS latency.
//
// Request (li, k): source li's walk k (0: its primary plate, 1 + ci: the
// plate in collision slot ci, when valid and not the primary).  A target is
// kept as its cell and as its offset from the source before wrapping: the
// reference tells positions apart by unwrapped coordinates (two offsets can
// wrap to one cell on worlds narrower than two walks).
#define TT_PROP_INVALID 0xFFFFFFFFu
#define TT_PROP_SLOTS TT_PROP_SLOT_WORDS    // distinct target cells held in LDS
#define TT_PROP_CHUNK 512     // sources staged in LDS at a time

__global__ __launch_bounds__(256) void k_prop_targets(tt_dev d, const uint32_t *list, const uint32_t *count,
                                                      const uint32_t *word, const tt_uplift *uplift,
                                                      const float *height, const tt_plate_step *ps, int cap,
                                                      uint32_t *tcell, uint16_t *toff)
{
    const uint32_t t = blockIdx.x * 256u + threadIdx.x;
    const uint32_t li = t / 5u, k = t - 5u * li;
    const uint32_t n = *count;
    if (n > (uint32_t)cap || li >= n) return;
    const uint32_t c = list[li];
    int x, y;
    tt_cell_xy(c, d, x, y);
    const uint32_t cv = word[c];
    const tt_uplift U = uplift[c];
    const int prim = tt_slot_plate(cv, (int)((U.types >> 1) & 3u));
    const int pl = k == 0 ? prim : tt_slot_plate(cv, (int)k - 1);
    const size_t r = (size_t)k * cap + li;
    if (k > 0 && (pl == TT_MAXP || pl == prim)) {
        tcell[r] = TT_PROP_INVALID;
        return;
    }
    int tx, ty;
    tt_cont_prop(d, height, x, y, ps[pl].dirx, ps[pl].diry, ps[pl].pcx, ps[pl].pcy, height[c], tx, ty);
    tcell[r] = (uint32_t)tt_wrapg(ty, d.H) * (uint32_t)d.W + (uint32_t)tt_wrapg(tx, d.W);
    // a walk moves at most continentalPropMaxIters cells; offsets beyond
    // +-127 (an iteration limit above 127) leave the LDS form (tt_prop_apply
    // then falls back to the global pass)
    const int ox = tx - x + 128, oy = ty - y + 128;
    toff[r] = (ox < 0 || ox > 255 || oy < 0 || oy > 255) ? (uint16_t)0xFFFFu : (uint16_t)(ox | (oy << 8));
}

__global__ __launch_bounds__(1024) void k_prop_apply(tt_dev d, const uint32_t *list, const uint32_t *count,
                                                     const uint32_t *word, const tt_uplift *uplift,
                                                     const float *height, const tt_plate_step *ps, int cap,
                                                     const uint32_t *tcell, const uint16_t *toff, uint32_t *tslot,
                                                     uint32_t *slotOf, uint32_t *slotCell, tt_prop *prop)
{
    __shared__ uint32_t stInv[TT_PROP_SLOTS];
    __shared__ float stCum[TT_PROP_SLOTS];
    __shared__ float cCum[TT_PROP_CHUNK];
    __shared__ uint32_t cSec[TT_PROP_CHUNK];      // plate ids of slots 0..3 (0xFF: no request)
    __shared__ uint8_t cPrim[TT_PROP_CHUNK];
    __shared__ uint16_t cSlot[5][TT_PROP_CHUNK];
    __shared__ uint16_t cOff[5][TT_PROP_CHUNK];
    __shared__ uint32_t nSlots;
    __shared__ int overflow;
    const int tid = threadIdx.x;
    const uint32_t n = *count;
    if (n == 0) return;
    if (n > (uint32_t)cap) {
        if (tid == 0) tt_propagate_serial_global(d, list, n, word, uplift, height, ps, prop);
        return;
    }
    if (tid == 0) { nSlots = 0; overflow = 0; }
    __syncthreads();
    // 1. one owner request per distinct target cell (the smallest index)
    for (uint32_t k = 0; k < 5; ++k)
        for (uint32_t li = tid; li < n; li += 1024) {
            const size_t r = (size_t)k * cap + li;
            const uint32_t T = tcell[r];
            if (T == TT_PROP_INVALID) continue;
            if (toff[r] == 0xFFFFu) overflow = 1;
            atomicMin(&slotOf[T], (uint32_t)r);
        }
    __syncthreads();
    // 2. owners number the slots
    for (uint32_t k = 0; k < 5; ++k)
        for (uint32_t li = tid; li < n; li += 1024) {
            const size_t r = (size_t)k * cap + li;
            const uint32_t T = tcell[r];
            if (T == TT_PROP_INVALID || slotOf[T] != (uint32_t)r) continue;
            const uint32_t sl = atomicAdd(&nSlots, 1u);
            tslot[r] = sl;
            if (sl < TT_PROP_SLOTS) {
                slotCell[sl] = T;
                stInv[sl] = 0u;     // contested cells start from a cleared record
                stCum[sl] = 0.0f;
            }
        }
    __syncthreads();
    // 3. the other requests take their owner's slot
    for (uint32_t k = 0; k < 5; ++k)
        for (uint32_t li = tid; li < n; li += 1024) {
            const size_t r = (size_t)k * cap + li;
            const uint32_t T = tcell[r];
            if (T == TT_PROP_INVALID) continue;
            const uint32_t o = slotOf[T];
            if (o != (uint32_t)r) tslot[r] = tslot[o];
        }
    __syncthreads();
    const bool lds = nSlots <= TT_PROP_SLOTS && !overflow;
    if (lds) {
        for (uint32_t base = 0; base < n; base += TT_PROP_CHUNK) {
            // stage a chunk of sources
            if (tid < TT_PROP_CHUNK && base + tid < n) {
                const uint32_t li = base + tid;
                const uint32_t c = list[li];
                const uint32_t cv = word[c];
                const tt_uplift U = uplift[c];
                cCum[tid] = U.cum;
                cPrim[tid] = (uint8_t)tt_slot_plate(cv, (int)((U.types >> 1) & 3u));
                uint32_t sec = 0;
                for (int k = 0; k < 5; ++k) {
                    const size_t r = (size_t)k * cap + li;
                    const bool v = tcell[r] != TT_PROP_INVALID;
                    cSlot[k][tid] = v ? (uint16_t)tslot[r] : (uint16_t)0;
                    cOff[k][tid] = v ? toff[r] : (uint16_t)0;
                    if (k > 0) sec |= (v ? (uint32_t)tt_slot_plate(cv, k - 1) : 0xFFu) << (8 * (k - 1));
                }
                cSec[tid] = sec;
            }
            __syncthreads();
            if (tid == 0) {
                const uint32_t m = min((uint32_t)TT_PROP_CHUNK, n - base);
                for (uint32_t j = 0; j < m; ++j) {
                    const int prim = cPrim[j];
                    const uint32_t Spp = cSlot[0][j], kpp = cOff[0][j];
                    // positions by key in up to 6 slots; every index is a
                    // constant after unrolling, so the slots stay in registers
                    uint32_t pk[6] = {0, 0, 0, 0, 0, 0}, ps_[6] = {0, 0, 0, 0, 0, 0};
                    int cnt[6] = {0, 0, 0, 0, 0, 0}, np = 0, total = 0;
                    auto count_at = [&](uint32_t key, uint32_t sl, int inc) {
                        bool done = false;
#pragma unroll
                        for (int i = 0; i < 6; ++i)
                            if (!done && i < np && pk[i] == key) { cnt[i] += inc; done = true; }
                        if (!done && np < 6) {
#pragma unroll
                            for (int i = 0; i < 6; ++i)
                                if (i == np) { pk[i] = key; ps_[i] = sl; cnt[i] = inc; }
                            ++np;
                            done = true;
                        }
                        if (done) total += inc;
                    };
                    const uint32_t secs = cSec[j];
                    for (int ci = 0; ci < 4; ++ci) {
                        const uint32_t sec = (secs >> (8 * ci)) & 0xFFu;
                        if (sec == 0xFFu) continue;
                        const uint32_t Ssp = cSlot[1 + ci][j], ksp = cOff[1 + ci][j];
                        if (kpp == ksp) {
                            uint32_t inv = stInv[Spp];
                            if (tt_insert_pair(inv, prim, (int)sec)) { stInv[Spp] = inv; count_at(kpp, Spp, 2); }
                        } else {
                            uint32_t inv = stInv[Spp];
                            if (tt_insert_pair(inv, prim, (int)sec)) { stInv[Spp] = inv; count_at(kpp, Spp, 1); }
                            uint32_t inv2 = stInv[Ssp];
                            if (tt_insert_pair(inv2, prim, (int)sec)) { stInv[Ssp] = inv2; count_at(ksp, Ssp, 1); }
                        }
                        count_at(kpp, Spp, 1);
                        count_at(ksp, Ssp, 1);
                    }
                    const float cum = cCum[j];
#pragma unroll
                    for (int i = 0; i < 6; ++i)
                        if (i < np) stCum[ps_[i]] = stCum[ps_[i]] + (cum / (float)total) * (float)cnt[i];
                }
            }
            __syncthreads();
        }
        for (uint32_t sl = tid; sl < nSlots; sl += 1024) {
            tt_prop *t = &prop[slotCell[sl]];
            t->inv = stInv[sl];
            t->cum = stCum[sl];
        }
    } else if (tid == 0) {
        tt_propagate_serial_global(d, list, n, word, uplift, height, ps, prop);
    }
    __syncthreads();
    // leave the slot map cleared for the next step
    for (uint32_t k = 0; k < 5; ++k)
        for (uint32_t li = tid; li < n; li += 1024) {
            const uint32_t T = tcell[(size_t)k * cap + li];
            if (T != TT_PROP_INVALID) slotOf[T] = TT_PROP_INVALID;
        }
}

int tt_prop_cap(long long cells) { return (int)std::max(1024LL, std::min(16384LL, cells / 64)); }

hipError_t tt_launch_propagate_serial(hipStream_t s, const tt_dev &d, const uint32_t *list, const uint32_t *count,
                                      const uint32_t *word, const tt_uplift *uplift, const float *height,
                                      const tt_plate_step *ps, tt_prop *prop, int cap, uint32_t *tcell,
                                      uint16_t *toff, uint32_t *tslot, uint32_t *slotOf, uint32_t *slotCell)
{
    const unsigned blocks = (unsigned)(((long long)cap * 5 + 255) / 256);
    hipLaunchKernelGGL(k_prop_targets, dim3(blocks), dim3(256), 0, s, d, list, count, word, uplift, height, ps, cap,
                       tcell, toff);
    hipLaunchKernelGGL(k_prop_apply, dim3(1), dim3(1024), 0, s, d, list, count, word, uplift, height, ps, cap, tcell,
                       toff, tslot, slotOf, slotCell, prop);
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_combine_contested(tt_dev d, const uint32_t *word, const tt_uplift *uplift,
                                                           const uint8_t *contested, const uint8_t *srcFlag,
                                                           tt_prop *prop, int *jfa, tt_step_ctl *ctl)
{
    const long long c = (long long)blockIdx.x * 256 + threadIdx.x;
    int remote = 0;
    if (c < d.N) {
        remote = srcFlag[c] == 2;
        if (contested[c]) {
            int x, y;
    tt_cell_xy((uint32_t)c, d, x, y);
            const tt_prop P = tt_combine(prop[c], uplift[c], word[c]);
            prop[c] = P;
            jfa[c] = tt_seed_of(P, x, y);
        }
    }
    // wavefront-aggregated count of sources with remote targets
    const unsigned long long b = __ballot(remote);
    if ((threadIdx.x & 63) == 0 && b) atomicAdd(&ctl->complex_sources, __popcll(b));
}

hipError_t tt_launch_combine_contested(hipStream_t s, const tt_dev &d, const uint32_t *word, const tt_uplift *uplift,
                                       const uint8_t *contested, const uint8_t *srcFlag, tt_prop *prop, int *jfa,
                                       tt_step_ctl *ctl)
{
    const unsigned blocks = (unsigned)(((long long)d.N + 255) / 256);
    hipLaunchKernelGGL(k_combine_contested, dim3(blocks), dim3(256), 0, s, d, word, uplift, contested, srcFlag, prop,
                       jfa, ctl);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// jump flood: seeds are packed (y << 16 | x), -1 = none
// ---------------------------------------------------------------------------
__device__ __forceinline__ int tt_tdist2(int ax, int ay, int bx, int by, int W, int H)
{
    int dx = abs(ax - bx), dy = abs(ay - by);
    dx = min(dx, W - dx);
    dy = min(dy, H - dy);
    return dx * dx + dy * dy;
}

// Squared toroidal distance between two cells packed as (y << 16) | x, with
// packed 16-bit arithmetic: |a - b| per component, min(d, size - d), then
// dx * dx + dy * dy by one 2-way dot product (exact: components < 2^14).
typedef short tt_s2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ int tt_tdist2_pk(int p, int s, int wh)
{
    const tt_s2 a = __builtin_bit_cast(tt_s2, p), b = __builtin_bit_cast(tt_s2, s);
    const tt_s2 n = __builtin_bit_cast(tt_s2, wh);
    tt_s2 dd = a - b;
    dd = __builtin_elementwise_max(dd, -dd);
    dd = __builtin_elementwise_min(dd, n - dd);
    return __builtin_amdgcn_sdot2(dd, dd, 0, false);
}

// Power-of-two sides: u = (a - b) mod n is the 16-bit difference masked to
// n - 1, and min(u, n - u) equals min(|a - b|, n - |a - b|) (for a - b < 0,
// u = n - |a - b|), one instruction fewer than the absolute value.
typedef unsigned short tt_u2 __attribute__((ext_vector_type(2)));
template <bool P2>
__device__ __forceinline__ int tt_tdist2(int p, int s, int wh)
{
    if (!P2) return tt_tdist2_pk(p, s, wh);
    const tt_u2 a = __builtin_bit_cast(tt_u2, p), b = __builtin_bit_cast(tt_u2, s);
    const tt_u2 n = __builtin_bit_cast(tt_u2, wh);
    const unsigned m = __builtin_bit_cast(unsigned, n - (unsigned short)1);
    const tt_u2 u = __builtin_bit_cast(tt_u2, __builtin_bit_cast(unsigned, a - b) & m);
    const tt_s2 t = __builtin_bit_cast(tt_s2, __builtin_elementwise_min(u, n - u));
    return __builtin_amdgcn_sdot2(t, t, 0, false);
}

// Toroidal wrap, remainder and quotient for the jump flood.  On power-of-two
// sides (P2) they are masks and shifts; otherwise the general forms, whose
// integer divisions the compiler expands on the vector unit in every thread.
template <bool P2> __device__ __forceinline__ int tt_jwrap(int v, int n) { return P2 ? (v & (n - 1)) : tt_wrapg(v, n); }
template <bool P2> __device__ __forceinline__ int tt_jmod(int v, int n) { return P2 ? (v & (n - 1)) : v % n; }
template <bool P2> __device__ __forceinline__ int tt_jdiv(int v, int n) { return P2 ? (v >> __builtin_ctz((unsigned)n)) : v / n; }

// JFAPass (plate_tectonics_kernel.cu:1512-1571): candidates at (ox, oy) * step
// in scan order oy = -1..1, ox = -1..1; a strictly smaller squared distance
// replaces the current seed.  4 consecutive cells of a row per thread: every
// tap is an aligned 16-byte load, so a wavefront moves four times the bytes per load
// instruction.  When the column offset (step mod W) is a multiple of 4 the
// taps at x - s, x, x + s are whole groups; for steps 1..3 each row loads the
// groups at x - 4, x and x + 4 and the taps are picked from that 12-cell
// window.  Each cell scans its own candidates in the same order.
// SX: 0 when the column offset is a multiple of 4 (taps are whole groups);
// 1, 2 or 3 for the window form (the offset fixed at compile time, so each tap
// is a fixed window entry)
template <int SX, bool P2>
__global__ __launch_bounds__(256) void k_jfa_pass4(tt_dev d, const int4 *__restrict__ in, int4 *__restrict__ out,
                                                   int step)
{
    const uint32_t q = blockIdx.x * 256u + threadIdx.x;
    if (q >= (uint32_t)(d.N >> 2)) return;
    int x, y;
    tt_cell_xy(q << 2, d, x, y);
    const int W = d.W, H = d.H;
    const int wh = (H << 16) | W;
    const int sx = tt_jmod<P2>(step, W), sy = tt_jmod<P2>(step, H);
    const int rows[3] = {tt_wrap1(y - sy, H) * W, y * W, tt_wrap1(y + sy, H) * W};
    // window: groups at x - 4, x, x + 4; otherwise at x - s, x, x + s
    constexpr bool WINDOW = SX > 0;
    const int gs = WINDOW ? 4 : sx;
    const int cols[3] = {tt_wrap1(x - gs, W), x, tt_wrap1(x + gs, W)};
    int t[3][12];
#pragma unroll
    for (int oy = 0; oy < 3; ++oy)
#pragma unroll
        for (int g = 0; g < 3; ++g) {
            const int4 v = in[(rows[oy] + cols[g]) >> 2];
            t[oy][4 * g + 0] = v.x;
            t[oy][4 * g + 1] = v.y;
            t[oy][4 * g + 2] = v.z;
            t[oy][4 * g + 3] = v.w;
        }
    int res[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int me = (y << 16) | (x + k);
        int best = t[1][4 + k];
        int bd = 0x7FFFFFFF;
        if (best >= 0) bd = tt_tdist2<P2>(me, best, wh);
#pragma unroll
        for (int oy = 0; oy < 3; ++oy) {
#pragma unroll
            for (int ox = 0; ox < 3; ++ox) {
                if (ox == 1 && oy == 1) continue;
                int cand;
                if (WINDOW) {
                    // window index 4 + k + (ox - 1) * SX, in [1, 10]
                    cand = t[oy][4 + k + (ox - 1) * SX];
                } else {
                    cand = t[oy][4 * ox + k];
                }
                if (cand < 0) continue;
                const int dd = tt_tdist2<P2>(me, cand, wh);
                if (dd < bd) { bd = dd; best = cand; }
            }
        }
        res[k] = best;
    }
    out[q] = make_int4(res[0], res[1], res[2], res[3]);
}

// The same pass for R cells of a column chain (rows y0, y0 + s, ...,
// y0 + (R - 1) s) times 4 cells of a row per thread, when the column offset
// is a multiple of 4 and H is a multiple of R s: the chain reads R + 2 rows
// of taps instead of 3 R, each row's taps serving up to three of its cells.
// Rows are grouped into bands of R s rows; chain (band b, offset r0) starts at
// row b R s + r0, so the chains cover every row once.
template <int R, bool P2>
__global__ __launch_bounds__(256) void k_jfa_chain(tt_dev d, const int4 *__restrict__ in, int4 *__restrict__ out,
                                                   int step)
{
    const uint32_t q = blockIdx.x * 256u + threadIdx.x;
    const int W = d.W, H = d.H;
    const uint32_t g4 = (uint32_t)(W >> 2);
    if (q >= (uint32_t)(d.N >> 2) / R) return;
    const int x = tt_jmod<P2>((int)q, (int)g4) * 4;
    const int chain = tt_jdiv<P2>((int)q, (int)g4);
    const int sy = tt_jmod<P2>(step, H), sx = tt_jmod<P2>(step, W);
    const int y0 = tt_jdiv<P2>(chain, sy) * (R * sy) + tt_jmod<P2>(chain, sy);
    const int wh = (H << 16) | W;
    const int cols[3] = {tt_wrap1(x - sx, W), x, tt_wrap1(x + sx, W)};
    // rows y0 - s, y0, ..., y0 + R s
    int t[R + 2][3][4];
#pragma unroll
    for (int j = 0; j < R + 2; ++j) {
        const int row = tt_jwrap<P2>(y0 + (j - 1) * sy, H) * W;
#pragma unroll
        for (int g = 0; g < 3; ++g) {
            const int4 v = in[(row + cols[g]) >> 2];
            t[j][g][0] = v.x;
            t[j][g][1] = v.y;
            t[j][g][2] = v.z;
            t[j][g][3] = v.w;
        }
    }
#pragma unroll
    for (int j = 0; j < R; ++j) {
        const int y = y0 + j * sy;   // < H: the chain stays inside its band
        int res[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int me = (y << 16) | (x + k);
            int best = t[j + 1][1][k];
            int bd = 0x7FFFFFFF;
            if (best >= 0) bd = tt_tdist2<P2>(me, best, wh);
#pragma unroll
            for (int oy = 0; oy < 3; ++oy) {
#pragma unroll
                for (int ox = 0; ox < 3; ++ox) {
                    if (ox == 1 && oy == 1) continue;
                    const int cand = t[j + oy][ox][k];
                    if (cand < 0) continue;
                    const int dd = tt_tdist2<P2>(me, cand, wh);
                    if (dd < bd) { bd = dd; best = cand; }
                }
            }
            res[k] = best;
        }
        out[((size_t)y * W + x) >> 2] = make_int4(res[0], res[1], res[2], res[3]);
    }
}

template <bool P2>
static hipError_t launch_jfa(hipStream_t s, const tt_dev &d, const int *in, int *out, int step, int sx, int sy,
                             unsigned blocks)
{
    if (sx % 4 == 0 && sy > 0 && d.H % (4 * sy) == 0)
        hipLaunchKernelGGL((k_jfa_chain<4, P2>), dim3((blocks + 3) / 4), dim3(256), 0, s, d, (const int4 *)in,
                           (int4 *)out, step);
    else if (sx % 4 == 0 && sy > 0 && d.H % (2 * sy) == 0)
        hipLaunchKernelGGL((k_jfa_chain<2, P2>), dim3((blocks + 1) / 2), dim3(256), 0, s, d, (const int4 *)in,
                           (int4 *)out, step);
    else if (sx % 4 == 0)
        hipLaunchKernelGGL((k_jfa_pass4<0, P2>), dim3(blocks), dim3(256), 0, s, d, (const int4 *)in, (int4 *)out, step);
    else if (sx == 1)
        hipLaunchKernelGGL((k_jfa_pass4<1, P2>), dim3(blocks), dim3(256), 0, s, d, (const int4 *)in, (int4 *)out, step);
    else if (sx == 2)
        hipLaunchKernelGGL((k_jfa_pass4<2, P2>), dim3(blocks), dim3(256), 0, s, d, (const int4 *)in, (int4 *)out, step);
    else if (sx == 3)
        hipLaunchKernelGGL((k_jfa_pass4<3, P2>), dim3(blocks), dim3(256), 0, s, d, (const int4 *)in, (int4 *)out, step);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t tt_launch_jfa_pass(hipStream_t s, const tt_dev &d, const int *in, int *out, int step)
{
    // N is a multiple of 4 (W is a multiple of 64); 16-byte aligned buffers.
    // Steps are powers of two: step mod W is a multiple of 4 for steps >= 4
    // (W is a multiple of 64) and is 1 or 2 otherwise.
    const unsigned blocks = (unsigned)(((long long)(d.N >> 2) + 255) / 256);
    const int sx = step % d.W, sy = step % d.H;
    const bool p2 = (d.W & (d.W - 1)) == 0 && (d.H & (d.H - 1)) == 0;
    return p2 ? launch_jfa<true>(s, d, in, out, step, sx, sy, blocks)
              : launch_jfa<false>(s, d, in, out, step, sx, sy, blocks);
}

// ---------------------------------------------------------------------------
// JFAComputeDistance + applyUplift
// ---------------------------------------------------------------------------
__device__ __forceinline__ float tt_cont_formula(float x)
{
    const float ax = tt_fabsf(x);
    return tt_expf((-0.02f * x) * x) + tt_expf((-0.2f * x) * x) * (0.5f - tt_fabsf(tt_fmodf(ax, 1.0f) - 0.5f));
}

__device__ __forceinline__ float tt_sub_under(float x)
{
    const float l = tt_log10f(1.0f + 8.0f * (x - 1.0f));
    return tt_fminf(l / 1.908485019f - 1.0f, 0.0f);
}

__global__ __launch_bounds__(256) void k_uplift(tt_dev d, const int *jfa, const tt_prop *prop, const uint8_t *idW,
                                                const tt_plate_step *ps, float *hW, tt_dist *dist)
{
    const long long c = (long long)blockIdx.x * 256 + threadIdx.x;
    if (c >= d.N) return;
    int x, y;
    tt_cell_xy((uint32_t)c, d, x, y);
    const int s = jfa[c];
    if (s < 0) {
        tt_dist z;
        z.dist = -1;
        z.orig = 0u;
        dist[c] = z;
        return;
    }
    const int sx = s & 0xFFFF, sy = s >> 16;
    const int d2 = tt_tdist2(x, y, sx, sy, d.W, d.H);
    tt_dist db;
    db.dist = (int)sqrtf((float)d2);
    db.orig = (uint32_t)sy * (uint32_t)d.W + (uint32_t)sx;
    dist[c] = db;
    const tt_prop U = prop[db.orig];
    const int pid = idW[c];
    const int type = U.types & 1;
    const int pi = (U.types >> 1) & 3, si = (U.types >> 3) & 3;
    const int prim = tt_slot_plate(U.inv, pi);
    const int sec = tt_slot_plate(U.inv, si);
    float h = hW[c];
    if (type == 0 && sec == pid) {
        const float v = tt_sub_under((float)db.dist);
        const float m = 1.0f + v * d.ks.subductionUnderUpliftMultiplier;
        h = tt_fmaxf(h * m, d.ks.minUpliftClamp);
    } else if (type == 0 && prim == pid) {
        const float sizeX = (float)d.W, sizeY = (float)d.H;
        const float dx = tt_fmodf((float)(x - sx) + sizeX * 0.5f, sizeX) - sizeX * 0.5f;
        const float dy = tt_fmodf((float)(y - sy) + sizeY * 0.5f, sizeY) - sizeY * 0.5f;
        const float mg = tt_fmaxf(sqrtf(dx * dx + dy * dy), 1e-20f);
        const float nx = dx / mg, ny = dy / mg;
        const float dot = tt_fmaxf(nx * ps[sec].dirx + ny * ps[sec].diry, 0.0f);
        const float v = tt_cont_formula(((float)db.dist - (float)d.ks.subductionOverRange) - 10.0f);
        const float m = 1.0f + (v * dot) * d.ks.subductionOverUpliftMultiplier;
        h = tt_fmaxf(h * m, d.ks.minUpliftClamp);
    } else if (type == 1) {
        const float v = tt_cont_formula((float)db.dist);
        const float hc = (v * U.cum) * d.ks.continentalUpliftMultiplier;
        h = tt_fminf(h + hc, d.ks.maxUpliftClamp);
    } else {
        return;
    }
    hW[c] = h;
}

hipError_t tt_launch_uplift(hipStream_t s, const tt_dev &d, const int *jfa, const tt_prop *prop, const uint8_t *idW,
                            const tt_plate_step *ps, float *hW, tt_dist *dist)
{
    const unsigned blocks = (unsigned)(((long long)d.N + 255) / 256);
    hipLaunchKernelGGL(k_uplift, dim3(blocks), dim3(256), 0, s, d, jfa, prop, idW, ps, hW, dist);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// bounded-radius separable uplift (uplift_mode 1): VerticalBlur
// (plate_tectonics_kernel.cu:1603-1643) and HorizontalBlur (:1740-1843)
// ---------------------------------------------------------------------------
// nearest seed cell (>= 2 involved plates) within R rows of the column; first
// in offset order -R..R on ties
__global__ __launch_bounds__(256) void k_vblur(tt_dev d, const tt_prop *__restrict__ prop, tt_dist *__restrict__ dist)
{
    const uint32_t c = blockIdx.x * 256u + threadIdx.x;
    if (c >= (uint32_t)d.N) return;
    int x, y;
    tt_cell_xy(c, d, x, y);
    const int R = d.ks.upliftRange;
    int best = -1, by = 0;
    for (int i = -R; i <= R; ++i) {
        const int sy = tt_wrapg(y + i, d.H);
        if ((prop[(uint32_t)sy * (uint32_t)d.W + (uint32_t)x].inv >> 8) != 0u) {
            const int sd = i < 0 ? -i : i;
            if (best == -1 || sd < best) { best = sd; by = sy; }
        }
    }
    tt_dist o;
    o.dist = best;
    o.orig = best != -1 ? (uint32_t)by * (uint32_t)d.W + (uint32_t)x : 0u;
    dist[c] = o;
}

// the vertical results within R columns of the row whose collision involves
// the cell's plate (isPartOfCollision, :1409-1416), plus the column offset;
// first strictly smallest total; then the uplift formulas with
// subductionOverFormula(dist)
__global__ __launch_bounds__(256) void k_hblur(tt_dev d, const tt_dist *__restrict__ dist,
                                               const tt_prop *__restrict__ prop, const uint8_t *__restrict__ idW,
                                               const tt_plate_step *__restrict__ ps, float *__restrict__ hW)
{
    const uint32_t c = blockIdx.x * 256u + threadIdx.x;
    if (c >= (uint32_t)d.N) return;
    int x, y;
    tt_cell_xy(c, d, x, y);
    const int R = d.ks.upliftRange;
    const int pid = idW[c];
    int cd = -1;
    uint32_t corig = 0u;
    tt_prop U;
    U.cum = 0.0f;
    U.inv = 0u;
    U.types = 0;
    const uint32_t row = (uint32_t)y * (uint32_t)d.W;
    for (int i = -R; i <= R; ++i) {
        const tt_dist b = dist[row + (uint32_t)tt_wrapg(x + i, d.W)];
        if (b.dist == -1) continue;
        const tt_prop P = prop[b.orig];
        bool part = false;
#pragma unroll
        for (int k = 0; k != 4; ++k)
            if ((int)(((uint32_t)TT_MAXP - (P.inv >> (8 * k))) & 0xFFu) == pid) part = true;
        if (!part) continue;
        const int td = b.dist + (i < 0 ? -i : i);
        if (cd == -1 || td < cd) { cd = td; corig = b.orig; U = P; }
    }
    if (cd <= -1) return;
    const int type = U.types & 1;
    const int pi = (U.types >> 1) & 3, si = (U.types >> 3) & 3;
    const int prim = tt_slot_plate(U.inv, pi);
    const int sec = tt_slot_plate(U.inv, si);
    float h = hW[c];
    if (type == 0 && sec == pid) {
        const float v = tt_sub_under((float)cd);
        const float m = 1.0f + v * d.ks.subductionUnderUpliftMultiplier;
        h = tt_fmaxf(h * m, d.ks.minUpliftClamp);
    } else if (type == 0 && prim == pid) {
        int ox, oy;
        tt_cell_xy(corig, d, ox, oy);
        const float sizeX = (float)d.W, sizeY = (float)d.H;
        const float dx = tt_fmodf((float)(x - ox) + sizeX * 0.5f, sizeX) - sizeX * 0.5f;
        const float dy = tt_fmodf((float)(y - oy) + sizeY * 0.5f, sizeY) - sizeY * 0.5f;
        const float mg = tt_fmaxf(sqrtf(dx * dx + dy * dy), 1e-20f);
        const float nx = dx / mg, ny = dy / mg;
        const float dot = tt_fmaxf(nx * ps[sec].dirx + ny * ps[sec].diry, 0.0f);
        const float v = tt_cont_formula((float)cd - 10.0f);
        const float m = 1.0f + (v * dot) * d.ks.subductionOverUpliftMultiplier;
        h = tt_fmaxf(h * m, d.ks.minUpliftClamp);
    } else if (type == 1) {
        const float v = tt_cont_formula((float)cd);
        const float hc = (v * U.cum) * d.ks.continentalUpliftMultiplier;
        h = tt_fminf(h + hc, d.ks.maxUpliftClamp);
    } else {
        return;
    }
    hW[c] = h;
}

hipError_t tt_launch_separable_uplift(hipStream_t s, const tt_dev &d, const tt_prop *prop, const uint8_t *idW,
                                      const tt_plate_step *ps, float *hW, tt_dist *dist)
{
    const unsigned blocks = (unsigned)(((long long)d.N + 255) / 256);
    hipLaunchKernelGGL(k_vblur, dim3(blocks), dim3(256), 0, s, d, prop, dist);
    hipLaunchKernelGGL(k_hblur, dim3(blocks), dim3(256), 0, s, d, dist, prop, idW, ps, hW);
    return hipGetLastError();
}
