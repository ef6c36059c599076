// k_collide.hip -- stage A (advection, (cell, plate) key sort, collision
// words) and stage B (collision typing and resolution) on gfx950.
//
// Reference: getPixelMovements / registerPlateCollisions
// (plate_tectonics_kernel.cu:232-298) with the sort-by-key design of
// plate_tectonic_sim.cu:188-201; determineCollisionType / writeCollision /
// createCollisionTypeMatrix (:782-972); processCollisions and its helpers
// (:301-745); flipDivergedBitmap (:772-780).
//
// Advection as a counting sort: every plate moves rigidly by an integer offset
// o_p, so the plates arriving at destination d are exactly the plates p with
// ids[d - o_p] == p.  Each distinct offset contributes at most one plate, so a
// destination needs one LDS lookup per distinct offset (typically < 40) instead
// of an atomic scatter.  The per-destination counts are prefix-summed with
// wavefront scans (one wavefront = one 64-cell row segment) to place the
// sorted (dest << 8 | plate) keys; slots of the collision word are the four
// smallest plate ids in ascending order.
#include "tt_kernels.h"
#include "tt_stage.h"

#define ADV_R 8
#define ADV_TW (TT_TILE_W + 2 * ADV_R)
#define ADV_TH (TT_TILE_H + 2 * ADV_R)
#define ADV_K (2 * ADV_R + 1)   // offsets per axis within ADV_R

// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_plate_prep(const tt_plate_data *pd, tt_plate_step *ps, tt_step_ctl *ctl,
                                                    long long *fric)
{
    __shared__ int sox[256], soy[256];
    __shared__ int nd, maxr, zero;
    const int p = threadIdx.x;
    if (p == 0) { nd = 0; maxr = 0; zero = 0; }
    tt_plate_step st;
    if (p < TT_MAXP) {
        const tt_plate_data q = pd[p];
        const float mx = q.direction[0] * q.velocity;
        const float my = q.direction[1] * q.velocity;
        st.ox = (int)tt_floorf(q.pixelCenter[0] + mx);
        st.oy = (int)tt_floorf(q.pixelCenter[1] + my);
        st.vel = q.velocity;
        st.mass = q.mass;
        st.dirx = q.direction[0];
        st.diry = q.direction[1];
        st.pcx = q.pixelCenter[0];
        st.pcy = q.pixelCenter[1];
        fric[p] = 0;
    } else {
        // entry 255 stands for "no plate": the default PlateData of the reference
        st.ox = 0; st.oy = 0; st.vel = 0.0f; st.mass = 0.0f;
        st.dirx = 0.0f; st.diry = 0.0f; st.pcx = 0.0f; st.pcy = 0.0f;
    }
    ps[p] = st;
    sox[p] = st.ox;
    soy[p] = st.oy;
    __syncthreads();
    if (p < TT_MAXP) {
        bool first = true;
        for (int q = 0; q < p; ++q)
            if (sox[q] == st.ox && soy[q] == st.oy) { first = false; break; }
        if (first) {
            const int k = atomicAdd(&nd, 1);
            ctl->dox[k] = st.ox;
            ctl->doy[k] = st.oy;
            const int r = max(abs(st.ox), abs(st.oy));
            atomicMax(&maxr, r);
            if (r == 0) zero = 1;
        }
    }
    __syncthreads();
    if (p == 0) {
        ctl->n_distinct = nd;
        ctl->max_r = maxr;
        ctl->zero_distinct = zero;
        ctl->overflow_cells = 0;
        ctl->complex_sources = 0;
    }
}

hipError_t tt_launch_plate_prep(hipStream_t s, const tt_plate_data *pd, tt_plate_step *ps, tt_step_ctl *ctl,
                                long long *fric)
{
    hipLaunchKernelGGL(k_plate_prep, dim3(1), dim3(256), 0, s, pd, ps, ctl, fric);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// collision typing helpers
// ---------------------------------------------------------------------------
// writeCollision: the counter (pair * 5 + kind) that the pair's collision
// increments, or -1 when the pair's type is kept (the cooldown early exit)
__device__ __forceinline__ int tt_collision_key(const tt_dev &d, const uint8_t *bitmap, float hA, float hB, int pA,
                                                int pB)
{
    const bool cond = pA < pB;
    const int mn = cond ? pA : pB, mx = cond ? pB : pA;
    const float hmn = cond ? hA : hB, hmx = cond ? hB : hA;
    const unsigned t = tt_tri((unsigned)mn, (unsigned)mx);
    const unsigned packed = bitmap[t];
    // `packed >> 0 & 1 && packed >> 4 & 0x0F < cooldown` (plate_tectonics_kernel.cu:813)
    const unsigned cmpbit = (0x0F < d.ks.collisionTypeUpdateCooldown) ? 1u : 0u;
    if ((packed & 1u) && ((packed >> 4) & cmpbit)) return -1;
    const bool minCont = hmn >= d.ks.continentalCrustThreshold;
    const bool maxCont = hmx >= d.ks.continentalCrustThreshold;
    int k;
    if (minCont && maxCont) k = 4;       // continental
    else if (minCont) k = 2;             // subductionsBContinental
    else if (maxCont) k = 0;             // subductionsAContinental
    else if (hmn <= hmx) k = 1;          // subductionsAOceanic
    else k = 3;                          // subductionsBOceanic
    return (int)(t * 5u) + k;
}

// counts[key] += 1 for every lane with key >= 0, as one atomic per distinct key
// of the wavefront (neighbouring collision cells mostly share the pair and the
// kind; integer sums do not depend on the order).  Called by the whole
// wavefront.
__device__ __forceinline__ void tt_wave_count(int key, int *counts)
{
    unsigned long long pending = __builtin_amdgcn_ballot_w64(key >= 0);
    const int lane = (int)(threadIdx.x & 63u);
    while (pending) {
        const int first = (int)__builtin_ctzll(pending);
        const int leader = __builtin_amdgcn_readlane(key, first);
        const unsigned long long same = __builtin_amdgcn_ballot_w64(key == leader);
        if (lane == first) atomicAdd(&counts[leader], (int)__popcll(same));
        pending &= ~same;
    }
}

// arrivals at destination (gx, gy): the set of plates p with ids[g - o_p] == p
// (one candidate per distinct offset); returns the count.
__device__ __forceinline__ int tt_arrivals(const tt_dev &d, const uint8_t *ids, int gx, int gy, const int *dox,
                                           const int *doy, int nd, const int *pox, const int *poy, tt_min4 &set)
{
    int n = 0;
    set.clear();
    for (int j = 0; j < nd; ++j) {
        const int ox = dox[j], oy = doy[j];
        const int p = ids[(size_t)tt_wrapg(gy - oy, d.H) * d.W + tt_wrapg(gx - ox, d.W)];
        if (pox[p] == ox && poy[p] == oy) {
            set.add((unsigned)p);
            ++n;
        }
    }
    return n;
}

// offset key of |ox|, |oy| <= ADV_R
__device__ __forceinline__ uint32_t tt_adv_key(int ox, int oy) { return ((uint32_t)(oy + 128) << 8) | (uint32_t)(ox + 128); }

// The same set from the LDS tile, whose entries hold (offset key of the
// plate << 8) | plate: the source of offset j is entry base - del[j], and its
// plate arrives exactly when the plate's own offset key is key[j].
__device__ __forceinline__ int tt_arrivals_lds(const uint32_t *tileK, int base, const uint32_t *key, const int *del,
                                               int nd, tt_min4 &set)
{
    int n = 0;
    set.clear();
    for (int j = 0; j < nd; ++j) {
        const uint32_t e = tileK[base - del[j]];
        if ((e >> 8) == key[j]) {
            set.add(e & 0xFFu);
            ++n;
        }
    }
    return n;
}

__global__ __launch_bounds__(256) void k_advect(tt_dev d, const uint8_t *ids, const float *height,
                                                const tt_plate_step *ps, const tt_step_ctl *ctl,
                                                const uint8_t *typeBitmap, uint32_t *word, uint8_t *cnt,
                                                uint32_t *segCount, int *typeCounts, tt_step_ctl *ctlw)
{
    __shared__ uint32_t tileK[ADV_TH * ADV_TW];
    __shared__ int pox[256], poy[256];
    __shared__ int dox[256], doy[256];
    __shared__ uint32_t dkey[ADV_K * ADV_K];
    __shared__ int ddel[ADV_K * ADV_K];
    __shared__ uint8_t offUsed[ADV_K * ADV_K];
    __shared__ int nuse;
    const int tid = threadIdx.x;
    int nd = ctl->n_distinct;
    const int R = ctl->max_r;
    const int zeroDistinct = ctl->zero_distinct;
    pox[tid] = ps[tid].ox;
    poy[tid] = ps[tid].oy;
    for (int i = tid; i < ADV_K * ADV_K; i += 256) offUsed[i] = 0;
    if (tid == 0) nuse = 0;
    if (tid < nd) { dox[tid] = ctl->dox[tid]; doy[tid] = ctl->doy[tid]; }
    const int x0 = blockIdx.x * TT_TILE_W, y0 = blockIdx.y * TT_TILE_H;
    const bool lds = R <= ADV_R;
    __syncthreads();
    if (lds) {
        // Every plate's offset is within ADV_R here (R <= ADV_R), so its key
        // is defined.  A plate arrives at a cell of the tile only from a
        // source cell at its offset, which lies in the tile plus halo: only
        // the offsets of the plates present there need to be tried.  These
        // are among the distinct offsets of the plates 0..254; the "no plate"
        // entry 255 (offset (0, 0)) adds its offset only when that is one of
        // them.  The order of the list does not matter (the arrivals are a set).
        for (int i = tid; i < ADV_TH * ADV_TW; i += 256) {
            const int ly = i / ADV_TW, lx = i - ly * ADV_TW;
            // |x0 + lx - ADV_R| < 2W and |y0 + ly - ADV_R| < 2H: one conditional wrap
            const int gx = tt_wrap1(x0 + lx - ADV_R, d.W), gy = tt_wrap1(y0 + ly - ADV_R, d.H);
            const uint8_t p = ids[(size_t)gy * d.W + gx];
            const int ox = pox[p], oy = poy[p];
            tileK[i] = (tt_adv_key(ox, oy) << 8) | p;
            if (p != TT_MAXP || zeroDistinct) offUsed[(oy + ADV_R) * ADV_K + ox + ADV_R] = 1;
        }
        __syncthreads();
        for (int i = tid; i < ADV_K * ADV_K; i += 256)
            if (offUsed[i]) {
                const int ox = i % ADV_K - ADV_R, oy = i / ADV_K - ADV_R;
                const int m = atomicAdd(&nuse, 1);
                dkey[m] = tt_adv_key(ox, oy);
                ddel[m] = oy * ADV_TW + ox;
            }
    }
    __syncthreads();
    if (lds) nd = nuse;
    const int lane = tid & 63, wv = tid >> 6;
    const int segPerRow = d.W / 64;
    int overflow = 0;
    for (int k = 0; k < 4; ++k) {
        const int ly = wv * 4 + k;
        const int gx = x0 + lane, gy = y0 + ly;
        const size_t c = (size_t)gy * d.W + gx;
        tt_min4 set;
        const int n = lds ? tt_arrivals_lds(tileK, (ly + ADV_R) * ADV_TW + lane + ADV_R, dkey, ddel, nd, set)
                          : tt_arrivals(d, ids, gx, gy, dox, doy, nd, pox, poy, set);
        const unsigned sl[4] = {set.s0, set.s1, set.s2, set.s3};
        uint32_t wd = 0;
#pragma unroll
        for (int s = 0; s < 4; ++s)
            if (sl[s] < 256u) wd |= (0xFFu - sl[s]) << (8 * s);
        word[c] = wd;
        cnt[c] = (uint8_t)(n < 255 ? n : 255);
        if (n > 4) ++overflow;
        int sum = n;
        for (int off = 32; off > 0; off >>= 1) sum += __shfl_xor(sum, off, 64);
        if (lane == 0) segCount[(size_t)gy * segPerRow + blockIdx.x] = (uint32_t)sum;
        int key[6] = {-1, -1, -1, -1, -1, -1};
        if (wd > 0xFFu && ((wd >> 8) & 0xFFu) != 0u) {
            // B1: pair counts (determineCollisionType, plate_tectonics_kernel.cu:838-888)
            int P[4];
            for (int s = 0; s < 4; ++s) P[s] = tt_slot_plate(wd, s);
            float hh[4];
            for (int s = 0; s < 4; ++s)
                hh[s] = P[s] != TT_MAXP
                            ? height[(size_t)tt_wrapg(gy - poy[P[s]], d.H) * d.W + tt_wrapg(gx - pox[P[s]], d.W)]
                            : 0.0f;
            if (P[3] != TT_MAXP) {
                key[0] = tt_collision_key(d, typeBitmap, hh[0], hh[1], P[0], P[1]);
                key[1] = tt_collision_key(d, typeBitmap, hh[0], hh[2], P[0], P[2]);
                key[2] = tt_collision_key(d, typeBitmap, hh[0], hh[3], P[0], P[3]);
                key[3] = tt_collision_key(d, typeBitmap, hh[1], hh[2], P[1], P[2]);
                key[4] = tt_collision_key(d, typeBitmap, hh[1], hh[3], P[1], P[3]);
                key[5] = tt_collision_key(d, typeBitmap, hh[2], hh[3], P[2], P[3]);
            } else if (P[2] != TT_MAXP) {
                key[0] = tt_collision_key(d, typeBitmap, hh[0], hh[1], P[0], P[1]);
                key[1] = tt_collision_key(d, typeBitmap, hh[0], hh[2], P[0], P[2]);
                key[2] = tt_collision_key(d, typeBitmap, hh[1], hh[2], P[1], P[2]);
            } else {
                key[0] = tt_collision_key(d, typeBitmap, hh[0], hh[1], P[0], P[1]);
            }
        }
#pragma unroll
        for (int s = 0; s < 6; ++s) tt_wave_count(key[s], typeCounts);
    }
    if (overflow) atomicAdd(&ctlw->overflow_cells, overflow);
}

hipError_t tt_launch_advect(hipStream_t s, const tt_dev &d, const uint8_t *ids, const float *height,
                            const tt_plate_step *ps, const tt_step_ctl *ctl_dev, int max_r_host, const uint8_t *typeBitmap,
                            uint32_t *word, uint8_t *cnt, uint32_t *segCount, int *typeCounts, tt_step_ctl *ctl)
{
    (void)max_r_host;
    dim3 grid(d.W / TT_TILE_W, d.H / TT_TILE_H);
    hipLaunchKernelGGL(k_advect, grid, dim3(256), 0, s, d, ids, height, ps, ctl_dev, typeBitmap, word, cnt, segCount,
                       typeCounts, ctl);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// sorted keys: one wavefront per 64-cell row segment; the exclusive prefix of
// the arrival counts inside the segment comes from a wavefront scan.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_keys(tt_dev d, const uint32_t *word, const uint8_t *cnt,
                                              const uint32_t *segBase, const uint8_t *ids, const tt_plate_step *ps,
                                              const tt_step_ctl *ctl, unsigned long long *keys)
{
    const int lane = threadIdx.x & 63;
    const long long seg = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
    const long long c = seg * 64 + lane;
    if (c >= d.N) return;
    const uint32_t wd = word[c];
    const int n = cnt[c];
    int incl = n;
    for (int off = 1; off < 64; off <<= 1) {
        const int v = __shfl_up(incl, off, 64);
        if (lane >= off) incl += v;
    }
    const unsigned long long base = (unsigned long long)segBase[seg] + (unsigned long long)(incl - n);
    const unsigned long long dk = (unsigned long long)c << 8;
    if (n <= 4) {
        for (int s = 0; s < n; ++s) keys[base + s] = dk | (unsigned long long)(uint8_t)(0xFFu - ((wd >> (8 * s)) & 0xFFu));
        return;
    }
    // more than four plates: enumerate all arrivals in ascending order
    int gx, gy;
    tt_cell_xy((uint32_t)c, d, gx, gy);
    const int nd = ctl->n_distinct;
    int last = -1;
    for (int s = 0; s < n; ++s) {
        int best = 256;
        for (int j = 0; j < nd; ++j) {
            const int ox = ctl->dox[j], oy = ctl->doy[j];
            const int p = ids[(size_t)tt_wrapg(gy - oy, d.H) * d.W + tt_wrapg(gx - ox, d.W)];
            if (ps[p].ox == ox && ps[p].oy == oy && p > last && p < best) best = p;
        }
        keys[base + s] = dk | (unsigned long long)best;
        last = best;
    }
}

hipError_t tt_launch_keys(hipStream_t s, const tt_dev &d, const uint32_t *word, const uint8_t *cnt,
                          const uint32_t *segBase, const uint8_t *ids, const tt_plate_step *ps, const tt_step_ctl *ctl,
                          unsigned long long *keys)
{
    const long long segs = (long long)d.N / 64;
    hipLaunchKernelGGL(k_keys, dim3((unsigned)((segs + 3) / 4)), dim3(256), 0, s, d, word, cnt, segBase, ids, ps, ctl,
                       keys);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// B2: createCollisionTypeMatrix (plate_tectonics_kernel.cu:890-972); the pair
// decode via double sqrt (:900-905) is unused there and omitted.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_type_matrix(tt_dev d, const int *counts, uint8_t *bitmap)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= TT_NUM_TRI) return;
    const unsigned packed = bitmap[i];
    const bool hasType = packed & 1u;
    const unsigned count = (packed >> 4) & 0x0Fu;
    const int sAC = counts[i * 5 + 0], sAO = counts[i * 5 + 1], sBC = counts[i * 5 + 2], sBO = counts[i * 5 + 3],
              cont = counts[i * 5 + 4];
    const uint32_t total = (uint32_t)sAC + (uint32_t)sAO + (uint32_t)sBC + (uint32_t)sBO + (uint32_t)cont;
    if (total == 0) return;
    if (hasType && (int)count < d.ks.collisionTypeUpdateCooldown) {
        bitmap[i] = (uint8_t)((packed & 0x0Fu) | (((count + 1u) & 0x0Fu) << 4));
        return;
    }
    if (total < (uint32_t)d.ks.minCollisionSize) return;
    unsigned out = packed & 0x0Fu;
    const bool hasPol = (packed >> 2) & 1u;
    const bool polA = (packed >> 3) & 1u;
    const int contOppA = sBC + cont, contOppB = sAC + cont;
    const int powA = sAC + sAO - contOppA, powB = sBC + sBO - contOppB;
    if (cont >= sAC + sAO && cont >= sBC + sBO) out |= 0x03u;
    else if (powA > 0 && powA > powB && (!hasPol || polA)) out = (out & ~0x02u) | 0x0Du;
    else if (powB > 0 && powB > powA && (!hasPol || !polA)) out = (out & ~0x0Au) | 0x05u;
    bitmap[i] = (uint8_t)out;
}

hipError_t tt_launch_type_matrix(hipStream_t s, const tt_dev &d, const int *typeCounts, uint8_t *typeBitmap)
{
    hipLaunchKernelGGL(k_type_matrix, dim3((TT_NUM_TRI + 255) / 256), dim3(256), 0, s, d, typeCounts, typeBitmap);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// B3: processCollisions
// ---------------------------------------------------------------------------
__device__ __forceinline__ unsigned tt_det_type(const tt_dev &d, const uint8_t *bitmap, int pa, int pb, float ha,
                                                float hb)
{
    const bool cond = pa < pb;
    const int mn = cond ? pa : pb, mx = cond ? pb : pa;
    const float hmn = cond ? ha : hb, hmx = cond ? hb : ha;
    const unsigned packed = bitmap[tt_tri((unsigned)mn, (unsigned)mx)];
    const bool hasType = packed & 1u, type = (packed >> 1) & 1u, pol = (packed >> 3) & 1u;
    unsigned out = 0;
    const float thr = d.ks.continentalCrustThreshold;
    if (hasType && !type) {
        if (pol) out = hmn >= thr ? (mn == pa ? 3u : 4u) : (mn == pa ? 1u : 2u);
        else out = hmx >= thr ? (mx == pa ? 3u : 4u) : (mx == pa ? 1u : 2u);
    }
    return out;
}

__global__ __launch_bounds__(256) void k_resolve(tt_dev d, const uint32_t *iterp, const uint32_t *word, const uint8_t *ids,
                                                 const float *height, const tt_plate_step *ps,
                                                 const uint32_t *divBitmap, const uint8_t *typeBitmap, float *hW,
                                                 uint8_t *idW, tt_uplift *uplift, uint32_t *accw, uint32_t *hasDiv,
                                                 long long *fric, uint8_t *srcFlag, uint8_t *contested, int do_uplift)
{
    __shared__ unsigned long long sf[256];
    __shared__ tt_plate_step sps[256];
    const int tid = threadIdx.x;
    sf[tid] = 0ull;
    sps[tid] = ps[tid];
    // the four cells' collision words are loaded together
    uint32_t cvs[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const long long c = (long long)blockIdx.x * 1024 + k * 256 + tid;
        cvs[k] = c < d.N ? word[c] : 0u;
    }
    __syncthreads();
    ps = sps;
    const int W = d.W, H = d.H;
    for (int k = 0; k < 4; ++k) {
        const long long c = (long long)blockIdx.x * 1024 + k * 256 + tid;
        if (c >= d.N) break;
        int x, y;
        tt_cell_xy((uint32_t)c, d, x, y);
        const uint32_t cv = cvs[k];
        if (cv == 0) {
            // divergence (processDivergence, plate_tectonics_kernel.cu:301-362)
            const int prev = ids[c];
            const tt_plate_step pp = ps[prev];
            const uint32_t ov = word[(size_t)tt_wrap_near(y - pp.oy, H) * W + tt_wrap_near(x - pp.ox, W)];
            const int opp = (int)(uint8_t)((TT_MAXP - ov) & 0xFFu);
            int nid = prev;
            if (ov != 0 && ov <= 0xFFu && opp != prev && prev != TT_MAXP) {
                const int mn = prev < opp ? prev : opp, mx = prev < opp ? opp : prev;
                const unsigned bit = tt_tri((unsigned)mn, (unsigned)mx);
                nid = (divBitmap[bit >> 5] & (1u << (bit & 31u))) ? mn : mx;
                atomicOr(&hasDiv[bit >> 5], 1u << (bit & 31u));
            }
            const float nz = tt_simplex((float)x, (float)y, (float)*iterp, 0.01f, d.seed);
            const float target = d.ks.divergence_height_target_min +
                                 nz * (d.ks.divergence_height_target_max - d.ks.divergence_height_target_min);
            hW[c] = tt_lerp(height[c], target, d.ks.divergence_interpolation_factor);
            idW[c] = (uint8_t)nid;
            if (do_uplift) srcFlag[c] = 0;
            continue;
        }
        int P[4];
        for (int s = 0; s < 4; ++s) P[s] = tt_slot_plate(cv, s);
        if (P[1] == TT_MAXP) {
            // movement (processMovement, :586-609)
            const tt_plate_step pp = ps[P[0]];
            idW[c] = (uint8_t)P[0];
            hW[c] = height[(size_t)tt_wrap_near(y - pp.oy, H) * W + tt_wrap_near(x - pp.ox, W)];
            if (do_uplift) srcFlag[c] = 0;
            continue;
        }
        // convergence
        tt_plate_step PD[4];
        for (int s = 0; s < 4; ++s) PD[s] = ps[P[s]];  // ps[255] is the default plate
        float prevH[4];
        prevH[0] = height[(size_t)tt_wrap_near(y - PD[0].oy, H) * W + tt_wrap_near(x - PD[0].ox, W)];
        prevH[1] = height[(size_t)tt_wrap_near(y - PD[1].oy, H) * W + tt_wrap_near(x - PD[1].ox, W)];
        prevH[2] = P[2] != TT_MAXP ? height[(size_t)tt_wrap_near(y - PD[2].oy, H) * W + tt_wrap_near(x - PD[2].ox, W)] : 0.0f;
        prevH[3] = P[3] != TT_MAXP ? height[(size_t)tt_wrap_near(y - PD[3].oy, H) * W + tt_wrap_near(x - PD[3].ox, W)] : 0.0f;
        unsigned types[6] = {0, 0, 0, 0, 0, 0};
        types[0] = tt_det_type(d, typeBitmap, P[0], P[1], prevH[0], prevH[1]);
        if (P[2] != TT_MAXP) {
            types[1] = tt_det_type(d, typeBitmap, P[0], P[2], prevH[0], prevH[2]);
            types[2] = tt_det_type(d, typeBitmap, P[1], P[2], prevH[1], prevH[2]);
            if (P[3] != TT_MAXP) {
                types[3] = tt_det_type(d, typeBitmap, P[0], P[3], prevH[0], prevH[3]);
                types[4] = tt_det_type(d, typeBitmap, P[1], P[3], prevH[1], prevH[3]);
                types[5] = tt_det_type(d, typeBitmap, P[2], P[3], prevH[2], prevH[3]);
            }
        }
        const int nColl = P[3] != TT_MAXP ? 4 : P[2] != TT_MAXP ? 3 : 2;
        // friction part of applyInelasticCollision (:397-467 -> :371-394)
        unsigned contB = 0, accB = 0;
#pragma unroll
        for (int j = 1; j < 4; ++j)
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                if (i >= j || j >= nColl) continue;
                const unsigned t = types[(j * (j - 1)) / 2 + i];
                contB |= (unsigned)(t == 0) << i;
                accB |= (unsigned)(t == 3 || t == 4) << i;
                contB |= (unsigned)(t == 0) << j;
                accB |= (unsigned)(t == 3 || t == 4) << j;
            }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            if (q >= nColl) continue;
            const float massDiff = prevH[q] / PD[q].mass;
            const bool hc = (contB >> q) & 1u, ha = (accB >> q) & 1u;
            const float coef = hc ? d.ks.frictionCoefficientContinental
                                  : ha ? d.ks.frictionCoefficientAccretion : d.ks.frictionCoefficientSubduction;
            const float f = (PD[q].vel * coef) * massDiff;
            atomicAdd(&sf[P[q]], (unsigned long long)tt_fix(f, TT_FRIC_BITS));
        }
        // processConvergence (:469-584)
        const int pairLen = nColl == 4 ? 6 : nColl == 3 ? 3 : 1;
        const int pa_[6] = {0, 0, 1, 0, 1, 2};
        const int pb_[6] = {1, 2, 2, 3, 3, 3};
        bool sub[4] = {false, false, false, false};
#pragma unroll
        for (int a = 0; a < 6; ++a) {
            if (a >= pairLen) continue;
            const unsigned t = types[a];
            if (t == 1 || t == 3) sub[pa_[a]] = true;
            else if (t == 2 || t == 4) sub[pb_[a]] = true;
        }
        int own = 0;
#pragma unroll
        for (int i = 0; i < 4; ++i)
            if (i < nColl && !sub[i]) own = i;
        const int newOwner = own == 0 ? P[0] : own == 1 ? P[1] : own == 2 ? P[2] : P[3];
        float writeH = own == 0 ? prevH[0] : own == 1 ? prevH[1] : own == 2 ? prevH[2] : prevH[3];
        float add = 0.0f;
        unsigned typeByte = 0;
#pragma unroll
        for (int a = 0; a < 6; ++a) {
            if (a >= pairLen) continue;
            const int ia = pa_[a], ib = pb_[a];
            float h;
            if (ia == own) h = prevH[ib];
            else if (ib == own) h = prevH[ia];
            else continue;
            const unsigned t = types[a];
            if (h > writeH) { const float tmp = h; h = writeH; writeH = tmp; }
            add = add + h;
            // accretion marks: canonical "highest writer cell wins"
            if (ia == own && t == 4) {
                const int bx = x - PD[ib].ox, by = y - PD[ib].oy;
                size_t tgt;
                if (bx != x || by != y) tgt = (size_t)tt_wrap_near(by, H) * W + tt_wrap_near(bx, W);
                else tgt = (size_t)tt_wrap_near(y + PD[ia].oy, H) * W + tt_wrap_near(x + PD[ia].ox, W);
                atomicMax(&accw[tgt], (uint32_t)(c + 1));
            } else if (ib == own && t == 3) {
                const int ax_ = x - PD[ia].ox, ay_ = y - PD[ia].oy;
                size_t tgt;
                if (ax_ != x || ay_ != y) tgt = (size_t)tt_wrap_near(ay_, H) * W + tt_wrap_near(ax_, W);
                else tgt = (size_t)tt_wrap_near(y + PD[ib].oy, H) * W + tt_wrap_near(x + PD[ib].ox, W);
                atomicMax(&accw[tgt], (uint32_t)(c + 1));
            }
            if (ia == own) {
                typeByte |= (t == 2 ? 0u : 1u);
                typeByte |= (unsigned)(ia & 3) << 1;
                typeByte |= (unsigned)(ib & 3) << 3;
            } else {
                typeByte |= (t == 1 ? 0u : 1u);
                typeByte |= (unsigned)(ib & 3) << 1;
                typeByte |= (unsigned)(ia & 3) << 3;
            }
        }
        hW[c] = writeH;
        tt_uplift U;
        U.cum = add;
        U.types = (uint8_t)typeByte;
        U.pad[0] = U.pad[1] = U.pad[2] = 0;
        uplift[c] = U;
        idW[c] = (uint8_t)newOwner;
        if (!do_uplift) continue;
        // classify continental-propagation sources (moveContinentalCollisions,
        // :1087-1185): "simple" sources touch only their own cell
        if (typeByte & 1u) {
            const int prim = tt_slot_plate(cv, (int)((typeByte >> 1) & 3u));
            const float hc = height[c];
            bool remote = false;
            if (hc >= d.ks.continentalPropThreshold) {
                int tx, ty;
                tt_cont_prop(d, height, x, y, ps[prim].dirx, ps[prim].diry, ps[prim].pcx, ps[prim].pcy, hc, tx, ty);
                if (tx != x || ty != y) { remote = true; contested[(size_t)tt_wrap_near(ty, H) * W + tt_wrap_near(tx, W)] = 1; }
                for (int ci = 0; ci < 4; ++ci) {
                    const int sec = tt_slot_plate(cv, ci);
                    if (sec == TT_MAXP || sec == prim) continue;
                    tt_cont_prop(d, height, x, y, ps[sec].dirx, ps[sec].diry, ps[sec].pcx, ps[sec].pcy, hc, tx, ty);
                    if (tx != x || ty != y) { remote = true; contested[(size_t)tt_wrap_near(ty, H) * W + tt_wrap_near(tx, W)] = 1; }
                }
            }
            srcFlag[c] = remote ? 2 : 1;
            if (remote) contested[c] = 1;
        } else {
            srcFlag[c] = 0;
        }
    }
    __syncthreads();
    if (sf[tid] != 0ull) atomicAdd((unsigned long long *)&fric[tid], sf[tid]);
}

hipError_t tt_launch_resolve(hipStream_t s, const tt_dev &d, const uint32_t *iterp, const uint32_t *word, const uint8_t *ids,
                             const float *height, const tt_plate_step *ps, const uint32_t *divBitmap,
                             const uint8_t *typeBitmap, float *hW, uint8_t *idW, tt_uplift *uplift, uint32_t *accw,
                             uint32_t *hasDiv, long long *fric, uint8_t *srcFlag, uint8_t *contested, int do_uplift)
{
    const unsigned blocks = (unsigned)(((long long)d.N + 1023) / 1024);
    hipLaunchKernelGGL(k_resolve, dim3(blocks), dim3(256), 0, s, d, iterp, word, ids, height, ps, divBitmap, typeBitmap, hW,
                       idW, uplift, accw, hasDiv, fric, srcFlag, contested, do_uplift);
    return hipGetLastError();
}

__global__ void k_flip(uint32_t *divBitmap, const uint32_t *hasDiv)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < TT_NUM_TRI_WORDS) divBitmap[i] ^= hasDiv[i];
}

hipError_t tt_launch_flip(hipStream_t s, uint32_t *divBitmap, const uint32_t *hasDiv)
{
    hipLaunchKernelGGL(k_flip, dim3((TT_NUM_TRI_WORDS + 255) / 256), dim3(256), 0, s, divBitmap, hasDiv);
    return hipGetLastError();
}
