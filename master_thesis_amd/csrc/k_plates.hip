// k_plates.hip -- stages F and G on gfx950: plate kinematics, accretion, CCL of
// plate masks with relabelling, merging, per-plate sums and plate splitting.
//
// Reference: applyPlateMovementChanges (plate_tectonics_kernel.cu:1863-1925),
// applyAccretion (:747-770), ccl.cu:35-177 with applyCCL
// (plate_tectonic_sim.cu:375-468), createLabelIdLookup / postCClIdReassign /
// assignUnassignedIdsToNeighbor (:3130-3206, :2052-2105), mergePlates
// (plate_tectonic_sim.cu:470-494, :1283-1392, :3208-3225, :102-197),
// computeBreakScore (:2706-2733), accumulatePlateAngularCoords /
// calculatePlateCenters (:2154-2226), statisticsPass (:2377-2402) and
// SplitPlateV2 (plate_tectonic_sim.cu:933-1040, :2824-3112, :3227-3247).
//
// Float sums that feed integer decisions are exact int64 fixed-point sums
// (order-free), so every grid decomposition gives the serial result.
#include "tt_kernels.h"
#include "tt_stage.h"

// ---------------------------------------------------------------------------
// G1: applyPlateMovementChanges
// ---------------------------------------------------------------------------
__global__ void k_plate_movement(tt_dev d, tt_plate_data *pd, const long long *fric, const uint8_t *noise)
{
    const int p = threadIdx.x;
    if (p >= TT_MAXP) return;
    const tt_kernel_settings &K = d.ks;
    tt_plate_data cur = pd[p];
    const float cx = cur.pixelCenter[0] + cur.direction[0] * cur.velocity;
    const float cy = cur.pixelCenter[1] + cur.direction[1] * cur.velocity;
    const float nx = cx - tt_floorf(cx), ny = cy - tt_floorf(cy);
    cur.hasMoved = (nx != cx || ny != cy) ? 1 : 0;
    cur.pixelCenter[0] = nx;
    cur.pixelCenter[1] = ny;
    const float frictionLoss = tt_unfix(fric[p], TT_FRIC_BITS);
    const float ovx = cur.direction[0] * cur.velocity, ovy = cur.direction[1] * cur.velocity;
    const int sx = (int)(cur.geometricCenter[0] + 0.5f), sy = (int)(cur.geometricCenter[1] + 0.5f);
    float asx, asy;
    tt_asth_at(noise, tt_wrapg(sx, d.W), tt_wrapg(sy, d.H), d.axs, d.ays, asx, asy);
    const float avx = tt_fmaxf(tt_fminf((float)((double)(asx * K.curlNoiseMultiplier) + (asx < 0 ? -0.2 : 0.2)), 1.0f), -1.0f);
    const float avy = tt_fmaxf(tt_fminf((float)((double)(asy * K.curlNoiseMultiplier) + (asy < 0 ? -0.2 : 0.2)), 1.0f), -1.0f);
    const float m0 = tt_fmaxf(sqrtf(ovx * ovx + ovy * ovy), 1e-20f);
    const float post = tt_fmaxf(m0 - frictionLoss, 0.0f);
    const float pvx = cur.direction[0] * post, pvy = cur.direction[1] * post;
    const float rx = avx - pvx, ry = avy - pvy;
    const float rm = sqrtf(rx * rx + ry * ry);
    const float dd = (K.environmentalDragCoefficient * rm) * rm;
    const float rmz = tt_fmaxf(sqrtf(rx * rx + ry * ry), 1e-20f);
    const float rnx = rx / rmz, rny = ry / rmz;
    const float nvx = pvx + rnx * dd, nvy = pvy + rny * dd;
    const float nm = tt_fmaxf(sqrtf(nvx * nvx + nvy * nvy), 1e-20f);
    cur.velocity = nm;
    cur.direction[0] = nvx / nm;
    cur.direction[1] = nvy / nm;
    cur.velocitySmoothed = cur.velocitySmoothed * K.velocitySmoothingFactor + cur.velocity * (1.0f - K.velocitySmoothingFactor);
    const float sdx = cur.directionSmoothed[0] * K.directionSmoothingFactor + cur.direction[0] * (1.0f - K.directionSmoothingFactor);
    const float sdy = cur.directionSmoothed[1] * K.directionSmoothingFactor + cur.direction[1] * (1.0f - K.directionSmoothingFactor);
    const float sm = tt_fmaxf(sqrtf(sdx * sdx + sdy * sdy), 1e-20f);
    cur.directionSmoothed[0] = sdx / sm;
    cur.directionSmoothed[1] = sdy / sm;
    cur.mergeWaitTime = cur.mergeWaitTime > 0 ? cur.mergeWaitTime - 1 : 0;
    pd[p] = cur;
}

hipError_t tt_launch_plate_movement(hipStream_t s, const tt_dev &d, tt_plate_data *pd, const long long *fric,
                                    const uint8_t *noise)
{
    hipLaunchKernelGGL(k_plate_movement, dim3(1), dim3(256), 0, s, d, pd, fric, noise);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// B5: applyAccretion.  accw holds (highest writer cell + 1); the value the
// writer recorded is MAX_PLATE_COUNT - its new owner, i.e. the accreted id is
// the writer's owner after resolution.
// ---------------------------------------------------------------------------
// Four cells per thread (N is a multiple of 4, the fields 16-byte aligned):
// the word, id and accretion accesses are 16- and 4-byte vectors.
__global__ __launch_bounds__(256) void k_accretion_apply(tt_dev d, const uint32_t *accw, const float *height,
                                                         const uint8_t *idsIn, uint8_t *idsOut, uint8_t *accretion)
{
    const long long c = ((long long)blockIdx.x * 256 + threadIdx.x) * 4;
    if (c >= d.N) return;
    const uint4 a4 = *reinterpret_cast<const uint4 *>(accw + c);
    const uchar4 i4 = *reinterpret_cast<const uchar4 *>(idsIn + c);
    const uint32_t a[4] = {a4.x, a4.y, a4.z, a4.w};
    uint8_t id[4] = {i4.x, i4.y, i4.z, i4.w}, val[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        val[k] = 0;
        if (a[k] != 0u) {
            const uint8_t owner = idsIn[a[k] - 1u];
            val[k] = (uint8_t)(TT_MAXP - owner);
            const uint8_t aid = (uint8_t)(TT_MAXP - val[k]);
            if (aid != TT_MAXP && height[c + k] >= d.ks.continentalCrustThreshold) id[k] = aid;
        }
    }
    *reinterpret_cast<uchar4 *>(idsOut + c) = make_uchar4(id[0], id[1], id[2], id[3]);
    *reinterpret_cast<uchar4 *>(accretion + c) = make_uchar4(val[0], val[1], val[2], val[3]);
}

hipError_t tt_launch_accretion_apply(hipStream_t s, const tt_dev &d, const uint32_t *accw, const float *height,
                                     const uint8_t *idsIn, uint8_t *idsOut, uint8_t *accretion)
{
    if (d.N & 3) return hipErrorInvalidValue;
    const unsigned blocks = (unsigned)(((long long)d.N / 4 + 255) / 256);
    hipLaunchKernelGGL(k_accretion_apply, dim3(blocks), dim3(256), 0, s, d, accw, height, idsIn, idsOut, accretion);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// F1-F4: connected components of the reference's edge set.  For cell c with
// neighbours tl, t, tr, l (wrapped): the first same-id neighbour in that order
// (init, ccl.cu:107-116), plus tl and l, plus t and tr on row 0 (reduce,
// ccl.cu:161-176).  Label = minimum cell index of the component.
// ---------------------------------------------------------------------------
// edge bits: 1 = tl, 2 = t, 4 = tr, 8 = l
__device__ __forceinline__ unsigned tt_ccl_edges(uint8_t id, uint8_t tl, uint8_t t, uint8_t tr, uint8_t l, bool row0)
{
    unsigned m = 0;
    if (id == tl) m |= 1u;
    else if (id == t) m |= 2u;
    else if (id == tr) m |= 4u;
    else if (id == l) m |= 8u;
    if (id == tl) m |= 1u;
    if (id == l) m |= 8u;
    if (row0) {
        if (id == t) m |= 2u;
        if (id == tr) m |= 4u;
    }
    return m;
}

// union-find in LDS (32-bit parents), linking the larger root under the
// smaller with atomicMin as in unionStep (ccl.cu:60-85).  A parent is never
// larger than its child and always lies in the child's component, so the root
// of a component is its smallest cell whatever the order of the links.
//
// Find with path halving: each visited node is pointed at its grandparent, an
// ancestor in the same component, so chains stay inside their components and
// shrink.  A concurrent atomicMin on the node (by a thread that found it as a
// root before it was linked) may be overwritten; that thread then retries its
// union from the node's old parent, so the link is made there.  Without
// halving, a row of one id links into a chain of its whole length and the
// tile's finds walk chains of up to a thousand cells.
__device__ __forceinline__ unsigned tt_find32h(unsigned *par, unsigned a)
{
    for (;;) {
        const unsigned p = par[a];
        if (p == a) return a;
        const unsigned gp = par[p];
        if (gp == p) return p;
        par[a] = gp;
        a = gp;
    }
}

__device__ __forceinline__ void tt_union32(unsigned *par, unsigned a, unsigned b)
{
    bool done = false;
    while (!done) {
        a = tt_find32h(par, a);
        b = tt_find32h(par, b);
        if (a < b) {
            const unsigned old = atomicMin(&par[b], a);
            done = old == b;
            b = old;
        } else if (b < a) {
            const unsigned old = atomicMin(&par[a], b);
            done = old == a;
            a = old;
        } else {
            done = true;
        }
    }
}

// Local labelling of one 64x16 tile.  labels[c] = global index of the
// smallest cell of c's component inside the tile (its local root);
// rootFlag[c] = 1 at local roots; rootCount[r] = cells of local root r.  A tile
// holding one plate id is one component (every row is chained by l edges and
// rows are joined by tl edges), rooted at its first cell; other tiles run an
// LDS union-find over the reference's edge set.
__global__ __launch_bounds__(256) void k_ccl_local(tt_dev d, const uint8_t *ids, uint32_t *labels, uint8_t *rootFlag,
                                                   uint32_t *rootCount)
{
    __shared__ unsigned par[TT_TILE_H * TT_TILE_W];
    __shared__ unsigned cnt[TT_TILE_H * TT_TILE_W];
    __shared__ uint8_t tid8[(TT_TILE_H + 1) * (TT_TILE_W + 2)];
    const int tid = threadIdx.x;
    const int x0 = blockIdx.x * TT_TILE_W, y0 = blockIdx.y * TT_TILE_H;
    const int W = d.W;
    // fast path: one plate id over the whole tile.  Each thread takes 4
    // consecutive cells of a row (one 4-byte load, one 16-byte label store and
    // one 4-byte flag store).
    {
        const int fr = tid >> 4, fc = (tid & 15) * 4;
        const size_t c4 = (size_t)(y0 + fr) * W + x0 + fc;
        const uint32_t own = *reinterpret_cast<const uint32_t *>(ids + c4);
        __shared__ uint32_t first;
        if (tid == 0) first = (own & 0xFFu) * 0x01010101u;
        __syncthreads();
        if (__syncthreads_and(own == first)) {
            const uint32_t root = (uint32_t)((size_t)y0 * W + x0);
            *reinterpret_cast<uint4 *>(labels + c4) = make_uint4(root, root, root, root);
            *reinterpret_cast<uint32_t *>(rootFlag + c4) = tid == 0 ? 1u : 0u;
            if (tid == 0) rootCount[root] = TT_TILE_W * TT_TILE_H;
            return;
        }
    }
    const int TW2 = TT_TILE_W + 2;
    for (int i = tid; i < (TT_TILE_H + 1) * TW2; i += 256) {
        const int ly = i / TW2, lx = i - ly * TW2;
        const int gx = tt_wrap1(x0 + lx - 1, d.W), gy = tt_wrap1(y0 + ly - 1, d.H);
        tid8[i] = ids[(size_t)gy * d.W + gx];
    }
    for (int i = tid; i < TT_TILE_H * TT_TILE_W; i += 256) {
        par[i] = (unsigned)i;
        cnt[i] = 0u;
    }
    __syncthreads();
    for (int i = tid; i < TT_TILE_H * TT_TILE_W; i += 256) {
        const int ly = i / TT_TILE_W, lx = i - ly * TT_TILE_W;
        const int r = ly + 1, cc = lx + 1;
        const uint8_t id = tid8[r * TW2 + cc];
        const unsigned m = tt_ccl_edges(id, tid8[(r - 1) * TW2 + cc - 1], tid8[(r - 1) * TW2 + cc],
                                        tid8[(r - 1) * TW2 + cc + 1], tid8[r * TW2 + cc - 1], (y0 + ly) == 0);
        if ((m & 1u) && ly > 0 && lx > 0) tt_union32(par, (unsigned)i, (unsigned)(i - TT_TILE_W - 1));
        if ((m & 2u) && ly > 0) tt_union32(par, (unsigned)i, (unsigned)(i - TT_TILE_W));
        if ((m & 4u) && ly > 0 && lx < TT_TILE_W - 1) tt_union32(par, (unsigned)i, (unsigned)(i - TT_TILE_W + 1));
        if ((m & 8u) && lx > 0) tt_union32(par, (unsigned)i, (unsigned)(i - 1));
    }
    __syncthreads();
    for (int i = tid; i < TT_TILE_H * TT_TILE_W; i += 256) {
        const unsigned r = tt_find32h(par, (unsigned)i);
        atomicAdd(&cnt[r], 1u);
        const int ry = (int)(r / TT_TILE_W), rx = (int)(r % TT_TILE_W);
        const int ly = i / TT_TILE_W, lx = i - ly * TT_TILE_W;
        const size_t c = (size_t)(y0 + ly) * d.W + x0 + lx;
        labels[c] = (uint32_t)((size_t)(y0 + ry) * d.W + x0 + rx);
        rootFlag[c] = (uint8_t)(r == (unsigned)i ? 1 : 0);
    }
    __syncthreads();
    for (int i = tid; i < TT_TILE_H * TT_TILE_W; i += 256)
        if (cnt[i] != 0u) {
            const int ly = i / TT_TILE_W, lx = i - ly * TT_TILE_W;
            rootCount[(size_t)(y0 + ly) * d.W + x0 + lx] = cnt[i];
        }
}

hipError_t tt_launch_ccl_local(hipStream_t s, const tt_dev &d, const uint8_t *ids, uint32_t *labels, uint8_t *rootFlag,
                               uint32_t *rootCount)
{
    hipLaunchKernelGGL(k_ccl_local, dim3(d.W / TT_TILE_W, d.H / TT_TILE_H), dim3(256), 0, s, d, ids, labels, rootFlag,
                       rootCount);
    return hipGetLastError();
}

// global union-find over local roots (labels[] of a local root is its parent;
// the parent of a root is itself; atomicMin links the larger root under the
// smaller as in unionStep, ccl.cu:60-85)
__device__ __forceinline__ uint32_t tt_gfind(const uint32_t *lab, uint32_t a)
{
    uint32_t p = lab[a];
    while (p != a) { a = p; p = lab[a]; }
    return a;
}

__device__ __forceinline__ void tt_gunion(uint32_t *lab, uint32_t a, uint32_t b)
{
    bool done = false;
    while (!done) {
        a = tt_gfind(lab, a);
        b = tt_gfind(lab, b);
        if (a < b) {
            const uint32_t old = atomicMin(&lab[b], a);
            done = old == b;
            b = old;
        } else if (b < a) {
            const uint32_t old = atomicMin(&lab[a], b);
            done = old == a;
            a = old;
        } else {
            done = true;
        }
    }
}

// Edges that cross tile borders join local roots: for each tile its top row
// (tl, t, tr), left column (tl, l) and right column (tr).  128 threads per tile
// cover the 94 border cells; a lane skips a pair of roots its left neighbour
// lane already submitted.
__global__ __launch_bounds__(128) void k_ccl_merge(tt_dev d, const uint8_t *ids, uint32_t *labels)
{
    const int tid = threadIdx.x;
    const int x0 = blockIdx.x * TT_TILE_W, y0 = blockIdx.y * TT_TILE_H;
    int lx, ly;
    if (tid < 64) { lx = tid; ly = 0; }
    else if (tid < 79) { lx = 0; ly = tid - 63; }
    else if (tid < 94) { lx = TT_TILE_W - 1; ly = tid - 78; }
    else { lx = -1; ly = -1; }
    const int W = d.W, H = d.H;
    uint32_t pa[4], pb[4];
    int np = 0;
    if (lx >= 0) {
        const int x = x0 + lx, y = y0 + ly;
        const int xl = tt_wrap1(x - 1, W), xr = tt_wrap1(x + 1, W), yu = tt_wrap1(y - 1, H);
        const uint32_t c = (uint32_t)y * W + x;
        const uint8_t id = ids[c];
        const uint32_t ntl = (uint32_t)yu * W + xl, nt = (uint32_t)yu * W + x, ntr = (uint32_t)yu * W + xr;
        const uint32_t nl = (uint32_t)y * W + xl;
        const unsigned m = tt_ccl_edges(id, ids[ntl], ids[nt], ids[ntr], ids[nl], y == 0);
        const uint32_t lc = labels[c];
        if ((m & 1u) && (ly == 0 || lx == 0)) { pa[np] = lc; pb[np] = labels[ntl]; ++np; }
        if ((m & 2u) && ly == 0) { pa[np] = lc; pb[np] = labels[nt]; ++np; }
        if ((m & 4u) && (ly == 0 || lx == TT_TILE_W - 1)) { pa[np] = lc; pb[np] = labels[ntr]; ++np; }
        if ((m & 8u) && lx == 0) { pa[np] = lc; pb[np] = labels[nl]; ++np; }
    }
    // first pair of the left lane, for de-duplication along the top row
    const uint32_t la = __shfl_up(np > 0 ? pa[0] : 0xFFFFFFFFu, 1, 64);
    const uint32_t lb = __shfl_up(np > 0 ? pb[0] : 0xFFFFFFFFu, 1, 64);
    for (int k = 0; k < np; ++k) {
        if ((tid & 63) != 0 && pa[k] == la && pb[k] == lb) continue;
        if (pa[k] != pb[k]) tt_gunion(labels, pa[k], pb[k]);
    }
}

hipError_t tt_launch_ccl_merge(hipStream_t s, const tt_dev &d, const uint8_t *ids, uint32_t *labels)
{
    hipLaunchKernelGGL(k_ccl_merge, dim3(d.W / TT_TILE_W, d.H / TT_TILE_H), dim3(128), 0, s, d, ids, labels);
    return hipGetLastError();
}

// Final roots of the local roots by pointer jumping; a local root that is not
// final adds its cell count to its final root and clears its root flag.
// Grid-stride over the local-root list (its length is read on the device).
__global__ __launch_bounds__(256) void k_ccl_flatten(const uint32_t *locList, const uint32_t *nloc, uint32_t *labels,
                                                     uint32_t *rootCount, uint8_t *rootFlag)
{
    const uint32_t n = *nloc;
    for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        const uint32_t r = locList[i];
        uint32_t p = labels[r];
        if (p == r) continue;
        for (;;) {
            const uint32_t gp = labels[p];
            if (gp == p) break;
            labels[r] = gp;
            p = gp;
        }
        labels[r] = p;
        atomicAdd(&rootCount[p], rootCount[r]);
        rootFlag[r] = 0;
    }
}

hipError_t tt_launch_ccl_flatten(hipStream_t s, const uint32_t *locList, const uint32_t *nloc, uint32_t *labels,
                                 uint32_t *rootCount, uint8_t *rootFlag)
{
    hipLaunchKernelGGL(k_ccl_flatten, dim3(1024), dim3(256), 0, s, locList, nloc, labels, rootCount, rootFlag);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// F5: relabel.  Fragment key sorts ascending as (count desc, label desc).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_frag_build(tt_dev d, const uint32_t *rootList, const uint32_t *nroots,
                                                    const uint32_t *rootCount, const uint8_t *ids,
                                                    unsigned long long *fragKey, int *maxCount, uint8_t *newId)
{
    const uint32_t n = *nroots;
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const uint32_t r = rootList[i];
    const uint32_t cnt = rootCount[r];
    fragKey[i] = ((unsigned long long)(0xFFFFFFFFu - cnt) << 32) | (unsigned long long)(0xFFFFFFFFu - r);
    atomicMax(&maxCount[ids[r]], (int)cnt);
    newId[r] = (uint8_t)TT_MAXP;
}

hipError_t tt_launch_frag_build(hipStream_t s, const tt_dev &d, const uint32_t *rootList, const uint32_t *nroots,
                                uint32_t nroots_host, const uint32_t *rootCount, const uint8_t *ids,
                                unsigned long long *fragKey, int *maxCount, uint8_t *newId)
{
    if (nroots_host == 0) return hipSuccess;
    const unsigned blocks = (nroots_host + 255) / 256;
    hipLaunchKernelGGL(k_frag_build, dim3(blocks), dim3(256), 0, s, d, rootList, nroots, rootCount, ids, fragKey,
                       maxCount, newId);
    return hipGetLastError();
}

// single-workgroup bitonic sort of up to `capacity` (power of two) keys
__global__ __launch_bounds__(1024) void k_frag_sort(unsigned long long *key, const uint32_t *nroots, int capacity)
{
    extern __shared__ unsigned long long sk[];
    const uint32_t n = *nroots;
    for (int i = threadIdx.x; i < capacity; i += 1024) sk[i] = (uint32_t)i < n ? key[i] : ~0ull;
    __syncthreads();
    for (int k = 2; k <= capacity; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = threadIdx.x; i < capacity; i += 1024) {
                const int ixj = i ^ j;
                if (ixj > i) {
                    const unsigned long long a = sk[i], b = sk[ixj];
                    const bool up = (i & k) == 0;
                    if ((a > b) == up) { sk[i] = b; sk[ixj] = a; }
                }
            }
            __syncthreads();
        }
    }
    for (int i = threadIdx.x; i < capacity; i += 1024)
        if ((uint32_t)i < n) key[i] = sk[i];
}

hipError_t tt_launch_frag_sort(hipStream_t s, unsigned long long *fragKey, const uint32_t *nroots, int capacity)
{
    hipLaunchKernelGGL(k_frag_sort, dim3(1), dim3(1024), (size_t)capacity * 8, s, fragKey, nroots, capacity);
    return hipGetLastError();
}

// One phase of the all-ascending bitonic network over the first n keys; keys
// at indices >= n count as +infinity.  Every compare-exchange puts the smaller
// key at the lower index, so a real key never moves to an index >= n and the
// virtual ones never move below it.  flip: the pairs (b k + o, b k + k - 1 - o)
// that merge two sorted halves of a block of k; otherwise the pairs (i, i + j)
// in blocks of 2j.
__global__ __launch_bounds__(256) void k_sort_phase(unsigned long long *a, const uint32_t *np, long long pairs,
                                                    long long k, long long j, int flip)
{
    const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
    if (t >= pairs) return;
    const long long n = (long long)*np;
    long long i, l;
    if (flip) {
        const long long h = k >> 1, b = t / h, o = t - b * h;
        i = b * k + o;
        l = b * k + k - 1 - o;
    } else {
        const long long b = t / j, o = t - b * j;
        i = b * 2 * j + o;
        l = i + j;
    }
    if (l >= n) return;
    const unsigned long long x = a[i], y = a[l];
    if (y < x) {
        a[i] = y;
        a[l] = x;
    }
}

hipError_t tt_sort_keys_large(hipStream_t s, unsigned long long *key, const uint32_t *nroots, uint32_t n)
{
    long long cap = 2;
    while (cap < (long long)n) cap <<= 1;
    const long long pairs = cap >> 1;
    const unsigned blocks = (unsigned)((pairs + 255) / 256);
    for (long long k = 2; k <= cap; k <<= 1) {
        hipLaunchKernelGGL(k_sort_phase, dim3(blocks), dim3(256), 0, s, key, nroots, pairs, k, 0LL, 1);
        for (long long j = k >> 2; j > 0; j >>= 1)
            hipLaunchKernelGGL(k_sort_phase, dim3(blocks), dim3(256), 0, s, key, nroots, pairs, k, j, 0);
    }
    return hipGetLastError();
}

// createLabelIdLookup in canonical serial order
__global__ __launch_bounds__(256) void k_frag_assign(tt_dev d, const unsigned long long *fragKey,
                                                     const uint32_t *nroots, const int *maxCount, const uint8_t *ids,
                                                     tt_plate_data *pd, uint8_t *newId)
{
    __shared__ uint8_t avail[256];
    __shared__ int navail;
    const uint32_t n = *nroots;
    const int M = n < (uint32_t)TT_MAXP ? (int)n : TT_MAXP;
    const int i = threadIdx.x;
    // available ids: i < M with size == 0, ascending (`!size > 0`, :3151)
    const bool av = i < M && pd[i].size == 0;
    const unsigned long long b = __ballot(av);
    const int w = i >> 6, lane = i & 63;
    __shared__ int wcount[4];
    if (lane == 0) wcount[w] = __popcll(b);
    __syncthreads();
    int base = 0;
    for (int k = 0; k < w; ++k) base += wcount[k];
    if (av) avail[base + __popcll(b & ((1ull << lane) - 1ull))] = (uint8_t)i;
    if (i == 0) navail = wcount[0] + wcount[1] + wcount[2] + wcount[3];
    __syncthreads();
    if (i != 0) return;
    int ai = 0;
    for (int k = 0; k < M; ++k) {
        const unsigned long long key = fragKey[k];
        const uint32_t cnt = 0xFFFFFFFFu - (uint32_t)(key >> 32);
        const uint32_t label = 0xFFFFFFFFu - (uint32_t)(key & 0xFFFFFFFFu);
        const int pid = ids[label];
        if ((int)cnt == maxCount[pid]) {
            newId[label] = (uint8_t)pid;
        } else if ((int)cnt >= d.ks.minPlateSize) {
            const int idx = ai++;
            if (idx < navail) {
                const int nid = avail[idx];
                pd[nid] = pd[pid];
                newId[label] = (uint8_t)nid;
            }
        }
    }
}

hipError_t tt_launch_frag_assign(hipStream_t s, const tt_dev &d, const unsigned long long *fragKey,
                                 const uint32_t *nroots, const int *maxCount, const uint8_t *ids, tt_plate_data *pd,
                                 uint8_t *newId)
{
    hipLaunchKernelGGL(k_frag_assign, dim3(1), dim3(256), 0, s, d, fragKey, nroots, maxCount, ids, pd, newId);
    return hipGetLastError();
}

// postCClIdReassign; also completes the labels: labels[c] held c's local
// root, whose entry now holds the final root (the minimum cell index of the
// component)
__global__ __launch_bounds__(256) void k_relabel(tt_dev d, uint32_t *labels, const uint8_t *newId, uint8_t *ids,
                                                 uint8_t *unFlag)
{
    const long long c = (long long)blockIdx.x * 256 + threadIdx.x;
    if (c >= d.N) return;
    const uint32_t f = labels[labels[c]];
    labels[c] = f;
    const uint8_t id = newId[f];
    ids[c] = id;
    unFlag[c] = id == TT_MAXP ? 1 : 0;
}

hipError_t tt_launch_relabel(hipStream_t s, const tt_dev &d, uint32_t *labels, const uint8_t *newId,
                             uint8_t *ids, uint8_t *unFlag)
{
    const unsigned blocks = (unsigned)(((long long)d.N + 255) / 256);
    hipLaunchKernelGGL(k_relabel, dim3(blocks), dim3(256), 0, s, d, labels, newId, ids, unFlag);
    return hipGetLastError();
}

// neighbour vote of assignUnassignedIdsToNeighbor, Jacobi rounds
__global__ __launch_bounds__(256) void k_fill_vote(tt_dev d, const uint32_t *list, const uint32_t *n,
                                                   const uint8_t *ids, uint8_t *vote)
{
    const uint32_t k = blockIdx.x * 256 + threadIdx.x;
    if (k >= *n) return;
    const uint32_t c = list[k];
    int x, y;
    tt_cell_xy((uint32_t)c, d, x, y);
    const int NO[8][2] = {{1, 0}, {-1, 0}, {0, 1}, {0, -1}, {1, 1}, {1, -1}, {-1, 1}, {-1, -1}};
    uint8_t nb[8];
    int cnt = 0;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const uint8_t v = ids[(size_t)tt_wrap1(y + NO[q][1], d.H) * d.W + tt_wrap1(x + NO[q][0], d.W)];
        nb[q] = v;
        cnt += v != TT_MAXP;
    }
    uint8_t sel = TT_MAXP;
    if (cnt > 0) {
        int best = 0;
        sel = 0;
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            if (nb[i] == TT_MAXP) continue;
            int lc = 0;
#pragma unroll
            for (int j = 0; j < 8; ++j) lc += nb[j] == nb[i];
            if (lc > best) { best = lc; sel = nb[i]; }
        }
    }
    vote[k] = sel;
}

hipError_t tt_launch_fill_vote(hipStream_t s, const tt_dev &d, const uint32_t *list, const uint32_t *n,
                               const uint8_t *ids, uint8_t *vote)
{
    const unsigned blocks = (unsigned)(((long long)d.N + 255) / 256);
    hipLaunchKernelGGL(k_fill_vote, dim3(blocks), dim3(256), 0, s, d, list, n, ids, vote);
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_fill_apply(tt_dev d, const uint32_t *list, const uint32_t *n,
                                                    const uint8_t *vote, uint8_t *ids, uint32_t *nextList,
                                                    uint32_t *nextN)
{
    const uint32_t k = blockIdx.x * 256 + threadIdx.x;
    if (k >= *n) return;
    const uint8_t v = vote[k];
    if (v != TT_MAXP) ids[list[k]] = v;
    else nextList[atomicAdd(nextN, 1u)] = list[k];
}

hipError_t tt_launch_fill_apply(hipStream_t s, const tt_dev &d, const uint32_t *list, const uint32_t *n,
                                const uint8_t *vote, uint8_t *ids, uint32_t *nextList, uint32_t *nextN)
{
    const unsigned blocks = (unsigned)(((long long)d.N + 255) / 256);
    hipLaunchKernelGGL(k_fill_apply, dim3(blocks), dim3(256), 0, s, d, list, n, vote, ids, nextList, nextN);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// G2: merge
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_neighbor_matrix(tt_dev d, const uint8_t *ids, uint8_t *nbm)
{
    const long long c = (long long)blockIdx.x * 256 + threadIdx.x;
    if (c >= d.N) return;
    int x, y;
    tt_cell_xy((uint32_t)c, d, x, y);
    const uint8_t cur = ids[c];
    const uint8_t r = ids[(size_t)y * d.W + tt_wrap1(x + 1, d.W)];
    const uint8_t t = ids[(size_t)tt_wrap1(y + 1, d.H) * d.W + x];
    if (cur != r) {
        const int mn = r < cur ? r : cur, mx = r < cur ? cur : r;
        nbm[(size_t)(mx % TT_MAXP) * TT_MAXP + (mn % TT_MAXP)] = 1;
    }
    if (cur != t) {
        const int mn = t < cur ? t : cur, mx = t < cur ? cur : t;
        nbm[(size_t)(mx % TT_MAXP) * TT_MAXP + (mn % TT_MAXP)] = 1;
    }
}

hipError_t tt_launch_neighbor_matrix(hipStream_t s, const tt_dev &d, const uint8_t *ids, uint8_t *nbm)
{
    const unsigned blocks = (unsigned)(((long long)d.N + 255) / 256);
    hipLaunchKernelGGL(k_neighbor_matrix, dim3(blocks), dim3(256), 0, s, d, ids, nbm);
    return hipGetLastError();
}

#define TT_NBM (TT_MAXP * TT_MAXP)
#define TT_NBM_WORDS ((TT_NBM + 31) / 32)

// getPlateMerges + clearCollisionTypes: candidates evaluated in parallel,
// pairing applied by one thread in ascending matrix index.  The plate fields
// the test reads are staged in LDS and the adjacency matrix is read 16 bytes
// per load, so the candidate pass costs a few memory round trips.
struct tt_merge_pl {
    float dsx, dsy, vs, v, mass;
    int size, wait;
};

__device__ __forceinline__ void tt_merge_test(int i, const tt_merge_pl *P, const tt_kernel_settings &K, unsigned *cand)
{
    const int x = i % TT_MAXP, y = i / TT_MAXP;
    if (x >= y) return;
    const tt_merge_pl &X = P[x], &Y = P[y];
    const float dot = X.dsx * Y.dsx + X.dsy * Y.dsy;
    const float vd = tt_fabsf(X.vs - Y.vs);
    if (!(X.v <= K.mergeMinVelocity && Y.v <= K.mergeMinVelocity) &&
        !(dot >= K.mergeDotDirectionThreshold && vd <= K.mergeVelocityDiffThreshold))
        return;
    if (X.wait > 0 || Y.wait > 0) return;
    if (X.size + Y.size > K.targetMaximumPlateArea) return;
    atomicOr(&cand[i >> 5], 1u << (i & 31));
}

__device__ __forceinline__ void tt_merge_word(uint32_t wd, int i0, const tt_merge_pl *P, const tt_kernel_settings &K,
                                              unsigned *cand)
{
    if (wd == 0u) return;
#pragma unroll
    for (int b = 0; b < 4; ++b)
        if (((wd >> (8 * b)) & 0xFFu) && i0 + b < TT_NBM) tt_merge_test(i0 + b, P, K, cand);
}

__global__ __launch_bounds__(1024) void k_merge_decide(tt_dev d, const uint8_t *nbm, const tt_plate_data *pd,
                                                       int *mergeIds, uint8_t *typeBitmap, tt_step_ctl *ctl)
{
    __shared__ unsigned cand[TT_NBM_WORDS];
    __shared__ int mids[256];
    __shared__ tt_merge_pl P[TT_MAXP];
    const tt_kernel_settings &K = d.ks;
    for (int i = threadIdx.x; i < TT_NBM_WORDS; i += 1024) cand[i] = 0u;
    for (int i = threadIdx.x; i < 256; i += 1024) mids[i] = 0;
    if (threadIdx.x < TT_MAXP) {
        const tt_plate_data &Q = pd[threadIdx.x];
        tt_merge_pl q;
        q.dsx = Q.directionSmoothed[0];
        q.dsy = Q.directionSmoothed[1];
        q.vs = Q.velocitySmoothed;
        q.v = Q.velocity;
        q.size = Q.size;
        q.mass = Q.mass;
        q.wait = Q.mergeWaitTime;
        P[threadIdx.x] = q;
    }
    // the matrix as 16-byte groups: 1024 threads x 4 groups = 65536 bytes; the
    // buffer holds at least TT_NBM + 64 bytes, zero past TT_NBM + 4
    uint4 g[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int q = threadIdx.x + 1024 * k;
        g[k] = q * 16 < TT_NBM ? reinterpret_cast<const uint4 *>(nbm)[q] : make_uint4(0u, 0u, 0u, 0u);
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int i0 = (threadIdx.x + 1024 * k) * 16;
        tt_merge_word(g[k].x, i0, P, K, cand);
        tt_merge_word(g[k].y, i0 + 4, P, K, cand);
        tt_merge_word(g[k].z, i0 + 8, P, K, cand);
        tt_merge_word(g[k].w, i0 + 12, P, K, cand);
    }
    __syncthreads();
    // the pairing is serial in ascending matrix index; wavefront 0 finds the
    // non-empty candidate words 64 at a time, lane 0 walks their bits
    if (threadIdx.x < 64) {
        int merges = 0;
        for (int w0 = 0; w0 < TT_NBM_WORDS; w0 += 64) {
            const int wl = w0 + (int)threadIdx.x;
            unsigned long long nz = __ballot(wl < TT_NBM_WORDS && cand[wl] != 0u);
            if (threadIdx.x != 0) continue;
            while (nz) {
                const int wi = w0 + __ffsll((long long)nz) - 1;
                nz &= nz - 1ull;
                unsigned bits = cand[wi];
                while (bits) {
                    const int bit = __ffs((int)bits) - 1;
                    bits &= bits - 1u;
                    const int i = wi * 32 + bit;
                    const int x = i % TT_MAXP, y = i / TT_MAXP;
                    const bool xIsParent = P[y].mass <= P[x].mass;
                    if (mids[x] != 0) continue;
                    mids[x] = xIsParent ? -1 : TT_MAXP - y;
                    if (mids[y] != 0) { mids[x] = 0; continue; }
                    mids[y] = xIsParent ? TT_MAXP - x : -1;
                    ++merges;
                    const int parent = xIsParent ? x : y;
                    // clearCollisionTypes writes entry min(parent, j) of the 1-row table
                    const int hi = parent < TT_MAXP - 1 ? parent : parent - 1;
                    for (int j = 0; j <= hi; ++j) typeBitmap[j] = 0;
                }
            }
        }
        if (threadIdx.x == 0) ctl->merges = merges;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < TT_MAXP; i += 1024) mergeIds[i] = mids[i];
}

hipError_t tt_launch_merge_decide(hipStream_t s, const tt_dev &d, const uint8_t *nbm, const tt_plate_data *pd,
                                  int *mergeIds, uint8_t *typeBitmap, tt_step_ctl *ctl)
{
    hipLaunchKernelGGL(k_merge_decide, dim3(1), dim3(1024), 0, s, d, nbm, pd, mergeIds, typeBitmap, ctl);
    return hipGetLastError();
}

// per-plate accumulator layout (int64 each)
#define ACC_MASS 0
#define ACC_AX 1
#define ACC_AY 2
#define ACC_S0 3
#define ACC_S1 4
#define ACC_S2 5
#define ACC_S3 6
#define ACC_SIZE 7
#define ACC_PER 8
#define ACC_ACNT 9
#define ACC_N 10

// mergeAndCountSizeMass (mode 0) / countSizePostSplit (mode 1), plus the
// angular samples of accumulatePlateAngularCoords in mode 0.  Each thread
// takes 16 consecutive cells of one row (vector loads), sums runs of equal
// plate id in registers and adds each run to the per-plate LDS totals; 4096
// cells per workgroup.  All sums are integers (fixed point), so the order of
// the additions does not matter.
struct tt_count_run {
    long long mass, ax, ay, s0, s1, s2, s3;
    unsigned size, per, acnt;
};

__device__ __forceinline__ void tt_count_flush(unsigned long long *sa, int pid, tt_count_run &R, int mode)
{
    if (R.size == 0u) return;
    unsigned long long *a = &sa[pid * ACC_N];
    atomicAdd(&a[ACC_MASS], (unsigned long long)R.mass);
    atomicAdd(&a[ACC_SIZE], (unsigned long long)R.size);
    if (R.per) atomicAdd(&a[ACC_PER], (unsigned long long)R.per);
    if (mode == 0) {
        atomicAdd(&a[ACC_AX], (unsigned long long)R.ax);
        atomicAdd(&a[ACC_AY], (unsigned long long)R.ay);
        if (R.acnt) {
            atomicAdd(&a[ACC_S0], (unsigned long long)R.s0);
            atomicAdd(&a[ACC_S1], (unsigned long long)R.s1);
            atomicAdd(&a[ACC_S2], (unsigned long long)R.s2);
            atomicAdd(&a[ACC_S3], (unsigned long long)R.s3);
            atomicAdd(&a[ACC_ACNT], (unsigned long long)R.acnt);
        }
    }
    R = tt_count_run{0, 0, 0, 0, 0, 0, 0, 0u, 0u, 0u};
}

__device__ __forceinline__ uint8_t tt_byte(const uint4 &v, int j)
{
    const uint32_t w = j < 4 ? v.x : j < 8 ? v.y : j < 12 ? v.z : v.w;
    return (uint8_t)(w >> (8 * (j & 3)));
}

__global__ __launch_bounds__(256) void k_count(tt_dev d, const int *mergeIds, const uint8_t *idsIn, uint8_t *idsOut,
                                               const float *height, const uint8_t *noise, long long *acc, int mode)
{
    __shared__ unsigned long long sa[256 * ACC_N];
    __shared__ int mmap[256];
    const int tid = threadIdx.x;
    for (int i = tid; i < 256 * ACC_N; i += 256) sa[i] = 0ull;
    mmap[tid] = (mode == 0 && tid < TT_MAXP) ? mergeIds[tid] : 0;
    __syncthreads();
    const int W = d.W, H = d.H;
    const long long c0 = (long long)blockIdx.x * 4096 + (long long)tid * 16;
    if (c0 < d.N) {
        int x0, y;
        tt_cell_xy((uint32_t)c0, d, x0, y);
        // merged id of a raw id (uint8 arithmetic of the reference)
        auto mid = [&](int raw) -> int {
            if (mode != 0) return raw;
            const int mg = mmap[raw];
            return mg > 0 ? TT_MAXP - mg : raw;
        };
        auto nid = [&](int raw) -> int {
            if (mode != 0) return raw;
            const uint8_t mm = (uint8_t)mmap[raw];
            return mm != 0 ? (int)(uint8_t)(TT_MAXP - mm) : raw;
        };
        const uint4 own = *(const uint4 *)(idsIn + c0);
        const uint4 up = *(const uint4 *)(idsIn + (size_t)tt_wrap1(y - 1, H) * W + x0);
        const uint4 dn = *(const uint4 *)(idsIn + (size_t)tt_wrap1(y + 1, H) * W + x0);
        const int lft = idsIn[(size_t)y * W + tt_wrap1(x0 - 1, W)];
        const int rgt = idsIn[(size_t)y * W + tt_wrap1(x0 + 16, W)];
        float hv[16];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const float4 f = *(const float4 *)(height + c0 + 4 * q);
            hv[4 * q] = f.x; hv[4 * q + 1] = f.y; hv[4 * q + 2] = f.z; hv[4 * q + 3] = f.w;
        }
        uint32_t ow[4] = {0u, 0u, 0u, 0u};
        tt_count_run R = tt_count_run{0, 0, 0, 0, 0, 0, 0, 0u, 0u, 0u};
        int cur = -1;
        const double twopi = 6.283185307179586;
        // neighbouring cells sample the same noise texel (16 cells per texel
        // at 4096 columns): the texel's vector is computed when it changes
        const int sy = (int)((double)y * d.ays);
        int csx = -1;
        float cax = 0.0f, cay = 0.0f;
        long long fax = 0, fay = 0;  // the texel's vector in fixed point
        // the row's angle terms of the centre sums, the same for all 16 cells
        long long fs2 = 0, fs3 = 0;
        if (mode == 0 && y % 10 == 0) {
            const double agy = (twopi * (double)y) / (double)H;
            fs2 = tt_fix_d(tt_sin_d(agy), TT_ANG_BITS);
            fs3 = tt_fix_d(tt_cos_d(agy), TT_ANG_BITS);
        }
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            const int x = x0 + j;
            const int pid = mid(tt_byte(own, j));
            ow[j >> 2] |= (uint32_t)pid << (8 * (j & 3));
            if (pid != cur) {
                if (cur >= 0) tt_count_flush(sa, cur, R, mode);
                cur = pid;
            }
            R.mass += tt_fix(hv[j], TT_MASS_BITS);
            R.size += 1u;
            const int nb0 = nid(tt_byte(up, j));
            const int nb1 = nid(j < 15 ? (int)tt_byte(own, j + 1) : rgt);
            const int nb2 = nid(tt_byte(dn, j));
            const int nb3 = nid(j > 0 ? (int)tt_byte(own, j - 1) : lft);
            const unsigned cnt = (unsigned)(nb0 != pid) + (unsigned)(nb1 != pid) + (unsigned)(nb2 != pid) +
                                 (unsigned)(nb3 != pid);
            if (mode == 0) {
                R.per += cnt;
                const int sx = (int)((double)x * d.axs);
                if (sx != csx) {
                    tt_asth_texel(noise, sx, sy, cax, cay);
                    fax = tt_fix(cax, TT_ASTH_BITS);
                    fay = tt_fix(cay, TT_ASTH_BITS);
                    csx = sx;
                }
                R.ax += fax;
                R.ay += fay;
                if (x % 10 == 0 && y % 10 == 0) {
                    const double agx = (twopi * (double)x) / (double)W;
                    R.s0 += tt_fix_d(tt_sin_d(agx), TT_ANG_BITS);
                    R.s1 += tt_fix_d(tt_cos_d(agx), TT_ANG_BITS);
                    R.s2 += fs2;
                    R.s3 += fs3;
                    R.acnt += 1u;
                }
            } else if (cnt) {
                R.per += 1u;
            }
        }
        tt_count_flush(sa, cur, R, mode);
        *(uint4 *)(idsOut + c0) = make_uint4(ow[0], ow[1], ow[2], ow[3]);
    }
    __syncthreads();
    for (int i = tid; i < 256 * ACC_N; i += 256)
        if (sa[i] != 0ull) atomicAdd((unsigned long long *)&acc[i], sa[i]);
}

hipError_t tt_launch_count(hipStream_t s, const tt_dev &d, const int *mergeIds, const uint8_t *idsIn, uint8_t *idsOut,
                           const float *height, const uint8_t *noise, long long *acc, int split_mode)
{
    const unsigned blocks = (unsigned)(((long long)d.N + 4095) / 4096);
    hipLaunchKernelGGL(k_count, dim3(blocks), dim3(256), 0, s, d, mergeIds, idsIn, idsOut, height, noise, acc,
                       split_mode);
    return hipGetLastError();
}

// mode 0: resetPlateDataPreCount + sums, computeBreakScore, plate centres,
//         statisticsPass and the split decision (mode 2: without the split);
// mode 1: resetPlateDataPreCountPostSplit + sums of countSizePostSplit;
// mode 3: sizes and masses of initPixelDependantPlateData.
__global__ __launch_bounds__(256) void k_finalize(tt_dev d, tt_plate_data *pd, const long long *acc, int mode,
                                                  tt_step_ctl *ctl)
{
    const int p = threadIdx.x;
    const tt_kernel_settings &K = d.ks;
    if (p < TT_MAXP && mode == 3) {
        // initPixelDependantPlateData (:199-230): sizes and masses added
        tt_plate_data c = pd[p];
        const long long *a = &acc[p * ACC_N];
        if (a[ACC_SIZE] != 0) {
            c.size += (int)a[ACC_SIZE];
            c.mass = c.mass + tt_unfix(a[ACC_MASS], TT_MASS_BITS);
        }
        pd[p] = c;
    } else if (p < TT_MAXP) {
        tt_plate_data c = pd[p];
        const long long *a = &acc[p * ACC_N];
        c.asthenosphereVelocity[0] = 0.0f;
        c.asthenosphereVelocity[1] = 0.0f;
        c.mass = 0.0f;
        c.size = 0;
        c.perimeter = 0;
        c.breakScore = 0.0f;
        c.used = 0;
        if (mode == 1 && (p == ctl->split_plate || p == ctl->split_new)) c.mergeWaitTime = K.mergeWaitAfterSplit;
        if (a[ACC_SIZE] != 0) {
            c.mass = tt_unfix(a[ACC_MASS], TT_MASS_BITS);
            c.size = (int)a[ACC_SIZE];
            c.perimeter = (int)a[ACC_PER];
            if (mode == 0 || mode == 2) {
                c.asthenosphereVelocity[0] = tt_unfix(a[ACC_AX], TT_ASTH_BITS);
                c.asthenosphereVelocity[1] = tt_unfix(a[ACC_AY], TT_ASTH_BITS);
            }
        }
        if (mode == 0 || mode == 2) {
            // computeBreakScore
            if (c.size > 0) {
                const float area = (float)c.size, per = (float)c.perimeter;
                c.circularity = 1.0f - (((2.0f * 6.2831855f) * area) / tt_fmaxf(per * per, 1.0f));
                const float mn = (float)K.targetMinimumPlateArea, mx = (float)K.targetMaximumPlateArea;
                const float as = tt_fmaxf(tt_fminf((area - mn) / (mx - mn), 1.0f), 0.0f);
                c.breakScore = as + c.circularity;
            } else {
                c.breakScore = 0.0f;
            }
            // calculatePlateCenters
            if (a[ACC_ACNT] > 0) {
                const double n = (double)a[ACC_ACNT];
                const double a0 = tt_unfix_d(a[ACC_S0], TT_ANG_BITS) / n, a1 = tt_unfix_d(a[ACC_S1], TT_ANG_BITS) / n;
                const double a2 = tt_unfix_d(a[ACC_S2], TT_ANG_BITS) / n, a3 = tt_unfix_d(a[ACC_S3], TT_ANG_BITS) / n;
                const double twopi = 6.283185307179586;
                double cx = (tt_atan2(a0, a1) * (double)d.W) / twopi;
                double cy = (tt_atan2(a2, a3) * (double)d.H) / twopi;
                if (cx < 0.0) cx = cx + (double)d.W;
                if (cy < 0.0) cy = cy + (double)d.H;
                c.geometricCenter[0] = (float)cx;
                c.geometricCenter[1] = (float)cy;
            }
        }
        pd[p] = c;
    }
    __syncthreads();
    if (p != 0) return;
    if (mode == 1 || mode == 3) return;
    // statisticsPass
    float heavy = 0.0f;
    int largest = 0, hid = 0, lid = 0, np = 0;
    for (int i = 0; i < TT_MAXP; ++i) {
        if (heavy < pd[i].mass) { heavy = pd[i].mass; hid = i; }
        if (largest < pd[i].size) { largest = pd[i].size; lid = i; }
        pd[i].used = pd[i].size > 0 ? 1 : 0;
        np += pd[i].used;
    }
    ctl->heaviest_value = heavy;
    ctl->heaviest_id = hid;
    ctl->largest_value = largest;
    ctl->largest_id = lid;
    ctl->n_plates = np;
    if (mode == 2) {
        ctl->split_plate = -1;
        ctl->split_new = -1;
        return;
    }
    // processPlateSplitting: first used plate with breakScore >= 1.2, and the
    // id selectUnusedPlateId would pick
    int sp = -1, sn = -1;
    for (int i = 0; i < TT_MAXP; ++i)
        if (pd[i].used && pd[i].breakScore >= 1.2f) { sp = i; break; }
    if (sp >= 0)
        for (int i = 0; i < TT_MAXP; ++i)
            if (!pd[i].used) { sn = i; break; }
    if (sn < 0) sp = -1;  // no free id: no split
    if (sp >= 0) pd[sn].used = 1;  // selectUnusedPlateId marks the id used
    ctl->split_plate = sp;
    ctl->split_new = sn;
}

hipError_t tt_launch_finalize(hipStream_t s, const tt_dev &d, tt_plate_data *pd, const long long *acc, int mode,
                              tt_step_ctl *ctl)
{
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(256), 0, s, d, pd, acc, mode, ctl);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// G6: split
// ---------------------------------------------------------------------------
// initEffortToBoundary (:2824-2836): +inf on the plate, 0 elsewhere; the
// plate's cells are flagged for the compaction, and a plate cell whose height
// (the relaxation's cost) is negative or NaN is reported: such a cost has no
// fixed point, and the split is abandoned.
__global__ __launch_bounds__(256) void k_split_prepare(tt_dev d, const uint8_t *ids, int plate, const float *height,
                                                       float *eff, uint8_t *flag, tt_step_ctl *ctl)
{
    const long long c = (long long)blockIdx.x * 256 + threadIdx.x;
    if (c >= d.N) return;
    const bool in = ids[c] == plate;
    eff[c] = in ? tt_u2f(0x7f800000u) : 0.0f;
    flag[c] = in ? 1 : 0;
    if (in && !(height[c] >= 0.0f)) ctl->split_negative = 1;
}

hipError_t tt_launch_split_prepare(hipStream_t s, const tt_dev &d, const uint8_t *ids, int plate, const float *height,
                                   float *eff, uint8_t *flag, tt_step_ctl *ctl)
{
    const unsigned blocks = (unsigned)(((long long)d.N + 255) / 256);
    hipLaunchKernelGGL(k_split_prepare, dim3(blocks), dim3(256), 0, s, d, ids, plate, height, eff, flag, ctl);
    return hipGetLastError();
}

// Tiles of the split stage: 64 columns by TH rows, 256 threads; thread t owns
// column t % 64 and rows t / 64 + 4 k.  k_split_tiles flags the tiles that hold
// cells of the plate.
template <int TH>
__global__ __launch_bounds__(256) void k_split_tiles(tt_dev d, const uint8_t *ids, int plate, uint8_t *flag)
{
    const int tpr = d.W / 64;
    const int tx0 = (int)(blockIdx.x % (unsigned)tpr) * 64, ty0 = (int)(blockIdx.x / (unsigned)tpr) * TH;
    const int lx = threadIdx.x & 63, wv = threadIdx.x >> 6;
    int any = 0;
#pragma unroll
    for (int j = 0; j < TH / 4; ++j) any |= ids[(size_t)(ty0 + wv + 4 * j) * d.W + tx0 + lx] == plate;
    any = __syncthreads_or(any);
    if (threadIdx.x == 0) flag[blockIdx.x] = any ? 1 : 0;
}

// Round bookkeeping of the relaxation.  act[p][t]
// (p = round parity) says that tile t must run again: it wrote cells a
// neighbour reads, or it stopped at the sweep cap.  A tile runs in round r > 0
// only if it or one of its 8 neighbouring tiles (toroidal) is marked in
// act[(r - 1) & 1]; it always rewrites its own act[r & 1] entry.  rchg[r]
// counts the tiles that wrote in round r; a round with rchg == 0 is a fixed
// point and later rounds exit at once.  rchg is a ring of TT_RCHG counters
// (two host batches), so a round can read its predecessor across batches.
struct tt_round {
    uint32_t *rchg;
    uint8_t *act;       // 2 x tiles
    int ntiles;         // tiles in the whole grid
    int tpr;            // tiles per row
    int tpc;            // tiles per column
    uint32_t tprMagic;  // ceil(2^32 / tpr) (0 for tpr = 1): tile / tpr as a multiply-high,
                        // exact for every tile index below 2^19 and tpr up to 256
    int round;
};

// (column, row) of a tile of the split grid, on the scalar unit
__device__ __forceinline__ void tt_round_tile_xy(const tt_round &R, uint32_t tile, int &tx, int &ty)
{
    ty = R.tpr == 1 ? (int)tile : (int)__umulhi(tile, R.tprMagic);
    tx = (int)tile - ty * R.tpr;
}

// The activity of tile and its 8 neighbours in the previous round, one
// flag per calling thread (threads 0..8; the others return 0): combined with
// __syncthreads_or, a block learns in one memory round trip whether to run.
__device__ __forceinline__ int tt_round_neighbour_active(const tt_round &R, uint32_t tile, int t)
{
    if (R.round == 0 || t >= 9) return 0;
    const uint8_t *prev = R.act + (size_t)((R.round - 1) & 1) * R.ntiles;
    int tx, ty;
    tt_round_tile_xy(R, tile, tx, ty);
    const int dy = t / 3 - 1, dx = t % 3 - 1;
    return prev[tt_wrap1(ty + dy, R.tpc) * R.tpr + tt_wrap1(tx + dx, R.tpr)];
}

__device__ __forceinline__ void tt_round_mark(const tt_round &R, uint32_t tile, int wrote, int capped)
{
    uint8_t *cur = R.act + (size_t)(R.round & 1) * R.ntiles;
    cur[tile] = (uint8_t)((wrote || capped) ? 1 : 0);
    if (wrote) atomicAdd(&R.rchg[R.round % TT_RCHG], 1u);
}

// Sweeps of a tile per round.  A tile that still changes after them writes
// what it has and stays active for the next round, which reads fresher halo
// values; a round then ends with its slowest tile after two barrier intervals
// instead of at that tile's own fixed point.  On the 4096^2 benchmark world
// 8 sweeps per round took the split from 1.32 to 0.86 ms per step (16: 0.90,
// 4: 1.04, unbounded: 1.32).
#define TT_SWEEP_CAP 8

// propagateEffortToBoundary (plate_tectonics_kernel.cu:2838-2881) as rounds of
// tile-local relaxation: each active tile relaxes its cells in LDS (halo
// values read at the start of the round) for at most TT_SWEEP_CAP sweeps,
// then writes back the cells it lowered.  The update is monotone and starts
// from an upper bound, so the fixed point does not depend on the order of the
// updates: it equals the oracle's worklist result.
//
// Inside a tile each wavefront owns TH / 4 consecutive rows and relaxes them
// top-down and then bottom-up (Gauss-Seidel in the column direction): a row
// reads the row the same wavefront has just lowered, so a value crosses the
// wavefront's rows in one pass instead of one row per pass.  Between two
// barriers every wavefront makes the down and up passes twice; a barrier
// interval in which no cell of the tile changed is the tile's fixed point.
template <int TH>
__global__ __launch_bounds__(256) void k_relax_tiles(tt_dev d, const uint32_t *tiles, const uint8_t *ids, int plate,
                                                     const float *height, float *eff, tt_round R)
{
    const uint32_t tile = tiles[blockIdx.x];
    // Whether the tile runs this round (round 0, or a change in the previous
    // round and an active tile among it and its neighbours) and the tile's
    // inputs are loaded together: one memory round trip before the decision.
    const uint32_t prevChg = R.round > 0 ? R.rchg[(R.round - 1) % TT_RCHG] : 1u;
    const int nbActive = tt_round_neighbour_active(R, tile, (int)threadIdx.x);
    constexpr int PR = TH / 4;
    constexpr int NE = ((TH + 2) * 66 + 255) / 256;
    __shared__ float e[TH + 2][67];
    int ttx, tty;
    tt_round_tile_xy(R, tile, ttx, tty);
    const int tx0 = ttx * 64, ty0 = tty * TH;
    const int W = d.W, H = d.H;
    float ev[NE];
#pragma unroll
    for (int k = 0; k < NE; ++k) {
        const int i = (int)threadIdx.x + 256 * k;
        if (i < (TH + 2) * 66) {
            const int r = i / 66, q = i - r * 66;
            ev[k] = eff[(size_t)tt_wrap1(ty0 + r - 1, H) * W + tt_wrap1(tx0 + q - 1, W)];
        }
    }
    const int lx = threadIdx.x & 63, wv = threadIdx.x >> 6;
    float c1[PR], c2[PR], init[PR];
    bool in[PR];
    const float SQRT2 = 1.41421356237f;
#pragma unroll
    for (int j = 0; j < PR; ++j) {
        const size_t c = (size_t)(ty0 + wv * PR + j) * W + tx0 + lx;
        in[j] = ids[c] == plate;
        c1[j] = height[c] * 1.0f;
        c2[j] = height[c] * SQRT2;
    }
    const int active = __syncthreads_or(nbActive);
    if (R.round > 0 && (prevChg == 0u || !active)) {
        if (threadIdx.x == 0) R.act[(size_t)(R.round & 1) * R.ntiles + tile] = 0;
        return;
    }
#pragma unroll
    for (int k = 0; k < NE; ++k) {
        const int i = (int)threadIdx.x + 256 * k;
        if (i < (TH + 2) * 66) e[i / 66][i % 66] = ev[k];
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < PR; ++j) init[j] = e[wv * PR + j + 1][lx + 1];
    int capped = 1;
    for (int it = 0; it < TT_SWEEP_CAP; it += 4) {
        int ch = 0;
#pragma unroll
        for (int jj = 0; jj < 4 * PR; ++jj) {
            const int j = (jj % (2 * PR)) < PR ? (jj % (2 * PR)) : 2 * PR - 1 - (jj % (2 * PR));
            if (in[j]) {
                const int r = wv * PR + j + 1, q = lx + 1;
                const float old = e[r][q];
                float md = old;
                // neighbour order (-1,0),(1,0),(0,-1),(0,1),(-1,-1),(-1,1),(1,-1),(1,1)
                float t;
                t = e[r][q - 1] + c1[j]; if (t < md) md = t;
                t = e[r][q + 1] + c1[j]; if (t < md) md = t;
                t = e[r - 1][q] + c1[j]; if (t < md) md = t;
                t = e[r + 1][q] + c1[j]; if (t < md) md = t;
                t = e[r - 1][q - 1] + c2[j]; if (t < md) md = t;
                t = e[r + 1][q - 1] + c2[j]; if (t < md) md = t;
                t = e[r - 1][q + 1] + c2[j]; if (t < md) md = t;
                t = e[r + 1][q + 1] + c2[j]; if (t < md) md = t;
                if (md != old) { e[r][q] = md; ch = 1; }
            }
            // the next row's reads follow this row's writes (also of the
            // other lanes): wait for the LDS queue, and keep the compiler
            // from moving those reads up
            asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
        }
        if (!__syncthreads_or(ch)) { capped = 0; break; }
    }
    int wrote = 0;
#pragma unroll
    for (int j = 0; j < PR; ++j) {
        const float v = e[wv * PR + j + 1][lx + 1];
        if (in[j] && v != init[j]) {
            eff[(size_t)(ty0 + wv * PR + j) * W + tx0 + lx] = v;
            wrote = 1;
        }
    }
    wrote = __syncthreads_or(wrote);
    if (threadIdx.x == 0) tt_round_mark(R, tile, wrote, capped);
}

// Rows of the split tiles (16, 32 and 64 are instantiated).  16 measures
// fastest: a taller tile needs fewer rounds, but its in-tile convergence costs
// more than the rounds it saves.
int tt_split_tile_h(const tt_dev &d) { return 16; }

hipError_t tt_launch_split_tiles(hipStream_t s, const tt_dev &d, const uint8_t *ids, int plate, uint8_t *flag)
{
    const int th = tt_split_tile_h(d);
    const unsigned nt = (unsigned)((d.W / 64) * (d.H / th));
    if (th == 64) hipLaunchKernelGGL(k_split_tiles<64>, dim3(nt), dim3(256), 0, s, d, ids, plate, flag);
    else if (th == 32) hipLaunchKernelGGL(k_split_tiles<32>, dim3(nt), dim3(256), 0, s, d, ids, plate, flag);
    else hipLaunchKernelGGL(k_split_tiles<16>, dim3(nt), dim3(256), 0, s, d, ids, plate, flag);
    return hipGetLastError();
}

static tt_round make_round(const tt_dev &d, uint32_t *rchg, uint8_t *act, int r)
{
    tt_round R;
    R.rchg = rchg;
    R.act = act;
    R.tpr = d.W / 64;
    R.tpc = d.H / tt_split_tile_h(d);
    R.ntiles = R.tpr * R.tpc;
    R.tprMagic = R.tpr >= 2 ? 0xFFFFFFFFu / (uint32_t)R.tpr + 1u : 0u;
    R.round = r;
    return R;
}

hipError_t tt_launch_relax_tiles(hipStream_t s, const tt_dev &d, const uint32_t *tiles, unsigned ntiles,
                                 const uint8_t *ids, int plate, const float *height, float *eff, uint32_t *rchg,
                                 uint8_t *act, int round0, int rounds)
{
    if (ntiles == 0) return hipSuccess;
    const int th = tt_split_tile_h(d);
    for (int r = round0; r < round0 + rounds; ++r) {
        const tt_round R = make_round(d, rchg, act, r);
        if (th == 64)
            hipLaunchKernelGGL(k_relax_tiles<64>, dim3(ntiles), dim3(256), 0, s, d, tiles, ids, plate, height, eff, R);
        else if (th == 32)
            hipLaunchKernelGGL(k_relax_tiles<32>, dim3(ntiles), dim3(256), 0, s, d, tiles, ids, plate, height, eff, R);
        else
            hipLaunchKernelGGL(k_relax_tiles<16>, dim3(ntiles), dim3(256), 0, s, d, tiles, ids, plate, height, eff, R);
    }
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// floodFillPlate (plate_tectonics_kernel.cu:2985-3024) as connected
// components: the flood from the seed through 4-connected cells of the old
// plate reaches exactly the components (4-connectivity, toroidal) of old-plate
// cells that touch the seed cell, so they are labelled once and relabelled.
// lab[c] = parent in a union-find over old-plate cells (0xFFFFFFFF elsewhere).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_fl_local(tt_dev d, const uint8_t *ids, int oldId, uint32_t *lab,
                                                  uint8_t *rootFlag)
{
    __shared__ unsigned par[TT_TILE_H * TT_TILE_W];
    const int tid = threadIdx.x;
    const int x0 = blockIdx.x * TT_TILE_W, y0 = blockIdx.y * TT_TILE_H;
    const int W = d.W;
    bool in[4];
    int any = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int i = tid + 256 * k;
        const int ly = i >> 6, lx = i & 63;
        in[k] = ids[(size_t)(y0 + ly) * W + x0 + lx] == oldId;
        any |= in[k];
        par[i] = (unsigned)i;
    }
    if (!__syncthreads_or(any)) {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int i = tid + 256 * k;
            const size_t c = (size_t)(y0 + (i >> 6)) * W + x0 + (i & 63);
            lab[c] = 0xFFFFFFFFu;
            rootFlag[c] = 0;
        }
        return;
    }
    // in-tile edges: left and up neighbours of the same plate
    __shared__ uint8_t inL[TT_TILE_H * TT_TILE_W];
#pragma unroll
    for (int k = 0; k < 4; ++k) inL[tid + 256 * k] = in[k] ? 1 : 0;
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int i = tid + 256 * k;
        if (!in[k]) continue;
        const int ly = i >> 6, lx = i & 63;
        if (lx > 0 && inL[i - 1]) tt_union32(par, (unsigned)i, (unsigned)(i - 1));
        if (ly > 0 && inL[i - TT_TILE_W]) tt_union32(par, (unsigned)i, (unsigned)(i - TT_TILE_W));
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int i = tid + 256 * k;
        const size_t c = (size_t)(y0 + (i >> 6)) * W + x0 + (i & 63);
        if (!in[k]) {
            lab[c] = 0xFFFFFFFFu;
            rootFlag[c] = 0;
            continue;
        }
        const unsigned r = tt_find32h(par, (unsigned)i);
        lab[c] = (uint32_t)((size_t)(y0 + (int)(r >> 6)) * W + x0 + (int)(r & 63));
        rootFlag[c] = r == (unsigned)i ? 1 : 0;
    }
}

// unions across tile borders: the left column joins its left neighbours, the
// top row its upper neighbours (toroidal)
__global__ __launch_bounds__(128) void k_fl_merge(tt_dev d, const uint8_t *ids, int oldId, uint32_t *lab)
{
    const int tid = threadIdx.x;
    const int x0 = blockIdx.x * TT_TILE_W, y0 = blockIdx.y * TT_TILE_H;
    int x, y, nx, ny;
    if (tid < 64) { x = x0 + tid; y = y0; nx = x; ny = tt_wrap1(y - 1, d.H); }
    else if (tid < 64 + TT_TILE_H) { x = x0; y = y0 + tid - 64; nx = tt_wrap1(x - 1, d.W); ny = y; }
    else return;
    const uint32_t c = (uint32_t)y * d.W + x, n = (uint32_t)ny * d.W + nx;
    if (ids[c] != oldId || ids[n] != oldId) return;
    const uint32_t a = lab[c], b = lab[n];
    if (a != b) tt_gunion(lab, a, b);
}

// final roots of the local roots (pointer jumping; grid-stride over the list)
__global__ __launch_bounds__(256) void k_fl_flatten(const uint32_t *list, const uint32_t *n, uint32_t *lab)
{
    const uint32_t cnt = *n;
    for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < cnt; i += gridDim.x * 256) {
        const uint32_t r = list[i];
        uint32_t p = lab[r];
        if (p == r) continue;
        for (;;) {
            const uint32_t gp = lab[p];
            if (gp == p) break;
            lab[r] = gp;
            p = gp;
        }
        lab[r] = p;
    }
}

// cells whose component touches the seed take the new id
__global__ __launch_bounds__(256) void k_fl_apply(tt_dev d, const uint32_t *lab, const tt_step_ctl *ctl,
                                                  uint8_t *ids, int oldId, int newId)
{
    const int seed = ctl->split_seed;
    if (seed < 0) return;
    const uint32_t c = blockIdx.x * 256u + threadIdx.x;
    if (c >= (uint32_t)d.N || ids[c] != oldId) return;
    int sx, sy;
    tt_cell_xy((uint32_t)seed, d, sx, sy);
    const uint32_t nb[4] = {(uint32_t)sy * d.W + tt_wrap1(sx - 1, d.W), (uint32_t)sy * d.W + tt_wrap1(sx + 1, d.W),
                            (uint32_t)tt_wrap1(sy - 1, d.H) * d.W + sx, (uint32_t)tt_wrap1(sy + 1, d.H) * d.W + sx};
    const uint32_t r = lab[lab[c]];
    bool hit = false;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const uint32_t l = lab[nb[k]];
        if (l != 0xFFFFFFFFu && lab[l] == r) hit = true;
    }
    if (hit) ids[c] = (uint8_t)newId;
}

hipError_t tt_launch_flood_cc(hipStream_t s, const tt_dev &d, uint8_t *ids, int oldId, int newId, uint32_t *lab,
                              uint8_t *rootFlag, uint32_t *list, uint32_t *nlist, uint32_t *scanTmp,
                              const tt_step_ctl *ctl)
{
    const dim3 tiles(d.W / TT_TILE_W, d.H / TT_TILE_H);
    hipLaunchKernelGGL(k_fl_local, tiles, dim3(256), 0, s, d, ids, oldId, lab, rootFlag);
    hipLaunchKernelGGL(k_fl_merge, tiles, dim3(128), 0, s, d, ids, oldId, lab);
    hipError_t e = tt_compact_u8_any(s, rootFlag, (long long)d.N, list, nlist);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_fl_flatten, dim3(1024), dim3(256), 0, s, list, nlist, lab);
    hipLaunchKernelGGL(k_fl_apply, dim3((unsigned)(((long long)d.N + 255) / 256)), dim3(256), 0, s, d, lab, ctl, ids,
                       oldId, newId);
    return hipGetLastError();
}

// BacktrackPath (plate_tectonics_kernel.cu:2883-2983).  The walks are serial:
// each step depends on the last.  A step reads the cell and its eight
// neighbours, so instead of one memory round trip per step one wavefront keeps
// a 64 x 64 window of efforts and plate ids in LDS (one round trip of 16-byte
// and 4-byte loads fills it) and walks inside it; the window is re-centred on
// the walk when the walk reaches its border, about every 30 steps.  All 64
// lanes walk the same path in lockstep from the same LDS values; every lane
// writes a mark to the window, lane 0 alone to global memory.  The window's
// left column is a multiple of 4, so a 4-cell group never straddles the wrap
// (W is a multiple of 64).  Requires W, H >= 64 (no cell appears twice in a
// window).
#define TT_BT_WIN 64
struct tt_bt_win {
    float e[TT_BT_WIN][TT_BT_WIN];
    uint8_t id[TT_BT_WIN][TT_BT_WIN];
};

// fill the window whose top-left cell is the wrapped cell (wxw, wyw), wxw % 4 == 0
__device__ __forceinline__ void tt_bt_fill(tt_bt_win &S, const tt_dev &d, const float *eff, const uint8_t *ids,
                                           int wxw, int wyw)
{
    __syncthreads();  // the marks stored so far are visible to the loads
    const int lane = threadIdx.x;
    const int g = lane & 15, r0 = lane >> 4;
    int col = wxw + 4 * g;
    col -= (col >= d.W) ? d.W : 0;
    // global loads and LDS stores do not alias: the compiler issues all the
    // loads before the first store
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        int row = wyw + r0 + 4 * k;
        row -= (row >= d.H) ? d.H : 0;
        const size_t q = (size_t)row * d.W + col;
        *(float4 *)&S.e[r0 + 4 * k][4 * g] = *(const float4 *)(eff + q);
        *(uint32_t *)&S.id[r0 + 4 * k][4 * g] = *(const uint32_t *)(ids + q);
    }
    __syncthreads();
}

// One wavefront runs alone: it may use every register of its SIMD, so the
// window fill's 16 staged loads stay in registers (bounded to the default
// occupancy target, they spilled to scratch, a memory round trip per refill).
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(1, 1))) void k_backtrack(
    tt_dev d, const float *eff, uint8_t *ids, int px, int py, int newId, tt_step_ctl *ctl)
{
    __shared__ tt_bt_win S;
    const int W = d.W, H = d.H;
    const bool t0 = threadIdx.x == 0;
    const float INF = tt_u2f(0x7f800000u);
    const int OFF[8][2] = {{-1, 0}, {1, 0}, {0, -1}, {0, 1}, {-1, -1}, {-1, 1}, {1, -1}, {1, 1}};
    auto at = [&](int x, int y) { return (size_t)tt_wrapg(y, H) * W + tt_wrapg(x, W); };
    const uint8_t oldAtPoint = ids[at(px, py)];
    int cx = px, cy = py;
    // window origin: unwrapped (wx, wy) and wrapped (wxw, wyw)
    int wx = 0, wy = 0, wxw = 0, wyw = 0;
    auto recentre = [&]() {
        wx = (cx - TT_BT_WIN / 2) & ~3;
        wy = cy - TT_BT_WIN / 2;
        wxw = tt_wrapg(wx, W);
        wyw = tt_wrapg(wy, H);
        tt_bt_fill(S, d, eff, ids, wxw, wyw);
    };
    recentre();
    // re-centre the window when (cx, cy) and its neighbours are not inside it
    auto cover = [&]() {
        const int lx = cx - wx, ly = cy - wy;
        if (lx < 1 || lx > TT_BT_WIN - 2 || ly < 1 || ly > TT_BT_WIN - 2) recentre();
    };
    auto mark = [&](int lx, int ly) {
        S.id[ly][lx] = (uint8_t)TT_MAXP;
        if (t0) {
            int c = wxw + lx, r = wyw + ly;
            c -= (c >= W) ? W : 0;
            r -= (r >= H) ? H : 0;
            ids[(size_t)r * W + c] = (uint8_t)TT_MAXP;
        }
    };
    bool dead = false;
    int seed = -1;
    // The first walk is memoryless: reaching a cell it already marked means an
    // endless cycle (zero-height plateau); canonical outcome: dead end.
    for (;;) {
        cover();
        const int lx = cx - wx, ly = cy - wy;
        const float ec = S.e[ly][lx];
        const uint8_t ic = S.id[ly][lx];
        if (!(ec > 0.0f)) break;
        if (ic == TT_MAXP) { dead = true; break; }
        mark(lx, ly);
        float best = INF;
        int nx = cx, ny = cy;
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const float en = S.e[ly + OFF[i][1]][lx + OFF[i][0]];
            const bool take = en < best;
            best = take ? en : best;
            nx = take ? cx + OFF[i][0] : nx;
            ny = take ? cy + OFF[i][1] : ny;
        }
        if (best == INF) { dead = true; break; }
        cx = nx;
        cy = ny;
    }
    float pxf = 0.0f, pyf = 0.0f;
    if (!dead) {
        const float fbx = (float)(cx - px), fby = (float)(cy - py);
        const float obx = fbx * -1.0f, oby = fby * -1.0f;
        const float om = tt_fmaxf(sqrtf(obx * obx + oby * oby), 1e-20f);
        const float odx = obx / om, ody = oby / om;
        pxf = -ody;
        pyf = odx;
        // the direction bias of each neighbour is the same at every step
        float bias[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const float ofx = (float)OFF[i][0], ofy = (float)OFF[i][1];
            const float ofm = sqrtf(ofx * ofx + ofy * ofy);
            const float tdx = ofx / ofm, tdy = ofy / ofm;
            const float dot = tdx * odx + tdy * ody;
            bias[i] = -dot * 50.0f;
        }
        cx = px;
        cy = py;
        for (;;) {
            cover();
            const int lx = cx - wx, ly = cy - wy;
            const float ec = S.e[ly][lx];
            if (!(ec > 0.0f)) break;
            mark(lx, ly);
            float best = INF;
            int nx = cx, ny = cy;
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                const int qx = cx + OFF[i][0], qy = cy + OFF[i][1];
                const uint8_t iq = S.id[ly + OFF[i][1]][lx + OFF[i][0]];
                const float dv = S.e[ly + OFF[i][1]][lx + OFF[i][0]];
                const float tnx = (float)(qx - px), tny = (float)(qy - py);
                const float pdot = tnx * odx + tny * ody;
                const float score = dv + bias[i];
                const bool take = (iq != TT_MAXP) & !(pdot < 0.0f) & (score < best);
                best = take ? score : best;
                nx = take ? qx : nx;
                ny = take ? qy : ny;
            }
            if (best != INF) { cx = nx; cy = ny; }
            else { dead = true; break; }
        }
    }
    if (!t0) return;
    if (!dead) {
        int sx = px, sy = py;
        for (int i = 1; i <= 100 && ids[at(sx, sy)] != oldAtPoint; ++i) {
            const int xo = (i % 2 == 1) ? (i + 1) / 2 : -(i / 2);
            const float fx = (float)px + pxf * (float)xo;
            const float fy = (float)py + pyf * (float)xo;
            sx = (int)(fx + 0.5f);
            sy = (int)(fy + 0.5f);
        }
        if (ids[at(sx, sy)] == oldAtPoint) {
            ids[at(sx, sy)] = (uint8_t)newId;
            seed = (int)at(sx, sy);
        }
    }
    ctl->dead_end = dead ? 1 : 0;
    ctl->split_seed = seed;
}

hipError_t tt_launch_backtrack(hipStream_t s, const tt_dev &d, const float *eff, uint8_t *ids, int px, int py,
                               int newId, tt_step_ctl *ctl)
{
    if (d.W < TT_BT_WIN || d.H < TT_BT_WIN) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_backtrack, dim3(1), dim3(64), 0, s, d, eff, ids, px, py, newId, ctl);
    return hipGetLastError();
}

// finalizePlateSplit (:3026-3045): cut-line cells take the new id
__global__ __launch_bounds__(256) void k_split_finalize(tt_dev d, const uint32_t *list, const uint32_t *n,
                                                        uint8_t *ids, int oldId, int newId, tt_plate_data *pd)
{
    const uint32_t k = blockIdx.x * 256 + threadIdx.x;
    if (k < *n) {
        const uint32_t c = list[k];
        if (ids[c] == TT_MAXP) ids[c] = (uint8_t)newId;
    }
    if (k == 0 && newId != oldId) {
        pd[newId].velocity = pd[oldId].velocity;
        pd[newId].velocitySmoothed = pd[oldId].velocitySmoothed;
        pd[newId].direction[0] = pd[oldId].direction[0];
        pd[newId].direction[1] = pd[oldId].direction[1];
        pd[newId].directionSmoothed[0] = pd[oldId].directionSmoothed[0];
        pd[newId].directionSmoothed[1] = pd[oldId].directionSmoothed[1];
    }
}

hipError_t tt_launch_split_finalize(hipStream_t s, const tt_dev &d, const uint32_t *list, const uint32_t *n,
                                    uint8_t *ids, int oldId, int newId, tt_plate_data *pd)
{
    const unsigned blocks = (unsigned)(((long long)d.N + 255) / 256);
    hipLaunchKernelGGL(k_split_finalize, dim3(blocks), dim3(256), 0, s, d, list, n, ids, oldId, newId, pd);
    return hipGetLastError();
}
