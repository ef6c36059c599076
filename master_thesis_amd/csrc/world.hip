// world.hip -- host orchestration of one world (PlateTectonicSim,
// src/cuda_compute/plate_tectonic_sim.{h,cu}) and the C ABI of libtectonic.
//
// Step schedule = executeIteration (plate_tectonic_sim.cu:54-131):
//   collisions -> uplift -> thermal -> hydraulic -> copy-back ->
//   plate movement -> accretion -> CCL (every iterations_for_ccl steps) ->
//   merge + per-plate sums + break score + centres + statistics -> split.
// All device work of a step goes to the world's own stream; the host reads
// back small control values where the reference's control flow depends on
// device data (CCL fragment count, neighbour-fill rounds, split decision and
// relaxation convergence).
#include <unistd.h>
#include <zlib.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>
#include <new>
#include <string>
#include <vector>

#include "tt_kernels.h"
#include "tt_world.h"

namespace {

thread_local std::string g_last_error;

void set_error(const char *what, hipError_t e)
{
    g_last_error = std::string(what) + ": " + hipGetErrorString(e);
}

void set_error(const std::string &what) { g_last_error = what; }

}  // namespace

// error text of the calling thread, shared with batch.hip and server.cpp
void tt_set_error(const std::string &msg) { g_last_error = msg; }
std::string tt_get_error() { return g_last_error; }

namespace {

#define CK(expr)                                        \
    do {                                                \
        hipError_t e_ = (expr);                         \
        if (e_ != hipSuccess) {                         \
            set_error(#expr, e_);                       \
            return TT_ERR_DEVICE;                       \
        }                                               \
    } while (0)

#define CKS(expr)                   \
    do {                            \
        int st_ = (expr);           \
        if (st_ != TT_OK) return st_; \
    } while (0)

constexpr int MAXP = TT_MAXP;

struct SplitMix {
    uint64_t s;
    uint64_t next()
    {
        s += 0x9E3779B97F4A7C15ull;
        uint64_t z = s;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    float u01() { return (float)((double)(next() >> 40) * (1.0 / 16777216.0)); }
};

uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

tt_plate_data default_plate()
{
    tt_plate_data p;
    memset(&p, 0, sizeof(p));
    p.breakScore = 1.0f;
    return p;
}

size_t rup(size_t b) { return (b + 255) & ~(size_t)255; }

}  // namespace

// ---------------------------------------------------------------------------
// allocation
// ---------------------------------------------------------------------------
size_t tt_world_layout(const tt_sim_config &c, bool alloc, tt_world *w)
{
    const size_t N = (size_t)c.width * (size_t)c.height;
    size_t total = 0;
    auto take = [&](void **p, size_t bytes) {
        bytes = rup(bytes);
        total += bytes;
        if (alloc) w->allocs.push_back({p, bytes});
    };
    if (alloc) w->allocs.clear();
    for (int i = 0; i < 2; ++i) {
        take((void **)&w->h[i], N * 4);
        take((void **)&w->ids[i], N);
        take((void **)&w->water[i], N * 4);
        take((void **)&w->flux[i], N * 16);
        take((void **)&w->sed[i], N * 4);
        take((void **)&w->jfa[i], N * 4);
        take((void **)&w->list[i], N * 4);
        take((void **)&w->listN[i], 16);
    }
    take((void **)&w->vel, N * 8);
    take((void **)&w->wn, N * 4);
    take((void **)&w->word, N * 4);
    take((void **)&w->cnt, N);
    take((void **)&w->segCount, (N / 64 + 1) * 4);
    take((void **)&w->segBase, (N / 64 + 1) * 4);
    take((void **)&w->keys, N * 8);
    take((void **)&w->typeCounts, (size_t)TT_NUM_TRI * 5 * 4);
    take((void **)&w->typeBitmap, TT_NUM_TRI);
    take((void **)&w->divBitmap, TT_NUM_TRI_WORDS * 4);
    take((void **)&w->hasDiv, TT_NUM_TRI_WORDS * 4);
    take((void **)&w->accw, N * 4);
    take((void **)&w->accretion, N);
    take((void **)&w->uplift, N * 8);
    take((void **)&w->prop, N * 12);
    take((void **)&w->srcFlag, N);
    take((void **)&w->contested, N);
    take((void **)&w->serialFlag, N);
    take((void **)&w->serialList, N * 4);
    take((void **)&w->serialCount, 16);
    {
        const size_t cap = (size_t)tt_prop_cap((long long)N);
        take((void **)&w->propCell, 5 * cap * 4);
        take((void **)&w->propOff, 5 * cap * 2);
        take((void **)&w->propSlot, 5 * cap * 4);
        take((void **)&w->slotOf, N * 4);
        take((void **)&w->slotCell, TT_PROP_SLOT_WORDS * 4);
    }
    take((void **)&w->dist, N * 8);
    take((void **)&w->pd, 256 * sizeof(tt_plate_data));
    take((void **)&w->ps, 256 * sizeof(tt_plate_step));
    take((void **)&w->ctl, sizeof(tt_step_ctl));
    take((void **)&w->fric, 256 * 8);
    take((void **)&w->noise, 65536);
    take((void **)&w->labels, N * 4);
    take((void **)&w->rootCount, N * 4);
    take((void **)&w->rootFlag, N);
    take((void **)&w->rootList, N * 4);
    take((void **)&w->locList, N * 4);
    take((void **)&w->nloc, 16);
    take((void **)&w->nroots, 16);
    take((void **)&w->fragKey, N * 8);
    take((void **)&w->maxCount, 256 * 4);
    take((void **)&w->newId, N);
    take((void **)&w->unFlag, N);
    take((void **)&w->vote, N);
    take((void **)&w->nbm, (size_t)MAXP * MAXP + 64);
    take((void **)&w->mergeIds, 256 * 4);
    take((void **)&w->acc, 256 * 10 * 8);
    take((void **)&w->eff, N * 4);
    take((void **)&w->rchg, TT_RCHG * 4);
    take((void **)&w->act, 2 * (N / 1024 + 64));
    take((void **)&w->tileMM, (N / 1024 + 64) * 8);
    take((void **)&w->tileFrozen, N / 1024 + 64);
    take((void **)&w->tileChg, 2 * (N / 1024 + 64));
    take((void **)&w->dIter, 4);
    if (c.precision == TT_PRECISION_COMPENSATED) {
        take((void **)&w->ch, N * 4);
        take((void **)&w->cs[0], N * 4);
        take((void **)&w->cs[1], N * 4);
    }
    take((void **)&w->scanTmp, (N / 1024 + tt_scan_tmp_words((long long)N) + 4096) * 4);
    take((void **)&w->initSeeds, TT_MAX_SEEDS * 12);
    take((void **)&w->initWts, 128 * 8);
    return total;
}

static int zero(tt_world *w, void *p, size_t bytes)
{
    CK(tt_fill_u32(w->stream, (uint32_t *)p, 0u, (long long)((bytes + 3) / 4)));
    return TT_OK;
}

static tt_dev make_dev(const tt_world *w)
{
    tt_dev d;
    d.W = w->W;
    d.H = w->H;
    d.N = w->N;
    d.iter = w->iter;
    d.seed = (int)w->cfg.seed;
    d.axs = 256.0 / (double)w->W;
    d.ays = 256.0 / (double)w->H;
    d.ks = w->ks;
    d.rL = w->rL;
    d.divL = w->divL;
    d.sqrtOk = w->sqrtOk;
    d.cvtOk = w->cvtOk;
    tt_dev_set_divisor(d);
    return d;
}

// The hydraulic flux divides by the pipe length.  tt_div_const replaces that
// IEEE division once it has been checked, on the device, to give the same bits
// for every numerator it may receive; the verdict is kept per divisor for the
// process.  Divisors outside [2^-20, 2^20] keep the IEEE division.  Likewise
// tt_sqrt_normal replaces the IEEE square root once it has been checked over
// its whole range, and the rain's hardware float -> uint32 conversion
// (tt_f2u32_hw) replaces the branches of tt_f2u32_sat once it has been checked
// over every float (one verdict each for the process).
static std::mutex g_divMu;
static std::map<uint32_t, int> g_divOk;
static int g_sqrtOk = -1, g_cvtOk = -1;

static int set_divisor(tt_world *w)
{
    const float L = w->ks.hydrationPipeLength;
    volatile float one = 1.0f;
    w->rL = one / L;
    w->divL = 0;
    std::lock_guard<std::mutex> lk(g_divMu);
    if (g_sqrtOk < 0) {
        uint32_t bad[2] = {1, 1};
        CK(tt_check_sqrt(w->stream, bad));
        g_sqrtOk = bad[0] == 0 ? 1 : 0;
        g_cvtOk = bad[1] == 0 ? 1 : 0;
    }
    w->sqrtOk = g_sqrtOk;
    w->cvtOk = g_cvtOk;
    if (!(L >= 0x1p-20f && L <= 0x1p20f)) return TT_OK;
    uint32_t key;
    std::memcpy(&key, &L, 4);
    auto it = g_divOk.find(key);
    if (it == g_divOk.end()) {
        uint32_t bad = 1;
        CK(tt_check_divisor(w->stream, L, w->rL, &bad));
        it = g_divOk.emplace(key, bad == 0 ? 1 : 0).first;
    }
    w->divL = it->second;
    return TT_OK;
}

static int sync_ctl(tt_world *w)
{
    CK(hipMemcpyAsync(w->hctl, w->ctl, sizeof(tt_step_ctl), hipMemcpyDeviceToHost, w->stream));
    CK(hipStreamSynchronize(w->stream));
    return TT_OK;
}

static int read_u32(tt_world *w, const uint32_t *dev, uint32_t *out)
{
    CK(hipMemcpyAsync(w->hu32, dev, 4, hipMemcpyDeviceToHost, w->stream));
    CK(hipStreamSynchronize(w->stream));
    *out = *w->hu32;
    return TT_OK;
}

// ---------------------------------------------------------------------------
// CCL + relabel (applyCCL, plate_tectonic_sim.cu:375-468)
// ---------------------------------------------------------------------------
static int run_ccl(tt_world *w)
{
    hipStream_t s = w->stream;
    const tt_dev d = make_dev(w);
    uint8_t *ids = w->ids[w->ic];
    const size_t N = (size_t)w->N;
    // local roots per tile, unions across tile borders, final roots of the
    // local roots; per-cell labels are completed by the relabel pass
    CK(tt_launch_ccl_local(s, d, ids, w->labels, w->rootFlag, w->rootCount));
    CK(tt_launch_ccl_merge(s, d, ids, w->labels));
    CK(tt_compact_u8_any(s, w->rootFlag, (long long)N, w->locList, w->nloc));
    CK(tt_launch_ccl_flatten(s, w->locList, w->nloc, w->labels, w->rootCount, w->rootFlag));
    CK(tt_compact_u8_any(s, w->rootFlag, (long long)N, w->rootList, w->nroots));
    uint32_t nroots = 0;
    CKS(read_u32(w, w->nroots, &nroots));
    w->stats.num_labels = (int)nroots;
    CKS(zero(w, w->maxCount, 256 * 4));
    CK(tt_launch_frag_build(s, d, w->rootList, w->nroots, nroots, w->rootCount, ids, w->fragKey, w->maxCount,
                            w->newId));
    if (nroots <= 4096) {
        int cap = 2;
        while ((uint32_t)cap < nroots) cap <<= 1;
        CK(tt_launch_frag_sort(s, w->fragKey, w->nroots, cap));
    } else {
        CK(tt_sort_keys_large(s, w->fragKey, w->nroots, nroots));
    }
    CK(tt_launch_frag_assign(s, d, w->fragKey, w->nroots, w->maxCount, ids, w->pd, w->newId));
    CK(tt_launch_relabel(s, d, w->labels, w->newId, ids, w->unFlag));
    CK(tt_compact_u8_any(s, w->unFlag, (long long)N, w->list[0], w->listN[0]));
    uint32_t n = 0;
    CKS(read_u32(w, w->listN[0], &n));
    w->stats.unassigned_cells = (int)n;
    int rounds = 0, cur = 0;
    while (n > 0) {
        ++rounds;
        CK(tt_launch_fill_vote(s, d, w->list[cur], w->listN[cur], ids, w->vote));
        CKS(zero(w, w->listN[1 - cur], 4));
        CK(tt_launch_fill_apply(s, d, w->list[cur], w->listN[cur], w->vote, ids, w->list[1 - cur], w->listN[1 - cur]));
        uint32_t n2 = 0;
        CKS(read_u32(w, w->listN[1 - cur], &n2));
        cur = 1 - cur;
        if (n2 == n) break;  // no cell had an assigned neighbour
        n = n2;
    }
    w->stats.fill_rounds = rounds;
    return TT_OK;
}

// ---------------------------------------------------------------------------
// split (SplitPlateV2, plate_tectonic_sim.cu:933-1040)
// ---------------------------------------------------------------------------
// Runs rounds of a tiled fixed-point kernel in batches, with one host read per
// batch: a round that changed nothing ends the loop (later rounds of the batch
// exit at once).  Round r counts its changes in rchg[r % TT_RCHG]; a batch
// clears its own counters (never the previous round's) before it runs.  The
// first batch has the length the world's last relaxation needed plus two;
// later batches double, up to TT_ROUND_BATCH rounds.
template <typename Launch>
static int run_rounds(tt_world *w, Launch launch, int *rounds_out)
{
    const long long ntiles_all = (long long)(w->W / 64) * (w->H / tt_split_tile_h(make_dev(w)));
    CKS(zero(w, w->act, (size_t)(2 * ntiles_all)));
    int B = std::max(4, std::min(TT_ROUND_BATCH, w->lastRounds + 2));
    for (int r0 = 0; r0 < 3200000; r0 += B, B = std::min(TT_ROUND_BATCH, 2 * B)) {
        const int a = r0 % TT_RCHG, n1 = std::min(B, TT_RCHG - a);
        CKS(zero(w, w->rchg + a, (size_t)n1 * 4));
        if (B > n1) CKS(zero(w, w->rchg, (size_t)(B - n1) * 4));
        CK(launch(r0, B));
        CK(hipMemcpyAsync(w->hu32, w->rchg, TT_RCHG * 4, hipMemcpyDeviceToHost, w->stream));
        CK(hipStreamSynchronize(w->stream));
        for (int r = 0; r < B; ++r)
            if (w->hu32[(r0 + r) % TT_RCHG] == 0u) {
                *rounds_out = r0 + r + 1;
                w->lastRounds = r0 + r + 1;
                return TT_OK;
            }
    }
    set_error("split fixed point did not converge");
    return TT_ERR_STATE;
}

static int run_split(tt_world *w)
{
    hipStream_t s = w->stream;
    const tt_dev d = make_dev(w);
    const int plate = w->hctl->split_plate, nid = w->hctl->split_new;
    w->stats.split_plate = plate;
    w->stats.split_new_plate = nid;
    uint8_t *ids = w->ids[w->ic];
    const size_t N = (size_t)w->N;
    tt_plate_data hp;
    CK(hipMemcpyAsync(&hp, w->pd + plate, sizeof(hp), hipMemcpyDeviceToHost, s));
    CKS(zero(w, &w->ctl->split_negative, 4));
    CK(tt_launch_split_prepare(s, d, ids, plate, w->h[w->hc], w->eff, w->unFlag, w->ctl));
    CK(tt_compact_u8_any(s, w->unFlag, (long long)N, w->list[0], w->listN[0]));
    // tiles holding cells of the plate
    const int th = tt_split_tile_h(d);
    const long long ntiles_all = (long long)(w->W / 64) * (w->H / th);
    CK(tt_launch_split_tiles(s, d, ids, plate, w->vote));
    CK(tt_compact_u8_any(s, w->vote, ntiles_all, w->list[1], w->listN[1]));
    // one read-back: the tile count and the negative-cost flag
    CK(hipMemcpyAsync(w->hu32, w->listN[1], 4, hipMemcpyDeviceToHost, s));
    CK(hipMemcpyAsync(w->hu32 + 1, &w->ctl->split_negative, 4, hipMemcpyDeviceToHost, s));
    CK(hipStreamSynchronize(s));
    const uint32_t ntiles = w->hu32[0];
    if (w->hu32[1]) return TT_OK;   // no fixed point of the relaxation: the split is abandoned
    const uint32_t *tiles = w->list[1];
    const float *height = w->h[w->hc];
    int rounds = 0;
    CKS(run_rounds(w, [&](int r0, int B) {
        return tt_launch_relax_tiles(s, d, tiles, ntiles, ids, plate, height, w->eff, w->rchg, w->act, r0, B);
    }, &rounds));
    w->stats.split_iterations = rounds;
    const int px = (int)(hp.geometricCenter[0] + 0.5f), py = (int)(hp.geometricCenter[1] + 0.5f);
    CK(tt_launch_backtrack(s, d, w->eff, ids, px, py, nid, w->ctl));
    CKS(sync_ctl(w));
    if (w->hctl->dead_end) {
        CK(tt_launch_split_finalize(s, d, w->list[0], w->listN[0], ids, plate, plate, w->pd));
        return TT_OK;
    }
    // flood fill from the seed: components of the old plate touching the seed
    // (CCL scratch buffers are free again at this point of the step)
    CK(tt_launch_flood_cc(s, d, ids, plate, nid, w->rootCount, w->rootFlag, w->locList, w->nloc, w->scanTmp, w->ctl));
    CK(tt_launch_split_finalize(s, d, w->list[0], w->listN[0], ids, plate, nid, w->pd));
    CKS(zero(w, w->acc, 256 * 10 * 8));
    CK(tt_launch_count(s, d, w->mergeIds, ids, w->ids[1 - w->ic], w->h[w->hc], w->noise, w->acc, 1));
    w->ic = 1 - w->ic;
    CK(tt_launch_finalize(s, d, w->pd, w->acc, 1, w->ctl));
    return TT_OK;
}

// ---------------------------------------------------------------------------
// one executeIteration
// ---------------------------------------------------------------------------
static int record(tt_world *w, int slot)
{
    if (w->timing) CK(hipEventRecord(w->ev[slot], w->stream));
    return TT_OK;
}

// executeIteration in three phases: before erosion (plate movement,
// collisions, uplift), erosion (thermal, hydraulic, closing transport) and
// after erosion (copy-back, kinematics, accretion, CCL, merge, split,
// ledger).  A batch of worlds runs the first and last phases per world and the
// erosion of all its worlds together (batch.hip).
// the launches of the pre-erosion phase (collisions, propagation, uplift);
// they depend on the world only through its buffer parities, so a capture
// per parity replays them
static int pre_launch(tt_world *w)
{
    hipStream_t s = w->stream;
    const tt_dev d = make_dev(w);
    const int mask = w->cfg.stage_mask;
    const size_t N = (size_t)w->N;
    CKS(record(w, 0));
    float *hOld = w->h[w->hc], *hNew = w->h[1 - w->hc];
    uint8_t *idOld = w->ids[w->ic], *idNew = w->ids[1 - w->ic];
    CK(tt_launch_plate_prep(s, w->pd, w->ps, w->ctl, w->fric));
    const bool upl = (mask & TT_STAGE_UPLIFT) != 0;
    if (mask & TT_STAGE_COLLIDE) {
        CKS(zero(w, w->typeCounts, (size_t)TT_NUM_TRI * 5 * 4));
        CKS(zero(w, w->hasDiv, TT_NUM_TRI_WORDS * 4));
        CKS(zero(w, w->accw, N * 4));
        if (upl) CKS(zero(w, w->contested, N));
        CK(tt_launch_advect(s, d, idOld, hOld, w->ps, w->ctl, 0, w->typeBitmap, w->word, w->cnt, w->segCount,
                            w->typeCounts, w->ctl));
        CK(tt_scan_u32(s, w->segCount, w->segBase, (long long)(N / 64), w->scanTmp, nullptr));
        CK(tt_launch_keys(s, d, w->word, w->cnt, w->segBase, idOld, w->ps, w->ctl, w->keys));
        CK(tt_launch_type_matrix(s, d, w->typeCounts, w->typeBitmap));
        CK(tt_launch_resolve(s, d, w->dIter, w->word, idOld, hOld, w->ps, w->divBitmap, w->typeBitmap, hNew, idNew,
                             w->uplift, w->accw, w->hasDiv, w->fric, w->srcFlag, w->contested, upl ? 1 : 0));
        CK(tt_launch_flip(s, w->divBitmap, w->hasDiv));
        CKS(record(w, 1));
        if (upl) {
            CK(tt_launch_propagate_local(s, d, w->word, w->uplift, w->srcFlag, w->contested, w->prop, w->jfa[0],
                                         w->serialFlag));
            CK(tt_compact_u8(s, w->serialFlag, (long long)N, w->serialList, w->serialCount, w->scanTmp));
            CK(tt_launch_propagate_serial(s, d, w->serialList, w->serialCount, w->word, w->uplift, hOld, w->ps,
                                          w->prop, tt_prop_cap((long long)N), w->propCell, w->propOff,
                                          w->propSlot, w->slotOf, w->slotCell));
            CK(tt_launch_combine_contested(s, d, w->word, w->uplift, w->contested, w->srcFlag, w->prop, w->jfa[0],
                                           w->ctl));
            if (w->cfg.uplift_mode == TT_UPLIFT_SEPARABLE) {
                CK(tt_launch_separable_uplift(s, d, w->prop, idNew, w->ps, hNew, w->dist));
            } else {
                const int maxDim = std::max(w->W, w->H);
                int step = 1;
                while (step < maxDim) step <<= 1;
                step >>= 1;
                int jc = 0;
                while (step > 0) {
                    CK(tt_launch_jfa_pass(s, d, w->jfa[jc], w->jfa[1 - jc], step));
                    jc = 1 - jc;
                    step >>= 1;
                }
                CK(tt_launch_uplift(s, d, w->jfa[jc], w->prop, idNew, w->ps, hNew, w->dist));
            }
        }
    } else {
        CK(hipMemcpyAsync(hNew, hOld, N * 4, hipMemcpyDeviceToDevice, s));
        CK(hipMemcpyAsync(idNew, idOld, N, hipMemcpyDeviceToDevice, s));
        CKS(zero(w, w->accw, N * 4));
        CKS(zero(w, w->fric, 256 * 8));
        CKS(record(w, 1));
    }
    CKS(record(w, 2));
    // the height compensation belongs to a cell's height: it restarts when the
    // collision stage has moved heights between cells
    if (w->ch && (mask & TT_STAGE_COLLIDE)) CKS(zero(w, w->ch, N * 4));
    return TT_OK;
}

static void drop_pre_graphs(tt_world *w)
{
    for (auto &g : w->preExec) {
        if (g) (void)hipGraphExecDestroy(g);
        g = nullptr;
    }
}

static int capture_pre(tt_world *w, hipGraphExec_t *out)
{
    hipGraph_t g = nullptr;
    if (hipStreamBeginCapture(w->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) return TT_ERR_DEVICE;
    const int st = pre_launch(w);
    const hipError_t e = hipStreamEndCapture(w->stream, &g);
    int res = (st == TT_OK && e == hipSuccess && g) ? TT_OK : TT_ERR_DEVICE;
    if (res == TT_OK && hipGraphInstantiate(out, g, nullptr, nullptr, 0) != hipSuccess) res = TT_ERR_DEVICE;
    if (g) (void)hipGraphDestroy(g);
    if (res != TT_OK) {
        (void)hipGetLastError();
        *out = nullptr;
    }
    return res;
}

int tt_step_pre(tt_world *w)
{
    w->stats.overflow_cells = 0;
    w->stats.complex_sources = 0;
    w->stats.num_labels = 0;
    w->stats.unassigned_cells = 0;
    w->stats.fill_rounds = 0;
    w->stats.merges = 0;
    w->stats.split_plate = -1;
    w->stats.split_new_plate = -1;
    w->stats.split_iterations = 0;
    // the step's iteration for the device (the collision resolution's
    // divergence noise reads it there, so captured launches stay valid)
    *w->hIter = w->iter;
    CK(hipMemcpyAsync(w->dIter, w->hIter, 4, hipMemcpyHostToDevice, w->stream));
    if (w->preGraphs && !w->timing) {
        hipGraphExec_t &g = w->preExec[w->hc * 2 + w->ic];
        if (!g && capture_pre(w, &g) != TT_OK) w->preGraphs = false;   // launches from now on
        if (g) {
            CK(hipGraphLaunch(g, w->stream));
        } else {
            CKS(pre_launch(w));
        }
    } else {
        CKS(pre_launch(w));
    }
    w->ehc = 1 - w->hc;   // erosion works on the written height field
    return TT_OK;
}

int tt_step_thermal_count(const tt_world *w)
{
    return (w->cfg.stage_mask & TT_STAGE_THERMAL) ? w->cfg.thermal_iterations : 0;
}

int tt_step_hydraulic_count(const tt_world *w)
{
    return (w->cfg.stage_mask & TT_STAGE_HYDRAULIC) ? w->cfg.hydraulic_iterations : 0;
}

// buffer indices after T thermal passes (tt_thermal_launches of them) and K
// hydraulic sub-iterations; returns the sediment buffer holding the last
// sub-iteration's output (input of the closing transport) in *sedLast
void tt_step_erode_indices(tt_world *w, int T, int K, int *sedLast)
{
    w->ehc ^= (tt_thermal_launches(T, w->ch != nullptr) + K) & 1;
    w->wc = tt_hyd_water_index_after(w->wc, K);
    w->fc ^= K & 1;
    *sedLast = K > 0 ? (w->sc ^ ((K - 1) & 1)) : w->sc;
}

// the closing sediment transport of a step whose hydraulic loop ran K > 0
// sub-iterations (sediment output in sed[sedLast])
int tt_step_transport(tt_world *w, int sedLast)
{
    const tt_dev d = make_dev(w);
    CK(tt_launch_transport(w->stream, d, w->sed[sedLast], w->vel, w->sed[1 - sedLast], w->ch ? w->cs[sedLast] : nullptr,
                           w->ch ? w->cs[1 - sedLast] : nullptr));
    w->sc = 1 - sedLast;
    return TT_OK;
}

// per-tile frozen flags of this step's thermal passes (tiles that no pass changes)
int tt_step_thermal_prep(tt_world *w)
{
    const int T = tt_step_thermal_count(w);
    if (T <= 0) return TT_OK;
    const tt_dev d = make_dev(w);
    CK(tt_launch_thermal_frozen(w->stream, d, T, w->h[w->ehc], w->tileMM, w->tileFrozen));
    return TT_OK;
}

static int step_erode(tt_world *w)
{
    hipStream_t s = w->stream;
    const tt_dev d = make_dev(w);
    const int T = tt_step_thermal_count(w), K = tt_step_hydraulic_count(w);
    int hc = w->ehc;
    CKS(tt_step_thermal_prep(w));
    // an odd first pass alone, then pairs of passes per launch
    const int TL = tt_thermal_launches(T, w->ch != nullptr), singles = TL - tt_thermal_pairs(T, w->ch != nullptr);
    for (int l = 0; l < TL; ++l) {
        CK(tt_launch_thermal(s, d, TT_TILE_H, w->h[hc], w->h[1 - hc], w->ch, w->tileFrozen, w->tileChg, l,
                             l >= singles));
        hc = 1 - hc;
    }
    CKS(record(w, 3));
    const int rseed = (int)(w->cfg.seed + w->iter);
    // online tile tuning: groups of G middle sub-iterations per candidate height
    int G = 0, ncand = 0;
    if (w->tuning) {
        for (int rows : {16, 8, 32})
            if (tt_tile_rows_valid(rows, w->H)) w->tuneRows[ncand++] = rows;
        G = ncand > 0 ? std::min(16, (K - 2) / ncand) : 0;
        if (G < 4) G = 0;
    }
    int fc = w->fc;
    for (int it = 0; it < K; ++it) {
        const int si = it == 0 ? w->sc : (w->sc ^ ((it - 1) & 1)), so = it == 0 ? w->sc : (w->sc ^ (it & 1));
        float *sprev = w->sed[si], *sout = w->sed[so];
        const float *win;
        float *wout;
        tt_hyd_water_bufs(w->water[0], w->water[1], w->wn, w->wc, it, K, &win, &wout);
        const int grp = (G > 0 && it >= 1 && it < 1 + G * ncand) ? (it - 1) / G : -1;
        const int rows = grp >= 0 ? w->tuneRows[grp] : w->rows;
        if (grp >= 0 && (it - 1) % G == 0) CK(hipEventRecord(w->tuneEv[2 * grp], s));
        CK(tt_launch_hydraulic(s, d, rows, it == 0, it == K - 1, rseed, w->h[hc], w->h[1 - hc], win, wout,
                               w->flux[fc], w->flux[1 - fc], w->vel, sprev, sout, w->ch, w->ch ? w->cs[si] : nullptr,
                               w->ch ? w->cs[so] : nullptr));
        if (grp >= 0 && (it - 1) % G == G - 1) {
            CK(hipEventRecord(w->tuneEv[2 * grp + 1], s));
            w->tuneCount[grp] = G;
        }
        hc = 1 - hc;
        fc = 1 - fc;
    }
    w->tuneGroups = G > 0 ? ncand : 0;
    CKS(record(w, 4));
    int sedLast = 0;
    tt_step_erode_indices(w, T, K, &sedLast);
    if (K > 0) CKS(tt_step_transport(w, sedLast));
    CKS(record(w, 5));
    return TT_OK;
}

static int tune_finish(tt_world *w);

int tt_step_post(tt_world *w)
{
    hipStream_t s = w->stream;
    const tt_dev d = make_dev(w);
    const int mask = w->cfg.stage_mask;
    const size_t N = (size_t)w->N;
    const bool upl = (mask & TT_STAGE_UPLIFT) != 0;
    // copy-back (plate_tectonic_sim.cu:77-78) is a change of buffer index
    w->hc = w->ehc;
    w->ic = 1 - w->ic;
    if (mask & TT_STAGE_KINEMATICS) {
        CK(tt_launch_plate_movement(s, d, w->pd, w->fric, w->noise));
        CK(tt_launch_accretion_apply(s, d, w->accw, w->h[w->hc], w->ids[w->ic], w->ids[1 - w->ic], w->accretion));
        w->ic = 1 - w->ic;
    }
    if ((mask & TT_STAGE_CCL) && (w->iter % (uint32_t)w->cfg.iterations_for_ccl) == 0) CKS(run_ccl(w));
    if (mask & TT_STAGE_MERGE) {
        CKS(zero(w, w->nbm, (size_t)MAXP * MAXP + 4));
        CK(tt_launch_neighbor_matrix(s, d, w->ids[w->ic], w->nbm));
        CK(tt_launch_merge_decide(s, d, w->nbm, w->pd, w->mergeIds, w->typeBitmap, w->ctl));
        CKS(zero(w, w->acc, 256 * 10 * 8));
        CK(tt_launch_count(s, d, w->mergeIds, w->ids[w->ic], w->ids[1 - w->ic], w->h[w->hc], w->noise, w->acc, 0));
        w->ic = 1 - w->ic;
        CK(tt_launch_finalize(s, d, w->pd, w->acc, (mask & TT_STAGE_SPLIT) ? 0 : 2, w->ctl));
    }
    CKS(record(w, 6));
    CKS(sync_ctl(w));
    if ((mask & TT_STAGE_MERGE) && (mask & TT_STAGE_SPLIT) && w->hctl->split_plate >= 0) CKS(run_split(w));
    CK(tt_launch_ledger(s, w->h[w->hc], (long long)N, &w->ctl->ledger));
    CKS(record(w, 7));
    CKS(sync_ctl(w));
    if (w->tuneGroups > 0) CKS(tune_finish(w));
    const tt_step_ctl &c = *w->hctl;
    w->stats.overflow_cells = (mask & TT_STAGE_COLLIDE) ? c.overflow_cells : 0;
    w->stats.complex_sources = ((mask & TT_STAGE_COLLIDE) && upl) ? c.complex_sources : 0;
    if (mask & TT_STAGE_MERGE) {
        w->stats.merges = c.merges;
        w->stats.largest_value = c.largest_value;
        w->stats.largest_plate_id = c.largest_id;
        w->stats.heaviest_value = c.heaviest_value;
        w->stats.heaviest_plate_id = c.heaviest_id;
        w->stats.number_of_plates = c.n_plates;
    }
    w->stats.mass_ledger = tt_unfix_d(c.ledger, TT_MASS_BITS);
    w->iter++;
    w->stats.iteration = w->iter;
    w->stats.tile_rows = w->rows;
    w->stats.hydraulic_div_const = w->divL;
    w->stats.hydraulic_sqrt_fast = w->sqrtOk;
    if (w->timing) {
        float t[8] = {0};
        for (int i = 1; i <= 7; ++i) CK(hipEventElapsedTime(&t[i], w->ev[i - 1], w->ev[i]));
        w->times.ms_collide += t[1];
        w->times.ms_uplift += t[2];
        w->times.ms_thermal += t[3];
        w->times.ms_hydraulic += t[4];
        w->times.ms_transport += t[5];
        w->times.ms_post += t[6];
        w->times.ms_split += t[7];
        float tot = 0;
        CK(hipEventElapsedTime(&tot, w->ev[0], w->ev[7]));
        w->times.ms_total += tot;
        w->times.hydraulic_launches += (float)tt_step_hydraulic_count(w);
        w->times.thermal_launches += (float)tt_thermal_launches(tt_step_thermal_count(w), w->ch != nullptr);
        w->times.steps += 1.0f;
    }
    return TT_OK;
}

static int step_once(tt_world *w)
{
    CKS(tt_step_pre(w));
    CKS(step_erode(w));
    return tt_step_post(w);
}

// ---------------------------------------------------------------------------
// initialisation (initializeTectonics, plate_tectonic_sim.cu:705-773)
// ---------------------------------------------------------------------------
static int upload_seeds(tt_world *w, const std::vector<float> &sx, const std::vector<float> &sy,
                        const std::vector<uint8_t> &sid)
{
    const size_t n = sx.size();
    if (n > TT_MAX_SEEDS) { set_error("too many Voronoi seeds"); return TT_ERR_UNSUPPORTED; }
    std::vector<uint8_t> blob(TT_MAX_SEEDS * 12, 0);
    memcpy(blob.data(), sx.data(), n * 4);
    memcpy(blob.data() + TT_MAX_SEEDS * 4, sy.data(), n * 4);
    memcpy(blob.data() + TT_MAX_SEEDS * 8, sid.data(), n);
    CK(hipMemcpyAsync(w->initSeeds, blob.data(), blob.size(), hipMemcpyHostToDevice, w->stream));
    CK(hipStreamSynchronize(w->stream));
    const tt_dev d = make_dev(w);
    const float *dsx = (const float *)w->initSeeds;
    const float *dsy = (const float *)(w->initSeeds + TT_MAX_SEEDS * 4);
    const uint8_t *dsid = (const uint8_t *)(w->initSeeds + TT_MAX_SEEDS * 8);
    CK(tt_launch_init_plate_ids(w->stream, d, dsx, dsy, dsid, (int)n, w->ids[w->ic]));
    return TT_OK;
}

// generatePlateCenters (plate_tectonic_sim.cu:873-890): plate i sits at
// (U[0, height), U[0, width)) -- x drawn over the height range and y over the
// width range, as in the reference -- with id i.
static void gen_centers(SplitMix &rng, int n, int W, int H, std::vector<tt_voronoi_seed> &out)
{
    out.assign((size_t)n, tt_voronoi_seed{});
    for (int i = 0; i < n; ++i) {
        out[i].position[0] = rng.u01() * (float)H;
        out[i].position[1] = rng.u01() * (float)W;
        out[i].id = (uint8_t)i;
    }
}

// generateVoronoiSeeds (plate_tectonic_sim.cu:892-928): each new seed takes the
// id of its nearest seed of the previous generation (toroidal squared
// distance, first index on ties).
static void gen_seeds(SplitMix &rng, const std::vector<tt_voronoi_seed> &prev, int n, int W, int H,
                      std::vector<tt_voronoi_seed> &out)
{
    out.assign((size_t)n, tt_voronoi_seed{});
    for (int i = 0; i < n; ++i) {
        const float px = rng.u01() * (float)H;
        const float py = rng.u01() * (float)W;
        float best = 3.402823466e38f;
        uint8_t cl = 0;
        for (const tt_voronoi_seed &q : prev) {
            const float dx = tt_fabsf(px - q.position[0]), dy = tt_fabsf(py - q.position[1]);
            const float wx = tt_fminf(dx, (float)W - dx), wy = tt_fminf(dy, (float)H - dy);
            const float d2 = wx * wx + wy * wy;
            if (d2 < best) { cl = q.id; best = d2; }
        }
        out[i].position[0] = px;
        out[i].position[1] = py;
        out[i].id = cl;
    }
}

// generatePlateData (plate_tectonic_sim.cu:816-842): pixel centre = fractional
// part of the centre, speed (U(-1,1) + 1) / 2 (or U[lo, hi] for the synthetic
// worlds), direction = normalised pair of U(-1,1); 256 entries (the
// reference's 255 plus the spare slot of the device table).
static void gen_plate_data(SplitMix &rng, const std::vector<tt_voronoi_seed> &centers, int P, bool reference_speed,
                           float lo, float hi, std::vector<tt_plate_data> &pd)
{
    pd.assign(256, default_plate());
    for (int i = 0; i < P; ++i) {
        tt_plate_data &p = pd[i];
        p.pixelCenter[0] = centers[i].position[0] - tt_floorf(centers[i].position[0]);
        p.pixelCenter[1] = centers[i].position[1] - tt_floorf(centers[i].position[1]);
        if (reference_speed) p.velocity = ((rng.u01() * 2.0f - 1.0f) + 1.0f) / 2.0f;
        else p.velocity = lo + (hi - lo) * rng.u01();
        p.velocitySmoothed = p.velocity;
        const float xd = rng.u01() * 2.0f - 1.0f;
        const float yd = rng.u01() * 2.0f - 1.0f;
        const float m = sqrtf(xd * xd + yd * yd);
        if (m > 0.0f) { p.direction[0] = xd / m; p.direction[1] = yd / m; }
        else { p.direction[0] = 1.0f; p.direction[1] = 0.0f; }
        p.directionSmoothed[0] = p.direction[0];
        p.directionSmoothed[1] = p.direction[1];
    }
}

static int init_common_zero(tt_world *w)
{
    const size_t N = (size_t)w->N;
    for (int i = 0; i < 2; ++i) {
        CKS(zero(w, w->water[i], N * 4));
        CKS(zero(w, w->flux[i], N * 16));
        CKS(zero(w, w->sed[i], N * 4));
    }
    CKS(zero(w, w->vel, N * 8));
    CKS(zero(w, w->wn, N * 4));
    if (w->ch) {
        CKS(zero(w, w->ch, N * 4));
        CKS(zero(w, w->cs[0], N * 4));
        CKS(zero(w, w->cs[1], N * 4));
    }
    CKS(zero(w, w->word, N * 4));
    CKS(zero(w, w->accretion, N));
    CKS(zero(w, w->accw, N * 4));
    CKS(zero(w, w->uplift, N * 8));
    CKS(zero(w, w->prop, N * 12));
    CKS(zero(w, w->divBitmap, TT_NUM_TRI_WORDS * 4));
    CKS(zero(w, w->typeBitmap, TT_NUM_TRI));
    CKS(zero(w, w->keys, N * 8));
    CKS(zero(w, w->labels, N * 4));
    CKS(zero(w, w->dist, N * 8));
    return TT_OK;
}

static int init_counts(tt_world *w)
{
    const tt_dev d = make_dev(w);
    CKS(zero(w, w->acc, 256 * 10 * 8));
    CK(tt_launch_count(w->stream, d, w->mergeIds, w->ids[w->ic], w->ids[1 - w->ic], w->h[w->hc], w->noise, w->acc, 1));
    w->ic = 1 - w->ic;
    CK(tt_launch_finalize(w->stream, d, w->pd, w->acc, 3, w->ctl));
    return TT_OK;
}

int tt_world_do_init(tt_world *w)
{
    hipStream_t s = w->stream;
    drop_pre_graphs(w);   // a reset may change the seed, an argument of the captured launches
    CKS(set_divisor(w));
    w->iter = 0;
    w->hc = w->ic = w->wc = w->fc = w->sc = 0;
    CKS(init_common_zero(w));
    const tt_dev d = make_dev(w);
    const int mode = w->cfg.init_mode;
    if (mode == TT_INIT_EMPTY) {
        std::vector<tt_plate_data> pd(256, default_plate());
        CK(hipMemcpyAsync(w->pd, pd.data(), 256 * sizeof(tt_plate_data), hipMemcpyHostToDevice, s));
        CKS(zero(w, w->h[0], (size_t)w->N * 4));
        CKS(zero(w, w->ids[0], (size_t)w->N));
        CK(hipStreamSynchronize(s));
        w->initialized = true;
        return TT_OK;
    }
    SplitMix rng{(uint64_t)w->cfg.seed};
    const int P = w->cfg.num_plates;
    std::vector<tt_voronoi_seed> centers, seeds, next;
    gen_centers(rng, P, w->W, w->H, centers);
    seeds = centers;
    if (mode == TT_INIT_REFERENCE) {
        // every generation of numVoronoiSeeds in turn; fewer seeds than plates
        // fall back to the plate centres (plate_tectonic_sim.cu:720-728)
        for (int n : w->gens) {
            gen_seeds(rng, seeds, n, w->W, w->H, next);
            seeds.swap(next);
        }
        if (seeds.size() < centers.size()) seeds = centers;
    }
    std::vector<tt_plate_data> pd;
    gen_plate_data(rng, centers, P, mode == TT_INIT_REFERENCE, w->cfg.speed_min, w->cfg.speed_max, pd);
    CK(hipMemcpyAsync(w->pd, pd.data(), 256 * sizeof(tt_plate_data), hipMemcpyHostToDevice, s));
    std::vector<float> sx(seeds.size()), sy(seeds.size());
    std::vector<uint8_t> sid(seeds.size());
    for (size_t i = 0; i < seeds.size(); ++i) {
        sx[i] = seeds[i].position[0];
        sy[i] = seeds[i].position[1];
        sid[i] = seeds[i].id;
    }
    if (mode == TT_INIT_SYNTHETIC) {
        double wts[65], tot = 0.0;
        for (int k = -32; k <= 32; ++k) { wts[k + 32] = tt_exp_core(-(double)(k * k) / 128.0); tot = tot + wts[k + 32]; }
        for (int k = 0; k < 65; ++k) wts[k] = wts[k] / tot;
        CK(hipMemcpyAsync(w->initWts, wts, sizeof(wts), hipMemcpyHostToDevice, s));
        const uint64_t hseed = mix64((uint64_t)w->cfg.seed ^ 0x6A09E667F3BCC909ull);
        // the two fp64 blur buffers borrow the flux arrays (16 B/cell each)
        CK(tt_launch_init_synthetic_heights(s, d, hseed, (double *)w->flux[0], (double *)w->flux[1],
                                            (const double *)w->initWts, w->h[w->hc]));
    } else {
        CK(tt_launch_init_reference_heights(s, d, w->h[w->hc]));
    }
    CKS(upload_seeds(w, sx, sy, sid));
    CKS(init_counts(w));
    for (int i = 0; i < 2; ++i) CKS(zero(w, w->flux[i], (size_t)w->N * 16));
    CKS(run_ccl(w));
    CK(hipStreamSynchronize(s));
    w->initialized = true;
    return TT_OK;
}

// ---------------------------------------------------------------------------
// field access
// ---------------------------------------------------------------------------
static bool field_ptr(tt_world *w, int f, void **p, size_t *bytes, bool write)
{
    const size_t N = (size_t)w->N;
    switch (f) {
        case TT_FIELD_HEIGHT: *p = w->h[w->hc]; *bytes = N * 4; return true;
        case TT_FIELD_PLATE_IDS: *p = w->ids[w->ic]; *bytes = N; return true;
        case TT_FIELD_WATER: *p = w->water[w->wc]; *bytes = N * 4; return true;
        case TT_FIELD_FLUX: *p = w->flux[w->fc]; *bytes = N * 16; return true;
        case TT_FIELD_VELOCITY: *p = w->vel; *bytes = N * 8; return true;
        case TT_FIELD_SEDIMENT: *p = w->sed[w->sc]; *bytes = N * 4; return true;
        case TT_FIELD_ASTHENOSPHERE: if (write) return false; *p = nullptr; *bytes = N * 8; return true;
        case TT_FIELD_COLLISIONS: *p = w->word; *bytes = N * 4; return true;
        case TT_FIELD_ACCRETION: *p = w->accretion; *bytes = N; return true;
        case TT_FIELD_UPLIFT: *p = w->uplift; *bytes = N * 8; return true;
        case TT_FIELD_PROPAGATED: *p = w->prop; *bytes = N * 12; return true;
        case TT_FIELD_PLATE_DATA: *p = w->pd; *bytes = sizeof(tt_plate_data) * MAXP; return true;
        case TT_FIELD_DIVERGENCE_BITMAP: *p = w->divBitmap; *bytes = TT_NUM_TRI_WORDS * 4; return true;
        case TT_FIELD_COLLISION_TYPES: *p = w->typeBitmap; *bytes = TT_NUM_TRI; return true;
        case TT_FIELD_SORTED_KEYS: *p = w->keys; *bytes = N * 8; return true;
        case TT_FIELD_LABELS: *p = w->labels; *bytes = N * 4; return true;
        case TT_FIELD_DISTANCE: *p = w->dist; *bytes = N * 8; return true;
        default: return false;
    }
}

int tt_world_geometry(tt_world *w, int32_t *width, int32_t *height, tt_kernel_settings *ks)
{
    if (!w || !width || !height || !ks) return TT_ERR_INVALID;
    *width = w->W;
    *height = w->H;
    *ks = w->ks;
    return TT_OK;
}

// ---------------------------------------------------------------------------
// tile autotuner: rows of the erosion tiles per size class
// ---------------------------------------------------------------------------
// The hydraulic sub-iteration is timed for each tile height that divides the
// world's height (8, 16, 32 rows), on the world's own buffers before
// initialisation overwrites them; the fastest is kept per (device, size class),
// with size class = floor(log2(cells)).  Results are cached in the process and,
// when TECTONIC_TUNE_FILE names a file, in a flat JSON object
// {"<arch>/<class>": rows, ...} that later processes read.
namespace {
std::mutex g_tune_mu;
std::map<std::string, int> g_tune;
bool g_tune_loaded = false;

std::string tune_key(const tt_world *w)
{
    hipDeviceProp_t prop;
    std::string arch = "gpu";
    if (hipGetDeviceProperties(&prop, w->device) == hipSuccess) arch = prop.gcnArchName;
    const std::string::size_type colon = arch.find(':');
    if (colon != std::string::npos) arch = arch.substr(0, colon);
    int cls = 0;
    const long long n = (long long)w->W * w->H;
    while ((1ll << (cls + 1)) <= n) ++cls;
    return arch + "/" + std::to_string(cls);
}

void tune_load()
{
    g_tune_loaded = true;
    const char *path = std::getenv("TECTONIC_TUNE_FILE");
    if (!path) return;
    std::ifstream f(path);
    if (!f) return;
    std::stringstream ss;
    ss << f.rdbuf();
    const std::string t = ss.str();
    size_t i = 0;
    while ((i = t.find('"', i)) != std::string::npos) {
        const size_t j = t.find('"', i + 1);
        if (j == std::string::npos) break;
        const std::string key = t.substr(i + 1, j - i - 1);
        const size_t colon = t.find(':', j);
        if (colon == std::string::npos) break;
        const int v = std::atoi(t.c_str() + colon + 1);
        if (v == 8 || v == 16 || v == 32) g_tune[key] = v;
        i = t.find_first_of(",}", colon);
        if (i == std::string::npos) break;
    }
}

// written to a temporary file and renamed over the old one, so that another
// process never reads a half-written file
void tune_save()
{
    const char *path = std::getenv("TECTONIC_TUNE_FILE");
    if (!path) return;
    const std::string tmp = std::string(path) + ".tmp" + std::to_string((long long)getpid());
    {
        std::ofstream f(tmp);
        if (!f) return;
        f << "{";
        bool first = true;
        for (const auto &kv : g_tune) {
            f << (first ? "" : ", ") << "\"" << kv.first << "\": " << kv.second;
            first = false;
        }
        f << "}\n";
        if (!f) {
            f.close();
            std::remove(tmp.c_str());
            return;
        }
    }
    if (std::rename(tmp.c_str(), path) != 0) std::remove(tmp.c_str());
}
}  // namespace

// A cached choice for the world's size class sets the tile height at once.
// Without one (or with tile_rows = -1) the world starts on 16-row tiles and
// tunes online: its first step with enough hydraulic sub-iterations runs
// groups of its middle sub-iterations at each candidate height (all heights
// give the same bits), timed with events on the world's stream, and
// tune_finish keeps the fastest per launch for the size class.
static int tune_rows(tt_world *w, bool force)
{
    std::lock_guard<std::mutex> lock(g_tune_mu);
    if (!g_tune_loaded) tune_load();
    const std::string key = tune_key(w);
    const auto it = g_tune.find(key);
    if (!force && it != g_tune.end() && tt_tile_rows_valid(it->second, w->H)) {
        w->rows = it->second;
        w->tuning = false;
        return TT_OK;
    }
    w->rows = 16;
    w->tuning = true;
    return TT_OK;
}

static int tune_finish(tt_world *w)
{
    int best = w->rows;
    float bestMs = 1e30f;
    for (int i = 0; i < w->tuneGroups; ++i) {
        float ms = 0.0f;
        CK(hipEventElapsedTime(&ms, w->tuneEv[2 * i], w->tuneEv[2 * i + 1]));
        const float per = ms / (float)w->tuneCount[i];
        if (per < bestMs) { bestMs = per; best = w->tuneRows[i]; }
    }
    w->tuneGroups = 0;
    w->tuning = false;
    w->rows = best;
    std::lock_guard<std::mutex> lock(g_tune_mu);
    g_tune[tune_key(w)] = best;
    tune_save();
    return TT_OK;
}

int tt_validate_config(const tt_sim_config *c)
{
    if (c->width < 64 || c->height < 64 || c->width % TT_TILE_W || c->height % TT_TILE_H) {
        set_error("width must be a multiple of 64 and height a multiple of 16, both >= 64");
        return TT_ERR_INVALID;
    }
    if (c->width > 16384 || c->height > 16384) { set_error("sides above 16384 are not supported"); return TT_ERR_UNSUPPORTED; }
    if (c->num_plates < 1 || c->num_plates > 254) { set_error("num_plates must be in 1..254"); return TT_ERR_INVALID; }
    if (c->iterations_for_ccl < 1) { set_error("iterations_for_ccl must be >= 1"); return TT_ERR_INVALID; }
    if (c->hydraulic_iterations < 0 || c->thermal_iterations < 0) { set_error("negative iteration count"); return TT_ERR_INVALID; }
    if (c->num_voronoi_gens < 0 || c->num_voronoi_gens > 4) { set_error("num_voronoi_gens must be in 0..4"); return TT_ERR_INVALID; }
    if (c->init_mode < 0 || c->init_mode > 2) { set_error("bad init_mode"); return TT_ERR_INVALID; }
    if (c->uplift_mode != TT_UPLIFT_JFA && c->uplift_mode != TT_UPLIFT_SEPARABLE) {
        set_error("uplift_mode must be TT_UPLIFT_JFA or TT_UPLIFT_SEPARABLE");
        return TT_ERR_INVALID;
    }
    if (c->precision != TT_PRECISION_FP32 && c->precision != TT_PRECISION_COMPENSATED) {
        set_error("precision must be TT_PRECISION_FP32 or TT_PRECISION_COMPENSATED");
        return TT_ERR_INVALID;
    }
    if (c->tile_rows != 0 && c->tile_rows != -1 && !tt_tile_rows_valid(c->tile_rows, c->height)) {
        set_error("tile_rows must be 0 or -1 (tuned), or 8, 16 or 32 dividing the height");
        return TT_ERR_INVALID;
    }
    return TT_OK;
}

// Seed counts of the Voronoi generations: each at least 1; the last one is
// uploaded to the device's seed table of TT_MAX_SEEDS entries.
static int validate_gens(const std::vector<int32_t> &gens)
{
    for (int32_t n : gens)
        if (n < 1) { set_error("every Voronoi generation needs at least one seed"); return TT_ERR_INVALID; }
    if (!gens.empty() && gens.back() > TT_MAX_SEEDS) {
        set_error("the last Voronoi generation has more than " + std::to_string(TT_MAX_SEEDS) + " seeds");
        return TT_ERR_UNSUPPORTED;
    }
    return TT_OK;
}

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

const char *tt_status_string(int status)
{
    switch (status) {
        case TT_OK: return "ok";
        case TT_ERR_INVALID: return "invalid argument";
        case TT_ERR_DEVICE: return "device error";
        case TT_ERR_NOMEM: return "out of device memory";
        case TT_ERR_STATE: return "call out of order";
        case TT_ERR_UNSUPPORTED: return "unsupported";
        case TT_ERR_BUDGET: return "request exceeds the HBM budget";
        default: return "unknown status";
    }
}

const char *tt_last_error(void) { return g_last_error.c_str(); }

int tt_device_count(int *count)
{
    if (!count) return TT_ERR_INVALID;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) n = 0;
    *count = n;
    return TT_OK;
}

int tt_settings_default(tt_kernel_settings *k)
{
    if (!k) return TT_ERR_INVALID;
    tt_settings_fill_defaults(k);
    return TT_OK;
}

int tt_config_default(tt_sim_config *c)
{
    if (!c) return TT_ERR_INVALID;
    memset(c, 0, sizeof(*c));
    c->width = 1024;
    c->height = 1024;
    c->seed = 1000;
    c->num_plates = 16;
    c->num_voronoi_gens = 2;
    c->voronoi_gens[0] = 200;
    c->voronoi_gens[1] = 500;
    c->iterations_for_ccl = 10;
    c->hydraulic_iterations = 10;
    c->thermal_iterations = 1;
    c->init_mode = TT_INIT_REFERENCE;
    c->speed_min = 0.0f;
    c->speed_max = 1.0f;
    c->stage_mask = TT_STAGE_ALL;
    return TT_OK;
}

int tt_world_bytes(const tt_sim_config *cfg, size_t *bytes)
{
    if (!cfg || !bytes) return TT_ERR_INVALID;
    tt_world tmp;
    *bytes = tt_world_layout(*cfg, false, &tmp);
    return TT_OK;
}

int tt_world_create(const tt_sim_config *cfg, const tt_kernel_settings *ks, tt_world **out)
{
    if (!cfg || !ks || !out) return TT_ERR_INVALID;
    *out = nullptr;
    CKS(tt_validate_config(cfg));
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) { set_error("no HIP device"); return TT_ERR_DEVICE; }
    tt_world *w = new (std::nothrow) tt_world();
    if (!w) return TT_ERR_NOMEM;
    w->cfg = *cfg;
    w->ks = *ks;
    w->W = cfg->width;
    w->H = cfg->height;
    w->N = cfg->width * cfg->height;
    w->rows = cfg->tile_rows;
    w->gens.assign(cfg->voronoi_gens, cfg->voronoi_gens + cfg->num_voronoi_gens);
    if (validate_gens(w->gens) != TT_OK) {
        delete w;
        return TT_ERR_INVALID;
    }
    CK(hipGetDevice(&w->device));
    tt_world_layout(*cfg, true, w);
    size_t total = 0;
    for (auto &a : w->allocs) total += a.bytes;
    char *base = nullptr;
    if (hipMalloc((void **)&base, total) != hipSuccess) {
        delete w;
        set_error("hipMalloc of " + std::to_string(total) + " bytes failed");
        return TT_ERR_NOMEM;
    }
    w->base = base;
    size_t off = 0;
    for (auto &a : w->allocs) { *a.ptr = base + off; off += a.bytes; }
    if (hipStreamCreateWithFlags(&w->stream, hipStreamNonBlocking) != hipSuccess ||
        hipHostMalloc((void **)&w->hctl, sizeof(tt_step_ctl), hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void **)&w->hu32, TT_RCHG * 4, hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void **)&w->hIter, 64, hipHostMallocDefault) != hipSuccess) {
        tt_world_destroy(w);
        set_error("stream / pinned memory allocation failed");
        return TT_ERR_DEVICE;
    }
    for (int i = 0; i < 8; ++i)
        if (hipEventCreate(&w->ev[i]) != hipSuccess) { tt_world_destroy(w); return TT_ERR_DEVICE; }
    for (int i = 0; i < 6; ++i)
        if (hipEventCreate(&w->tuneEv[i]) != hipSuccess) { tt_world_destroy(w); return TT_ERR_DEVICE; }
    memset(&w->stats, 0, sizeof(w->stats));
    memset(&w->times, 0, sizeof(w->times));
    w->stats.split_plate = w->stats.split_new_plate = -1;
    const std::vector<uint8_t> &noise = tt_noise_table();
    if (noise.size() != 65536) {
        tt_world_destroy(w);
        set_error("noise table missing (master_thesis_amd/data/noise_r8.bin)");
        return TT_ERR_STATE;
    }
    if (hipMemcpy(w->noise, noise.data(), 65536, hipMemcpyHostToDevice) != hipSuccess ||
        tt_fill_u32(w->stream, w->slotOf, 0xFFFFFFFFu, (long long)w->N) != hipSuccess ||
        hipStreamSynchronize(w->stream) != hipSuccess) {
        tt_world_destroy(w);
        return TT_ERR_DEVICE;
    }
    *out = w;
    return TT_OK;
}

int tt_world_init(tt_world *w)
{
    if (!w) return TT_ERR_INVALID;
    if (w->rows <= 0) CKS(tune_rows(w, w->cfg.tile_rows < 0));
    CKS(tt_world_do_init(w));
    w->stats.tile_rows = w->rows;
    w->stats.hydraulic_div_const = w->divL;
    w->stats.hydraulic_sqrt_fast = w->sqrtOk;
    return TT_OK;
}

int tt_world_reset(tt_world *w, uint32_t seed, int32_t num_plates, const int32_t *voronoi_gens, int32_t num_gens)
{
    if (!w || num_gens < 0 || (num_gens > 0 && !voronoi_gens)) return TT_ERR_INVALID;
    tt_sim_config c = w->cfg;
    c.seed = seed;
    c.num_plates = num_plates;
    c.num_voronoi_gens = 0;   // the generations live in w->gens (any number)
    CKS(tt_validate_config(&c));
    std::vector<int32_t> gens(voronoi_gens, voronoi_gens + num_gens);
    CKS(validate_gens(gens));
    c.num_voronoi_gens = num_gens <= 4 ? num_gens : 4;
    for (int i = 0; i < c.num_voronoi_gens; ++i) c.voronoi_gens[i] = gens[(size_t)i];
    w->cfg = c;
    w->gens = gens;
    if (w->rows <= 0) CKS(tune_rows(w, w->cfg.tile_rows < 0));
    CKS(tt_world_do_init(w));
    w->stats.tile_rows = w->rows;
    w->stats.hydraulic_div_const = w->divL;
    w->stats.hydraulic_sqrt_fast = w->sqrtOk;
    return TT_OK;
}

int tt_generate_plate_centers(uint64_t *rng_state, int32_t num_plates, int32_t width, int32_t height,
                              tt_voronoi_seed *out)
{
    if (!rng_state || num_plates < 0 || (num_plates > 0 && !out) || width < 1 || height < 1) return TT_ERR_INVALID;
    SplitMix rng{*rng_state};
    std::vector<tt_voronoi_seed> v;
    gen_centers(rng, num_plates, width, height, v);
    std::copy(v.begin(), v.end(), out);
    *rng_state = rng.s;
    return TT_OK;
}

int tt_generate_voronoi_seeds(uint64_t *rng_state, const tt_voronoi_seed *centers, int32_t num_centers,
                              int32_t num_seeds, int32_t width, int32_t height, tt_voronoi_seed *out)
{
    if (!rng_state || num_centers < 0 || num_seeds < 0 || (num_centers > 0 && !centers) || (num_seeds > 0 && !out) ||
        width < 1 || height < 1)
        return TT_ERR_INVALID;
    SplitMix rng{*rng_state};
    const std::vector<tt_voronoi_seed> prev(centers, centers + num_centers);
    std::vector<tt_voronoi_seed> v;
    gen_seeds(rng, prev, num_seeds, width, height, v);
    std::copy(v.begin(), v.end(), out);
    *rng_state = rng.s;
    return TT_OK;
}

int tt_generate_plate_data(uint64_t *rng_state, const tt_voronoi_seed *centers, int32_t num_plates,
                           tt_plate_data *out)
{
    if (!rng_state || !out || num_plates < 0 || num_plates > MAXP || (num_plates > 0 && !centers)) return TT_ERR_INVALID;
    SplitMix rng{*rng_state};
    const std::vector<tt_voronoi_seed> c(centers, centers + num_plates);
    std::vector<tt_plate_data> pd;
    gen_plate_data(rng, c, num_plates, true, 0.0f, 0.0f, pd);
    std::copy(pd.begin(), pd.begin() + MAXP, out);
    *rng_state = rng.s;
    return TT_OK;
}

int tt_world_config(tt_world *w, tt_sim_config *cfg, tt_kernel_settings *ks)
{
    if (!w) return TT_ERR_INVALID;
    if (cfg) *cfg = w->cfg;
    if (ks) *ks = w->ks;
    return TT_OK;
}

int tt_world_step(tt_world *w, int32_t n)
{
    if (!w || n < 0) return TT_ERR_INVALID;
    if (!w->initialized) { set_error("tt_world_step before tt_world_init"); return TT_ERR_STATE; }
    if (w->batches > 0) { set_error("tt_world_step on a batch member (remove it from the batch first)"); return TT_ERR_STATE; }
    for (int i = 0; i < n; ++i) CKS(step_once(w));
    CK(hipStreamSynchronize(w->stream));
    return TT_OK;
}

int tt_world_set_settings(tt_world *w, const tt_kernel_settings *ks)
{
    if (!w || !ks) return TT_ERR_INVALID;
    w->ks = *ks;
    drop_pre_graphs(w);   // the settings are arguments of the captured launches
    return set_divisor(w);
}

int tt_world_download(tt_world *w, int32_t field, void *host, size_t bytes)
{
    if (!w || !host) return TT_ERR_INVALID;
    void *p;
    size_t b;
    if (!field_ptr(w, field, &p, &b, false) || b != bytes) { set_error("bad field or size"); return TT_ERR_INVALID; }
    if (field == TT_FIELD_ASTHENOSPHERE) {
        float2 *tmp = (float2 *)w->fragKey;  // N * 8 bytes of CCL scratch
        const tt_dev d = make_dev(w);
        CK(tt_launch_asthenosphere(w->stream, d, w->noise, tmp));
        CK(hipMemcpyAsync(host, tmp, b, hipMemcpyDeviceToHost, w->stream));
        CK(hipStreamSynchronize(w->stream));
        return TT_OK;
    }
    CK(hipMemcpyAsync(host, p, b, hipMemcpyDeviceToHost, w->stream));
    CK(hipStreamSynchronize(w->stream));
    return TT_OK;
}

int tt_world_upload(tt_world *w, int32_t field, const void *host, size_t bytes)
{
    if (!w || !host) return TT_ERR_INVALID;
    void *p;
    size_t b;
    if (!field_ptr(w, field, &p, &b, true) || b != bytes) { set_error("bad field, read-only field or size"); return TT_ERR_INVALID; }
    CK(hipMemcpyAsync(p, host, b, hipMemcpyHostToDevice, w->stream));
    CK(hipStreamSynchronize(w->stream));
    w->initialized = true;
    return TT_OK;
}

// ---- snapshot file ------------------------------------------------------
// "TTSNAP2\n", u32 header bytes, tt_sim_config, tt_kernel_settings, u32
// iteration, tt_step_stats, u32 field count; per field: i32 id, u64 bytes, u32
// CRC32 of the bytes (zlib crc32_z over the whole length), the bytes.  Worlds
// in TT_PRECISION_COMPENSATED also store their Kahan compensation terms (ids
// SNAP_HEIGHT_COMP and SNAP_SEDIMENT_COMP), which carry over between steps.
// Version-1 files (no compensation fields) still load.
static const int kSnapFields[] = {TT_FIELD_HEIGHT, TT_FIELD_PLATE_IDS, TT_FIELD_WATER, TT_FIELD_FLUX,
                                  TT_FIELD_VELOCITY, TT_FIELD_SEDIMENT, TT_FIELD_UPLIFT, TT_FIELD_PLATE_DATA,
                                  TT_FIELD_DIVERGENCE_BITMAP, TT_FIELD_COLLISION_TYPES, TT_FIELD_LABELS};
enum { SNAP_HEIGHT_COMP = 1000, SNAP_SEDIMENT_COMP = 1001 };
static const char kSnapMagic[8] = {'T', 'T', 'S', 'N', 'A', 'P', '2', '\n'};
static const char kSnapMagicV1[8] = {'T', 'T', 'S', 'N', 'A', 'P', '1', '\n'};

static uint32_t snap_crc(const uint8_t *p, size_t n) { return (uint32_t)crc32_z(0L, p, (z_size_t)n); }

// device buffer of a snapshot field id (public fields and the compensation terms)
static bool snap_ptr(tt_world *w, int id, void **p, size_t *bytes)
{
    const size_t N = (size_t)w->N;
    if (id == SNAP_HEIGHT_COMP || id == SNAP_SEDIMENT_COMP) {
        if (!w->ch) return false;
        *p = id == SNAP_HEIGHT_COMP ? (void *)w->ch : (void *)w->cs[w->sc];
        *bytes = N * 4;
        return true;
    }
    return field_ptr(w, id, p, bytes, false);
}

struct snap_header {
    tt_sim_config cfg;
    tt_kernel_settings ks;
    uint32_t iteration;
    tt_step_stats stats;
    uint32_t nfields;
};

int tt_world_save(tt_world *w, const char *path)
{
    if (!w || !path) return TT_ERR_INVALID;
    if (!w->initialized) { set_error("tt_world_save before tt_world_init"); return TT_ERR_STATE; }
    FILE *f = fopen(path, "wb");
    if (!f) { set_error(std::string("cannot open ") + path); return TT_ERR_INVALID; }
    snap_header hd;
    memset(&hd, 0, sizeof hd);
    hd.cfg = w->cfg;
    hd.ks = w->ks;
    hd.iteration = w->iter;
    hd.stats = w->stats;
    std::vector<int> ids(std::begin(kSnapFields), std::end(kSnapFields));
    if (w->ch) {
        ids.push_back(SNAP_HEIGHT_COMP);
        ids.push_back(SNAP_SEDIMENT_COMP);
    }
    hd.nfields = (uint32_t)ids.size();
    const uint32_t hb = (uint32_t)sizeof hd;
    bool ok = fwrite(kSnapMagic, 8, 1, f) == 1 && fwrite(&hb, 4, 1, f) == 1 && fwrite(&hd, sizeof hd, 1, f) == 1;
    std::vector<uint8_t> buf;
    for (int fid : ids) {
        if (!ok) break;
        void *p;
        size_t b;
        if (!snap_ptr(w, fid, &p, &b)) { ok = false; break; }
        buf.resize(b);
        if (fid == SNAP_HEIGHT_COMP || fid == SNAP_SEDIMENT_COMP) {
            if (hipMemcpyAsync(buf.data(), p, b, hipMemcpyDeviceToHost, w->stream) != hipSuccess ||
                hipStreamSynchronize(w->stream) != hipSuccess) {
                fclose(f);
                set_error("downloading a compensation field failed");
                return TT_ERR_DEVICE;
            }
        } else {
            const int st = tt_world_download(w, fid, buf.data(), b);
            if (st != TT_OK) { fclose(f); return st; }
        }
        const int32_t id = fid;
        const uint64_t bytes = b;
        const uint32_t crc = snap_crc(buf.data(), b);
        ok = fwrite(&id, 4, 1, f) == 1 && fwrite(&bytes, 8, 1, f) == 1 && fwrite(&crc, 4, 1, f) == 1 &&
             fwrite(buf.data(), 1, b, f) == b;
    }
    if (fclose(f) != 0) ok = false;
    if (!ok) { set_error(std::string("writing ") + path + " failed"); return TT_ERR_INVALID; }
    return TT_OK;
}

int tt_world_load(const char *path, tt_world **out)
{
    if (!path || !out) return TT_ERR_INVALID;
    *out = nullptr;
    FILE *f = fopen(path, "rb");
    if (!f) { set_error(std::string("cannot open ") + path); return TT_ERR_INVALID; }
    char magic[8];
    uint32_t hb = 0;
    snap_header hd;
    if (fread(magic, 8, 1, f) != 1 || (memcmp(magic, kSnapMagic, 8) != 0 && memcmp(magic, kSnapMagicV1, 8) != 0) ||
        fread(&hb, 4, 1, f) != 1 || hb != sizeof hd || fread(&hd, sizeof hd, 1, f) != 1) {
        fclose(f);
        set_error(std::string(path) + " is not a world snapshot (version 1 or 2)");
        return TT_ERR_INVALID;
    }
    // read and check every field before any device work; a field may not
    // claim more bytes than the file still holds (nothing is allocated for a
    // length the file cannot back)
    const long here = ftell(f);
    uint64_t left = 0;
    if (here >= 0 && fseek(f, 0, SEEK_END) == 0) {
        const long end = ftell(f);
        if (end >= here) left = (uint64_t)(end - here);
    }
    if (here < 0 || fseek(f, here, SEEK_SET) != 0) {
        fclose(f);
        set_error(std::string(path) + ": cannot seek");
        return TT_ERR_INVALID;
    }
    struct blob { int32_t id; std::vector<uint8_t> data; };
    std::vector<blob> blobs;
    for (uint32_t k = 0; k < hd.nfields; ++k) {
        int32_t id = 0;
        uint64_t bytes = 0;
        uint32_t crc = 0;
        if (left < 16 || fread(&id, 4, 1, f) != 1 || fread(&bytes, 8, 1, f) != 1 || fread(&crc, 4, 1, f) != 1 ||
            bytes > (uint64_t)16384 * 16384 * 16 || bytes > left - 16) {
            fclose(f);
            set_error(std::string(path) + ": truncated field header");
            return TT_ERR_INVALID;
        }
        blob b;
        b.id = id;
        b.data.resize((size_t)bytes);
        if (fread(b.data.data(), 1, (size_t)bytes, f) != (size_t)bytes ||
            snap_crc(b.data.data(), (size_t)bytes) != crc) {
            fclose(f);
            set_error(std::string(path) + ": field " + std::to_string(id) + " is truncated or corrupt");
            return TT_ERR_INVALID;
        }
        left -= 16 + bytes;
        blobs.push_back(std::move(b));
    }
    fclose(f);
    CKS(tt_validate_config(&hd.cfg));
    tt_sim_config c = hd.cfg;
    c.init_mode = TT_INIT_EMPTY;
    tt_world *w = nullptr;
    CKS(tt_world_create(&c, &hd.ks, &w));
    int st = tt_world_init(w);
    for (const blob &b : blobs) {
        if (st != TT_OK) break;
        if (b.id == SNAP_HEIGHT_COMP || b.id == SNAP_SEDIMENT_COMP) {
            void *p;
            size_t bytes;
            if (!snap_ptr(w, b.id, &p, &bytes) || bytes != b.data.size()) {
                set_error(std::string(path) + ": compensation field of a world that is not compensated, or bad size");
                st = TT_ERR_INVALID;
            } else if (hipMemcpyAsync(p, b.data.data(), bytes, hipMemcpyHostToDevice, w->stream) != hipSuccess ||
                       hipStreamSynchronize(w->stream) != hipSuccess) {
                set_error("uploading a compensation field failed");
                st = TT_ERR_DEVICE;
            }
            continue;
        }
        st = tt_world_upload(w, b.id, b.data.data(), b.data.size());
    }
    if (st != TT_OK) {
        tt_world_destroy(w);
        return st;
    }
    w->cfg = hd.cfg;
    w->iter = hd.iteration;
    w->stats = hd.stats;
    *out = w;
    return TT_OK;
}

int tt_world_stats(tt_world *w, tt_step_stats *out)
{
    if (!w || !out) return TT_ERR_INVALID;
    *out = w->stats;
    return TT_OK;
}

int tt_world_iteration(tt_world *w, uint32_t *it)
{
    if (!w || !it) return TT_ERR_INVALID;
    *it = w->iter;
    return TT_OK;
}

int tt_world_set_iteration(tt_world *w, uint32_t it)
{
    if (!w) return TT_ERR_INVALID;
    w->iter = it;
    return TT_OK;
}

int tt_world_set_timing(tt_world *w, int32_t enabled)
{
    if (!w) return TT_ERR_INVALID;
    w->timing = enabled != 0;
    memset(&w->times, 0, sizeof(w->times));
    return TT_OK;
}

int tt_world_stage_times(tt_world *w, tt_stage_times *out)
{
    if (!w || !out) return TT_ERR_INVALID;
    *out = w->times;
    return TT_OK;
}

int tt_world_destroy(tt_world *w)
{
    if (!w) return TT_ERR_INVALID;
    if (w->batches > 0) {
        set_error("tt_world_destroy of a batch member (remove it from the batch first)");
        return TT_ERR_STATE;
    }
    if (w->stream) (void)hipStreamSynchronize(w->stream);
    for (int i = 0; i < 8; ++i)
        if (w->ev[i]) (void)hipEventDestroy(w->ev[i]);
    for (int i = 0; i < 6; ++i)
        if (w->tuneEv[i]) (void)hipEventDestroy(w->tuneEv[i]);
    if (w->stream) (void)hipStreamDestroy(w->stream);
    drop_pre_graphs(w);
    if (w->hctl) (void)hipHostFree(w->hctl);
    if (w->hu32) (void)hipHostFree(w->hu32);
    if (w->hIter) (void)hipHostFree(w->hIter);
    if (w->base) (void)hipFree(w->base);
    delete w;
    return TT_OK;
}

int tt_math_eval_device(int32_t fn, const float *in, float *out, int64_t n)
{
    if (!in || !out || n < 0 || fn < TT_MATH_EXPF || fn > TT_MATH_DIV_RCP) return TT_ERR_INVALID;
    if (n == 0) return TT_OK;
    const int per = (fn == TT_MATH_SIMPLEX || fn == TT_MATH_DISCRETE) ? 3
                    : (fn == TT_MATH_FMODF || fn == TT_MATH_ATAN2 || fn == TT_MATH_DIV_RCP) ? 2 : 1;
    float *din = nullptr, *dout = nullptr;
    CK(hipMalloc((void **)&din, (size_t)n * per * 4));
    if (hipMalloc((void **)&dout, (size_t)n * 4) != hipSuccess) { (void)hipFree(din); return TT_ERR_NOMEM; }
    int st = TT_OK;
    if (hipMemcpy(din, in, (size_t)n * per * 4, hipMemcpyHostToDevice) != hipSuccess ||
        tt_launch_math_eval(0, fn, din, dout, n) != hipSuccess || hipDeviceSynchronize() != hipSuccess ||
        hipMemcpy(out, dout, (size_t)n * 4, hipMemcpyDeviceToHost) != hipSuccess) {
        set_error("device math evaluation failed");
        st = TT_ERR_DEVICE;
    }
    (void)hipFree(din);
    (void)hipFree(dout);
    return st;
}

int tt_probe_hbm(int64_t bytes, int32_t reps, tt_hbm_probe *out)
{
    if (!out || bytes < 16 || (bytes % 16) != 0 || reps < 1 || bytes > (4LL << 30)) {
        set_error("tt_probe_hbm: bytes must be a multiple of 16 in [16, 4 GiB] and reps >= 1");
        return TT_ERR_INVALID;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) { set_error("no HIP device"); return TT_ERR_DEVICE; }
    memset(out, 0, sizeof *out);
    out->bytes = bytes;
    out->reps = reps;
    float4 *a = nullptr, *b = nullptr;
    float *sink = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int st = TT_OK;
    const long long n = bytes / 16;
    if (hipMalloc((void **)&a, (size_t)bytes) != hipSuccess || hipMalloc((void **)&b, (size_t)bytes) != hipSuccess ||
        hipMalloc((void **)&sink, 256) != hipSuccess) {
        set_error("tt_probe_hbm: device allocation failed");
        st = TT_ERR_NOMEM;
    }
    if (st == TT_OK && (hipStreamCreate(&s) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
                        hipEventCreate(&e1) != hipSuccess))
        st = TT_ERR_DEVICE;
    if (st == TT_OK && (tt_fill_u32(s, (uint32_t *)a, 0u, n * 4) != hipSuccess ||
                        tt_fill_u32(s, (uint32_t *)b, 0u, n * 4) != hipSuccess))
        st = TT_ERR_DEVICE;
    double best[4] = {0.0, 0.0, 0.0, 0.0};
    for (int which = 0; which < 4 && st == TT_OK; ++which) {
        // one untimed launch, then reps timed launches one at a time
        if (tt_launch_probe(s, which, a, b, sink, n) != hipSuccess) { st = TT_ERR_DEVICE; break; }
        for (int r = 0; r < reps; ++r) {
            float ms = 0.0f;
            if (hipEventRecord(e0, s) != hipSuccess || tt_launch_probe(s, which, a, b, sink, n) != hipSuccess ||
                hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
                st = TT_ERR_DEVICE;
                break;
            }
            // the stream probe moves 64 bytes per cell of (bytes / 32) & ~3 cells
            const double moved = which == 3 ? 64.0 * (double)((bytes / 32) & ~3LL)
                                            : (which == 0 ? 2.0 : 1.0) * (double)bytes;
            if (ms > 0.0f) best[which] = std::max(best[which], moved / (ms * 1e-3) / 1e9);
        }
    }
    if (st == TT_ERR_DEVICE) set_error("tt_probe_hbm: device error");
    out->copy_gbps = best[0];
    out->read_gbps = best[1];
    out->write_gbps = best[2];
    out->streams_gbps = best[3];
    if (s) (void)hipStreamSynchronize(s);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (s) (void)hipStreamDestroy(s);
    (void)hipFree(a);
    (void)hipFree(b);
    (void)hipFree(sink);
    return st;
}

}  // extern "C"
