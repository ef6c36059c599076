// batch.hip -- many worlds stepped together (tt_batch_* in tectonic.h).
//
// A step of a batch runs, for every member:
//   1. the pre-erosion phase (plate movement, collisions, uplift) on the
//      member's own stream, from a pool of host threads;
//   2. the erosion of all members at once: thermal passes 2j, 2j + 1 (after
//      an odd first pass) and hydraulic sub-iteration k of every member in
//      one launch each (k_batch.hip), the
//      whole pass sequence captured once as a hipGraph and replayed every
//      step (its kernels read the per-step buffer indices from a descriptor
//      table copied to the device before the replay);
//   3. the post-erosion phase (transport, kinematics, CCL, merge, split,
//      ledger) on the member's stream, from the thread pool.
// Streams are ordered by events: the batch stream waits for every member's
// pre phase, every member stream waits for the erosion.  A member's result is
// bit-identical to stepping it alone: the batched kernels run the same tile
// bodies, and nothing is shared between members.
#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <functional>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "tt_batch.h"
#include "tt_kernels.h"
#include "tt_world.h"

void tt_set_error(const std::string &msg);
std::string tt_get_error();

#define BCK(x)                                                                  \
    do {                                                                        \
        hipError_t e_ = (x);                                                    \
        if (e_ != hipSuccess) {                                                 \
            tt_set_error(std::string(#x) + ": " + hipGetErrorString(e_));       \
            return TT_ERR_DEVICE;                                               \
        }                                                                       \
    } while (0)
#define BCKS(x)                   \
    do {                          \
        int s_ = (x);             \
        if (s_ != TT_OK) return s_; \
    } while (0)

// Host threads that live as long as the batch: run(f) has every thread (and
// the caller) execute f once and returns when all have finished.
struct tt_pool {
    std::vector<std::thread> th;
    std::mutex mu;
    std::condition_variable wake, idle;
    std::function<void()> job;
    uint64_t gen = 0;
    int busy = 0;
    bool stop = false;

    // A new worker starts from the current generation: after a restart (the
    // member count changed) gen is not 0, and a worker starting from 0 would
    // run the last job again, whose lambda refers to the stack of a call
    // that has returned.
    void start(int n, int device)
    {
        uint64_t g0;
        {
            std::lock_guard<std::mutex> lk(mu);
            g0 = gen;
            busy = 0;
        }
        for (int i = 0; i < n; ++i)
            th.emplace_back([this, device, g0]() {
                (void)hipSetDevice(device);
                uint64_t seen = g0;
                for (;;) {
                    std::function<void()> f;
                    {
                        std::unique_lock<std::mutex> lk(mu);
                        wake.wait(lk, [&] { return stop || gen != seen; });
                        if (stop) return;
                        seen = gen;
                        f = job;
                    }
                    f();
                    std::lock_guard<std::mutex> lk(mu);
                    if (--busy == 0) idle.notify_all();
                }
            });
    }
    void run(const std::function<void()> &f)
    {
        {
            std::lock_guard<std::mutex> lk(mu);
            job = f;
            busy = (int)th.size();
            ++gen;
        }
        wake.notify_all();
        f();
        std::unique_lock<std::mutex> lk(mu);
        idle.wait(lk, [&] { return busy == 0; });
        job = nullptr;   // its captures die with the caller's frame
    }
    void shutdown()
    {
        {
            std::lock_guard<std::mutex> lk(mu);
            stop = true;
        }
        wake.notify_all();
        for (auto &t : th) t.join();
        th.clear();
        stop = false;
    }
    ~tt_pool() { shutdown(); }
};

struct tt_batch {
    int device = 0;
    tt_pool pool;               // threads - 1 workers beside the calling thread
    hipStream_t stream = nullptr;
    hipEvent_t erosionDone = nullptr;
    hipEvent_t tStart = nullptr, tPre = nullptr, tEro = nullptr, tEnd = nullptr;   // phase timing
    tt_batch_times times{};
    std::vector<tt_world *> worlds;
    std::vector<hipEvent_t> preDone, postDone;
    std::vector<int> sedLast, Ts, Ks;
    int threads = 16;
    bool dirty = true;
    // device tables
    tt_bworld *dDesc = nullptr;
    tt_bworld *hDesc = nullptr;   // pinned
    size_t descCap = 0;
    uint32_t *dThermTiles = nullptr;
    int nThermTiles = 0;
    uint32_t *dHydTiles[3] = {nullptr, nullptr, nullptr};   // rows 8, 16, 32
    int nHydTiles[3] = {0, 0, 0};
    // captured erosion
    hipGraphExec_t exec = nullptr;
    int graphT = -1, graphK = -1;
    bool useGraph = true;
};

namespace {

int rows_slot(int rows) { return rows == 8 ? 0 : rows == 32 ? 2 : 1; }
int slot_rows(int slot) { return slot == 0 ? 8 : slot == 2 ? 32 : 16; }

void drop_graph(tt_batch *b)
{
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    b->exec = nullptr;
    b->graphT = b->graphK = -1;
}

void free_tables(tt_batch *b)
{
    if (b->dDesc) (void)hipFree(b->dDesc);
    if (b->hDesc) (void)hipHostFree(b->hDesc);
    if (b->dThermTiles) (void)hipFree(b->dThermTiles);
    for (auto &p : b->dHydTiles)
        if (p) (void)hipFree(p);
    b->dDesc = nullptr;
    b->hDesc = nullptr;
    b->dThermTiles = nullptr;
    for (auto &p : b->dHydTiles) p = nullptr;
    b->descCap = 0;
}

int upload_table(tt_batch *b, const std::vector<uint32_t> &v, uint32_t **dst)
{
    if (v.empty()) return TT_OK;
    BCK(hipMalloc((void **)dst, v.size() * 4));
    BCK(hipMemcpyAsync(*dst, v.data(), v.size() * 4, hipMemcpyHostToDevice, b->stream));
    BCK(hipStreamSynchronize(b->stream));
    return TT_OK;
}

// descriptor tables and tile tables for the current membership
int rebuild(tt_batch *b)
{
    drop_graph(b);
    BCK(hipStreamSynchronize(b->stream));
    free_tables(b);
    const size_t n = b->worlds.size();
    b->descCap = n;
    BCK(hipMalloc((void **)&b->dDesc, n * sizeof(tt_bworld)));
    BCK(hipHostMalloc((void **)&b->hDesc, n * sizeof(tt_bworld), hipHostMallocDefault));
    std::vector<uint32_t> therm, hyd[3];
    for (size_t i = 0; i < n; ++i) {
        const tt_world *w = b->worlds[i];
        const uint32_t tag = (uint32_t)i << 20;
        const int nbx = w->W / TT_TILE_W;   // at most 256 (8 bits); rows at most 2048 (12 bits)
        for (int ty = 0; ty < w->H / TT_TILE_H; ++ty)
            for (int tx = 0; tx < nbx; ++tx) therm.push_back(tag | ((uint32_t)ty << 8) | (uint32_t)tx);
        const int slot = rows_slot(w->rows);
        for (int ty = 0; ty < w->H / w->rows; ++ty)
            for (int tx = 0; tx < nbx; ++tx) hyd[slot].push_back(tag | ((uint32_t)ty << 8) | (uint32_t)tx);
    }
    BCKS(upload_table(b, therm, &b->dThermTiles));
    b->nThermTiles = (int)therm.size();
    for (int s = 0; s < 3; ++s) {
        BCKS(upload_table(b, hyd[s], &b->dHydTiles[s]));
        b->nHydTiles[s] = (int)hyd[s].size();
    }
    b->preDone.resize(n, nullptr);
    for (auto &e : b->preDone)
        if (!e) BCK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    b->postDone.resize(n, nullptr);
    for (auto &e : b->postDone)
        if (!e) BCK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    b->sedLast.assign(n, 0);
    b->Ts.assign(n, 0);
    b->Ks.assign(n, 0);
    b->dirty = false;
    return TT_OK;
}

// runs f(world index) for every member on the batch's host threads
template <typename F>
int parallel(tt_batch *b, F f)
{
    const int n = (int)b->worlds.size();
    const int nt = std::max(1, std::min(b->threads, n));
    if ((int)b->pool.th.size() != nt - 1) {
        b->pool.shutdown();
        b->pool.start(nt - 1, b->device);
    }
    std::atomic<int> next{0};
    std::atomic<int> status{TT_OK};
    std::mutex mu;
    std::string msg;
    auto work = [&]() {
        for (;;) {
            const int i = next.fetch_add(1);
            if (i >= n || status.load() != TT_OK) break;
            const int st = f(i);
            if (st != TT_OK) {
                int expected = TT_OK;
                if (status.compare_exchange_strong(expected, st)) {
                    std::lock_guard<std::mutex> lock(mu);
                    msg = "batch member " + std::to_string(i) + ": " + tt_get_error();
                }
            }
        }
    };
    if (nt == 1) work();
    else b->pool.run(work);
    if (status.load() != TT_OK) tt_set_error(msg);
    return status.load();
}

void fill_desc(tt_bworld &B, const tt_world *w, int T, int K)
{
    const tt_kernel_settings &k = w->ks;
    B.W = w->W;
    B.H = w->H;
    B.T = T;
    B.K = K;
    B.hc0 = w->ehc;
    B.hc1 = w->ehc ^ (tt_thermal_launches(T, false) & 1);
    B.wc = w->wc;
    B.fc = w->fc;
    B.sc = w->sc;
    B.rseed = (int)(w->cfg.seed + w->iter);
    B.talus = k.thermalThresholdAngle * k.thermalCellSize;
    B.amp = k.thermalErosionAmplitude;
    B.str = k.thermalErosionStrength;
    B.rain = k.hydrationRainfall;
    B.evap = k.hydrationEvaporation;
    B.cap = k.sedimentCapacity;
    B.ke = k.sedimentErosionRate;
    B.kd = k.sedimentDepositionRate;
    B.fk = 1.0f * k.hydrationPipeCrossSection;
    B.g = k.gravity;
    B.L = k.hydrationPipeLength;
    B.rL = w->rL;
    B.divL = w->divL;
    B.sqrtOk = w->sqrtOk;
    B.cvtOk = w->cvtOk;
    B.sin003 = tt_sinf(0.03f);
    for (int i = 0; i < 2; ++i) {
        B.h[i] = w->h[i];
        B.water[i] = w->water[i];
        B.flux[i] = w->flux[i];
        B.sed[i] = w->sed[i];
    }
    B.vel = w->vel;
    B.wn = w->wn;
    B.frozen = w->tileFrozen;
    B.chg = w->tileChg;
}

// Tcode = 2 * (most pairs of thermal passes of a member) + (some member has an
// odd pass count)
int launch_erosion(tt_batch *b, int Tcode, int Kmax)
{
    if (Tcode & 1) BCK(tt_launch_thermal_batch(b->stream, b->dDesc, b->dThermTiles, b->nThermTiles, 0, false));
    for (int j = 0; j < Tcode / 2; ++j)
        BCK(tt_launch_thermal_batch(b->stream, b->dDesc, b->dThermTiles, b->nThermTiles, j, true));
    for (int k = 0; k < Kmax; ++k)
        for (int s = 0; s < 3; ++s)
            BCK(tt_launch_hydraulic_batch(b->stream, slot_rows(s), b->dDesc, b->dHydTiles[s], b->nHydTiles[s], k));
    return TT_OK;
}

// replays the captured pass sequence; captures it when the pass counts changed
int run_erosion(tt_batch *b, int Tcode, int Kmax)
{
    if (b->useGraph && (b->graphT != Tcode || b->graphK != Kmax)) {
        drop_graph(b);
        hipGraph_t g = nullptr;
        bool ok = hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal) == hipSuccess;
        if (ok) {
            const int st = launch_erosion(b, Tcode, Kmax);
            ok = hipStreamEndCapture(b->stream, &g) == hipSuccess && st == TT_OK && g != nullptr;
        }
        if (ok) ok = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0) == hipSuccess;
        if (g) (void)hipGraphDestroy(g);
        if (ok) {
            b->graphT = Tcode;
            b->graphK = Kmax;
        } else {
            (void)hipGetLastError();
            drop_graph(b);
            b->useGraph = false;   // direct launches from now on
        }
    }
    if (b->useGraph && b->exec) {
        BCK(hipGraphLaunch(b->exec, b->stream));
        return TT_OK;
    }
    return launch_erosion(b, Tcode, Kmax);
}

int batch_step_once(tt_batch *b)
{
    const int n = (int)b->worlds.size();
    if (n == 0) return TT_OK;
    if (b->dirty) BCKS(rebuild(b));
    BCK(hipEventRecord(b->tStart, b->stream));
    BCKS(parallel(b, [&](int i) -> int {
        tt_world *w = b->worlds[i];
        w->timing = false;
        BCKS(tt_step_pre(w));
        BCKS(tt_step_thermal_prep(w));
        BCK(hipEventRecord(b->preDone[i], w->stream));
        return TT_OK;
    }));
    int Pmax = 0, anyOdd = 0, Kmax = 0;
    for (int i = 0; i < n; ++i) {
        tt_world *w = b->worlds[i];
        BCK(hipStreamWaitEvent(b->stream, b->preDone[i], 0));
        b->Ts[i] = tt_step_thermal_count(w);
        b->Ks[i] = tt_step_hydraulic_count(w);
        Pmax = std::max(Pmax, tt_thermal_pairs(b->Ts[i], false));
        anyOdd |= b->Ts[i] & 1;
        Kmax = std::max(Kmax, b->Ks[i]);
        fill_desc(b->hDesc[i], w, b->Ts[i], b->Ks[i]);
    }
    BCK(hipEventRecord(b->tPre, b->stream));
    BCK(hipMemcpyAsync(b->dDesc, b->hDesc, (size_t)n * sizeof(tt_bworld), hipMemcpyHostToDevice, b->stream));
    BCKS(run_erosion(b, 2 * Pmax + anyOdd, Kmax));
    BCK(hipEventRecord(b->erosionDone, b->stream));
    BCK(hipEventRecord(b->tEro, b->stream));
    for (int i = 0; i < n; ++i) {
        tt_world *w = b->worlds[i];
        tt_step_erode_indices(w, b->Ts[i], b->Ks[i], &b->sedLast[i]);
        BCK(hipStreamWaitEvent(w->stream, b->erosionDone, 0));
    }
    BCKS(parallel(b, [&](int i) -> int {
        tt_world *w = b->worlds[i];
        if (b->Ks[i] > 0) BCKS(tt_step_transport(w, b->sedLast[i]));
        BCKS(tt_step_post(w));
        BCK(hipEventRecord(b->postDone[i], w->stream));
        return TT_OK;
    }));
    for (int i = 0; i < n; ++i) BCK(hipStreamWaitEvent(b->stream, b->postDone[i], 0));
    BCK(hipEventRecord(b->tEnd, b->stream));
    // the next step overwrites the pinned descriptors: the copy must be done
    BCK(hipStreamSynchronize(b->stream));
    float a = 0.0f, e = 0.0f, c = 0.0f;
    BCK(hipEventElapsedTime(&a, b->tStart, b->tPre));
    BCK(hipEventElapsedTime(&e, b->tPre, b->tEro));
    BCK(hipEventElapsedTime(&c, b->tEro, b->tEnd));
    b->times.ms_pre += a;
    b->times.ms_erosion += e;
    b->times.ms_post += c;
    b->times.steps += 1;
    return TT_OK;
}

}  // namespace

extern "C" {

int tt_batch_create(tt_batch **out)
{
    if (!out) return TT_ERR_INVALID;
    *out = nullptr;
    tt_batch *b = new (std::nothrow) tt_batch();
    if (!b) return TT_ERR_NOMEM;
    if (hipGetDevice(&b->device) != hipSuccess || hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&b->erosionDone, hipEventDisableTiming) != hipSuccess ||
        hipEventCreate(&b->tStart) != hipSuccess || hipEventCreate(&b->tPre) != hipSuccess ||
        hipEventCreate(&b->tEro) != hipSuccess || hipEventCreate(&b->tEnd) != hipSuccess) {
        delete b;
        tt_set_error("no HIP device for the batch");
        return TT_ERR_DEVICE;
    }
    const unsigned hw = std::thread::hardware_concurrency();
    b->threads = (int)std::max(1u, std::min(16u, hw ? hw : 4u));
    *out = b;
    return TT_OK;
}

int tt_batch_add(tt_batch *b, tt_world *w)
{
    if (!b || !w) return TT_ERR_INVALID;
    if (!w->initialized) { tt_set_error("tt_batch_add: world not initialised"); return TT_ERR_STATE; }
    if (w->device != b->device) { tt_set_error("tt_batch_add: world on another device"); return TT_ERR_INVALID; }
    if (w->ch) { tt_set_error("tt_batch_add: compensated-precision worlds step alone"); return TT_ERR_UNSUPPORTED; }
    if (std::find(b->worlds.begin(), b->worlds.end(), w) != b->worlds.end()) return TT_OK;
    if ((int)b->worlds.size() >= TT_BATCH_MAX_WORLDS) { tt_set_error("batch is full"); return TT_ERR_UNSUPPORTED; }
    b->worlds.push_back(w);
    ++w->batches;
    b->dirty = true;
    return TT_OK;
}

int tt_batch_remove(tt_batch *b, tt_world *w)
{
    if (!b || !w) return TT_ERR_INVALID;
    auto it = std::find(b->worlds.begin(), b->worlds.end(), w);
    if (it == b->worlds.end()) return TT_ERR_INVALID;
    b->worlds.erase(it);
    --w->batches;
    b->dirty = true;
    return TT_OK;
}

int tt_batch_size(tt_batch *b, int32_t *n)
{
    if (!b || !n) return TT_ERR_INVALID;
    *n = (int32_t)b->worlds.size();
    return TT_OK;
}

int tt_batch_set_threads(tt_batch *b, int32_t threads)
{
    if (!b || threads < 1) return TT_ERR_INVALID;
    b->threads = threads;
    return TT_OK;
}

int tt_batch_set_graph(tt_batch *b, int32_t enabled)
{
    if (!b) return TT_ERR_INVALID;
    drop_graph(b);
    b->useGraph = enabled != 0;
    return TT_OK;
}

int tt_batch_step(tt_batch *b, int32_t steps)
{
    if (!b || steps < 0) return TT_ERR_INVALID;
    if (hipSetDevice(b->device) != hipSuccess) return TT_ERR_DEVICE;
    for (int s = 0; s < steps; ++s) BCKS(batch_step_once(b));
    return TT_OK;
}

int tt_batch_stage_times(tt_batch *b, tt_batch_times *out)
{
    if (!b || !out) return TT_ERR_INVALID;
    *out = b->times;
    return TT_OK;
}

int tt_batch_destroy(tt_batch *b)
{
    if (!b) return TT_ERR_INVALID;
    (void)hipSetDevice(b->device);
    (void)hipStreamSynchronize(b->stream);
    b->pool.shutdown();
    for (tt_world *w : b->worlds) --w->batches;   // members stay valid
    drop_graph(b);
    free_tables(b);
    for (auto &e : b->preDone)
        if (e) (void)hipEventDestroy(e);
    for (auto &e : b->postDone)
        if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : {b->erosionDone, b->tStart, b->tPre, b->tEro, b->tEnd})
        if (e) (void)hipEventDestroy(e);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return TT_OK;
}

}  // extern "C"
