// tt_world.h -- the world object behind the C ABI (host side).
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "tt_device.h"


// entries of a world's Voronoi seed table (the last generation of
// numVoronoiSeeds, or the plate centres)
constexpr int TT_MAX_SEEDS = 65536;

struct tt_alloc {
    void **ptr;
    size_t bytes;
};

// One simulated world: device state of PlateTectonicSim
// (plate_tectonic_sim.h:62-108) in one device allocation, plus the world's
// stream and pinned read-back buffers.
struct tt_world {
    tt_sim_config cfg{};
    tt_kernel_settings ks{};
    std::vector<int32_t> gens;   // numVoronoiSeeds: seed counts of the Voronoi generations
    int W = 0, H = 0, N = 0;
    uint32_t iter = 0;
    bool initialized = false;
    int batches = 0;            // batches holding the world (it steps only through them)
    int device = 0;
    int rows = 0;            // rows of the hydraulic tiles (0 or -1 until resolved)
    bool tuning = false;     // the tile height is tuned on the next step with enough sub-iterations
    hipEvent_t tuneEv[6] = {};
    int tuneRows[3] = {}, tuneCount[3] = {}, tuneGroups = 0;
    int ehc = 0;             // height buffer index during erosion
    hipStream_t stream = nullptr;
    char *base = nullptr;
    std::vector<tt_alloc> allocs;

    // state (ping-pong pairs; the index names the current buffer)
    float *h[2] = {nullptr, nullptr};
    uint8_t *ids[2] = {nullptr, nullptr};
    float *water[2] = {nullptr, nullptr};
    float4 *flux[2] = {nullptr, nullptr};
    float *sed[2] = {nullptr, nullptr};
    float2 *vel = nullptr;
    float *wn = nullptr;        // hydraulic: water after the flow, between sub-iterations
    int hc = 0, ic = 0, wc = 0, fc = 0, sc = 0;
    tt_plate_data *pd = nullptr;
    uint32_t *divBitmap = nullptr;
    uint8_t *typeBitmap = nullptr;
    tt_uplift *uplift = nullptr;
    uint8_t *noise = nullptr;

    // per-step products
    uint32_t *word = nullptr;
    uint8_t *cnt = nullptr;
    uint32_t *segCount = nullptr, *segBase = nullptr;
    unsigned long long *keys = nullptr;
    int *typeCounts = nullptr;
    uint32_t *hasDiv = nullptr;
    uint32_t *accw = nullptr;
    uint8_t *accretion = nullptr;
    tt_prop *prop = nullptr;
    uint8_t *srcFlag = nullptr, *contested = nullptr, *serialFlag = nullptr;
    uint32_t *serialList = nullptr, *serialCount = nullptr;
    // ordered continental propagation: walk targets of the listed sources,
    // their LDS slots, the per-cell slot map (kept all-ones) and slot cells
    uint32_t *propCell = nullptr, *propSlot = nullptr, *slotOf = nullptr, *slotCell = nullptr;
    uint16_t *propOff = nullptr;
    int *jfa[2] = {nullptr, nullptr};
    tt_dist *dist = nullptr;
    tt_plate_step *ps = nullptr;
    tt_step_ctl *ctl = nullptr;
    long long *fric = nullptr;

    // CCL / relabel
    uint32_t *labels = nullptr, *rootCount = nullptr, *rootList = nullptr, *nroots = nullptr;
    uint32_t *locList = nullptr, *nloc = nullptr;  // local roots of the tiles
    uint8_t *rootFlag = nullptr, *newId = nullptr, *unFlag = nullptr, *vote = nullptr;
    unsigned long long *fragKey = nullptr;
    int *maxCount = nullptr;
    uint32_t *list[2] = {nullptr, nullptr};
    uint32_t *listN[2] = {nullptr, nullptr};

    // merge / split
    uint8_t *nbm = nullptr;
    int *mergeIds = nullptr;
    long long *acc = nullptr;
    float *eff = nullptr;
    uint32_t *rchg = nullptr;   // per-round change counters of the split fixed points (TT_RCHG)
    int lastRounds = 28;        // rounds the last split relaxation needed (first batch length)
    uint8_t *act = nullptr;     // per-tile activity flags of the split rounds (2 x tiles)
    float2 *tileMM = nullptr;   // per 64x16 tile: min / max height (thermal)
    uint8_t *tileFrozen = nullptr;  // per 64x16 tile: unchanged by this step's thermal passes
    uint8_t *tileChg = nullptr;     // per 64x16 tile and pass parity: the thermal pass changed a cell
    float *ch = nullptr;        // precision mode: height compensation
    float *cs[2] = {nullptr, nullptr};  // precision mode: sediment compensation (ping-pong like sed)

    // scratch
    uint32_t *scanTmp = nullptr;
    uint8_t *initSeeds = nullptr;
    double *initWts = nullptr;

    uint32_t *dIter = nullptr;  // the step's iteration on the device (read by collision resolution)
    float rL = 0.0f;            // RN(1 / hydrationPipeLength)
    int divL = 0;               // 1: the hydraulic division by the pipe length may take tt_div_const
    int sqrtOk = 0;             // 1: hydraulic square roots of normal arguments may take tt_sqrt_normal
    int cvtOk = 0;              // 1: the rain's saturating conversion may take tt_f2u32_hw

    // host mirrors
    tt_step_ctl *hctl = nullptr;
    uint32_t *hu32 = nullptr;
    uint32_t *hIter = nullptr;  // pinned source of dIter

    // the pre-erosion launches captured as graphs, one per (height, id)
    // buffer parity; replayed while the event timing is off
    hipGraphExec_t preExec[4] = {};
    bool preGraphs = true;
    tt_step_stats stats{};
    bool timing = false;
    hipEvent_t ev[8] = {};
    tt_stage_times times{};
};

// byte layout of one world; with alloc = true the pointer slots of w are
// registered for the single device allocation
size_t tt_world_layout(const tt_sim_config &c, bool alloc, tt_world *w);

// step phases (world.hip), used by the batched step (batch.hip)
int tt_step_pre(tt_world *w);
int tt_step_post(tt_world *w);
int tt_step_thermal_count(const tt_world *w);
int tt_step_hydraulic_count(const tt_world *w);
void tt_step_erode_indices(tt_world *w, int T, int K, int *sedLast);
int tt_step_transport(tt_world *w, int sedLast);
int tt_step_thermal_prep(tt_world *w);
int tt_world_do_init(tt_world *w);
int tt_validate_config(const tt_sim_config *c);

// host.cpp
void tt_settings_fill_defaults(tt_kernel_settings *k);
const std::vector<uint8_t> &tt_noise_table();
