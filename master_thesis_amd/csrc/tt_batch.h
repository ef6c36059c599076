// tt_batch.h -- device descriptor of a batch member and the batched erosion
// launchers (k_batch.hip), used by the batch host code (batch.hip).
#pragma once

#include "tt_device.h"

// One member world as the batched erosion kernels see it for one step.
struct tt_bworld {
    int W, H;
    int T, K;              // thermal passes, hydraulic sub-iterations this step
    int hc0;               // height buffer index at the first thermal pass
    int hc1;               // height buffer index at the first hydraulic sub-iteration
    int wc, fc, sc;        // water, flux, sediment buffer indices at the first sub-iteration
    int rseed;             // seed + iteration (rain noise)
    float talus, amp, str; // thermal settings
    float rain, evap, cap, ke, kd, fk, g, L;  // hydraulic settings
    float rL;                                 // RN(1 / L)
    int divL;                                 // the division by L may take tt_div_const
    int sqrtOk;                               // square roots may take tt_sqrt_normal
    int cvtOk;                                // the rain's conversion may take tt_f2u32_hw
    float sin003;                             // tt_sinf(0.03f)
    float *h[2];
    float *water[2];
    float4 *flux[2];
    float *sed[2];
    float2 *vel;
    float *wn;              // water after the flow between sub-iterations
    const uint8_t *frozen;  // per 64x16 tile: unchanged by this step's thermal passes
    uint8_t *chg;           // per 64x16 tile and launch parity: the thermal launch changed a cell
};

// tile table entries: (world index << 20) | (tile row << 8) | tile column
#define TT_BATCH_MAX_WORLDS 4096

// two: pair j of every member's thermal passes; otherwise the single first pass
// of the members with an odd pass count
hipError_t tt_launch_thermal_batch(hipStream_t s, const tt_bworld *worlds, const uint32_t *tiles, int ntiles, int j,
                                   bool two);
hipError_t tt_launch_hydraulic_batch(hipStream_t s, int rows, const tt_bworld *worlds, const uint32_t *tiles,
                                     int ntiles, int k);
