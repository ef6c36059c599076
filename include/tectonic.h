/* tectonic.h -- C ABI of libtectonic, the MI355X (gfx950) implementation of the
 * per-step terrain pipeline of MarWo22/master_thesis.
 *
 * The reference has no FFI layer: its boundary for this path is the C++ class
 * PlateTectonicSim (src/cuda_compute/plate_tectonic_sim.h:56-179) plus the
 * global settings objects it reads (KernelSettings, kernel_settings.cuh:4-81;
 * SimulationSettings, generation_settings.h:67-76; Config, config.h:15-29).
 * Every entry point below names the reference interface it replaces.
 *
 * Conventions: plain C types only; every function returns an int status
 * (TT_OK == 0, negative on error) and never exits the process (the reference
 * calls exit() from CUDA_ERROR_CHECK, plate_tectonic_sim.h:40-54).  The caller
 * owns host buffers; the library owns device memory.  One world handle may be
 * used by one thread at a time; different handles may be used concurrently.
 */
#ifndef TECTONIC_H
#define TECTONIC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------- */
enum {
    TT_OK = 0,
    TT_ERR_INVALID = -1,      /* bad argument (null pointer, bad size, bad enum) */
    TT_ERR_DEVICE = -2,       /* HIP runtime error (message in tt_last_error) */
    TT_ERR_NOMEM = -3,        /* device allocation failed */
    TT_ERR_STATE = -4,        /* call out of order (e.g. step before init) */
    TT_ERR_UNSUPPORTED = -5,  /* feature or size not supported */
    TT_ERR_BUDGET = -6        /* serving admission: request exceeds the HBM budget */
};

/* ---- parameters --------------------------------------------------------- */

/* Mirrors struct KernelSettings (src/cuda_compute/kernel_settings.cuh:4-81):
 * same 50 fields, same order, same names and defaults (tt_settings_default).
 * The reference keeps these in __constant__ memory that is only written when
 * the GUI changes a slider (kernel_settings.cu:7-9, gui.cpp:434-441); this
 * library always uses the values passed here. */
typedef struct tt_kernel_settings {
    float inelasticCollisionMultiplierContinental;
    float inelasticCollisionMultiplierSubduction;
    float inelasticCollisionMultiplierAccretion;
    float frictionCoefficientContinental;
    float frictionCoefficientSubduction;
    float frictionCoefficientAccretion;
    float environmentalDragCoefficient;
    float velocitySmoothingFactor;
    float directionSmoothingFactor;
    float breakScoreSmoothingFactor;
    float mergeDotDirectionThreshold;
    float mergeVelocityDiffThreshold;
    float mergeMinVelocity;
    int32_t mergeWaitAfterSplit;
    int32_t minPlateSize;
    float continentalCrustThreshold;
    float divergence_height_target_min;
    float divergence_height_target_max;
    float divergence_interpolation_factor;
    int32_t collisionTypeUpdateCooldown;
    int32_t minCollisionSize;
    float subductionOverUpliftMultiplier;
    float subductionUnderUpliftMultiplier;
    float continentalUpliftMultiplier;
    float minUpliftClamp;
    float maxUpliftClamp;
    int32_t upliftRange;
    int32_t subductionOverRange;
    int32_t continentalPropMaxIters;
    float continentalPropThreshold;
    float hydrationPipeCrossSection;
    float hydrationPipeLength;
    float hydrationRainfall;
    float hydrationEvaporation;
    float gravity;
    float sedimentCapacity;
    float sedimentErosionRate;
    float sedimentDepositionRate;
    float pressureAccumulation;
    float pressureDecayRate;
    int32_t pressureBlurRange;
    float pressureMultiplier;
    float curlNoiseMultiplier;
    float stressSplitThreshold;
    int32_t targetMinimumPlateArea;
    int32_t targetMaximumPlateArea;
    float thermalErosionStrength;
    float thermalErosionAmplitude;
    float thermalCellSize;
    float thermalThresholdAngle;
} tt_kernel_settings;

/* How the initial state is produced (untimed). */
enum {
    /* synthetic benchmark world: heights U[0,1) smoothed by a toroidal
     * Gaussian (sigma 8, radius 32, fp64), plates = nearest of num_plates
     * uniform sites, speeds U[speed_min, speed_max]. */
    TT_INIT_SYNTHETIC = 0,
    /* the reference's initializeTectonics (plate_tectonic_sim.cu:705-773):
     * simplex fBm heights in [5000, 15000] (initHeightmap,
     * plate_tectonics_kernel.cu:68-99), nested Voronoi seed generations, speed
     * (U(-1,1)+1)/2; the MSVC RNG stream is replaced by SplitMix64. */
    TT_INIT_REFERENCE = 1,
    /* all fields zero; the caller uploads a state with tt_world_upload. */
    TT_INIT_EMPTY = 2
};

/* Stage mask bits (tests may switch stages off to localise a difference;
 * the default TT_STAGE_ALL is the reference step). */
enum {
    TT_STAGE_COLLIDE = 1 << 0,   /* advection, collision typing and resolution */
    TT_STAGE_UPLIFT = 1 << 1,    /* continental propagation, JFA, applyUplift */
    TT_STAGE_THERMAL = 1 << 2,
    TT_STAGE_HYDRAULIC = 1 << 3,
    TT_STAGE_KINEMATICS = 1 << 4, /* applyPlateMovementChanges + applyAccretion */
    TT_STAGE_CCL = 1 << 5,
    TT_STAGE_MERGE = 1 << 6,     /* mergePlates + break score + centres + stats */
    TT_STAGE_SPLIT = 1 << 7,
    TT_STAGE_ALL = 0xff
};

/* World geometry and schedule.  Replaces Config (config.h:15-29:
 * heightmap_dimensions) and SimulationSettings (generation_settings.h:67-76:
 * seed, numStartingPlates, numVoronoiSeeds, iterationsForCCL), plus explicit
 * fields for counts the reference hard-codes (10 hydraulic sub-iterations at
 * plate_tectonic_sim.cu:336, one thermal pass at :73). */
typedef struct tt_sim_config {
    int32_t width;                 /* multiple of 64, >= 64 */
    int32_t height;                /* multiple of 16, >= 64 */
    uint32_t seed;
    int32_t num_plates;            /* 1..254 */
    int32_t num_voronoi_gens;      /* TT_INIT_REFERENCE only, 0..4 (longer lists: tt_world_reset) */
    int32_t voronoi_gens[4];
    int32_t iterations_for_ccl;    /* CCL when step % iterations_for_ccl == 0 */
    int32_t hydraulic_iterations;  /* default 10 */
    int32_t thermal_iterations;    /* default 1 */
    int32_t init_mode;             /* TT_INIT_* */
    float speed_min;               /* TT_INIT_SYNTHETIC plate speeds */
    float speed_max;
    int32_t stage_mask;            /* TT_STAGE_* */
    int32_t tile_rows;             /* rows of the hydraulic tiles: 8, 16 or 32 (dividing height);
                                      0 = chosen by the tile autotuner (cached per size class; without
                                      a cached choice the first step times its sub-iterations at each
                                      height and keeps the fastest); -1 = tune even when cached.
                                      Thermal tiles have 16 rows (their skip flags are per 64x16 tile). */
    int32_t uplift_mode;           /* TT_UPLIFT_JFA (the reference's live path) or TT_UPLIFT_SEPARABLE */
    int32_t precision;             /* TT_PRECISION_FP32 (reference arithmetic, bit-exact with the oracle)
                                      or TT_PRECISION_COMPENSATED (Kahan height / sediment updates) */
    int32_t reserved[4];
} tt_sim_config;

/* Uplift spread.  JFA: jump flood + applyUplift (plate_tectonic_sim.cu:283-317,
 * the reference's live path).  SEPARABLE: the bounded-radius VerticalBlur +
 * HorizontalBlur pair (plate_tectonics_kernel.cu:1603-1643, 1740-1843; its
 * call is commented out at plate_tectonic_sim.cu:319-327), radius
 * KernelSettings.upliftRange. */
enum { TT_UPLIFT_JFA = 0, TT_UPLIFT_SEPARABLE = 1 };

/* Arithmetic of the erosion stages.  FP32: the reference's float arithmetic
 * (results bit-identical to the CPU oracle).  COMPENSATED: the height and
 * sediment updates of thermal and hydraulic erosion carry per-cell Kahan
 * compensation terms (the sediment term travels with the sediment transport);
 * the height term restarts whenever collisions move heights between cells.
 * Not bit-comparable with the oracle; see tt_precision_study. */
enum { TT_PRECISION_FP32 = 0, TT_PRECISION_COMPENSATED = 1 };

/* struct PlateData (src/cuda_compute/types/plate_data.h:5-38), 84 bytes, same
 * field order and layout; the plate table has 255 entries (MAX_PLATE_COUNT,
 * plate_tectonics_kernel.cuh:13). */
typedef struct tt_plate_data {
    float velocity;
    float velocitySmoothed;
    float direction[2];
    float directionSmoothed[2];
    float pixelCenter[2];
    float velocityChange[2];
    float geometricCenter[2];
    float asthenosphereVelocity[2];
    int32_t size;
    float mass;
    int32_t perimeter;
    float breakScore;
    float circularity;
    uint8_t used;
    uint8_t hasMoved;
    uint8_t pad_[2];
    int32_t mergeWaitTime;
} tt_plate_data;

#define TT_MAX_PLATES 255

/* Fields that can be moved between host and device.  Row-major, index y*W+x. */
enum {
    TT_FIELD_HEIGHT = 0,        /* float[N]        m_heightMapTexture */
    TT_FIELD_PLATE_IDS = 1,     /* uint8[N]        m_plateIdsTexture */
    TT_FIELD_WATER = 2,         /* float[N]        m_hydrationLevel */
    TT_FIELD_FLUX = 3,          /* float4[N]       m_hydrationFlux (x=-x, y=+y, z=+x, w=-y) */
    TT_FIELD_VELOCITY = 4,      /* float2[N]       m_hydrationVelocity */
    TT_FIELD_SEDIMENT = 5,      /* float[N]        m_sedimentLevel */
    TT_FIELD_ASTHENOSPHERE = 6, /* float2[N]       m_asthenosphere (read-only) */
    TT_FIELD_COLLISIONS = 7,    /* uint32[N]       m_plateCollisions of the last step */
    TT_FIELD_ACCRETION = 8,     /* uint8[N]        m_accretionTexture of the last step */
    TT_FIELD_UPLIFT = 9,        /* {float, uint8, pad[3]}[N]   m_upliftData */
    TT_FIELD_PROPAGATED = 10,   /* {float, uint32, uint8, pad[3]}[N] m_propagatedUpliftData */
    TT_FIELD_PLATE_DATA = 11,   /* tt_plate_data[255] m_plateDataLookup */
    TT_FIELD_DIVERGENCE_BITMAP = 12, /* uint32[1013]  m_divergenceBitmap */
    TT_FIELD_COLLISION_TYPES = 13,   /* uint8[32385]  m_collisionTypeBitmap */
    TT_FIELD_SORTED_KEYS = 14,  /* uint64[N] sorted (dest << 8 | plate) keys of the last step */
    TT_FIELD_LABELS = 15,       /* uint32[N] CCL labels (min cell index per component) of the last CCL */
    TT_FIELD_DISTANCE = 16,     /* {int32 dist, uint32 origIndex}[N] JFA distance field of the last step */
    TT_FIELD_COUNT_
};

/* Per-step statistics (replaces IterationStatistics, iteration_statistics.h,
 * and the per-step stdout prints of the reference). */
typedef struct tt_step_stats {
    uint32_t iteration;          /* m_iterations after the step */
    int32_t largest_value;       /* IterationStatistics */
    int32_t largest_plate_id;
    float heaviest_value;
    int32_t heaviest_plate_id;
    int32_t number_of_plates;    /* plates with size > 0 */
    int32_t overflow_cells;      /* cells reached by > 4 plates (4 smallest ids kept) */
    int32_t num_labels;          /* CCL components in the last CCL */
    int32_t unassigned_cells;    /* cells filled by neighbour vote in the last CCL */
    int32_t fill_rounds;
    int32_t merges;              /* merge pairs accepted in the last step */
    int32_t split_plate;         /* plate split in the last step, -1 if none */
    int32_t split_new_plate;
    int32_t split_iterations;    /* relaxation sweeps until convergence */
    int32_t complex_sources;     /* continental sources handled in canonical serial order */
    int32_t tile_rows;           /* rows of the hydraulic tiles in use (tt_sim_config.tile_rows or tuned) */
    int32_t hydraulic_div_const; /* 1: the flux division by the pipe length uses the checked reciprocal form */
    int32_t hydraulic_sqrt_fast; /* 1: square roots of normal arguments use the checked hardware-root form */
    double mass_ledger;          /* sum of height over the grid, exact fixed-point, fp64 */
} tt_step_stats;

typedef struct tt_world tt_world;

/* ---- library -------------------------------------------------------------- */
const char *tt_status_string(int status);
/* Text of the last error on the calling thread. */
const char *tt_last_error(void);
/* Device count visible to the library (0 on a host without GPU). */
int tt_device_count(int *count);

/* KernelSettings defaults, kernel_settings.cuh:6-71. */
int tt_settings_default(tt_kernel_settings *out);
/* SimulationSettings / Config defaults (generation_settings.h:67-76,
 * config.json:1-6): 1024x1024, seed 1000, 16 plates, gens {200, 500}, CCL every
 * 10 steps, 10 hydraulic and 1 thermal iteration, reference init. */
int tt_config_default(tt_sim_config *out);
/* Parse config.json / generation-settings JSON text into the two structs.
 * Recognised keys: heightmap_dimensions.{x,y} (config.json:3-5), seed,
 * numStartingPlates, numVoronoiSeeds, iterationsForCCL, hydraulicIterations,
 * thermalIterations, and any KernelSettings field by name. */
int tt_config_parse_json(const char *text, tt_sim_config *cfg, tt_kernel_settings *ks);

/* ---- world (replaces PlateTectonicSim) ---------------------------------- */
/* PlateTectonicSim(int width, int height, unsigned seed, CudaGlInteropManager*)
 * plate_tectonic_sim.h:111; allocates device state for one world on the
 * current device. */
int tt_world_create(const tt_sim_config *cfg, const tt_kernel_settings *ks, tt_world **out);
/* initialize(numStartingPlates, numVoronoiSeeds), plate_tectonic_sim.h:115 */
int tt_world_init(tt_world *w);
/* resetSim(seed, numStartingPlates, numVoronoiSeeds), plate_tectonic_sim.h:119:
 * re-initialises the world with a new seed, plate count and list of Voronoi
 * generations (num_gens entries, any number; each at least 1 seed, the last at
 * most 65536).  With the world's own seed this is also
 * initialize(numStartingPlates, numVoronoiSeeds), plate_tectonic_sim.h:115, for
 * lists longer than the 4 slots of tt_sim_config. */
int tt_world_reset(tt_world *w, uint32_t seed, int32_t num_plates, const int32_t *voronoi_gens, int32_t num_gens);
/* executeIteration() x n, plate_tectonic_sim.h:117; returns after the device
 * work is complete (the reference synchronises at sim.cu:117-119).
 * TT_ERR_STATE for a member of a batch (it steps through the batch). */
int tt_world_step(tt_world *w, int32_t n);
/* The world's configuration and settings (e.g. of a world made by
 * tt_world_load or a server job); either pointer may be NULL. */
int tt_world_config(tt_world *w, tt_sim_config *cfg, tt_kernel_settings *ks);
/* Upload of new KernelSettings (copyHostKernelSettingsToDevice, kernel_settings.cu:7-9). */
int tt_world_set_settings(tt_world *w, const tt_kernel_settings *ks);
int tt_world_download(tt_world *w, int32_t field, void *host, size_t bytes);
int tt_world_upload(tt_world *w, int32_t field, const void *host, size_t bytes);
int tt_world_stats(tt_world *w, tt_step_stats *out);
int tt_world_iteration(tt_world *w, uint32_t *iteration);
/* Sets m_iterations (used when a state is moved between worlds). */
int tt_world_set_iteration(tt_world *w, uint32_t iteration);
/* Device bytes held by the world (persistent + transient), for admission. */
int tt_world_bytes(const tt_sim_config *cfg, size_t *bytes);
/* TT_ERR_STATE (and nothing freed) while the world is a member of a batch. */
int tt_world_destroy(tt_world *w);
/* Headless export (replaces saveFloatCudaTextureToDiskAsGray16 and
 * saveGrayscale8BitCudaTextureToDiskAsRgb, texture_save.cuh:12-18): a 16-bit
 * grey PNG of height normalised to [minUpliftClamp, maxUpliftClamp], or the
 * RGB plate palette.  Untimed. */
int tt_world_export_png(tt_world *w, int32_t what, const char *path);
/* HEIGHT16 and PLATES_RGB as above; COLLISION_CLASSES and PROPAGATED_CLASSES:
 * the createVisualizationCollisionTexture / createVisualizationPropagated-
 * CollisionTexture views (plate_tectonics_kernel.cu:3250-3313), classes 0..5
 * (none, divergence, subduction, continental with 2 / 3 / 4 plates) in RGB. */
enum { TT_EXPORT_HEIGHT16 = 0, TT_EXPORT_PLATES_RGB = 1, TT_EXPORT_COLLISION_CLASSES = 2,
       TT_EXPORT_PROPAGATED_CLASSES = 3 };
/* Diagnostics as JSON text: step statistics, field summaries, counts of the
 * collision classes, and one record per live plate with the fields of
 * copyPlateDataGuiKernel (plate_tectonics_kernel.cu:1220-1281) including its
 * continental partners and signed subduction partners.  With buf == NULL only
 * *len (the text length) is set; otherwise cap must exceed *len. */
int tt_world_diagnostics_json(tt_world *w, char *buf, size_t cap, size_t *len);
/* Snapshot: every state field, the configuration, the settings and the
 * iteration in one file (magic "TTSNAP2", a CRC32 per field over its whole
 * length; worlds in TT_PRECISION_COMPENSATED also keep their compensation
 * terms).  tt_world_load checks every CRC before it creates the world and
 * also reads version-1 files.  A world loaded from a snapshot continues
 * bit-identically. */
int tt_world_save(tt_world *w, const char *path);
int tt_world_load(const char *path, tt_world **out);

/* ---- host-side generators (the static members of PlateTectonicSim,
 * plate_tectonic_sim.h:121-130) ------------------------------------------------
 * The reference draws from std::default_random_engine, whose stream is
 * implementation-defined (MSVC); these draw from SplitMix64 instead.
 * *rng_state is the generator state, advanced by every draw; a world's
 * initialisation starts from *rng_state = seed and calls
 * tt_generate_plate_centers, then tt_generate_voronoi_seeds per generation,
 * then tt_generate_plate_data, in that order.  Each U[0, b) draw is
 * (next() >> 40) * 2^-24 * b in float.  No device work. */

/* struct VoronoiSeed (types/voronoi_seed.h): position (x, y) and plate id. */
typedef struct tt_voronoi_seed {
    float position[2];
    uint8_t id;
    uint8_t pad_[3];
} tt_voronoi_seed;

/* generatePlateCenters(generator, numPlates, width, height),
 * plate_tectonic_sim.cu:873-890: out[i] = ((U[0,height), U[0,width)), id i),
 * x drawn over the height range as in the reference. */
int tt_generate_plate_centers(uint64_t *rng_state, int32_t num_plates, int32_t width, int32_t height,
                              tt_voronoi_seed *out);
/* generateVoronoiSeeds(generator, centerSeeds, numSeeds, width, height),
 * plate_tectonic_sim.cu:892-928: num_seeds new seeds, each with the id of its
 * nearest (toroidal, first on ties) seed among centers. */
int tt_generate_voronoi_seeds(uint64_t *rng_state, const tt_voronoi_seed *centers, int32_t num_centers,
                              int32_t num_seeds, int32_t width, int32_t height, tt_voronoi_seed *out);
/* generatePlateData(generator, plateCenters, numStartingPlates),
 * plate_tectonic_sim.cu:816-842: out[0..254] = PlateData() defaults, with
 * pixelCenter = frac(center), velocity (U(-1,1)+1)/2 and a normalised random
 * direction for the first num_plates entries. */
int tt_generate_plate_data(uint64_t *rng_state, const tt_voronoi_seed *centers, int32_t num_plates,
                           tt_plate_data *out);

/* ---- batches: many worlds stepped together (no reference counterpart; the
 * batched form of executeIteration for many independent worlds) ------------
 * A step of a batch advances every member by one executeIteration.  The
 * members' erosion passes run as one launch per pass over all members (replayed
 * from a captured hipGraph); the other stages run per member on the member's
 * stream from a pool of host threads.  Each member's result is bit-identical to
 * stepping it alone.  Members must be initialised worlds of the batch's
 * device; while in a batch a member cannot be stepped on its own or destroyed
 * (TT_ERR_STATE).  Removing a member, or destroying the batch, leaves it
 * valid. */
typedef struct tt_batch tt_batch;
int tt_batch_create(tt_batch **out);
int tt_batch_add(tt_batch *b, tt_world *w);
int tt_batch_remove(tt_batch *b, tt_world *w);
int tt_batch_size(tt_batch *b, int32_t *n);
int tt_batch_step(tt_batch *b, int32_t steps);
/* host threads driving the per-member stages (default: min(16, cores)) */
int tt_batch_set_threads(tt_batch *b, int32_t threads);
/* 1 (default): replay the erosion passes from a captured hipGraph; 0: launch them directly */
int tt_batch_set_graph(tt_batch *b, int32_t enabled);
/* Device time of the batch's phases summed over its steps (HIP events): from
 * the start of a step until every member's pre-erosion phase is done, the
 * batched erosion, and from the end of the erosion until every member's
 * post-erosion phase is done. */
typedef struct tt_batch_times {
    double ms_pre, ms_erosion, ms_post;
    int64_t steps;
    double reserved[4];
} tt_batch_times;
int tt_batch_stage_times(tt_batch *b, tt_batch_times *out);
int tt_batch_destroy(tt_batch *b);

/* ---- serving: world-generation jobs under an HBM budget (no reference
 * counterpart; a request loop around the step for many independent worlds) --
 * Jobs are admitted in submission order while the summed tt_world_bytes() of
 * the resident worlds stays within the budget and at most max_active worlds
 * are resident.  Resident worlds are stepped together (a tt_batch) one step at
 * a time; a world leaves at the step boundary where its steps are done, and
 * waiting jobs are admitted into the freed space (continuous batching).  Times
 * are seconds on the server's monotonic clock. */
typedef struct tt_server tt_server;
typedef struct tt_job_result {
    int64_t id;
    int32_t status;        /* 1 while queued or running, TT_OK when done, < 0 on failure */
    uint32_t iteration;    /* iterations of the finished world */
    double mass_ledger;    /* the world's final mass ledger (tt_step_stats) */
    double t_submit, t_admit, t_done;
    int64_t bytes;         /* device bytes of the world (admission size) */
    int32_t width, height, steps;
    int32_t reserved[3];
} tt_job_result;
typedef struct tt_server_report {
    int64_t submitted, admitted, completed, failed, rejected;
    int64_t active, pending;
    int64_t bytes_in_use, bytes_peak, budget;
    int64_t batch_steps;
    double wall_s;         /* first submission to last completion */
    double worlds_per_s;   /* completed / wall_s */
    double latency_p50_s, latency_p99_s, latency_mean_s, latency_max_s;  /* completion - submission */
    double cell_updates;   /* sum over completed jobs of cells x steps */
    int64_t reserved[4];
} tt_server_report;
int tt_server_create(uint64_t hbm_budget_bytes, int32_t max_active, tt_server **out);
/* 1: keep finished worlds for tt_server_take (default 0: released at once) */
int tt_server_set_keep(tt_server *s, int32_t keep);
/* Queue a job; TT_ERR_BUDGET if the world alone exceeds the budget. */
int tt_server_submit(tt_server *s, const tt_sim_config *cfg, const tt_kernel_settings *ks, int32_t steps,
                     int64_t *id);
/* Result of a job; waits up to wait_ms (-1: until done).  TT_ERR_STATE if not done yet. */
int tt_server_result(tt_server *s, int64_t id, int32_t wait_ms, tt_job_result *out);
/* Ownership of a finished world (keep mode); the caller destroys it. */
int tt_server_take(tt_server *s, int64_t id, tt_world **w);
/* Wait until every submitted job has finished (timeout_ms -1: no limit). */
int tt_server_drain(tt_server *s, int32_t timeout_ms);
int tt_server_stats(tt_server *s, tt_server_report *out);
/* Stops the scheduler; queued and running jobs end with TT_ERR_STATE. */
int tt_server_destroy(tt_server *s);

/* ---- precision study (no reference counterpart) ----------------------------
 * Runs `steps` erosion-only steps (thermal + hydraulic stages, so every
 * variant takes the same discrete decisions) of the configured world three
 * times from one initial state: TT_PRECISION_FP32, TT_PRECISION_COMPENSATED and
 * an fp64 twin of the same formulas, and reports each float variant's error
 * against the fp64 twin and the drift of sum(height + sediment). */
typedef struct tt_precision_field {
    double max_abs_fp32, rms_fp32;   /* TT_PRECISION_FP32 - fp64 twin */
    double max_abs_comp, rms_comp;   /* TT_PRECISION_COMPENSATED - fp64 twin */
    double rms_ref;                  /* RMS of the fp64 field (scale) */
} tt_precision_field;
typedef struct tt_precision_report {
    int32_t steps, width, height_cells, reserved0;
    tt_precision_field height, water, sediment;
    double mass_start;               /* sum of height + sediment before the first step */
    double mass_ref, mass_fp32, mass_comp;
    double drift_ref, drift_fp32, drift_comp;   /* (end - start) / start */
    double ms_fp32, ms_comp, ms_ref;            /* device time per step */
    double reserved[4];
} tt_precision_report;
int tt_precision_study(const tt_sim_config *cfg, const tt_kernel_settings *ks, int32_t steps,
                       tt_precision_report *out);

/* ---- timing (hipEvent around the stages of the last tt_world_step) -------- */
typedef struct tt_stage_times {
    float ms_total;
    float ms_collide;
    float ms_uplift;
    float ms_thermal;
    float ms_hydraulic;    /* the fused hydraulic sub-iteration kernels only */
    float ms_transport;    /* the closing sediment transport of each step */
    float ms_post;         /* kinematics, accretion, CCL, merge, stats */
    float ms_split;
    float hydraulic_launches;
    float thermal_launches;
    float steps;
    float reserved[5];
} tt_stage_times;
int tt_world_set_timing(tt_world *w, int32_t enabled);
int tt_world_stage_times(tt_world *w, tt_stage_times *out);

/* ---- device self-test of the deterministic math (no reference counterpart:
 * used by the parity tests to show device and host math agree bit for bit) -- */
enum { TT_MATH_EXPF = 0, TT_MATH_LOG10F = 1, TT_MATH_SINF = 2, TT_MATH_COSF = 3,
       TT_MATH_SIMPLEX = 4, TT_MATH_DISCRETE = 5, TT_MATH_FMODF = 6, TT_MATH_ATAN2 = 7,
       TT_MATH_DIV_RCP = 8 /* pairs (a, b): a / b from RN(1 / b) as the thermal passes divide */ };
int tt_math_eval_device(int32_t fn, const float *in, float *out, int64_t n);
/* Allocation-free variant of tt_math_eval_device (no counterpart in the
 * reference): for callers that cannot allocate device memory (code under HIP
 * graph capture, serving loops, harnesses that forbid hipMalloc).  Stages the
 * host arrays through a static device buffer of 262144 elements per GPU
 * (4 MiB, held until the process ends); n above that is TT_ERR_INVALID.  Same
 * results as tt_math_eval_device.  Calls that overlap on one GPU race. */
int tt_math_eval_device_static(int32_t fn, const float *in, float *out, int64_t n);

/* ---- HBM ceiling probe (no reference counterpart: gives the roofline a
 * measured ceiling beside the published peak).  Times `reps` launches each of
 * a float4 copy (bytes read + bytes written), a float4 read-only sum, a
 * float4 fill, and a ten-stream copy with the fused hydraulic kernel's bytes
 * per cell (reads and writes of 4, 4, 16, 4 and 4 bytes) over two device
 * buffers of `bytes` each (a multiple of 16, at most 4 GiB), and reports the
 * best GB/s of each.  Allocates 2 x bytes of device memory for the duration
 * of the call. */
typedef struct tt_hbm_probe {
    double copy_gbps;
    double read_gbps;
    double write_gbps;
    double streams_gbps;   /* the ten-stream copy (64 bytes per cell) */
    int64_t bytes;
    int32_t reps;
    int32_t reserved;
} tt_hbm_probe;
int tt_probe_hbm(int64_t bytes, int32_t reps, tt_hbm_probe *out);

#ifdef __cplusplus
}
#endif

#endif /* TECTONIC_H */
