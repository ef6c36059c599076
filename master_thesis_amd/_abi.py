"""ctypes mirror of include/tectonic.h (the C ABI of libtectonic).

Plain data only: structure layouts, enum values and the loader.  The layouts
must match tectonic.h byte for byte; tests check the sizes against the C
library.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

PKG_DIR = Path(__file__).resolve().parent
LIB_PATH = PKG_DIR / "libtectonic.so"
NOISE_PATH = PKG_DIR / "data" / "noise_r8.bin"

TT_OK = 0
TT_ERR_INVALID, TT_ERR_DEVICE, TT_ERR_NOMEM, TT_ERR_STATE, TT_ERR_UNSUPPORTED, TT_ERR_BUDGET = -1, -2, -3, -4, -5, -6
STATUS_NAMES = {0: "ok", -1: "invalid argument", -2: "device error", -3: "out of device memory",
                -4: "bad state", -5: "unsupported", -6: "over HBM budget"}

TT_INIT_SYNTHETIC, TT_INIT_REFERENCE, TT_INIT_EMPTY = 0, 1, 2
TT_UPLIFT_JFA, TT_UPLIFT_SEPARABLE = 0, 1
TT_PRECISION_FP32, TT_PRECISION_COMPENSATED = 0, 1
TT_EXPORT_HEIGHT16, TT_EXPORT_PLATES_RGB, TT_EXPORT_COLLISION_CLASSES, TT_EXPORT_PROPAGATED_CLASSES = 0, 1, 2, 3

TT_STAGE_COLLIDE = 1 << 0
TT_STAGE_UPLIFT = 1 << 1
TT_STAGE_THERMAL = 1 << 2
TT_STAGE_HYDRAULIC = 1 << 3
TT_STAGE_KINEMATICS = 1 << 4
TT_STAGE_CCL = 1 << 5
TT_STAGE_MERGE = 1 << 6
TT_STAGE_SPLIT = 1 << 7
TT_STAGE_ALL = 0xFF

MAX_PLATES = 255
NUM_TRI = MAX_PLATES * (MAX_PLATES - 1) // 2
NUM_TRI_WORDS = (NUM_TRI + 31) // 32

# field id -> (numpy dtype string, values per cell or fixed count)
FIELD_HEIGHT = 0
FIELD_PLATE_IDS = 1
FIELD_WATER = 2
FIELD_FLUX = 3
FIELD_VELOCITY = 4
FIELD_SEDIMENT = 5
FIELD_ASTHENOSPHERE = 6
FIELD_COLLISIONS = 7
FIELD_ACCRETION = 8
FIELD_UPLIFT = 9
FIELD_PROPAGATED = 10
FIELD_PLATE_DATA = 11
FIELD_DIVERGENCE_BITMAP = 12
FIELD_COLLISION_TYPES = 13
FIELD_SORTED_KEYS = 14
FIELD_LABELS = 15
FIELD_DISTANCE = 16

TT_MATH_EXPF, TT_MATH_LOG10F, TT_MATH_SINF, TT_MATH_COSF = 0, 1, 2, 3
TT_MATH_SIMPLEX, TT_MATH_DISCRETE, TT_MATH_FMODF, TT_MATH_ATAN2 = 4, 5, 6, 7
TT_MATH_DIV_RCP = 8

KS_FIELDS = [
    ("inelasticCollisionMultiplierContinental", "f", 1.0),
    ("inelasticCollisionMultiplierSubduction", "f", 0.25),
    ("inelasticCollisionMultiplierAccretion", "f", 0.5),
    ("frictionCoefficientContinental", "f", 1.0),
    ("frictionCoefficientSubduction", "f", 0.2),
    ("frictionCoefficientAccretion", "f", 0.5),
    ("environmentalDragCoefficient", "f", 0.25),
    ("velocitySmoothingFactor", "f", 0.9),
    ("directionSmoothingFactor", "f", 0.9),
    ("breakScoreSmoothingFactor", "f", 0.9),
    ("mergeDotDirectionThreshold", "f", 0.998),
    ("mergeVelocityDiffThreshold", "f", 0.01),
    ("mergeMinVelocity", "f", 0.05),
    ("mergeWaitAfterSplit", "i", 10),
    ("minPlateSize", "i", 10),
    ("continentalCrustThreshold", "f", 10000.0),
    ("divergence_height_target_min", "f", 1000.0),
    ("divergence_height_target_max", "f", 2000.0),
    ("divergence_interpolation_factor", "f", 0.025),
    ("collisionTypeUpdateCooldown", "i", 4),
    ("minCollisionSize", "i", 50),
    ("subductionOverUpliftMultiplier", "f", 0.005),
    ("subductionUnderUpliftMultiplier", "f", 0.01),
    ("continentalUpliftMultiplier", "f", 0.01),
    ("minUpliftClamp", "f", 0.0),
    ("maxUpliftClamp", "f", 20000.0),
    ("upliftRange", "i", 50),
    ("subductionOverRange", "i", 20),
    ("continentalPropMaxIters", "i", 50),
    ("continentalPropThreshold", "f", 13000.0),
    ("hydrationPipeCrossSection", "f", 1000.0),
    ("hydrationPipeLength", "f", 1000.0),
    ("hydrationRainfall", "f", 0.101),
    ("hydrationEvaporation", "f", 0.303),
    ("gravity", "f", 9.81),
    ("sedimentCapacity", "f", 0.9),
    ("sedimentErosionRate", "f", 0.5),
    ("sedimentDepositionRate", "f", 0.01),
    ("pressureAccumulation", "f", 0.02),
    ("pressureDecayRate", "f", 0.8),
    ("pressureBlurRange", "i", 25),
    ("pressureMultiplier", "f", 1.0),
    ("curlNoiseMultiplier", "f", 30.0),
    ("stressSplitThreshold", "f", 30.0),
    ("targetMinimumPlateArea", "i", 100),
    ("targetMaximumPlateArea", "i", 100000),
    ("thermalErosionStrength", "f", 1.0),
    ("thermalErosionAmplitude", "f", 0.5),
    ("thermalCellSize", "f", 1000.0),
    ("thermalThresholdAngle", "f", 0.5),
]


class KernelSettings(C.Structure):
    """struct KernelSettings (kernel_settings.cuh:4-81)."""
    _fields_ = [(n, C.c_float if t == "f" else C.c_int32) for n, t, _ in KS_FIELDS]

    @classmethod
    def defaults(cls) -> "KernelSettings":
        ks = cls()
        for n, _, v in KS_FIELDS:
            setattr(ks, n, v)
        return ks

    def copy(self) -> "KernelSettings":
        out = KernelSettings()
        C.memmove(C.byref(out), C.byref(self), C.sizeof(self))
        return out


class SimConfig(C.Structure):
    _fields_ = [
        ("width", C.c_int32),
        ("height", C.c_int32),
        ("seed", C.c_uint32),
        ("num_plates", C.c_int32),
        ("num_voronoi_gens", C.c_int32),
        ("voronoi_gens", C.c_int32 * 4),
        ("iterations_for_ccl", C.c_int32),
        ("hydraulic_iterations", C.c_int32),
        ("thermal_iterations", C.c_int32),
        ("init_mode", C.c_int32),
        ("speed_min", C.c_float),
        ("speed_max", C.c_float),
        ("stage_mask", C.c_int32),
        ("tile_rows", C.c_int32),
        ("uplift_mode", C.c_int32),
        ("precision", C.c_int32),
        ("reserved", C.c_int32 * 4),
    ]

    @classmethod
    def defaults(cls) -> "SimConfig":
        """SimulationSettings + Config defaults (generation_settings.h:67-76, config.json)."""
        c = cls()
        c.width, c.height, c.seed, c.num_plates = 1024, 1024, 1000, 16
        c.num_voronoi_gens = 2
        c.voronoi_gens[0], c.voronoi_gens[1] = 200, 500
        c.iterations_for_ccl = 10
        c.hydraulic_iterations, c.thermal_iterations = 10, 1
        c.init_mode = TT_INIT_REFERENCE
        c.speed_min, c.speed_max = 0.0, 1.0
        c.stage_mask = TT_STAGE_ALL
        return c

    def copy(self) -> "SimConfig":
        out = SimConfig()
        C.memmove(C.byref(out), C.byref(self), C.sizeof(self))
        return out


class PlateData(C.Structure):
    """struct PlateData (plate_data.h:5-38), 84 bytes."""
    _fields_ = [
        ("velocity", C.c_float),
        ("velocitySmoothed", C.c_float),
        ("direction", C.c_float * 2),
        ("directionSmoothed", C.c_float * 2),
        ("pixelCenter", C.c_float * 2),
        ("velocityChange", C.c_float * 2),
        ("geometricCenter", C.c_float * 2),
        ("asthenosphereVelocity", C.c_float * 2),
        ("size", C.c_int32),
        ("mass", C.c_float),
        ("perimeter", C.c_int32),
        ("breakScore", C.c_float),
        ("circularity", C.c_float),
        ("used", C.c_uint8),
        ("hasMoved", C.c_uint8),
        ("pad_", C.c_uint8 * 2),
        ("mergeWaitTime", C.c_int32),
    ]


class StepStats(C.Structure):
    _fields_ = [
        ("iteration", C.c_uint32),
        ("largest_value", C.c_int32),
        ("largest_plate_id", C.c_int32),
        ("heaviest_value", C.c_float),
        ("heaviest_plate_id", C.c_int32),
        ("number_of_plates", C.c_int32),
        ("overflow_cells", C.c_int32),
        ("num_labels", C.c_int32),
        ("unassigned_cells", C.c_int32),
        ("fill_rounds", C.c_int32),
        ("merges", C.c_int32),
        ("split_plate", C.c_int32),
        ("split_new_plate", C.c_int32),
        ("split_iterations", C.c_int32),
        ("complex_sources", C.c_int32),
        ("tile_rows", C.c_int32),
        ("hydraulic_div_const", C.c_int32),
        ("hydraulic_sqrt_fast", C.c_int32),
        ("mass_ledger", C.c_double),
    ]


class StageTimes(C.Structure):
    _fields_ = [
        ("ms_total", C.c_float),
        ("ms_collide", C.c_float),
        ("ms_uplift", C.c_float),
        ("ms_thermal", C.c_float),
        ("ms_hydraulic", C.c_float),
        ("ms_transport", C.c_float),
        ("ms_post", C.c_float),
        ("ms_split", C.c_float),
        ("hydraulic_launches", C.c_float),
        ("thermal_launches", C.c_float),
        ("steps", C.c_float),
        ("reserved", C.c_float * 5),
    ]


class JobResult(C.Structure):
    _fields_ = [
        ("id", C.c_int64),
        ("status", C.c_int32),
        ("iteration", C.c_uint32),
        ("mass_ledger", C.c_double),
        ("t_submit", C.c_double),
        ("t_admit", C.c_double),
        ("t_done", C.c_double),
        ("bytes", C.c_int64),
        ("width", C.c_int32),
        ("height", C.c_int32),
        ("steps", C.c_int32),
        ("reserved", C.c_int32 * 3),
    ]


class ServerReport(C.Structure):
    _fields_ = [(n, C.c_int64) for n in ("submitted", "admitted", "completed", "failed", "rejected", "active",
                                           "pending", "bytes_in_use", "bytes_peak", "budget", "batch_steps")] + [
        (n, C.c_double) for n in ("wall_s", "worlds_per_s", "latency_p50_s", "latency_p99_s", "latency_mean_s",
                                  "latency_max_s", "cell_updates")] + [("reserved", C.c_int64 * 4)]


class PrecisionField(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("max_abs_fp32", "rms_fp32", "max_abs_comp", "rms_comp", "rms_ref")]


class PrecisionReport(C.Structure):
    _fields_ = [("steps", C.c_int32), ("width", C.c_int32), ("height_cells", C.c_int32), ("reserved0", C.c_int32),
                ("height", PrecisionField), ("water", PrecisionField), ("sediment", PrecisionField)] + [
        (n, C.c_double) for n in ("mass_start", "mass_ref", "mass_fp32", "mass_comp", "drift_ref", "drift_fp32",
                                  "drift_comp", "ms_fp32", "ms_comp", "ms_ref")] + [("reserved", C.c_double * 4)]


class HbmProbe(C.Structure):
    _fields_ = [("copy_gbps", C.c_double), ("read_gbps", C.c_double), ("write_gbps", C.c_double),
                ("streams_gbps", C.c_double), ("bytes", C.c_int64), ("reps", C.c_int32), ("reserved", C.c_int32)]


# numpy-compatible record layouts of the per-cell structs
import numpy as np  # noqa: E402

class BatchTimes(C.Structure):
    _fields_ = [("ms_pre", C.c_double), ("ms_erosion", C.c_double), ("ms_post", C.c_double),
                ("steps", C.c_int64), ("reserved", C.c_double * 4)]


class VoronoiSeed(C.Structure):
    """struct VoronoiSeed (types/voronoi_seed.h) = tt_voronoi_seed."""
    _fields_ = [("position", C.c_float * 2), ("id", C.c_uint8), ("pad_", C.c_uint8 * 3)]


VORONOI_SEED_DTYPE = np.dtype([("position", "<f4", 2), ("id", "u1"), ("pad", "u1", 3)])
assert VORONOI_SEED_DTYPE.itemsize == 12 and C.sizeof(VoronoiSeed) == 12

UPLIFT_DTYPE = np.dtype([("cum", "<f4"), ("types", "u1"), ("pad", "u1", 3)])
PROPAGATED_DTYPE = np.dtype([("cum", "<f4"), ("inv", "<u4"), ("types", "u1"), ("pad", "u1", 3)])
DISTANCE_DTYPE = np.dtype([("dist", "<i4"), ("orig", "<u4")])
PLATE_DTYPE = np.dtype([
    ("velocity", "<f4"), ("velocitySmoothed", "<f4"), ("direction", "<f4", 2),
    ("directionSmoothed", "<f4", 2), ("pixelCenter", "<f4", 2), ("velocityChange", "<f4", 2),
    ("geometricCenter", "<f4", 2), ("asthenosphereVelocity", "<f4", 2), ("size", "<i4"),
    ("mass", "<f4"), ("perimeter", "<i4"), ("breakScore", "<f4"), ("circularity", "<f4"),
    ("used", "u1"), ("hasMoved", "u1"), ("pad_", "u1", 2), ("mergeWaitTime", "<i4")])
assert PLATE_DTYPE.itemsize == 84 and C.sizeof(PlateData) == 84


def field_spec(field: int, n_cells: int):
    """(numpy dtype, shape) of a field for a world with n_cells cells."""
    f4 = np.dtype("<f4")
    table = {
        FIELD_HEIGHT: (f4, (n_cells,)),
        FIELD_PLATE_IDS: (np.dtype("u1"), (n_cells,)),
        FIELD_WATER: (f4, (n_cells,)),
        FIELD_FLUX: (f4, (n_cells, 4)),
        FIELD_VELOCITY: (f4, (n_cells, 2)),
        FIELD_SEDIMENT: (f4, (n_cells,)),
        FIELD_ASTHENOSPHERE: (f4, (n_cells, 2)),
        FIELD_COLLISIONS: (np.dtype("<u4"), (n_cells,)),
        FIELD_ACCRETION: (np.dtype("u1"), (n_cells,)),
        FIELD_UPLIFT: (UPLIFT_DTYPE, (n_cells,)),
        FIELD_PROPAGATED: (PROPAGATED_DTYPE, (n_cells,)),
        FIELD_PLATE_DATA: (PLATE_DTYPE, (MAX_PLATES,)),
        FIELD_DIVERGENCE_BITMAP: (np.dtype("<u4"), (NUM_TRI_WORDS,)),
        FIELD_COLLISION_TYPES: (np.dtype("u1"), (NUM_TRI,)),
        FIELD_SORTED_KEYS: (np.dtype("<u8"), (n_cells,)),
        FIELD_LABELS: (np.dtype("<u4"), (n_cells,)),
        FIELD_DISTANCE: (DISTANCE_DTYPE, (n_cells,)),
    }
    return table[field]


FIELD_NAMES = {
    FIELD_HEIGHT: "height", FIELD_PLATE_IDS: "plate_ids", FIELD_WATER: "water", FIELD_FLUX: "flux",
    FIELD_VELOCITY: "velocity", FIELD_SEDIMENT: "sediment", FIELD_ASTHENOSPHERE: "asthenosphere",
    FIELD_COLLISIONS: "collisions", FIELD_ACCRETION: "accretion", FIELD_UPLIFT: "uplift",
    FIELD_PROPAGATED: "propagated", FIELD_PLATE_DATA: "plate_data",
    FIELD_DIVERGENCE_BITMAP: "divergence_bitmap", FIELD_COLLISION_TYPES: "collision_types",
    FIELD_SORTED_KEYS: "sorted_keys", FIELD_LABELS: "labels", FIELD_DISTANCE: "distance",
}


def load_noise() -> bytes:
    data = NOISE_PATH.read_bytes()
    if len(data) != 65536:
        raise RuntimeError(f"{NOISE_PATH} has {len(data)} bytes, expected 65536")
    return data


_lib = None


def lib() -> C.CDLL:
    """Load libtectonic.so (built in-tree).  Raises if it is missing: there is no
    fallback path."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build() first")
    L = C.CDLL(str(LIB_PATH))
    vp = C.c_void_p
    L.tt_status_string.restype = C.c_char_p
    L.tt_status_string.argtypes = [C.c_int]
    L.tt_last_error.restype = C.c_char_p
    L.tt_last_error.argtypes = []
    L.tt_device_count.argtypes = [C.POINTER(C.c_int)]
    L.tt_settings_default.argtypes = [C.POINTER(KernelSettings)]
    L.tt_config_default.argtypes = [C.POINTER(SimConfig)]
    L.tt_config_parse_json.argtypes = [C.c_char_p, C.POINTER(SimConfig), C.POINTER(KernelSettings)]
    L.tt_world_create.argtypes = [C.POINTER(SimConfig), C.POINTER(KernelSettings), C.POINTER(vp)]
    L.tt_world_init.argtypes = [vp]
    L.tt_world_reset.argtypes = [vp, C.c_uint32, C.c_int32, C.POINTER(C.c_int32), C.c_int32]
    L.tt_generate_plate_centers.argtypes = [C.POINTER(C.c_uint64), C.c_int32, C.c_int32, C.c_int32,
                                            C.POINTER(VoronoiSeed)]
    L.tt_generate_voronoi_seeds.argtypes = [C.POINTER(C.c_uint64), C.POINTER(VoronoiSeed), C.c_int32, C.c_int32,
                                            C.c_int32, C.c_int32, C.POINTER(VoronoiSeed)]
    L.tt_generate_plate_data.argtypes = [C.POINTER(C.c_uint64), C.POINTER(VoronoiSeed), C.c_int32,
                                         C.POINTER(PlateData)]
    L.tt_world_step.argtypes = [vp, C.c_int32]
    L.tt_world_config.argtypes = [vp, C.POINTER(SimConfig), C.POINTER(KernelSettings)]
    L.tt_world_set_settings.argtypes = [vp, C.POINTER(KernelSettings)]
    L.tt_world_download.argtypes = [vp, C.c_int32, vp, C.c_size_t]
    L.tt_world_upload.argtypes = [vp, C.c_int32, vp, C.c_size_t]
    L.tt_world_stats.argtypes = [vp, C.POINTER(StepStats)]
    L.tt_world_iteration.argtypes = [vp, C.POINTER(C.c_uint32)]
    L.tt_world_bytes.argtypes = [C.POINTER(SimConfig), C.POINTER(C.c_size_t)]
    L.tt_world_destroy.argtypes = [vp]
    L.tt_world_export_png.argtypes = [vp, C.c_int32, C.c_char_p]
    L.tt_world_set_timing.argtypes = [vp, C.c_int32]
    L.tt_world_stage_times.argtypes = [vp, C.POINTER(StageTimes)]
    L.tt_math_eval_device.argtypes = [C.c_int32, vp, vp, C.c_int64]
    L.tt_math_eval_device_static.argtypes = [C.c_int32, vp, vp, C.c_int64]
    L.tt_world_set_iteration.argtypes = [vp, C.c_uint32]
    L.tt_batch_create.argtypes = [C.POINTER(vp)]
    L.tt_batch_add.argtypes = [vp, vp]
    L.tt_batch_remove.argtypes = [vp, vp]
    L.tt_batch_size.argtypes = [vp, C.POINTER(C.c_int32)]
    L.tt_batch_step.argtypes = [vp, C.c_int32]
    L.tt_batch_set_threads.argtypes = [vp, C.c_int32]
    L.tt_batch_set_graph.argtypes = [vp, C.c_int32]
    L.tt_batch_destroy.argtypes = [vp]
    L.tt_batch_stage_times.argtypes = [vp, C.POINTER(BatchTimes)]
    L.tt_server_create.argtypes = [C.c_uint64, C.c_int32, C.POINTER(vp)]
    L.tt_server_set_keep.argtypes = [vp, C.c_int32]
    L.tt_server_submit.argtypes = [vp, C.POINTER(SimConfig), C.POINTER(KernelSettings), C.c_int32,
                                   C.POINTER(C.c_int64)]
    L.tt_server_result.argtypes = [vp, C.c_int64, C.c_int32, C.POINTER(JobResult)]
    L.tt_server_take.argtypes = [vp, C.c_int64, C.POINTER(vp)]
    L.tt_server_drain.argtypes = [vp, C.c_int32]
    L.tt_server_stats.argtypes = [vp, C.POINTER(ServerReport)]
    L.tt_server_destroy.argtypes = [vp]
    L.tt_world_diagnostics_json.argtypes = [vp, C.c_char_p, C.c_size_t, C.POINTER(C.c_size_t)]
    L.tt_world_save.argtypes = [vp, C.c_char_p]
    L.tt_world_load.argtypes = [C.c_char_p, C.POINTER(vp)]
    L.tt_precision_study.argtypes = [C.POINTER(SimConfig), C.POINTER(KernelSettings), C.c_int32,
                                     C.POINTER(PrecisionReport)]
    L.tt_probe_hbm.argtypes = [C.c_int64, C.c_int32, C.POINTER(HbmProbe)]
    _lib = L
    return L


def check(status: int, what: str = "") -> None:
    if status != TT_OK:
        msg = ""
        try:
            msg = lib().tt_last_error().decode(errors="replace")
        except Exception:  # pragma: no cover - error path while reporting an error
            pass
        raise RuntimeError(f"{what}: {STATUS_NAMES.get(status, status)} {msg}".strip())
