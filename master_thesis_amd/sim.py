"""Python host mirror of the reference's step interface over the C ABI.

`World` is a thin handle on a libtectonic world (tectonic.h).
`PlateTectonicSim` keeps the public surface of the reference class
(src/cuda_compute/plate_tectonic_sim.h:111-130): constructor
(width, height, seed), initialize(numStartingPlates, numVoronoiSeeds),
executeIteration(), resetSim(seed, numStartingPlates, numVoronoiSeeds).
Every call goes to the gfx950 kernels; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
from typing import Sequence

import numpy as np

from . import _abi
from ._abi import check, lib


class World:
    def __init__(self, cfg: _abi.SimConfig, ks: _abi.KernelSettings):
        self.cfg = cfg.copy()
        self.ks = ks.copy()
        self.n = cfg.width * cfg.height
        h = C.c_void_p()
        check(lib().tt_world_create(C.byref(self.cfg), C.byref(self.ks), C.byref(h)), "tt_world_create")
        self.h = h

    @classmethod
    def _adopt(cls, handle: C.c_void_p, cfg: _abi.SimConfig) -> "World":
        """Wrap a world handle created by the library (a finished server job)."""
        w = cls.__new__(cls)
        w.cfg = cfg.copy()
        w.ks = _abi.KernelSettings.defaults()
        w.n = cfg.width * cfg.height
        w.h = handle
        return w

    # lifecycle -----------------------------------------------------------
    def init(self) -> "World":
        check(lib().tt_world_init(self.h), "tt_world_init")
        return self

    def reset(self, seed: int, plates: int, voronoi_gens: Sequence[int] = ()) -> None:
        """resetSim(seed, numStartingPlates, numVoronoiSeeds): re-initialise
        with any number of Voronoi generations."""
        gens = [int(g) for g in voronoi_gens]
        arr = (C.c_int32 * max(1, len(gens)))(*gens)
        check(lib().tt_world_reset(self.h, seed, plates, arr, len(gens)), "tt_world_reset")
        self.cfg.seed, self.cfg.num_plates = seed, plates
        self.cfg.num_voronoi_gens = min(4, len(gens))
        for i in range(self.cfg.num_voronoi_gens):
            self.cfg.voronoi_gens[i] = gens[i]

    def step(self, n: int = 1) -> None:
        check(lib().tt_world_step(self.h, n), "tt_world_step")

    def close(self) -> None:
        if getattr(self, "h", None):
            check(lib().tt_world_destroy(self.h), "tt_world_destroy")
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # state ----------------------------------------------------------------
    def download(self, field: int) -> np.ndarray:
        dt, shape = _abi.field_spec(field, self.n)
        a = np.empty(shape, dtype=dt)
        check(lib().tt_world_download(self.h, field, a.ctypes.data, a.nbytes), f"download {field}")
        return a

    def upload(self, field: int, a: np.ndarray) -> None:
        dt, shape = _abi.field_spec(field, self.n)
        a = np.ascontiguousarray(a, dtype=dt).reshape(shape)
        check(lib().tt_world_upload(self.h, field, a.ctypes.data, a.nbytes), f"upload {field}")

    def set_settings(self, ks: _abi.KernelSettings) -> None:
        check(lib().tt_world_set_settings(self.h, C.byref(ks)), "set_settings")
        self.ks = ks.copy()

    def stats(self) -> _abi.StepStats:
        s = _abi.StepStats()
        check(lib().tt_world_stats(self.h, C.byref(s)), "stats")
        return s

    @property
    def iteration(self) -> int:
        it = C.c_uint32()
        check(lib().tt_world_iteration(self.h, C.byref(it)), "iteration")
        return it.value

    def set_iteration(self, it: int) -> None:
        check(lib().tt_world_set_iteration(self.h, it), "set_iteration")

    def set_timing(self, on: bool) -> None:
        check(lib().tt_world_set_timing(self.h, 1 if on else 0), "set_timing")

    def stage_times(self) -> _abi.StageTimes:
        t = _abi.StageTimes()
        check(lib().tt_world_stage_times(self.h, C.byref(t)), "stage_times")
        return t

    def export_png(self, what: int, path: str) -> None:
        check(lib().tt_world_export_png(self.h, what, path.encode()), "export_png")

    def diagnostics(self) -> dict:
        """Step statistics, field summaries, collision-class counts and one
        record per live plate (tt_world_diagnostics_json)."""
        import json
        n = C.c_size_t()
        check(lib().tt_world_diagnostics_json(self.h, None, 0, C.byref(n)), "diagnostics")
        buf = C.create_string_buffer(n.value + 1)
        check(lib().tt_world_diagnostics_json(self.h, buf, n.value + 1, C.byref(n)), "diagnostics")
        return json.loads(buf.value.decode())

    def save(self, path: str) -> None:
        check(lib().tt_world_save(self.h, str(path).encode()), "tt_world_save")

    @classmethod
    def load(cls, path: str) -> "World":
        h = C.c_void_p()
        check(lib().tt_world_load(str(path).encode(), C.byref(h)), "tt_world_load")
        w = cls._adopt(h, _abi.SimConfig())
        return w._refresh_geometry()

    def _refresh_geometry(self) -> "World":
        check(lib().tt_world_config(self.h, C.byref(self.cfg), C.byref(self.ks)), "tt_world_config")
        self.n = self.cfg.width * self.cfg.height
        return self


class Batch:
    """Many worlds stepped together (tt_batch_*): one launch per erosion pass
    over all members, replayed from a captured hipGraph; the other stages per
    member on its own stream.  Members keep their own results bit for bit."""

    def __init__(self, worlds=(), threads: int | None = None, graph: bool = True):
        h = C.c_void_p()
        check(lib().tt_batch_create(C.byref(h)), "tt_batch_create")
        self.h = h
        self.members: list[World] = []
        if threads is not None:
            check(lib().tt_batch_set_threads(self.h, threads), "tt_batch_set_threads")
        if not graph:
            check(lib().tt_batch_set_graph(self.h, 0), "tt_batch_set_graph")
        for w in worlds:
            self.add(w)

    def add(self, w: World) -> None:
        check(lib().tt_batch_add(self.h, w.h), "tt_batch_add")
        self.members.append(w)

    def remove(self, w: World) -> None:
        check(lib().tt_batch_remove(self.h, w.h), "tt_batch_remove")
        self.members.remove(w)

    def step(self, n: int = 1) -> None:
        check(lib().tt_batch_step(self.h, n), "tt_batch_step")

    def stage_times(self) -> dict:
        """Device time (ms, summed over steps) of the pre-erosion phases, the
        batched erosion and the post-erosion phases."""
        t = _abi.BatchTimes()
        check(lib().tt_batch_stage_times(self.h, C.byref(t)), "tt_batch_stage_times")
        return {"ms_pre": t.ms_pre, "ms_erosion": t.ms_erosion, "ms_post": t.ms_post, "steps": int(t.steps)}

    def __len__(self) -> int:
        n = C.c_int32()
        check(lib().tt_batch_size(self.h, C.byref(n)), "tt_batch_size")
        return n.value

    def close(self) -> None:
        if getattr(self, "h", None):
            lib().tt_batch_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """World-generation jobs under an HBM budget (tt_server_*): continuous
    batching of the resident worlds, FIFO admission by tt_world_bytes()."""

    def __init__(self, budget_bytes: int, max_active: int = 256, keep: bool = False):
        h = C.c_void_p()
        check(lib().tt_server_create(int(budget_bytes), int(max_active), C.byref(h)), "tt_server_create")
        self.h = h
        if keep:
            check(lib().tt_server_set_keep(self.h, 1), "tt_server_set_keep")

    def submit(self, cfg: _abi.SimConfig, ks: _abi.KernelSettings, steps: int) -> int:
        jid = C.c_int64()
        check(lib().tt_server_submit(self.h, C.byref(cfg), C.byref(ks), int(steps), C.byref(jid)), "tt_server_submit")
        return jid.value

    def result(self, jid: int, wait_ms: int = -1) -> _abi.JobResult | None:
        r = _abi.JobResult()
        st = lib().tt_server_result(self.h, int(jid), int(wait_ms), C.byref(r))
        if st == _abi.TT_ERR_STATE:
            return None
        check(st, "tt_server_result")
        return r

    def take(self, jid: int) -> World:
        h = C.c_void_p()
        check(lib().tt_server_take(self.h, int(jid), C.byref(h)), "tt_server_take")
        r = self.result(jid, 0)
        cfg = _abi.SimConfig()
        cfg.width, cfg.height = r.width, r.height
        return World._adopt(h, cfg)._refresh_geometry()

    def drain(self, timeout_ms: int = -1) -> None:
        check(lib().tt_server_drain(self.h, int(timeout_ms)), "tt_server_drain")

    def report(self) -> _abi.ServerReport:
        r = _abi.ServerReport()
        check(lib().tt_server_stats(self.h, C.byref(r)), "tt_server_stats")
        return r

    def close(self) -> None:
        if getattr(self, "h", None):
            lib().tt_server_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


STATE_FIELDS = (
    _abi.FIELD_HEIGHT, _abi.FIELD_PLATE_IDS, _abi.FIELD_WATER, _abi.FIELD_FLUX, _abi.FIELD_VELOCITY,
    _abi.FIELD_SEDIMENT, _abi.FIELD_UPLIFT, _abi.FIELD_PLATE_DATA, _abi.FIELD_DIVERGENCE_BITMAP,
    _abi.FIELD_COLLISION_TYPES,
)


def world_bytes(cfg: _abi.SimConfig) -> int:
    b = C.c_size_t()
    check(lib().tt_world_bytes(C.byref(cfg), C.byref(b)), "tt_world_bytes")
    return b.value


def device_count() -> int:
    n = C.c_int()
    check(lib().tt_device_count(C.byref(n)), "tt_device_count")
    return n.value


def math_eval_device(fn: int, x: np.ndarray) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.float32)
    per = {_abi.TT_MATH_SIMPLEX: 3, _abi.TT_MATH_DISCRETE: 3, _abi.TT_MATH_FMODF: 2, _abi.TT_MATH_ATAN2: 2,
           _abi.TT_MATH_DIV_RCP: 2}.get(fn, 1)
    n = x.size // per
    out = np.empty(n, dtype=np.float32)
    check(lib().tt_math_eval_device(fn, x.ctypes.data, out.ctypes.data, n), "tt_math_eval_device")
    return out


def precision_study(cfg: _abi.SimConfig, ks: _abi.KernelSettings, steps: int) -> dict:
    """Erosion-only steps in fp32, compensated fp32 and an fp64 twin from one
    initial state (tt_precision_study): per-field errors and mass drift."""
    r = _abi.PrecisionReport()
    check(lib().tt_precision_study(C.byref(cfg), C.byref(ks), int(steps), C.byref(r)), "tt_precision_study")
    out = {"steps": r.steps, "width": r.width, "height": r.height_cells}
    for f in ("height", "water", "sediment"):
        pf = getattr(r, f)
        out[f] = {k: getattr(pf, k) for k, _ in _abi.PrecisionField._fields_}
    for k in ("mass_start", "mass_ref", "mass_fp32", "mass_comp", "drift_ref", "drift_fp32", "drift_comp",
              "ms_fp32", "ms_comp", "ms_ref"):
        out[k] = getattr(r, k)
    return out


def probe_hbm(nbytes: int = 2 << 30, reps: int = 5) -> dict:
    """Measured HBM ceilings (tt_probe_hbm): best GB/s of a float4 copy (read
    plus write bytes), a read-only pass, a fill, and a ten-stream copy with the
    hydraulic kernel's 64 bytes per cell."""
    r = _abi.HbmProbe()
    check(lib().tt_probe_hbm(int(nbytes), int(reps), C.byref(r)), "tt_probe_hbm")
    return {"copy_gbps": r.copy_gbps, "read_gbps": r.read_gbps, "write_gbps": r.write_gbps,
            "streams_gbps": r.streams_gbps, "bytes": r.bytes, "reps": r.reps}


def parse_json_config(text: str, cfg: _abi.SimConfig | None = None, ks: _abi.KernelSettings | None = None):
    cfg = cfg.copy() if cfg is not None else _abi.SimConfig.defaults()
    ks = ks.copy() if ks is not None else _abi.KernelSettings.defaults()
    check(lib().tt_config_parse_json(text.encode(), C.byref(cfg), C.byref(ks)), "tt_config_parse_json")
    return cfg, ks


class PlateTectonicSim:
    """Mirror of PlateTectonicSim (plate_tectonic_sim.h:56-179) without the GL
    interop manager (headless; plate_tectonic_sim.cu:49,121 accept nullptr)."""

    def __init__(self, width: int, height: int, seed: int, interop_manager=None,
                 settings: _abi.KernelSettings | None = None, config: _abi.SimConfig | None = None):
        if interop_manager is not None:
            raise ValueError("the OpenGL interop path is not part of this library (headless only)")
        self._cfg = config.copy() if config is not None else _abi.SimConfig.defaults()
        self._cfg.width, self._cfg.height, self._cfg.seed = width, height, seed
        self._ks = settings.copy() if settings is not None else _abi.KernelSettings.defaults()
        self._world: World | None = None

    def initialize(self, numStartingPlates: int, numVoronoiSeeds: Sequence[int]) -> None:
        if self._world is not None:
            self._world.close()
        cfg = self._cfg.copy()
        cfg.num_plates = numStartingPlates
        cfg.num_voronoi_gens = 0
        self._world = World(cfg, self._ks)
        self._world.reset(self._cfg.seed, numStartingPlates, numVoronoiSeeds)
        self._cfg = self._world.cfg.copy()

    def executeIteration(self) -> None:
        if self._world is None:
            raise RuntimeError("initialize() first")
        self._world.step(1)

    def resetSim(self, seed: int, numStartingPlates: int, numVoronoiSeeds: Sequence[int]) -> None:
        self._cfg.seed = seed
        if self._world is None:
            self.initialize(numStartingPlates, numVoronoiSeeds)
        else:
            self._world.reset(seed, numStartingPlates, numVoronoiSeeds)
            self._cfg = self._world.cfg.copy()

    # static generators of the reference class (plate_tectonic_sim.h:121-130);
    # the generator is a SplitMix64 state, a one-element list advanced in place
    @staticmethod
    def generatePlateCenters(generator: list, numPlates: int, width: int, height: int) -> np.ndarray:
        st = C.c_uint64(generator[0])
        out = (_abi.VoronoiSeed * max(1, numPlates))()
        check(lib().tt_generate_plate_centers(C.byref(st), numPlates, width, height, out), "tt_generate_plate_centers")
        generator[0] = st.value
        return np.frombuffer(bytes(out), _abi.VORONOI_SEED_DTYPE)[:numPlates].copy()

    @staticmethod
    def generateVoronoiSeeds(generator: list, centerSeeds: np.ndarray, numSeeds: int, width: int,
                             height: int) -> np.ndarray:
        st = C.c_uint64(generator[0])
        c = np.ascontiguousarray(centerSeeds, _abi.VORONOI_SEED_DTYPE)
        out = (_abi.VoronoiSeed * max(1, numSeeds))()
        check(lib().tt_generate_voronoi_seeds(C.byref(st), c.ctypes.data_as(C.POINTER(_abi.VoronoiSeed)), len(c),
                                              numSeeds, width, height, out), "tt_generate_voronoi_seeds")
        generator[0] = st.value
        return np.frombuffer(bytes(out), _abi.VORONOI_SEED_DTYPE)[:numSeeds].copy()

    @staticmethod
    def generatePlateData(generator: list, plateCenters: np.ndarray, numStartingPlates: int) -> np.ndarray:
        st = C.c_uint64(generator[0])
        c = np.ascontiguousarray(plateCenters, _abi.VORONOI_SEED_DTYPE)
        out = np.zeros(_abi.MAX_PLATES, _abi.PLATE_DTYPE)
        check(lib().tt_generate_plate_data(C.byref(st), c.ctypes.data_as(C.POINTER(_abi.VoronoiSeed)),
                                           numStartingPlates, out.ctypes.data_as(C.POINTER(_abi.PlateData))),
              "tt_generate_plate_data")
        generator[0] = st.value
        return out

    @property
    def world(self) -> World:
        if self._world is None:
            raise RuntimeError("initialize() first")
        return self._world
