"""Named workloads: the reference defaults and the benchmark worlds.

Each preset returns (SimConfig, KernelSettings).  Parameter mapping from the
benchmark knobs to reference parameters (kernel_settings.cuh:6-71,
generation_settings.h:67-76):

* rain -> hydrationRainfall, evaporation -> hydrationEvaporation,
  capacity -> sedimentCapacity, deposition -> sedimentDepositionRate;
* talus 0.6 -> thermalThresholdAngle * thermalCellSize = 0.6 * 1.0;
* "uplift 0.01 per overlap" -> continentalUpliftMultiplier = 0.01;
* "blur radius r" -> upliftRange = r (read only by the bounded-radius uplift
  variant; the live path is the jump flood);
* hydraulic / thermal iteration counts -> hydraulic_iterations /
  thermal_iterations (hard-coded 10 and 1 in the reference,
  plate_tectonic_sim.cu:336 and :73);
* CCL period -> iterations_for_ccl.
"""
from __future__ import annotations

import numpy as np

from ._abi import KernelSettings, SimConfig, TT_INIT_SYNTHETIC, TT_STAGE_ALL


def reference_defaults(width: int = 1024, height: int = 1024):
    return SimConfig.defaults(), KernelSettings.defaults()


def _erosion(ks: KernelSettings, uplift_range: int) -> KernelSettings:
    ks.hydrationRainfall = 0.01
    ks.hydrationEvaporation = 0.02
    ks.sedimentCapacity = 0.05
    ks.sedimentDepositionRate = 0.3
    ks.thermalThresholdAngle = 0.6
    ks.thermalCellSize = 1.0
    ks.continentalUpliftMultiplier = 0.01
    ks.upliftRange = uplift_range
    return ks


def synthetic(width: int, height: int, plates: int, speed: tuple, seed: int = 1, *,
              hydraulic: int = 50, thermal: int = 20, ccl_every: int = 10, uplift_range: int = 3):
    cfg = SimConfig.defaults()
    cfg.width, cfg.height, cfg.seed, cfg.num_plates = width, height, seed, plates
    cfg.num_voronoi_gens = 0
    cfg.iterations_for_ccl = ccl_every
    cfg.hydraulic_iterations, cfg.thermal_iterations = hydraulic, thermal
    cfg.init_mode = TT_INIT_SYNTHETIC
    cfg.speed_min, cfg.speed_max = float(speed[0]), float(speed[1])
    cfg.stage_mask = TT_STAGE_ALL
    return cfg, _erosion(KernelSettings.defaults(), uplift_range)


def world_512(seed: int = 1):
    """512x512, 8 plates, speed U[0.5, 2], K=50, T=20, blur radius 3, CCL every 10 (100 steps)."""
    return synthetic(512, 512, 8, (0.5, 2.0), seed, ccl_every=10, uplift_range=3)


def world_4096(seed: int = 1):
    """4096x4096, 32 plates, speed U[0.5, 3], K=50, T=20, blur radius 5, CCL every step (500 steps)."""
    return synthetic(4096, 4096, 32, (0.5, 3.0), seed, ccl_every=1, uplift_range=5)


def world_8192(seed: int = 1):
    """8192x8192, 64 plates (precision study), erosion as world_4096 (300 steps)."""
    return synthetic(8192, 8192, 64, (0.5, 3.0), seed, ccl_every=1, uplift_range=5)


def round64(x: float) -> int:
    return max(64, int(round(x / 64.0)) * 64)


def mixed_world(rng, lo: int, hi: int, pmin: int, pmax: int, seed: int):
    """A square world with side log-uniform in [lo, hi] rounded to a multiple
    of 64 and plates uniform in {pmin..pmax}; the erosion of world_4096."""
    side = round64(float(np.exp(rng.uniform(np.log(lo), np.log(hi)))))
    plates = int(rng.integers(pmin, pmax + 1))
    return synthetic(side, side, plates, (0.5, 3.0), seed, hydraulic=50, thermal=20, ccl_every=1, uplift_range=5)


def batch_mix(n: int = 256, seed: int = 1):
    """The batched workload: n worlds, sides log-uniform in [256, 2048]
    (multiples of 64), plates U{4..40}, world seeds seed + i (200 steps)."""
    rng = np.random.default_rng(seed)
    return [mixed_world(rng, 256, 2048, 4, 40, seed + i) for i in range(n)]


STEPS = {"world_512": 100, "world_4096": 500, "world_8192": 300, "batch_mix": 200}
