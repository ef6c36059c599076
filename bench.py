#!/usr/bin/env python3
"""Benchmark of the plain terrain step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]

Workload ("plain method on one MI355X"): a 4096x4096
fp32 world with 32 Voronoi plates, plate speeds U[0.5, 3] cells/step, 50
hydraulic sub-iterations (rain 0.01, evaporation 0.02, capacity 0.05,
deposition 0.3), 20 thermal iterations (talus 0.6), uplift 0.01, blur radius 5
(upliftRange), CCL every step.  Synthetic seeded initial state (uniform
heights smoothed by a sigma-8 Gaussian), created before timing; the timed
region is W untimed warm-up steps followed by exactly K steps with every input
already resident in HBM.

Prints one JSON line with the throughput (cell-updates/s), the roofline of the
dominant kernel (the fused hydraulic sub-iteration, HBM-bound) timed with HIP
events on the world's stream, and the CPU oracle timed on this host on one
step of the same world from the same initial state.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
PROFILES = ROOT / "profiles"


def hydraulic_bytes_per_cell(first: bool) -> int:
    """Compulsory HBM bytes of one fused hydraulic launch per cell.

    reads: height 4, water 4, flux 16, and either the step's sediment 4 (first
    sub-iteration) or velocity 8 + the transported sediment gather 4; writes:
    height 4, water 4, flux 16, velocity 8, sediment 4.
    """
    reads = 4 + 4 + 16 + (4 if first else 8 + 4)
    writes = 4 + 4 + 16 + 8 + 4
    return reads + writes


def model_bytes_per_cell_step(K: int, T: int, side: int, C: int) -> float:
    """Algorithmic bytes per cell per step at reference-pass granularity:
    B = 135 + 8P + 8T + 128K + 10/C with P = ceil(log2(side))."""
    P = (side - 1).bit_length()
    return 135 + 8 * P + 8 * T + 128 * K + 10.0 / C


def load_traffic():
    p = PROFILES / "hydraulic_traffic.json"
    if p.exists():
        try:
            return json.loads(p.read_text()).get("bytes_per_launch")
        except Exception:
            return None
    return None


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--side", type=int, default=4096)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--tile-rows", type=int, default=0, help="erosion tile rows (0: autotuned)")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="run one plain device copy of this size first (calibrates HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported (worlds are independent; "
              "a multi-GPU run would be N copies of this measurement)", file=sys.stderr)
        return 2

    from master_thesis_amd import _abi, configs
    from master_thesis_amd.sim import World, STATE_FIELDS

    if args.copy_probe_mib > 0:
        # a plain copy kernel of known size: its FETCH_SIZE / WRITE_SIZE counters
        # give the scale factor of the hardware byte counters
        import torch
        n = args.copy_probe_mib * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    cfg, ks = configs.world_4096(seed=args.seed)
    if args.side != 4096:
        cfg.width = cfg.height = args.side
    cfg.tile_rows = args.tile_rows
    N = cfg.width * cfg.height
    world = World(cfg, ks).init()
    init_state = None
    if not args.no_cpu_baseline:
        init_state = {f: world.download(f) for f in STATE_FIELDS}
    world.step(args.warmup)
    world.set_timing(True)
    t0 = time.perf_counter()
    world.step(args.steps)  # returns after the device work is complete
    t1 = time.perf_counter()
    times = world.stage_times()
    stats = world.stats()
    wall = t1 - t0
    value = N * args.steps / wall

    K = cfg.hydraulic_iterations
    launches = max(times.hydraulic_launches, 1.0)
    avg_ms = times.ms_hydraulic / launches
    bpl = N * (hydraulic_bytes_per_cell(True) + (K - 1) * hydraulic_bytes_per_cell(False)) / max(K, 1)
    achieved = bpl / (avg_ms * 1e-3) / 1e9
    model_b = model_bytes_per_cell_step(K, cfg.thermal_iterations, max(cfg.width, cfg.height),
                                        cfg.iterations_for_ccl)
    result = {
        "metric": "cell-updates/s per full step at 4096²; served worlds/s + p50/p99 latency",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1000.0 * wall / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded uniform heights, Gaussian sigma 8; Voronoi plates)",
        "config": {
            "workload": f"plain step, {cfg.width}x{cfg.height}, {cfg.num_plates} plates, speed U[0.5,3], "
                        f"K={K} hydraulic, T={cfg.thermal_iterations} thermal, CCL every "
                        f"{cfg.iterations_for_ccl} step, blur radius {ks.upliftRange}",
            "side": cfg.width, "plates": cfg.num_plates, "hydraulic_iterations": K,
            "thermal_iterations": cfg.thermal_iterations, "ccl_every": cfg.iterations_for_ccl,
            "seed": cfg.seed, "parallelism": "single world per GPU",
        },
        "roofline": {
            "kernel": "k_hydraulic (fused rain/flux/flow/evaporate/erosion + transport gather)",
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": load_traffic(),
            "bytes_per_launch": bpl,
            "avg_launch_ms": avg_ms,
        },
        "stages_ms_per_step": {
            "collide": times.ms_collide / args.steps, "uplift": times.ms_uplift / args.steps,
            "thermal": times.ms_thermal / args.steps, "hydraulic": times.ms_hydraulic / args.steps,
            "transport": times.ms_transport / args.steps, "post": times.ms_post / args.steps,
            "split": times.ms_split / args.steps, "events_total": times.ms_total / args.steps,
        },
        "step_model": {
            "bytes_per_cell_step": model_b,
            "frac_of_hbm": model_b * N * args.steps / wall / (HBM_PEAK_GBS * 1e9),
        },
        "final_stats": {"plates": stats.number_of_plates, "labels": stats.num_labels,
                        "iteration": stats.iteration, "tile_rows": stats.tile_rows},
    }
    world.close()

    if init_state is not None:
        from oracle.oracle_py import OracleWorld
        threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
        ow = OracleWorld(cfg, ks)
        for f, a in init_state.items():
            ow.set(f, a)
        ow.set_iteration(0)
        c0 = time.perf_counter()
        ow.step(1)
        c1 = time.perf_counter()
        result["cpu_baseline"] = {
            "value": N / (c1 - c0),
            "unit": "cell-updates/s",
            "cores": threads,
            "kind": "port",
            "sample": f"1 step of the same {cfg.width}x{cfg.height} world from the same initial state "
                      f"(CPU restatement, OpenMP, {threads} threads)",
        }
    print(json.dumps(result))
    return 0


if __name__ == "__main__":
    sys.exit(main())
