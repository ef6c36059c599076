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
already resident in HBM.  The defaults (W = 0, K = 500) time the whole
500-step run of the workload.

Prints one JSON line with the throughput (cell-updates/s), the roofline of the
dominant kernel (the fused hydraulic sub-iteration, HBM-bound) timed with HIP
events on the world's stream, and the CPU oracle timed on this host on one
step of the same world from the same initial state.

By default (--mode all) the line also carries two bounded samples of the
other system forms: "batched" (256 mixed-size worlds stepped as one batch and
one at a time) and "serving" (Poisson world-generation requests under a 200 GB
HBM budget, worlds/s and p50/p99 latency).  --mode batch / serve print those
alone; --mode plain prints the 4096^2 measurement only.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
PROFILES = ROOT / "profiles"


def hydraulic_bytes_per_cell(first: bool, last: bool) -> int:
    """Compulsory HBM bytes of one fused hydraulic launch per cell.

    reads: height 4, water 4 (the step's state water on the first
    sub-iteration, else the previous sub-iteration's water after its flow),
    flux 16, and the step's sediment 4 (first) or the transported-sediment
    gather 4; writes: height 4, water 4 (the flow water, or the state water on
    the last sub-iteration), flux 16, sediment 4, and the velocity 8 on the
    last sub-iteration.
    """
    del first
    reads = 4 + 4 + 16 + 4
    writes = 4 + 4 + 16 + 4 + (8 if last else 0)
    return reads + writes


def hydraulic_bytes_per_launch_avg(K: int) -> float:
    """Average of hydraulic_bytes_per_cell over the K launches of a step."""
    if K <= 0:
        return 0.0
    return sum(hydraulic_bytes_per_cell(k == 0, k == K - 1) for k in range(K)) / K


def model_bytes_per_cell_step(K: int, T: int, side: int, C: int) -> float:
    """Algorithmic bytes per cell per step at reference-pass granularity:
    B = 135 + 8P + 8T + 128K + 10/C with P = ceil(log2(side))."""
    P = (side - 1).bit_length()
    return 135 + 8 * P + 8 * T + 128 * K + 10.0 / C


def load_traffic():
    """HBM bytes per launch of the dominant kernel from the committed counter
    profile (rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE runs, scaled by the
    counters' reading of a copy kernel of known size); counters cannot be read
    from inside this process."""
    p = PROFILES / "hydraulic_traffic.json"
    if p.exists():
        try:
            return json.loads(p.read_text()).get("bytes_per_launch"), "profiles/hydraulic_traffic.json"
        except Exception:
            return None, None
    return None, None


def load_step_bytes():
    """Measured HBM bytes of one whole step (all kernels of the step, from the
    committed counter profile of the same workload), or None."""
    p = PROFILES / "step_4096_hbm_bytes.json"
    if p.exists():
        try:
            return json.loads(p.read_text()).get("bytes_per_step"), "profiles/step_4096_hbm_bytes.json"
        except Exception:
            return None, None
    return None, None


def cpu_model() -> str:
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def run_batch_mode(args) -> dict:
    """Mixed-size worlds (sides log-uniform in [256, 2048], multiples of 64,
    plates U{4..40}): the same worlds stepped as one batch and one at a time."""
    from master_thesis_amd import configs
    from master_thesis_amd.sim import Batch, World
    specs = configs.batch_mix(args.batch_worlds, args.seed)
    cells = sum(c.width * c.height for c, _ in specs)
    worlds = [World(c, k).init() for c, k in specs]
    warm = max(args.warmup, 1)   # the first batched step captures the erosion graph
    with Batch(worlds) as b:
        b.step(warm)
        t_warm = b.stage_times()
        t0 = time.perf_counter()
        for a in range(0, args.batch_steps, 25):
            b.step(min(25, args.batch_steps - a))
            print(f"batched: {min(a + 25, args.batch_steps)}/{args.batch_steps} steps", file=sys.stderr, flush=True)
        tb = time.perf_counter() - t0
        t_all = b.stage_times()
    phases = {k: (t_all[k] - t_warm[k]) / args.batch_steps for k in ("ms_pre", "ms_erosion", "ms_post")}
    for w in worlds:
        w.close()
    worlds = [World(c, k).init() for c, k in specs]
    for w in worlds:
        w.step(warm)
    t0 = time.perf_counter()
    for i, w in enumerate(worlds):
        w.step(args.batch_steps)
        if (i + 1) % 32 == 0:
            print(f"one at a time: {i + 1}/{len(worlds)} worlds stepped", file=sys.stderr, flush=True)
    tp = time.perf_counter() - t0
    for w in worlds:
        w.close()
    n = len(specs)
    return {
        "workload": f"{n} worlds, sides log-uniform [256,2048] (multiples of 64), plates U{{4..40}}, "
                    f"K=50 hydraulic, T=20 thermal, CCL every step; {args.batch_steps} timed steps each",
        "worlds": n, "cells": cells, "steps": args.batch_steps,
        "batched_s": tb, "plain_s": tp,
        "batched_cell_updates_per_s": cells * args.batch_steps / tb,
        "plain_cell_updates_per_s": cells * args.batch_steps / tp,
        "batched_worlds_per_s": n * args.batch_steps / tb / args.batch_steps,
        "plain_worlds_per_s": n * args.batch_steps / tp / args.batch_steps,
        "speedup": tp / tb,
        "batched_phases_ms_per_step": phases,
    }


def run_serve_mode(args) -> dict:
    """Poisson arrivals of world-generation jobs (side log-uniform in
    [256, 4096], plates U{4..48}, steps U[50, 1000]) served under an HBM
    budget; a bounded sample of the request stream."""
    import threading
    from master_thesis_amd import configs
    from master_thesis_amd.sim import Server
    rng = np.random.default_rng(args.seed + 1000)
    reqs = []
    t = 0.0
    for i in range(args.serve_requests):
        t += float(rng.exponential(1.0 / args.serve_rate))
        cfg, ks = configs.mixed_world(rng, 256, 4096, 4, 48, args.seed + 10000 + i)
        steps = int(rng.integers(50, 1001))
        reqs.append((t, cfg, ks, steps))
    budget = int(args.serve_budget_gb * (1 << 30))
    with Server(budget) as srv:
        ids = []
        t0 = time.perf_counter()

        def submitter():
            for (ta, cfg, ks, steps) in reqs:
                dt = ta - (time.perf_counter() - t0)
                if dt > 0:
                    time.sleep(dt)
                ids.append(srv.submit(cfg, ks, steps))

        th = threading.Thread(target=submitter)
        th.start()
        th.join()
        srv.drain(-1)
        rep = srv.report()
        res = [srv.result(i, 0) for i in ids]
    work = sum(r.width * r.height * r.steps for r in res)
    return {
        "workload": f"{len(reqs)} requests, Poisson lambda={args.serve_rate}/s, side log-uniform [256,4096] "
                    f"(multiples of 64), plates U{{4..48}}, steps U[50,1000], K=50, T=20, budget "
                    f"{args.serve_budget_gb:g} GB",
        "requests": len(reqs), "completed": int(rep.completed), "failed": int(rep.failed),
        "worlds_per_s": rep.worlds_per_s, "latency_p50_s": rep.latency_p50_s, "latency_p99_s": rep.latency_p99_s,
        "latency_mean_s": rep.latency_mean_s, "wall_s": rep.wall_s, "cell_updates_per_s": work / rep.wall_s,
        "bytes_peak_gb": rep.bytes_peak / (1 << 30), "batch_steps": int(rep.batch_steps),
        "offered_worlds_per_s": args.serve_rate,
        "offered_over_capacity": args.serve_rate / rep.worlds_per_s if rep.worlds_per_s > 0 else None,
        "arrival_span_s": reqs[-1][0] if reqs else 0.0,
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500,
                    help="timed steps (default: the 500 tectonic steps of the 4096^2 workload)")
    ap.add_argument("--warmup", type=int, default=0)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--side", type=int, default=4096)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-cpu-512", action="store_true", help="skip the CPU run of the 512^2 workload")
    ap.add_argument("--tile-rows", type=int, default=0, help="erosion tile rows (0: autotuned)")
    ap.add_argument("--uplift", choices=["jfa", "separable"], default="jfa",
                    help="uplift spread: the jump flood (the reference's live path) or the bounded-radius "
                         "separable pass (radius = blur radius)")
    ap.add_argument("--mode", choices=["plain", "batch", "serve", "precision", "all"], default="all",
                    help="plain: the 4096^2 step (the metric); batch / serve: the batched and served forms; "
                         "all: the plain line with batched and served samples added")
    ap.add_argument("--batch-worlds", type=int, default=256)
    ap.add_argument("--batch-steps", type=int, default=10)
    ap.add_argument("--serve-requests", type=int, default=100)
    ap.add_argument("--serve-rate", type=float, default=20.0)
    ap.add_argument("--serve-budget-gb", type=float, default=200.0)
    ap.add_argument("--precision-side", type=int, default=8192)
    ap.add_argument("--precision-plates", type=int, default=64)
    ap.add_argument("--precision-steps", type=int, default=300)
    ap.add_argument("--precision-hydraulic", type=int, default=50, help="hydraulic sub-iterations per step")
    ap.add_argument("--precision-thermal", type=int, default=20, help="thermal passes per step")
    ap.add_argument("--no-hbm-probe", action="store_true",
                    help="skip the measured HBM copy/read/fill ceilings (2 GiB buffers)")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="run one plain device copy of this size first (calibrates HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported (worlds are independent; "
              "a multi-GPU run would be N copies of this measurement)", file=sys.stderr)
        return 2

    from master_thesis_amd import _abi, configs
    from master_thesis_amd.sim import World, STATE_FIELDS

    if args.mode == "precision":
        from master_thesis_amd.sim import precision_study
        pc, pk = configs.synthetic(args.precision_side, args.precision_side, args.precision_plates, (0.5, 3.0),
                                   args.seed, hydraulic=args.precision_hydraulic,
                                   thermal=args.precision_thermal, ccl_every=1, uplift_range=5)
        rep = precision_study(pc, pk, args.precision_steps)
        line = {"metric": "erosion error vs an fp64 twin (fp32 and compensated fp32), mass drift",
                "value": rep["height"]["rms_comp"], "unit": "height RMS error (compensated fp32)",
                "n_gpus": 1, "steps": args.precision_steps, "warmup": 0, "higher_is_better": False,
                "scaling": "weak", "vs_baseline": None, "dtype": "f32+compensated / f64",
                "data": "synthetic (seeded uniform heights, Gaussian sigma 8)",
                "config": {"workload": f"{args.precision_side}x{args.precision_side}, {args.precision_plates} plates, "
                                       f"{args.precision_steps} erosion-only steps (K={args.precision_hydraulic}, "
                                       f"T={args.precision_thermal})"}}
        line.update(rep)
        print(json.dumps(line))
        return 0

    if args.mode in ("batch", "serve"):
        out = run_batch_mode(args) if args.mode == "batch" else run_serve_mode(args)
        line = {
            "metric": "batched worlds/s (mixed sizes, same worlds one at a time beside)" if args.mode == "batch"
            else "served worlds/s with p50/p99 completion latency",
            "value": out["batched_worlds_per_s"] if args.mode == "batch" else out["worlds_per_s"],
            "unit": "worlds/s", "n_gpus": 1,
            "steps": args.batch_steps if args.mode == "batch" else None, "warmup": args.warmup,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
            "data": "synthetic (seeded worlds)", "config": {"workload": out["workload"]},
        }
        line.update({k: v for k, v in out.items() if k != "workload"})
        print(json.dumps(line))
        return 0

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
    cfg.uplift_mode = _abi.TT_UPLIFT_SEPARABLE if args.uplift == "separable" else _abi.TT_UPLIFT_JFA
    N = cfg.width * cfg.height
    world = World(cfg, ks).init()
    init_state = None
    if not args.no_cpu_baseline:
        init_state = {f: world.download(f) for f in STATE_FIELDS}
    world.step(args.warmup)
    world.set_timing(True)
    # the K timed steps in up to 5 equal segments (one timed region; the
    # segment times give the spread)
    nseg = max(1, min(5, args.steps))
    bounds = [round(i * args.steps / nseg) for i in range(nseg + 1)]
    seg_ms = []
    t0 = time.perf_counter()
    for a, b in zip(bounds[:-1], bounds[1:]):
        ts = time.perf_counter()
        world.step(b - a)  # returns after the device work is complete
        if b > a:
            seg_ms.append(1000.0 * (time.perf_counter() - ts) / (b - a))
    t1 = time.perf_counter()
    times = world.stage_times()
    stats = world.stats()
    wall = t1 - t0
    value = N * args.steps / wall

    K = cfg.hydraulic_iterations
    launches = max(times.hydraulic_launches, 1.0)
    avg_ms = times.ms_hydraulic / launches
    bpl = N * hydraulic_bytes_per_launch_avg(K)
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
                        f"{cfg.iterations_for_ccl} step, blur radius {ks.upliftRange}"
                        + (", separable uplift" if args.uplift == "separable" else ""),
            "uplift": args.uplift,
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
            "traffic": None,
            "bytes_per_launch": bpl,
            "avg_launch_ms": avg_ms,
        },
        "ms_per_step_segments": seg_ms,
        "ms_per_step_median_segment": float(np.median(seg_ms)) if seg_ms else None,
        "stages_ms_per_step": {
            "collide": times.ms_collide / args.steps, "uplift": times.ms_uplift / args.steps,
            "thermal": times.ms_thermal / args.steps, "hydraulic": times.ms_hydraulic / args.steps,
            "transport": times.ms_transport / args.steps, "post": times.ms_post / args.steps,
            "split": times.ms_split / args.steps, "events_total": times.ms_total / args.steps,
        },
        "step_model": {
            "reference_pass_bytes_per_cell_step": model_b,
            "note": "byte model of the reference's separate passes (B = 135 + 8P + 8T + 128K + 10/C); the fused "
                    "kernels move less, so it is not a measure of HBM use",
        },
        "final_stats": {"plates": stats.number_of_plates, "labels": stats.num_labels,
                        "iteration": stats.iteration, "tile_rows": stats.tile_rows},
    }
    world.close()
    traffic, tsrc = load_traffic()
    result["roofline"]["traffic"] = traffic
    result["roofline"]["traffic_source"] = tsrc
    sb, ssrc = load_step_bytes()
    if sb:
        step_s = wall / args.steps
        result["step_hbm"] = {"bytes_per_step": sb, "source": ssrc, "achieved_gbps": sb / step_s / 1e9,
                              "frac_of_peak": sb / step_s / 1e9 / HBM_PEAK_GBS}
    if not args.no_hbm_probe:
        # measured ceilings beside the published peak: a float4 copy has the
        # same 50/50 read/write mix as the hydraulic kernel
        from master_thesis_amd.sim import probe_hbm
        probe = probe_hbm(2 << 30, 5)
        result["roofline"]["measured_ceiling"] = probe
        result["roofline"]["frac_of_copy"] = achieved / probe["copy_gbps"] if probe["copy_gbps"] > 0 else None
        # the same bytes and streams as the kernel, without its stencil and gather
        result["roofline"]["frac_of_streams"] = (achieved / probe["streams_gbps"]
                                                 if probe["streams_gbps"] > 0 else None)

    if init_state is not None:
        # the CPU restatement timed on this host: 3 steps of the same world from
        # the same initial state (the metric's workload), and the 512^2 workload
        # in full (100 steps), median of 3 runs
        from oracle.oracle_py import OracleWorld
        threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
        ow = OracleWorld(cfg, ks)
        for f, a in init_state.items():
            ow.set(f, a)
        ow.set_iteration(0)
        csteps = 3
        c0 = time.perf_counter()
        ow.step(csteps)
        c1 = time.perf_counter()
        ow.close()
        result["cpu_baseline"] = {
            "value": N * csteps / (c1 - c0),
            "unit": "cell-updates/s",
            "cores": threads,
            "kind": "port",
            "cpu_model": cpu_model(),
            "sample": f"{csteps} steps of the same {cfg.width}x{cfg.height} world from the same initial state "
                      f"(CPU restatement, OpenMP, {threads} threads)",
        }
        if not args.no_cpu_512:
            c512, k512 = configs.world_512(seed=args.seed)
            n512, s512 = c512.width * c512.height, configs.STEPS["world_512"]
            runs = []
            for _ in range(3):
                o5 = OracleWorld(c512, k512)
                o5.init()
                r0 = time.perf_counter()
                o5.step(s512)
                runs.append(n512 * s512 / (time.perf_counter() - r0))
                o5.close()
            result["cpu_world_512"] = {
                "value": float(np.median(runs)), "runs": runs, "unit": "cell-updates/s", "cores": threads,
                "kind": "port", "cpu_model": cpu_model(),
                "sample": f"512x512, 8 plates, {s512} steps (K=50, T=20, CCL every 10), median of 3 runs",
            }
    if args.mode == "all":
        for key, fn in (("batched", run_batch_mode), ("serving", run_serve_mode)):
            try:
                result[key] = fn(args)
            except Exception as e:  # the plain measurement stands on its own
                result[key] = {"error": f"{type(e).__name__}: {e}"}
    print(json.dumps(result))
    return 0


if __name__ == "__main__":
    sys.exit(main())
