"""Build libtectonic.so in-tree: hand-written gfx950 HIP kernels + host code.

    python -m master_thesis_amd.build            # build (incremental)
    python -m master_thesis_amd.build --force    # rebuild everything

Every HIP source is compiled for gfx950 with strict IEEE float semantics
(-ffp-contract=off, no fast-math) so that device results match the host
arithmetic bit for bit, and with packed FP32 instructions disabled.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
ROOT = PKG.parent
OBJ = PKG / "_build"
LIB = PKG / "libtectonic.so"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("TT_OFFLOAD_ARCH", "gfx950")
HIP_SOURCES = ["k_collide.hip", "k_uplift.hip", "k_erosion.hip", "k_plates.hip", "k_misc.hip", "k_batch.hip",
               "world.hip", "batch.hip", "precision.hip"]
HOST_SOURCES = ["host.cpp", "server.cpp"]
HEADERS = ["tt_math.h", "tt_noise.h", "tt_device.h", "tt_stage.h", "tt_kernels.h", "tt_world.h", "tt_erosion.h",
           "tt_batch.h"]

HIP_FLAGS = [
    "-O3", f"--offload-arch={ARCH}", "-fPIC", "-std=c++17",
    "-ffp-contract=off", "-fno-fast-math", "-fno-gpu-rdc",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-result",
    f"-I{ROOT / 'include'}",
]
HOST_FLAGS = ["-O2", "-fPIC", "-std=c++17", "-pthread", "-ffp-contract=off", "-fno-fast-math", "-Wall",
              f"-I{ROOT / 'include'}"]


def _stale(src: Path, obj: Path, deps) -> bool:
    if not obj.exists():
        return True
    t = obj.stat().st_mtime
    return any(p.stat().st_mtime > t for p in [src, *deps] if p.exists())


def _run(cmd) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(map(str, cmd))}\n{r.stdout}\n{r.stderr}")
    # hipcc's host pass does not know the packed-fp32 feature; that warning is harmless
    noise = [ln for ln in r.stderr.splitlines() if ln.strip() and "packed-fp32-ops" not in ln
             and "argument unused during compilation" not in ln]
    if noise:
        sys.stderr.write("\n".join(noise) + "\n")


def build(force: bool = False, jobs: int | None = None) -> Path:
    if shutil.which(HIPCC) is None and not Path(HIPCC).exists():
        raise RuntimeError(f"hipcc not found at {HIPCC}")
    OBJ.mkdir(exist_ok=True)
    deps = [CSRC / h for h in HEADERS] + [ROOT / "include" / "tectonic.h"]
    tasks = []
    objs = []
    for s in HIP_SOURCES:
        src, obj = CSRC / s, OBJ / (s + ".o")
        objs.append(obj)
        if force or _stale(src, obj, deps):
            tasks.append([HIPCC, *HIP_FLAGS, "-c", str(src), "-o", str(obj)])
    for s in HOST_SOURCES:
        src, obj = CSRC / s, OBJ / (s + ".o")
        objs.append(obj)
        if force or _stale(src, obj, deps):
            tasks.append(["g++", *HOST_FLAGS, "-c", str(src), "-o", str(obj)])
    jobs = jobs or min(8, os.cpu_count() or 4)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        for f in [ex.submit(_run, t) for t in tasks]:
            f.result()
    if force or tasks or not LIB.exists():
        tmp = LIB.with_suffix(".so.tmp")
        _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *map(str, objs), "-o", str(tmp), "-lz", "-ldl",
              "-lpthread"])
        os.replace(tmp, LIB)
    return LIB


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--force", action="store_true")
    ap.add_argument("-j", type=int, default=None)
    a = ap.parse_args()
    print(build(force=a.force, jobs=a.j))


if __name__ == "__main__":
    main()
