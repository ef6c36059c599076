"""Precision study: float and compensated-float erosion against an fp64 twin."""
from __future__ import annotations

import math
import sys

import pytest

from master_thesis_amd import _abi, configs
from tests.helpers import STATE_FIELDS, STEP_OUTPUTS, diff_fields

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sim(tectonic_lib):
    from master_thesis_amd import sim as s
    return s


def test_compensated_erosion_is_closer_to_fp64(sim):
    """Reference-scale heights (5000..15000 m): float updates of size ~1e-4
    are mostly lost against ulp(1e4) ~ 1e-3; the compensated arithmetic keeps
    them, so its height error against the fp64 twin is smaller."""
    cfg, ks = configs.reference_defaults()
    cfg.width, cfg.height, cfg.num_plates, cfg.seed = 256, 256, 8, 5
    cfg.hydraulic_iterations, cfg.thermal_iterations = 10, 1
    print("[precision] 256x256 reference heights", flush=True)
    r = sim.precision_study(cfg, ks, 8)
    print(r, flush=True)
    for f in ("height", "water", "sediment"):
        for k, v in r[f].items():
            assert math.isfinite(v), (f, k)
    assert r["height"]["rms_ref"] > 1000.0
    # the maximum is set by cells where a threshold test (talus, erosion vs
    # deposition) decides differently in float and double, for either variant;
    # the RMS measures the accumulated rounding
    assert r["height"]["rms_comp"] <= r["height"]["rms_fp32"]
    assert abs(r["drift_comp"] - r["drift_ref"]) <= abs(r["drift_fp32"] - r["drift_ref"]) + 1e-12


def test_precision_mode_runs_and_rejects_batches(sim):
    cfg, ks = configs.synthetic(128, 64, 5, (0.5, 3.0), 3, hydraulic=3, thermal=2)
    cfg.precision = _abi.TT_PRECISION_COMPENSATED
    with sim.World(cfg, ks).init() as w:
        w.step(2)
        assert w.iteration == 2
        with sim.Batch() as b:
            with pytest.raises(RuntimeError):
                b.add(w)


@pytest.fixture(scope="module")
def O(oracle_lib):
    return oracle_lib


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("mode", ["reference", "synthetic"])
def test_compensated_mode_matches_oracle(sim, O, mode):
    """Precision mode through full steps (collisions reset the height
    compensation, the sediment compensation travels with the sediment):
    every field bit for bit against the oracle's compensated mode."""
    if mode == "reference":
        cfg, ks = configs.reference_defaults()
        cfg.width, cfg.height, cfg.num_plates, cfg.seed = 192, 128, 6, 4
        cfg.hydraulic_iterations, cfg.thermal_iterations = 6, 2
    else:
        cfg, ks = configs.synthetic(128, 128, 5, (0.5, 3.0), 9, hydraulic=5, thermal=4, ccl_every=1)
        ks.thermalThresholdAngle = 1e-4
    cfg.precision = _abi.TT_PRECISION_COMPENSATED
    _say(f"[compensated {mode}] {cfg.width}x{cfg.height}")
    g = sim.World(cfg, ks).init()
    o = O.OracleWorld(cfg, ks)
    o.init()
    try:
        for s in range(3):
            _say(f"[compensated {mode}] step {s}")
            g.step(1)
            o.step(1)
            bad = diff_fields(g, o, tuple(STATE_FIELDS) + tuple(STEP_OUTPUTS))
            assert not bad, f"step {s}: {bad}"
            assert g.stats().mass_ledger == o.stats().mass_ledger
    finally:
        g.close()
        o.close()


def test_precision_study_matches_oracle(sim, O):
    """The device study against the oracle's: float and compensated worlds are
    bit-identical to the oracle's and the fp64 twins perform the same double
    operations, so the maxima agree exactly and the sums to reduction order."""
    cfg, ks = configs.synthetic(128, 64, 5, (0.5, 3.0), 3, hydraulic=6, thermal=2, ccl_every=1)
    ks.thermalThresholdAngle = 1e-4
    _say("[precision study] 128x64 K=6 T=2, 3 steps")
    rg = sim.precision_study(cfg, ks, 3)
    ro = O.precision_study(cfg, ks, 3)
    for f in ("height", "water", "sediment"):
        for k in ("max_abs_fp32", "max_abs_comp"):
            assert rg[f][k] == ro[f][k], (f, k, rg[f][k], ro[f][k])
        for k in ("rms_fp32", "rms_comp", "rms_ref"):
            assert rg[f][k] == pytest.approx(ro[f][k], rel=1e-9, abs=1e-300), (f, k)
    for k in ("mass_start", "mass_ref", "mass_fp32", "mass_comp"):
        assert rg[k] == pytest.approx(ro[k], rel=1e-12), k
    for k in ("drift_ref", "drift_fp32", "drift_comp"):
        assert rg[k] == pytest.approx(ro[k], rel=1e-9, abs=1e-15), k


def test_precision_study_at_8192(sim):
    """The precision workload's geometry (8192x8192, 64 plates, K = 50,
    T = 20) for two erosion-only steps: every statistic finite, both float
    variants still well correlated with the fp64 twin (errors far below the
    fields' own spread) and the same mass drift as the twin."""
    cfg, ks = configs.world_8192(seed=1)
    _say("[precision study] 8192x8192, 64 plates, 2 steps")
    r = sim.precision_study(cfg, ks, 2)
    _say(r)
    for f in ("height", "water", "sediment"):
        for k, v in r[f].items():
            assert math.isfinite(v), (f, k)
    h = r["height"]
    assert h["rms_ref"] > 0.0
    assert h["rms_fp32"] < 0.5 * h["rms_ref"] and h["rms_comp"] < 0.5 * h["rms_ref"]
    assert abs(r["drift_fp32"] - r["drift_ref"]) < 1e-3 and abs(r["drift_comp"] - r["drift_ref"]) < 1e-3
    assert r["ms_fp32"] > 0 and r["ms_comp"] > 0 and r["ms_ref"] > 0
