"""The reference-shaped boundary on the device: resetSim with any number of
Voronoi generations (plate_tectonic_sim.h:119), snapshots of compensated
worlds, and the canonical outcome of a split whose plate has a negative cost."""
from __future__ import annotations

import sys

import numpy as np
import pytest

from master_thesis_amd import _abi, configs
from tests.helpers import STATE_FIELDS, STEP_OUTPUTS, diff_fields

pytestmark = pytest.mark.gpu
F32 = np.float32


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def sim(tectonic_lib):
    from master_thesis_amd import sim as s
    return s


@pytest.fixture(scope="module")
def O(oracle_lib):
    return oracle_lib


def test_reset_with_five_generations_matches_the_oracle(sim, O):
    """resetSim(seed, plates, {30, 70, 150, 260, 400}) on a world that has
    already stepped: the state equals the oracle's reference initialisation
    with the same five generations, and two steps stay bit-exact."""
    cfg, ks = configs.reference_defaults()
    cfg.width, cfg.height, cfg.num_plates, cfg.seed, cfg.iterations_for_ccl = 256, 128, 5, 3, 1
    cfg.num_voronoi_gens = 1
    cfg.voronoi_gens[0] = 40
    gens = [30, 70, 150, 260, 400]
    say("[reset-5-gens] init, 2 steps, reset")
    g = sim.World(cfg, ks).init()
    g.step(2)
    g.reset(77, 9, gens)
    ocfg = cfg.copy()
    ocfg.seed, ocfg.num_plates, ocfg.num_voronoi_gens = 77, 9, 0
    o = O.OracleWorld(ocfg, ks)
    o.set_gens(gens)
    o.init()
    try:
        assert g.iteration == 0
        bad = diff_fields(g, o, STATE_FIELDS + (_abi.FIELD_LABELS,))
        assert not bad, bad
        for s in range(2):
            say(f"[reset-5-gens] step {s}")
            g.step(1)
            o.step(1)
            bad = diff_fields(g, o, STATE_FIELDS + STEP_OUTPUTS)
            assert not bad, f"step {s}: {bad}"
    finally:
        g.close()
        o.close()


def test_facade_initialize_equals_reset(sim):
    """PlateTectonicSim.initialize(P, gens) and resetSim on an existing world
    give the same state as a world created with the same generations."""
    cfg, ks = configs.reference_defaults()
    cfg.width, cfg.height, cfg.seed = 192, 128, 8
    say("[facade-mirror]")
    p = sim.PlateTectonicSim(192, 128, 8, config=cfg, settings=ks)
    p.initialize(6, [20, 50])
    p.executeIteration()
    p.resetSim(8, 6, [20, 50])
    cfg2 = cfg.copy()
    cfg2.num_plates, cfg2.num_voronoi_gens = 6, 2
    cfg2.voronoi_gens[0], cfg2.voronoi_gens[1] = 20, 50
    with sim.World(cfg2, ks).init() as w:
        assert not diff_fields(p.world, w, STATE_FIELDS)
    p.world.close()


def test_compensated_snapshot_resumes_bit_identically(sim, tmp_path):
    """A compensated world saved after 3 steps and loaded again carries its
    Kahan terms: 3 more steps equal 3 more steps of the original."""
    cfg, ks = configs.synthetic(192, 128, 6, (0.5, 2.0), 5, hydraulic=8, thermal=4, ccl_every=1)
    cfg.precision = _abi.TT_PRECISION_COMPENSATED
    say("[snapshot-compensated]")
    a = sim.World(cfg, ks).init()
    a.step(3)
    path = tmp_path / "c.snap"
    a.save(str(path))
    assert path.read_bytes()[:8] == b"TTSNAP2\n"
    b = sim.World.load(str(path))
    try:
        assert b.cfg.precision == _abi.TT_PRECISION_COMPENSATED
        a.step(3)
        b.step(3)
        bad = diff_fields(a, b, STATE_FIELDS)
        assert not bad, bad
    finally:
        a.close()
        b.close()


def _split_world(sim, O, negative):
    """Plate 1 is a 64x8 band (break score 1 + circularity 0.69 with the
    area target below its size) on plate 0; random positive costs, one of
    them negative when asked."""
    W, H = 128, 64
    cfg, ks = configs.synthetic(W, H, 2, (0.5, 1.0), 1, hydraulic=1, thermal=1, ccl_every=1)
    cfg.init_mode = _abi.TT_INIT_EMPTY
    cfg.stage_mask = _abi.TT_STAGE_MERGE | _abi.TT_STAGE_SPLIT
    ks.targetMaximumPlateArea = 400
    ids = np.zeros((H, W), np.uint8)
    ids[20:28, 30:94] = 1
    h = np.random.default_rng(4).uniform(1.0, 10.0, (H, W)).astype(F32)
    if negative:
        h[23, 70] = F32(-0.5)
    out = []
    for w in (sim.World(cfg, ks).init(), O.OracleWorld(cfg, ks)):
        if hasattr(w, "init") and not hasattr(w, "download"):
            w.init()
        put = w.upload if hasattr(w, "download") else w.set
        get = w.download if hasattr(w, "download") else w.get
        put(_abi.FIELD_PLATE_IDS, ids.ravel())
        put(_abi.FIELD_HEIGHT, h.ravel())
        pd = get(_abi.FIELD_PLATE_DATA)
        for p in (0, 1):
            pd[p]["used"], pd[p]["mergeWaitTime"] = 1, 5
        pd[1]["velocity"], pd[1]["direction"] = 0.75, (0.6, 0.8)
        put(_abi.FIELD_PLATE_DATA, pd)
        w.step(1)
        out.append((get(_abi.FIELD_PLATE_IDS), get(_abi.FIELD_PLATE_DATA), w.stats()))
        w.close()
    return ids, out


@pytest.mark.parametrize("negative", [False, True])
def test_split_with_a_negative_cost_is_abandoned(sim, O, negative):
    say(f"[split negative={negative}]")
    ids, ((gi, gp, gs), (oi, op, os_)) = _split_world(sim, O, negative)
    assert gi.tobytes() == oi.tobytes()
    assert gp.tobytes() == op.tobytes()
    assert (gs.split_plate, gs.split_new_plate) == (os_.split_plate, os_.split_new_plate) == (1, 2)
    if negative:
        assert np.array_equal(gi, ids.ravel())
        assert gp[2]["used"] == 1 and gp[2]["size"] == 0
    else:
        assert (gi == 2).sum() > 50 and (gi == 1).sum() > 50
