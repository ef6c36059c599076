"""Snapshot files, diagnostics JSON and the collision-class views."""
from __future__ import annotations

import sys
import zlib

import numpy as np
import pytest

from master_thesis_amd import _abi, configs
from tests.helpers import STATE_FIELDS, diff_fields

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def sim(tectonic_lib):
    from master_thesis_amd import sim as s
    return s


def reference_world():
    cfg, ks = configs.reference_defaults()
    cfg.width, cfg.height, cfg.num_plates, cfg.seed, cfg.iterations_for_ccl = 192, 128, 10, 23, 1
    cfg.num_voronoi_gens = 0
    return cfg, ks


def test_snapshot_resume_is_bit_identical(sim, tmp_path):
    cfg, ks = reference_world()
    say("[snapshot]")
    a = sim.World(cfg, ks).init()
    a.step(3)
    path = tmp_path / "w.snap"
    a.save(str(path))
    b = sim.World.load(str(path))
    assert b.iteration == 3 and b.cfg.width == 192
    a.step(3)
    b.step(3)
    assert not diff_fields(a, b, STATE_FIELDS + (_abi.FIELD_LABELS,))
    raw = bytearray(path.read_bytes())
    raw[-10] ^= 0xFF
    path.write_bytes(bytes(raw))
    with pytest.raises(RuntimeError):
        sim.World.load(str(path))
    a.close()
    b.close()


def test_diagnostics_json(sim):
    cfg, ks = reference_world()
    say("[diagnostics]")
    with sim.World(cfg, ks).init() as w:
        w.step(2)
        d = w.diagnostics()
        s = w.stats()
        ids = w.download(_abi.FIELD_PLATE_IDS)
        pd = w.download(_abi.FIELD_PLATE_DATA)
        h = w.download(_abi.FIELD_HEIGHT)
        word = w.download(_abi.FIELD_COLLISIONS)
    assert d["iteration"] == 2 and d["stats"]["number_of_plates"] == s.number_of_plates
    assert len(d["plates"]) == int((pd["size"] > 0).sum())
    assert sum(p["size"] for p in d["plates"]) == 192 * 128
    for p in d["plates"]:
        assert p["size"] == int(pd["size"][p["id"]])
    assert abs(d["fields"]["height_max"] - float(h.max())) <= 1e-3 * abs(float(h.max()))
    cc = d["collision_classes"]
    assert sum(cc.values()) == 192 * 128
    assert cc["divergence"] == int((word == 0).sum())


def test_class_view_export(sim, tmp_path):
    cfg, ks = reference_world()
    with sim.World(cfg, ks).init() as w:
        w.step(1)
        for what in (_abi.TT_EXPORT_COLLISION_CLASSES, _abi.TT_EXPORT_PROPAGATED_CLASSES):
            p = tmp_path / f"v{what}.png"
            w.export_png(what, str(p))
            data = p.read_bytes()
            assert data[:8] == b"\x89PNG\r\n\x1a\n"
            assert int.from_bytes(data[16:20], "big") == 192 and int.from_bytes(data[20:24], "big") == 128


def test_hbm_probe(sim):
    """The ceiling probe reports positive bandwidths of a plausible size."""
    print("[hbm probe] 64 MiB", flush=True)
    r = sim.probe_hbm(64 << 20, 2)
    assert r["bytes"] == 64 << 20 and r["reps"] == 2
    for k in ("copy_gbps", "read_gbps", "write_gbps", "streams_gbps"):
        assert 10.0 < r[k] < 100000.0, (k, r[k])
