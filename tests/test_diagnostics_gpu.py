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


# --------------------------------------------------------------------------
# the exports against the reference formulas
# --------------------------------------------------------------------------
def _rgb_image(path):
    from tests.test_parity_gpu import _png_rows
    w, h, depth, ctype, raw = _png_rows(path)
    assert depth == 8 and ctype == 2
    stride = 1 + 3 * w
    rows = [raw[y * stride:(y + 1) * stride] for y in range(h)]
    assert all(r[0] == 0 for r in rows)          # filter type 0 on every row
    return np.frombuffer(b"".join(r[1:] for r in rows), np.uint8).reshape(h, w, 3)


def reference_palette():
    """convertGrayscaleToRgbKernel (texture_save.cu:10-78) for every id, in
    float32 as the kernel computes it: hue 0.125 (id % 8) + id / 8 / 256,
    saturation 0.5 + 0.5 sin(0.1 id), value 0.5 + 0.5 cos(0.1 id), the HSV
    sextant chosen by double comparisons, channels lround((c + m) 255); id
    255 is the background (230, 235, 240).  sinf / cosf are taken as the
    correctly rounded float of the double result."""
    f = np.float32
    out = np.zeros((256, 3), np.uint8)
    for val in range(256):
        if val == 255:
            out[val] = (230, 235, 240)
            continue
        h = f(f(0.125) * f(val % 8)) + f(f(f(val) / f(8.0)) / f(256.0))
        a = f(f(val) * f(0.1))
        s = f(f(0.5) + f(f(0.5) * f(np.sin(np.float64(a)))))
        v = f(f(0.5) + f(f(0.5) * f(np.cos(np.float64(a)))))
        c = f(v * s)
        x = f(c * f(f(1.0) - f(abs(f(np.fmod(f(h * f(6.0)), f(2.0))) - f(1.0)))))
        m = f(v - c)
        hd = float(h)
        if 0.0 <= hd < 1.0 / 6.0:
            r, g, b = c, x, f(0.0)
        elif 1.0 / 6.0 <= hd < 2.0 / 6.0:
            r, g, b = x, c, f(0.0)
        elif 2.0 / 6.0 <= hd < 3.0 / 6.0:
            r, g, b = f(0.0), c, x
        elif 3.0 / 6.0 <= hd < 4.0 / 6.0:
            r, g, b = f(0.0), x, c
        elif 4.0 / 6.0 <= hd < 5.0 / 6.0:
            r, g, b = x, f(0.0), c
        else:
            r, g, b = c, f(0.0), x
        out[val] = [int(np.floor(float(f(f(ch + m) * f(255.0))) + 0.5)) for ch in (r, g, b)]
    return out


def test_plate_palette_follows_the_reference_formula(sim, tmp_path):
    """Every id 0..255 exported as a plate map and decoded: each pixel has
    the reference's HSV colour for its id."""
    cfg, ks = configs.synthetic(256, 64, 2, (0.5, 1.0), 1, hydraulic=1, thermal=1)
    cfg.init_mode = _abi.TT_INIT_EMPTY
    ids = np.tile(np.arange(256, dtype=np.uint8), 64)
    with sim.World(cfg, ks).init() as w:
        w.upload(_abi.FIELD_PLATE_IDS, ids)
        p = tmp_path / "palette.png"
        w.export_png(_abi.TT_EXPORT_PLATES_RGB, str(p))
    img = _rgb_image(p).reshape(-1, 3)
    pal = reference_palette()
    assert np.array_equal(img, pal[ids])
    assert len({tuple(c) for c in pal[:255]}) > 200


# class colours of the exported views (the legend of this library's PNGs)
CLASS_COLOR = np.array([[20, 20, 24], [70, 130, 220], [220, 80, 60], [240, 200, 80], [250, 150, 40],
                        [255, 255, 255]], np.uint8)


def reference_classes(packed, types):
    """createVisualizationCollisionTexture / ...PropagatedCollisionTexture
    (plate_tectonics_kernel.cu:3250-3313): 1 divergence (no plate), 0 a single
    plate, 2 subduction (type bit 0 clear), 3 / 4 / 5 continental with 2 / 3
    / 4 plates (the packed word's second, third and fourth bytes)."""
    packed = packed.astype(np.uint64)
    cls = np.zeros(packed.shape, np.uint8)
    multi = (packed >> 8) != 0
    cont = (types & 1) != 0
    cls[packed == 0] = 1
    cls[multi & ~cont] = 2
    cls[multi & cont & ((packed >> 16) == 0)] = 3
    cls[multi & cont & ((packed >> 16) != 0) & ((packed >> 24) == 0)] = 4
    cls[multi & cont & ((packed >> 24) != 0)] = 5
    return cls


def _classes_of(img):
    flat = img.reshape(-1, 3)
    out = np.full(len(flat), 255, np.uint8)
    for k, c in enumerate(CLASS_COLOR):
        out[(flat == c).all(axis=1)] = k
    assert (out != 255).all()
    return out


def _continental_world():
    cfg, ks = configs.reference_defaults()
    cfg.width, cfg.height, cfg.num_plates, cfg.seed, cfg.iterations_for_ccl = 256, 192, 12, 5, 1
    cfg.num_voronoi_gens = 0
    return cfg, ks


def test_class_views_follow_the_reference_classification(sim, tmp_path):
    cfg, ks = _continental_world()
    with sim.World(cfg, ks).init() as w:
        w.step(3)
        word = w.download(_abi.FIELD_COLLISIONS)
        up = w.download(_abi.FIELD_UPLIFT)
        pr = w.download(_abi.FIELD_PROPAGATED)
        pc, pp = tmp_path / "c.png", tmp_path / "p.png"
        w.export_png(_abi.TT_EXPORT_COLLISION_CLASSES, str(pc))
        w.export_png(_abi.TT_EXPORT_PROPAGATED_CLASSES, str(pp))
        d = w.diagnostics()
    exp_c = reference_classes(word, up["types"])
    exp_p = reference_classes(pr["inv"], pr["types"])
    assert np.array_equal(_classes_of(_rgb_image(pc)), exp_c)
    assert np.array_equal(_classes_of(_rgb_image(pp)), exp_p)
    # every class the views can show occurs, and the JSON counts agree
    assert {1, 2, 3}.issubset(set(np.unique(exp_c)))
    names = ["none", "divergence", "subduction", "continental_2", "continental_3", "continental_4"]
    assert [d["collision_classes"][n] for n in names] == [int((exp_c == k).sum()) for k in range(6)]
    assert [d["propagated_classes"][n] for n in names] == [int((exp_p == k).sum()) for k in range(6)]


def test_plate_records_follow_copy_plate_data_gui(sim):
    """copyPlateDataGuiKernel (plate_tectonics_kernel.cu:1220-1281): for each
    other plate j with a typed pair entry, a continental pair lists j; a
    subduction lists +j or -j by the pair's polarity bit and by which of the
    two is the smaller id; per-plate fields are the plate table's."""
    cfg, ks = _continental_world()
    with sim.World(cfg, ks).init() as w:
        w.step(3)
        d = w.diagnostics()
        types = w.download(_abi.FIELD_COLLISION_TYPES)
        pd = w.download(_abi.FIELD_PLATE_DATA)

    def tri(i, j):  # upperTriangleIndexUnchecked (cuda_helper.cuh:98-101), matrix size 255
        return i * (2 * 255 - i - 1) // 2 + (j - i - 1)

    seen_cont = seen_sub = 0
    for rec in d["plates"]:
        p = rec["id"]
        cont, sub = [], []
        for i in range(255):
            if i == p:
                continue
            lo, hi = min(p, i), max(p, i)
            t = int(types[tri(lo, hi)])
            if not t & 1:
                continue
            if (t >> 1) & 1:
                cont.append(hi if lo == p else lo)
            elif (t >> 3) & 1:
                sub.append(hi if lo == p else -lo)
            else:
                sub.append(-hi if lo == p else lo)
        assert rec["continental"] == cont, p
        assert rec["subductions"] == sub, p
        seen_cont += len(cont)
        seen_sub += len(sub)
        assert rec["size"] == int(pd[p]["size"]) and rec["perimeter"] == int(pd[p]["perimeter"])
        assert rec["has_moved"] == int(pd[p]["hasMoved"]) and rec["merge_wait_time"] == int(pd[p]["mergeWaitTime"])
        for key, field in (("mass", "mass"), ("velocity", "velocity"), ("velocity_smoothed", "velocitySmoothed"),
                           ("break_score", "breakScore"), ("circularity", "circularity")):
            assert rec[key] == pytest.approx(float(pd[p][field]), rel=1e-6, abs=1e-9), (p, key)
    assert seen_cont > 0 and seen_sub > 0
