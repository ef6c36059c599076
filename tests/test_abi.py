"""The C-ABI library: loads without a GPU, exports every declared entry point,
struct layouts and defaults match the reference parameter surface."""
import ctypes as C
import re
import struct
from pathlib import Path

import numpy as np
import pytest

from master_thesis_amd import _abi

ROOT = Path(__file__).resolve().parents[1]


def header_functions():
    text = (ROOT / "include" / "tectonic.h").read_text()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*\*?\s*(tt_[a-z0-9_]+)\s*\(", text, re.M)))


def test_exports_every_declared_symbol(tectonic_lib):
    names = header_functions()
    assert len(names) >= 20
    for n in names:
        assert hasattr(tectonic_lib, n), n


def test_precision_struct_sizes():
    assert C.sizeof(_abi.PrecisionReport) == 16 + 3 * 5 * 8 + 10 * 8 + 4 * 8


def test_hbm_probe_struct_size():
    assert C.sizeof(_abi.HbmProbe) == 48


def test_server_struct_sizes():
    assert C.sizeof(_abi.JobResult) == 8 + 4 + 4 + 8 * 4 + 8 + 4 * 6
    assert C.sizeof(_abi.ServerReport) == 11 * 8 + 7 * 8 + 4 * 8


def test_struct_sizes():
    assert C.sizeof(_abi.KernelSettings) == 200          # 50 x 4-byte fields (kernel_settings.cuh:4-81)
    assert C.sizeof(_abi.PlateData) == 84                # sizeof(PlateData) (plate_data.h:5-38)
    assert C.sizeof(_abi.SimConfig) == 23 * 4
    assert C.sizeof(_abi.StepStats) == 18 * 4 + 8
    assert C.sizeof(_abi.StageTimes) == 16 * 4
    assert _abi.UPLIFT_DTYPE.itemsize == 8 and _abi.PROPAGATED_DTYPE.itemsize == 12


def test_kernel_settings_defaults(tectonic_lib):
    ks = _abi.KernelSettings()
    assert tectonic_lib.tt_settings_default(C.byref(ks)) == 0
    py = _abi.KernelSettings.defaults()
    assert bytes(ks) == bytes(py)
    assert ks.hydrationRainfall == np.float32(0.101) and ks.maxUpliftClamp == 20000.0


def test_config_defaults(tectonic_lib):
    c = _abi.SimConfig()
    assert tectonic_lib.tt_config_default(C.byref(c)) == 0
    assert bytes(c) == bytes(_abi.SimConfig.defaults())


def test_json_parameter_surface(tectonic_lib):
    text = """{ "use_renderer": true, "heightmap_dimensions": {"x": 512, "y": 256},
                "seed": 7, "numStartingPlates": 12, "numVoronoiSeeds": [100, 300, 900],
                "iterationsForCCL": 3, "hydraulicIterations": 50, "thermalIterations": 20,
                "kernelSettings": {"hydrationRainfall": 0.01, "minPlateSize": 25, "upliftRange": 5} }"""
    c, k = _abi.SimConfig.defaults(), _abi.KernelSettings.defaults()
    assert tectonic_lib.tt_config_parse_json(text.encode(), C.byref(c), C.byref(k)) == 0
    assert (c.width, c.height, c.seed, c.num_plates) == (512, 256, 7, 12)
    assert c.num_voronoi_gens == 3 and list(c.voronoi_gens)[:3] == [100, 300, 900]
    assert (c.iterations_for_ccl, c.hydraulic_iterations, c.thermal_iterations) == (3, 50, 20)
    assert k.hydrationRainfall == np.float32(0.01) and k.minPlateSize == 25 and k.upliftRange == 5
    assert tectonic_lib.tt_config_parse_json(b"{ not json", C.byref(c), C.byref(k)) != 0


def test_reference_config_json(tectonic_lib):
    # the reference's own config.json layout (config.json:1-6)
    text = '{\n  "use_renderer" : true,\n  "heightmap_dimensions" : {\n    "x": 1024,\n    "y": 1024\n  }\n}\n'
    c, k = _abi.SimConfig.defaults(), _abi.KernelSettings.defaults()
    c.width = c.height = 64
    assert tectonic_lib.tt_config_parse_json(text.encode(), C.byref(c), C.byref(k)) == 0
    assert (c.width, c.height) == (1024, 1024)


def test_world_bytes_and_validation(tectonic_lib):
    from master_thesis_amd import configs
    cfg, ks = configs.world_4096()
    b = C.c_size_t()
    assert tectonic_lib.tt_world_bytes(C.byref(cfg), C.byref(b)) == 0
    assert 100 * 4096 * 4096 < b.value < 300 * 4096 * 4096
    bad = cfg.copy()
    bad.width = 100
    h = C.c_void_p()
    st = tectonic_lib.tt_world_create(C.byref(bad), C.byref(ks), C.byref(h))
    assert st == -1 and not h.value
    assert b"multiple of 64" in tectonic_lib.tt_last_error()


def test_status_strings(tectonic_lib):
    for s in range(0, -7, -1):
        assert tectonic_lib.tt_status_string(s)


def test_no_device_is_an_error_not_a_crash(tectonic_lib):
    n = C.c_int(-1)
    assert tectonic_lib.tt_device_count(C.byref(n)) == 0
    if n.value == 0:
        from master_thesis_amd import configs
        cfg, ks = configs.synthetic(64, 64, 2, (0.5, 1.0))
        h = C.c_void_p()
        assert tectonic_lib.tt_world_create(C.byref(cfg), C.byref(ks), C.byref(h)) == -2


def test_hbm_probe_rejects_bad_sizes(tectonic_lib):
    """Argument checks come before any device call."""
    r = _abi.HbmProbe()
    assert tectonic_lib.tt_probe_hbm(24, 1, C.byref(r)) == _abi.TT_ERR_INVALID
    assert tectonic_lib.tt_probe_hbm(16, 0, C.byref(r)) == _abi.TT_ERR_INVALID
    assert tectonic_lib.tt_probe_hbm((4 << 30) + 16, 1, C.byref(r)) == _abi.TT_ERR_INVALID
    assert b"multiple of 16" in tectonic_lib.tt_last_error()


def test_snapshot_load_rejects_bad_files(tectonic_lib, tmp_path):
    """tt_world_load checks the magic, the header and every field's CRC before
    any device work, so bad files fail without a GPU."""
    vp = C.c_void_p()
    p = tmp_path / "junk.snap"
    p.write_bytes(b"not a snapshot at all")
    assert tectonic_lib.tt_world_load(str(p).encode(), C.byref(vp)) == _abi.TT_ERR_INVALID
    assert b"snapshot" in tectonic_lib.tt_last_error()
    assert tectonic_lib.tt_world_load(str(tmp_path / "missing.snap").encode(), C.byref(vp)) == _abi.TT_ERR_INVALID


def test_snapshot_load_rejects_a_field_longer_than_the_file(tectonic_lib, tmp_path):
    """A field header that claims more bytes than the file still holds is
    rejected before anything is allocated for it."""
    class Header(C.Structure):
        _fields_ = [("cfg", _abi.SimConfig), ("ks", _abi.KernelSettings), ("iteration", C.c_uint32),
                    ("stats", _abi.StepStats), ("nfields", C.c_uint32)]
    hd = Header()
    hd.nfields = 1
    body = bytes(hd)
    field = struct.pack("<iQI", 0, 4 << 30, 0) + b"\0" * 64   # claims 4 GiB, holds 64 bytes
    p = tmp_path / "short.snap"
    p.write_bytes(b"TTSNAP2\n" + struct.pack("<I", len(body)) + body + field)
    vp = C.c_void_p()
    assert tectonic_lib.tt_world_load(str(p).encode(), C.byref(vp)) == _abi.TT_ERR_INVALID
    assert b"truncated field header" in tectonic_lib.tt_last_error()
