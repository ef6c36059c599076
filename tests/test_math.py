"""Pins of the shared deterministic math (tt_math.h) and noise (tt_noise.h).

exp / log10 / sin / cos / atan2 are checked against numpy's float64 libm
rounded to float32 (error bound: at most 1 ulp, and equal in all but a tiny
fraction); fmod against numpy's exact fmod; the noise against an independent
restatement of the published cudaNoise formulas (tests/golden).
"""
import json
from pathlib import Path

import numpy as np
import pytest

from master_thesis_amd import _abi
from tests.helpers import ulp_diff

GOLD = Path(__file__).parent / "golden"


@pytest.fixture(scope="module")
def om(oracle_lib):
    return oracle_lib


def _check_close(got, ref, max_ulp=1, exact_frac=0.9999):
    d = ulp_diff(got, ref)
    assert d.max() <= max_ulp, f"max ulp {d.max()}"
    assert (d == 0).mean() >= exact_frac


def test_expf(om):
    x = np.random.default_rng(1).uniform(-103, 88.5, 300000).astype(np.float32)
    _check_close(om.math_eval(_abi.TT_MATH_EXPF, x), np.exp(x.astype(np.float64)).astype(np.float32))
    sp = np.array([0.0, -0.0, 1.0, -1.0, 88.0, -103.0, -200.0, 100.0, np.inf, -np.inf], np.float32)
    got = om.math_eval(_abi.TT_MATH_EXPF, sp)
    assert got[0] == 1.0 and got[1] == 1.0
    assert got[6] == 0.0 and np.isinf(got[7]) and np.isinf(got[8]) and got[9] == 0.0
    assert np.isnan(om.math_eval(_abi.TT_MATH_EXPF, np.array([np.nan], np.float32)))[0]


def test_log10f(om):
    x = np.exp(np.random.default_rng(2).uniform(-85, 85, 300000)).astype(np.float32)
    _check_close(om.math_eval(_abi.TT_MATH_LOG10F, x), np.log10(x.astype(np.float64)).astype(np.float32))
    powers = np.array([10.0 ** k for k in range(0, 11)], np.float32)
    assert np.array_equal(om.math_eval(_abi.TT_MATH_LOG10F, powers), np.arange(11, dtype=np.float32))
    sp = om.math_eval(_abi.TT_MATH_LOG10F, np.array([0.0, -7.0, 1e-45, np.inf], np.float32))
    assert np.isneginf(sp[0]) and np.isnan(sp[1]) and abs(sp[2] + 44.85) < 0.01 and np.isposinf(sp[3])


@pytest.mark.parametrize("fn,ref", [(_abi.TT_MATH_SINF, np.sin), (_abi.TT_MATH_COSF, np.cos)])
def test_sin_cos(om, fn, ref):
    rng = np.random.default_rng(3)
    x = np.concatenate([rng.uniform(-4, 4, 100000), rng.uniform(-3e4, 3e4, 200000)]).astype(np.float32)
    _check_close(om.math_eval(fn, x), ref(x.astype(np.float64)).astype(np.float32))


def test_atan2(om):
    rng = np.random.default_rng(4)
    yx = rng.uniform(-10, 10, (200000, 2)).astype(np.float32)
    got = om.math_eval(_abi.TT_MATH_ATAN2, yx.ravel())
    ref = np.arctan2(yx[:, 0].astype(np.float64), yx[:, 1].astype(np.float64)).astype(np.float32)
    _check_close(got, ref)
    quad = np.array([[0, 1], [1, 0], [0, -1], [-1, 0], [1, 1], [-1, -1]], np.float32)
    assert np.allclose(om.math_eval(_abi.TT_MATH_ATAN2, quad.ravel()),
                       [0, np.pi / 2, np.pi, -np.pi / 2, np.pi / 4, -3 * np.pi / 4], atol=1e-7)


def test_fmodf_exact(om):
    rng = np.random.default_rng(5)
    a = rng.uniform(-1e5, 1e5, 200000).astype(np.float32)
    b = rng.uniform(1e-3, 5e3, 200000).astype(np.float32)
    got = om.math_eval(_abi.TT_MATH_FMODF, np.stack([a, b], 1).ravel())
    assert np.array_equal(got, np.fmod(a, b))


def test_hash_known_answers(om):
    vec = json.loads((GOLD / "noise_vectors.json").read_text())
    inp = np.array([h for h, _ in vec["hash"]], np.uint32)
    exp = np.array([v for _, v in vec["hash"]], np.uint32)
    assert np.array_equal(om.hash_eval(inp), exp)


@pytest.mark.parametrize("fn,key", [(_abi.TT_MATH_SIMPLEX, "simplex"), (_abi.TT_MATH_DISCRETE, "discrete")])
def test_noise_known_answers(om, fn, key):
    vec = json.loads((GOLD / "noise_vectors.json").read_text())
    pts = np.array([r[:3] for r in vec[key]], np.float32)
    exp = np.array([r[3] for r in vec[key]], np.float32)
    got = om.math_eval(fn, pts.ravel())
    assert np.array_equal(got.view(np.uint32), exp.view(np.uint32))


def test_noise_restatement_random(om):
    from tests.golden import noise_restatement as nr
    rng = np.random.default_rng(6)
    pts = rng.uniform(-300, 300, (400, 3)).astype(np.float32)
    got = om.math_eval(_abi.TT_MATH_SIMPLEX, pts.ravel())
    exp = np.array([nr.simplex(*p, 0.01, 1234) for p in pts], np.float32)
    assert np.array_equal(got.view(np.uint32), exp.view(np.uint32))
    # simplex stays inside [-1, 1]
    many = om.math_eval(_abi.TT_MATH_SIMPLEX, rng.uniform(0, 5000, 300000).astype(np.float32))
    assert np.abs(many).max() <= 1.0


def test_sine_quadrant_count_is_zero_below_the_wavefront_bound():
    """The hydraulic kernel's wavefront sine takes the bare polynomial when every
    lane's argument is below 0.78f, which needs rint(x * 2/pi) == 0 there (the
    reduction then leaves x unchanged).  The double product and rint are both
    monotone in x, so the largest float below the bound decides it; the bound
    must keep a margin (pi/4 itself would not)."""
    two_over_pi = 6.36619772367581382433e-01
    x_max = float(np.nextafter(np.float32(0.78), np.float32(0.0)))
    assert np.rint(x_max * two_over_pi) == 0.0
    assert x_max * two_over_pi < 0.497
    assert np.rint(float(np.float32(np.pi / 2)) * two_over_pi) == 1.0
