"""Two exact rewrites of the device rain noise, checked exhaustively on the
host from the product header (tests/rain_terms_check.cpp):

* the row's lattice term (float)y * 149812 built from integer operations on
  the scalar unit (tt_lattice_z_term) equals the float product for every row
  of every supported world (y < 16384, checked up to 28668);
* the noise value tt_random_float(u) * 2 - 1 written as one fused
  multiply-add by 2^-31 equals the three-step form for every float the hash
  conversion can produce.

A truncating variant of the row term must fail, so the comparison has teeth.
The device side (the kernel's rain with these forms) is compared with the
oracle, bit for bit, by the GPU parity tests of the water field."""
from __future__ import annotations

import shutil
import subprocess
from pathlib import Path

import pytest

HERE = Path(__file__).resolve().parent


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    exe = tmp_path_factory.mktemp("rain") / "rain_terms_check"
    subprocess.run(["g++", "-O2", "-mfma", "-ffp-contract=off", "-std=c++17", "-o", str(exe),
                    str(HERE / "rain_terms_check.cpp")], check=True)
    return exe


def _run(exe, mode):
    out = subprocess.run([str(exe), mode], check=True, capture_output=True, text=True).stdout.split()
    return int(out[0]), int(out[1])


def test_row_term_from_integers_is_the_float_product(checker):
    n, bad = _run(checker, "zterm")
    assert n == 28669 and bad == 0


def test_truncated_row_term_is_not(checker):
    n, bad = _run(checker, "trunc")
    assert bad > 1000


def test_noise_scale_as_one_fma_is_exact(checker):
    n, bad = _run(checker, "fma")
    assert n == (0x4F800000 - 0x3F800000 + 1) + 1 and bad == 0
