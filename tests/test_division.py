"""The division the thermal passes use on the device, a / b from a stored
y = RN(1 / b) with two corrections (tt_div_by_rcp), restated in C on the host
and compared with IEEE division over every divisor significand and 50 M random
pairs with exponents in [-40, 39].  The same check of the bare product a y
fails (the comparison detects a wrong quotient).  The device code itself is
compared with numpy's division in test_division_by_stored_reciprocal_is_ieee."""
from __future__ import annotations

import shutil
import subprocess
from pathlib import Path

import pytest

HERE = Path(__file__).resolve().parent


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    exe = tmp_path_factory.mktemp("div") / "division_check"
    subprocess.run(["gcc", "-O2", "-mfma", "-ffp-contract=off", "-o", str(exe), str(HERE / "division_check.c"),
                    "-lm"], check=True)
    return exe


def _run(exe, mode):
    out = subprocess.run([str(exe), mode], check=True, capture_output=True, text=True).stdout.split()
    return int(out[0]), int(out[1])


def test_division_from_the_reciprocal_is_ieee(checker):
    n, bad = _run(checker, "full")
    assert n > 3e8 and bad == 0


def test_the_bare_product_is_not(checker):
    n, bad = _run(checker, "product")
    assert bad > n // 10
