"""tt_math_eval_device_static, the allocation-free variant of
tt_math_eval_device: it stages through a static device buffer and must give
the same bits as the allocating entry point, call after call (the second call
smaller than the first, so stale contents of the shared buffer would show)."""
import ctypes as C
import sys

import numpy as np
import pytest

from master_thesis_amd import _abi

pytestmark = pytest.mark.gpu
MAX = 262144


def run(L, entry, fn, x, n):
    out = np.full(n, np.nan, np.float32)
    st = getattr(L, entry)(fn, x.ctypes.data, out.ctypes.data, n)
    assert st == _abi.TT_OK, (entry, st, L.tt_last_error())
    return out


def inputs(fn, n, rng):
    per = {_abi.TT_MATH_SIMPLEX: 3, _abi.TT_MATH_DISCRETE: 3, _abi.TT_MATH_FMODF: 2, _abi.TT_MATH_ATAN2: 2}.get(fn, 1)
    return np.ascontiguousarray(rng.uniform(-300.0, 300.0, n * per).astype(np.float32))


@pytest.mark.parametrize("fn", [_abi.TT_MATH_EXPF, _abi.TT_MATH_SINF, _abi.TT_MATH_SIMPLEX, _abi.TT_MATH_ATAN2])
def test_static_entry_matches_the_allocating_one(tectonic_lib, fn):
    L = tectonic_lib
    rng = np.random.default_rng(fn)
    for n in (MAX, 1000):
        print(f"[no-alloc] fn={fn} n={n}", flush=True)
        sys.stdout.flush()
        x = inputs(fn, n, rng)
        a = run(L, "tt_math_eval_device", fn, x, n)
        b = run(L, "tt_math_eval_device_static", fn, x, n)
        assert a.tobytes() == b.tobytes()


def test_static_entry_rejects_sizes_beyond_its_buffer(tectonic_lib):
    L = tectonic_lib
    x = np.zeros(MAX + 1, np.float32)
    out = np.zeros(MAX + 1, np.float32)
    assert L.tt_math_eval_device_static(_abi.TT_MATH_EXPF, x.ctypes.data, out.ctypes.data, MAX + 1) == _abi.TT_ERR_INVALID
    assert b"exceeds" in L.tt_last_error()
    assert L.tt_math_eval_device_static(_abi.TT_MATH_EXPF, x.ctypes.data, out.ctypes.data, 0) == _abi.TT_OK
