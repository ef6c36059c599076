"""Shared helpers of the test suite (state transfer and field comparison)."""
from __future__ import annotations

import numpy as np

from master_thesis_amd import _abi

STATE_FIELDS = (
    _abi.FIELD_HEIGHT, _abi.FIELD_PLATE_IDS, _abi.FIELD_WATER, _abi.FIELD_FLUX, _abi.FIELD_VELOCITY,
    _abi.FIELD_SEDIMENT, _abi.FIELD_UPLIFT, _abi.FIELD_PLATE_DATA, _abi.FIELD_DIVERGENCE_BITMAP,
    _abi.FIELD_COLLISION_TYPES,
)
STEP_OUTPUTS = (
    _abi.FIELD_COLLISIONS, _abi.FIELD_SORTED_KEYS, _abi.FIELD_ACCRETION, _abi.FIELD_PROPAGATED,
    _abi.FIELD_DISTANCE, _abi.FIELD_LABELS,
)


def get(world, field):
    return world.download(field) if hasattr(world, "download") else world.get(field)


def put(world, field, arr):
    return world.upload(field, arr) if hasattr(world, "upload") else world.set(field, arr)


def transfer_state(src, dst):
    for f in STATE_FIELDS:
        put(dst, f, get(src, f))
    it = src.iteration if hasattr(src, "iteration") else None
    if it is not None:
        dst.set_iteration(it)


def diff_fields(a, b, fields):
    """Names and byte-mismatch counts of the fields that differ."""
    out = []
    for f in fields:
        x, y = get(a, f), get(b, f)
        if x.tobytes() != y.tobytes():
            xb = np.frombuffer(x.tobytes(), np.uint8)
            yb = np.frombuffer(y.tobytes(), np.uint8)
            out.append((_abi.FIELD_NAMES[f], int(np.count_nonzero(xb != yb))))
    return out


def ulp_diff(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    ai = a.astype(np.float32).view(np.int32).astype(np.int64)
    bi = b.astype(np.float32).view(np.int32).astype(np.int64)
    ai = np.where(ai < 0, np.int64(-2**31) - ai, ai)
    bi = np.where(bi < 0, np.int64(-2**31) - bi, bi)
    return np.abs(ai - bi)


def banded_continental_heights(rng, a, b, H=64, W=64):
    """Continental crust at or above the propagation threshold (13000) on
    columns 31-a[y]..31 of plate 0 and 32..32+b[y] of plate 1, with one cell of
    each band exactly at the threshold; continental crust just below it
    (12990..12999) elsewhere."""
    h = rng.uniform(12990, 12999, (H, W)).astype(np.float32)
    for y in range(H):
        h[y, 31 - a[y]:32] = rng.uniform(13500, 14500, a[y] + 1)
        h[y, 32:33 + b[y]] = rng.uniform(13500, 14500, b[y] + 1)
        h[y, 31 - rng.integers(0, a[y] + 1)] = 13000.0
        h[y, 32 + rng.integers(0, b[y] + 1)] = 13000.0
    return h
