"""Independent pure-Python / numpy-float32 restatement of the cudaNoise formulas
the reference vendors (src/cuda_compute/cuda_noise.cuh:44-71, 112-115,
158-160, 209-223, 228-319, 470-477), written from the published algorithm and
not from the library's C code.  Used to pin the shared noise header, and to
regenerate tests/golden/noise_vectors.json:

    python tests/golden/noise_restatement.py
"""
from __future__ import annotations

import json
from pathlib import Path

import numpy as np

F = np.float32
M32 = 0xFFFFFFFF


def jenkins_hash(seed: int) -> int:
    a = seed & M32
    a = ((a + 0x7ED55D16) + (a << 12)) & M32
    a = ((a ^ 0xC761C23C) ^ (a >> 19)) & M32
    a = ((a + 0x165667B1) + (a << 5)) & M32
    a = ((a + 0xD3A2646C) ^ (a << 9)) & M32
    a = ((a + 0xFD7046C5) + (a << 3)) & M32
    a = ((a ^ 0xB55A4F09) ^ (a >> 16)) & M32
    return a


def random_float(seed: int) -> np.float32:
    return F(F(jenkins_hash(seed)) / F(4294967296.0))


def f2u_sat(v: np.float32) -> int:
    if not (v > 0):
        return 0
    if v >= F(4294967296.0):
        return M32
    return int(v)


def random_grid(x: int, y: int, z: int, seed: int) -> np.float32:
    s = F(F(F(F(F(x) * F(1723.0)) + F(F(y) * F(93241.0))) + F(F(z) * F(149812.0))) + F(3824.0))
    s = F(s + F(seed))
    return F(F(random_float(f2u_sat(s)) * F(2.0)) - F(1.0))


def discrete_noise(px: float, py: float, pz: float, scale: float, seed: int) -> np.float32:
    return random_grid(int(F(F(px) * F(scale))), int(F(F(py) * F(scale))), int(F(F(pz) * F(scale))), seed)


GRAD = [(1, 1, 0), (-1, 1, 0), (1, -1, 0), (-1, -1, 0), (1, 0, 1), (-1, 0, 1), (1, 0, -1), (-1, 0, -1),
        (0, 1, 1), (0, -1, 1), (0, 1, -1), (0, -1, -1)]


def _corner(x, y, z, gi):
    t = F(F(F(F(0.6) - F(x * x)) - F(y * y)) - F(z * z))
    if t < 0:
        return F(0.0)
    t = F(t * t)
    g = GRAD[gi]
    dot = F(F(F(F(g[0]) * x) + F(F(g[1]) * y)) + F(F(g[2]) * z))
    return F(F(t * t) * dot)


def simplex(px: float, py: float, pz: float, scale: float, seed: int) -> np.float32:
    xin, yin, zin = F(F(px) * F(scale)), F(F(py) * F(scale)), F(F(pz) * F(scale))
    F3, G3 = F(F(1.0) / F(3.0)), F(F(1.0) / F(6.0))
    s = F(F(F(xin + yin) + zin) * F3)
    i, j, k = int(np.floor(F(xin + s))), int(np.floor(F(yin + s))), int(np.floor(F(zin + s)))
    t = F(F(i + j + k) * G3)
    x0, y0, z0 = F(xin - F(F(i) - t)), F(yin - F(F(j) - t)), F(zin - F(F(k) - t))
    if x0 >= y0:
        if y0 >= z0:
            o1, o2 = (1, 0, 0), (1, 1, 0)
        elif x0 >= z0:
            o1, o2 = (1, 0, 0), (1, 0, 1)
        else:
            o1, o2 = (0, 0, 1), (1, 0, 1)
    else:
        if y0 < z0:
            o1, o2 = (0, 0, 1), (0, 1, 1)
        elif x0 < z0:
            o1, o2 = (0, 1, 0), (0, 1, 1)
        else:
            o1, o2 = (0, 1, 0), (1, 1, 0)
    c1 = [F(F(v - F(o)) + G3) for v, o in zip((x0, y0, z0), o1)]
    c2 = [F(F(v - F(o)) + F(F(2.0) * G3)) for v, o in zip((x0, y0, z0), o2)]
    c3 = [F(F(v - F(1.0)) + F(F(3.0) * G3)) for v in (x0, y0, z0)]

    def perm(a, b, c):
        return jenkins_hash((seed + a * 607495 + b * 359609 + c * 654846) & M32) % 12

    n0 = _corner(x0, y0, z0, perm(i, j, k))
    n1 = _corner(*c1, perm(i + o1[0], j + o1[1], k + o1[2]))
    n2 = _corner(*c2, perm(i + o2[0], j + o2[1], k + o2[2]))
    n3 = _corner(*c3, perm(i + 1, j + 1, k + 1))
    return F(F(32.0) * F(F(F(n0 + n1) + n2) + n3))


def make_vectors():
    rng = np.random.default_rng(20240607)
    hs = [int(v) for v in rng.integers(0, 2**32, 64, dtype=np.uint64)]
    pts = [(float(F(a)), float(F(b)), float(F(c))) for a, b, c in rng.uniform(0, 4096, (64, 3))]
    pts += [(float(x), float(y), float(z)) for x, y, z in [(0, 0, 0), (1, 2, 3), (4095, 4095, 499), (100.5, 7.25, 3)]]
    return {
        "hash": [[h, jenkins_hash(h)] for h in hs],
        "simplex_scale": 0.01, "simplex_seed": 1234,
        "simplex": [[p[0], p[1], p[2], float(simplex(p[0], p[1], p[2], 0.01, 1234))] for p in pts],
        "discrete_seed": 1001,
        "discrete": [[p[0], p[1], p[2], float(discrete_noise(p[0], p[1], p[2], 1.0, 1001))] for p in pts],
    }


if __name__ == "__main__":
    out = Path(__file__).with_name("noise_vectors.json")
    out.write_text(json.dumps(make_vectors(), indent=1))
    print("wrote", out)
