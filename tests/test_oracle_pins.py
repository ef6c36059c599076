"""Pins of oracle functions that the step-level tests leave open, each against
what the mathematics or a brute-force search fixes:

* computeCurl (plate_tectonics_kernel.cu:2788-2822): centred differences of
  the committed noise texture, rotated, at the scaled texel of each cell;
* the plate centres (accumulatePlateAngularCoords / calculatePlateCenters,
  :2154-2226): the circular mean of every tenth cell in x and y, for plates
  that straddle the seams of the torus;
* applyAccretion (:747-770): a marked cell changes owner exactly when its
  height is continental;
* the split (SplitPlateV2, plate_tectonic_sim.cu:933-1040, with
  propagateEffortToBoundary :2838-2881, BacktrackPath :2883-2983,
  floodFillPlate :2985-3024, finalizePlateSplit :3026-3045): the effort field
  by Dijkstra (generalised to the float32 path sums the relaxation evaluates),
  the two cut-line walks, the seed and the 4-connected flood, all restated
  here from the reference kernels and compared with the oracle's ids."""
import heapq

import numpy as np
import pytest

from master_thesis_amd import _abi
from tests.test_oracle import _continental_front, empty_world, set_plates

F32 = np.float32


@pytest.fixture(scope="module")
def O(oracle_lib):
    return oracle_lib


# --------------------------------------------------------------------------
# computeCurl
# --------------------------------------------------------------------------
@pytest.mark.parametrize("W,H", [(256, 256), (128, 64), (512, 192), (320, 64)])
def test_asthenosphere_is_the_rotated_gradient_of_the_noise(O, W, H):
    """The texture value is R/255; cell (x, y) reads texel
    (int(x*256/W), int(y*256/H)); dir.y = (p(x+1) - p(x-1)) / 2 and
    dir.x = -(p(y+1) - p(y-1)) / 2, wrapped on the 256x256 texture."""
    noise = np.frombuffer(_abi.load_noise(), np.uint8).reshape(256, 256)
    p = noise.astype(F32) / F32(255.0)
    dx = (np.roll(p, -1, axis=1) - np.roll(p, 1, axis=1)) * F32(0.5)     # d/dx at each texel
    dy = (np.roll(p, -1, axis=0) - np.roll(p, 1, axis=0)) * F32(0.5)     # d/dy at each texel
    sx = (np.arange(W) * (256.0 / W)).astype(np.int64)
    sy = (np.arange(H) * (256.0 / H)).astype(np.int64)
    exp = np.stack([-dy[sy][:, sx], dx[sy][:, sx]], axis=-1)
    w = empty_world(O, W, H, 0)
    a = w.get(_abi.FIELD_ASTHENOSPHERE).reshape(H, W, 2)
    w.close()
    assert np.array_equal(a, exp)
    assert (np.abs(a) > 0).mean() > 0.5


# --------------------------------------------------------------------------
# plate centres
# --------------------------------------------------------------------------
def circular_centre(cells, n):
    """Circular mean of integer coordinates on a ring of n cells, in [0, n)."""
    ang = 2 * np.pi * np.asarray(cells, np.float64) / n
    c = np.arctan2(np.sin(ang).sum(), np.cos(ang).sum()) * n / (2 * np.pi)
    return c + n if c < 0 else c


def test_plate_centres_are_circular_means_across_the_seams(O):
    """Plates 1-3 straddle the x seam, the y seam and both; plate 4 sits
    inside.  Only cells with x % 10 == 0 and y % 10 == 0 are sampled; the
    centre is the circular mean of their coordinates on each axis."""
    W, H = 160, 120
    ids = np.zeros((H, W), np.uint8)
    ys, xs = np.mgrid[0:H, 0:W]
    ids[(ys >= 30) & (ys < 61) & ((xs >= W - 25) | (xs < 15))] = 1
    ids[((ys >= H - 33) | (ys < 12)) & (xs >= 40) & (xs < 77)] = 2
    ids[((ys >= H - 21) | (ys < 26)) & ((xs >= W - 41) | (xs < 8))] = 3
    ids[(ys >= 70) & (ys < 95) & (xs >= 90) & (xs < 131)] = 4
    w = empty_world(O, W, H, _abi.TT_STAGE_MERGE)
    w.set(_abi.FIELD_PLATE_IDS, ids.ravel())
    set_plates(w, **{f"p{p}": dict(used=1, size=int((ids == p).sum()), mergeWaitTime=5) for p in range(5)})
    w.step(1)
    pd = w.get(_abi.FIELD_PLATE_DATA)
    w.close()
    samp = (xs % 10 == 0) & (ys % 10 == 0)
    for p in range(1, 5):
        sel = samp & (ids == p)
        cx = circular_centre(xs[sel], W)
        cy = circular_centre(ys[sel], H)
        assert pd[p]["geometricCenter"][0] == pytest.approx(cx, abs=1e-4), p
        assert pd[p]["geometricCenter"][1] == pytest.approx(cy, abs=1e-4), p
    # the straddling plates' centres lie on the far side of the seam from
    # where an arithmetic mean would put them
    assert pd[1]["geometricCenter"][0] > W - 10
    assert pd[2]["geometricCenter"][1] > H - 15


# --------------------------------------------------------------------------
# applyAccretion
# --------------------------------------------------------------------------
def test_accretion_moves_only_continental_cells(O):
    """The subduction front of the friction test, whose accretion rows mark
    cells of the overriding plate's source: one step with the kinematics
    (plate movement + applyAccretion) and one without must differ exactly in
    the marked cells whose final height is at least the continental crust
    threshold, which take the id 255 - mark."""
    rng = np.random.default_rng(14)
    h = rng.uniform(12000, 13000, (64, 64)).astype(F32)
    acc_rows = np.arange(64) % 4 == 0
    h[:, :32] = np.where(acc_rows[:, None], rng.uniform(10500, 11500, (64, 32)),
                         rng.uniform(1000, 5000, (64, 32))).astype(F32)
    # the cells that move onto some marked cells are just below the threshold
    h[::8, 29] = F32(9990.0)
    runs = {}
    for tag, mask in (("without", _abi.TT_STAGE_COLLIDE), ("with", _abi.TT_STAGE_COLLIDE | _abi.TT_STAGE_KINEMATICS)):
        w = _continental_front(O, h, mask, frictionCoefficientContinental=0.7, frictionCoefficientSubduction=0.3,
                               frictionCoefficientAccretion=0.45, environmentalDragCoefficient=0.0)
        runs[tag] = (w.get(_abi.FIELD_PLATE_IDS), w.get(_abi.FIELD_ACCRETION), w.get(_abi.FIELD_HEIGHT))
        thr = w.ks.continentalCrustThreshold
        w.close()
    ids0, acc0, h0 = runs["without"]
    ids1, acc1, h1 = runs["with"]
    assert np.array_equal(acc0, acc1) and np.array_equal(h0, h1)
    marked = acc1 != 0
    moves = marked & (h1 >= thr)
    assert moves.sum() > 10 and (marked & (h1 < thr)).sum() > 2
    exp = np.where(moves, (255 - acc1.astype(np.int64)).astype(np.uint8), ids0)
    assert np.array_equal(ids1, exp)
    assert (ids1 != ids0).sum() > 5


# --------------------------------------------------------------------------
# split: effort field, cut line, seed and flood
# --------------------------------------------------------------------------
OFFS = [(-1, 0), (1, 0), (0, -1), (0, 1), (-1, -1), (-1, 1), (1, -1), (1, 1)]
SQRT2 = F32(1.41421356237)
INF = F32(np.inf)


def effort_by_dijkstra(ids, h, plate):
    """Fixed point of propagateEffortToBoundary: cells outside the plate hold
    0; a plate cell's effort is the least, over its 8 neighbours n, of
    effort(n) + h(cell) * (1 or sqrt2), in float32.  Float addition of a
    non-negative term is monotone, so Dijkstra from the outside cells finds
    the same least fixed point."""
    H, W = ids.shape
    E = np.where(ids == plate, INF, F32(0.0)).astype(F32)
    done = np.zeros((H, W), bool)
    pq = [(0.0, int(y), int(x)) for y, x in zip(*np.nonzero(ids != plate))]
    heapq.heapify(pq)
    while pq:
        e, y, x = heapq.heappop(pq)
        if done[y, x] or F32(e) != E[y, x]:
            continue
        done[y, x] = True
        for i, (ox, oy) in enumerate(OFFS):
            # the cell c = (x - ox, y - oy) relaxes from its neighbour (x, y) at offset (ox, oy)
            cx, cy = (x - ox) % W, (y - oy) % H
            if ids[cy, cx] != plate or done[cy, cx]:
                continue
            mul = SQRT2 if i >= 4 else F32(1.0)
            t = F32(E[y, x] + F32(h[cy, cx] * mul))
            if t < E[cy, cx]:
                E[cy, cx] = t
                heapq.heappush(pq, (float(t), cy, cx))
    return E


def split_by_restatement(ids, h, plate, new_id, point):
    """BacktrackPath, floodFillPlate and finalizePlateSplit on host arrays;
    coordinates stay unwrapped as in the reference (the texture wraps them)."""
    H, W = ids.shape
    ids = ids.copy()
    E = effort_by_dijkstra(ids, h, plate)
    at = lambda p: (p[1] % H, p[0] % W)
    old = ids[at(point)]
    cur = point
    while E[at(cur)] > 0:
        assert ids[at(cur)] != 255, "first walk revisits a cell"
        ids[at(cur)] = 255
        best, nxt = INF, None
        for ox, oy in OFFS:
            v = E[at((cur[0] + ox, cur[1] + oy))]
            if v < best:
                best, nxt = v, (cur[0] + ox, cur[1] + oy)
        cur = nxt
    fb = np.array([cur[0] - point[0], cur[1] - point[1]], F32)
    o = fb * F32(-1.0)
    m = max(F32(np.sqrt(F32(o[0] * o[0] + o[1] * o[1]))), F32(1e-20))
    od = np.array([o[0] / m, o[1] / m], F32)
    perp = np.array([-od[1], od[0]], F32)
    cur = point
    while E[at(cur)] > 0:
        ids[at(cur)] = 255
        best, nxt = INF, None
        for ox, oy in OFFS:
            nb = (cur[0] + ox, cur[1] + oy)
            if ids[at(nb)] == 255:
                continue
            tn = np.array([nb[0] - point[0], nb[1] - point[1]], F32)
            if F32(F32(tn[0] * od[0]) + F32(tn[1] * od[1])) < 0:
                continue
            om = F32(np.sqrt(F32(ox * ox + oy * oy)))
            d = F32(F32(F32(ox) / om) * od[0]) + F32(F32(F32(oy) / om) * od[1])
            score = F32(E[at(nb)] + F32(F32(-d) * F32(50.0)))
            if score < best:
                best, nxt = score, nb
        assert best != INF, "second walk reaches a dead end"
        cur = nxt
    s = point
    for i in range(1, 101):
        if ids[at(s)] == old:
            break
        k = (i + 1) // 2 if i % 2 == 1 else -(i // 2)
        fx = F32(F32(point[0]) + F32(perp[0] * F32(k)))
        fy = F32(F32(point[1]) + F32(perp[1] * F32(k)))
        s = (int(np.trunc(F32(fx + F32(0.5)))), int(np.trunc(F32(fy + F32(0.5)))))
    seeded = ids[at(s)] == old
    if seeded:
        ids[at(s)] = new_id
        stack = [at(s)]
        while stack:
            y, x = stack.pop()
            for ox, oy in OFFS[:4]:
                q = ((y + oy) % H, (x + ox) % W)
                if ids[q] == old:
                    ids[q] = new_id
                    stack.append(q)
    ids[ids == 255] = new_id
    return ids, seeded


def _split_world(O, ids, h, plate, centre):
    H, W = ids.shape
    w = empty_world(O, W, H, _abi.TT_STAGE_SPLIT)
    w.set(_abi.FIELD_PLATE_IDS, ids.ravel())
    w.set(_abi.FIELD_HEIGHT, h.ravel())
    plates = {f"p{p}": dict(used=1, size=int((ids == p).sum()), breakScore=0.5) for p in np.unique(ids)}
    plates[f"p{plate}"].update(breakScore=1.5, geometricCenter=centre, velocity=1.25, direction=(0.6, 0.8))
    set_plates(w, **plates)
    w.step(1)
    out = w.get(_abi.FIELD_PLATE_IDS).reshape(H, W), w.stats(), w.get(_abi.FIELD_PLATE_DATA)
    w.close()
    return out


@pytest.mark.parametrize("case", ["ellipse", "seam", "blob"])
def test_split_cut_line_matches_the_restatement(O, case):
    """Plate 1 on background plate 0 with random positive costs: the oracle's
    split equals the Dijkstra effort field + the two walks + the seed + the
    flood restated here, and partitions plate 1 into itself and plate 2 (the
    first unused id) along the cut line."""
    rng = np.random.default_rng({"ellipse": 1, "seam": 2, "blob": 3}[case])
    W, H = 96, 64
    ys, xs = np.mgrid[0:H, 0:W]
    if case == "ellipse":
        inside = ((xs - 50) / 30.0) ** 2 + ((ys - 30) / 12.0) ** 2 <= 1
        centre = (50.3, 29.8)
    elif case == "seam":
        dx = np.minimum(np.abs(xs - 2), W - np.abs(xs - 2))
        inside = (dx / 26.0) ** 2 + ((ys - 40) / 14.0) ** 2 <= 1
        centre = (1.6, 40.2)
    else:
        inside = ((xs - 48) / 22.0) ** 2 + ((ys - 32) / 20.0) ** 2 <= 1
        inside &= ~(((xs - 40) / 9.0) ** 2 + ((ys - 26) / 7.0) ** 2 <= 1)   # a hole of plate 0
        centre = (61.7, 30.2)
    ids = np.where(inside, 1, 0).astype(np.uint8)
    h = rng.uniform(1.0, 10.0, (H, W)).astype(F32)
    point = (int(F32(centre[0]) + F32(0.5)), int(F32(centre[1]) + F32(0.5)))
    exp, seeded = split_by_restatement(ids, h, 1, 2, point)
    got, st, pd = _split_world(O, ids, h, 1, centre)
    assert seeded
    assert st.split_plate == 1 and st.split_new_plate == 2
    assert np.array_equal(got, exp)
    n1, n2 = (got == 1).sum(), (got == 2).sum()
    assert n1 > 50 and n2 > 50 and n1 + n2 == inside.sum()
    assert pd[2]["velocity"] == F32(1.25) and tuple(pd[2]["direction"]) == (F32(0.6), F32(0.8))
