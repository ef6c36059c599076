"""Batched stepping (tt_batch_*): every member of a mixed batch ends with the
same bits as the same world stepped alone (and as the oracle), with the
erosion replayed from a captured hipGraph or launched directly."""
from __future__ import annotations

import sys

import pytest

from master_thesis_amd import _abi, configs
from tests.helpers import STATE_FIELDS, diff_fields

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def mixed_configs():
    out = []
    for (W, H, P, seed, K, T, rows) in [(128, 128, 6, 1, 6, 3, 0), (256, 64, 9, 2, 4, 2, 8), (64, 64, 3, 3, 5, 1, 16),
                                          (192, 96, 7, 4, 3, 4, 32), (320, 160, 12, 5, 6, 0, 0)]:
        cfg, ks = configs.synthetic(W, H, P, (0.5, 3.0), seed, hydraulic=K, thermal=T, ccl_every=1)
        ks.thermalThresholdAngle = 2e-3
        cfg.tile_rows = rows if rows and H % rows == 0 else 0
        out.append((cfg, ks))
    cfg, ks = configs.reference_defaults()
    cfg.width, cfg.height, cfg.num_plates, cfg.seed, cfg.iterations_for_ccl = 128, 128, 8, 9, 2
    out.append((cfg, ks))
    return out


@pytest.fixture(scope="module")
def sim(tectonic_lib):
    from master_thesis_amd import sim as s
    return s


@pytest.mark.parametrize("graph", [True, False])
def test_batch_members_match_solo_worlds(sim, graph):
    cfgs = mixed_configs()
    say(f"[batch graph={graph}] {len(cfgs)} worlds")
    members = [sim.World(c, k).init() for c, k in cfgs]
    solo = [sim.World(c, k).init() for c, k in cfgs]
    with sim.Batch(members, graph=graph) as b:
        assert len(b) == len(cfgs)
        for s in range(4):
            say(f"[batch graph={graph}] step {s}")
            b.step(1)
            for w in solo:
                w.step(1)
            for i, (m, w) in enumerate(zip(members, solo)):
                bad = diff_fields(m, w, STATE_FIELDS)
                assert not bad, f"member {i} step {s}: {bad}"
                assert m.iteration == w.iteration == s + 1
    for w in members + solo:
        w.close()


def test_batch_member_matches_oracle(sim, oracle_lib):
    cfg, ks = configs.synthetic(128, 64, 5, (0.5, 3.0), 7, hydraulic=4, thermal=2, ccl_every=1)
    other = configs.synthetic(256, 256, 9, (0.5, 3.0), 8, hydraulic=7, thermal=1, ccl_every=1)
    say("[batch oracle]")
    a = sim.World(cfg, ks).init()
    b2 = sim.World(*other).init()
    o = oracle_lib.OracleWorld(cfg, ks)
    o.init()
    with sim.Batch([a, b2]) as b:
        b.step(3)
    o.step(3)
    assert not diff_fields(a, o, STATE_FIELDS)
    a.close()
    b2.close()
    o.close()


def test_batch_membership_changes(sim):
    """Worlds join and leave between steps (continuous batching); a world that
    spent steps in batches of different composition matches its solo twin."""
    cfgs = mixed_configs()[:4]
    say("[batch membership]")
    ws = [sim.World(c, k).init() for c, k in cfgs]
    solo = [sim.World(c, k).init() for c, k in cfgs]
    with sim.Batch(ws[:2]) as b:
        b.step(2)
        b.add(ws[2])
        b.step(1)
        b.remove(ws[0])
        b.add(ws[3])
        b.step(2)
    steps = [3, 5, 3, 2]
    for w, n in zip(solo, steps):
        w.step(n)
    for i, (m, w) in enumerate(zip(ws, solo)):
        assert not diff_fields(m, w, STATE_FIELDS), f"world {i}"
    for w in ws + solo:
        w.close()


def test_member_cannot_step_alone_or_be_destroyed(sim):
    """A batch member steps only through its batch: stepping or destroying it
    on its own is refused with TT_ERR_STATE and leaves it intact.  Removing it,
    or closing the batch, hands it back."""
    (c0, k0), (c1, k1) = mixed_configs()[:2]
    say("[batch member guard]")
    a, b_ = sim.World(c0, k0).init(), sim.World(c1, k1).init()
    with sim.Batch([a, b_]) as b:
        b.step(1)
        for call in (lambda: a.step(1), a.close):
            with pytest.raises(RuntimeError, match="bad state"):
                call()
        assert a.h and a.iteration == 1
        b.remove(a)
        a.step(1)
        assert a.iteration == 2
    b_.step(1)
    assert b_.iteration == 2
    a.close()
    b_.close()


def test_host_threads_restart_as_membership_grows_and_shrinks(sim):
    """The batch's host threads are restarted whenever the member count moves
    the number of threads in use (here on every step: 1, 2, 3, 4, 3, 2, 1
    members with up to 4 threads).  Every member still ends equal to its
    solo twin: a restarted thread runs only the jobs issued after it started."""
    cfgs = mixed_configs()[:4]
    say("[batch restarts]")
    ws = [sim.World(c, k).init() for c, k in cfgs]
    solo = [sim.World(c, k).init() for c, k in cfgs]
    steps = [0] * 4
    with sim.Batch([ws[0]], threads=4) as b:
        plan = [+1, +2, +3, -3, -2, -1, 0]
        for s, change in enumerate(plan):
            b.step(1)
            for i in range(4):
                if ws[i] in b.members:
                    steps[i] += 1
            say(f"[batch restarts] step {s}: {len(b)} members")
            if change > 0:
                b.add(ws[change])
            elif change < 0:
                b.remove(ws[-change])
    for w, n in zip(solo, steps):
        w.step(n)
    for i, (m, w) in enumerate(zip(ws, solo)):
        assert not diff_fields(m, w, STATE_FIELDS), f"world {i} after {steps[i]} steps"
    for w in ws + solo:
        w.close()
