"""Serving (tt_server_*): admission under an HBM budget, continuous batching,
latency accounting, and the same result for a job alone and in a mix."""
from __future__ import annotations

import sys
import time

import pytest

from master_thesis_amd import _abi, configs
from tests.helpers import STATE_FIELDS, diff_fields

pytestmark = pytest.mark.gpu

GB = 1 << 30


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def sim(tectonic_lib):
    from master_thesis_amd import sim as s
    return s


def job(W, H, P, seed, K=4, T=2):
    return configs.synthetic(W, H, P, (0.5, 3.0), seed, hydraulic=K, thermal=T, ccl_every=1)


def test_job_alone_and_in_a_mix(sim):
    target = job(192, 128, 7, 11)
    say("[serve] alone")
    with sim.Server(8 * GB, keep=True) as s:
        jid = s.submit(*target, steps=6)
        s.drain(120000)
        alone = s.take(jid)
    say("[serve] mix")
    with sim.Server(8 * GB, keep=True) as s:
        others = [s.submit(*job(256, 256, 9, 1), steps=9), s.submit(*job(128, 64, 4, 2), steps=3)]
        time.sleep(0.05)
        jid = s.submit(*target, steps=6)          # arrives while the others run
        others.append(s.submit(*job(320, 192, 12, 3, K=6, T=1), steps=4))
        s.drain(120000)
        mixed = s.take(jid)
        rep = s.report()
        assert rep.completed == 4 and rep.failed == 0
        for o in others:
            assert s.result(o, 0).status == 0
    assert not diff_fields(alone, mixed, STATE_FIELDS)
    assert alone.iteration == mixed.iteration == 6
    alone.close()
    mixed.close()


def test_admission_respects_the_budget(sim):
    cfg, ks = job(512, 512, 8, 5)
    one = sim.world_bytes(cfg)
    budget = int(one * 2.5)                        # two worlds of this size at a time
    say(f"[serve] budget {budget}")
    with sim.Server(budget) as s:
        ids = [s.submit(*job(512, 512, 8, 5 + i), steps=3) for i in range(5)]
        big = job(1024, 1024, 8, 1)
        with pytest.raises(RuntimeError):
            s.submit(*big, steps=1)                # larger than the whole budget
        s.drain(300000)
        rep = s.report()
        assert rep.completed == 5 and rep.rejected == 1
        assert rep.bytes_peak <= budget and rep.bytes_peak >= 2 * one
        res = [s.result(i, 0) for i in ids]
        # FIFO admission: the third job waits for a free slot
        assert res[2].t_admit >= min(res[0].t_done, res[1].t_done)
        assert all(r.iteration == 3 for r in res)
        assert rep.latency_p50_s > 0 and rep.latency_p99_s >= rep.latency_p50_s
        assert rep.worlds_per_s > 0


def test_result_wait_and_zero_steps(sim):
    with sim.Server(4 * GB) as s:
        jid = s.submit(*job(64, 64, 2, 9), steps=0)
        r = s.result(jid, 60000)
        assert r is not None and r.status == 0 and r.iteration == 0
