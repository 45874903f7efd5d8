"""Serving-loop host logic (admission by footprint, back-fill, retirement, latency) on a
fake engine and a fake clock: no device involved."""
from collections import deque

import pytest

from pihm_amd.serve import (PIHM_ENS_DONE, PIHM_ENS_FAILED, PIHM_ENS_RUNNING, Request, Scheduler,
                            latency_summary, slot_bytes)


class FakeMesh:
    def __init__(self, ne, nr):
        self.nelem, self.nriver = ne, nr


class FakeForcing:
    def __init__(self, nrec):
        self.nrec = nrec


class FakeEngine:
    """A member finishes after `work` ticks of its own (work = t_end / 60 here)."""

    def __init__(self, caps, clock):
        self.caps = caps
        self.nslots = len(caps)
        self.left = {}
        self.loads = []
        self.clock = clock
        self.fail_ids = set()

    def fits(self, k, mesh, forcing):
        ne, nr, nrec = self.caps[k]
        return mesh.nelem <= ne and mesh.nriver <= nr and forcing.nrec <= nrec

    def load(self, k, mesh, forcing, y0, t_end):
        assert k not in self.left
        self.left[k] = [int(t_end // 60), y0]
        self.loads.append((k, mesh.nelem))

    def advance(self, max_ticks):
        fin = 0
        for _ in range(max_ticks):
            self.clock.t += 1.0
            for k in list(self.left):
                self.left[k][0] -= 1
                if self.left[k][0] <= 0:
                    fin += 1
            if fin:
                break
        return fin

    def status(self, k):
        if k not in self.left:
            return 0
        if self.left[k][0] > 0:
            return PIHM_ENS_RUNNING
        return PIHM_ENS_FAILED if self.left[k][1] in self.fail_ids else PIHM_ENS_DONE

    def result(self, k):
        y0 = self.left.pop(k)[1]
        return y0, None


class Clock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        return self.t


def _req(rid, arrival, ne, steps, nrec=10):
    return Request(rid, arrival, steps * 60.0, (FakeMesh(ne, ne // 50), FakeForcing(nrec), rid))


def test_smallest_fitting_slot_and_backfill():
    clock = Clock()
    eng = FakeEngine([(1000, 20, 10), (5000, 100, 10), (20000, 400, 10)], clock)
    s = Scheduler(eng, clock, ticks_per_round=1)
    reqs = [_req(0, 0.0, 900, 5), _req(1, 0.0, 18000, 50), _req(2, 0.0, 15000, 3),
            _req(3, 0.0, 800, 2), _req(4, 0.0, 4000, 2)]
    s.run(reqs)
    order = eng.loads
    # 900 -> smallest slot (0); 18000 -> slot 2; 15000 waits (only slot 2 fits) while
    # 800 back-fills the next smallest free slot and 4000 follows when slot 1 frees
    assert order[:3] == [(0, 900), (2, 18000), (1, 800)]
    pos = {ne: i for i, (_, ne) in enumerate(order)}
    assert pos[800] < pos[15000] and pos[4000] < pos[15000]
    assert all(r.t_done is not None and r.result == r.rid for r in reqs)
    assert reqs[2].t_admit >= reqs[1].t_done       # the big request waited for the big slot


def test_arrivals_over_time_and_latency():
    clock = Clock()
    eng = FakeEngine([(1000, 20, 10)] * 2, clock)
    s = Scheduler(eng, clock, ticks_per_round=4)
    reqs = [_req(i, 3.0 * i, 500, 4) for i in range(6)]

    def sleep(dt):
        clock.t += max(dt, 1e-3)

    s.run(reqs, sleep=sleep)
    lat = latency_summary(reqs)
    assert lat["n"] == 6
    for r in reqs:
        assert r.t_admit >= r.arrival and r.t_done > r.t_admit
    assert lat["p50_s"] <= lat["p99_s"] <= lat["max_s"]


def test_failed_member_frees_its_slot():
    clock = Clock()
    eng = FakeEngine([(1000, 20, 10)], clock)
    eng.fail_ids = {1}
    s = Scheduler(eng, clock, ticks_per_round=2)
    reqs = [_req(i, 0.0, 500, 3) for i in range(3)]
    s.run(reqs)
    assert [r.failed for r in reqs] == [False, True, False]
    assert all(r.t_done is not None for r in reqs)


def test_request_that_fits_no_slot_is_an_error():
    clock = Clock()
    eng = FakeEngine([(1000, 20, 10)], clock)
    s = Scheduler(eng, clock)
    with pytest.raises(RuntimeError):
        s.run([_req(0, 0.0, 5000, 3)])


def test_forcing_capacity_counts():
    clock = Clock()
    eng = FakeEngine([(1000, 20, 10), (1000, 20, 100)], clock)
    s = Scheduler(eng, clock)
    s.run([_req(0, 0.0, 500, 3, nrec=50)])
    assert eng.loads == [(1, 500)]


def test_slot_bytes_grow_with_size():
    assert slot_bytes(2000, 40) < slot_bytes(4000, 80)
    assert slot_bytes(100000, 2047) > 100000 * 20 * 8 * 3
