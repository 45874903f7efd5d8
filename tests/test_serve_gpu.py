"""Serving path on the device: solves admitted into slots of a slot context while others
run, slots reused for later requests.

* Every request's result is bit-identical to the same request served alone, and to the
  same solve in the batched ensemble path (its arithmetic depends on its own data only).
* Every request agrees with the single-solve device path within the solver tolerance
  (WRMS of the difference, weighted with the run's tolerances, below 1).
* Loading a slot again gives the same result as the first time (no state leaks between
  the solves a slot serves).
"""
import math
import time

import numpy as np
import pytest

from conftest import mesh_cached
from pihm_amd.ensemble import Ensemble
from pihm_amd.forcing import StormForcing, run_config
from pihm_amd.meshgen import make_watershed
from pihm_amd.rhs import run_gpu
from pihm_amd.serve import PIHM_ENS_DONE, Request, Scheduler, SlotEngine, latency_summary

pytestmark = pytest.mark.gpu

T_END = 1800.0


def _wrms(a, b, rtol=1e-3, atol=1e-4):
    w = 1.0 / (rtol * np.abs(b) + atol)
    return math.sqrt(np.mean(((a - b) * w) ** 2))


def _requests():
    specs = [("c0", lambda: make_watershed(535, seed=1, n_seg=20, network="valley")),
             ("ens_t3000", lambda: make_watershed(3000, seed=2, n_seg=80, network="tree")),
             ("ens_t1200", lambda: make_watershed(1200, seed=4, n_seg=30, network="tree",
                                                  param_mult=dict(ksat=1.4, porosity=0.8, rough=1.2))),
             ("srv_t2000", lambda: make_watershed(2000, seed=8, n_seg=40, network="tree",
                                                  param_mult=dict(ksat=0.6, porosity=1.3, rough=0.7)))]
    out = []
    for i, (key, f) in enumerate(specs):
        m = mesh_cached(key, f)
        fo = StormForcing(4, seed=40 + i, nelem=m.nelem, p_wet=0.5)
        out.append((m, fo, m.initial_state()))
    return out


CAPS = [(1500, 40, 12), (4000, 100, 12)]


def _serve(payloads, caps=CAPS, arrivals=None):
    eng = SlotEngine(caps, run_config(T_END))
    try:
        t0 = time.perf_counter()
        clock = lambda: time.perf_counter() - t0
        reqs = [Request(i, 0.0 if arrivals is None else arrivals[i], T_END, p) for i, p in enumerate(payloads)]
        Scheduler(eng, clock, ticks_per_round=16).run(reqs, sleep=time.sleep)
        cnt = eng.counters()
    finally:
        eng.close()
    return reqs, cnt


def test_served_equals_alone_and_ensemble():
    pls = _requests()
    print("[serve] 4 requests, 2 slots", flush=True)
    reqs, cnt = _serve(pls)
    print(cnt, latency_summary(reqs), flush=True)
    assert all(not r.failed for r in reqs)
    assert all(r.stats.model_steps == T_END / 60 for r in reqs)
    for i, p in enumerate(pls):
        print(f"[serve] request {i} alone", flush=True)
        alone, _ = _serve([p], caps=[CAPS[1]])
        assert np.array_equal(alone[0].result, reqs[i].result), f"request {i}"
        assert alone[0].stats.nst == reqs[i].stats.nst and alone[0].stats.nfe == reqs[i].stats.nfe
        e = Ensemble([p[0]], [p[1]], run_config(T_END))
        try:
            e.set_state(0, p[2])
            assert e.run(T_END) == 0
            ye = e.get_state(0)
        finally:
            e.close()
        assert np.array_equal(ye, reqs[i].result), f"request {i} vs ensemble"


def test_served_matches_single_solve_path():
    pls = _requests()
    reqs, _ = _serve(pls)
    for i, (m, f, y0) in enumerate(pls):
        rc, yg, stg = run_gpu(m, f, run_config(T_END), y0)
        assert rc == 0
        d = _wrms(reqs[i].result, yg)
        print(i, "wrms", d, reqs[i].stats.nst, stg.nst, flush=True)
        assert d < 1.0


def test_slot_reuse_is_clean():
    """Six requests through one slot (every later one loads the slot a previous one
    used) give the same results as the first pass of the same requests."""
    pls = _requests()
    seq = [pls[0], pls[2], pls[0], pls[3], pls[2], pls[3]]
    reqs, _ = _serve(seq, caps=[CAPS[1]])
    assert np.array_equal(reqs[0].result, reqs[2].result)
    assert np.array_equal(reqs[1].result, reqs[4].result)
    assert np.array_equal(reqs[3].result, reqs[5].result)


def test_staggered_arrivals_overlap():
    """Requests arriving while others run are admitted into free slots and their work
    shares launches with the running ones; results equal the all-at-once results."""
    pls = _requests()
    ref, _ = _serve(pls)
    reqs, cnt = _serve(pls, caps=CAPS + [(4000, 100, 12)], arrivals=[0.0, 0.0, 0.05, 0.1])
    for a, b in zip(ref, reqs):
        assert np.array_equal(a.result, b.result)
    running_together = [r for r in reqs if r.t_admit > 0.0]
    assert running_together, "late arrivals were admitted"
