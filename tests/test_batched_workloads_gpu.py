"""The batched ensemble and the serving loop at the shapes of their workloads, against the
oracle's host run.

* Ensemble: a stratified 16-member subset of the 256-member ensemble (log-uniform
  2,000-50,000 triangles, per-member Ksat/porosity/Manning multipliers and storm seeds)
  over 20 simulated minutes.  Every member finishes; the smallest, a middle and the
  largest member agree with the oracle's host run (same integrator on the oracle RHS and
  host vectors) within the solver tolerance (WRMS of the difference weighted with the
  run's tolerances below 1, step counts within 10%); two members run alone give the bits
  they gave in the mix.
* Serving: a one-second Poisson stream at 20 requests/s of sizes drawn log-uniformly from
  1,000-100,000 triangles, horizons U(1, 30) minutes, served with continuous batching.
  Every request finishes; the smallest, the median and the largest request served alone
  give the bits they gave in the mix; the smallest agrees with the oracle's host run.
"""
import math
import time

import numpy as np
import pytest

import oracle as O
from pihm_amd.ensemble import Ensemble
from pihm_amd.forcing import run_config
from pihm_amd.serve import Request, Scheduler, SlotEngine, latency_summary
from pihm_amd.workloads import _build, member_forcing, member_spec, request_stream, serve_payload, tree_segments

pytestmark = pytest.mark.gpu

ENS_T = 1200.0


def _wrms(a, b, rtol=1e-3, atol=1e-4):
    w = 1.0 / (rtol * np.abs(b) + atol)
    return math.sqrt(np.mean(((a - b) * w) ** 2))


def _host(m, f, y0, t_end):
    rc, y, st = O.OracleRun(m, f, nthreads=8).run(y0, run_config(t_end))
    assert rc == 0
    return y, st


def _ens(meshes, forc, t_end):
    e = Ensemble(meshes, forc, run_config(t_end))
    try:
        for k, m in enumerate(meshes):
            e.set_state(k, m.initial_state())
        nfail = e.run(t_end)
        return nfail, [e.get_state(k) for k in range(len(meshes))], [e.stats(k) for k in range(len(meshes))]
    finally:
        e.close()


def test_ensemble_workload_subset_against_host_runs():
    specs = member_spec(256)
    order = np.argsort([s["n_tri"] for s in specs])
    pick = order[np.linspace(0, 255, 16).round().astype(int)]
    sub = [specs[k] for k in pick]
    meshes = [_build(s) for s in sub]          # in this process: no worker forked from it
    forc = [member_forcing(s, m, 3, p_wet=0.5) for s, m in zip(sub, meshes)]
    sizes = [m.nelem for m in meshes]
    print(f"[c3 subset] 16 members, {min(sizes)}-{max(sizes)} triangles, {sum(sizes)} in all", flush=True)
    nfail, ys, st = _ens(meshes, forc, ENS_T)
    assert nfail == 0, [s.as_dict() for s in st]
    assert all(s.model_steps == ENS_T / 60 for s in st)
    assert min(sizes) < 2500 and max(sizes) > 40000
    for k in (0, 8, 15):
        m = meshes[k]
        print(f"[c3 subset] member {k} ({m.nelem} triangles) against the host run", flush=True)
        yh, sth = _host(m, forc[k], m.initial_state(), ENS_T)
        d = _wrms(ys[k], yh)
        print(f"  wrms {d:.3g} steps {st[k].nst} vs {sth.nst}", flush=True)
        assert d < 1.0
        assert abs(st[k].nst - sth.nst) <= 0.1 * sth.nst + 2
    for k in (0, 15):
        print(f"[c3 subset] member {k} alone", flush=True)
        nf1, y1, st1 = _ens([meshes[k]], [forc[k]], ENS_T)
        assert nf1 == 0
        assert np.array_equal(y1[0], ys[k])
        assert st1[0].nst == st[k].nst and st1[0].nfe == st[k].nfe


SLOTS = [(8192, 255, 4), (32768, 1023, 4), (131072, 2047, 4)]


def _serve(payloads, t_ends, arrivals, caps):
    eng = SlotEngine(caps, run_config(3600.0))
    try:
        t0 = time.perf_counter()
        clock = lambda: time.perf_counter() - t0
        reqs = [Request(i, arrivals[i], t_ends[i], p) for i, p in enumerate(payloads)]
        Scheduler(eng, clock, ticks_per_round=8).run(reqs, sleep=time.sleep)
    finally:
        eng.close()
    return reqs


def test_serving_stream_alone_equals_mix_and_host_run():
    stream = request_stream(duration_s=1.0, rate=20.0, seed=77, hmin_days=1.0, hmax_days=30.0)
    # horizons U(1, 30) minutes, in whole 60 s model steps
    t_ends = [max(60.0, round(rq["horizon_days"] * 60.0 / 60.0) * 60.0) for rq in stream]
    pls = [serve_payload(rq, t / 3600.0 + 1) for rq, t in zip(stream, t_ends)]
    for (m, f, _), rq in zip(pls, stream):
        assert m.nelem == rq["n_tri"] and m.nriver == tree_segments(rq["n_tri"])
    caps = []
    for ne, nr, cnt in SLOTS:
        caps += [(ne, nr, max(f.nrec for _, f, _ in pls))] * cnt
    sizes = [p[0].nelem for p in pls]
    print(f"[c4 stream] {len(pls)} requests, {min(sizes)}-{max(sizes)} triangles", flush=True)
    reqs = _serve(pls, t_ends, [rq["arrival"] for rq in stream], caps)
    assert all(r.t_done is not None and not r.failed for r in reqs)
    for r, t in zip(reqs, t_ends):
        assert r.stats.model_steps == t / 60
    print("[c4 stream]", latency_summary(reqs), flush=True)
    order = np.argsort(sizes)
    for i in (order[0], order[len(order) // 2], order[-1]):
        print(f"[c4 stream] request {i} ({sizes[i]} triangles) alone", flush=True)
        alone = _serve([pls[i]], [t_ends[i]], [0.0], [c for c in caps if c[0] >= sizes[i]][:1])
        assert np.array_equal(alone[0].result, reqs[i].result), i
        assert alone[0].stats.nst == reqs[i].stats.nst
    i = order[0]
    m, f, y0 = pls[i]
    print(f"[c4 stream] request {i} against the host run", flush=True)
    yh, sth = _host(m, f, y0, t_ends[i])
    d = _wrms(reqs[i].result, yh)
    print(f"  wrms {d:.3g} steps {reqs[i].stats.nst} vs {sth.nst}", flush=True)
    assert d < 1.0
    assert abs(reqs[i].stats.nst - sth.nst) <= 0.1 * sth.nst + 2
