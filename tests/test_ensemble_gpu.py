"""Ensemble path (pihm_ensemble.h): many members advanced together on the device.

* Each member's trajectory equals its trajectory when it runs alone in an ensemble,
  bit for bit, with identical counters (its reductions depend on its own data only).
* Each member agrees with the single-solve device path (pihm_run_gpu) within the solver
  tolerance: the integrators take the same steps and differ only in the summation order
  of the inner products, so the WRMS of the state difference, weighted with the run's own
  tolerances, stays below 1 and the step counts agree within 10%.
* A member whose derivative turns NaN is retired with PIHM_ERR_NAN; the others finish
  with results equal to their runs without it.
"""
import math

import numpy as np
import pytest

from conftest import mesh_cached
from pihm_amd.ensemble import Ensemble
from pihm_amd.forcing import StormForcing, run_config
from pihm_amd.meshgen import make_watershed
from pihm_amd.rhs import run_gpu

pytestmark = pytest.mark.gpu

PIHM_ERR_NAN = 1
HOURS = 2


def _wrms(a, b, rtol=1e-3, atol=1e-4):
    w = 1.0 / (rtol * np.abs(b) + atol)
    return math.sqrt(np.mean(((a - b) * w) ** 2))


def _members():
    ms = [
        mesh_cached("c0", lambda: make_watershed(535, seed=1, n_seg=20, network="valley")),
        mesh_cached("ens_t3000", lambda: make_watershed(3000, seed=2, n_seg=80, network="tree")),
        mesh_cached("ens_t1200", lambda: make_watershed(1200, seed=4, n_seg=30, network="tree",
                                                        param_mult=dict(ksat=1.4, porosity=0.8,
                                                                        rough=1.2))),
    ]
    fs = [StormForcing(HOURS + 4, seed=20 + k, nelem=m.nelem, p_wet=0.5) for k, m in enumerate(ms)]
    y0 = [m.initial_state() for m in ms]
    return ms, fs, y0


def _run_ens(ms, fs, y0, hours=HOURS):
    cfg = run_config(hours * 3600.0)
    e = Ensemble(ms, fs, cfg)
    try:
        for k, y in enumerate(y0):
            e.set_state(k, y)
        nfail = e.run(hours * 3600.0)
        out = [e.get_state(k) for k in range(len(ms))]
        st = [e.stats(k) for k in range(len(ms))]
        cnt = e.counters()
    finally:
        e.close()
    return nfail, out, st, cnt


def _same_stats(a, b):
    return all(getattr(a, f) == getattr(b, f) for f in ("nst", "nfe", "nfes", "nni", "nli", "ncfn",
                                                         "netf", "model_steps", "flag"))


def test_ensemble_alone_equals_mix():
    ms, fs, y0 = _members()
    print("[ens] mix of 3", flush=True)
    nfail, ymix, stmix, cnt = _run_ens(ms, fs, y0)
    assert nfail == 0, [s.as_dict() for s in stmix]
    print("counters", cnt, flush=True)
    for k in range(len(ms)):
        print(f"[ens] member {k} alone", flush=True)
        nf1, y1, st1, _ = _run_ens([ms[k]], [fs[k]], [y0[k]])
        assert nf1 == 0
        assert np.array_equal(y1[0], ymix[k]), f"member {k}"
        assert _same_stats(st1[0], stmix[k]), (st1[0].as_dict(), stmix[k].as_dict())
        assert stmix[k].model_steps == HOURS * 60


def test_ensemble_matches_single_solve_path():
    ms, fs, y0 = _members()
    print("[ens] mix of 3 vs pihm_run_gpu", flush=True)
    nfail, ymix, stmix, _ = _run_ens(ms, fs, y0)
    assert nfail == 0
    for k, m in enumerate(ms):
        rc, yg, stg = run_gpu(m, fs[k], run_config(HOURS * 3600.0), y0[k])
        assert rc == 0
        d = _wrms(ymix[k], yg)
        print(k, "wrms", d, "ens", stmix[k].as_dict(), "\nsingle", stg.as_dict(), flush=True)
        assert d < 1.0
        assert abs(stmix[k].nst - stg.nst) <= 0.1 * stg.nst + 2


def test_ensemble_deterministic():
    ms, fs, y0 = _members()
    r1 = _run_ens(ms, fs, y0, hours=1)
    r2 = _run_ens(ms, fs, y0, hours=1)
    for k in range(len(ms)):
        assert np.array_equal(r1[1][k], r2[1][k])
        assert _same_stats(r1[2][k], r2[2][k])


def _adverse_outlet(m):
    """Raise the outlet node 3 m above the bed upstream: the zero-depth-gradient outlet
    formula then takes the square root of a negative slope, so ydot is NaN there."""
    r_out = int(np.nonzero(m.rivers["down"] < 0)[0][0])
    to = m.rivers["tonode"][r_out] - 1
    m.node_zmax[to] += 3.0
    m.node_zmin[to] += 3.0
    m.a = None
    m.init()
    return m


def test_ensemble_nan_member_is_retired():
    ms, fs, y0 = _members()
    bad = _adverse_outlet(make_watershed(1200, seed=4, n_seg=30, network="tree",
                                         param_mult=dict(ksat=1.4, porosity=0.8, rough=1.2)))
    mix = [ms[0], bad, ms[1]]
    fmix = [fs[0], StormForcing(5, seed=9, nelem=bad.nelem, p_wet=0.5), fs[1]]
    print("[ens] NaN member", flush=True)
    nfail, ys, st, _ = _run_ens(mix, fmix, [y0[0], bad.initial_state(), y0[1]], hours=1)
    assert nfail == 1
    assert st[1].flag == PIHM_ERR_NAN
    ok, yok, stok, _ = _run_ens(ms[:2], fs[:2], y0[:2], hours=1)
    assert ok == 0
    assert np.array_equal(ys[0], yok[0]) and np.array_equal(ys[2], yok[1])
    assert st[0].flag == 0 and st[2].flag == 0
    assert _same_stats(st[0], stok[0]) and _same_stats(st[2], stok[1])


def test_ensemble_integrator_failure_is_isolated():
    """A NaN in a member's initial state (clamped away by the RHS, but not by the
    integrator's norms) ends that member with an integrator error; the others finish."""
    ms, fs, y0 = _members()
    bad = [y.copy() for y in y0]
    bad[1][5] = np.nan
    nfail, ys, st, _ = _run_ens(ms, fs, bad, hours=1)
    assert nfail == 1
    assert st[1].flag < 0
    ok, yok, _, _ = _run_ens([ms[0], ms[2]], [fs[0], fs[2]], [y0[0], y0[2]], hours=1)
    assert ok == 0
    assert np.array_equal(ys[0], yok[0]) and np.array_equal(ys[2], yok[1])


def test_ensemble_resumes_across_runs():
    """Two runs of one hour each equal one run of two hours (the integrator state persists)."""
    ms, fs, y0 = _members()
    cfg = run_config(2 * 3600.0)
    e = Ensemble(ms[:2], fs[:2], cfg)
    try:
        for k in range(2):
            e.set_state(k, y0[k])
        assert e.run(3600.0) == 0
        assert e.run(2 * 3600.0) == 0
        ya = [e.get_state(k) for k in range(2)]
    finally:
        e.close()
    _, yb, _, _ = _run_ens(ms[:2], fs[:2], y0[:2], hours=2)
    for k in range(2):
        assert np.array_equal(ya[k], yb[k])


def test_ensemble_op_stats():
    """The operation kernel's launches are timed with events and priced at their
    algorithmic bytes (pihm_ens_op_stats): counts grow with each run, and every timed
    launch moves at least one vector of the smallest member and at most every member's
    whole queue (ops of at most 14 vectors, eight launches a tick)."""
    ms, fs, y0 = _members()
    cfg = run_config(HOURS * 3600.0)
    e = Ensemble(ms, fs, cfg)
    try:
        for k, y in enumerate(y0):
            e.set_state(k, y)
        s0 = e.op_stats()
        assert e.run(3600.0) == 0
        s1 = e.op_stats()
        assert e.run(HOURS * 3600.0) == 0
        s2 = e.op_stats()
        cnt = e.counters()
    finally:
        e.close()
    assert s0["launches"] < s1["launches"] < s2["launches"] <= cnt["op_launches"]
    assert s0["ms"] < s1["ms"] < s2["ms"] and s0["bytes"] < s1["bytes"] < s2["bytes"]
    per = s2["bytes"] / s2["launches"]
    neq = [m.neq for m in ms]
    assert 8.0 * min(neq) <= per <= 8.0 * sum(neq) * 14 * 64
