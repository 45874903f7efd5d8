"""The autotuned launch shapes (pihm_ctx_autotune, pihm_ens_autotune) change timing only:

* every element-tile placement over the XCDs gives the same RHS bit for bit, and the
  single-solve tuner leaves the carried river-edge flux as it found it;
* ensemble runs under every (tile placement, chunks per workgroup) choice, including the
  tuner's own, give the same member states and counters bit for bit.
"""
import ctypes as C

import numpy as np
import pytest

from pihm_amd import _lib
from pihm_amd.ensemble import Ensemble
from pihm_amd.forcing import run_config
from pihm_amd.rhs import HipRHS
from test_ensemble_gpu import _members, _same_stats
from test_oracle_pins import _inputs, _states

pytestmark = pytest.mark.gpu


def test_tile_placement_is_neutral_and_tuner_restores_state(tree_small):
    m = tree_small
    print(f"[tune] single solve ne={m.nelem}", flush=True)
    rng = np.random.default_rng(4)
    h = HipRHS(m)
    h.set_step_inputs(**_inputs(m, rng))
    y = _states(m, rng, 2)[2]
    h.ode(y)                                      # a carried river-edge flux
    carried = h.get_river_edge_ovlflow()
    ref = None
    for span in _lib.PIHM_TUNE_SPANS:
        assert h.L.pihm_ctx_set_tile_span(h.ctx, span) == 0
        h.set_river_edge_ovlflow(carried)
        rc, d = h.ode(y)
        assert rc == 0
        if ref is None:
            ref = d
        assert np.array_equal(d, ref), span
    h.set_river_edge_ovlflow(carried)
    assert h.L.pihm_ctx_state_to_device(h.ctx, y.ctypes.data_as(_lib.c_double_p), h.y) == 0
    best = C.c_int()
    ms = np.zeros(len(_lib.PIHM_TUNE_SPANS))
    ydev = h.L.N_VGetDeviceArrayPointer_Hip(h.y)
    yddev = h.L.N_VGetDeviceArrayPointer_Hip(h.yd)
    assert h.L.pihm_ctx_autotune(h.ctx, ydev, yddev, 10, C.byref(best), ms.ctypes.data_as(_lib.c_double_p)) == 0
    print("[tune] ms by span", dict(zip(_lib.PIHM_TUNE_SPANS, ms.round(5))), "best", best.value, flush=True)
    assert best.value in _lib.PIHM_TUNE_SPANS and np.all(ms > 0)
    assert h.L.pihm_ctx_get_tile_span(h.ctx) == best.value
    assert h.rhs_path() in (_lib.PIHM_RHS_ONE_LAUNCH, _lib.PIHM_RHS_TWO_KERNELS)
    assert np.array_equal(h.get_river_edge_ovlflow(), carried)
    rc, d = h.ode(y)
    assert np.array_equal(d, ref)
    h.close()


def _run(ms, fs, y0, tuning, hours=1):
    e = Ensemble(ms, fs, run_config(hours * 3600.0))
    try:
        for k, y in enumerate(y0):
            e.set_state(k, y)
        info = e.autotune(5) if tuning == "auto" else (e.set_tuning(*tuning) or None)
        assert e.run(hours * 3600.0) == 0
        return [e.get_state(k) for k in range(len(ms))], [e.stats(k) for k in range(len(ms))], info
    finally:
        e.close()


def test_ensemble_tuning_is_neutral():
    ms, fs, y0 = _members()
    print("[tune] ensemble default", flush=True)
    y_ref, st_ref, _ = _run(ms, fs, y0, (0, 1))
    for tuning in ((16, 8), (1, 2), "auto"):
        print(f"[tune] ensemble {tuning}", flush=True)
        y, st, info = _run(ms, fs, y0, tuning)
        if info:
            print("[tune] chose", info, flush=True)
        for k in range(len(ms)):
            assert np.array_equal(y[k], y_ref[k]), (tuning, k)
            assert _same_stats(st[k], st_ref[k]), (tuning, k)


def test_ensemble_autotune_refuses_running_members():
    from pihm_amd.serve import SlotEngine
    from pihm_amd.workloads import tree_segments
    ms, fs, y0 = _members()
    m = ms[0]
    eng = SlotEngine([(1024, tree_segments(1024), 16)], run_config(3600.0))
    try:
        t = _lib.EnsTuning()
        assert eng.L.pihm_ens_autotune(eng.ens, 2, C.byref(t)) == _lib.PIHM_ERR_ARG   # nothing loaded
        eng.load(0, m, fs[0], y0[0], 3600.0)                                        # running from now on
        assert eng.L.pihm_ens_autotune(eng.ens, 2, C.byref(t)) == _lib.PIHM_ERR_ARG
    finally:
        eng.close()
