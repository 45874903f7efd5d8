"""Water-balance closure (the accumulators of pihm_ctx_set_closure).

* Per call: the storage rate implied by ydot, sum area (dsurf + porosity (dunsat + dgw))
  + sum area_r (dstage + porosity_r drivgw), equals the accumulated boundary rates
  q_in - q_et - q_bnd - q_riv to rounding (internal exchanges cancel: lateral fluxes are
  antisymmetric, bank fluxes mirror between element and river equations).  The state
  entries of ydot are bit-identical with and without the accumulators.
* Over a free run: storage change minus net inflow is below 1e-6 of the precipitation
  volume (the integrator carries the accumulators with the same steps, and its linear
  algebra preserves the linear invariant up to rounding).
"""
import math

import numpy as np
import pytest

from pihm_amd.balance import closure, storage
from pihm_amd.forcing import StormForcing, run_config
from pihm_amd.rhs import HipRHS, run_gpu
from test_oracle_pins import _inputs, _states

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("which", ["c0", "tree_small"])
def test_closure_rates_per_call(which, request):
    m = request.getfixturevalue(which)
    rng = np.random.default_rng(31)
    h0 = HipRHS(m)
    h1 = HipRHS(m, closure=True)
    inp = _inputs(m, rng)
    for h in (h0, h1):
        h.set_step_inputs(**inp)
    a = m.a
    ne, nr = m.nelem, m.nriver
    for y in _states(m, rng)[:3]:
        rc0, d0 = h0.ode(y)
        rc1, d1 = h1.ode(y)
        assert rc0 == 0 and rc1 == 0
        assert np.array_equal(d0, d1[:m.neq])
        rate = math.fsum((a["area"] * (d1[:ne] + a["porosity"] * (d1[ne:2 * ne] + d1[2 * ne:3 * ne]))).tolist()) + \
            math.fsum((a["rarea"] * (d1[3 * ne:3 * ne + nr] + a["rporosity"] * d1[3 * ne + nr:3 * ne + 2 * nr])).tolist())
        q = d1[m.neq:m.neq + 4]
        net = q[0] - q[1] - q[2] - q[3]
        scale = abs(q[0]) + abs(q[1]) + abs(q[2]) + abs(q[3]) + \
            math.fsum(np.abs(a["area"] * d1[:ne]).tolist())
        print(which, "rate", rate, "net", net, "scale", scale, flush=True)
        assert abs(rate - net) <= 1e-11 * scale
        assert np.isclose(q[0], math.fsum((inp["pcpdrp"] * a["area"]).tolist()), rtol=1e-12)
    h0.close()
    h1.close()


def test_closure_free_run(c0):
    m = c0
    f = StormForcing(30, seed=5, nelem=m.nelem, p_wet=0.5)
    y0 = m.initial_state()
    print("[closure] c0 24 h", flush=True)
    rc, y1, st = run_gpu(m, f, run_config(24 * 3600.0), y0, closure=True)
    assert rc == 0
    c = closure(m, y0, y1)
    print(c, st.as_dict(), flush=True)
    assert c["precip_m3"] > 0.0
    assert c["rel"] < 1e-6
    # the model's own state agrees with the run without accumulators
    rc2, y2, _ = run_gpu(m, f, run_config(24 * 3600.0), y0)
    w = 1.0 / (1e-3 * np.abs(y2) + 1e-4)
    assert math.sqrt(np.mean(((y1[:m.neq] - y2) * w) ** 2)) < 1.0


def test_closure_free_run_tree(tree_small):
    m = tree_small
    f = StormForcing(6, seed=8, nelem=m.nelem, p_wet=0.5)
    y0 = m.initial_state()
    rc, y1, _ = run_gpu(m, f, run_config(2 * 3600.0), y0, closure=True)
    assert rc == 0
    c = closure(m, y0, y1)
    print(c, flush=True)
    assert c["rel"] < 1e-6
