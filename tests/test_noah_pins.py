"""The _NOAH_ branches of the oracle RHS (the reference built for Flux-PIHM), on the CPU.

With _NOAH_ the reference changes three things in the hydrology RHS:
* EtExtract (src/hydrol.c:43-56,79-84) never takes direct evaporation from ponded water,
  and splits transpiration with GwTransp (src/noah/lsm_func.c:163-185), the share of the
  layers from the water-table layer down to the root layer;
* Infil (src/vert_flow.c:94-98) takes the unsaturated branch's saturation from the top
  soil layer's liquid water, (sh2o[0] - smcmin) / (smcmax - smcmin);
* VerticalFlow (src/vert_flow.c:16-19) multiplies infiltration by the frozen-soil factor
  fcr before Recharge reads it.

The reference ships no tests for these; they are pinned here by closed forms of GwTransp,
by the routing rules, by reduction to the plain build when the top layer's saturation is
the unsaturated zone's own, and by the linearity of the fcr factor.
"""
import numpy as np
import pytest

import oracle as O
from test_oracle_pins import _inputs, _states

OL = O.lib()


def _gw(ett, et, nwtbl, nroot):
    et = np.ascontiguousarray(et, dtype=np.float64)
    return OL.oracle_GwTransp(float(ett), et.ctypes.data_as(O.dp), int(nwtbl), int(nroot))


def test_gwtransp_closed_forms():
    ett = 3e-8
    for nroot in (1, 4, 11):
        et = np.full(11, ett / nroot)
        for nwtbl in range(-1, nroot + 1):
            first = max(nwtbl, 1)
            want = (nroot - first + 1) / nroot
            assert abs(_gw(ett, et, nwtbl, nroot) - want) <= 4e-16, (nroot, nwtbl)
        # water table below the roots: nothing from the saturated zone
        assert _gw(ett, et, nroot + 1, nroot) == 0.0
    et = np.zeros(11)
    et[2] = 1e-8                        # only the third layer transpires
    assert _gw(ett, et, 3, 5) == pytest.approx(1.0 / 3.0, rel=1e-15)
    assert _gw(ett, et, 4, 5) == 0.0    # layer 3 is above the water-table layer 4
    # no transpiration, or none asked for: zero
    assert _gw(0.0, np.full(11, 1e-8), 1, 4) == 0.0
    assert _gw(-1e-9, np.full(11, 1e-8), 1, 4) == 0.0
    # clamped to [0, 1]
    assert _gw(ett, np.full(11, ett), 1, 4) == 1.0
    assert _gw(ett, np.full(11, -ett), 1, 4) == 0.0


def _noah_inputs(m, rng, nlayer=11):
    ne = m.nelem
    por = m.a["porosity"]
    smcmin = rng.uniform(0.01, 0.06, ne)
    sh2o0 = smcmin + por * rng.uniform(-0.05, 1.05, ne)
    fcr = np.where(rng.random(ne) < 0.5, 1.0, rng.uniform(0.0, 1.0, ne))
    fcr[rng.random(ne) < 0.05] = 0.0
    et = rng.uniform(-0.1, 1.0, (ne, nlayer)) * 8e-8 / nlayer
    nroot = rng.integers(1, nlayer + 1, ne).astype(np.int32)
    nwtbl = rng.integers(-1, nlayer + 2, ne).astype(np.int32)
    return dict(smcmin=smcmin, sh2o0=sh2o0, fcr=fcr, et=et, nwtbl=nwtbl, nroot=nroot)


def _run(m, y, inp, noah=None):
    o = O.OracleRHS(m)
    o.set_step_inputs(**inp)
    if noah is not None:
        o.set_noah(**noah)
    rc, dy = o.ode(y)
    f = {k: o.field(k).copy() for k in ("edir_surf", "edir_unsat", "edir_gw", "ett_unsat", "ett_gw",
                                        "infil", "rechg", "surfh")}
    return rc, dy, f


def test_noah_etextract_routing(c0):
    m = c0
    rng = np.random.default_rng(31)
    inp = _inputs(m, rng)
    nz = _noah_inputs(m, rng)
    ne = m.nelem
    wet = 0
    for y in _states(m, rng, 4)[1:]:
        _, _, f = _run(m, y, inp, nz)
        gw = np.maximum(y[2 * ne:3 * ne], 0.0)
        wet += int(np.sum(f["surfh"] >= 1e-4))
        # never from ponded water
        assert np.all(f["edir_surf"] == 0.0)
        deep = gw > m.a["depth"] - m.a["dinf"]
        assert np.array_equal(f["edir_gw"], np.where(deep, inp["edir"], 0.0))
        assert np.array_equal(f["edir_unsat"], np.where(deep, 0.0, inp["edir"]))
        ett = inp["ett"]
        # the two shares make up the transpiration
        assert np.all(np.abs(f["ett_unsat"] + f["ett_gw"] - ett) <= 2.3e-16 * np.abs(ett))
        below = nz["nwtbl"] > nz["nroot"]
        assert np.all(f["ett_gw"][below] == 0.0) and np.array_equal(f["ett_unsat"][below], ett[below])
        # share over the layers nwtbl..nroot of the element's own et
        first = np.maximum(nz["nwtbl"], 1) - 1
        share = np.array([nz["et"][i, first[i]:nz["nroot"][i]].sum() for i in range(ne)]) / ett
        share = np.clip(np.where(below, 0.0, share), 0.0, 1.0)
        assert np.all(np.abs(f["ett_gw"] - share * ett) <= 1e-13 * ett)
    assert wet > 0          # ponded elements were there to route


def test_noah_infil_reduces_to_plain_build(c0, example):
    """With fcr = 1 and the top layer's saturation equal to the unsaturated zone's
    (sh2o[0] = smcmin + porosity * unsat / deficit), infiltration and recharge are the
    plain build's to rounding (bitwise where the saturated branch is taken)."""
    for m in (c0, example):
        rng = np.random.default_rng(32)
        inp = _inputs(m, rng)
        ne = m.nelem
        por, depth = m.a["porosity"], m.a["depth"]
        for y in _states(m, rng, 4)[1:]:
            _, _, fp = _run(m, y, inp)
            u = np.maximum(y[ne:2 * ne], 0.0)
            g = np.maximum(y[2 * ne:3 * ne], 0.0)
            smcmin = rng.uniform(0.01, 0.06, ne)
            sh2o0 = smcmin + por * (u / (depth - g))
            nz = dict(smcmin=smcmin, sh2o0=sh2o0, fcr=np.ones(ne), et=np.zeros((ne, 4)),
                      nwtbl=np.full(ne, 5, np.int32), nroot=np.full(ne, 4, np.int32))
            _, _, fn = _run(m, y, inp, nz)
            scale = np.maximum(np.abs(fp["infil"]), 1e-300)
            assert np.all(np.abs(fn["infil"] - fp["infil"]) <= 1e-12 * scale)
            deep = g > depth - m.a["dinf"]
            assert np.array_equal(fn["infil"][deep], fp["infil"][deep])
            assert np.all(np.abs(fn["rechg"] - fp["rechg"]) <= 1e-12 * np.maximum(np.abs(fp["rechg"]), 1e-300))
            assert np.array_equal(fn["rechg"][~deep], fp["rechg"][~deep])


def test_noah_top_layer_saturation_is_read(c0):
    """Infiltration follows sh2o[0], not the unsaturated storage: with the top layer at
    saturation s, Noah's infiltration equals the plain build's for a state whose
    unsaturated storage is s * deficit (and differs from the plain build at the
    original storage)."""
    m = c0
    rng = np.random.default_rng(33)
    inp = _inputs(m, rng)
    ne = m.nelem
    por, depth = m.a["porosity"], m.a["depth"]
    y = _states(m, rng, 2)[1]
    y[2 * ne:3 * ne] = np.abs(y[2 * ne:3 * ne])
    g = y[2 * ne:3 * ne]
    s = rng.uniform(0.15, 0.95, ne)
    smcmin = rng.uniform(0.01, 0.06, ne)
    nz = dict(smcmin=smcmin, sh2o0=smcmin + por * s, fcr=np.ones(ne), et=np.zeros((ne, 1)),
              nwtbl=np.full(ne, 2, np.int32), nroot=np.ones(ne, np.int32))
    _, _, fn = _run(m, y, inp, nz)
    _, _, fp = _run(m, y, inp)
    y2 = y.copy()
    y2[ne:2 * ne] = s * (depth - g)
    _, _, f2 = _run(m, y2, inp)
    u = np.maximum(y[ne:2 * ne], 0.0)
    unsat_branch = ~(g > depth - m.a["dinf"]) & (u + g <= depth)
    assert unsat_branch.sum() > 50
    sel = unsat_branch & (f2["infil"] > 0)
    assert sel.sum() > 10
    err = np.abs(fn["infil"] - f2["infil"])[unsat_branch]
    assert np.all(err <= 1e-12 * np.abs(f2["infil"])[unsat_branch])
    assert np.any(fn["infil"][sel] != fp["infil"][sel])


def test_noah_fcr_scales_infiltration_before_recharge(c0):
    m = c0
    rng = np.random.default_rng(34)
    inp = _inputs(m, rng)
    nz = _noah_inputs(m, rng)
    ne = m.nelem
    for y in _states(m, rng, 3)[1:]:
        _, _, f1 = _run(m, y, inp, dict(nz, fcr=np.ones(ne)))
        _, _, ff = _run(m, y, inp, nz)
        assert np.array_equal(ff["infil"], f1["infil"] * nz["fcr"])
        g = np.maximum(y[2 * ne:3 * ne], 0.0)
        deep = g > m.a["depth"] - m.a["dinf"]
        # the saturated branch of Recharge passes the frozen-soil infiltration on
        assert np.array_equal(ff["rechg"][deep], ff["infil"][deep])
        assert np.array_equal(ff["rechg"][~deep], f1["rechg"][~deep])


def test_noah_switch_off_restores_plain_build(c0):
    m = c0
    rng = np.random.default_rng(35)
    inp = _inputs(m, rng)
    nz = _noah_inputs(m, rng)
    y = _states(m, rng, 2)[1]
    _, d_plain, _ = _run(m, y, inp)
    o = O.OracleRHS(m)
    o.set_step_inputs(**inp)
    o.set_noah(**nz)
    _, d_noah = o.ode(y)
    o.set_noah(None)
    # the carried river-edge flux of the first call is what the second reads
    o2 = O.OracleRHS(m)
    o2.set_step_inputs(**inp)
    o2.ode(y)
    _, want = o2.ode(y)
    _, got = o.ode(y)
    assert not np.array_equal(d_noah, d_plain)
    assert np.array_equal(got, want)
