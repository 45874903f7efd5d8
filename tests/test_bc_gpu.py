"""Parity of the device boundary-condition series (pihm_bc_*) with the oracle's
ApplyElemBc/ApplyRiverBc, their hand-over to the RHS, and runs that apply them every model
step.

The interpolation and the copy are the same operations in the same order on both sides
(no transcendental functions), so the boundary values agree bit for bit; the RHS after
them is held to the RHS tolerance (1e-12 of the assembled terms) and the free runs to the
solver tolerance (WRMS of the difference below 1).
"""
import ctypes as C

import numpy as np
import pytest

import oracle as O
from conftest import assert_rhs_close
from pihm_amd import _lib
from pihm_amd.bc import HipBC, code_array, ts_struct
from pihm_amd.rhs import HipRHS
from test_rhs_gpu import _variant
from test_oracle_pins import _inputs, _states

pytestmark = pytest.mark.gpu
T0 = 1.2e9


def _series(nts, nrec, base, amp, seed, dt=3600.0):
    rng = np.random.default_rng(seed)
    ft = np.tile(T0 + dt * np.arange(nrec, dtype=np.float64), (nts, 1))
    ft[:, 1:] += np.cumsum(rng.integers(0, 600, (nts, nrec - 1)), axis=1)   # uneven spacing
    ph = rng.uniform(0, 2 * np.pi, (nts, 1))
    v = np.asarray(base, np.float64).reshape(-1, 1) + np.asarray(amp, np.float64).reshape(-1, 1) * np.sin(
        ft / 20000.0 + ph)
    return ft, v


def _bc_setup(m, seed=0):
    """Codes and series for a mesh whose boundary edges carry Dirichlet (+) and Neumann (-)
    types: one head series per Dirichlet edge, two shared flux series, one head series
    per Dirichlet river outlet and one shared flux series."""
    a = m.a
    bt = a["bc_type"].reshape(m.nelem, 3)
    code = np.zeros_like(bt)
    dir_edges = np.argwhere(bt > 0)
    code[bt > 0] = 1 + np.arange(len(dir_edges))
    neu = np.argwhere(bt < 0)
    code[bt < 0] = -(len(dir_edges) + 1 + np.arange(len(neu)) % 2)
    head = (a["zmin"] + 0.7 * a["depth"])[dir_edges[:, 0]]
    ets_d = _series(len(dir_edges), 40, head, 0.2, seed)
    ets_n = _series(2, 40, [2e-4, 1e-4], [1e-4, 5e-5], seed + 1)
    ets_n = (ets_n[0] * 0 + ets_d[0][0], ets_n[1])      # same times for every series
    ets = (np.vstack([ets_d[0], ets_n[0]]), np.vstack([ets_d[1], ets_n[1]]))
    rbt = a["riverbc_type"]
    rcode = np.zeros_like(rbt)
    rd = np.nonzero(rbt > 0)[0]
    rcode[rd] = 1 + np.arange(len(rd))
    rcode[rbt < 0] = -(len(rd) + 1)
    rts_d = _series(len(rd), 30, a["zbed"][rd] - 0.5, 0.1, seed + 2, dt=5400.0)
    rts_n = _series(1, 30, [0.05], [0.02], seed + 3, dt=5400.0)
    rts = (np.vstack([rts_d[0], rts_d[0][:1] * 0 + rts_d[0][0]]), np.vstack([rts_d[1], rts_n[1]]))
    return code, ets, rcode, rts


def test_bc_apply_matches_oracle_bitwise(c0):
    m = c0
    ne, nr = m.nelem, m.nriver
    print(f"[bc] c0 ne={ne} nr={nr}", flush=True)
    rng = np.random.default_rng(1)
    code = rng.integers(-5, 6, (ne, 3)).astype(np.int32)
    code[rng.random((ne, 3)) < 0.5] = 0
    rcode = rng.integers(-3, 4, nr).astype(np.int32)
    ets = _series(5, 50, rng.uniform(100, 200, 5), 3.0, 4)
    ets = (np.tile(ets[0][:1], (5, 1)), ets[1])
    rts = _series(3, 20, [1.0, 2.0, 3.0], 0.5, 5, dt=7200.0)
    h = HipRHS(m)
    e0, r0 = rng.normal(size=3 * ne), rng.normal(size=nr)
    h.set_step_inputs(elem_bc=e0, river_bc=r0)
    b = HipBC(h, code, ets, rcode, rts)
    eb, rb = e0.reshape(ne, 3).copy(), r0.copy()
    t_hi = min(ets[0][:, -1].min(), rts[0][:, -1].min())
    times = list(ets[0][0, :12]) + list(np.linspace(T0, t_hi, 37)) + [t_hi]
    for t in times:
        assert b.apply(t) == 0
        ge, gr = b.values()
        rc, eb, rb = O.bc_apply(ne, nr, t, code, ets, rcode, rts, eb, rb)
        assert rc == 0
        assert np.array_equal(ge, eb), t
        assert np.array_equal(gr, rb), t
    # edges with code 0 kept the values set before
    assert np.array_equal(ge[code == 0], e0.reshape(ne, 3)[code == 0])
    assert b.apply(T0 - 1.0) == _lib.PIHM_ERR_ARG
    assert b.apply(t_hi + 1.0) == _lib.PIHM_ERR_ARG
    b.close()
    h.close()


def test_bc_feeds_rhs():
    m = _variant("bc")
    print(f"[bc->rhs] ne={m.nelem} nr={m.nriver}", flush=True)
    code, ets, rcode, rts = _bc_setup(m)
    rng = np.random.default_rng(2)
    h, o = HipRHS(m), O.OracleRHS(m)
    inp = _inputs(m, rng)
    h.set_step_inputs(**inp)
    b = HipBC(h, code, ets, rcode, rts)
    for t in (T0, T0 + 5000.0, T0 + 20 * 3600.0):
        assert b.apply(t) == 0
        rc, eb, rb = O.bc_apply(m.nelem, m.nriver, t, code, ets, rcode, rts)
        o.set_step_inputs(elem_bc=eb.ravel(), river_bc=rb, **inp)
        y = _states(m, rng, 2)[1]
        rc_h, dh = h.ode(y, 0.0)
        rc_o, do, mag = o.ode(y, 0.0, with_mag=True)
        assert rc_h == rc_o
        assert_rhs_close(dh, do, mag, 1e-12, f"ydot after the boundary series at t={t}")
    b.close()
    h.close()


def test_bc_argument_errors(c0):
    m = c0
    L = _lib.lib()
    h = HipRHS(m)
    keep = []
    ets = _series(2, 10, [1.0, 2.0], 0.1, 7)
    es = ts_struct(ets, keep)
    out = C.c_void_p()
    bad = np.zeros((m.nelem, 3), np.int32)
    bad[0, 1] = -3                                                   # names series 3 of 2
    assert L.pihm_bc_create(h.ctx, code_array(bad, 3 * m.nelem, keep), C.byref(es), None, None,
                            C.byref(out)) == _lib.PIHM_ERR_ARG
    ft = ets[0].copy()
    ft[1, 3] = ft[1, 2]
    es2 = ts_struct((ft, ets[1]), keep)
    ok = np.zeros((m.nelem, 3), np.int32)
    assert L.pihm_bc_create(h.ctx, code_array(ok, 3 * m.nelem, keep), C.byref(es2), None, None,
                            C.byref(out)) == _lib.PIHM_ERR_ARG
    # no series at all: apply is a no-op at any time
    b = HipBC(h)
    assert b.apply(-1e30) == 0
    b.close()
    h.close()


def test_run_with_bc_and_lsm_matches_host_run():
    from pihm_amd.forcing import run_config
    from pihm_amd.lsm import run_gpu_lsm, synthetic_meteo, synthetic_params
    m = _variant("bc")
    ne = m.nelem
    print(f"[bc run] ne={ne} nr={m.nriver} 3 h", flush=True)
    code, ets, rcode, rts = _bc_setup(m, seed=9)
    p = synthetic_params(ne, 2, seed=4)
    for k in ("depth", "porosity", "rzd", "zmax"):
        p[k] = m.a[k].copy()
    p["smcmin"] = np.minimum(p["smcmin"], 0.5 * p["porosity"])
    meteo = synthetic_meteo(2, 30, t0=T0 - 3600.0, seed=4)
    rng = np.random.default_rng(4)
    veg = (rng.uniform(0.5, 3.0, ne), rng.uniform(0.05, 0.3, ne), 1e-7)
    y0 = m.initial_state()
    cfg = run_config(T0 + 3 * 3600.0, t_start=T0)
    bc = (code, ets, rcode, rts)
    rc_g, y_g, st_g, cmc_g, sn_g = run_gpu_lsm(m, p, meteo, veg, cfg, y0, bc=bc)
    assert rc_g == 0, st_g.as_dict()
    o = O.OracleLsmRun(m, p, meteo, veg[0], veg[1], veg[2], np.zeros(ne), np.zeros(ne), cfg.etstep, bc=bc)
    rc_c, y_c, st_c = o.run(y0, cfg)
    assert rc_c == 0
    print("gpu", st_g.as_dict(), "\ncpu", st_c.as_dict(), flush=True)
    w = 1.0 / (1e-3 * np.abs(y_c) + 1e-4)
    assert float(np.sqrt(np.mean(((y_g - y_c) * w) ** 2))) < 1.0
    # the boundary series changed the answer: the same run without them differs
    rc_n, y_n, _, _, _ = run_gpu_lsm(m, p, meteo, veg, cfg, y0)
    assert rc_n == 0 and not np.allclose(y_n, y_g, rtol=1e-6, atol=0)
