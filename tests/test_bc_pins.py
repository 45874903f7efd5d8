"""Pins of the oracle's boundary-condition series (oracle_bc_apply: ApplyElemBc and
ApplyRiverBc, src/forcing.c:127-244, over IntrplForc, src/forcing.c:482-530) and of the
readers of the reference's .bc file and .riv river-series blocks: numpy's piecewise-linear
interpolation, the sign convention of the codes (k > 0 head from series k, k < 0 flux
from series -k, 0 untouched) and the out-of-range exit."""
import os
import textwrap

import numpy as np

import oracle as O
from pihm_amd import inputs


def _series(nts, nrec, seed):
    rng = np.random.default_rng(seed)
    ft = np.tile(1000.0 + 600.0 * np.arange(nrec), (nts, 1))
    return ft, rng.normal(size=(nts, nrec))


def test_codes_and_interpolation():
    ne, nr = 50, 7
    rng = np.random.default_rng(0)
    ets, rts = _series(3, 12, 1), _series(2, 12, 2)
    code = rng.integers(-3, 4, (ne, 3)).astype(np.int32)
    rcode = rng.integers(-2, 3, nr).astype(np.int32)
    e0, r0 = rng.normal(size=(ne, 3)), rng.normal(size=nr)
    t = 1000.0 + 600.0 * 4.25
    rc, eb, rb = O.bc_apply(ne, nr, t, code, ets, rcode, rts, e0, r0)
    assert rc == 0
    ev = [np.interp(t, ets[0][k], ets[1][k]) for k in range(3)]
    rv = [np.interp(t, rts[0][k], rts[1][k]) for k in range(2)]
    for i in range(ne):
        for j in range(3):
            c = code[i, j]
            want = e0[i, j] if c == 0 else ev[abs(c) - 1]
            assert abs(eb[i, j] - want) <= 4e-16 * max(1.0, abs(want))
    for r in range(nr):
        c = rcode[r]
        want = r0[r] if c == 0 else rv[abs(c) - 1]
        assert abs(rb[r] - want) <= 4e-16 * max(1.0, abs(want))


def test_out_of_range_and_absent_sets():
    ets = _series(2, 5, 3)
    code = np.ones((4, 3), np.int32)
    assert O.bc_apply(4, 0, 999.0, code, ets)[0] == 1
    assert O.bc_apply(4, 0, 1000.0 + 600.0 * 4 + 1, code, ets)[0] == 1
    rc, eb, _ = O.bc_apply(4, 0, 1000.0 + 600.0 * 4, code, ets)
    assert rc == 0 and np.allclose(eb, ets[1][0, -1], rtol=2.3e-16, atol=0)
    rc, eb, rb = O.bc_apply(4, 3, 5.0)            # nothing to apply
    assert rc == 0 and not eb.any() and not rb.any()


def test_read_bc_example():
    ft, v = inputs.read_bc(os.path.join(os.path.dirname(__file__), "golden", "example", "example.bc"))
    assert ft.shape == v.shape == (1, 2)
    assert ft[0, 0] == inputs.str_time("2008-01-01 00:00") and v[0].tolist() == [270.0, 270.0]


def test_read_riv_series(tmp_path):
    src = os.path.join(os.path.dirname(__file__), "golden", "example", "example.riv")
    txt = open(src).read()
    i = txt.index("\nBC")
    series = textwrap.dedent("""
        BC      2
        RIV_TS  1   # Dirichlet
        TIME                HEAD
        TS                  m
        2009-01-01 00:00    1.5
        2009-01-02 00:00    2.5
        RIV_TS  2   # Neumann
        TIME                FLUX
        TS                  m3/s
        2009-01-01 00:00    -0.1
        2009-01-02 00:00    -0.3
        RES     0
        """)
    p = tmp_path / "x.riv"
    p.write_text(txt[:i] + series)
    r = inputs.read_riv(str(p))
    ft, v = r["bc_series"]
    assert ft.shape == (2, 2) and v.tolist() == [[1.5, 2.5], [-0.1, -0.3]]
    assert ft[1, 1] - ft[1, 0] == 86400.0
