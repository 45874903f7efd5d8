"""Reader for the reference's input formats, on the reference's own example watershed
(benchmark/input/example: 535 triangles, 20 river segments; files copied unchanged into
tests/golden/example as fixtures), and the oracle's invariants on that real mesh."""
import os

import numpy as np

import oracle as O
from conftest import GOLDEN
from test_oracle_pins import _inputs, _states


def test_example_sizes_and_tables(example):
    m = example
    assert (m.nelem, m.nriver, m.neq) == (535, 20, 1645)
    a = m.a
    # counter-clockwise triangles, soil depth = zmax - zmin (src/init_soil.c:28)
    assert np.all(a["area"] > 0.0)
    assert np.array_equal(a["depth"], a["zmax"] - a["zmin"])
    # calibration multipliers of example.calib applied once (KSATH 5.0 on table values)
    assert np.isclose(a["ksath"].min(), 5.0 * 9.802e-6) and np.isclose(a["ksath"].max(), 5.0 * 6.962e-5)
    assert np.allclose(a["rrough"], 5.0 * 0.04)
    # every shape has OINT 1: rectangular channels of width CWID
    assert np.all(a["intrpl_ord"] == 1) and np.allclose(a["width"], 1.5)
    assert m.para["stepsize"] == 60.0 and m.para["reltol"] == 1e-3 and m.para["abstol"] == 1e-4


def test_example_river_tagging(example):
    a = example.a
    for r in range(example.nriver):
        for e in (a["leftele"][r] - 1, a["rightele"][r] - 1):
            assert (a["nabr"][e] == -(r + 1)).sum() == 1
    # one outlet, zero-depth-gradient (DOWN = -3)
    assert (a["down"] < 0).sum() == 1 and a["down"][a["down"] < 0][0] == -3


def test_example_oracle_invariants(example):
    """Interior fluxes antisymmetric bit for bit and global conservation on real data."""
    m = example
    rng = np.random.default_rng(12)
    r = O.OracleRHS(m)
    inp = _inputs(m, rng)
    r.set_step_inputs(**inp)
    a = m.a
    ne, nr = m.nelem, m.nriver
    for y in _states(m, rng, 3):
        rc, dy, mag = r.ode(y, with_mag=True)
        assert rc == 0
        ovl = r.field("ovlflow").reshape(-1, 3)
        sub = r.field("subsurf").reshape(-1, 3)
        nb = a["nabr"]
        for i in range(ne):
            for j in range(3):
                k = nb[i, j]
                if k > 0:
                    jj = int(np.nonzero(nb[k - 1] == i + 1)[0][0])
                    assert ovl[i, j] == -ovl[k - 1, jj] and sub[i, j] == -sub[k - 1, jj]
        rate = (np.sum(a["area"] * (dy[:ne] + a["porosity"] * (dy[ne:2 * ne] + dy[2 * ne:3 * ne])))
                + np.sum(a["rarea"] * (dy[3 * ne:3 * ne + nr] + a["rporosity"] * dy[3 * ne + nr:])))
        rf = r.field("rivflow").reshape(nr, 11)
        bnd = nb == 0
        src = (np.sum(a["area"] * (inp["pcpdrp"] - inp["edir"] - inp["ett"]))
               - np.sum(sub[bnd]) - rf[a["down"] < 0, 1].sum())
        scale = np.sum(a["area"] * (inp["pcpdrp"] + inp["edir"] + inp["ett"]) + a["area"] * mag[:ne])
        assert abs(rate - src) <= 1e-11 * scale


def test_print_controls_of_the_example_para():
    """ReadPrtCtrl's keywords of the example's .para (src/read_para.c:111-180): every
    control HOURLY, SUBFLX and SURFFLX one file per edge."""
    from pihm_amd.output import EXTS, HOURLY, read_print_ctrl
    v = read_print_ctrl(os.path.join(GOLDEN, "example.para"))
    names = [EXTS[k] for k, _ in v]
    assert len(v) == 29 and all(iv == HOURLY for _, iv in v)
    assert names[:5] == ["surf", "unsat", "gw", "stage", "rivgw"]
    assert [f"subflx{k}" for k in range(3)] == [n for n in names if n.startswith("subflx")]
    assert "rivflx10" in names and "is" in names
