"""Model output on the device (pihm_output.h) against a restatement of the reference's
UpdPrintVar / PrintData / PrintNow (src/print.c:190-211, 545-603, 814-856) in this test:
the averaged values, their caller order, the .dat records and the .txt lines must be
byte-identical, for state, land-surface and flux variables at hourly, daily and
every-N-seconds intervals; the restart file has the layout of src/print.c:250-300; and a
producer-driven run writes one record per print time."""
import calendar
import ctypes as C
import os
import struct
import time

import numpy as np
import pytest

from pihm_amd import _lib
from pihm_amd.inputs import str_time
from pihm_amd.lsm import HipLSM, synthetic_meteo, synthetic_params
from pihm_amd.output import DAILY, HOURLY, HYDROL_STEP, KIND, LS_STEP, HipOutput, read_dat
from pihm_amd.rhs import HipRHS

pytestmark = pytest.mark.gpu
T0 = str_time("2009-01-01 00:00")


def _print_now(intvl, lapse, t):
    tm = time.gmtime(t)
    if intvl == HOURLY:
        return tm.tm_min == 0
    if intvl == DAILY:
        return tm.tm_hour == 0 and tm.tm_min == 0
    return intvl > 0 and int(lapse) % intvl == 0


class Checker:
    """The reference's print bookkeeping on host arrays (caller order)."""

    def __init__(self, variables):
        self.v = [dict(kind=k, intvl=iv, sum=None, n=0, dat=b"", txt="") for k, iv in variables]

    def update(self, step, values):
        for x in self.v:
            upd = LS_STEP if x["kind"] in {KIND[e] for e in ("snow", "is", "ec", "ett", "edir")} else HYDROL_STEP
            if upd != step:
                continue
            val = values[x["kind"]]
            # the sums start at 0.0 (calloc), so 0.0 + (-0.0) leaves +0.0
            x["sum"] = (np.zeros_like(val) if x["sum"] is None else x["sum"]) + val
            x["n"] += 1

    def print(self, t, lapse):
        stamp = time.strftime("%Y-%m-%d %H:%M", time.gmtime(t))
        for x in self.v:
            if not _print_now(x["intvl"], lapse, t):
                continue
            avg = x["sum"] / float(x["n"]) if x["n"] > 0 else x["sum"]
            x["dat"] += struct.pack("<d", float(t)) + avg.astype("<f8").tobytes()
            x["txt"] += '"' + stamp + '"' + "".join("\t%f" % v for v in avg) + "\n"
            x["sum"] = np.zeros_like(avg)
            x["n"] = 0


def _mass_balance(a, prev, y, sub, edir, ett, dt):
    """MassBalance (src/update.c:107-160) with EtExtract's partition (src/hydrol.c:33-97)
    at the state y, operation for operation."""
    ne = len(a["area"])
    depth = a["depth"]
    s0 = prev[2 * ne:3 * ne] + prev[ne:2 * ne]
    s0 = np.where(s0 > depth, depth, s0)
    s0 = np.where(s0 < 0.0, 0.0, s0)
    s1 = y[2 * ne:3 * ne] + y[ne:2 * ne]
    s1 = np.where(s1 > depth, depth, s1)
    s1 = np.where(s1 < 0.0, 0.0, s1)
    sr = np.zeros(ne)
    for j in range(3):
        sr = sr + sub[:, j] / a["area"]
    sc = np.maximum(y[:ne], 0.0)
    dep = 1e-4
    surfh = np.where(sc <= 0.5 * dep, np.sqrt(2.0 * dep * sc), dep + (sc - 0.5 * dep))
    g = np.maximum(y[2 * ne:3 * ne], 0.0)
    to_surf = surfh >= dep
    to_gw = ~to_surf & (g > depth - a["dinf"])
    eu = np.where(~to_surf & ~to_gw, edir, 0.0)
    eg = np.where(to_gw, edir, 0.0)
    tg = np.where(g > depth - a["rzd"], ett, 0.0)
    tu = np.where(g > depth - a["rzd"], 0.0, ett)
    inf = (s1 - s0) * a["porosity"] / dt + sr + eu + eg + tu + tg
    return np.where(inf < 0.0, 0.0, inf)


def test_output_matches_print_bookkeeping(c0, tmp_path):
    m = c0
    ne, nr = m.nelem, m.nriver
    print(f"[out] c0 ne={ne}", flush=True)
    rng = np.random.default_rng(3)
    variables = [(KIND["surf"], HOURLY), (KIND["gw"], DAILY), (KIND["stage"], 1800), (KIND["snow"], HOURLY),
                 (KIND["is"], HOURLY), (KIND["edir"], HOURLY), (KIND["infil"], HOURLY),
                 (KIND["rivflx1"], HOURLY), (KIND["subflx0"], HOURLY), (KIND["surfflx2"], 1200)]
    h = HipRHS(m)
    ref = HipRHS(m, record=True)                  # the same flux evaluation, for the checker
    p = synthetic_params(ne, 2, seed=1)
    for k in ("depth", "porosity", "rzd", "zmax"):
        p[k] = m.a[k].copy()
    p["smcmin"] = np.minimum(p["smcmin"], 0.5 * p["porosity"])
    meteo = synthetic_meteo(2, 40, t0=T0 - 3600.0, seed=2)
    lsm = HipLSM(h, p, meteo)
    lsm.set_veg(rng.uniform(0.5, 3, ne), rng.uniform(0.05, 0.3, ne), 1e-7)
    lsm.set_state(np.zeros(ne), rng.uniform(0, 0.01, ne))
    out = HipOutput(h, str(tmp_path / "ex"), variables, ascii=True, lsm=lsm)
    chk = Checker(variables)
    zero_carried = np.zeros(3 * ne)
    dt, et = 600.0, 1800.0
    y = m.initial_state()
    out.start(y)
    prev = y.copy()
    a = m.a
    for k in range(30):                           # five simulated hours
        t = T0 + dt * k
        if k % int(et / dt) == 0:
            assert lsm.step(t, et, y) == 0
            out.update(LS_STEP)
            cmc, sneqv = lsm.get_state()
            f = lsm.fluxes()
            chk.update(LS_STEP, {KIND["snow"]: sneqv, KIND["is"]: cmc, KIND["edir"]: f["edir"]})
        y = y.copy()
        y[:ne] = np.abs(y[:ne] + rng.normal(0, 1e-4, ne))
        y[3 * ne:3 * ne + nr] += rng.uniform(0, 0.01, nr)
        out.update(HYDROL_STEP, y)
        pc, ed, et_ = np.zeros(ne), np.zeros(ne), np.zeros(ne)
        h.L.pihm_ctx_get_step_inputs(h.ctx, *(a.ctypes.data_as(_lib.c_double_p) for a in (pc, ed, et_)),
                                     None, None, None)
        ref.set_step_inputs(pcpdrp=pc, edir=ed, ett=et_)
        ref.set_river_edge_ovlflow(zero_carried)
        ref.ode(y)
        fl = ref.fluxes()
        rf = fl["rivflow"].reshape(nr, 11)
        infil = _mass_balance(a, prev, y, fl["subsurf"].reshape(ne, 3), ed, et_, m.stepsize)
        prev = y.copy()
        chk.update(HYDROL_STEP, {KIND["surf"]: y[:ne], KIND["gw"]: y[2 * ne:3 * ne],
                                 KIND["stage"]: y[3 * ne:3 * ne + nr], KIND["infil"]: infil,
                                 KIND["rivflx1"]: rf[:, 1], KIND["subflx0"]: fl["subsurf"].reshape(ne, 3)[:, 0],
                                 KIND["surfflx2"]: fl["ovlflow"].reshape(ne, 3)[:, 2]})
        tn = t + dt
        out.print(tn, tn - T0)
        chk.print(tn, tn - T0)
    out.close()
    for x in chk.v:
        from pihm_amd.output import EXTS
        name = str(tmp_path / ("ex." + EXTS[x["kind"]]))
        got = open(name + ".dat", "rb").read()
        assert got == x["dat"], EXTS[x["kind"]]
        assert open(name + ".txt").read() == x["txt"], EXTS[x["kind"]]
        assert len(got) > 0 or x["intvl"] == DAILY
    times, vals = read_dat(str(tmp_path / "ex.surf.dat"), ne)
    assert np.array_equal(times, T0 + 3600.0 * np.arange(1, 6))
    lsm.close()
    h.close()
    ref.close()


def test_restart_file_layout(c0, tmp_path):
    m = c0
    ne, nr = m.nelem, m.nriver
    h = HipRHS(m)
    p = synthetic_params(ne, 1, seed=5)
    lsm = HipLSM(h, p, synthetic_meteo(1, 4, t0=T0, seed=5))
    rng = np.random.default_rng(5)
    cmc, sneqv = rng.uniform(0, 1e-4, ne), rng.uniform(0, 0.02, ne)
    lsm.set_state(cmc, sneqv)
    out = HipOutput(h, str(tmp_path / "x"), [], lsm=lsm)
    y = rng.normal(size=m.neq)
    path = str(tmp_path / "x.ic")
    out.write_ic(path, y)
    raw = np.fromfile(path, dtype="<f8")
    assert raw.size == 5 * ne + 2 * nr
    e = raw[:5 * ne].reshape(ne, 5)
    assert np.array_equal(e[:, 0], cmc) and np.array_equal(e[:, 1], sneqv)
    assert np.array_equal(e[:, 2], y[:ne]) and np.array_equal(e[:, 3], y[ne:2 * ne])
    assert np.array_equal(e[:, 4], y[2 * ne:3 * ne])
    r = raw[5 * ne:].reshape(nr, 2)
    assert np.array_equal(r[:, 0], y[3 * ne:3 * ne + nr]) and np.array_equal(r[:, 1], y[3 * ne + nr:])
    out.close()
    lsm.close()
    h.close()


def test_producers_run_writes_outputs(c0, tmp_path):
    from pihm_amd.forcing import run_config
    from pihm_amd.output import read_print_ctrl
    from conftest import GOLDEN
    m = c0
    ne = m.nelem
    print(f"[out] producers run ne={ne}", flush=True)
    variables = read_print_ctrl(os.path.join(GOLDEN, "example.para"))
    h = HipRHS(m, nan_check=False)
    p = synthetic_params(ne, 1, seed=6)
    for k in ("depth", "porosity", "rzd", "zmax"):
        p[k] = m.a[k].copy()
    p["smcmin"] = np.minimum(p["smcmin"], 0.5 * p["porosity"])
    lsm = HipLSM(h, p, synthetic_meteo(1, 12, t0=T0 - 3600.0, seed=6))
    lsm.set_veg(np.full(ne, 2.0), np.full(ne, 0.1), 1e-7)
    lsm.set_state(np.zeros(ne), np.zeros(ne))
    out = HipOutput(h, str(tmp_path / "run"), variables, ascii=False, lsm=lsm)
    prod = (C.c_void_p * 3)(lsm.h, None, out.h)
    y = np.ascontiguousarray(m.initial_state())
    st = _lib.RunStats()
    cfg = run_config(T0 + 3 * 3600.0, t_start=T0)
    rc = h.L.pihm_run_gpu_producers(h.ctx, C.cast(prod, C.c_void_p), C.byref(cfg),
                                    y.ctypes.data_as(_lib.c_double_p), C.byref(st))
    assert rc == 0 and st.model_steps == 180
    out.close()
    from pihm_amd.output import EXTS, RIVER_KINDS
    for k, iv in variables:
        n = m.nriver if k in RIVER_KINDS else ne
        times, vals = read_dat(str(tmp_path / f"run.{EXTS[k]}.dat"), n)
        assert np.array_equal(times, T0 + 3600.0 * np.arange(1, 4)), EXTS[k]
        assert np.all(np.isfinite(vals)), EXTS[k]
    # the last hour's surface average lies between the hour's extremes of the final state
    times, surf = read_dat(str(tmp_path / "run.surf.dat"), ne)
    assert np.all(surf[-1] >= 0.0)
    lsm.close()
    h.close()
