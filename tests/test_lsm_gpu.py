"""Parity of the device land-surface step (pihm_lsm_*) with the oracle's IntrplForc and
IntcpSnowEt, and its hand-over of pcpdrp/edir/ett to the RHS.

Forcing: the reference example watershed's own station records (the first two weeks of
its simulation, tests/golden/example/example.meteo) and two copies of them shifted to
+3 K and +8 K, so snowfall, mixed precipitation, melt and rain all occur.

Tolerance, per call (each step starts both sides from the device's carried state):
|device - oracle| <= 1e-12 * scale, where scale is the magnitude of the terms the value is
assembled from (the canopy store: store + rain caught + evaporation + drip over the step;
the snow store likewise; pcpdrp: rain + drip + melt; ett: ett + ec, which bounds
shdfac * etp where 1 - wet^cfactr cancels).  exp, log, pow and cos are the device
library's on one side and glibc's on the other; everything else is the same operations
in the same order.
"""
import ctypes as C
import os

import numpy as np
import pytest

import oracle as O
from conftest import GOLDEN, assert_rhs_close
from pihm_amd import _lib
from pihm_amd.inputs import read_meteo, str_time
from pihm_amd.lsm import HipLSM, lsm_structs, synthetic_params
from pihm_amd.rhs import HipRHS

pytestmark = pytest.mark.gpu
RTOL = 1e-12
T0 = str_time("2009-01-01 00:00")


def _meteo():
    ft, data, _ = read_meteo(os.path.join(GOLDEN, "example.meteo"))
    ft3 = np.repeat(ft, 3, axis=0)
    d3 = np.repeat(data, 3, axis=0)
    d3[1, :, 1] += 8.0
    d3[2, :, 1] += 3.0
    d3[2, :, 0] *= 2.0
    return ft3, d3


def _setup(m, seed):
    rng = np.random.default_rng(seed)
    ne = m.nelem
    p = synthetic_params(ne, 3, seed=seed)
    for k in ("depth", "porosity", "rzd", "zmax"):     # the mesh's own soil and topography
        p[k] = m.a[k].copy()
    p["smcmin"] = np.minimum(p["smcmin"], 0.5 * p["porosity"])
    meteo = _meteo()
    lai = rng.uniform(0.0, 4.0, ne)
    lai[rng.random(ne) < 0.1] = 0.0
    z0 = rng.uniform(0.01, 0.5, ne)
    meltf = 1.0e-7
    intcp = p["cmcfactr"] * lai * p["shdfac"]
    cmc0 = intcp * rng.uniform(0.0, 1.0, ne)
    sneqv0 = rng.uniform(0.0, 0.02, ne)
    sneqv0[rng.random(ne) < 0.3] = 0.0
    return p, meteo, lai, z0, meltf, cmc0, sneqv0


def _state(m, p, rng):
    """A soil state: surface, unsat, gw per element, then the river states."""
    ne = m.nelem
    y = np.zeros(m.neq)
    y[:ne] = rng.uniform(0.0, 0.01, ne)
    gw = p["depth"] * rng.uniform(0.05, 1.0, ne)
    unsat = (p["depth"] - gw) * rng.uniform(-0.05, 1.2, ne)
    y[ne:2 * ne] = unsat
    y[2 * ne:3 * ne] = gw
    y[3 * ne:] = rng.uniform(0.0, 0.5, m.neq - 3 * ne)
    return y


def _check_call(dev, ref, cmc_in, sneqv_in, cmc_d, sneqv_d, cmc_o, sneqv_o, prcp, dt, what):
    ec, drip = ref["ec"], ref["drip"]
    scale = dict(
        pcpdrp=np.abs(ref["pcpdrp"]) + prcp + drip + sneqv_in / dt,
        edir=np.abs(ref["edir"]),
        ett=np.abs(ref["ett"]) + np.abs(ec),
        ec=np.abs(ec) + prcp,
        drip=np.abs(drip) + prcp + cmc_in / dt,
    )
    worst = {}
    for k, s in scale.items():
        err = np.abs(dev[k] - ref[k])
        bad = ~(err <= RTOL * s)
        worst[k] = float(np.max(np.where(s > 0, err / np.where(s > 0, s, 1.0), 0.0)))
        assert not bad.any(), f"{what} {k}: {np.nonzero(bad)[0][:5]} dev {dev[k][bad][:3]} ref {ref[k][bad][:3]}"
    cs = cmc_in + prcp * dt + (ec + drip) * dt
    ss = sneqv_in + prcp * dt
    assert np.all(np.abs(cmc_d - cmc_o) <= RTOL * cs), what
    assert np.all(np.abs(sneqv_d - sneqv_o) <= RTOL * ss), what
    return worst


@pytest.mark.parametrize("which,dt,nsteps", [("c0", 900.0, 96), ("example", 60.0, 40)])
def test_lsm_parity_per_call(which, dt, nsteps, request):
    m = request.getfixturevalue(which)
    print(f"[lsm] {which} ne={m.nelem} dt={dt} steps={nsteps}", flush=True)
    rng = np.random.default_rng(5)
    p, meteo, lai, z0, meltf, cmc0, sneqv0 = _setup(m, 3)
    h = HipRHS(m)
    lsm = HipLSM(h, p, meteo)
    lsm.set_veg(lai, z0, meltf)
    lsm.set_state(cmc0, sneqv0)
    ne = m.nelem
    worst = {}
    t = T0
    for step in range(nsteps):
        y = _state(m, p, rng)
        cmc_in, sneqv_in = lsm.get_state()
        assert lsm.step(t, dt, y) == 0
        cmc_d, sneqv_d = lsm.get_state()
        dev = lsm.fluxes()
        rc, met = O.meteo_interp(meteo, t)
        assert rc == 0
        cmc_o, sneqv_o, ref = O.intcp_snow_et(p, met, lai, z0, meltf, y[ne:2 * ne], y[2 * ne:3 * ne], dt,
                                              cmc_in, sneqv_in)
        prcp = met[0, p["station"]] / 1000.0
        w = _check_call(dev, ref, cmc_in, sneqv_in, cmc_d, sneqv_d, cmc_o, sneqv_o, prcp, dt,
                        f"step {step} t={t}")
        for k, v in w.items():
            worst[k] = max(worst.get(k, 0.0), v)
        t += dt + (7.0 if step % 5 == 2 else 0.0)     # times between records, not only on them
    print("[lsm] worst err/scale", {k: f"{v:.2e}" for k, v in worst.items()}, flush=True)
    cmc, sneqv = lsm.get_state()
    assert (sneqv > 0).any() and (sneqv == 0).any()      # snow came and went somewhere
    lsm.close()
    h.close()


def test_lsm_feeds_rhs(c0):
    """After a device LSM step the RHS reads the step's pcpdrp/edir/ett: the HIP RHS equals
    the oracle RHS given the same three fields."""
    m = c0
    print(f"[lsm->rhs] ne={m.nelem}", flush=True)
    rng = np.random.default_rng(9)
    p, meteo, lai, z0, meltf, cmc0, sneqv0 = _setup(m, 4)
    h = HipRHS(m)
    o = O.OracleRHS(m)
    lsm = HipLSM(h, p, meteo)
    lsm.set_veg(lai, z0, meltf)
    lsm.set_state(cmc0, sneqv0)
    y = _state(m, p, rng)
    assert lsm.step(T0 + 12 * 3600.0, 900.0, y) == 0
    f = lsm.fluxes()
    assert f["pcpdrp"].max() > 0 and f["edir"].max() > 0
    o.set_step_inputs(pcpdrp=f["pcpdrp"], edir=f["edir"], ett=f["ett"])
    rc_h, dh = h.ode(y, 0.0)
    rc_o, do, mag = o.ode(y, 0.0, with_mag=True)
    assert rc_h == rc_o
    assert_rhs_close(dh, do, mag, 1e-12, "ydot after the LSM step")
    lsm.close()
    h.close()


def test_lsm_argument_errors(c0):
    m = c0
    L = _lib.lib()
    p, meteo, lai, z0, meltf, cmc0, sneqv0 = _setup(m, 6)
    h = HipRHS(m)
    lsm = HipLSM(h, p, meteo)
    # no step before the vegetation inputs are set
    assert lsm.step(T0, 900.0) == _lib.PIHM_ERR_ARG
    lsm.set_veg(lai, z0, meltf)
    ft = meteo[0]
    assert lsm.step(ft[0, 0] - 1.0, 900.0) == _lib.PIHM_ERR_ARG      # before the first record
    assert lsm.step(ft[0, -1] + 1.0, 900.0) == _lib.PIHM_ERR_ARG     # after the last record
    assert lsm.step(T0, 0.0) == _lib.PIHM_ERR_ARG
    y = _state(m, p, np.random.default_rng(1))
    assert lsm.step(ft[0, -1], 900.0, y) == 0                        # the last record itself
    lsm.close()
    keep = []
    bad = dict(p, station=np.full(m.nelem, 3, np.int32))
    d, met = lsm_structs(bad, meteo, keep)
    out = C.c_void_p()
    assert L.pihm_lsm_create(h.ctx, C.byref(d), C.byref(met), C.byref(out)) == _lib.PIHM_ERR_ARG
    ft2 = meteo[0].copy()
    ft2[1, 5] = ft2[1, 4]                                             # times not increasing
    d, met = lsm_structs(p, (ft2, meteo[1]), keep)
    assert L.pihm_lsm_create(h.ctx, C.byref(d), C.byref(met), C.byref(out)) == _lib.PIHM_ERR_ARG
    # a batched context is refused (one solve per LSM)
    descs = (C.POINTER(_lib.MeshDesc) * 2)(C.pointer(m.desc()), C.pointer(m.desc()))
    bctx = C.c_void_p()
    assert L.pihm_ctx_create_batch(2, descs, C.byref(bctx)) == 0
    d, met = lsm_structs(p, meteo, keep)
    assert L.pihm_lsm_create(bctx, C.byref(d), C.byref(met), C.byref(out)) == _lib.PIHM_ERR_ARG
    L.pihm_ctx_destroy(bctx)
    h.close()


def _wrms(a, b, rtol=1e-3, atol=1e-4):
    w = 1.0 / (rtol * np.abs(b) + atol)
    return float(np.sqrt(np.mean(((a - b) * w) ** 2)))


def test_lsm_run_matches_host_run(c0):
    """pihm_run_gpu_lsm (device LSM every 900 s, then the device solve) against the same
    model-step driver over the oracle RHS and the oracle LSM on host vectors: the states
    agree within the solver tolerance, and the canopy and snow storages, which the step
    computes from the forcing alone, agree to rounding.  Two device runs agree bit for bit."""
    from pihm_amd.forcing import run_config
    from pihm_amd.lsm import run_gpu_lsm
    m = c0
    p, meteo, lai, z0, meltf, cmc0, sneqv0 = _setup(m, 7)
    y0 = m.initial_state()
    cfg = run_config(T0 + 6 * 3600.0, t_start=T0)
    print("[lsm run] c0 6 h device", flush=True)
    rc_g, y_g, st_g, cmc_g, sn_g = run_gpu_lsm(m, p, meteo, (lai, z0, meltf), cfg, y0, cmc0, sneqv0)
    assert rc_g == 0, st_g.as_dict()
    assert st_g.model_steps == 360
    print("[lsm run] c0 6 h host", flush=True)
    o = O.OracleLsmRun(m, p, meteo, lai, z0, meltf, cmc0, sneqv0, cfg.etstep)
    rc_c, y_c, st_c = o.run(y0, cfg)
    assert rc_c == 0
    cmc_c, sn_c = o.storages()
    print("gpu", st_g.as_dict(), "\ncpu", st_c.as_dict(), flush=True)
    assert _wrms(y_g, y_c) < 1.0
    np.testing.assert_allclose(cmc_g, cmc_c, rtol=1e-10, atol=1e-18)
    np.testing.assert_allclose(sn_g, sn_c, rtol=1e-10, atol=1e-18)
    r2 = run_gpu_lsm(m, p, meteo, (lai, z0, meltf), cfg, y0, cmc0, sneqv0)
    assert np.array_equal(r2[1], y_g) and np.array_equal(r2[3], cmc_g) and np.array_equal(r2[4], sn_g)


def test_lsm_run_stops_outside_records(c0):
    from pihm_amd.forcing import run_config
    from pihm_amd.lsm import run_gpu_lsm
    m = c0
    p, meteo, lai, z0, meltf, cmc0, sneqv0 = _setup(m, 8)
    t_last = meteo[0][0, -1]
    cfg = run_config(t_last + 1800.0, t_start=t_last - 1800.0)
    rc, y, st, _, _ = run_gpu_lsm(m, p, meteo, (lai, z0, meltf), cfg, m.initial_state())
    # LSM steps at t_last - 1800, - 900, t_last itself, then t_last + 900: outside
    assert rc == _lib.PIHM_ERR_ARG
    assert st.model_steps == 45
