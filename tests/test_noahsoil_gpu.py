"""The _NOAH_ build's soil-water step on MI355X (pihm_noah_soil_step) against the oracle's
restatement (oracle/noahsoil_oracle.c), step after step, and the RHS that follows it.

The water-table search and the lateral-runoff split involve no transcendental function and
must agree bit for bit; the column update goes through WDfCnd's powers (the device's xpow
against the host's pow) and is held to 1e-11 of each quantity's magnitude.  The inputs are
drawn to reach every branch: water tables below, inside and above the column, layers
clamped at both moisture bounds, frozen layers (the frozen-soil factor and the diffusivity's
ice weighting), negative subsurface runoff.
"""
import numpy as np
import pytest

import oracle as O
from conftest import assert_rhs_close
from pihm_amd.noahsoil import MAXLYR, HipNoahSoil
from pihm_amd.rhs import HipRHS
from test_oracle_pins import _inputs, _states

pytestmark = pytest.mark.gpu


def _setup(m, rng):
    ne = m.nelem
    por = m.a["porosity"]
    smcmin = rng.uniform(0.01, 0.06, ne)
    smcmax = smcmin + por
    nsoil = rng.integers(2, MAXLYR + 1, ne).astype(np.int32)
    sld = np.zeros((ne, MAXLYR))
    for i in range(ne):
        sld[i, :nsoil[i]] = rng.uniform(0.05, 0.8, nsoil[i])
    frzx = rng.uniform(0.05, 0.3, ne)
    et = np.zeros((ne, MAXLYR))
    for i in range(ne):
        et[i, :nsoil[i]] = rng.uniform(-0.1, 1.0, nsoil[i]) * 3e-8 / nsoil[i]
    nroot = np.minimum(rng.integers(1, MAXLYR + 1, ne), nsoil).astype(np.int32)
    sh2o = np.zeros((ne, MAXLYR))
    smc = np.zeros((ne, MAXLYR))
    for i in range(ne):
        n = nsoil[i]
        sh2o[i, :n] = smcmin[i] + por[i] * rng.uniform(-0.05, 1.05, n)
        ice = np.where(rng.random(n) < 0.2, rng.uniform(0.0, 0.2, n), 0.0)
        smc[i, :n] = sh2o[i, :n] + ice
    soil = dict(smcmax=smcmax, smcmin=smcmin, ksatv=m.a["ksatv"], alpha=m.a["alpha"], beta=m.a["beta"],
                frzx=frzx)
    return dict(smcmin=smcmin, nsoil=nsoil, sld=sld, et=et, nroot=nroot, sh2o=sh2o, smc=smc, soil=soil)


def _compare(got, ref, what, flux_terms):
    """flux_terms[i, k]: the magnitude of the terms that make up the layer flux smflxv[i, k]
    (the storage change of the layer below at full moisture, its runoff and transpiration,
    the flux below it), against which its rounding is measured: smflxv is a difference of
    storages divided by dt and can cancel to far below its terms."""
    ne = len(ref["fcr"])
    assert np.array_equal(got["nwtbl"], ref["nwtbl"]), what
    assert np.array_equal(got["satdpth"], ref["satdpth"]), what
    for k in ("runoff2_lyr", "smc", "sh2o", "smflxv", "runoff3", "fcr"):
        g, r = np.asarray(got[k]).reshape(ne, -1), np.asarray(ref[k]).reshape(ne, -1)
        scale = np.maximum(np.abs(r), np.max(np.abs(r), axis=1, keepdims=True) * 1e-6 + 1e-300)
        if k == "smflxv":
            scale = np.maximum(scale, flux_terms)
        err = np.abs(g - r) / scale
        assert np.all(err <= 1e-11), (what, k, float(err.max()), np.unravel_index(np.argmax(err), err.shape))


@pytest.mark.parametrize("which", ["c0", "example", "tree_small"])
def test_noah_soil_step_parity(which, request):
    m = request.getfixturevalue(which)
    ne = m.nelem
    print(f"[noah soil] {which} ne={ne}", flush=True)
    rng = np.random.default_rng(51)
    z = _setup(m, rng)
    h, o = HipRHS(m), O.OracleRHS(m)
    inp = _inputs(m, rng)
    h.set_step_inputs(**inp)
    o.set_step_inputs(**inp)
    h.set_noah(z["smcmin"])
    h.set_noah_step(et=z["et"], nwtbl=np.ones(ne, np.int32), nroot=z["nroot"])
    s = HipNoahSoil(h, z["nsoil"], z["sld"], z["soil"]["smcmax"], z["soil"]["frzx"])
    s.set_state(z["smc"], z["sh2o"])
    ref = dict(smc=z["smc"].copy(), sh2o=z["sh2o"].copy(), fcr=np.ones(ne), smflxv=np.zeros((ne, MAXLYR)),
               runoff3=np.zeros(ne), runoff2_lyr=np.zeros((ne, MAXLYR)))
    depth = z["sld"].sum(axis=1)
    for step, y in enumerate(_states(m, rng, 3)):
        y = y.copy()
        # water tables below, inside and above the soil column
        y[2 * ne:3 * ne] = depth * rng.choice([-0.1, 0.3, 0.7, 0.97, 1.2], ne)
        infil = rng.uniform(0.0, 3e-6, ne)
        runoff2 = rng.uniform(-2e-8, 2e-7, ne)
        s.step(60.0, infil, runoff2, y)
        ref = O.noah_soil_step(60.0, z["nsoil"], z["sld"], z["soil"], y[2 * ne:3 * ne], infil, runoff2,
                               inp["edir"], z["et"], ref)
        terms = np.zeros((ne, MAXLYR))
        terms[:, :-1] = (z["soil"]["smcmax"][:, None] * z["sld"][:, 1:] / 60.0 + np.abs(ref["runoff2_lyr"][:, 1:])
                         + np.abs(z["et"][:, 1:]) + np.abs(ref["smflxv"][:, 1:]))
        _compare(s.get_state(), ref, f"{which} step {step}", terms)
        # the next model step's RHS reads what the step left
        o.set_noah(z["smcmin"], sh2o0=ref["sh2o"][:, 0], fcr=ref["fcr"], et=z["et"], nwtbl=ref["nwtbl"],
                   nroot=z["nroot"])
        rc_h, dh = h.ode(y)
        rc_o, do, mag = o.ode(y, with_mag=True)
        assert rc_h == rc_o
        assert_rhs_close(dh, do, mag, 1e-12, f"{which} RHS after soil step {step}")
    s.close()
    h.close()


def test_noah_soil_argument_errors(c0):
    from pihm_amd import _lib
    import ctypes as C
    L = _lib.lib()
    m = c0
    ne = m.nelem
    h = HipRHS(m)
    rng = np.random.default_rng(52)
    z = _setup(m, rng)
    with pytest.raises(RuntimeError):            # Noah branches not on
        HipNoahSoil(h, z["nsoil"], z["sld"], z["soil"]["smcmax"], z["soil"]["frzx"])
    h.set_noah(z["smcmin"])
    bad = z["nsoil"].copy()
    bad[3] = 1                                   # one layer: SRT reads a second
    with pytest.raises(RuntimeError):
        HipNoahSoil(h, bad, z["sld"], z["soil"]["smcmax"], z["soil"]["frzx"])
    sld = z["sld"].copy()
    sld[5, 0] = 0.0
    with pytest.raises(RuntimeError):
        HipNoahSoil(h, z["nsoil"], sld, z["soil"]["smcmax"], z["soil"]["frzx"])
    s = HipNoahSoil(h, z["nsoil"], z["sld"], z["soil"]["smcmax"], z["soil"]["frzx"])
    one = np.ones(ne)
    assert L.pihm_noah_soil_step(s.h, 0.0, L.N_VGetDeviceArrayPointer_Hip(h.y),
                                 one.ctypes.data_as(_lib.c_double_p), one.ctypes.data_as(_lib.c_double_p)) \
        == _lib.PIHM_ERR_ARG
    s.close()
    h.close()


def _wrms(a, b, rtol=1e-3, atol=1e-4):
    w = 1.0 / (rtol * np.abs(b) + atol)
    return float(np.sqrt(np.mean(((a - b) * w) ** 2)))


@pytest.mark.parametrize("which", ["c0", "example"])
def test_coupled_noah_run_matches_host_run(which, request):
    """pihm_run_gpu_noah (the _NOAH_ RHS inside the integrator, the soil-water step with its
    MassBalance at every Summary on the device) against the same coupled run on the host:
    the hydrologic state and the soil moisture within the solver tolerance."""
    from pihm_amd.forcing import StormForcing, run_config
    from pihm_amd.noahsoil import run_gpu_noah
    from test_noahsoil_pins import noah_run_setup
    m = request.getfixturevalue(which)
    rng = np.random.default_rng(61)
    y0 = m.initial_state()
    nz = noah_run_setup(m, rng, y0)
    f = StormForcing(8, seed=5, nelem=m.nelem, p_wet=0.5)
    cfg = lambda: run_config(3 * 3600.0)
    print(f"[noah run] {which} 3 h gpu", flush=True)
    rc_g, y_g, st_g, smc_g, sh2o_g = run_gpu_noah(m, f, cfg(), y0, nz)
    assert rc_g == 0, st_g.as_dict()
    run = O.OracleRun(m, f, noah=nz)
    rc_c, y_c, st_c = run.run(y0, cfg())
    assert rc_c == 0
    smc_c, sh2o_c = run.noah_soil_state()
    print("gpu", st_g.as_dict(), "\ncpu", st_c.as_dict(), "\nwrms", _wrms(y_g, y_c),
          "max d sh2o", float(np.max(np.abs(sh2o_g - sh2o_c))), flush=True)
    assert _wrms(y_g, y_c) < 1.0
    assert abs(st_g.nst - st_c.nst) <= 0.1 * st_c.nst
    # the columns follow the groundwater of their run, so they inherit its solver-tolerance
    # differences (a water table that moves by dz re-saturates dz of a layer): held to the
    # same weighted RMS bound, with no layer further off than 5e-3 m3 m-3
    used = np.arange(11)[None, :] < nz["nsoil"][:, None]
    assert _wrms(sh2o_g[used], sh2o_c[used]) < 1.0
    assert _wrms(smc_g[used], smc_c[used]) < 1.0
    assert np.max(np.abs(sh2o_g - sh2o_c)) < 5e-3 and np.max(np.abs(smc_g - smc_c)) < 5e-3
    assert not np.array_equal(sh2o_g, nz["sh2o"])
