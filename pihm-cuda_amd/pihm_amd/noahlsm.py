"""The Noah land-surface step of the _NOAH_ build on the device (pihm_noahlsm.h), and the
synthetic Flux-PIHM inputs of the 100,000-triangle coupled workload.

``synthetic_noah(mesh, seed)`` draws per-element land cover and soil parameters from the
ranges of the reference's vegetation and soil tables (benchmark/input/vegprmt.tbl,
example.soil) and the universal parameters of its example .lsm file (SBETA -2, FXEXP 2,
CSOIL 2e6, SALP 2.6, FRZK 0.15, ZBOT -8 m, TBOT 282.15 K, CZIL 0.1, LVCOEF 0.5); the soil
column has 11 layers growing downwards by a factor 1.15 and summing to the element's
soil depth, with the root layer found as FindLayer does (src/noah/lsm_func.c:56-84).

``synthetic_noah_meteo(nstation, hours, seed)`` is the workload's hourly station forcing:
T = 10 + 8 sin(diurnal) degC, RH U(0.4, 0.9), wind Weibull (k = 2, scale 3 m/s),
SW = max(0, 800 sin(diurnal)) W/m2, LW 300 +/- 30 W/m2, pressure 101 kPa, and the
storm precipitation of the plain workload (10% wet hours, Exp(5 mm/h)).
"""
import ctypes as C
import math

import numpy as np

from . import _lib

MAXLYR = 11
MET_NVAR = 7      # PRCP (kg m-2 s-1) SFCTMP (K) RH (%) SFCSPD (m/s) SOLAR (W/m2) LONGWV (W/m2) PRES (Pa)
UNIVERSAL = dict(csoil=2e6, sbeta=-2.0, salp=2.6, fxexp=2.0, czil=0.1, lvcoef=0.5, zbot=-8.0, tbot=282.15,
                 snoalb=0.75, zlvl=3.0)
FRZK = 0.15


def layer_depths(depth, nsoil=MAXLYR, growth=1.15):
    """Layer thicknesses [ne, MAXLYR] growing by `growth` and summing to each depth."""
    w = growth ** np.arange(nsoil)
    w = w / w.sum()
    out = np.zeros((len(depth), MAXLYR))
    out[:, :nsoil] = np.asarray(depth)[:, None] * w[None, :]
    return out


def find_layer(sldpth, nsoil, depth):
    """FindLayer (src/noah/lsm_func.c:56-84): the 1-based layer that contains `depth`."""
    if depth <= 0.0:
        return 0
    dsum, j, ind = 0.0, 0, 0
    while dsum < depth:
        if j >= len(sldpth) or sldpth[j] < 0.0:
            break
        dsum += sldpth[j]
        ind = j
        j += 1
    return min(ind + 1, nsoil)


def synthetic_noah(mesh, seed=0, nsoil=MAXLYR):
    """Per-element parameters (caller order) for the Noah step and the soil-water step:
    returns a dict of arrays plus the universal scalars."""
    rng = np.random.default_rng(seed)
    ne = mesh.nelem
    a = mesh.a
    sld = layer_depths(a["depth"], nsoil)
    smcmin = rng.uniform(0.03, 0.06, ne)
    smcmax = smcmin + a["porosity"]
    lai_lo = rng.uniform(0.5, 2.0, ne)
    p = dict(
        nsoil=np.full(ne, nsoil, np.int32), sldpth=sld, smcmin=smcmin, smcmax=smcmax,
        nroot=np.array([find_layer(sld[i], nsoil, a["rzd"][i]) for i in range(ne)], np.int32),
        shdfac=rng.uniform(0.5, 0.95, ne), cmcfactr=np.full(ne, 2e-4), laimin=lai_lo,
        laimax=lai_lo + rng.uniform(1.0, 4.0, ne), emissmin=rng.uniform(0.92, 0.94, ne),
        albedomin=rng.uniform(0.12, 0.18, ne), z0min=rng.uniform(0.01, 0.2, ne),
        snup=rng.uniform(0.02, 0.08, ne), cfactr=np.full(ne, 0.5), rgl=rng.choice([30.0, 65.0, 100.0], ne),
        rsmin=rng.uniform(40.0, 300.0, ne), rsmax=np.full(ne, 5000.0), topt=np.full(ne, 298.0),
        hs=rng.uniform(36.0, 55.0, ne), quartz=rng.uniform(0.1, 0.6, ne),
        zlvl_wind=np.full(ne, 10.0), station=np.zeros(ne, np.int32))
    p["emissmax"] = p["emissmin"] + rng.uniform(0.0, 0.04, ne)
    p["albedomax"] = p["albedomin"] + rng.uniform(0.0, 0.08, ne)
    p["z0max"] = p["z0min"] + rng.uniform(0.0, 0.4, ne)
    p["smcwlt"] = smcmin + a["porosity"] * rng.uniform(0.15, 0.3, ne)
    p["smcref"] = smcmin + a["porosity"] * rng.uniform(0.6, 0.8, ne)
    p["smcdry"] = p["smcwlt"].copy()
    # ps.frzx (src/noah/lsm_init.c:39-43)
    p["frzx"] = FRZK * ((smcmax / p["smcref"]) * (0.412 / 0.468))
    p.update(UNIVERSAL)
    return p


def synthetic_noah_state(mesh, p, seed=0, t0=283.15):
    """An initial state: soil moisture between wilting point and saturation, a linear
    temperature profile from t0 at the surface to TBOT at ZBOT, no snow, an empty
    canopy, and the reference's initial exchange coefficients and albedo."""
    rng = np.random.default_rng(seed)
    ne = mesh.nelem
    frac = rng.uniform(0.3, 0.9, (ne, MAXLYR))
    smc = p["smcwlt"][:, None] + (p["smcmax"] - p["smcwlt"])[:, None] * frac
    zs = -np.cumsum(p["sldpth"], 1)
    zmid = zs + 0.5 * p["sldpth"]
    stc = t0 + (p["tbot"] - t0) * (zmid / p["zbot"])
    return dict(smc=smc, sh2o=smc.copy(), stc=stc, t1=np.full(ne, t0), sneqv=np.zeros(ne), snowh=np.zeros(ne),
                cmc=np.zeros(ne))


def synthetic_noah_meteo(nstation, hours, seed=0, p_wet=0.1, mean_mm_h=5.0, t_mean_c=10.0):
    """Hourly records [nstation, hours + 2, MET_NVAR] and their times [s]."""
    rng = np.random.default_rng(seed)
    nrec = int(hours) + 2
    hod = np.arange(nrec) % 24
    diurnal = np.sin(2.0 * np.pi * (hod - 6.0) / 24.0)
    out = np.zeros((nstation, nrec, MET_NVAR))
    wet = rng.random((nstation, nrec)) < p_wet
    out[:, :, 0] = np.where(wet, rng.exponential(mean_mm_h, (nstation, nrec)), 0.0) / 3600.0   # kg m-2 s-1
    out[:, :, 1] = 273.15 + t_mean_c + 8.0 * diurnal[None, :]
    out[:, :, 2] = 100.0 * rng.uniform(0.4, 0.9, (nstation, nrec))
    out[:, :, 3] = 3.0 * rng.weibull(2.0, (nstation, nrec))
    out[:, :, 4] = np.maximum(0.0, 800.0 * diurnal)[None, :].repeat(nstation, 0)
    out[:, :, 5] = 300.0 + rng.uniform(-30.0, 30.0, (nstation, nrec))
    out[:, :, 6] = 101000.0
    ftime = np.tile(3600.0 * np.arange(nrec, dtype=np.float64), (nstation, 1))
    return ftime, out


def interp_meteo(ftime, data, t):
    """IntrplForc (src/forcing.c:482-530) at t for every station: [nstation, MET_NVAR]."""
    out = np.zeros((data.shape[0], MET_NVAR))
    for s in range(data.shape[0]):
        tt = ftime[s]
        k = int(np.searchsorted(tt, t, side="right")) - 1
        k = min(max(k, 0), len(tt) - 2)
        ta, tb = tt[k], tt[k + 1]
        out[s] = ((tb - t) * data[s, k] + (t - ta) * data[s, k + 1]) / (tb - ta)
    return out
