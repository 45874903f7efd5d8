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


def _dp(a):
    return a.ctypes.data_as(_lib.c_double_p)


class HipNoahLSM:
    """The Noah step (pihm_noah_lsm_*) on the soil column of a ``HipNoahSoil``, whose
    context has its Noah branches on.  ``p`` is a dict from ``synthetic_noah`` (or the
    same fields from the reference's tables); ``ftime``/``data`` the station records."""

    def __init__(self, soil, p, ftime, data):
        self.L = _lib.lib()
        self.soil = soil
        self.ne = soil.ne
        f = lambda a: np.ascontiguousarray(a, dtype=np.float64)
        i = lambda a: np.ascontiguousarray(a, dtype=np.int32)
        self._keep = {k: f(p[k]) for k in _lib.NOAH_LSM_ARRAYS}
        self._keep["nroot"], self._keep["station"] = i(p["nroot"]), i(p["station"])
        d = _lib.NoahLsmDesc()
        d.nroot = self._keep["nroot"].ctypes.data_as(_lib.c_int32_p)
        d.station = self._keep["station"].ctypes.data_as(_lib.c_int32_p)
        for k in _lib.NOAH_LSM_ARRAYS:
            setattr(d, k, _dp(self._keep[k]))
        for k in _lib.NOAH_LSM_SCALARS:
            setattr(d, k, float(p[k]))
        self._ft, self._data = f(ftime), f(data)
        m = _lib.Meteo(self._data.shape[0], self._data.shape[1], _dp(self._ft), _dp(self._data))
        h = C.c_void_p()
        rc = self.L.pihm_noah_lsm_create(soil.h, C.byref(d), C.byref(m), C.byref(h))
        if rc:
            raise RuntimeError(f"pihm_noah_lsm_create failed ({rc})")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.L.pihm_noah_lsm_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_lai(self, lai):
        lai = np.ascontiguousarray(lai, dtype=np.float64)
        if self.L.pihm_noah_lsm_set_lai(self.h, _dp(lai)):
            raise RuntimeError("pihm_noah_lsm_set_lai failed")

    def set_state(self, stc=None, t1=None, sneqv=None, snowh=None, cmc=None):
        f = lambda a: None if a is None else np.ascontiguousarray(a, dtype=np.float64)
        arrs = [f(stc), f(t1), f(sneqv), f(snowh), f(cmc)]
        if self.L.pihm_noah_lsm_set_state(self.h, *[None if a is None else _dp(a) for a in arrs]):
            raise RuntimeError("pihm_noah_lsm_set_state failed")

    def get_state(self):
        stc = np.empty((self.ne, MAXLYR))
        sc = np.empty((self.ne, _lib.PIHM_NOAH_NSCALAR))
        if self.L.pihm_noah_lsm_get_state(self.h, _dp(stc), _dp(sc)):
            raise RuntimeError("pihm_noah_lsm_get_state failed")
        out = {k: sc[:, j].copy() for j, k in enumerate(_lib.NOAH_STATE_NAMES)}
        out["stc"] = stc
        return out

    def step(self, t, dt):
        rc = self.L.pihm_noah_lsm_step(self.h, float(t), float(dt))
        if rc:
            raise RuntimeError(f"pihm_noah_lsm_step failed ({rc})")

    def fluxes(self):
        fx = np.empty((self.ne, _lib.PIHM_NOAH_NFLUX))
        et = np.empty((self.ne, MAXLYR))
        if self.L.pihm_noah_lsm_get_fluxes(self.h, _dp(fx), _dp(et)):
            raise RuntimeError("pihm_noah_lsm_get_fluxes failed")
        out = {k: fx[:, j].copy() for j, k in enumerate(_lib.NOAH_FLUX_NAMES)}
        out["et"] = et
        return out

    def time(self, t, dt, iters=20):
        ms = C.c_double()
        if self.L.pihm_noah_lsm_time(self.h, float(t), float(dt), int(iters), C.byref(ms)):
            raise RuntimeError("pihm_noah_lsm_time failed")
        return ms.value


def flux_pihm_setup(mesh, p, state, ftime, data, lai):
    """A single-solve context with the _NOAH_ branches, the soil-water step and the Noah
    step over it, in their initial state.  Returns (rhs, soil, lsm)."""
    from .noahsoil import HipNoahSoil
    from .rhs import HipRHS
    h = HipRHS(mesh, nan_check=False)
    h.set_noah(p["smcmin"])
    soil = HipNoahSoil(h, p["nsoil"], p["sldpth"], p["smcmax"], p["frzx"])
    soil.set_state(state["smc"], state["sh2o"])
    lsm = HipNoahLSM(soil, p, ftime, data)
    lsm.set_lai(lai)
    lsm.set_state(stc=state["stc"], t1=state["t1"], sneqv=state["sneqv"], snowh=state["snowh"], cmc=state["cmc"])
    return h, soil, lsm


def run_gpu_flux(lsm, cfg, y0):
    """The Flux-PIHM model-step loop on the device (pihm_run_gpu_flux). Returns (rc, y, stats)."""
    y = np.ascontiguousarray(y0, dtype=np.float64).copy()
    st = _lib.RunStats()
    rc = lsm.L.pihm_run_gpu_flux(lsm.h, C.byref(cfg), _dp(y), C.byref(st))
    return rc, y, st
