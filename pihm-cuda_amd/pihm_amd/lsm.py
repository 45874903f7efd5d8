"""Host-side mirror of the reference's simple land-surface step over the HIP library.

Every LSM step the reference interpolates the meteorological stations' records to the
current time (IntrplForc, src/forcing.c:482-530; ApplyMeteoForc, src/forcing.c:305-344)
and evaluates interception, snow and evapotranspiration per element (IntcpSnowEt,
src/is_sm_et.c:5-226), which leaves the pcpdrp, edir and ett that ODE() reads.
``HipLSM(rhs, params, meteo)`` runs both on the device through the C ABI
(``pihm_lsm_*``) and writes the three fields straight into the RHS context; there is no
host implementation here.

``params`` maps each name of ``LSM_FIELDS`` to a per-element array in the caller's order
(the element_struct fields InitLc / InitSoil / InitTopo fill, src/init_lc.c:22-37,
src/init_soil.c:24-74), plus ``station`` (0-based meteorological station per element) and
the calibration multipliers ``cal_edir``, ``cal_ec``, ``cal_ett`` (src/read_calib.c).
``meteo`` is ``(ftime[nstation, nrec], data[nstation, nrec, 5])`` with the variables
PRCP [kg m-2 s-1], SFCTMP [K], RH [%], SFCSPD [m s-1], SOLAR [W m-2].
"""
import ctypes as C

import numpy as np

from . import _lib

LSM_FIELDS = ("albedomin", "albedomax", "shdfac", "cmcfactr", "cfactr", "zlvl_wind", "rzd",
              "rgl", "rsmin", "rsmax", "topt", "depth", "porosity", "smcmin", "smcwlt", "smcref", "zmax")
MET_VARS = ("prcp", "sfctmp", "rh", "sfcspd", "solar")


def _dp(a):
    return a.ctypes.data_as(_lib.c_double_p)


def lsm_structs(params, meteo, keep):
    """(LsmDesc, Meteo) ctypes structures over contiguous copies kept alive in ``keep``."""
    d = _lib.LsmDesc()
    for f in LSM_FIELDS:
        a = np.ascontiguousarray(params[f], dtype=np.float64)
        keep.append(a)
        setattr(d, f, _dp(a))
    st = np.ascontiguousarray(params["station"], dtype=np.int32)
    keep.append(st)
    d.station = st.ctypes.data_as(_lib.c_int32_p)
    d.cal_edir = float(params.get("cal_edir", 1.0))
    d.cal_ec = float(params.get("cal_ec", 1.0))
    d.cal_ett = float(params.get("cal_ett", 1.0))
    ftime, data = meteo
    ft = np.ascontiguousarray(ftime, dtype=np.float64)
    dt = np.ascontiguousarray(data, dtype=np.float64)
    if ft.ndim != 2 or dt.shape != ft.shape + (_lib.PIHM_MET_NVAR,):
        raise ValueError("meteo must be ftime[nstation, nrec] and data[nstation, nrec, 5]")
    keep += [ft, dt]
    m = _lib.Meteo()
    m.nstation, m.nrec = ft.shape
    m.ftime = _dp(ft)
    m.data = _dp(dt)
    return d, m


class HipLSM:
    """The land-surface step bound to a single-solve ``HipRHS`` context."""

    def __init__(self, rhs, params, meteo):
        self.L = _lib.lib()
        self.rhs = rhs
        self.ne = rhs.mesh.nelem
        keep = []
        desc, met = lsm_structs(params, meteo, keep)
        h = C.c_void_p()
        rc = self.L.pihm_lsm_create(rhs.ctx, C.byref(desc), C.byref(met), C.byref(h))
        if rc:
            raise RuntimeError(f"pihm_lsm_create failed ({rc})")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.L.pihm_lsm_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _elem(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        if a.shape != (self.ne,):
            raise ValueError(f"expected {self.ne} per-element values")
        return a

    def set_veg(self, lai, z0, meltf):
        """Projected LAI, roughness length [m] per element and the melt factor [m s-1 K-1]."""
        lai, z0 = self._elem(lai), self._elem(z0)
        if self.L.pihm_lsm_set_veg(self.h, _dp(lai), _dp(z0), float(meltf)):
            raise RuntimeError("pihm_lsm_set_veg failed")

    def set_state(self, cmc, sneqv):
        cmc, sneqv = self._elem(cmc), self._elem(sneqv)
        if self.L.pihm_lsm_set_state(self.h, _dp(cmc), _dp(sneqv)):
            raise RuntimeError("pihm_lsm_set_state failed")

    def get_state(self):
        cmc, sneqv = np.empty(self.ne), np.empty(self.ne)
        if self.L.pihm_lsm_get_state(self.h, _dp(cmc), _dp(sneqv)):
            raise RuntimeError("pihm_lsm_get_state failed")
        return cmc, sneqv

    def step(self, t, stepsize, y=None):
        """One LSM step at time t over stepsize.  ``y`` (caller order) is uploaded into the
        context's state vector first; without it the vector already on the device is used.
        Returns the C ABI's code (PIHM_ERR_ARG when t is outside the records)."""
        if y is not None:
            y = np.ascontiguousarray(y, dtype=np.float64)
            if len(y) < self.rhs.neq:
                y = np.concatenate([y, np.zeros(self.rhs.neq - len(y))])
            if self.L.pihm_ctx_state_to_device(self.rhs.ctx, _dp(y), self.rhs.y):
                raise RuntimeError("state_to_device failed")
        return self.L.pihm_lsm_step(self.h, float(t), float(stepsize), self.rhs.y)

    def fluxes(self):
        out = {k: np.empty(self.ne) for k in ("pcpdrp", "edir", "ett", "ec", "drip")}
        rc = self.L.pihm_lsm_get_fluxes(self.h, *(_dp(out[k]) for k in ("pcpdrp", "edir", "ett", "ec", "drip")))
        if rc:
            raise RuntimeError(f"pihm_lsm_get_fluxes failed ({rc})")
        return out


def run_gpu_lsm(mesh, params, meteo, veg, cfg, y0, cmc0=None, sneqv0=None, bc=None):
    """Integrate on the device with the land-surface step as the forcing
    (pihm_run_gpu_lsm).  ``veg`` is (lai, z0, meltf); ``bc`` the HipBC arguments
    (elem_type, elem_series, river_type, river_series) or None.  Returns
    (rc, y, stats, cmc, sneqv)."""
    from .bc import HipBC
    from .rhs import HipRHS
    h = HipRHS(mesh, nan_check=False)
    lsm = HipLSM(h, params, meteo)
    hb = HipBC(h, *bc) if bc is not None else None
    try:
        lsm.set_veg(*veg)
        ne = mesh.nelem
        lsm.set_state(np.zeros(ne) if cmc0 is None else cmc0, np.zeros(ne) if sneqv0 is None else sneqv0)
        y = np.ascontiguousarray(y0, dtype=np.float64).copy()
        st = _lib.RunStats()
        rc = h.L.pihm_run_gpu_lsm(h.ctx, lsm.h, hb.h if hb else None, C.byref(cfg), _dp(y), C.byref(st))
        cmc, sneqv = lsm.get_state()
        return rc, y, st, cmc, sneqv
    finally:
        if hb:
            hb.close()
        lsm.close()
        h.close()


def synthetic_params(ne, nstation, seed=0):
    """Seeded per-element land-surface parameters spanning the reference's land-cover and
    soil tables (vegprmt.tbl, .soil): shading 0.05-0.95, root depth 0.05-1 m, minimum
    stomatal resistance 40-300 s/m, soil depth 2-8 m."""
    rng = np.random.default_rng(seed)
    amin = rng.uniform(0.08, 0.25, ne)
    smcmin = rng.uniform(0.01, 0.06, ne)
    smcmax = rng.uniform(0.35, 0.5, ne)
    smcwlt = smcmin + rng.uniform(0.02, 0.08, ne)
    p = dict(albedomin=amin, albedomax=amin + rng.uniform(0.0, 0.08, ne),
             shdfac=rng.uniform(0.05, 0.95, ne), cmcfactr=np.full(ne, 2e-4),
             cfactr=rng.uniform(0.5, 1.0, ne), zlvl_wind=np.full(ne, 10.0),
             rzd=rng.uniform(0.05, 1.0, ne), rgl=rng.uniform(30.0, 100.0, ne),
             rsmin=rng.uniform(40.0, 300.0, ne), rsmax=np.full(ne, 5000.0), topt=np.full(ne, 298.0),
             depth=rng.uniform(2.0, 8.0, ne), porosity=smcmax - smcmin, smcmin=smcmin, smcwlt=smcwlt,
             smcref=smcwlt + rng.uniform(0.1, 0.2, ne), zmax=rng.uniform(200.0, 800.0, ne),
             station=rng.integers(0, nstation, ne).astype(np.int32),
             cal_edir=1.0, cal_ec=1.0, cal_ett=1.0)
    return p


def synthetic_meteo(nstation, nrec, t0=0.0, dt=3600.0, seed=0):
    """Seeded hourly station records: a diurnal temperature swing across the freezing
    range, intermittent rain, humidity 30-100 %, wind 0.5-9 m/s, daytime radiation."""
    rng = np.random.default_rng(seed)
    ftime = np.tile(t0 + dt * np.arange(nrec, dtype=np.float64), (nstation, 1))
    hour = (ftime / 3600.0) % 24.0
    base = rng.uniform(266.0, 285.0, (nstation, 1))
    tmp = base + 6.0 * np.sin((hour - 9.0) / 24.0 * 2 * np.pi) + rng.normal(0.0, 1.0, ftime.shape)
    wet = rng.random(ftime.shape) < 0.35
    prcp = np.where(wet, rng.exponential(4e-4, ftime.shape), 0.0)
    rh = np.clip(rng.uniform(30.0, 100.0, ftime.shape) + 20.0 * wet, 30.0, 100.0)
    spd = rng.uniform(0.5, 9.0, ftime.shape)
    sol = np.maximum(0.0, 800.0 * np.sin((hour - 6.0) / 12.0 * np.pi)) * rng.uniform(0.3, 1.0, ftime.shape)
    data = np.stack([prcp, tmp, rh, spd, sol], axis=-1)
    return ftime, data
