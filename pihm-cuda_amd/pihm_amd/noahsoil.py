"""Host-side mirror of the _NOAH_ build's soil-water step (Summary's FindWaterTable and
CalcLatFlx, src/update.c:37-43; NoahHydrol, src/noah/noah.c:57-94) over the HIP library.

``HipNoahSoil(rhs, nsoil, sldpth, smcmax, frzx)`` binds the step to a single-solve
``HipRHS`` context whose Noah branches are on (``rhs.set_noah``).  Per-layer arrays are
``[nelem, 11]`` (the reference's ``[MAXLYR]`` per element), caller order.  ``step`` runs
one model step's column update on the device and leaves sh2o[0], fcr, the water-table
layer and GwTransp's share in the context for the next RHS calls; there is no host
implementation here.
"""
import ctypes as C

import numpy as np

from . import _lib

MAXLYR = _lib.PIHM_NOAH_MAXLYR


def _dp(a):
    return a.ctypes.data_as(_lib.c_double_p)


class HipNoahSoil:
    def __init__(self, rhs, nsoil, sldpth, smcmax, frzx):
        self.L = _lib.lib()
        self.rhs = rhs
        self.ne = rhs.mesh.nelem
        self._ns = np.ascontiguousarray(nsoil, dtype=np.int32)
        self._sl = self._layers(sldpth)
        self._sm = np.ascontiguousarray(smcmax, dtype=np.float64)
        self._fz = np.ascontiguousarray(frzx, dtype=np.float64)
        d = _lib.NoahSoilDesc(self._ns.ctypes.data_as(_lib.c_int32_p), _dp(self._sl), _dp(self._sm), _dp(self._fz))
        h = C.c_void_p()
        rc = self.L.pihm_noah_soil_create(rhs.ctx, C.byref(d), C.byref(h))
        if rc:
            raise RuntimeError(f"pihm_noah_soil_create failed ({rc})")
        self.h = h

    def _layers(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        if a.shape != (self.ne, MAXLYR):
            raise ValueError(f"expected [{self.ne}, {MAXLYR}] per-layer values")
        return a

    def close(self):
        if getattr(self, "h", None):
            self.L.pihm_noah_soil_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_state(self, smc, sh2o, fcr=None):
        smc, sh2o = self._layers(smc), self._layers(sh2o)
        f = None if fcr is None else np.ascontiguousarray(fcr, dtype=np.float64)
        if self.L.pihm_noah_soil_set_state(self.h, _dp(smc), _dp(sh2o), None if f is None else _dp(f)):
            raise RuntimeError("pihm_noah_soil_set_state failed")

    def get_state(self):
        ne = self.ne
        out = dict(smc=np.empty((ne, MAXLYR)), sh2o=np.empty((ne, MAXLYR)), fcr=np.empty(ne),
                   nwtbl=np.empty(ne, np.int32), satdpth=np.empty((ne, MAXLYR)),
                   runoff2_lyr=np.empty((ne, MAXLYR)), smflxv=np.empty((ne, MAXLYR)), runoff3=np.empty(ne))
        rc = self.L.pihm_noah_soil_get_state(self.h, _dp(out["smc"]), _dp(out["sh2o"]), _dp(out["fcr"]),
                                             out["nwtbl"].ctypes.data_as(_lib.c_int32_p), _dp(out["satdpth"]),
                                             _dp(out["runoff2_lyr"]), _dp(out["smflxv"]), _dp(out["runoff3"]))
        if rc:
            raise RuntimeError(f"pihm_noah_soil_get_state failed ({rc})")
        return out

    def _ydev(self, y):
        if y is not None:
            y = np.ascontiguousarray(y, dtype=np.float64)
            if len(y) < self.rhs.neq:
                y = np.concatenate([y, np.zeros(self.rhs.neq - len(y))])
            if self.L.pihm_ctx_state_to_device(self.rhs.ctx, _dp(y), self.rhs.y):
                raise RuntimeError("state_to_device failed")
        return self.L.N_VGetDeviceArrayPointer_Hip(self.rhs.y)

    def step(self, dt, infil, runoff2, y=None):
        """One soil-water step at state y (caller order; None: the context's state vector)."""
        infil = np.ascontiguousarray(infil, dtype=np.float64)
        runoff2 = np.ascontiguousarray(runoff2, dtype=np.float64)
        rc = self.L.pihm_noah_soil_step(self.h, float(dt), self._ydev(y), _dp(infil), _dp(runoff2))
        if rc:
            raise RuntimeError(f"pihm_noah_soil_step failed ({rc})")

    def time(self, dt, iters=100):
        ms = C.c_double()
        if self.L.pihm_noah_soil_time(self.h, float(dt), self._ydev(None), int(iters), C.byref(ms)):
            raise RuntimeError("pihm_noah_soil_time failed")
        return ms.value


def run_gpu_noah(mesh, forcing, cfg, y0, noah):
    """Integrate on the device with the _NOAH_ RHS branches and the soil-water step at every
    Summary (pihm_run_gpu_noah).  ``noah`` holds smcmin, et [nelem, 11], nwtbl, nroot (the
    initial hand-over) and nsoil, sldpth, smcmax, frzx, smc, sh2o (the soil columns).
    Returns (rc, y, stats, smc, sh2o)."""
    from .rhs import HipRHS
    h = HipRHS(mesh, nan_check=False)
    try:
        h.set_noah(noah["smcmin"])
        h.set_noah_step(et=noah["et"], nwtbl=noah["nwtbl"], nroot=noah["nroot"])
        soil = HipNoahSoil(h, noah["nsoil"], noah["sldpth"], noah["smcmax"], noah["frzx"])
        try:
            soil.set_state(noah["smc"], noah["sh2o"])
            y = np.ascontiguousarray(y0, dtype=np.float64).copy()
            st = _lib.RunStats()
            fs = forcing.struct()
            rc = h.L.pihm_run_gpu_noah(h.ctx, C.byref(fs), soil.h, C.byref(cfg), _dp(y), C.byref(st))
            s = soil.get_state()
            return rc, y, st, s["smc"], s["sh2o"]
        finally:
            soil.close()
    finally:
        h.close()
