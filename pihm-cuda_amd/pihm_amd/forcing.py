"""Synthetic storm forcing and the run configuration of the model-step driver.

Precipitation: hourly records per station, each hour wet with probability ``p_wet``
and then of intensity Exp(mean ``mean_mm_h``) mm/h.  Evaporation: a diurnal potential rate
(sine over daylight hours, peak ``pet_peak_mm_h``) split into soil evaporation
(``edir_frac``) and transpiration.  Records are interpolated linearly in time by the
driver, as IntrplForc does (src/forcing.c:482-530).
"""
import ctypes as C

import numpy as np

from . import _lib

MM_H = 1.0 / 1000.0 / 3600.0   # mm/h -> m/s


class StormForcing:
    def __init__(self, hours, nstation=1, seed=0, p_wet=0.1, mean_mm_h=5.0, pet_peak_mm_h=0.3,
                 edir_frac=0.3, station=None, nelem=None, rec_dt=3600.0):
        rng = np.random.default_rng(seed)
        nrec = int(hours) + 2
        wet = rng.random((nstation, nrec)) < p_wet
        inten = rng.exponential(mean_mm_h, (nstation, nrec))
        self.prcp = np.ascontiguousarray(np.where(wet, inten, 0.0) * MM_H)
        hod = (np.arange(nrec) * rec_dt / 3600.0) % 24.0
        pet = np.maximum(0.0, np.sin(np.pi * (hod - 6.0) / 12.0)) * pet_peak_mm_h * MM_H
        self.edir = np.ascontiguousarray(edir_frac * pet)
        self.ett = np.ascontiguousarray((1.0 - edir_frac) * pet)
        self.nstation, self.nrec, self.rec_dt = nstation, nrec, rec_dt
        if station is None:
            station = np.zeros(nelem if nelem is not None else 0, np.int32)
        self.station = np.ascontiguousarray(station, dtype=np.int32)

    def for_mesh(self, mesh):
        if len(self.station) != mesh.nelem:
            self.station = np.zeros(mesh.nelem, np.int32)
        return self

    def struct(self):
        f = _lib.Forcing()
        f.nstation, f.nrec, f.rec_dt = self.nstation, self.nrec, self.rec_dt
        f.prcp = self.prcp.ctypes.data_as(_lib.c_double_p)
        f.edir = self.edir.ctypes.data_as(_lib.c_double_p)
        f.ett = self.ett.ctypes.data_as(_lib.c_double_p)
        f.station = self.station.ctypes.data_as(_lib.c_int32_p)
        return f

    def at(self, t):
        """Interpolated (pcpdrp per station, edir, ett) at time t (same arithmetic as the driver)."""
        k = int(np.floor(t / self.rec_dt))
        k = min(max(k, 0), self.nrec - 2)
        ta, tb = k * self.rec_dt, (k + 1) * self.rec_dt
        f = lambda a, b: ((tb - t) * a + (t - ta) * b) / (tb - ta)
        return f(self.prcp[:, k], self.prcp[:, k + 1]), f(self.edir[k], self.edir[k + 1]), f(self.ett[k], self.ett[k + 1])


def run_config(t_end, t_start=0.0, stepsize=60.0, etstep=900.0, reltol=1e-3, abstol=1e-4,
               initstep=5e-5, maxl=0, max_steps_per_call=0, graphs=1, precond=0):
    """Driver settings of the reference example (benchmark/input/example/example.para):
    60 s model step, 900 s LSM step, reltol 1e-3, abstol 1e-4, initial step 5e-5,
    NUM_NONCOV_FAIL 0, MAX_NONLIN_ITER 3, MIN_NONLIN_ITER 1, DECR/INCR 1.2, MIN_MAXSTEP 1.
    precond=1 adds the block-Jacobi right preconditioner to SPGMR (the reference has none)."""
    c = _lib.RunConfig()
    c.t_start, c.t_end, c.stepsize, c.etstep = t_start, t_end, stepsize, etstep
    c.reltol, c.abstol, c.initstep = reltol, abstol, initstep
    c.nncfn, c.nnimax, c.nnimin, c.decr, c.incr, c.stmin = 0.0, 3.0, 1.0, 1.2, 1.2, 1.0
    c.maxl = maxl
    c.max_steps_per_call = max_steps_per_call
    c.graphs = graphs
    c.precond = precond
    return c


class StepRecord:
    """Storage for the size, order and Newton/Krylov counts of up to ``cap`` accepted
    integrator steps (PihmCVodeSetStepRecord through pihm_run_config)."""

    def __init__(self, cap):
        self.h = np.zeros(cap)
        self.q = np.zeros(cap, dtype=np.int32)
        self.iters = np.zeros((cap, _lib.PIHM_REC_ITERS), dtype=np.int32)
        self.n = C.c_long(0)

    def attach(self, cfg):
        cfg.rec_h = self.h.ctypes.data_as(_lib.c_double_p)
        cfg.rec_q = self.q.ctypes.data_as(_lib.c_int32_p)
        cfg.rec_iters = self.iters.ctypes.data_as(_lib.c_int32_p)
        cfg.rec_cap = len(self.h)
        cfg.rec_n = C.pointer(self.n)
        cfg._record = self            # the arrays stay alive with the configuration
        return cfg

    def steps(self):
        """(h, q, iters) of the recorded steps."""
        n = self.n.value
        return self.h[:n].copy(), self.q[:n].copy(), self.iters[:n].copy()


def set_replay(cfg, h, q, iters=None):
    """Make the run take the recorded steps (h[k], q[k]) in place of the step-size and
    order selection, and with ``iters`` (StepRecord.steps' third array) also the
    recorded Newton and Krylov counts (PihmCVodeSetStepReplay through pihm_run_config)."""
    h = np.ascontiguousarray(h, dtype=np.float64)
    q = np.ascontiguousarray(q, dtype=np.int32)
    if h.shape != q.shape or h.ndim != 1:
        raise ValueError("h and q must be 1-D arrays of one length")
    cfg.replay_h = h.ctypes.data_as(_lib.c_double_p)
    cfg.replay_q = q.ctypes.data_as(_lib.c_int32_p)
    keep = [h, q]
    if iters is not None:
        iters = np.ascontiguousarray(iters, dtype=np.int32)
        if iters.shape != (len(h), _lib.PIHM_REC_ITERS):
            raise ValueError(f"iters must have shape ({len(h)}, {_lib.PIHM_REC_ITERS})")
        cfg.replay_iters = iters.ctypes.data_as(_lib.c_int32_p)
        keep.append(iters)
    cfg.replay_n = len(h)
    cfg._replay = keep
    return cfg
