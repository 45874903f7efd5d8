"""Host-side mirror of the reference RHS interface over the HIP library.

``HipRHS(mesh).ode(y, t)`` has the meaning of the reference's
``ODE(t, CV_Y, CV_Ydot, pihm)`` (src/ode.c:4-399): y and the returned ydot are in the
caller's order; the carried river-edge overland flux (read by Infil one call later) lives
in the device context between calls.  Every evaluation runs the HIP kernels through the
C ABI (``PIHM_ODE_Hip``); there is no host implementation here.
"""
import ctypes as C

import numpy as np

from . import _lib


def _dp(a):
    return a.ctypes.data_as(_lib.c_double_p)


class HipRHS:
    def __init__(self, mesh, nan_check=True, record=False, closure=False, rhs_path=None):
        self.L = _lib.lib()
        self.mesh = mesh
        self.desc = mesh.desc()
        ctx = C.c_void_p()
        rc = self.L.pihm_ctx_create(C.byref(self.desc), C.byref(ctx))
        if rc != 0:
            raise RuntimeError(f"pihm_ctx_create failed ({rc})")
        self.ctx = ctx
        if closure and self.L.pihm_ctx_set_closure(ctx, 1):
            raise RuntimeError("pihm_ctx_set_closure failed")
        self.neq = self.L.pihm_ctx_neq(ctx)
        self.y = self.L.N_VNew_Hip(self.neq)
        self.yd = self.L.N_VNew_Hip(self.neq)
        if not self.y or not self.yd:
            raise RuntimeError("N_VNew_Hip failed")
        self.L.pihm_ctx_set_nan_check(ctx, 1 if nan_check else 0)
        self.L.pihm_ctx_set_record(ctx, 1 if record else 0)
        self.record = record
        if rhs_path is not None:
            self.set_rhs_path(rhs_path)

    def set_rhs_path(self, path):
        """_lib.PIHM_RHS_ONE_LAUNCH (the default) or _lib.PIHM_RHS_TWO_KERNELS."""
        if self.L.pihm_ctx_set_rhs_path(self.ctx, int(path)):
            raise RuntimeError(f"pihm_ctx_set_rhs_path({path}) failed")

    def rhs_path(self):
        return self.L.pihm_ctx_get_rhs_path(self.ctx)

    def close(self):
        if getattr(self, "ctx", None):
            self.L.N_VDestroy_Hip(self.y)
            self.L.N_VDestroy_Hip(self.yd)
            self.L.pihm_ctx_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def perm(self):
        p = np.zeros(self.mesh.nelem, np.int32)
        self.L.pihm_ctx_get_perm(self.ctx, p.ctypes.data_as(_lib.c_int32_p))
        return p

    def set_step_inputs(self, pcpdrp=None, edir=None, ett=None, ws0_surf=None, elem_bc=None,
                        river_bc=None):
        keep = []
        s = _lib.StepInputs()
        for name, arr in (("pcpdrp", pcpdrp), ("edir", edir), ("ett", ett), ("ws0_surf", ws0_surf),
                          ("elem_bc", elem_bc), ("river_bc", river_bc)):
            if arr is not None:
                arr = np.ascontiguousarray(arr, dtype=np.float64)
                keep.append(arr)
                setattr(s, name, _dp(arr))
        rc = self.L.pihm_ctx_set_step_inputs(self.ctx, C.byref(s))
        if rc:
            raise RuntimeError(f"set_step_inputs failed ({rc})")

    def get_step_inputs(self):
        """(pcpdrp, edir, ett, ws0_surf, elem_bc [nelem, 3], river_bc) as the context holds
        them (pihm_ctx_get_step_inputs), caller order."""
        ne, nr = self.mesh.nelem, self.mesh.nriver
        out = [np.empty(ne) for _ in range(4)] + [np.empty(3 * ne), np.empty(max(nr, 1))]
        if self.L.pihm_ctx_get_step_inputs(self.ctx, *[_dp(a) for a in out]):
            raise RuntimeError("pihm_ctx_get_step_inputs failed")
        return out[0], out[1], out[2], out[3], out[4].reshape(ne, 3), out[5][:nr]

    def set_noah(self, smcmin):
        """Switch on the reference's _NOAH_ branches of EtExtract and Infil with each
        element's smcmin (pihm_ctx_set_noah); None switches them off."""
        if smcmin is None:
            rc = self.L.pihm_ctx_set_noah(self.ctx, None)
        else:
            a = np.ascontiguousarray(smcmin, dtype=np.float64)
            rc = self.L.pihm_ctx_set_noah(self.ctx, _dp(a))
        if rc:
            raise RuntimeError(f"pihm_ctx_set_noah failed ({rc})")

    def set_noah_step(self, sh2o0=None, fcr=None, et=None, nwtbl=None, nroot=None):
        """The Noah step's hand-over (pihm_ctx_set_noah_step): top-layer liquid water,
        frozen-soil factor, and per-layer transpiration et[nelem, nlayer] with 1-based
        water-table and root layers.  Call after set_step_inputs (GwTransp reads ett)."""
        f = lambda a: None if a is None else np.ascontiguousarray(a, dtype=np.float64)
        i = lambda a: None if a is None else np.ascontiguousarray(a, dtype=np.int32)
        sh2o0, fcr, et, nwtbl, nroot = f(sh2o0), f(fcr), f(et), i(nwtbl), i(nroot)
        nl = 0 if et is None else et.shape[1]
        rc = self.L.pihm_ctx_set_noah_step(
            self.ctx, None if sh2o0 is None else _dp(sh2o0), None if fcr is None else _dp(fcr),
            None if et is None else _dp(et), nl,
            None if nwtbl is None else nwtbl.ctypes.data_as(_lib.c_int32_p),
            None if nroot is None else nroot.ctypes.data_as(_lib.c_int32_p))
        if rc:
            raise RuntimeError(f"pihm_ctx_set_noah_step failed ({rc})")
        return rc

    def ode(self, y, t=0.0):
        """ydot = f(t, y) (caller order).  Returns (rc, ydot); rc 1 = NaN in ydot."""
        y = np.ascontiguousarray(y, dtype=np.float64)
        if len(y) < self.neq:
            y = np.concatenate([y, np.zeros(self.neq - len(y))])
        if self.L.pihm_ctx_state_to_device(self.ctx, _dp(y), self.y):
            raise RuntimeError("state_to_device failed")
        rc = self.L.PIHM_ODE_Hip(t, self.y, self.yd, self.ctx)
        out = np.empty(self.neq)
        if self.L.pihm_ctx_state_from_device(self.ctx, self.yd, _dp(out)):
            raise RuntimeError("state_from_device failed")
        return rc, out

    def get_river_edge_ovlflow(self):
        o = np.zeros(3 * self.mesh.nelem)
        self.L.pihm_ctx_get_river_edge_ovlflow(self.ctx, _dp(o))
        return o

    def set_river_edge_ovlflow(self, ovl3):
        o = np.ascontiguousarray(ovl3, dtype=np.float64)
        self.L.pihm_ctx_set_river_edge_ovlflow(self.ctx, _dp(o))

    def fluxes(self):
        ne, nr = self.mesh.nelem, self.mesh.nriver
        ovl, sub = np.zeros(3 * ne), np.zeros(3 * ne)
        infil, rechg = np.zeros(ne), np.zeros(ne)
        rf = np.zeros(11 * max(nr, 1))
        rc = self.L.pihm_ctx_get_fluxes(self.ctx, _dp(ovl), _dp(sub), _dp(infil), _dp(rechg), _dp(rf))
        if rc:
            raise RuntimeError("flux read-back needs record=True")
        return dict(ovlflow=ovl, subsurf=sub, infil=infil, rechg=rechg, rivflow=rf[:11 * nr])


def run_gpu(mesh, forcing, cfg, y0, closure=False):
    """Integrate on the device with the model-step driver (pihm_run_gpu); returns (rc, y, stats).
    With closure=True the state carries the 4 water-balance accumulators after the
    reference's entries (y0 may include them or not; they start at 0 if not)."""
    L = _lib.lib()
    desc = mesh.desc()
    ctx = C.c_void_p()
    rc = L.pihm_ctx_create(C.byref(desc), C.byref(ctx))
    if rc:
        raise RuntimeError(f"pihm_ctx_create failed ({rc})")
    try:
        y = np.ascontiguousarray(y0, dtype=np.float64).copy()
        if closure:
            if L.pihm_ctx_set_closure(ctx, 1):
                raise RuntimeError("pihm_ctx_set_closure failed")
            if len(y) == mesh.neq:
                y = np.concatenate([y, np.zeros(4)])
        st = _lib.RunStats()
        fs = forcing.struct()
        rc = L.pihm_run_gpu(ctx, C.byref(fs), C.byref(cfg), _dp(y), C.byref(st))
        return rc, y, st
    finally:
        L.pihm_ctx_destroy(ctx)
