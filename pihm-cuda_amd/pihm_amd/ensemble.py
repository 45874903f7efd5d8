"""Many MM-PIHM solves advanced together on one device (pihm_ensemble.h).

``Ensemble(meshes, forcings, cfg)`` builds one batch context over the meshes and one
ensemble over it; every member has its own forcing, initial state and integrator.
"""
import ctypes as C

import numpy as np

from . import _lib


def _dp(a):
    return a.ctypes.data_as(_lib.c_double_p)


class Ensemble:
    def __init__(self, meshes, forcings, cfg):
        if len(meshes) != len(forcings) or not meshes:
            raise ValueError("one forcing per mesh")
        self.L = _lib.lib()
        self.meshes = list(meshes)
        self.n = len(meshes)
        self._descs = [m.desc() for m in meshes]
        arr = (C.POINTER(_lib.MeshDesc) * self.n)(*[C.pointer(d) for d in self._descs])
        self.ctx = C.c_void_p()
        rc = self.L.pihm_ctx_create_batch(self.n, arr, C.byref(self.ctx))
        if rc:
            raise RuntimeError(f"pihm_ctx_create_batch failed ({rc})")
        self._forcings = [f.for_mesh(m) for f, m in zip(forcings, meshes)]
        self._fs = (_lib.Forcing * self.n)(*[f.struct() for f in self._forcings])
        self.cfg = cfg
        self.ens = C.c_void_p()
        rc = self.L.pihm_ens_create(self.ctx, self._fs, C.byref(cfg), C.byref(self.ens))
        if rc:
            self.L.pihm_ctx_destroy(self.ctx)
            self.ctx = None
            raise RuntimeError(f"pihm_ens_create failed ({rc})")

    def set_state(self, k, y):
        y = np.ascontiguousarray(y, dtype=np.float64)
        if len(y) != self.meshes[k].neq:
            raise ValueError("state length")
        rc = self.L.pihm_ens_set_state(self.ens, k, _dp(y))
        if rc:
            raise RuntimeError(f"pihm_ens_set_state failed ({rc})")

    def get_state(self, k):
        y = np.empty(self.meshes[k].neq)
        rc = self.L.pihm_ens_get_state(self.ens, k, _dp(y))
        if rc:
            raise RuntimeError(f"pihm_ens_get_state failed ({rc})")
        return y

    def enable(self, k, on=True):
        return self.L.pihm_ens_set_member_enabled(self.ens, k, 1 if on else 0)

    def run(self, t_end):
        """Advance every enabled member to solver time t_end; returns the number of members
        that failed during the call (negative: an error of the call itself)."""
        return self.L.pihm_ens_run(self.ens, float(t_end))

    def stats(self, k):
        st = _lib.RunStats()
        self.L.pihm_ens_member_stats(self.ens, k, C.byref(st))
        return st

    def autotune(self, reps=20):
        """Pick the launch shapes (pihm_ens_autotune) over the ready members; returns the
        choices and the timings of every candidate."""
        t = _lib.EnsTuning()
        rc = self.L.pihm_ens_autotune(self.ens, int(reps), C.byref(t))
        if rc:
            raise RuntimeError(f"pihm_ens_autotune failed ({rc})")
        return t.as_dict()

    def set_tuning(self, tile_span, ops_chunks_per_wg):
        rc = self.L.pihm_ens_set_tuning(self.ens, int(tile_span), int(ops_chunks_per_wg))
        if rc:
            raise RuntimeError(f"pihm_ens_set_tuning failed ({rc})")

    def counters(self):
        a, b, c = C.c_long(), C.c_long(), C.c_long()
        self.L.pihm_ens_counters(self.ens, C.byref(a), C.byref(b), C.byref(c))
        return dict(ticks=a.value, op_launches=b.value, rhs_launches=c.value)

    def op_stats(self):
        """The vector-operation kernel's timed launches so far: count, algorithmic bytes,
        device ms (pihm_ens_op_stats)."""
        n, b, ms = C.c_long(), C.c_double(), C.c_double()
        self.L.pihm_ens_op_stats(self.ens, C.byref(n), C.byref(b), C.byref(ms))
        return dict(launches=n.value, bytes=b.value, ms=ms.value)

    def close(self):
        if getattr(self, "ens", None):
            self.L.pihm_ens_destroy(self.ens)
            self.ens = None
        if getattr(self, "ctx", None):
            self.L.pihm_ctx_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
