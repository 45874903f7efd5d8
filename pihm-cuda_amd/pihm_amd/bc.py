"""Host-side mirror of the reference's boundary-condition series over the HIP library.

Every model step the reference interpolates each boundary-condition series to the current
time and copies the values into the element edges' and river segments' bc fields
(ApplyBc, src/forcing.c:8-37; ApplyElemBc, src/forcing.c:127-177; ApplyRiverBc,
src/forcing.c:214-244).  ``HipBC(rhs, elem_type, elem_series, river_type, river_series)``
does it on the device through the C ABI (``pihm_bc_*``) into the RHS context.

``elem_type`` is [nelem, 3] (the .att file's BC0..BC2 columns): k > 0 takes head from
series k, k < 0 flux from series -k, 0 nothing; ``river_type`` is [nriver] (the .riv
file's BC column).  A series set is ``(ftime[nts, nrec], value[nts, nrec])`` or None.
"""
import ctypes as C

import numpy as np

from . import _lib


def _dp(a):
    return a.ctypes.data_as(_lib.c_double_p)


def ts_struct(series, keep):
    """pihm_ts over contiguous copies kept alive in ``keep`` (None -> None)."""
    if series is None:
        return None
    ft = np.ascontiguousarray(series[0], dtype=np.float64)
    v = np.ascontiguousarray(series[1], dtype=np.float64)
    if ft.ndim != 2 or v.shape != ft.shape:
        raise ValueError("a series set is ftime[nts, nrec] and value[nts, nrec]")
    keep += [ft, v]
    s = _lib.Ts()
    s.nts, s.nrec = ft.shape
    s.ftime, s.value = _dp(ft), _dp(v)
    return s


def code_array(a, n, keep):
    if a is None:
        return None
    a = np.ascontiguousarray(np.asarray(a, dtype=np.int32).reshape(n))
    keep.append(a)
    return a.ctypes.data_as(_lib.c_int32_p)


class HipBC:
    """Boundary series bound to a single-solve ``HipRHS`` context."""

    def __init__(self, rhs, elem_type=None, elem_series=None, river_type=None, river_series=None):
        self.L = _lib.lib()
        self.rhs = rhs
        m = rhs.mesh
        keep = []
        es, rs = ts_struct(elem_series, keep), ts_struct(river_series, keep)
        h = C.c_void_p()
        rc = self.L.pihm_bc_create(rhs.ctx, code_array(elem_type, 3 * m.nelem, keep),
                                   C.byref(es) if es is not None else None,
                                   code_array(river_type, m.nriver, keep),
                                   C.byref(rs) if rs is not None else None, C.byref(h))
        if rc:
            raise RuntimeError(f"pihm_bc_create failed ({rc})")
        self.h = h

    def apply(self, t):
        """Returns the C ABI's code (PIHM_ERR_ARG when t is outside a series)."""
        return self.L.pihm_bc_apply(self.h, float(t))

    def values(self):
        """The context's boundary values now: (elem_bc[nelem, 3], river_bc[nriver])."""
        m = self.rhs.mesh
        eb, rb = np.zeros(3 * m.nelem), np.zeros(max(m.nriver, 1))
        if self.L.pihm_ctx_get_step_inputs(self.rhs.ctx, None, None, None, None, _dp(eb), _dp(rb)):
            raise RuntimeError("pihm_ctx_get_step_inputs failed")
        return eb.reshape(m.nelem, 3), rb[:m.nriver]

    def close(self):
        if getattr(self, "h", None):
            self.L.pihm_bc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
