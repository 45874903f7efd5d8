"""pihm_nv_dotprod_noalloc: N_VDotProd_Hip on raw device arrays with the library's fixed
device workspace (no allocation).  Two calls in a row, the second on a smaller input,
each equal bit for bit to N_VDotProd_Hip on the same vectors (same summation tree), and
within 1e-14 of the terms of an exactly rounded sum; argument errors launch nothing."""
import ctypes as C
import math

import numpy as np
import pytest

from pihm_amd import _lib

pytestmark = pytest.mark.gpu


def _dev(L, a):
    v = L.N_VNew_Hip(len(a))
    assert v
    a = np.ascontiguousarray(a, dtype=np.float64)
    assert L.N_VCopyToDevice_Hip(v, a.ctypes.data_as(_lib.c_double_p)) == 0
    return v


def test_dotprod_noalloc_matches_dotprod():
    L = _lib.lib()
    rng = np.random.default_rng(2026)
    out = L.N_VNew_Hip(1)
    res = np.zeros(1)
    for n in (1_000_003, 1_234, 0):
        x = rng.standard_normal(n) * rng.uniform(0.1, 10.0, n) if n else np.zeros(0)
        y = rng.standard_normal(n) if n else np.zeros(0)
        vx, vy = _dev(L, x), _dev(L, y)
        print(f"[noalloc] n={n}", flush=True)
        want = L.N_VDotProd_Hip(vx, vy)
        rc = L.pihm_nv_dotprod_noalloc(n, L.N_VGetDeviceArrayPointer_Hip(vx), L.N_VGetDeviceArrayPointer_Hip(vy),
                                       L.N_VGetDeviceArrayPointer_Hip(out), None)
        assert rc == 0
        assert L.N_VCopyFromDevice_Hip(out, res.ctypes.data_as(_lib.c_double_p)) == 0
        assert res[0] == want or (n == 0 and res[0] == 0.0), (n, res[0], want)
        exact = math.fsum((x * y).tolist())
        assert abs(res[0] - exact) <= 1e-14 * float(np.sum(np.abs(x * y))) + 0.0
        L.N_VDestroy_Hip(vx)
        L.N_VDestroy_Hip(vy)
    # argument errors: nothing launched, result untouched
    res_before = res.copy()
    p = L.N_VGetDeviceArrayPointer_Hip(out)
    assert L.pihm_nv_dotprod_noalloc(-1, p, p, p, None) == -1
    assert L.pihm_nv_dotprod_noalloc((1 << 33) + 1, p, p, p, None) == -1
    assert L.pihm_nv_dotprod_noalloc(4, None, p, p, None) == -1
    assert L.pihm_nv_dotprod_noalloc(4, p, p, None, None) == -1
    assert L.N_VCopyFromDevice_Hip(out, res.ctypes.data_as(_lib.c_double_p)) == 0
    assert np.array_equal(res, res_before)
    L.N_VDestroy_Hip(out)
