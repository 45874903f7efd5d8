"""Device N_Vector operations against the oracle's restatement of nvector_serial.

Element-wise operations must agree bit for bit (including N_VLinearSum's special cases,
cvode/src/nvec_ser/nvector_serial.c:421-506).  Reductions differ from the serial
left-to-right sum only in summation order: |dev - ser| <= 1e-14 * sum of |terms|.
Max/min reductions are exact.  The host vector is also pinned against numpy (CPU test).
"""
import ctypes as C
import math

import numpy as np
import pytest

import oracle as O
from pihm_amd import _lib

FN = C.CFUNCTYPE
vp = C.c_void_p
d = C.c_double


class Ops(C.Structure):
    _fields_ = [
        ("getid", FN(C.c_int, vp)), ("clone", FN(vp, vp)), ("cloneempty", FN(vp, vp)),
        ("destroy", FN(None, vp)), ("space", FN(None, vp, vp, vp)), ("getarray", FN(vp, vp)),
        ("setarray", FN(None, vp, vp)), ("linearsum", FN(None, d, vp, d, vp, vp)),
        ("const", FN(None, d, vp)), ("prod", FN(None, vp, vp, vp)), ("div", FN(None, vp, vp, vp)),
        ("scale", FN(None, d, vp, vp)), ("abs", FN(None, vp, vp)), ("inv", FN(None, vp, vp)),
        ("addconst", FN(None, vp, d, vp)), ("dot", FN(d, vp, vp)), ("maxnorm", FN(d, vp)),
        ("wrms", FN(d, vp, vp)), ("wrmsmask", FN(d, vp, vp, vp)), ("min", FN(d, vp)),
        ("wl2", FN(d, vp, vp)), ("l1", FN(d, vp)), ("compare", FN(None, d, vp, vp)),
        ("invtest", FN(C.c_int, vp, vp)), ("constrmask", FN(C.c_int, vp, vp, vp)),
        ("minquotient", FN(d, vp, vp)),
    ]


class NV(C.Structure):
    _fields_ = [("content", vp), ("ops", C.POINTER(Ops))]


def ops_of(v):
    return C.cast(v, C.POINTER(NV)).contents.ops.contents


class HostVecs:
    def __init__(self, n, k):
        self.L = O.lib()
        self.v = [self.L.oracle_nv_new(n, 1) for _ in range(k)]
        self.n = n

    def arr(self, i):
        return np.ctypeslib.as_array(C.cast(self.L.oracle_nv_data(self.v[i]), _lib.c_double_p), shape=(self.n,))

    def set(self, i, a):
        self.arr(i)[:] = a

    def get(self, i):
        return self.arr(i).copy()


def test_host_vector_against_numpy():
    n = 1000
    rng = np.random.default_rng(0)
    H = HostVecs(n, 3)
    x, y = rng.normal(size=n), rng.normal(size=n)
    H.set(0, x)
    H.set(1, y)
    o = ops_of(H.v[0])
    o.linearsum(0.3, H.v[0], -1.7, H.v[1], H.v[2])
    assert np.array_equal(H.get(2), 0.3 * x + (-1.7) * y)
    o.linearsum(0.5, H.v[0], 0.5, H.v[1], H.v[2])            # a == b: a*(x+y)
    assert np.array_equal(H.get(2), 0.5 * (x + y))
    assert o.dot(H.v[0], H.v[1]) == pytest.approx(math.fsum(x * y), rel=1e-13)
    assert o.wrms(H.v[0], H.v[1]) == pytest.approx(math.sqrt(math.fsum((x * y) ** 2) / n), rel=1e-13)
    assert o.min(H.v[0]) == x.min() and o.maxnorm(H.v[0]) == np.abs(x).max()


@pytest.mark.gpu
def test_device_vector_ops_match_serial():
    L = _lib.lib()
    n = 200_003
    rng = np.random.default_rng(1)
    X = [rng.normal(size=n) for _ in range(3)]
    X[1][::97] = 0.0
    w = rng.uniform(0.1, 2.0, n)
    H = HostVecs(n, 5)
    D = [L.N_VNew_Hip(n) for _ in range(5)]
    ho, do = ops_of(H.v[0]), ops_of(D[0])

    def load(i, a):
        H.set(i, a)
        assert L.N_VCopyToDevice_Hip(D[i], a.ctypes.data_as(_lib.c_double_p)) == 0

    def dget(i):
        a = np.empty(n)
        assert L.N_VCopyFromDevice_Hip(D[i], a.ctypes.data_as(_lib.c_double_p)) == 0
        return a

    for i in range(3):
        load(i, X[i])
    load(3, w)
    cases = [(1.0, 1.0), (1.0, -1.0), (-1.0, 1.0), (0.7, 1.0), (1.0, 0.7), (-1.0, 0.3), (0.3, -1.0),
             (0.25, 0.25), (0.25, -0.25), (1.3, -0.2), (-1.0, -1.0)]
    for a, b in cases:
        ho.linearsum(a, H.v[0], b, H.v[1], H.v[4])
        do.linearsum(a, D[0], b, D[1], D[4])
        assert np.array_equal(dget(4), H.get(4)), (a, b)
    # in-place forms (axpy paths of the serial vector)
    for a, b, zi in [(0.4, 1.0, 1), (1.0, 0.4, 0), (-1.0, 1.0, 1), (1.0, -1.0, 0)]:
        load(0, X[0]); load(1, X[1])
        ho.linearsum(a, H.v[0], b, H.v[1], H.v[zi])
        do.linearsum(a, D[0], b, D[1], D[zi])
        assert np.array_equal(dget(zi), H.get(zi)), (a, b, zi)
    load(0, X[0]); load(1, X[1])
    for name, args_h, args_d in [
        ("prod", (H.v[0], H.v[1], H.v[4]), (D[0], D[1], D[4])),
        ("div", (H.v[0], H.v[3], H.v[4]), (D[0], D[3], D[4])),
        ("scale", (2.5, H.v[0], H.v[4]), (2.5, D[0], D[4])),
        ("scale", (-1.0, H.v[0], H.v[4]), (-1.0, D[0], D[4])),
        ("scale", (1.0, H.v[0], H.v[4]), (1.0, D[0], D[4])),
        ("abs", (H.v[0], H.v[4]), (D[0], D[4])),
        ("inv", (H.v[3], H.v[4]), (D[3], D[4])),
        ("addconst", (H.v[0], 0.125, H.v[4]), (D[0], 0.125, D[4])),
        ("compare", (0.5, H.v[0], H.v[4]), (0.5, D[0], D[4])),
        ("const", (3.25, H.v[4]), (3.25, D[4])),
    ]:
        getattr(ho, name)(*args_h)
        getattr(do, name)(*args_d)
        assert np.array_equal(dget(4), H.get(4)), name
    # reductions: both sums carry rounding error; the serial left-to-right sum is bounded
    # by (n-1)*u*sum|terms| (Higham, Accuracy and Stability, eq. 4.4), the device tree by
    # log2(n)*u*sum|terms|; their difference must stay inside the sum of both bounds.
    u = 2.0 ** -53
    bound = (n + math.log2(n)) * u
    terms = {"dot": np.abs(X[0] * X[1]).sum(), "l1": np.abs(X[0]).sum(),
             "wsq": ((X[0] * w) ** 2).sum()}
    assert abs(do.dot(D[0], D[1]) - ho.dot(H.v[0], H.v[1])) <= bound * terms["dot"]
    assert abs(do.l1(D[0]) - ho.l1(H.v[0])) <= bound * terms["l1"]
    assert abs(do.wl2(D[0], D[3]) ** 2 - ho.wl2(H.v[0], H.v[3]) ** 2) <= 1.01 * bound * terms["wsq"]
    assert abs(do.wrms(D[0], D[3]) ** 2 - ho.wrms(H.v[0], H.v[3]) ** 2) * n <= 1.01 * bound * terms["wsq"]
    load(2, (X[2] > 0).astype(float))
    assert do.wrmsmask(D[0], D[3], D[2]) == pytest.approx(ho.wrmsmask(H.v[0], H.v[3], H.v[2]), rel=bound)
    assert do.maxnorm(D[0]) == ho.maxnorm(H.v[0])
    assert do.min(D[0]) == ho.min(H.v[0])
    assert do.minquotient(D[0], D[1]) == ho.minquotient(H.v[0], H.v[1])
    assert do.invtest(D[1], D[4]) == ho.invtest(H.v[1], H.v[4]) == 0
    assert do.invtest(D[3], D[4]) == ho.invtest(H.v[3], H.v[4]) == 1
    assert np.array_equal(dget(4), H.get(4))
    c = np.where(np.arange(n) % 3 == 0, 2.0, np.where(np.arange(n) % 3 == 1, -1.0, 0.0))
    load(2, c)
    assert do.constrmask(D[2], D[0], D[4]) == ho.constrmask(H.v[2], H.v[0], H.v[4])
    assert np.array_equal(dget(4), H.get(4))
    # determinism of reductions
    vals = {do.dot(D[0], D[1]) for _ in range(5)}
    assert len(vals) == 1
    # fused: dot-multi equals separate dots bitwise; linear combination
    arr = (C.c_void_p * 3)(D[0], D[1], D[3])
    res = np.zeros(3)
    assert L.N_VDotProdMulti_Hip(3, D[0], arr, res.ctypes.data_as(_lib.c_double_p)) == 0
    assert res[0] == do.dot(D[0], D[0]) and res[1] == do.dot(D[0], D[1])
    for v in D:
        L.N_VDestroy_Hip(v)
