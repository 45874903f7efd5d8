"""The power routine of the device flux functions (x^y = exp(y ln x), built into the
library for host and device from one source) against the host libm pow.

The reference calls libm pow (src/soil.c:8-13, src/vert_flow.c:272-278,
src/lat_flow.c:279-283, src/river_flow.c:520-583).  The bounds pinned here: within a
few ulp for the exponents the flux functions use, exactly 1 for 1^y, exactly +0 for
(+0)^y with y > 0, never below 1 for x >= 1 and y >= 0 (the van Genuchten head takes
(1/s)^m - 1 to a fractional power), and the libm result itself wherever the fast path
does not apply.  The CPU tests run the host build; the GPU test checks that the device
kernel returns the same bits.
"""
import numpy as np
import pytest

from pihm_amd import _lib


def xpow(x, y, device=False):
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(np.broadcast_to(y, x.shape), dtype=np.float64)
    out = np.empty_like(x)
    rc = _lib.lib().pihm_math_pow(len(x), x.ctypes.data_as(_lib.c_double_p),
                                  y.ctypes.data_as(_lib.c_double_p),
                                  out.ctypes.data_as(_lib.c_double_p), 1 if device else 0)
    assert rc == 0
    return out


def _ulps(a, b):
    ia = a.view(np.int64)
    ib = b.view(np.int64)
    return np.abs(ia - ib)


def _cases(n=20000, seed=0):
    rng = np.random.default_rng(seed)
    beta = rng.uniform(1.05, 3.0, n)
    m = beta / (beta - 1.0)
    satn = rng.uniform(0.1, 1.0, n)
    return {
        # Manning depth^(2/3) with the reference's literal exponent
        "manning": (10 ** rng.uniform(-12, 2, n), np.full(n, 0.6666667), 2),
        # van Genuchten: s^m, (1/s)^m (m up to 21), x^(1/beta), x^((beta-1)/beta)
        "vg_s_m": (satn, m, 16),
        "vg_inv_s_m": (1.0 / satn, m, 16),
        "vg_root": (10 ** rng.uniform(-6, 6, n), 1.0 / beta, 2),
        "vg_root2": (rng.uniform(0.0, 1.0, n), (beta - 1.0) / beta, 2),
        # channel geometry: depth^(4/3), coeff^(1/3), coeff^(2/3), x^0.5
        "channel": (10 ** rng.uniform(-4, 3, n), rng.choice([4.0 / 3.0, 1.0 / 3.0, 2.0 / 3.0, 0.5], n), 2),
        # saturation within 1e-16..1e-1 of 1
        "near_one": (1.0 - 10 ** rng.uniform(-16, -1, n), rng.uniform(0.1, 20.0, n), 3),
        # the whole exponent range of x with moderate y
        "wide": (10 ** rng.uniform(-300, 300, n), rng.uniform(-2.0, 2.0, n), 3),
    }


@pytest.mark.parametrize("name", list(_cases(10)))
def test_xpow_ulp_bound_against_libm(name):
    x, y, bound = _cases()[name]
    got = xpow(x, y)
    with np.errstate(over="ignore"):
        ref = np.power(x, y)
    u = _ulps(got, ref)
    assert u.max() <= bound, (name, int(u.max()), x[np.argmax(u)], y[np.argmax(u)])
    # most results are the libm's own bits
    assert np.mean(u == 0) > 0.7


def test_xpow_exact_cases():
    ys = np.array([0.3, 0.6666667, 1.0 / 3.0, 2.0, 21.0, 1e-9, 0.0])
    assert np.all(xpow(np.ones_like(ys), ys) == 1.0)
    z = xpow(np.zeros_like(ys[:-1]), ys[:-1])
    assert np.all(z == 0.0) and not np.any(np.signbit(z))
    assert xpow(np.array([5.0]), 0.0)[0] == 1.0
    # kr_func(beta, 1): 1 - (1 - 1^m)^(1/m) == 1 exactly
    beta = np.random.default_rng(2).uniform(1.01, 5.0, 1000)
    m = beta / (beta - 1.0)
    inner = 1.0 - xpow(np.ones_like(m), m)
    assert np.all(inner == 0.0)
    assert np.all(1.0 - xpow(inner, 1.0 / m) == 1.0)


def test_xpow_not_below_one_for_bases_at_least_one():
    rng = np.random.default_rng(3)
    x = np.concatenate([1.0 + np.arange(1, 2000) * 2.0 ** -52, 1.0 / rng.uniform(0.1, 1.0, 20000),
                        np.nextafter(1.0, 2.0) * np.ones(10)])
    y = rng.uniform(1.0, 21.0, len(x))
    assert np.all(xpow(x, y) >= 1.0)
    # and not above one for bases in (0, 1]
    x2 = 1.0 - np.arange(0, 2000) * 2.0 ** -53
    assert np.all(xpow(x2, rng.uniform(0.1, 21.0, len(x2))) <= 1.0)


def test_xpow_outside_the_fast_path_is_libm():
    x = np.array([-2.0, -8.0, -0.0, -0.0, np.inf, np.nan, 2.0, 1e-300, 1e300, 0.0, 3.0])
    y = np.array([0.5, 1.0 / 3.0, 1.0, 0.5, 0.5, 2.0, np.inf, 3.0, -3.0, -1.0, np.nan])
    with np.errstate(all="ignore"):
        ref = np.power(x, y)
    got = xpow(x, y)
    same = (got == ref) | (np.isnan(got) & np.isnan(ref))
    assert np.all(same), (got, ref)
    assert np.array_equal(np.signbit(got[~np.isnan(got)]), np.signbit(ref[~np.isnan(ref)]))


@pytest.mark.gpu
def test_xpow_device_matches_host_bitwise():
    """Same bits on host and device on the fast path (finite x > 0, |y ln x| < 708);
    outside it each side calls its own library pow, which agree to an ulp or two."""
    parts = [np.stack([x, np.broadcast_to(y, x.shape)]) for x, y, _ in _cases(50000, seed=7).values()]
    xy = np.concatenate(parts, axis=1)
    host = xpow(xy[0], xy[1])
    dev = xpow(xy[0], xy[1], device=True)
    with np.errstate(divide="ignore"):
        fast = (xy[0] > 0) & np.isfinite(xy[0]) & (np.abs(xy[1] * np.log(xy[0])) < 700.0)
    assert fast.sum() > 0.9 * len(fast)
    assert np.array_equal(host[fast].view(np.int64), dev[fast].view(np.int64))
    slow = ~fast & np.isfinite(host)
    assert np.all(_ulps(host[slow], dev[slow]) <= 2)
