"""The RHS with the reference's _NOAH_ branches (pihm_ctx_set_noah / pihm_ctx_set_noah_step)
against the oracle's, per call, on MI355X: ydot within 1e-12 of the terms' magnitude, and
the recorded infiltration and recharge likewise.  The Noah hand-over is drawn to cover
every branch: top-layer saturation below SATMIN and above 1, frozen-soil factors 0, 1 and
between, water-table layers above, inside and below the root zone (including nwtbl <= 0),
negative layer transpiration and non-positive ett."""
import numpy as np
import pytest

import oracle as O
from conftest import assert_rhs_close
from pihm_amd.rhs import HipRHS
from test_noah_pins import _noah_inputs
from test_oracle_pins import _inputs, _states

pytestmark = pytest.mark.gpu
RTOL = 1e-12


def _setup(m, rng, nlayer=11):
    h, o = HipRHS(m, record=True), O.OracleRHS(m)
    inp = _inputs(m, rng)
    inp["ett"][rng.random(m.nelem) < 0.1] = 0.0
    inp["ett"][rng.random(m.nelem) < 0.05] = -1e-9
    nz = _noah_inputs(m, rng, nlayer)
    h.set_step_inputs(**inp)
    o.set_step_inputs(**inp)
    h.set_noah(nz["smcmin"])
    h.set_noah_step(nz["sh2o0"], nz["fcr"], nz["et"], nz["nwtbl"], nz["nroot"])
    o.set_noah(**nz)
    return h, o, nz


def _check(h, o, y, what):
    rc_h, dh = h.ode(y)
    rc_o, do, mag = o.ode(y, with_mag=True)
    assert rc_h == rc_o
    assert_rhs_close(dh, do, mag, RTOL, what)
    f = h.fluxes()
    for k in ("infil", "rechg"):
        ref = o.field(k)
        err = np.abs(f[k] - ref)
        assert np.all(err <= RTOL * np.maximum(np.abs(ref), 1e-30)), (what, k, err.max())


@pytest.mark.parametrize("which", ["c0", "tree_small", "example"])
def test_noah_rhs_parity(which, request):
    m = request.getfixturevalue(which)
    print(f"[noah parity] {which} ne={m.nelem}", flush=True)
    rng = np.random.default_rng(41)
    h, o, _ = _setup(m, rng)
    for k, y in enumerate(_states(m, rng, 5)):
        _check(h, o, y, f"{which} noah state {k}")
    h.close()


def test_noah_step_updates_and_switch_off(c0):
    """A second hand-over replaces the first (fewer layers this time), and switching
    Noah off gives the plain RHS again; the carried river-edge flux runs through all."""
    m = c0
    rng = np.random.default_rng(42)
    h, o, nz0 = _setup(m, rng)
    ys = _states(m, rng, 3)
    _check(h, o, ys[1], "first hand-over")
    nz = _noah_inputs(m, rng, nlayer=4)
    nz["smcmin"] = nz0["smcmin"]          # static: set once with pihm_ctx_set_noah
    h.set_noah_step(nz["sh2o0"], nz["fcr"], nz["et"], nz["nwtbl"], nz["nroot"])
    o.set_noah(**nz)
    _check(h, o, ys[2], "second hand-over")
    # fcr only: the rest stays
    fcr = rng.uniform(0.0, 1.0, m.nelem)
    h.set_noah_step(fcr=fcr)
    o.set_noah(nz["smcmin"], fcr=fcr)
    _check(h, o, ys[3], "fcr only")
    h.set_noah(None)
    o.set_noah(None)
    _check(h, o, ys[1], "noah off")
    h.close()


def test_noah_argument_errors(c0):
    from pihm_amd import _lib
    L = _lib.lib()
    h = HipRHS(c0)
    ne = c0.nelem
    one = np.ones(ne)
    assert L.pihm_ctx_set_noah_step(h.ctx, one.ctypes.data_as(_lib.c_double_p), None, None, 0, None, None) \
        == _lib.PIHM_ERR_ARG                         # Noah not switched on
    h.set_noah(np.full(ne, 0.05))
    et = np.zeros((ne, 12))
    nw, nr = np.ones(ne, np.int32), np.ones(ne, np.int32)
    with pytest.raises(RuntimeError):
        h.set_noah_step(et=et, nwtbl=nw, nroot=nr)   # 12 layers
    with pytest.raises(RuntimeError):
        h.set_noah_step(et=np.zeros((ne, 3)), nwtbl=nw, nroot=np.full(ne, 4, np.int32))   # root below the layers
    with pytest.raises(RuntimeError):
        h.set_noah_step(et=np.zeros((ne, 3)), nwtbl=nw)                                  # nroot missing
    h.close()


def test_noah_rejected_hand_over_leaves_context_unchanged(c0):
    """A hand-over refused for a root layer past the given layers uploads nothing: the
    RHS after it gives the bits it gave before, top-layer water and fcr included."""
    m = c0
    rng = np.random.default_rng(43)
    h, o, nz = _setup(m, rng)
    y = _states(m, rng, 2)[1]
    hid0 = h.get_river_edge_ovlflow()
    before = h.ode(y)[1]
    bad = _noah_inputs(m, rng, 3)
    bad["nroot"][m.nelem // 2] = 4
    with pytest.raises(RuntimeError):
        h.set_noah_step(bad["sh2o0"], bad["fcr"], bad["et"], bad["nwtbl"], bad["nroot"])
    h.set_river_edge_ovlflow(hid0)
    assert np.array_equal(h.ode(y)[1], before)
    h.close()
