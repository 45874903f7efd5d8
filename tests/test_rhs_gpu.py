"""Parity of the HIP RHS (through the C ABI, PIHM_ODE_Hip) with the CPU oracle.

Tolerance: per component |hip - oracle| <= 1e-12 * max(|oracle|, sum of |terms| of that
component) (see conftest.assert_rhs_close).  Only the device pow/log may round differently
from glibc; every other operation is in the reference's order with no FMA contraction.
Renumbering maps are exact: states round-trip bit for bit.
"""
import sys

import numpy as np
import pytest

import oracle as O
from conftest import assert_rhs_close
from pihm_amd.meshgen import make_watershed
from pihm_amd.rhs import HipRHS
from test_oracle_pins import _inputs, _states

pytestmark = pytest.mark.gpu
RTOL = 1e-12


def _pair(m, record=False):
    return HipRHS(m, record=record), O.OracleRHS(m)


def _check(h, o, y, t=0.0, what=""):
    rc_h, dh = h.ode(y, t)
    rc_o, do, mag = o.ode(y, t, with_mag=True)
    assert rc_h == rc_o
    assert_rhs_close(dh, do, mag, RTOL, what)
    return dh, do


def test_state_permutation_round_trip(c0):
    h = HipRHS(c0)
    p = h.perm()
    assert np.array_equal(np.sort(p), np.arange(c0.nelem))
    y = np.random.default_rng(0).normal(size=c0.neq)
    import ctypes as C
    from pihm_amd import _lib
    L = _lib.lib()
    out = np.empty_like(y)
    assert L.pihm_ctx_state_to_device(h.ctx, y.ctypes.data_as(_lib.c_double_p), h.y) == 0
    assert L.pihm_ctx_state_from_device(h.ctx, h.y, out.ctypes.data_as(_lib.c_double_p)) == 0
    assert np.array_equal(out, y)
    # device order really is the Hilbert order (a non-trivial permutation)
    assert not np.array_equal(p, np.arange(c0.nelem))
    h.close()


@pytest.mark.parametrize("which", ["c0", "tree_small", "example"])
def test_rhs_parity_states(which, request):
    m = request.getfixturevalue(which)
    print(f"[parity] {which} ne={m.nelem} nr={m.nriver}", flush=True)
    rng = np.random.default_rng(11)
    h, o = _pair(m)
    inp = _inputs(m, rng)
    h.set_step_inputs(**inp)
    o.set_step_inputs(**inp)
    for k, y in enumerate(_states(m, rng, 6)):
        _check(h, o, y, what=f"{which} state {k}")
    h.close()


def test_rhs_parity_call_sequence_carries_river_edge_flux(tree_small):
    """Same call sequence on both sides: the river-edge overland flux written in call k
    is read by Infil in call k+1 (src/river_flow.c:159-182 then src/vert_flow.c:59-62)."""
    m = tree_small
    rng = np.random.default_rng(12)
    h, o = _pair(m)
    inp = _inputs(m, rng)
    h.set_step_inputs(**inp)
    o.set_step_inputs(**inp)
    ys = _states(m, rng, 4)
    for y in ys:
        y[3 * m.nelem:3 * m.nelem + m.nriver] *= 2.0    # rivers over their banks
    for k, y in enumerate(ys + ys[::-1]):
        _check(h, o, y, what=f"call {k}")
    hid_o = o.field("ovlflow").reshape(-1, 3)
    hid_h = h.get_river_edge_ovlflow().reshape(-1, 3)
    riv = m.a["nabr"] < 0
    np.testing.assert_allclose(hid_h[riv], hid_o[riv], rtol=1e-12, atol=0)
    # set/get of the carried flux round-trips and drives Infil the same way
    h.set_river_edge_ovlflow(np.where(riv, 0.0, 0.0).ravel())
    o.field("ovlflow")[:] = 0.0
    _check(h, o, ys[1], what="after reset")
    h.close()


def _variant(kind):
    m = make_watershed(3000, seed=21, n_seg=80, network="tree")
    rv = m.rivers
    nr = m.nriver
    if kind == "shapes":
        rv["intrpl_ord"] = (1 + np.arange(nr) % 4).astype(np.int32)
        rv["coeff"] = np.array([3.0, 0.8, 0.6, 1.5])[np.arange(nr) % 4]
    elif kind == "kinematic":
        m.surf_mode = m.riv_mode = 1
    elif kind == "bc":
        bnd = m.nabr_raw == 0
        codes = np.where(np.random.default_rng(3).random(bnd.shape) < 0.5, 1, -1)
        m.bc_type = np.where(bnd, codes, 0).astype(np.int32)
        leaves = np.setdiff1d(np.arange(nr), rv["down"][rv["down"] > 0] - 1)
        bt = np.zeros(nr, np.int32)
        bt[leaves[::2]] = 1
        bt[leaves[1::2]] = -1
        rv["bc_type"] = bt
    elif kind.startswith("outlet"):
        code = int(kind[6:])
        rv["down"] = np.where(rv["down"] < 0, code, rv["down"]).astype(np.int32)
    m.a = None
    return m.init()


@pytest.mark.parametrize("kind", ["shapes", "kinematic", "bc", "outlet-1", "outlet-2", "outlet-4"])
def test_rhs_parity_branch_variants(kind):
    m = _variant(kind)
    print(f"[parity] variant {kind}", flush=True)
    rng = np.random.default_rng(13)
    h, o = _pair(m)
    inp = _inputs(m, rng)
    ne, nr = m.nelem, m.nriver
    ebc = np.where(m.a["bc_type"] > 0, (m.a["zmin"] + 0.7 * m.a["depth"])[:, None], 2e-4).ravel()
    # Dirichlet river heads below the bed: the reference's formula takes the square root
    # of the head gradient (src/river_flow.c:405-419, :348-362), so only outflow is defined.
    rbc = np.where(m.a["riverbc_type"] > 0, m.a["zbed"] - 0.5, 0.05)
    if kind.startswith("outlet"):
        rbc = np.where(m.a["down"] < 0, m.a["zbed"] - 0.5 if kind == "outlet-1" else 0.3, rbc)
    inp.update(elem_bc=ebc, river_bc=rbc)
    h.set_step_inputs(**inp)
    o.set_step_inputs(**inp)
    for k, y in enumerate(_states(m, rng, 4)):
        _check(h, o, y, what=f"{kind} state {k}")
    h.close()


def test_rhs_flux_readback_matches_oracle(tree_small):
    m = tree_small
    rng = np.random.default_rng(14)
    h, o = _pair(m, record=True)
    inp = _inputs(m, rng)
    h.set_step_inputs(**inp)
    o.set_step_inputs(**inp)
    y = _states(m, rng, 2)[2]
    _check(h, o, y)
    f = h.fluxes()
    for k in ("infil", "rechg", "ovlflow", "subsurf", "rivflow"):
        ref = o.field(k)
        scale = np.maximum(np.abs(ref), np.max(np.abs(ref)) * 1e-3)
        assert np.all(np.abs(f[k] - ref) <= 1e-11 * scale), k
    h.close()


def test_rhs_nan_reported_not_exited():
    """A NaN in ydot returns the recoverable code 1 where the reference exits
    (src/ode.c:401-410): a zero-depth-gradient outlet on an adverse bed slope."""
    m = make_watershed(535, seed=1, n_seg=20)
    r_out = int(np.nonzero(m.rivers["down"] < 0)[0][0])
    to = m.rivers["tonode"][r_out] - 1
    m.node_zmax[to] += 3.0
    m.node_zmin[to] += 3.0
    m.a = None
    m.init()
    h, o = _pair(m)
    y = m.initial_state()
    rc_o, _ = o.ode(y)
    rc_h, _ = h.ode(y)
    assert rc_o == 1 and rc_h == 1
    # the flag is cleared: a following healthy call returns 0
    h2 = HipRHS(make_watershed(535, seed=1, n_seg=20))
    assert h2.ode(h2.mesh.initial_state())[0] == 0
    h.close()
    h2.close()


def test_rhs_deterministic_bitwise(tree_small):
    m = tree_small
    rng = np.random.default_rng(15)
    inp = _inputs(m, rng)
    y = _states(m, rng, 1)[1]
    outs = []
    for _ in range(2):
        h = HipRHS(m)
        h.set_step_inputs(**inp)
        outs.append([h.ode(y)[1] for _ in range(3)])
        h.close()
    for a, b in zip(*outs):
        assert np.array_equal(a, b)


def test_rhs_parity_full_size_200k():
    """The 200,000-triangle, 4,095-segment workload, per call."""
    from pihm_amd.meshgen import config_mesh
    m = config_mesh("c1", seed=1)
    print(f"[parity] c1 ne={m.nelem} nr={m.nriver}", flush=True)
    rng = np.random.default_rng(16)
    h, o = _pair(m)
    inp = _inputs(m, rng)
    h.set_step_inputs(**inp)
    o.set_step_inputs(**inp)
    for k, y in enumerate(_states(m, rng, 2)):
        _check(h, o, y, what=f"c1 state {k}")
    h.close()


def test_batch_members_equal_solo_bitwise():
    """Several watersheds of different sizes in one batch context: each member's ydot is
    the same bits as its own single-solve context, and inactive members are untouched."""
    import ctypes as C
    from pihm_amd import _lib
    L = _lib.lib()
    meshes = [make_watershed(n, seed=s, n_seg=k, network=net) for n, s, k, net in
              ((535, 1, 20, "valley"), (3000, 21, 80, "tree"), (1200, 5, 30, "tree"))]
    rng = np.random.default_rng(17)
    ins = [_inputs(m, rng) for m in meshes]
    ys = [_states(m, rng, 1)[1] for m in meshes]
    solo = []
    for m, inp, y in zip(meshes, ins, ys):
        h = HipRHS(m)
        h.set_step_inputs(**inp)
        solo.append([h.ode(y)[1] for _ in range(2)])     # two calls: carried state too
        h.close()
    descs = [m.desc() for m in meshes]
    arr = (C.POINTER(_lib.MeshDesc) * 3)(*[C.pointer(d) for d in descs])
    ctx = C.c_void_p()
    assert L.pihm_ctx_create_batch(3, arr, C.byref(ctx)) == 0
    assert L.pihm_ctx_nmembers(ctx) == 3
    neq = L.pihm_ctx_neq(ctx)
    assert neq == sum(m.neq for m in meshes)
    cat = lambda k: np.concatenate([i[k] for i in ins])
    keep = [cat(k) for k in ("pcpdrp", "edir", "ett", "ws0_surf")]
    s = _lib.StepInputs()
    s.pcpdrp, s.edir, s.ett, s.ws0_surf = [a.ctypes.data_as(_lib.c_double_p) for a in keep]
    assert L.pihm_ctx_set_step_inputs(ctx, C.byref(s)) == 0
    L.pihm_ctx_set_nan_check(ctx, 1)
    yv, ydv = L.N_VNew_Hip(neq), L.N_VNew_Hip(neq)
    ycat = np.concatenate(ys)
    assert L.pihm_ctx_state_to_device(ctx, ycat.ctypes.data_as(_lib.c_double_p), yv) == 0
    offs = np.cumsum([0] + [m.neq for m in meshes])
    for call in range(2):
        assert L.PIHM_ODE_Hip(0.0, yv, ydv, ctx) == 0
        out = np.empty(neq)
        L.pihm_ctx_state_from_device(ctx, ydv, out.ctypes.data_as(_lib.c_double_p))
        for k in range(3):
            assert np.array_equal(out[offs[k]:offs[k + 1]], solo[k][call]), (call, k)
    # deactivate member 1: its ydot stays as it was
    act = (C.c_int * 3)(1, 0, 1)
    L.pihm_ctx_set_active(ctx, act)
    before = out.copy()
    assert L.PIHM_ODE_Hip(0.0, yv, ydv, ctx) == 0
    L.pihm_ctx_state_from_device(ctx, ydv, out.ctypes.data_as(_lib.c_double_p))
    assert np.array_equal(out[offs[1]:offs[2]], before[offs[1]:offs[2]])
    L.N_VDestroy_Hip(yv)
    L.N_VDestroy_Hip(ydv)
    L.pihm_ctx_destroy(ctx)
