"""The one-launch RHS of single solves (k_rhs_tile: each element block stages the records
of a tile of elements and of its halo in LDS, and the river blocks split every segment's
work over four wavefronts) against the two-kernel RHS (neighbour records, then fluxes).
Both evaluate the same expressions on the same operands, so every output is compared bit
for bit: the derivative, the recorded fluxes, the carried river-edge flux over a sequence
of calls, the water-balance accumulators and the NaN report, on every mesh and branch
variant the oracle parity tests use, with and without the _NOAH_ branches, and on the
200,000-triangle workload.  The oracle parity tests themselves run on the one-launch path
(the default of single-solve contexts)."""
import ctypes as C

import numpy as np
import pytest

from pihm_amd import _lib
from pihm_amd.meshgen import make_watershed
from pihm_amd.rhs import HipRHS
from test_oracle_pins import _inputs, _states
from test_rhs_gpu import _variant

pytestmark = pytest.mark.gpu
ONE, TWO = _lib.PIHM_RHS_ONE_LAUNCH, _lib.PIHM_RHS_TWO_KERNELS


def _pair(m, **kw):
    a = HipRHS(m, **kw)
    b = HipRHS(m, **kw)
    assert a.rhs_path() == ONE, "single solves default to the one-launch RHS"
    b.set_rhs_path(TWO)
    return a, b


def _same(a, b, what):
    a, b = np.asarray(a), np.asarray(b)
    bad = ~((a == b) | (np.isnan(a) & np.isnan(b)))
    # the same bits, the sign of zero included
    bad |= (np.signbit(a) != np.signbit(b)) & ~np.isnan(a)
    if bad.any():
        i = np.nonzero(bad.ravel())[0][:5]
        raise AssertionError(f"{what}: {bad.sum()} entries differ, first {i}: {a.ravel()[i]} vs {b.ravel()[i]}")


def _compare(m, rng, what, inputs=None, noah=None, nstates=4, closure=False):
    a, b = _pair(m, record=True, closure=closure)
    try:
        inp = inputs if inputs is not None else _inputs(m, rng)
        for h in (a, b):
            h.set_step_inputs(**inp)
            if noah is not None:
                h.set_noah(noah["smcmin"])
                h.set_noah_step(**noah["step"])
        ys = _states(m, rng, nstates)
        if closure:
            ys = [np.concatenate([y, np.zeros(4)]) for y in ys]
        for k, y in enumerate(ys + ys[::-1]):       # the sequence carries river-edge fluxes
            rc_a, da = a.ode(y)
            rc_b, db = b.ode(y)
            assert rc_a == rc_b, (what, k)
            _same(da, db, f"{what} call {k} ydot")
            fa, fb = a.fluxes(), b.fluxes()
            for key in fa:
                _same(fa[key], fb[key], f"{what} call {k} {key}")
            _same(a.get_river_edge_ovlflow(), b.get_river_edge_ovlflow(), f"{what} call {k} carried")
    finally:
        a.close()
        b.close()


@pytest.mark.parametrize("which", ["c0", "tree_small", "example"])
def test_paths_bitwise(which, request):
    m = request.getfixturevalue(which)
    print(f"[paths] {which} ne={m.nelem} nr={m.nriver}", flush=True)
    _compare(m, np.random.default_rng(31), which)


@pytest.mark.parametrize("kind", ["shapes", "kinematic", "bc", "outlet-1", "outlet-2", "outlet-4"])
def test_paths_bitwise_branch_variants(kind):
    m = _variant(kind)
    print(f"[paths] variant {kind}", flush=True)
    rng = np.random.default_rng(32)
    inp = _inputs(m, rng)
    ebc = np.where(m.a["bc_type"] > 0, (m.a["zmin"] + 0.7 * m.a["depth"])[:, None], 2e-4).ravel()
    rbc = np.where(m.a["riverbc_type"] > 0, m.a["zbed"] - 0.5, 0.05)
    if kind.startswith("outlet"):
        rbc = np.where(m.a["down"] < 0, m.a["zbed"] - 0.5 if kind == "outlet-1" else 0.3, rbc)
    inp.update(elem_bc=ebc, river_bc=rbc)
    _compare(m, rng, kind, inputs=inp)


def test_paths_bitwise_noah(tree_small):
    m = tree_small
    rng = np.random.default_rng(33)
    ne, nl = m.nelem, 11
    smcmin = rng.uniform(0.01, 0.06, ne)
    noah = dict(smcmin=smcmin,
                step=dict(sh2o0=smcmin + m.a["porosity"] * rng.uniform(0.1, 0.9, ne), fcr=rng.uniform(0.5, 1.0, ne),
                          et=rng.uniform(0.0, 8e-9, (ne, nl)), nwtbl=rng.integers(1, nl + 1, ne).astype(np.int32),
                          nroot=np.full(ne, 4, np.int32)))
    _compare(m, rng, "noah", noah=noah)


def test_paths_bitwise_closure(tree_small):
    _compare(tree_small, np.random.default_rng(34), "closure", closure=True)


def test_paths_bitwise_nan():
    """A NaN in the derivative (a zero-depth-gradient outlet on an adverse bed slope) is
    reported by both paths, with the same derivative."""
    m = make_watershed(535, seed=1, n_seg=20)
    r_out = int(np.nonzero(m.rivers["down"] < 0)[0][0])
    to = m.rivers["tonode"][r_out] - 1
    m.node_zmax[to] += 3.0
    m.node_zmin[to] += 3.0
    m.a = None
    m.init()
    a, b = _pair(m)
    try:
        y = m.initial_state()
        rc_a, da = a.ode(y)
        rc_b, db = b.ode(y)
        assert rc_a == rc_b == 1
        _same(da, db, "nan ydot")
    finally:
        a.close()
        b.close()


def test_paths_bitwise_full_size_200k():
    from pihm_amd.meshgen import config_mesh
    m = config_mesh("c1", seed=1)
    print(f"[paths] c1 ne={m.nelem} nr={m.nriver}", flush=True)
    _compare(m, np.random.default_rng(35), "c1", nstates=2)


def test_rhs_path_selection(c0):
    L = _lib.lib()
    h = HipRHS(c0)
    try:
        assert h.rhs_path() == ONE
        h.set_rhs_path(TWO)
        assert h.rhs_path() == TWO
        h.set_rhs_path(ONE)
        assert h.rhs_path() == ONE
        assert L.pihm_ctx_set_rhs_path(h.ctx, 7) != 0
        assert h.rhs_path() == ONE
    finally:
        h.close()
    # a batch keeps the two-kernel RHS
    meshes = [make_watershed(n, seed=s, n_seg=k) for n, s, k in ((535, 1, 20), (800, 2, 10))]
    descs = [mm.desc() for mm in meshes]
    arr = (C.POINTER(_lib.MeshDesc) * 2)(*[C.pointer(d) for d in descs])
    ctx = C.c_void_p()
    assert L.pihm_ctx_create_batch(2, arr, C.byref(ctx)) == 0
    try:
        assert L.pihm_ctx_get_rhs_path(ctx) == TWO
        assert L.pihm_ctx_set_rhs_path(ctx, ONE) != 0
    finally:
        L.pihm_ctx_destroy(ctx)
