"""Step record and replay of the integrator (PihmCVodeSetStepRecord / PihmCVodeSetStepReplay,
and their pihm_run_config fields), on the CPU.

A replayed run takes the recorded (h, q) of every step in place of CVODE's step-size and
order selection (cvode.c:3029-3184) and skips its error test, so two integrations of one
sequence differ only by the rounding of their right-hand sides.  Pins:
* a replayed run takes exactly the given sizes and orders: its own record equals the
  sequence bit for bit, including order rises, a drop of three orders at once and
  sizes that no selection would choose;
* replaying a free run's record through the same right-hand side gives that run's
  result and step count (the Nordsieck array is rescaled by h'/h where the free run used
  the selected ratio, which can differ in the last bit);
* with the Newton and Krylov counts replayed as well, a model run follows its record
  exactly, and the storages and outlet discharge stay within 1e-8 relative of the
  recorded run over ten minutes when the vector reductions round differently;
* on a linear system y' = A y the replayed BDF solution matches the matrix exponential
  to the accuracy the step sizes allow;
* a sequence with an order rise of more than one is refused (PIHM_CV_ILL_INPUT), and a
  run past the end of a replayed prefix continues under the normal selection.
"""
import ctypes as C

import numpy as np
import pytest
from scipy.linalg import expm

import oracle as O
from pihm_amd import _lib
from pihm_amd.forcing import StepRecord, StormForcing, run_config, set_replay

L = _lib.lib()
OL = O.lib()


def _arr(v, n):
    return np.ctypeslib.as_array(C.cast(OL.oracle_nv_data(v), _lib.c_double_p), shape=(n,))


def _linear_system(n=24, seed=0):
    rng = np.random.default_rng(seed)
    Q, _ = np.linalg.qr(rng.normal(size=(n, n)))
    A = -(Q * np.logspace(0, 2, n)) @ Q.T
    return A, rng.normal(size=n)


def _solve_linear(A, y0, t_end, replay=None, cap=4096):
    """Integrate y' = A y to t_end; returns (flag, y, (h, q) recorded, nst)."""
    n = len(y0)

    def rhs(t, y, yd, ud):
        _arr(yd, n)[:] = A @ _arr(y, n)
        return 0

    cb = _lib.RhsFn(rhs)
    v = OL.oracle_nv_new(n, 1)
    _arr(v, n)[:] = y0
    mem = L.PihmCVodeCreate(2, 2)
    try:
        assert L.PihmCVodeInit(mem, C.cast(cb, C.c_void_p), 0.0, v) == 0
        assert L.PihmCVodeSStolerances(mem, 1e-6, 1e-9) == 0
        assert L.PihmCVSpgmr(mem, 0, 0) == 0
        L.PihmCVodeSetInitStep(mem, 1e-4)
        L.PihmCVodeSetMaxNumSteps(mem, 100000)
        rec_h, rec_q, rec_n = np.zeros(cap), np.zeros(cap, np.int32), C.c_long(0)
        assert L.PihmCVodeSetStepRecord(mem, rec_h.ctypes.data_as(_lib.c_double_p),
                                        rec_q.ctypes.data_as(_lib.c_int32_p), None, cap, C.byref(rec_n)) == 0
        if replay is not None:
            rh = np.ascontiguousarray(replay[0], dtype=np.float64)
            rq = np.ascontiguousarray(replay[1], dtype=np.int32)
            assert L.PihmCVodeSetStepReplay(mem, rh.ctypes.data_as(_lib.c_double_p),
                                            rq.ctypes.data_as(_lib.c_int32_p), None, len(rh)) == 0
        L.PihmCVodeSetStopTime(mem, t_end)
        tret = C.c_double()
        flag = L.PihmCVode(mem, t_end, v, C.byref(tret), 1)
        nst = C.c_long()
        L.PihmCVodeGetNumSteps(mem, C.byref(nst))
        k = rec_n.value
        return flag, _arr(v, n).copy(), (rec_h[:k].copy(), rec_q[:k].copy()), nst.value
    finally:
        p = C.c_void_p(mem)
        L.PihmCVodeFree(C.byref(p))
        OL.oracle_nv_destroy(v)


def test_replay_follows_given_sequence_exactly():
    A, y0 = _linear_system()
    # order 1 to 5 one rise at a time, a drop from 5 to 2, sizes that grow and shrink
    q = np.array([1, 2, 3, 4, 5, 5, 5, 5, 2, 3, 3, 4, 4, 4, 1, 1, 2, 2, 3, 3] * 3, dtype=np.int32)
    h = 2e-3 * (1.0 + 0.3 * np.sin(np.arange(len(q)))) * np.where(q >= 4, 1.5, 1.0)
    t_end = float(np.sum(h))
    flag, y, (rh, rq), nst = _solve_linear(A, y0, t_end, replay=(h, q))
    assert flag >= 0
    assert nst == len(h)
    assert np.array_equal(rq, q)
    assert np.array_equal(rh, h)
    exact = expm(A * t_end) @ y0
    # step sizes up to 4e-3 at orders 1-5 on a unit-scale solution: first-order steps bound the error
    assert np.max(np.abs(y - exact)) < 2e-3 * np.max(np.abs(y0))


def test_replay_of_a_free_run_reproduces_it():
    A, y0 = _linear_system(seed=3)
    t_end = 0.5
    f1, y1, (h1, q1), n1 = _solve_linear(A, y0, t_end)
    assert f1 >= 0 and n1 == len(h1) and q1.max() >= 3
    f2, y2, (h2, q2), n2 = _solve_linear(A, y0, t_end, replay=(h1, q1))
    assert f2 >= 0 and n2 == n1
    assert np.array_equal(h2, h1) and np.array_equal(q2, q1)
    assert np.max(np.abs(y2 - y1)) <= 1e-12 * np.max(np.abs(y1))


def test_replay_refuses_an_order_jump():
    A, y0 = _linear_system()
    flag, _, (rh, rq), _ = _solve_linear(A, y0, 1.0, replay=([1e-3, 1e-3, 1e-3], [1, 3, 3]))
    assert flag == -22                      # PIHM_CV_ILL_INPUT after the first step
    assert len(rh) == 1


def test_replay_prefix_then_normal_selection():
    A, y0 = _linear_system(seed=4)
    f1, y1, (h1, q1), _ = _solve_linear(A, y0, 0.5)
    k = len(h1) // 2
    f2, y2, (h2, q2), _ = _solve_linear(A, y0, 0.5, replay=(h1[:k], q1[:k]))
    assert f2 >= 0
    assert np.array_equal(h2[:k], h1[:k]) and np.array_equal(q2[:k], q1[:k])
    # after the prefix the selection runs again: same decisions as the free run
    assert np.max(np.abs(y2 - y1)) <= 1e-9 * np.max(np.abs(y1))


def _stores(mesh, y):
    """Surface, unsaturated, groundwater, river and river-bed storages [m3]."""
    a, ne, nr = mesh.a, mesh.nelem, mesh.nriver
    ys, yu, yg = y[:ne], y[ne:2 * ne], y[2 * ne:3 * ne]
    st, rg = y[3 * ne:3 * ne + nr], y[3 * ne + nr:3 * ne + 2 * nr]
    return np.array([np.sum(a["area"] * ys), np.sum(a["area"] * a["porosity"] * yu),
                     np.sum(a["area"] * a["porosity"] * yg), np.sum(a["rarea"] * st),
                     np.sum(a["rarea"] * a["rporosity"] * rg)])


def _outlet(mesh, y):
    """Discharge of the outlet segments [m3 s-1] (OutletFlux, src/river_flow.c:333-390),
    from the oracle right-hand side at state y."""
    o = O.OracleRHS(mesh)
    o.ode(y)
    rf = o.field("rivflow").reshape(mesh.nriver, 11)
    return rf[np.asarray(mesh.rivers["down"]) < 0, 1].copy()


def test_model_run_replay_reproduces_record(c0):
    """MM-PIHM under storm forcing through the model-step driver (pihm_drive on the
    oracle backend): replaying the run's own record (steps, orders, Newton and Krylov
    counts) takes the same path; over the first ten minutes it reproduces the storages
    and the outlet discharge within 1e-8 relative, and the state within 5e-8 of each
    component (1e-4 m at least), also when the vector reductions round differently
    (eight OpenMP threads against one: 1.5e-8 at most), five orders of magnitude under
    the solver's relative tolerance."""
    m = c0
    f = StormForcing(24, seed=5, nelem=m.nelem, p_wet=0.5)
    y0 = m.initial_state()
    rec = StepRecord(200000)
    rc1, y1, s1 = O.OracleRun(m, f).run(y0, rec.attach(run_config(600.0)))
    assert rc1 == 0
    h, q, it = rec.steps()
    assert len(h) == s1.nst and q.max() >= 2 and it[:, 0].min() >= 1
    for nth in (1, 8):
        rec2 = StepRecord(200000)
        cfg = set_replay(rec2.attach(run_config(600.0)), h, q, it)
        rc2, y2, s2 = O.OracleRun(m, f, nthreads=nth).run(y0, cfg)
        assert rc2 == 0 and s2.nst == s1.nst
        h2, q2, it2 = rec2.steps()
        assert np.array_equal(h2, h) and np.array_equal(q2, q) and np.array_equal(it2, it)
        # Newton iterations of the accepted steps only: the record's sum
        assert s2.nni == int(it[:, 0].sum())
        assert np.max(np.abs(y2 - y1) / np.maximum(np.abs(y1), 1e-4)) <= 5e-8
        S1, S2 = _stores(m, y1), _stores(m, y2)
        live = np.abs(S1) > 0
        assert np.all(np.abs(S2 - S1)[live] <= 1e-8 * np.abs(S1)[live])
        o1, o2 = _outlet(m, y1), _outlet(m, y2)
        assert np.all(np.abs(o2 - o1) <= 1e-8 * np.abs(o1))


def test_record_and_replay_argument_errors():
    h = np.zeros(4)
    q = np.zeros(4, np.int32)
    n = C.c_long()
    dp, ip = h.ctypes.data_as(_lib.c_double_p), q.ctypes.data_as(_lib.c_int32_p)
    assert L.PihmCVodeSetStepRecord(None, dp, ip, None, 4, C.byref(n)) == -21     # PIHM_CV_MEM_NULL
    assert L.PihmCVodeSetStepReplay(None, dp, ip, None, 4) == -21
    mem = L.PihmCVodeCreate(2, 2)
    try:
        assert L.PihmCVodeSetStepRecord(mem, None, ip, None, 4, C.byref(n)) == -22   # PIHM_CV_ILL_INPUT
        assert L.PihmCVodeSetStepRecord(mem, dp, ip, None, -1, C.byref(n)) == -22
        assert L.PihmCVodeSetStepRecord(mem, dp, ip, None, 4, None) == -22
        assert L.PihmCVodeSetStepRecord(mem, None, None, None, 0, None) == 0          # switched off
        assert L.PihmCVodeSetStepReplay(mem, dp, None, None, 4) == -22
        assert L.PihmCVodeSetStepReplay(mem, dp, ip, None, -2) == -22
        assert L.PihmCVodeSetStepReplay(mem, None, None, None, 0) == 0                # switched off
    finally:
        p = C.c_void_p(mem)
        L.PihmCVodeFree(C.byref(p))
    with pytest.raises(ValueError):
        set_replay(run_config(60.0), np.zeros(3), np.zeros(4, np.int32))
    with pytest.raises(ValueError):
        set_replay(run_config(60.0), np.zeros(3), np.zeros(3, np.int32), np.zeros((3, 5), np.int32))


def test_reduction_order_alone_breaks_the_bound_within_the_hour(c0):
    """Why the 1e-8 bound is a ten-minute bound: two host runs of one replayed sequence
    whose only difference is the reduction order (eight OpenMP threads against one)
    agree to 1e-10 over ten minutes, but after one hour their storages differ by more
    than 1e-6 relative and the outlet discharge by more than 1e-4 (branch crossings of
    ponding depth and river stage amplify the rounding), while the storages stay within
    the solver's relative tolerance of 1e-3."""
    m = c0
    f = StormForcing(24, seed=5, nelem=m.nelem, p_wet=0.5)
    y0 = m.initial_state()
    div = {}
    for T in (600.0, 3600.0):
        rec = StepRecord(200000)
        rc1, y1, _ = O.OracleRun(m, f).run(y0, rec.attach(run_config(T)))
        h, q, it = rec.steps()
        rc2, y2, _ = O.OracleRun(m, f, nthreads=8).run(y0, set_replay(StepRecord(200000).attach(run_config(T)),
                                                                      h, q, it))
        assert rc1 == 0 and rc2 == 0
        S1, S2 = _stores(m, y1), _stores(m, y2)
        live = np.abs(S1) > 0
        o1, o2 = _outlet(m, y1), _outlet(m, y2)
        div[T] = (np.max(np.abs(S2 - S1)[live] / np.abs(S1)[live]), np.max(np.abs(o2 - o1) / np.abs(o1)))
    assert div[600.0][0] <= 1e-10 and div[600.0][1] <= 1e-10
    assert div[3600.0][0] > 1e-6 and div[3600.0][1] > 1e-4
    assert div[3600.0][0] < 1e-3
