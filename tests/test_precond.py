"""Right preconditioning of SPGMR (PihmCVSpgmr with PIHM_PREC_RIGHT and
PihmCVSpilsSetPreconditioner; cvode_spgmr.c:289-420, cvode.c:2651-2832) and the
block-Jacobi preconditioner of the host backend (oracle/precond_oracle.c), on the CPU.

Pins:
* with the exact preconditioner P = (I - gamma A)^-1 of a linear system every SPGMR solve
  converges in one iteration (the plumbing applies P on the right and to the solution),
  and the solution matches the matrix exponential;
* the colouring: nodes of a colour are more than two links apart, and the coloured
  difference-quotient blocks equal single-node difference quotients bit for bit (the
  RHS reads no state further than two links away, and the carried river-edge flux is
  the same for every call);
* a preconditioned MM-PIHM run stays within the solver tolerance of a tight run, as the
  unpreconditioned one does.
"""
import ctypes as C
import math

import numpy as np
from scipy.linalg import expm

import oracle as O
from pihm_amd import _lib
from pihm_amd.forcing import StormForcing, run_config

L = _lib.lib()
OL = O.lib()


def _arr(v, n):
    return np.ctypeslib.as_array(C.cast(OL.oracle_nv_data(v), _lib.c_double_p), shape=(n,))


def test_exact_preconditioner_one_iteration():
    n = 40
    rng = np.random.default_rng(0)
    # stiff SPD-negative system: eigenvalues spread over four decades
    Q, _ = np.linalg.qr(rng.normal(size=(n, n)))
    A = -(Q * np.logspace(0, 4, n)) @ Q.T
    A += 0.01 * rng.normal(size=(n, n))
    calls = {"setup": 0, "solve": 0}

    def rhs(t, y, yd, ud):
        _arr(yd, n)[:] = A @ _arr(y, n)
        return 0

    def pset(t, y, fy, jok, jcur, gamma, ud, t1, t2, t3):
        calls["setup"] += 1
        jcur[0] = 1
        return 0

    def psol(t, y, fy, r, z, gamma, delta, lr, ud, tmp):
        calls["solve"] += 1
        _arr(z, n)[:] = np.linalg.solve(np.eye(n) - gamma * A, _arr(r, n))
        return 0

    cb_f, cb_s, cb_p = _lib.RhsFn(rhs), _lib.PrecSetupFn(pset), _lib.PrecSolveFn(psol)
    y0 = rng.normal(size=n)
    v = OL.oracle_nv_new(n, 1)
    _arr(v, n)[:] = y0
    mem = L.PihmCVodeCreate(2, 2)
    assert L.PihmCVodeInit(mem, C.cast(cb_f, C.c_void_p), 0.0, v) == 0
    assert L.PihmCVodeSStolerances(mem, 1e-6, 1e-9) == 0
    assert L.PihmCVSpgmr(mem, 1, 5) == -22            # PIHM_CV_ILL_INPUT: left preconditioning is not provided
    assert L.PihmCVSpgmr(mem, _lib.PIHM_PREC_RIGHT, 5) == 0
    assert L.PihmCVSpilsSetPreconditioner(mem, C.cast(cb_s, C.c_void_p), C.cast(cb_p, C.c_void_p)) == 0
    L.PihmCVodeSetInitStep(mem, 1e-6)
    L.PihmCVodeSetMaxNumSteps(mem, 100000)
    tret = C.c_double()
    assert L.PihmCVode(mem, 1.0, v, C.byref(tret), 1) >= 0
    got = _arr(v, n).copy()
    cnt = {}
    for name, fn in (("nni", L.PihmCVodeGetNumNonlinSolvIters), ("nli", L.PihmCVodeGetNumLinIters),
                     ("nsetups", L.PihmCVodeGetNumLinSolvSetups), ("npe", L.PihmCVSpilsGetNumPrecEvals),
                     ("nps", L.PihmCVSpilsGetNumPrecSolves)):
        x = C.c_long()
        fn(mem, C.byref(x))
        cnt[name] = x.value
    p = C.c_void_p(mem)
    L.PihmCVodeFree(C.byref(p))
    OL.oracle_nv_destroy(v)
    assert cnt["nli"] <= cnt["nni"]                  # one Krylov vector per Newton iteration at most
    assert cnt["nsetups"] == calls["setup"] > 0 and cnt["npe"] > 0
    assert cnt["nps"] == calls["solve"] >= cnt["nli"]
    want = expm(A * tret.value) @ y0
    assert np.max(np.abs(got - want)) < 1e-4 * max(1.0, np.max(np.abs(want)))


def _graph(m):
    """Node graph of the preconditioner, built here from the mesh arrays: elements
    (edge neighbours), river segments (downstream), banks (segment-element and
    left-right)."""
    ne, nr = m.nelem, m.nriver
    a = m.a
    nabr = a["nabr"].reshape(ne, 3)
    adj = [set() for _ in range(ne + nr)]
    for i in range(ne):
        for c in nabr[i]:
            if c > 0:
                adj[i].add(c - 1); adj[c - 1].add(i)
            elif c < 0:
                adj[i].add(ne - c - 1); adj[ne - c - 1].add(i)
    for r in range(nr):
        Lb, Rb, d = a["leftele"][r] - 1, a["rightele"][r] - 1, a["down"][r]
        for e in (Lb, Rb):
            adj[ne + r].add(e); adj[e].add(ne + r)
        adj[Lb].add(Rb); adj[Rb].add(Lb)
        if d > 0:
            adj[ne + r].add(ne + d - 1); adj[ne + d - 1].add(ne + r)
    return adj


def test_colouring_distance_two_and_exact_blocks(c0, tree_small):
    for m in (c0, tree_small):
        ne, nr, n = m.nelem, m.nriver, m.neq
        o = O.OracleRHS(m)
        rng = np.random.default_rng(3)
        o.set_step_inputs(pcpdrp=rng.uniform(0, 3e-6, ne), edir=rng.uniform(0, 2e-8, ne),
                          ett=rng.uniform(0, 4e-8, ne), ws0_surf=rng.uniform(0, 2e-3, ne))
        y = m.initial_state()
        y[:ne] = rng.uniform(0.0, 3e-3, ne)
        y[3 * ne:3 * ne + nr] += rng.uniform(0.0, 0.5, nr)
        o.ode(y)                                   # a carried river-edge flux
        bj = OL.oracle_bj_new(C.byref(o.desc), o.st, n)
        try:
            ncol = OL.oracle_bj_ncolor(bj)
            col = np.ctypeslib.as_array(OL.oracle_bj_colors(bj), shape=(ne + nr,)).copy()
            adj = _graph(m)
            for u in range(ne + nr):
                ring = set(adj[u])
                for v in adj[u]:
                    ring |= adj[v]
                ring.discard(u)
                assert all(col[w] != col[u] for w in ring), u
            assert ncol <= 16
            if m is not c0:
                continue                           # one-at-a-time quotients on the small mesh only
            ovl = o.field("ovlflow")
            save = ovl.copy()
            assert OL.oracle_bj_jacobian(bj, 0.0, y.ctypes.data_as(O.dp)) == 0
            assert np.array_equal(ovl, save)
            Jc = np.ctypeslib.as_array(OL.oracle_bj_jac(bj), shape=(9 * ne + 4 * nr,)).copy()
            _, f0 = o.ode(y)
            ovl[:] = save

            def idx(u, v):
                return v * ne + u if u < ne else 3 * ne + v * nr + (u - ne)
            for u in range(ne + nr):
                nv = 3 if u < ne else 2
                for v in range(nv):
                    k = idx(u, v)
                    yp = y.copy()
                    yp[k] = y[k] + math.sqrt(np.finfo(float).eps) * max(abs(y[k]), 1e-2)
                    inc = yp[k] - y[k]
                    _, fp = o.ode(yp)
                    ovl[:] = save
                    for w in range(nv):
                        q = idx(u, w)
                        at = 9 * u + 3 * w + v if u < ne else 9 * ne + 4 * (u - ne) + 2 * w + v
                        assert Jc[at] == (fp[q] - f0[q]) / inc, (u, v, w)
        finally:
            OL.oracle_bj_free(bj)


def test_preconditioned_run_within_tolerance(c0):
    m = c0
    f = StormForcing(8, seed=5, nelem=m.nelem, p_wet=0.5)
    y0 = m.initial_state()
    hours = 2.0
    rc_t, y_t, _ = O.OracleRun(m, f).run(y0, run_config(hours * 3600.0, reltol=1e-5, abstol=1e-7,
                                                          max_steps_per_call=10 ** 6))
    rc_n, y_n, st_n = O.OracleRun(m, f).run(y0, run_config(hours * 3600.0))
    rc_p, y_p, st_p = O.OracleRun(m, f, precond=True).run(y0, run_config(hours * 3600.0, precond=1))
    assert rc_t == rc_n == rc_p == 0
    assert st_p.nsetups > 0 and st_p.npe > 0 and st_p.nps >= st_p.nli
    assert st_n.nsetups == st_n.npe == st_n.nps == 0
    w = 1.0 / (1e-3 * np.abs(y_t) + 1e-4)
    err_n = math.sqrt(np.mean(((y_n - y_t) * w) ** 2))
    err_p = math.sqrt(np.mean(((y_p - y_t) * w) ** 2))
    assert err_p < 2.0 * err_n + 1.0, (err_n, err_p)
