"""ctypes wrapper of the CPU oracle (oracle/build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the cpu_baseline
leg of bench.py, as the checker or the timed CPU baseline; never by the product path.
Build with ``make -C oracle``.
"""
import ctypes as C
import os
import sys

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# PIHM_ORACLE_LIB points the tests at another build of the same sources (a copy with a
# planted mistake, to show that the pins catch it)
LIB_PATH = os.environ.get("PIHM_ORACLE_LIB") or os.path.join(_HERE, "build", "liboracle.so")
sys.path.insert(0, os.path.join(os.path.dirname(_HERE), "pihm-cuda_amd"))
from pihm_amd import _lib as P  # noqa: E402  (structure definitions only)

dp = P.c_double_p
ip = P.c_int32_p
_o = None


def lib():
    global _o
    if _o is not None:
        return _o
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"oracle not built: {LIB_PATH} (run `make -C oracle`)")
    L = C.CDLL(LIB_PATH, mode=C.RTLD_GLOBAL)
    protos = [
        ("oracle_state_new", C.c_void_p, [C.c_int, C.c_int]),
        ("oracle_state_free", None, [C.c_void_p]),
        ("oracle_set_step_inputs", None, [C.c_void_p, C.POINTER(P.StepInputs)]),
        ("oracle_set_noah", None, [C.c_void_p, dp, dp, dp, dp, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
        ("oracle_GwTransp", C.c_double, [C.c_double, dp, C.c_int, C.c_int]),
        ("oracle_FindWaterTable", C.c_int, [dp, C.c_int, C.c_double, dp]),
        ("oracle_CalcLatFlx", None, [C.c_int, dp, C.c_double, dp]),
        ("oracle_WDfCnd", None, [dp, dp, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                 C.c_double]),
        ("oracle_Rosr12", None, [dp, dp, dp, dp, dp, dp, C.c_int]),
        ("oracle_noah_soil_step", None, [C.c_int, C.c_double, C.POINTER(C.c_int)] + [dp] * 12 +
         [dp, dp, dp, C.POINTER(C.c_int), dp, dp, dp, dp]),
        ("oracle_ode", C.c_int, [C.POINTER(P.MeshDesc), C.c_void_p, C.c_double, dp, dp, dp]),
        ("oracle_SurfH", C.c_double, [C.c_double]),
        ("oracle_KrFunc", C.c_double, [C.c_double, C.c_double]),
        ("oracle_Psi", C.c_double, [C.c_double, C.c_double, C.c_double]),
        ("oracle_EffKh", C.c_double, [C.c_double] * 6),
        ("oracle_AvgKv", C.c_double, [C.c_double] * 7),
        ("oracle_EffKinf", C.c_double, [C.c_double] * 8),
        ("oracle_DhByDl", C.c_double, [dp, dp, dp]),
        ("oracle_OverLandFlow", C.c_double, [C.c_double] * 5),
        ("oracle_RiverCroSectArea", C.c_double, [C.c_int, C.c_double, C.c_double]),
        ("oracle_RiverPerim", C.c_double, [C.c_int, C.c_double, C.c_double]),
        ("oracle_RiverEqWid", C.c_double, [C.c_int, C.c_double, C.c_double]),
        ("oracle_OvlFlowElemToRiver_raw", C.c_double, [C.c_double] * 7),
        ("oracle_init_topo", None, [C.c_int, ip, dp, dp, dp, dp, dp, dp, dp, dp, dp, dp]),
        ("oracle_init_river_tags", None, [C.c_int, C.c_int, ip, ip, ip]),
        ("oracle_init_surfl", None, [C.c_int, ip, ip, dp, dp, dp, dp, dp, dp, dp, dp, dp, dp, dp]),
        ("oracle_init_river_topo", None, [C.c_int, ip, ip, ip, ip, dp, dp, dp, dp, dp, dp, dp,
                                          dp, dp, dp, dp, dp, dp, dp, dp]),
        ("oracle_nv_new", C.c_void_p, [C.c_long, C.c_int]),
        ("oracle_nv_destroy", None, [C.c_void_p]),
        ("oracle_nv_data", C.c_void_p, [C.c_void_p]),
        ("oracle_nv_length", C.c_long, [C.c_void_p]),
        ("oracle_backend_new", C.c_void_p, [C.POINTER(P.MeshDesc), C.c_void_p, C.POINTER(P.Forcing), C.c_int]),
        ("oracle_backend_get", C.POINTER(P.Backend), [C.c_void_p]),
        ("oracle_backend_free", None, [C.c_void_p]),
        ("oracle_backend_enable_precond", None, [C.c_void_p, C.c_long]),
        ("oracle_backend_set_noah_soil", None, [C.c_void_p, C.POINTER(C.c_int), dp, dp, dp, dp, dp]),
        ("oracle_backend_noah_soil_state", None, [C.c_void_p, dp, dp]),
        ("oracle_backend_start", C.c_int, [C.c_void_p, C.c_void_p]),
        ("oracle_backend_bj", C.c_void_p, [C.c_void_p]),
        ("oracle_bj_new", C.c_void_p, [C.POINTER(P.MeshDesc), C.c_void_p, C.c_long]),
        ("oracle_bj_free", None, [C.c_void_p]),
        ("oracle_bj_ncolor", C.c_int, [C.c_void_p]),
        ("oracle_bj_colors", ip, [C.c_void_p]),
        ("oracle_bj_jac", dp, [C.c_void_p]),
        ("oracle_bj_jacobian", C.c_int, [C.c_void_p, C.c_double, dp]),
        ("oracle_meteo_interp", C.c_int, [C.POINTER(P.Meteo), C.c_double, dp]),
        ("oracle_lsm_backend_new", C.c_void_p, [C.POINTER(P.MeshDesc), C.c_void_p, C.POINTER(P.LsmDesc),
                                                C.POINTER(P.Meteo), dp, dp, C.c_double, dp, dp, C.c_double]),
        ("oracle_lsm_backend_start", C.c_int, [C.c_void_p, C.c_void_p]),
        ("oracle_lsm_backend_get", C.POINTER(P.Backend), [C.c_void_p]),
        ("oracle_lsm_backend_state", None, [C.c_void_p, dp, dp]),
        ("oracle_lsm_backend_free", None, [C.c_void_p]),
        ("oracle_lsm_backend_set_bc", None, [C.c_void_p, C.POINTER(P.Ts), ip, C.POINTER(P.Ts), ip]),
        ("oracle_bc_apply", C.c_int, [C.POINTER(P.Ts), ip, C.c_int, dp, C.POINTER(P.Ts), ip, C.c_int, dp,
                                      C.c_double]),
        ("oracle_intcp_snow_et", None, [C.c_int, C.POINTER(P.LsmDesc), dp, C.c_int, dp, dp, C.c_double,
                                        dp, dp, C.c_double, dp, dp, dp, dp, dp, dp, dp]),
    ]
    for name, res, args in protos:
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _o = L
    return L


def _dptr(a):
    return a.ctypes.data_as(dp)


def step_inputs(pcpdrp=None, edir=None, ett=None, ws0_surf=None, elem_bc=None, river_bc=None, keep=None):
    s = P.StepInputs()
    for name, arr in (("pcpdrp", pcpdrp), ("edir", edir), ("ett", ett), ("ws0_surf", ws0_surf),
                      ("elem_bc", elem_bc), ("river_bc", river_bc)):
        if arr is not None:
            arr = np.ascontiguousarray(arr, dtype=np.float64)
            if keep is not None:
                keep.append(arr)
            setattr(s, name, _dptr(arr))
    return s


class OracleRHS:
    """The reference RHS on the host (original order), with its carried state."""

    def __init__(self, mesh):
        self.L = lib()
        self.mesh = mesh
        self.desc = mesh.desc()
        self.st = self.L.oracle_state_new(mesh.nelem, mesh.nriver)
        self._keep = []

    def __del__(self):
        try:
            self.L.oracle_state_free(self.st)
        except Exception:
            pass

    def set_step_inputs(self, **kw):
        keep = []
        s = step_inputs(keep=keep, **kw)
        self.L.oracle_set_step_inputs(self.st, C.byref(s))

    def set_noah(self, smcmin, sh2o0=None, fcr=None, et=None, nwtbl=None, nroot=None):
        """The _NOAH_ branches of EtExtract and Infil (smcmin None: off).  et is
        [nelem, nlayer]; nwtbl, nroot 1-based layers per element."""
        if smcmin is None:
            self.L.oracle_set_noah(self.st, None, None, None, None, 0, None, None)
            return
        f = lambda a: None if a is None else np.ascontiguousarray(a, dtype=np.float64)
        i = lambda a: None if a is None else np.ascontiguousarray(a, dtype=np.intc)
        smcmin, sh2o0, fcr, et, nwtbl, nroot = f(smcmin), f(sh2o0), f(fcr), f(et), i(nwtbl), i(nroot)
        nl = 0 if et is None else et.shape[1]
        ip = lambda a: None if a is None else a.ctypes.data_as(C.POINTER(C.c_int))
        self.L.oracle_set_noah(self.st, _dptr(smcmin), _dptr(sh2o0) if sh2o0 is not None else None,
                               _dptr(fcr) if fcr is not None else None, _dptr(et) if et is not None else None,
                               nl, ip(nwtbl), ip(nroot))

    def _field(self, name, n):
        class _S(C.Structure):
            pass
        # the oracle_state layout: two ints then pointers in declaration order
        names = ["pcpdrp", "edir", "ett", "ws0_surf", "elem_bc", "river_bc", "surf", "unsat", "gw",
                 "surfh", "edir_surf", "edir_unsat", "edir_gw", "ett_unsat", "ett_gw", "infil",
                 "rechg", "ovlflow", "subsurf", "dhbydx", "dhbydy", "stage", "rgw", "rivflow"]
        _S._fields_ = [("nelem", C.c_int), ("nriver", C.c_int)] + [(k, dp) for k in names]
        s = C.cast(self.st, C.POINTER(_S)).contents
        ptr = getattr(s, name)
        return np.ctypeslib.as_array(ptr, shape=(n,))

    def field(self, name):
        ne, nr = self.mesh.nelem, self.mesh.nriver
        sizes = {"ovlflow": 3 * ne, "subsurf": 3 * ne, "elem_bc": 3 * ne, "river_bc": nr,
                 "stage": nr, "rgw": nr, "rivflow": 11 * nr}
        return self._field(name, sizes.get(name, ne))

    def ode(self, y, t=0.0, with_mag=False):
        y = np.ascontiguousarray(y, dtype=np.float64)
        dy = np.zeros_like(y)
        mag = np.zeros_like(y) if with_mag else None
        rc = self.L.oracle_ode(C.byref(self.desc), self.st, t, _dptr(y), _dptr(dy),
                               _dptr(mag) if with_mag else None)
        if with_mag:
            return rc, dy, mag
        return rc, dy


def init_geometry(mesh):
    """Oracle restatement of InitTopo/InitRiver/InitSurfL on the mesh's raw topology."""
    L = lib()
    ne, nr = mesh.nelem, mesh.nriver
    node = np.ascontiguousarray(mesh.tri, dtype=np.int32)
    out = {k: np.zeros(ne) for k in ("area", "x", "y", "zmin", "zmax")}
    edge = np.zeros(3 * ne)
    L.oracle_init_topo(ne, node.ctypes.data_as(ip), _dptr(mesh.node_x), _dptr(mesh.node_y),
                       _dptr(mesh.node_zmin), _dptr(mesh.node_zmax), _dptr(out["area"]),
                       _dptr(out["x"]), _dptr(out["y"]), _dptr(out["zmin"]), _dptr(out["zmax"]),
                       _dptr(edge))
    out["edge"] = edge
    nabr = np.ascontiguousarray(mesh.nabr_raw, dtype=np.int32).copy()
    left = np.ascontiguousarray(mesh.rivers["leftele"], dtype=np.int32)
    right = np.ascontiguousarray(mesh.rivers["rightele"], dtype=np.int32)
    L.oracle_init_river_tags(ne, nr, nabr.ctypes.data_as(ip), left.ctypes.data_as(ip), right.ctypes.data_as(ip))
    out["nabr"] = nabr
    rk = {k: np.zeros(max(nr, 1)) for k in ("rx", "ry", "rzmax", "rzmin", "node_zmax", "dist_left",
                                              "dist_right", "length")}
    fn = np.ascontiguousarray(mesh.rivers["fromnode"], dtype=np.int32)
    tn = np.ascontiguousarray(mesh.rivers["tonode"], dtype=np.int32)
    L.oracle_init_river_topo(nr, fn.ctypes.data_as(ip), tn.ctypes.data_as(ip), left.ctypes.data_as(ip),
                             right.ctypes.data_as(ip), _dptr(mesh.node_x), _dptr(mesh.node_y),
                             _dptr(mesh.node_zmax), _dptr(out["zmax"]), _dptr(out["zmin"]),
                             _dptr(out["x"]), _dptr(out["y"]), *[_dptr(rk[k]) for k in
                             ("rx", "ry", "rzmax", "rzmin", "node_zmax", "dist_left", "dist_right",
                              "length")])
    for k, v in rk.items():
        out[k] = v[:nr]
    nx_, ny_, nd_ = np.zeros(3 * ne), np.zeros(3 * ne), np.zeros(3 * ne)
    L.oracle_init_surfl(ne, node.ctypes.data_as(ip), nabr.ctypes.data_as(ip), _dptr(mesh.node_x),
                        _dptr(mesh.node_y), _dptr(out["x"]), _dptr(out["y"]), _dptr(out["area"]),
                        _dptr(edge), _dptr(rk["rx"]), _dptr(rk["ry"]), _dptr(nx_), _dptr(ny_), _dptr(nd_))
    out["nabr_x"], out["nabr_y"], out["nabrdist"] = nx_, ny_, nd_
    return out


class OracleRun:
    """The model-step driver (product integrator) over the oracle RHS on host vectors."""

    def __init__(self, mesh, forcing, nthreads=1, precond=False, noah=None):
        self.L = lib()
        self.PL = P.lib()
        self.mesh = mesh
        self.desc = mesh.desc()
        self.st = self.L.oracle_state_new(mesh.nelem, mesh.nriver)
        self.forcing = forcing
        self.fs = forcing.struct()
        self.be = self.L.oracle_backend_new(C.byref(self.desc), self.st, C.byref(self.fs), nthreads)
        self.nthreads = nthreads
        if precond:
            self.L.oracle_backend_enable_precond(self.be, mesh.neq)
        self.noah = noah
        if noah is not None:
            # the _NOAH_ RHS branches (11-layer et) and the soil-water step at every Summary
            f = lambda a: np.ascontiguousarray(a, dtype=np.float64)
            ci = lambda a: np.ascontiguousarray(a, dtype=np.intc)
            self._nk = [f(noah["smcmin"]), f(noah["sh2o"][:, 0]), f(np.ones(mesh.nelem)), f(noah["et"]),
                        ci(noah["nwtbl"]), ci(noah["nroot"]), ci(noah["nsoil"]), f(noah["sldpth"]),
                        f(noah["smcmax"]), f(noah["frzx"]), f(noah["smc"]), f(noah["sh2o"])]
            k = self._nk
            assert k[3].shape == (mesh.nelem, 11)
            ip = lambda a: a.ctypes.data_as(C.POINTER(C.c_int))
            self.L.oracle_set_noah(self.st, _dptr(k[0]), _dptr(k[1]), _dptr(k[2]), _dptr(k[3]), 11, ip(k[4]), ip(k[5]))
            self.L.oracle_backend_set_noah_soil(self.be, ip(k[6]), _dptr(k[7]), _dptr(k[8]), _dptr(k[9]),
                                                _dptr(k[10]), _dptr(k[11]))

    def set_noah_lsm(self, par, state, station, ftime, data, lai, etstep):
        """The Noah step at every LSM step (oracle_backend_set_noah_lsm): par and state are
        NOAH_PAR_DT / NOAH_STATE_DT arrays (state updated in place as the run goes)."""
        L = self.L
        if not hasattr(L, "_noah_lsm_proto"):
            L.oracle_backend_set_noah_lsm.restype = None
            L.oracle_backend_set_noah_lsm.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(C.c_int),
                                                      C.c_int, C.c_int, dp, dp, dp, C.c_double]
            L._noah_lsm_proto = True
        assert par.dtype == NOAH_PAR_DT and state.dtype == NOAH_STATE_DT
        st = np.ascontiguousarray(np.asarray(station), dtype=np.intc)
        ft = np.ascontiguousarray(ftime, dtype=np.float64)
        dd = np.ascontiguousarray(data, dtype=np.float64)
        la = np.ascontiguousarray(lai, dtype=np.float64)
        self._lsm_keep = (par, state, st, ft, dd, la)
        L.oracle_backend_set_noah_lsm(self.be, par.ctypes.data, state.ctypes.data,
                                      st.ctypes.data_as(C.POINTER(C.c_int)), dd.shape[0], dd.shape[1],
                                      _dptr(ft), _dptr(dd), _dptr(la), float(etstep))

    def noah_soil_state(self):
        ne = self.mesh.nelem
        smc, sh2o = np.zeros((ne, 11)), np.zeros((ne, 11))
        self.L.oracle_backend_noah_soil_state(self.be, _dptr(smc), _dptr(sh2o))
        return smc, sh2o

    def run(self, y0, cfg):
        n = self.mesh.neq
        v = self.L.oracle_nv_new(n, self.nthreads)
        try:
            data = np.ctypeslib.as_array(C.cast(self.L.oracle_nv_data(v), dp), shape=(n,))
            data[:] = y0
            # initial ws0 = initial surface state
            ws0 = self._ws0()
            ws0[:] = y0[:self.mesh.nelem]
            if self.noah is not None:
                self.L.oracle_backend_start(self.be, v)
            st = P.RunStats()
            rc = self.PL.pihm_drive(self.L.oracle_backend_get(self.be), v, C.byref(cfg), C.byref(st))
            return rc, data.copy(), st
        finally:
            self.L.oracle_nv_destroy(v)

    def _ws0(self):
        o = OracleRHS.__new__(OracleRHS)
        o.st, o.mesh = self.st, self.mesh
        return o._field("ws0_surf", self.mesh.nelem)

    def __del__(self):
        try:
            self.L.oracle_backend_free(self.be)
            self.L.oracle_state_free(self.st)
        except Exception:
            pass


def meteo_interp(meteo, t):
    """IntrplForc for every station: (rc, met[5, nstation]); rc 1 when t is outside."""
    from pihm_amd.lsm import lsm_structs
    keep = []
    ft = np.asarray(meteo[0])
    dummy = {f: np.zeros(1) for f in _lsm_fields()}
    dummy["station"] = np.zeros(1, np.int32)
    _, m = lsm_structs(dummy, meteo, keep)
    out = np.zeros((P.PIHM_MET_NVAR, ft.shape[0]))
    rc = lib().oracle_meteo_interp(C.byref(m), float(t), _dptr(out))
    return rc, out


def _lsm_fields():
    from pihm_amd.lsm import LSM_FIELDS
    return LSM_FIELDS


def intcp_snow_et(params, met, lai, z0, meltf, unsat, gw, stepsize, cmc, sneqv):
    """IntcpSnowEt on the host for every element (caller order).  ``met`` is [5, nstation]
    as meteo_interp returns it.  Returns (cmc, sneqv, fluxes) with fluxes keyed pcpdrp,
    edir, ett, ec, drip; the inputs are not modified."""
    from pihm_amd.lsm import lsm_structs
    keep = []
    met = np.ascontiguousarray(met, dtype=np.float64)
    ns = met.shape[1]
    fake_meteo = (np.zeros((ns, 2)), np.zeros((ns, 2, P.PIHM_MET_NVAR)))
    d, _ = lsm_structs(params, fake_meteo, keep)
    ne = len(params["station"])
    f64 = lambda a: np.ascontiguousarray(np.broadcast_to(np.asarray(a, np.float64), (ne,))).copy()
    lai, z0, unsat, gw = f64(lai), f64(z0), f64(unsat), f64(gw)
    cmc, sneqv = f64(cmc), f64(sneqv)
    out = {k: np.zeros(ne) for k in ("pcpdrp", "edir", "ett", "ec", "drip")}
    lib().oracle_intcp_snow_et(ne, C.byref(d), _dptr(met), ns, _dptr(lai), _dptr(z0), float(meltf),
                               _dptr(unsat), _dptr(gw), float(stepsize), _dptr(cmc), _dptr(sneqv),
                               *(_dptr(out[k]) for k in ("pcpdrp", "edir", "ett", "ec", "drip")))
    return cmc, sneqv, out


def bc_apply(ne, nr, t, elem_type=None, elem_series=None, river_type=None, river_series=None,
             elem_bc=None, river_bc=None):
    """ApplyElemBc/ApplyRiverBc on the host: (rc, elem_bc[ne, 3], river_bc[nr]) starting from
    the given values (zeros by default)."""
    from pihm_amd.bc import code_array, ts_struct
    keep = []
    es, rs = ts_struct(elem_series, keep), ts_struct(river_series, keep)
    eb = np.zeros(3 * ne) if elem_bc is None else np.ascontiguousarray(elem_bc, np.float64).reshape(3 * ne).copy()
    rb = np.zeros(max(nr, 1)) if river_bc is None else np.ascontiguousarray(river_bc, np.float64).copy()
    rc = lib().oracle_bc_apply(C.byref(es) if es is not None else None, code_array(elem_type, 3 * ne, keep), ne,
                               _dptr(eb), C.byref(rs) if rs is not None else None,
                               code_array(river_type, nr, keep), nr, _dptr(rb), float(t))
    return rc, eb.reshape(ne, 3), rb[:nr]


class OracleLsmRun:
    """The model-step driver over the oracle RHS with the oracle land-surface step as the
    forcing, on host vectors (the host leg of pihm_run_gpu_lsm's parity test).  ``bc`` is
    (elem_type, elem_series, river_type, river_series) or None."""

    def __init__(self, mesh, params, meteo, lai, z0, meltf, cmc0, sneqv0, etstep, bc=None):
        from pihm_amd.lsm import lsm_structs
        self.L = lib()
        self.PL = P.lib()
        self.mesh = mesh
        self.desc = mesh.desc()
        self.keep = []
        self.d, self.m = lsm_structs(params, meteo, self.keep)
        arr = lambda a: self.keep.append(np.ascontiguousarray(a, dtype=np.float64)) or self.keep[-1]
        lai, z0, cmc0, sneqv0 = arr(lai), arr(z0), arr(cmc0), arr(sneqv0)
        self.st = self.L.oracle_state_new(mesh.nelem, mesh.nriver)
        self.b = self.L.oracle_lsm_backend_new(C.byref(self.desc), self.st, C.byref(self.d), C.byref(self.m),
                                               _dptr(lai), _dptr(z0), float(meltf), _dptr(cmc0),
                                               _dptr(sneqv0), float(etstep))
        if bc is not None:
            from pihm_amd.bc import code_array, ts_struct
            et, es, rt, rs = bc
            self.es, self.rs = ts_struct(es, self.keep), ts_struct(rs, self.keep)
            self.L.oracle_lsm_backend_set_bc(
                self.b, C.byref(self.es) if self.es is not None else None,
                code_array(et, 3 * mesh.nelem, self.keep), C.byref(self.rs) if self.rs is not None else None,
                code_array(rt, mesh.nriver, self.keep))

    def run(self, y0, cfg):
        n = self.mesh.neq
        v = self.L.oracle_nv_new(n, 1)
        try:
            data = np.ctypeslib.as_array(C.cast(self.L.oracle_nv_data(v), dp), shape=(n,))
            data[:] = y0
            self.L.oracle_lsm_backend_start(self.b, v)
            st = P.RunStats()
            rc = self.PL.pihm_drive(self.L.oracle_lsm_backend_get(self.b), v, C.byref(cfg), C.byref(st))
            return rc, data.copy(), st
        finally:
            self.L.oracle_nv_destroy(v)

    def storages(self):
        cmc, sneqv = np.zeros(self.mesh.nelem), np.zeros(self.mesh.nelem)
        self.L.oracle_lsm_backend_state(self.b, _dptr(cmc), _dptr(sneqv))
        return cmc, sneqv

    def __del__(self):
        try:
            self.L.oracle_lsm_backend_free(self.b)
            self.L.oracle_state_free(self.st)
        except Exception:
            pass


NS_MAXLYR = 11


def noah_soil_step(dt, nsoil, sldpth, soil, gw, infil, runoff2, edir, et, state):
    """The soil-water step of the _NOAH_ build for every element (noahsoil_oracle.c).
    Per-layer arrays are [nelem, 11]; soil holds smcmax, smcmin, ksatv, alpha, beta, frzx;
    state holds smc, sh2o, fcr, smflxv, runoff3 and runoff2_lyr (updated in place) and
    receives nwtbl and satdpth."""
    L = lib()
    ne = len(gw)
    f = lambda a: np.ascontiguousarray(a, dtype=np.float64)
    ns = np.ascontiguousarray(nsoil, dtype=np.intc)
    arrs = [f(sldpth)] + [f(soil[k]) for k in ("smcmax", "smcmin", "ksatv", "alpha", "beta", "frzx")] + \
        [f(gw), f(infil), f(runoff2), f(edir), f(et)]
    for k in ("smc", "sh2o", "smflxv", "runoff2_lyr"):
        state[k] = f(state[k]).copy()
    for k in ("fcr", "runoff3"):
        state[k] = f(state[k]).copy()
    state["nwtbl"] = np.zeros(ne, dtype=np.intc)
    state["satdpth"] = np.zeros((ne, NS_MAXLYR))
    L.oracle_noah_soil_step(ne, float(dt), ns.ctypes.data_as(C.POINTER(C.c_int)), *[_dptr(a) for a in arrs],
                            _dptr(state["smc"]), _dptr(state["sh2o"]), _dptr(state["fcr"]),
                            state["nwtbl"].ctypes.data_as(C.POINTER(C.c_int)), _dptr(state["satdpth"]),
                            _dptr(state["runoff2_lyr"]), _dptr(state["smflxv"]), _dptr(state["runoff3"]))
    return state


# ----------------------------------------------------------------------------- Noah land-surface step
_NP_SCALARS = ["shdfac", "cmcfactr", "laimin", "laimax", "emissmin", "emissmax", "albedomin", "albedomax",
               "z0min", "z0max", "snup", "cfactr", "rgl", "rsmin", "rsmax", "topt", "hs", "smcmax", "smcmin",
               "smcref", "smcwlt", "smcdry", "quartz", "csoil", "alpha", "beta", "sbeta", "salp", "fxexp", "czil",
               "lvcoef", "zbot", "tbot", "snoalb", "zlvl", "zlvl_wind"]
NOAH_PAR_DT = np.dtype([("nsoil", np.intc), ("nroot", np.intc), ("sldpth", np.float64, 11)] +
                       [(k, np.float64) for k in _NP_SCALARS], align=True)
NOAH_STATE_DT = np.dtype([("smc", np.float64, 11), ("sh2o", np.float64, 11), ("stc", np.float64, 11)] +
                         [(k, np.float64) for k in ("t1", "sneqv", "snowh", "sndens", "cmc", "ch", "cm",
                                                    "snotime1", "albedo", "emissi", "q1")], align=True)
NOAH_FORC_DT = np.dtype([(k, np.float64) for k in ("sfctmp", "rh", "sfcspd", "soldn", "longwave", "sfcprs",
                                                   "prcp", "proj_lai")], align=True)
NOAH_OUT_DT = np.dtype([(k, np.float64) for k in ("edir", "ett", "ec")] + [("et", np.float64, 11)] +
                       [(k, np.float64) for k in ("pcpdrp", "drip", "esnow", "etp", "snomlt", "ssoil", "sheat",
                                                  "eta", "fdown", "sncovr", "beta", "rc", "pc", "flx1", "flx2",
                                                  "flx3", "dew", "soilw", "soilm", "cmcmax", "rr", "rch", "epsca",
                                                  "q2", "q2sat", "th2")], align=True)


def noah_lsm_step(dt, par, state, forc):
    """Noah() over every element (noahlsm_oracle.c): par, state and forc are structured
    arrays of NOAH_PAR_DT, NOAH_STATE_DT (updated in place) and NOAH_FORC_DT; returns
    the NOAH_OUT_DT outputs."""
    L = lib()
    if not hasattr(L, "_noah_protos"):
        L.oracle_noah_step.restype = None
        L.oracle_noah_step.argtypes = [C.c_int, C.c_double, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L._noah_protos = True
    assert par.dtype == NOAH_PAR_DT and state.dtype == NOAH_STATE_DT and forc.dtype == NOAH_FORC_DT
    assert state.flags.c_contiguous and par.flags.c_contiguous and forc.flags.c_contiguous
    out = np.zeros(len(par), NOAH_OUT_DT)
    L.oracle_noah_step(len(par), float(dt), par.ctypes.data, state.ctypes.data, forc.ctypes.data, out.ctypes.data)
    return out


def noah_fn(name):
    """A single function of the Noah restatement, for the pins."""
    L = lib()
    f = getattr(L, name)
    d = C.c_double
    sig = {"oracle_CSnow": (d, [d]), "oracle_SnFrac": (d, [d, d, d]), "oracle_TDfCnd": (d, [d] * 5),
           "oracle_Pspmu": (d, [d]), "oracle_Psphu": (d, [d]),
           "oracle_SfcDifOff": (None, [d] * 7 + [C.POINTER(d), C.POINTER(d)]),
           "oracle_TBnd": (d, [d, d, dp, d, C.c_int, C.c_int]), "oracle_TmpAvg": (d, [d, d, d, dp, C.c_int]),
           "oracle_FrH2O": (d, [d] * 7),
           "oracle_SnowPack": (None, [d, d, C.POINTER(d), C.POINTER(d), d, d])}[name]
    f.restype, f.argtypes = sig
    return f


def noah_arrays(mesh, p, state, lai=None):
    """The oracle's per-element Noah records from the per-element dicts of
    pihm_amd.noahlsm (synthetic_noah / synthetic_noah_state) and the mesh's van Genuchten
    parameters.  Returns (par, st) structured arrays."""
    ne = mesh.nelem
    par = np.zeros(ne, NOAH_PAR_DT)
    par["nsoil"], par["nroot"], par["sldpth"] = p["nsoil"], p["nroot"], p["sldpth"]
    for k in _NP_SCALARS:
        if k in ("alpha", "beta"):
            par[k] = mesh.a[k]
        else:
            par[k] = p[k]
    st = np.zeros(ne, NOAH_STATE_DT)
    st["smc"], st["sh2o"], st["stc"] = state["smc"], state["sh2o"], state["stc"]
    st["t1"], st["sneqv"], st["snowh"], st["cmc"] = state["t1"], state["sneqv"], state["snowh"], state["cmc"]
    st["ch"] = st["cm"] = 1.0e-4
    st["albedo"], st["emissi"], st["q1"] = 0.18, 0.96, -999.0
    return par, st


def noah_meteo_interp(ftime, data, t):
    """The stations' records at t (IntrplForc): [7, nstation]."""
    L = lib()
    if not hasattr(L, "_interp_n_proto"):
        L.oracle_meteo_interp_n.restype = C.c_int
        L.oracle_meteo_interp_n.argtypes = [dp, dp, C.c_int, C.c_int, C.c_int, C.c_double, dp]
        L._interp_n_proto = True
    ft = np.ascontiguousarray(ftime, dtype=np.float64)
    dd = np.ascontiguousarray(data, dtype=np.float64)
    met = np.zeros((dd.shape[2], dd.shape[0]))
    if L.oracle_meteo_interp_n(_dptr(ft), _dptr(dd), dd.shape[0], dd.shape[1], dd.shape[2], float(t), _dptr(met)):
        raise ValueError("t outside the records")
    return met


def noah_forcing(met, station, lai):
    """ApplyMeteoForc with _NOAH_ (src/forcing.c:340-353) per element."""
    ne = len(station)
    f = np.zeros(ne, NOAH_FORC_DT)
    f["prcp"] = met[0][station] / 1000.0
    f["sfctmp"], f["rh"], f["sfcspd"] = met[1][station], met[2][station], met[3][station]
    f["soldn"] = np.maximum(met[4][station], 0.0)
    f["longwave"], f["sfcprs"] = met[5][station], met[6][station]
    f["proj_lai"] = lai
    return f
