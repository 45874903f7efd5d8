"""Watershed mesh in the reference's index conventions, and its initialisation.

A ``Mesh`` holds what MM-PIHM reads from ``.mesh``/``.att``/``.riv``/``.soil`` (nodes,
triangles with their neighbours, river segments, per-element soil and land-cover
parameters), and ``Mesh.init()`` derives the per-element and per-segment quantities the
RHS uses, restating the reference's initialisation:

* ``InitTopo``  (src/init_topo.c:6-47): area, centroid, mean zmin/zmax, edge lengths;
* ``InitRiver`` (src/init_river.c:6-97): river-edge tagging ``nabr = -(seg+1)``, river
  topology, shape, material, width and area;
* ``InitSurfL`` (src/initialize.c:385-458): neighbour coordinates and distances;
* ``InitSoil``  (src/init_soil.c:4-74): soil depth, porosity, macropore depth cap.

All arrays use the caller's (original) order; indices are 1-based where the reference's
are (``node``, ``nabr``, ``leftele``/``rightele``, ``down``).
"""
import ctypes as C

import numpy as np

from . import _lib

RECTANGLE, TRIANGLE, QUADRATIC, CUBIC = 1, 2, 3, 4
DIRICHLET, NEUMANN, ZERO_DPTH_GRAD, CRIT_DPTH = -1, -2, -3, -4
KINEMATIC, DIFF_WAVE = 1, 2
RIVDPTHMIN = 0.05


def river_eq_wid(order, depth, coeff):
    """RiverEqWid (src/init_river.c:99-124), vectorised."""
    order = np.asarray(order)
    depth = np.maximum(np.asarray(depth, dtype=np.float64), 0.0)
    depth = np.where(depth > 0.0, depth, 0.0)
    coeff = np.asarray(coeff, dtype=np.float64)
    out = np.empty(np.broadcast(order, depth, coeff).shape, dtype=np.float64)
    rect = order == RECTANGLE
    out[rect] = np.broadcast_to(coeff, out.shape)[rect]
    other = ~rect
    if other.any():
        o = np.broadcast_to(order, out.shape)[other].astype(np.float64)
        d = np.broadcast_to(depth, out.shape)[other]
        c = np.broadcast_to(coeff, out.shape)[other]
        out[other] = 2.0 * np.power(d + RIVDPTHMIN, 1.0 / (o - 1.0)) / np.power(c, 1.0 / (o - 1.0))
    return out


class Mesh:
    """Raw watershed description plus the initialised RHS arrays (after ``init()``)."""

    def __init__(self, node_xy, node_zmin, node_zmax, tri, nabr, rivers, soil, surf_mode=DIFF_WAVE,
                 riv_mode=DIFF_WAVE, stepsize=60.0, bc_type=None):
        self.node_x = np.ascontiguousarray(node_xy[:, 0], dtype=np.float64)
        self.node_y = np.ascontiguousarray(node_xy[:, 1], dtype=np.float64)
        self.node_zmin = np.ascontiguousarray(node_zmin, dtype=np.float64)
        self.node_zmax = np.ascontiguousarray(node_zmax, dtype=np.float64)
        self.tri = np.ascontiguousarray(tri, dtype=np.int32)          # [ne,3] 1-based node ids
        self.nabr_raw = np.ascontiguousarray(nabr, dtype=np.int32)    # [ne,3] 1-based, 0 = boundary
        self.rivers = rivers        # dict of per-segment arrays (see meshgen)
        self.soil = soil            # dict of per-element arrays
        self.surf_mode = surf_mode
        self.riv_mode = riv_mode
        self.stepsize = float(stepsize)
        ne = self.tri.shape[0]
        self.bc_type = (np.zeros((ne, 3), np.int32) if bc_type is None
                        else np.ascontiguousarray(bc_type, dtype=np.int32))
        self.a = None
        self._keep = []

    @property
    def nelem(self):
        return self.tri.shape[0]

    @property
    def nriver(self):
        return len(self.rivers["down"])

    @property
    def neq(self):
        return 3 * self.nelem + 2 * self.nriver

    # ------------------------------------------------------------------ init
    def init(self):
        """Derive the RHS arrays (InitTopo, InitRiver, InitSurfL, InitSoil)."""
        ne, nr = self.nelem, self.nriver
        t = self.tri - 1
        X = self.node_x[t]
        Y = self.node_y[t]
        Zn = self.node_zmin[t]
        Zx = self.node_zmax[t]
        a = {}
        # InitTopo: src/init_topo.c:30-46
        a["area"] = 0.5 * ((X[:, 1] - X[:, 0]) * (Y[:, 2] - Y[:, 0]) - (Y[:, 1] - Y[:, 0]) * (X[:, 2] - X[:, 0]))
        a["x"] = (X[:, 0] + X[:, 1] + X[:, 2]) / 3.0
        a["y"] = (Y[:, 0] + Y[:, 1] + Y[:, 2]) / 3.0
        a["zmin"] = (Zn[:, 0] + Zn[:, 1] + Zn[:, 2]) / 3.0
        a["zmax"] = (Zx[:, 0] + Zx[:, 1] + Zx[:, 2]) / 3.0
        edge = np.empty((ne, 3))
        edge[:, 0] = np.sqrt((X[:, 1] - X[:, 2]) ** 2 + (Y[:, 1] - Y[:, 2]) ** 2)
        edge[:, 1] = np.sqrt((X[:, 2] - X[:, 0]) ** 2 + (Y[:, 2] - Y[:, 0]) ** 2)
        edge[:, 2] = np.sqrt((X[:, 0] - X[:, 1]) ** 2 + (Y[:, 0] - Y[:, 1]) ** 2)
        a["edge"] = edge

        # InitRiver: src/init_river.c:13-96
        rv = self.rivers
        nabr = self.nabr_raw.copy()
        left = np.asarray(rv["leftele"], np.int32)
        right = np.asarray(rv["rightele"], np.int32)
        for i in range(nr):
            L, R = left[i] - 1, right[i] - 1
            for j in range(3):
                if nabr[L, j] == right[i]:
                    nabr[L, j] = -(i + 1)
                if nabr[R, j] == left[i]:
                    nabr[R, j] = -(i + 1)
        a["nabr"] = nabr
        fn = np.asarray(rv["fromnode"], np.int64) - 1
        tn = np.asarray(rv["tonode"], np.int64) - 1
        Li, Ri = left.astype(np.int64) - 1, right.astype(np.int64) - 1
        rx = 0.5 * (self.node_x[fn] + self.node_x[tn])
        ry = 0.5 * (self.node_y[fn] + self.node_y[tn])
        rzmax = 0.5 * (self.node_zmax[fn] + self.node_zmax[tn])
        rzmin = rzmax - (0.5 * (a["zmax"][Li] + a["zmax"][Ri]) - 0.5 * (a["zmin"][Li] + a["zmin"][Ri]))
        a["rx"], a["ry"], a["rzmax"], a["rzmin"] = rx, ry, rzmax, rzmin
        a["node_zmax"] = self.node_zmax[tn].copy()
        a["dist_left"] = np.sqrt((rx - a["x"][Li]) * (rx - a["x"][Li]) + (ry - a["y"][Li]) * (ry - a["y"][Li]))
        a["dist_right"] = np.sqrt((rx - a["x"][Ri]) * (rx - a["x"][Ri]) + (ry - a["y"][Ri]) * (ry - a["y"][Ri]))
        ord_ = np.asarray(rv["intrpl_ord"], np.int32)
        rdepth = np.asarray(rv["depth"], np.float64)
        coeff = np.asarray(rv["coeff"], np.float64)
        a["rdepth"], a["intrpl_ord"], a["coeff"] = rdepth, ord_, coeff
        a["length"] = np.sqrt((self.node_x[fn] - self.node_x[tn]) ** 2 + (self.node_y[fn] - self.node_y[tn]) ** 2)
        a["width"] = river_eq_wid(ord_, rdepth, coeff) if nr else np.zeros(0)
        a["zbed"] = rzmax - rdepth
        for k in ("rough", "cwr", "ksath", "ksatv", "bedthick"):
            a["r" + k if k in ("rough", "ksath", "ksatv") else k] = np.asarray(rv[k], np.float64)

        # InitSoil: src/init_soil.c:26-67
        s = self.soil
        a["depth"] = a["zmax"] - a["zmin"]
        a["porosity"] = np.asarray(s["smcmax"], np.float64) - np.asarray(s["smcmin"], np.float64)
        dmac = np.asarray(s["dmac"], np.float64)
        a["dmac"] = np.where(dmac > a["depth"], a["depth"], dmac)
        for k in ("ksath", "ksatv", "kinfv", "dinf", "alpha", "beta", "kmach", "kmacv", "areafv",
                  "areafh", "rough", "rzd"):
            a[k] = np.ascontiguousarray(np.broadcast_to(np.asarray(s[k], np.float64), (ne,)))
        a["rporosity"] = 0.5 * (a["porosity"][Li] + a["porosity"][Ri]) if nr else np.zeros(0)
        a["rarea"] = a["length"] * river_eq_wid(ord_, rdepth, coeff) if nr else np.zeros(0)

        # InitSurfL: src/initialize.c:411-456
        nabr_x = np.empty((ne, 3))
        nabr_y = np.empty((ne, 3))
        nabrdist = np.empty((ne, 3))
        mids = [(1, 2), (2, 0), (0, 1)]
        circ2 = (edge[:, 0] * edge[:, 1] * edge[:, 2] / (4.0 * a["area"])) ** 2
        for j in range(3):
            p, q = mids[j]
            distx = a["x"] - 0.5 * (X[:, p] + X[:, q])
            disty = a["y"] - 0.5 * (Y[:, p] + Y[:, q])
            nb = nabr[:, j]
            bnd = nb == 0
            el = nb > 0
            rv_ = nb < 0
            nx = np.empty(ne)
            ny = np.empty(ne)
            nx[bnd] = a["x"][bnd] - 2.0 * distx[bnd]
            ny[bnd] = a["y"][bnd] - 2.0 * disty[bnd]
            nx[el] = a["x"][nb[el] - 1]
            ny[el] = a["y"][nb[el] - 1]
            nx[rv_] = rx[-nb[rv_] - 1]
            ny[rv_] = ry[-nb[rv_] - 1]
            d = (a["x"] - nx) * (a["x"] - nx)
            d = d + (a["y"] - ny) * (a["y"] - ny)
            d = np.sqrt(d)
            d[bnd] = np.sqrt(circ2[bnd] - (edge[bnd, j] / 2.0) ** 2)
            nabr_x[:, j], nabr_y[:, j], nabrdist[:, j] = nx, ny, d
        a["nabr_x"], a["nabr_y"], a["nabrdist"] = nabr_x, nabr_y, nabrdist

        a["bc_type"] = self.bc_type
        a["down"] = np.asarray(rv["down"], np.int32)
        a["leftele"], a["rightele"] = left, right
        a["riverbc_type"] = np.asarray(rv.get("bc_type", np.zeros(nr, np.int32)), np.int32)
        self.a = {k: np.ascontiguousarray(v) for k, v in a.items()}
        return self

    # ------------------------------------------------------------------ C view
    def desc(self):
        """pihm_mesh_desc pointing at this mesh's arrays (kept alive by the mesh)."""
        if self.a is None:
            self.init()
        a = self.a
        d = _lib.MeshDesc()
        d.nelem, d.nriver = self.nelem, self.nriver
        d.surf_mode, d.riv_mode, d.stepsize = self.surf_mode, self.riv_mode, self.stepsize

        def dp(arr):
            arr = np.ascontiguousarray(arr, dtype=np.float64)
            self._keep.append(arr)
            return arr.ctypes.data_as(_lib.c_double_p)

        def ip(arr):
            arr = np.ascontiguousarray(arr, dtype=np.int32)
            self._keep.append(arr)
            return arr.ctypes.data_as(_lib.c_int32_p)

        d.nabr, d.bc_type = ip(a["nabr"]), ip(a["bc_type"])
        for k in ("x", "y", "area", "zmin", "zmax", "edge", "nabrdist", "nabr_x", "nabr_y", "depth",
                  "ksath", "ksatv", "kinfv", "dinf", "alpha", "beta", "porosity", "dmac", "kmach",
                  "kmacv", "areafv", "areafh", "rough", "rzd"):
            setattr(d, k, dp(a[k]))
        d.down, d.leftele, d.rightele = ip(a["down"]), ip(a["leftele"]), ip(a["rightele"])
        d.riverbc_type, d.intrpl_ord = ip(a["riverbc_type"]), ip(a["intrpl_ord"])
        d.rarea, d.rzmin, d.rzmax, d.zbed = dp(a["rarea"]), dp(a["rzmin"]), dp(a["rzmax"]), dp(a["zbed"])
        d.node_zmax, d.dist_left, d.dist_right = dp(a["node_zmax"]), dp(a["dist_left"]), dp(a["dist_right"])
        d.rdepth, d.coeff, d.length, d.width = dp(a["rdepth"]), dp(a["coeff"]), dp(a["length"]), dp(a["width"])
        d.rrough, d.cwr, d.rksath, d.rksatv = dp(a["rrough"]), dp(a["cwr"]), dp(a["rksath"]), dp(a["rksatv"])
        d.bedthick, d.rporosity = dp(a["bedthick"]), dp(a["rporosity"])
        return d

    # ------------------------------------------------------------------ helpers
    def initial_state(self, gw_frac=0.5, satn=0.5, surf=0.0, stage=0.1):
        """Initial state vector in caller order (surface, unsat, gw, stage, river gw)."""
        if self.a is None:
            self.init()
        a = self.a
        ne, nr = self.nelem, self.nriver
        y = np.zeros(self.neq)
        gw = gw_frac * a["depth"]
        y[:ne] = surf
        y[ne:2 * ne] = satn * (a["depth"] - gw)
        y[2 * ne:3 * ne] = gw
        if nr:
            y[3 * ne:3 * ne + nr] = stage
            y[3 * ne + nr:] = 0.5 * (a["zbed"] - a["rzmin"]).clip(min=0.0)
        return y

    def storage(self, y):
        """Water volume [m^3] held in state y (surface + porosity*(unsat+gw) + river)."""
        a = self.a
        ne, nr = self.nelem, self.nriver
        v = np.sum(a["area"] * (y[:ne] + a["porosity"] * (y[ne:2 * ne] + y[2 * ne:3 * ne])))
        if nr:
            v += np.sum(a["rarea"] * (y[3 * ne:3 * ne + nr] + a["rporosity"] * y[3 * ne + nr:]))
        return float(v)
