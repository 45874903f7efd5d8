"""Synthetic watershed generator (seeded, exact triangle and segment counts).

Geometry: a rectangular catchment meshed by a Delaunay triangulation (scipy/qhull) of a
jittered grid of interior points plus evenly spaced boundary points.  The point counts
are chosen so the triangulation has exactly ``n_tri`` triangles (a triangulation of n
points with h on the hull has 2n - 2 - h triangles).  Node elevation: a plane rising at
``cross_slope`` away from a central valley line (the "V"), rising ``along_slope`` away
from the outlet, plus N(0, noise) per node; soil depth is uniform.

Rivers run along mesh edges (each segment is one edge, its banks the two triangles that
share it, as the reference requires, src/init_river.c:24-35).  ``network="valley"`` lays
a single reach of ``n_seg`` segments up the valley from the outlet; ``network="tree"``
grows a binary-branching network (every junction joins exactly two upstream reaches): a
main stem up the valley, primary tributaries leaving it alternately to either side, and
secondary tributaries leaving the primaries alternately up- and down-valley; leaf reaches
are then trimmed so that the total is exactly ``n_seg``.

Per-element parameters: Ksat lognormal (median ``ksat_median``, sigma ``ksat_sigma``),
porosity U(0.35, 0.5), Manning n 0.05 overland / 0.04 in channels, and the remaining
van Genuchten / macropore parameters fixed (values in ``SOIL_DEFAULTS``).
"""
import math

import numpy as np
from scipy.spatial import Delaunay

from .mesh import Mesh, RECTANGLE, ZERO_DPTH_GRAD

SOIL_DEFAULTS = dict(alpha=2.0, beta=1.3, dinf=0.10, dmac=1.0, areafh=0.01, areafv=0.01,
                     kmacv_ro=100.0, kmach_ro=1000.0, smcmin=0.05, rzd=0.6, rough=0.05)
RIVER_DEFAULTS = dict(rough=0.04, cwr=0.6, ksath=1e-5, ksatv=1e-5, bedthick=0.5)


def _points(n_tri, lx, ly, rng):
    area = lx * ly
    d = math.sqrt(area / (n_tri / 2.0))
    h = max(4, int(round(2.0 * (lx + ly) / d)))
    if (n_tri - h) % 2:
        h += 1
    n_i = (n_tri - h + 2) // 2
    if n_i < 1:
        raise ValueError("mesh too small")
    # boundary: 4 corners + the rest spread over the edges by length
    rest = h - 4
    lens = np.array([lx, ly, lx, ly])
    counts = np.floor(rest * lens / lens.sum()).astype(int)
    while counts.sum() < rest:
        counts[np.argmax(rest * lens / lens.sum() - counts)] += 1
    corners = [(0.0, 0.0), (lx, 0.0), (lx, ly), (0.0, ly)]
    bpts = []
    for e in range(4):
        (x0, y0), (x1, y1) = corners[e], corners[(e + 1) % 4]
        bpts.append((x0, y0))
        k = counts[e]
        for i in range(k):
            f = (i + 1) / (k + 1) + rng.uniform(-0.2, 0.2) / (k + 1)
            bpts.append((x0 + f * (x1 - x0), y0 + f * (y1 - y0)))
    bpts = np.array(bpts)
    # interior: jittered grid, extra points removed at random
    nx = max(1, int(round(math.sqrt(n_i * lx / ly))))
    ny = int(math.ceil(n_i / nx))
    while nx * ny < n_i:
        ny += 1
    gx, gy = np.meshgrid((np.arange(nx) + 0.5) / nx * lx, (np.arange(ny) + 0.5) / ny * ly)
    ipts = np.stack([gx.ravel(), gy.ravel()], 1)
    ipts[:, 0] += rng.uniform(-0.25, 0.25, len(ipts)) * lx / nx
    ipts[:, 1] += rng.uniform(-0.25, 0.25, len(ipts)) * ly / ny
    keep = rng.permutation(len(ipts))[:n_i]
    ipts = ipts[np.sort(keep)]
    return np.concatenate([bpts, ipts]), h


def _triangulate(n_tri, lx, ly, rng):
    pts, h = _points(n_tri, lx, ly, rng)
    tri = Delaunay(pts)
    simp = tri.simplices.copy()
    nb = tri.neighbors.copy()
    if len(simp) != n_tri:
        raise RuntimeError(f"triangulation has {len(simp)} triangles, wanted {n_tri}")
    # counter-clockwise node order (src/include/elem_struct.h:1559-1560)
    P = pts[simp]
    cross = (P[:, 1, 0] - P[:, 0, 0]) * (P[:, 2, 1] - P[:, 0, 1]) - (P[:, 1, 1] - P[:, 0, 1]) * (P[:, 2, 0] - P[:, 0, 0])
    flip = cross < 0
    simp[flip] = simp[flip][:, [0, 2, 1]]
    nb[flip] = nb[flip][:, [0, 2, 1]]
    if np.any(np.abs(cross) < 1e-9 * lx * ly / n_tri):
        raise RuntimeError("degenerate triangle")
    return pts, simp, nb, h


class _EdgeMap:
    """Node graph of the triangulation with the triangles on each side of every edge."""

    def __init__(self, pts, simp):
        self.pts = pts
        self.adj = [[] for _ in range(len(pts))]
        self.edge_tris = {}
        for t, (a, b, c) in enumerate(simp):
            for u, v in ((a, b), (b, c), (c, a)):
                key = (min(u, v), max(u, v))
                self.edge_tris.setdefault(key, []).append(t)
        self.boundary_node = np.zeros(len(pts), bool)
        for (u, v), ts in self.edge_tris.items():
            self.adj[u].append(v)
            self.adj[v].append(u)
            if len(ts) == 1:
                self.boundary_node[u] = self.boundary_node[v] = True

    def walk(self, path, target_xy, blocked, n_more):
        """Extend ``path`` (a node list, last entry the tip) by up to ``n_more`` interior
        edges, each step to the free neighbour closest to ``target_xy`` that is closer than
        the tip (greedy routing on the Delaunay graph)."""
        pts = self.pts
        tx, ty = target_xy
        added = 0
        while added < n_more:
            cur = path[-1]
            best, bd = -1, math.hypot(pts[cur][0] - tx, pts[cur][1] - ty)
            for v in self.adj[cur]:
                if v in blocked:
                    continue
                if len(self.edge_tris[(min(cur, v), max(cur, v))]) != 2:
                    continue
                dv = math.hypot(pts[v][0] - tx, pts[v][1] - ty)
                if dv < bd:
                    best, bd = v, dv
            if best < 0:
                break
            path.append(best)
            blocked.add(best)
            added += 1
        return added


def _nearest(pts, xy, mask=None):
    d = np.hypot(pts[:, 0] - xy[0], pts[:, 1] - xy[1])
    if mask is not None:
        d = np.where(mask, d, np.inf)
    return int(np.argmin(d))


def _build_network(em, lx, ly, n_seg, network, rng):
    """Reaches as node paths listed downstream -> upstream, with parent links
    (parent reach, position of the junction node in the parent's path)."""
    pts = em.pts
    yc = ly / 2.0
    outlet = _nearest(pts, (0.0, yc), mask=em.boundary_node)
    blocked = set(np.nonzero(em.boundary_node)[0].tolist())
    blocked.add(outlet)
    reaches = []
    far = 1e9

    def reach(start, direction, parent, n):
        path = [start]
        tgt = (pts[start][0] + far * direction[0], pts[start][1] + far * direction[1])
        em.walk(path, tgt, blocked, n)
        reaches.append(dict(nodes=path, parent=parent, dir=direction))
        return len(reaches) - 1

    dmean = math.sqrt(lx * ly / len(pts)) * 1.07
    if network == "valley":
        r = reach(outlet, (1.0, 0.0), None, n_seg)
        if len(reaches[r]["nodes"]) - 1 != n_seg:
            raise RuntimeError("valley reach too short")
        return reaches
    # binary-branching tree: main stem, primaries alternating sides, secondaries
    # alternating up/down valley; every junction joins exactly two upstream reaches
    main_cap = int(0.95 * lx / dmean)
    main = reach(outlet, (1.0, 0.0), None, main_cap)
    mnodes = reaches[main]["nodes"]
    n_p = max(2, int(round(n_seg / 256.0)))
    prim_cap = int((ly / 2.0 - 3 * dmean) / dmean)
    prims = []
    for k in range(n_p):
        pos = int(round((k + 0.5) * (len(mnodes) - 1) / n_p))
        pos = min(max(pos, 1), len(mnodes) - 2)
        side = 1.0 if k % 2 == 0 else -1.0
        prims.append(reach(mnodes[pos], (0.0, side), (main, pos), prim_cap))
    used = sum(len(r["nodes"]) - 1 for r in reaches)
    plan = []
    sp2 = 10
    for p in prims:
        pn = reaches[p]["nodes"]
        s2 = 1.0
        for pos2 in range(5, len(pn) - 2, sp2):
            plan.append((p, pos2, (s2, 0.0)))
            s2 = -s2
    remaining = n_seg - used
    if remaining < 2 * len(plan):
        plan = plan[: max(0, remaining // 2)]
    per = remaining // max(1, len(plan))
    extra = remaining - per * len(plan)
    for i, (p, pos2, dr) in enumerate(plan):
        reach(reaches[p]["nodes"][pos2], dr, (p, pos2), per + (1 if i < extra else 0))
    # settle the total: extend tips (any reach) or trim leaf tips above their junctions
    top_junction = {}
    for i, r in enumerate(reaches):
        if r["parent"] is not None:
            pr, pp = r["parent"]
            top_junction[pr] = max(top_junction.get(pr, 0), pp)
    total = sum(len(r["nodes"]) - 1 for r in reaches)
    order = list(range(len(reaches)))
    rng.shuffle(order)
    while total != n_seg:
        progressed = False
        for i in order:
            if total == n_seg:
                break
            r = reaches[i]
            d = r["dir"]
            if total < n_seg:
                tip = r["nodes"][-1]
                tgt = (pts[tip][0] + far * d[0], pts[tip][1] + far * d[1])
                if em.walk(r["nodes"], tgt, blocked, 1):
                    total += 1
                    progressed = True
            else:
                if len(r["nodes"]) - 1 > max(1, top_junction.get(i, 0) + 1):
                    r["nodes"].pop()
                    total -= 1
                    progressed = True
        if not progressed:
            raise RuntimeError(f"cannot settle the network at {n_seg} segments (have {total})")
    return reaches


def _rivers_from_reaches(em, reaches, simp, depth_by_level, coeff_by_level):
    """Segment arrays in the reference's .riv meaning (FROM, TO, DOWN, LEFT, RIGHT ...)."""
    pts = em.pts
    seg_of = {}     # (reach, k) -> segment index; k = position of FROM node in reach path
    frm, to, lvl = [], [], []
    for ri, r in enumerate(reaches):
        nodes = r["nodes"]
        level = 0 if r["parent"] is None else 1 + (0 if reaches[r["parent"][0]]["parent"] is None else 1)
        for k in range(len(nodes) - 1, 0, -1):   # upstream to downstream, like the example .riv
            seg_of[(ri, k)] = len(frm)
            frm.append(nodes[k])
            to.append(nodes[k - 1])
            lvl.append(level)
    n = len(frm)
    down = np.empty(n, np.int32)
    for ri, r in enumerate(reaches):
        nodes = r["nodes"]
        for k in range(len(nodes) - 1, 0, -1):
            s = seg_of[(ri, k)]
            if k > 1:
                down[s] = seg_of[(ri, k - 1)] + 1
            elif r["parent"] is None:
                down[s] = ZERO_DPTH_GRAD
            else:
                pr, pos = r["parent"]
                down[s] = seg_of[(pr, pos)] + 1   # parent segment whose FROM is the junction
    left = np.empty(n, np.int32)
    right = np.empty(n, np.int32)
    for s in range(n):
        u, v = frm[s], to[s]
        ts = em.edge_tris[(min(u, v), max(u, v))]
        assert len(ts) == 2
        t0 = ts[0]
        c = [x for x in simp[t0] if x != u and x != v][0]
        dvec = pts[v] - pts[u]
        w = pts[c] - pts[u]
        if dvec[0] * w[1] - dvec[1] * w[0] > 0:
            left[s], right[s] = t0 + 1, ts[1] + 1
        else:
            left[s], right[s] = ts[1] + 1, t0 + 1
    lvl = np.array(lvl)
    riv = dict(fromnode=np.array(frm, np.int32) + 1, tonode=np.array(to, np.int32) + 1, down=down,
               leftele=left, rightele=right,
               intrpl_ord=np.full(n, RECTANGLE, np.int32),
               depth=np.array([depth_by_level[min(l, 2)] for l in lvl]),
               coeff=np.array([coeff_by_level[min(l, 2)] for l in lvl]),
               bc_type=np.zeros(n, np.int32))
    for k, v in RIVER_DEFAULTS.items():
        riv[k] = np.full(n, v)
    return riv


def make_watershed(n_tri, seed=0, lx=None, ly=None, n_seg=20, network="valley",
                   cross_slope=0.02, along_slope=0.005, noise=0.5, soil_depth=1.5,
                   ksat_median=1e-5, ksat_sigma=1.0, stepsize=60.0, param_mult=None):
    """Build a synthetic watershed ``Mesh`` with exactly ``n_tri`` triangles and ``n_seg``
    river segments.  Default domain: 1 km x 0.6 km at 535 triangles, scaled with the
    triangle count at the same element size.  ``param_mult`` (dict with keys ``ksat``,
    ``porosity``, ``rough``) multiplies the generated parameters (ensemble members)."""
    rng = np.random.default_rng(seed)
    if lx is None or ly is None:
        s = math.sqrt(n_tri / 535.0)
        lx, ly = 1000.0 * s, 600.0 * s
    for attempt in range(20):
        try:
            pts, simp, nb, h = _triangulate(n_tri, lx, ly, rng)
            break
        except RuntimeError:
            if attempt == 19:
                raise
    yc = ly / 2.0
    zmax = (100.0 + cross_slope * np.abs(pts[:, 1] - yc) + along_slope * pts[:, 0]
            + rng.normal(0.0, noise, len(pts)))
    em = _EdgeMap(pts, simp)
    reaches = _build_network(em, lx, ly, n_seg, network, rng) if n_seg > 0 else []
    # The zero-depth-gradient outlet takes sqrt of the bed slope between the outlet
    # segment's mid-point and its downstream node (src/river_flow.c:369-375): carve the
    # outlet node 0.2 m below its upstream neighbour so that slope is positive.
    if reaches:
        o_node, o_up = reaches[0]["nodes"][0], reaches[0]["nodes"][1]
        zmax[o_node] = min(zmax[o_node], zmax[o_up] - 0.2)
    zmin = zmax - soil_depth
    riv = _rivers_from_reaches(em, reaches, simp, depth_by_level=(1.0, 0.7, 0.5),
                               coeff_by_level=(3.0, 2.0, 1.5))
    ne = n_tri
    ksat = ksat_median * np.exp(rng.normal(0.0, ksat_sigma, ne))
    por = rng.uniform(0.35, 0.5, ne)
    rough = np.full(ne, SOIL_DEFAULTS["rough"])
    if param_mult:
        ksat = ksat * param_mult.get("ksat", 1.0)
        por = np.clip(por * param_mult.get("porosity", 1.0), 0.05, 0.9)
        rough = rough * param_mult.get("rough", 1.0)
        riv["rough"] = riv["rough"] * param_mult.get("rough", 1.0)
    S = SOIL_DEFAULTS
    soil = dict(ksath=ksat, ksatv=ksat, kinfv=ksat, smcmin=np.full(ne, S["smcmin"]),
                smcmax=S["smcmin"] + por, alpha=np.full(ne, S["alpha"]), beta=np.full(ne, S["beta"]),
                dinf=np.full(ne, S["dinf"]), dmac=np.full(ne, S["dmac"]),
                areafh=np.full(ne, S["areafh"]), areafv=np.full(ne, S["areafv"]),
                kmacv=S["kmacv_ro"] * ksat, kmach=S["kmach_ro"] * ksat, rough=rough,
                rzd=np.full(ne, S["rzd"]))
    nabr = np.where(nb >= 0, nb + 1, 0).astype(np.int32)
    mesh = Mesh(pts, zmin, zmax, simp + 1, nabr, riv, soil, stepsize=stepsize)
    mesh.seed = seed
    return mesh.init()


def config_mesh(name, seed=0):
    """The named workloads: 'c0' (535 triangles, 20 segments along the valley),
    'c1' (200,000 triangles, 4,095-segment binary-branching network) and 'c2'
    (100,000 triangles, 2,047 segments: the Noah-coupled workload's mesh)."""
    if name == "c0":
        return make_watershed(535, seed=seed, n_seg=20, network="valley")
    if name == "c1":
        return make_watershed(200_000, seed=seed, n_seg=4095, network="tree")
    if name == "c2":
        return make_watershed(100_000, seed=seed, n_seg=2047, network="tree")
    raise ValueError(name)
