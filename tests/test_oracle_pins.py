"""Pins of the CPU oracle against what the mathematics fixes.

The reference ships no tests, golden vectors or fixtures for its RHS, and its own build is
not made here (its headers need CMake-generated SUNDIALS configuration), so every oracle
function is checked against an independent route: closed forms, textbook formulas,
numerical calculus, invariants of the discrete system, or brute force on small inputs.
Each check is chosen so a plausible mistake (a dropped term, a flipped sign, a swapped
index or operand) fails it.
"""
import math

import numpy as np
import pytest
from scipy import integrate

import oracle as O

L = O.lib()
DEPRSTG = 1e-4
GRAV = 9.80665


# ----------------------------------------------------------------------------- SurfH
def test_surfh_parabola_and_continuity():
    # inverse of the depression-storage parabola surfeqv = surfh^2 / (2*DEPRSTG)
    for s in np.linspace(0.0, DEPRSTG / 2, 11):
        h = L.oracle_SurfH(s)
        assert abs(h * h / (2 * DEPRSTG) - s) <= 1e-18
    # above DEPRSTG/2 the slope is one and the value continuous at DEPRSTG/2
    assert L.oracle_SurfH(DEPRSTG / 2) == pytest.approx(DEPRSTG, rel=1e-15)
    eps = 1e-9
    left = (L.oracle_SurfH(DEPRSTG / 2) - L.oracle_SurfH(DEPRSTG / 2 - eps)) / eps
    right = (L.oracle_SurfH(DEPRSTG / 2 + eps) - L.oracle_SurfH(DEPRSTG / 2)) / eps
    assert left == pytest.approx(1.0, rel=1e-4) and right == pytest.approx(1.0, rel=1e-6)
    assert L.oracle_SurfH(0.3) == pytest.approx(0.3 + DEPRSTG / 2, rel=1e-15)
    assert L.oracle_SurfH(-1.0) == 0.0


# ------------------------------------------------------------------- van Genuchten / Mualem
def _vg_satn(h, alpha, n):
    m = 1.0 - 1.0 / n
    return (1.0 + (alpha * h) ** n) ** (-m)


@pytest.mark.parametrize("alpha,n", [(2.0, 1.3), (8.8, 1.24), (0.5, 2.5)])
def test_psi_inverts_van_genuchten(alpha, n):
    for h in [0.01, 0.1, 1.0, 5.0, 20.0]:
        s = _vg_satn(h, alpha, n)
        if s < 0.1:
            continue  # Psi clamps saturation at SATMIN
        assert L.oracle_Psi(s, alpha, n) == pytest.approx(-h, rel=1e-9)
    # clamp at SATMIN: Psi(0.05) == Psi(0.1)
    assert L.oracle_Psi(0.05, alpha, n) == L.oracle_Psi(0.1, alpha, n)


@pytest.mark.parametrize("alpha,n", [(2.0, 1.3), (8.8, 1.24), (0.5, 2.5)])
def test_krfunc_matches_mualem_in_head_form(alpha, n):
    m = 1.0 - 1.0 / n
    for h in [0.01, 0.1, 1.0, 5.0]:
        ah = alpha * h
        kr_h = (1.0 - ah ** (n - 1) * (1.0 + ah ** n) ** (-m)) ** 2 / (1.0 + ah ** n) ** (m / 2)
        s = _vg_satn(h, alpha, n)
        assert L.oracle_KrFunc(n, s) == pytest.approx(kr_h, rel=1e-9)
    assert L.oracle_KrFunc(n, 1.0) == 1.0


# ------------------------------------------------------------------- conductivities
def test_effkh_thickness_weighted():
    depth, dmac, kmach, areafv, ksath = 1.5, 1.0, 1e-2, 0.01, 1e-5
    kmac_layer = kmach * areafv + ksath * (1 - areafv)
    assert L.oracle_EffKh(depth, dmac, kmach, areafv, ksath, 0.3) == ksath       # below macropores
    for gw in (0.6, 1.0, 1.4):
        wet_mac = gw - (depth - dmac)
        ref = np.average([kmac_layer, ksath], weights=[wet_mac, depth - dmac])
        assert L.oracle_EffKh(depth, dmac, kmach, areafv, ksath, gw) == pytest.approx(ref, rel=1e-14)
    ref_full = np.average([kmac_layer, ksath], weights=[dmac, depth - dmac])
    assert L.oracle_EffKh(depth, dmac, kmach, areafv, ksath, 2.0) == pytest.approx(ref_full, rel=1e-14)
    assert L.oracle_EffKh(depth, dmac, kmach, areafv, ksath, -0.5) == ksath       # gw clamped at 0


def test_avgkv_arithmetic_layer_mean():
    dmac, ksatv, kmacv, areafh, kr = 1.0, 1e-5, 1e-3, 0.01, 0.3
    # deficit > dmac: macropore layer dry
    deficit, gw = 1.2, 0.3
    ref = np.average([kr * ksatv, kr * ksatv, ksatv], weights=[dmac, deficit - dmac, gw])
    assert L.oracle_AvgKv(dmac, ksatv, kmacv, areafh, deficit, gw, kr) == pytest.approx(ref, rel=1e-14)
    deficit, gw = 0.4, 1.1
    kmid = kmacv * areafh + ksatv * (1 - areafh)
    ref = np.average([kr * ksatv, kmid, ksatv], weights=[deficit, dmac - deficit, gw - (dmac - deficit)])
    assert L.oracle_AvgKv(dmac, ksatv, kmacv, areafh, deficit, gw, kr) == pytest.approx(ref, rel=1e-14)


def test_effkinf_regimes():
    kinfv, kmacv, areafh = 1e-5, 1e-3, 0.01
    kr, satn, dz = 0.4, 0.7, 2.0
    matrix = kinfv * kr
    mac = kinfv * (1 - areafh) * kr + kmacv * areafh
    assert L.oracle_EffKinf(kinfv, kmacv, 0.0, dz, kr, satn, 1.0, 0.0) == pytest.approx(kinfv * kr)
    assert L.oracle_EffKinf(kinfv, kmacv, areafh, dz, kr, satn, 0.0, 2 * DEPRSTG) == pytest.approx(mac)
    # application below matrix capacity -> matrix control
    assert L.oracle_EffKinf(kinfv, kmacv, areafh, dz, kr, satn, 0.5 * dz * matrix, 0.0) == pytest.approx(matrix)
    # above the macropore capacity -> macropore control
    assert L.oracle_EffKinf(kinfv, kmacv, areafh, dz, kr, satn, 10 * dz * mac, 0.0) == pytest.approx(mac)
    # in between: application control with crack conductivity Kr(2, satn)
    app = 0.5 * dz * (matrix + mac)
    ref = kinfv * (1 - areafh) * kr + kmacv * areafh * L.oracle_KrFunc(2.0, satn)
    assert L.oracle_EffKinf(kinfv, kmacv, areafh, dz, kr, satn, app, 0.0) == pytest.approx(ref)


# ------------------------------------------------------------------- gradients and flow laws
def test_dhbydl_recovers_plane_gradient():
    rng = np.random.default_rng(0)
    for _ in range(20):
        x = rng.uniform(0, 100, 3)
        y = rng.uniform(0, 100, 3)
        b, c, a0 = rng.normal(size=3)
        h = a0 + b * x + c * y
        X = (C := __import__("ctypes")).c_double * 3
        dx = L.oracle_DhByDl(X(*y), X(*x), X(*h))
        dy = L.oracle_DhByDl(X(*x), X(*y), X(*h))
        assert dx == pytest.approx(b, rel=1e-9, abs=1e-12)
        assert dy == pytest.approx(c, rel=1e-9, abs=1e-12)


def test_overland_flow_is_manning():
    for hh, s, w, n in [(0.01, 0.02, 30.0, 0.05), (0.2, 1e-4, 5.0, 0.1)]:
        manning = (hh * w) * hh ** (2.0 / 3.0) * math.sqrt(s) / n
        got = L.oracle_OverLandFlow(hh, s, s, hh * w, n)
        assert got == pytest.approx(manning, rel=1e-6)   # literal exponent 0.6666667


def test_weir_formula():
    cwr, length, zbed, zbank_r, ezmax = 0.6, 40.0, 99.0, 100.0, 100.0
    # free-flowing weir from the river: Q = (2/3) C sqrt(2g) L H^1.5, H over the bank
    stage = 1.3
    H = zbed + stage - zbank_r
    q = L.oracle_OvlFlowElemToRiver_raw(zbank_r, ezmax, 0.0, zbed, stage, cwr, length)
    assert q == pytest.approx(2.0 / 3.0 * cwr * math.sqrt(2 * GRAV) * length * H ** 1.5, rel=1e-14)
    # towards the river, ponded element over a low river: negative sign
    q2 = L.oracle_OvlFlowElemToRiver_raw(zbank_r, ezmax, 0.05, zbed, 0.5, cwr, length)
    assert q2 == pytest.approx(-2.0 / 3.0 * cwr * math.sqrt(2 * GRAV) * length * 0.05 ** 1.5, rel=1e-14)
    # submerged: sqrt of the head difference times the head over the bank
    q3 = L.oracle_OvlFlowElemToRiver_raw(zbank_r, ezmax, 0.1, zbed, 1.3, cwr, length)
    ref = 2.0 / 3.0 * cwr * math.sqrt(2 * GRAV) * length * math.sqrt(1.3 - 1.1) * (0.3)
    assert q3 == pytest.approx(ref, rel=1e-12)
    # element water below the depression storage and river low: no flow
    assert L.oracle_OvlFlowElemToRiver_raw(zbank_r, ezmax, 5e-5, zbed, 0.5, cwr, length) == 0.0


# ------------------------------------------------------------------- channel geometry
@pytest.mark.parametrize("order,coeff", [(1, 3.0), (2, 0.8), (3, 0.6), (4, 1.5)])
def test_cross_section_area_closed_form_and_top_width(order, coeff):
    for d in (0.05, 0.3, 1.2):
        a = L.oracle_RiverCroSectArea(order, d, coeff)
        if order == 1:
            ref = d * coeff
        elif order == 2:
            ref = 2 * integrate.quad(lambda x: d - coeff * x, 0, d / coeff)[0]
        elif order == 3:
            ref = 2 * integrate.quad(lambda x: d - coeff * x * x, 0, math.sqrt(d / coeff))[0]
        else:
            ref = 2 * integrate.quad(lambda x: d - coeff * x ** 3, 0, (d / coeff) ** (1 / 3))[0]
        assert a == pytest.approx(ref, rel=1e-9)
        # dA/d(depth) = top width; RiverEqWid is the top width RIVDPTHMIN above the depth
        eps = 1e-7
        dadd = (L.oracle_RiverCroSectArea(order, d + eps, coeff) - L.oracle_RiverCroSectArea(order, d - eps, coeff)) / (2 * eps)
        assert L.oracle_RiverEqWid(order, d - 0.05, coeff) == pytest.approx(dadd, rel=1e-6)
    assert L.oracle_RiverCroSectArea(order, -1.0, coeff) == 0.0


@pytest.mark.parametrize("order,coeff", [(1, 3.0), (2, 0.8), (3, 0.6)])
def test_wetted_perimeter_is_arc_length(order, coeff):
    for d in (0.05, 0.3, 1.2):
        p = L.oracle_RiverPerim(order, d, coeff)
        if order == 1:
            ref = 2 * d + coeff
        elif order == 2:
            ref = 2 * integrate.quad(lambda x: math.sqrt(1 + coeff ** 2), 0, d / coeff)[0]
        else:
            ref = 2 * integrate.quad(lambda x: math.sqrt(1 + 4 * coeff ** 2 * x * x), 0, math.sqrt(d / coeff))[0]
        assert p == pytest.approx(ref, rel=1e-9)


def test_cubic_perimeter_limits():
    # The CUBIC perimeter of the reference (src/river_flow.c:569-575) is not the arc length
    # of a closed-form section; pinned here by its limits: zero at zero depth, increasing,
    # and between the top width and top width + 2*depth.
    assert L.oracle_RiverPerim(4, 0.0, 1.5) == 0.0
    prev = 0.0
    for d in (0.01, 0.1, 0.5, 1.0, 2.0):
        p = L.oracle_RiverPerim(4, d, 1.5)
        assert p > prev
        prev = p


# ------------------------------------------------------------------- geometry initialisation
def test_geometry_numpy_equals_oracle_bitwise(c0, tree_small):
    for m in (c0, tree_small):
        g = O.init_geometry(m)
        a = m.a
        for k in ("area", "x", "y", "zmin", "zmax"):
            assert np.array_equal(a[k], g[k]), k
        assert np.array_equal(a["edge"].ravel(), g["edge"])
        assert np.array_equal(a["nabr"].ravel(), g["nabr"].ravel())
        for k in ("nabr_x", "nabr_y", "nabrdist"):
            assert np.array_equal(a[k].ravel(), g[k]), k
        for k in ("rzmax", "rzmin", "node_zmax", "dist_left", "dist_right", "length"):
            assert np.array_equal(a[k], g[k]), k


def test_geometry_against_first_principles(c0):
    m = c0
    a = m.a
    t = m.tri - 1
    P = np.stack([m.node_x[t], m.node_y[t]], -1)
    # area by the cross product, positive for counter-clockwise nodes
    cr = np.cross(P[:, 1] - P[:, 0], P[:, 2] - P[:, 0]) / 2
    assert np.allclose(a["area"], cr, rtol=1e-13) and (a["area"] > 0).all()
    # boundary neighbour distance = distance from circumcentre to the edge mid-point
    ax, ay, bx, by, cx, cy = P[:, 0, 0], P[:, 0, 1], P[:, 1, 0], P[:, 1, 1], P[:, 2, 0], P[:, 2, 1]
    dd = 2 * (ax * (by - cy) + bx * (cy - ay) + cx * (ay - by))
    ux = ((ax**2 + ay**2) * (by - cy) + (bx**2 + by**2) * (cy - ay) + (cx**2 + cy**2) * (ay - by)) / dd
    uy = ((ax**2 + ay**2) * (cx - bx) + (bx**2 + by**2) * (ax - cx) + (cx**2 + cy**2) * (bx - ax)) / dd
    mids = [(1, 2), (2, 0), (0, 1)]
    for j in range(3):
        bnd = a["nabr"][:, j] == 0
        p, q = mids[j]
        mx = 0.5 * (P[:, p, 0] + P[:, q, 0])
        my = 0.5 * (P[:, p, 1] + P[:, q, 1])
        ref = np.hypot(ux - mx, uy - my)
        assert np.allclose(a["nabrdist"][bnd, j], ref[bnd], rtol=1e-8)
        el = a["nabr"][:, j] > 0
        n = a["nabr"][el, j] - 1
        assert np.allclose(a["nabrdist"][el, j], np.hypot(a["x"][el] - a["x"][n], a["y"][el] - a["y"][n]), rtol=1e-14)
    # river tagging: each segment tags exactly one edge on each bank
    nr = m.nriver
    for r in range(nr):
        for e in (m.rivers["leftele"][r], m.rivers["rightele"][r]):
            assert (a["nabr"][e - 1] == -(r + 1)).sum() == 1


# ------------------------------------------------------------------- RHS invariants
def _states(m, rng, n=4):
    base = m.initial_state()
    ne, nr = m.nelem, m.nriver
    out = [base]
    for k in range(n):
        y = base.copy()
        y[:ne] = rng.choice([0.0, 2e-5, 8e-5, 3e-3, 0.05], ne) * rng.uniform(0.5, 1.5, ne)
        d = m.a["depth"]
        y[2 * ne:3 * ne] = rng.uniform(0.0, 1.1, ne) * d
        y[ne:2 * ne] = rng.uniform(0.0, 1.0, ne) * np.maximum(d - y[2 * ne:3 * ne], 0.0)
        if nr:
            y[3 * ne:3 * ne + nr] = rng.uniform(0.0, 2.0, nr) * m.a["rdepth"]
            y[3 * ne + nr:] = rng.uniform(0.0, 1.5, nr) * (m.a["zbed"] - m.a["rzmin"])
        neg = rng.random(len(y)) < 0.03
        y[neg] = -rng.uniform(0, 1e-4, neg.sum())
        out.append(y)
    return out


def _inputs(m, rng):
    ne = m.nelem
    return dict(pcpdrp=rng.uniform(0, 3e-6, ne), edir=rng.uniform(0, 2e-8, ne),
                ett=rng.uniform(0, 4e-8, ne), ws0_surf=rng.uniform(-1e-4, 2e-3, ne))


def test_rhs_global_mass_conservation(c0, tree_small):
    rng = np.random.default_rng(5)
    for m in (c0, tree_small):
        r = O.OracleRHS(m)
        inp = _inputs(m, rng)
        r.set_step_inputs(**inp)
        a = m.a
        ne, nr = m.nelem, m.nriver
        for y in _states(m, rng):
            rc, dy, mag = r.ode(y, with_mag=True)
            assert rc == 0
            rate = (np.sum(a["area"] * (dy[:ne] + a["porosity"] * (dy[ne:2 * ne] + dy[2 * ne:3 * ne])))
                    + np.sum(a["rarea"] * (dy[3 * ne:3 * ne + nr] + a["rporosity"] * dy[3 * ne + nr:])))
            rf = r.field("rivflow").reshape(nr, 11)
            sub = r.field("subsurf").reshape(ne, 3)
            bnd = a["nabr"] == 0
            src = (np.sum(a["area"] * (inp["pcpdrp"] - inp["edir"] - inp["ett"]))
                   - np.sum(sub[bnd]) - rf[a["down"] < 0, 1].sum())
            scale = np.sum(a["area"] * (inp["pcpdrp"] + inp["edir"] + inp["ett"]) + a["area"] * mag[:ne])
            assert abs(rate - src) <= 1e-11 * scale


def test_rhs_interior_fluxes_antisymmetric_bitwise(c0):
    m = c0
    rng = np.random.default_rng(2)
    r = O.OracleRHS(m)
    r.set_step_inputs(**_inputs(m, rng))
    a = m.a
    for y in _states(m, rng, 2):
        r.ode(y)
        ovl = r.field("ovlflow").reshape(-1, 3).copy()
        sub = r.field("subsurf").reshape(-1, 3).copy()
        for i in range(m.nelem):
            for j in range(3):
                n = a["nabr"][i, j]
                if n > 0:
                    k = int(np.nonzero(a["nabr"][n - 1] == i + 1)[0][0])
                    assert ovl[i, j] == -ovl[n - 1, k]
                    assert sub[i, j] == -sub[n - 1, k]


def test_rhs_hydrostatic_rest_state(c0):
    """Flat water table, no surface water, no forcing: zero lateral groundwater flux."""
    m = c0
    a = m.a
    ne, nr = m.nelem, m.nriver
    r = O.OracleRHS(m)
    r.set_step_inputs(pcpdrp=np.zeros(ne), edir=np.zeros(ne), ett=np.zeros(ne), ws0_surf=np.zeros(ne))
    H = float(np.min(a["zmax"])) - 0.1
    y = m.initial_state()
    y[:ne] = 0.0
    y[2 * ne:3 * ne] = np.clip(H - a["zmin"], 0.0, None)
    r.ode(y)
    sub = r.field("subsurf").reshape(ne, 3)
    el = a["nabr"] > 0
    assert np.all(sub[el][(y[2 * ne:3 * ne][np.nonzero(el)[0]] > 0)] == 0.0)


def test_rhs_bank_fluxes_mirror_river_terms(c0):
    m = c0
    rng = np.random.default_rng(3)
    r = O.OracleRHS(m)
    r.set_step_inputs(**_inputs(m, rng))
    a = m.a
    y = _states(m, rng, 1)[1]
    r.ode(y)
    ovl = r.field("ovlflow").reshape(-1, 3)
    sub = r.field("subsurf").reshape(-1, 3)
    rf = r.field("rivflow").reshape(-1, 11)
    for k in range(m.nriver):
        for side, e in enumerate((a["leftele"][k], a["rightele"][k])):
            j = int(np.nonzero(a["nabr"][e - 1] == -(k + 1))[0][0])
            assert ovl[e - 1, j] == -rf[k, 2 + side]
            assert sub[e - 1, j] == -(rf[k, 4 + side] + rf[k, 7 + side])
    # upstream inflow = minus the sum of the upstream segments' outflow
    for k in range(m.nriver):
        ups = np.nonzero(a["down"] == k + 1)[0]
        assert rf[k, 0] == pytest.approx(-rf[ups, 1].sum(), rel=1e-15, abs=0.0)


def test_rhs_infil_reads_previous_river_edge_flux(c0):
    """Infil uses the river-edge overland flux left by the previous call (carried state):
    a second call at the same state differs from the first exactly on bank elements."""
    m = c0
    rng = np.random.default_rng(4)
    r = O.OracleRHS(m)
    inp = _inputs(m, rng)
    inp["pcpdrp"][:] = 1e-7
    r.set_step_inputs(**inp)
    y = m.initial_state()
    y[:m.nelem] = 2e-5                       # wet, infiltration-limited surface
    y[3 * m.nelem:3 * m.nelem + m.nriver] = 1.5   # rivers over their banks -> flux onto banks
    _, d1 = r.ode(y)
    i1 = r.field("infil").copy()
    _, d2 = r.ode(y)
    i2 = r.field("infil").copy()
    bank = np.any(m.a["nabr"] < 0, axis=1)
    assert np.array_equal(i1[~bank], i2[~bank])
    assert np.any(i1[bank] != i2[bank])
