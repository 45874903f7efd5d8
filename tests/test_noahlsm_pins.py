"""Pins of the oracle's Noah land-surface step (oracle/noahlsm_oracle.c) against what the
physics fixes, independently of the restatement's code:

* snow: Dyachkova's conductivity, the depletion curve's limits, Anderson's compaction as
  rho * (e^x - 1) / x;
* soil thermal conductivity: Johansen / Peters-Lidard at saturation (geometric mean of
  the constituents), dry soil, and frozen soil;
* surface exchange: the neutral log law, C_m = k^2 U / ln^2(z/z0) and
  C_h = k^2 U / (ln(z/z0) ln(z/zt)), and Paulson's functions through their defining
  integral psi'(zeta) = (1 - phi(zeta)) / zeta;
* soil temperature: layer-boundary interpolation exact on linear profiles, the layer
  mean of the sub-freezing part of a piecewise-linear profile, the supercooled-water
  equation of Koren et al. (1999) solved by a bracketing root finder;
* the whole step: the canopy and snow water balances, the transpiration total of the
  layer redistribution, soil evaporation's moisture control, and the heat budget of the
  soil column (heat content change = surface flux in - bottom flux out + latent heat
  of the water that froze) over one step.
"""
import ctypes as C
import math

import numpy as np
import pytest
from scipy import integrate, optimize

import oracle as O

TFREEZ = 273.15
LSUBF = 3.335e5
GRAV = 9.80665
VKRM = 0.40
MAXLYR = 11


def test_snow_conductivity_and_cover():
    cs = O.noah_fn("oracle_CSnow")
    for rho in (0.05, 0.2, 0.4):
        # Dyachkova (1960), doubled, in W m-1 K-1 (0.11631 W/m/K per cal/(cm h C))
        assert cs(rho) == pytest.approx(2 * 0.11631 * 0.328 * 10 ** (2.25 * rho), rel=1e-14)
    sf = O.noah_fn("oracle_SnFrac")
    snup, salp = 0.04, 2.6
    assert sf(0.0, snup, salp) == 0.0
    assert sf(snup * (1 - 1e-12), snup, salp) == pytest.approx(1.0, abs=1e-10)   # continuous at snup
    assert sf(2 * snup, snup, salp) == 1.0
    v = [sf(s, snup, salp) for s in np.linspace(0, snup, 20)]
    assert np.all(np.diff(v) > 0)


def test_snow_compaction_is_anderson():
    sp = O.noah_fn("oracle_SnowPack")
    for esd, rho, tsnow, tsoil in ((0.05, 0.1, 265.0, 270.0), (0.2, 0.25, 270.0, 272.0), (0.01, 0.3, 260.0, 262.0)):
        h, r = C.c_double(esd / rho), C.c_double(rho)
        dt = 900.0
        sp(esd, dt, C.byref(h), C.byref(r), tsnow, tsoil)
        x = (dt / 3600.0) * 0.01 * math.exp(0.08 * (0.5 * (tsnow + tsoil) - 273.15) - 21.0 * rho) * \
            max(esd * 100.0, 1e-2)
        ref = rho * (math.exp(x) - 1.0) / x
        assert r.value == pytest.approx(min(max(ref, 0.05), 0.40), rel=1e-9)   # 5-term series: x^5/720 left out
        assert h.value == pytest.approx(esd / r.value, rel=1e-14)
    # melting surface: 13% of liquid water per day retained
    h, r = C.c_double(0.5), C.c_double(0.2)
    sp(0.1, 3600.0, C.byref(h), C.byref(r), 274.0, 273.5)
    assert r.value > 0.2 and h.value == pytest.approx(0.1 / r.value, rel=1e-14)


@pytest.mark.parametrize("qz,phi", [(0.25, 0.4), (0.6, 0.45), (0.1, 0.35)])
def test_thermal_conductivity_johansen(qz, phi):
    td = O.noah_fn("oracle_TDfCnd")
    smcmin = 0.05
    ks = 7.7 ** qz * 2.0 ** (1 - qz)                       # solids: quartz and other minerals
    gd = (1 - phi) * 2700.0
    kdry = (0.135 * gd + 64.7) / (2700.0 - 0.947 * gd)
    # saturated, unfrozen: Kersten number 1, geometric mean of solids and water
    assert td(phi, qz, phi, smcmin, phi) == pytest.approx(ks ** (1 - phi) * 0.57 ** phi, rel=1e-13)
    # dry (relative saturation below 0.1): the dry conductivity
    s = smcmin + 0.05 * (phi - smcmin)
    assert td(s, qz, phi, smcmin, s) == pytest.approx(kdry, rel=1e-13)
    # unfrozen, partly saturated: Ke = log10(Sr) + 1
    s = smcmin + 0.5 * (phi - smcmin)
    ksat = ks ** (1 - phi) * 0.57 ** phi
    assert td(s, qz, phi, smcmin, s) == pytest.approx((math.log10(0.5) + 1) * (ksat - kdry) + kdry, rel=1e-12)
    # frozen: Ke = Sr, saturated conductivity with ice in the frozen pore fraction
    liq = 0.3 * s
    xu = liq / s * phi
    ksat_f = ks ** (1 - phi) * 2.2 ** (phi - xu) * 0.57 ** xu
    assert td(s, qz, phi, smcmin, liq) == pytest.approx(0.5 * (ksat_f - kdry) + kdry, rel=1e-12)


def test_exchange_coefficients_neutral_log_law():
    sd = O.noah_fn("oracle_SfcDifOff")
    zlvl, zlvl_wind, z0, czil = 3.0, 10.0, 0.1, 0.1
    for U in (1.5, 4.0, 9.0):
        lu = math.log((zlvl_wind + z0) / z0)
        cm0 = VKRM ** 2 * U / lu ** 2                      # the iteration's fixed point
        cm, ch = C.c_double(cm0), C.c_double(1e-3)
        sd(zlvl, zlvl_wind, z0, czil, U, 285.0, 285.0, C.byref(cm), C.byref(ch))
        ustar = VKRM * U / lu
        zt = math.exp(-czil * VKRM * 258.2 * math.sqrt(ustar * z0)) * z0      # Zilitinkevich
        lt = math.log((zlvl + zt) / zt)
        assert cm.value == pytest.approx(cm0, rel=1e-12)
        assert ch.value == pytest.approx(VKRM ** 2 * U / (lu * lt), rel=1e-12)
    # stable (warm air over a cold surface) transfers less heat than neutral, unstable more
    out = []
    for t1v in (280.0, 285.0, 290.0):
        cm, ch = C.c_double(0.01), C.c_double(0.01)
        sd(zlvl, zlvl_wind, z0, czil, 4.0, t1v, 285.0, C.byref(cm), C.byref(ch))
        out.append(ch.value)
    assert out[0] < out[1] < out[2]


def test_paulson_functions_are_their_integrals():
    pm, ph = O.noah_fn("oracle_Pspmu"), O.noah_fn("oracle_Psphu")
    assert pm(1.0) == pytest.approx(0.0, abs=1e-8) and ph(1.0) == pytest.approx(0.0, abs=1e-15)
    for zeta in (-0.01, -0.3, -2.0):
        x = lambda z: (1 - 16 * z) ** 0.25
        # psi(zeta) = int_0^zeta (1 - phi(z)) / z dz, phi_m = x^-1, phi_h = x^-2; the routines return -psi
        psim = integrate.quad(lambda z: (1 - 1 / x(z)) / z, zeta, 0)[0]
        psih = integrate.quad(lambda z: (1 - 1 / x(z) ** 2) / z, zeta, 0)[0]
        assert -pm(x(zeta)) == pytest.approx(-psim, rel=1e-6, abs=1e-7)
        assert -ph(x(zeta)) == pytest.approx(-psih, rel=1e-9)


def _zsoil(sld):
    return -np.cumsum(sld)


def test_boundary_temperature_exact_on_linear_profiles():
    tb = O.noah_fn("oracle_TBnd")
    sld = np.array([0.1, 0.2, 0.3, 0.4, 0.5])
    zs = np.ascontiguousarray(_zsoil(sld))
    zbot = -8.0
    T = lambda z: 283.0 + 0.7 * z
    zup = np.concatenate([[0.0], zs[:-1]])
    zmid = 0.5 * (zup + zs)
    zp = zs.ctypes.data_as(O.dp)
    for k in range(len(sld)):
        below = T(zmid[k + 1]) if k < len(sld) - 1 else T(zbot)
        assert tb(T(zmid[k]), below, zp, zbot, k, len(sld)) == pytest.approx(T(zs[k]), rel=1e-14)


def _clipped_integral(ta, tb, length):
    """Integral of min(T, TFREEZ) over a segment where T runs linearly from ta to tb."""
    if ta <= TFREEZ and tb <= TFREEZ:
        return length * (ta + tb) / 2
    if ta >= TFREEZ and tb >= TFREEZ:
        return length * TFREEZ
    x = length * (TFREEZ - ta) / (tb - ta)                       # where T crosses freezing
    cold, warm = (x, length - x) if ta < TFREEZ else (length - x, x)
    tcold = ta if ta < TFREEZ else tb
    return cold * (tcold + TFREEZ) / 2 + warm * TFREEZ


def test_layer_mean_of_the_subfreezing_part():
    ta = O.noah_fn("oracle_TmpAvg")
    sld = np.array([0.2, 0.3, 0.5])
    zs = np.ascontiguousarray(_zsoil(sld))
    zp = zs.ctypes.data_as(O.dp)
    rng = np.random.default_rng(1)
    for k in range(3):
        zup = 0.0 if k == 0 else zs[k - 1]
        dz = zup - zs[k]
        for case in range(8):
            signs = [(case >> b) & 1 for b in range(3)]          # 1: above freezing
            tup, tm, tdn = [TFREEZ + (s * 2 - 1) * rng.uniform(0.5, 5.0) for s in signs]
            if all(signs):
                ref = (tup + 2 * tm + tdn) / 4                   # all above freezing: the plain mean
            else:
                ref = (_clipped_integral(tup, tm, dz / 2) + _clipped_integral(tm, tdn, dz / 2)) / dz
            assert ta(tup, tm, tdn, zp, k) == pytest.approx(ref, rel=1e-10), (k, signs)


def _psi_vg(satn, alpha, n):
    m = 1 - 1 / n
    return (satn ** (-1 / m) - 1) ** (1 / n) / alpha


@pytest.mark.parametrize("tk", [272.0, 270.0, 265.0, 255.0])
def test_supercooled_water_solves_koren_equation(tk):
    fr = O.noah_fn("oracle_FrH2O")
    smcmax, smcmin, alpha, n = 0.45, 0.05, 3.0, 1.4
    smc = 0.35

    def resid(liq):        # |psi(liquid)| (1 + 8 ice)^2 = (Lf / g) (-dT / T)
        s = (liq - smcmin) / (smcmax - smcmin)
        return math.log(_psi_vg(s, alpha, n) * (1 + 8 * (smc - liq)) ** 2) - math.log(LSUBF / GRAV * (TFREEZ - tk) / tk)
    # less liquid water means more suction and more ice: the residual falls as liquid rises
    root = optimize.brentq(resid, smcmin + 1e-9, smc, xtol=1e-14) if resid(smc) < 0 else smc
    got = fr(tk, smc, smc, smcmax, smcmin, alpha, n)
    assert got == pytest.approx(root, abs=3e-3)      # Newton stops when the update is below 0.005
    assert 0.02 <= got <= smc
    assert fr(TFREEZ - 1e-4, smc, 0.1, smcmax, smcmin, alpha, n) == smc   # not below freezing: all liquid


# ------------------------------------------------------------------ whole steps
def _elements(ne, rng, cold=False, snow=False):
    par = np.zeros(ne, O.NOAH_PAR_DT)
    par["nsoil"] = 6
    par["nroot"] = rng.integers(1, 5, ne)
    par["sldpth"][:, :6] = np.array([0.1, 0.15, 0.2, 0.3, 0.4, 0.5])
    for k, lo, hi in (("shdfac", 0.3, 0.9), ("laimin", 0.5, 1.5), ("emissmin", 0.92, 0.94), ("albedomin", 0.12, 0.18),
                      ("z0min", 0.02, 0.2), ("snup", 0.02, 0.08), ("rgl", 30, 100), ("rsmin", 40, 300),
                      ("hs", 36, 55), ("quartz", 0.1, 0.6), ("smcmin", 0.03, 0.06), ("alpha", 1.0, 6.0),
                      ("beta", 1.2, 1.6)):
        par[k] = rng.uniform(lo, hi, ne)
    par["laimax"] = par["laimin"] + rng.uniform(1, 4, ne)
    par["emissmax"] = par["emissmin"] + 0.03
    par["albedomax"] = par["albedomin"] + 0.05
    par["z0max"] = par["z0min"] + 0.2
    par["smcmax"] = par["smcmin"] + rng.uniform(0.35, 0.45, ne)
    por = par["smcmax"] - par["smcmin"]
    par["smcwlt"] = par["smcmin"] + 0.2 * por
    par["smcref"] = par["smcmin"] + 0.7 * por
    par["smcdry"] = par["smcwlt"]
    par["cmcfactr"], par["cfactr"], par["rsmax"], par["topt"] = 2e-4, 0.5, 5000.0, 298.0
    for k, v in dict(csoil=2e6, sbeta=-2.0, salp=2.6, fxexp=2.0, czil=0.1, lvcoef=0.5, zbot=-8.0, tbot=282.15,
                     snoalb=0.75, zlvl=3.0, zlvl_wind=10.0).items():
        par[k] = v
    st = np.zeros(ne, O.NOAH_STATE_DT)
    frac = rng.uniform(0.25, 0.95, (ne, MAXLYR))
    st["smc"] = par["smcwlt"][:, None] + (par["smcmax"] - par["smcwlt"])[:, None] * frac
    st["sh2o"] = st["smc"]
    t0 = 268.0 if cold else 288.0
    st["stc"] = t0 + rng.uniform(-2, 2, (ne, MAXLYR))
    st["t1"] = t0
    st["cmc"] = rng.uniform(0, 1, ne) * 2e-4 * par["shdfac"] * 2.0
    st["cmc"][rng.random(ne) < 0.15] = 1e-9          # nearly empty canopies: drainage limited by the store
    st["ch"] = st["cm"] = 1e-3
    st["albedo"], st["emissi"], st["q1"] = 0.18, 0.96, -999.0
    if snow:
        st["sneqv"] = rng.uniform(0.005, 0.1, ne)
        st["sndens"] = rng.uniform(0.1, 0.3, ne)
        st["snowh"] = st["sneqv"] / st["sndens"]
    f = np.zeros(ne, O.NOAH_FORC_DT)
    f["sfctmp"] = (265.0 if cold else 290.0) + rng.uniform(-5, 5, ne)
    f["rh"] = rng.uniform(40, 95, ne)
    f["sfcspd"] = rng.uniform(0.5, 8, ne)
    f["soldn"] = np.where(rng.random(ne) < 0.5, rng.uniform(100, 800, ne), 0.0)
    f["longwave"] = rng.uniform(270, 330, ne)
    f["sfcprs"] = 101000.0
    f["prcp"] = np.where(rng.random(ne) < 0.4, rng.exponential(5.0, ne) / 3.6e6, 0.0)
    f["proj_lai"] = rng.uniform(0.3, 5, ne)
    return par, st, f


def test_canopy_water_balance_without_snow():
    rng = np.random.default_rng(3)
    par, st, f = _elements(400, rng)
    cmc0 = st["cmc"].copy()
    out = O.noah_lsm_step(900.0, par, st, f)
    dt = 900.0
    inside = (st["cmc"] > 0) & (st["cmc"] < out["cmcmax"]) & (st["sneqv"] == 0)
    assert inside.sum() > 50
    # rain (and dew) reaching the canopy, minus canopy evaporation and drip
    rain = f["prcp"] + out["dew"]
    lhs = (st["cmc"] - cmc0)[inside] / dt
    rhs = (par["shdfac"] * rain - out["ec"] - out["drip"])[inside]
    np.testing.assert_allclose(lhs, rhs, rtol=1e-9, atol=1e-18)
    # what reaches the ground: the part of the rain the canopy does not cover plus drip
    np.testing.assert_allclose(out["pcpdrp"], (1 - par["shdfac"]) * rain + out["drip"], rtol=1e-14, atol=1e-20)
    # drainage never takes more than the canopy holds
    avail = cmc0 + dt * (par["shdfac"] * rain - out["ec"])
    assert np.all(out["drip"] * dt <= np.maximum(avail, 0.0) * (1 + 1e-12) + 1e-30)
    assert np.any((cmc0 == 1e-9) & (out["drip"] > 0))


def test_transpiration_redistribution_keeps_the_total():
    rng = np.random.default_rng(4)
    par, st, f = _elements(400, rng)
    cmc0 = st["cmc"].copy()
    out = O.noah_lsm_step(900.0, par, st, f)
    et = out["et"]
    assert np.allclose(et.sum(1), out["ett"], rtol=1e-12, atol=1e-25)
    for i in range(len(par)):
        assert np.all(et[i, par["nroot"][i]:] == 0.0)
    # with positive demand the layers share the canopy's transpiration demand
    # shdfac * pc * etp * (1 - (cmc/cmcmax)^cfactr) whenever any root layer is above wilting
    pos = out["etp"] > 0
    demand = par["shdfac"] * out["pc"] * out["etp"] * (1 - np.where(cmc0 > 0, (cmc0 / out["cmcmax"]) ** 0.5, 0.0))
    wet = np.array([np.any(st["smc"][i, :par["nroot"][i]] > par["smcwlt"][i]) for i in range(len(par))])
    sel = pos & wet
    assert sel.sum() > 50
    np.testing.assert_allclose(out["ett"][sel] / 1.0, demand[sel], rtol=1e-9)
    assert np.all(out["ett"][~pos] == 0.0)


def test_layers_with_more_water_transpire_more():
    """With equal root fractions (equal layer thicknesses), a layer's share of the
    transpiration grows with its available water."""
    rng = np.random.default_rng(8)
    par, st, f = _elements(300, rng)
    par["sldpth"][:, :6] = 0.2
    par["nroot"] = 4
    cmc0 = st["cmc"].copy()
    out = O.noah_lsm_step(900.0, par, st, f)
    gx = np.clip((st["smc"][:, :4] - par["smcwlt"][:, None]) / (par["smcref"] - par["smcwlt"])[:, None], 0, 1)
    # positive demand: positive etp and a canopy store below its capacity (a store above the
    # capacity of this step's LAI makes the reference's demand negative)
    pos = (out["etp"] > 0) & (cmc0 < out["cmcmax"])
    n = 0
    for i in np.nonzero(pos)[0]:
        o = np.argsort(gx[i])
        if np.ptp(gx[i]) > 0.05:
            assert np.all(np.diff(out["et"][i, :4][o]) >= -1e-30), i
            n += 1
    assert n > 50


def test_soil_evaporation_moisture_control():
    rng = np.random.default_rng(5)
    par, st, f = _elements(300, rng)
    st["sh2o"][:, 0] = np.where(rng.random(300) < 0.5, 0.7 * st["smc"][:, 0], st["smc"][:, 0])   # partly frozen tops
    sh2o0 = st["sh2o"][:, 0].copy()
    out = O.noah_lsm_step(900.0, par, st, f)
    sr = (sh2o0 - par["smcdry"]) / (par["smcmax"] - par["smcdry"])
    fx = np.clip(np.where(sr > 0, np.maximum(sr, 0) ** par["fxexp"], 0.0), 0, 1)
    pos = out["etp"] > 0
    np.testing.assert_allclose(out["edir"][pos], (fx * (1 - par["shdfac"]) * out["etp"])[pos], rtol=1e-12)
    assert np.all(out["edir"][~pos] == 0.0)


def test_snow_water_balance():
    rng = np.random.default_rng(6)
    par, st, f = _elements(400, rng, cold=True, snow=True)
    swe0 = st["sneqv"].copy()
    dt = 900.0
    out = O.noah_lsm_step(dt, par, st, f)
    snowing = (f["prcp"] > 0) & (f["sfctmp"] < TFREEZ)
    keep = st["sneqv"] > 0
    assert keep.sum() > 100
    snowfall = np.where(snowing, f["prcp"], 0.0)
    # snowpack: + snowfall - sublimation (or + frost) - melt
    lhs = (st["sneqv"] - swe0)[keep] / dt
    rhs = (snowfall - out["esnow"] - out["snomlt"])[keep]
    frost = (out["etp"] <= 0)[keep]
    rhs = np.where(frost, (snowfall - out["etp"] * 1.0 - out["snomlt"])[keep], rhs)
    np.testing.assert_allclose(lhs, rhs, rtol=1e-9, atol=1e-16)
    # melt energy is the latent heat of the melted water
    np.testing.assert_allclose(out["flx3"], out["snomlt"] * 1000.0 * LSUBF, rtol=1e-12, atol=1e-12)


def _heat_capacity(par, smc, sh2o):
    # volumetric heat capacity of the mixture: water, solids, air (CP) and ice
    return sh2o * 4.2e6 + (1 - par["smcmax"][:, None]) * par["csoil"][:, None] + \
        (par["smcmax"][:, None] - smc) * 1004.0 + (smc - sh2o) * 2.106e3


@pytest.mark.parametrize("cold", [False, True])
def test_soil_column_heat_budget(cold):
    rng = np.random.default_rng(7 + cold)
    ne = 300
    par, st, f = _elements(ne, rng, cold=cold)
    if cold:        # some layers start partly frozen
        st["sh2o"] = np.where(rng.random((ne, MAXLYR)) < 0.5, 0.6 * st["smc"], st["smc"])
    smc0, sh2o0, stc0 = st["smc"].copy(), st["sh2o"].copy(), st["stc"].copy()
    dt = 900.0
    out = O.noah_lsm_step(dt, par, st, f)
    ns = 6
    sld = par["sldpth"][:, :ns]
    cap = _heat_capacity(par, smc0[:, :ns], sh2o0[:, :ns])
    heat = np.sum(cap * sld * (st["stc"][:, :ns] - stc0[:, :ns]), 1)
    latent = np.sum(1000.0 * LSUBF * sld * (sh2o0[:, :ns] - st["sh2o"][:, :ns]), 1)   # water that froze
    zs = -np.cumsum(sld, 1)
    zmid = zs[:, -1] + 0.5 * sld[:, -1]
    td = O.noah_fn("oracle_TDfCnd")
    kbot = np.array([td(smc0[i, ns - 1], par["quartz"][i], par["smcmax"][i], par["smcmin"][i], sh2o0[i, ns - 1])
                     for i in range(ne)])
    f_bot = kbot * (stc0[:, ns - 1] - par["tbot"]) / (zmid - par["zbot"])
    f_top = -out["ssoil"]              # SFlx returns the flux with the sign turned (positive: warms the surface)
    sel = st["sneqv"] == 0
    assert sel.sum() > 100
    scale = np.abs(f_top) + np.abs(f_bot) + np.abs(latent) / dt + 1.0
    np.testing.assert_array_less(np.abs(heat / dt - (f_top - f_bot + latent / dt))[sel], 1e-8 * scale[sel])
    if cold:
        assert np.any(latent != 0.0)


@pytest.mark.parametrize("warming", [False, True])
def test_phase_change_follows_the_heat_flow(warming):
    """The soil's frozen water moves towards equilibrium in the direction the heat flows:
    a column just below freezing under cold, dark, windy air loses heat at the top and its
    top layer's liquid water freezes, never thaws; a partly frozen column at equilibrium
    under warm sunshine gains heat and its top layer's ice thaws, never grows.  (The heat
    budget closes whatever the direction, since the latent term is taken from the water
    that changed phase.)"""
    rng = np.random.default_rng(31 + warming)
    ne = 200
    par, st, f = _elements(ne, rng, cold=True)
    frh2o = O.noah_fn("oracle_FrH2O")
    t0 = TFREEZ - (0.3 if warming else 0.5)
    st["stc"] = t0
    st["t1"] = t0
    if warming:      # liquid water at the supercooled-water equilibrium of t0
        for i in range(ne):
            for k in range(MAXLYR):
                st["sh2o"][i, k] = frh2o(t0, st["smc"][i, k], st["smc"][i, k], par["smcmax"][i],
                                         par["smcmin"][i], par["alpha"][i], par["beta"][i])
        f["sfctmp"], f["soldn"], f["longwave"] = TFREEZ + 20.0, 800.0, 360.0
    else:
        f["sfctmp"], f["soldn"], f["longwave"] = TFREEZ - 25.0, 0.0, 180.0
    f["prcp"] = 0.0
    f["sfcspd"] = 6.0
    sh2o0 = st["sh2o"][:, 0].copy()
    O.noah_lsm_step(900.0, par, st, f)
    d = st["sh2o"][:, 0] - sh2o0
    if warming:
        assert np.all(d >= 0.0) and np.mean(d > 0.0) > 0.9
    else:
        assert np.all(d <= 0.0) and np.mean(d < 0.0) > 0.9
