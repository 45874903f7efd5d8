// noahlsm.hip - the Noah land-surface step of the _NOAH_ build on the device.
//
// One thread per element runs Noah() (src/noah/noah.c:3-55): the humidity
// terms, the moisture clamps and SFlx (:166-609) with every routine it calls,
// in the reference's operation order.  The soil column (smc, sh2o, at most 11
// layers) is shared with the soil-water step and lives in HBM layer-major
// ([k * ne + p]) so that a wavefront reads one layer of 64 neighbouring
// elements together; the temperatures stc are stored the same way.  Inside
// the thread the column is worked in private arrays.  The step ends by
// handing the RHS what it reads (pcpdrp, edir, ett, the layer transpiration,
// sh2o[0]) and by refreshing GwTransp's share.  exp, log, log10, atan, sqrt
// and pow are the device library's.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/pihm_noahlsm.h"
#include "common.hpp"
#include "ctx.hpp"
#include "tsinterp.hpp"

using namespace pihm;

namespace {

constexpr int ML = PIHM_NOAH_MAXLYR;
// src/include/pihm_const.h:6-37,81
constexpr double NPI = 3.14159265;
constexpr double NGRAV = 9.80665;
constexpr double NCP = 1004.0;
constexpr double LVH2O = 2.501e6;
constexpr double SIGMA = 5.67e-8;
constexpr double RD = 287.04;
constexpr double CPH2O = 4.218e3;
constexpr double CPICE = 2.106e3;
constexpr double LSUBF = 3.335e5;
constexpr double EMISSI_S = 0.95;
constexpr double TFREEZ = 273.15;
constexpr double LSUBS = 2.83e6;
constexpr double BADVAL = -999.0;
constexpr double SH2OMIN = 0.02;
constexpr int NSC = PIHM_NOAH_NSCALAR;
constexpr int NFX = PIHM_NOAH_NFLUX;

struct NoahDev {
  // static, device order
  int32_t *nroot, *station;
  double *shdfac, *cmcfactr, *laimin, *laimax, *emissmin, *emissmax, *albedomin, *albedomax, *z0min, *z0max;
  double *snup, *cfactr, *rgl, *rsmin, *rsmax, *topt, *hs, *smcref, *smcwlt, *smcdry, *quartz, *zlvl_wind;
  double *lai;
  // carried: stc [ML * ne]; sc [NSC * ne] (t1, sneqv, snowh, sndens, cmc, ch, cm, snotime1, albedo, emissi)
  double *stc, *sc, *q1;
  // the last step's fluxes [NFX * ne]
  double *flux;
  double *met;          // [PIHM_NOAH_MET_NVAR * nstation]
};

struct NoahUni {
  double csoil, sbeta, salp, fxexp, czil, lvcoef, zbot, tbot, snoalb, zlvl;
};

// the soil column and the context fields the step reads and writes
struct ColDev {
  int32_t *nsoil;
  double *sldpth, *smcmax, *smc, *sh2o;
  double *smcmin, *alpha, *beta;               // context statics
  double *pcpdrp, *edir, *ett, *et, *sh2o0;    // context step inputs and Noah hand-over
};

// One element's working record: the reference's ps/ws/wf/es/ef fields the step touches.
struct Col {
  int nsoil, nroot;
  double zsoil[ML], rtdis[ML], smc[ML], sh2o[ML], stc[ML], et[ML];
  // parameters
  double shdfac, cmcfactr, laimin, laimax, emissmin, emissmax, albedomin, albedomax, z0min, z0max, snup, cfactr;
  double rgl, rsmin, rsmax, topt, hs, smcmax, smcmin, smcref, smcwlt, smcdry, quartz, alpha, beta, zlvl_wind;
  // carried
  double t1, sneqv, snowh, sndens, cmc, ch, cm, snotime1, albedo, emissi, q1;
  // forcing
  double sfctmp, rh, sfcspd, soldn, longwave, sfcprs, prcp, proj_lai;
  // diagnostics
  double q2, q2sat, dqsdt2, th2, ffrozp, alb, embrd, z0brd, z0, cmcmax, sncovr, sncond;
  double rr, rch, epsca, rc, pc, beta_, eta_kinematic;
  double pcpdrp, snomlt, edir, ec, ett, etp, eta, etns, esnow, dew, drip;
  double solnet, lwdn, fdown, ssoil, flx1, flx2, flx3, sheat, ef_edir, ef_ec, ef_ett, ef_esnow, ef_etp, ef_eta;
};

// CalHum: src/noah/noah.c:659-702
__device__ void cal_hum(Col &c, const NoahUni &u) {
  const double A2 = 17.67, A3 = 273.15, A4 = 29.65, ELWV = 2.501e6, E0 = 611.0, RVV = 461.0, EPSILON = 0.622;
  const double SVP1 = 611.2, SVP2 = 17.67, SVP3 = 29.65, SVPT0 = 273.15;
  const double rh = c.rh / 100.0;
  const double svp = SVP1 * exp(SVP2 * (c.sfctmp - SVPT0) / (c.sfctmp - SVP3));
  const double e = rh * svp;
  c.q2 = (0.622 * e) / (c.sfcprs - (1.0 - 0.622) * e);
  c.th2 = c.sfctmp + (0.0098 * u.zlvl);
  const double a23m4 = A2 * (A3 - A4);
  const double esat = E0 * exp(ELWV / RVV * (1.0 / A3 - 1.0 / c.sfctmp));
  c.q2sat = EPSILON * esat / (c.sfcprs - (1.0 - EPSILON) * esat);
  c.dqsdt2 = c.q2sat * a23m4 / ((c.sfctmp - A4) * (c.sfctmp - A4));
}

// CSnow: src/noah/noah.c:811-839
__device__ double c_snow(double dsnow) {
  const double UNIT = 0.11631;
  const double cc = 0.328 * pow(10.0, 2.25 * dsnow);
  return 2.0 * UNIT * cc;
}

// SnowNew: src/noah/noah.c:2222-2268
__device__ void snow_new(Col &c, double newsn) {
  double snowhc = c.snowh * 100.0;
  const double newsnc = newsn * 100.0;
  const double tempc = c.sfctmp - 273.15;
  const double dsnew = (tempc <= -15.0) ? 0.05 : 0.05 + 0.0017 * pow(tempc + 15.0, 1.5);
  const double hnewc = newsnc / dsnew;
  if (snowhc + hnewc < 0.001) c.sndens = (dsnew > c.sndens) ? dsnew : c.sndens;
  else c.sndens = (snowhc * c.sndens + hnewc * dsnew) / (snowhc + hnewc);
  snowhc = snowhc + hnewc;
  c.snowh = snowhc * 0.01;
}

// SnFrac: src/noah/noah.c:1713-1744
__device__ double sn_frac(double sneqv, double snup, double salp) {
  if (sneqv < snup) {
    const double rsnow = sneqv / snup;
    return 1.0 - (exp(-salp * rsnow) - rsnow * exp(-salp));
  }
  return 1.0;
}

// AlCalc: src/noah/noah.c:610-657
__device__ void al_calc(Col &c, const NoahUni &u, double dt, int snowng) {
  const double SNACCA = 0.94, SNACCB = 0.58;
  const double snoalb = u.snoalb;
  c.albedo = c.alb + c.sncovr * (snoalb - c.alb);
  c.emissi = c.embrd + c.sncovr * (EMISSI_S - c.embrd);
  const double snoalb1 = snoalb + u.lvcoef * (0.85 - snoalb);
  double snoalb2 = snoalb1;
  if (snowng) {
    c.snotime1 = 0.0;
  } else {
    c.snotime1 += dt;
    snoalb2 = snoalb1 * pow(SNACCA, pow(c.snotime1 / 86400.0, SNACCB));
  }
  snoalb2 = (snoalb2 > c.alb) ? snoalb2 : c.alb;
  c.albedo = c.alb + c.sncovr * (snoalb2 - c.alb);
  c.albedo = (c.albedo > snoalb2) ? snoalb2 : c.albedo;
}

// TDfCnd: src/noah/noah.c:2520-2619
__device__ double tdf_cnd(double smc, double qz, double smcmax, double smcmin, double sh2o) {
  const double THKICE = 2.2, THKQTZ = 7.7, THKW = 0.57;
  const double satratio = (smc - smcmin) / (smcmax - smcmin);
  const double thko = 2.0;
  const double thks = pow(THKQTZ, qz) * pow(thko, 1.0 - qz);
  const double xunfroz = sh2o / smc;
  const double xu = xunfroz * smcmax;
  const double thksat = pow(thks, 1.0 - smcmax) * pow(THKICE, smcmax - xu) * pow(THKW, xu);
  const double gammd = (1.0 - smcmax) * 2700.0;
  const double thkdry = (0.135 * gammd + 64.7) / (2700.0 - 0.947 * gammd);
  double ake;
  if ((sh2o + 0.0005) < smc) ake = satratio;
  else if (satratio > 0.1) ake = log10(satratio) + 1.0;
  else ake = 0.0;
  return ake * (thksat - thkdry) + thkdry;
}

// Snowz0: src/noah/noah.c:2202-2220
__device__ double snow_z0(double sncovr, double z0brd, double snowh) {
  const double Z0S = 0.001;
  const double burial = 7.0 * z0brd - snowh;
  const double z0eff = (burial <= 0.0007) ? Z0S : (burial / 7.0);
  return (1.0 - sncovr) * z0brd + sncovr * z0eff;
}

// Paulson's stability functions: src/noah/noah.c:3204-3222
__device__ double psp_mu(double xx) {
  return -2.0 * log((xx + 1.0) * 0.5) - log((xx * xx + 1.0) * 0.5) + 2.0 * atan(xx) - NPI * 0.5;
}
__device__ double psp_hu(double xx) { return -2.0 * log((xx * xx + 1.0) * 0.5); }

// SfcDifOff: src/noah/noah.c:2862-3079 with iz0tlnd = 0 and Paulson's
// functions (ilech = 0; the Lech branches are never taken)
__device__ void sfc_dif_off(double zlvl, double zlvl_wind, double z0, double czil, double sfcspd, double t1v,
                            double th2v, double &cm, double &ch) {
  const double WWST = 1.2, VKRM = 0.40, EXCM = 0.001, BETA = 1.0 / 270.0;
  const double WOLD = 0.15;
  const int ITRMX = 5;
  const double EPSU2 = 1.0e-4, EPSUST = 0.07, ZTMIN = -5.0, ZTMAX = 1.0, HPBL = 1000.0, SQVISC = 258.2;
  const double wwst2 = WWST * WWST;
  const double btg = BETA * NGRAV;
  const double elfc = VKRM * btg;
  const double wnew = 1.0 - WOLD;
  const double zilfc = -czil * VKRM * SQVISC;
  const double zu = z0;
  const double rdz = 1.0 / zlvl_wind;
  const double cxch = EXCM * rdz;
  const double dthv = th2v - t1v;
  double du2 = sfcspd * sfcspd;
  du2 = (du2 > EPSU2) ? du2 : EPSU2;
  const double btgh = btg * HPBL;
  double wstar2 = (btgh * ch * dthv != 0.0) ? wwst2 * pow(fabs(btgh * ch * dthv), 2.0 / 3.0) : 0.0;
  double ustar = sqrt(cm * sqrt(du2 + wstar2));
  ustar = (ustar > EPSUST) ? ustar : EPSUST;
  double zt = exp(zilfc * sqrt(ustar * z0)) * z0;
  const double zslu = zlvl_wind + zu;
  double zslt = zlvl + zt;
  const double rlogu = log(zslu / zu);
  double rlogt = log(zslt / zt);
  double rlmo = elfc * ch * dthv / pow(ustar, 3.0);
  for (int itr = 0; itr < ITRMX; itr++) {
    double zetalt = zslt * rlmo;
    zetalt = zetalt > ZTMIN ? zetalt : ZTMIN;
    rlmo = zetalt / zslt;
    double zetalu = zslu * rlmo;
    const double zetau = zu * rlmo;
    const double zetat = zt * rlmo;
    double simm, simh;
    if (rlmo < 0.0) {
      const double xlu4 = 1.0 - 16.0 * zetalu;
      const double xlt4 = 1.0 - 16.0 * zetalt;
      const double xu4 = 1.0 - 16.0 * zetau;
      const double xt4 = 1.0 - 16.0 * zetat;
      const double xlu = sqrt(sqrt(xlu4));
      const double xlt = sqrt(sqrt(xlt4));
      const double xu = sqrt(sqrt(xu4));
      const double xt = sqrt(sqrt(xt4));
      const double psmz = psp_mu(xu);
      simm = psp_mu(xlu) - psmz + rlogu;
      const double pshz = psp_hu(xt);
      simh = psp_hu(xlt) - pshz + rlogt;
    } else {
      zetalu = (zetalu < ZTMAX) ? zetalu : ZTMAX;
      zetalt = (zetalt < ZTMAX) ? zetalt : ZTMAX;
      const double psmz = 5.0 * zetau;
      simm = 5.0 * zetalu - psmz + rlogu;
      const double pshz = 5.0 * zetat;
      simh = 5.0 * zetalt - pshz + rlogt;
    }
    ustar = sqrt(cm * sqrt(du2 + wstar2));
    ustar = (ustar > EPSUST) ? ustar : EPSUST;
    zt = exp(zilfc * sqrt(ustar * z0)) * z0;
    zslt = zlvl + zt;
    rlogt = log(zslt / zt);
    const double ustark = ustar * VKRM;
    cm = (ustark / simm > cxch) ? ustark / simm : cxch;
    ch = (ustark / simh > cxch) ? ustark / simh : cxch;
    wstar2 = (btgh * ch * dthv != 0.0) ? wwst2 * pow(fabs(btgh * ch * dthv), 2.0 / 3.0) : 0.0;
    const double rlmn = elfc * ch * dthv / pow(ustar, 3.0);
    const double rlma = rlmo * WOLD + rlmn * wnew;
    rlmo = rlma;
  }
}

// Penman: src/noah/noah.c:1508-1570
__device__ void penman(Col &c, double &t24, double t2v, int snowng, int frzgra) {
  const double ELCP = 2.4888e+3, LSUBC = 2.501000e+6;
  const double emissi = c.emissi;
  const double elcp1 = (1.0 - c.sncovr) * ELCP + c.sncovr * ELCP * LSUBS / LSUBC;
  const double lvs = (1.0 - c.sncovr) * LSUBC + c.sncovr * LSUBS;
  c.flx2 = 0.0;
  const double delta = elcp1 * c.dqsdt2;
  t24 = c.sfctmp * c.sfctmp * c.sfctmp * c.sfctmp;
  c.rr = emissi * t24 * 6.48e-8 / (c.sfcprs * c.ch) + 1.0;
  const double rho = c.sfcprs / (RD * t2v);
  c.rch = rho * NCP * c.ch;
  if (!snowng) {
    if (c.prcp > 0.0) c.rr += CPH2O * c.prcp * 1000.0 / c.rch;
  } else {
    c.rr += CPICE * c.prcp * 1000.0 / c.rch;
  }
  double fnet = c.fdown - emissi * SIGMA * t24 - c.ssoil;
  if (frzgra) {
    c.flx2 = -LSUBF * c.prcp * 1000.0;
    fnet -= c.flx2;
  }
  const double rad = fnet / c.rch + c.th2 - c.sfctmp;
  const double a = elcp1 * (c.q2sat - c.q2);
  c.epsca = (a * c.rr + rad * delta) / (delta + c.rr);
  c.etp = c.epsca * c.rch / lvs / 1000.0;
}

// CanRes: src/noah/noah.c:704-809
__device__ void can_res(Col &c) {
  const double SLV = 2.501000e6;
  c.rc = 0.0;
  const double ff = 0.55 * 2.0 * c.soldn / (c.rgl * c.proj_lai);
  double rcs = (ff + c.rsmin / c.rsmax) / (1.0 + ff);
  rcs = (rcs > 0.0001) ? rcs : 0.0001;
  double rct = 1.0 - 0.0016 * pow(c.topt - c.sfctmp, 2.0);
  rct = (rct > 0.0001) ? rct : 0.0001;
  double rcq = 1.0 / (1.0 + c.hs * (c.q2sat - c.q2));
  rcq = (rcq > 0.01) ? rcq : 0.01;
  double part[ML];
  double gx = (c.sh2o[0] - c.smcwlt) / (c.smcref - c.smcwlt);
  gx = (gx > 1.0) ? 1.0 : gx;
  gx = (gx < 0.0) ? 0.0 : gx;
  part[0] = (c.zsoil[0] / c.zsoil[c.nroot - 1]) * gx;
  for (int k = 1; k < c.nroot; k++) {
    gx = (c.sh2o[k] - c.smcwlt) / (c.smcref - c.smcwlt);
    gx = (gx > 1.0) ? 1.0 : gx;
    gx = (gx < 0.0) ? 0.0 : gx;
    part[k] = ((c.zsoil[k] - c.zsoil[k - 1]) / c.zsoil[c.nroot - 1]) * gx;
  }
  double rcsoil = 0.0;
  for (int k = 0; k < c.nroot; k++) rcsoil += part[k];
  rcsoil = (rcsoil > 0.0001) ? rcsoil : 0.0001;
  c.rc = c.rsmin / (c.proj_lai * rcs * rct * rcq * rcsoil);
  const double rr = (4.0 * c.emissi * SIGMA * RD / NCP) * pow(c.sfctmp, 4.0) / (c.sfcprs * c.ch) + 1.0;
  const double delta = (SLV / NCP) * c.dqsdt2;
  c.pc = (rr + delta) / (rr * (1.0 + c.rc * c.ch) + delta);
}

// DEvap: src/noah/noah.c:841-870
__device__ void d_evap(Col &c, const NoahUni &u) {
  const double sratio = (c.sh2o[0] - c.smcdry) / (c.smcmax - c.smcdry);
  double fx = 0.0;
  if (sratio > 0.0) {
    fx = pow(sratio, u.fxexp);
    fx = (fx > 1.0) ? 1.0 : fx;
    fx = (fx < 0.0) ? 0.0 : fx;
  }
  c.edir = fx * (1.0 - c.shdfac) * c.etp;
}

// Transp: src/noah/noah.c:2725-2811
__device__ void transp(Col &c) {
  for (int k = 0; k < c.nsoil; k++) c.et[k] = 0.0;
  double etpa;
  if (c.cmc != 0.0) etpa = c.shdfac * c.pc * c.etp * (1.0 - pow(c.cmc / c.cmcmax, c.cfactr));
  else etpa = c.shdfac * c.pc * c.etp;
  double gx[ML];
  double sgx = 0.0;
  for (int i = 0; i < c.nroot; i++) {
    gx[i] = (c.smc[i] - c.smcwlt) / (c.smcref - c.smcwlt);
    gx[i] = (gx[i] < 0.0) ? 0.0 : gx[i];
    gx[i] = (gx[i] > 1.0) ? 1.0 : gx[i];
    sgx += gx[i];
  }
  sgx /= (double)c.nroot;
  double denom = 0.0;
  for (int i = 0; i < c.nroot; i++) {
    const double rtx = c.rtdis[i] + gx[i] - sgx;
    gx[i] *= ((rtx > 0.0) ? rtx : 0.0);
    denom += gx[i];
  }
  denom = (denom <= 0.0) ? 1.0 : denom;
  for (int i = 0; i < c.nroot; i++) c.et[i] = etpa * gx[i] / denom;
}

// Evapo: src/noah/noah.c:872-959
__device__ void evapo(Col &c, const NoahUni &u, double dt) {
  c.edir = 0.0;
  c.ec = 0.0;
  c.ett = 0.0;
  for (int k = 0; k < c.nsoil; k++) c.et[k] = 0.0;
  if (c.etp > 0.0) {
    if (c.shdfac < 1.0) d_evap(c, u);
    if (c.shdfac > 0.0) {
      transp(c);
      for (int k = 0; k < c.nsoil; k++) c.ett += c.et[k];
      if (c.cmc > 0.0)
        c.ec = c.shdfac * pow((c.cmc / c.cmcmax > 1.0) ? 1.0 : c.cmc / c.cmcmax, c.cfactr) * c.etp;
      else
        c.ec = 0.0;
      const double cmc2ms = c.cmc / dt;
      c.ec = (cmc2ms < c.ec) ? cmc2ms : c.ec;
    }
  }
  c.etns = c.edir + c.ett + c.ec;
}

// PcpDrp: src/noah/noah.c:1451-1506
__device__ void pcp_drp(Col &c, double prcp, double dt) {
  const double KD = 6.54e-7, BFACTR = 3.89;
  double rhsct = c.shdfac * prcp - c.ec;
  c.drip = 0.0;
  const double trhsct = dt * rhsct;
  const double excess = c.cmc + trhsct;
  if (excess > 0.0) {
    if (excess >= c.cmcmax) c.drip = (KD * c.cmcmax * exp(BFACTR)) + (excess - c.cmcmax) / dt;
    else c.drip = (KD * c.cmcmax * exp(BFACTR * excess / c.cmcmax));
    c.drip = (c.drip > excess / dt) ? excess / dt : c.drip;
  }
  rhsct -= c.drip;
  c.pcpdrp = (1.0 - c.shdfac) * prcp + c.drip;
  c.cmc += dt * rhsct;
  if (c.cmc < 1.0e-20) c.cmc = 0.0;
  c.cmc = (c.cmc < c.cmcmax) ? c.cmc : c.cmcmax;
}

// TBnd: src/noah/noah.c:2482-2518
__device__ double t_bnd(double tu, double tb, const double *zsoil, double zbot, int k, int nsoil) {
  const double zup = (k == 0) ? 0.0 : zsoil[k - 1];
  const double zb = (k == nsoil - 1) ? 2.0 * zbot - zsoil[k] : zsoil[k + 1];
  return tu + (tb - tu) * (zup - zsoil[k]) / (zup - zb);
}

// TmpAvg: src/noah/noah.c:2621-2723
__device__ double tmp_avg(double tup, double tm, double tdn, const double *zsoil, int k) {
  const double dz = (k == 0) ? -zsoil[0] : zsoil[k - 1] - zsoil[k];
  const double dzh = dz * 0.5;
  double tavg;
  if (tup < TFREEZ) {
    if (tm < TFREEZ) {
      if (tdn < TFREEZ) {
        tavg = (tup + 2.0 * tm + tdn) / 4.0;
      } else {
        const double x0 = (TFREEZ - tm) * dzh / (tdn - tm);
        tavg = 0.5 * (tup * dzh + tm * (dzh + x0) + TFREEZ * (2.0 * dzh - x0)) / dz;
      }
    } else {
      if (tdn < TFREEZ) {
        const double xup = (TFREEZ - tup) * dzh / (tm - tup);
        const double xdn = dzh - (TFREEZ - tm) * dzh / (tdn - tm);
        tavg = 0.5 * (tup * xup + TFREEZ * (2.0 * dz - xup - xdn) + tdn * xdn) / dz;
      } else {
        const double xup = (TFREEZ - tup) * dzh / (tm - tup);
        tavg = 0.5 * (tup * xup + TFREEZ * (2.0 * dz - xup)) / dz;
      }
    }
  } else {
    if (tm < TFREEZ) {
      if (tdn < TFREEZ) {
        const double xup = dzh - (TFREEZ - tup) * dzh / (tm - tup);
        tavg = 0.5 * (TFREEZ * (dz - xup) + tm * (dzh + xup) + tdn * dzh) / dz;
      } else {
        const double xup = dzh - (TFREEZ - tup) * dzh / (tm - tup);
        const double xdn = (TFREEZ - tm) * dzh / (tdn - tm);
        tavg = 0.5 * (TFREEZ * (2.0 * dz - xup - xdn) + tm * (xup + xdn)) / dz;
      }
    } else {
      if (tdn < TFREEZ) {
        const double xdn = dzh - (TFREEZ - tm) * dzh / (tdn - tm);
        tavg = (TFREEZ * (dz - xdn) + 0.5 * (TFREEZ + tdn) * xdn) / dz;
      } else {
        tavg = (tup + 2.0 * tm + tdn) / 4.0;
      }
    }
  }
  return tavg;
}

// FrH2O: src/noah/noah.c:978-1090 (van Genuchten form)
__device__ double fr_h2o(double tkelv, double smc, double sh2o, double smcmax, double smcmin, double alpha,
                         double beta) {
  const double CK = 8.0, ERROR = 0.005;
  if (tkelv > (TFREEZ - 1.0e-3)) return smc;
  double swl = smc - sh2o;
  swl = (swl > smc - SH2OMIN) ? (smc - SH2OMIN) : swl;
  swl = (swl < 0.0) ? 0.0 : swl;
  int nlog = 0, kcount = 0;
  double mx = 0.0;
  while (nlog < 10 && kcount == 0) {
    nlog++;
    const double satn = (smc - swl - smcmin) / (smcmax - smcmin);
    mx = beta / (1.0 - beta);
    const double df = log(NGRAV / alpha / LSUBF) + 1.0 / beta * log(pow(satn, mx) - 1.0) +
                      2.0 * log(1.0 + CK * swl) - log(-(tkelv - TFREEZ) / tkelv);
    const double denom = 1.0 / (beta - 1.0) / (smcmax - smcmin) * pow(satn, mx - 1.0) / (pow(satn, mx) - 1.0) +
                         2.0 * CK / (1.0 + CK * swl);
    double swlk = swl - df / denom;
    swlk = (swlk > smc - SH2OMIN) ? (smc - SH2OMIN) : swlk;
    swlk = (swlk < 0.0) ? 0.0 : swlk;
    const double dswl = fabs(swlk - swl);
    swl = swlk;
    if (dswl <= ERROR) kcount++;
  }
  double freew = smc - swl;
  if (kcount == 0) {
    double fk = pow(pow(-(tkelv - TFREEZ) / tkelv * alpha * LSUBF / NGRAV, beta), 1.0 / mx) * (smcmax - smcmin) -
                smcmin;
    fk = (fk < SH2OMIN) ? SH2OMIN : fk;
    freew = (fk < smc) ? fk : smc;
  }
  return freew;
}

// SnkSrc: src/noah/noah.c:1746-1811 (updates sh2o[k])
__device__ double snk_src(Col &c, double tavg, int k, double dt, double qtot) {
  const double DH2O = 1.0000e3;
  const double dz = (k == 0) ? -c.zsoil[0] : c.zsoil[k - 1] - c.zsoil[k];
  const double sh2o = c.sh2o[k], smc = c.smc[k];
  const double freew = fr_h2o(tavg, smc, sh2o, c.smcmax, c.smcmin, c.alpha, c.beta);
  double xh2o = sh2o + qtot * dt / (DH2O * LSUBF * dz);
  if (xh2o < sh2o && xh2o < freew) xh2o = (freew > sh2o) ? sh2o : freew;
  if (xh2o > sh2o && xh2o > freew) xh2o = (freew < sh2o) ? sh2o : freew;
  xh2o = (xh2o < 0.0) ? 0.0 : xh2o;
  xh2o = (xh2o > smc) ? smc : xh2o;
  const double tsnsr = -DH2O * LSUBF * dz * (xh2o - sh2o) / dt;
  c.sh2o[k] = xh2o;
  return tsnsr;
}

// ShFlx: src/noah/noah.c:1621-1652 with HRT (:1092-1299, itavg on), HStep
// (:1301-1340) and Rosr12 (:1572-1620) inline
__device__ void sh_flx(Col &c, const NoahUni &u, double dt, double yy, double zz1, double df1) {
  const double CH2O = 4.2e6;
  const int nsoil = c.nsoil;
  const double csoil_loc = u.csoil;
  double ai[ML], bi[ML], ci[ML], rhsts[ML];
  // HRT: top layer
  double hcpct = c.sh2o[0] * CH2O + (1.0 - c.smcmax) * csoil_loc + (c.smcmax - c.smc[0]) * NCP +
                 (c.smc[0] - c.sh2o[0]) * CPICE;
  double ddz = 1.0 / (-0.5 * c.zsoil[1]);
  ai[0] = 0.0;
  ci[0] = (df1 * ddz) / (c.zsoil[0] * hcpct);
  bi[0] = -ci[0] + df1 / (0.5 * c.zsoil[0] * c.zsoil[0] * hcpct * zz1);
  double dtsdz = (c.stc[0] - c.stc[1]) / (-0.5 * c.zsoil[1]);
  c.ssoil = df1 * (c.stc[0] - yy) / (0.5 * c.zsoil[0] * zz1);
  double denom = c.zsoil[0] * hcpct;
  rhsts[0] = (df1 * dtsdz - c.ssoil) / denom;
  double qtot = -1.0 * rhsts[0] * denom;
  double sice = c.smc[0] - c.sh2o[0];
  const double tsurf = (yy + (zz1 - 1.0) * c.stc[0]) / zz1;
  double tbk = t_bnd(c.stc[0], c.stc[1], c.zsoil, u.zbot, 0, nsoil);
  if ((sice > 0.0) || (c.stc[0] < TFREEZ) || (tsurf < TFREEZ) || (tbk < TFREEZ)) {
    const double tavg = tmp_avg(tsurf, c.stc[0], tbk, c.zsoil, 0);
    const double tsnsr = snk_src(c, tavg, 0, dt, qtot);
    rhsts[0] -= tsnsr / denom;
  }
  // HRT: the layers below
  double ddz2 = 0.0, df1k = df1, dtsdz2 = 0.0, tbk1 = 0.0;
  for (int k = 1; k < nsoil; k++) {
    hcpct = c.sh2o[k] * CH2O + (1.0 - c.smcmax) * csoil_loc + (c.smcmax - c.smc[k]) * NCP +
            (c.smc[k] - c.sh2o[k]) * CPICE;
    double df1n;
    if (k != nsoil - 1) {
      df1n = tdf_cnd(c.smc[k], c.quartz, c.smcmax, c.smcmin, c.sh2o[k]);
      denom = 0.5 * (c.zsoil[k - 1] - c.zsoil[k + 1]);
      dtsdz2 = (c.stc[k] - c.stc[k + 1]) / denom;
      ddz2 = 2.0 / (c.zsoil[k - 1] - c.zsoil[k + 1]);
      ci[k] = -df1n * ddz2 / ((c.zsoil[k - 1] - c.zsoil[k]) * hcpct);
      tbk1 = t_bnd(c.stc[k], c.stc[k + 1], c.zsoil, u.zbot, k, nsoil);
    } else {
      df1n = tdf_cnd(c.smc[k], c.quartz, c.smcmax, c.smcmin, c.sh2o[k]);
      denom = 0.5 * (c.zsoil[k - 1] + c.zsoil[k]) - u.zbot;
      dtsdz2 = (c.stc[k] - u.tbot) / denom;
      ci[k] = 0.0;
      tbk1 = t_bnd(c.stc[k], u.tbot, c.zsoil, u.zbot, k, nsoil);
    }
    denom = (c.zsoil[k] - c.zsoil[k - 1]) * hcpct;
    rhsts[k] = (df1n * dtsdz2 - df1k * dtsdz) / denom;
    qtot = -1.0 * denom * rhsts[k];
    sice = c.smc[k] - c.sh2o[k];
    const double tavg = tmp_avg(tbk, c.stc[k], tbk1, c.zsoil, k);
    if ((sice > 0.0) || (c.stc[k] < TFREEZ) || (tbk < TFREEZ) || (tbk1 < TFREEZ)) {
      const double tsnsr = snk_src(c, tavg, k, dt, qtot);
      rhsts[k] = rhsts[k] - tsnsr / denom;
    }
    ai[k] = -df1k * ddz / ((c.zsoil[k - 1] - c.zsoil[k]) * hcpct);
    bi[k] = -(ai[k] + ci[k]);
    tbk = tbk1;
    df1k = df1n;
    dtsdz = dtsdz2;
    ddz = ddz2;
  }
  // HStep
  double rin[ML], cin[ML];
  for (int k = 0; k < nsoil; k++) {
    rhsts[k] *= dt;
    ai[k] *= dt;
    bi[k] = 1.0 + bi[k] * dt;
    ci[k] *= dt;
  }
  for (int k = 0; k < nsoil; k++) rin[k] = rhsts[k];
  for (int k = 0; k < nsoil; k++) cin[k] = ci[k];
  // Rosr12 (p = ci, a = ai, b = bi, c = cin, d = rin, delta = rhsts)
  {
    double *p = ci, *delta = rhsts;
    cin[nsoil - 1] = 0.0;
    p[0] = -cin[0] / bi[0];
    delta[0] = rin[0] / bi[0];
    for (int k = 1; k < nsoil; k++) {
      p[k] = -cin[k] * (1.0 / (bi[k] + ai[k] * p[k - 1]));
      delta[k] = (rin[k] - ai[k] * delta[k - 1]) * (1.0 / (bi[k] + ai[k] * p[k - 1]));
    }
    p[nsoil - 1] = delta[nsoil - 1];
    for (int k = nsoil - 2; k >= 0; k--) p[k] = p[k] * p[k + 1] + delta[k];
  }
  for (int k = 0; k < nsoil; k++) c.stc[k] += ci[k];
  c.t1 = (yy + (zz1 - 1.0) * c.stc[0]) / zz1;
  c.ssoil = df1 * (c.stc[0] - c.t1) / (0.5 * c.zsoil[0]);
}

// NoPac: src/noah/noah.c:1343-1449
__device__ void no_pac(Col &c, const NoahUni &u, double dt, double t24) {
  double prcpf = c.prcp;
  c.dew = 0.0;
  c.edir = 0.0;
  c.ec = 0.0;
  for (int k = 0; k < c.nsoil; k++) c.et[k] = 0.0;
  c.ett = 0.0;
  if (c.etp > 0.0) {
    evapo(c, u, dt);
    c.eta = c.etns;
  } else {
    c.dew = -c.etp;
    prcpf += c.dew;
  }
  pcp_drp(c, prcpf, dt);
  if (c.etp <= 0.0) {
    c.beta_ = 0.0;
    c.eta = c.etp;
    if (c.etp < 0.0) c.beta_ = 1.0;
  } else {
    c.beta_ = c.eta / c.etp;
  }
  double df1 = tdf_cnd(c.smc[0], c.quartz, c.smcmax, c.smcmin, c.sh2o[0]);
  df1 *= exp(u.sbeta * c.shdfac);
  const double yynum = c.fdown - c.emissi * SIGMA * t24;
  const double yy = c.sfctmp + (yynum / c.rch + c.th2 - c.sfctmp - c.beta_ * c.epsca) / c.rr;
  const double zz1 = df1 / (-0.5 * c.zsoil[0] * c.rch * c.rr) + 1.0;
  sh_flx(c, u, dt, yy, zz1, df1);
  c.flx1 = CPH2O * c.prcp * 1000.0 * (c.t1 - c.sfctmp);
  c.flx3 = 0.0;
}

// SnowPack: src/noah/noah.c:2100-2200
__device__ void snow_pack(double esd, double dtsec, double &snowh, double &sndens, double tsnow, double tsoil) {
  const double C1 = 0.01, C2 = 21.0;
  const double esdc = esd * 100.0;
  const double dthr = dtsec / 3600.0;
  const double tsnowc = tsnow - 273.15;
  const double tsoilc = tsoil - 273.15;
  const double tavgc = 0.5 * (tsnowc + tsoilc);
  const double esdcx = (esdc > 1.0e-2) ? esdc : 1.0e-2;
  const double bfac = dthr * C1 * exp(0.08 * tavgc - C2 * sndens);
  double pexp = 0.0;
  for (int j = 4; j > 0; j--) pexp = (1.0 + pexp) * bfac * esdcx / (double)(j + 1);
  pexp += 1.0;
  double dsx = sndens * pexp;
  dsx = (dsx > 0.40) ? 0.40 : dsx;
  dsx = (dsx < 0.05) ? 0.05 : dsx;
  sndens = dsx;
  if (tsnowc >= 0.0) {
    const double dw = 0.13 * dthr / 24.0;
    sndens = sndens * (1.0 - dw) + dw;
    if (sndens >= 0.40) sndens = 0.40;
  }
  const double snowhc = esdc / sndens;
  snowh = snowhc * 0.01;
}

// SnoPac: src/noah/noah.c:1813-2098
__device__ void sno_pac(Col &c, const NoahUni &u, int snowng, double dt, double t24, double prcpf, double df1) {
  const double ESDMIN = 1.0e-6, SNOEXP = 2.0;
  c.dew = 0.0;
  c.edir = 0.0;
  c.ec = 0.0;
  for (int k = 0; k < c.nsoil; k++) c.et[k] = 0.0;
  c.ett = 0.0;
  c.esnow = 0.0;
  double esnow2 = 0.0, etanrg;
  c.beta_ = 1.0;
  if (c.etp <= 0.0) {
    // the bulk Richardson number is 0: no frost adjustment (src/noah/noah.c:1870-1876)
    if (c.etp == 0.0) c.beta_ = 0.0;
    c.dew = -c.etp;
    esnow2 = c.etp * dt;
    etanrg = c.etp * 1000.0 * ((1.0 - c.sncovr) * LVH2O + c.sncovr * LSUBS);
  } else {
    if (c.sncovr < 1.0) {
      evapo(c, u, dt);
      c.edir *= (1.0 - c.sncovr);
      c.ec *= (1.0 - c.sncovr);
      for (int k = 0; k < c.nsoil; k++) c.et[k] = c.et[k] * (1.0 - c.sncovr);
      c.ett *= (1.0 - c.sncovr);
      c.etns *= (1.0 - c.sncovr);
    }
    c.esnow = c.etp * c.sncovr;
    esnow2 = c.esnow * dt;
    etanrg = c.esnow * 1000.0 * LSUBS + c.etns * 1000.0 * LVH2O;
  }
  c.flx1 = 0.0;
  if (snowng) c.flx1 = CPICE * c.prcp * 1000.0 * (c.t1 - c.sfctmp);
  else if (c.prcp > 0.0) c.flx1 = CPH2O * c.prcp * 1000.0 * (c.t1 - c.sfctmp);
  const double dsoil = -0.5 * c.zsoil[0];
  const double dtot = c.snowh + dsoil;
  const double denom = 1.0 + df1 / (dtot * c.rr * c.rch);
  const double t12a = ((c.fdown - c.flx1 - c.flx2 - c.emissi * SIGMA * t24) / c.rch + c.th2 - c.sfctmp -
                       etanrg / c.rch) / c.rr;
  const double t12b = df1 * c.stc[0] / (dtot * c.rr * c.rch);
  const double t12 = (c.sfctmp + t12a + t12b) / denom;
  double ex;
  if (t12 <= TFREEZ) {
    c.t1 = t12;
    c.ssoil = df1 * (c.t1 - c.stc[0]) / dtot;
    if (c.sneqv - esnow2 > 0.0) {
      c.sneqv -= esnow2;
    } else {
      c.sneqv = 0.0;
      esnow2 = c.sneqv;
      c.esnow = esnow2 / dt;
    }
    c.flx3 = 0.0;
    ex = 0.0;
    c.snomlt = 0.0;
  } else {
    c.t1 = TFREEZ * pow(c.sncovr, SNOEXP) + t12 * (1.0 - pow(c.sncovr, SNOEXP));
    c.beta_ = 1.0;
    c.ssoil = df1 * (c.t1 - c.stc[0]) / dtot;
    if (c.sneqv - esnow2 <= ESDMIN) {
      c.sneqv = 0.0;
      ex = 0.0;
      c.snomlt = 0.0;
      c.flx3 = 0.0;
    } else {
      c.sneqv -= esnow2;
      const double seh = c.rch * (c.t1 - c.th2);
      double t14 = c.t1 * c.t1;
      t14 = t14 * t14;
      c.flx3 = c.fdown - c.flx1 - c.flx2 - c.emissi * SIGMA * t14 - c.ssoil - seh - etanrg;
      c.flx3 = (c.flx3 <= 0.0) ? 0.0 : c.flx3;
      ex = c.flx3 * 0.001 / LSUBF;
      c.snomlt = ex;
      if (c.sneqv - c.snomlt * dt >= ESDMIN) {
        c.sneqv -= c.snomlt * dt;
      } else {
        ex = c.sneqv / dt;
        c.flx3 = ex * 1000.0 * LSUBF;
        c.snomlt = c.sneqv / dt;
        c.sneqv = 0.0;
      }
    }
    prcpf += ex;
  }
  pcp_drp(c, prcpf, dt);
  const double zz1 = 1.0;
  const double yy = c.stc[0] - 0.5 * c.ssoil * c.zsoil[0] * zz1 / df1;
  const double t11 = c.t1;
  const double ssoil1 = c.ssoil;
  sh_flx(c, u, dt, yy, zz1, df1);
  c.t1 = t11;
  c.ssoil = ssoil1;
  if (c.sneqv > 0.0) {
    snow_pack(c.sneqv, dt, c.snowh, c.sndens, c.t1, yy);
  } else {
    c.sneqv = 0.0;
    c.snowh = 0.0;
    c.sndens = 0.0;
    c.sncovr = 0.0;
  }
}

// SFlx: src/noah/noah.c:166-609
__device__ void s_flx(Col &c, const NoahUni &u, double dt) {
  const double shdfac = c.shdfac;
  double t24 = 0.0;
  c.snomlt = 0.0;
  c.pcpdrp = 0.0;
  if (shdfac > 0.0) c.proj_lai = (c.proj_lai > 0.5) ? c.proj_lai : 0.5;
  c.cmcmax = shdfac * c.cmcfactr * c.proj_lai;
  if (c.proj_lai >= c.laimax) {
    c.embrd = c.emissmax;
    c.alb = c.albedomin;
    c.z0brd = c.z0max;
  } else if (c.proj_lai <= c.laimin) {
    c.embrd = c.emissmin;
    c.alb = c.albedomax;
    c.z0brd = c.z0min;
  } else if (c.laimax > c.laimin) {
    double f = (c.proj_lai - c.laimin) / (c.laimax - c.laimin);
    f = (f < 1.0) ? f : 1.0;
    f = (f > 0.0) ? f : 0.0;
    c.embrd = ((1.0 - f) * c.emissmin) + (f * c.emissmax);
    c.alb = ((1.0 - f) * c.albedomax) + (f * c.albedomin);
    c.z0brd = ((1.0 - f) * c.z0min) + (f * c.z0max);
  } else {
    c.embrd = 0.5 * c.emissmin + 0.5 * c.emissmax;
    c.alb = 0.5 * c.albedomin + 0.5 * c.albedomax;
    c.z0brd = 0.5 * c.z0min + 0.5 * c.z0max;
  }
  int snowng = 0, frzgra = 0;
  if (c.sneqv <= 1.0e-7) {
    c.sneqv = 0.0;
    c.sndens = 0.0;
    c.snowh = 0.0;
    c.sncond = 1.0;
  } else {
    c.sndens = c.sneqv / c.snowh;
    c.sncond = c_snow(c.sndens);
  }
  if (c.prcp > 0.0) {
    if (c.ffrozp > 0.5) snowng = 1;
    else if (c.t1 <= TFREEZ) frzgra = 1;
  }
  double prcpf;
  if (snowng || frzgra) {
    const double sn_new = c.prcp * dt;
    c.sneqv += sn_new;
    prcpf = 0.0;
    snow_new(c, sn_new);
    c.sncond = c_snow(c.sndens);
  } else {
    prcpf = c.prcp;
  }
  if (c.sneqv == 0.0) {
    c.sncovr = 0.0;
    c.albedo = c.alb;
    c.emissi = c.embrd;
  } else {
    c.sncovr = sn_frac(c.sneqv, c.snup, u.salp);
    c.sncovr = (c.sncovr < 0.98) ? c.sncovr : 0.98;
    al_calc(c, u, dt, snowng);
  }
  double df1 = tdf_cnd(c.smc[0], c.quartz, c.smcmax, c.smcmin, c.sh2o[0]);
  df1 *= exp(u.sbeta * shdfac);
  if (c.sncovr > 0.97) df1 = c.sncond;
  const double dsoil = -(0.5 * c.zsoil[0]);
  if (c.sneqv == 0.0) {
    c.ssoil = df1 * (c.t1 - c.stc[0]) / dsoil;
  } else {
    const double dtot = c.snowh + dsoil;
    const double frcsno = c.snowh / dtot;
    const double frcsoi = dsoil / dtot;
    const double df1a = frcsno * c.sncond + frcsoi * df1;
    df1 = df1a * c.sncovr + df1 * (1.0 - c.sncovr);
    c.ssoil = df1 * (c.t1 - c.stc[0]) / dtot;
  }
  c.z0 = (c.sncovr > 0.0) ? snow_z0(c.sncovr, c.z0brd, c.snowh) : c.z0brd;
  const double t1v = c.t1 * (1.0 + 0.61 * c.q2);
  const double th2v = c.th2 * (1.0 + 0.61 * c.q2);
  sfc_dif_off(u.zlvl, c.zlvl_wind, c.z0, u.czil, c.sfcspd, t1v, th2v, c.cm, c.ch);
  c.fdown = c.solnet + c.lwdn;
  const double t2v = c.sfctmp * (1.0 + 0.61 * c.q2);
  penman(c, t24, t2v, snowng, frzgra);
  if (shdfac > 0.0) can_res(c);
  else c.rc = 0.0;
  c.esnow = 0.0;
  if (c.sneqv == 0.0) {
    no_pac(c, u, dt, t24);
    c.eta_kinematic = c.eta * 1000.0;
  } else {
    sno_pac(c, u, snowng, dt, t24, prcpf, df1);
    c.eta_kinematic = (c.esnow + c.etns) * 1000.0;
  }
  c.q1 = c.q2 + c.eta_kinematic * NCP / c.rch;
  c.sheat = -(c.ch * NCP * c.sfcprs) / (RD * t2v) * (c.th2 - c.t1);
  c.ef_edir = c.edir * 1000.0 * LVH2O;
  c.ef_ec = c.ec * 1000.0 * LVH2O;
  c.ef_ett = c.ett * 1000.0 * LVH2O;
  c.ef_esnow = c.esnow * 1000.0 * LSUBS;
  c.ef_etp = c.etp * 1000.0 * ((1.0 - c.sncovr) * LVH2O + c.sncovr * LSUBS);
  if (c.ef_etp > 0.0) c.ef_eta = c.ef_edir + c.ef_ec + c.ef_ett + c.ef_esnow;
  else c.ef_eta = c.ef_etp;
  c.beta_ = (c.ef_etp == 0.0) ? 0.0 : (c.ef_eta / c.ef_etp);
  c.ssoil *= -1.0;
}

__global__ __launch_bounds__(kBlock) void k_noah_meteo(int nstation, int nrec, const double *ftime, const double *data,
                                                       double t, double *met) {
  const int k = blockIdx.x * kBlock + threadIdx.x;
  if (k >= nstation) return;
  ts_interp<PIHM_NOAH_MET_NVAR>(ftime + (long)k * nrec, data + (long)k * nrec * PIHM_NOAH_MET_NVAR, nrec, t,
                                met + k, nstation);
}

// Noah(): src/noah/noah.c:3-55, one element per thread.
__global__ __launch_bounds__(kBlock) void k_noah(int ne, int nstation, NoahDev d, NoahUni u, ColDev s, double dt) {
  const int p = blockIdx.x * kBlock + threadIdx.x;
  if (p >= ne) return;
  auto at = [ne, p](double *f, int k) -> double & { return f[(long)k * ne + p]; };
  Col c;
  c.nsoil = s.nsoil[p];
  c.nroot = d.nroot[p];
  for (int k = 0; k < ML; k++) {
    c.smc[k] = 0.0;
    c.sh2o[k] = 0.0;
    c.stc[k] = 0.0;
    c.et[k] = 0.0;
    c.zsoil[k] = 0.0;
    c.rtdis[k] = 0.0;
  }
  // DefSldpth's depths and RootDist (src/noah/lsm_func.c:86-205)
  c.zsoil[0] = -at(s.sldpth, 0);
  for (int k = 1; k < c.nsoil; k++) c.zsoil[k] = -at(s.sldpth, k) + c.zsoil[k - 1];
  for (int j = 0; j < c.nroot; j++) c.rtdis[j] = -at(s.sldpth, j) / c.zsoil[c.nroot - 1];
  for (int k = 0; k < c.nsoil; k++) {
    c.smc[k] = at(s.smc, k);
    c.sh2o[k] = at(s.sh2o, k);
    c.stc[k] = at(d.stc, k);
  }
  c.shdfac = d.shdfac[p]; c.cmcfactr = d.cmcfactr[p]; c.laimin = d.laimin[p]; c.laimax = d.laimax[p];
  c.emissmin = d.emissmin[p]; c.emissmax = d.emissmax[p]; c.albedomin = d.albedomin[p];
  c.albedomax = d.albedomax[p]; c.z0min = d.z0min[p]; c.z0max = d.z0max[p]; c.snup = d.snup[p];
  c.cfactr = d.cfactr[p]; c.rgl = d.rgl[p]; c.rsmin = d.rsmin[p]; c.rsmax = d.rsmax[p]; c.topt = d.topt[p];
  c.hs = d.hs[p]; c.smcref = d.smcref[p]; c.smcwlt = d.smcwlt[p]; c.smcdry = d.smcdry[p];
  c.quartz = d.quartz[p]; c.zlvl_wind = d.zlvl_wind[p];
  c.smcmax = s.smcmax[p]; c.smcmin = s.smcmin[p]; c.alpha = s.alpha[p]; c.beta = s.beta[p];
  c.t1 = at(d.sc, 0); c.sneqv = at(d.sc, 1); c.snowh = at(d.sc, 2); c.sndens = at(d.sc, 3);
  c.cmc = at(d.sc, 4); c.ch = at(d.sc, 5); c.cm = at(d.sc, 6); c.snotime1 = at(d.sc, 7);
  c.albedo = at(d.sc, 8); c.emissi = at(d.sc, 9); c.q1 = d.q1[p];
  c.proj_lai = d.lai[p];
  // ApplyMeteoForc with _NOAH_ (src/forcing.c:340-353)
  const int st = d.station[p];
  const double *m = d.met + st;
  c.prcp = m[0] / 1000.0;
  c.sfctmp = m[(long)1 * nstation];
  c.rh = m[(long)2 * nstation];
  c.sfcspd = m[(long)3 * nstation];
  c.soldn = m[(long)4 * nstation];
  c.soldn = (c.soldn > 0.0) ? c.soldn : 0.0;
  c.longwave = m[(long)5 * nstation];
  c.sfcprs = m[(long)6 * nstation];
  c.ssoil = 0.0;

  cal_hum(c, u);
  c.ffrozp = (c.prcp > 0.0 && c.sfctmp < TFREEZ) ? 1.0 : 0.0;   // FrozRain (:961-976)
  c.alb = BADVAL;
  if (c.q1 == BADVAL) c.q1 = c.q2;
  c.solnet = c.soldn * (1.0 - c.albedo);
  c.lwdn = c.longwave * c.emissi;
  for (int k = 0; k < c.nsoil; k++) {
    c.smc[k] = (c.smc[k] > c.smcmin + SH2OMIN) ? c.smc[k] : c.smcmin + SH2OMIN;
    c.smc[k] = (c.smc[k] < c.smcmax) ? c.smc[k] : c.smcmax;
    c.sh2o[k] = (c.sh2o[k] < c.smc[k]) ? c.sh2o[k] : c.smc[k];
  }
  s_flx(c, u, dt);

  // ET from W m-2 back to m s-1 (src/noah/noah.c:50-53)
  const double ec = c.ef_ec / LVH2O / 1000.0;
  const double ett = c.ef_ett / LVH2O / 1000.0;
  const double edir = c.ef_edir / LVH2O / 1000.0;
  for (int k = 0; k < c.nsoil; k++) {
    at(s.smc, k) = c.smc[k];
    at(s.sh2o, k) = c.sh2o[k];
    at(d.stc, k) = c.stc[k];
  }
  at(d.sc, 0) = c.t1; at(d.sc, 1) = c.sneqv; at(d.sc, 2) = c.snowh; at(d.sc, 3) = c.sndens;
  at(d.sc, 4) = c.cmc; at(d.sc, 5) = c.ch; at(d.sc, 6) = c.cm; at(d.sc, 7) = c.snotime1;
  at(d.sc, 8) = c.albedo; at(d.sc, 9) = c.emissi; d.q1[p] = c.q1;
  const double fx[NFX] = {edir, ett, ec, c.pcpdrp, c.drip, c.esnow, c.etp, c.snomlt, c.ssoil, c.sheat, c.ef_eta,
                          c.fdown, c.sncovr, c.beta_, c.rc, c.pc};
  for (int k = 0; k < NFX; k++) at(d.flux, k) = fx[k];
  // hand-over to the RHS: step inputs, layer transpiration (every layer: the
  // soil-water step reads all nsoil), the top layer's liquid water
  s.pcpdrp[p] = c.pcpdrp;
  s.edir[p] = edir;
  s.ett[p] = ett;
  for (int k = 0; k < ML; k++) at(s.et, k) = (k < c.nsoil) ? c.et[k] : 0.0;
  s.sh2o0[p] = c.sh2o[0];
}

}  // namespace

struct pihm_noah_lsm {
  pihm_noah_soil *soil;
  pihm_ctx *ctx;
  int ne, nstation, nrec;
  NoahDev d;
  NoahUni u;
  ColDev s;
  double *ftime, *data;      // station records on the device
  std::vector<double> ftime_host;
  void *arena;
};

namespace {

int upload_caller(pihm_noah_lsm *n, double *dst, const double *src, int per) {
  const int ne = n->ne;
  std::vector<double> h((size_t)per * ne);
  for (int p = 0; p < ne; p++) {
    const long ci = n->ctx->perm[p];
    for (int k = 0; k < per; k++) h[(size_t)k * ne + p] = src[ci * per + k];
  }
  PIHM_HIP_TRY(hipMemcpyAsync(dst, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, n->ctx->stream));
  PIHM_HIP_TRY(hipStreamSynchronize(n->ctx->stream));
  return PIHM_OK;
}

int upload_caller_i(pihm_noah_lsm *n, int32_t *dst, const int32_t *src) {
  const int ne = n->ne;
  std::vector<int32_t> h(ne);
  for (int p = 0; p < ne; p++) h[p] = src[n->ctx->perm[p]];
  PIHM_HIP_TRY(hipMemcpyAsync(dst, h.data(), ne * sizeof(int32_t), hipMemcpyHostToDevice, n->ctx->stream));
  PIHM_HIP_TRY(hipStreamSynchronize(n->ctx->stream));
  return PIHM_OK;
}

int download_caller(pihm_noah_lsm *n, const double *src, double *dst, int per) {
  const int ne = n->ne;
  std::vector<double> h((size_t)per * ne);
  PIHM_HIP_TRY(hipMemcpyAsync(h.data(), src, h.size() * sizeof(double), hipMemcpyDeviceToHost, n->ctx->stream));
  PIHM_HIP_TRY(hipStreamSynchronize(n->ctx->stream));
  for (int p = 0; p < ne; p++) {
    const long ci = n->ctx->perm[p];
    for (int k = 0; k < per; k++) dst[ci * per + k] = h[(size_t)k * ne + p];
  }
  return PIHM_OK;
}

}  // namespace

extern "C" {

int pihm_noah_lsm_create(pihm_noah_soil *soil, const pihm_noah_lsm_desc *desc, const pihm_noah_meteo *met,
                         pihm_noah_lsm **out) {
  if (!soil || !desc || !met || !out) return PIHM_ERR_ARG;
  const NoahSoilView v = noah_soil_view(soil);
  pihm_ctx *c = v.ctx;
  if (!c || !c->noah || c->nmem != 1) return PIHM_ERR_ARG;
  const int ne = v.ne;
  if (met->nstation < 1 || met->nrec < 2 || !met->ftime || !met->data) return PIHM_ERR_ARG;
  const double *arrs[] = {desc->shdfac, desc->cmcfactr, desc->laimin, desc->laimax, desc->emissmin,
                          desc->emissmax, desc->albedomin, desc->albedomax, desc->z0min, desc->z0max,
                          desc->snup, desc->cfactr, desc->rgl, desc->rsmin, desc->rsmax, desc->topt,
                          desc->hs, desc->smcref, desc->smcwlt, desc->smcdry, desc->quartz, desc->zlvl_wind};
  constexpr int NARR = sizeof(arrs) / sizeof(arrs[0]);
  for (const double *a : arrs)
    if (!a) return PIHM_ERR_ARG;
  if (!desc->nroot || !desc->station) return PIHM_ERR_ARG;
  // the column's layer counts bound the root layer (checked before any upload)
  std::vector<int32_t> nsoil(ne);
  PIHM_HIP_TRY(hipMemcpy(nsoil.data(), v.nsoil, ne * sizeof(int32_t), hipMemcpyDeviceToHost));
  for (int p = 0; p < ne; p++) {
    const long ci = c->perm[p];
    if (desc->nroot[ci] < 1 || desc->nroot[ci] > nsoil[p]) return PIHM_ERR_ARG;
    if (desc->station[ci] < 0 || desc->station[ci] >= met->nstation) return PIHM_ERR_ARG;
  }
  for (int k = 0; k < met->nstation; k++)
    for (int r = 1; r < met->nrec; r++)
      if (!(met->ftime[(long)k * met->nrec + r] > met->ftime[(long)k * met->nrec + r - 1])) return PIHM_ERR_ARG;
  pihm_noah_lsm *n = new pihm_noah_lsm();
  n->soil = soil;
  n->ctx = c;
  n->ne = ne;
  n->nstation = met->nstation;
  n->nrec = met->nrec;
  n->ftime_host.assign(met->ftime, met->ftime + (long)met->nstation * met->nrec);
  // one allocation: NARR + lai + q1 doubles, stc, sc, flux, 2 int arrays, met, records
  const size_t nd = (size_t)(NARR + 2) * ne + (size_t)ML * ne + (size_t)NSC * ne + (size_t)NFX * ne +
                    (size_t)PIHM_NOAH_MET_NVAR * met->nstation + (size_t)met->nstation * met->nrec +
                    (size_t)met->nstation * met->nrec * PIHM_NOAH_MET_NVAR;
  const size_t bytes = nd * sizeof(double) + 2 * (size_t)ne * sizeof(int32_t) + 256;
  if (hipMalloc(&n->arena, bytes) != hipSuccess) {
    delete n;
    return PIHM_ERR_HIP;
  }
  double *q = (double *)n->arena;
  double **dst[] = {&n->d.shdfac, &n->d.cmcfactr, &n->d.laimin, &n->d.laimax, &n->d.emissmin, &n->d.emissmax,
                    &n->d.albedomin, &n->d.albedomax, &n->d.z0min, &n->d.z0max, &n->d.snup, &n->d.cfactr,
                    &n->d.rgl, &n->d.rsmin, &n->d.rsmax, &n->d.topt, &n->d.hs, &n->d.smcref, &n->d.smcwlt,
                    &n->d.smcdry, &n->d.quartz, &n->d.zlvl_wind};
  for (int k = 0; k < NARR; k++) { *dst[k] = q; q += ne; }
  n->d.lai = q; q += ne;
  n->d.q1 = q; q += ne;
  n->d.stc = q; q += (size_t)ML * ne;
  n->d.sc = q; q += (size_t)NSC * ne;
  n->d.flux = q; q += (size_t)NFX * ne;
  n->d.met = q; q += (size_t)PIHM_NOAH_MET_NVAR * met->nstation;
  n->ftime = q; q += (size_t)met->nstation * met->nrec;
  n->data = q; q += (size_t)met->nstation * met->nrec * PIHM_NOAH_MET_NVAR;
  n->d.nroot = (int32_t *)q;
  n->d.station = n->d.nroot + ne;
  int rc = PIHM_OK;
  for (int k = 0; k < NARR && !rc; k++) rc = upload_caller(n, *dst[k], arrs[k], 1);
  if (!rc) rc = upload_caller_i(n, n->d.nroot, desc->nroot);
  if (!rc) rc = upload_caller_i(n, n->d.station, desc->station);
  if (!rc && hipMemcpy(n->ftime, met->ftime, (size_t)met->nstation * met->nrec * sizeof(double),
                       hipMemcpyHostToDevice) != hipSuccess) rc = PIHM_ERR_HIP;
  if (!rc && hipMemcpy(n->data, met->data, (size_t)met->nstation * met->nrec * PIHM_NOAH_MET_NVAR * sizeof(double),
                       hipMemcpyHostToDevice) != hipSuccess) rc = PIHM_ERR_HIP;
  // InitVar's values (src/initialize.c:660-700): no snow or canopy water,
  // albedo 0.18, emissivity 0.96, exchange coefficients 1e-4, q1 unset
  if (!rc) {
    std::vector<double> h((size_t)NSC * ne, 0.0);
    for (int p = 0; p < ne; p++) {
      h[(size_t)0 * ne + p] = TFREEZ + 10.0;
      h[(size_t)5 * ne + p] = 1.0e-4;
      h[(size_t)6 * ne + p] = 1.0e-4;
      h[(size_t)8 * ne + p] = 0.18;
      h[(size_t)9 * ne + p] = 0.96;
    }
    if (hipMemcpy(n->d.sc, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) rc = PIHM_ERR_HIP;
    std::vector<double> qb(ne, BADVAL), lai(ne, 1.0);
    if (!rc && hipMemcpy(n->d.q1, qb.data(), ne * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
      rc = PIHM_ERR_HIP;
    if (!rc && hipMemcpy(n->d.lai, lai.data(), ne * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
      rc = PIHM_ERR_HIP;
    if (!rc && hipMemset(n->d.stc, 0, (size_t)ML * ne * sizeof(double)) != hipSuccess) rc = PIHM_ERR_HIP;
    if (!rc && hipMemset(n->d.flux, 0, (size_t)NFX * ne * sizeof(double)) != hipSuccess) rc = PIHM_ERR_HIP;
  }
  // the root layers also drive GwTransp in the RHS
  if (!rc && hipMemcpy(c->args.e.nroot, n->d.nroot, ne * sizeof(int32_t), hipMemcpyDeviceToDevice) != hipSuccess)
    rc = PIHM_ERR_HIP;
  if (rc) {
    (void)hipFree(n->arena);
    delete n;
    return rc;
  }
  n->u = NoahUni{desc->csoil, desc->sbeta, desc->salp, desc->fxexp, desc->czil, desc->lvcoef, desc->zbot,
                 desc->tbot, desc->snoalb, desc->zlvl};
  const ElemDev &E = c->args.e;
  n->s = ColDev{v.nsoil, v.sldpth, v.smcmax, v.smc, v.sh2o, E.smcmin, E.alpha, E.beta,
                E.pcpdrp, E.edir, E.ett, E.et, E.sh2o0};
  *out = n;
  return PIHM_OK;
}

void pihm_noah_lsm_destroy(pihm_noah_lsm *n) {
  if (!n) return;
  (void)hipStreamSynchronize(n->ctx->stream);
  (void)hipFree(n->arena);
  delete n;
}

int pihm_noah_lsm_set_lai(pihm_noah_lsm *n, const double *lai) {
  if (!n || !lai) return PIHM_ERR_ARG;
  return upload_caller(n, n->d.lai, lai, 1);
}

int pihm_noah_lsm_set_state(pihm_noah_lsm *n, const double *stc, const double *t1, const double *sneqv,
                            const double *snowh, const double *cmc) {
  if (!n) return PIHM_ERR_ARG;
  int rc = PIHM_OK;
  if (stc && (rc = upload_caller(n, n->d.stc, stc, ML))) return rc;
  const double *f[] = {t1, sneqv, snowh, nullptr, cmc};
  for (int k = 0; k < 5; k++)
    if (f[k] && (rc = upload_caller(n, n->d.sc + (size_t)k * n->ne, f[k], 1))) return rc;
  return PIHM_OK;
}

int pihm_noah_lsm_get_state(pihm_noah_lsm *n, double *stc, double *scalars) {
  if (!n) return PIHM_ERR_ARG;
  int rc = PIHM_OK;
  if (stc && (rc = download_caller(n, n->d.stc, stc, ML))) return rc;
  if (scalars && (rc = download_caller(n, n->d.sc, scalars, NSC))) return rc;
  return PIHM_OK;
}

int pihm_noah_lsm_step(pihm_noah_lsm *n, double t, double dt) {
  if (!n || !(dt > 0.0)) return PIHM_ERR_ARG;
  for (int k = 0; k < n->nstation; k++) {
    const double *ft = n->ftime_host.data() + (long)k * n->nrec;
    if (t < ft[0] || t > ft[n->nrec - 1]) return PIHM_ERR_ARG;   // the reference exits (src/forcing.c:487-497)
  }
  hipStream_t st = n->ctx->stream;
  hipLaunchKernelGGL(k_noah_meteo, dim3(div_up(n->nstation, kBlock)), dim3(kBlock), 0, st, n->nstation, n->nrec,
                     n->ftime, n->data, t, n->d.met);
  hipLaunchKernelGGL(k_noah, dim3(div_up(n->ne, kBlock)), dim3(kBlock), 0, st, n->ne, n->nstation, n->d, n->u, n->s,
                     dt);
  PIHM_HIP_TRY(hipGetLastError());
  return pihm_ctx_noah_refresh(n->ctx);
}

int pihm_noah_lsm_get_fluxes(pihm_noah_lsm *n, double *flux, double *et) {
  if (!n) return PIHM_ERR_ARG;
  int rc = PIHM_OK;
  if (flux && (rc = download_caller(n, n->d.flux, flux, NFX))) return rc;
  if (et && (rc = download_caller(n, n->s.et, et, ML))) return rc;
  return PIHM_OK;
}

int pihm_noah_lsm_time(pihm_noah_lsm *n, double t, double dt, int iters, double *ms_per_step) {
  if (!n || iters <= 0 || !ms_per_step) return PIHM_ERR_ARG;
  hipEvent_t ev[2];
  for (auto &e : ev) PIHM_HIP_TRY(hipEventCreate(&e));
  int rc = PIHM_OK;
  if (hipEventRecord(ev[0], n->ctx->stream) != hipSuccess) rc = PIHM_ERR_HIP;
  for (int k = 0; k < iters && !rc; k++) rc = pihm_noah_lsm_step(n, t, dt);
  if (!rc && hipEventRecord(ev[1], n->ctx->stream) != hipSuccess) rc = PIHM_ERR_HIP;
  if (!rc && hipEventSynchronize(ev[1]) != hipSuccess) rc = PIHM_ERR_HIP;
  float ms = 0.0f;
  if (!rc && hipEventElapsedTime(&ms, ev[0], ev[1]) != hipSuccess) rc = PIHM_ERR_HIP;
  for (auto &e : ev) (void)hipEventDestroy(e);
  if (!rc) *ms_per_step = ms / iters;
  return rc;
}

}  // extern "C"

namespace pihm {
pihm_noah_soil *noah_lsm_soil(pihm_noah_lsm *n) { return n->soil; }
pihm_ctx *noah_lsm_context(pihm_noah_lsm *n) { return n->ctx; }
}  // namespace pihm
