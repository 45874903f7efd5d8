/*
 * noahlsm_oracle.c - CPU restatement of the Noah land-surface step of the
 * reference's _NOAH_ build (Flux-PIHM): Noah() (src/noah/noah.c:3-55) and
 * SFlx with everything it calls (:166-609), run once per LSM step per element.
 *
 * TEST INFRASTRUCTURE ONLY (see pihm_oracle.h): the checker for the HIP Noah
 * step and the host twin of the coupled runs.  Nothing in the product links
 * it.  Configuration: _NOAH_ on, _CYCLES_ off, rad_mode 0 (the station's
 * SOLAR record is the downward shortwave), no urban land cover (lc.isurban 0:
 * the urban overrides of src/noah/noah.c:206-219 and the fixed conductivity
 * 3.24 are not taken), and SfcDifOff's bulk Richardson number ribb stays at
 * its initial 0 (src/initialize.c:665; nothing in this configuration updates
 * it).  Every function keeps the reference's operation order.
 */
#include "noahlsm_oracle.h"

#include <math.h>
#include <string.h>

/* src/include/pihm_const.h:6-37,81 */
#define PI        3.14159265
#define GRAV      9.80665
#define CP        1004.0
#define LVH2O     2.501e6
#define SIGMA     5.67e-8
#define RD        287.04
#define CPH2O     4.218e3
#define CPICE     2.106e3
#define LSUBF     3.335e5
#define EMISSI_S  0.95
#define TFREEZ    273.15
#define LSUBS     2.83e6
#define BADVAL    -999
#define SH2OMIN   0.02
#define MAXLYR    11

void oracle_Rosr12(double *p, const double *a, const double *b, double *c, const double *d, double *delta,
                   int nsoil);   /* noahsoil_oracle.c */

/* working record of one element: the reference's ps/ws/wf/es/ef fields the
 * step touches */
typedef struct {
    const oracle_noahlsm_par *par;
    oracle_noahlsm_state *st;
    /* ps */
    double zsoil[MAXLYR], rtdis[MAXLYR];
    int nsoil, nroot;
    double rh, sfcspd, sfcprs, q2, q2sat, dqsdt2, ffrozp, alb, embrd, z0brd, z0, cmcmax, sncovr;
    double rr, rch, epsca, rc, pc, rcs, rct, rcq, rcsoil, beta, eta_kinematic, soilw, proj_lai, sncond;
    /* ws */
    double soilm;
    /* wf */
    double prcp, pcpdrp, snomlt, edir, ec, ett, et[MAXLYR], etp, eta, etns, esnow, dew, drip;
    /* es */
    double sfctmp, th2;
    /* ef */
    double soldn, longwave, solnet, lwdn, fdown, ssoil, flx1, flx2, flx3, sheat, ef_edir, ef_ec, ef_ett,
        ef_esnow, ef_etp, ef_eta;
} Col;

/* CalHum: src/noah/noah.c:659-702 */
static void CalHum(Col *c)
{
    const double A2 = 17.67, A3 = 273.15, A4 = 29.65, ELWV = 2.501e6, E0 = 611.0, RVV = 461.0,
        EPSILON = 0.622, SVP1 = 611.2, SVP2 = 17.67, SVP3 = 29.65, SVPT0 = 273.15;
    double a23m4, e, esat, svp, rh;
    rh = c->rh / 100.0;
    svp = SVP1 * exp(SVP2 * (c->sfctmp - SVPT0) / (c->sfctmp - SVP3));
    e = rh * svp;
    c->q2 = (0.622 * e) / (c->sfcprs - (1.0 - 0.622) * e);
    c->th2 = c->sfctmp + (0.0098 * c->par->zlvl);
    a23m4 = A2 * (A3 - A4);
    esat = E0 * exp(ELWV / RVV * (1.0 / A3 - 1.0 / c->sfctmp));
    c->q2sat = EPSILON * esat / (c->sfcprs - (1.0 - EPSILON) * esat);
    c->dqsdt2 = c->q2sat * a23m4 / ((c->sfctmp - A4) * (c->sfctmp - A4));
}

/* FrozRain: src/noah/noah.c:961-976 */
static double FrozRain(double prcp, double sfctmp)
{
    return (prcp > 0.0 && sfctmp < TFREEZ) ? 1.0 : 0.0;
}

/* CSnow: src/noah/noah.c:811-839 */
double oracle_CSnow(double dsnow)
{
    const double UNIT = 0.11631;
    double c = 0.328 * pow(10, 2.25 * dsnow);
    return 2.0 * UNIT * c;
}

/* SnowNew: src/noah/noah.c:2222-2268 */
static void SnowNew(Col *c, double newsn)
{
    double dsnew, hnewc, snowhc, newsnc, tempc;
    oracle_noahlsm_state *st = c->st;
    snowhc = st->snowh * 100.0;
    newsnc = newsn * 100.0;
    tempc = c->sfctmp - 273.15;
    if (tempc <= -15.0)
        dsnew = 0.05;
    else
        dsnew = 0.05 + 0.0017 * pow(tempc + 15.0, 1.5);
    hnewc = newsnc / dsnew;
    if (snowhc + hnewc < 0.001)
        st->sndens = (dsnew > st->sndens) ? dsnew : st->sndens;
    else
        st->sndens = (snowhc * st->sndens + hnewc * dsnew) / (snowhc + hnewc);
    snowhc = snowhc + hnewc;
    st->snowh = snowhc * 0.01;
}

/* SnFrac: src/noah/noah.c:1713-1744 */
double oracle_SnFrac(double sneqv, double snup, double salp)
{
    double rsnow;
    if (sneqv < snup) {
        rsnow = sneqv / snup;
        return 1.0 - (exp(-salp * rsnow) - rsnow * exp(-salp));
    }
    return 1.0;
}

/* AlCalc: src/noah/noah.c:610-657 */
static void AlCalc(Col *c, double dt, int snowng)
{
    const double SNACCA = 0.94, SNACCB = 0.58;
    oracle_noahlsm_state *st = c->st;
    double snoalb1, snoalb2, snoalb = c->par->snoalb;
    st->albedo = c->alb + c->sncovr * (snoalb - c->alb);
    st->emissi = c->embrd + c->sncovr * (EMISSI_S - c->embrd);
    snoalb1 = snoalb + c->par->lvcoef * (0.85 - snoalb);
    snoalb2 = snoalb1;
    if (snowng) {
        st->snotime1 = 0.0;
    } else {
        st->snotime1 += dt;
        snoalb2 = snoalb1 * pow(SNACCA, pow(st->snotime1 / 86400.0, SNACCB));
    }
    snoalb2 = (snoalb2 > c->alb) ? snoalb2 : c->alb;
    st->albedo = c->alb + c->sncovr * (snoalb2 - c->alb);
    st->albedo = (st->albedo > snoalb2) ? snoalb2 : st->albedo;
}

/* TDfCnd: src/noah/noah.c:2520-2619 */
double oracle_TDfCnd(double smc, double qz, double smcmax, double smcmin, double sh2o)
{
    const double THKICE = 2.2, THKQTZ = 7.7, THKW = 0.57;
    double df, ake, gammd, thkdry, thko, thksat, thks, satratio, xu, xunfroz;
    satratio = (smc - smcmin) / (smcmax - smcmin);
    thko = 2.0;
    thks = pow(THKQTZ, qz) * pow(thko, 1.0 - qz);
    xunfroz = sh2o / smc;
    xu = xunfroz * smcmax;
    thksat = pow(thks, 1.0 - smcmax) * pow(THKICE, smcmax - xu) * pow(THKW, xu);
    gammd = (1.0 - smcmax) * 2700.0;
    thkdry = (0.135 * gammd + 64.7) / (2700.0 - 0.947 * gammd);
    if ((sh2o + 0.0005) < smc)
        ake = satratio;
    else if (satratio > 0.1)
        ake = log10(satratio) + 1.0;
    else
        ake = 0.0;
    df = ake * (thksat - thkdry) + thkdry;
    return df;
}

/* Snowz0: src/noah/noah.c:2202-2220 */
static double Snowz0(double sncovr, double z0brd, double snowh)
{
    const double Z0S = 0.001;
    double burial = 7.0 * z0brd - snowh;
    double z0eff = (burial <= 0.0007) ? Z0S : (burial / 7.0);
    return (1.0 - sncovr) * z0brd + sncovr * z0eff;
}

/* Paulson's and Lech's stability functions: src/noah/noah.c:3173-3222 */
static double Pslmu(double zz) { return -0.96 * log(1.0 - 4.5 * zz); }
static double Pslms(double zz)
{
    const double RIC = 0.183;
    double rric = 1.0 / RIC;
    return zz * rric - 2.076 * (1.0 - 1.0 / (zz + 1.0));
}
static double Pslhu(double zz) { return -0.96 * log(1.0 - 4.5 * zz); }
static double Pslhs(double zz)
{
    const double RIC = 0.183, FHNEU = 0.8, RFC = 0.191;
    double rfac = RIC / (FHNEU * RFC * RFC);
    return zz * rfac - 2.076 * (1.0 - exp(-1.2 * zz));
}
double oracle_Pspmu(double xx)
{
    return -2.0 * log((xx + 1.0) * 0.5) - log((xx * xx + 1.0) * 0.5) + 2.0 * atan(xx) - PI * 0.5;
}
static double Pspms(double yy) { return 5.0 * yy; }
double oracle_Psphu(double xx) { return -2.0 * log((xx * xx + 1.0) * 0.5); }
static double Psphs(double yy) { return 5.0 * yy; }

/* SfcDifOff: src/noah/noah.c:2862-3079 (iz0tlnd = 0; ilech = 0) */
void oracle_SfcDifOff(double zlvl, double zlvl_wind, double z0, double czil, double sfcspd, double t1v,
                      double th2v, double *cm, double *ch)
{
    double zilfc, zu, zt, rdz, cxch, dthv, du2, btgh, wstar2, ustar, zslu, zslt, rlogu, rlogt, rlmo;
    double zetalt, zetalu, zetau, zetat, xlu4, xlt4, xu4, xt4, xlu, xlt, xu, xt, psmz, simm, pshz, simh;
    double ustark, rlmn, rlma;
    int ilech, itr;
    const double WWST = 1.2, VKRM = 0.40, EXCM = 0.001, BETA = 1.0 / 270.0;
    double wwst2, btg, elfc, WOLD = 0.15, wnew;
    const int ITRMX = 5;
    const double EPSU2 = 1.0e-4, EPSUST = 0.07, ZTMIN = -5.0, ZTMAX = 1.0, HPBL = 1000.0, SQVISC = 258.2;

    wwst2 = WWST * WWST;
    btg = BETA * GRAV;
    elfc = VKRM * btg;
    wnew = 1.0 - WOLD;
    ilech = 0;
    /* iz0tlnd == 0: the original czil */
    zilfc = -czil * VKRM * SQVISC;
    zu = z0;
    rdz = 1.0 / zlvl_wind;
    cxch = EXCM * rdz;
    dthv = th2v - t1v;
    du2 = sfcspd * sfcspd;
    du2 = (du2 > EPSU2) ? du2 : EPSU2;
    btgh = btg * HPBL;
    if (btgh * *ch * dthv != 0.0)
        wstar2 = wwst2 * pow(fabs(btgh * *ch * dthv), 2.0 / 3.0);
    else
        wstar2 = 0.0;
    ustar = sqrt(*cm * sqrt(du2 + wstar2));
    ustar = (ustar > EPSUST) ? ustar : EPSUST;
    zt = exp(zilfc * sqrt(ustar * z0)) * z0;
    zslu = zlvl_wind + zu;
    zslt = zlvl + zt;
    rlogu = log(zslu / zu);
    rlogt = log(zslt / zt);
    rlmo = elfc * *ch * dthv / pow(ustar, 3);
    for (itr = 0; itr < ITRMX; itr++) {
        zetalt = zslt * rlmo;
        zetalt = zetalt > ZTMIN ? zetalt : ZTMIN;
        rlmo = zetalt / zslt;
        zetalu = zslu * rlmo;
        zetau = zu * rlmo;
        zetat = zt * rlmo;
        if (ilech == 0) {
            if (rlmo < 0.0) {
                xlu4 = 1.0 - 16.0 * zetalu;
                xlt4 = 1.0 - 16.0 * zetalt;
                xu4 = 1.0 - 16.0 * zetau;
                xt4 = 1.0 - 16.0 * zetat;
                xlu = sqrt(sqrt(xlu4));
                xlt = sqrt(sqrt(xlt4));
                xu = sqrt(sqrt(xu4));
                xt = sqrt(sqrt(xt4));
                psmz = oracle_Pspmu(xu);
                simm = oracle_Pspmu(xlu) - psmz + rlogu;
                pshz = oracle_Psphu(xt);
                simh = oracle_Psphu(xlt) - pshz + rlogt;
            } else {
                zetalu = (zetalu < ZTMAX) ? zetalu : ZTMAX;
                zetalt = (zetalt < ZTMAX) ? zetalt : ZTMAX;
                psmz = Pspms(zetau);
                simm = Pspms(zetalu) - psmz + rlogu;
                pshz = Psphs(zetat);
                simh = Psphs(zetalt) - pshz + rlogt;
            }
        } else {
            if (rlmo < 0.0) {
                psmz = Pslmu(zetau);
                simm = Pslmu(zetalu) - psmz + rlogu;
                pshz = Pslhu(zetat);
                simh = Pslhu(zetalt) - pshz + rlogt;
            } else {
                zetalu = zetalu < ZTMAX ? zetalu : ZTMAX;
                zetalt = zetalt < ZTMAX ? zetalt : ZTMAX;
                psmz = Pslms(zetau);
                simm = Pslms(zetalu) - psmz + rlogu;
                pshz = Pslhs(zetat);
                simh = Pslhs(zetalt) - pshz + rlogt;
            }
        }
        ustar = sqrt(*cm * sqrt(du2 + wstar2));
        ustar = (ustar > EPSUST) ? ustar : EPSUST;
        zt = exp(zilfc * sqrt(ustar * z0)) * z0;
        zslt = zlvl + zt;
        rlogt = log(zslt / zt);
        ustark = ustar * VKRM;
        *cm = (ustark / simm > cxch) ? ustark / simm : cxch;
        *ch = (ustark / simh > cxch) ? ustark / simh : cxch;
        if (btgh * *ch * dthv != 0.0)
            wstar2 = wwst2 * pow(fabs(btgh * *ch * dthv), 2.0 / 3.0);
        else
            wstar2 = 0.0;
        rlmn = elfc * *ch * dthv / pow(ustar, 3.0);
        rlma = rlmo * WOLD + rlmn * wnew;
        rlmo = rlma;
    }
}

/* Penman: src/noah/noah.c:1508-1570 */
static void Penman(Col *c, double *t24, double t2v, int snowng, int frzgra)
{
    const double ELCP = 2.4888e+3, LSUBC = 2.501000e+6;
    double a, delta, fnet, rad, rho, emissi, elcp1, lvs;
    oracle_noahlsm_state *st = c->st;
    emissi = st->emissi;
    elcp1 = (1.0 - c->sncovr) * ELCP + c->sncovr * ELCP * LSUBS / LSUBC;
    lvs = (1.0 - c->sncovr) * LSUBC + c->sncovr * LSUBS;
    c->flx2 = 0.0;
    delta = elcp1 * c->dqsdt2;
    *t24 = c->sfctmp * c->sfctmp * c->sfctmp * c->sfctmp;
    c->rr = emissi * *t24 * 6.48e-8 / (c->sfcprs * st->ch) + 1.0;
    rho = c->sfcprs / (RD * t2v);
    c->rch = rho * CP * st->ch;
    if (!snowng) {
        if (c->prcp > 0.0) c->rr += CPH2O * c->prcp * 1000.0 / c->rch;
    } else {
        c->rr += CPICE * c->prcp * 1000.0 / c->rch;
    }
    fnet = c->fdown - emissi * SIGMA * *t24 - c->ssoil;
    if (frzgra) {
        c->flx2 = -LSUBF * c->prcp * 1000.0;
        fnet -= c->flx2;
    }
    rad = fnet / c->rch + c->th2 - c->sfctmp;
    a = elcp1 * (c->q2sat - c->q2);
    c->epsca = (a * c->rr + rad * delta) / (delta + c->rr);
    c->etp = c->epsca * c->rch / lvs / 1000.0;
}

/* CanRes: src/noah/noah.c:704-809 (_CYCLES_ off) */
static void CanRes(Col *c)
{
    const oracle_noahlsm_par *p = c->par;
    oracle_noahlsm_state *st = c->st;
    double delta, rr, ff, gx, part[MAXLYR];
    int k;
    const double SLV = 2.501000e6;
    c->rcs = 0.0;
    c->rct = 0.0;
    c->rcq = 0.0;
    c->rcsoil = 0.0;
    c->rc = 0.0;
    ff = 0.55 * 2.0 * c->soldn / (p->rgl * c->proj_lai);
    c->rcs = (ff + p->rsmin / p->rsmax) / (1.0 + ff);
    c->rcs = (c->rcs > 0.0001) ? c->rcs : 0.0001;
    c->rct = 1.0 - 0.0016 * pow(p->topt - c->sfctmp, 2.0);
    c->rct = (c->rct > 0.0001) ? c->rct : 0.0001;
    c->rcq = 1.0 / (1.0 + p->hs * (c->q2sat - c->q2));
    c->rcq = (c->rcq > 0.01) ? c->rcq : 0.01;
    gx = (st->sh2o[0] - p->smcwlt) / (p->smcref - p->smcwlt);
    gx = (gx > 1.0) ? 1.0 : gx;
    gx = (gx < 0.0) ? 0.0 : gx;
    part[0] = (c->zsoil[0] / c->zsoil[c->nroot - 1]) * gx;
    for (k = 1; k < c->nroot; k++) {
        gx = (st->sh2o[k] - p->smcwlt) / (p->smcref - p->smcwlt);
        gx = (gx > 1.0) ? 1.0 : gx;
        gx = (gx < 0.0) ? 0.0 : gx;
        part[k] = ((c->zsoil[k] - c->zsoil[k - 1]) / c->zsoil[c->nroot - 1]) * gx;
    }
    for (k = 0; k < c->nroot; k++) c->rcsoil += part[k];
    c->rcsoil = (c->rcsoil > 0.0001) ? c->rcsoil : 0.0001;
    c->rc = p->rsmin / (c->proj_lai * c->rcs * c->rct * c->rcq * c->rcsoil);
    rr = (4.0 * st->emissi * SIGMA * RD / CP) * pow(c->sfctmp, 4.0) / (c->sfcprs * st->ch) + 1.0;
    delta = (SLV / CP) * c->dqsdt2;
    c->pc = (rr + delta) / (rr * (1.0 + c->rc * st->ch) + delta);
}

/* DEvap: src/noah/noah.c:841-870 */
static void DEvap(Col *c)
{
    const oracle_noahlsm_par *p = c->par;
    double fx, sratio;
    sratio = (c->st->sh2o[0] - p->smcdry) / (p->smcmax - p->smcdry);
    if (sratio > 0.0) {
        fx = pow(sratio, p->fxexp);
        fx = (fx > 1.0) ? 1.0 : fx;
        fx = (fx < 0.0) ? 0.0 : fx;
    } else {
        fx = 0.0;
    }
    c->edir = fx * (1.0 - p->shdfac) * c->etp;
}

/* Transp: src/noah/noah.c:2725-2811 */
static void Transp(Col *c)
{
    const oracle_noahlsm_par *p = c->par;
    oracle_noahlsm_state *st = c->st;
    int i, k;
    double denom, etpa, gx[MAXLYR], rtx, sgx;
    for (k = 0; k < c->nsoil; k++) c->et[k] = 0.0;
    if (st->cmc != 0.0)
        etpa = p->shdfac * c->pc * c->etp * (1.0 - pow(st->cmc / c->cmcmax, p->cfactr));
    else
        etpa = p->shdfac * c->pc * c->etp;
    sgx = 0.0;
    for (i = 0; i < c->nroot; i++) {
        gx[i] = (st->smc[i] - p->smcwlt) / (p->smcref - p->smcwlt);
        gx[i] = (gx[i] < 0.0) ? 0.0 : gx[i];
        gx[i] = (gx[i] > 1.0) ? 1.0 : gx[i];
        sgx += gx[i];
    }
    sgx /= (double)c->nroot;
    denom = 0.0;
    for (i = 0; i < c->nroot; i++) {
        rtx = c->rtdis[i] + gx[i] - sgx;
        gx[i] *= ((rtx > 0.0) ? rtx : 0.0);
        denom += gx[i];
    }
    denom = (denom <= 0.0) ? 1.0 : denom;
    for (i = 0; i < c->nroot; i++) c->et[i] = etpa * gx[i] / denom;
}

/* Evapo: src/noah/noah.c:872-959 (_CYCLES_ off) */
static void Evapo(Col *c, double dt)
{
    const oracle_noahlsm_par *p = c->par;
    oracle_noahlsm_state *st = c->st;
    int k;
    double cmc2ms;
    c->edir = 0.0;
    c->ec = 0.0;
    c->ett = 0.0;
    for (k = 0; k < c->nsoil; k++) c->et[k] = 0.0;
    if (c->etp > 0.0) {
        if (p->shdfac < 1.0) DEvap(c);
        if (p->shdfac > 0.0) {
            Transp(c);
            for (k = 0; k < c->nsoil; k++) c->ett += c->et[k];
            if (st->cmc > 0.0)
                c->ec = p->shdfac * pow((st->cmc / c->cmcmax > 1.0) ? 1.0 : st->cmc / c->cmcmax, p->cfactr) * c->etp;
            else
                c->ec = 0.0;
            cmc2ms = st->cmc / dt;
            c->ec = (cmc2ms < c->ec) ? cmc2ms : c->ec;
        }
    }
    c->etns = c->edir + c->ett + c->ec;
}

/* PcpDrp: src/noah/noah.c:1451-1506 */
static void PcpDrp(Col *c, double prcp, double dt)
{
    const oracle_noahlsm_par *p = c->par;
    oracle_noahlsm_state *st = c->st;
    double excess, rhsct, trhsct;
    const double KD = 6.54e-7, BFACTR = 3.89;
    rhsct = p->shdfac * prcp - c->ec;
    c->drip = 0.0;
    trhsct = dt * rhsct;
    excess = st->cmc + trhsct;
    if (excess > 0.0) {
        if (excess >= c->cmcmax)
            c->drip = (KD * c->cmcmax * exp(BFACTR)) + (excess - c->cmcmax) / dt;
        else
            c->drip = (KD * c->cmcmax * exp(BFACTR * excess / c->cmcmax));
        c->drip = (c->drip > excess / dt) ? excess / dt : c->drip;
    }
    rhsct -= c->drip;
    c->pcpdrp = (1.0 - p->shdfac) * prcp + c->drip;
    st->cmc += dt * rhsct;
    if (st->cmc < 1.0e-20) st->cmc = 0.0;
    st->cmc = (st->cmc < c->cmcmax) ? st->cmc : c->cmcmax;
}

/* TBnd: src/noah/noah.c:2482-2518 */
double oracle_TBnd(double tu, double tb, const double *zsoil, double zbot, int k, int nsoil)
{
    double zb, zup;
    zup = (k == 0) ? 0.0 : zsoil[k - 1];
    if (k == nsoil - 1)
        zb = 2.0 * zbot - zsoil[k];
    else
        zb = zsoil[k + 1];
    return tu + (tb - tu) * (zup - zsoil[k]) / (zup - zb);
}

/* TmpAvg: src/noah/noah.c:2621-2723 */
double oracle_TmpAvg(double tup, double tm, double tdn, const double *zsoil, int k)
{
    double dz, dzh, x0, xdn, xup, tavg;
    dz = (k == 0) ? -zsoil[0] : zsoil[k - 1] - zsoil[k];
    dzh = dz * 0.5;
    if (tup < TFREEZ) {
        if (tm < TFREEZ) {
            if (tdn < TFREEZ) {
                tavg = (tup + 2.0 * tm + tdn) / 4.0;
            } else {
                x0 = (TFREEZ - tm) * dzh / (tdn - tm);
                tavg = 0.5 * (tup * dzh + tm * (dzh + x0) + TFREEZ * (2.0 * dzh - x0)) / dz;
            }
        } else {
            if (tdn < TFREEZ) {
                xup = (TFREEZ - tup) * dzh / (tm - tup);
                xdn = dzh - (TFREEZ - tm) * dzh / (tdn - tm);
                tavg = 0.5 * (tup * xup + TFREEZ * (2.0 * dz - xup - xdn) + tdn * xdn) / dz;
            } else {
                xup = (TFREEZ - tup) * dzh / (tm - tup);
                tavg = 0.5 * (tup * xup + TFREEZ * (2.0 * dz - xup)) / dz;
            }
        }
    } else {
        if (tm < TFREEZ) {
            if (tdn < TFREEZ) {
                xup = dzh - (TFREEZ - tup) * dzh / (tm - tup);
                tavg = 0.5 * (TFREEZ * (dz - xup) + tm * (dzh + xup) + tdn * dzh) / dz;
            } else {
                xup = dzh - (TFREEZ - tup) * dzh / (tm - tup);
                xdn = (TFREEZ - tm) * dzh / (tdn - tm);
                tavg = 0.5 * (TFREEZ * (2.0 * dz - xup - xdn) + tm * (xup + xdn)) / dz;
            }
        } else {
            if (tdn < TFREEZ) {
                xdn = dzh - (TFREEZ - tm) * dzh / (tdn - tm);
                tavg = (TFREEZ * (dz - xdn) + 0.5 * (TFREEZ + tdn) * xdn) / dz;
            } else {
                tavg = (tup + 2.0 * tm + tdn) / 4.0;
            }
        }
    }
    return tavg;
}

/* FrH2O: src/noah/noah.c:978-1090 */
double oracle_FrH2O(double tkelv, double smc, double sh2o, double smcmax, double smcmin, double alpha,
                    double beta)
{
    double denom, df, dswl, fk, swl, swlk, satn, mx = 0.0, freew;
    int nlog = 0, kcount = 0;
    const double CK = 8.0, ERROR = 0.005;
    if (tkelv > (TFREEZ - 1.0e-3)) {
        freew = smc;
    } else {
        if (CK != 0.0) {
            swl = smc - sh2o;
            swl = (swl > smc - SH2OMIN) ? (smc - SH2OMIN) : swl;
            swl = (swl < 0.0) ? 0.0 : swl;
            while (nlog < 10 && kcount == 0) {
                nlog++;
                satn = (smc - swl - smcmin) / (smcmax - smcmin);
                mx = beta / (1.0 - beta);
                df = log(GRAV / alpha / LSUBF) + 1.0 / beta * log(pow(satn, mx) - 1.0) +
                    2.0 * log(1.0 + CK * swl) - log(-(tkelv - TFREEZ) / tkelv);
                denom = 1.0 / (beta - 1.0) / (smcmax - smcmin) * pow(satn, mx - 1.0) / (pow(satn, mx) - 1.0) +
                    2.0 * CK / (1.0 + CK * swl);
                swlk = swl - df / denom;
                swlk = (swlk > smc - SH2OMIN) ? (smc - SH2OMIN) : swlk;
                swlk = (swlk < 0.0) ? 0.0 : swlk;
                dswl = fabs(swlk - swl);
                swl = swlk;
                if (dswl <= ERROR) kcount++;
            }
            freew = smc - swl;
        }
        if (kcount == 0) {
            fk = pow(pow(-(tkelv - TFREEZ) / tkelv * alpha * LSUBF / GRAV, beta), 1.0 / mx) * (smcmax - smcmin) -
                smcmin;
            fk = (fk < SH2OMIN) ? SH2OMIN : fk;
            freew = (fk < smc) ? fk : smc;
        }
    }
    return freew;
}

/* SnkSrc: src/noah/noah.c:1746-1811 */
static void SnkSrc(const Col *c, double *tsnsr, double tavg, double smc, double *sh2o, double dt, int k,
                   double qtot)
{
    const oracle_noahlsm_par *p = c->par;
    double dz, freew, xh2o;
    double DH2O = 1.0000e3;
    dz = (0 == k) ? -c->zsoil[0] : c->zsoil[k - 1] - c->zsoil[k];
    freew = oracle_FrH2O(tavg, smc, *sh2o, p->smcmax, p->smcmin, p->alpha, p->beta);
    xh2o = *sh2o + qtot * dt / (DH2O * LSUBF * dz);
    if (xh2o < *sh2o && xh2o < freew) xh2o = (freew > *sh2o) ? *sh2o : freew;
    if (xh2o > *sh2o && xh2o > freew) xh2o = (freew < *sh2o) ? *sh2o : freew;
    xh2o = (xh2o < 0.0) ? 0.0 : xh2o;
    xh2o = (xh2o > smc) ? smc : xh2o;
    *tsnsr = -DH2O * LSUBF * dz * (xh2o - *sh2o) / dt;
    *sh2o = xh2o;
}

/* HRT: src/noah/noah.c:1092-1299 (itavg true) */
static void HRT(Col *c, double *rhsts, double yy, double zz1, double dt, double df1, double *ai, double *bi,
                double *ci)
{
    const oracle_noahlsm_par *p = c->par;
    oracle_noahlsm_state *st = c->st;
    int k;
    double ddz, ddz2, denom, df1k, dtsdz, dtsdz2 = 0.0, hcpct, sice, csoil_loc, df1n, qtot, tavg, tbk,
        tbk1 = 0.0, tsnsr, tsurf;
    const double CH2O = 4.2e6;
    csoil_loc = p->csoil;
    hcpct = st->sh2o[0] * CH2O + (1.0 - p->smcmax) * csoil_loc + (p->smcmax - st->smc[0]) * CP +
        (st->smc[0] - st->sh2o[0]) * CPICE;
    ddz = 1.0 / (-0.5 * c->zsoil[1]);
    ai[0] = 0.0;
    ci[0] = (df1 * ddz) / (c->zsoil[0] * hcpct);
    bi[0] = -ci[0] + df1 / (0.5 * c->zsoil[0] * c->zsoil[0] * hcpct * zz1);
    dtsdz = (st->stc[0] - st->stc[1]) / (-0.5 * c->zsoil[1]);
    c->ssoil = df1 * (st->stc[0] - yy) / (0.5 * c->zsoil[0] * zz1);
    denom = c->zsoil[0] * hcpct;
    rhsts[0] = (df1 * dtsdz - c->ssoil) / denom;
    qtot = -1.0 * rhsts[0] * denom;
    sice = st->smc[0] - st->sh2o[0];
    tsurf = (yy + (zz1 - 1.0) * st->stc[0]) / zz1;
    tbk = oracle_TBnd(st->stc[0], st->stc[1], c->zsoil, p->zbot, 0, c->nsoil);
    if ((sice > 0.0) || (st->stc[0] < TFREEZ) || (tsurf < TFREEZ) || (tbk < TFREEZ)) {
        tavg = oracle_TmpAvg(tsurf, st->stc[0], tbk, c->zsoil, 0);
        SnkSrc(c, &tsnsr, tavg, st->smc[0], &st->sh2o[0], dt, 0, qtot);
        rhsts[0] -= tsnsr / denom;
    }
    ddz2 = 0.0;
    df1k = df1;
    for (k = 1; k < c->nsoil; k++) {
        hcpct = st->sh2o[k] * CH2O + (1.0 - p->smcmax) * csoil_loc + (p->smcmax - st->smc[k]) * CP +
            (st->smc[k] - st->sh2o[k]) * CPICE;
        if (k != c->nsoil - 1) {
            df1n = oracle_TDfCnd(st->smc[k], p->quartz, p->smcmax, p->smcmin, st->sh2o[k]);
            denom = 0.5 * (c->zsoil[k - 1] - c->zsoil[k + 1]);
            dtsdz2 = (st->stc[k] - st->stc[k + 1]) / denom;
            ddz2 = 2.0 / (c->zsoil[k - 1] - c->zsoil[k + 1]);
            ci[k] = -df1n * ddz2 / ((c->zsoil[k - 1] - c->zsoil[k]) * hcpct);
            tbk1 = oracle_TBnd(st->stc[k], st->stc[k + 1], c->zsoil, p->zbot, k, c->nsoil);
        } else {
            df1n = oracle_TDfCnd(st->smc[k], p->quartz, p->smcmax, p->smcmin, st->sh2o[k]);
            denom = 0.5 * (c->zsoil[k - 1] + c->zsoil[k]) - p->zbot;
            dtsdz2 = (st->stc[k] - p->tbot) / denom;
            ci[k] = 0.0;
            tbk1 = oracle_TBnd(st->stc[k], p->tbot, c->zsoil, p->zbot, k, c->nsoil);
        }
        denom = (c->zsoil[k] - c->zsoil[k - 1]) * hcpct;
        rhsts[k] = (df1n * dtsdz2 - df1k * dtsdz) / denom;
        qtot = -1.0 * denom * rhsts[k];
        sice = st->smc[k] - st->sh2o[k];
        tavg = oracle_TmpAvg(tbk, st->stc[k], tbk1, c->zsoil, k);
        if ((sice > 0.0) || (st->stc[k] < TFREEZ) || (tbk < TFREEZ) || (tbk1 < TFREEZ)) {
            SnkSrc(c, &tsnsr, tavg, st->smc[k], &st->sh2o[k], dt, k, qtot);
            rhsts[k] = rhsts[k] - tsnsr / denom;
        }
        ai[k] = -df1k * ddz / ((c->zsoil[k - 1] - c->zsoil[k]) * hcpct);
        bi[k] = -(ai[k] + ci[k]);
        tbk = tbk1;
        df1k = df1n;
        dtsdz = dtsdz2;
        ddz = ddz2;
    }
}

/* HStep: src/noah/noah.c:1301-1340 */
static void HStep(Col *c, double *rhsts, double dt, double *ai, double *bi, double *ci)
{
    int k, nsoil = c->nsoil;
    double rhstsin[MAXLYR], ciin[MAXLYR];
    for (k = 0; k < nsoil; k++) {
        rhsts[k] *= dt;
        ai[k] *= dt;
        bi[k] = 1.0 + bi[k] * dt;
        ci[k] *= dt;
    }
    for (k = 0; k < nsoil; k++) rhstsin[k] = rhsts[k];
    for (k = 0; k < nsoil; k++) ciin[k] = ci[k];
    oracle_Rosr12(ci, ai, bi, ciin, rhstsin, rhsts, nsoil);
    for (k = 0; k < nsoil; k++) c->st->stc[k] += ci[k];
}

/* ShFlx: src/noah/noah.c:1621-1652 */
static void ShFlx(Col *c, double dt, double yy, double zz1, double df1)
{
    double ai[MAXLYR], bi[MAXLYR], ci[MAXLYR], rhsts[MAXLYR];
    oracle_noahlsm_state *st = c->st;
    HRT(c, rhsts, yy, zz1, dt, df1, ai, bi, ci);
    HStep(c, rhsts, dt, ai, bi, ci);
    st->t1 = (yy + (zz1 - 1.0) * st->stc[0]) / zz1;
    c->ssoil = df1 * (st->stc[0] - st->t1) / (0.5 * c->zsoil[0]);
}

/* NoPac: src/noah/noah.c:1343-1449 (_CYCLES_ off) */
static void NoPac(Col *c, double dt, double t24)
{
    const oracle_noahlsm_par *p = c->par;
    oracle_noahlsm_state *st = c->st;
    int k;
    double df1, yy, zz1, yynum, prcpf;
    prcpf = c->prcp;
    c->dew = 0.0;
    c->edir = 0.0;
    c->ec = 0.0;
    for (k = 0; k < c->nsoil; k++) c->et[k] = 0.0;
    c->ett = 0.0;
    if (c->etp > 0.0) {
        Evapo(c, dt);
        c->eta = c->etns;
    } else {
        c->dew = -c->etp;
        prcpf += c->dew;
    }
    PcpDrp(c, prcpf, dt);
    if (c->etp <= 0.0) {
        c->beta = 0.0;
        c->eta = c->etp;
        if (c->etp < 0.0) c->beta = 1.0;
    } else {
        c->beta = c->eta / c->etp;
    }
    df1 = oracle_TDfCnd(st->smc[0], p->quartz, p->smcmax, p->smcmin, st->sh2o[0]);
    df1 *= exp(p->sbeta * p->shdfac);
    yynum = c->fdown - st->emissi * SIGMA * t24;
    yy = c->sfctmp + (yynum / c->rch + c->th2 - c->sfctmp - c->beta * c->epsca) / c->rr;
    zz1 = df1 / (-0.5 * c->zsoil[0] * c->rch * c->rr) + 1.0;
    ShFlx(c, dt, yy, zz1, df1);
    c->flx1 = CPH2O * c->prcp * 1000.0 * (st->t1 - c->sfctmp);
    c->flx3 = 0.0;
}

/* SnowPack: src/noah/noah.c:2100-2200 */
void oracle_SnowPack(double esd, double dtsec, double *snowh, double *sndens, double tsnow, double tsoil)
{
    int ipol, j;
    double bfac, dsx, dthr, dw, snowhc, pexp, tavgc, tsnowc, tsoilc, esdc, esdcx;
    const double C1 = 0.01, C2 = 21.0;
    snowhc = *snowh * 100.0;
    esdc = esd * 100.0;
    dthr = dtsec / 3600.0;
    tsnowc = tsnow - 273.15;
    tsoilc = tsoil - 273.15;
    tavgc = 0.5 * (tsnowc + tsoilc);
    esdcx = (esdc > 1.0e-2) ? esdc : 1.0e-2;
    bfac = dthr * C1 * exp(0.08 * tavgc - C2 * *sndens);
    ipol = 4;
    pexp = 0.0;
    for (j = ipol; j > 0; j--) pexp = (1.0 + pexp) * bfac * esdcx / (double)(j + 1);
    pexp += 1.0;
    dsx = *sndens * pexp;
    dsx = (dsx > 0.40) ? 0.40 : dsx;
    dsx = (dsx < 0.05) ? 0.05 : dsx;
    *sndens = dsx;
    if (tsnowc >= 0.0) {
        dw = 0.13 * dthr / 24.0;
        *sndens = *sndens * (1.0 - dw) + dw;
        if (*sndens >= 0.40) *sndens = 0.40;
    }
    (void)snowhc;
    snowhc = esdc / *sndens;
    *snowh = snowhc * 0.01;
}

/* SnoPac: src/noah/noah.c:1813-2098 (_CYCLES_ off) */
static void SnoPac(Col *c, int snowng, double dt, double t24, double prcpf, double df1)
{
    oracle_noahlsm_state *st = c->st;
    int k;
    double denom, dsoil, dtot, esnow2, etp3, etanrg, ex, seh, t12, t12a, t12b, t14, ssoil1, t11, yy, zz1;
    const double ESDMIN = 1.0e-6, SNOEXP = 2.0;
    c->dew = 0.0;
    c->edir = 0.0;
    c->ec = 0.0;
    for (k = 0; k < c->nsoil; k++) c->et[k] = 0.0;
    c->ett = 0.0;
    c->esnow = 0.0;
    esnow2 = 0.0;
    c->beta = 1.0;
    if (c->etp <= 0.0) {
        /* ribb is 0 here, so the frost adjustment of src/noah/noah.c:1870-1876 does not apply */
        if (c->etp == 0.0) c->beta = 0.0;
        c->dew = -c->etp;
        esnow2 = c->etp * dt;
        etanrg = c->etp * 1000.0 * ((1.0 - c->sncovr) * LVH2O + c->sncovr * LSUBS);
    } else {
        if (c->sncovr < 1.0) {
            Evapo(c, dt);
            c->edir *= (1.0 - c->sncovr);
            c->ec *= (1.0 - c->sncovr);
            for (k = 0; k < c->nsoil; k++) c->et[k] = c->et[k] * (1.0 - c->sncovr);
            c->ett *= (1.0 - c->sncovr);
            c->etns *= (1.0 - c->sncovr);
        }
        c->esnow = c->etp * c->sncovr;
        esnow2 = c->esnow * dt;
        etanrg = c->esnow * 1000.0 * LSUBS + c->etns * 1000.0 * LVH2O;
    }
    c->flx1 = 0.0;
    if (snowng) {
        c->flx1 = CPICE * c->prcp * 1000.0 * (st->t1 - c->sfctmp);
    } else if (c->prcp > 0.0) {
        c->flx1 = CPH2O * c->prcp * 1000.0 * (st->t1 - c->sfctmp);
    }
    dsoil = -0.5 * c->zsoil[0];
    dtot = st->snowh + dsoil;
    denom = 1.0 + df1 / (dtot * c->rr * c->rch);
    t12a = ((c->fdown - c->flx1 - c->flx2 - st->emissi * SIGMA * t24) / c->rch + c->th2 - c->sfctmp -
            etanrg / c->rch) / c->rr;
    t12b = df1 * st->stc[0] / (dtot * c->rr * c->rch);
    t12 = (c->sfctmp + t12a + t12b) / denom;
    if (t12 <= TFREEZ) {
        st->t1 = t12;
        c->ssoil = df1 * (st->t1 - st->stc[0]) / dtot;
        if (st->sneqv - esnow2 > 0.0) {
            st->sneqv -= esnow2;
        } else {
            st->sneqv = 0.0;
            esnow2 = st->sneqv;
            c->esnow = esnow2 / dt;
        }
        c->flx3 = 0.0;
        ex = 0.0;
        c->snomlt = 0.0;
    } else {
        st->t1 = TFREEZ * pow(c->sncovr, SNOEXP) + t12 * (1.0 - pow(c->sncovr, SNOEXP));
        c->beta = 1.0;
        c->ssoil = df1 * (st->t1 - st->stc[0]) / dtot;
        if (st->sneqv - esnow2 <= ESDMIN) {
            st->sneqv = 0.0;
            ex = 0.0;
            c->snomlt = 0.0;
            c->flx3 = 0.0;
        } else {
            st->sneqv -= esnow2;
            etp3 = c->etp * 1000.0 * LVH2O;
            (void)etp3;
            seh = c->rch * (st->t1 - c->th2);
            t14 = st->t1 * st->t1;
            t14 = t14 * t14;
            c->flx3 = c->fdown - c->flx1 - c->flx2 - st->emissi * SIGMA * t14 - c->ssoil - seh - etanrg;
            c->flx3 = (c->flx3 <= 0.0) ? 0.0 : c->flx3;
            ex = c->flx3 * 0.001 / LSUBF;
            c->snomlt = ex;
            if (st->sneqv - c->snomlt * dt >= ESDMIN) {
                st->sneqv -= c->snomlt * dt;
            } else {
                ex = st->sneqv / dt;
                c->flx3 = ex * 1000.0 * LSUBF;
                c->snomlt = st->sneqv / dt;
                st->sneqv = 0.0;
            }
        }
        prcpf += ex;
    }
    PcpDrp(c, prcpf, dt);
    zz1 = 1.0;
    yy = st->stc[0] - 0.5 * c->ssoil * c->zsoil[0] * zz1 / df1;
    t11 = st->t1;
    ssoil1 = c->ssoil;
    ShFlx(c, dt, yy, zz1, df1);
    st->t1 = t11;
    c->ssoil = ssoil1;
    if (st->sneqv > 0.0) {
        oracle_SnowPack(st->sneqv, dt, &st->snowh, &st->sndens, st->t1, yy);
    } else {
        st->sneqv = 0.0;
        st->snowh = 0.0;
        st->sndens = 0.0;
        c->sncovr = 0.0;
    }
}

/* SFlx: src/noah/noah.c:166-609 (_CYCLES_ off) */
static void SFlx(Col *c, double dt)
{
    const oracle_noahlsm_par *p = c->par;
    oracle_noahlsm_state *st = c->st;
    int frzgra, snowng, k;
    const int IZ0TLND = 0;
    double df1, df1a, dsoil, dtot, frcsno, frcsoi, t1v, th2v, t2v, t24 = 0.0, interp_fraction, sn_new, prcpf,
        soilwm, soilww, smav[MAXLYR];
    double shdfac = p->shdfac;
    (void)IZ0TLND;

    c->snomlt = 0.0;
    c->pcpdrp = 0.0;
    if (shdfac > 0.0) c->proj_lai = (c->proj_lai > 0.5) ? c->proj_lai : 0.5;
    c->cmcmax = shdfac * p->cmcfactr * c->proj_lai;
    if (c->proj_lai >= p->laimax) {
        c->embrd = p->emissmax;
        c->alb = p->albedomin;
        c->z0brd = p->z0max;
    } else if (c->proj_lai <= p->laimin) {
        c->embrd = p->emissmin;
        c->alb = p->albedomax;
        c->z0brd = p->z0min;
    } else {
        if (p->laimax > p->laimin) {
            interp_fraction = (c->proj_lai - p->laimin) / (p->laimax - p->laimin);
            interp_fraction = (interp_fraction < 1.0) ? interp_fraction : 1.0;
            interp_fraction = (interp_fraction > 0.0) ? interp_fraction : 0.0;
            c->embrd = ((1.0 - interp_fraction) * p->emissmin) + (interp_fraction * p->emissmax);
            c->alb = ((1.0 - interp_fraction) * p->albedomax) + (interp_fraction * p->albedomin);
            c->z0brd = ((1.0 - interp_fraction) * p->z0min) + (interp_fraction * p->z0max);
        } else {
            c->embrd = 0.5 * p->emissmin + 0.5 * p->emissmax;
            c->alb = 0.5 * p->albedomin + 0.5 * p->albedomax;
            c->z0brd = 0.5 * p->z0min + 0.5 * p->z0max;
        }
    }
    snowng = 0;
    frzgra = 0;
    if (st->sneqv <= 1.0e-7) {
        st->sneqv = 0.0;
        st->sndens = 0.0;
        st->snowh = 0.0;
        c->sncond = 1.0;
    } else {
        st->sndens = st->sneqv / st->snowh;
        /* the reference exits when sndens > 1 (src/noah/noah.c:272-279) */
        c->sncond = oracle_CSnow(st->sndens);
    }
    if (c->prcp > 0.0) {
        if (c->ffrozp > 0.5)
            snowng = 1;
        else if (st->t1 <= TFREEZ)
            frzgra = 1;
    }
    if (snowng || frzgra) {
        sn_new = c->prcp * dt;
        st->sneqv += sn_new;
        prcpf = 0.0;
        SnowNew(c, sn_new);
        c->sncond = oracle_CSnow(st->sndens);
    } else {
        prcpf = c->prcp;
    }
    if (st->sneqv == 0.0) {
        c->sncovr = 0.0;
        st->albedo = c->alb;
        st->emissi = c->embrd;
    } else {
        c->sncovr = oracle_SnFrac(st->sneqv, p->snup, p->salp);
        c->sncovr = (c->sncovr < 0.98) ? c->sncovr : 0.98;
        AlCalc(c, dt, snowng);
    }
    df1 = oracle_TDfCnd(st->smc[0], p->quartz, p->smcmax, p->smcmin, st->sh2o[0]);
    df1 *= exp(p->sbeta * shdfac);
    if (c->sncovr > 0.97) df1 = c->sncond;
    dsoil = -(0.5 * c->zsoil[0]);
    if (st->sneqv == 0.0) {
        c->ssoil = df1 * (st->t1 - st->stc[0]) / dsoil;
    } else {
        dtot = st->snowh + dsoil;
        frcsno = st->snowh / dtot;
        frcsoi = dsoil / dtot;
        df1a = frcsno * c->sncond + frcsoi * df1;
        df1 = df1a * c->sncovr + df1 * (1.0 - c->sncovr);
        c->ssoil = df1 * (st->t1 - st->stc[0]) / dtot;
    }
    if (c->sncovr > 0.0)
        c->z0 = Snowz0(c->sncovr, c->z0brd, st->snowh);
    else
        c->z0 = c->z0brd;
    t1v = st->t1 * (1.0 + 0.61 * c->q2);
    th2v = c->th2 * (1.0 + 0.61 * c->q2);
    oracle_SfcDifOff(p->zlvl, p->zlvl_wind, c->z0, p->czil, c->sfcspd, t1v, th2v, &st->cm, &st->ch);
    c->fdown = c->solnet + c->lwdn;
    t2v = c->sfctmp * (1.0 + 0.61 * c->q2);
    Penman(c, &t24, t2v, snowng, frzgra);
    if (shdfac > 0.0)
        CanRes(c);
    else
        c->rc = 0.0;
    c->esnow = 0.0;
    if (st->sneqv == 0.0) {
        NoPac(c, dt, t24);
        c->eta_kinematic = c->eta * 1000.0;
    } else {
        SnoPac(c, snowng, dt, t24, prcpf, df1);
        c->eta_kinematic = (c->esnow + c->etns) * 1000.0;
    }
    st->q1 = c->q2 + c->eta_kinematic * CP / c->rch;
    c->sheat = -(st->ch * CP * c->sfcprs) / (RD * t2v) * (c->th2 - st->t1);
    c->ef_edir = c->edir * 1000.0 * LVH2O;
    c->ef_ec = c->ec * 1000.0 * LVH2O;
    c->ef_ett = c->ett * 1000.0 * LVH2O;
    c->ef_esnow = c->esnow * 1000.0 * LSUBS;
    c->ef_etp = c->etp * 1000.0 * ((1.0 - c->sncovr) * LVH2O + c->sncovr * LSUBS);
    if (c->ef_etp > 0.0)
        c->ef_eta = c->ef_edir + c->ef_ec + c->ef_ett + c->ef_esnow;
    else
        c->ef_eta = c->ef_etp;
    c->beta = (c->ef_etp == 0.0) ? 0.0 : (c->ef_eta / c->ef_etp);
    c->ssoil *= -1.0;
    c->soilm = -1.0 * st->smc[0] * c->zsoil[0];
    for (k = 1; k < c->nsoil; k++) c->soilm += st->smc[k] * (c->zsoil[k - 1] - c->zsoil[k]);
    soilwm = -1.0 * (p->smcmax - p->smcwlt) * c->zsoil[0];
    soilww = -1.0 * (st->smc[0] - p->smcwlt) * c->zsoil[0];
    for (k = 0; k < c->nsoil; k++) smav[k] = (st->smc[k] - p->smcwlt) / (p->smcmax - p->smcwlt);
    (void)smav;
    if (c->nroot > 1) {
        for (k = 1; k < c->nroot; k++) {
            soilwm += (p->smcmax - p->smcwlt) * (c->zsoil[k - 1] - c->zsoil[k]);
            soilww += (st->smc[k] - p->smcwlt) * (c->zsoil[k - 1] - c->zsoil[k]);
        }
    }
    if (soilwm < 1.0e-6) {
        soilwm = 0.0;
        c->soilw = 0.0;
        c->soilm = 0.0;
    } else {
        c->soilw = soilww / soilwm;
    }
}

/* Noah: src/noah/noah.c:3-55, one element; DefSldpth's zsoil and RootDist's
 * rtdis (src/noah/lsm_func.c:86-205) from the layer thicknesses */
void oracle_noah_element(const oracle_noahlsm_par *p, oracle_noahlsm_state *st, const oracle_noahlsm_forc *f,
                         double dt, oracle_noahlsm_out *o)
{
    Col c;
    int j, kz;
    memset(&c, 0, sizeof(c));
    c.par = p;
    c.st = st;
    c.nsoil = p->nsoil;
    c.nroot = p->nroot;
    c.zsoil[0] = -p->sldpth[0];
    for (kz = 1; kz < p->nsoil; kz++) c.zsoil[kz] = -p->sldpth[kz] + c.zsoil[kz - 1];
    for (j = 0; j < p->nroot; j++) c.rtdis[j] = -p->sldpth[j] / c.zsoil[p->nroot - 1];
    c.sfctmp = f->sfctmp;
    c.rh = f->rh;
    c.sfcspd = f->sfcspd;
    c.soldn = f->soldn;
    c.longwave = f->longwave;
    c.sfcprs = f->sfcprs;
    c.prcp = f->prcp;
    c.proj_lai = f->proj_lai;

    CalHum(&c);
    c.ffrozp = FrozRain(c.prcp, c.sfctmp);
    c.alb = BADVAL;
    if (st->q1 == BADVAL) st->q1 = c.q2;
    c.solnet = c.soldn * (1.0 - st->albedo);
    c.lwdn = c.longwave * st->emissi;
    for (j = 0; j < p->nsoil; j++) {
        st->smc[j] = (st->smc[j] > p->smcmin + SH2OMIN) ? st->smc[j] : p->smcmin + SH2OMIN;
        st->smc[j] = (st->smc[j] < p->smcmax) ? st->smc[j] : p->smcmax;
        st->sh2o[j] = (st->sh2o[j] < st->smc[j]) ? st->sh2o[j] : st->smc[j];
    }
    SFlx(&c, dt);

    memset(o, 0, sizeof(*o));
    /* ET: W m-2 back to m s-1 (src/noah/noah.c:50-53) */
    o->ec = c.ef_ec / LVH2O / 1000.0;
    o->ett = c.ef_ett / LVH2O / 1000.0;
    o->edir = c.ef_edir / LVH2O / 1000.0;
    for (j = 0; j < MAXLYR; j++) o->et[j] = (j < p->nsoil) ? c.et[j] : 0.0;
    o->pcpdrp = c.pcpdrp;
    o->drip = c.drip;
    o->esnow = c.esnow;
    o->etp = c.etp;
    o->snomlt = c.snomlt;
    o->ssoil = c.ssoil;
    o->sheat = c.sheat;
    o->eta = c.ef_eta;
    o->fdown = c.fdown;
    o->sncovr = c.sncovr;
    o->beta = c.beta;
    o->rc = c.rc;
    o->pc = c.pc;
    o->flx1 = c.flx1;
    o->flx2 = c.flx2;
    o->flx3 = c.flx3;
    o->dew = c.dew;
    o->soilw = c.soilw;
    o->soilm = c.soilm;
    o->cmcmax = c.cmcmax;
    o->rr = c.rr;
    o->rch = c.rch;
    o->epsca = c.epsca;
    o->q2 = c.q2;
    o->q2sat = c.q2sat;
    o->th2 = c.th2;
}

/* The whole mesh: per-element arrays, caller order. */
void oracle_noah_step(int ne, double dt, const oracle_noahlsm_par *par, oracle_noahlsm_state *st,
                      const oracle_noahlsm_forc *forc, oracle_noahlsm_out *out)
{
    int i;
#if defined(_OPENMP)
# pragma omp parallel for
#endif
    for (i = 0; i < ne; i++) oracle_noah_element(&par[i], &st[i], &forc[i], dt, &out[i]);
}
