// lsm.hip - the reference's simple land-surface step on the device.
//
// Per LSM step: one thread per meteorological station interpolates the
// station's records to the current time (IntrplForc, src/forcing.c:482-530,
// with ApplyMeteoForc's unit conversion and clamp, src/forcing.c:340-349),
// then one thread per element evaluates interception, snow and
// evapotranspiration (IntcpSnowEt, src/is_sm_et.c:5-226) and writes the
// RHS's per-step inputs pcpdrp, edir and ett into the context.  Every
// expression keeps the reference's operation order, with two changes of form:
// terms that depend only on static parameters or on the vegetation inputs
// (albedo, surface pressure, the aerodynamic resistance's numerator) are
// evaluated once, when those are set, with the same expressions; and integer
// powers (T^4, T^2, betas^2) are products.  exp, log, pow and cos are the
// device library's.
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "../../include/pihm_lsm.h"
#include "common.hpp"
#include "ctx.hpp"
#include "tsinterp.hpp"

using namespace pihm;

namespace {

// src/include/pihm_const.h:9-15,80
constexpr double CP = 1004.0;
constexpr double LVH2O = 2.501e6;
constexpr double SIGMA = 5.67e-8;
constexpr double RD = 287.04;
constexpr double RV = 461.5;

enum { MET_PRCP = 0, MET_SFCTMP = 1, MET_RH = 2, MET_SFCSPD = 3, MET_SOLAR = 4 };

struct LsmDev {
  double *albedomin, *albedomax, *shdfac, *cmcfactr, *cfactr, *zlvl_wind, *rzd;
  double *rgl, *rsmin, *rsmax, *topt;
  double *depth, *porosity, *smcmin, *smcwlt, *smcref, *zmax;
  double *albedo, *pres;    // from albedomin/albedomax and zmax, once at creation
  double *ra_num;           // log(zlvl / z0) * log(10 zlvl / z0), at every vegetation update
  int32_t *station;
  double *lai, *z0;
  double *cmc, *sneqv;      // carried
  double *ec, *drip;        // the last step's canopy evaporation and drip
  double *met;              // [PIHM_MET_NVAR * nstation] interpolated station values
};

// IntrplForc (src/forcing.c:499-529) for every station, one thread each.
__global__ __launch_bounds__(kBlock) void k_meteo(int nstation, int nrec, const double *ftime, const double *data,
                                                  double t, double *met) {
  const int k = blockIdx.x * kBlock + threadIdx.x;
  if (k >= nstation) return;
  ts_interp<PIHM_MET_NVAR>(ftime + (long)k * nrec, data + (long)k * nrec * PIHM_MET_NVAR, nrec, t, met + k, nstation);
}

// Per-element values that change only with the parameters or the vegetation
// inputs, with the reference's expressions (src/is_sm_et.c:47,55-56,95-96).
__global__ __launch_bounds__(kBlock) void k_lsm_static(int ne, LsmDev l) {
  const int p = blockIdx.x * kBlock + threadIdx.x;
  if (p >= ne) return;
  l.albedo[p] = 0.5 * (l.albedomin[p] + l.albedomax[p]);
  l.pres[p] = 101.325 * 1.0e3 * pow((293.0 - 0.0065 * l.zmax[p]) / 293.0, 5.26);
}

__global__ __launch_bounds__(kBlock) void k_lsm_veg(int ne, LsmDev l) {
  const int p = blockIdx.x * kBlock + threadIdx.x;
  if (p >= ne) return;
  const double z0 = l.z0[p], zlvl = l.zlvl_wind[p];
  l.ra_num[p] = log(zlvl / z0) * log(10.0 * zlvl / z0);
}

struct StepArgs {
  LsmDev l;
  int ne, nstation;
  double stepsize, meltf, cal_edir, cal_ec, cal_ett;
  const double *y;                        // device state (unsat at ne + p, gw at 2 ne + p)
  double *pcpdrp, *edir, *ett;            // the context's step inputs
};

// IntcpSnowEt: src/is_sm_et.c:16-224, with ApplyMeteoForc's per-element
// values (src/forcing.c:340-349).
__global__ __launch_bounds__(kBlock) void k_intcp_snow_et(StepArgs a) {
  const int p = blockIdx.x * kBlock + threadIdx.x;
  if (p >= a.ne) return;
  const LsmDev &l = a.l;
  const double TSNOW = -3.0, TRAIN = 1.0, T0 = 0.0;
  const int st = l.station[p];
  const double prcp = a.l.met[MET_PRCP * a.nstation + st] / 1000.0;
  const double sfctmp_k = a.l.met[MET_SFCTMP * a.nstation + st];
  const double rh_pct = a.l.met[MET_RH * a.nstation + st];
  const double sfcspd = a.l.met[MET_SFCSPD * a.nstation + st];
  double soldn = a.l.met[MET_SOLAR * a.nstation + st];
  soldn = (soldn > 0.0) ? soldn : 0.0;
  const double unsat = a.y[a.ne + p], gw = a.y[2 * a.ne + p];
  const double stepsize = a.stepsize;
  const double shdfac = l.shdfac[p], depth = l.depth[p], rzd = l.rzd[p];
  double cmc = l.cmc[p], sneqv = l.sneqv[p];

  const double albedo = l.albedo[p];
  const double radnet = soldn * (1.0 - albedo);
  const double sfctmp = sfctmp_k - 273.15;
  const double wind = sfcspd;
  const double rh = rh_pct / 100.0;
  const double vp = 611.2 * exp(17.67 * sfctmp / (sfctmp + 243.5)) * rh;
  const double pres = l.pres[p];
  const double qv = 0.622 * vp / pres;
  const double qvsat = 0.622 * (vp / rh) / pres;
  const double lai = l.lai[p];
  const double meltf = a.meltf;

  // snow accumulation and melt
  const double frac_snow =
      (sfctmp < TSNOW) ? 1.0 : ((sfctmp > TRAIN) ? 0.0 : (TRAIN - sfctmp) / (TRAIN - TSNOW));
  const double snow_rate = frac_snow * prcp;
  sneqv += snow_rate * stepsize;
  double melt_rate = (sfctmp > T0) ? (sfctmp - T0) * meltf : 0.0;
  if (sneqv > melt_rate * stepsize) {
    sneqv -= melt_rate * stepsize;
  } else {
    melt_rate = sneqv / stepsize;
    sneqv = 0.0;
  }

  // throughfall and evaporation from the canopy
  const double intcp_max = l.cmcfactr[p] * lai * shdfac;
  const double ra = l.ra_num[p] / (wind * 0.16);
  // pow(T, 4) and pow(T, 2) as products (within an ulp of the library's)
  const double tk = sfctmp + 273.15, tk2 = tk * tk;
  const double gamma = 4.0 * 0.7 * SIGMA * RD / CP * (tk2 * tk2) / (pres / ra) + 1.0;
  const double delta = LVH2O * LVH2O * 0.622 / RV / CP / tk2 * qvsat;
  const double etp = (radnet * delta + gamma * (1.2 * LVH2O * (qvsat - qv) / ra)) / (1000.0 * LVH2O * (delta + gamma));

  double satn;
  if (depth - gw < rzd) {
    satn = 1.0;
  } else {
    satn = ((unsat / (depth - gw)) > 1.0) ? 1.0
           : ((unsat / (depth - gw)) < 0.0) ? 0.0
                                            : 0.5 * (1.0 - cos(3.14 * (unsat / (depth - gw))));
  }
  double betas = (satn * l.porosity[p] + l.smcmin[p] - l.smcwlt[p]) / (l.smcref[p] - l.smcwlt[p]);
  betas = (betas < 0.0001) ? 0.0001 : ((betas > 1.0) ? 1.0 : betas);
  double edir = (1.0 - shdfac) * (betas * betas) * etp;
  edir *= a.cal_edir;
  edir = (edir < 0.0) ? 0.0 : edir;

  double ec, ett, drip;
  if (lai > 0.0) {
    const double cfactr = l.cfactr[p];
    ec = shdfac * pow(((cmc < 0.0) ? 0.0 : ((cmc > intcp_max) ? intcp_max : cmc)) / intcp_max, cfactr) * etp;
    ec *= a.cal_ec;
    ec = (ec < 0.0) ? 0.0 : ec;
    double fr = 1.1 * radnet / (l.rgl[p] * lai);
    fr = (fr < 0.0) ? 0.0 : fr;
    const double rsmin = l.rsmin[p], rsmax = l.rsmax[p];
    double alphar = (1.0 + fr) / (fr + (rsmin / rsmax));
    alphar = (alphar > 10000.0) ? 10000.0 : alphar;
    const double dt_opt = l.topt[p] - 273.15 - sfctmp;
    double etas = 1.0 - 0.0016 * (dt_opt * dt_opt);
    etas = (etas < 0.0001) ? 0.0001 : etas;
    double gammas = 1.0 / (1.0 + 0.00025 * (vp / rh - vp));
    gammas = (gammas < 0.01) ? 0.01 : gammas;
    double rs = rsmin * alphar / (betas * lai * etas * gammas);
    rs = (rs > rsmax) ? rsmax : rs;
    const double pc = (1.0 + delta / gamma) / (1.0 + rs / ra + delta / gamma);
    ett = shdfac * pc *
          (1.0 - pow((cmc < 0.0) ? 0.0 : ((cmc > intcp_max) ? intcp_max : cmc) / intcp_max, cfactr)) * etp;
    ett *= a.cal_ett;
    ett = (ett < 0.0) ? 0.0 : ett;
    ett = ((gw < (depth - rzd)) && unsat <= 0.0) ? 0.0 : ett;
    // drip (Rutter and Morton 1977): D0 = 6.52E-7 m/s
    drip = (cmc <= 0.0) ? 0.0 : 6.52E-7 * intcp_max * exp(3.89 * cmc / intcp_max);
  } else {
    ett = 0.0;
    ec = 0.0;
    drip = 0.0;
  }
  if (drip < 0.0) drip = 0.0;
  if (drip * stepsize > cmc) drip = cmc / stepsize;

  const double isval =
      cmc + (1.0 - frac_snow) * prcp * shdfac * stepsize - ec * stepsize - drip * stepsize;
  if (isval > intcp_max) {
    cmc = intcp_max;
    drip += (isval - intcp_max) / stepsize;
  } else if (isval < 0.0) {
    // the reference rescales with cmc already set to 0 and with the rescaled
    // ec in drip's denominator (src/is_sm_et.c:204-216); kept as it is
    cmc = 0.0;
    if (ec + drip > 0.0) {
      ec = ec / (ec + drip) * (cmc + (1.0 - frac_snow) * prcp * shdfac * stepsize);
      drip = drip / (ec + drip) * (cmc + (1.0 - frac_snow) * prcp * shdfac * stepsize);
    }
  } else {
    cmc = isval;
  }
  const double pcpdrp = (1.0 - shdfac) * (1.0 - frac_snow) * prcp + drip + melt_rate;

  l.cmc[p] = cmc;
  l.sneqv[p] = sneqv;
  l.ec[p] = ec;
  l.drip[p] = drip;
  a.pcpdrp[p] = pcpdrp;
  a.edir[p] = edir;
  a.ett[p] = ett;
}

}  // namespace

struct pihm_lsm {
  pihm_ctx *ctx;
  int ne, nstation, nrec;
  LsmDev d;
  double *ftime, *data;
  std::vector<double> t_first, t_last;   // each station's record range (host copy)
  double meltf, cal_edir, cal_ec, cal_ett;
  int veg_set;
  void *arena;
};

namespace {

// Caller-order host array -> device order (single solve: device p holds
// caller element perm[p]).
int upload_elem(pihm_lsm *l, double *dst, const double *src) {
  std::vector<double> h(l->ne);
  for (int p = 0; p < l->ne; p++) h[p] = src[l->ctx->perm[p]];
  PIHM_HIP_TRY(hipMemcpyAsync(dst, h.data(), l->ne * sizeof(double), hipMemcpyHostToDevice, l->ctx->stream));
  PIHM_HIP_TRY(hipStreamSynchronize(l->ctx->stream));
  return PIHM_OK;
}

int download_elem(pihm_lsm *l, const double *src, double *dst) {
  std::vector<double> h(l->ne);
  PIHM_HIP_TRY(hipMemcpyAsync(h.data(), src, l->ne * sizeof(double), hipMemcpyDeviceToHost, l->ctx->stream));
  PIHM_HIP_TRY(hipStreamSynchronize(l->ctx->stream));
  for (int p = 0; p < l->ne; p++) dst[l->ctx->perm[p]] = h[p];
  return PIHM_OK;
}

}  // namespace

namespace pihm {
pihm_ctx *lsm_context(const pihm_lsm *l) { return l ? l->ctx : nullptr; }
// Device arrays of the carried storages and the last step's canopy
// evaporation, device order: 0 cmc, 1 sneqv, 2 ec.
const double *lsm_field(const pihm_lsm *l, int which) {
  if (!l) return nullptr;
  return which == 0 ? l->d.cmc : which == 1 ? l->d.sneqv : which == 2 ? l->d.ec : nullptr;
}
}

extern "C" {

int pihm_lsm_create(pihm_ctx *ctx, const pihm_lsm_desc *desc, const pihm_meteo *met, pihm_lsm **out) {
  if (!out) return PIHM_ERR_ARG;
  *out = nullptr;
  if (!ctx || !desc || !met || ctx->nmem != 1) return PIHM_ERR_ARG;
  if (met->nstation < 1 || met->nrec < 2 || !met->ftime || !met->data) return PIHM_ERR_ARG;
  const double *ep[] = {desc->albedomin, desc->albedomax, desc->shdfac, desc->cmcfactr, desc->cfactr,
                        desc->zlvl_wind, desc->rzd, desc->rgl, desc->rsmin, desc->rsmax, desc->topt,
                        desc->depth, desc->porosity, desc->smcmin, desc->smcwlt, desc->smcref, desc->zmax};
  for (auto q : ep) if (!q) return PIHM_ERR_ARG;
  if (!desc->station) return PIHM_ERR_ARG;
  const int ne = ctx->ne, ns = met->nstation, nrec = met->nrec;
  for (int i = 0; i < ne; i++)
    if (desc->station[i] < 0 || desc->station[i] >= ns) return PIHM_ERR_ARG;
  for (int k = 0; k < ns; k++)
    for (int r = 1; r < nrec; r++)
      if (!(met->ftime[(long)k * nrec + r] > met->ftime[(long)k * nrec + r - 1])) return PIHM_ERR_ARG;

  pihm_lsm *l = new pihm_lsm();
  l->ctx = ctx;
  l->ne = ne;
  l->nstation = ns;
  l->nrec = nrec;
  l->meltf = 0.0;
  l->cal_edir = desc->cal_edir;
  l->cal_ec = desc->cal_ec;
  l->cal_ett = desc->cal_ett;
  l->veg_set = 0;
  l->arena = nullptr;
  LsmDev &D = l->d;
  double **fields[] = {&D.albedomin, &D.albedomax, &D.shdfac, &D.cmcfactr, &D.cfactr, &D.zlvl_wind, &D.rzd,
                       &D.rgl, &D.rsmin, &D.rsmax, &D.topt, &D.depth, &D.porosity, &D.smcmin, &D.smcwlt,
                       &D.smcref, &D.zmax, &D.lai, &D.z0, &D.cmc, &D.sneqv, &D.ec, &D.drip,
                       &D.albedo, &D.pres, &D.ra_num};
  const int nf = (int)(sizeof(fields) / sizeof(fields[0]));
  const size_t fbytes = ((size_t)ne * sizeof(double) + 255) & ~size_t(255);
  const size_t sbytes = ((size_t)ne * sizeof(int32_t) + 255) & ~size_t(255);
  const size_t tbytes = ((size_t)ns * nrec * sizeof(double) + 255) & ~size_t(255);
  const size_t dbytes = ((size_t)ns * nrec * PIHM_MET_NVAR * sizeof(double) + 255) & ~size_t(255);
  const size_t mbytes = ((size_t)ns * PIHM_MET_NVAR * sizeof(double) + 255) & ~size_t(255);
  const size_t total = nf * fbytes + sbytes + tbytes + dbytes + mbytes;
  if (hipMalloc(&l->arena, total) != hipSuccess) { delete l; return PIHM_ERR_HIP; }
  char *base = (char *)l->arena;
  for (int f = 0; f < nf; f++) { *fields[f] = (double *)base; base += fbytes; }
  D.station = (int32_t *)base; base += sbytes;
  l->ftime = (double *)base; base += tbytes;
  l->data = (double *)base; base += dbytes;
  D.met = (double *)base;
  int rc = PIHM_OK;
  if (hipMemset(l->arena, 0, total) != hipSuccess) rc = PIHM_ERR_HIP;
  for (int f = 0; f < 17 && !rc; f++) rc = upload_elem(l, *fields[f], ep[f]);
  if (!rc) {
    std::vector<int32_t> s(ne);
    for (int p = 0; p < ne; p++) s[p] = desc->station[ctx->perm[p]];
    if (hipMemcpy(D.station, s.data(), ne * sizeof(int32_t), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(l->ftime, met->ftime, (size_t)ns * nrec * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(l->data, met->data, (size_t)ns * nrec * PIHM_MET_NVAR * sizeof(double), hipMemcpyHostToDevice) !=
            hipSuccess)
      rc = PIHM_ERR_HIP;
  }
  if (!rc) {
    hipLaunchKernelGGL(k_lsm_static, dim3(div_up(ne, kBlock)), dim3(kBlock), 0, ctx->stream, ne, l->d);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(ctx->stream) != hipSuccess) rc = PIHM_ERR_HIP;
  }
  if (rc) { pihm_lsm_destroy(l); return rc; }
  l->t_first.resize(ns);
  l->t_last.resize(ns);
  for (int k = 0; k < ns; k++) {
    l->t_first[k] = met->ftime[(long)k * nrec];
    l->t_last[k] = met->ftime[(long)k * nrec + nrec - 1];
  }
  *out = l;
  return PIHM_OK;
}

void pihm_lsm_destroy(pihm_lsm *l) {
  if (!l) return;
  if (l->arena) (void)hipFree(l->arena);
  delete l;
}

int pihm_lsm_set_veg(pihm_lsm *l, const double *lai, const double *z0, double meltf) {
  if (!l || !lai || !z0) return PIHM_ERR_ARG;
  int rc = upload_elem(l, l->d.lai, lai);
  if (!rc) rc = upload_elem(l, l->d.z0, z0);
  if (rc) return rc;
  hipLaunchKernelGGL(k_lsm_veg, dim3(div_up(l->ne, kBlock)), dim3(kBlock), 0, l->ctx->stream, l->ne, l->d);
  PIHM_HIP_TRY(hipGetLastError());
  l->meltf = meltf;
  l->veg_set = 1;
  return PIHM_OK;
}

int pihm_lsm_set_state(pihm_lsm *l, const double *cmc, const double *sneqv) {
  if (!l || !cmc || !sneqv) return PIHM_ERR_ARG;
  int rc = upload_elem(l, l->d.cmc, cmc);
  return rc ? rc : upload_elem(l, l->d.sneqv, sneqv);
}

int pihm_lsm_get_state(pihm_lsm *l, double *cmc, double *sneqv) {
  if (!l) return PIHM_ERR_ARG;
  int rc = cmc ? download_elem(l, l->d.cmc, cmc) : PIHM_OK;
  if (!rc && sneqv) rc = download_elem(l, l->d.sneqv, sneqv);
  return rc;
}

static int check_step(pihm_lsm *l, double t, double stepsize, N_Vector y) {
  if (!l || !y || !(stepsize > 0.0) || !l->veg_set) return PIHM_ERR_ARG;
  if (N_VGetLength_Hip(y) != l->ctx->neq) return PIHM_ERR_ARG;
  for (int k = 0; k < l->nstation; k++)
    if (t < l->t_first[k] || t > l->t_last[k]) return PIHM_ERR_ARG;
  return PIHM_OK;
}

static void launch_step(pihm_lsm *l, double t, double stepsize, N_Vector y) {
  hipStream_t st = l->ctx->stream;
  hipLaunchKernelGGL(k_meteo, dim3(div_up(l->nstation, kBlock)), dim3(kBlock), 0, st, l->nstation, l->nrec,
                     l->ftime, l->data, t, l->d.met);
  StepArgs a;
  a.l = l->d;
  a.ne = l->ne;
  a.nstation = l->nstation;
  a.stepsize = stepsize;
  a.meltf = l->meltf;
  a.cal_edir = l->cal_edir;
  a.cal_ec = l->cal_ec;
  a.cal_ett = l->cal_ett;
  a.y = N_VGetDeviceArrayPointer_Hip(y);
  a.pcpdrp = l->ctx->args.e.pcpdrp;
  a.edir = l->ctx->args.e.edir;
  a.ett = l->ctx->args.e.ett;
  hipLaunchKernelGGL(k_intcp_snow_et, dim3(div_up(l->ne, kBlock)), dim3(kBlock), 0, st, a);
}

int pihm_lsm_step(pihm_lsm *l, double t, double stepsize, N_Vector y) {
  if (int rc = check_step(l, t, stepsize, y)) return rc;
  launch_step(l, t, stepsize, y);
  PIHM_HIP_TRY(hipGetLastError());
  return PIHM_OK;
}

int pihm_lsm_time(pihm_lsm *l, double t, double stepsize, N_Vector y, int iters, double *ms_per_step) {
  if (iters <= 0 || !ms_per_step) return PIHM_ERR_ARG;
  if (int rc = check_step(l, t, stepsize, y)) return rc;
  hipEvent_t ev[2];
  for (auto &e : ev) PIHM_HIP_TRY(hipEventCreate(&e));
  PIHM_HIP_TRY(hipEventRecord(ev[0], l->ctx->stream));
  for (int it = 0; it < iters; it++) launch_step(l, t, stepsize, y);
  PIHM_HIP_TRY(hipGetLastError());
  PIHM_HIP_TRY(hipEventRecord(ev[1], l->ctx->stream));
  PIHM_HIP_TRY(hipEventSynchronize(ev[1]));
  float ms = 0;
  PIHM_HIP_TRY(hipEventElapsedTime(&ms, ev[0], ev[1]));
  for (auto &e : ev) (void)hipEventDestroy(e);
  *ms_per_step = ms / iters;
  return PIHM_OK;
}

int pihm_lsm_get_fluxes(pihm_lsm *l, double *pcpdrp, double *edir, double *ett, double *ec, double *drip) {
  if (!l) return PIHM_ERR_ARG;
  const pihm::ElemDev &E = l->ctx->args.e;
  const double *src[] = {E.pcpdrp, E.edir, E.ett, l->d.ec, l->d.drip};
  double *dst[] = {pcpdrp, edir, ett, ec, drip};
  for (int f = 0; f < 5; f++)
    if (dst[f]) {
      const int rc = download_elem(l, src[f], dst[f]);
      if (rc) return rc;
    }
  return PIHM_OK;
}

}  // extern "C"
