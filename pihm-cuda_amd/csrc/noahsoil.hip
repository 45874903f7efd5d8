// noahsoil.hip - the _NOAH_ build's soil-water step after each model step.
//
// One thread per element runs the element's soil column in the reference's
// order: Summary's FindWaterTable and CalcLatFlx (src/update.c:37-43), then
// NoahHydrol (src/noah/noah.c:57-94): FindWaterTable again, the moisture
// clamps, and SmFlx with SRT, SStep, Rosr12 and AdjSmProf.  The layer arrays
// live in HBM layer-major ([k * ne + p]) so that the threads of a wavefront
// read one layer of 64 neighbouring elements together; inside the thread the
// column (at most 11 layers) is worked in private arrays.  Every expression
// keeps the reference's operation order; powers use xpow (as the RHS does) and
// exp is the device library's.  At the end the thread leaves in the context
// what the next model step's RHS reads: sh2o[0], fcr, nwtbl and GwTransp's
// share (src/noah/lsm_func.c:163-185) for the new water-table layer.
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "../../include/pihm_noahsoil.h"
#include "common.hpp"
#include "ctx.hpp"
#include "physics.hpp"

using namespace pihm;

namespace {

constexpr int ML = PIHM_NOAH_MAXLYR;
constexpr double SH2OMIN = 0.02;   // src/include/pihm_const.h:81

struct SoilDev {
  int32_t *nsoil;
  double *sldpth, *smcmax, *frzx;               // static
  double *smc, *sh2o, *smflxv, *runoff2_lyr;    // carried, [ML * ne]
  double *satdpth;                              // [ML * ne]
  double *runoff3;                              // carried
  double *infil, *runoff2;                      // the step's MassBalance values
  double *prev;                                 // [2 * ne] the previous Summary's unsat, gw
};

// MassBalance (src/update.c:107-160) with the _NOAH_ EtExtract split of the
// fluxes evaluated at y: infiltration re-derived from the storage change since
// the previous Summary, and the subsurface runoff that CalcLatFlx spreads.
__global__ __launch_bounds__(kBlock) void k_mass_balance_noah(int ne, SoilDev s, ElemDev e, const double *y,
                                                              const double *sub, double dt) {
  const int p = blockIdx.x * kBlock + threadIdx.x;
  if (p >= ne) return;
  const double depth = e.depth[p], area = e.area[p];
  const double u = y[ne + p], g = y[2L * ne + p];
  double soilw0 = s.prev[ne + p] + s.prev[p];
  soilw0 = (soilw0 > depth) ? depth : soilw0;
  soilw0 = (soilw0 < 0.0) ? 0.0 : soilw0;
  double soilw1 = g + u;
  soilw1 = (soilw1 > depth) ? depth : soilw1;
  soilw1 = (soilw1 < 0.0) ? 0.0 : soilw1;
  double subrunoff = 0.0;
  for (int j = 0; j < 3; j++) subrunoff += sub[(long)j * ne + p] / area;
  // EtExtract with _NOAH_ (src/hydrol.c:43-56,79-84) at the clamped groundwater
  const double gc = (g > 0.0) ? g : 0.0;
  const double edir = e.edir[p], ett = e.ett[p], gwet = e.gwet[p];
  double edir_u, edir_g;
  if (gc > depth - e.dinf[p]) { edir_u = 0.0; edir_g = edir; }
  else { edir_u = edir; edir_g = 0.0; }
  const double ett_u = (1.0 - gwet) * ett, ett_g = gwet * ett;
  double infil = (soilw1 - soilw0) * e.porosity[p] / dt + subrunoff + edir_u + edir_g + ett_u + ett_g;
  if (infil < 0.0) {
    subrunoff -= infil;
    infil = 0.0;
  }
  s.infil[p] = infil;
  s.runoff2[p] = subrunoff;
  s.prev[p] = u;
  s.prev[ne + p] = g;
}

__global__ __launch_bounds__(kBlock) void k_soil_prev(int ne, SoilDev s, const double *y) {
  const int p = blockIdx.x * kBlock + threadIdx.x;
  if (p >= ne) return;
  s.prev[p] = y[ne + p];
  s.prev[ne + p] = y[2L * ne + p];
}

// WDfCnd: src/noah/noah.c:2813-2861
__device__ void wdf_cnd(double *wdf, double *wcnd, double smc, double sicemax, double smcmax, double smcmin,
                        double ksatv, double alpha, double beta) {
  double factr1 = 0.05 / (smcmax - smcmin);
  double factr2 = (smc - smcmin) / (smcmax - smcmin);
  factr2 = (factr2 > 1.0 - 5.0e-4) ? 1.0 - 5.0e-4 : factr2;
  factr2 = (factr2 < 0.0 + 5.0e-4) ? 5.0e-4 : factr2;
  factr1 = (factr1 < factr2) ? factr1 : factr2;
  const double expon = 1.0 - 1.0 / beta;
  double satkfunc = kr_func(beta, factr2);
  double dpsidsm = (1.0 - expon) / alpha / expon / (smcmax - smcmin) *
                   xpow(xpow(factr2, -1.0 / expon) - 1.0, -expon) * xpow(factr2, -(1.0 / expon + 1.0));
  *wcnd = ksatv * satkfunc;
  *wdf = *wcnd * dpsidsm;
  if (sicemax > 0.0) {
    const double vkwgt = 1.0 / (1.0 + xpow(500.0 * sicemax, 3.0));
    satkfunc = kr_func(beta, factr1);
    dpsidsm = (1.0 - expon) / alpha / expon / (smcmax - smcmin) *
              xpow(xpow(factr1, -1.0 / expon) - 1.0, -expon) * xpow(factr1, -(1.0 / expon + 1.0));
    *wdf = vkwgt * *wdf + (1.0 - vkwgt) * dpsidsm * satkfunc * ksatv;
  }
}

// Every loop over the layers runs over the fixed 11 with a test for the
// column's own nsoil, so each array index is a constant after unrolling and
// the column stays in registers.  The water table, the lateral split and the
// layer fluxes go straight to their HBM fields, which the thread reads back
// where the reference reads them again.
__global__ __launch_bounds__(kBlock) void k_noah_soil(int ne, SoilDev s, ElemDev e, const double *y, double dt) {
  const int p = blockIdx.x * kBlock + threadIdx.x;
  if (p >= ne) return;
  const int nsoil = s.nsoil[p];
  const double smcmax = s.smcmax[p], smcmin = e.smcmin[p];
  const double ksatv = e.ksatv[p], alpha = e.alpha[p], beta = e.beta[p];
  const double gw = y[2L * ne + p];
  auto at = [ne, p](double *f, int k) -> double & { return f[(long)k * ne + p]; };
  double zsoil[ML], smc[ML], sh2o[ML], sice[ML], ai[ML], bi[ML], ci[ML], rhstt[ML];
  // layer depths below the surface (src/noah/lsm_func.c:150-160)
#pragma unroll
  for (int k = 0; k < ML; k++) {
    zsoil[k] = 0.0;
    if (k < nsoil) zsoil[k] = (k == 0) ? -at(s.sldpth, 0) : -at(s.sldpth, k) + zsoil[k - 1];
    smc[k] = at(s.smc, k);
    sh2o[k] = at(s.sh2o, k);
  }
  double fcr = e.fcr[p];

  // FindWaterTable (src/noah/lsm_func.c:4-55); Summary and NoahHydrol call it
  // with the same arguments (src/update.c:40, src/noah/noah.c:72), so once
  int nwtbl = -999;
  {
    double depth = 0.0;
#pragma unroll
    for (int k = 0; k < ML; k++)
      if (k < nsoil) depth += at(s.sldpth, k);
    if (gw <= 0.0) {
      nwtbl = nsoil;
#pragma unroll
      for (int k = 0; k < ML; k++) at(s.satdpth, k) = (k == nsoil - 1) ? 1.0E-3 : 0.0;
    } else if (gw > depth) {
      nwtbl = 0;
#pragma unroll
      for (int k = 0; k < ML; k++) at(s.satdpth, k) = (k < nsoil) ? at(s.sldpth, k) : 0.0;
    } else {
      double dsum = 0.0;
      bool found = false;
#pragma unroll
      for (int k = ML - 1; k >= 0; k--) {
        double v = 0.0;
        if (k < nsoil && !found) {
          const double sl = at(s.sldpth, k);
          if (dsum + sl > gw) {
            v = gw - dsum;
            nwtbl = k + 1;
            found = true;
          } else {
            v = sl;
            dsum += sl;
          }
        }
        at(s.satdpth, k) = v;
      }
    }
  }
  // CalcLatFlx: src/noah/lsm_func.c:226-255 (runoff2 is MassBalance's subsurface runoff)
  {
    double sattot = 0.0;
#pragma unroll
    for (int k = 0; k < ML; k++)
      if (k < nsoil) sattot += at(s.satdpth, k);
    const double runoff2 = s.runoff2[p];
#pragma unroll
    for (int k = 0; k < ML; k++) {
      if (sattot <= 0.0) {
        if (k == nsoil - 1) at(s.runoff2_lyr, k) = runoff2;
      } else if (k < nsoil) {
        at(s.runoff2_lyr, k) = at(s.satdpth, k) / sattot * runoff2;
      }
    }
  }
  // NoahHydrol's clamps: src/noah/noah.c:74-83
#pragma unroll
  for (int k = 0; k < ML; k++) {
    if (k < nsoil) {
      smc[k] = (smc[k] > smcmin + SH2OMIN) ? smc[k] : smcmin + SH2OMIN;
      smc[k] = (smc[k] < smcmax) ? smc[k] : smcmax;
      sh2o[k] = (sh2o[k] < smc[k]) ? sh2o[k] : smc[k];
    }
    sice[k] = smc[k] - sh2o[k];   // SmFlx: src/noah/noah.c:1676-1679
  }
  if (0 == nwtbl) {
#pragma unroll
    for (int k = 0; k < ML; k++) {
      if (k < nsoil) {
        smc[k] = smcmax;
        sh2o[k] = smc[k] - sice[k];
      }
    }
  } else {
    // SRT: src/noah/noah.c:2315-2425
    {
      const double CVFRZ = 3.0;
      double sicemax = 0.0, dice = 0.0;
#pragma unroll
      for (int k = 0; k < ML; k++) {
        if (k < nsoil) {
          sicemax = (sice[k] > sicemax) ? sice[k] : sicemax;
          dice = (k == 0) ? -zsoil[0] * sice[0] : dice + (zsoil[k - 1] - zsoil[k]) * sice[k];
        }
      }
      fcr = 1.0;
      if (dice > 1.0e-2) {
        const double acrt = CVFRZ * s.frzx[p] / dice;
        double sum = 1.0;
        const double ialp1 = (int)round(CVFRZ) - 1;
        for (int j = 1; j < ialp1 + 1; j++) {
          int kf = 1;
          for (int jj = j + 1; jj < ialp1; jj++) kf *= jj;
          sum += xpow(acrt, CVFRZ - (double)j) / (double)kf;
        }
        fcr = 1.0 - exp(-acrt) * sum;
      }
      const double pddum = s.infil[p];
      double wdf, wcnd;
      double dsmdz = (sh2o[0] - sh2o[1]) / (-0.5 * zsoil[1]);
      wdf_cnd(&wdf, &wcnd, sh2o[0], sicemax, smcmax, smcmin, ksatv, alpha, beta);
      double ddz = 1.0 / (-0.5 * zsoil[1]);
      ai[0] = 0.0;
      bi[0] = wdf * ddz / (-zsoil[0]);
      ci[0] = -bi[0];
      rhstt[0] = (wdf * dsmdz + wcnd - pddum + e.edir[p] + at(e.et, 0)) / zsoil[0];
      rhstt[0] += at(s.runoff2_lyr, 0) / zsoil[0];
#pragma unroll
      for (int k = 1; k < ML; k++) {
        ai[k] = bi[k] = ci[k] = rhstt[k] = 0.0;
        if (k < nsoil) {
          const double denom2 = (zsoil[k - 1] - zsoil[k]);
          double wdf2 = 0.0, wcnd2 = 0.0, dsmdz2 = 0.0, ddz2 = 0.0;
          if (k + 1 < ML && k < nsoil - 1) {
            const double denom = zsoil[k - 1] - zsoil[k + 1];
            dsmdz2 = (sh2o[k] - sh2o[k + 1]) / (denom * 0.5);
            wdf_cnd(&wdf2, &wcnd2, sh2o[k], sicemax, smcmax, smcmin, ksatv, alpha, beta);
            ddz2 = 2.0 / denom;
            ci[k] = -wdf2 * ddz2 / denom2;
          } else {
            ci[k] = 0.0;
          }
          double numer = (wdf2 * dsmdz2) + wcnd2 - (wdf * dsmdz) - wcnd + at(e.et, k);
          numer += at(s.runoff2_lyr, k);
          rhstt[k] = numer / (-denom2);
          ai[k] = -wdf * ddz / denom2;
          bi[k] = -(ai[k] + ci[k]);
          if (k != nsoil - 1) {
            wdf = wdf2;
            wcnd = wcnd2;
            dsmdz = dsmdz2;
            ddz = ddz2;
          }
        }
      }
    }
    // SStep: src/noah/noah.c:2437-2480
#pragma unroll
    for (int k = 0; k < ML; k++) {
      if (k < nsoil) {
        rhstt[k] *= dt;
        ai[k] *= dt;
        bi[k] = 1.0 + bi[k] * dt;
        ci[k] *= dt;
        at(s.smflxv, k) = 0.0;
      }
    }
    // Rosr12 (src/noah/noah.c:1572-1620) in place: ci becomes p, rhstt becomes
    // delta; each entry is read before it is overwritten, so the arithmetic is
    // the reference's, which works on copies
#pragma unroll
    for (int k = 0; k < ML; k++)
      if (k == nsoil - 1) ci[k] = 0.0;
    ci[0] = -ci[0] / bi[0];
    rhstt[0] = rhstt[0] / bi[0];
#pragma unroll
    for (int k = 1; k < ML; k++) {
      if (k < nsoil) {
        const double inv = 1.0 / (bi[k] + ai[k] * ci[k - 1]);
        const double pk = -ci[k] * inv;
        rhstt[k] = (rhstt[k] - ai[k] * rhstt[k - 1]) * inv;
        ci[k] = pk;
      }
    }
#pragma unroll
    for (int k = 0; k < ML; k++)
      if (k == nsoil - 1) ci[k] = rhstt[k];
#pragma unroll
    for (int k = ML - 2; k >= 0; k--)
      if (k <= nsoil - 2) ci[k] = ci[k] * ci[k + 1] + rhstt[k];
    // the solution is added to sh2o; rhstt keeps the old values (sh2o0 of SStep)
#pragma unroll
    for (int k = 0; k < ML; k++) {
      if (k < nsoil) {
        rhstt[k] = sh2o[k];
        sh2o[k] += ci[k];
      }
    }
    // AdjSmProf: src/noah/noah.c:3081-3172 (ai holds sh2omid)
    {
      double wplus = 0.0;
#pragma unroll
      for (int k = ML - 1; k >= 0; k--) {
        if (k < nsoil) {
          const double ddz = (k != 0) ? zsoil[k - 1] - zsoil[k] : -zsoil[0];
          ai[k] = sh2o[k] + wplus / ddz;
          const double stot = ai[k] + sice[k];
          if (stot > smcmax) {
            smc[k] = smcmax;
            wplus = (stot - smcmax) * ddz;
          } else {
            const double sat = at(s.satdpth, k), sl = at(s.sldpth, k);
            double stotmin = (sat * smcmax + (sl - sat) * (smcmin + SH2OMIN)) / sl;
            stotmin = (stotmin > smcmax) ? smcmax : stotmin;
            stotmin = (stotmin < smcmin + SH2OMIN) ? (smcmin + SH2OMIN) : stotmin;
            if (stot < stotmin) {
              smc[k] = stotmin;
              wplus = (stot - stotmin) * ddz;
            } else {
              smc[k] = stot;
              wplus = 0.0;
            }
          }
          ai[k] = smc[k] - sice[k];
        }
      }
#pragma unroll
      for (int k = 0; k < ML; k++) {
        if (k < nsoil) {
          const double ddz = (k != 0) ? zsoil[k - 1] - zsoil[k] : -zsoil[0];
          sh2o[k] = ai[k] + wplus / ddz;
          const double stot = sh2o[k] + sice[k];
          wplus = (stot > smcmax) ? (stot - smcmax) * ddz : 0.0;
          smc[k] = (stot < smcmax) ? stot : smcmax;
          smc[k] = (smc[k] > smcmin + SH2OMIN) ? smc[k] : (smcmin + SH2OMIN);
          sh2o[k] = smc[k] - sice[k];
          sh2o[k] = (sh2o[k] > 0.0) ? sh2o[k] : 0.0;
        }
      }
      s.runoff3[p] = wplus / dt;
    }
    // layer fluxes, bottom up (smflxv[nsoil - 1] stays 0)
    double below = 0.0;
#pragma unroll
    for (int k = ML - 1; k >= 1; k--) {
      if (k <= nsoil - 1) {
        below = (sh2o[k] - rhstt[k]) * at(s.sldpth, k) / dt + at(s.runoff2_lyr, k) + at(e.et, k) + below;
        at(s.smflxv, k - 1) = below;
      }
    }
  }

#pragma unroll
  for (int k = 0; k < ML; k++) {
    if (k < nsoil) {
      at(s.smc, k) = smc[k];
      at(s.sh2o, k) = sh2o[k];
    }
  }
  // what the next model step's RHS reads
  e.sh2o0[p] = sh2o[0];
  e.fcr[p] = fcr;
  e.nwtbl[p] = nwtbl;
  const double ett = e.ett[p];
  const int nroot = e.nroot[p];
  double g = 0.0;
  if (ett > 0.0 && nwtbl <= nroot) {
    for (int j = (nwtbl <= 0 ? 0 : nwtbl - 1); j < nroot; j++) g += at(e.et, j);
    g = g / ett;
    g = (g > 1.0) ? 1.0 : g;
    g = (g < 0.0) ? 0.0 : g;
  }
  e.gwet[p] = g;
}

}  // namespace

struct pihm_noah_soil {
  pihm_ctx *ctx;
  int ne;
  SoilDev d;
  void *arena;
  double *scratch;     // [neq + 2 nr]: fluxes evaluated at the Summary state, carried flux saved
  int started;
};

namespace pihm {
// The soil column of a soil-water step, for the Noah land-surface step that
// shares it (noahlsm.hip): device order, layer arrays layer-major.
NoahSoilView noah_soil_view(pihm_noah_soil *s) {
  NoahSoilView v;
  v.ctx = s->ctx;
  v.ne = s->ne;
  v.nsoil = s->d.nsoil;
  v.sldpth = s->d.sldpth;
  v.smcmax = s->d.smcmax;
  v.smc = s->d.smc;
  v.sh2o = s->d.sh2o;
  return v;
}
}  // namespace pihm

namespace {

// caller-order host values of one field -> device order at dst (layers: ML
// per element, element-major on the host, layer-major on the device)
int upload(pihm_noah_soil *s, double *dst, const double *src, int layers) {
  const int ne = s->ne;
  std::vector<double> h((size_t)layers * ne);
  for (int p = 0; p < ne; p++) {
    const long ci = s->ctx->perm[p];
    for (int k = 0; k < layers; k++) h[(size_t)k * ne + p] = src[ci * layers + k];
  }
  PIHM_HIP_TRY(hipMemcpyAsync(dst, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, s->ctx->stream));
  PIHM_HIP_TRY(hipStreamSynchronize(s->ctx->stream));
  return PIHM_OK;
}

template <class T>
int download(pihm_noah_soil *s, const T *src, T *dst, int layers) {
  const int ne = s->ne;
  std::vector<T> h((size_t)layers * ne);
  PIHM_HIP_TRY(hipMemcpyAsync(h.data(), src, h.size() * sizeof(T), hipMemcpyDeviceToHost, s->ctx->stream));
  PIHM_HIP_TRY(hipStreamSynchronize(s->ctx->stream));
  for (int p = 0; p < ne; p++) {
    const long ci = s->ctx->perm[p];
    for (int k = 0; k < layers; k++) dst[ci * layers + k] = h[(size_t)k * ne + p];
  }
  return PIHM_OK;
}

int launch(pihm_noah_soil *s, double dt, const double *y_dev) {
  pihm_ctx *c = s->ctx;
  hipLaunchKernelGGL(k_noah_soil, dim3(div_up(s->ne, kBlock)), dim3(kBlock), 0, c->stream, s->ne, s->d, c->args.e,
                     y_dev, dt);
  PIHM_HIP_TRY(hipGetLastError());
  return PIHM_OK;
}

}  // namespace

extern "C" {

int pihm_noah_soil_create(pihm_ctx *ctx, const pihm_noah_soil_desc *desc, pihm_noah_soil **out) {
  if (!out) return PIHM_ERR_ARG;
  *out = nullptr;
  if (!ctx || !desc || ctx->nmem != 1 || !ctx->noah) return PIHM_ERR_ARG;
  if (!desc->nsoil || !desc->sldpth || !desc->smcmax || !desc->frzx) return PIHM_ERR_ARG;
  const int ne = ctx->ne;
  for (int i = 0; i < ne; i++) {
    const int n = desc->nsoil[i];
    if (n < 2 || n > ML) return PIHM_ERR_ARG;
    for (int k = 0; k < n; k++)
      if (!(desc->sldpth[(long)i * ML + k] > 0.0)) return PIHM_ERR_ARG;
  }
  pihm_noah_soil *s = new pihm_noah_soil();
  s->ctx = ctx;
  s->ne = ne;
  SoilDev &D = s->d;
  const size_t b1 = ((size_t)ne * sizeof(double) + 255) & ~size_t(255);
  const size_t bl = ((size_t)ML * ne * sizeof(double) + 255) & ~size_t(255);
  const size_t bi = ((size_t)ne * sizeof(int32_t) + 255) & ~size_t(255);
  double **one[] = {&D.smcmax, &D.frzx, &D.runoff3, &D.infil, &D.runoff2};
  double **lay[] = {&D.sldpth, &D.smc, &D.sh2o, &D.smflxv, &D.runoff2_lyr, &D.satdpth};
  const size_t bs = ((size_t)(ctx->neq + 2L * ctx->nr + 1) * sizeof(double) + 255) & ~size_t(255);
  const size_t total = 7 * b1 + 6 * bl + bi + bs;
  if (hipMalloc(&s->arena, total) != hipSuccess) { delete s; return PIHM_ERR_HIP; }
  char *q = (char *)s->arena;
  for (auto f : one) { *f = (double *)q; q += b1; }
  D.prev = (double *)q; q += 2 * b1;
  for (auto f : lay) { *f = (double *)q; q += bl; }
  D.nsoil = (int32_t *)q; q += bi;
  s->scratch = (double *)q;
  int rc = PIHM_OK;
  if (hipMemsetAsync(s->arena, 0, total, ctx->stream) != hipSuccess) rc = PIHM_ERR_HIP;
  std::vector<int32_t> hn(ne);
  for (int p = 0; p < ne; p++) hn[p] = desc->nsoil[ctx->perm[p]];
  if (!rc && hipMemcpyAsync(D.nsoil, hn.data(), ne * sizeof(int32_t), hipMemcpyHostToDevice, ctx->stream) != hipSuccess)
    rc = PIHM_ERR_HIP;
  if (!rc && hipStreamSynchronize(ctx->stream) != hipSuccess) rc = PIHM_ERR_HIP;
  if (!rc) rc = upload(s, D.sldpth, desc->sldpth, ML);
  if (!rc) rc = upload(s, D.smcmax, desc->smcmax, 1);
  if (!rc) rc = upload(s, D.frzx, desc->frzx, 1);
  if (rc) { pihm_noah_soil_destroy(s); return rc; }
  *out = s;
  return PIHM_OK;
}

void pihm_noah_soil_destroy(pihm_noah_soil *s) {
  if (!s) return;
  if (s->arena) (void)hipFree(s->arena);
  delete s;
}

int pihm_noah_soil_set_state(pihm_noah_soil *s, const double *smc, const double *sh2o, const double *fcr) {
  if (!s || !smc || !sh2o) return PIHM_ERR_ARG;
  int rc;
  if ((rc = upload(s, s->d.smc, smc, ML))) return rc;
  if ((rc = upload(s, s->d.sh2o, sh2o, ML))) return rc;
  ElemDev &E = s->ctx->args.e;
  if (fcr && (rc = upload(s, E.fcr, fcr, 1))) return rc;
  // the RHS's top-layer water: sh2o[0] of each element
  std::vector<double> top(s->ne);
  for (int i = 0; i < s->ne; i++) top[i] = sh2o[(long)i * ML];
  return upload(s, E.sh2o0, top.data(), 1);
}

int pihm_noah_soil_get_state(pihm_noah_soil *s, double *smc, double *sh2o, double *fcr, int32_t *nwtbl,
                             double *satdpth, double *runoff2_lyr, double *smflxv, double *runoff3) {
  if (!s) return PIHM_ERR_ARG;
  const ElemDev &E = s->ctx->args.e;
  int rc = PIHM_OK;
  if (!rc && smc) rc = download(s, s->d.smc, smc, ML);
  if (!rc && sh2o) rc = download(s, s->d.sh2o, sh2o, ML);
  if (!rc && fcr) rc = download(s, (const double *)E.fcr, fcr, 1);
  if (!rc && nwtbl) rc = download(s, (const int32_t *)E.nwtbl, nwtbl, 1);
  if (!rc && satdpth) rc = download(s, s->d.satdpth, satdpth, ML);
  if (!rc && runoff2_lyr) rc = download(s, s->d.runoff2_lyr, runoff2_lyr, ML);
  if (!rc && smflxv) rc = download(s, s->d.smflxv, smflxv, ML);
  if (!rc && runoff3) rc = download(s, s->d.runoff3, runoff3, 1);
  return rc;
}

int pihm_noah_soil_step(pihm_noah_soil *s, double dt, const double *y_dev, const double *infil,
                        const double *runoff2) {
  if (!s || !y_dev || !infil || !runoff2 || !(dt > 0.0) || !s->ctx->noah) return PIHM_ERR_ARG;
  int rc;
  if ((rc = upload(s, s->d.infil, infil, 1))) return rc;
  if ((rc = upload(s, s->d.runoff2, runoff2, 1))) return rc;
  return launch(s, dt, y_dev);
}

int pihm_noah_soil_start(pihm_noah_soil *s, const double *y_dev) {
  if (!s || !y_dev) return PIHM_ERR_ARG;
  hipLaunchKernelGGL(k_soil_prev, dim3(div_up(s->ne, kBlock)), dim3(kBlock), 0, s->ctx->stream, s->ne, s->d, y_dev);
  PIHM_HIP_TRY(hipGetLastError());
  s->started = 1;
  return PIHM_OK;
}

int pihm_noah_soil_summary(pihm_noah_soil *s, double dt, const double *y_dev) {
  if (!s || !y_dev || !(dt > 0.0) || !s->ctx->noah || !s->started) return PIHM_ERR_ARG;
  pihm_ctx *c = s->ctx;
  int rc = pihm_ctx_eval_fluxes(c, y_dev, s->scratch, s->scratch + c->neq);
  if (rc) return rc;
  hipLaunchKernelGGL(k_mass_balance_noah, dim3(div_up(s->ne, kBlock)), dim3(kBlock), 0, c->stream, s->ne, s->d,
                     c->args.e, y_dev, (const double *)c->args.rec.sub, dt);
  PIHM_HIP_TRY(hipGetLastError());
  return launch(s, dt, y_dev);
}

int pihm_noah_soil_time(pihm_noah_soil *s, double dt, const double *y_dev, int iters, double *ms) {
  if (!s || !y_dev || iters < 1 || !ms || !(dt > 0.0) || !s->ctx->noah) return PIHM_ERR_ARG;
  hipStream_t st = s->ctx->stream;
  hipEvent_t ev[2];
  for (auto &e : ev) PIHM_HIP_TRY(hipEventCreate(&e));
  int rc = PIHM_OK;
  float t = 0.0f;
  if (hipEventRecord(ev[0], st) != hipSuccess) rc = PIHM_ERR_HIP;
  for (int it = 0; !rc && it < iters; it++) rc = launch(s, dt, y_dev);
  if (!rc && (hipEventRecord(ev[1], st) != hipSuccess || hipEventSynchronize(ev[1]) != hipSuccess ||
              hipEventElapsedTime(&t, ev[0], ev[1]) != hipSuccess))
    rc = PIHM_ERR_HIP;
  for (auto &e : ev) (void)hipEventDestroy(e);
  if (!rc) *ms = (double)t / iters;
  return rc;
}

}  // extern "C"
