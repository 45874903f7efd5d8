// output.hip - the reference's print variables, summed in HBM, averaged and
// written in its formats at print time (UpdPrintVar src/print.c:190-211,
// PrintData src/print.c:545-603, PrintNow src/print.c:814-856, the restart
// file src/print.c:250-300).
//
// Each variable's running sum lives on the device in device order; an update
// is one add per entry (the same double additions, in the same order, as the
// reference's host loop), and only a print moves data: the sums come back,
// are divided by the update count on the host and written in caller order.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <ctime>
#include <string>
#include <vector>

#include "../../include/pihm_output.h"
#include "common.hpp"
#include "ctx.hpp"
#include "nvector_impl.hpp"
#include "physics.hpp"

using namespace pihm;

namespace pihm {
const double *lsm_field(const pihm_lsm *l, int which);   // lsm.hip
}

namespace {

__global__ __launch_bounds__(kBlock) void k_out_add(long n, const double *src, double *sum) {
  const long i = (long)blockIdx.x * kBlock + threadIdx.x;
  if (i < n) sum[i] += src[i];
}

// MassBalance (src/update.c:107-160): the infiltration the step's storage
// change and outflows imply, with the RHS's fluxes at the step's final state
// (subsurface outflow, EtExtract's partition, src/hydrol.c:33-97); prev holds
// unsat, gw of the previous model step (ws0), replaced by this step's.
__global__ __launch_bounds__(kBlock) void k_mass_balance(int ne, ElemDev e, const double *y, double *prev,
                                                        const double *sub, double dt, double *infil) {
  const int p = blockIdx.x * kBlock + threadIdx.x;
  if (p >= ne) return;
  const double depth = e.depth[p], area = e.area[p];
  const double u = y[ne + p], g = y[2L * ne + p];
  double soilw0 = prev[ne + p] + prev[p];
  soilw0 = (soilw0 > depth) ? depth : soilw0;
  soilw0 = (soilw0 < 0.0) ? 0.0 : soilw0;
  double soilw1 = g + u;
  soilw1 = (soilw1 > depth) ? depth : soilw1;
  soilw1 = (soilw1 < 0.0) ? 0.0 : soilw1;
  double subrunoff = 0.0;
  for (int j = 0; j < 3; j++) subrunoff += sub[(long)j * ne + p] / area;
  // EtExtract at the final state (the clamped surface and groundwater)
  const double surfh = surf_h((y[p] > 0.0) ? y[p] : 0.0);
  const double gc = (g > 0.0) ? g : 0.0;
  const double edir = e.edir[p], ett = e.ett[p];
  double edir_u = 0.0, edir_g = 0.0, ett_u = 0.0, ett_g = 0.0;
  if (surfh >= DEPRSTG) {
  } else if (gc > depth - e.dinf[p]) {
    edir_g = edir;
  } else {
    edir_u = edir;
  }
  if (gc > depth - e.rzd[p]) ett_g = ett;
  else ett_u = ett;
  double inf = (soilw1 - soilw0) * e.porosity[p] / dt + subrunoff + edir_u + edir_g + ett_u + ett_g;
  if (inf < 0.0) inf = 0.0;
  infil[p] = inf;
  prev[p] = u;
  prev[ne + p] = g;
}

struct OutVar {
  int kind, intvl, upd;
  bool river, flux;
  long n;
  double *sum;           // device, device order
  long counter;
  FILE *dat, *txt;
};

// src/map_output.c:30-255
const char *ext_of(int kind, char *buf) {
  static const char *base[] = {"surf", "unsat", "gw", "stage", "rivgw", "snow", "is",
                               "infil", "recharge", "ec", "ett", "edir"};
  if (kind < PIHM_OUT_RIVFLX0) return base[kind];
  if (kind < PIHM_OUT_SUBFLX0) { std::snprintf(buf, 32, "rivflx%d", kind - PIHM_OUT_RIVFLX0); return buf; }
  if (kind < PIHM_OUT_SURFFLX0) { std::snprintf(buf, 32, "subflx%d", kind - PIHM_OUT_SUBFLX0); return buf; }
  std::snprintf(buf, 32, "surfflx%d", kind - PIHM_OUT_SURFFLX0);
  return buf;
}

bool is_ls(int kind) {
  return kind == PIHM_OUT_SNOW || kind == PIHM_OUT_CMC || kind == PIHM_OUT_EC || kind == PIHM_OUT_ETT ||
         kind == PIHM_OUT_EDIR;
}

bool is_river(int kind) {
  return kind == PIHM_OUT_RIVSTG || kind == PIHM_OUT_RIVGW ||
         (kind >= PIHM_OUT_RIVFLX0 && kind < PIHM_OUT_SUBFLX0);
}

bool is_flux(int kind) {
  return kind == PIHM_OUT_INFIL || kind == PIHM_OUT_RECHARGE || kind >= PIHM_OUT_RIVFLX0;
}

// PrintNow (src/print.c:814-856) on the UTC calendar of t
bool print_now(int intvl, long lapse, const std::tm &tm) {
  switch (intvl) {
    case 0: return false;
    case PIHM_OUT_YEARLY: return tm.tm_mon == 0 && tm.tm_mday == 1 && tm.tm_hour == 0 && tm.tm_min == 0;
    case PIHM_OUT_MONTHLY: return tm.tm_mday == 1 && tm.tm_hour == 0 && tm.tm_min == 0;
    case PIHM_OUT_DAILY: return tm.tm_hour == 0 && tm.tm_min == 0;
    case PIHM_OUT_HOURLY: return tm.tm_min == 0;
    default: return intvl > 0 && lapse % intvl == 0;
  }
}

}  // namespace

struct pihm_output {
  pihm_ctx *ctx;
  pihm_lsm *lsm;
  std::vector<OutVar> v;
  bool any_flux, need_mb, started;
  double *scratch;       // RHS output and the carried-flux copy of a flux evaluation
  double *prev;          // MassBalance: unsat, gw of the previous model step [2 ne]
  double *infil_mb;      // MassBalance's infiltration [ne]
  std::vector<double> host;
};

extern "C" {

void pihm_output_destroy(pihm_output *o) {
  if (!o) return;
  for (auto &x : o->v) {
    if (x.sum) (void)hipFree(x.sum);
    if (x.dat) std::fclose(x.dat);
    if (x.txt) std::fclose(x.txt);
  }
  if (o->scratch) (void)hipFree(o->scratch);
  if (o->prev) (void)hipFree(o->prev);
  if (o->infil_mb) (void)hipFree(o->infil_mb);
  delete o;
}

int pihm_output_start(pihm_output *o, N_Vector yv) {
  if (!o || !yv || N_VGetLength_Hip(yv) != o->ctx->neq) return PIHM_ERR_ARG;
  if (o->prev) {
    const double *y = N_VGetDeviceArrayPointer_Hip(yv);
    const int ne = o->ctx->ne;
    // prev = (unsat, gw) of the initial state: ws0 = ws after initialisation
    PIHM_HIP_TRY(hipMemcpyAsync(o->prev, y + ne, 2L * ne * sizeof(double), hipMemcpyDeviceToDevice,
                                o->ctx->stream));
  }
  o->started = true;
  return PIHM_OK;
}

int pihm_output_create(pihm_ctx *ctx, pihm_lsm *lsm, const char *prefix, const pihm_output_var *vars, int nvars,
                       int ascii, pihm_output **out) {
  if (!out) return PIHM_ERR_ARG;
  *out = nullptr;
  if (!ctx || ctx->nmem != 1 || !prefix || (nvars > 0 && !vars) || nvars < 0) return PIHM_ERR_ARG;
  for (int i = 0; i < nvars; i++) {
    const int k = vars[i].kind;
    if (k < 0 || k >= PIHM_OUT_NKIND || vars[i].intvl < PIHM_OUT_HOURLY) return PIHM_ERR_ARG;
    if (is_ls(k) && (k == PIHM_OUT_SNOW || k == PIHM_OUT_CMC || k == PIHM_OUT_EC) && !lsm) return PIHM_ERR_ARG;
  }
  pihm_output *o = new pihm_output();
  o->ctx = ctx;
  o->lsm = lsm;
  o->any_flux = false;
  o->need_mb = false;
  o->started = false;
  o->scratch = nullptr;
  o->prev = nullptr;
  o->infil_mb = nullptr;
  long nmax = 1;
  for (int i = 0; i < nvars; i++) {
    OutVar x{};
    x.kind = vars[i].kind;
    x.intvl = vars[i].intvl;
    x.upd = is_ls(x.kind) ? PIHM_OUT_LS_STEP : PIHM_OUT_HYDROL_STEP;
    x.river = is_river(x.kind);
    x.flux = is_flux(x.kind);
    x.n = x.river ? ctx->nr : ctx->ne;
    o->any_flux |= x.flux;
    o->need_mb |= x.kind == PIHM_OUT_INFIL;
    nmax = std::max(nmax, x.n);
    char buf[32];
    const std::string name = std::string(prefix) + "." + ext_of(x.kind, buf);
    o->v.push_back(x);
    OutVar &y = o->v.back();
    if (hipMalloc(&y.sum, std::max(1L, y.n) * sizeof(double)) != hipSuccess ||
        hipMemset(y.sum, 0, std::max(1L, y.n) * sizeof(double)) != hipSuccess) {
      pihm_output_destroy(o);
      return PIHM_ERR_HIP;
    }
    y.dat = std::fopen((name + ".dat").c_str(), "wb");
    y.txt = ascii ? std::fopen((name + ".txt").c_str(), "w") : nullptr;
    if (!y.dat || (ascii && !y.txt)) {
      pihm_output_destroy(o);
      return PIHM_ERR_ARG;
    }
  }
  if (o->any_flux &&
      hipMalloc(&o->scratch, ((size_t)ctx->neq + 2L * ctx->nr + 1) * sizeof(double)) != hipSuccess) {
    pihm_output_destroy(o);
    return PIHM_ERR_HIP;
  }
  if (o->need_mb && (hipMalloc(&o->prev, 2L * std::max(1, ctx->ne) * sizeof(double)) != hipSuccess ||
                     hipMalloc(&o->infil_mb, std::max(1, ctx->ne) * sizeof(double)) != hipSuccess)) {
    pihm_output_destroy(o);
    return PIHM_ERR_HIP;
  }
  o->host.resize(nmax);
  *out = o;
  return PIHM_OK;
}

int pihm_output_update(pihm_output *o, int step, N_Vector yv) {
  if (!o || (step != PIHM_OUT_HYDROL_STEP && step != PIHM_OUT_LS_STEP)) return PIHM_ERR_ARG;
  pihm_ctx *c = o->ctx;
  const double *y = nullptr;
  if (step == PIHM_OUT_HYDROL_STEP) {
    if (!yv || N_VGetLength_Hip(yv) != c->neq) return PIHM_ERR_ARG;
    y = N_VGetDeviceArrayPointer_Hip(yv);
  }
  bool flux_done = false;
  const int ne = c->ne, nr = c->nr;
  for (auto &x : o->v) {
    if (x.upd != step) continue;
    if (x.flux && !flux_done) {
      const int rc = pihm_ctx_eval_fluxes(c, y, o->scratch, o->scratch + c->neq);
      if (rc) return rc;
      flux_done = true;
      if (o->need_mb) {
        if (!o->started) return PIHM_ERR_ARG;   // the previous step's storage is needed
        hipLaunchKernelGGL(k_mass_balance, dim3(div_up(ne, kBlock)), dim3(kBlock), 0, c->stream, ne, c->args.e, y,
                           o->prev, c->args.rec.sub, c->args.dt, o->infil_mb);
      }
    }
    const double *src = nullptr;
    const int k = x.kind;
    if (k == PIHM_OUT_SURF) src = y;
    else if (k == PIHM_OUT_UNSAT) src = y + ne;
    else if (k == PIHM_OUT_GW) src = y + 2L * ne;
    else if (k == PIHM_OUT_RIVSTG) src = y + 3L * ne;
    else if (k == PIHM_OUT_RIVGW) src = y + 3L * ne + nr;
    else if (k == PIHM_OUT_SNOW) src = lsm_field(o->lsm, 1);
    else if (k == PIHM_OUT_CMC) src = lsm_field(o->lsm, 0);
    else if (k == PIHM_OUT_EC) src = lsm_field(o->lsm, 2);
    else if (k == PIHM_OUT_ETT) src = c->args.e.ett;
    else if (k == PIHM_OUT_EDIR) src = c->args.e.edir;
    else if (k == PIHM_OUT_INFIL) src = o->infil_mb;
    else if (k == PIHM_OUT_RECHARGE) src = c->args.rec.rechg;
    else if (k < PIHM_OUT_SUBFLX0) src = c->args.r.rf + (long)(k - PIHM_OUT_RIVFLX0) * nr;
    else if (k < PIHM_OUT_SURFFLX0) src = c->args.rec.sub + (long)(k - PIHM_OUT_SUBFLX0) * ne;
    else src = c->args.rec.ovl + (long)(k - PIHM_OUT_SURFFLX0) * ne;
    if (!src) return PIHM_ERR_ARG;
    if (x.n > 0)
      hipLaunchKernelGGL(k_out_add, dim3(div_up(x.n, kBlock)), dim3(kBlock), 0, c->stream, x.n, src, x.sum);
    x.counter++;
  }
  PIHM_HIP_TRY(hipGetLastError());
  return PIHM_OK;
}

int pihm_output_print(pihm_output *o, double t, double lapse) {
  if (!o) return PIHM_ERR_ARG;
  pihm_ctx *c = o->ctx;
  const std::time_t tt = (std::time_t)t;
  std::tm tm{};
  gmtime_r(&tt, &tm);
  char stamp[32];
  std::strftime(stamp, sizeof(stamp), "%Y-%m-%d %H:%M", &tm);
  std::vector<double> vals;
  for (auto &x : o->v) {
    if (!print_now(x.intvl, (long)lapse, tm)) continue;
    vals.assign(x.n, 0.0);
    if (x.n > 0) {
      PIHM_HIP_TRY(hipMemcpyAsync(o->host.data(), x.sum, x.n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
      PIHM_HIP_TRY(hipStreamSynchronize(c->stream));
    }
    // caller order; sum / count as the reference divides (src/print.c:566-592)
    for (long p = 0; p < x.n; p++) {
      const long q = x.river ? caller_river(c, (int)p) : caller_elem(c, (int)p);
      if (q < 0) continue;
      vals[q] = (x.counter > 0) ? o->host[p] / (double)x.counter : o->host[p];
    }
    if (x.txt) {
      std::fprintf(x.txt, "\"%s\"", stamp);
      for (long j = 0; j < x.n; j++) std::fprintf(x.txt, "\t%lf", vals[j]);
      std::fprintf(x.txt, "\n");
      std::fflush(x.txt);
    }
    const double outtime = t;
    std::fwrite(&outtime, sizeof(double), 1, x.dat);
    if (x.n > 0) std::fwrite(vals.data(), sizeof(double), x.n, x.dat);
    std::fflush(x.dat);
    if (x.n > 0) PIHM_HIP_TRY(hipMemsetAsync(x.sum, 0, x.n * sizeof(double), c->stream));
    x.counter = 0;
  }
  return PIHM_OK;
}

int pihm_output_write_ic(pihm_output *o, const char *path, N_Vector yv) {
  if (!o || !path || !yv || N_VGetLength_Hip(yv) != o->ctx->neq) return PIHM_ERR_ARG;
  pihm_ctx *c = o->ctx;
  const int ne = c->ne, nr = c->nr;
  std::vector<double> y(c->neq), cmc(ne, 0.0), sneqv(ne, 0.0);
  int rc = pihm_ctx_state_from_device(c, yv, y.data());
  if (!rc && o->lsm) rc = pihm_lsm_get_state(o->lsm, cmc.data(), sneqv.data());
  if (rc) return rc;
  FILE *f = std::fopen(path, "wb");
  if (!f) return PIHM_ERR_ARG;
  for (int i = 0; i < ne; i++) {
    const double rec[5] = {cmc[i], sneqv[i], y[i], y[(long)ne + i], y[2L * ne + i]};
    std::fwrite(rec, sizeof(double), 5, f);
  }
  for (int r = 0; r < nr; r++) {
    const double rec[2] = {y[3L * ne + r], y[3L * ne + nr + r]};
    std::fwrite(rec, sizeof(double), 2, f);
  }
  const int bad = std::fclose(f);
  return bad ? PIHM_ERR_ARG : PIHM_OK;
}

}  // extern "C"
