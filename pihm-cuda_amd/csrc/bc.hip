// bc.hip - boundary-condition time series applied on the device.
//
// Two launches on the context's stream per apply: one thread per series
// interpolates it (IntrplForc, src/forcing.c:482-530) into a small table, then
// one thread per element edge and river segment copies its series' value
// (ApplyElemBc, src/forcing.c:127-177; ApplyRiverBc, src/forcing.c:214-244).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "../../include/pihm_bc.h"
#include "common.hpp"
#include "ctx.hpp"
#include "tsinterp.hpp"

using namespace pihm;

namespace {

__global__ __launch_bounds__(kBlock) void k_bc_interp(int nts_e, int nrec_e, const double *ft_e, const double *v_e,
                                                      int nts_r, int nrec_r, const double *ft_r, const double *v_r,
                                                      double t, double *val) {
  const int k = blockIdx.x * kBlock + threadIdx.x;
  if (k < nts_e) {
    ts_interp<1>(ft_e + (long)k * nrec_e, v_e + (long)k * nrec_e, nrec_e, t, val + k, 1);
  } else if (k < nts_e + nts_r) {
    const int q = k - nts_e;
    ts_interp<1>(ft_r + (long)q * nrec_r, v_r + (long)q * nrec_r, nrec_r, t, val + k, 1);
  }
}

// code > 0: series code-1 (head); code < 0: series -code-1 (flux); 0: unchanged.
// Element edges first (edge-major, device order), then river segments, whose
// series sit after the element series in val.
__global__ __launch_bounds__(kBlock) void k_bc_scatter(long n_edge, int nr, const int32_t *code_e,
                                                       const int32_t *code_r, int nts_e, const double *val,
                                                       double *ebc, double *rbc) {
  const long i = (long)blockIdx.x * kBlock + threadIdx.x;
  if (i < n_edge) {
    const int c = code_e[i];
    if (c != 0) ebc[i] = val[(c > 0 ? c : -c) - 1];
  } else if (i < n_edge + nr) {
    const long r = i - n_edge;
    const int c = code_r[r];
    if (c != 0) rbc[r] = val[nts_e + (c > 0 ? c : -c) - 1];
  }
}

bool times_increase(const pihm_ts *s) {
  for (int k = 0; k < s->nts; k++)
    for (int r = 1; r < s->nrec; r++)
      if (!(s->ftime[(long)k * s->nrec + r] > s->ftime[(long)k * s->nrec + r - 1])) return false;
  return true;
}

}  // namespace

struct pihm_bc {
  pihm_ctx *ctx;
  int nts_e, nrec_e, nts_r, nrec_r;
  long n_edge;
  int nr;
  int32_t *code_e, *code_r;
  double *ft_e, *v_e, *ft_r, *v_r, *val;
  double t_lo, t_hi;    // the time range every series covers
  void *arena;
};

namespace pihm {
pihm_ctx *bc_context(const pihm_bc *b) { return b ? b->ctx : nullptr; }
}

extern "C" {

int pihm_bc_create(pihm_ctx *ctx, const int32_t *elem_type, const pihm_ts *ets, const int32_t *river_type,
                   const pihm_ts *rts, pihm_bc **out) {
  if (!out) return PIHM_ERR_ARG;
  *out = nullptr;
  if (!ctx || ctx->nmem != 1) return PIHM_ERR_ARG;
  const int ne = ctx->ne, nr = ctx->nr;
  const int nts_e = ets ? ets->nts : 0, nts_r = rts ? rts->nts : 0;
  for (const pihm_ts *s : {ets, rts})
    if (s && (s->nts < 0 || (s->nts > 0 && (s->nrec < 2 || !s->ftime || !s->value || !times_increase(s)))))
      return PIHM_ERR_ARG;
  std::vector<int32_t> ce(3L * ne, 0), cr(nr > 0 ? nr : 1, 0);
  if (elem_type)
    for (int p = 0; p < ne; p++)
      for (int j = 0; j < 3; j++) {
        const int c = elem_type[3L * ctx->perm[p] + j];
        if (c > nts_e || -c > nts_e) return PIHM_ERR_ARG;
        ce[(long)j * ne + p] = c;
      }
  if (river_type)
    for (int r = 0; r < nr; r++) {
      const long q = caller_river(ctx, r);
      const int c = q >= 0 ? river_type[q] : 0;
      if (c > nts_r || -c > nts_r) return PIHM_ERR_ARG;
      cr[r] = c;
    }
  pihm_bc *b = new pihm_bc();
  b->ctx = ctx;
  b->nts_e = nts_e;
  b->nrec_e = nts_e ? ets->nrec : 0;
  b->nts_r = nts_r;
  b->nrec_r = nts_r ? rts->nrec : 0;
  b->n_edge = 3L * ne;
  b->nr = nr;
  b->t_lo = -HUGE_VAL;
  b->t_hi = HUGE_VAL;
  for (const pihm_ts *s : {ets, rts})
    if (s)
      for (int k = 0; k < s->nts; k++) {
        b->t_lo = std::max(b->t_lo, s->ftime[(long)k * s->nrec]);
        b->t_hi = std::min(b->t_hi, s->ftime[(long)k * s->nrec + s->nrec - 1]);
      }
  auto al = [](size_t n) { return (n + 255) & ~size_t(255); };
  const size_t se = (size_t)b->nts_e * b->nrec_e * sizeof(double), sr = (size_t)b->nts_r * b->nrec_r * sizeof(double);
  const size_t total = al(3L * ne * 4) + al(cr.size() * 4) + 2 * al(se) + 2 * al(sr) + al((nts_e + nts_r + 1) * 8);
  if (hipMalloc(&b->arena, total) != hipSuccess) { delete b; return PIHM_ERR_HIP; }
  char *base = (char *)b->arena;
  b->code_e = (int32_t *)base; base += al(3L * ne * 4);
  b->code_r = (int32_t *)base; base += al(cr.size() * 4);
  b->ft_e = (double *)base; base += al(se);
  b->v_e = (double *)base; base += al(se);
  b->ft_r = (double *)base; base += al(sr);
  b->v_r = (double *)base; base += al(sr);
  b->val = (double *)base;
  bool ok = hipMemcpy(b->code_e, ce.data(), 3L * ne * 4, hipMemcpyHostToDevice) == hipSuccess &&
            hipMemcpy(b->code_r, cr.data(), cr.size() * 4, hipMemcpyHostToDevice) == hipSuccess;
  if (ok && se)
    ok = hipMemcpy(b->ft_e, ets->ftime, se, hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(b->v_e, ets->value, se, hipMemcpyHostToDevice) == hipSuccess;
  if (ok && sr)
    ok = hipMemcpy(b->ft_r, rts->ftime, sr, hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(b->v_r, rts->value, sr, hipMemcpyHostToDevice) == hipSuccess;
  if (!ok) { pihm_bc_destroy(b); return PIHM_ERR_HIP; }
  *out = b;
  return PIHM_OK;
}

void pihm_bc_destroy(pihm_bc *b) {
  if (!b) return;
  if (b->arena) (void)hipFree(b->arena);
  delete b;
}

static void launch_apply(pihm_bc *b, double t) {
  const int nts = b->nts_e + b->nts_r;
  pihm_ctx *c = b->ctx;
  hipLaunchKernelGGL(k_bc_interp, dim3(div_up(nts, kBlock)), dim3(kBlock), 0, c->stream, b->nts_e, b->nrec_e,
                     b->ft_e, b->v_e, b->nts_r, b->nrec_r, b->ft_r, b->v_r, t, b->val);
  const long n = b->n_edge + b->nr;
  hipLaunchKernelGGL(k_bc_scatter, dim3((unsigned)div_up(n, (long)kBlock)), dim3(kBlock), 0, c->stream, b->n_edge,
                     b->nr, b->code_e, b->code_r, b->nts_e, b->val, c->args.e.ebc, c->args.r.bc);
}

int pihm_bc_apply(pihm_bc *b, double t) {
  if (!b) return PIHM_ERR_ARG;
  if (b->nts_e + b->nts_r == 0) return PIHM_OK;
  if (t < b->t_lo || t > b->t_hi) return PIHM_ERR_ARG;
  launch_apply(b, t);
  PIHM_HIP_TRY(hipGetLastError());
  return PIHM_OK;
}

int pihm_bc_time(pihm_bc *b, double t, int iters, double *ms_per_apply) {
  if (!b || iters <= 0 || !ms_per_apply || b->nts_e + b->nts_r == 0) return PIHM_ERR_ARG;
  if (t < b->t_lo || t > b->t_hi) return PIHM_ERR_ARG;
  hipEvent_t ev[2];
  for (auto &e : ev) PIHM_HIP_TRY(hipEventCreate(&e));
  PIHM_HIP_TRY(hipEventRecord(ev[0], b->ctx->stream));
  for (int it = 0; it < iters; it++) launch_apply(b, t);
  PIHM_HIP_TRY(hipGetLastError());
  PIHM_HIP_TRY(hipEventRecord(ev[1], b->ctx->stream));
  PIHM_HIP_TRY(hipEventSynchronize(ev[1]));
  float ms = 0;
  PIHM_HIP_TRY(hipEventElapsedTime(&ms, ev[0], ev[1]));
  for (auto &e : ev) (void)hipEventDestroy(e);
  *ms_per_apply = ms / iters;
  return PIHM_OK;
}

}  // extern "C"
