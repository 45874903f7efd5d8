// bdf.cpp - variable-order BDF (Nordsieck form) with Newton/SPGMR iteration.
//
// The integrator the MM-PIHM glue drives (src/ode.c:413-661): CVODE 2.9's
// BDF/Newton/SPGMR method.  The host-vector step logic is a close
// restatement of the vendored CVODE routines (same locals, constants and
// control flow; file:line at each routine) so that step counts and
// trajectories behave like the reference's; stability-limit detection is not
// performed.  Those routines are SUNDIALS code under the BSD-3 licence of
// Lawrence Livermore National Security (Copyright (c) 2002-2016, LLNS;
// UCRL-CODE-155951); its full text and disclaimer are in NOTICE.SUNDIALS at
// the package root.
//
// Every vector operation goes through the N_Vector operation table.  When
// the vectors are device vectors (pihm_nvector.h), groups of operations are
// fused into one launch each; the fused kernels round exactly like the
// operation sequences they replace.
#include <hip/hip_runtime_api.h>

#include <cfloat>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <new>

#include "../../include/pihm_cvode.h"
#include "bdf_core.hpp"
#include "nvector_impl.hpp"

namespace {

using namespace pihm::bdf;

struct Mem {
  PihmRhsFn f = nullptr;
  void *ud = nullptr;
  int qmax = 5;
  long mxstep = MXSTEP_DEFAULT;
  double hin = 0.0, hmin = 0.0, hmax_inv = 0.0;
  int maxcor = NLS_MAXCOR, maxnef = MXNEF, maxncf = MXNCF;
  double nlscoef = CORTES;
  double reltol = 0.0, abstol = 0.0;
  bool tol_set = false, malloc_done = false, ls_set = false, sldet = false;
  bool tstopset = false;
  double tstop = 0.0;
  int maxl = 5;
  double eplifac = 0.05;
  // right preconditioning (PihmCVSpilsSetPreconditioner)
  int pretype = 0;
  PihmPrecSetupFn pset = nullptr;
  PihmPrecSolveFn psolve = nullptr;
  bool jcur = false;
  long nstlp = 0, nstlpre = 0, npe = 0, nps = 0, nsetups = 0;
  double delta_ls = 0.0;
  // step record and replay (PihmCVodeSetStepRecord / SetStepReplay)
  double *rec_h = nullptr;
  int *rec_q = nullptr, *rec_it = nullptr;   // rec_it: PIHM_REC_ITERS counts per step
  long rec_cap = 0, *rec_n = nullptr;
  const double *rp_h = nullptr;
  const int *rp_q = nullptr, *rp_it = nullptr;
  long rp_n = 0;
  int att_it[PIHM_REC_ITERS] = {};   // Newton and Krylov counts of the step attempt in progress
  bool fused = false;
  // Krylov iterations replayed from graphs (device vectors on a non-null
  // stream and an RHS declared capturable, PihmCVodeSetGraphRhs)
  bool graph_rhs = false, graph_off = false;
  hipGraphExec_t gx[MAXL_CAP] = {};
  hipStream_t gx_stream = nullptr;   // the stream the executable graphs are launched on
  double *kry = nullptr;        // device: copies of the Krylov sums and gamma

  N_Vector zn[6] = {};
  N_Vector ewt = nullptr, y = nullptr, acor = nullptr, tempv = nullptr, ftemp = nullptr;
  N_Vector V[MAXL_CAP + 1] = {};
  N_Vector xcor = nullptr, vtemp = nullptr, ytemp = nullptr, xs = nullptr;
  N_Vector ycur = nullptr, fcur = nullptr;
  long neq = 0;
  double sqrtN = 1.0;

  double tn = 0, h = 0, hu = 0, hprime = 0, hscale = 0, eta = 1, etamax = ETAMX1;
  double etaq = 0, etaqm1 = 0, etaqp1 = 0, next_h = 0, h0u = 0, tretlast = 0, toutc = 0;
  double tolsf = 1.0;
  double tau[7] = {}, tq[6] = {}, l[6] = {};
  double rl1 = 0, gamma = 0, gammap = 0, gamrat = 1, crate = 1, acnrm = 0, saved_tq5 = 0;
  double ynrm_cached = 0.0;
  int q = 1, qprime = 1, next_q = 0, qu = 0, L = 2, qwait = 2, indx_acor = 5;
  long nst = 0, nfe = 0, ncfn = 0, netf = 0, nni = 0, nli = 0, nfes = 0, nhnil = 0, nscon = 0;
  int mnewt = 0;
  double Hes[MAXL_CAP + 1][MAXL_CAP] = {};
  double givens[2 * MAXL_CAP] = {};
  double yg[MAXL_CAP + 1] = {};
};

// ---------------------------------------------------------------- vector ops
inline void VLin(double a, N_Vector x, double b, N_Vector y, N_Vector z) { x->ops->nvlinearsum(a, x, b, y, z); }
inline void VConst(double c, N_Vector z) { z->ops->nvconst(c, z); }
inline void VScale(double c, N_Vector x, N_Vector z) { x->ops->nvscale(c, x, z); }
inline void VProd(N_Vector x, N_Vector y, N_Vector z) { x->ops->nvprod(x, y, z); }
inline void VDiv(N_Vector x, N_Vector y, N_Vector z) { x->ops->nvdiv(x, y, z); }
inline double VDot(N_Vector x, N_Vector y) { return x->ops->nvdotprod(x, y); }
inline double VWrms(N_Vector x, N_Vector w) { return x->ops->nvwrmsnorm(x, w); }
inline N_Vector VClone(N_Vector w) { return w->ops->nvclone(w); }
inline void VDestroy(N_Vector v) { if (v) v->ops->nvdestroy(v); }

bool is_hip(N_Vector v) { return v->ops->nvgetvectorid == N_VGetVectorID_Hip; }

// cv_setupNonNull (cvode_spgmr.c CVSpgmr / CVSpilsSetPreconditioner)
// The recorded Newton and Krylov counts of the step being replayed, or null.
const int *replay_iters(const Mem *m) {
  return (m->rp_it && m->nst < m->rp_n) ? m->rp_it + m->nst * PIHM_REC_ITERS : nullptr;
}
constexpr int FREE = -100;   // no forced Krylov count
// Krylov count forced for Newton iteration mi of a replayed step (FREE past the record)
int forced_krylov(const int *fi, int mi) { return (fi && mi + 1 < PIHM_REC_ITERS) ? fi[1 + mi] : FREE; }
void note_krylov(Mem *m, int mi, int code) { if (mi + 1 < PIHM_REC_ITERS) m->att_it[1 + mi] = code; }

bool setup_nonnull(const Mem *m) { return m->pretype != PIHM_PREC_NONE && m->pset != nullptr; }
bool pre_right(const Mem *m) { return m->pretype == PIHM_PREC_RIGHT && m->psolve != nullptr; }

// CVSpilsPSolve (cvode_spils.c:639-655): z = P^-1 r, right side
int psolve(Mem *m, N_Vector r, N_Vector z) {
  const int rv = m->psolve(m->tn, m->ycur, m->fcur, r, z, m->gamma, m->delta_ls, PIHM_PREC_RIGHT, m->ud, m->ytemp);
  m->nps++;
  return rv;
}

// Newton's retry with fresh preconditioner data (cvode.c:2760-2830)
inline int retry_or(const Mem *m, int flag) { return (!m->jcur && setup_nonnull(m)) ? TRY_AGAIN : flag; }

// A replay's last kernel hands its sums to the host before the runtime has
// retired the launch, so the host can reach this point with launches of these
// executable graphs still in flight: wait for their stream first.
void free_graphs(Mem *m) {
  bool any = false;
  for (auto g : m->gx) any = any || g;
  if (any && m->gx_stream) (void)hipStreamSynchronize(m->gx_stream);
  for (auto &g : m->gx) {
    if (g) (void)hipGraphExecDestroy(g);
    g = nullptr;
  }
  m->gx_stream = nullptr;
}

void free_vectors(Mem *m) {
  free_graphs(m);
  if (m->kry) (void)hipFree(m->kry);
  m->kry = nullptr;
  for (auto &v : m->zn) { VDestroy(v); v = nullptr; }
  N_Vector *vs[] = {&m->ewt, &m->y, &m->acor, &m->tempv, &m->ftemp, &m->xcor, &m->vtemp, &m->ytemp, &m->xs};
  for (auto p : vs) { VDestroy(*p); *p = nullptr; }
  for (auto &v : m->V) { VDestroy(v); v = nullptr; }
}

bool alloc_vectors(Mem *m, N_Vector tmpl) {
  for (int j = 0; j <= m->qmax; j++) if (!(m->zn[j] = VClone(tmpl))) return false;
  N_Vector *vs[] = {&m->ewt, &m->y, &m->acor, &m->tempv, &m->ftemp};
  for (auto p : vs) if (!(*p = VClone(tmpl))) return false;
  return true;
}

bool alloc_spgmr(Mem *m) {
  for (int j = 0; j <= m->maxl; j++) if (!(m->V[j] = VClone(m->zn[0]))) return false;
  N_Vector *vs[] = {&m->xcor, &m->vtemp, &m->ytemp, &m->xs};
  for (auto p : vs) if (!(*p = VClone(m->zn[0]))) return false;
  if (m->fused && !m->kry) {
    if (hipMalloc((void **)&m->kry, PIHM_KRY_SIZE * sizeof(double)) != hipSuccess) return false;
    if (hipMemset(m->kry, 0, PIHM_KRY_SIZE * sizeof(double)) != hipSuccess) return false;
  }
  return true;
}

// cvEwtSetSS (cvode.c:4081-4098); with device vectors the same arithmetic in
// one launch, which also returns WRMS(ycur, ewt).
int ewt_set(Mem *m, N_Vector ycur, double *ynrm) {
  if (m->fused) {
    double mn = 0, nrm = 0;
    if (pihm_nv_ewt_ss(ycur, m->reltol, m->abstol, m->ewt, &mn, &nrm)) return -1;
    if (ynrm) *ynrm = nrm;
    return (mn <= 0.0) ? -1 : 0;
  }
  ycur->ops->nvabs(ycur, m->tempv);
  VScale(m->reltol, m->tempv, m->tempv);
  m->tempv->ops->nvaddconst(m->tempv, m->abstol, m->tempv);
  if (m->tempv->ops->nvmin(m->tempv) <= 0.0) return -1;
  m->tempv->ops->nvinv(m->tempv, m->ewt);
  if (ynrm) *ynrm = VWrms(ycur, m->ewt);
  return 0;
}

// ---------------------------------------------------------------- Nordsieck
void rescale(Mem *m) {   // cvode.c:2251-2275
  double f[6] = {1, 1, 1, 1, 1, 1};
  double factor = m->eta;
  for (int j = 1; j <= m->q; j++) { f[j] = factor; factor *= m->eta; }
  if (m->fused) pihm_nv_rescale(m->q, f, m->zn);
  else for (int j = 1; j <= m->q; j++) VScale(f[j], m->zn[j], m->zn[j]);
  m->h = m->hscale * m->eta;
  m->next_h = m->h;
  m->hscale = m->h;
  m->nscon = 0;
}

void predict(Mem *m) {   // cvode.c:2277-2306
  m->tn += m->h;
  if (m->tstopset && (m->tn - m->tstop) * m->h > 0.0) m->tn = m->tstop;
  if (m->fused) { pihm_nv_predict(m->q, m->zn); return; }
  for (int k = 1; k <= m->q; k++)
    for (int j = m->q; j >= k; j--) VLin(1.0, m->zn[j - 1], 1.0, m->zn[j], m->zn[j - 1]);
}

void restore(Mem *m, double saved_t) {   // cvode.c:2882-2913
  m->tn = saved_t;
  if (m->fused) { pihm_nv_restore(m->q, m->zn); return; }
  for (int k = 1; k <= m->q; k++)
    for (int j = m->q; j >= k; j--) VLin(1.0, m->zn[j - 1], -1.0, m->zn[j], m->zn[j - 1]);
}

void increase_bdf(Mem *m) {   // cvode.c:2189-2223
  const int q = m->q;
  for (int i = 0; i <= m->qmax; i++) m->l[i] = 0.0;
  double alpha1 = 1.0, prod = 1.0, xiold = 1.0, alpha0 = -1.0;
  m->l[2] = 1.0;
  double hsum = m->hscale;
  if (q > 1) {
    for (int j = 1; j < q; j++) {
      hsum += m->tau[j + 1];
      const double xi = hsum / m->hscale;
      prod *= xi;
      alpha0 -= 1.0 / (j + 1);
      alpha1 += 1.0 / xi;
      for (int i = j + 2; i >= 2; i--) m->l[i] = m->l[i] * xiold + m->l[i - 1];
      xiold = xi;
    }
  }
  const double A1 = (-alpha0 - alpha1) / prod;
  VScale(A1, m->zn[m->indx_acor], m->zn[m->L]);
  for (int j = 2; j <= q; j++) VLin(m->l[j], m->zn[m->L], 1.0, m->zn[j], m->zn[j]);
}

void decrease_bdf(Mem *m) {   // cvode.c:2225-2249
  const int q = m->q;
  for (int i = 0; i <= m->qmax; i++) m->l[i] = 0.0;
  m->l[2] = 1.0;
  double hsum = 0.0;
  for (int j = 1; j <= q - 2; j++) {
    hsum += m->tau[j];
    const double xi = hsum / m->hscale;
    for (int i = j + 2; i >= 2; i--) m->l[i] = m->l[i] * xi + m->l[i - 1];
  }
  for (int j = 2; j < q; j++) VLin(-m->l[j], m->zn[q], 1.0, m->zn[j], m->zn[j]);
}

void adjust_order(Mem *m, int deltaq) {   // cvode.c:2099-2118, 2165-2187
  if (m->q == 2 && deltaq != 1) return;
  if (deltaq == 1) increase_bdf(m);
  else if (deltaq == -1) decrease_bdf(m);
}

void adjust_params(Mem *m) {   // cvode.c:2078-2097
  if (m->qprime != m->q) {
    adjust_order(m, m->qprime - m->q);
    m->q = m->qprime;
    m->L = m->q + 1;
    m->qwait = m->L;
  }
  rescale(m);
}

// ---------------------------------------------------------------- SPGMR
// ModifiedGS with the reorthogonalisation test of sundials_iterative.c:44-90
void modified_gs(Mem *m, int k, double *new_vk_norm) {
  N_Vector *v = m->V;
  const double vk_norm = std::sqrt(VDot(v[k], v[k]));
  const int k1 = k - 1;
  const int i0 = (k - m->maxl > 0) ? k - m->maxl : 0;
  for (int i = i0; i < k; i++) {
    m->Hes[i][k1] = VDot(v[i], v[k]);
    VLin(1.0, v[k], -m->Hes[i][k1], v[i], v[k]);
  }
  *new_vk_norm = std::sqrt(VDot(v[k], v[k]));
  double temp = GS_FACTOR * vk_norm;
  if ((temp + (*new_vk_norm)) != temp) return;
  double new_norm_2 = 0.0;
  for (int i = i0; i < k; i++) {
    const double np = VDot(v[i], v[k]);
    temp = GS_FACTOR * m->Hes[i][k1];
    if ((temp + np) == temp) continue;
    m->Hes[i][k1] += np;
    VLin(1.0, v[k], -np, v[i], v[k]);
    new_norm_2 += np * np;
  }
  if (new_norm_2 != 0.0) {
    const double np = (*new_vk_norm) * (*new_vk_norm) - new_norm_2;
    *new_vk_norm = (np > 0.0) ? std::sqrt(np) : 0.0;
  }
}

// CVSpilsDQJtimes + CVSpilsAtimes (cvode_spils.c:600-700): z = v - gamma*J*v
int atimes(Mem *m, N_Vector v, N_Vector z) {
  double sig = 1.0 / VWrms(v, m->ewt);
  int rv = 0;
  for (int it = 0; it < DQ_MAX_ITERS; it++) {
    VLin(sig, v, 1.0, m->ycur, m->ytemp);
    rv = m->f(m->tn, m->ytemp, z, m->ud);
    m->nfes++;
    if (rv == 0) break;
    if (rv < 0) return -1;
    sig *= 0.25;
  }
  if (rv > 0) return 1;
  const double siginv = 1.0 / sig;
  VLin(siginv, z, -siginv, m->fcur, z);
  VLin(1.0, v, -m->gamma, z, z);
  return 0;
}

// SpgmrSolve with x0 = 0, no restarts, s1 = s2 = ewt (sundials_spgmr.c:165-425)
// force: FREE, or the exact iteration count of a replayed solve (-1: the zero solution)
int spgmr(Mem *m, N_Vector x, N_Vector b, double delta, int *nli, int force) {
  N_Vector *V = m->V;
  N_Vector s = m->ewt;
  *nli = 0;
  VProd(s, b, V[0]);
  const double beta = std::sqrt(VDot(V[0], V[0]));
  double r_norm = beta, rho = beta;
  if (force == FREE ? r_norm <= delta : force < 0) { *nli = -1; return SP_SUCCESS; }
  VConst(0.0, m->xcor);
  for (int i = 0; i <= m->maxl; i++)
    for (int j = 0; j < m->maxl; j++) m->Hes[i][j] = 0.0;
  double rotation_product = 1.0;
  VScale(1.0 / r_norm, V[0], V[0]);
  bool converged = false;
  int krydim = 0;
  const bool pre = pre_right(m);
  for (int l = 0; l < m->maxl; l++) {
    (*nli)++;
    krydim = l + 1;
    VDiv(V[l], s, m->vtemp);
    if (pre) {
      VScale(1.0, m->vtemp, V[l + 1]);
      const int ps = psolve(m, V[l + 1], m->vtemp);
      if (ps != 0) return (ps < 0) ? SP_PSOLVE_FAIL_UNREC : SP_PSOLVE_FAIL_REC;
    }
    const int ier = atimes(m, m->vtemp, V[l + 1]);
    if (ier != 0) return (ier < 0) ? SP_ATIMES_FAIL_UNREC : SP_ATIMES_FAIL_REC;
    VProd(s, V[l + 1], V[l + 1]);
    modified_gs(m, l + 1, &m->Hes[l + 1][l]);
    if (qr_update(m, krydim) != 0) return SP_QRFACT_FAIL;
    rotation_product *= m->givens[2 * l + 1];
    rho = std::fabs(rotation_product * r_norm);
    if (force == FREE ? rho <= delta : (l + 1 >= force || rho == 0.0)) { converged = true; break; }
    VScale(1.0 / m->Hes[l + 1][l], V[l + 1], V[l + 1]);
  }
  m->yg[0] = r_norm;
  for (int i = 1; i <= krydim; i++) m->yg[i] = 0.0;
  if (qr_solve(m, krydim) != 0) return SP_QRSOL_FAIL;
  for (int k = 0; k < krydim; k++) VLin(m->yg[k], V[k], 1.0, m->xcor, m->xcor);
  if (converged || rho < beta) {
    VDiv(m->xcor, s, m->xcor);
    if (pre) {
      const int ps = psolve(m, m->xcor, m->vtemp);
      if (ps != 0) return (ps < 0) ? SP_PSOLVE_FAIL_UNREC : SP_PSOLVE_FAIL_REC;
      VLin(1.0, x, 1.0, m->vtemp, x);
    } else {
      VLin(1.0, x, 1.0, m->xcor, x);
    }
    return converged ? SP_SUCCESS : SP_RES_REDUCED;
  }
  return SP_CONV_FAIL;
}

// CVSpgmrSolve (cvode_spgmr.c:355-420)
// *code: Krylov iterations taken, 0 for the small-residual exit, -1 for the zero solution
int lsolve(Mem *m, N_Vector b, N_Vector ynow, N_Vector fnow, int force, int *code) {
  const double deltar = m->eplifac * m->tq[4];
  const double bnorm = VWrms(b, m->ewt);
  *code = 0;
  if (force == FREE ? bnorm <= deltar : force == 0) {
    if (m->mnewt > 0) VConst(0.0, b);
    return 0;
  }
  m->ycur = ynow;
  m->fcur = fnow;
  const double delta = deltar * m->sqrtN;
  m->delta_ls = delta;
  VConst(0.0, m->xs);
  int nli = 0;
  const int rv = spgmr(m, m->xs, b, delta, &nli, force);
  VScale(1.0, m->xs, b);
  *code = nli;
  if (nli > 0) m->nli += nli;
  switch (rv) {
    case SP_SUCCESS: return 0;
    case SP_RES_REDUCED: return (m->mnewt == 0) ? 0 : 1;
    case SP_CONV_FAIL: case SP_QRFACT_FAIL: case SP_ATIMES_FAIL_REC: case SP_PSOLVE_FAIL_REC: return 1;
    default: return -1;
  }
}

// CVSpgmrSetup (cvode_spgmr.c:289-330)
int lsetup(Mem *m, int convfail, N_Vector ypred, N_Vector fpred) {
  const double dgamma = std::fabs((m->gamma / m->gammap) - 1.0);
  const bool jbad = (m->nst == 0) || (m->nst > m->nstlpre + SPILS_MSBPRE) ||
                    ((convfail == CV_FAIL_BAD_J) && (dgamma < SPILS_DGMAX)) || (convfail == CV_FAIL_OTHER);
  int jcur = jbad ? 1 : 0;
  const int rv = m->pset(m->tn, ypred, fpred, jbad ? 0 : 1, &jcur, m->gamma, m->ud, m->acor, m->y, m->tempv);
  if (jbad) jcur = 1;
  if (jcur) {
    m->npe++;
    m->nstlpre = m->nst;
  }
  m->jcur = jcur != 0;
  return rv;
}

// ---------------------------------------------------------------- fused (device vectors)
// SPGMR as above with classical Gram-Schmidt: the inner products of an Arnoldi
// step come from the same launch that forms the new vector, and the
// orthogonalisation is applied when the next basis vector is formed
// (pihm_gm_prep).  The norm of the orthogonalised vector follows from
// ||w||^2 - sum h_k^2; when that loses more than six digits an explicit
// second pass (pihm_gm_orth) is made.  The difference-quotient increment is
// sig = 1/||V_l/s||_WRMS(s) = sqrt(N) for the unit basis vector V_l.
enum { SP_ZERO = 10 };

// Capture Krylov iteration l (prep from the device sums, the RHS at y + sig*v,
// the Arnoldi step) on the vectors' stream.  False (and graphs off for this
// integrator) when the stream is the null stream or capture fails.
bool build_graph(Mem *m, int l) {
  hipStream_t st = (hipStream_t)pihm_nv_stream(m->ewt);
  if (!st || !m->kry) { m->graph_off = true; return false; }
  hipGraph_t g = nullptr;
  if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) != hipSuccess) { m->graph_off = true; return false; }
  const bool pre = pre_right(m);
  int bad = pihm_gm_prep_dev(l, m->V, m->kry, m->y, m->ewt, m->sqrtN, m->ytemp, pre ? m->xs : nullptr);
  if (!bad && pre) {
    // the preconditioner's solve is captured with the iteration: its data is
    // read from device memory when the graph replays
    bad = psolve(m, m->xs, m->xcor);
    m->nps--;
    if (!bad) VLin(m->sqrtN, m->xcor, 1.0, m->y, m->ytemp);
  }
  if (!bad) bad = m->f(m->tn, m->ytemp, m->vtemp, m->ud);
  if (!bad) bad = pihm_gm_arnoldi(l, m->V, m->ewt, m->vtemp, m->ftemp, 1.0 / m->sqrtN, 0.0, nullptr, m->kry,
                                  pre ? m->xcor : nullptr);
  const hipError_t ec = hipStreamEndCapture(st, &g);
  if (bad || ec != hipSuccess || !g) {
    if (g) (void)hipGraphDestroy(g);
    (void)hipGetLastError();
    m->graph_off = true;
    return false;
  }
  const hipError_t ei = hipGraphInstantiate(&m->gx[l], g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (ei != hipSuccess) { m->gx[l] = nullptr; m->graph_off = true; return false; }
  m->gx_stream = st;
  return true;
}

int spgmr_fused(Mem *m, double sumsq, double delta, int *krydim_out, int force) {
  N_Vector *V = m->V;
  N_Vector s = m->ewt;
  const double beta = std::sqrt(sumsq);
  double r_norm = beta, rho = beta;
  *krydim_out = 0;
  if (force == FREE ? r_norm <= delta : force < 0) return SP_ZERO;
  for (int i = 0; i <= m->maxl; i++)
    for (int j = 0; j < m->maxl; j++) m->Hes[i][j] = 0.0;
  double rotation_product = 1.0;
  double lazy[MAXL_CAP + 1] = {};
  double cscale = 1.0 / r_norm;
  const double sig = m->sqrtN;
  const double siginv = 1.0 / sig;
  bool converged = false;
  int krydim = 0;
  bool graph = m->graph_rhs && !m->graph_off && m->kry;   // gamma was left in kry by pihm_gm_start
  // graphs captured on another stream (the vectors' stream was changed) are rebuilt
  if (graph && m->gx_stream && m->gx_stream != (hipStream_t)pihm_nv_stream(s)) free_graphs(m);
  const bool pre = pre_right(m);
  for (int l = 0; l < m->maxl; l++) {
    krydim = l + 1;
    double sums[MAXL_CAP + 2];
    if (graph && !m->gx[l]) graph = build_graph(m, l);
    if (graph) {
      // the replayed iteration computes the same coefficients on the device
      hipStream_t st = (hipStream_t)pihm_nv_stream(s);
      if (hipGraphLaunch(m->gx[l], st) != hipSuccess) return SP_ATIMES_FAIL_UNREC;
      m->nfes++;
      if (pre) m->nps++;
      pihm_redwork w = pihm_nv_redwork(s);
      if (pihm_red_wait(w, st)) return SP_ATIMES_FAIL_UNREC;
      for (int k = 0; k <= l + 1; k++) sums[k] = w.result[k];
    } else {
      if (pihm_gm_prep(l, V, lazy, cscale, m->ycur, s, sig, m->ytemp, pre ? m->xs : nullptr)) return SP_ATIMES_FAIL_UNREC;
      if (pre) {
        const int ps = psolve(m, m->xs, m->xcor);
        if (ps != 0) return (ps < 0) ? SP_PSOLVE_FAIL_UNREC : SP_PSOLVE_FAIL_REC;
        VLin(sig, m->xcor, 1.0, m->ycur, m->ytemp);
      }
      const int rv = m->f(m->tn, m->ytemp, m->vtemp, m->ud);
      m->nfes++;
      if (rv < 0) return SP_ATIMES_FAIL_UNREC;
      if (rv > 0) return SP_ATIMES_FAIL_REC;
      if (pihm_gm_arnoldi(l, V, s, m->vtemp, m->fcur, siginv, m->gamma, sums, m->kry, pre ? m->xcor : nullptr))
        return SP_ATIMES_FAIL_UNREC;
    }
    const double ww = sums[l + 1];
    double hsq = 0.0;
    for (int k = 0; k <= l; k++) {
      m->Hes[k][l] = sums[k];
      lazy[k] = sums[k];
      hsq += sums[k] * sums[k];
    }
    double nn = ww - hsq;
    if (!(nn > 1e-6 * ww)) {
      double s2[MAXL_CAP + 2];
      if (pihm_gm_orth(l, V, lazy, s2, m->kry)) return SP_GS_FAIL;
      double hsq2 = 0.0;
      for (int k = 0; k <= l; k++) {
        m->Hes[k][l] += s2[k];
        lazy[k] = s2[k];
        hsq2 += s2[k] * s2[k];
      }
      nn = s2[l + 1] - hsq2;
      if (nn < 0.0) nn = 0.0;
    }
    m->Hes[l + 1][l] = std::sqrt(nn);
    if (qr_update(m, krydim) != 0) return SP_QRFACT_FAIL;
    rotation_product *= m->givens[2 * l + 1];
    rho = std::fabs(rotation_product * r_norm);
    if (force == FREE ? rho <= delta : (l + 1 >= force || rho == 0.0)) { converged = true; break; }
    cscale = 1.0 / m->Hes[l + 1][l];
  }
  m->yg[0] = r_norm;
  for (int i = 1; i <= krydim; i++) m->yg[i] = 0.0;
  if (qr_solve(m, krydim) != 0) return SP_QRSOL_FAIL;
  *krydim_out = krydim;
  if (converged) return SP_SUCCESS;
  if (rho < beta) return SP_RES_REDUCED;
  return SP_CONV_FAIL;
}

int newton_iteration_fused(Mem *m) {   // cvode.c:2729-2832 with CVSpgmrSolve fused in
  m->mnewt = 0;
  int mi = 0;
  double del = 0.0, delp = 0.0;
  const double N = (double)m->neq;
  const int *fi = replay_iters(m);
  for (int k = 0; k < PIHM_REC_ITERS; k++) m->att_it[k] = 0;
  for (;;) {
    double sumsq = 0.0;
    if (pihm_gm_start(m->rl1, m->zn[1], m->acor, m->gamma, m->ftemp, m->ewt, m->tempv, m->V[0], &sumsq, m->kry))
      return PIHM_CV_LSOLVE_FAIL;
    const double deltar = m->eplifac * m->tq[4];
    const double bnorm = std::sqrt(sumsq / N);
    int mode = 0, krydim = 0, lret = 0;
    const int force = forced_krylov(fi, mi);
    if (force == FREE ? bnorm <= deltar : force == 0) {
      mode = (m->mnewt > 0) ? 2 : 1;
      note_krylov(m, mi, 0);
    } else {
      m->ycur = m->y;
      m->fcur = m->ftemp;
      m->delta_ls = deltar * m->sqrtN;
      const int sp = spgmr_fused(m, sumsq, deltar * m->sqrtN, &krydim, force);
      m->nli += krydim;
      note_krylov(m, mi, sp == SP_ZERO ? -1 : krydim);
      switch (sp) {
        case SP_ZERO: mode = 2; break;
        case SP_SUCCESS: mode = 0; break;
        case SP_RES_REDUCED: mode = 0; lret = (m->mnewt == 0) ? 0 : 1; break;
        case SP_CONV_FAIL: case SP_QRFACT_FAIL: case SP_ATIMES_FAIL_REC: case SP_PSOLVE_FAIL_REC: lret = 1; break;
        default: lret = -1; break;
      }
    }
    m->nni++;
    if (lret < 0) return PIHM_CV_LSOLVE_FAIL;
    if (lret > 0) return retry_or(m, CONV_FAIL);
    double sums[2];
    const double *xsrc = nullptr;
    if (mode == 0 && pre_right(m)) {
      // x = P^-1 (sum_j y_j V_j / s), then as the unpreconditioned finish
      if (pihm_gm_combine(krydim, m->yg, m->V, m->ewt, m->xs)) return PIHM_CV_LSOLVE_FAIL;
      const int ps = psolve(m, m->xs, m->xcor);
      if (ps < 0) return PIHM_CV_LSOLVE_FAIL;
      if (ps > 0) return retry_or(m, CONV_FAIL);
      mode = 3;
      xsrc = N_VGetDeviceArrayPointer_Hip(m->xcor);
    }
    if (pihm_gm_finish(mode, krydim, m->yg, m->V, m->ewt, m->tempv, m->zn[0], m->acor, m->y, xsrc, sums))
      return PIHM_CV_LSOLVE_FAIL;
    del = std::sqrt(sums[0] / N);
    if (mi > 0) m->crate = std::fmax(CRDOWN * m->crate, del / delp);
    const double dcon = del * std::fmin(1.0, m->crate) / m->tq[4];
    if (fi ? mi + 1 >= fi[0] : dcon <= 1.0) {   // a replayed step takes its recorded count
      m->acnrm = (mi == 0) ? del : std::sqrt(sums[1] / N);
      m->att_it[0] = mi + 1;
      return PIHM_CV_SUCCESS;
    }
    m->mnewt = ++mi;
    if (mi == m->maxcor || (!fi && mi >= 2 && del > RDIV * delp)) return retry_or(m, CONV_FAIL);
    delp = del;
    const int fr = m->f(m->tn, m->y, m->ftemp, m->ud);
    m->nfe++;
    if (fr < 0) return PIHM_CV_RHSFUNC_FAIL;
    if (fr > 0) return retry_or(m, RHSFUNC_RECVR);
  }
}

// ---------------------------------------------------------------- Newton
int newton_iteration(Mem *m) {   // cvode.c:2729-2832
  m->mnewt = 0;
  int mi = 0;
  double del = 0.0, delp = 0.0;
  const int *fi = replay_iters(m);
  for (int k = 0; k < PIHM_REC_ITERS; k++) m->att_it[k] = 0;
  for (;;) {
    if (m->fused) pihm_nv_newton_rhs(m->rl1, m->zn[1], m->acor, m->gamma, m->ftemp, m->tempv);
    else {
      VLin(m->rl1, m->zn[1], 1.0, m->acor, m->tempv);
      VLin(m->gamma, m->ftemp, -1.0, m->tempv, m->tempv);
    }
    N_Vector b = m->tempv;
    int code = 0;
    const int rv = lsolve(m, b, m->y, m->ftemp, forced_krylov(fi, mi), &code);
    note_krylov(m, mi, code);
    m->nni++;
    if (rv < 0) return PIHM_CV_LSOLVE_FAIL;
    if (rv > 0) return retry_or(m, CONV_FAIL);
    del = VWrms(b, m->ewt);
    VLin(1.0, m->acor, 1.0, b, m->acor);
    VLin(1.0, m->zn[0], 1.0, m->acor, m->y);
    if (mi > 0) m->crate = std::fmax(CRDOWN * m->crate, del / delp);
    const double dcon = del * std::fmin(1.0, m->crate) / m->tq[4];
    if (fi ? mi + 1 >= fi[0] : dcon <= 1.0) {   // a replayed step takes its recorded count
      m->acnrm = (mi == 0) ? del : VWrms(m->acor, m->ewt);
      m->att_it[0] = mi + 1;
      return PIHM_CV_SUCCESS;
    }
    m->mnewt = ++mi;
    if (mi == m->maxcor || (!fi && mi >= 2 && del > RDIV * delp)) return retry_or(m, CONV_FAIL);
    delp = del;
    const int fr = m->f(m->tn, m->y, m->ftemp, m->ud);
    m->nfe++;
    if (fr < 0) return PIHM_CV_RHSFUNC_FAIL;
    if (fr > 0) return retry_or(m, RHSFUNC_RECVR);
  }
}

int nls_newton(Mem *m, int nflag) {   // cvode.c:2651-2727
  int convfail = ((nflag == FIRST_CALL) || (nflag == PREV_ERR_FAIL)) ? CV_NO_FAILURES : CV_FAIL_OTHER;
  bool call_setup = false;
  if (setup_nonnull(m)) {
    call_setup = (nflag == PREV_CONV_FAIL) || (nflag == PREV_ERR_FAIL) || (m->nst == 0) ||
                 (m->nst >= m->nstlp + MSBP) || (std::fabs(m->gamrat - 1.0) > DGMAX);
  } else {
    m->crate = 1.0;
  }
  for (;;) {
    const int rv = m->f(m->tn, m->zn[0], m->ftemp, m->ud);
    m->nfe++;
    if (rv < 0) return PIHM_CV_RHSFUNC_FAIL;
    if (rv > 0) return RHSFUNC_RECVR;
    if (call_setup) {
      const int ier = lsetup(m, convfail, m->zn[0], m->ftemp);
      m->nsetups++;
      call_setup = false;
      m->gamrat = m->crate = 1.0;
      m->gammap = m->gamma;
      m->nstlp = m->nst;
      if (ier < 0) return PIHM_CV_LSETUP_FAIL;
      if (ier > 0) return CONV_FAIL;
    }
    VConst(0.0, m->acor);
    VScale(1.0, m->zn[0], m->y);
    const int ier = m->fused ? newton_iteration_fused(m) : newton_iteration(m);
    if (ier != TRY_AGAIN) return ier;
    call_setup = true;
    convfail = CV_FAIL_BAD_J;
  }
}

int handle_nflag(Mem *m, int *nflag, double saved_t, int *ncf) {   // cvode.c:2834-2880
  const int nf = *nflag;
  if (nf == PIHM_CV_SUCCESS) return DO_ERROR_TEST;
  m->ncfn++;
  restore(m, saved_t);
  if (nf == PIHM_CV_LSETUP_FAIL) return PIHM_CV_LSETUP_FAIL;
  if (nf == PIHM_CV_LSOLVE_FAIL) return PIHM_CV_LSOLVE_FAIL;
  if (nf == PIHM_CV_RHSFUNC_FAIL) return PIHM_CV_RHSFUNC_FAIL;
  (*ncf)++;
  m->etamax = 1.0;
  if ((std::fabs(m->h) <= m->hmin * ONEPSM) || (*ncf == m->maxncf)) {
    if (nf == CONV_FAIL) return PIHM_CV_CONV_FAILURE;
    if (nf == RHSFUNC_RECVR) return PIHM_CV_REPTD_RHSFUNC_ERR;
  }
  m->eta = std::fmax(ETACF, m->hmin / std::fabs(m->h));
  *nflag = PREV_CONV_FAIL;
  rescale(m);
  return PREDICT_AGAIN;
}

int do_error_test(Mem *m, int *nflag, double saved_t, int *nef, double *dsmp) {   // cvode.c:2915-2994
  const double dsm = m->acnrm * m->tq[2];
  *dsmp = dsm;
  if (dsm <= 1.0) return PIHM_CV_SUCCESS;
  if (m->rp_h && m->nst < m->rp_n) return PIHM_CV_SUCCESS;   // a replayed step is taken as recorded
  (*nef)++;
  m->netf++;
  *nflag = PREV_ERR_FAIL;
  restore(m, saved_t);
  if ((std::fabs(m->h) <= m->hmin * ONEPSM) || (*nef == m->maxnef)) return PIHM_CV_ERR_FAILURE;
  m->etamax = 1.0;
  if (*nef <= MXNEF1) {
    m->eta = 1.0 / (std::pow(BIAS2 * dsm, 1.0 / m->L) + ADDON);
    m->eta = std::fmax(ETAMIN, std::fmax(m->eta, m->hmin / std::fabs(m->h)));
    if (*nef >= SMALL_NEF) m->eta = std::fmin(m->eta, ETAMXF);
    rescale(m);
    return TRY_AGAIN;
  }
  if (m->q > 1) {
    m->eta = std::fmax(ETAMIN, m->hmin / std::fabs(m->h));
    adjust_order(m, -1);
    m->L = m->q;
    m->q--;
    m->qwait = m->L;
    rescale(m);
    return TRY_AGAIN;
  }
  m->eta = std::fmax(ETAMIN, m->hmin / std::fabs(m->h));
  m->h *= m->eta;
  m->next_h = m->h;
  m->hscale = m->h;
  m->qwait = LONG_WAIT;
  m->nscon = 0;
  const int rv = m->f(m->tn, m->zn[0], m->tempv, m->ud);
  m->nfe++;
  if (rv < 0) return PIHM_CV_RHSFUNC_FAIL;
  if (rv > 0) return PIHM_CV_UNREC_RHSFUNC_ERR;
  VScale(m->h, m->tempv, m->zn[1]);
  return TRY_AGAIN;
}

void complete_step(Mem *m) {   // cvode.c:2996-3027
  if (m->rec_h && m->nst < m->rec_cap) {
    m->rec_h[m->nst] = m->h;
    m->rec_q[m->nst] = m->q;
    if (m->rec_it)
      for (int k = 0; k < PIHM_REC_ITERS; k++) m->rec_it[m->nst * PIHM_REC_ITERS + k] = m->att_it[k];
    *m->rec_n = m->nst + 1;
  }
  m->nst++;
  m->nscon++;
  m->hu = m->h;
  m->qu = m->q;
  for (int i = m->q; i >= 2; i--) m->tau[i] = m->tau[i - 1];
  if ((m->q == 1) && (m->nst > 1)) m->tau[2] = m->tau[1];
  m->tau[1] = m->h;
  if (m->fused) pihm_nv_correct(m->q, m->l, m->acor, m->zn);
  else for (int j = 0; j <= m->q; j++) VLin(m->l[j], m->acor, 1.0, m->zn[j], m->zn[j]);
  m->qwait--;
  if ((m->qwait == 1) && (m->q != m->qmax)) {
    VScale(1.0, m->acor, m->zn[m->qmax]);
    m->saved_tq5 = m->tq[5];
    m->indx_acor = m->qmax;
  }
}

double compute_etaqm1(Mem *m) {   // cvode.c:3091-3108
  m->etaqm1 = 0.0;
  if (m->q > 1) {
    const double ddn = VWrms(m->zn[m->q], m->ewt) * m->tq[1];
    m->etaqm1 = 1.0 / (std::pow(BIAS1 * ddn, 1.0 / m->q) + ADDON);
  }
  return m->etaqm1;
}

double compute_etaqp1(Mem *m) {   // cvode.c:3110-3135
  m->etaqp1 = 0.0;
  if (m->q != m->qmax) {
    if (m->saved_tq5 == 0.0) return m->etaqp1;
    double pw = 1.0;
    const double r = m->h / m->tau[2];
    for (int i = 0; i < m->L; i++) pw *= r;   // SUNRpowerI(h/tau[2], L)
    const double cquot = (m->tq[5] / m->saved_tq5) * pw;
    VLin(-cquot, m->zn[m->qmax], 1.0, m->acor, m->tempv);
    const double dup = VWrms(m->tempv, m->ewt) * m->tq[3];
    m->etaqp1 = 1.0 / (std::pow(BIAS3 * dup, 1.0 / (m->L + 1)) + ADDON);
  }
  return m->etaqp1;
}

void choose_eta(Mem *m) {   // cvode.c:3137-3184
  const double etam = std::fmax(m->etaqm1, std::fmax(m->etaq, m->etaqp1));
  if (etam < THRESH) {
    m->eta = 1.0;
    m->qprime = m->q;
    return;
  }
  if (etam == m->etaq) {
    m->eta = m->etaq;
    m->qprime = m->q;
  } else if (etam == m->etaqm1) {
    m->eta = m->etaqm1;
    m->qprime = m->q - 1;
  } else {
    m->eta = m->etaqp1;
    m->qprime = m->q + 1;
    VScale(1.0, m->acor, m->zn[m->qmax]);
  }
}

void prepare_next_step(Mem *m, double dsm) {   // cvode.c:3029-3066
  if (m->etamax == 1.0) {
    m->qwait = (m->qwait > 2) ? m->qwait : 2;
    m->qprime = m->q;
    m->hprime = m->h;
    m->eta = 1.0;
    return;
  }
  m->etaq = 1.0 / (std::pow(BIAS2 * dsm, 1.0 / m->L) + ADDON);
  if (m->qwait != 0) {
    m->eta = m->etaq;
    m->qprime = m->q;
    set_eta(m);
    return;
  }
  m->qwait = 2;
  m->etaqm1 = compute_etaqm1(m);
  m->etaqp1 = compute_etaqp1(m);
  choose_eta(m);
  set_eta(m);
}

int step(Mem *m) {   // cvode.c:2005-2066
  const double saved_t = m->tn;
  int ncf = 0, nef = 0, nflag = FIRST_CALL;
  double dsm = 0.0;
  if ((m->nst > 0) && (m->hprime != m->h || (m->rp_h && m->qprime != m->q))) adjust_params(m);
  if (m->rp_h && m->nst > 0 && m->nst < m->rp_n) {   // the recorded size exactly, not h * (h'/h)
    m->h = m->rp_h[m->nst];
    m->hscale = m->h;
    m->next_h = m->h;
  }
  for (;;) {
    predict(m);
    set_bdf(m);
    nflag = nls_newton(m, nflag);
    const int kflag = handle_nflag(m, &nflag, saved_t, &ncf);
    if (kflag == PREDICT_AGAIN && m->rp_h && m->nst < m->rp_n) return PIHM_CV_CONV_FAILURE;   // cannot follow
    if (kflag == PREDICT_AGAIN) continue;
    if (kflag != DO_ERROR_TEST) return kflag;
    const int eflag = do_error_test(m, &nflag, saved_t, &nef, &dsm);
    if (eflag == TRY_AGAIN) continue;
    if (eflag != PIHM_CV_SUCCESS) return eflag;
    break;
  }
  complete_step(m);
  prepare_next_step(m, dsm);
  if (m->rp_h && m->nst < m->rp_n) {
    // the next step as recorded: its order and size replace the selection
    const int qn = m->rp_q[m->nst];
    if (qn < 1 || qn > m->qmax || qn > m->q + 1) return PIHM_CV_ILL_INPUT;
    if (qn == m->q + 1 && m->qprime != qn) VScale(1.0, m->acor, m->zn[m->qmax]);   // as choose_eta
    while (m->q > qn + 1) {   // a drop of several orders, as repeated error-test failures make it
      adjust_order(m, -1);
      m->q--;
      m->L = m->q + 1;
    }
    m->qprime = qn;
    m->hprime = m->rp_h[m->nst];
    m->eta = m->hprime / m->h;
  }
  m->etamax = (m->nst <= SMALL_NST) ? ETAMX2 : ETAMX3;
  VScale(m->tq[2], m->acor, m->acor);
  return PIHM_CV_SUCCESS;
}

int get_dky(Mem *m, double t, int k, N_Vector dky) {   // cvode.c:1509-1569
  if (k < 0 || k > m->q) return PIHM_CV_BAD_K;
  double tfuzz = FUZZ_FACTOR * UROUND * (std::fabs(m->tn) + std::fabs(m->hu));
  if (m->hu < 0.0) tfuzz = -tfuzz;
  const double tp = m->tn - m->hu - tfuzz;
  const double tn1 = m->tn + tfuzz;
  if ((t - tp) * (t - tn1) > 0.0) return PIHM_CV_BAD_T;
  const double s = (t - m->tn) / m->h;
  for (int j = m->q; j >= k; j--) {
    double c = 1.0;
    for (int i = j; i >= j - k + 1; i--) c *= i;
    if (j == m->q) VScale(c, m->zn[m->q], dky);
    else VLin(c, m->zn[j], s, dky, dky);
  }
  if (k == 0) return PIHM_CV_SUCCESS;
  double r = 1.0;
  for (int i = 0; i < k; i++) r /= m->h;
  VScale(r, dky, dky);
  return PIHM_CV_SUCCESS;
}

}  // namespace

// ---------------------------------------------------------------- C API
extern "C" {

void *PihmCVodeCreate(int lmm, int iter) {
  if (lmm != PIHM_CV_BDF || iter != PIHM_CV_NEWTON) return nullptr;
  return new (std::nothrow) Mem();
}

int PihmCVodeInit(void *mem, PihmRhsFn f, realtype t0, N_Vector y0) {   // cvode.c:438-578
  if (!mem) return PIHM_CV_MEM_NULL;
  Mem *m = (Mem *)mem;
  if (!y0 || !f) return PIHM_CV_ILL_INPUT;
  free_vectors(m);
  if (!alloc_vectors(m, y0)) return PIHM_CV_MEM_FAIL;
  long lrw = 0, liw = 0;
  y0->ops->nvspace(y0, &lrw, &liw);
  m->neq = lrw;
  m->sqrtN = std::sqrt((double)lrw);
  m->fused = is_hip(y0);
  m->f = f;
  m->tn = t0;
  m->q = 1; m->L = 2; m->qwait = m->L; m->etamax = ETAMX1;
  m->qu = 0; m->hu = 0.0; m->tolsf = 1.0;
  VScale(1.0, y0, m->zn[0]);
  m->nst = m->nfe = m->ncfn = m->netf = m->nni = m->nli = m->nfes = m->nhnil = m->nscon = 0;
  m->nstlp = m->nstlpre = m->npe = m->nps = m->nsetups = 0;
  m->jcur = false;
  m->h0u = 0.0; m->next_h = 0.0; m->next_q = 0;
  m->indx_acor = m->qmax;
  for (double &t : m->tau) t = 0.0;
  m->malloc_done = true;
  if (m->ls_set && !m->V[0]) {
    if (!alloc_spgmr(m)) return PIHM_CV_MEM_FAIL;
  }
  return PIHM_CV_SUCCESS;
}

int PihmCVodeReInit(void *mem, realtype t0, N_Vector y0) {   // cvode.c:580-675
  if (!mem) return PIHM_CV_MEM_NULL;
  Mem *m = (Mem *)mem;
  if (!m->malloc_done) return PIHM_CV_NO_MALLOC;
  m->tn = t0;
  m->q = 1; m->L = 2; m->qwait = m->L; m->etamax = ETAMX1;
  m->qu = 0; m->hu = 0.0; m->tolsf = 1.0;
  VScale(1.0, y0, m->zn[0]);
  m->nst = m->nfe = m->ncfn = m->netf = m->nni = m->nli = m->nfes = m->nhnil = m->nscon = 0;
  m->nstlp = m->nstlpre = m->npe = m->nps = m->nsetups = 0;
  m->jcur = false;
  m->h0u = 0.0; m->next_h = 0.0; m->next_q = 0;
  m->indx_acor = m->qmax;
  for (double &t : m->tau) t = 0.0;
  return PIHM_CV_SUCCESS;
}

int PihmCVodeSStolerances(void *mem, realtype reltol, realtype abstol) {
  if (!mem) return PIHM_CV_MEM_NULL;
  if (reltol < 0.0 || abstol < 0.0) return PIHM_CV_ILL_INPUT;
  Mem *m = (Mem *)mem;
  m->reltol = reltol; m->abstol = abstol; m->tol_set = true;
  return PIHM_CV_SUCCESS;
}
int PihmCVodeSetUserData(void *mem, void *ud) { if (!mem) return PIHM_CV_MEM_NULL; ((Mem *)mem)->ud = ud; return 0; }
int PihmCVodeSetInitStep(void *mem, realtype hin) { if (!mem) return PIHM_CV_MEM_NULL; ((Mem *)mem)->hin = hin; return 0; }
int PihmCVodeSetMaxStep(void *mem, realtype hmax) {
  if (!mem) return PIHM_CV_MEM_NULL;
  if (hmax < 0.0) return PIHM_CV_ILL_INPUT;
  Mem *m = (Mem *)mem;
  m->hmax_inv = (hmax == 0.0) ? 0.0 : 1.0 / hmax;
  return 0;
}
int PihmCVodeSetMinStep(void *mem, realtype hmin) {
  if (!mem) return PIHM_CV_MEM_NULL;
  if (hmin < 0.0) return PIHM_CV_ILL_INPUT;
  ((Mem *)mem)->hmin = hmin;
  return 0;
}
int PihmCVodeSetMaxNumSteps(void *mem, long int mx) {
  if (!mem) return PIHM_CV_MEM_NULL;
  ((Mem *)mem)->mxstep = (mx == 0) ? MXSTEP_DEFAULT : mx;
  return 0;
}
int PihmCVodeSetMaxOrd(void *mem, int maxord) {
  if (!mem) return PIHM_CV_MEM_NULL;
  Mem *m = (Mem *)mem;
  if (maxord <= 0 || maxord > 5) return PIHM_CV_ILL_INPUT;
  if (m->malloc_done && maxord > m->qmax) return PIHM_CV_ILL_INPUT;
  m->qmax = maxord;
  return 0;
}
int PihmCVodeSetStopTime(void *mem, realtype tstop) {
  if (!mem) return PIHM_CV_MEM_NULL;
  Mem *m = (Mem *)mem;
  if (m->nst > 0 && (tstop - m->tn) * m->h < 0.0) return PIHM_CV_ILL_INPUT;
  m->tstop = tstop; m->tstopset = true;
  return 0;
}
int PihmCVodeSetGraphRhs(void *mem, int onoff) {
  if (!mem) return PIHM_CV_MEM_NULL;
  Mem *m = (Mem *)mem;
  m->graph_rhs = onoff != 0;
  m->graph_off = false;
  free_graphs(m);
  return PIHM_CV_SUCCESS;
}
int PihmCVodeSetStepRecord(void *mem, double *h, int *q, int *iters, long cap, long *n) {
  if (!mem) return PIHM_CV_MEM_NULL;
  if (cap < 0 || (cap > 0 && (!h || !q || !n))) return PIHM_CV_ILL_INPUT;
  Mem *m = (Mem *)mem;
  m->rec_h = cap > 0 ? h : nullptr;
  m->rec_q = q;
  m->rec_it = cap > 0 ? iters : nullptr;
  m->rec_cap = cap;
  m->rec_n = n;
  if (n) *n = 0;
  return PIHM_CV_SUCCESS;
}

int PihmCVodeSetStepReplay(void *mem, const double *h, const int *q, const int *iters, long n) {
  if (!mem) return PIHM_CV_MEM_NULL;
  if (n < 0 || (n > 0 && (!h || !q))) return PIHM_CV_ILL_INPUT;
  Mem *m = (Mem *)mem;
  m->rp_h = n > 0 ? h : nullptr;
  m->rp_q = q;
  m->rp_it = n > 0 ? iters : nullptr;
  m->rp_n = n;
  return PIHM_CV_SUCCESS;
}

int PihmCVSpilsSetPreconditioner(void *mem, PihmPrecSetupFn pset, PihmPrecSolveFn psolve) {
  if (!mem) return PIHM_CV_MEM_NULL;
  Mem *m = (Mem *)mem;
  m->pset = pset;
  m->psolve = psolve;
  free_graphs(m);
  return PIHM_CV_SUCCESS;
}
int PihmCVodeSetStabLimDet(void *mem, int onoff) { if (!mem) return PIHM_CV_MEM_NULL; ((Mem *)mem)->sldet = onoff != 0; return 0; }

int PihmCVSpgmr(void *mem, int pretype, int maxl) {
  if (!mem) return PIHM_CV_MEM_NULL;
  if (pretype != PIHM_PREC_NONE && pretype != PIHM_PREC_RIGHT) return PIHM_CV_ILL_INPUT;
  Mem *m = (Mem *)mem;
  m->pretype = pretype;
  free_graphs(m);
  m->maxl = (maxl <= 0) ? 5 : (maxl > MAXL_CAP ? MAXL_CAP : maxl);
  m->ls_set = true;
  if (m->zn[0]) {
    free_graphs(m);
    for (auto &v : m->V) { VDestroy(v); v = nullptr; }
    N_Vector *vs[] = {&m->xcor, &m->vtemp, &m->ytemp, &m->xs};
    for (auto p : vs) { VDestroy(*p); *p = nullptr; }
    if (!alloc_spgmr(m)) return PIHM_CV_MEM_FAIL;
  }
  return 0;
}

int PihmCVode(void *mem, realtype tout, N_Vector yout, realtype *tret, int itask) {   // cvode.c:1074-1507
  if (!mem) return PIHM_CV_MEM_NULL;
  Mem *m = (Mem *)mem;
  if (!m->malloc_done) return PIHM_CV_NO_MALLOC;
  if (!yout || !tret) return PIHM_CV_ILL_INPUT;
  if (itask != PIHM_CV_NORMAL && itask != PIHM_CV_ONE_STEP) return PIHM_CV_ILL_INPUT;
  if (itask == PIHM_CV_NORMAL) m->toutc = tout;
  bool have_nrm = false;
  double nrm = 0.0;

  if (m->nst == 0) {
    m->tretlast = *tret = m->tn;
    if (!m->tol_set || !m->ls_set) return PIHM_CV_ILL_INPUT;
    if (ewt_set(m, m->zn[0], &nrm) != 0) return PIHM_CV_ILL_INPUT;
    have_nrm = m->fused;
    const int rv = m->f(m->tn, m->zn[0], m->zn[1], m->ud);
    m->nfe++;
    if (rv < 0) return PIHM_CV_RHSFUNC_FAIL;
    if (rv > 0) return PIHM_CV_FIRST_RHSFUNC_ERR;
    if (m->tstopset && (m->tstop - m->tn) * (tout - m->tn) <= 0.0) return PIHM_CV_ILL_INPUT;
    m->h = m->hin;
    if (m->h != 0.0 && (tout - m->tn) * m->h < 0.0) return PIHM_CV_ILL_INPUT;
    if (m->h == 0.0) {
      // MM-PIHM always sets the initial step (src/ode.c:480); without one, a
      // conservative estimate from the first derivative is used.
      const double tdist = std::fabs(tout - m->tn);
      const double ydn = VWrms(m->zn[1], m->ewt);
      double hg = (ydn > 0.0) ? 0.5 / ydn : 0.01 * tdist;
      if (hg > 0.1 * tdist) hg = 0.1 * tdist;
      if (hg <= 0.0) hg = 1e-6;
      m->h = (tout - m->tn >= 0.0) ? hg : -hg;
    }
    const double rh = std::fabs(m->h) * m->hmax_inv;
    if (rh > 1.0) m->h /= rh;
    if (std::fabs(m->h) < m->hmin) m->h *= m->hmin / std::fabs(m->h);
    if (m->tstopset && (m->tn + m->h - m->tstop) * m->h > 0.0)
      m->h = (m->tstop - m->tn) * (1.0 - 4.0 * UROUND);
    if (m->rp_h && m->rp_n > 0) m->h = m->rp_h[0];   // the recorded first step
    m->hscale = m->h;
    m->h0u = m->h;
    m->hprime = m->h;
    VScale(m->h, m->zn[1], m->zn[1]);
  }

  if (m->nst > 0) {
    const double troundoff = FUZZ_FACTOR * UROUND * (std::fabs(m->tn) + std::fabs(m->h));
    if (itask == PIHM_CV_NORMAL && (m->tn - tout) * m->h >= 0.0) {
      m->tretlast = *tret = tout;
      if (get_dky(m, tout, 0, yout) != 0) return PIHM_CV_ILL_INPUT;
      return PIHM_CV_SUCCESS;
    }
    if (itask == PIHM_CV_ONE_STEP && std::fabs(m->tn - m->tretlast) > troundoff) {
      m->tretlast = *tret = m->tn;
      VScale(1.0, m->zn[0], yout);
      return PIHM_CV_SUCCESS;
    }
    if (m->tstopset) {
      if (std::fabs(m->tn - m->tstop) <= troundoff) {
        if (get_dky(m, m->tstop, 0, yout) != 0) return PIHM_CV_ILL_INPUT;
        m->tretlast = *tret = m->tstop;
        m->tstopset = false;
        return PIHM_CV_TSTOP_RETURN;
      }
      if ((m->tn + m->hprime - m->tstop) * m->h > 0.0) {
        m->hprime = (m->tstop - m->tn) * (1.0 - 4.0 * UROUND);
        m->eta = m->hprime / m->h;
      }
    }
  }

  long nstloc = 0;
  int istate = PIHM_CV_SUCCESS;
  for (;;) {
    m->next_h = m->h;
    m->next_q = m->q;
    if (m->nst > 0) {
      if (ewt_set(m, m->zn[0], &nrm) != 0) {
        istate = PIHM_CV_ILL_INPUT;
        m->tretlast = *tret = m->tn;
        VScale(1.0, m->zn[0], yout);
        break;
      }
      have_nrm = m->fused;
    }
    if (m->mxstep > 0 && nstloc >= m->mxstep) {
      istate = PIHM_CV_TOO_MUCH_WORK;
      m->tretlast = *tret = m->tn;
      VScale(1.0, m->zn[0], yout);
      break;
    }
    if (!have_nrm) nrm = VWrms(m->zn[0], m->ewt);
    have_nrm = false;
    m->tolsf = UROUND * nrm;
    if (m->tolsf > 1.0) {
      istate = PIHM_CV_TOO_MUCH_ACC;
      m->tretlast = *tret = m->tn;
      VScale(1.0, m->zn[0], yout);
      m->tolsf *= 2.0;
      break;
    }
    m->tolsf = 1.0;
    if (m->tn + m->h == m->tn) m->nhnil++;

    const int kflag = step(m);
    if (kflag != PIHM_CV_SUCCESS) {
      istate = kflag;
      m->tretlast = *tret = m->tn;
      VScale(1.0, m->zn[0], yout);
      break;
    }
    nstloc++;
    if (itask == PIHM_CV_NORMAL && (m->tn - tout) * m->h >= 0.0) {
      istate = PIHM_CV_SUCCESS;
      m->tretlast = *tret = tout;
      (void)get_dky(m, tout, 0, yout);
      m->next_q = m->qprime;
      m->next_h = m->hprime;
      break;
    }
    if (m->tstopset) {
      const double troundoff = FUZZ_FACTOR * UROUND * (std::fabs(m->tn) + std::fabs(m->h));
      if (std::fabs(m->tn - m->tstop) <= troundoff) {
        (void)get_dky(m, m->tstop, 0, yout);
        m->tretlast = *tret = m->tstop;
        m->tstopset = false;
        istate = PIHM_CV_TSTOP_RETURN;
        break;
      }
      if ((m->tn + m->hprime - m->tstop) * m->h > 0.0) {
        m->hprime = (m->tstop - m->tn) * (1.0 - 4.0 * UROUND);
        m->eta = m->hprime / m->h;
      }
    }
    if (itask == PIHM_CV_ONE_STEP) {
      istate = PIHM_CV_SUCCESS;
      m->tretlast = *tret = m->tn;
      VScale(1.0, m->zn[0], yout);
      m->next_q = m->qprime;
      m->next_h = m->hprime;
      break;
    }
  }
  return istate;
}

int PihmCVodeGetDky(void *mem, realtype t, int k, N_Vector dky) {
  if (!mem) return PIHM_CV_MEM_NULL;
  if (!dky) return PIHM_CV_ILL_INPUT;
  return get_dky((Mem *)mem, t, k, dky);
}

void PihmCVodeFree(void **mem) {
  if (!mem || !*mem) return;
  Mem *m = (Mem *)*mem;
  free_vectors(m);
  delete m;
  *mem = nullptr;
}

#define GETTER(name, type, field)                                  \
  int name(void *mem, type *out) {                                 \
    if (!mem) return PIHM_CV_MEM_NULL;                             \
    *out = (type)((Mem *)mem)->field;                              \
    return PIHM_CV_SUCCESS;                                        \
  }
GETTER(PihmCVodeGetNumSteps, long int, nst)
GETTER(PihmCVodeGetNumRhsEvals, long int, nfe)
GETTER(PihmCVodeGetNumLinIters, long int, nli)
GETTER(PihmCVodeGetNumJtimesRhsEvals, long int, nfes)
GETTER(PihmCVodeGetNumNonlinSolvIters, long int, nni)
GETTER(PihmCVodeGetNumNonlinSolvConvFails, long int, ncfn)
GETTER(PihmCVodeGetNumErrTestFails, long int, netf)
GETTER(PihmCVodeGetLastStep, realtype, hu)
GETTER(PihmCVodeGetLastOrder, int, qu)
GETTER(PihmCVodeGetCurrentTime, realtype, tn)
GETTER(PihmCVodeGetCurrentStep, realtype, next_h)
GETTER(PihmCVodeGetNumLinSolvSetups, long int, nsetups)
GETTER(PihmCVSpilsGetNumPrecEvals, long int, npe)
GETTER(PihmCVSpilsGetNumPrecSolves, long int, nps)
#undef GETTER

}  // extern "C"
