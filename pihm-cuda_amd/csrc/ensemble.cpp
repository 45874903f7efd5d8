// ensemble.cpp - many MM-PIHM solves interleaved on one device.
//
// Each member runs the model-step loop and the BDF/Newton/SPGMR integrator as
// a coroutine.  Device work is queued, not launched: vector operations and
// right-hand-side evaluations go into the member's queue, and a member
// suspends when it needs the value of an inner product.  When every member
// has suspended, the scheduler runs all queued work in rounds (one launch of
// the batched vector-operation kernel for every member whose next item is a
// vector operation, then one batched RHS evaluation for every member whose
// next item is an RHS call, and so on), synchronises once, and resumes the
// members.  A member's queue order is its program order, so each member sees
// exactly the sequence of operations the single-solve integrator performs.
//
// Integrator steps follow CVODE 2.9 (cvode/src/cvode/cvode.c, cvode_spgmr.c,
// sundials_spgmr.c; file:line at each routine), in the form of the fused
// single-solve device path: classical Gram-Schmidt with lazily applied
// orthogonalisation, one synchronisation per Krylov iteration, error weights
// computed in the same pass as the first prediction of each step.  The step
// logic restates SUNDIALS routines, BSD-3 licensed by Lawrence Livermore
// National Security (Copyright (c) 2002-2016, LLNS; UCRL-CODE-155951); the
// licence text and disclaimer are in NOTICE.SUNDIALS at the package root.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <coroutine>
#include <cstring>
#include <exception>
#include <new>
#include <vector>

#include "../../include/pihm_ensemble.h"
#include "bdf_core.hpp"
#include "common.hpp"
#include "ctx.hpp"
#include "ens.hpp"

using namespace pihm;
using namespace pihm::ens;
using namespace pihm::bdf;

namespace {

// ------------------------------------------------------------------ coroutine task
template <class T>
struct Task {
  struct promise_type {
    T value{};
    std::coroutine_handle<> cont;
    Task get_return_object() { return Task{std::coroutine_handle<promise_type>::from_promise(*this)}; }
    std::suspend_always initial_suspend() noexcept { return {}; }
    struct Final {
      bool await_ready() noexcept { return false; }
      std::coroutine_handle<> await_suspend(std::coroutine_handle<promise_type> h) noexcept {
        auto c = h.promise().cont;
        return c ? c : std::noop_coroutine();
      }
      void await_resume() noexcept {}
    };
    Final final_suspend() noexcept { return {}; }
    void return_value(T v) { value = v; }
    void unhandled_exception() { std::terminate(); }
  };
  std::coroutine_handle<promise_type> h;
  Task() : h(nullptr) {}
  explicit Task(std::coroutine_handle<promise_type> hh) : h(hh) {}
  Task(Task &&o) noexcept : h(o.h) { o.h = nullptr; }
  Task &operator=(Task &&o) noexcept {
    if (this != &o) { if (h) h.destroy(); h = o.h; o.h = nullptr; }
    return *this;
  }
  Task(const Task &) = delete;
  ~Task() { if (h) h.destroy(); }
  bool done() const { return !h || h.done(); }
  T result() const { return h.promise().value; }
  // awaiting a task starts it and resumes the awaiting coroutine when it ends
  bool await_ready() { return false; }
  std::coroutine_handle<> await_suspend(std::coroutine_handle<> c) {
    h.promise().cont = c;
    return h;
  }
  T await_resume() { return h.promise().value; }
};

// Integrator outcomes beyond the CVODE flags: error weights invalid or
// accuracy unattainable, found at the first synchronisation of a step.
enum { EWT_BAD = 20, ACC_BAD = 21 };

struct Item {
  bool rhs;
  bool red;
  int in, out;
  Op op;
};

struct Member {
  int id = 0;
  bool enabled = true, failed = false;
  // configuration (SetCVodeParam, src/ode.c:413-529)
  double reltol = 0, abstol = 0, hin = 0, hmin = 0, hmax_inv = 0;
  long mxstep = MXSTEP_DEFAULT;
  int qmax = kMaxQ, maxl = kMaxL, maxcor = NLS_MAXCOR, maxnef = MXNEF, maxncf = MXNCF;
  double nlscoef = CORTES, eplifac = 0.05;
  bool tstopset = false;
  double tstop = 0;
  long neq = 0;
  double sqrtN = 1;
  // integrator state (the names of cvode.c)
  double tn = 0, h = 0, hu = 0, hprime = 0, hscale = 0, eta = 1, etamax = ETAMX1;
  double etaq = 0, etaqm1 = 0, etaqp1 = 0, next_h = 0, h0u = 0, tretlast = 0, toutc = 0;
  double tolsf = 1;
  double tau[7] = {}, tq[6] = {}, l[6] = {};
  double rl1 = 0, gamma = 0, gammap = 0, gamrat = 1, crate = 1, acnrm = 0, saved_tq5 = 0;
  int q = 1, qprime = 1, next_q = 0, qu = 0, L = 2, qwait = 2, indx_acor = kMaxQ;
  long nst = 0, nfe = 0, ncfn = 0, netf = 0, nni = 0, nli = 0, nfes = 0, nhnil = 0, nscon = 0;
  int mnewt = 0;
  double Hes[MAXL_CAP + 1][MAXL_CAP] = {};
  double givens[2 * MAXL_CAP] = {};
  double yg[MAXL_CAP + 1] = {};
  bool ewt_check = false;       // the queued prediction also sets the error weights
  // model-step loop (pihm_drive)
  long ksteps = 0;
  double maxstep = 0;
  long nst0 = 0, ncfn0 = 0, nni0 = 0;
  int flag = 0;
  long model_steps = 0;
  // scheduling
  std::vector<Item> q_items;
  std::coroutine_handle<> waiting;
  const double *res = nullptr;  // result slots (pinned)
};

struct SyncAw {
  Member *m;
  bool await_ready() noexcept { return false; }
  void await_suspend(std::coroutine_handle<> h) noexcept { m->waiting = h; }
  void await_resume() noexcept {}
};
inline SyncAw sync(Member &m) { return SyncAw{&m}; }

// ------------------------------------------------------------------ queued work
inline Op mk(int kind) {
  Op o;
  std::memset(&o, 0, sizeof(o));
  o.kind = kind;
  return o;
}
inline void put(Member &m, const Op &o, bool red) { m.q_items.push_back(Item{false, red, 0, 0, o}); }
inline void rhs(Member &m, int in, int out) {
  Item it;
  std::memset(&it, 0, sizeof(it));
  it.rhs = true;
  it.in = in;
  it.out = out;
  m.q_items.push_back(it);
}
inline void v_scale(Member &m, double c, int x, int z) {
  Op o = mk(OP_SCALE);
  o.a0 = x; o.a2 = z; o.s[0] = c;
  put(m, o, false);
}
inline void v_lin(Member &m, double a, int x, double b, int y, int z) {
  Op o = mk(OP_LIN);
  o.a0 = x; o.a1 = y; o.a2 = z; o.s[0] = a; o.s[1] = b;
  put(m, o, false);
}

// ------------------------------------------------------------------ Nordsieck array
void rescale(Member &m) {   // cvode.c:2251-2275
  Op o = mk(OP_RESCALE);
  o.a0 = m.q;
  double factor = m.eta;
  for (int j = 1; j <= m.q; j++) { o.s[j] = factor; factor *= m.eta; }
  put(m, o, false);
  m.h = m.hscale * m.eta;
  m.next_h = m.h;
  m.hscale = m.h;
  m.nscon = 0;
}

void predict(Member &m, bool with_ewt) {   // cvode.c:2277-2306 (+ cvEwtSetSS, cvNlsNewton init)
  m.tn += m.h;
  if (m.tstopset && (m.tn - m.tstop) * m.h > 0.0) m.tn = m.tstop;
  Op o = mk(OP_PREDICT);
  o.a0 = m.q;
  o.a1 = (with_ewt ? 1 : 0) | 2;
  o.s[0] = m.reltol;
  o.s[1] = m.abstol;
  if (with_ewt) { o.nsum = 2; o.minfirst = 1; o.rslot = 8; }
  put(m, o, with_ewt);
  m.ewt_check = with_ewt;
}

void restore(Member &m, double saved_t) {   // cvode.c:2882-2913
  m.tn = saved_t;
  Op o = mk(OP_RESTORE);
  o.a0 = m.q;
  put(m, o, false);
}

void increase_bdf(Member &m) {   // cvode.c:2189-2223
  const int q = m.q;
  for (int i = 0; i <= m.qmax; i++) m.l[i] = 0.0;
  double alpha1 = 1.0, prod = 1.0, xiold = 1.0, alpha0 = -1.0;
  m.l[2] = 1.0;
  double hsum = m.hscale;
  if (q > 1) {
    for (int j = 1; j < q; j++) {
      hsum += m.tau[j + 1];
      const double xi = hsum / m.hscale;
      prod *= xi;
      alpha0 -= 1.0 / (j + 1);
      alpha1 += 1.0 / xi;
      for (int i = j + 2; i >= 2; i--) m.l[i] = m.l[i] * xiold + m.l[i - 1];
      xiold = xi;
    }
  }
  const double A1 = (-alpha0 - alpha1) / prod;
  v_scale(m, A1, ZN0 + m.indx_acor, ZN0 + m.L);
  for (int j = 2; j <= q; j++) v_lin(m, m.l[j], ZN0 + m.L, 1.0, ZN0 + j, ZN0 + j);
}

void decrease_bdf(Member &m) {   // cvode.c:2225-2249
  const int q = m.q;
  for (int i = 0; i <= m.qmax; i++) m.l[i] = 0.0;
  m.l[2] = 1.0;
  double hsum = 0.0;
  for (int j = 1; j <= q - 2; j++) {
    hsum += m.tau[j];
    const double xi = hsum / m.hscale;
    for (int i = j + 2; i >= 2; i--) m.l[i] = m.l[i] * xi + m.l[i - 1];
  }
  for (int j = 2; j < q; j++) v_lin(m, -m.l[j], ZN0 + q, 1.0, ZN0 + j, ZN0 + j);
}

void adjust_order(Member &m, int deltaq) {   // cvode.c:2099-2118, 2165-2187
  if (m.q == 2 && deltaq != 1) return;
  if (deltaq == 1) increase_bdf(m);
  else if (deltaq == -1) decrease_bdf(m);
}

void adjust_params(Member &m) {   // cvode.c:2078-2097
  if (m.qprime != m.q) {
    adjust_order(m, m.qprime - m.q);
    m.q = m.qprime;
    m.L = m.q + 1;
    m.qwait = m.L;
  }
  rescale(m);
}

void get_dky(Member &m, double t) {   // cvode.c:1509-1569, k = 0, into the output vector
  Op o = mk(OP_DKY);
  o.a0 = m.q;
  o.s[0] = (t - m.tn) / m.h;
  put(m, o, false);
}

bool dky_t_ok(const Member &m, double t) {
  double tfuzz = FUZZ_FACTOR * UROUND * (std::fabs(m.tn) + std::fabs(m.hu));
  if (m.hu < 0.0) tfuzz = -tfuzz;
  const double tp = m.tn - m.hu - tfuzz;
  const double tn1 = m.tn + tfuzz;
  return !((t - tp) * (t - tn1) > 0.0);
}

// ------------------------------------------------------------------ SPGMR and Newton
// SpgmrSolve with x0 = 0, no restarts, s1 = s2 = ewt (sundials_spgmr.c:165-425),
// classical Gram-Schmidt: the inner products of an Arnoldi step come from the
// launch that forms the new vector and the orthogonalisation is applied when
// the next basis vector is formed; the norm of the orthogonalised vector is
// ||w||^2 - sum h_k^2, with an explicit second pass when that loses more than
// six digits.  DQ increment sig = 1/||V_l/s||_WRMS(s) = sqrt(N).
enum { SP_ZERO = 10 };

Task<int> spgmr(Member &m, double sumsq, double delta, int *krydim_out) {
  const double beta = std::sqrt(sumsq);
  double r_norm = beta, rho = beta;
  *krydim_out = 0;
  if (r_norm <= delta) co_return SP_ZERO;
  for (int i = 0; i <= m.maxl; i++)
    for (int j = 0; j < m.maxl; j++) m.Hes[i][j] = 0.0;
  double rotation_product = 1.0;
  double lazy[MAXL_CAP + 1] = {};
  double cscale = 1.0 / r_norm;
  const double sig = m.sqrtN;
  const double siginv = 1.0 / sig;
  bool converged = false;
  int krydim = 0;
  for (int l = 0; l < m.maxl; l++) {
    krydim = l + 1;
    Op p = mk(OP_GM_PREP);
    p.a0 = l; p.s[0] = cscale; p.s[1] = sig;
    for (int k = 0; k < l; k++) p.s[2 + k] = lazy[k];
    put(m, p, false);
    rhs(m, YTEMP, VTEMP);
    m.nfes++;
    Op a = mk(OP_GM_ARNOLDI);
    a.a0 = l; a.s[0] = siginv; a.s[1] = m.gamma; a.nsum = l + 2;
    put(m, a, true);
    co_await sync(m);
    const double ww = m.res[l + 1];
    double hsq = 0.0;
    for (int k = 0; k <= l; k++) {
      m.Hes[k][l] = m.res[k];
      lazy[k] = m.res[k];
      hsq += m.res[k] * m.res[k];
    }
    double nn = ww - hsq;
    if (!(nn > 1e-6 * ww)) {
      Op o = mk(OP_GM_ORTH);
      o.a0 = l; o.nsum = l + 2;
      for (int k = 0; k <= l; k++) o.s[k] = lazy[k];
      put(m, o, true);
      co_await sync(m);
      double hsq2 = 0.0;
      for (int k = 0; k <= l; k++) {
        m.Hes[k][l] += m.res[k];
        lazy[k] = m.res[k];
        hsq2 += m.res[k] * m.res[k];
      }
      nn = m.res[l + 1] - hsq2;
      if (nn < 0.0) nn = 0.0;
    }
    m.Hes[l + 1][l] = std::sqrt(nn);
    if (qr_update(&m, krydim) != 0) co_return SP_QRFACT_FAIL;
    rotation_product *= m.givens[2 * l + 1];
    rho = std::fabs(rotation_product * r_norm);
    if (rho <= delta) { converged = true; break; }
    cscale = 1.0 / m.Hes[l + 1][l];
  }
  m.yg[0] = r_norm;
  for (int i = 1; i <= krydim; i++) m.yg[i] = 0.0;
  if (qr_solve(&m, krydim) != 0) co_return SP_QRSOL_FAIL;
  *krydim_out = krydim;
  if (converged) co_return SP_SUCCESS;
  if (rho < beta) co_return SP_RES_REDUCED;
  co_return SP_CONV_FAIL;
}

bool ewt_results_ok(Member &m, int *why) {
  const double mn = m.res[8];
  const double nrm = std::sqrt(m.res[9] / (double)m.neq);
  if (mn <= 0.0) { *why = EWT_BAD; return false; }   // cvode.c:1354-1362
  m.tolsf = UROUND * nrm;                              // cvode.c:1377-1385
  if (m.tolsf > 1.0) { m.tolsf *= 2.0; *why = ACC_BAD; return false; }
  m.tolsf = 1.0;
  return true;
}

Task<int> newton_iteration(Member &m) {   // cvode.c:2729-2832 with CVSpgmrSolve (cvode_spgmr.c:355-420)
  m.mnewt = 0;
  int mi = 0;
  double del = 0.0, delp = 0.0;
  const double N = (double)m.neq;
  for (;;) {
    Op st = mk(OP_GM_START);
    st.s[0] = m.rl1; st.s[1] = m.gamma; st.nsum = 1;
    put(m, st, true);
    co_await sync(m);
    if (m.ewt_check) {
      m.ewt_check = false;
      int why = 0;
      if (!ewt_results_ok(m, &why)) co_return why;
    }
    const double sumsq = m.res[0];
    const double deltar = m.eplifac * m.tq[4];
    const double bnorm = std::sqrt(sumsq / N);
    int mode = 0, krydim = 0, lret = 0;
    if (bnorm <= deltar) {
      mode = (m.mnewt > 0) ? 2 : 1;
    } else {
      const int sp = co_await spgmr(m, sumsq, deltar * m.sqrtN, &krydim);
      m.nli += krydim;
      switch (sp) {
        case SP_ZERO: mode = 2; break;
        case SP_SUCCESS: mode = 0; break;
        case SP_RES_REDUCED: mode = 0; lret = (m.mnewt == 0) ? 0 : 1; break;
        case SP_CONV_FAIL: case SP_QRFACT_FAIL: case SP_ATIMES_FAIL_REC: lret = 1; break;
        default: lret = -1; break;
      }
    }
    m.nni++;
    if (lret < 0) co_return PIHM_CV_LSOLVE_FAIL;
    if (lret > 0) co_return CONV_FAIL;
    Op f = mk(OP_GM_FINISH);
    f.a0 = mode; f.a1 = krydim; f.nsum = 2;
    for (int j = 0; j < krydim; j++) f.s[j] = m.yg[j];
    put(m, f, true);
    co_await sync(m);
    del = std::sqrt(m.res[0] / N);
    if (mi > 0) m.crate = std::fmax(CRDOWN * m.crate, del / delp);
    const double dcon = del * std::fmin(1.0, m.crate) / m.tq[4];
    if (dcon <= 1.0) {
      m.acnrm = (mi == 0) ? del : std::sqrt(m.res[1] / N);
      co_return PIHM_CV_SUCCESS;
    }
    m.mnewt = ++mi;
    if (mi == m.maxcor || (mi >= 2 && del > RDIV * delp)) co_return CONV_FAIL;
    delp = del;
    rhs(m, YCOR, FTEMP);
    m.nfe++;
  }
}

Task<int> nls_newton(Member &m) {   // cvode.c:2651-2727 (PREC_NONE: no setup)
  m.crate = 1.0;
  rhs(m, ZN0, FTEMP);   // acor = 0 and y = zn0 were set by the prediction
  m.nfe++;
  co_return co_await newton_iteration(m);
}

int handle_nflag(Member &m, int *nflag, double saved_t, int *ncf) {   // cvode.c:2834-2880
  const int nf = *nflag;
  if (nf == PIHM_CV_SUCCESS) return DO_ERROR_TEST;
  m.ncfn++;
  restore(m, saved_t);
  if (nf == PIHM_CV_LSOLVE_FAIL) return PIHM_CV_LSOLVE_FAIL;
  if (nf == PIHM_CV_RHSFUNC_FAIL) return PIHM_CV_RHSFUNC_FAIL;
  (*ncf)++;
  m.etamax = 1.0;
  if ((std::fabs(m.h) <= m.hmin * ONEPSM) || (*ncf == m.maxncf)) {
    if (nf == CONV_FAIL) return PIHM_CV_CONV_FAILURE;
    if (nf == RHSFUNC_RECVR) return PIHM_CV_REPTD_RHSFUNC_ERR;
  }
  m.eta = std::fmax(ETACF, m.hmin / std::fabs(m.h));
  *nflag = PREV_CONV_FAIL;
  rescale(m);
  return PREDICT_AGAIN;
}

int do_error_test(Member &m, int *nflag, double saved_t, int *nef, double *dsmp) {   // cvode.c:2915-2994
  const double dsm = m.acnrm * m.tq[2];
  *dsmp = dsm;
  if (dsm <= 1.0) return PIHM_CV_SUCCESS;
  (*nef)++;
  m.netf++;
  *nflag = PREV_ERR_FAIL;
  restore(m, saved_t);
  if ((std::fabs(m.h) <= m.hmin * ONEPSM) || (*nef == m.maxnef)) return PIHM_CV_ERR_FAILURE;
  m.etamax = 1.0;
  if (*nef <= MXNEF1) {
    m.eta = 1.0 / (std::pow(BIAS2 * dsm, 1.0 / m.L) + ADDON);
    m.eta = std::fmax(ETAMIN, std::fmax(m.eta, m.hmin / std::fabs(m.h)));
    if (*nef >= SMALL_NEF) m.eta = std::fmin(m.eta, ETAMXF);
    rescale(m);
    return TRY_AGAIN;
  }
  if (m.q > 1) {
    m.eta = std::fmax(ETAMIN, m.hmin / std::fabs(m.h));
    adjust_order(m, -1);
    m.L = m.q;
    m.q--;
    m.qwait = m.L;
    rescale(m);
    return TRY_AGAIN;
  }
  m.eta = std::fmax(ETAMIN, m.hmin / std::fabs(m.h));
  m.h *= m.eta;
  m.next_h = m.h;
  m.hscale = m.h;
  m.qwait = LONG_WAIT;
  m.nscon = 0;
  rhs(m, ZN0, TEMPV);
  m.nfe++;
  v_scale(m, m.h, TEMPV, ZN0 + 1);
  return TRY_AGAIN;
}

void complete_step(Member &m) {   // cvode.c:2996-3027
  m.nst++;
  m.nscon++;
  m.hu = m.h;
  m.qu = m.q;
  for (int i = m.q; i >= 2; i--) m.tau[i] = m.tau[i - 1];
  if ((m.q == 1) && (m.nst > 1)) m.tau[2] = m.tau[1];
  m.tau[1] = m.h;
  Op o = mk(OP_CORRECT);
  o.a0 = m.q;
  for (int j = 0; j <= m.q; j++) o.s[j] = m.l[j];
  put(m, o, false);
  m.qwait--;
  if ((m.qwait == 1) && (m.q != m.qmax)) {
    v_scale(m, 1.0, ACOR, ZN0 + m.qmax);
    m.saved_tq5 = m.tq[5];
    m.indx_acor = m.qmax;
  }
}

void choose_eta(Member &m) {   // cvode.c:3137-3184
  const double etam = std::fmax(m.etaqm1, std::fmax(m.etaq, m.etaqp1));
  if (etam < THRESH) {
    m.eta = 1.0;
    m.qprime = m.q;
    return;
  }
  if (etam == m.etaq) {
    m.eta = m.etaq;
    m.qprime = m.q;
  } else if (etam == m.etaqm1) {
    m.eta = m.etaqm1;
    m.qprime = m.q - 1;
  } else {
    m.eta = m.etaqp1;
    m.qprime = m.q + 1;
    v_scale(m, 1.0, ACOR, ZN0 + m.qmax);
  }
}

// cvPrepareNextStep (cvode.c:3029-3066) with cvComputeEtaqm1/cvComputeEtaqp1
// (:3091-3135) sharing one pass over the vectors.
Task<int> prepare_next_step(Member &m, double dsm) {
  if (m.etamax == 1.0) {
    m.qwait = (m.qwait > 2) ? m.qwait : 2;
    m.qprime = m.q;
    m.hprime = m.h;
    m.eta = 1.0;
    co_return 0;
  }
  m.etaq = 1.0 / (std::pow(BIAS2 * dsm, 1.0 / m.L) + ADDON);
  if (m.qwait != 0) {
    m.eta = m.etaq;
    m.qprime = m.q;
    set_eta(&m);
    co_return 0;
  }
  m.qwait = 2;
  const bool need_m1 = m.q > 1;
  const bool need_p1 = (m.q != m.qmax) && (m.saved_tq5 != 0.0);
  double cquot = 0.0;
  if (need_p1) {
    double pw = 1.0;
    const double r = m.h / m.tau[2];
    for (int i = 0; i < m.L; i++) pw *= r;   // SUNRpowerI(h/tau[2], L)
    cquot = (m.tq[5] / m.saved_tq5) * pw;
  }
  if (need_m1 || need_p1) {
    Op o = mk(OP_ORDSEL);
    o.a0 = m.q;
    o.a1 = (need_m1 ? 1 : 0) | (need_p1 ? 2 : 0);
    o.s[0] = cquot;
    o.nsum = 2;
    put(m, o, true);
    co_await sync(m);
  }
  const double N = (double)m.neq;
  m.etaqm1 = 0.0;
  if (need_m1) {
    const double ddn = std::sqrt(m.res[0] / N) * m.tq[1];
    m.etaqm1 = 1.0 / (std::pow(BIAS1 * ddn, 1.0 / m.q) + ADDON);
  }
  m.etaqp1 = 0.0;
  if (need_p1) {
    const double dup = std::sqrt(m.res[1] / N) * m.tq[3];
    m.etaqp1 = 1.0 / (std::pow(BIAS3 * dup, 1.0 / (m.L + 1)) + ADDON);
  }
  choose_eta(m);
  set_eta(&m);
  co_return 0;
}

Task<int> step(Member &m) {   // cvode.c:2005-2066
  const double saved_t = m.tn;
  int ncf = 0, nef = 0, nflag = FIRST_CALL;
  double dsm = 0.0;
  if ((m.nst > 0) && (m.hprime != m.h)) adjust_params(m);
  bool first = true;
  for (;;) {
    // the error weights of this step (cvode.c:1354) come from the zn0 the
    // first prediction starts from
    predict(m, first && m.nst > 0);
    first = false;
    set_bdf(&m);
    nflag = co_await nls_newton(m);
    if (nflag == EWT_BAD || nflag == ACC_BAD) {
      restore(m, saved_t);
      co_return nflag == EWT_BAD ? PIHM_CV_ILL_INPUT : PIHM_CV_TOO_MUCH_ACC;
    }
    const int kflag = handle_nflag(m, &nflag, saved_t, &ncf);
    if (kflag == PREDICT_AGAIN) continue;
    if (kflag != DO_ERROR_TEST) co_return kflag;
    const int eflag = do_error_test(m, &nflag, saved_t, &nef, &dsm);
    if (eflag == TRY_AGAIN) continue;
    if (eflag != PIHM_CV_SUCCESS) co_return eflag;
    break;
  }
  complete_step(m);
  co_await prepare_next_step(m, dsm);
  m.etamax = (m.nst <= SMALL_NST) ? ETAMX2 : ETAMX3;
  v_scale(m, m.tq[2], ACOR, ACOR);
  co_return PIHM_CV_SUCCESS;
}

// CVode in CV_NORMAL mode with the stop time set (cvode.c:1074-1507); the
// solution at the return time goes to the member's output vector.
Task<int> cvode(Member &m, double tout) {
  m.toutc = tout;
  if (m.nst == 0) {
    m.tretlast = m.tn;
    Op e = mk(OP_PREDICT);   // q = 0: error weights of y0 only
    e.a1 = 1; e.s[0] = m.reltol; e.s[1] = m.abstol;
    e.nsum = 2; e.minfirst = 1; e.rslot = 8;
    put(m, e, true);
    co_await sync(m);
    int why = 0;
    if (!ewt_results_ok(m, &why)) co_return why == EWT_BAD ? PIHM_CV_ILL_INPUT : PIHM_CV_TOO_MUCH_ACC;
    rhs(m, ZN0, ZN0 + 1);
    m.nfe++;
    if (m.tstopset && (m.tstop - m.tn) * (tout - m.tn) <= 0.0) co_return PIHM_CV_ILL_INPUT;
    m.h = m.hin;
    if (m.h == 0.0 || (tout - m.tn) * m.h < 0.0) co_return PIHM_CV_ILL_INPUT;
    const double rh = std::fabs(m.h) * m.hmax_inv;
    if (rh > 1.0) m.h /= rh;
    if (std::fabs(m.h) < m.hmin) m.h *= m.hmin / std::fabs(m.h);
    if (m.tstopset && (m.tn + m.h - m.tstop) * m.h > 0.0) m.h = (m.tstop - m.tn) * (1.0 - 4.0 * UROUND);
    m.hscale = m.h;
    m.h0u = m.h;
    m.hprime = m.h;
    v_scale(m, m.h, ZN0 + 1, ZN0 + 1);
  } else {
    const double troundoff = FUZZ_FACTOR * UROUND * (std::fabs(m.tn) + std::fabs(m.h));
    if ((m.tn - tout) * m.h >= 0.0) {
      m.tretlast = tout;
      if (!dky_t_ok(m, tout)) co_return PIHM_CV_ILL_INPUT;
      get_dky(m, tout);
      co_return PIHM_CV_SUCCESS;
    }
    if (m.tstopset) {
      if (std::fabs(m.tn - m.tstop) <= troundoff) {
        if (!dky_t_ok(m, m.tstop)) co_return PIHM_CV_ILL_INPUT;
        get_dky(m, m.tstop);
        m.tretlast = m.tstop;
        m.tstopset = false;
        co_return PIHM_CV_TSTOP_RETURN;
      }
      if ((m.tn + m.hprime - m.tstop) * m.h > 0.0) {
        m.hprime = (m.tstop - m.tn) * (1.0 - 4.0 * UROUND);
        m.eta = m.hprime / m.h;
      }
    }
  }
  long nstloc = 0;
  int istate = PIHM_CV_SUCCESS;
  for (;;) {
    m.next_h = m.h;
    m.next_q = m.q;
    if (m.mxstep > 0 && nstloc >= m.mxstep) {
      istate = PIHM_CV_TOO_MUCH_WORK;
      m.tretlast = m.tn;
      v_scale(m, 1.0, ZN0, YOUT);
      break;
    }
    if (m.tn + m.h == m.tn) m.nhnil++;
    const int kflag = co_await step(m);
    if (kflag != PIHM_CV_SUCCESS) {
      istate = kflag;
      m.tretlast = m.tn;
      v_scale(m, 1.0, ZN0, YOUT);
      break;
    }
    nstloc++;
    if ((m.tn - tout) * m.h >= 0.0) {
      istate = PIHM_CV_SUCCESS;
      m.tretlast = tout;
      get_dky(m, tout);
      m.next_q = m.qprime;
      m.next_h = m.hprime;
      break;
    }
    if (m.tstopset) {
      const double troundoff = FUZZ_FACTOR * UROUND * (std::fabs(m.tn) + std::fabs(m.h));
      if (std::fabs(m.tn - m.tstop) <= troundoff) {
        get_dky(m, m.tstop);
        m.tretlast = m.tstop;
        m.tstopset = false;
        istate = PIHM_CV_TSTOP_RETURN;
        break;
      }
      if ((m.tn + m.hprime - m.tstop) * m.h > 0.0) {
        m.hprime = (m.tstop - m.tn) * (1.0 - 4.0 * UROUND);
        m.eta = m.hprime / m.h;
      }
    }
  }
  co_return istate;
}

}  // namespace

// ------------------------------------------------------------------ the ensemble
enum MemberStatus { ST_EMPTY = 0, ST_READY = 1, ST_RUNNING = 2, ST_DONE = 3, ST_FAILED = -1 };

struct pihm_ensemble {
  pihm_ctx *ctx = nullptr;
  hipStream_t st = nullptr;
  pihm_run_config cfg{};
  int nmem = 0;
  long stride = 0;
  int total_chunks = 0;
  Dev dev{};
  std::vector<MemberInfo> info;
  std::vector<int> nrec_cap, nst_cap;
  MemberInfo *info_dev = nullptr;
  double *vecs = nullptr;
  double *partials = nullptr;
  double *result = nullptr;    // pinned
  int *mnan = nullptr;         // pinned
  int32_t *station = nullptr;
  double *prcp = nullptr, *fedir = nullptr, *fett = nullptr;
  char *stage_h = nullptr, *stage_d = nullptr;
  size_t stage_cap = 0;
  std::vector<Member> mem;
  std::vector<int> status;     // MemberStatus per slot
  std::vector<long> k_end;     // model steps each member runs to
  std::vector<Task<int>> tops; // running model-step loops
  long ticks = 0, op_launches = 0, rhs_launches = 0;
  // the operation kernel's launches timed with HIP events (the first kEvPairs
  // of each tick) and their algorithmic bytes
  static constexpr int kEvPairs = 8;
  hipEvent_t ev_ops[2 * kEvPairs] = {};
  long ops_timed = 0;
  double ops_bytes = 0.0, ops_ms = 0.0;
  unsigned *sig_ctr = nullptr;           // device tick count
  volatile unsigned *sig_flag = nullptr; // pinned host copy
  unsigned sig_expect = 0;
  int ops_g = 1;               // chunks per workgroup of k_ens_ops (pihm_ens_autotune)

  double *P(int v, int m) const { return vecs + (long)v * stride + info[m].yoff; }
  int ensure_stage(size_t bytes);
  int flush();
  int upload_info(int k);
};

int pihm_ensemble::ensure_stage(size_t bytes) {
  if (bytes <= stage_cap) return PIHM_OK;
  size_t cap = std::max(bytes, stage_cap * 2);
  if (stage_h) (void)hipHostFree(stage_h);
  if (stage_d) (void)hipFree(stage_d);
  stage_h = nullptr;
  stage_d = nullptr;
  stage_cap = 0;
  PIHM_HIP_TRY(hipHostMalloc((void **)&stage_h, cap, hipHostMallocDefault));
  PIHM_HIP_TRY(hipMalloc((void **)&stage_d, cap));
  stage_cap = cap;
  return PIHM_OK;
}

int pihm_ensemble::upload_info(int k) {
  PIHM_HIP_TRY(hipMemcpyAsync(info_dev + k, &info[k], sizeof(MemberInfo), hipMemcpyHostToDevice, st));
  PIHM_HIP_TRY(hipStreamSynchronize(st));
  return PIHM_OK;
}

namespace {
struct RoundPlan {
  bool rhs;
  size_t off_part, off_cpre, off_obeg, off_ops, off_io, off_rmem, off_rinfo, off_bm, off_bmm;
  int np, nblocks, nred;
  int nact, nbe, nbr;
  double bytes;      // algorithmic bytes of the round's operations
};

// Algorithmic bytes of operation o over member M: the vectors it reads and
// writes, 8 bytes per entry each (what k_ens_ops loads and stores).
double op_bytes(const Op &o, const MemberInfo &M) {
  const double n = 8.0 * M.neq, ne = (double)M.ne;
  switch (o.kind) {
    case OP_LIN: return 3 * n;
    case OP_SCALE: return 2 * n;
    case OP_CONST: return n;
    case OP_PREDICT: return n * (2 * o.a0 + 1 + ((o.a1 & 1) ? 1 : 0) + ((o.a1 & 2) ? 2 : 0));
    case OP_RESTORE: return n * (2 * o.a0 + 1);
    case OP_RESCALE: return n * (2 * o.a0);
    case OP_CORRECT: return n * (1 + 2 * (o.a0 + 1));
    case OP_GM_START: return 6 * n;
    case OP_GM_PREP: return n * (o.a0 + 5);
    case OP_GM_ARNOLDI: return n * (o.a0 + 5);
    case OP_GM_ORTH: return n * (o.a0 + 3);
    case OP_GM_FINISH: return n * ((o.a0 == 0 ? o.a1 + 3 : (o.a0 == 1 ? 4 : 3)) + 2);
    case OP_ORDSEL: return n * (1 + ((o.a1 & 1) ? 1 : 0) + ((o.a1 & 2) ? 3 : 0));
    case OP_DKY: return n * (o.a0 + 2);
    case OP_FORCING: return 28.0 * ne;   // station index, three step inputs written
    case OP_WS0: return 16.0 * ne;
    default: return 0.0;
  }
}
inline size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }
}  // namespace

// Runs every queued item of every member in rounds, then synchronises.
int pihm_ensemble::flush() {
  std::vector<size_t> pos(nmem, 0);
  std::vector<RoundPlan> plans;
  std::vector<char> buf;
  auto reserve = [&](size_t bytes) {
    const size_t off = align_up(buf.size(), 64);
    buf.resize(off + bytes);
    return off;
  };
  for (;;) {
    // vector operations of every member whose next item is one
    std::vector<int> part, cpre{0}, obeg{0}, rmem, rinfo;
    std::vector<Op> ops;
    double round_bytes = 0.0;
    for (int m = 0; m < nmem; m++) {
      auto &Q = mem[m].q_items;
      if (pos[m] >= Q.size() || Q[pos[m]].rhs) continue;
      part.push_back(m);
      bool red = false;
      while (pos[m] < Q.size() && !Q[pos[m]].rhs) {
        if (Q[pos[m]].red && red) break;   // one reduction per member per launch
        if (Q[pos[m]].red) {
          const Op &o = Q[pos[m]].op;
          rmem.push_back(m);
          rinfo.push_back(o.nsum);
          rinfo.push_back(o.minfirst);
          rinfo.push_back(o.rslot);
          red = true;
        }
        ops.push_back(Q[pos[m]].op);
        round_bytes += op_bytes(Q[pos[m]].op, info[m]);
        pos[m]++;
      }
      obeg.push_back((int)ops.size());
      cpre.push_back(cpre.back() + info[m].nch);
    }
    bool any = false;
    if (!part.empty()) {
      RoundPlan r{};
      r.rhs = false;
      r.bytes = round_bytes;
      r.np = (int)part.size();
      r.nblocks = cpre.back();
      r.off_part = reserve(part.size() * sizeof(int));
      std::memcpy(buf.data() + r.off_part, part.data(), part.size() * sizeof(int));
      r.off_cpre = reserve(cpre.size() * sizeof(int));
      std::memcpy(buf.data() + r.off_cpre, cpre.data(), cpre.size() * sizeof(int));
      r.off_obeg = reserve(obeg.size() * sizeof(int));
      std::memcpy(buf.data() + r.off_obeg, obeg.data(), obeg.size() * sizeof(int));
      r.off_ops = reserve(ops.size() * sizeof(Op));
      std::memcpy(buf.data() + r.off_ops, ops.data(), ops.size() * sizeof(Op));
      r.nred = (int)rmem.size();
      r.off_rmem = reserve(std::max<size_t>(1, rmem.size()) * sizeof(int));
      if (!rmem.empty()) std::memcpy(buf.data() + r.off_rmem, rmem.data(), rmem.size() * sizeof(int));
      r.off_rinfo = reserve(std::max<size_t>(1, rinfo.size()) * sizeof(int));
      if (!rinfo.empty()) std::memcpy(buf.data() + r.off_rinfo, rinfo.data(), rinfo.size() * sizeof(int));
      plans.push_back(r);
      any = true;
    }
    // RHS evaluations of every member whose next item is one
    std::vector<MemberIO> io(nmem);
    std::memset(io.data(), 0, nmem * sizeof(MemberIO));
    std::vector<int> act;
    for (int m = 0; m < nmem; m++) {
      auto &Q = mem[m].q_items;
      if (pos[m] >= Q.size() || !Q[pos[m]].rhs) continue;
      io[m].y = P(Q[pos[m]].in, m);
      io[m].ydot = P(Q[pos[m]].out, m);
      io[m].active = 1;
      pos[m]++;
      act.push_back(m);
    }
    if (!act.empty()) {
      // block map of the active members: element and river blocks
      const int na = (int)act.size();
      std::vector<int> bm(2 * (na + 1), 0);
      for (int i = 0; i < na; i++) {
        const auto &M = ctx->members[act[i]];
        bm[i + 1] = bm[i] + div_up(M.ne, kBlock);
        bm[(na + 1) + i + 1] = bm[(na + 1) + i] + div_up(M.nr, kBlock);
      }
      RoundPlan r{};
      r.rhs = true;
      r.nact = na;
      r.nbe = bm[na];
      r.nbr = bm[(na + 1) + na];
      r.off_io = reserve(nmem * sizeof(MemberIO));
      std::memcpy(buf.data() + r.off_io, io.data(), nmem * sizeof(MemberIO));
      r.off_bm = reserve(bm.size() * sizeof(int));
      std::memcpy(buf.data() + r.off_bm, bm.data(), bm.size() * sizeof(int));
      r.off_bmm = reserve(act.size() * sizeof(int));
      std::memcpy(buf.data() + r.off_bmm, act.data(), act.size() * sizeof(int));
      plans.push_back(r);
      any = true;
    }
    if (!any) break;
  }
  for (auto &M : mem) M.q_items.clear();
  if (plans.empty()) return PIHM_OK;
  int rc = ensure_stage(buf.size());
  if (rc) return rc;
  std::memcpy(stage_h, buf.data(), buf.size());
  PIHM_HIP_TRY(hipMemcpyAsync(stage_d, stage_h, buf.size(), hipMemcpyHostToDevice, st));
  int ntimed = 0;
  double timed_bytes[kEvPairs];
  for (const auto &r : plans) {
    if (r.rhs) {
      rc = pihm_rhs_launch_compact(ctx, (const MemberIO *)(stage_d + r.off_io), mnan,
                                   (const int *)(stage_d + r.off_bm), (const int *)(stage_d + r.off_bmm),
                                   r.nact, r.nbe, r.nbr);
      if (rc) return rc;
      rhs_launches++;
    } else {
      Round rd;
      rd.part = (const int *)(stage_d + r.off_part);
      rd.cpre = (const int *)(stage_d + r.off_cpre);
      rd.obeg = (const int *)(stage_d + r.off_obeg);
      rd.ops = (const Op *)(stage_d + r.off_ops);
      rd.np = r.np;
      rd.rmem = (const int *)(stage_d + r.off_rmem);
      rd.rinfo = (const int *)(stage_d + r.off_rinfo);
      rd.nred = r.nred;
      rd.nchunks = r.nblocks;
      rd.g = ops_g;
      const bool timed = ntimed < kEvPairs && ev_ops[2 * ntimed];
      if (timed) PIHM_HIP_TRY(hipEventRecord(ev_ops[2 * ntimed], st));
      launch_ops(dev, rd, st, timed ? ev_ops[2 * ntimed + 1] : nullptr);
      PIHM_HIP_TRY(hipGetLastError());
      if (timed) timed_bytes[ntimed++] = r.bytes;
      op_launches++;
    }
  }
  // wait for the tick: spin on the count the signal kernel writes to host
  // memory, handing over to the runtime's wait after 2 ms
  launch_signal(sig_ctr, sig_flag, st);
  PIHM_HIP_TRY(hipGetLastError());
  const unsigned want = ++sig_expect;
  const auto t0 = std::chrono::steady_clock::now();
  for (unsigned it = 1; *sig_flag != want; it++) {
    if ((it & 255u) == 0 && std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(2)) {
      PIHM_HIP_TRY(hipStreamSynchronize(st));
      if (*sig_flag != want) return PIHM_ERR_HIP;
      break;
    }
    __builtin_ia32_pause();
  }
  std::atomic_thread_fence(std::memory_order_acquire);
  for (int k = 0; k < ntimed; k++) {
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, ev_ops[2 * k], ev_ops[2 * k + 1]) == hipSuccess) {
      ops_ms += ms;
      ops_bytes += timed_bytes[k];
      ops_timed++;
    }
  }
  ticks++;
  return PIHM_OK;
}

namespace {

// The model-step loop of pihm_drive for one member (src/main.c:217-254,
// src/pihm.c:10-187, AdjCVodeMaxStep src/ode.c:601-661), up to model step k_end.
Task<int> drive(Member &m, const pihm_run_config &cfg, long k_end) {
  const long et_every = std::max(1L, (long)std::llround(cfg.etstep / cfg.stepsize));
  int flag = 0;
  while (m.ksteps < k_end) {
    const long k = m.ksteps;
    const double t = k * cfg.stepsize;
    if (k % et_every == 0) {
      Op f = mk(OP_FORCING);
      f.s[0] = t;
      put(m, f, false);
    }
    const double tout = (double)((k + 1) * cfg.stepsize);
    if (m.nst > 0 && (tout - m.tn) * m.h < 0.0) { flag = PIHM_CV_ILL_INPUT; break; }
    m.tstop = tout;
    m.tstopset = true;
    flag = co_await cvode(m, tout);
    if (flag < 0) break;
    put(m, mk(OP_WS0), false);
    m.ksteps++;
    m.model_steps++;
    const double dsteps = (double)(m.nst - m.nst0);
    const double nfails = (double)(m.ncfn - m.ncfn0) / dsteps;
    const double niters = (double)(m.nni - m.nni0) / dsteps;
    if (nfails > cfg.nncfn || niters >= cfg.nnimax) m.maxstep /= cfg.decr;
    if (nfails == 0.0 && niters <= cfg.nnimin) m.maxstep *= cfg.incr;
    m.maxstep = (m.maxstep < cfg.stepsize) ? m.maxstep : cfg.stepsize;
    m.maxstep = (m.maxstep > cfg.stmin) ? m.maxstep : cfg.stmin;
    m.hmax_inv = (m.maxstep == 0.0) ? 0.0 : 1.0 / m.maxstep;
    m.nst0 = m.nst; m.ncfn0 = m.ncfn; m.nni0 = m.nni;
  }
  if (flag > 0) flag = 0;   // CV_TSTOP_RETURN ends each model step
  co_return flag;
}

void reset_member(Member &m, const pihm_run_config &cfg) {
  const int id = m.id;
  const long neq = m.neq;
  const double *res = m.res;
  const bool en = m.enabled;
  m = Member();
  m.id = id; m.neq = neq; m.res = res; m.enabled = en;
  m.sqrtN = std::sqrt((double)neq);
  m.reltol = cfg.reltol;
  m.abstol = cfg.abstol;
  m.hin = cfg.initstep;
  m.maxstep = cfg.stepsize;
  m.hmax_inv = 1.0 / cfg.stepsize;
  m.mxstep = cfg.max_steps_per_call > 0 ? cfg.max_steps_per_call : (long)(cfg.stepsize * 10);
  m.maxl = (cfg.maxl <= 0 || cfg.maxl > kMaxL) ? kMaxL : cfg.maxl;
}

bool cfg_ok(const pihm_run_config *cfg) {
  return cfg && cfg->stepsize > 0.0 && cfg->etstep > 0.0 && cfg->initstep > 0.0 && cfg->maxl <= kMaxL;
}

bool forcing_ok(const pihm_forcing &f, int ne) {
  if (f.nrec < 2 || f.nstation < 1 || !(f.rec_dt > 0.0) || !f.prcp || !f.edir || !f.ett || !f.station) return false;
  for (int i = 0; i < ne; i++)
    if (f.station[i] < 0 || f.station[i] >= f.nstation) return false;
  return true;
}

// Allocate an ensemble over every slot of a context, slot k able to hold
// forcing records of up to nrec_cap[k] records and nst_cap[k] stations.
int ens_alloc(pihm_ctx *c, const int *nrec_cap, const int *nst_cap, const pihm_run_config *cfg,
              pihm_ensemble **out) {
  const int nmem = c->nmem;
  pihm_ensemble *e = new (std::nothrow) pihm_ensemble();
  if (!e) return PIHM_ERR_ARG;
  e->ctx = c;
  e->st = c->stream;
  e->cfg = *cfg;
  e->nmem = nmem;
  e->info.resize(nmem);
  e->nrec_cap.assign(nrec_cap, nrec_cap + nmem);
  e->nst_cap.assign(nst_cap, nst_cap + nmem);
  long fb = 0;
  int rb = 0, cb = 0;
  for (int k = 0; k < nmem; k++) {
    const auto &M = c->members[k];
    MemberInfo &I = e->info[k];
    std::memset(&I, 0, sizeof(I));
    I.yoff = M.yoff;
    I.eoff = M.eoff;
    I.cbase = cb;
    cb += (3 * c->caps[k].ne + 2 * c->caps[k].nr + kChunk - 1) / kChunk;
    I.fbase = fb;
    fb += (long)nst_cap[k] * nrec_cap[k];
    I.rec_off = rb;
    rb += nrec_cap[k];
    I.rec_dt = 1.0;
  }
  e->total_chunks = cb;
  e->stride = ((long)c->neq + 31) / 32 * 32 + 32;   // 256-byte aligned vectors
  auto fail = [&](int r) { pihm_ens_destroy(e); return r; };
  if (hipMalloc((void **)&e->vecs, (size_t)NVEC * e->stride * sizeof(double)) != hipSuccess) return fail(PIHM_ERR_HIP);
  if (hipMemsetAsync(e->vecs, 0, (size_t)NVEC * e->stride * sizeof(double), e->st) != hipSuccess) return fail(PIHM_ERR_HIP);
  if (hipMalloc((void **)&e->partials, (size_t)std::max(cb, 1) * kSums * sizeof(double)) != hipSuccess) return fail(PIHM_ERR_HIP);
  if (hipHostMalloc((void **)&e->result, (size_t)nmem * kRes * sizeof(double), hipHostMallocCoherent) != hipSuccess)
    return fail(PIHM_ERR_HIP);
  std::memset(e->result, 0, (size_t)nmem * kRes * sizeof(double));
  if (hipHostMalloc((void **)&e->mnan, nmem * sizeof(int), hipHostMallocCoherent) != hipSuccess) return fail(PIHM_ERR_HIP);
  if (hipMalloc((void **)&e->sig_ctr, sizeof(unsigned)) != hipSuccess) return fail(PIHM_ERR_HIP);
  for (auto &ev : e->ev_ops)
    if (hipEventCreate(&ev) != hipSuccess) return fail(PIHM_ERR_HIP);
  if (hipMemset(e->sig_ctr, 0, sizeof(unsigned)) != hipSuccess) return fail(PIHM_ERR_HIP);
  {
    unsigned *f = nullptr;
    if (hipHostMalloc((void **)&f, sizeof(unsigned), hipHostMallocCoherent) != hipSuccess) return fail(PIHM_ERR_HIP);
    *f = 0u;
    e->sig_flag = f;
  }
  std::memset(e->mnan, 0, nmem * sizeof(int));
  if (hipMalloc((void **)&e->info_dev, nmem * sizeof(MemberInfo)) != hipSuccess) return fail(PIHM_ERR_HIP);
  if (hipMemcpy(e->info_dev, e->info.data(), nmem * sizeof(MemberInfo), hipMemcpyHostToDevice) != hipSuccess)
    return fail(PIHM_ERR_HIP);
  if (hipMalloc((void **)&e->station, (size_t)std::max(c->ne, 1) * sizeof(int32_t)) != hipSuccess ||
      hipMalloc((void **)&e->prcp, (size_t)std::max(fb, 1L) * sizeof(double)) != hipSuccess ||
      hipMalloc((void **)&e->fedir, (size_t)std::max(rb, 1) * sizeof(double)) != hipSuccess ||
      hipMalloc((void **)&e->fett, (size_t)std::max(rb, 1) * sizeof(double)) != hipSuccess)
    return fail(PIHM_ERR_HIP);
  if (hipMemsetAsync(e->station, 0, (size_t)std::max(c->ne, 1) * sizeof(int32_t), e->st) != hipSuccess)
    return fail(PIHM_ERR_HIP);
  Dev &d = e->dev;
  d.base = e->vecs;
  d.stride = e->stride;
  d.mem = e->info_dev;
  d.partials = e->partials;
  d.result = e->result;
  d.pcpdrp = c->args.e.pcpdrp;
  d.edir = c->args.e.edir;
  d.ett = c->args.e.ett;
  d.ws0s = c->args.e.ws0s;
  d.station = e->station;
  d.prcp = e->prcp;
  d.fedir = e->fedir;
  d.fett = e->fett;
  e->mem.resize(nmem);
  e->status.assign(nmem, ST_EMPTY);
  e->k_end.assign(nmem, 0);
  e->tops.resize(nmem);
  for (int k = 0; k < nmem; k++) {
    Member &m = e->mem[k];
    m.id = k;
    m.res = e->result + (long)k * kRes;
  }
  int rc = e->ensure_stage(1 << 20);
  if (rc) return fail(rc);
  if (hipStreamSynchronize(e->st) != hipSuccess) return fail(PIHM_ERR_HIP);
  *out = e;
  return PIHM_OK;
}

// Forcing records of slot k (station of each element in device order).
int ens_set_forcing(pihm_ensemble *e, int k, const pihm_forcing &f) {
  pihm_ctx *c = e->ctx;
  const auto &M = c->members[k];
  if (!forcing_ok(f, M.ne)) return PIHM_ERR_ARG;
  if (f.nrec > e->nrec_cap[k] || f.nstation > e->nst_cap[k]) return PIHM_ERR_TOO_LARGE;
  MemberInfo &I = e->info[k];
  I.nrec = f.nrec;
  I.nstation = f.nstation;
  I.rec_dt = f.rec_dt;
  std::vector<int32_t> station(std::max(M.ne, 1));
  for (int lp = 0; lp < M.ne; lp++) station[lp] = f.station[c->perm[M.eoff + lp]];
  hipStream_t st = e->st;
  if (M.ne > 0)
    PIHM_HIP_TRY(hipMemcpyAsync(e->station + M.eoff, station.data(), M.ne * sizeof(int32_t), hipMemcpyHostToDevice, st));
  PIHM_HIP_TRY(hipMemcpyAsync(e->prcp + I.fbase, f.prcp, (size_t)f.nstation * f.nrec * sizeof(double),
                              hipMemcpyHostToDevice, st));
  PIHM_HIP_TRY(hipMemcpyAsync(e->fedir + I.rec_off, f.edir, f.nrec * sizeof(double), hipMemcpyHostToDevice, st));
  PIHM_HIP_TRY(hipMemcpyAsync(e->fett + I.rec_off, f.ett, f.nrec * sizeof(double), hipMemcpyHostToDevice, st));
  PIHM_HIP_TRY(hipStreamSynchronize(st));
  return PIHM_OK;
}

// Sizes of slot k after its context slot was (re)loaded.
void ens_sync_sizes(pihm_ensemble *e, int k) {
  const auto &M = e->ctx->members[k];
  MemberInfo &I = e->info[k];
  I.neq = M.neq;
  I.ne = M.ne;
  I.nch = (M.neq + kChunk - 1) / kChunk;
  e->mem[k].neq = M.neq;
}

// Start (or continue) member k's model-step loop up to model step ke.
void ens_start(pihm_ensemble *e, int k, long ke) {
  Member &m = e->mem[k];
  e->k_end[k] = ke;
  e->tops[k] = drive(m, e->cfg, ke);
  m.waiting = e->tops[k].h;
  e->status[k] = ST_RUNNING;
}

}  // namespace

// Advance the running members by at most max_ticks synchronisations; stops
// early after a tick in which a member finished when stop_on_finish is set.
// Returns the number of members that finished or failed, or an error < 0.
static int ens_advance(pihm_ensemble *e, long max_ticks, bool stop_on_finish, int *nfailed) {
  int nfin = 0, nfail = 0;
  for (long t = 0; t < max_ticks; t++) {
    bool any = false;
    for (int k = 0; k < e->nmem; k++) {
      Member &m = e->mem[k];
      if (e->status[k] != ST_RUNNING || !m.waiting) continue;
      auto h = m.waiting;
      m.waiting = nullptr;
      h.resume();
      any = true;
      if (e->tops[k].done()) {
        m.flag = e->tops[k].result();
        e->tops[k] = Task<int>();
        if (m.flag < 0) { m.failed = true; e->status[k] = ST_FAILED; nfail++; }
        else e->status[k] = ST_DONE;
        nfin++;
      }
    }
    bool queued = false;
    for (int k = 0; k < e->nmem; k++) queued = queued || !e->mem[k].q_items.empty();
    if (!any && !queued) break;
    const int rc = e->flush();
    if (rc) return rc;
    for (int k = 0; k < e->nmem; k++) {
      if (!e->mnan[k]) continue;
      e->mnan[k] = 0;
      Member &m = e->mem[k];
      if (e->status[k] == ST_FAILED) continue;
      if (e->status[k] == ST_RUNNING) { e->tops[k] = Task<int>(); nfin++; }
      m.failed = true;
      m.flag = PIHM_ERR_NAN;
      m.waiting = nullptr;
      e->status[k] = ST_FAILED;
      nfail++;
    }
    if (stop_on_finish && nfin > 0) break;
  }
  if (nfailed) *nfailed = nfail;
  return nfin;
}

extern "C" {

int pihm_ens_create(pihm_ctx *c, const pihm_forcing *forc, const pihm_run_config *cfg, pihm_ensemble **out) {
  if (!c || !forc || !out || !cfg_ok(cfg)) return PIHM_ERR_ARG;
  const int nmem = c->nmem;
  std::vector<int> nrec(nmem), nst(nmem);
  for (int k = 0; k < nmem; k++) {
    if (c->members[k].ne == 0 || !forcing_ok(forc[k], c->members[k].ne)) return PIHM_ERR_ARG;
    nrec[k] = forc[k].nrec;
    nst[k] = forc[k].nstation;
  }
  pihm_ensemble *e = nullptr;
  int rc = ens_alloc(c, nrec.data(), nst.data(), cfg, &e);
  if (rc) return rc;
  for (int k = 0; k < nmem; k++) {
    ens_sync_sizes(e, k);
    if ((rc = ens_set_forcing(e, k, forc[k])) || (rc = e->upload_info(k))) { pihm_ens_destroy(e); return rc; }
    reset_member(e->mem[k], e->cfg);
    e->status[k] = ST_READY;
  }
  *out = e;
  return PIHM_OK;
}

int pihm_ens_create_slots(pihm_ctx *c, const int32_t *nrec_cap, const int32_t *nstation_cap,
                          const pihm_run_config *cfg, pihm_ensemble **out) {
  if (!c || !nrec_cap || !nstation_cap || !out || !cfg_ok(cfg)) return PIHM_ERR_ARG;
  std::vector<int> nrec(c->nmem), nst(c->nmem);
  for (int k = 0; k < c->nmem; k++) {
    if (nrec_cap[k] < 2 || nstation_cap[k] < 1) return PIHM_ERR_ARG;
    nrec[k] = nrec_cap[k];
    nst[k] = nstation_cap[k];
  }
  return ens_alloc(c, nrec.data(), nst.data(), cfg, out);
}

void pihm_ens_destroy(pihm_ensemble *e) {
  if (!e) return;
  (void)hipStreamSynchronize(e->st);
  e->tops.clear();
  (void)hipFree(e->vecs);
  (void)hipFree(e->partials);
  if (e->result) (void)hipHostFree(e->result);
  if (e->mnan) (void)hipHostFree(e->mnan);
  if (e->sig_ctr) (void)hipFree(e->sig_ctr);
  if (e->sig_flag) (void)hipHostFree((void *)e->sig_flag);
  for (auto &ev : e->ev_ops)
    if (ev) (void)hipEventDestroy(ev);
  (void)hipFree(e->info_dev);
  (void)hipFree(e->station);
  (void)hipFree(e->prcp);
  (void)hipFree(e->fedir);
  (void)hipFree(e->fett);
  if (e->stage_h) (void)hipHostFree(e->stage_h);
  (void)hipFree(e->stage_d);
  delete e;
}

int pihm_ens_set_tuning(pihm_ensemble *e, int tile_span, int g) {
  if (!e || tile_span < 0 || (g != 1 && g != 2 && g != 4 && g != 8)) return PIHM_ERR_ARG;
  e->ctx->args.tile_span = tile_span;
  e->ops_g = g;
  return PIHM_OK;
}

int pihm_ens_autotune(pihm_ensemble *e, int reps, pihm_ens_tuning *out) {
  if (!e || reps <= 0) return PIHM_ERR_ARG;
  std::vector<int> act;
  for (int m = 0; m < e->nmem; m++) {
    if (e->status[m] == ST_RUNNING) return PIHM_ERR_ARG;
    if (e->status[m] == ST_READY) act.push_back(m);
  }
  if (act.empty()) return PIHM_ERR_ARG;
  pihm_ctx *c = e->ctx;
  const int na = (int)act.size();
  // staged tables: the RHS member table and block map, then one round of
  // vector operations (VTEMP = 1*ZN0; YTEMP = VTEMP - ZN0; YTEMP = .5 YTEMP + .5 VTEMP)
  std::vector<MemberIO> io(e->nmem);
  std::memset(io.data(), 0, e->nmem * sizeof(MemberIO));
  std::vector<int> bm(2 * (na + 1), 0), cpre{0}, obeg{0};
  std::vector<Op> ops;
  for (int i = 0; i < na; i++) {
    const int m = act[i];
    io[m].y = e->P(ZN0, m);
    io[m].ydot = e->P(VTEMP, m);
    io[m].active = 1;
    const auto &M = c->members[m];
    bm[i + 1] = bm[i] + div_up(M.ne, kBlock);
    bm[(na + 1) + i + 1] = bm[(na + 1) + i] + div_up(M.nr, kBlock);
    Op o{};
    o.kind = OP_SCALE; o.a0 = ZN0; o.a2 = VTEMP; o.s[0] = 1.0;
    ops.push_back(o);
    o = Op{};
    o.kind = OP_LIN; o.a0 = VTEMP; o.a1 = ZN0; o.a2 = YTEMP; o.s[0] = 1.0; o.s[1] = -1.0;
    ops.push_back(o);
    o = Op{};
    o.kind = OP_LIN; o.a0 = YTEMP; o.a1 = VTEMP; o.a2 = YTEMP; o.s[0] = 0.5; o.s[1] = 0.5;
    ops.push_back(o);
    obeg.push_back((int)ops.size());
    cpre.push_back(cpre.back() + e->info[m].nch);
  }
  std::vector<char> buf;
  auto put = [&](const void *src, size_t bytes) {
    const size_t off = (buf.size() + 63) / 64 * 64;
    buf.resize(off + std::max<size_t>(bytes, 1));
    if (bytes) std::memcpy(buf.data() + off, src, bytes);
    return off;
  };
  const size_t o_io = put(io.data(), io.size() * sizeof(MemberIO));
  const size_t o_bm = put(bm.data(), bm.size() * sizeof(int));
  const size_t o_act = put(act.data(), act.size() * sizeof(int));
  const size_t o_cpre = put(cpre.data(), cpre.size() * sizeof(int));
  const size_t o_obeg = put(obeg.data(), obeg.size() * sizeof(int));
  const size_t o_ops = put(ops.data(), ops.size() * sizeof(Op));
  int rc = e->ensure_stage(buf.size());
  if (rc) return rc;
  std::memcpy(e->stage_h, buf.data(), buf.size());
  PIHM_HIP_TRY(hipMemcpyAsync(e->stage_d, e->stage_h, buf.size(), hipMemcpyHostToDevice, e->st));
  double *save = nullptr;
  if (c->nr > 0) PIHM_HIP_TRY(hipMalloc(&save, 2L * c->nr * sizeof(double)));
  hipStream_t cs = c->stream;
  c->stream = e->st;
  rc = pihm_ctx_carried_copy(c, save, 0);
  hipEvent_t ev[2] = {nullptr, nullptr};
  for (auto &x : ev)
    if (!rc && hipEventCreate(&x) != hipSuccess) rc = PIHM_ERR_HIP;
  auto timed = [&](auto launch, double *ms) -> int {
    for (int w = 0; w < 2; w++) if (int r = launch()) return r;
    PIHM_HIP_TRY(hipEventRecord(ev[0], e->st));
    for (int it = 0; it < reps; it++) if (int r = launch()) return r;
    PIHM_HIP_TRY(hipEventRecord(ev[1], e->st));
    PIHM_HIP_TRY(hipEventSynchronize(ev[1]));
    float f = 0.0f;
    PIHM_HIP_TRY(hipEventElapsedTime(&f, ev[0], ev[1]));
    *ms = f / reps;
    return PIHM_OK;
  };
  static const int spans[PIHM_TUNE_NSPAN] = PIHM_TUNE_SPANS;
  pihm_ens_tuning t{};
  double best = HUGE_VAL;
  t.tile_span = c->args.tile_span;
  for (int k = 0; k < PIHM_TUNE_NSPAN && !rc; k++) {
    c->args.tile_span = spans[k];
    rc = timed([&]() {
      return pihm_rhs_launch_compact(c, (const MemberIO *)(e->stage_d + o_io), e->mnan,
                                     (const int *)(e->stage_d + o_bm), (const int *)(e->stage_d + o_act), na,
                                     bm[na], bm[(na + 1) + na]);
    }, &t.rhs_ms[k]);
    if (!rc && t.rhs_ms[k] < best) { best = t.rhs_ms[k]; t.tile_span = spans[k]; }
  }
  c->args.tile_span = t.tile_span;
  Round rd{};
  rd.part = (const int *)(e->stage_d + o_act);
  rd.cpre = (const int *)(e->stage_d + o_cpre);
  rd.obeg = (const int *)(e->stage_d + o_obeg);
  rd.ops = (const Op *)(e->stage_d + o_ops);
  rd.np = na;
  rd.rmem = rd.part;
  rd.rinfo = rd.part;
  rd.nred = 0;
  rd.nchunks = cpre.back();
  best = HUGE_VAL;
  t.ops_chunks_per_wg = e->ops_g;
  for (int k = 0; k < kTuneNG && !rc; k++) {
    rd.g = kTuneG[k];
    rc = timed([&]() {
      launch_ops(e->dev, rd, e->st);
      return hipGetLastError() == hipSuccess ? (int)PIHM_OK : (int)PIHM_ERR_HIP;
    }, &t.ops_ms[k]);
    if (!rc && t.ops_ms[k] < best) { best = t.ops_ms[k]; t.ops_chunks_per_wg = kTuneG[k]; }
  }
  e->ops_g = t.ops_chunks_per_wg;
  if (!rc) rc = pihm_ctx_carried_copy(c, save, 1);
  if (!rc && hipStreamSynchronize(e->st) != hipSuccess) rc = PIHM_ERR_HIP;
  c->stream = cs;
  for (int m : act) e->mnan[m] = 0;
  for (auto &x : ev) if (x) (void)hipEventDestroy(x);
  if (save) (void)hipFree(save);
  if (out) *out = t;
  return rc;
}

int pihm_ens_set_state(pihm_ensemble *e, int k, const double *y) {
  if (!e || !y || k < 0 || k >= e->nmem) return PIHM_ERR_ARG;
  if (e->status[k] == ST_RUNNING || e->status[k] == ST_EMPTY) return PIHM_ERR_ARG;
  pihm_ctx *c = e->ctx;
  const auto &M = c->members[k];
  std::vector<double> d(M.neq);
  for (int b = 0; b < 3; b++)
    for (int lp = 0; lp < M.ne; lp++) d[(long)b * M.ne + lp] = y[(long)b * M.ne + c->perm[M.eoff + lp]];
  for (long i = 3L * M.ne; i < M.neq; i++) d[i] = y[i];
  PIHM_HIP_TRY(hipMemcpyAsync(e->P(YOUT, k), d.data(), M.neq * sizeof(double), hipMemcpyHostToDevice, e->st));
  PIHM_HIP_TRY(hipMemcpyAsync(e->P(ZN0, k), d.data(), M.neq * sizeof(double), hipMemcpyHostToDevice, e->st));
  // ws0 = initial surface state; carried river-edge fluxes start at 0 (InitVar)
  PIHM_HIP_TRY(hipMemcpyAsync(c->args.e.ws0s + M.eoff, d.data(), M.ne * sizeof(double), hipMemcpyHostToDevice, e->st));
  if (M.nr > 0) {
    PIHM_HIP_TRY(hipMemsetAsync(c->args.r.hid_ovl + 2L * M.roff, 0, 2L * M.nr * sizeof(double), e->st));
  }
  PIHM_HIP_TRY(hipStreamSynchronize(e->st));
  reset_member(e->mem[k], e->cfg);
  e->status[k] = ST_READY;
  return PIHM_OK;
}

int pihm_ens_load_member(pihm_ensemble *e, int k, const pihm_mesh_desc *mesh, const pihm_forcing *forc,
                         const double *y0, double t_end) {
  if (!e || !mesh || !forc || !y0 || k < 0 || k >= e->nmem) return PIHM_ERR_ARG;
  if (e->status[k] == ST_RUNNING) return PIHM_ERR_ARG;
  const double ks = t_end / e->cfg.stepsize;
  const long ke = std::llround(ks);
  if (ke < 1 || std::fabs(ks - (double)ke) > 1e-9 * ks) return PIHM_ERR_ARG;
  if (mesh->stepsize != e->cfg.stepsize) return PIHM_ERR_ARG;
  e->status[k] = ST_EMPTY;
  int rc = pihm_ctx_load_member(e->ctx, k, mesh);
  if (rc) return rc;
  ens_sync_sizes(e, k);
  if ((rc = ens_set_forcing(e, k, *forc))) return rc;
  if ((rc = e->upload_info(k))) return rc;
  e->status[k] = ST_READY;
  if ((rc = pihm_ens_set_state(e, k, y0))) return rc;
  e->mem[k].enabled = true;
  e->mem[k].failed = false;
  ens_start(e, k, ke);
  return PIHM_OK;
}

int pihm_ens_get_state(pihm_ensemble *e, int k, double *y) {
  if (!e || !y || k < 0 || k >= e->nmem || e->status[k] == ST_EMPTY) return PIHM_ERR_ARG;
  pihm_ctx *c = e->ctx;
  const auto &M = c->members[k];
  std::vector<double> d(M.neq);
  PIHM_HIP_TRY(hipMemcpyAsync(d.data(), e->P(YOUT, k), M.neq * sizeof(double), hipMemcpyDeviceToHost, e->st));
  PIHM_HIP_TRY(hipStreamSynchronize(e->st));
  for (int b = 0; b < 3; b++)
    for (int lp = 0; lp < M.ne; lp++) y[(long)b * M.ne + c->perm[M.eoff + lp]] = d[(long)b * M.ne + lp];
  for (long i = 3L * M.ne; i < M.neq; i++) y[i] = d[i];
  return PIHM_OK;
}

int pihm_ens_set_member_enabled(pihm_ensemble *e, int k, int on) {
  if (!e || k < 0 || k >= e->nmem) return PIHM_ERR_ARG;
  e->mem[k].enabled = on != 0;
  return PIHM_OK;
}

int pihm_ens_run(pihm_ensemble *e, double t_end) {
  if (!e) return PIHM_ERR_ARG;
  const double ks = t_end / e->cfg.stepsize;
  const long ke = std::llround(ks);
  if (std::fabs(ks - (double)ke) > 1e-9 * std::max(1.0, std::fabs(ks)) || ke < 0) return PIHM_ERR_ARG;
  for (int k = 0; k < e->nmem; k++) {
    Member &m = e->mem[k];
    if (e->status[k] == ST_RUNNING || e->status[k] == ST_EMPTY || e->status[k] == ST_FAILED) continue;
    if (!m.enabled || m.failed || m.ksteps >= ke) continue;
    ens_start(e, k, ke);
  }
  int nfail_total = 0;
  for (;;) {
    int nf = 0;
    const int r = ens_advance(e, 1L << 40, false, &nf);
    if (r < 0) return r;
    nfail_total += nf;
    bool running = false;
    for (int k = 0; k < e->nmem; k++) running = running || e->status[k] == ST_RUNNING;
    if (!running) break;
  }
  return nfail_total;
}

int pihm_ens_advance(pihm_ensemble *e, long max_ticks, int *nfailed) {
  if (!e || max_ticks < 1) return PIHM_ERR_ARG;
  return ens_advance(e, max_ticks, true, nfailed);
}

int pihm_ens_member_status(const pihm_ensemble *e, int k) {
  if (!e || k < 0 || k >= e->nmem) return PIHM_ERR_ARG;
  return e->status[k];
}

int pihm_ens_member_stats(pihm_ensemble *e, int k, pihm_run_stats *s) {
  if (!e || !s || k < 0 || k >= e->nmem) return PIHM_ERR_ARG;
  const Member &m = e->mem[k];
  std::memset(s, 0, sizeof(*s));
  s->nst = m.nst; s->nfe = m.nfe; s->nfes = m.nfes; s->nni = m.nni; s->nli = m.nli;
  s->ncfn = m.ncfn; s->netf = m.netf; s->model_steps = m.model_steps; s->flag = m.flag;
  return PIHM_OK;
}

int pihm_ens_op_stats(pihm_ensemble *e, long *launches, double *bytes, double *ms) {
  if (!e) return PIHM_ERR_ARG;
  if (launches) *launches = e->ops_timed;
  if (bytes) *bytes = e->ops_bytes;
  if (ms) *ms = e->ops_ms;
  return PIHM_OK;
}

int pihm_ens_counters(pihm_ensemble *e, long *ticks, long *op_launches, long *rhs_launches) {
  if (!e) return PIHM_ERR_ARG;
  if (ticks) *ticks = e->ticks;
  if (op_launches) *op_launches = e->op_launches;
  if (rhs_launches) *rhs_launches = e->rhs_launches;
  return PIHM_OK;
}

}  // extern "C"
