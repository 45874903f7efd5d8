// bdf_core.hpp - host-side coefficient logic of the BDF/Newton/SPGMR
// integrator, shared by the single-solve integrator and the ensemble.
//
// Constants and routines restate CVODE 2.9 (cvode/src/cvode/cvode.c,
// sundials_iterative.c; file:line at each routine).  The routines are
// templates over the integrator's state record, which provides the fields
// they name (q, h, tau, l, tq, Hes, givens, yg, ...).  The restated routines
// are SUNDIALS code, BSD-3 licensed by Lawrence Livermore National Security
// (Copyright (c) 2002-2016, LLNS; UCRL-CODE-155951); the licence text and
// disclaimer are in NOTICE.SUNDIALS at the package root.
#pragma once
#include <cfloat>
#include <cmath>

namespace pihm {
namespace bdf {

constexpr double UROUND = DBL_EPSILON;
// cvode.c:206-240
constexpr double FUZZ_FACTOR = 100.0;
constexpr double CORTES = 0.1;
constexpr double THRESH = 1.5;
constexpr double ETAMX1 = 10000.0, ETAMX2 = 10.0, ETAMX3 = 10.0, ETAMXF = 0.2, ETAMIN = 0.1,
                 ETACF = 0.25;
constexpr double ADDON = 0.000001;
constexpr double BIAS1 = 6.0, BIAS2 = 6.0, BIAS3 = 10.0;
constexpr double ONEPSM = 1.000001;
constexpr int SMALL_NST = 10, MXNCF = 10, MXNEF = 7, MXNEF1 = 3, SMALL_NEF = 2, LONG_WAIT = 10;
constexpr int NLS_MAXCOR = 3;
constexpr double CRDOWN = 0.3, RDIV = 2.0;
constexpr long MXSTEP_DEFAULT = 500;
constexpr int MAXL_CAP = 10;
constexpr double GS_FACTOR = 1000.0;   // sundials_iterative.c ModifiedGS
constexpr int DQ_MAX_ITERS = 3;        // cvode_spils.c:36

// internal flags (cvode.c:104-114)
enum { DO_ERROR_TEST = 2, PREDICT_AGAIN = 3, CONV_FAIL = 4, TRY_AGAIN = 5, FIRST_CALL = 6,
       PREV_CONV_FAIL = 7, PREV_ERR_FAIL = 8, RHSFUNC_RECVR = 9 };
// SPGMR outcomes (sundials_spgmr.h)
enum { SP_SUCCESS = 0, SP_RES_REDUCED = 1, SP_CONV_FAIL = 2, SP_QRFACT_FAIL = 3, SP_PSOLVE_FAIL_REC = 4,
       SP_ATIMES_FAIL_REC = 5, SP_ATIMES_FAIL_UNREC = -2, SP_PSOLVE_FAIL_UNREC = -3, SP_GS_FAIL = -4,
       SP_QRSOL_FAIL = -5 };
// linear-solver setup decisions: cvode.c:237-240 (MSBP, DGMAX), cvode_spils.h:69-70
// (CVSPILS_MSBPRE, CVSPILS_DGMAX), convfail values cvode_impl.h
constexpr int MSBP = 20;
constexpr double DGMAX = 0.3;
constexpr long SPILS_MSBPRE = 50;
constexpr double SPILS_DGMAX = 0.2;
enum { CV_NO_FAILURES = 0, CV_FAIL_BAD_J = 1, CV_FAIL_OTHER = 2 };

template <class M>
void set_tq_bdf(M *m, double hsum, double alpha0, double alpha0_hat, double xi_inv,
                double xistar_inv) {   // cvode.c:2501-2537
  const int q = m->q;
  const double A1 = 1.0 - alpha0_hat + alpha0;
  const double A2 = 1.0 + q * A1;
  m->tq[2] = std::fabs(A1 / (alpha0 * A2));
  m->tq[5] = std::fabs(A2 * xistar_inv / (m->l[q] * xi_inv));
  if (m->qwait == 1) {
    if (q > 1) {
      const double C = xistar_inv / m->l[q];
      const double A3 = alpha0 + 1.0 / q;
      const double A4 = alpha0_hat + xi_inv;
      const double Cpinv = (1.0 - A4 + A3) / A3;
      m->tq[1] = std::fabs(C * Cpinv);
    } else {
      m->tq[1] = 1.0;
    }
    hsum += m->tau[q];
    xi_inv = m->h / hsum;
    const double A5 = alpha0 - (1.0 / (q + 1));
    const double A6 = alpha0_hat - xi_inv;
    const double Cppinv = (1.0 - A6 + A5) / A2;
    m->tq[3] = std::fabs(Cppinv / (xi_inv * (q + 2) * A5));
  }
  m->tq[4] = m->nlscoef / m->tq[2];
}

template <class M>
void set_bdf(M *m) {   // cvode.c:2464-2499 then cvSet :2308-2340
  const int q = m->q;
  double xi_inv = 1.0, xistar_inv = 1.0;
  m->l[0] = m->l[1] = 1.0;
  for (int i = 2; i <= q; i++) m->l[i] = 0.0;
  double alpha0 = -1.0, alpha0_hat = -1.0;
  double hsum = m->h;
  if (q > 1) {
    for (int j = 2; j < q; j++) {
      hsum += m->tau[j - 1];
      xi_inv = m->h / hsum;
      alpha0 -= 1.0 / j;
      for (int i = j; i >= 1; i--) m->l[i] += m->l[i - 1] * xi_inv;
    }
    alpha0 -= 1.0 / q;
    xistar_inv = -m->l[1] - alpha0;
    hsum += m->tau[q - 1];
    xi_inv = m->h / hsum;
    alpha0_hat = -m->l[1] - xi_inv;
    for (int i = q; i >= 1; i--) m->l[i] += m->l[i - 1] * xistar_inv;
  }
  set_tq_bdf(m, hsum, alpha0, alpha0_hat, xi_inv, xistar_inv);
  m->rl1 = 1.0 / m->l[1];
  m->gamma = m->h * m->rl1;
  if (m->nst == 0) m->gammap = m->gamma;
  m->gamrat = (m->nst > 0) ? m->gamma / m->gammap : 1.0;
}

template <class M>
int qr_update(M *m, int n) {   // QRfact, column n-1 update (sundials_iterative.c)
  double (*h)[MAXL_CAP] = m->Hes;
  double *qv = m->givens;
  const int n1 = n - 1;
  for (int k = 0; k < n1; k++) {
    const int i = 2 * k;
    const double t1 = h[k][n1], t2 = h[k + 1][n1], c = qv[i], s = qv[i + 1];
    h[k][n1] = c * t1 - s * t2;
    h[k + 1][n1] = s * t1 + c * t2;
  }
  const double t1 = h[n1][n1], t2 = h[n][n1];
  double c, s;
  if (t2 == 0.0) { c = 1.0; s = 0.0; }
  else if (std::fabs(t2) >= std::fabs(t1)) {
    const double t3 = t1 / t2;
    s = -1.0 / std::sqrt(1.0 + t3 * t3);
    c = -s * t3;
  } else {
    const double t3 = t2 / t1;
    c = 1.0 / std::sqrt(1.0 + t3 * t3);
    s = -c * t3;
  }
  qv[2 * n1] = c;
  qv[2 * n1 + 1] = s;
  if ((h[n1][n1] = c * t1 - s * t2) == 0.0) return n;
  return 0;
}

template <class M>
int qr_solve(M *m, int n) {   // QRsol (sundials_iterative.c)
  double *b = m->yg;
  for (int k = 0; k < n; k++) {
    const double c = m->givens[2 * k], s = m->givens[2 * k + 1];
    const double t1 = b[k], t2 = b[k + 1];
    b[k] = c * t1 - s * t2;
    b[k + 1] = s * t1 + c * t2;
  }
  for (int k = n - 1; k >= 0; k--) {
    if (m->Hes[k][k] == 0.0) return k + 1;
    b[k] /= m->Hes[k][k];
    for (int i = 0; i < k; i++) b[i] -= b[k] * m->Hes[i][k];
  }
  return 0;
}

template <class M>
void set_eta(M *m) {   // cvode.c:3068-3089
  if (m->eta < THRESH) {
    m->eta = 1.0;
    m->hprime = m->h;
  } else {
    m->eta = std::fmin(m->eta, m->etamax);
    m->eta /= std::fmax(1.0, std::fabs(m->h) * m->hmax_inv * m->eta);
    m->hprime = m->h * m->eta;
    if (m->qprime < m->q) m->nscon = 0;
  }
}


}  // namespace bdf
}  // namespace pihm
