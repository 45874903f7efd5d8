// physics.hpp - MM-PIHM hydrology flux functions for the device.
//
// Each function follows the reference serial C code in operation order (file
// and lines at each definition) so that, compiled with -ffp-contract=off,
// every add/mul/div/sqrt rounds exactly as the reference does on the host.
// Powers use xpow (exp(y ln x), within a few ulp of the host pow) and the
// channel-geometry logarithm the device log.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

#include "xmath.hpp"

namespace pihm {

// src/include/pihm_const.h:8,74-79
constexpr double GRAV = 9.80665;
constexpr double PSIMIN = -70.0;
constexpr double DEPRSTG = 1E-4;
constexpr double GRADMIN = 5E-8;
constexpr double SATMIN = 0.1;
constexpr double RIVDPTHMIN = 0.05;
constexpr double RIVGRADMIN = 0.05;
constexpr int KINEMATIC = 1;
constexpr int DIFF_WAVE = 2;
// river outlet conditions (src/include/pihm_const.h:140-143)
constexpr int DIRICHLET = -1, NEUMANN = -2, ZERO_DPTH_GRAD = -3, CRIT_DPTH = -4;
// river shapes (src/include/pihm_const.h:149-152)
constexpr int RECTANGLE = 1, TRIANGLE = 2, QUADRATIC = 3, CUBIC = 4;

#define PIHM_DEV __device__ __forceinline__

// src/ode.c:31-33 clamp: `(y >= 0.0) ? y : 0.0` (keeps -0.0 as the reference)
PIHM_DEV double clamp0(double v) { return (v >= 0.0) ? v : 0.0; }

// SurfH: src/hydrol.c:100-134
PIHM_DEV double surf_h(double surfeqv) {
  double surfh;
  if (surfeqv < 0.0) surfh = 0.0;
  else if (surfeqv <= 0.5 * DEPRSTG) surfh = sqrt(2.0 * DEPRSTG * surfeqv);
  else surfh = DEPRSTG + (surfeqv - 0.5 * DEPRSTG);
  return surfh;
}

// KrFunc: src/soil.c:8-13.  The two bracketed factors are the same
// expression; it is evaluated once and squared in the written order
// (a * b * b with b == b), which rounds identically.
PIHM_DEV double kr_func(double beta, double satn) {
  double f = 1.0 - xpow(1.0 - xpow(satn, beta / (beta - 1.0)), (beta - 1.0) / beta);
  return sqrt(satn) * f * f;
}

// Psi: src/vert_flow.c:272-278
PIHM_DEV double psi_vg(double satn, double alpha, double beta) {
  satn = (satn < SATMIN) ? SATMIN : satn;
  return -xpow(xpow(1.0 / satn, beta / (beta - 1.0)) - 1.0, 1.0 / beta) / alpha;
}

// EffKh: src/lat_flow.c:248-277
PIHM_DEV double eff_kh(double depth, double dmac, double kmach, double areafv,
                       double ksath, double gw) {
  gw = (gw > 0.0) ? gw : 0.0;
  if (gw > depth - dmac) {
    double k1 = kmach * areafv + ksath * (1.0 - areafv);
    double k2 = ksath;
    double d1, d2;
    if (gw > depth) { d1 = dmac; d2 = depth - dmac; }
    else { d1 = gw - (depth - dmac); d2 = depth - dmac; }
    return (k1 * d1 + k2 * d2) / (d1 + d2);
  }
  return ksath;
}

// AvgKv, arithmetic mean (_ARITH_): src/vert_flow.c:172-209
PIHM_DEV double avg_kv(double dmac, double ksatv, double kmacv, double areafh,
                       double deficit, double gw, double satkfunc) {
  double k1, k2, k3, d1, d2, d3;
  if (deficit > dmac) {
    k1 = satkfunc * ksatv; d1 = dmac;
    k2 = satkfunc * ksatv; d2 = deficit - dmac;
    k3 = ksatv; d3 = gw;
  } else {
    k1 = satkfunc * ksatv; d1 = deficit;
    k2 = (areafh > 0.0) ? kmacv * areafh + ksatv * (1.0 - areafh) : ksatv;
    d2 = dmac - deficit;
    k3 = ksatv; d3 = gw - (dmac - deficit);
  }
  return (k1 * d1 + k2 * d2 + k3 * d3) / (d1 + d2 + d3);
}

// EffKinf: src/vert_flow.c:211-270
PIHM_DEV double eff_kinf(double kinfv, double kmacv, double areafh, double dh_by_dz,
                         double ksatfunc, double elemsatn, double applrate,
                         double surfh) {
  const double BETA_CRACK = 2.0;
  double keff;
  if (areafh == 0.0) {
    keff = kinfv * ksatfunc;
  } else if (surfh > DEPRSTG) {
    keff = kinfv * (1.0 - areafh) * ksatfunc + kmacv * areafh;
  } else if (applrate <= dh_by_dz * kinfv * ksatfunc) {
    keff = kinfv * ksatfunc;
  } else {
    double kmax = dh_by_dz * (kmacv * areafh + kinfv * (1.0 - areafh) * ksatfunc);
    if (applrate < kmax)
      keff = kinfv * (1.0 - areafh) * ksatfunc +
             kmacv * areafh * kr_func(BETA_CRACK, elemsatn);
    else
      keff = kinfv * (1.0 - areafh) * ksatfunc + kmacv * areafh;
  }
  return keff;
}

// DhByDl: src/lat_flow.c:239-246
PIHM_DEV double dh_by_dl(double a0, double a1, double a2, double b0, double b1,
                         double b2, double h0, double h1, double h2) {
  return -1.0 * (a2 * (h1 - h0) + a1 * (h0 - h2) + a0 * (h2 - h1)) /
         (b2 * (a1 - a0) + b1 * (a0 - a2) + b0 * (a2 - a1));
}

// OverLandFlow: src/lat_flow.c:279-283 (literal exponent 0.6666667)
PIHM_DEV double overland_flow(double avg_h, double grad_h, double avg_sf,
                              double crossa, double avg_rough) {
  return crossa * xpow(avg_h, 0.6666667) * grad_h / (sqrt(avg_sf) * avg_rough);
}

// AvgHsurf: src/lat_flow.c:187-215
PIHM_DEV double avg_hsurf(double diff, double hsurf, double hnabr) {
  if (diff > 0.0) return (hsurf > DEPRSTG) ? 1.0 * (hsurf - DEPRSTG) : 0.0;
  return (hnabr > DEPRSTG) ? 1.0 * (hnabr - DEPRSTG) : 0.0;
}

// AvgH: src/lat_flow.c:217-237
PIHM_DEV double avg_h_sub(double diff, double hsub, double hnabr) {
  double avg_h = 0.0;
  if (diff > 0.0) { if (hsub > 0.0) avg_h = hsub; }
  else { if (hnabr > 0.0) avg_h = hnabr; }
  return avg_h;
}

// RiverCroSectArea: src/river_flow.c:520-548 (unknown shapes are rejected
// when the context is built, where the reference exits at run time)
PIHM_DEV double riv_cs_area(int order, double depth, double coeff) {
  depth = (depth > 0.0) ? depth : 0.0;
  switch (order) {
    case RECTANGLE: return depth * coeff;
    case TRIANGLE: return depth * depth / coeff;
    case QUADRATIC: return 4.0 * depth * sqrt(depth) / (3.0 * sqrt(coeff));
    default: return 3.0 * xpow(depth, 4.0 / 3.0) / (2.0 * xpow(coeff, 1.0 / 3.0));
  }
}

// RiverPerim: src/river_flow.c:550-583
PIHM_DEV double riv_perim(int order, double depth, double coeff) {
  depth = (depth > 0.0) ? depth : 0.0;
  switch (order) {
    case RECTANGLE: return 2.0 * depth + coeff;
    case TRIANGLE: return 2.0 * depth * sqrt(1.0 + coeff * coeff) / coeff;
    case QUADRATIC:
      return sqrt(depth * (1.0 + 4.0 * coeff * depth) / coeff) +
             log(2.0 * sqrt(coeff * depth) + sqrt(1.0 + 4.0 * coeff * depth)) /
                 (2.0 * coeff);
    default:
      return 2.0 * ((xpow(depth * (1.0 + 9.0 * xpow(coeff, 2.0 / 3.0) * depth), 0.5) / 3.0) +
                    (log(3.0 * xpow(coeff, 1.0 / 3.0) * sqrt(depth) +
                         xpow(1.0 + 9.0 * xpow(coeff, 2.0 / 3.0) * depth, 0.5)) /
                     (9.0 * xpow(coeff, 1.0 / 3.0))));
  }
}

// OvlFlowElemToRiver: src/river_flow.c:185-255
PIHM_DEV double ovl_elem_to_river(double rzmax, double ezmax, double surfh, double zbed,
                                  double stage, double cwr, double length) {
  double zbank = (rzmax > ezmax) ? rzmax : ezmax;
  double elem_h = ezmax + surfh;
  double rivseg_h = zbed + stage;
  double flux;
  if (rivseg_h > elem_h) {
    if (elem_h > zbank)
      flux = cwr * 2.0 * sqrt(2.0 * GRAV) * length * sqrt(rivseg_h - elem_h) *
             (rivseg_h - zbank) / 3.0;
    else if (zbank < rivseg_h)
      flux = cwr * 2.0 * sqrt(2.0 * GRAV) * length * sqrt(rivseg_h - zbank) *
             (rivseg_h - zbank) / 3.0;
    else
      flux = 0.0;
  } else if (surfh > DEPRSTG) {
    if (rivseg_h > zbank)
      flux = -cwr * 2.0 * sqrt(2.0 * GRAV) * length * sqrt(elem_h - rivseg_h) *
             (elem_h - zbank) / 3.0;
    else if (zbank < elem_h)
      flux = -cwr * 2.0 * sqrt(2.0 * GRAV) * length * sqrt(elem_h - zbank) *
             (elem_h - zbank) / 3.0;
    else
      flux = 0.0;
  } else {
    flux = 0.0;
  }
  return flux;
}

// ChanFlowElemToRiver: src/river_flow.c:429-460 (bank element -> channel).
// Evaluated by the river segment for both banks and by each bank element for
// its own river edge, from the same operands, so both get the same bits.
PIHM_DEV double chan_elem_to_river(double stage, double zbed, double length, double rksath, double ec_gw,
                                   double ec_gwz, double ezmin, double effk, double distance) {
  const double diff_h = (stage + zbed) - ec_gwz;
  double avg_h;
  if (ezmin > zbed) avg_h = ec_gw;
  else if (ec_gwz > zbed) avg_h = ec_gwz - zbed;
  else avg_h = 0.0;
  avg_h = avg_h_sub(diff_h, stage, avg_h);
  const double grad_h = diff_h / distance;
  const double avg_ksat = 0.5 * (effk + rksath);
  return length * avg_ksat * grad_h * avg_h;
}

// SubFlowElemToRiver (_ARITH_): src/river_flow.c:462-498 (bank aquifer ->
// river aquifer), shared in the same way.
PIHM_DEV double sub_elem_to_river(double rgw, double rzmin, double zbed, double length, double ec_gw,
                                  double ec_gwz, double ezmin, double effk, double effk_riv, double distance) {
  const double diff_h = (rgw + rzmin) - ec_gwz;
  double avg_h;
  if (ezmin > zbed) avg_h = 0.0;
  else if (ec_gwz > zbed) avg_h = zbed - ezmin;
  else avg_h = ec_gw;
  avg_h = avg_h_sub(diff_h, rgw, avg_h);
  const double avg_ksat = 0.5 * (effk + effk_riv);
  const double grad_h = diff_h / distance;
  return length * avg_ksat * grad_h * avg_h;
}

}  // namespace pihm
