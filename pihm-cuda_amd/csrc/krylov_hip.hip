// krylov_hip.hip - fused Newton/SPGMR vector work for device vectors.
//
// The SPGMR iteration of CVODE (cvode_spgmr.c:355-420, sundials_spgmr.c:165-425)
// restructured for the device: each Arnoldi step is one element-wise kernel
// fused with all of its inner products (classical Gram-Schmidt, the
// orthogonalisation itself applied lazily when the next basis vector is
// formed), so a Krylov iteration costs one synchronisation instead of one per
// inner product.  Element-wise arithmetic keeps the operation order of the
// N_Vector sequence it replaces.
#include <hip/hip_runtime.h>

#include "redux.hpp"

using namespace pihm;

namespace {

inline double *D(N_Vector v) { return ((N_VectorContent_Hip)v->content)->data; }
inline long L(N_Vector v) { return ((N_VectorContent_Hip)v->content)->length; }
inline hipStream_t S(N_Vector v) { return (hipStream_t)((N_VectorContent_Hip)v->content)->stream; }

constexpr int kMaxBasis = 11;   // maxl + 1 basis vectors (maxl <= 10)

// The fused functors come in two parts, load(i) and apply(i, loaded, acc),
// so that k_fused_sum can issue two entries' loads before either entry's
// stores (a store ahead of a load through another pointer keeps the load
// from being hoisted above it).
struct FStart {
  double rl1, gamma;
  const double *zn1, *acor, *f, *s;
  double *b, *v0;
  struct In { double zn1, acor, f, s; };
  __device__ __forceinline__ In load(long i) const { return In{zn1[i], acor[i], f[i], s[i]}; }
  __device__ __forceinline__ void apply(long i, const In &a, double (&acc)[1]) const {
    const double t = (rl1 * a.zn1) + a.acor;        // LinearSum(rl1, zn1, 1, acor)
    const double bi = (gamma * a.f) - t;            // LinearSum(gamma, f, -1, t)
    b[i] = bi;
    const double v = a.s * bi;                      // Prod(s1, vtemp)
    v0[i] = v;
    acc[0] += v * v;
  }
  __device__ __forceinline__ void operator()(long i, double (&acc)[1]) const { apply(i, load(i), acc); }
};

struct PrepArgs {
  int l;
  double c, sig;
  double *vl;
  const double *vk[kMaxBasis];
  double coef[kMaxBasis];
  const double *y, *s;
  double *work;
  double *vt_out;                                   // right preconditioning: V_l / s, no work vector
};

__global__ void k_prep(long n, PrepArgs a) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double v = a.vl[i];
  for (int k = 0; k < a.l; k++) v = (-a.coef[k] * a.vk[k][i]) + v;
  v = a.c * v;                                      // Scale(1/h, V_l)
  a.vl[i] = v;
  const double vt = v / a.s[i];                     // Div(V_l, s2)
  if (a.vt_out) a.vt_out[i] = vt;
  else a.work[i] = (a.sig * vt) + a.y[i];           // LinearSum(sig, v, 1, y)
}

// k_prep with its coefficients read from the device copy of the previous
// step's sums (kry): the same values the host computes for pihm_gm_prep, in
// the same order, so a captured Krylov iteration needs no host arguments.
__global__ void k_prep_dev(long n, int l, const double *kry, PrepArgs a) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double c;
  double coef[kMaxBasis];
  if (l == 0) {
    c = 1.0 / sqrt(kry[PIHM_KRY_START]);            // 1 / r_norm
  } else {
    const double *h = kry + (long)(l - 1) * PIHM_KRY_LD;
    double hsq = 0.0;
    for (int k = 0; k < l; k++) { coef[k] = h[k]; hsq += h[k] * h[k]; }
    double nn = h[l] - hsq;
    if (nn < 0.0) nn = 0.0;
    c = 1.0 / sqrt(nn);                             // 1 / H[l][l-1]
  }
  double v = a.vl[i];
  for (int k = 0; k < l; k++) v = (-coef[k] * a.vk[k][i]) + v;
  v = c * v;
  a.vl[i] = v;
  const double vt = v / a.s[i];
  if (a.vt_out) a.vt_out[i] = vt;
  else a.work[i] = (a.sig * vt) + a.y[i];
}

template <int NV>
struct FArnoldi {
  const double *vl, *s, *jf, *fy;
  const double *zp;                                 // preconditioned vector (right preconditioning)
  double siginv, gamma;
  const double *gp;                                 // gamma from device memory when set
  double *w;
  const double *vk[kMaxBasis];
  struct In { double g, si, v, jf, fy, vk[NV - 1]; };
  __device__ __forceinline__ In load(long i) const {
    In a;
    a.g = gp ? *gp : gamma;
    a.si = s[i];
    a.v = zp ? zp[i] : vl[i];
    a.jf = jf[i];
    a.fy = fy[i];
#pragma unroll
    for (int k = 0; k < NV - 1; k++) a.vk[k] = vk[k][i];
    return a;
  }
  __device__ __forceinline__ void apply(long i, const In &a, double (&acc)[NV]) const {
    const double vt = zp ? a.v : a.v / a.si;        // the vector atimes is applied to
    const double jv = siginv * (a.jf - a.fy);       // LinearSum(siginv, Jv, -siginv, fy)
    const double z = (-a.g * jv) + vt;              // LinearSum(1, v, -gamma, Jv)
    const double wi = a.si * z;                     // Prod(s1, z)
    w[i] = wi;
#pragma unroll
    for (int k = 0; k < NV - 1; k++) acc[k] += a.vk[k] * wi;
    acc[NV - 1] += wi * wi;
  }
  __device__ __forceinline__ void operator()(long i, double (&acc)[NV]) const { apply(i, load(i), acc); }
};

template <int NV>
struct FOrth {
  double *w;
  const double *vk[kMaxBasis];
  double coef[kMaxBasis];
  __device__ __forceinline__ void operator()(long i, double (&acc)[NV]) const {
    double wi = w[i];
#pragma unroll
    for (int k = 0; k < NV - 1; k++) wi = (-coef[k] * vk[k][i]) + wi;
    w[i] = wi;
#pragma unroll
    for (int k = 0; k < NV - 1; k++) acc[k] += vk[k][i] * wi;
    acc[NV - 1] += wi * wi;
  }
};

struct FFinish {
  int mode, k;
  double yc[kMaxBasis];
  const double *vk[kMaxBasis];
  const double *s, *b, *zn0, *xsrc;
  double *acor, *y;
  // basis entries loaded ahead: those of the default SPGMR dimension (5);
  // a larger dimension reads the rest where they are used
  static constexpr int kPre = 6;
  struct In { double xin, s, acor, zn0, vk[kPre]; };
  __device__ __forceinline__ In load(long i) const {
    In a;
    a.xin = (mode == 3) ? xsrc[i] : (mode == 1 ? b[i] : 0.0);
#pragma unroll
    for (int j = 0; j < kPre; j++) a.vk[j] = (mode == 0 && j < k) ? vk[j][i] : 0.0;
    a.s = s[i];
    a.acor = acor[i];
    a.zn0 = zn0[i];
    return a;
  }
  __device__ __forceinline__ void apply(long i, const In &a, double (&acc)[2]) const {
    double x;
    if (mode == 3) {
      x = 0.0 + a.xin;                                       // x = 0 + P^-1 (xcor / s2)
    } else if (mode == 0) {
      double xc = 0.0;
#pragma unroll
      for (int j = 0; j < kMaxBasis; j++)
        if (j < k) xc += yc[j] * (j < kPre ? a.vk[j] : vk[j][i]);   // LinearSum(yg, V, 1, xcor, xcor)
      xc = xc / a.s;                                         // Div(xcor, s2)
      x = 0.0 + xc;                                          // x = 0 + vtemp
    } else if (mode == 1) {
      x = a.xin;
    } else {
      x = 0.0;
    }
    const double si = a.s;
    const double ac = a.acor + x;                            // LinearSum(1, acor, 1, b, acor)
    acor[i] = ac;
    y[i] = a.zn0 + ac;                                       // LinearSum(1, zn0, 1, acor, y)
    const double p = x * si, q = ac * si;
    acc[0] += p * p;
    acc[1] += q * q;
  }
  __device__ __forceinline__ void operator()(long i, double (&acc)[2]) const { apply(i, load(i), acc); }
};

// xcor = sum_j y_j V_j (LinearSum(y_j, V_j, 1, xcor, xcor) from xcor = 0), then
// Div(xcor, s2): the operand of the right preconditioner's final solve.
struct CombineArgs {
  int k;
  double yc[kMaxBasis];
  const double *vk[kMaxBasis];
  const double *s;
  double *out;
};

__global__ void k_combine(long n, CombineArgs a) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double xc = 0.0;
  for (int j = 0; j < a.k; j++) xc += a.yc[j] * a.vk[j][i];
  a.out[i] = xc / a.s[i];
}

// sums == nullptr: launch only (gamma from kry, nothing awaited)
template <int NV>
int arnoldi_n(N_Vector *V, int l, N_Vector s, N_Vector jf, N_Vector fy, double siginv, double gamma,
              double *sums, double *kry, N_Vector z) {
  FArnoldi<NV> f;
  f.vl = D(V[l]); f.s = D(s); f.jf = D(jf); f.fy = D(fy);
  f.zp = z ? D(z) : nullptr;
  f.siginv = siginv; f.gamma = gamma;
  f.gp = sums ? nullptr : kry + PIHM_KRY_GAMMA;
  f.w = D(V[l + 1]);
  for (int k = 0; k <= l; k++) f.vk[k] = D(V[k]);
  double *dc = kry ? kry + (long)l * PIHM_KRY_LD : nullptr;
  if (!sums) {
    fused_sum_launch<NV>(L(s), f, pihm_nv_redwork(s), S(s), dc);
    return hipGetLastError() == hipSuccess ? 0 : -2;
  }
  return fused_sum<NV>(L(s), f, pihm_nv_redwork(s), S(s), sums, dc);
}

template <int NV>
int orth_n(N_Vector *V, int l, const double *coef, double *sums, double *kry) {
  FOrth<NV> f;
  f.w = D(V[l + 1]);
  for (int k = 0; k <= l; k++) { f.vk[k] = D(V[k]); f.coef[k] = coef[k]; }
  double *dc = kry ? kry + (long)l * PIHM_KRY_LD : nullptr;
  return fused_sum<NV>(L(V[0]), f, pihm_nv_redwork(V[0]), S(V[0]), sums, dc);
}

}  // namespace

extern "C" {

int pihm_gm_start(double rl1, N_Vector zn1, N_Vector acor, double gamma, N_Vector f, N_Vector s,
                  N_Vector b, N_Vector v0, double *sumsq, double *kry) {
  FStart k{rl1, gamma, D(zn1), D(acor), D(f), D(s), D(b), D(v0)};
  // the second pass also leaves gamma in kry for the replayed Krylov iterations
  return fused_sum<1>(L(s), k, pihm_nv_redwork(s), S(s), sumsq, kry ? kry + PIHM_KRY_START : nullptr,
                      kry ? kry + PIHM_KRY_GAMMA : nullptr, gamma);
}

int pihm_gm_prep_dev(int l, N_Vector *V, const double *kry, N_Vector y, N_Vector s, double sig, N_Vector work,
                     N_Vector vt) {
  if (l < 0 || l >= kMaxBasis) return -1;
  PrepArgs a;
  a.l = l; a.c = 0.0; a.sig = sig;
  a.vl = D(V[l]);
  for (int k = 0; k < l; k++) { a.vk[k] = D(V[k]); a.coef[k] = 0.0; }
  a.y = D(y); a.s = D(s); a.work = D(work);
  a.vt_out = vt ? D(vt) : nullptr;
  const long n = L(s);
  hipLaunchKernelGGL(k_prep_dev, dim3(div_up(n, kBlock)), dim3(kBlock), 0, S(s), n, l, kry, a);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

int pihm_gm_prep(int l, N_Vector *V, const double *coef, double c, N_Vector y, N_Vector s, double sig,
                 N_Vector work, N_Vector vt) {
  if (l < 0 || l >= kMaxBasis) return -1;
  PrepArgs a;
  a.l = l; a.c = c; a.sig = sig;
  a.vl = D(V[l]);
  for (int k = 0; k < l; k++) { a.vk[k] = D(V[k]); a.coef[k] = coef[k]; }
  a.y = D(y); a.s = D(s); a.work = D(work);
  a.vt_out = vt ? D(vt) : nullptr;
  const long n = L(s);
  hipLaunchKernelGGL(k_prep, dim3(div_up(n, kBlock)), dim3(kBlock), 0, S(s), n, a);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

int pihm_gm_arnoldi(int l, N_Vector *V, N_Vector s, N_Vector jf, N_Vector fy, double siginv,
                    double gamma, double *sums, double *kry, N_Vector z) {
  if (!sums && !kry) return -1;
  switch (l + 2) {
    case 2: return arnoldi_n<2>(V, l, s, jf, fy, siginv, gamma, sums, kry, z);
    case 3: return arnoldi_n<3>(V, l, s, jf, fy, siginv, gamma, sums, kry, z);
    case 4: return arnoldi_n<4>(V, l, s, jf, fy, siginv, gamma, sums, kry, z);
    case 5: return arnoldi_n<5>(V, l, s, jf, fy, siginv, gamma, sums, kry, z);
    case 6: return arnoldi_n<6>(V, l, s, jf, fy, siginv, gamma, sums, kry, z);
    case 7: return arnoldi_n<7>(V, l, s, jf, fy, siginv, gamma, sums, kry, z);
    case 8: return arnoldi_n<8>(V, l, s, jf, fy, siginv, gamma, sums, kry, z);
    case 9: return arnoldi_n<9>(V, l, s, jf, fy, siginv, gamma, sums, kry, z);
    case 10: return arnoldi_n<10>(V, l, s, jf, fy, siginv, gamma, sums, kry, z);
    case 11: return arnoldi_n<11>(V, l, s, jf, fy, siginv, gamma, sums, kry, z);
    case 12: return arnoldi_n<12>(V, l, s, jf, fy, siginv, gamma, sums, kry, z);
    default: return -1;
  }
}

int pihm_gm_orth(int l, N_Vector *V, const double *coef, double *sums, double *kry) {
  switch (l + 2) {
    case 2: return orth_n<2>(V, l, coef, sums, kry);
    case 3: return orth_n<3>(V, l, coef, sums, kry);
    case 4: return orth_n<4>(V, l, coef, sums, kry);
    case 5: return orth_n<5>(V, l, coef, sums, kry);
    case 6: return orth_n<6>(V, l, coef, sums, kry);
    case 7: return orth_n<7>(V, l, coef, sums, kry);
    case 8: return orth_n<8>(V, l, coef, sums, kry);
    case 9: return orth_n<9>(V, l, coef, sums, kry);
    case 10: return orth_n<10>(V, l, coef, sums, kry);
    case 11: return orth_n<11>(V, l, coef, sums, kry);
    case 12: return orth_n<12>(V, l, coef, sums, kry);
    default: return -1;
  }
}

int pihm_gm_finish(int mode, int k, const double *ycoef, N_Vector *V, N_Vector s, N_Vector b,
                   N_Vector zn0, N_Vector acor, N_Vector y, const double *xsrc, double *sums) {
  if (k < 0 || k > kMaxBasis || (mode == 3 && !xsrc)) return -1;
  FFinish f;
  f.mode = mode; f.k = k;
  for (int j = 0; j < k; j++) { f.yc[j] = ycoef[j]; f.vk[j] = D(V[j]); }
  f.s = D(s); f.b = D(b); f.zn0 = D(zn0); f.acor = D(acor); f.y = D(y); f.xsrc = xsrc;
  return fused_sum<2>(L(s), f, pihm_nv_redwork(s), S(s), sums);
}

int pihm_gm_combine(int k, const double *ycoef, N_Vector *V, N_Vector s, N_Vector out) {
  if (k < 0 || k > kMaxBasis) return -1;
  CombineArgs a;
  a.k = k;
  for (int j = 0; j < k; j++) { a.yc[j] = ycoef[j]; a.vk[j] = D(V[j]); }
  a.s = D(s); a.out = D(out);
  const long n = L(s);
  hipLaunchKernelGGL(k_combine, dim3(div_up(n, kBlock)), dim3(kBlock), 0, S(s), n, a);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

}  // extern "C"
