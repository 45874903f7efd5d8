// ens_kernels.hip - the batched vector-operation kernel of the ensemble
// integrator.
//
// One workgroup processes one chunk (kChunk consecutive entries) of one
// member and runs that member's queued operations over it in order.  All
// operations are index-local, so a chunk never waits on another chunk; the
// one reduction a member may have per launch is summed in a fixed tree: each
// thread over its kChunkPer entries in order, wavefront butterfly, the four
// waves in order, then (second launch, k_ens_final) the member's chunk
// partials in chunk order.  The tree depends only on the member's length.  Element-wise arithmetic repeats the rounding
// of the N_Vector operation sequences of the single-solve integrator.
#include <hip/hip_runtime.h>

#include <cfloat>

#include "common.hpp"
#include "ens.hpp"

namespace pihm {
namespace ens {
namespace {

__device__ __forceinline__ double wsum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
__device__ __forceinline__ double wmin(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmin(v, __shfl_xor(v, off, 64));
  return v;
}

// N_VLinearSum rounding: a*(x+y) if a == b, a*(x-y) if a == -b, else (a*x)+(b*y).
__device__ __forceinline__ double lin(double a, double x, double b, double y) {
  if (a == b) return a * (x + y);
  if (a == -b) return a * (x - y);
  return (a * x) + (b * y);
}

struct Ctx {
  const Dev *d;
  MemberInfo M;
  int m, c;
  long i0;
  __device__ double *P(int v) const { return d->base + (long)v * d->stride + M.yoff; }
};

// The chunk's NV per-thread sums (slot 0 a minimum when MIN0) combined in the
// fixed order wavefront butterfly, waves in order, into the chunk's partials.
template <int NV, bool MIN0>
__device__ __forceinline__ void chunk_sums(const Ctx &x, double (&acc)[NV], double *lds) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < NV; k++) {
    const double v = (MIN0 && k == 0) ? wmin(acc[k]) : wsum(acc[k]);
    if (lane == 0) lds[wave * kSums + k] = v;
  }
  __syncthreads();
  if (threadIdx.x < NV) {
    const int k = threadIdx.x;
    double v;
    if (MIN0 && k == 0) v = fmin(fmin(fmin(lds[k], lds[kSums + k]), lds[2 * kSums + k]), lds[3 * kSums + k]);
    else v = ((lds[k] + lds[kSums + k]) + lds[2 * kSums + k]) + lds[3 * kSums + k];
    x.d->partials[(long)(x.M.cbase + x.c) * kSums + k] = v;
  }
  __syncthreads();
}

// An operation over the chunk, kLoadAhead entries per thread at a time: ld(i)
// returns entry i's inputs, st(i, in) (reductions: st(i, in, acc)) computes
// and stores.  All loads of a group are issued before any of its stores.
// Operations store into one vector where another is read only at the same
// index, so loading later entries first changes no value; written this way
// the compiler need not prove it, and a wavefront keeps the group's loads in
// flight together instead of waiting for memory once per entry.
constexpr int kLoadAhead = 2;
static_assert(kChunkPer % kLoadAhead == 0, "whole groups of entries per thread");

template <class Ld, class St>
__device__ __forceinline__ void each(const Ctx &x, Ld ld, St st) {
  using In = decltype(ld(0L));
#pragma unroll 1
  for (int j0 = 0; j0 < kChunkPer; j0 += kLoadAhead) {
    In in[kLoadAhead];
#pragma unroll
    for (int u = 0; u < kLoadAhead; u++) {
      const long i = x.i0 + (long)(j0 + u) * kChunkThreads + threadIdx.x;
      if (i < x.M.neq) in[u] = ld(i);
    }
#pragma unroll
    for (int u = 0; u < kLoadAhead; u++) {
      const long i = x.i0 + (long)(j0 + u) * kChunkThreads + threadIdx.x;
      if (i < x.M.neq) st(i, in[u]);
    }
  }
}

// NV sums over the chunk (slot 0 a minimum when MIN0), each thread over its
// entries in order, into the chunk's partials.
template <int NV, bool MIN0, class Ld, class St>
__device__ __forceinline__ void reduce(const Ctx &x, Ld ld, St st, double *lds) {
  using In = decltype(ld(0L));
  double acc[NV];
#pragma unroll
  for (int k = 0; k < NV; k++) acc[k] = (MIN0 && k == 0) ? DBL_MAX : 0.0;
#pragma unroll 1
  for (int j0 = 0; j0 < kChunkPer; j0 += kLoadAhead) {
    In in[kLoadAhead];
#pragma unroll
    for (int u = 0; u < kLoadAhead; u++) {
      const long i = x.i0 + (long)(j0 + u) * kChunkThreads + threadIdx.x;
      if (i < x.M.neq) in[u] = ld(i);
    }
#pragma unroll
    for (int u = 0; u < kLoadAhead; u++) {
      const long i = x.i0 + (long)(j0 + u) * kChunkThreads + threadIdx.x;
      if (i < x.M.neq) st(i, in[u], acc);
    }
  }
  chunk_sums<NV, MIN0>(x, acc, lds);
}

struct Zn { double v[kMaxQ + 1]; };

// zn[0..q] at entry i (the others 0)
__device__ __forceinline__ Zn load_zn(double *const (&zn)[kMaxQ + 1], int q, long i) {
  Zn a;
#pragma unroll
  for (int j = 0; j <= kMaxQ; j++) a.v[j] = (j <= q) ? zn[j][i] : 0.0;
  return a;
}

// ------------------------------------------------------------------ operations
__device__ __forceinline__ void op_predict(const Ctx &x, const Op &op, double *lds) {
  const int q = op.a0;
  const bool ewt = op.a1 & 1, init = op.a1 & 2;
  double *zn[kMaxQ + 1];
#pragma unroll
  for (int j = 0; j <= kMaxQ; j++) zn[j] = x.P(ZN0 + j);
  double *w = x.P(EWT), *yc = x.P(YCOR), *ac = x.P(ACOR);
  const double rtol = op.s[0], atol = op.s[1];
  auto ld = [&](long i) { return load_zn(zn, q, i); };
  auto body = [&](long i, const Zn &a) {
    double v[kMaxQ + 1];
#pragma unroll
    for (int j = 0; j <= kMaxQ; j++) v[j] = a.v[j];
    // cvode.c:2277-2306: for k = 1..q, j = q..k: zn[j-1] += zn[j]
#pragma unroll
    for (int k = 1; k <= kMaxQ; k++)
#pragma unroll
      for (int j = kMaxQ; j >= k; j--)
        if (k <= q && j <= q) v[j - 1] = v[j - 1] + v[j];
#pragma unroll
    for (int j = 0; j < kMaxQ; j++)
      if (j < q) zn[j][i] = v[j];
    if (init) { yc[i] = v[0]; ac[i] = 0.0; }
  };
  if (ewt) {
    reduce<2, true>(x, ld, [&](long i, const Zn &a, double (&acc)[2]) {
      // cvEwtSetSS (cvode.c:4081-4098) of the zn0 the step starts from
      const double y0 = a.v[0];
      const double t = (rtol * fabs(y0)) + atol;
      const double wi = 1.0 / t;
      w[i] = wi;
      acc[0] = fmin(acc[0], t);
      const double p = y0 * wi;
      acc[1] += p * p;
      body(i, a);
    }, lds);
  } else {
    each(x, ld, body);
  }
}

__device__ __forceinline__ void op_restore(const Ctx &x, const Op &op) {
  const int q = op.a0;
  double *zn[kMaxQ + 1];
#pragma unroll
  for (int j = 0; j <= kMaxQ; j++) zn[j] = x.P(ZN0 + j);
  each(x, [&](long i) { return load_zn(zn, q, i); }, [&](long i, const Zn &a) {
    double v[kMaxQ + 1];
#pragma unroll
    for (int j = 0; j <= kMaxQ; j++) v[j] = a.v[j];
#pragma unroll
    for (int k = 1; k <= kMaxQ; k++)
#pragma unroll
      for (int j = kMaxQ; j >= k; j--)
        if (k <= q && j <= q) v[j - 1] = v[j - 1] - v[j];
#pragma unroll
    for (int j = 0; j < kMaxQ; j++)
      if (j < q) zn[j][i] = v[j];
  });
}

__device__ __forceinline__ void op_rescale(const Ctx &x, const Op &op) {
  const int q = op.a0;
  double *zn[kMaxQ + 1];
#pragma unroll
  for (int j = 0; j <= kMaxQ; j++) zn[j] = x.P(ZN0 + j);
  double f[kMaxQ + 1];
#pragma unroll
  for (int j = 0; j <= kMaxQ; j++) f[j] = op.s[j];
  each(x, [&](long i) {
    Zn a;
#pragma unroll
    for (int j = 0; j <= kMaxQ; j++) a.v[j] = (j >= 1 && j <= q) ? zn[j][i] : 0.0;
    return a;
  }, [&](long i, const Zn &a) {
#pragma unroll
    for (int j = 1; j <= kMaxQ; j++)
      if (j <= q) zn[j][i] = f[j] * a.v[j];
  });
}

__device__ __forceinline__ void op_correct(const Ctx &x, const Op &op) {
  const int q = op.a0;
  double *zn[kMaxQ + 1];
#pragma unroll
  for (int j = 0; j <= kMaxQ; j++) zn[j] = x.P(ZN0 + j);
  const double *ac = x.P(ACOR);
  double l[kMaxQ + 1];
#pragma unroll
  for (int j = 0; j <= kMaxQ; j++) l[j] = op.s[j];
  struct In { double a; Zn z; };
  each(x, [&](long i) { return In{ac[i], load_zn(zn, q, i)}; }, [&](long i, const In &in) {
    const double a = in.a;
#pragma unroll
    for (int j = 0; j <= kMaxQ; j++) {
      if (j > q) break;
      // LinearSum(l[j], acor, 1, zn[j], zn[j])
      const double lj = l[j];
      double z = in.z.v[j];
      if (lj == 1.0) z += a;
      else if (lj == -1.0) z -= a;
      else z += lj * a;
      zn[j][i] = z;
    }
  });
}

__device__ __forceinline__ void op_gm_start(const Ctx &x, const Op &op, double *lds) {
  const double rl1 = op.s[0], gamma = op.s[1];
  const double *zn1 = x.P(ZN0 + 1), *ac = x.P(ACOR), *f = x.P(FTEMP), *s = x.P(EWT);
  double *b = x.P(TEMPV), *v0 = x.P(V0);
  struct In { double zn1, ac, f, s; };
  reduce<1, false>(x, [&](long i) { return In{zn1[i], ac[i], f[i], s[i]}; },
                   [&](long i, const In &a, double (&acc)[1]) {
    const double t = (rl1 * a.zn1) + a.ac;       // LinearSum(rl1, zn1, 1, acor)
    const double bi = (gamma * a.f) - t;         // LinearSum(gamma, f, -1, t)
    b[i] = bi;
    const double v = a.s * bi;
    v0[i] = v;
    acc[0] += v * v;
  }, lds);
}

__device__ __forceinline__ void op_gm_prep(const Ctx &x, const Op &op) {
  const int l = op.a0;
  const double c = op.s[0], sig = op.s[1];
  double *vl = x.P(V0 + l);
  const double *s = x.P(EWT), *y = x.P(YCOR);
  double *work = x.P(YTEMP);
  const double *vk[kMaxL + 1];
#pragma unroll
  for (int k = 0; k <= kMaxL; k++) vk[k] = x.P(V0 + k);
  double coef[kMaxL];
#pragma unroll
  for (int k = 0; k < kMaxL; k++) coef[k] = op.s[2 + k];
  struct In { double v, vk[kMaxL], s, y; };
  each(x, [&](long i) {
    In a{};
    a.v = vl[i];
#pragma unroll
    for (int k = 0; k < kMaxL; k++)
      if (k < l) a.vk[k] = vk[k][i];
    a.s = s[i];
    a.y = y[i];
    return a;
  }, [&](long i, const In &a) {
    double v = a.v;
#pragma unroll
    for (int k = 0; k < kMaxL; k++)
      if (k < l) v = (-coef[k] * a.vk[k]) + v;
    v = c * v;                                   // Scale(1/h, V_l)
    vl[i] = v;
    const double vt = v / a.s;                   // Div(V_l, s2)
    work[i] = (sig * vt) + a.y;                  // LinearSum(sig, v, 1, y)
  });
}

template <int L>
__device__ __forceinline__ void op_gm_arnoldi(const Ctx &x, const Op &op, double *lds) {
  const double siginv = op.s[0], gamma = op.s[1];
  const double *vl = x.P(V0 + L), *s = x.P(EWT), *jf = x.P(VTEMP), *fy = x.P(FTEMP);
  double *w = x.P(V0 + L + 1);
  const double *vk[L + 1];
#pragma unroll
  for (int k = 0; k <= L; k++) vk[k] = x.P(V0 + k);
  struct In { double s, vl, jf, fy, vk[L + 1]; };
  reduce<L + 2, false>(x, [&](long i) {
    In a;
    a.s = s[i];
    a.vl = vl[i];
    a.jf = jf[i];
    a.fy = fy[i];
#pragma unroll
    for (int k = 0; k <= L; k++) a.vk[k] = vk[k][i];
    return a;
  }, [&](long i, const In &a, double (&acc)[L + 2]) {
    const double si = a.s;
    const double vt = a.vl / si;
    const double jv = siginv * (a.jf - a.fy);    // LinearSum(siginv, Jv, -siginv, fy)
    const double z = (-gamma * jv) + vt;         // LinearSum(1, v, -gamma, Jv)
    const double wi = si * z;
    w[i] = wi;
#pragma unroll
    for (int k = 0; k <= L; k++) acc[k] += a.vk[k] * wi;
    acc[L + 1] += wi * wi;
  }, lds);
}

template <int L>
__device__ __forceinline__ void op_gm_orth(const Ctx &x, const Op &op, double *lds) {
  double *w = x.P(V0 + L + 1);
  double coef[L + 1];
#pragma unroll
  for (int k = 0; k <= L; k++) coef[k] = op.s[k];
  const double *vk[L + 1];
#pragma unroll
  for (int k = 0; k <= L; k++) vk[k] = x.P(V0 + k);
  struct In { double w, vk[L + 1]; };
  reduce<L + 2, false>(x, [&](long i) {
    In a;
    a.w = w[i];
#pragma unroll
    for (int k = 0; k <= L; k++) a.vk[k] = vk[k][i];
    return a;
  }, [&](long i, const In &a, double (&acc)[L + 2]) {
    double wi = a.w;
#pragma unroll
    for (int k = 0; k <= L; k++) wi = (-coef[k] * a.vk[k]) + wi;
    w[i] = wi;
#pragma unroll
    for (int k = 0; k <= L; k++) acc[k] += a.vk[k] * wi;
    acc[L + 1] += wi * wi;
  }, lds);
}

__device__ __forceinline__ void op_gm_finish(const Ctx &x, const Op &op, double *lds) {
  const int mode = op.a0, kd = op.a1;
  const double *vk[kMaxL + 1];
#pragma unroll
  for (int k = 0; k <= kMaxL; k++) vk[k] = x.P(V0 + k);
  const double *s = x.P(EWT), *b = x.P(TEMPV), *zn0 = x.P(ZN0);
  double *ac = x.P(ACOR), *y = x.P(YCOR);
  double yc[kMaxL + 1];
#pragma unroll
  for (int j = 0; j <= kMaxL; j++) yc[j] = op.s[j];
  struct In { double vk[kMaxL + 1], s, b, ac, zn0; };
  reduce<2, false>(x, [&](long i) {
    In a{};
    if (mode == 0) {
#pragma unroll
      for (int j = 0; j <= kMaxL; j++)
        if (j < kd) a.vk[j] = vk[j][i];
    } else if (mode == 1) {
      a.b = b[i];
    }
    a.s = s[i];
    a.ac = ac[i];
    a.zn0 = zn0[i];
    return a;
  }, [&](long i, const In &in, double (&acc)[2]) {
    double xv;
    if (mode == 0) {
      double xc = 0.0;
#pragma unroll
      for (int j = 0; j <= kMaxL; j++)
        if (j < kd) xc += yc[j] * in.vk[j];                     // LinearSum(yg, V, 1, xcor, xcor)
      xc = xc / in.s;                                          // Div(xcor, s2)
      xv = 0.0 + xc;
    } else if (mode == 1) {
      xv = in.b;
    } else {
      xv = 0.0;
    }
    const double si = in.s;
    const double a = in.ac + xv;                                // LinearSum(1, acor, 1, b, acor)
    ac[i] = a;
    y[i] = in.zn0 + a;                                          // LinearSum(1, zn0, 1, acor, y)
    const double p = xv * si, r = a * si;
    acc[0] += p * p;
    acc[1] += r * r;
  }, lds);
}

__device__ __forceinline__ void op_ordsel(const Ctx &x, const Op &op, double *lds) {
  const int q = op.a0;
  const bool ddn = op.a1 & 1, dup = op.a1 & 2;
  const double cquot = op.s[0];
  const double *znq = x.P(ZN0 + q), *znm = x.P(ZN0 + kMaxQ), *ac = x.P(ACOR), *s = x.P(EWT);
  double *t = x.P(TEMPV);
  struct In { double znq, znm, ac, s; };
  reduce<2, false>(x, [&](long i) {
    In a{};
    if (ddn) a.znq = znq[i];
    if (dup) { a.znm = znm[i]; a.ac = ac[i]; }
    a.s = s[i];
    return a;
  }, [&](long i, const In &in, double (&acc)[2]) {
    const double si = in.s;
    if (ddn) { const double p = in.znq * si; acc[0] += p * p; }
    if (dup) {
      const double v = lin(-cquot, in.znm, 1.0, in.ac);        // LinearSum(-cquot, zn[qmax], 1, acor)
      t[i] = v;
      const double p = v * si;
      acc[1] += p * p;
    }
  }, lds);
}

__device__ __forceinline__ void op_dky(const Ctx &x, const Op &op) {
  const int q = op.a0;
  const double s = op.s[0];
  double *zn[kMaxQ + 1];
#pragma unroll
  for (int j = 0; j <= kMaxQ; j++) zn[j] = x.P(ZN0 + j);
  double *out = x.P(YOUT);
  each(x, [&](long i) { return load_zn(zn, q, i); }, [&](long i, const Zn &a) {
    double v = 0.0;
#pragma unroll
    for (int j = kMaxQ; j >= 0; j--) {
      if (j == q) v = a.v[j];                                   // Scale(1, zn[q], dky)
      else if (j < q) v = lin(1.0, a.v[j], s, v);               // LinearSum(1, zn[j], s, dky)
    }
    out[i] = v;
  });
}

__device__ __forceinline__ void op_forcing(const Ctx &x, const Op &op) {
  const Dev &d = *x.d;
  const MemberInfo &M = x.M;
  const double t = op.s[0];
  long k = (long)floor(t / M.rec_dt);
  if (k < 0) k = 0;
  if (k > M.nrec - 2) k = M.nrec - 2;
  const double ta = (double)k * M.rec_dt, tb = (double)(k + 1) * M.rec_dt;
  const double *E = d.fedir + M.rec_off, *T = d.fett + M.rec_off;
#pragma unroll 2
  for (int j = 0; j < kChunkPer; j++) {
    const long i = x.i0 + (long)j * kChunkThreads + threadIdx.x;
    if (i >= M.ne) continue;
    const long p = M.eoff + i;
    const double *P = d.prcp + M.fbase + (long)d.station[p] * M.nrec;
    // IntrplForc: src/forcing.c:514-519
    d.pcpdrp[p] = ((tb - t) * P[k] + (t - ta) * P[k + 1]) / (tb - ta);
    d.edir[p] = ((tb - t) * E[k] + (t - ta) * E[k + 1]) / (tb - ta);
    d.ett[p] = ((tb - t) * T[k] + (t - ta) * T[k + 1]) / (tb - ta);
  }
}

__device__ __forceinline__ void op_ws0(const Ctx &x) {
  const Dev &d = *x.d;
  const double *y = x.P(YOUT);
#pragma unroll 2
  for (int j = 0; j < kChunkPer; j++) {
    const long i = x.i0 + (long)j * kChunkThreads + threadIdx.x;
    if (i < x.M.ne) d.ws0s[x.M.eoff + i] = y[i];
  }
}

__global__ __launch_bounds__(kChunkThreads) void k_ens_ops(Dev d, Round r) {
  __shared__ double lds[4 * kSums];
  for (int gi = 0; gi < r.g; gi++) {
  const int b = blockIdx.x * r.g + gi;
  if (b >= r.nchunks) return;
  if (gi > 0) __syncthreads();   // the previous chunk's reduction is done with lds
  int lo = 0, hi = r.np - 1;   // cpre[lo] <= b < cpre[lo+1]
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (r.cpre[mid] <= b) lo = mid;
    else hi = mid - 1;
  }
  Ctx x;
  x.d = &d;
  x.m = r.part[lo];
  x.M = d.mem[x.m];
  x.c = b - r.cpre[lo];
  x.i0 = (long)x.c * kChunk;
  __shared__ long chunk_i0;
  if (threadIdx.x == 0) chunk_i0 = x.i0;
  __syncthreads();
  for (int o = r.obeg[lo]; o < r.obeg[lo + 1]; o++) {
    x.i0 = *(volatile long *)&chunk_i0;
    const Op &op = r.ops[o];
    switch (op.kind) {
      case OP_LIN: {
        const double *xa = x.P(op.a0), *ya = x.P(op.a1);
        double *z = x.P(op.a2);
        const double a = op.s[0], bb = op.s[1];
        struct In { double x, y; };
        each(x, [&](long i) { return In{xa[i], ya[i]}; },
             [&](long i, const In &v) { z[i] = lin(a, v.x, bb, v.y); });
        break;
      }
      case OP_SCALE: {
        const double *xa = x.P(op.a0);
        double *z = x.P(op.a2);
        const double c = op.s[0];
        each(x, [&](long i) { return xa[i]; }, [&](long i, double v) { z[i] = c * v; });
        break;
      }
      case OP_CONST: {
        double *z = x.P(op.a2);
        const double c = op.s[0];
        each(x, [&](long) { return 0; }, [&](long i, int) { z[i] = c; });
        break;
      }
      case OP_PREDICT: op_predict(x, op, lds); break;
      case OP_RESTORE: op_restore(x, op); break;
      case OP_RESCALE: op_rescale(x, op); break;
      case OP_CORRECT: op_correct(x, op); break;
      case OP_GM_START: op_gm_start(x, op, lds); break;
      case OP_GM_PREP: op_gm_prep(x, op); break;
      case OP_GM_ARNOLDI:
        switch (op.a0) {
          case 0: op_gm_arnoldi<0>(x, op, lds); break;
          case 1: op_gm_arnoldi<1>(x, op, lds); break;
          case 2: op_gm_arnoldi<2>(x, op, lds); break;
          case 3: op_gm_arnoldi<3>(x, op, lds); break;
          default: op_gm_arnoldi<4>(x, op, lds); break;
        }
        break;
      case OP_GM_ORTH:
        switch (op.a0) {
          case 0: op_gm_orth<0>(x, op, lds); break;
          case 1: op_gm_orth<1>(x, op, lds); break;
          case 2: op_gm_orth<2>(x, op, lds); break;
          case 3: op_gm_orth<3>(x, op, lds); break;
          default: op_gm_orth<4>(x, op, lds); break;
        }
        break;
      case OP_GM_FINISH: op_gm_finish(x, op, lds); break;
      case OP_ORDSEL: op_ordsel(x, op, lds); break;
      case OP_DKY: op_dky(x, op); break;
      case OP_FORCING: op_forcing(x, op); break;
      case OP_WS0: op_ws0(x); break;
      default: break;
    }
  }
  }
}

// Member totals: the chunk partials of member rmem[b] in chunk order.  Each
// thread loads its first partial of every sum before combining any (one
// memory latency, not one per sum); further rounds of chunks follow in order.
__global__ __launch_bounds__(kChunkThreads) void k_ens_final(Dev d, Round r) {
  __shared__ double lds[4];
  const int m = r.rmem[blockIdx.x];
  const int nsum = r.rinfo[3 * blockIdx.x], minfirst = r.rinfo[3 * blockIdx.x + 1];
  const int rslot = r.rinfo[3 * blockIdx.x + 2];
  const MemberInfo M = d.mem[m];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int c0 = threadIdx.x;
  double first[kSums];
#pragma unroll
  for (int k = 0; k < kSums; k++)
    first[k] = (k < nsum && c0 < M.nch) ? d.partials[(long)(M.cbase + c0) * kSums + k] : 0.0;
#pragma unroll
  for (int k = 0; k < kSums; k++) {
    if (k >= nsum) break;
    const bool mn = minfirst && k == 0;
    double v = mn ? DBL_MAX : 0.0;
    if (c0 < M.nch) v = mn ? fmin(v, first[k]) : v + first[k];
    for (int cc = c0 + kChunkThreads; cc < M.nch; cc += kChunkThreads) {
      const double p = d.partials[(long)(M.cbase + cc) * kSums + k];
      v = mn ? fmin(v, p) : v + p;
    }
    v = mn ? wmin(v) : wsum(v);
    if (lane == 0) lds[wave] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
      const double res = mn ? fmin(fmin(fmin(lds[0], lds[1]), lds[2]), lds[3])
                            : ((lds[0] + lds[1]) + lds[2]) + lds[3];
      d.result[(long)m * kRes + rslot + k] = res;
    }
    __syncthreads();
  }
}

}  // namespace

// End-of-tick signal: the running count of ticks goes to host memory after
// everything queued before it on the stream.
__global__ void k_ens_signal(unsigned *ctr, volatile unsigned *flag) {
  if (threadIdx.x == 0) {
    __threadfence_system();
    const unsigned s = *ctr + 1u;
    *ctr = s;
    *flag = s;   // made visible by the end of the kernel
  }
}

void launch_signal(unsigned *ctr, volatile unsigned *flag, hipStream_t st) {
  hipLaunchKernelGGL(k_ens_signal, dim3(1), dim3(64), 0, st, ctr, flag);
}

void launch_ops(const Dev &d, const Round &r, hipStream_t st, hipEvent_t after_ops) {
  const int g = r.g > 0 ? r.g : 1;
  hipLaunchKernelGGL(k_ens_ops, dim3((r.nchunks + g - 1) / g), dim3(kChunkThreads), 0, st, d, r);
  if (after_ops) (void)hipEventRecord(after_ops, st);
  if (r.nred > 0) hipLaunchKernelGGL(k_ens_final, dim3(r.nred), dim3(kChunkThreads), 0, st, d, r);
}

}  // namespace ens
}  // namespace pihm
