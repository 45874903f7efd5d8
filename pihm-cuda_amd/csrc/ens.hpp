// ens.hpp - batched vector operations of the ensemble integrator.
//
// An ensemble keeps every integrator vector (Nordsieck array, error weights,
// Newton and Krylov work vectors) of every member in one HBM allocation:
// vector v of member m is base + v*stride + yoff[m], member-major inside each
// vector, the state layout of each member unchanged.  One launch of k_ens_ops
// applies, for each participating member, that member's list of queued
// operations to its slice; the grid is the concatenation of fixed-size chunks
// of the participating members, so every member's partial sums, and therefore
// its reductions, depend on that member's data alone and never on which other
// members share the launch.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace pihm {
namespace ens {

// Vector roles (index of the vector in the ensemble allocation).
enum Vec : int {
  ZN0 = 0,            // ZN0 + j: Nordsieck column j, j = 0..5
  EWT = 6, YCOR = 7, ACOR = 8, TEMPV = 9, FTEMP = 10,
  V0 = 11,            // V0 + l: Krylov basis vector l, l = 0..5
  YTEMP = 17, VTEMP = 18, YOUT = 19,
  NVEC = 20
};
constexpr int kMaxQ = 5;
constexpr int kMaxL = 5;          // SPGMR dimension (MM-PIHM uses the default 5)
constexpr int kSums = 8;          // partial sums per chunk
constexpr int kRes = 16;          // result slots per member: 0..7 the last reduction, 8..9 error weights
constexpr int kChunkThreads = 256;
constexpr int kChunkPer = 8;      // entries per thread per chunk
constexpr int kChunk = kChunkThreads * kChunkPer;

enum OpKind : int {
  OP_LIN = 1,        // z = a*x + b*y with the N_VLinearSum rounding      (a0=x, a1=y, a2=z)
  OP_SCALE,          // z = c*x                                          (a0=x, a2=z)
  OP_CONST,          // z = c                                            (a2=z)
  OP_PREDICT,        // ewt/min/|zn0|_wrms of the old zn0, then the Pascal update of zn,
                     // then ycor = zn0, acor = 0                        (q=a0, rtol=s0, atol=s1)
  OP_RESTORE,        // inverse Pascal update                             (q=a0)
  OP_RESCALE,        // zn[j] *= f[j], j = 1..q                           (q=a0, f = s1..s5)
  OP_CORRECT,        // zn[j] += l[j]*acor, j = 0..q                      (q=a0, l = s0..s5)
  OP_GM_START,       // b = gamma*f - (rl1*zn1 + acor) into TEMPV; V0 = ewt*b; sum (ewt*b)^2
  OP_GM_PREP,        // V_l <- c*(V_l - sum_k coef_k V_k); YTEMP = sig*(V_l/ewt) + ycor
  OP_GM_ARNOLDI,     // V_{l+1} = ewt*((-gamma*(siginv*(VTEMP - FTEMP))) + V_l/ewt); dots
  OP_GM_ORTH,        // V_{l+1} -= sum_k coef_k V_k; dots
  OP_GM_FINISH,      // x (mode 0: sum yc_k V_k / ewt, 1: TEMPV, 2: 0); acor += x;
                     // ycor = zn0 + acor; sums (x*ewt)^2, (acor*ewt)^2
  OP_ORDSEL,         // sums (zn_q*ewt)^2 and, with TEMPV = acor - cquot*zn_qmax, (TEMPV*ewt)^2
  OP_DKY,            // YOUT = sum_j s^j zn_j (Horner, N_VLinearSum rounding), q = a0, s = s0
  OP_FORCING,        // per element: pcpdrp/edir/ett interpolated at t = s0
  OP_WS0             // per element: ws0.surf = YOUT[surf]
};

struct alignas(16) Op {
  int kind;
  int a0, a1, a2, a3;
  int nsum;          // reductions written by this op (0 = none)
  int rslot;         // first result slot
  int minfirst;      // slot 0 of the reduction is a minimum
  double s[16];
};

struct MemberInfo {
  long yoff;         // first entry of the member in each vector
  int neq, ne;       // state length, element count
  int eoff;          // first device element of the member in the batch context
  int nch;           // chunks of kChunk entries
  int cbase;         // first chunk of the member in the partials table
  long fbase;        // first precipitation record of the member
  int nrec, nstation;
  int rec_off;       // first record of edir/ett of the member
  double rec_dt;
};

struct Dev {
  double *base;
  long stride;
  const MemberInfo *mem;
  double *partials;          // [total chunks * kSums]
  double *result;            // pinned host [nmem * kRes]
  // element per-step inputs of the batch context (device order)
  double *pcpdrp, *edir, *ett, *ws0s;
  // forcing records
  const int32_t *station;    // [ne total], station of each device element within its member
  const double *prcp, *fedir, *fett;
};

// Launch descriptors for one round of k_ens_ops (device copies).
struct Round {
  const int *part;    // [np] participating members
  const int *cpre;    // [np+1] chunk prefix
  const int *obeg;    // [np+1] op list ranges
  const Op *ops;
  int np;
  const int *rmem;    // [nred] members with a reduction in this round
  const int *rinfo;   // [3*nred] nsum, minfirst, rslot
  int nred;
  int nchunks;        // chunks of all participants (cpre[np])
  int g;              // chunks per workgroup, run one after another (tuned; the arithmetic does not depend on it)
};

// The round's operations (nblocks = chunks of all participants), then the
// per-member totals of its reductions (a second launch, one workgroup per
// member with a reduction).
// after_ops, when not null, is recorded between the operation kernel and the
// reduction kernel (timing of the operation kernel alone).
void launch_ops(const Dev &d, const Round &r, hipStream_t st, hipEvent_t after_ops = nullptr);
constexpr int kTuneNG = 4;
constexpr int kTuneG[kTuneNG] = {1, 2, 4, 8};
// One-lane kernel writing ++(*ctr) to *flag (host memory) after the work queued before it.
void launch_signal(unsigned *ctr, volatile unsigned *flag, hipStream_t st);

}  // namespace ens
}  // namespace pihm
