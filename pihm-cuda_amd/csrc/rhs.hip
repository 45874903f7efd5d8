// rhs.hip - MM-PIHM CVODE right-hand side on MI355X (gfx950).
//
// Replaces ODE() (src/ode.c:4-399) and everything it calls: SurfH, EtExtract
// (src/hydrol.c), LateralFlow/FrictSlope (src/lat_flow.c), VerticalFlow
// (src/vert_flow.c) and RiverFlow (src/river_flow.c).
//
// One RHS call is two launches on one stream:
//   k_prep     per element: clamp the state, surface depth, EffKh, and the
//              FrictSlope gradient magnitude (neighbour heads recomputed
//              from the state); writes one 64-byte neighbour record.
//   k_flux     river blocks first: per river segment, all segment fluxes,
//              the upstream inflow (each upstream segment's outflow
//              evaluated again here, subtracted in ascending caller index
//              as the reference's serial loop does, src/river_flow.c:96-110)
//              and the river ydot; then the element blocks: per element,
//              lateral fluxes (one 64-byte gather per neighbour), EtExtract,
//              Infil, Recharge, the element's own river-edge fluxes (the
//              same expressions and operands as the river's) and the
//              element ydot.
// No atomics and no cross-thread hand-off inside k_flux: a flux needed by
// two threads is evaluated by both from the same operands, so every output
// has one writer and results are deterministic.  Apart from the power
// routine (xmath.hpp), every operation rounds as the reference's does.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <vector>

#include "common.hpp"
#include "ctx.hpp"
#include "physics.hpp"
#include "nvector_impl.hpp"

using namespace pihm;

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------

// Where element/river entries of one solve live in the state vector.  A batch
// concatenates solves (member-major state, contiguous element and river
// ranges); a single solve (B = false) has all offsets zero.
struct Ix {
  const double *y;  // state of the solve
  double *ydot;     // and its time derivative
  int ne, eoff;     // element count and first device element
  int nr, roff;     // river count and first river
  double dt;
  int active;
  int m;            // member
};

// A batch reads each member's state either at its offset in one whole-batch
// vector (mio == nullptr) or through the per-call table mio (state and
// derivative pointers plus the active flag of every member).
template <bool B>
__device__ __forceinline__ Ix ix_of(const RhsArgs &a, int member) {
  if (!B) return Ix{a.y, a.ydot, a.ne, 0, a.nr, 0, a.dt, 1, 0};
  const MemberDev M = a.mem[member];
  if (a.mio) {
    const MemberIO io = a.mio[member];
    return Ix{io.y, io.ydot, M.ne, M.eoff, M.nr, M.roff, M.dt, io.active, member};
  }
  return Ix{a.y + M.yoff, a.ydot + M.yoff, M.ne, M.eoff, M.nr, M.roff, M.dt, M.active, member};
}

template <bool B>
__device__ __forceinline__ void flag_nan(const RhsArgs &a, const Ix &ix) {
  *a.nanflag = 1;
  if (B && a.mnan) a.mnan[ix.m] = 1;
}
template <bool B>
__device__ __forceinline__ Ix ix_elem(const RhsArgs &a, int p) { return ix_of<B>(a, B ? a.emem[p] : 0); }
template <bool B>
__device__ __forceinline__ Ix ix_river(const RhsArgs &a, int r) { return ix_of<B>(a, B ? a.rmem[r] : 0); }

// Largest i in [0, n) with pre[i] <= b (pre ascending).
__device__ __forceinline__ int find_seg(const int *pre, int n, int b) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (pre[mid] <= b) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// Position handled by this thread of block `b` of a compact launch: kind 0
// elements, 1 rivers; -1 past the member's range.
__device__ __forceinline__ int compact_index(const RhsArgs &a, int kind, int b) {
  const int *pre = a.bm + kind * (a.nact + 1);
  const int i = find_seg(pre, a.nact, b);
  const MemberDev M = a.mem[a.bm_mem[i]];
  const int l = (b - pre[i]) * kBlock + threadIdx.x;
  if (kind == 0) return l < M.ne ? M.eoff + l : -1;
  return l < M.nr ? M.roff + l : -1;
}

// Loads come first, in two rounds (the element's own entries and edge codes,
// then what the codes point at), so a wavefront waits for memory twice
// instead of once per neighbour; the arithmetic follows them unchanged.
template <bool B>
__device__ __forceinline__ void river_flux(const RhsArgs &a, int r, int part);

// Grid: 3*nbr river blocks first (the three parts of every segment's fluxes,
// part-major, the longest threads of the launch), then nbe element blocks
// (the neighbour records).
template <bool B>
__global__ __launch_bounds__(kBlock) void k_prep(RhsArgs a, int nbe) {
  const int nbr3 = (int)gridDim.x - nbe, nbr = nbr3 / 3;
  if ((int)blockIdx.x < nbr3) {
    const int part = (int)blockIdx.x / nbr, rb = (int)blockIdx.x - part * nbr;
    const int r = (B && a.bm) ? compact_index(a, 1, rb) : rb * kBlock + threadIdx.x;
    if (r >= 0 && r < a.nr) river_flux<B>(a, r, part);
    return;
  }
  const int tile = xcd_tile((int)blockIdx.x - nbr3, nbe, a.tile_span);
  const int p = (B && a.bm) ? compact_index(a, 0, tile) : tile * kBlock + threadIdx.x;
  const int ne = a.ne;
  if (p < 0 || p >= ne) return;
  const Ix ix = ix_elem<B>(a, p);
  if (!ix.active) return;
  const double *y = ix.y;
  const int lp = p - ix.eoff;
  if (B && lp >= ix.ne) return;     // unused capacity of a slot
  const ElemDev &e = a.e;
  const bool diff = a.surf_mode == DIFF_WAVE;

  // first group of loads
  const double ys = y[lp], yg = y[2 * ix.ne + lp];
  const double zmax = e.zmax[p], zmin = e.zmin[p], depth = e.depth[p], dmac = e.dmac[p];
  const double kmach = e.kmach[p], areafv = e.areafv[p], ksath = e.ksath[p];
  int c[3] = {0, 0, 0};
  double nx[3] = {0.0, 0.0, 0.0}, ny[3] = {0.0, 0.0, 0.0};
  if (diff) {
#pragma unroll
    for (int j = 0; j < 3; j++) {
      c[j] = e.nb[j * ne + p];
      nx[j] = e.nabr_x[j * ne + p];
      ny[j] = e.nabr_y[j * ne + p];
    }
  }
  // second group of loads: element neighbour (zmax, surface storage), boundary value,
  // river edge (stage, depth, zbed, zmax).  Each case loads into
  // its own registers, left unset in the other cases (each is read only in
  // its own case): a register shared by two cases, or a default written into
  // it, would make a load wait for another lane's load to the same register.
  double ez[3], es[3], bv[3], rs[3], rd[3], rb[3], rz[3];
  if (diff) {
#pragma unroll
    for (int j = 0; j < 3; j++) {
      if (c[j] >= 0) {
        ez[j] = e.zmax[c[j]];
        es[j] = y[c[j] - ix.eoff];
      } else if (nb_is_bound(c[j])) {
        bv[j] = e.ebc[j * ne + p];
      } else {
        const int r = (-c[j] - 1) >> 1;
        rs[j] = y[3 * ix.ne + (r - ix.roff)];
        rd[j] = a.r.depth[r];
        rb[j] = a.r.zbed[r];
        rz[j] = a.r.zmax[r];
      }
    }
  }

  const double s = clamp0(ys);
  const double g = clamp0(yg);
  const double surfh = surf_h(s);
  const double head = zmax + surfh;
  const double gwz = g + zmin;
  const double effkh = eff_kh(depth, dmac, kmach, areafv, ksath, g);

  double sfmag = 0.0;
  if (diff) {
    // FrictSlope: src/lat_flow.c:130-185
    double h[3];
#pragma unroll
    for (int j = 0; j < 3; j++) {
      if (c[j] >= 0) {
        h[j] = ez[j] + surf_h(clamp0(es[j]));
      } else if (nb_is_bound(c[j])) {
        h[j] = (c[j] == NB_BOUND) ? zmax + surfh : bv[j];
      } else {
        const double st = clamp0(rs[j]);
        h[j] = (st > rd[j]) ? rb[j] + st : rz[j];
      }
    }
    const double dhdx = dh_by_dl(ny[0], ny[1], ny[2], nx[0], nx[1], nx[2], h[0], h[1], h[2]);
    const double dhdy = dh_by_dl(nx[0], nx[1], nx[2], ny[0], ny[1], ny[2], h[0], h[1], h[2]);
    sfmag = sqrt(dhdx * dhdx + dhdy * dhdy);
  }
  // the six per-call fields of the record, as three 16-byte stores
  double2 *rp = reinterpret_cast<double2 *>(&a.e.rec[p]);
  rp[0] = make_double2(surfh, head);
  rp[1] = make_double2(g, gwz);
  rp[2] = make_double2(effkh, sfmag);
}

// What a river segment reads of one bank element: its record fields, evaluated
// from the element's state with the same expressions as the record's.
struct BankView {
  double surfh, gw, gwz, effkh, zmax, zmin;
};

// RiverToElem for one bank (src/river_flow.c:113-183): overland (rivflow 2/3),
// channel exchange (4/5) and aquifer exchange (7/8) between segment r and the
// bank element bv on `side` (0 left, 1 right); effk_riv is the mean EffKh of
// the two banks.  Evaluated by the segment and by the bank element, from the
// same operands, so both get the same bits.
struct BankFlux {
  double ovl, chan, sub;
};
__device__ __forceinline__ BankFlux bank_flux(const RiverDev &R, int r, int side, double stage, double rgw,
                                              const BankView &bv, double effk_riv) {
  const double zbed = R.zbed[r], length = R.length[r];
  const double dist = side ? R.dist_r[r] : R.dist_l[r];
  BankFlux f;
  f.ovl = ovl_elem_to_river(R.zmax[r], bv.zmax, bv.surfh, zbed, stage, R.cwr[r], length);
  f.chan = chan_elem_to_river(stage, zbed, length, R.ksath[r], bv.gw, bv.gwz, bv.zmin, bv.effkh, dist);
  f.sub = sub_elem_to_river(rgw, R.zmin[r], zbed, length, bv.gw, bv.gwz, bv.zmin, bv.effkh, effk_riv, dist);
  return f;
}

template <bool B>
__device__ __forceinline__ double effkh_of(const RhsArgs &a, const Ix &ix, int q) {
  const ElemDev &e = a.e;
  const double g = clamp0(ix.y[2 * ix.ne + (q - ix.eoff)]);
  return eff_kh(e.depth[q], e.dmac[q], e.kmach[q], e.areafv[q], e.ksath[q], g);
}

// Where an element's per-call neighbour values come from.
//   GlobalSrc: the records k_prep wrote (codes: device element indices) and
//              the bank fluxes its river blocks left.
//   TileSrc:   the tile's records in LDS (codes: LDS slots) and the bank
//              fluxes evaluated here, from the element's own record.
struct GlobalSrc {
  static constexpr bool kGatherFirst = true;   // neighbour records from HBM: all three loads at once
  const RhsArgs &a;
  int p;
  __device__ __forceinline__ NbrRec self() const { return a.e.rec[p]; }
  __device__ __forceinline__ int code(int j) const { return a.e.nb[j * a.ne + p]; }
  __device__ __forceinline__ NbrRec nbr(int c) const { return a.e.rec[c]; }
  __device__ __forceinline__ void bank(const Ix &, int, int slot, const NbrRec &, double, double, double &ovl,
                                       double &sub) const {
    ovl = a.r.hid_ovl[2 * a.nr + slot];
    sub = a.r.bank_sub[slot];
  }
};

// The tile's records in LDS, one array of cap doubles per NbrRec field (slots
// 0 .. tile size - 1: the tile's elements; then the halo), so that lanes
// touching consecutive slots touch consecutive banks.
struct LdsRecs {
  double *f;
  int cap;
  __device__ __forceinline__ NbrRec get(int s) const {
    NbrRec r;
    r.surfh = f[s]; r.head = f[cap + s]; r.gw = f[2 * cap + s]; r.gwz = f[3 * cap + s];
    r.effkh = f[4 * cap + s]; r.sfmag = f[5 * cap + s]; r.rough = f[6 * cap + s]; r.zmax = f[7 * cap + s];
    return r;
  }
  __device__ __forceinline__ void put(int s, const NbrRec &r) const {
    f[s] = r.surfh; f[cap + s] = r.head; f[2 * cap + s] = r.gw; f[3 * cap + s] = r.gwz;
    f[4 * cap + s] = r.effkh; f[5 * cap + s] = r.sfmag; f[6 * cap + s] = r.rough; f[7 * cap + s] = r.zmax;
  }
  __device__ __forceinline__ void put_sfmag(int s, double v) const { f[5 * cap + s] = v; }
};

constexpr int kTile = 128;            // elements per tile = threads per block of the one-launch RHS (two wavefronts)

struct TileSrc {
  static constexpr bool kGatherFirst = false;  // from LDS: each edge's record when its turn comes
  const RhsArgs &a;
  LdsRecs srec;         // LDS records of the tile and its halo
  const int32_t *lnb;   // neighbour slots of the tile's elements
  int p, i;
  const double *bk;     // LDS: the bank fluxes of each tile element's first river edge (ovl, then sub)
  const int *bkj;       // LDS: that edge (-1: none)
  __device__ __forceinline__ NbrRec self() const { return srec.get(i); }
  __device__ __forceinline__ int code(int j) const { return lnb[j * a.ne + p]; }
  __device__ __forceinline__ NbrRec nbr(int c) const { return srec.get(c); }
  // edge j's bank fluxes: the first river edge's were evaluated during the
  // staging, from the same operands; a further river edge's are evaluated here
  __device__ __forceinline__ void bank(const Ix &ix, int j, int slot, const NbrRec &me, double zmax,
                                       double zmin, double &ovl, double &sub) const {
    if (bkj[i] == j) {
      ovl = bk[i];
      sub = bk[kTile + i];
      return;
    }
    bank_eval(ix, slot, me, zmax, zmin, ovl, sub);
  }
  __device__ __forceinline__ void bank_eval(const Ix &ix, int slot, const NbrRec &me, double zmax, double zmin,
                                            double &ovl, double &sub) const {
    const RiverDev &R = a.r;
    const int r = slot >> 1, side = slot & 1;
    const double *ys = ix.y + 3L * ix.ne - ix.roff;
    const double stage = clamp0(ys[r]);
    const double rgw = clamp0(ys[ix.nr + r]);
    const double other = effkh_of<false>(a, ix, side ? R.left[r] : R.right[r]);
    const BankView bv{me.surfh, me.gw, me.gwz, me.effkh, zmax, zmin};
    const double effk_riv = side ? 0.5 * (other + me.effkh) : 0.5 * (me.effkh + other);
    const BankFlux f = bank_flux(R, r, side, stage, rgw, bv, effk_riv);
    ovl = -f.ovl;
    sub = -(f.chan + f.sub);
  }
};

// An element's own inputs to elem_core: its unsaturated storage, statics and
// per-step inputs, its edges (codes from the Src, distances, lengths) and,
// by edge kind, the boundary value or the carried river-edge overland flux.
struct ElemIn {
  double yu, area, zmax, zmin, depth, dinf, rzd, edir, ett, pcpdrp, ws0;
  double alpha, beta, kinfv, kmacv, areafh, dmac, ksatv, porosity;
  double gwet, sh2o0, smcmin, fcr;   // the _NOAH_ branches only
  int c[3];
  double dist[3], edge[3];
  double bcv[3], hov[3];
};

// Loads in two groups: the element's own entries and edge codes, then what the
// codes point at (a wavefront waits for memory twice, not once per edge).
template <bool N, class Src>
__device__ __forceinline__ ElemIn load_elem_in(const RhsArgs &a, int p, const Ix &ix, const Src &src) {
  const int ne = a.ne;
  const ElemDev &e = a.e;
  ElemIn in;
  in.yu = ix.y[ix.ne + (p - ix.eoff)];
  in.area = e.area[p]; in.zmax = e.zmax[p]; in.zmin = e.zmin[p];
  in.depth = e.depth[p]; in.dinf = e.dinf[p]; in.rzd = e.rzd[p];
  in.edir = e.edir[p]; in.ett = e.ett[p]; in.pcpdrp = e.pcpdrp[p]; in.ws0 = e.ws0s[p];
  in.alpha = e.alpha[p]; in.beta = e.beta[p]; in.kinfv = e.kinfv[p]; in.kmacv = e.kmacv[p];
  in.areafh = e.areafh[p]; in.dmac = e.dmac[p]; in.ksatv = e.ksatv[p]; in.porosity = e.porosity[p];
  if (N) {
    in.gwet = e.gwet[p]; in.sh2o0 = e.sh2o0[p]; in.smcmin = e.smcmin[p]; in.fcr = e.fcr[p];
  }
#pragma unroll
  for (int j = 0; j < 3; j++) {
    in.c[j] = src.code(j);
    in.dist[j] = e.nabrdist[j * ne + p];
    in.edge[j] = e.edge[j * ne + p];
  }
  // each read only in its own case (a register shared by two cases, or a
  // default written into it, would make a load wait for another lane's)
#pragma unroll
  for (int j = 0; j < 3; j++) {
    if (in.c[j] >= 0) {
    } else if (nb_is_bound(in.c[j])) {
      in.bcv[j] = e.ebc[j * ne + p];
    } else {
      in.hov[j] = a.r.hid_ovl[-in.c[j] - 1];     // the carried overland flux
    }
  }
  return in;
}

// An element's fluxes come in three parts, which the two-kernel RHS runs one
// after the other in one thread and the one-launch RHS in two wavefronts:
//   elem_lateral   LateralFlow over the three edges (src/lat_flow.c:29-63),
//                  with the overland flux carried from the previous call on
//                  river edges, the sum Infil's application rate starts from,
//                  then this call's bank fluxes on river edges;
//   elem_vert_pre  EtExtract's split and the unsaturated zone's saturation,
//                  pressure head and conductivity;
//   elem_vert_post Infil, Recharge and the element ODEs.
// N: the reference built with _NOAH_ (src/hydrol.c:43-56,79-84;
// src/vert_flow.c:16-19,94-98), single solves only.
struct LatOut {
  double ovl[3], sub[3];   // this call's edge fluxes
  double asum;             // sum of -ovl / area with the carried fluxes on river edges
  double out;              // ovl + sub over the domain-boundary edges (closure)
};

template <bool B, class Src>
__device__ __forceinline__ LatOut elem_lateral(const RhsArgs &a, int p, const Ix &ix, const Src &src,
                                               const ElemIn &in, const NbrRec &me) {
  const int ne = a.ne;
  NbrRec nrec[3];
  if constexpr (Src::kGatherFirst) {
#pragma unroll
    for (int j = 0; j < 3; j++)
      if (in.c[j] >= 0) nrec[j] = src.nbr(in.c[j]);
  }
  const double area = in.area, zmax = in.zmax, zmin = in.zmin;
  const double g = me.gw;
  const double surfh = me.surfh;
  LatOut o;
  double *ovl = o.ovl, *sub = o.sub;
  bool bank = false;
  bool bnd[3] = {false, false, false};
#pragma unroll
  for (int j = 0; j < 3; j++) {
    const int c = in.c[j];
    bnd[j] = nb_is_bound(c);
    if (c >= 0) {
      if constexpr (!Src::kGatherFirst) nrec[j] = src.nbr(c);
      const NbrRec &nr = nrec[j];
      // SubFlowElemToElem: src/lat_flow.c:285-309
      {
        const double diff_h = me.gwz - nr.gwz;
        const double avg_h = avg_h_sub(diff_h, g, nr.gw);
        const double grad_h = diff_h / in.dist[j];
        const double avg_ksat = 0.5 * (me.effkh + nr.effkh);
        sub[j] = avg_ksat * grad_h * avg_h * in.edge[j];
      }
      // OvlFlowElemToElem: src/lat_flow.c:311-338
      {
        double avg_sf = 0.5 * (me.sfmag + nr.sfmag);
        const double diff_h = (a.surf_mode == KINEMATIC) ? zmax - nr.zmax : me.head - nr.head;
        const double avg_h = avg_hsurf(diff_h, surfh, nr.surfh);
        const double grad_h = diff_h / in.dist[j];
        if (a.surf_mode == KINEMATIC) avg_sf = (grad_h > 0.0) ? grad_h : GRADMIN;
        else avg_sf = (avg_sf > GRADMIN) ? avg_sf : GRADMIN;
        const double avg_rough = 0.5 * (me.rough + nr.rough);
        const double crossa = avg_h * in.edge[j];
        ovl[j] = overland_flow(avg_h, grad_h, avg_sf, crossa, avg_rough);
      }
    } else if (nb_is_bound(c)) {
      // BoundFluxElem: src/lat_flow.c:340-382
      ovl[j] = 0.0;
      const int bt = nb_bc_sign(c);
      if (bt == 0) {
        sub[j] = 0.0;
      } else if (bt > 0) {
        const double diff_h = me.gwz - in.bcv[j];
        const double avg_h = avg_h_sub(diff_h, g, in.bcv[j] - zmin);
        const double grad_h = diff_h / in.dist[j];
        sub[j] = me.effkh * grad_h * avg_h * in.edge[j];
      } else {
        sub[j] = -in.bcv[j];
      }
    } else {
      // river edge: the overland flux carried from the previous call is what
      // Infil reads (src/vert_flow.c:59-62); both fluxes are replaced below.
      bank = true;
      ovl[j] = in.hov[j];
      sub[j] = 0.0;
    }
  }
  // Infil's application rate starts from the lateral inflow (src/vert_flow.c:56-66)
  o.asum = 0.0;
#pragma unroll
  for (int j = 0; j < 3; j++) o.asum += -ovl[j] / area;
  // River edges: this call's bank fluxes (src/river_flow.c:113-183) replace
  // the carried overland flux that Infil uses, and the new overland flux is
  // carried to the next call (src/river_flow.c:159-182).
  if (bank) {
#pragma unroll
    for (int j = 0; j < 3; j++) {
      if (nb_is_river(in.c[j])) {
        const int slot = -in.c[j] - 1;
        src.bank(ix, j, slot, me, zmax, zmin, ovl[j], sub[j]);
        a.r.hid_ovl[slot] = ovl[j];
      }
    }
  }
  if (a.record) {
#pragma unroll
    for (int j = 0; j < 3; j++) {
      a.rec.ovl[j * ne + p] = ovl[j];
      a.rec.sub[j * ne + p] = sub[j];
    }
  }
  // volume rates crossing the element's outer boundary
  o.out = 0.0;
#pragma unroll
  for (int j = 0; j < 3; j++)
    if (bnd[j]) o.out += ovl[j] + sub[j];
  return o;
}

struct VertPre {
  double edir_s, edir_u, edir_g, ett_u, ett_g;
  double deficit, satn, psi_raw, kr;
  bool wet_bottom;
};

template <bool N>
__device__ __forceinline__ VertPre elem_vert_pre(const ElemIn &in, const NbrRec &me) {
  const double depth = in.depth, dinf = in.dinf, rzd = in.rzd;
  const double edir = in.edir, ett = in.ett;
  const double u = clamp0(in.yu);
  const double g = me.gw;
  const double surfh = me.surfh;
  VertPre v;
  // EtExtract: src/hydrol.c:33-97
  if (N) {
    // no direct evaporation from ponded water; the saturated zone's share of
    // transpiration, GwTransp (src/noah/lsm_func.c:163-185), is fixed per step
    v.edir_s = 0.0;
    if (g > depth - dinf) { v.edir_u = 0.0; v.edir_g = edir; }
    else { v.edir_u = edir; v.edir_g = 0.0; }
    const double gwet = in.gwet;
    v.ett_u = (1.0 - gwet) * ett;
    v.ett_g = gwet * ett;
  } else {
    if (surfh >= DEPRSTG) { v.edir_s = edir; v.edir_u = 0.0; v.edir_g = 0.0; }
    else if (g > depth - dinf) { v.edir_s = 0.0; v.edir_u = 0.0; v.edir_g = edir; }
    else { v.edir_s = 0.0; v.edir_u = edir; v.edir_g = 0.0; }
    if (g > depth - rzd) { v.ett_u = 0.0; v.ett_g = ett; }
    else { v.ett_u = ett; v.ett_g = 0.0; }
  }
  // Unsaturated-zone terms shared by Infil and Recharge (same deficit and
  // saturation in both, src/vert_flow.c:93-102 and :149-156).
  v.wet_bottom = g > depth - dinf;
  v.deficit = 0.0; v.satn = 0.0; v.psi_raw = 0.0; v.kr = 0.0;
  if (!v.wet_bottom) {
    v.deficit = depth - g;
    double satn = u / v.deficit;
    satn = (satn > 1.0) ? 1.0 : satn;
    satn = (satn < SATMIN) ? SATMIN : satn;
    v.satn = satn;
    v.psi_raw = psi_vg(satn, in.alpha, in.beta);
    v.kr = kr_func(in.beta, satn);
  }
  return v;
}

template <bool B, bool N>
__device__ __forceinline__ void elem_vert_post(const RhsArgs &a, int p, const Ix &ix, const ElemIn &in,
                                               const NbrRec &me, const VertPre &v, const LatOut &lat) {
  const int ne = a.ne;
  const int lp = p - ix.eoff;
  double *ydot = ix.ydot;
  const double area = in.area, zmax = in.zmax, zmin = in.zmin;
  const double depth = in.depth, dinf = in.dinf;
  const double pcpdrp = in.pcpdrp, ws0 = in.ws0;
  const double beta = in.beta, kinfv = in.kinfv, kmacv = in.kmacv;
  const double areafh = in.areafh, dmac = in.dmac, ksatv = in.ksatv, porosity = in.porosity;
  const double u = clamp0(in.yu);
  const double g = me.gw;
  const double surfh = me.surfh;
  const bool wet_bottom = v.wet_bottom;
  const double deficit = v.deficit, satn = v.satn, psi_raw = v.psi_raw, kr = v.kr;

  // Infil: src/vert_flow.c:33-130
  double infil;
  if (u + g > depth) {
    infil = 0.0;
  } else {
    double applrate = lat.asum;
    applrate = (applrate > 0.0) ? applrate : 0.0;
    applrate += pcpdrp;
    double wetfrac = surfh / DEPRSTG;
    wetfrac = (wetfrac > 0.0) ? wetfrac : 0.0;
    wetfrac = (wetfrac < 1.0) ? wetfrac : 1.0;
    double dh_by_dz, kinf;
    if (wet_bottom) {
      dh_by_dz = (surfh + zmax - (g + zmin)) / (0.5 * (surfh + dinf));
      dh_by_dz = (surfh <= 0.0 && dh_by_dz > 0.0) ? 0.0 : dh_by_dz;
      // KrFunc(beta, 1) is exactly 1 for beta > 1 (1^m = 1, 0^(1/m) = 0)
      const double satkfunc = (beta > 1.0) ? 1.0 : kr_func(beta, 1.0);
      kinf = eff_kinf(kinfv, kmacv, areafh, dh_by_dz, satkfunc, 1.0, applrate, surfh);
      infil = kinf * dh_by_dz;
    } else {
      double isatn = satn, ipsi = psi_raw, ikr = kr;
      if (N) {
        // saturation of the top soil layer; smcmax - smcmin is the porosity
        // (src/init_soil.c:34-36)
        isatn = (in.sh2o0 - in.smcmin) / porosity;
        isatn = (isatn > 1.0) ? 1.0 : isatn;
        isatn = (isatn < SATMIN) ? SATMIN : isatn;
        ipsi = psi_vg(isatn, in.alpha, beta);
        ikr = kr_func(beta, isatn);
      }
      const double psi_u = (ipsi > PSIMIN) ? ipsi : PSIMIN;
      const double h_u = psi_u + zmax - 0.5 * dinf;
      dh_by_dz = (surfh + zmax - h_u) / (0.5 * (surfh + dinf));
      dh_by_dz = (surfh <= 0.0 && dh_by_dz > 0.0) ? 0.0 : dh_by_dz;
      kinf = eff_kinf(kinfv, kmacv, areafh, dh_by_dz, ikr, isatn, applrate, surfh);
      infil = kinf * dh_by_dz;
      infil = (infil > 0.0) ? infil : 0.0;
    }
    const double infil_max = applrate + ((ws0 > 0.0) ? ws0 / ix.dt : 0.0);
    infil = (infil > infil_max) ? infil_max : infil;
    infil *= wetfrac;
  }
  if (N) infil *= in.fcr;   // frozen top soil (src/vert_flow.c:16-19), before Recharge reads it

  // Recharge: src/vert_flow.c:132-170
  double rechg;
  if (wet_bottom) {
    rechg = infil;
  } else {
    const double dh_by_dz = (0.5 * deficit + psi_raw) / (0.5 * (deficit + g));
    const double kavg = avg_kv(dmac, ksatv, kmacv, areafh, deficit, g, kr);
    rechg = kavg * dh_by_dz;
    rechg = (rechg > 0.0 && u <= 0.0) ? 0.0 : rechg;
    rechg = (rechg < 0.0 && g <= 0.0) ? 0.0 : rechg;
  }
  if (a.record) {
    a.rec.infil[p] = infil;
    a.rec.rechg[p] = rechg;
  }

  // element ODEs: src/ode.c:121-149
  double dys = 0.0, dyu = 0.0, dyg = 0.0;
  dys += pcpdrp - infil - v.edir_s;
  dyu += infil - rechg - v.edir_u - v.ett_u;
  dyg += rechg - v.edir_g - v.ett_g;
  dyu /= porosity;
#pragma unroll
  for (int j = 0; j < 3; j++) {
    dys -= lat.ovl[j] / area;
    dyg -= lat.sub[j] / area;
  }
  dyg /= porosity;
  ydot[lp] = dys;
  ydot[ix.ne + lp] = dyu;
  ydot[2 * ix.ne + lp] = dyg;
  if (isnan(dys) || isnan(dyu) || isnan(dyg)) flag_nan<B>(a, ix);
  if (!B && a.closure) {
    a.cl_e[p] = pcpdrp * area;
    a.cl_e[ne + p] = ((((v.edir_s + v.edir_u) + v.edir_g) + v.ett_u) + v.ett_g) * area;
    a.cl_e[2L * ne + p] = lat.out;
  }
}

template <bool B, bool N, class Src>
__device__ __forceinline__ void elem_core(const RhsArgs &a, int p, const Ix &ix, const Src &src, const ElemIn &in) {
  const NbrRec me = src.self();
  const LatOut lat = elem_lateral<B>(a, p, ix, src, in, me);
  const VertPre v = elem_vert_pre<N>(in, me);
  elem_vert_post<B, N>(a, p, ix, in, me, v, lat);
}

template <bool B, bool N = false>
__device__ __forceinline__ void elem_flux(const RhsArgs &a, int p) {
  const Ix ix = ix_elem<B>(a, p);
  if (!ix.active) return;
  if (B && p - ix.eoff >= ix.ne) return;     // unused capacity of a slot
  const GlobalSrc src{a, p};
  elem_core<B, N>(a, p, ix, src, load_elem_in<N>(a, p, ix, src));
}

template <bool B>
__device__ __forceinline__ BankView bank_view(const RhsArgs &a, const Ix &ix, int q) {
  const ElemDev &e = a.e;
  const int lq = q - ix.eoff;
  const double s = clamp0(ix.y[lq]);
  const double g = clamp0(ix.y[2 * ix.ne + lq]);
  BankView v;
  v.surfh = surf_h(s);
  v.gw = g;
  v.zmin = e.zmin[q];
  v.gwz = g + v.zmin;
  v.effkh = eff_kh(e.depth[q], e.dmac[q], e.kmach[q], e.areafv[q], e.ksath[q], g);
  v.zmax = e.zmax[q];
  return v;
}

// SubFlowRiverToRiver (_ARITH_): src/river_flow.c:302-331, aquifer flow from
// segment u into its downstream segment d (yg: river groundwater of the solve,
// by river index; effk, effk_nabr: the mean EffKh of u's and of d's banks).
// Evaluated by d for each upstream u and by u for itself, from the same
// operands, so both get the same bits.
__device__ __forceinline__ double sub_r2r_k(const RiverDev &R, const double *yg, int u, int d, double effk,
                                            double effk_nabr) {
  const double rgw = clamp0(yg[u]);
  const double dgw = clamp0(yg[d]);
  const double total_h = rgw + R.zmin[u];
  const double total_h_down = dgw + R.zmin[d];
  const double avg_wid = (R.width[u] + R.width[d]) / 2.0;
  const double diff_h = total_h - total_h_down;
  const double avg_h = avg_h_sub(diff_h, rgw, dgw);
  const double distance = 0.5 * (R.length[u] + R.length[d]);
  const double grad_h = diff_h / distance;
  const double avg_ksat = 0.5 * (effk + effk_nabr);
  return avg_ksat * grad_h * avg_h * avg_wid;
}

// The same with the banks' EffKh from the records k_prep wrote.
__device__ __forceinline__ double sub_r2r(const RhsArgs &a, const double *yg, int u, int d) {
  const RiverDev &R = a.r;
  const double effk = 0.5 * (a.e.rec[R.left[u]].effkh + a.e.rec[R.right[u]].effkh);
  const double effk_nabr = 0.5 * (a.e.rec[R.left[d]].effkh + a.e.rec[R.right[d]].effkh);
  return sub_r2r_k(R, yg, u, d, effk, effk_nabr);
}

// One segment's side of ChanFlowRiverToRiver (src/river_flow.c:257-300).
struct ChanSide {
  double stage, zbed, total_h, perim, crossa, rough, length;
};

__device__ __forceinline__ ChanSide chan_side(const RiverDev &R, const double *ys, int r) {
  ChanSide c;
  c.stage = clamp0(ys[r]);
  c.zbed = R.zbed[r];
  c.total_h = c.stage + c.zbed;
  c.perim = riv_perim(R.ord[r], c.stage, R.coeff[r]);
  c.crossa = riv_cs_area(R.ord[r], c.stage, R.coeff[r]);
  c.rough = R.rough[r];
  c.length = R.length[r];
  return c;
}

// ChanFlowRiverToRiver: channel flow from segment u into its downstream
// segment d, from their two sides.
__device__ __forceinline__ double chan_r2r(int riv_mode, const ChanSide &u, const ChanSide &d) {
  const double avg_perim = (u.perim + d.perim) / 2.0;
  const double avg_rough = (u.rough + d.rough) / 2.0;
  const double distance = 0.5 * (u.length + d.length);
  const double diff_h = (riv_mode == KINEMATIC) ? (u.zbed - d.zbed) : (u.total_h - d.total_h);
  const double grad_h = diff_h / distance;
  const double avg_sf = (grad_h > 0.0) ? grad_h : RIVGRADMIN;
  const double avg_crossa = 0.5 * (u.crossa + d.crossa);
  const double avg_h = (avg_perim == 0.0) ? 0.0 : (avg_crossa / avg_perim);
  return overland_flow(avg_h, grad_h, avg_sf, u.crossa, avg_rough);
}

// Segment r's channel outflow (rivflow 1: to the downstream segment or through
// the outlet), the boundary flux of its upstream node (upinit) and its channel
// leakage (rivflow 6): src/river_flow.c:10-87, 257-300, 333-427, 500-518.
// ys, yg: stage and river groundwater of the solve, by river index.
__device__ __forceinline__ void river_out(const RhsArgs &a, const double *ys, const double *yg, int r,
                                          double &upinit_out, double &rf1_out, double &rf6_out) {
  const RiverDev &R = a.r;
  const double stage = clamp0(ys[r]);
  const double rgw = clamp0(yg[r]);
  const double zbed = R.zbed[r], length = R.length[r];
  const double rzmin = R.zmin[r];
  const int ord = R.ord[r];
  const double coeff = R.coeff[r];
  const double bcv = R.bc[r];
  const int dcode = R.down[r];
  double upinit = 0.0, rf1;
  if (dcode >= 0) {
    const int d = dcode;
    if (R.bctype[r] != 0) {
      // BoundFluxRiver: src/river_flow.c:392-427
      double flux = 0.0;
      if (R.bctype[r] > 0) {
        const double total_h = stage + zbed;
        const double distance = 0.5 * length;
        const double grad_h = (total_h - bcv) / distance;
        const double perim = riv_perim(ord, stage, coeff);
        const double crossa = riv_cs_area(ord, stage, coeff);
        const double avg_h = (perim == 0.0) ? 0.0 : (crossa / perim);
        flux = overland_flow(avg_h, grad_h, grad_h, crossa, R.rough[r]);
      } else {
        flux = -bcv;
      }
      upinit += flux;
    }
    rf1 = chan_r2r(a.riv_mode, chan_side(R, ys, r), chan_side(R, ys, d));
  } else {
    // OutletFlux: src/river_flow.c:333-390
    double q = 0.0;
    switch (dcode) {
      case DIRICHLET: {
        const double total_h = stage + zbed;
        const double distance = 0.5 * length;
        const double grad_h = (total_h - bcv) / distance;
        const double perim = riv_perim(ord, stage, coeff);
        const double crossa = riv_cs_area(ord, stage, coeff);
        const double avg_h = (perim == 0.0) ? 0.0 : (crossa / perim);
        q = overland_flow(avg_h, grad_h, grad_h, crossa, R.rough[r]);
        break;
      }
      case NEUMANN: q = -bcv; break;
      case ZERO_DPTH_GRAD: {
        const double distance = 0.5 * length;
        const double grad_h = (zbed - (R.node_zmax[r] - R.depth[r])) / distance;
        const double perim = riv_perim(ord, stage, coeff);
        const double crossa = riv_cs_area(ord, stage, coeff);
        q = sqrt(grad_h) * crossa * ((perim > 0.0) ? xpow(crossa / perim, 2.0 / 3.0) : 0.0) / R.rough[r];
        break;
      }
      default: {  // CRIT_DPTH (other codes rejected at creation)
        const double crossa = riv_cs_area(ord, stage, coeff);
        q = crossa * sqrt(GRAV * stage);
        break;
      }
    }
    rf1 = q;
  }
  // ChanLeak: src/river_flow.c:500-518
  double rf6;
  {
    double diff_h;
    if (zbed - (rgw + rzmin) > 0.0) diff_h = stage;
    else diff_h = stage + zbed - (rgw + rzmin);
    const double grad_h = diff_h / R.bedthick[r];
    rf6 = R.ksatv[r] * R.width[r] * length * grad_h;
  }
  upinit_out = upinit;
  rf1_out = rf1;
  rf6_out = rf6;
}

// River segment fluxes (src/river_flow.c:10-87, 113-183, 257-300, 333-518),
// run by the river blocks of k_prep in three parts per segment, each its own
// thread so that no thread carries the whole chain of gathers:
//   part 0: the upstream-node boundary flux, the channel outflow (rivflow 1:
//           to the downstream segment or through the outlet) and the channel
//           leakage (rivflow 6);
//   part 1: the left bank's fluxes (rivflow 2, 4, 7) and its bank-edge
//           replacement values;
//   part 2: the same for the right bank (rivflow 3, 5, 8).
// The aquifer flow to the downstream segment (rivflow 9) needs the
// downstream banks' records and is left to k_flux.
template <bool B>
__device__ __forceinline__ void river_flux(const RhsArgs &a, int r, int part) {
  const int nr = a.nr;
  const RiverDev &R = a.r;
  const Ix ix = ix_river<B>(a, r);
  if (!ix.active) return;
  if (B && r - ix.roff >= ix.nr) return;            // unused capacity of a slot
  const double *ys = ix.y + 3L * ix.ne - ix.roff;   // ys[r']: stage of river r'
  const double *yg = ys + ix.nr;                    // yg[r']: river groundwater

  if (part == 0) {
    double upinit, rf1, rf6;
    river_out(a, ys, yg, r, upinit, rf1, rf6);
    if (!B && a.closure) a.cl_r[r] = upinit + ((R.down[r] < 0) ? rf1 : 0.0);   // leaves through river boundaries
    R.rf[1 * nr + r] = rf1;
    R.rf[6 * nr + r] = rf6;
    R.upinit[r] = upinit;
    return;
  }

  // RiverToElem for one bank: src/river_flow.c:113-183
  const int side = part - 1;
  const double stage = clamp0(ys[r]);
  const double rgw = clamp0(yg[r]);
  const int q = side ? R.right[r] : R.left[r];
  const BankView bv = bank_view<B>(a, ix, q);
  const double other = effkh_of<B>(a, ix, side ? R.left[r] : R.right[r]);
  const double effk_riv = side ? 0.5 * (other + bv.effkh) : 0.5 * (bv.effkh + other);
  const BankFlux f = bank_flux(R, r, side, stage, rgw, bv, effk_riv);
  R.rf[(2 + side) * nr + r] = f.ovl;
  R.rf[(4 + side) * nr + r] = f.chan;
  R.rf[(7 + side) * nr + r] = f.sub;
  // bank-edge replacement values (src/river_flow.c:159-182)
  R.hid_ovl[2 * nr + 2 * r + side] = -f.ovl;
  R.bank_sub[2 * r + side] = -(f.chan + f.sub);
}

// River ODEs (src/ode.c:238-251), run by the river blocks of k_flux: the
// aquifer flow to the downstream segment (rivflow 9), and the upstream inflow
// in ascending caller index, the order of the reference's serial loop
// (src/river_flow.c:96-110): each upstream segment's channel outflow as
// k_prep stored it, its aquifer outflow evaluated here.
template <bool B>
__device__ __forceinline__ void river_ode(const RhsArgs &a, int r) {
  const int nr = a.nr;
  const RiverDev &R = a.r;
  const Ix ix = ix_river<B>(a, r);
  if (!ix.active) return;
  if (B && r - ix.roff >= ix.nr) return;
  const double *yg = ix.y + 3L * ix.ne - ix.roff + ix.nr;   // yg[r']: river groundwater
  const int dcode = R.down[r];
  const double rf9 = (dcode >= 0) ? sub_r2r(a, yg, r, dcode) : 0.0;
  double up = R.upinit[r], upa = 0.0;
  for (int k = R.up_beg[r]; k < R.up_end[r]; k++) {
    const int u = R.up_idx[k];
    up -= R.rf[1 * nr + u];
    upa -= sub_r2r(a, yg, u, r);
  }
  const double area = R.area[r];
  const double rf6 = R.rf[6 * nr + r];
  double ds = 0.0;
  ds -= up / area;
#pragma unroll
  for (int j = 1; j <= 6; j++) ds -= R.rf[j * nr + r] / area;
  double dg = 0.0;
  dg += -R.rf[7 * nr + r] - R.rf[8 * nr + r] - rf9 - upa + rf6;
  dg /= R.porosity[r] * area;
  double *ydr = ix.ydot + 3L * ix.ne - ix.roff;
  ydr[r] = ds;
  ydr[ix.nr + r] = dg;
  if (a.record) {
    R.rf[0 * nr + r] = up;
    R.rf[9 * nr + r] = rf9;
    R.rf[10 * nr + r] = upa;
  }
  if (isnan(ds) || isnan(dg)) flag_nan<B>(a, ix);
}

// Grid: nbr river blocks (river ODEs, short), then nbe element blocks.  At
// most 168 VGPRs: three wavefronts per SIMD.
template <bool B, bool N = false>
__global__ __launch_bounds__(kBlock, 3) void k_flux(RhsArgs a, int nbe) {
  const int nbr = (int)gridDim.x - nbe;
  if ((int)blockIdx.x >= nbr) {
    const int tile = xcd_tile((int)blockIdx.x - nbr, nbe, a.tile_span);
    const int p = (B && a.bm) ? compact_index(a, 0, tile) : tile * kBlock + threadIdx.x;
    if (p >= 0 && p < a.ne) elem_flux<B, N>(a, p);
  } else {
    const int rb = blockIdx.x;
    const int r = (B && a.bm) ? compact_index(a, 1, rb) : rb * kBlock + threadIdx.x;
    if (r >= 0 && r < a.nr) river_ode<B>(a, r);
  }
}

// River segments of a single solve in the one-launch RHS: a block takes
// kRivSeg consecutive segments, and each quarter of it one part of every
// segment's work, so that no thread carries the whole chain of gathers:
//   part 0: river_out (channel outflow, upstream-node boundary flux, leakage);
//   part 1: the left bank's fluxes; part 2: the right bank's;
//   part 3: the aquifer flow to the downstream segment (rivflow 9) and the
//           upstream inflow: from the boundary flux of the upstream node,
//           each upstream segment's channel outflow subtracted in ascending
//           caller index (src/river_flow.c:96-110), and each one's aquifer
//           outflow, all evaluated here from the same operands that segment
//           uses for itself.
// The parts meet in LDS (xr: 9 doubles per segment); part 0 then writes the
// river ODEs (src/ode.c:238-251) in river_ode's order.
constexpr int kRivSeg = kTile / 4;

__device__ __forceinline__ void river_parts(const RhsArgs &a, int rb, double *xr) {
  const int nr = a.nr;
  const RiverDev &R = a.r;
  const int part = (int)threadIdx.x / kRivSeg, k = (int)threadIdx.x % kRivSeg;
  const int r = rb * kRivSeg + k;
  const Ix ix = ix_river<false>(a, 0);
  const double *ys = ix.y + 3L * ix.ne;
  const double *yg = ys + nr;
  double *x = xr + 9 * k;
  double upinit = 0.0, rf1 = 0.0, rf6 = 0.0;
  if (r < nr) {
    if (part == 0) {
      river_out(a, ys, yg, r, upinit, rf1, rf6);
    } else if (part <= 2) {
      const int side = part - 1;
      const double stage = clamp0(ys[r]);
      const double rgw = clamp0(yg[r]);
      const BankView bv = bank_view<false>(a, ix, side ? R.right[r] : R.left[r]);
      const double other = effkh_of<false>(a, ix, side ? R.left[r] : R.right[r]);
      const double effk_riv = side ? 0.5 * (other + bv.effkh) : 0.5 * (bv.effkh + other);
      const BankFlux f = bank_flux(R, r, side, stage, rgw, bv, effk_riv);
      x[3 * side + 0] = f.ovl;
      x[3 * side + 1] = f.chan;
      x[3 * side + 2] = f.sub;
    } else {
      double ui, u1, u6;
      river_out(a, ys, yg, r, ui, u1, u6);
      const double effk_r = 0.5 * (effkh_of<false>(a, ix, R.left[r]) + effkh_of<false>(a, ix, R.right[r]));
      const int dcode = R.down[r];
      double rf9 = 0.0;
      if (dcode >= 0) {
        const double effk_d =
            0.5 * (effkh_of<false>(a, ix, R.left[dcode]) + effkh_of<false>(a, ix, R.right[dcode]));
        rf9 = sub_r2r_k(R, yg, r, dcode, effk_r, effk_d);
      }
      double up = ui, upa = 0.0;
      const ChanSide me = chan_side(R, ys, r);
      for (int kk = R.up_beg[r]; kk < R.up_end[r]; kk++) {
        const int u = R.up_idx[kk];
        up -= chan_r2r(a.riv_mode, chan_side(R, ys, u), me);   // u's channel outflow: its downstream segment is r
        const double effk_u = 0.5 * (effkh_of<false>(a, ix, R.left[u]) + effkh_of<false>(a, ix, R.right[u]));
        upa -= sub_r2r_k(R, yg, u, r, effk_u, effk_r);
      }
      x[6] = up;
      x[7] = upa;
      x[8] = rf9;
    }
  }
  __syncthreads();
  if (part != 0 || r >= nr) return;
  const double flo = x[0], fro = x[3], flc = x[1], frc = x[4], fls = x[2], frs = x[5];
  const double up = x[6], upa = x[7], rf9 = x[8];
  const double area = R.area[r];
  const double rfs[7] = {0.0, rf1, flo, fro, flc, frc, rf6};
  double ds = 0.0;
  ds -= up / area;
#pragma unroll
  for (int j = 1; j <= 6; j++) ds -= rfs[j] / area;
  double dg = 0.0;
  dg += -fls - frs - rf9 - upa + rf6;
  dg /= R.porosity[r] * area;
  ix.ydot[3L * ix.ne + r] = ds;
  ix.ydot[3L * ix.ne + nr + r] = dg;
  if (a.closure) a.cl_r[r] = upinit + ((R.down[r] < 0) ? rf1 : 0.0);
  if (a.record) {
    R.rf[0 * nr + r] = up;
#pragma unroll
    for (int j = 1; j <= 6; j++) R.rf[j * nr + r] = rfs[j];
    R.rf[7 * nr + r] = fls;
    R.rf[8 * nr + r] = frs;
    R.rf[9 * nr + r] = rf9;
    R.rf[10 * nr + r] = upa;
  }
  if (isnan(ds) || isnan(dg)) flag_nan<false>(a, ix);
}

// A halo or tile entry's record fields (all but the FrictSlope gradient) and
// its head, as k_prep evaluates them.
__device__ __forceinline__ NbrRec tile_rec(double ys, double yg, double zmax, double zmin, double depth,
                                           double dmac, double kmach, double areafv, double ksath, double rough) {
  const double surfh = surf_h(clamp0(ys));
  const double g = clamp0(yg);
  NbrRec r;
  r.surfh = surfh;
  r.head = zmax + surfh;
  r.gw = g;
  r.gwz = g + zmin;
  r.effkh = eff_kh(depth, dmac, kmach, areafv, ksath, g);
  r.sfmag = 0.0;
  r.rough = rough;
  r.zmax = zmax;
  return r;
}

// FrictSlope (src/lat_flow.c:130-185) of the LDS entry at slot s (device
// element q): neighbour heads from LDS, or the boundary value, or the stage
// of the river edge.
__device__ __forceinline__ double tile_sfmag(const RhsArgs &a, int s, long q, const int c[3], const double nx[3],
                                             const double ny[3], const double *shead) {
  const long ne = a.ne;
  double h[3];
#pragma unroll
  for (int j = 0; j < 3; j++) {
    if (c[j] >= 0) {
      h[j] = shead[c[j]];
    } else if (nb_is_bound(c[j])) {
      h[j] = (c[j] == NB_BOUND) ? shead[s] : a.e.ebc[j * ne + q];
    } else {
      const int r = (-c[j] - 1) >> 1;
      const double st = clamp0(a.y[3 * ne + r]);
      h[j] = (st > a.r.depth[r]) ? a.r.zbed[r] + st : a.r.zmax[r];
    }
  }
  const double dhdx = dh_by_dl(ny[0], ny[1], ny[2], nx[0], nx[1], nx[2], h[0], h[1], h[2]);
  const double dhdy = dh_by_dl(nx[0], nx[1], nx[2], ny[0], ny[1], ny[2], h[0], h[1], h[2]);
  return sqrt(dhdx * dhdx + dhdy * dhdy);
}

// One RHS call of a single solve in one launch.  Grid: river blocks
// (river_parts, kRivSeg segments each), then one block per tile of kTile
// consecutive elements (Hilbert order), two wavefronts per block: blocks run
// independently, hiding each other's memory latency, and a tile of 128 has a
// smaller halo per element than one of 64 (at the 200,000- and
// 100,000-triangle meshes 128 measured 2% and 7% faster than 64; 192 and 256
// were no better).  Each element block stages in LDS what its elements'
// fluxes read: the heads of the tile and of its halo (the first ring of
// outside neighbours and their outside neighbours), and the records of the
// tile and the first ring, whose FrictSlope gradients need those heads; then
// it runs the element fluxes from them (elem_core with TileSrc).  The loads
// for the staging (its element's and its halo entry's record inputs, and the
// bank fluxes of its first river edge) are issued at the start, the element's
// flux inputs after the barriers, and each edge reads its neighbour's record
// from LDS when its turn comes: the kernel stays within 128 VGPRs, four
// wavefronts per SIMD, so that a 200,000-triangle RHS (3,381 wavefronts of
// element and river blocks) runs in one round of 4,096 wavefront slots
// instead of spilling into a second round of 3,072.  With the _NOAH_ branches
// the cap is three wavefronts (140 VGPRs): four would spill in the element
// path.  The halo's records are evaluated again by every tile that borders
// them, from the same operands: every output keeps one writer, and the
// arithmetic is that of k_prep + k_flux, bit for bit.
template <bool N>
__global__ __launch_bounds__(kTile, N ? 3 : 4) void k_rhs_tile(RhsArgs a, TileArgs t, int nbe) {
  extern __shared__ double tile_lds[];
  const LdsRecs srec{tile_lds, t.s1cap};
  double *shead = tile_lds + 8 * t.s1cap;
  double *sbank = shead + t.scap;                         // [2 * kTile]: ovl, then sub
  int *sbankj = reinterpret_cast<int *>(sbank + 2 * kTile);   // [kTile]
  const int nrb = (int)gridDim.x - nbe;
  if ((int)blockIdx.x < nrb) {
    river_parts(a, (int)blockIdx.x, tile_lds);
    return;
  }
  const int tile = xcd_tile((int)blockIdx.x - nrb, nbe, a.tile_span);
  const long ne = a.ne;
  const int lo = tile * kTile, i = (int)threadIdx.x, p = lo + i;
  const bool diff = a.surf_mode == DIFF_WAVE;
  const int hb = t.hbeg[tile], nh = t.hbeg[tile + 1] - hb, n1 = t.hn1[tile];
  const ElemDev &e = a.e;
  const double *y = a.y;
  const Ix ix = ix_elem<false>(a, 0);
  const TileSrc src{a, srec, t.lnb, p, i, sbank, sbankj};

  // own element: what its record and FrictSlope gradient need (the flux
  // inputs are loaded after the barriers, so that they do not hold registers
  // through the staging)
  const bool own = p < ne;
  double ys = 0.0, yg = 0.0, zmax = 0.0, zmin = 0.0, depth = 0.0, dmac = 0.0;
  double kmach = 0.0, areafv = 0.0, ksath = 0.0, rough = 0.0;
  int oc[3] = {0, 0, 0};
  double nx[3] = {0.0, 0.0, 0.0}, ny[3] = {0.0, 0.0, 0.0};
  if (own) {
    zmax = e.zmax[p]; zmin = e.zmin[p]; depth = e.depth[p]; dmac = e.dmac[p];
    ys = y[p];
    yg = y[2 * ne + p];
    kmach = e.kmach[p]; areafv = e.areafv[p]; ksath = e.ksath[p]; rough = e.rough[p];
#pragma unroll
    for (int j = 0; j < 3; j++) oc[j] = src.code(j);
    if (diff) {
#pragma unroll
      for (int j = 0; j < 3; j++) {
        nx[j] = e.nabr_x[j * ne + p];
        ny[j] = e.nabr_y[j * ne + p];
      }
    }
  }
  // the halo entry this thread stages (further entries, in halos longer than a
  // block, are staged in the loops below)
  const bool hv = i < nh, h1 = i < n1;
  long hq = 0;
  double hys = 0.0, hyg = 0.0, hzmax = 0.0, hzmin = 0.0, hdepth = 0.0, hdmac = 0.0;
  double hkmach = 0.0, hareafv = 0.0, hksath = 0.0, hrough = 0.0;
  int hc[3] = {0, 0, 0};
  double hnx[3] = {0.0, 0.0, 0.0}, hny[3] = {0.0, 0.0, 0.0};
  if (hv) {
    hq = t.hidx[hb + i];
    hys = y[hq];
    hzmax = e.zmax[hq];
    if (h1) {
      hyg = y[2 * ne + hq];
      hzmin = e.zmin[hq]; hdepth = e.depth[hq]; hdmac = e.dmac[hq];
      hkmach = e.kmach[hq]; hareafv = e.areafv[hq]; hksath = e.ksath[hq]; hrough = e.rough[hq];
      if (diff) {
#pragma unroll
        for (int j = 0; j < 3; j++) {
          hc[j] = t.hnb[3 * (hb + i) + j];
          hnx[j] = e.nabr_x[j * ne + hq];
          hny[j] = e.nabr_y[j * ne + hq];
        }
      }
    }
  }

  // records and heads; the bank fluxes of the element's first river edge,
  // whose gathers (segment, then the other bank) then overlap the halo's
  if (own) {
    const NbrRec r = tile_rec(ys, yg, zmax, zmin, depth, dmac, kmach, areafv, ksath, rough);
    srec.put(i, r);
    shead[i] = r.head;
    int bj = -1;
#pragma unroll
    for (int j = 2; j >= 0; j--)
      if (nb_is_river(oc[j])) bj = j;
    double bo = 0.0, bs = 0.0;
    if (bj >= 0) src.bank_eval(ix, -oc[bj] - 1, r, zmax, zmin, bo, bs);
    sbank[i] = bo;
    sbank[kTile + i] = bs;
    sbankj[i] = bj;
  }
  if (hv) {
    if (h1) {
      const NbrRec r = tile_rec(hys, hyg, hzmax, hzmin, hdepth, hdmac, hkmach, hareafv, hksath, hrough);
      srec.put(kTile + i, r);
      shead[kTile + i] = r.head;
    } else {
      shead[kTile + i] = hzmax + surf_h(clamp0(hys));
    }
  }
  for (int h = i + kTile; h < nh; h += kTile) {
    const long q = t.hidx[hb + h];
    if (h < n1) {
      const NbrRec r = tile_rec(y[q], y[2 * ne + q], e.zmax[q], e.zmin[q], e.depth[q], e.dmac[q], e.kmach[q],
                                e.areafv[q], e.ksath[q], e.rough[q]);
      srec.put(kTile + h, r);
      shead[kTile + h] = r.head;
    } else {
      shead[kTile + h] = e.zmax[q] + surf_h(clamp0(y[q]));
    }
  }
  __syncthreads();
  if (diff) {
    if (own) srec.put_sfmag(i, tile_sfmag(a, i, p, oc, nx, ny, shead));
    if (h1) srec.put_sfmag(kTile + i, tile_sfmag(a, kTile + i, hq, hc, hnx, hny, shead));
    for (int h = i + kTile; h < n1; h += kTile) {
      const long q = t.hidx[hb + h];
      int c[3];
      double qx[3], qy[3];
#pragma unroll
      for (int j = 0; j < 3; j++) {
        c[j] = t.hnb[3 * (hb + h) + j];
        qx[j] = e.nabr_x[j * ne + q];
        qy[j] = e.nabr_y[j * ne + q];
      }
      srec.put_sfmag(kTile + h, tile_sfmag(a, kTile + h, q, c, qx, qy, shead));
    }
    __syncthreads();
  }
  if (!own) return;
  elem_core<false, N>(a, p, ix, src, load_elem_in<N>(a, p, ix, src));
}

// k_flux of a single solve, with the _NOAH_ branches when the context has them.
static void launch_flux_single(const pihm_ctx *c, const RhsArgs &a, int nbe, int nbr) {
  if (c->noah) hipLaunchKernelGGL((k_flux<false, true>), dim3(nbe + nbr), dim3(kBlock), 0, c->stream, a, nbe);
  else hipLaunchKernelGGL((k_flux<false, false>), dim3(nbe + nbr), dim3(kBlock), 0, c->stream, a, nbe);
}

// One RHS evaluation of a single solve on the context's path.
static void launch_single(const pihm_ctx *c, const RhsArgs &a) {
  const int nbe = div_up(c->ne, kBlock);
  const int nbr = div_up(c->nr, kBlock);
  if (c->rhs_path == PIHM_RHS_ONE_LAUNCH) {
    const int nt = div_up(c->ne, kTile), nrb = div_up(c->nr, kRivSeg);
    if (c->noah)
      hipLaunchKernelGGL((k_rhs_tile<true>), dim3(nt + nrb), dim3(kTile), c->tile_lds, c->stream, a, c->tiles, nt);
    else
      hipLaunchKernelGGL((k_rhs_tile<false>), dim3(nt + nrb), dim3(kTile), c->tile_lds, c->stream, a, c->tiles, nt);
    return;
  }
  hipLaunchKernelGGL(k_prep<false>, dim3(nbe + 3 * nbr), dim3(kBlock), 0, c->stream, a, nbe);
  launch_flux_single(c, a, nbe, nbr);
}

// Water-balance accumulators: the four volume rates summed over the solve in
// a fixed order (strided per thread, wavefront butterfly, waves in order,
// then workgroups in order), written into the derivative entries that follow
// the reference's state.
constexpr int kClosureBlocks = 256;

__global__ __launch_bounds__(kBlock) void k_closure_part(RhsArgs a) {
  __shared__ double lds[4 * 4];
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  const int stride = gridDim.x * kBlock;
  for (int p = blockIdx.x * kBlock + threadIdx.x; p < a.ne; p += stride) {
    acc[0] += a.cl_e[p];
    acc[1] += a.cl_e[a.ne + p];
    acc[2] += a.cl_e[2L * a.ne + p];
  }
  for (int r = blockIdx.x * kBlock + threadIdx.x; r < a.nr; r += stride) acc[3] += a.cl_r[r];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    double v = acc[k];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if (lane == 0) lds[wave * 4 + k] = v;
  }
  __syncthreads();
  if (threadIdx.x < 4) {
    const int k = threadIdx.x;
    a.cl_part[k * kClosureBlocks + blockIdx.x] = ((lds[k] + lds[4 + k]) + lds[8 + k]) + lds[12 + k];
  }
}

__global__ __launch_bounds__(kBlock) void k_closure_final(RhsArgs a, int nb, double *out) {
  __shared__ double lds[4];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int k = 0; k < 4; k++) {
    double v = (threadIdx.x < nb) ? a.cl_part[k * kClosureBlocks + threadIdx.x] : 0.0;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if (lane == 0) lds[wave] = v;
    __syncthreads();
    if (threadIdx.x == 0) out[k] = ((lds[0] + lds[1]) + lds[2]) + lds[3];
    __syncthreads();
  }
}


// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

namespace {

uint64_t hilbert_key(uint32_t x, uint32_t y) {
  const uint32_t n = 1u << 16;
  uint64_t d = 0;
  for (uint32_t s = n / 2; s > 0; s /= 2) {
    const uint32_t rx = (x & s) ? 1u : 0u, ry = (y & s) ? 1u : 0u;
    d += (uint64_t)s * s * ((3u * rx) ^ ry);
    if (ry == 0) {
      if (rx == 1) { x = n - 1 - x; y = n - 1 - y; }
      std::swap(x, y);
    }
  }
  return d;
}

struct Arena {
  std::vector<std::pair<void **, size_t>> req;
  template <class T> void add(T **p, size_t n) { req.push_back({(void **)p, n * sizeof(T)}); }
  size_t total() const {
    size_t t = 0;
    for (auto &r : req) t += (r.second + 255) & ~size_t(255);
    return t;
  }
  void carve(char *base) {
    size_t off = 0;
    for (auto &r : req) { *r.first = base + off; off += (r.second + 255) & ~size_t(255); }
  }
};

int validate(const pihm_mesh_desc *m) {
  if (!m || m->nelem <= 0 || m->nriver < 0) return PIHM_ERR_ARG;
  if (m->surf_mode != PIHM_KINEMATIC && m->surf_mode != PIHM_DIFF_WAVE) return PIHM_ERR_ARG;
  if (m->riv_mode != PIHM_KINEMATIC && m->riv_mode != PIHM_DIFF_WAVE) return PIHM_ERR_ARG;
  if (!(m->stepsize > 0.0)) return PIHM_ERR_ARG;
  const void *ep[] = {m->nabr, m->bc_type, m->x, m->y, m->area, m->zmin, m->zmax, m->edge,
                      m->nabrdist, m->nabr_x, m->nabr_y, m->depth, m->ksath, m->ksatv, m->kinfv,
                      m->dinf, m->alpha, m->beta, m->porosity, m->dmac, m->kmach, m->kmacv,
                      m->areafv, m->areafh, m->rough, m->rzd};
  for (auto q : ep) if (!q) return PIHM_ERR_ARG;
  const int ne = m->nelem, nr = m->nriver;
  if (nr > 0) {
    const void *rp[] = {m->down, m->leftele, m->rightele, m->riverbc_type, m->intrpl_ord, m->rarea,
                        m->rzmin, m->rzmax, m->zbed, m->node_zmax, m->dist_left, m->dist_right,
                        m->rdepth, m->coeff, m->length, m->width, m->rrough, m->cwr, m->rksath,
                        m->rksatv, m->bedthick, m->rporosity};
    for (auto q : rp) if (!q) return PIHM_ERR_ARG;
  }
  for (long k = 0; k < 3L * ne; k++) {
    const int c = m->nabr[k];
    if (c > ne || c < -nr) return PIHM_ERR_ARG;
  }
  for (int r = 0; r < nr; r++) {
    const int d = m->down[r];
    if (d > nr || d == 0 || d < -4) return PIHM_ERR_ARG;
    if (d == r + 1) return PIHM_ERR_ARG;
    if (m->leftele[r] < 1 || m->leftele[r] > ne || m->rightele[r] < 1 || m->rightele[r] > ne)
      return PIHM_ERR_ARG;
    if (m->intrpl_ord[r] < 1 || m->intrpl_ord[r] > 4) return PIHM_ERR_ARG;
    // each bank must carry the river tag on one edge (src/init_river.c:24-35)
    for (int side = 0; side < 2; side++) {
      const int el = (side ? m->rightele[r] : m->leftele[r]) - 1;
      bool found = false;
      for (int j = 0; j < 3; j++) found = found || (m->nabr[3 * el + j] == -(r + 1));
      if (!found) return PIHM_ERR_ARG;
    }
  }
  // every river edge must belong to a bank of that river
  for (int i = 0; i < ne; i++)
    for (int j = 0; j < 3; j++) {
      const int c = m->nabr[3 * i + j];
      if (c < 0) {
        const int r = -c - 1;
        if (m->leftele[r] != i + 1 && m->rightele[r] != i + 1) return PIHM_ERR_ARG;
      }
    }
  return PIHM_OK;
}

}  // namespace

namespace {

// Allocate a context with one slot per entry of caps.  Element, river, bank,
// upstream-list and state ranges of the slots are concatenated at their
// capacities; slots start empty (no elements, no rivers) until a mesh is
// loaded into them (ctx_load).
int build_tiles(pihm_ctx *c);

int ctx_alloc(const std::vector<SlotCap> &caps, int surf_mode, int riv_mode, pihm_ctx **out) {
  *out = nullptr;
  const int nmem = (int)caps.size();
  if (nmem < 1) return PIHM_ERR_ARG;
  if (surf_mode != KINEMATIC && surf_mode != DIFF_WAVE) return PIHM_ERR_ARG;
  if (riv_mode != KINEMATIC && riv_mode != DIFF_WAVE) return PIHM_ERR_ARG;
  long ne = 0, nr = 0, neq = 0, nup = 0;
  for (const auto &s : caps) {
    if (s.ne < 1 || s.nr < 0) return PIHM_ERR_ARG;
    ne += s.ne; nr += s.nr; neq += 3L * s.ne + 2L * s.nr; nup += s.nr;
  }
  if (ne > INT32_MAX / 4 || nr > INT32_MAX / 16 || neq > INT32_MAX) return PIHM_ERR_TOO_LARGE;
  pihm_ctx *c = new pihm_ctx();
  c->nmem = nmem;
  c->caps = caps;
  c->members.resize(nmem);
  c->uoff.resize(nmem);
  long e0 = 0, r0 = 0, y0 = 0, u0 = 0;
  for (int k = 0; k < nmem; k++) {
    auto &M = c->members[k];
    M.eoff = (int)e0; M.ne = 0;
    M.roff = (int)r0; M.nr = 0;
    M.yoff = y0; M.neq = 0;
    M.ceoff = 0; M.croff = 0; M.cyoff = 0;
    M.dt = 60.0;
    c->uoff[k] = (int)u0;
    e0 += caps[k].ne; r0 += caps[k].nr; y0 += 3L * caps[k].ne + 2L * caps[k].nr;
    u0 += caps[k].nr;
  }
  c->ne = (int)ne; c->nr = (int)nr; c->neq = (int)neq;
  c->surf_mode = surf_mode; c->riv_mode = riv_mode; c->dt = 60.0;
  c->stream = nullptr; c->nan_check = 1; c->record = 0; c->ncalls = 0;
  c->arena = nullptr;
  c->nanflag_host = nullptr;
  c->rhs_path = PIHM_RHS_TWO_KERNELS;
  c->one_launch_ok = 0;
  c->tiles = TileArgs{};
  c->tile_lds = 0;
  c->tile_arena = nullptr;
  PIHM_HIP_TRY(hipGetDevice(&c->device));
  c->perm.assign(ne, -1);
  c->iperm.assign(ne, -1);
  c->emem_host.resize(ne);
  c->nb_host.assign(3L * ne, NB_BOUND);
  for (int k = 0; k < nmem; k++)
    for (int lp = 0; lp < caps[k].ne; lp++) c->emem_host[c->members[k].eoff + lp] = k;

  RhsArgs &A = c->args;
  ElemDev &E = A.e;
  RiverDev &R = A.r;
  Arena ar;
  double **ed[] = {&E.area, &E.zmin, &E.zmax, &E.depth, &E.ksath, &E.ksatv, &E.kinfv, &E.dinf,
                   &E.alpha, &E.beta, &E.porosity, &E.dmac, &E.kmach, &E.kmacv, &E.areafv,
                   &E.areafh, &E.rough, &E.rzd, &E.pcpdrp, &E.edir, &E.ett, &E.ws0s};
  for (auto p : ed) ar.add(p, ne);
  double **ed3[] = {&E.edge, &E.nabrdist, &E.nabr_x, &E.nabr_y, &E.ebc};
  for (auto p : ed3) ar.add(p, 3L * ne);
  ar.add(&E.nb, 3L * ne);
  ar.add(&E.rec, ne);
  const long nr1 = std::max<long>(nr, 1);
  int32_t **ri[] = {&R.down, &R.left, &R.right, &R.bctype, &R.ord, &R.up_beg, &R.up_end};
  for (auto p : ri) ar.add(p, nr1);
  ar.add(&R.up_idx, std::max<long>(nup, 1));
  double **rd[] = {&R.area, &R.zmin, &R.zmax, &R.zbed, &R.node_zmax, &R.dist_l, &R.dist_r,
                   &R.depth, &R.coeff, &R.length, &R.width, &R.rough, &R.cwr, &R.ksath,
                   &R.ksatv, &R.bedthick, &R.porosity, &R.bc, &R.upinit};
  for (auto p : rd) ar.add(p, nr1);
  ar.add(&R.rf, 11L * nr1);
  ar.add(&R.hid_ovl, 4L * nr1);
  ar.add(&R.bank_sub, 2L * nr1);
  ar.add(&A.cl_e, 3L * ne);
  ar.add(&A.cl_r, nr1);
  ar.add(&A.cl_part, 4L * kClosureBlocks);
  ar.add(&A.rec.ovl, 3L * ne);
  ar.add(&A.rec.sub, 3L * ne);
  ar.add(&A.rec.infil, ne);
  ar.add(&A.rec.rechg, ne);
  ar.add(&c->nanflag_dev, 1);
  ar.add(&c->mem_dev, nmem);
  int32_t *emem_dev = nullptr, *rmem_dev = nullptr;
  ar.add(&emem_dev, ne);
  ar.add(&rmem_dev, nr1);
  c->arena_bytes = ar.total();
  int rc = PIHM_OK;
  if (hipMalloc(&c->arena, c->arena_bytes) != hipSuccess) rc = PIHM_ERR_HIP;
  if (!rc && hipMemset(c->arena, 0, c->arena_bytes) != hipSuccess) rc = PIHM_ERR_HIP;
  if (!rc && hipHostMalloc((void **)&c->nanflag_host, sizeof(int), hipHostMallocDefault) != hipSuccess)
    rc = PIHM_ERR_HIP;
  if (rc) { pihm_ctx_destroy(c); return rc; }
  ar.carve((char *)c->arena);
  *c->nanflag_host = 0;
  // empty slots: empty upstream lists, boundary codes
  std::vector<int32_t> rm(nr1, 0), ub(nr1, 0);
  for (int k = 0; k < nmem; k++)
    for (int r = 0; r < caps[k].nr; r++) { rm[c->members[k].roff + r] = k; ub[c->members[k].roff + r] = c->uoff[k]; }
  if (hipMemcpy(E.nb, c->nb_host.data(), 3L * ne * 4, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(emem_dev, c->emem_host.data(), ne * 4, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(rmem_dev, rm.data(), nr1 * 4, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(R.up_beg, ub.data(), nr1 * 4, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(R.up_end, ub.data(), nr1 * 4, hipMemcpyHostToDevice) != hipSuccess) {
    pihm_ctx_destroy(c);
    return PIHM_ERR_HIP;
  }
  c->mem_host.resize(nmem);
  for (int k = 0; k < nmem; k++) {
    const auto &M = c->members[k];
    c->mem_host[k] = MemberDev{M.yoff, M.eoff, 0, M.roff, 0, M.dt, 0, 0};
  }
  if (hipMemcpy(c->mem_dev, c->mem_host.data(), nmem * sizeof(MemberDev), hipMemcpyHostToDevice) != hipSuccess) {
    pihm_ctx_destroy(c);
    return PIHM_ERR_HIP;
  }
  A.ne = (int)ne; A.nr = (int)nr;
  A.surf_mode = surf_mode; A.riv_mode = riv_mode; A.dt = c->dt;
  A.record = 0;
  A.nanflag = c->nanflag_dev;
  A.nmem = nmem;
  A.mem = c->mem_dev;
  A.emem = emem_dev;
  A.rmem = rmem_dev;
  A.mio = nullptr;
  A.mnan = nullptr;
  A.bm = nullptr;
  A.bm_mem = nullptr;
  A.nact = 0;
  A.closure = 0;
  c->closure = 0;
  c->stage_host = nullptr;
  *out = c;
  return PIHM_OK;
}

// Caller-order offsets of the loaded slots: members are concatenated in slot
// order at their actual sizes.
void update_caller_offsets(pihm_ctx *c) {
  long e = 0, r = 0, y = 0;
  for (auto &M : c->members) {
    M.ceoff = (int)e; M.croff = (int)r; M.cyoff = y;
    e += M.ne; r += M.nr; y += M.neq;
  }
  c->ne_loaded = (int)e; c->nr_loaded = (int)r; c->neq_loaded = y;
}

// Load one mesh into slot k: its elements are renumbered along the Hilbert
// curve of its own centroids and every array of the slot is written.  The
// carried river-edge fluxes of the slot restart at 0.  Asynchronous uploads
// on the context's stream; returns after they complete.
int ctx_load(pihm_ctx *c, int k, const pihm_mesh_desc *m) {
  if (!c || k < 0 || k >= c->nmem) return PIHM_ERR_ARG;
  int rc = validate(m);
  if (rc != PIHM_OK) return rc;
  if (m->surf_mode != c->surf_mode || m->riv_mode != c->riv_mode) return PIHM_ERR_ARG;
  const SlotCap cap = c->caps[k];
  const int n = m->nelem, nrk = m->nriver;
  if (n > cap.ne || nrk > cap.nr) return PIHM_ERR_TOO_LARGE;
  auto &M = c->members[k];
  const int off = M.eoff, roff = M.roff;
  const long ne = c->ne;

  // Hilbert order of the centroids
  double xmin = m->x[0], xmax = m->x[0], ymin = m->y[0], ymax = m->y[0];
  for (int i = 1; i < n; i++) {
    xmin = std::min(xmin, m->x[i]); xmax = std::max(xmax, m->x[i]);
    ymin = std::min(ymin, m->y[i]); ymax = std::max(ymax, m->y[i]);
  }
  const double span = std::max(std::max(xmax - xmin, ymax - ymin), 1e-300);
  std::vector<uint64_t> key(n);
  for (int i = 0; i < n; i++) {
    const double fx = (m->x[i] - xmin) / span, fy = (m->y[i] - ymin) / span;
    const uint32_t qx = (uint32_t)std::min(65535.0, std::max(0.0, std::floor(fx * 65535.0)));
    const uint32_t qy = (uint32_t)std::min(65535.0, std::max(0.0, std::floor(fy * 65535.0)));
    key[i] = hilbert_key(qx, qy);
  }
  std::vector<int32_t> order(n);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) { return key[a] < key[b]; });
  for (int lp = 0; lp < cap.ne; lp++) {
    c->iperm[off + lp] = -1;
    c->perm[off + lp] = (lp < n) ? order[lp] : -1;
  }
  for (int lp = 0; lp < n; lp++) c->iperm[off + order[lp]] = off + lp;

  // neighbour codes
  std::vector<int32_t> nbk(3L * cap.ne, NB_BOUND);
  for (int lp = 0; lp < n; lp++) {
    const int i = order[lp];
    for (int j = 0; j < 3; j++) {
      const int v = m->nabr[3 * i + j];
      int32_t code;
      if (v > 0) code = c->iperm[off + v - 1];
      else if (v == 0) {
        const int bt = m->bc_type[3 * i + j];
        code = NB_BOUND + (bt == 0 ? 0 : (bt > 0 ? 1 : 2));
      }
      else {
        const int r = -v - 1;
        const int side = (m->leftele[r] == i + 1) ? 0 : 1;
        code = -(2 * (roff + r) + side + 1);
      }
      nbk[(long)j * cap.ne + lp] = code;
    }
  }
  for (int j = 0; j < 3; j++)
    for (int lp = 0; lp < cap.ne; lp++) c->nb_host[(long)j * ne + off + lp] = nbk[(long)j * cap.ne + lp];
  // upstream lists in ascending caller index (src/river_flow.c:96-110)
  std::vector<int32_t> ub(std::max(cap.nr, 1)), ue(std::max(cap.nr, 1)), uidx(std::max(cap.nr, 1), 0);
  {
    std::vector<int32_t> cnt(cap.nr + 1, 0);
    for (int r = 0; r < nrk; r++)
      if (m->down[r] > 0) cnt[m->down[r]]++;
    for (int r = 0; r < cap.nr; r++) cnt[r + 1] += cnt[r];
    std::vector<int32_t> fill(cnt.begin(), cnt.end() - 1);
    for (int r = 0; r < nrk; r++)
      if (m->down[r] > 0) uidx[fill[m->down[r] - 1]++] = roff + r;
    for (int r = 0; r < cap.nr; r++) {
      ub[r] = c->uoff[k] + cnt[r];
      ue[r] = c->uoff[k] + cnt[r + 1];
    }
  }

  RhsArgs &A = c->args;
  ElemDev &E = A.e;
  RiverDev &R = A.r;
  hipStream_t st = c->stream;
  std::vector<double> hd(3L * cap.ne, 0.0);
  const double *es[] = {m->area, m->zmin, m->zmax, m->depth, m->ksath, m->ksatv, m->kinfv,
                        m->dinf, m->alpha, m->beta, m->porosity, m->dmac, m->kmach, m->kmacv,
                        m->areafv, m->areafh, m->rough, m->rzd};
  double *ed[] = {E.area, E.zmin, E.zmax, E.depth, E.ksath, E.ksatv, E.kinfv, E.dinf,
                  E.alpha, E.beta, E.porosity, E.dmac, E.kmach, E.kmacv, E.areafv,
                  E.areafh, E.rough, E.rzd};
  for (int f = 0; f < 18; f++) {
    for (int lp = 0; lp < n; lp++) hd[lp] = es[f][order[lp]];
    PIHM_HIP_TRY(hipMemcpyAsync(ed[f] + off, hd.data(), n * sizeof(double), hipMemcpyHostToDevice, st));
    PIHM_HIP_TRY(hipStreamSynchronize(st));
  }
  // per-step inputs of the slot start at 0
  std::fill(hd.begin(), hd.end(), 0.0);
  double *z1[] = {E.pcpdrp, E.edir, E.ett, E.ws0s};
  for (auto p : z1) PIHM_HIP_TRY(hipMemcpyAsync(p + off, hd.data(), cap.ne * sizeof(double), hipMemcpyHostToDevice, st));
  for (int j = 0; j < 3; j++)
    PIHM_HIP_TRY(hipMemcpyAsync(E.ebc + (long)j * ne + off, hd.data(), cap.ne * sizeof(double), hipMemcpyHostToDevice, st));
  PIHM_HIP_TRY(hipStreamSynchronize(st));
  const double *e3s[] = {m->edge, m->nabrdist, m->nabr_x, m->nabr_y};
  double *e3d[] = {E.edge, E.nabrdist, E.nabr_x, E.nabr_y};
  for (int f = 0; f < 4; f++) {
    for (int j = 0; j < 3; j++) {
      for (int lp = 0; lp < n; lp++) hd[lp] = e3s[f][3L * order[lp] + j];
      PIHM_HIP_TRY(hipMemcpyAsync(e3d[f] + (long)j * ne + off, hd.data(), n * sizeof(double), hipMemcpyHostToDevice, st));
      PIHM_HIP_TRY(hipStreamSynchronize(st));
    }
  }
  for (int j = 0; j < 3; j++) {
    PIHM_HIP_TRY(hipMemcpyAsync(E.nb + (long)j * ne + off, nbk.data() + (long)j * cap.ne, cap.ne * 4,
                                hipMemcpyHostToDevice, st));
    PIHM_HIP_TRY(hipStreamSynchronize(st));
  }
  {
    std::vector<NbrRec> rec(cap.ne);
    std::memset(rec.data(), 0, cap.ne * sizeof(NbrRec));
    for (int lp = 0; lp < n; lp++) {
      rec[lp].rough = m->rough[order[lp]];
      rec[lp].zmax = m->zmax[order[lp]];
    }
    PIHM_HIP_TRY(hipMemcpyAsync(E.rec + off, rec.data(), cap.ne * sizeof(NbrRec), hipMemcpyHostToDevice, st));
    PIHM_HIP_TRY(hipStreamSynchronize(st));
  }
  if (cap.nr > 0) {
    std::vector<int32_t> down(cap.nr, -3), left(cap.nr, 0), right(cap.nr, 0), bct(cap.nr, 0), ordv(cap.nr, 1);
    for (int r = 0; r < nrk; r++) {
      down[r] = (m->down[r] > 0) ? roff + m->down[r] - 1 : m->down[r];
      left[r] = c->iperm[off + m->leftele[r] - 1];
      right[r] = c->iperm[off + m->rightele[r] - 1];
      bct[r] = m->riverbc_type[r];
      ordv[r] = m->intrpl_ord[r];
    }
    int32_t *rdst[] = {R.down, R.left, R.right, R.bctype, R.ord, R.up_beg, R.up_end, R.up_idx + c->uoff[k]};
    const int32_t *rsrc[] = {down.data(), left.data(), right.data(), bct.data(), ordv.data(), ub.data(), ue.data(),
                             uidx.data()};
    for (int f = 0; f < 8; f++) {
      int32_t *dst = (f < 7) ? rdst[f] + roff : rdst[f];
      PIHM_HIP_TRY(hipMemcpyAsync(dst, rsrc[f], cap.nr * 4, hipMemcpyHostToDevice, st));
    }
    const double *rs[] = {m->rarea, m->rzmin, m->rzmax, m->zbed, m->node_zmax, m->dist_left,
                          m->dist_right, m->rdepth, m->coeff, m->length, m->width, m->rrough,
                          m->cwr, m->rksath, m->rksatv, m->bedthick, m->rporosity};
    double *rdd[] = {R.area, R.zmin, R.zmax, R.zbed, R.node_zmax, R.dist_l, R.dist_r,
                     R.depth, R.coeff, R.length, R.width, R.rough, R.cwr, R.ksath,
                     R.ksatv, R.bedthick, R.porosity};
    std::vector<double> rv(cap.nr, 1.0);
    for (int f = 0; f < 17; f++) {
      for (int r = 0; r < nrk; r++) rv[r] = rs[f][r];
      PIHM_HIP_TRY(hipMemcpyAsync(rdd[f] + roff, rv.data(), cap.nr * sizeof(double), hipMemcpyHostToDevice, st));
      PIHM_HIP_TRY(hipStreamSynchronize(st));
    }
    std::vector<double> zr(2L * cap.nr, 0.0);
    PIHM_HIP_TRY(hipMemcpyAsync(R.bc + roff, zr.data(), cap.nr * sizeof(double), hipMemcpyHostToDevice, st));
    PIHM_HIP_TRY(hipMemcpyAsync(R.hid_ovl + 2L * roff, zr.data(), 2L * cap.nr * sizeof(double), hipMemcpyHostToDevice, st));
  }
  M.ne = n;
  M.nr = nrk;
  M.neq = 3 * n + 2 * nrk;
  M.dt = m->stepsize;
  c->mem_host[k] = MemberDev{M.yoff, M.eoff, M.ne, M.roff, M.nr, M.dt, 1, 0};
  PIHM_HIP_TRY(hipMemcpyAsync(c->mem_dev + k, &c->mem_host[k], sizeof(MemberDev), hipMemcpyHostToDevice, st));
  PIHM_HIP_TRY(hipStreamSynchronize(st));
  if (k == 0) { c->dt = m->stepsize; c->args.dt = m->stepsize; }
  update_caller_offsets(c);
  return PIHM_OK;
}

// Build a context for one or more solves (members), each slot exactly the
// size of its mesh.
int build_ctx(const std::vector<const pihm_mesh_desc *> &ms, pihm_ctx **out) {
  *out = nullptr;
  if (ms.empty()) return PIHM_ERR_ARG;
  std::vector<SlotCap> caps;
  for (auto m : ms) {
    int rc = validate(m);
    if (rc != PIHM_OK) return rc;
    if (m->surf_mode != ms[0]->surf_mode || m->riv_mode != ms[0]->riv_mode) return PIHM_ERR_ARG;
    caps.push_back(SlotCap{m->nelem, m->nriver});
  }
  pihm_ctx *c = nullptr;
  int rc = ctx_alloc(caps, ms[0]->surf_mode, ms[0]->riv_mode, &c);
  if (rc) return rc;
  for (int k = 0; k < (int)ms.size(); k++) {
    if ((rc = ctx_load(c, k, ms[k]))) { pihm_ctx_destroy(c); return rc; }
  }
  if ((rc = build_tiles(c))) { pihm_ctx_destroy(c); return rc; }
  *out = c;
  return PIHM_OK;
}

// Tile tables of the one-launch RHS (TileArgs) for a single-solve context.
// Each element block holds s1cap records and scap heads in LDS; a mesh whose
// halos would need more than kTileLdsMax bytes keeps the two-kernel RHS.
constexpr size_t kTileLdsMax = 64 * 1024;

int build_tiles(pihm_ctx *c) {
  const long ne = c->ne;
  if (c->nmem != 1 || ne < 1) return PIHM_OK;
  const int ntile = div_up((int)ne, kTile);
  std::vector<int32_t> hbeg(ntile + 1, 0), hn1(ntile, 0), hidx, hnb, lnb(3 * ne);
  std::vector<int32_t> stamp(ne, -1), slot(ne, 0), r1, r2;
  const auto code = [&](int j, long q) { return c->nb_host[(long)j * ne + q]; };
  int n1max = 0, nhmax = 0;
  for (int t = 0; t < ntile; t++) {
    const long lo = (long)t * kTile, hi = std::min(ne, lo + kTile);
    for (long q = lo; q < hi; q++) { stamp[q] = t; slot[q] = (int)(q - lo); }
    r1.clear();
    r2.clear();
    for (long q = lo; q < hi; q++)
      for (int j = 0; j < 3; j++) {
        const int cc = code(j, q);
        if (cc >= 0 && stamp[cc] != t) { stamp[cc] = t; r1.push_back(cc); }
      }
    std::sort(r1.begin(), r1.end());
    for (size_t i = 0; i < r1.size(); i++) slot[r1[i]] = kTile + (int)i;
    for (int q : r1)
      for (int j = 0; j < 3; j++) {
        const int cc = code(j, q);
        if (cc >= 0 && stamp[cc] != t) { stamp[cc] = t; r2.push_back(cc); }
      }
    std::sort(r2.begin(), r2.end());
    for (size_t i = 0; i < r2.size(); i++) slot[r2[i]] = kTile + (int)(r1.size() + i);
    // every element neighbour of the tile and of its first ring now has a slot
    for (long q = lo; q < hi; q++)
      for (int j = 0; j < 3; j++) {
        const int cc = code(j, q);
        lnb[(long)j * ne + q] = (cc >= 0) ? slot[cc] : cc;
      }
    for (int q : r1)
      for (int j = 0; j < 3; j++) {
        const int cc = code(j, q);
        hnb.push_back((cc >= 0) ? slot[cc] : cc);
      }
    for (size_t i = 0; i < 3 * r2.size(); i++) hnb.push_back(NB_BOUND);
    hidx.insert(hidx.end(), r1.begin(), r1.end());
    hidx.insert(hidx.end(), r2.begin(), r2.end());
    hbeg[t + 1] = (int32_t)hidx.size();
    hn1[t] = (int32_t)r1.size();
    n1max = std::max(n1max, (int)r1.size());
    nhmax = std::max(nhmax, (int)(r1.size() + r2.size()));
  }
  const int s1cap = kTile + n1max, scap = kTile + nhmax;
  // 8 record fields, heads, bank fluxes of the first river edge (two fields), that edge
  const size_t lds = (8 * (size_t)s1cap + (size_t)scap + 2 * (size_t)kTile) * sizeof(double) + kTile * sizeof(int);
  if (lds > kTileLdsMax) return PIHM_OK;       // the two-kernel RHS stays
  const size_t nh = hidx.size();
  const size_t words = (ntile + 1) + ntile + nh + 3 * nh + 3 * ne;
  int32_t *d = nullptr;
  PIHM_HIP_TRY(hipMalloc((void **)&d, std::max<size_t>(words, 1) * sizeof(int32_t)));
  c->tile_arena = d;
  int32_t *w = d;
  const auto put = [&](const std::vector<int32_t> &v) -> const int32_t * {
    if (!v.empty() && hipMemcpy(w, v.data(), v.size() * sizeof(int32_t), hipMemcpyHostToDevice) != hipSuccess)
      return nullptr;
    const int32_t *at = w;
    w += v.size();
    return at;
  };
  TileArgs T{};
  T.hbeg = put(hbeg);
  T.hn1 = put(hn1);
  T.hidx = put(hidx);
  T.hnb = put(hnb);
  T.lnb = put(lnb);
  if (!T.hbeg || !T.hn1 || (nh && (!T.hidx || !T.hnb)) || !T.lnb) return PIHM_ERR_HIP;
  T.s1cap = s1cap;
  T.scap = scap;
  c->tiles = T;
  c->tile_lds = lds;
  c->one_launch_ok = 1;
  c->rhs_path = PIHM_RHS_ONE_LAUNCH;
  return PIHM_OK;
}

// Apply f(member, host_base, device_base) for the member-major caller-order
// layout of a concatenated state: perm maps device -> caller within members.
void state_permute(const pihm_ctx *c, const double *src, double *dst, bool to_device) {
  for (int k = 0; k < c->nmem; k++) {
    const auto &M = c->members[k];
    const double *s = src + (to_device ? M.cyoff : M.yoff);
    double *d = dst + (to_device ? M.yoff : M.cyoff);
    for (int b = 0; b < 3; b++)
      for (int lp = 0; lp < M.ne; lp++) {
        const long dv = (long)b * M.ne + lp, hv = (long)b * M.ne + c->perm[M.eoff + lp];
        if (to_device) d[dv] = s[hv];
        else d[hv] = s[dv];
      }
    for (long i = 3L * M.ne; i < M.neq; i++) d[i] = s[i];
  }
}

}  // namespace

extern "C" int pihm_ctx_create(const pihm_mesh_desc *m, pihm_ctx **out) {
  if (!out) return PIHM_ERR_ARG;
  std::vector<const pihm_mesh_desc *> ms{m};
  return build_ctx(ms, out);
}

extern "C" int pihm_ctx_create_batch(int nmem, const pihm_mesh_desc *const *meshes, pihm_ctx **out) {
  if (!out || nmem < 1 || !meshes) return PIHM_ERR_ARG;
  std::vector<const pihm_mesh_desc *> ms(meshes, meshes + nmem);
  return build_ctx(ms, out);
}

extern "C" int pihm_ctx_create_slots(int nslots, const int32_t *ne_cap, const int32_t *nr_cap, int surf_mode,
                                     int riv_mode, pihm_ctx **out) {
  if (!out || nslots < 1 || !ne_cap || !nr_cap) return PIHM_ERR_ARG;
  std::vector<SlotCap> caps(nslots);
  for (int k = 0; k < nslots; k++) caps[k] = SlotCap{ne_cap[k], nr_cap[k]};
  return ctx_alloc(caps, surf_mode, riv_mode, out);
}

extern "C" int pihm_ctx_load_member(pihm_ctx *c, int k, const pihm_mesh_desc *m) {
  return ctx_load(c, k, m);
}

extern "C" void pihm_ctx_destroy(pihm_ctx *c) {
  if (!c) return;
  if (c->arena) (void)hipFree(c->arena);
  if (c->noah_arena) (void)hipFree(c->noah_arena);
  if (c->tile_arena) (void)hipFree(c->tile_arena);
  if (c->nanflag_host) (void)hipHostFree(c->nanflag_host);
  delete c;
}

extern "C" int pihm_ctx_set_rhs_path(pihm_ctx *c, int path) {
  if (!c || c->nmem != 1) return PIHM_ERR_ARG;
  if (path == PIHM_RHS_ONE_LAUNCH && !c->one_launch_ok) return PIHM_ERR_ARG;
  if (path != PIHM_RHS_ONE_LAUNCH && path != PIHM_RHS_TWO_KERNELS) return PIHM_ERR_ARG;
  c->rhs_path = path;
  return PIHM_OK;
}

extern "C" int pihm_ctx_get_rhs_path(const pihm_ctx *c) { return c ? c->rhs_path : PIHM_ERR_ARG; }

extern "C" int pihm_ctx_set_stream(pihm_ctx *c, void *s) {
  if (!c) return PIHM_ERR_ARG;
  c->stream = (hipStream_t)s;
  return PIHM_OK;
}

extern "C" int pihm_stream_create(void **out) {
  if (!out) return PIHM_ERR_ARG;
  hipStream_t s = nullptr;
  PIHM_HIP_TRY(hipStreamCreate(&s));
  *out = s;
  return PIHM_OK;
}

extern "C" int pihm_stream_destroy(void *s) {
  if (!s) return PIHM_ERR_ARG;
  PIHM_HIP_TRY(hipStreamSynchronize((hipStream_t)s));
  PIHM_HIP_TRY(hipStreamDestroy((hipStream_t)s));
  return PIHM_OK;
}

extern "C" int pihm_ctx_neq(const pihm_ctx *c) { return c ? c->neq : -1; }
extern "C" int pihm_ctx_nmembers(const pihm_ctx *c) { return c ? c->nmem : -1; }

extern "C" int pihm_ctx_member_info(const pihm_ctx *c, int k, long *yoff, int *neq, int *ne, int *nr) {
  if (!c || k < 0 || k >= c->nmem) return PIHM_ERR_ARG;
  const auto &M = c->members[k];
  if (yoff) *yoff = M.yoff;
  if (neq) *neq = M.neq;
  if (ne) *ne = M.ne;
  if (nr) *nr = M.nr;
  return PIHM_OK;
}

extern "C" int pihm_ctx_set_active(pihm_ctx *c, const int *active) {
  if (!c || !active) return PIHM_ERR_ARG;
  for (int k = 0; k < c->nmem; k++) c->mem_host[k].active = active[k] ? 1 : 0;
  PIHM_HIP_TRY(hipMemcpyAsync(c->mem_dev, c->mem_host.data(), c->nmem * sizeof(MemberDev),
                              hipMemcpyHostToDevice, c->stream));
  return PIHM_OK;
}

extern "C" int pihm_ctx_set_nan_check(pihm_ctx *c, int on) {
  if (!c) return PIHM_ERR_ARG;
  c->nan_check = on ? 1 : 0;
  return PIHM_OK;
}

extern "C" int pihm_ctx_set_closure(pihm_ctx *c, int on) {
  if (!c || c->nmem != 1 || c->members[0].ne != c->caps[0].ne || c->members[0].nr != c->caps[0].nr)
    return PIHM_ERR_ARG;
  on = on ? 1 : 0;
  if (on == c->closure) return PIHM_OK;
  const int d = on ? 4 : -4;
  c->closure = on;
  c->args.closure = on;
  c->neq += d;
  c->members[0].neq += d;
  update_caller_offsets(c);
  return PIHM_OK;
}

extern "C" int pihm_ctx_set_record(pihm_ctx *c, int on) {
  if (!c) return PIHM_ERR_ARG;
  c->record = on ? 1 : 0;
  return PIHM_OK;
}

extern "C" int pihm_ctx_get_perm(const pihm_ctx *c, int32_t *out) {
  if (!c || !out) return PIHM_ERR_ARG;
  std::memcpy(out, c->perm.data(), c->ne * sizeof(int32_t));
  return PIHM_OK;
}

// GwTransp (src/noah/lsm_func.c:163-185) per element: the share of
// transpiration drawn from the saturated zone, from the layers between the
// water table layer nwtbl and the root layer nroot, summed in ascending order.
__global__ __launch_bounds__(kBlock) void k_gwtransp(ElemDev e, int ne) {
  const int p = blockIdx.x * kBlock + threadIdx.x;
  if (p >= ne) return;
  const double ett = e.ett[p];
  const int nwtbl = e.nwtbl[p], nroot = e.nroot[p];
  double gw = 0.0;
  if (ett > 0.0 && nwtbl <= nroot) {
    for (int j = (nwtbl <= 0 ? 0 : nwtbl - 1); j < nroot; j++) gw += e.et[(long)j * ne + p];
    gw = gw / ett;
    gw = (gw > 1.0) ? 1.0 : gw;
    gw = (gw < 0.0) ? 0.0 : gw;
  }
  e.gwet[p] = gw;
}

// The saturated zone's share of transpiration follows its inputs: ett (step
// inputs, device forcing), et and nroot (Noah hand-over), nwtbl (hand-over,
// soil-water step).  The reference evaluates it in every RHS call from the
// current values; evaluating it whenever one of them changes gives the same.
int pihm_ctx_noah_refresh(pihm_ctx *c) {
  if (!c->noah) return PIHM_OK;
  hipLaunchKernelGGL(k_gwtransp, dim3(div_up(c->ne, kBlock)), dim3(kBlock), 0, c->stream, c->args.e, c->ne);
  PIHM_HIP_TRY(hipGetLastError());
  return PIHM_OK;
}

extern "C" int pihm_ctx_set_step_inputs(pihm_ctx *c, const pihm_step_inputs *in) {
  if (!c || !in) return PIHM_ERR_ARG;
  const int ne = c->ne, nr = c->nr;
  std::vector<double> hd(3L * ne);
  // element arrays are concatenated over members in caller order: member k's
  // element i sits at members[k].eoff + i
  auto up1 = [&](double *dst, const double *src) -> int {
    for (int p = 0; p < ne; p++) {
      const long ci = caller_elem(c, p);
      hd[p] = (ci >= 0) ? src[ci] : 0.0;
    }
    PIHM_HIP_TRY(hipMemcpyAsync(dst, hd.data(), ne * sizeof(double), hipMemcpyHostToDevice, c->stream));
    PIHM_HIP_TRY(hipStreamSynchronize(c->stream));
    return PIHM_OK;
  };
  ElemDev &E = c->args.e;
  int rc;
  if (in->pcpdrp && (rc = up1(E.pcpdrp, in->pcpdrp))) return rc;
  if (in->edir && (rc = up1(E.edir, in->edir))) return rc;
  if (in->ett && (rc = up1(E.ett, in->ett))) return rc;
  if (in->ett && (rc = pihm_ctx_noah_refresh(c))) return rc;
  if (in->ws0_surf && (rc = up1(E.ws0s, in->ws0_surf))) return rc;
  if (in->elem_bc) {
    for (int j = 0; j < 3; j++)
      for (int p = 0; p < ne; p++)
        hd[(long)j * ne + p] = (caller_elem(c, p) >= 0) ? in->elem_bc[3L * caller_elem(c, p) + j] : 0.0;
    PIHM_HIP_TRY(hipMemcpyAsync(E.ebc, hd.data(), 3L * ne * sizeof(double), hipMemcpyHostToDevice, c->stream));
    PIHM_HIP_TRY(hipStreamSynchronize(c->stream));
  }
  if (in->river_bc && nr > 0) {
    std::vector<double> rb(nr, 0.0);
    for (int r = 0; r < nr; r++) {
      const long cr = caller_river(c, r);
      if (cr >= 0) rb[r] = in->river_bc[cr];
    }
    PIHM_HIP_TRY(hipMemcpyAsync(c->args.r.bc, rb.data(), nr * sizeof(double), hipMemcpyHostToDevice, c->stream));
    PIHM_HIP_TRY(hipStreamSynchronize(c->stream));
  }
  return PIHM_OK;
}

extern "C" int pihm_ctx_set_noah(pihm_ctx *c, const double *smcmin) {
  if (!c || c->nmem != 1) return PIHM_ERR_ARG;
  if (!smcmin) {
    c->noah = 0;
    return PIHM_OK;
  }
  const long ne = c->ne;
  ElemDev &E = c->args.e;
  if (!c->noah_arena) {
    Arena ar;
    double **f[] = {&E.smcmin, &E.sh2o0, &E.fcr, &E.gwet};
    for (auto q : f) ar.add(q, ne);
    ar.add(&E.et, (long)PIHM_NOAH_MAXLYR * ne);
    ar.add(&E.nwtbl, ne);
    ar.add(&E.nroot, ne);
    PIHM_HIP_TRY(hipMalloc(&c->noah_arena, ar.total()));
    ar.carve((char *)c->noah_arena);
  }
  // no layer transpiration, water table or roots until the first hand-over
  PIHM_HIP_TRY(hipMemsetAsync(E.et, 0, (size_t)PIHM_NOAH_MAXLYR * ne * sizeof(double), c->stream));
  PIHM_HIP_TRY(hipMemsetAsync(E.nwtbl, 0, ne * sizeof(int32_t), c->stream));
  PIHM_HIP_TRY(hipMemsetAsync(E.nroot, 0, ne * sizeof(int32_t), c->stream));
  std::vector<double> hd(ne);
  for (int p = 0; p < ne; p++) {
    const long ci = caller_elem(c, p);
    hd[p] = (ci >= 0) ? smcmin[ci] : 0.0;
  }
  PIHM_HIP_TRY(hipMemcpyAsync(E.smcmin, hd.data(), ne * sizeof(double), hipMemcpyHostToDevice, c->stream));
  PIHM_HIP_TRY(hipStreamSynchronize(c->stream));
  // until the first step's values: unfrozen soil, no saturated-zone transpiration
  std::fill(hd.begin(), hd.end(), 1.0);
  PIHM_HIP_TRY(hipMemcpyAsync(E.fcr, hd.data(), ne * sizeof(double), hipMemcpyHostToDevice, c->stream));
  PIHM_HIP_TRY(hipStreamSynchronize(c->stream));
  PIHM_HIP_TRY(hipMemsetAsync(E.gwet, 0, ne * sizeof(double), c->stream));
  PIHM_HIP_TRY(hipMemsetAsync(E.sh2o0, 0, ne * sizeof(double), c->stream));
  PIHM_HIP_TRY(hipStreamSynchronize(c->stream));
  c->noah = 1;
  return PIHM_OK;
}

extern "C" int pihm_ctx_set_noah_step(pihm_ctx *c, const double *sh2o0, const double *fcr, const double *et,
                                      int nlayer, const int32_t *nwtbl, const int32_t *nroot) {
  if (!c || !c->noah) return PIHM_ERR_ARG;
  const bool gw = et || nwtbl || nroot;
  if (gw && (!et || !nwtbl || !nroot || nlayer < 1 || nlayer > PIHM_NOAH_MAXLYR)) return PIHM_ERR_ARG;
  const long ne = c->ne;
  ElemDev &E = c->args.e;
  // every argument is checked before the first upload, so a rejected call
  // leaves the context as it was
  if (gw)
    for (long p = 0; p < ne; p++) {
      const long ci = caller_elem(c, (int)p);
      if (ci >= 0 && nroot[ci] > nlayer) return PIHM_ERR_ARG;   // a root layer past the given layers
    }
  std::vector<double> hd(ne);
  auto up = [&](double *dst, const double *src, long stride_src) -> int {
    for (int p = 0; p < ne; p++) {
      const long ci = caller_elem(c, p);
      hd[p] = (ci >= 0) ? src[ci * stride_src] : 0.0;
    }
    PIHM_HIP_TRY(hipMemcpyAsync(dst, hd.data(), ne * sizeof(double), hipMemcpyHostToDevice, c->stream));
    PIHM_HIP_TRY(hipStreamSynchronize(c->stream));
    return PIHM_OK;
  };
  int rc;
  if (sh2o0 && (rc = up(E.sh2o0, sh2o0, 1))) return rc;
  if (fcr && (rc = up(E.fcr, fcr, 1))) return rc;
  if (gw) {
    for (int j = 0; j < nlayer; j++)
      if ((rc = up(E.et + (long)j * ne, et + j, nlayer))) return rc;
    // layers past the given ones carry no transpiration (the soil-water step
    // reads every layer below nsoil)
    if (nlayer < PIHM_NOAH_MAXLYR)
      PIHM_HIP_TRY(hipMemsetAsync(E.et + (long)nlayer * ne, 0,
                                  (size_t)(PIHM_NOAH_MAXLYR - nlayer) * ne * sizeof(double), c->stream));
    std::vector<int32_t> hi(ne);
    const int32_t *srcs[] = {nwtbl, nroot};
    int32_t *dsts[] = {E.nwtbl, E.nroot};
    for (int f = 0; f < 2; f++) {
      for (int p = 0; p < ne; p++) {
        const long ci = caller_elem(c, p);
        hi[p] = (ci >= 0) ? srcs[f][ci] : 0;
      }
      PIHM_HIP_TRY(hipMemcpyAsync(dsts[f], hi.data(), ne * sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
      PIHM_HIP_TRY(hipStreamSynchronize(c->stream));
    }
    hipLaunchKernelGGL(k_gwtransp, dim3(div_up((int)ne, kBlock)), dim3(kBlock), 0, c->stream, E, (int)ne);
    PIHM_HIP_TRY(hipGetLastError());
  }
  return PIHM_OK;
}

extern "C" int pihm_ctx_get_step_inputs(pihm_ctx *c, double *pcpdrp, double *edir, double *ett,
                                        double *ws0_surf, double *elem_bc, double *river_bc) {
  if (!c) return PIHM_ERR_ARG;
  const int ne = c->ne, nr = c->nr;
  std::vector<double> hd(3L * ne);
  auto down = [&](const double *src, long n) -> int {
    PIHM_HIP_TRY(hipMemcpyAsync(hd.data(), src, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    PIHM_HIP_TRY(hipStreamSynchronize(c->stream));
    return PIHM_OK;
  };
  const ElemDev &E = c->args.e;
  const double *src[] = {E.pcpdrp, E.edir, E.ett, E.ws0s};
  double *dst[] = {pcpdrp, edir, ett, ws0_surf};
  for (int f = 0; f < 4; f++) {
    if (!dst[f]) continue;
    if (int rc = down(src[f], ne)) return rc;
    for (int p = 0; p < ne; p++) {
      const long ci = caller_elem(c, p);
      if (ci >= 0) dst[f][ci] = hd[p];
    }
  }
  if (elem_bc) {
    if (int rc = down(E.ebc, 3L * ne)) return rc;
    for (int j = 0; j < 3; j++)
      for (int p = 0; p < ne; p++) {
        const long ci = caller_elem(c, p);
        if (ci >= 0) elem_bc[3L * ci + j] = hd[(long)j * ne + p];
      }
  }
  if (river_bc && nr > 0) {
    if (int rc = down(c->args.r.bc, nr)) return rc;
    for (int r = 0; r < nr; r++) {
      const long cr = caller_river(c, r);
      if (cr >= 0) river_bc[cr] = hd[r];
    }
  }
  return PIHM_OK;
}

static int check_vec(const pihm_ctx *c, N_Vector v) {
  if (!v || !v->ops || v->ops->nvgetvectorid != N_VGetVectorID_Hip) return PIHM_ERR_ARG;
  if (N_VGetLength_Hip(v) != c->neq) return PIHM_ERR_ARG;
  return PIHM_OK;
}

extern "C" int pihm_ctx_state_to_device(pihm_ctx *c, const double *host, N_Vector dev) {
  if (!c || !host || check_vec(c, dev)) return PIHM_ERR_ARG;
  std::vector<double> tmp(c->neq);
  state_permute(c, host, tmp.data(), true);
  return N_VCopyToDevice_Hip(dev, tmp.data());
}

extern "C" int pihm_ctx_state_from_device(pihm_ctx *c, N_Vector dev, double *host) {
  if (!c || !host || check_vec(c, dev)) return PIHM_ERR_ARG;
  std::vector<double> tmp(c->neq);
  int rc = N_VCopyFromDevice_Hip(dev, tmp.data());
  if (rc) return rc;
  state_permute(c, tmp.data(), host, false);
  return PIHM_OK;
}

// Launch the two kernels of one RHS call (no synchronisation).
extern "C" int pihm_rhs_launch(pihm_ctx *c, const double *y, double *ydot) {
  RhsArgs a = c->args;
  a.y = y;
  a.ydot = ydot;
  a.record = c->record;
  const int nbe = div_up(c->ne, kBlock);
  const int nbr = div_up(c->nr, kBlock);
  if (c->nmem > 1) {
    hipLaunchKernelGGL(k_prep<true>, dim3(nbe + 3 * nbr), dim3(kBlock), 0, c->stream, a, nbe);
    hipLaunchKernelGGL(k_flux<true>, dim3(nbe + nbr), dim3(kBlock), 0, c->stream, a, nbe);
  } else {
    launch_single(c, a);
    if (a.closure) {
      const int nb = std::max(1, std::min(kClosureBlocks, div_up(std::max(c->ne, c->nr), kBlock)));
      hipLaunchKernelGGL(k_closure_part, dim3(nb), dim3(kBlock), 0, c->stream, a);
      hipLaunchKernelGGL(k_closure_final, dim3(1), dim3(kBlock), 0, c->stream, a, nb, ydot + (c->neq - 4));
    }
  }
  PIHM_HIP_TRY(hipGetLastError());
  c->ncalls++;
  return PIHM_OK;
}

int pihm_rhs_launch_io(pihm_ctx *c, const MemberIO *mio, int *mnan) {
  if (!c || !mio || c->nmem < 1) return PIHM_ERR_ARG;
  RhsArgs a = c->args;
  a.y = nullptr;
  a.ydot = nullptr;
  a.mio = mio;
  a.mnan = mnan;
  a.record = 0;
  const int nbe = div_up(c->ne, kBlock);
  const int nbr = div_up(c->nr, kBlock);
  hipLaunchKernelGGL(k_prep<true>, dim3(nbe + 3 * nbr), dim3(kBlock), 0, c->stream, a, nbe);
  hipLaunchKernelGGL(k_flux<true>, dim3(nbe + nbr), dim3(kBlock), 0, c->stream, a, nbe);
  PIHM_HIP_TRY(hipGetLastError());
  c->ncalls++;
  return PIHM_OK;
}

int pihm_rhs_launch_compact(pihm_ctx *c, const MemberIO *mio, int *mnan, const int *bm, const int *bm_mem,
                            int nact, int nbe, int nbr) {
  if (!c || !mio || !bm || !bm_mem || nact < 1) return PIHM_ERR_ARG;
  RhsArgs a = c->args;
  a.y = nullptr;
  a.ydot = nullptr;
  a.mio = mio;
  a.mnan = mnan;
  a.record = 0;
  a.bm = bm;
  a.bm_mem = bm_mem;
  a.nact = nact;
  if (nbe + nbr > 0) hipLaunchKernelGGL(k_prep<true>, dim3(nbe + 3 * nbr), dim3(kBlock), 0, c->stream, a, nbe);
  if (nbe + nbr > 0) hipLaunchKernelGGL(k_flux<true>, dim3(nbe + nbr), dim3(kBlock), 0, c->stream, a, nbe);
  PIHM_HIP_TRY(hipGetLastError());
  c->ncalls++;
  return PIHM_OK;
}

// Reads and clears the sticky NaN flag (synchronises the stream).
extern "C" int pihm_ctx_poll_nan(pihm_ctx *c) {
  if (!c) return PIHM_ERR_ARG;
  PIHM_HIP_TRY(hipMemcpyAsync(c->nanflag_host, c->nanflag_dev, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  PIHM_HIP_TRY(hipStreamSynchronize(c->stream));
  if (*c->nanflag_host) {
    *c->nanflag_host = 0;
    PIHM_HIP_TRY(hipMemcpyAsync(c->nanflag_dev, c->nanflag_host, sizeof(int), hipMemcpyHostToDevice, c->stream));
    PIHM_HIP_TRY(hipStreamSynchronize(c->stream));
    return PIHM_ERR_NAN;
  }
  return PIHM_OK;
}

extern "C" int PIHM_ODE_Hip(realtype t, N_Vector y, N_Vector ydot, void *user_data) {
  (void)t;
  pihm_ctx *c = (pihm_ctx *)user_data;
  if (!c || check_vec(c, y) || check_vec(c, ydot)) return PIHM_ERR_ARG;
  int rc = pihm_rhs_launch(c, N_VGetDeviceArrayPointer_Hip(y), N_VGetDeviceArrayPointer_Hip(ydot));
  if (rc) return rc;
  if (c->nan_check) return pihm_ctx_poll_nan(c);
  return PIHM_OK;
}


extern "C" int pihm_ctx_get_river_edge_ovlflow(pihm_ctx *c, double *ovl3) {
  if (!c || !ovl3) return PIHM_ERR_ARG;
  const int ne = c->ne, nr = c->nr;
  std::vector<double> hid(2L * std::max(nr, 1));
  if (nr > 0) {
    PIHM_HIP_TRY(hipMemcpyAsync(hid.data(), c->args.r.hid_ovl, 2L * nr * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    PIHM_HIP_TRY(hipStreamSynchronize(c->stream));
  }
  for (int p = 0; p < ne; p++)
    for (int j = 0; j < 3; j++) {
      const int code = c->nb_host[(long)j * ne + p];
      double v = 0.0;
      if (nb_is_river(code)) v = hid[-code - 1];
      if (caller_elem(c, p) >= 0) ovl3[3L * caller_elem(c, p) + j] = v;
    }
  return PIHM_OK;
}

extern "C" int pihm_ctx_set_river_edge_ovlflow(pihm_ctx *c, const double *ovl3) {
  if (!c || !ovl3) return PIHM_ERR_ARG;
  const int ne = c->ne, nr = c->nr;
  if (nr == 0) return PIHM_OK;
  std::vector<double> hid(2L * nr, 0.0);
  for (int p = 0; p < ne; p++)
    for (int j = 0; j < 3; j++) {
      const int code = c->nb_host[(long)j * ne + p];
      if (nb_is_river(code) && caller_elem(c, p) >= 0) hid[-code - 1] = ovl3[3L * caller_elem(c, p) + j];
    }
  PIHM_HIP_TRY(hipMemcpyAsync(c->args.r.hid_ovl, hid.data(), 2L * nr * sizeof(double), hipMemcpyHostToDevice, c->stream));
  PIHM_HIP_TRY(hipStreamSynchronize(c->stream));
  return PIHM_OK;
}

extern "C" int pihm_ctx_get_fluxes(pihm_ctx *c, double *ovl, double *sub, double *infil,
                                   double *rechg, double *rivflow) {
  if (!c) return PIHM_ERR_ARG;
  if (!c->record) return PIHM_ERR_ARG;  // fluxes are kept only in recording mode
  const int ne = c->ne, nr = c->nr;
  std::vector<double> h(std::max(3L * ne, 11L * nr));
  auto get = [&](const double *src, long n) -> int {
    PIHM_HIP_TRY(hipMemcpyAsync(h.data(), src, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    PIHM_HIP_TRY(hipStreamSynchronize(c->stream));
    return PIHM_OK;
  };
  int rc;
  if (ovl) {
    if ((rc = get(c->args.rec.ovl, 3L * ne))) return rc;
    for (int j = 0; j < 3; j++)
      for (int p = 0; p < ne; p++)
        if (caller_elem(c, p) >= 0) ovl[3L * caller_elem(c, p) + j] = h[(long)j * ne + p];
  }
  if (sub) {
    if ((rc = get(c->args.rec.sub, 3L * ne))) return rc;
    for (int j = 0; j < 3; j++)
      for (int p = 0; p < ne; p++)
        if (caller_elem(c, p) >= 0) sub[3L * caller_elem(c, p) + j] = h[(long)j * ne + p];
  }
  if (infil) {
    if ((rc = get(c->args.rec.infil, ne))) return rc;
    for (int p = 0; p < ne; p++)
      if (caller_elem(c, p) >= 0) infil[caller_elem(c, p)] = h[p];
  }
  if (rechg) {
    if ((rc = get(c->args.rec.rechg, ne))) return rc;
    for (int p = 0; p < ne; p++)
      if (caller_elem(c, p) >= 0) rechg[caller_elem(c, p)] = h[p];
  }
  if (rivflow && nr > 0) {
    if ((rc = get(c->args.r.rf, 11L * nr))) return rc;
    for (int r = 0; r < nr; r++) {
      const long cr = caller_river(c, r);
      if (cr < 0) continue;
      for (int k = 0; k < 11; k++) rivflow[11L * cr + k] = h[(long)k * nr + r];
    }
  }
  return PIHM_OK;
}

extern "C" int pihm_ctx_sync(pihm_ctx *c) {
  if (!c) return PIHM_ERR_ARG;
  PIHM_HIP_TRY(hipStreamSynchronize(c->stream));
  return PIHM_OK;
}

// Average duration of each of the two RHS kernels over `iters` evaluations,
// each kernel between its own pair of HIP events on the context's stream.
extern "C" int pihm_ctx_time_kernels(pihm_ctx *c, const double *y, double *ydot, int iters,
                                     double *ms_prep, double *ms_flux) {
  if (!c || !y || !ydot || iters <= 0) return PIHM_ERR_ARG;
  RhsArgs a = c->args;
  a.y = y;
  a.ydot = ydot;
  a.record = 0;
  a.closure = 0;
  const int nbe = div_up(c->ne, kBlock);
  const int nbr = div_up(c->nr, kBlock);
  hipEvent_t ev[3];
  for (auto &e : ev) PIHM_HIP_TRY(hipEventCreate(&e));
  float t[2] = {0, 0};
  for (int it = 0; it < iters; it++) {
    PIHM_HIP_TRY(hipEventRecord(ev[0], c->stream));
    if (c->nmem > 1) hipLaunchKernelGGL(k_prep<true>, dim3(nbe + 3 * nbr), dim3(kBlock), 0, c->stream, a, nbe);
    else hipLaunchKernelGGL(k_prep<false>, dim3(nbe + 3 * nbr), dim3(kBlock), 0, c->stream, a, nbe);
    PIHM_HIP_TRY(hipEventRecord(ev[1], c->stream));
    if (c->nmem > 1) hipLaunchKernelGGL(k_flux<true>, dim3(nbe + nbr), dim3(kBlock), 0, c->stream, a, nbe);
    else launch_flux_single(c, a, nbe, nbr);
    PIHM_HIP_TRY(hipEventRecord(ev[2], c->stream));
    PIHM_HIP_TRY(hipEventSynchronize(ev[2]));
    for (int k = 0; k < 2; k++) {
      float ms = 0;
      PIHM_HIP_TRY(hipEventElapsedTime(&ms, ev[k], ev[k + 1]));
      t[k] += ms;
    }
  }
  for (auto &e : ev) (void)hipEventDestroy(e);
  if (ms_prep) *ms_prep = t[0] / iters;
  if (ms_flux) *ms_flux = t[1] / iters;
  return PIHM_OK;
}

// ---------------------------------------------------------------------------
// the flux functions' power routine, exposed for its accuracy tests
// ---------------------------------------------------------------------------

namespace {
__global__ __launch_bounds__(kBlock) void k_math_pow(int n, const double *x, const double *y, double *out) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) out[i] = xpow(x[i], y[i]);
}
}  // namespace

extern "C" int pihm_math_pow(int n, const double *x, const double *y, double *out, int on_device) {
  if (n < 0 || (n > 0 && (!x || !y || !out))) return PIHM_ERR_ARG;
  if (n == 0) return PIHM_OK;
  if (!on_device) {
    for (int i = 0; i < n; i++) out[i] = xpow(x[i], y[i]);
    return PIHM_OK;
  }
  double *d = nullptr;
  const size_t bytes = (size_t)n * sizeof(double);
  if (hipMalloc((void **)&d, 3 * bytes) != hipSuccess) return PIHM_ERR_HIP;
  hipError_t e = hipMemcpy(d, x, bytes, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(d + n, y, bytes, hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_math_pow, dim3(div_up(n, kBlock)), dim3(kBlock), 0, 0, n, d, d + n, d + 2L * n);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpy(out, d + 2L * n, bytes, hipMemcpyDeviceToHost);
  (void)hipFree(d);
  return e == hipSuccess ? PIHM_OK : PIHM_ERR_HIP;
}

// Average device time of one whole RHS evaluation (both kernels back to
// back, as the integrator issues them) over `iters` evaluations: one HIP event
// before the first launch and one after the last, on the context's stream.
// One RHS evaluation at y with the fluxes recorded (rec.*, r.rf) and the
// carried river-edge flux put back afterwards (carried_save: 2 nr doubles of
// device scratch), so the next integrator call reads what it would have read.
int pihm_ctx_eval_fluxes(pihm_ctx *c, const double *y, double *ydot, double *carried_save) {
  if (!c || !y || !ydot || c->nmem != 1) return PIHM_ERR_ARG;
  int rc = pihm_ctx_carried_copy(c, carried_save, 0);
  if (rc) return rc;
  RhsArgs a = c->args;
  a.y = y;
  a.ydot = ydot;
  a.record = 1;
  a.closure = 0;
  launch_single(c, a);
  PIHM_HIP_TRY(hipGetLastError());
  return pihm_ctx_carried_copy(c, carried_save, 1);
}

extern "C" int pihm_ctx_set_tile_span(pihm_ctx *c, int span) {
  if (!c || span < 0) return PIHM_ERR_ARG;
  c->args.tile_span = span;
  return PIHM_OK;
}

extern "C" int pihm_ctx_get_tile_span(const pihm_ctx *c) { return c ? c->args.tile_span : PIHM_ERR_ARG; }

// Saves (dir 0) or restores (dir 1) the carried river-edge flux, hid_ovl[0, 2 nr).
int pihm_ctx_carried_copy(pihm_ctx *c, double *buf, int dir) {
  const size_t bytes = 2L * c->nr * sizeof(double);
  if (bytes == 0) return PIHM_OK;
  double *hid = c->args.r.hid_ovl;
  PIHM_HIP_TRY(hipMemcpyAsync(dir ? hid : buf, dir ? buf : hid, bytes, hipMemcpyDeviceToDevice, c->stream));
  return PIHM_OK;
}

extern "C" int pihm_ctx_autotune(pihm_ctx *c, const double *y, double *ydot, int reps, int *best_span, double *ms) {
  if (!c || !y || !ydot || reps <= 0 || c->nmem != 1) return PIHM_ERR_ARG;
  static const int spans[PIHM_TUNE_NSPAN] = PIHM_TUNE_SPANS;
  // candidates: every tile placement on each RHS path the context has
  const int npath = c->one_launch_ok ? 2 : 1;
  const int paths[2] = {c->one_launch_ok ? PIHM_RHS_ONE_LAUNCH : PIHM_RHS_TWO_KERNELS, PIHM_RHS_TWO_KERNELS};
  const int ncand = npath * PIHM_TUNE_NSPAN;
  std::vector<double> tmin(ncand, HUGE_VAL);
  double *save = nullptr;
  if (c->nr > 0) PIHM_HIP_TRY(hipMalloc(&save, 2L * c->nr * sizeof(double)));
  int rc = pihm_ctx_carried_copy(c, save, 0);
  // three interleaved rounds, the best of each candidate's rounds
  for (int round = 0; round < 3 && !rc; round++) {
    for (int k = 0; k < ncand && !rc; k++) {
      c->rhs_path = paths[k / PIHM_TUNE_NSPAN];
      c->args.tile_span = spans[k % PIHM_TUNE_NSPAN];
      double t = 0.0, w = 0.0;
      rc = pihm_ctx_time_rhs(c, y, ydot, 2, &w);
      if (!rc) rc = pihm_ctx_time_rhs(c, y, ydot, reps, &t);
      if (!rc) tmin[k] = std::min(tmin[k], t);
    }
  }
  // the default (candidate 0) stays unless another is more than 2% faster:
  // differences below that are within the spread of repeated timings
  int best = 0;
  for (int k = 1; k < ncand; k++)
    if (tmin[k] < tmin[best]) best = k;
  if (!(tmin[best] < 0.98 * tmin[0])) best = 0;
  c->rhs_path = paths[best / PIHM_TUNE_NSPAN];
  c->args.tile_span = spans[best % PIHM_TUNE_NSPAN];
  if (ms)
    for (int k = 0; k < PIHM_TUNE_NSPAN; k++) {
      ms[k] = tmin[k];
      for (int q = 1; q < npath; q++) ms[k] = std::min(ms[k], tmin[q * PIHM_TUNE_NSPAN + k]);
    }
  if (!rc) rc = pihm_ctx_carried_copy(c, save, 1);
  if (!rc && hipStreamSynchronize(c->stream) != hipSuccess) rc = PIHM_ERR_HIP;
  if (save) (void)hipFree(save);
  if (best_span) *best_span = c->args.tile_span;
  return rc;
}

extern "C" int pihm_ctx_time_rhs(pihm_ctx *c, const double *y, double *ydot, int iters, double *ms_per_eval) {
  if (!c || !y || !ydot || iters <= 0 || !ms_per_eval) return PIHM_ERR_ARG;
  RhsArgs a = c->args;
  a.y = y;
  a.ydot = ydot;
  a.record = 0;
  a.closure = 0;
  const int nbe = div_up(c->ne, kBlock);
  const int nbr = div_up(c->nr, kBlock);
  hipEvent_t ev[2];
  for (auto &e : ev) PIHM_HIP_TRY(hipEventCreate(&e));
  PIHM_HIP_TRY(hipEventRecord(ev[0], c->stream));
  for (int it = 0; it < iters; it++) {
    if (c->nmem > 1) {
      hipLaunchKernelGGL(k_prep<true>, dim3(nbe + 3 * nbr), dim3(kBlock), 0, c->stream, a, nbe);
      hipLaunchKernelGGL(k_flux<true>, dim3(nbe + nbr), dim3(kBlock), 0, c->stream, a, nbe);
    } else {
      launch_single(c, a);
    }
  }
  PIHM_HIP_TRY(hipGetLastError());
  PIHM_HIP_TRY(hipEventRecord(ev[1], c->stream));
  PIHM_HIP_TRY(hipEventSynchronize(ev[1]));
  float ms = 0;
  PIHM_HIP_TRY(hipEventElapsedTime(&ms, ev[0], ev[1]));
  for (auto &e : ev) (void)hipEventDestroy(e);
  *ms_per_eval = ms / iters;
  return PIHM_OK;
}
