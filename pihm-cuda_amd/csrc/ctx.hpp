// ctx.hpp - device-resident MM-PIHM mesh and RHS state.
//
// Layout in HBM (structure of arrays, element arrays in Hilbert order, river
// arrays in the caller's order):
//   element statics   18 x double[ne]      area .. rzd
//   edge statics       4 x double[3*ne]    edge, nabrdist, nabr_x, nabr_y  (index j*ne+p)
//   edge codes         int32[3*ne]         nb (neighbour code, boundary type folded in)
//   per-step inputs    4 x double[ne] + double[3*ne]
//   neighbour record   NbrRec[ne]          values a neighbour reads, one 64-byte load
//   river statics      17 x double[nr] + 6 x int32[nr], upstream CSR
//   river fluxes       double[11*nr] (rf[k*nr+r]), carried and new river-edge fluxes
#pragma once
#include <cstdint>
#include <vector>
#include "../../include/pihm_rhs.h"

namespace pihm {

// Neighbour code nb[j*ne+p]:
//   >= 0                 device index of the neighbouring element
//   NB_BOUND + t         domain boundary; t = 0 no flow, 1 Dirichlet (bc_type
//                        > 0), 2 Neumann (bc_type < 0): only the sign of the
//                        reference's bc_type is used (src/lat_flow.c:340-382)
//   other negative       river edge: -(2*r + side + 1), side 0 = left bank, 1 = right
constexpr int32_t NB_BOUND = INT32_MIN;
__host__ __device__ inline bool nb_is_bound(int32_t c) { return c <= NB_BOUND + 2; }
__host__ __device__ inline bool nb_is_river(int32_t c) { return c < 0 && c > NB_BOUND + 2; }
// the reference's bc_type sign of a boundary code: 0, 1 or -1
__host__ __device__ inline int nb_bc_sign(int32_t c) { return c == NB_BOUND ? 0 : (c == NB_BOUND + 1 ? 1 : -1); }

// What a neighbouring element needs from element p in one RHS call.
struct alignas(64) NbrRec {
  double surfh;   // ws.surfh
  double head;    // topo.zmax + ws.surfh
  double gw;      // clamped ws.gw
  double gwz;     // ws.gw + topo.zmin
  double effkh;   // EffKh(soil, gw)
  double sfmag;   // |grad h| from FrictSlope
  double rough;   // lc.rough (static)
  double zmax;    // topo.zmax (static)
};


struct ElemDev {
  double *area, *zmin, *zmax, *depth, *ksath, *ksatv, *kinfv, *dinf, *alpha,
      *beta, *porosity, *dmac, *kmach, *kmacv, *areafv, *areafh, *rough, *rzd;
  double *edge, *nabrdist, *nabr_x, *nabr_y;  // [3*ne]
  int32_t *nb;                                // [3*ne]
  // per-step inputs
  double *pcpdrp, *edir, *ett, *ws0s, *ebc;   // ebc [3*ne]
  NbrRec *rec;
  // Noah coupling (single solves, pihm_ctx_set_noah): static smcmin, per-step
  // top-layer liquid water sh2o0, frozen-soil factor fcr, and the saturated
  // zone's share of transpiration gwet (GwTransp of et[nlayer][ne], nwtbl, nroot)
  double *smcmin, *sh2o0, *fcr, *gwet, *et;
  int32_t *nwtbl, *nroot;
};

struct RiverDev {
  int32_t *down;        // >= 0 device index of the downstream segment, else outlet code
  int32_t *left, *right;  // device element index of the banks
  int32_t *bctype, *ord;
  int32_t *up_beg, *up_end, *up_idx;  // upstream segments of r: up_idx[up_beg[r] .. up_end[r]), ascending caller index
  double *area, *zmin, *zmax, *zbed, *node_zmax, *dist_l, *dist_r, *depth, *coeff,
      *length, *width, *rough, *cwr, *ksath, *ksatv, *bedthick, *porosity;
  double *bc;           // per-step river bc value
  double *rf;           // [11*nr], rf[k*nr + r] (0 and 10 in recording mode)
  double *upinit;       // 0.0 + boundary flux of the upstream node
  double *hid_ovl;      // [4*nr]: carried river-edge overland flux at slot 2r + side,
                        // this call's at 2*nr + slot
  double *bank_sub;     // [2*nr] this call's river-edge subsurface flux, by slot
};

// Element tiles of the one-launch RHS of a single solve (k_rhs_tile): tile t is
// the kTile consecutive device elements from t * kTile; its halo lists the
// outside neighbours of the tile's elements (first ring), then their outside
// neighbours not listed yet, each part in ascending device index.  In LDS,
// slot s < kTile is tile element s and the halo follows in list order.
struct TileArgs {
  const int32_t *hbeg;   // [ntile + 1] each tile's range of halo entries
  const int32_t *hn1;    // [ntile] first-ring entries at the front of the range
  const int32_t *hidx;   // device element of each halo entry
  const int32_t *hnb;    // [3 * entries] neighbour codes of the entries, elements as LDS slots
                         // (first ring only; the second ring's are not read)
  const int32_t *lnb;    // [3 * ne] (j * ne + p) the same for the tile elements
  int s1cap, scap;       // LDS records (tile + first ring) and heads (tile + halo) per block
};

struct RecordDev {      // optional flux recording (device order)
  double *ovl, *sub;    // [3*ne]
  double *infil, *rechg;
};

// One solve of a batch: its ranges in the concatenated element, river and
// state arrays.  A single solve is a batch of one with all offsets zero.
struct MemberDev {
  long yoff;            // first state entry (member-major: surf, unsat, gw, stage, river gw)
  int eoff, ne;         // device element range
  int roff, nr;         // river range
  double dt;            // ctrl.stepsize of this solve (Infil)
  int active;           // evaluated by the next RHS launch
  int pad;
};

// Per-call state and derivative of one batch member.
struct MemberIO {
  const double *y;
  double *ydot;
  int active;
  int pad;
};

struct RhsArgs {
  int ne, nr;
  int surf_mode, riv_mode;
  double dt;            // ctrl.stepsize
  int record;
  ElemDev e;
  RiverDev r;
  RecordDev rec;
  const double *y;
  double *ydot;
  int *nanflag;
  int nmem;
  const MemberDev *mem;        // [nmem]
  const int32_t *emem, *rmem;  // member of each device element / river
  const MemberIO *mio;         // [nmem] per-call member table, or nullptr
  int *mnan;                   // [nmem] per-member NaN flags, or nullptr
  // Water-balance accumulators (single solves): per-element precipitation,
  // evapotranspiration and boundary-edge outflow volume rates cl_e[3][ne],
  // per-river external outflow cl_r[nr]; summed into the 4 state entries
  // that follow the reference's state when closure is on.
  int closure;
  double *cl_e, *cl_r, *cl_part;
  // Compact launch over the active members only (nullptr: every position of
  // the batch): bm[0..nact] element-block and bm[nact+1..2nact+1] river-block
  // prefixes, bm_mem[i] the i-th active member.
  const int *bm, *bm_mem;
  int nact;
  // Element-tile placement over the XCDs (xcd_tile): 0 places one contiguous
  // range of tiles on each XCD; s >= 1 deals groups of s consecutive tiles to
  // the XCDs in turn.  The arithmetic does not depend on it.
  int tile_span;
};

}  // namespace pihm

namespace pihm {
// Capacity of one slot of a context: the largest mesh it can hold.
struct SlotCap {
  int ne, nr;
};
}  // namespace pihm

// One slot: device ranges at its capacity (eoff, roff, yoff), the loaded
// mesh's actual sizes (ne, nr, neq), and its offsets in caller order, where
// the loaded slots are concatenated at their actual sizes (ceoff, croff, cyoff).
struct pihm_member_host {
  long yoff;
  int eoff, ne, roff, nr, neq;
  int ceoff, croff;
  long cyoff;
  double dt;
};

// Launch one batched RHS evaluation whose member states come from the device
// table mio (inactive members untouched); NaN in a member's derivative sets
// mnan[member].  Asynchronous on the context's stream.
struct pihm_ctx;
int pihm_rhs_launch_io(pihm_ctx *c, const pihm::MemberIO *mio, int *mnan);
// The same over the nact members of a block map (see RhsArgs::bm); nbe and
// nbr are the map's element and river block totals.
// Saves (dir 0) or restores (dir 1) the carried river-edge flux of every slot
// (hid_ovl[0, 2 nr)) to or from the device buffer buf, on the context's stream.
int pihm_ctx_carried_copy(pihm_ctx *c, double *buf, int dir);
// One RHS evaluation of a single-solve context at y (device) with the fluxes
// recorded in args.rec / args.r.rf; the carried river-edge flux is restored.
int pihm_ctx_eval_fluxes(pihm_ctx *c, const double *y, double *ydot, double *carried_save);
// With the _NOAH_ branches on: the saturated zone's share of transpiration
// again from the current ett, et, nwtbl and nroot (after any of them changed).
int pihm_ctx_noah_refresh(pihm_ctx *c);
int pihm_rhs_launch_compact(pihm_ctx *c, const pihm::MemberIO *mio, int *mnan, const int *bm,
                            const int *bm_mem, int nact, int nbe, int nbr);

struct pihm_ctx {
  int ne, nr, neq;              // totals over the batch, at slot capacity
  int nmem;
  std::vector<pihm::SlotCap> caps;
  std::vector<int> uoff;        // first upstream entry of each slot
  int ne_loaded, nr_loaded;     // caller-order totals of the loaded meshes
  long neq_loaded;
  std::vector<pihm_member_host> members;
  std::vector<int32_t> emem_host;   // member of each device element
  std::vector<pihm::MemberDev> mem_host;
  pihm::MemberDev *mem_dev;
  int surf_mode, riv_mode;
  double dt;
  int device;
  hipStream_t stream;
  int nan_check;
  int record;
  int closure;                  // 4 water-balance accumulators after the state
  long long ncalls;
  std::vector<int32_t> perm;     // device position -> caller index within its member
  std::vector<int32_t> iperm;    // (member eoff + caller index) -> device position
  std::vector<int32_t> nb_host;  // device-order neighbour codes [3*ne]
  pihm::RhsArgs args;            // pointers filled at creation
  void *arena;                   // one device allocation for everything
  size_t arena_bytes;
  int *nanflag_dev;
  int *nanflag_host;             // pinned
  double *stage_host;            // scratch for host-side permutations
  int noah;                      // RHS with the reference's _NOAH_ branches
  void *noah_arena;              // the Noah fields of ElemDev
  int rhs_path;                  // PIHM_RHS_ONE_LAUNCH or PIHM_RHS_TWO_KERNELS (single solves)
  int one_launch_ok;             // tile tables built (single solve, halos fit the LDS budget)
  pihm::TileArgs tiles;
  size_t tile_lds;               // dynamic LDS bytes per element block
  void *tile_arena;
};

// Caller-order index of device element p / device river r (the loaded slots
// concatenated at their actual sizes), or -1 for unused slot capacity.
static inline long caller_elem(const pihm_ctx *c, int p) {
  if (c->perm[p] < 0) return -1;
  return (long)c->members[c->emem_host[p]].ceoff + c->perm[p];
}
static inline long caller_river(const pihm_ctx *c, int r) {
  for (int k = 0; k < c->nmem; k++) {
    const auto &M = c->members[k];
    if (r >= M.roff && r < M.roff + c->caps[k].nr) return (r - M.roff < M.nr) ? M.croff + (r - M.roff) : -1;
  }
  return -1;
}

struct pihm_noah_soil;
namespace pihm {
// The soil column a soil-water step holds (noahsoil.hip), shared with the Noah
// land-surface step: device order, layer arrays layer-major [k * ne + p].
struct NoahSoilView {
  pihm_ctx *ctx;
  int ne;
  int32_t *nsoil;
  double *sldpth, *smcmax, *smc, *sh2o;
};
NoahSoilView noah_soil_view(pihm_noah_soil *s);
}  // namespace pihm
