/*
 * pihm_rhs.h - C ABI of the MI355X MM-PIHM hydrology right-hand side.
 *
 * The entry point PIHM_ODE_Hip has the CVODE CVRhsFn signature
 * (reference: cvode/include/cvode/cvode.h:136-150) and replaces the
 * reference's `int ODE(realtype t, N_Vector CV_Y, N_Vector CV_Ydot,
 * void *pihm_data)` (src/ode.c:4-399), registered by
 * `CVodeInit(cvode_mem, ODE, 0.0, CV_Y)` (src/ode.c:436).  `user_data` is a
 * pihm_ctx made by pihm_ctx_create instead of the reference's pihm_struct
 * (src/ode.c:473).
 *
 * Index conventions of the mesh description follow the reference
 * structures after Initialize() (src/initialize.c:16):
 *   nabr[3*i+j]   > 0 : 1-based neighbour element across edge j
 *                 < 0 : -(river segment index, 1-based) (src/init_river.c:24-35)
 *                 = 0 : domain boundary, condition bc_type[3*i+j]
 *   river down    > 0 : 1-based downstream segment; <= 0 outlet condition
 *                       DIRICHLET -1, NEUMANN -2, ZERO_DPTH_GRAD -3,
 *                       CRIT_DPTH -4 (src/include/pihm_const.h:140-143)
 *   leftele/rightele : 1-based bank elements.
 * Arrays are in the caller's (original) order.  The state vector is
 * block-major: SURF(i)=i, UNSAT(i)=i+ne, GW(i)=i+2ne, RIVSTG(r)=r+3ne,
 * RIVGW(r)=r+3ne+nr (src/include/pihm_func.h:7-11).
 *
 * Device vectors given to PIHM_ODE_Hip hold the state in *device order*:
 * the same blocks, with elements inside each element block renumbered along
 * a Hilbert curve of the element centroids.  Use pihm_ctx_state_to_device /
 * pihm_ctx_state_from_device to move a state between the caller's order and
 * device order; those two maps are exact permutations.
 */
#ifndef PIHM_RHS_H
#define PIHM_RHS_H

#include <stdint.h>
#include "pihm_nvector.h"

#ifdef __cplusplus
extern "C" {
#endif

#define PIHM_KINEMATIC 1
#define PIHM_DIFF_WAVE 2

/* Return codes of the library (never exits the process). */
#define PIHM_OK              0
#define PIHM_ERR_ARG        -1   /* bad argument / inconsistent mesh */
#define PIHM_ERR_HIP        -2   /* HIP runtime error */
#define PIHM_ERR_NAN         1   /* NaN in ydot: recoverable per CVRhsFn */
#define PIHM_ERR_TOO_LARGE  -3   /* does not fit the limits */

/* Element and river description after the reference's Initialize().
 * Field meanings are those of elem_struct / river_struct
 * (src/include/elem_struct.h:1557-1618, src/include/river_struct.h:137-167). */
typedef struct pihm_mesh_desc {
  int32_t nelem;
  int32_t nriver;
  int32_t surf_mode;            /* ctrl.surf_mode: KINEMATIC or DIFF_WAVE */
  int32_t riv_mode;             /* ctrl.riv_mode */
  double  stepsize;             /* ctrl.stepsize [s], the dt of Infil */

  /* elements: per-element arrays [nelem], per-edge arrays [3*nelem] */
  const int32_t *nabr;
  const int32_t *bc_type;       /* attrib.bc_type */
  const double *x, *y;          /* topo.x, topo.y (centroid) */
  const double *area, *zmin, *zmax;
  const double *edge, *nabrdist, *nabr_x, *nabr_y;   /* [3*nelem] */
  const double *depth, *ksath, *ksatv, *kinfv, *dinf, *alpha, *beta,
               *porosity, *dmac, *kmach, *kmacv, *areafv, *areafh;
  const double *rough;          /* lc.rough */
  const double *rzd;            /* ps.rzd */

  /* river segments [nriver] */
  const int32_t *down, *leftele, *rightele;
  const int32_t *riverbc_type;  /* attrib.riverbc_type */
  const int32_t *intrpl_ord;    /* shp.intrpl_ord */
  const double *rarea, *rzmin, *rzmax, *zbed, *node_zmax, *dist_left,
               *dist_right;     /* river topo */
  const double *rdepth, *coeff, *length, *width;     /* shp */
  const double *rrough, *cwr, *rksath, *rksatv, *bedthick, *rporosity; /* matl */
} pihm_mesh_desc;

/* Per-step inputs read by the RHS (src/is_sm_et.c:127-224,
 * src/forcing.c:127-175,214-244, src/update.c:46), caller's order.
 * Any pointer may be NULL to leave that input unchanged. */
typedef struct pihm_step_inputs {
  const double *pcpdrp;         /* wf.pcpdrp [nelem], m/s */
  const double *edir;           /* wf.edir   [nelem], m/s */
  const double *ett;            /* wf.ett    [nelem], m/s */
  const double *ws0_surf;       /* ws0.surf  [nelem], m */
  const double *elem_bc;        /* bc.head / bc.flux union [3*nelem] */
  const double *river_bc;       /* river bc.head / bc.flux union [nriver] */
} pihm_step_inputs;

typedef struct pihm_ctx pihm_ctx;

/* Build a context on the current HIP device: validates the mesh, computes
 * the Hilbert renumbering and uploads the structure-of-arrays mesh.
 * Per-step inputs start at zero and the carried river-edge overland fluxes
 * (read by Infil one call later) start at zero, as in src/initialize.c:735-743. */
int  pihm_ctx_create(const pihm_mesh_desc *mesh, pihm_ctx **out);
void pihm_ctx_destroy(pihm_ctx *ctx);

/* Batch of independent solves (ensemble members, serving slots) in one
 * context: no counterpart in the reference, which runs one watershed per
 * process.  Members are concatenated; member k's state is the contiguous
 * range [yoff_k, yoff_k + neq_k) of a vector of length pihm_ctx_neq, laid
 * out as a single solve's state.  Host arrays of the other calls are the
 * members' caller-order arrays concatenated in member order.  One RHS launch
 * evaluates every active member; members share no data, so each member's
 * ydot is the same bits as in a context of its own. */
int  pihm_ctx_create_batch(int nmem, const pihm_mesh_desc *const *meshes, pihm_ctx **out);
int  pihm_ctx_nmembers(const pihm_ctx *ctx);
int  pihm_ctx_member_info(const pihm_ctx *ctx, int member, long *yoff, int *neq, int *nelem,
                          int *nriver);
int  pihm_ctx_set_active(pihm_ctx *ctx, const int *active);   /* [nmem], stream-ordered */

/* A context of nslots empty slots, slot k able to hold a mesh of up to
 * ne_cap[k] elements and nr_cap[k] river segments (device ranges are laid out
 * at these capacities; member_info's yoff/neq give a slot's state range and
 * the loaded mesh's length).  pihm_ctx_load_member (re)loads slot k with a
 * mesh, renumbering it and resetting the slot's carried river-edge fluxes and
 * per-step inputs; it is how a serving loop reuses memory without
 * reallocating.  Host-side arrays of the batch functions (step inputs, state
 * permutations, flux read-back) list the loaded meshes in slot order at their
 * actual sizes.  The reference has no counterpart (it runs one watershed). */
int  pihm_ctx_create_slots(int nslots, const int32_t *ne_cap, const int32_t *nr_cap, int surf_mode,
                           int riv_mode, pihm_ctx **out);
int  pihm_ctx_load_member(pihm_ctx *ctx, int slot, const pihm_mesh_desc *mesh);
int  pihm_ctx_set_stream(pihm_ctx *ctx, void *hip_stream);
/* A HIP stream of default priority for callers without a HIP toolchain
 * (graph replay needs a stream other than the null stream). */
int  pihm_stream_create(void **hip_stream);
int  pihm_stream_destroy(void *hip_stream);
int  pihm_ctx_set_step_inputs(pihm_ctx *ctx, const pihm_step_inputs *in);
/* The step inputs the RHS will read, in caller order (layouts as in
 * pihm_step_inputs); any pointer may be NULL.  Synchronises the stream.  Shows
 * what the device producers (land-surface step, boundary series) wrote. */
int  pihm_ctx_get_step_inputs(pihm_ctx *ctx, double *pcpdrp, double *edir, double *ett,
                              double *ws0_surf, double *elem_bc, double *river_bc);
/* The RHS of the reference built with _NOAH_ (Flux-PIHM), for single solves:
 *  - EtExtract (src/hydrol.c:43-56,79-84): no direct evaporation from ponded
 *    water, and transpiration split by GwTransp (src/noah/lsm_func.c:163-185);
 *  - Infil (src/vert_flow.c:94-98): the unsaturated branch takes the top
 *    layer's saturation (sh2o[0] - smcmin) / (smcmax - smcmin), where
 *    smcmax - smcmin is the mesh's porosity (src/init_soil.c:34-36);
 *  - VerticalFlow (src/vert_flow.c:16-19): infiltration times the frozen-soil
 *    factor fcr, which Recharge then reads.
 * pihm_ctx_set_noah switches it on with each element's smcmin (caller order;
 * NULL switches it off).  Until the first pihm_ctx_set_noah_step, fcr is 1,
 * sh2o[0] is 0 and no transpiration comes from the saturated zone.
 * pihm_ctx_set_noah_step hands over what the Noah step (src/noah/noah.c) left
 * for the next model steps: sh2o0[ne] (top layer liquid water, m3 m-3),
 * fcr[ne], and optionally et[ne*nlayer] (per-layer transpiration, element-
 * major as the reference's wf.et[MAXLYR]), nwtbl[ne] and nroot[ne] (1-based
 * water-table and root layers, nroot <= nlayer).  The saturated zone's share
 * of transpiration is evaluated on the device from these and the ett already
 * set for the step, so the call follows pihm_ctx_set_step_inputs.  NULL
 * pointers leave those values as they are.  Switch the branches on or off
 * before an integration starts: Krylov iterations captured as HIP graphs
 * (PihmCVodeSetGraphRhs) keep the kernel they were captured with. */
#define PIHM_NOAH_MAXLYR 11
int  pihm_ctx_set_noah(pihm_ctx *ctx, const double *smcmin);
int  pihm_ctx_set_noah_step(pihm_ctx *ctx, const double *sh2o0, const double *fcr, const double *et,
                            int nlayer, const int32_t *nwtbl, const int32_t *nroot);
int  pihm_ctx_neq(const pihm_ctx *ctx);

/* Exact permutations between caller order and device order. `perm_out`
 * receives, for each device element position p, the caller's element index. */
int  pihm_ctx_get_perm(const pihm_ctx *ctx, int32_t *perm_out);
int  pihm_ctx_state_to_device(pihm_ctx *ctx, const double *host_orig, N_Vector dev);
int  pihm_ctx_state_from_device(pihm_ctx *ctx, N_Vector dev, double *host_orig);

/* The RHS (CVRhsFn).  Returns 0, or PIHM_ERR_NAN (> 0, recoverable) when a
 * component of ydot is NaN, where the reference prints and exits
 * (src/ode.c:401-410), or < 0 on a runtime error.  The NaN check costs one
 * host synchronisation; pihm_ctx_set_nan_check(ctx, 0) turns it off. */
int  PIHM_ODE_Hip(realtype t, N_Vector y, N_Vector ydot, void *user_data);
int  pihm_ctx_set_nan_check(pihm_ctx *ctx, int on);

/* Asynchronous form: launch one RHS evaluation on raw device arrays (device
 * order, length neq) and return without synchronising.  A NaN raises a
 * sticky flag that pihm_ctx_poll_nan reads and clears (PIHM_ERR_NAN). */
int  pihm_rhs_launch(pihm_ctx *ctx, const double *y_dev, double *ydot_dev);
int  pihm_ctx_poll_nan(pihm_ctx *ctx);

/* Recording mode keeps every flux of the last call for pihm_ctx_get_fluxes
 * (extra HBM writes; off by default). */
int  pihm_ctx_set_record(pihm_ctx *ctx, int on);
/* Water-balance closure (single-solve contexts, set before creating
 * vectors): the state gains 4 entries after the reference's 3*nelem+2*nriver,
 * the cumulative volumes [m3] of precipitation in, evapotranspiration out,
 * outflow across the mesh boundary edges, and outflow through river
 * boundaries and outlets.  The RHS writes their rates, so the integrator
 * integrates them with the same steps as the state and
 *   sum_i area_i (surf_i + porosity_i (unsat_i + gw_i))
 *   + sum_r area_r (stage_r + porosity_r rivgw_r) - Q_in + Q_et + Q_bnd + Q_riv
 * is conserved up to rounding.  pihm_ctx_neq includes the 4 entries. */
int  pihm_ctx_set_closure(pihm_ctx *ctx, int on);

/* Wait for all work queued on the context's stream. */
int  pihm_ctx_sync(pihm_ctx *ctx);

/* How a single solve evaluates the RHS.  PIHM_RHS_ONE_LAUNCH (the default for
 * single-solve contexts): one kernel, each block staging the records of a
 * tile of elements and its halo in LDS.  PIHM_RHS_TWO_KERNELS: the
 * neighbour-record kernel, then the flux kernel (batches always use it).  Both
 * give the same bits.  pihm_ctx_set_rhs_path returns PIHM_ERR_ARG for a batch
 * context, or when the one-launch tables could not be built (a halo larger
 * than the LDS budget).  No counterpart in the reference. */
#define PIHM_RHS_ONE_LAUNCH 0
#define PIHM_RHS_TWO_KERNELS 1
int  pihm_ctx_set_rhs_path(pihm_ctx *ctx, int path);
int  pihm_ctx_get_rhs_path(const pihm_ctx *ctx);

/* Measurement: average device time [ms] of each of the two kernels of the
 * two-kernel RHS (neighbour records + slopes; all fluxes and the derivative)
 * over `iters` evaluations at state y_dev, each between its own pair of HIP
 * events on the context's stream, whichever path the context uses. */
int  pihm_ctx_time_kernels(pihm_ctx *ctx, const double *y_dev, double *ydot_dev, int iters,
                           double *ms_prep, double *ms_flux);

/* Measurement: average device time [ms] of one whole RHS evaluation on the
 * context's path (its launches issued back to back, as an integrator issues
 * them) over `iters` evaluations at state y_dev: one HIP event before the
 * first launch and one after the last, on the context's stream. */
int  pihm_ctx_time_rhs(pihm_ctx *ctx, const double *y_dev, double *ydot_dev, int iters, double *ms_per_eval);

/* Placement of the element tiles over the XCDs: 0 gives each XCD one
 * contiguous range of Hilbert-ordered tiles; s >= 1 deals groups of s
 * consecutive tiles to the XCDs in turn.  Results do not depend on it.
 * pihm_ctx_autotune (single-solve contexts) times reps back-to-back RHS
 * evaluations at y_dev for each placement of PIHM_TUNE_SPANS on each RHS path
 * the context has, three interleaved rounds, each candidate's best round
 * counting.  It keeps the default (the one-launch path where available,
 * placement 0) unless another candidate is more than 2% faster, sets the
 * chosen path and placement on the context and returns the placement (and,
 * when ms is not NULL, each placement's best ms per evaluation over the
 * paths).  The carried river-edge flux is restored afterwards, so the next
 * call reads what it would have read.  No counterpart in the reference. */
#define PIHM_TUNE_NSPAN 8
#define PIHM_TUNE_SPANS {0, 1, 2, 4, 8, 16, 32, 64}
int  pihm_ctx_set_tile_span(pihm_ctx *ctx, int span);
int  pihm_ctx_get_tile_span(const pihm_ctx *ctx);
int  pihm_ctx_autotune(pihm_ctx *ctx, const double *y_dev, double *ydot_dev, int reps, int *best_span,
                       double *ms);

/* Carried state: the overland flux on river edges written at the end of one
 * RHS call and read by Infil in the next (src/river_flow.c:159-182 then
 * src/vert_flow.c:59-62).  Arrays are [3*nelem] in caller order; entries of
 * non-river edges are written as 0 and ignored on set. */
int  pihm_ctx_get_river_edge_ovlflow(pihm_ctx *ctx, double *ovl3);
int  pihm_ctx_set_river_edge_ovlflow(pihm_ctx *ctx, const double *ovl3);

/* Flux read-back of the last RHS call, caller order (for Summary, water
 * balance and outputs):  ovlflow/subsurf [3*nelem], infil/rechg [nelem],
 * rivflow [11*nriver] (index r*11+k, k as src/include/pihm_const.h:127-137).
 * Any pointer may be NULL. */
int  pihm_ctx_get_fluxes(pihm_ctx *ctx, double *ovlflow, double *subsurf,
                         double *infil, double *rechg, double *rivflow);

/* The power routine of the flux functions (x^y as exp(y ln x), within a few
 * ulp of the host libm pow; exact for 1^y and (+0)^y), evaluated on the host
 * (on_device = 0) or by a device kernel (on_device = 1) over n host arrays.
 * Diagnostic: no counterpart in the reference, which calls libm pow
 * (src/soil.c:8-13, src/vert_flow.c:272-278, src/lat_flow.c:279-283,
 * src/river_flow.c:520-583); it lets tests bound the device results against
 * the host libm and check that host and device return the same bits. */
int  pihm_math_pow(int n, const double *x, const double *y, double *out, int on_device);

#ifdef __cplusplus
}
#endif
#endif /* PIHM_RHS_H */
