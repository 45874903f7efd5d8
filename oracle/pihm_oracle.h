/*
 * pihm_oracle.h - CPU restatement of the MM-PIHM hydrology RHS.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this code, and only as the checker
 * or as the timed CPU baseline.  The product library never links it.
 *
 * Every function restates the reference's serial/OpenMP C code in the
 * reference's operation order (file:line given at each definition), so that
 * the restatement is bit-identical to that code on the same host.
 */
#ifndef PIHM_ORACLE_H
#define PIHM_ORACLE_H

#include "../include/pihm_rhs.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Mutable per-call state of the reference structures that the RHS reads or
 * writes (elem.ws, elem.wf, river.ws, river.wf), in original order. */
typedef struct oracle_state {
  int nelem, nriver;
  /* per-step inputs */
  double *pcpdrp, *edir, *ett, *ws0_surf, *elem_bc, *river_bc;
  /* element states and fluxes */
  double *surf, *unsat, *gw, *surfh;
  double *edir_surf, *edir_unsat, *edir_gw, *ett_unsat, *ett_gw;
  double *infil, *rechg;
  double *ovlflow, *subsurf;                /* [3*nelem] */
  double *dhbydx, *dhbydy;
  /* river states and fluxes */
  double *stage, *rgw;
  double *rivflow;                          /* [11*nriver] */
  /* _NOAH_ build (oracle_set_noah): soil.smcmin, ws.sh2o[0], ps.fcr,
   * wf.et[MAXLYR] per element (element-major, nlayer used), ps.nwtbl, ps.nroot,
   * and the ps.gwet EtExtract leaves */
  int noah, nlayer;
  double *smcmin, *sh2o0, *fcr, *et, *gwet;
  int *nwtbl, *nroot;
} oracle_state;

oracle_state *oracle_state_new(int nelem, int nriver);
void oracle_state_free(oracle_state *s);
void oracle_set_step_inputs(oracle_state *s, const pihm_step_inputs *in);
/* Switch the _NOAH_ branches on (smcmin non-NULL; copies every array given,
 * et as et[i*nlayer + j], keeping the previous values and layer count of
 * those that are NULL) or off (smcmin NULL). */
void oracle_set_noah(oracle_state *s, const double *smcmin, const double *sh2o0, const double *fcr,
                     const double *et, int nlayer, const int *nwtbl, const int *nroot);
double oracle_GwTransp(double ett, const double *et, int nwtbl, int nroot);

/* The soil-water step of the _NOAH_ build after each model step
 * (noahsoil_oracle.c): one element's column. */
typedef struct oracle_noah_col {
  int nsoil;
  const double *sldpth;          /* [nsoil] layer thicknesses (m) */
  double smcmax, smcmin, ksatv, alpha, beta, frzx;
  double gw, infil, runoff2, edir;   /* runoff2: MassBalance's subsurface runoff (m s-1) */
  const double *et;              /* [nsoil] layer transpiration (m s-1) */
} oracle_noah_col;
typedef struct oracle_noah_state {
  double smc[11], sh2o[11];      /* carried */
  double fcr;                    /* carried: kept when every layer is saturated */
  double smflxv[11], runoff3;    /* carried likewise */
  int nwtbl;
  double satdpth[11], runoff2_lyr[11];
} oracle_noah_state;
void oracle_noah_soil_column(const oracle_noah_col *in, oracle_noah_state *st, double dt);
int oracle_FindWaterTable(const double *sldpth, int nsoil, double gw, double *satdpth);
void oracle_CalcLatFlx(int nsoil, const double *satdpth, double runoff2, double *runoff2_lyr);
void oracle_WDfCnd(double *wdf, double *wcnd, double smc, double sicemax, double smcmax, double smcmin,
                   double ksatv, double alpha, double beta);
void oracle_Rosr12(double *p, const double *a, const double *b, double *c, const double *d, double *delta,
                   int nsoil);
/* Every element: per-layer arrays element-major [i*11 + k]; smc, sh2o, fcr,
 * smflxv and runoff3 are read and written. */
void oracle_noah_soil_step(int ne, double dt, const int *nsoil, const double *sldpth, const double *smcmax,
                           const double *smcmin, const double *ksatv, const double *alpha, const double *beta,
                           const double *frzx, const double *gw, const double *infil, const double *runoff2,
                           const double *edir, const double *et, double *smc, double *sh2o, double *fcr,
                           int *nwtbl, double *satdpth, double *runoff2_lyr, double *smflxv, double *runoff3);

/* ODE (src/ode.c:4-399).  Returns 0, or 1 at the first NaN component (where
 * the reference prints and exits, src/ode.c:401-410).  When `mag` is not
 * NULL it receives, per component, the sum of magnitudes of the terms
 * assembled into that component (scale for the parity tolerance). */
int oracle_ode(const pihm_mesh_desc *m, oracle_state *s, double t,
               const double *y, double *dy, double *mag);

/* Individual physics functions, exported for unit tests. */
double oracle_SurfH(double surfeqv);
double oracle_KrFunc(double beta, double satn);
double oracle_Psi(double satn, double alpha, double beta);
double oracle_EffKh(double depth, double dmac, double kmach, double areafv,
                    double ksath, double gw);
double oracle_AvgKv(double dmac, double ksatv, double kmacv, double areafh,
                    double deficit, double gw, double satkfunc);
double oracle_EffKinf(double kinfv, double kmacv, double areafh,
                      double dh_by_dz, double ksatfunc, double elemsatn,
                      double applrate, double surfh);
double oracle_DhByDl(const double *l1, const double *l2, const double *surfh);
double oracle_OverLandFlow(double avg_h, double grad_h, double avg_sf,
                           double crossa, double avg_rough);
double oracle_RiverCroSectArea(int order, double depth, double coeff);
double oracle_RiverPerim(int order, double depth, double coeff);
double oracle_RiverEqWid(int order, double depth, double coeff);
double oracle_OvlFlowElemToRiver_raw(double rzmax, double ezmax, double surfh,
                                     double zbed, double stage, double cwr,
                                     double length);

/* Geometry initialisation (src/init_topo.c:6-47, src/initialize.c:385-458,
 * src/init_river.c:6-124) from nodes/triangles/rivers.  Inputs use the
 * reference's 1-based file indices.  Outputs are the mesh-description
 * arrays (caller allocates).  nabr is rewritten in place for river edges. */
void oracle_init_topo(int nelem, const int32_t *node /*3n*/, const double *nx,
                      const double *ny, const double *nzmin,
                      const double *nzmax, double *area, double *x, double *y,
                      double *zmin, double *zmax, double *edge /*3n*/);
void oracle_init_river_tags(int nelem, int nriver, int32_t *nabr /*3n*/,
                            const int32_t *leftele, const int32_t *rightele);
void oracle_init_surfl(int nelem, const int32_t *node, const int32_t *nabr,
                       const double *nx, const double *ny, const double *ex,
                       const double *ey, const double *area,
                       const double *edge, const double *rx, const double *ry,
                       double *nabr_x, double *nabr_y, double *nabrdist);
void oracle_init_river_topo(int nriver, const int32_t *fromnode,
                            const int32_t *tonode, const int32_t *leftele,
                            const int32_t *rightele, const double *nx,
                            const double *ny, const double *nzmax,
                            const double *ezmax, const double *ezmin,
                            const double *ex, const double *ey,
                            double *rx, double *ry, double *rzmax,
                            double *rzmin, double *node_zmax, double *dist_l,
                            double *dist_r, double *length);

#ifdef __cplusplus
}
#endif
#endif
