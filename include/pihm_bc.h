/*
 * pihm_bc.h - boundary-condition time series applied on the device.
 *
 * Every model step the reference interpolates each boundary-condition series
 * to the current time and copies the values into the elements' and river
 * segments' bc fields (ApplyBc, src/forcing.c:8-37; ApplyElemBc,
 * src/forcing.c:127-177; ApplyRiverBc, src/forcing.c:214-244; IntrplForc,
 * src/forcing.c:482-530).  An element edge with attrib.bc_type[j] = k > 0
 * takes head bc.head[j] from series k; k < 0 takes flux bc.flux[j] from series
 * -k; 0 keeps its value.  River segments do the same with riverbc_type and
 * the river series.  The series are the .bc file's BC_TS blocks
 * (src/read_bc.c) and the .riv file's RIV_TS blocks (src/read_river.c:118-175).
 *
 * pihm_bc_apply does this on the device for a single-solve context and writes
 * the values into the context's step inputs (the elem_bc / river_bc of
 * pihm_step_inputs): no upload per step.
 */
#ifndef PIHM_BC_H
#define PIHM_BC_H

#include <stdint.h>
#include "pihm_rhs.h"

#ifdef __cplusplus
extern "C" {
#endif

/* nts series of nrec records each, at ascending times ftime [s]. */
typedef struct pihm_ts {
  int32_t nts, nrec;
  const double *ftime;   /* [nts*nrec] */
  const double *value;   /* [nts*nrec] */
} pihm_ts;

typedef struct pihm_bc pihm_bc;

/* elem_bc_type: [3*nelem] caller order, element-major (the .att file's
 * BC0..BC2 columns, the mesh's bc_type); river_bc_type: [nriver] (the .riv
 * file's BC column).  Either series set may be NULL when no code refers to it.
 * PIHM_ERR_ARG for a batched context, a code naming a missing series, or
 * times that do not increase. */
int  pihm_bc_create(pihm_ctx *ctx, const int32_t *elem_bc_type, const pihm_ts *elem_ts,
                    const int32_t *river_bc_type, const pihm_ts *river_ts, pihm_bc **out);
void pihm_bc_destroy(pihm_bc *bc);

/* Interpolate every series at t and copy into the context's boundary values;
 * asynchronous on the context's stream.  PIHM_ERR_ARG, where the reference
 * exits, when t lies outside a series (src/forcing.c:487-497). */
int  pihm_bc_apply(pihm_bc *bc, double t);

/* Measurement: iters back-to-back applies at t on the context's stream, timed
 * with HIP events; *ms_per_apply = elapsed / iters.  No counterpart in the
 * reference. */
int  pihm_bc_time(pihm_bc *bc, double t, int iters, double *ms_per_apply);

#ifdef __cplusplus
}
#endif
#endif /* PIHM_BC_H */
