"""Water balance of a run with the closure accumulators (pihm_ctx_set_closure).

Storage S(y) = sum_i area_i (surf_i + porosity_i (unsat_i + gw_i))
             + sum_r area_r (stage_r + porosity_r rivgw_r)          [m3]
Net inflow   = Q_in - Q_et - Q_bnd - Q_riv                          [m3]
The closure residual of a run is (S(y1) - S(y0)) - (net inflow over the run), reported
relative to the precipitation volume of the run.
"""
import math

import numpy as np


def storage(mesh, y):
    a = mesh.a
    ne, nr = mesh.nelem, mesh.nriver
    ys, yu, yg = y[:ne], y[ne:2 * ne], y[2 * ne:3 * ne]
    st, rg = y[3 * ne:3 * ne + nr], y[3 * ne + nr:3 * ne + 2 * nr]
    elem = a["area"] * (ys + a["porosity"] * (yu + yg))
    riv = a["rarea"] * (st + a["rporosity"] * rg)
    return math.fsum(elem.tolist()) + math.fsum(riv.tolist())


def closure(mesh, y0, y1):
    """(residual [m3], precipitation volume [m3], relative residual, terms)."""
    n = mesh.neq
    q0 = np.zeros(4) if len(y0) == n else np.asarray(y0[n:n + 4])
    q1 = np.asarray(y1[n:n + 4])
    dq = q1 - q0
    ds = storage(mesh, y1) - storage(mesh, y0)
    net = dq[0] - dq[1] - dq[2] - dq[3]
    res = ds - net
    return dict(residual_m3=res, precip_m3=dq[0], rel=abs(res) / max(abs(dq[0]), 1e-300),
                storage_change_m3=ds, et_m3=dq[1], boundary_out_m3=dq[2], river_out_m3=dq[3])
