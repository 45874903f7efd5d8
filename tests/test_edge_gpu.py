"""Edge cases of the device path on MI355X: watersheds without a river network, and tiny
meshes (12 and 40 triangles, one workgroup, most elements on the boundary), against the
oracle per call (1e-12 of the terms) and, for a run, within the solver tolerance."""
import math

import numpy as np
import pytest

import oracle as O
from conftest import assert_rhs_close
from pihm_amd.forcing import StormForcing, run_config
from pihm_amd.meshgen import make_watershed
from pihm_amd.rhs import HipRHS, run_gpu
from test_oracle_pins import _inputs, _states

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n_tri,n_seg", [(535, 0), (40, 0), (12, 0), (12, 1), (40, 3)])
def test_rhs_parity_small_and_river_free(n_tri, n_seg):
    m = make_watershed(n_tri, seed=3, n_seg=n_seg, network="valley")
    print(f"[edge] {n_tri} triangles, {n_seg} segments", flush=True)
    rng = np.random.default_rng(71)
    h, o = HipRHS(m), O.OracleRHS(m)
    inp = _inputs(m, rng)
    h.set_step_inputs(**inp)
    o.set_step_inputs(**inp)
    for k, y in enumerate(_states(m, rng, 4)):
        rc_h, dh = h.ode(y)
        rc_o, do, mag = o.ode(y, with_mag=True)
        assert rc_h == rc_o
        assert_rhs_close(dh, do, mag, 1e-12, f"{n_tri}/{n_seg} state {k}")
    h.close()


def test_river_free_run_matches_host_run():
    m = make_watershed(535, seed=3, n_seg=0, network="valley")
    f = StormForcing(8, seed=5, nelem=m.nelem, p_wet=0.5)
    y0 = m.initial_state()
    cfg = run_config(2 * 3600.0)
    print("[edge] river-free 2 h run", flush=True)
    rc_g, y_g, st_g = run_gpu(m, f, cfg, y0)
    assert rc_g == 0, st_g.as_dict()
    rc_c, y_c, st_c = O.OracleRun(m, f).run(y0, cfg)
    assert rc_c == 0
    w = 1.0 / (1e-3 * np.abs(y_c) + 1e-4)
    assert math.sqrt(np.mean(((y_g - y_c) * w) ** 2)) < 1.0
    assert abs(st_g.nst - st_c.nst) <= 0.1 * st_c.nst
