"""Seeded workloads of the batched paths.

* Ensemble members: a log-uniform triangle count in [2,000, 50,000], a binary-branching
  river network with about one segment per 49 triangles (the 200,000-triangle /
  4,095-segment ratio), Ksat, porosity and Manning n each multiplied by a seeded
  U(0.5, 1.5), and an independent storm-forcing seed.
* Serving requests: Poisson arrivals, a log-uniform triangle count in [1,000, 100,000],
  a U(1, 30)-day horizon, and their own parameter and forcing seeds.

Meshes are generated in parallel worker processes (the generator is numpy/scipy work).
The workers are forked from a server process (multiprocessing's "forkserver"), never from
the calling process, which may hold an initialised GPU runtime; start_workers() starts
that server, and should be called before the first HIP call.
"""
import math
import multiprocessing
import os
from concurrent.futures import ProcessPoolExecutor

import numpy as np

from .forcing import StormForcing
from .meshgen import make_watershed


def tree_segments(n_tri):
    """Segments of a binary-branching network for n_tri triangles: 2^k - 1 closest (in
    log) to n_tri / 48.8, at least 31 (the smallest tree the generator lays out on a
    1,000-triangle mesh: main stem plus two primaries)."""
    k = max(5, int(round(math.log2(max(n_tri / 48.8, 1.0) + 1.0))))
    return (1 << k) - 1


def member_spec(nmem=256, seed=2024, nmin=2000, nmax=50000):
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(math.log(nmin), math.log(nmax), nmem)).astype(int)
    mult = rng.uniform(0.5, 1.5, (nmem, 3))
    seeds = rng.integers(0, 2**31 - 1, (nmem, 2))
    return [dict(n_tri=int(sizes[k]), n_seg=tree_segments(int(sizes[k])), mesh_seed=int(seeds[k, 0]),
                 forcing_seed=int(seeds[k, 1]),
                 param_mult=dict(ksat=float(mult[k, 0]), porosity=float(mult[k, 1]), rough=float(mult[k, 2])))
            for k in range(nmem)]


def _build(spec):
    # the network generator can fail to settle a tree of exactly n_seg segments on a given
    # triangulation; the next seeds are tried in order (deterministic)
    err = None
    for k in range(8):
        try:
            return make_watershed(spec["n_tri"], seed=spec["mesh_seed"] + k, n_seg=spec["n_seg"], network="tree",
                                  param_mult=spec["param_mult"])
        except RuntimeError as e:
            err = e
    raise err


def start_workers():
    """Start the server process that worker processes are forked from.  Called before the
    first HIP call, it keeps every worker free of the GPU runtime's state and threads."""
    from multiprocessing import forkserver
    # The server is a new interpreter that does not inherit this process's sys.path;
    # PYTHONPATH lets it import the module that resets foreign signal handlers.
    pkg_root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    paths = [p for p in os.environ.get("PYTHONPATH", "").split(os.pathsep) if p]
    if pkg_root not in paths:
        os.environ["PYTHONPATH"] = os.pathsep.join([pkg_root] + paths)
    forkserver.set_forkserver_preload(["pihm_amd._fs_init"])
    forkserver.ensure_running()


def worker_pool(workers=None):
    """A process pool whose workers fork from the forkserver, not from this process."""
    if workers is None:
        workers = min(16, len(os.sched_getaffinity(0)))
    return ProcessPoolExecutor(max_workers=workers, mp_context=multiprocessing.get_context("forkserver"))


def build_meshes(specs, workers=None):
    if workers is None:
        workers = min(16, len(os.sched_getaffinity(0)))
    if workers <= 1 or len(specs) < 4:
        return [_build(s) for s in specs]
    with worker_pool(workers) as ex:
        return list(ex.map(_build, specs, chunksize=max(1, len(specs) // (4 * workers))))


def member_forcing(spec, mesh, hours, p_wet=0.1):
    return StormForcing(hours, seed=spec["forcing_seed"], nelem=mesh.nelem, p_wet=p_wet)


def serve_payload(spec, hours):
    """A request's mesh, forcing (hours + 2 records) and initial state."""
    m = _build(spec)
    f = member_forcing(spec, m, hours + 2, p_wet=0.1)
    return m, f, m.initial_state()


def request_stream(duration_s=600.0, rate=20.0, seed=77, nmin=1000, nmax=100000, hmin_days=1.0,
                   hmax_days=30.0):
    """Poisson arrivals over duration_s; each request: arrival time, size, horizon, seeds."""
    rng = np.random.default_rng(seed)
    t, out = 0.0, []
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= duration_s:
            break
        n = int(math.exp(rng.uniform(math.log(nmin), math.log(nmax))))
        mult = rng.uniform(0.5, 1.5, 3)
        out.append(dict(arrival=t, n_tri=n, n_seg=tree_segments(n),
                        horizon_days=float(rng.uniform(hmin_days, hmax_days)),
                        mesh_seed=int(rng.integers(0, 2**31 - 1)), forcing_seed=int(rng.integers(0, 2**31 - 1)),
                        param_mult=dict(ksat=float(mult[0]), porosity=float(mult[1]), rough=float(mult[2]))))
    return out
