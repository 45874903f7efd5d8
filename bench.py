"""Benchmark: simulated days per second of the 200,000-triangle FP64 MM-PIHM solve on one
MI355X, with the RHS kernels' achieved HBM bandwidth and the CPU (oracle) baseline.

A step is one simulated hour (60 model steps of 60 s, forcing applied every 900 s) of the
200,000-triangle / 4,095-segment synthetic watershed (608,190 FP64 unknowns) under the
storm forcing; the integrator is BDF/Newton/SPGMR at reltol 1e-3, abstol 1e-4.  W hours
run untimed, then K hours are timed with the device synchronised on both sides.

Prints one JSON line.
"""
import argparse
import ctypes as C
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "pihm-cuda_amd"))

import numpy as np  # noqa: E402

from pihm_amd import _lib  # noqa: E402
from pihm_amd.forcing import StormForcing, run_config  # noqa: E402
from pihm_amd.meshgen import config_mesh  # noqa: E402

HBM_PEAK_GBS = 8000.0
GRAPHS = 1               # replay Krylov iterations as HIP graphs (off with --no-graphs)
BYTES_PER_ELEM = 344     # algorithmic bytes of one RHS per element
BYTES_PER_SEG = 200      # ... per river segment
# HBM bytes one RHS evaluation at the 200,000-triangle workload moves, from the rocprofv3
# FETCH_SIZE and WRITE_SIZE passes over `bench.py --profile-kernels 20` (KiB per dispatch),
# each scaled by what the counter reads for the calibration copy in the same run:
# FETCH_SIZE x 2.000 (262,157 KiB read back for 524,288 KiB copied), WRITE_SIZE x 1.000.
# One launch (k_rhs_tile): 40,867.1 x 2 + 5,038.1 KiB.  Two kernels (k_prep + k_flux):
# (14,789.9 + 28,239.8) x 2 + (12,919.4 + 4,818.1) KiB.
RHS_TRAFFIC_BYTES = {"one launch": (40867.1 * 2.0 + 5038.1) * 1024,
                     "two kernels": ((14789.9 + 28239.8) * 2.0 + (12919.4 + 4818.1)) * 1024}
RHS_KERNELS = {"one launch": "one RHS evaluation (k_rhs_tile: one launch, tile records in LDS)",
               "two kernels": "one RHS evaluation (k_prep + k_flux, back to back)"}
# Land-surface step (k_intcp_snow_et), algorithmic bytes per element: 15 FP64 parameters
# (albedo, shdfac, cmcfactr, cfactr, rzd, rgl, rsmin, rsmax, topt, depth, porosity, smcmin,
# smcwlt, smcref, surface pressure: 120) + station index (4) + lai, the aerodynamic
# resistance's numerator and the carried cmc, sneqv (32) + unsat, gw of the state (16)
# read; cmc, sneqv, ec, drip, pcpdrp, edir, ett (56) written.
LSM_BYTES_PER_ELEM = 228
# HBM bytes of one k_intcp_snow_et launch at 200,000 elements (FETCH_SIZE x 2 + WRITE_SIZE)
LSM_TRAFFIC_BYTES = 45800755.0
LSM_STATIONS = 64


def _log(msg):
    """Progress of the legs, on stderr (the JSON line stays alone on stdout)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def _cfg(t_end):
    return run_config(t_end, graphs=GRAPHS)


def gpu_run(mesh, forcing, hours_total, mark_hours, snap_hour=None):
    """Run the device path for hours_total simulated hours; return wall-clock marks at the
    requested hour boundaries (device synchronised at each), stats and kernel timings, and
    (snap_hour) the state at that hour, copied out before that hour's mark is taken."""
    L = _lib.lib()
    desc = mesh.desc()
    ctx = C.c_void_p()
    rc = L.pihm_ctx_create(C.byref(desc), C.byref(ctx))
    if rc:
        raise RuntimeError(f"pihm_ctx_create: {rc}")
    L.pihm_ctx_set_nan_check(ctx, 0)
    stream = C.c_void_p()
    assert L.pihm_stream_create(C.byref(stream)) == 0
    L.pihm_ctx_set_stream(ctx, stream)
    fs = forcing.struct()
    gb = C.c_void_p()
    assert L.pihm_gpu_backend_create(ctx, C.byref(fs), C.byref(gb)) == 0
    be = _lib.Backend()
    L.pihm_gpu_backend_get(gb, C.byref(be))
    y = L.N_VNew_Hip(mesh.neq)
    yd = L.N_VNew_Hip(mesh.neq)
    L.N_VSetStream_Hip(y, stream)
    L.N_VSetStream_Hip(yd, stream)
    y0 = mesh.initial_state()
    assert L.pihm_ctx_state_to_device(ctx, y0.ctypes.data_as(_lib.c_double_p), y) == 0
    _lib.SummaryFn(be.summary)(be.user, y)        # ws0 = initial surface state
    # launch shape of the RHS (element-tile placement over the XCDs), chosen on the
    # initial state before the timed run; the results do not depend on it
    span, span_ms = C.c_int(), np.zeros(len(_lib.PIHM_TUNE_SPANS))
    assert L.pihm_ctx_autotune(ctx, L.N_VGetDeviceArrayPointer_Hip(y), L.N_VGetDeviceArrayPointer_Hip(yd), 20,
                               C.byref(span), span_ms.ctypes.data_as(_lib.c_double_p)) == 0
    rhs_path = "one launch" if L.pihm_ctx_get_rhs_path(ctx) == _lib.PIHM_RHS_ONE_LAUNCH else "two kernels"
    marks = {}
    want = {int(round(h * 60)) for h in mark_hours}     # in model steps of 60 s
    snap_k = None if snap_hour is None else int(round(snap_hour * 60))
    snap = np.empty(mesh.neq)

    def step_end(user, t, yv):
        k = int(round(t / 60.0))
        if hours_total >= 24 and k % 1440 == 0:       # a line per simulated day of a long run
            _log(f"{mesh.nelem}-triangle run: day {k // 1440} of {hours_total / 24:g} at "
                 f"{time.perf_counter() - marks[0.0]:.1f} s")
        if k == snap_k:
            L.pihm_ctx_sync(ctx)
            L.pihm_ctx_state_from_device(ctx, y, snap.ctypes.data_as(_lib.c_double_p))
        if k in want:
            L.pihm_ctx_sync(ctx)
            marks[k / 60.0] = time.perf_counter()
        return 0

    cb = _lib.StepEndFn(step_end)
    be.model_step_end = C.cast(cb, C.c_void_p)
    cfg = _cfg(hours_total * 3600.0)
    st = _lib.RunStats()
    L.pihm_ctx_sync(ctx)
    marks[0.0] = time.perf_counter()
    rc = L.pihm_drive(C.byref(be), y, C.byref(cfg), C.byref(st))
    L.pihm_ctx_sync(ctx)
    nan = L.pihm_ctx_poll_nan(ctx)
    ydev = L.N_VGetDeviceArrayPointer_Hip(y)
    yddev = L.N_VGetDeviceArrayPointer_Hip(yd)
    t2 = [C.c_double(), C.c_double()]
    L.pihm_ctx_time_kernels(ctx, ydev, yddev, 5, *[C.byref(x) for x in t2])   # warm
    L.pihm_ctx_time_kernels(ctx, ydev, yddev, 100, *[C.byref(x) for x in t2])
    ms_eval = C.c_double()
    L.pihm_ctx_time_rhs(ctx, ydev, yddev, 200, C.byref(ms_eval))
    yend = np.empty(mesh.neq)
    L.pihm_ctx_state_from_device(ctx, y, yend.ctypes.data_as(_lib.c_double_p))
    L.N_VDestroy_Hip(y)
    L.N_VDestroy_Hip(yd)
    L.pihm_gpu_backend_destroy(gb)
    L.pihm_ctx_destroy(ctx)
    L.pihm_stream_destroy(stream)
    return dict(rc=rc, nan=nan, marks=marks, stats=st.as_dict(), kernel_ms=[x.value for x in t2],
                ms_eval=ms_eval.value, yend=yend, snap=snap if snap_k is not None else None,
                tuning={"tile_span": span.value,
                        "rhs_path": rhs_path,
                        "rhs_ms_by_span": dict(zip(_lib.PIHM_TUNE_SPANS, span_ms.round(6).tolist()))})


def ensemble_run(nmem, hours, sample, seed):
    """C3: nmem members advanced together (pihm_ensemble.h) over [0, hours], against a
    stratified sample of the same members run one at a time through the single-solve
    device path (pihm_run_gpu) over the same window."""
    from pihm_amd.ensemble import Ensemble
    from pihm_amd.workloads import build_meshes, member_forcing, member_spec
    L = _lib.lib()
    specs = member_spec(nmem, seed=seed)
    meshes = build_meshes(specs)
    forc = [member_forcing(sp, m, hours + 3) for sp, m in zip(specs, meshes)]
    cfg = _cfg(hours * 3600.0)
    e = Ensemble(meshes, forc, cfg)
    for k, m in enumerate(meshes):
        e.set_state(k, m.initial_state())
    tuning = e.autotune(10)          # launch shapes, before the timed run
    ops0 = e.op_stats()
    t0 = time.perf_counter()
    if hours <= 2.0:
        nfail = e.run(hours * 3600.0)
    else:
        # simulated hour by simulated hour, a progress line after each (the members'
        # model-step loops pause at each boundary and resume; results do not change)
        nfail, h = 0, 0.0
        while h < hours:
            h = min(h + 1.0, hours)
            nfail += e.run(h * 3600.0)
            _log(f"ensemble: hour {h:g} of {hours:g} at {time.perf_counter() - t0:.1f} s, {nfail} failed")
    wall = time.perf_counter() - t0
    st = [e.stats(k) for k in range(nmem)]
    cnt = e.counters()
    ops1 = e.op_stats()
    e.close()
    member_days = nmem * hours / 24.0
    # one at a time: every (nmem/sample)-th member in size order
    order = np.argsort([m.nelem for m in meshes])
    pick = order[np.linspace(0, nmem - 1, min(sample, nmem)).round().astype(int)]
    single_wall = []
    for n, k in enumerate(pick):
        m = meshes[k]
        _log(f"ensemble: one-at-a-time member {n + 1} of {len(pick)} ({m.nelem} triangles)")
        desc = m.desc()
        ctx = C.c_void_p()
        assert L.pihm_ctx_create(C.byref(desc), C.byref(ctx)) == 0
        y = np.ascontiguousarray(m.initial_state())
        rs = _lib.RunStats()
        fs = forc[k].struct()
        L.pihm_ctx_sync(ctx)
        t1 = time.perf_counter()
        rc = L.pihm_run_gpu(ctx, C.byref(fs), C.byref(cfg), y.ctypes.data_as(_lib.c_double_p), C.byref(rs))
        single_wall.append(time.perf_counter() - t1)
        L.pihm_ctx_destroy(ctx)
        if rc:
            raise RuntimeError(f"single-solve member {k} failed: {rc}")
    est_single_total = float(np.sum(single_wall)) * nmem / len(pick)
    return dict(
        metric="member-days/s", value=member_days / wall, unit="member-days/s", members=nmem,
        elements=int(sum(m.nelem for m in meshes)), unknowns=int(sum(m.neq for m in meshes)),
        window_hours=hours, wall_s=wall, failed=nfail,
        steps_total=int(sum(s.nst for s in st)), steps_max=int(max(s.nst for s in st)),
        rhs_evals_total=int(sum(s.nfe + s.nfes for s in st)), **cnt,
        one_at_a_time={"value": member_days / est_single_total, "unit": "member-days/s",
                       "sample": f"{len(pick)} members evenly spaced in size order, each timed alone "
                                 f"through pihm_run_gpu over the same window; total scaled by "
                                 f"{nmem}/{len(pick)}",
                       "sample_wall_s": float(np.sum(single_wall))},
        speedup_vs_one_at_a_time=(member_days / wall) / (member_days / est_single_total),
        tuning=tuning, ops_kernel=_ops_roofline(ops0, ops1))


def _ops_roofline(a, b):
    """k_ens_ops over the timed run: launches timed with HIP events, algorithmic bytes per
    launch, device time per launch and the achieved bandwidth against the HBM peak."""
    n, by, ms = b["launches"] - a["launches"], b["bytes"] - a["bytes"], b["ms"] - a["ms"]
    if n <= 0 or ms <= 0.0:
        return None
    gbps = by / (ms * 1e-3) / 1e9
    return {"kernel": "k_ens_ops (every ready member's queued vector operations, one launch)",
            "launches_timed": n, "bytes_per_launch": by / n, "us_per_launch": ms * 1e3 / n,
            "achieved": gbps, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": gbps / HBM_PEAK_GBS}


SERVE_CLASSES = [(8192, 48), (32768, 32), (131072, 12)]   # slot capacity (elements) x count


def serve_run(duration_s, rate, horizon_unit_s, check_alone, seed):
    """C4: a Poisson stream of solves served with continuous batching on a slot ensemble."""
    from pihm_amd.serve import Request, Scheduler, SlotEngine, latency_summary, slot_bytes
    from pihm_amd.workloads import request_stream, serve_payload, tree_segments, worker_pool
    stream = request_stream(duration_s=duration_s, rate=rate, seed=seed, hmin_days=1.0, hmax_days=30.0)
    # horizon: U(1, 30) units of horizon_unit_s, rounded to whole 60 s model steps
    for rq in stream:
        rq["t_end"] = max(60.0, round(rq["horizon_days"] * horizon_unit_s / 60.0) * 60.0)
    max_hours = max(rq["t_end"] for rq in stream) / 3600.0
    nrec_cap = int(max_hours) + 4
    caps = []
    for ne, cnt in SERVE_CLASSES:
        caps += [(ne, tree_segments(ne), nrec_cap)] * cnt
    hbm = sum(slot_bytes(ne, nr) for ne, nr, _ in caps)
    pool = worker_pool()      # forked from the server started before the first HIP call
    futs = [pool.submit(serve_payload, rq, rq["t_end"] / 3600.0) for rq in stream]
    reqs = [Request(i, rq["arrival"], rq["t_end"], futs[i]) for i, rq in enumerate(stream)]
    eng = SlotEngine(caps, _cfg(3600.0))
    t0 = time.perf_counter()
    clock = lambda: time.perf_counter() - t0
    sched = Scheduler(eng, clock, ticks_per_round=8)
    sched.run(reqs, sleep=time.sleep)
    wall = clock()
    cnt = eng.counters()
    done = [r for r in reqs if r.t_done is not None and not r.failed]
    member_days = sum(r.t_end for r in done) / 86400.0
    lat = latency_summary(done)
    # alone vs in the mix, bit for bit, for a few requests of different sizes
    eng.close()
    same = 0
    order = sorted(range(len(reqs)), key=lambda i: stream[i]["n_tri"])
    pick = [order[int(round(q * (len(order) - 1)))] for q in np.linspace(0.0, 1.0, check_alone)] if check_alone else []
    for i in pick:
        m, f, y0 = futs[i].result()
        cap = [c for c in caps if c[0] >= m.nelem][0]
        e1 = SlotEngine([cap], _cfg(3600.0))
        r1 = Request(0, 0.0, reqs[i].t_end, (m, f, y0))
        Scheduler(e1, time.perf_counter, ticks_per_round=64).run([r1])
        e1.close()
        same += int(np.array_equal(r1.result, reqs[i].result))
    pool.shutdown(cancel_futures=True)
    return dict(
        metric="request latency p50/p99 (s) and member-days/s under continuous batching",
        p50_s=lat.get("p50_s"), p99_s=lat.get("p99_s"), mean_s=lat.get("mean_s"), max_s=lat.get("max_s"),
        value=member_days / wall, unit="member-days/s", requests=len(reqs), completed=len(done),
        failed=sum(1 for r in reqs if r.failed), wall_s=wall,
        stream={"rate_per_s": rate, "duration_s": duration_s, "sizes": "log-uniform 1,000-100,000 triangles",
                "horizon": f"U(1, 30) x {horizon_unit_s:g} s (the request spec's days scaled to "
                           f"{horizon_unit_s:g} s so the offered load fits one GPU)"},
        slots={"classes": [{"elements": ne, "count": n} for ne, n in SERVE_CLASSES], "hbm_bytes": hbm},
        alone_equals_mix=f"{same}/{len(pick)} bit-identical", **cnt)


def closure_run(days):
    """Water-balance closure of the 535-triangle / 20-segment watershed over `days` simulated
    days on the device, with the closure accumulators (pihm_ctx_set_closure)."""
    from pihm_amd.balance import closure
    from pihm_amd.rhs import run_gpu
    m = config_mesh("c0", seed=1)
    f = StormForcing(int(days * 24) + 4, seed=11, nelem=m.nelem)
    y0 = m.initial_state()
    t0 = time.perf_counter()
    rc, y1, st = run_gpu(m, f, _cfg(days * 86400.0), y0, closure=True)
    wall = time.perf_counter() - t0
    if rc:
        raise RuntimeError(f"closure run failed: {rc}")
    c = closure(m, y0, y1)
    return {"workload": f"c0: 535 triangles, 20 segments, {days:g} simulated days, storm forcing",
            "residual_rel_to_precip": c["rel"], "precip_m3": c["precip_m3"],
            "storage_change_m3": c["storage_change_m3"], "residual_m3": c["residual_m3"],
            "sim_days_per_s": days / wall, "nst": st.nst}


def _producers(mesh, seed):
    """A single-solve context with the land-surface step (64 stations, 48 hourly records)
    and boundary series on every boundary edge and outlet (16 + 1 series) bound to it."""
    from pihm_amd.bc import HipBC
    from pihm_amd.lsm import HipLSM, synthetic_meteo, synthetic_params
    from pihm_amd.rhs import HipRHS
    ne = mesh.nelem
    h = HipRHS(mesh, nan_check=False)
    p = synthetic_params(ne, LSM_STATIONS, seed=seed)
    for k in ("depth", "porosity", "rzd", "zmax"):
        p[k] = mesh.a[k].copy()
    p["smcmin"] = np.minimum(p["smcmin"], 0.5 * p["porosity"])
    lsm = HipLSM(h, p, synthetic_meteo(LSM_STATIONS, 48, seed=seed))
    rng = np.random.default_rng(seed)
    lsm.set_veg(rng.uniform(0.0, 4.0, ne), rng.uniform(0.01, 0.5, ne), 1e-7)
    lsm.set_state(np.zeros(ne), rng.uniform(0.0, 0.01, ne))
    bnd = mesh.a["nabr"].reshape(ne, 3) == 0
    code = np.where(bnd, 1 + np.arange(3 * ne).reshape(ne, 3) % 16, 0).astype(np.int32)
    rcode = np.where(mesh.a["down"] < 0, -1, 0).astype(np.int32)
    ft = np.tile(3600.0 * np.arange(48.0), (16, 1))
    bc = HipBC(h, code, (ft, rng.uniform(100.0, 200.0, ft.shape)), rcode,
               (ft[:1], rng.uniform(0.01, 0.1, (1, 48))))
    y0 = mesh.initial_state()
    assert lsm.step(3600.0 * 5.5, 900.0, y0) == 0
    ncoded = int(bnd.sum()) + int((rcode != 0).sum())
    bc_bytes = 3 * ne * 4 + mesh.nriver * 4 + 8 * ncoded
    return h, lsm, bc, bc_bytes


def producers_run(mesh, seed, iters=200):
    """The per-step producers on the device at the 200,000-triangle workload, timed with
    HIP events: one land-surface step (k_meteo + k_intcp_snow_et) and one boundary-series
    apply (k_bc_interp + k_bc_scatter)."""
    L = _lib.lib()
    h, lsm, bc, bc_bytes = _producers(mesh, seed)
    ms_l, ms_b = C.c_double(), C.c_double()
    t = 3600.0 * 12.25
    assert L.pihm_lsm_time(lsm.h, t, 900.0, h.y, 5, C.byref(ms_l)) == 0
    assert L.pihm_lsm_time(lsm.h, t, 900.0, h.y, iters, C.byref(ms_l)) == 0
    assert L.pihm_bc_time(bc.h, t, 5, C.byref(ms_b)) == 0
    assert L.pihm_bc_time(bc.h, t, iters, C.byref(ms_b)) == 0
    ne = mesh.nelem
    lsm_bytes = LSM_BYTES_PER_ELEM * ne
    ach = lsm_bytes / (ms_l.value * 1e-3) / 1e9
    bc.close()
    lsm.close()
    h.close()
    return {
        "lsm_step": {"ms": ms_l.value, "elements": ne, "stations": LSM_STATIONS,
                     "roofline": {"kernel": "k_meteo + k_intcp_snow_et (one LSM step)", "bound": "hbm",
                                  "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                  "frac": ach / HBM_PEAK_GBS, "algorithmic_bytes": lsm_bytes,
                                  "traffic": LSM_TRAFFIC_BYTES}},
        "bc_apply": {"ms": ms_b.value, "algorithmic_bytes": bc_bytes,
                     "achieved_GBps": bc_bytes / (ms_b.value * 1e-3) / 1e9,
                     "note": "two launches over 600k edge codes: launch-latency bound"},
    }


NOAH_EXTRA_BYTES_PER_ELEM = 32   # smcmin, sh2o[0], fcr, gwet read by the _NOAH_ branches
# HBM bytes per launch at the 100,000-triangle mesh from the FETCH_SIZE and WRITE_SIZE passes
# (FETCH_SIZE x 2, calibrated on the copy kernel): one Noah RHS on the one-launch path
# (k_rhs_tile with the _NOAH_ branches: 19,646.4 x 2 + 2,411.0 KiB), one soil-water step
NOAH_RHS_TRAFFIC_BYTES = (19646.4 * 2.0 + 2411.0) * 1024
NOAH_SOIL_TRAFFIC_BYTES = 122959155.0
# algorithmic bytes of the soil-water step per element: reads 4 layer fields (sldpth, smc, sh2o, et;
# 352 B), 12 FP64 scalars and 2 int32 (104 B); writes 5 layer fields (satdpth, runoff2_lyr, smflxv,
# smc, sh2o; 440 B), 4 FP64 scalars and nwtbl (36 B)
NOAH_SOIL_BYTES_PER_ELEM = 932


def noah_rhs_run(seed, iters=200, coupled=True):
    """The RHS with the reference's _NOAH_ branches on the 100,000-triangle / 2,047-segment
    workload, timed with HIP events beside the plain RHS on the same mesh and state, the
    soil-water step, and (coupled=True) one simulated hour of the coupled and the plain run.
    The profile mode leaves the runs out: they replay Krylov iterations from HIP graphs,
    which has crashed the host process under the kernel tracer."""
    from pihm_amd.rhs import HipRHS
    L = _lib.lib()
    m = config_mesh("c2", seed=seed)
    ne = m.nelem
    rng = np.random.default_rng(seed)
    h = HipRHS(m, nan_check=False)
    h.set_step_inputs(pcpdrp=rng.uniform(0, 3e-6, ne), edir=rng.uniform(0, 2e-8, ne),
                      ett=rng.uniform(0, 4e-8, ne), ws0_surf=rng.uniform(0, 2e-3, ne))
    y0 = m.initial_state()
    rc, _ = h.ode(y0)
    ydev, yddev = L.N_VGetDeviceArrayPointer_Hip(h.y), L.N_VGetDeviceArrayPointer_Hip(h.yd)
    ms = {}
    for mode in ("plain", "noah"):
        if mode == "noah":
            smcmin = rng.uniform(0.01, 0.06, ne)
            h.set_noah(smcmin)
            nl = 11
            h.set_noah_step(smcmin + m.a["porosity"] * rng.uniform(0.1, 0.9, ne), rng.uniform(0.5, 1.0, ne),
                            rng.uniform(0.0, 8e-9, (ne, nl)), rng.integers(1, nl + 1, ne).astype(np.int32),
                            np.full(ne, 4, np.int32))
        t = C.c_double()
        L.pihm_ctx_time_rhs(h.ctx, ydev, yddev, 5, C.byref(t))
        L.pihm_ctx_time_rhs(h.ctx, ydev, yddev, iters, C.byref(t))
        ms[mode] = t.value
    # the soil-water step after each model step: 11 layers of 0.1-0.5 m per element
    from pihm_amd.noahsoil import MAXLYR, HipNoahSoil
    nsoil = np.full(ne, MAXLYR, np.int32)
    sld = rng.uniform(0.1, 0.5, (ne, MAXLYR))
    smcmax = smcmin + m.a["porosity"]
    soil = HipNoahSoil(h, nsoil, sld, smcmax, np.full(ne, 0.15))
    th = smcmin[:, None] + m.a["porosity"][:, None] * rng.uniform(0.3, 0.8, (ne, MAXLYR))
    soil.set_state(th, th)
    soil.step(60.0, rng.uniform(0, 2e-6, ne), rng.uniform(0, 1e-7, ne))
    soil.time(60.0, 5)
    soil_ms = soil.time(60.0, iters)
    soil_bytes = ne * NOAH_SOIL_BYTES_PER_ELEM
    soil.close()
    h.close()
    bytes_plain = BYTES_PER_ELEM * ne + BYTES_PER_SEG * m.nriver
    bytes_noah = bytes_plain + NOAH_EXTRA_BYTES_PER_ELEM * ne
    ach = bytes_noah / (ms["noah"] * 1e-3) / 1e9
    out = {"workload": f"c2 mesh: {ne} triangles, {m.nriver} river segments, NEQ {m.neq}",
           "ms_per_eval": ms["noah"], "plain_ms_per_eval": ms["plain"],
           "roofline": {"kernel": "one RHS evaluation with the _NOAH_ branches (k_rhs_tile<noah>: one launch)",
                        "bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": ach / HBM_PEAK_GBS, "algorithmic_bytes": bytes_noah,
                        "traffic": NOAH_RHS_TRAFFIC_BYTES},
           "soil_step": {"kernel": "k_noah_soil (Summary's water table and lateral split, NoahHydrol)",
                         "ms_per_step": soil_ms, "layers": MAXLYR, "algorithmic_bytes": soil_bytes,
                         "achieved_GBps": soil_bytes / (soil_ms * 1e-3) / 1e9,
                         "traffic": NOAH_SOIL_TRAFFIC_BYTES,
                         "note": "once per model step; one thread per element column"},
           "note": "per-LSM-step values (edir, ett, per-layer et) held fixed: the RHS and the soil-water "
                   "step alone; the coupled run with the Noah step is the flux_pihm object"}
    if not coupled:
        return out
    # one simulated hour of the coupled run (the _NOAH_ RHS in the integrator, the soil-water step at
    # every Summary) beside the plain run, same mesh and storm forcing
    from pihm_amd.noahsoil import run_gpu_noah
    from pihm_amd.rhs import run_gpu
    f = StormForcing(4, seed=11, nelem=ne)
    y0 = m.initial_state()
    gw0 = y0[2 * ne:3 * ne]
    depth = m.a["depth"]
    sld6 = np.zeros((ne, MAXLYR))
    sld6[:, :6] = depth[:, None] / 6.0
    nw = np.clip(6 - np.floor(gw0 / (depth / 6.0)).astype(np.int32), 0, 6).astype(np.int32)
    th6 = smcmin[:, None] + m.a["porosity"][:, None] * np.where(np.arange(MAXLYR) < 6, 0.5, 0.0)[None, :]
    noah = dict(smcmin=smcmin, et=np.where(np.arange(MAXLYR) < 3, 2e-8 / 3, 0.0)[None, :].repeat(ne, 0),
                nwtbl=nw, nroot=np.full(ne, 3, np.int32), nsoil=np.full(ne, 6, np.int32), sldpth=sld6,
                smcmax=smcmin + m.a["porosity"], frzx=np.full(ne, 0.15), smc=th6, sh2o=th6)
    t0 = time.perf_counter()
    rc_n, _, st_n, _, _ = run_gpu_noah(m, f, run_config(3600.0), y0, noah)
    wall_n = time.perf_counter() - t0
    t0 = time.perf_counter()
    rc_p, _, st_p = run_gpu(m, f, run_config(3600.0), y0)
    wall_p = time.perf_counter() - t0
    out["coupled_run"] = {"rc": rc_n, "sim_days_per_s": (1.0 / 24.0) / wall_n, "wall_s": wall_n, "nst": st_n.nst,
                          "plain_sim_days_per_s": (1.0 / 24.0) / wall_p, "plain_rc": rc_p, "plain_nst": st_p.nst,
                          "window": "first simulated hour, 60 model steps, storm forcing, 6 soil layers"}
    return out


# algorithmic bytes of one Noah energy-balance step per element with N soil layers: reads nsoil, nroot,
# station (12 B), 22 land-cover/soil parameters, smcmax, smcmin, alpha, beta, 10 carried scalars, q1 and
# LAI (264 B), and sldpth, smc, sh2o, stc per layer (32N B); writes smc, sh2o, stc per layer (24N B),
# the 11 carried values, 16 fluxes, pcpdrp/edir/ett, the 11 layers of et and sh2o[0] (336 B)
def noah_step_bytes_per_elem(nsoil):
    return 316 + 32 * nsoil + 336 + 24 * nsoil


def _flux_inputs(seed, hours):
    """C2's Flux-PIHM inputs: the 100,000-triangle mesh, 11 soil layers, one station with
    hourly synthetic forcing, LAI between each element's minimum and maximum."""
    from pihm_amd.noahlsm import synthetic_noah, synthetic_noah_meteo, synthetic_noah_state
    m = config_mesh("c2", seed=seed)
    p = synthetic_noah(m, seed=seed)
    state = synthetic_noah_state(m, p, seed=seed, t0=283.15)
    ft, data = synthetic_noah_meteo(1, int(math.ceil(hours)) + 1, seed=seed)
    rng = np.random.default_rng(seed)
    lai = p["laimin"] + (p["laimax"] - p["laimin"]) * rng.uniform(0.0, 1.0, m.nelem)
    return m, p, state, ft, data, lai


def _noah_step_ms(m, p, state, ft, data, lai, iters):
    """HIP-event time of one Noah step at noon (sunlit, every branch of the canopy and
    soil-heat physics live), `iters` steps back to back after 3 untimed ones."""
    from pihm_amd.noahlsm import flux_pihm_setup
    h, soil, lsm = flux_pihm_setup(m, p, state, ft, data, lai)
    try:
        lsm.time(12.0 * 3600.0, 900.0, 3)
        return lsm.time(12.0 * 3600.0, 900.0, iters)
    finally:
        lsm.close(); soil.close(); h.close()


def flux_pihm_run(seed, hours=24.0, cpu_hours=0.25, nthreads=0, iters=50):
    """C2, the Flux-PIHM coupled run, on the 100,000-triangle mesh with 11 soil layers: one
    Noah step timed with HIP events (the stations' interpolation, k_noah, the hand-over),
    the coupled model over the first `hours` simulated hours (pihm_run_gpu_flux: the Noah
    step every 15 minutes, the _NOAH_ RHS in the integrator, the soil-water step every
    model step), and (nthreads > 0) the oracle's host run over the first `cpu_hours`
    beside the device over the same window."""
    from pihm_amd.noahlsm import MAXLYR, flux_pihm_setup, run_gpu_flux
    m, p, state, ft, data, lai = _flux_inputs(seed, max(hours, cpu_hours, 13.0))
    ne = m.nelem
    ms = _noah_step_ms(m, p, state, ft, data, lai, iters)
    step_bytes = ne * noah_step_bytes_per_elem(MAXLYR)
    ach = step_bytes / (ms * 1e-3) / 1e9

    def device(hrs):
        h, soil, lsm = flux_pihm_setup(m, p, state, ft, data, lai)
        t0 = time.perf_counter()
        rc, y, st = run_gpu_flux(lsm, _cfg(hrs * 3600.0), m.initial_state())
        wall = time.perf_counter() - t0
        lsm.close(); soil.close(); h.close()
        if rc != 0:
            raise RuntimeError(f"Flux-PIHM device run failed: rc={rc} {st.as_dict()}")
        return wall, st

    wall, st = device(hours)
    out = {"workload": f"c2: {ne} triangles, {m.nriver} river segments, {MAXLYR} soil layers, NEQ {m.neq}; "
                       f"one station, hourly forcing T = 10 + 8 sin(diurnal) degC, storm precipitation",
           "window": f"first {hours:g} simulated hours from midnight: {int(hours * 4)} Noah steps, "
                     f"{int(hours * 60)} model steps",
           "sim_days_per_s": (hours / 24.0) / wall, "wall_s": wall, "nst": st.nst, "nfe": st.nfe,
           "nli": st.nli, "ncfn": st.ncfn,
           "noah_step": {"kernel": "one Noah step: k_noah_meteo + k_noah + the hand-over's k_gwtransp",
                         "ms_per_step": ms, "bound": "FP64 latency (exp/log/pow chains, fixed-point and Newton "
                                                     "iterations per element)",
                         "algorithmic_bytes": step_bytes, "achieved_GBps": ach,
                         "frac_of_hbm_peak": ach / HBM_PEAK_GBS, "per_simulated_hour": 4}}
    if nthreads > 0:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        gwall, _ = device(cpu_hours)
        noah = dict(smcmin=p["smcmin"], et=np.zeros((ne, MAXLYR)), nwtbl=np.zeros(ne, np.intc),
                    nroot=p["nroot"], nsoil=p["nsoil"], sldpth=p["sldpth"], smcmax=p["smcmax"], frzx=p["frzx"],
                    smc=state["smc"], sh2o=state["sh2o"])
        run = O.OracleRun(m, StormForcing(4, nelem=ne), nthreads=nthreads, noah=noah)
        par, ost = O.noah_arrays(m, p, state)
        run.set_noah_lsm(par, ost, p["station"], ft, data, lai, 900.0)
        t0 = time.perf_counter()
        rc_o, _, sto = run.run(m.initial_state(), _cfg(cpu_hours * 3600.0))
        cwall = time.perf_counter() - t0
        if rc_o != 0:
            raise RuntimeError(f"Flux-PIHM host run failed: rc={rc_o}")
        cpu_rate = (cpu_hours / 24.0) / cwall
        out["cpu_baseline"] = {"value": cpu_rate, "unit": "sim-days/s", "cores": nthreads, "kind": "port",
                               "sample": f"first {cpu_hours:g} simulated hours of C2 (oracle Noah step, RHS and "
                                         f"soil-water step, same integrator, host vectors, OpenMP {nthreads} "
                                         f"threads)",
                               "wall_s": cwall, "nst": sto.nst, "gpu_same_window": (cpu_hours / 24.0) / gwall,
                               "gpu_over_cpu_same_window": ((cpu_hours / 24.0) / gwall) / cpu_rate}
    return out


def fp64_valu(path, rhs_ms):
    """Share of the SIMDs' issue cycles the RHS's FP64 VALU instructions occupy during one
    evaluation (instruction counts from the profile, time from the live measurement)."""
    cycles = sum(insts * waves * 4.0 for insts, waves in RHS_VALU[path]) / SIMDS
    return {"issue_cycles_per_simd": cycles, "eval_cycles": rhs_ms * 1e3 * CLOCK_MHZ,
            "issue_frac": cycles / (rhs_ms * 1e3 * CLOCK_MHZ),
            "insts_per_wave": [x for x, _ in RHS_VALU[path]], "waves": [w for _, w in RHS_VALU[path]]}


def cpu_run(mesh, forcing, hours, nthreads, y0=None, t0=0.0):
    """The oracle's host run (same integrator and driver, host vectors, OpenMP nthreads)
    from state y0 at time t0 over `hours`; wall = the driver's own clock."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    run = O.OracleRun(mesh, forcing, nthreads=nthreads)
    cfg = run_config(t0 + hours * 3600.0, t_start=t0, graphs=GRAPHS)
    rc, y, st = run.run(mesh.initial_state() if y0 is None else y0, cfg)
    return dict(rc=rc, wall=st.wall_s, stats=st.as_dict(), yend=y)


def gpu_window(mesh, forcing, y0, t0, hours):
    """The device run (pihm_run_gpu) from state y0 at time t0 over `hours`: the driver's clock."""
    from pihm_amd.rhs import run_gpu
    rc, _, st = run_gpu(mesh, forcing, run_config(t0 + hours * 3600.0, t_start=t0, graphs=GRAPHS), y0)
    if rc != 0:
        raise RuntimeError(f"device window run failed: rc={rc}")
    return st.wall_s


def physical_cores():
    """Distinct (package, core) pairs among the CPUs this process may run on."""
    seen = set()
    for cpu in os.sched_getaffinity(0):
        base = f"/sys/devices/system/cpu/cpu{cpu}/topology/"
        try:
            seen.add((open(base + "physical_package_id").read().strip(), open(base + "core_id").read().strip()))
        except OSError:
            return None
    return len(seen)


# FP64 VALU issue: instructions per wavefront and wavefronts per RHS evaluation at the
# 200,000-triangle workload (SQ_INSTS_VALU / SQ_WAVES, the same profile passes as the traffic).
# A wavefront's FP64 instruction occupies a SIMD for 4 cycles (16 lanes per cycle); the chip
# has 1,024 SIMDs at 2.4 GHz.
RHS_VALU = {"one launch": [(1768.7, 3382)], "two kernels": [(268.2, 3320), (1669.2, 3192)]}
SIMDS, CLOCK_MHZ = 1024, 2400.0


def profile_kernels(iters, seed):
    """Short fixed workload for kernel traces and hardware-counter passes: `iters` RHS
    evaluations at C1 on a mid-storm state on the one-launch path and `iters` on the
    two-kernel path (each followed by `iters` more timed with HIP events), then
    `iters` copies of a 64 Mi-entry vector (N_VScale(1, x, z): 512 MiB read and 512 MiB
    written per launch), the calibration kernel for the byte counters.  Between the two,
    the first simulated quarter hour at C1 with Krylov iterations replayed from HIP graphs,
    the producers at C1, the plain and _NOAH_ RHS at C2 (100,000 triangles) and the Noah
    step at C2."""
    L = _lib.lib()
    mesh = config_mesh("c1", seed=seed)
    desc = mesh.desc()
    ctx = C.c_void_p()
    assert L.pihm_ctx_create(C.byref(desc), C.byref(ctx)) == 0
    L.pihm_ctx_set_nan_check(ctx, 0)
    rng = np.random.default_rng(3)
    y0 = mesh.initial_state()
    y0[:mesh.nelem] = rng.uniform(0.0, 2e-3, mesh.nelem)        # ponded surface water
    y = L.N_VNew_Hip(mesh.neq)
    yd = L.N_VNew_Hip(mesh.neq)
    assert L.pihm_ctx_state_to_device(ctx, y0.ctypes.data_as(_lib.c_double_p), y) == 0
    # the one-launch RHS (the default), then the two-kernel RHS, each iters times and timed
    ms_paths = {}
    for path in (_lib.PIHM_RHS_ONE_LAUNCH, _lib.PIHM_RHS_TWO_KERNELS):
        assert L.pihm_ctx_set_rhs_path(ctx, path) == 0
        for _ in range(iters):
            L.PIHM_ODE_Hip(0.0, y, yd, ctx)
        L.pihm_ctx_sync(ctx)
        ms = C.c_double()
        L.pihm_ctx_time_rhs(ctx, L.N_VGetDeviceArrayPointer_Hip(y), L.N_VGetDeviceArrayPointer_Hip(yd), iters,
                            C.byref(ms))
        ms_paths[path] = ms.value
    assert L.pihm_ctx_set_rhs_path(ctx, _lib.PIHM_RHS_ONE_LAUNCH) == 0
    ms_eval = C.c_double(ms_paths[_lib.PIHM_RHS_ONE_LAUNCH])
    # the integrator with its Krylov iterations replayed from HIP graphs: the first quarter hour at C1
    from pihm_amd.rhs import run_gpu
    _log("profile: C1 first quarter hour with graph replay")
    rc_run, _, st_run = run_gpu(mesh, StormForcing(2, seed=11, nelem=mesh.nelem), run_config(900.0, graphs=1),
                                mesh.initial_state())
    _log(f"profile: run rc {rc_run}, {st_run.nst} steps")
    h, lsm, bc, _ = _producers(mesh, seed)
    ms_l, ms_b = C.c_double(), C.c_double()
    L.pihm_lsm_time(lsm.h, 3600.0 * 12.25, 900.0, h.y, iters, C.byref(ms_l))
    L.pihm_bc_time(bc.h, 3600.0 * 12.25, iters, C.byref(ms_b))
    bc.close()
    lsm.close()
    h.close()
    noah = noah_rhs_run(seed, iters, coupled=False)
    from pihm_amd.noahlsm import MAXLYR
    fm, fp, fs, fft, fdata, flai = _flux_inputs(seed, 13.0)
    ms_noah = _noah_step_ms(fm, fp, fs, fft, fdata, flai, iters)
    n = 64 << 20
    a = L.N_VNew_Hip(n)
    b = L.N_VNew_Hip(n)
    L.N_VConst_Hip(1.0, a)
    for _ in range(iters):
        L.N_VScale_Hip(1.0, a, b)
    L.pihm_ctx_sync(ctx)
    for v in (y, yd, a, b):
        L.N_VDestroy_Hip(v)
    L.pihm_ctx_destroy(ctx)
    print(json.dumps({"profile_kernels": iters, "graph_run": {"rc": rc_run, "nst": st_run.nst, "nli": st_run.nli},
                      "rhs_bytes_algorithmic": BYTES_PER_ELEM * mesh.nelem + BYTES_PER_SEG * mesh.nriver,
                      "rhs_ms_per_eval_events": ms_eval.value,
                      "rhs_two_kernel_ms_per_eval_events": ms_paths[_lib.PIHM_RHS_TWO_KERNELS],
                      "copy_bytes": 2 * n * 8,
                      "lsm_ms_per_step_events": ms_l.value, "bc_ms_per_apply_events": ms_b.value,
                      "lsm_bytes_algorithmic": LSM_BYTES_PER_ELEM * mesh.nelem,
                      "noah_rhs_c2_ms_per_eval_events": noah["ms_per_eval"],
                      "plain_rhs_c2_ms_per_eval_events": noah["plain_ms_per_eval"],
                      "noah_rhs_c2_bytes_algorithmic": noah["roofline"]["algorithmic_bytes"],
                      "noah_step_c2_ms_events": ms_noah,
                      "noah_step_c2_bytes_algorithmic": fm.nelem * noah_step_bytes_per_elem(MAXLYR)}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=12, help="timed simulated hours")
    ap.add_argument("--warmup", type=int, default=1, help="untimed simulated hours")
    ap.add_argument("--cpu-hours", type=float, default=0.25, help="CPU baseline sample, all cores (simulated hours)")
    ap.add_argument("--cpu1-hours", type=float, default=0.05, help="CPU baseline sample, one thread (simulated hours)")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--ensemble-members", type=int, default=256)
    ap.add_argument("--ensemble-hours", type=float, default=1.0)
    ap.add_argument("--ensemble-sample", type=int, default=16)
    ap.add_argument("--no-closure", action="store_true")
    ap.add_argument("--no-noah", action="store_true")
    ap.add_argument("--closure-days", type=float, default=10.0)
    ap.add_argument("--no-flux", action="store_true")
    ap.add_argument("--flux-hours", type=float, default=24.0, help="C2 window (simulated hours)")
    ap.add_argument("--flux-cpu-hours", type=float, default=0.25, help="C2 CPU baseline sample (simulated hours)")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--serve-seconds", type=float, default=30.0)
    ap.add_argument("--serve-rate", type=float, default=20.0)
    ap.add_argument("--serve-horizon-unit", type=float, default=60.0,
                    help="seconds of simulated time per horizon unit (the request spec's unit is a day)")
    ap.add_argument("--serve-check", type=int, default=3)
    ap.add_argument("--no-graphs", action="store_true",
                    help="every device run launches each Krylov iteration's kernels one by one instead of "
                         "replaying graphs")
    ap.add_argument("--profile-kernels", type=int, default=0,
                    help="only run N RHS evaluations and N copy-kernel launches (counter passes)")
    args = ap.parse_args()
    if args.profile_kernels > 0:
        profile_kernels(args.profile_kernels, args.seed)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "bench.py measures one GPU; runs over several GPUs are not provided",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    # worker processes (mesh generation) fork from a server started before any HIP call
    from pihm_amd.workloads import start_workers
    start_workers()
    mesh = config_mesh("c1", seed=args.seed)
    W, K = args.warmup, args.steps
    forcing = StormForcing(W + K + 4, seed=11, nelem=mesh.nelem)
    S = args.cpu_hours
    global GRAPHS
    GRAPHS = 0 if args.no_graphs else 1
    _log(f"device run: {W} + {K} simulated hours at {mesh.nelem} triangles")
    g = gpu_run(mesh, forcing, W + K, mark_hours=[W, W + K], snap_hour=W)
    _log("device run done")
    if g["rc"] != 0 or g["nan"] != 0:
        raise RuntimeError(f"device run failed: rc={g['rc']} nan={g['nan']} {g['stats']}")
    elapsed = g["marks"][float(W + K)] - g["marks"][float(W)]
    value = (K / 24.0) / elapsed
    ms_prep, ms_flux = g["kernel_ms"]
    rhs_ms = g["ms_eval"]
    alg_bytes = BYTES_PER_ELEM * mesh.nelem + BYTES_PER_SEG * mesh.nriver
    achieved = alg_bytes / (rhs_ms * 1e-3) / 1e9
    st = g["stats"]
    out = {
        "metric": "sim-days/s @200k elems FP64 (RHS HBM GB/s); ensemble member-days/s; p50/p99",
        "value": value,
        "unit": "sim-days/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": elapsed / K * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded 200,000-triangle Delaunay watershed, 4,095-segment network, "
                "storm forcing: 10% wet hours, Exp(5 mm/h))",
        "config": {"workload": "c1: 200,000 triangles, 4,095 river segments, 608,190 FP64 unknowns; "
                               "step = 1 simulated hour; BDF+SPGMR reltol 1e-3 abstol 1e-4",
                   "nelem": mesh.nelem, "nriver": mesh.nriver, "neq": mesh.neq},
        "solver": {k: st[k] for k in ("nst", "nfe", "nfes", "nni", "nli", "ncfn", "netf")},
        "rhs": {"ms_per_eval": rhs_ms, "evals_per_s": 1e3 / rhs_ms,
                "fp64_valu": fp64_valu(g["tuning"]["rhs_path"], rhs_ms),
                "two_kernel_ms_each_timed_alone": {"k_prep": ms_prep, "k_flux": ms_flux}, "tuning": g["tuning"]},
        "roofline": {"kernel": RHS_KERNELS[g["tuning"]["rhs_path"]],
                     "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": RHS_TRAFFIC_BYTES[g["tuning"]["rhs_path"]],
                     "algorithmic_bytes": alg_bytes},
    }
    ms_ls, ms_dot = C.c_double(), C.c_double()
    assert _lib.lib().pihm_nv_time_ops(mesh.neq, 200, C.byref(ms_ls), C.byref(ms_dot)) == 0
    out["nvector"] = {"n": mesh.neq, "linear_sum_us": ms_ls.value * 1e3,
                      "linear_sum_GBps": 24.0 * mesh.neq / (ms_ls.value * 1e-3) / 1e9,
                      "linear_sum_frac_of_hbm_peak": 24.0 * mesh.neq / (ms_ls.value * 1e-3) / 1e9 / HBM_PEAK_GBS,
                      "dot_prod_us_per_call": ms_dot.value * 1e3,
                      "note": "N_VLinearSum_Hip (24 B per entry) and N_VDotProd_Hip (both passes and the "
                              "host's wait) on C1-length vectors, HIP events over 200 calls each"}
    if not args.no_cpu:
        nthreads = args.cpu_threads or len(os.sched_getaffinity(0))
        t0 = W * 3600.0
        legs = {}
        for nth, hrs in ((nthreads, S), (1, args.cpu1_hours)):
            _log(f"CPU baseline: hours {W} to {W} + {hrs}, {nth} threads, from the device state at hour {W}")
            c = cpu_run(mesh, forcing, hrs, nth, y0=g["snap"], t0=t0)
            if c["rc"] != 0:
                raise RuntimeError(f"CPU baseline run failed: rc={c['rc']}")
            gw = gpu_window(mesh, forcing, g["snap"], t0, hrs)
            rate, grate = (hrs / 24.0) / c["wall"], (hrs / 24.0) / gw
            legs[nth] = {"value": rate, "unit": "sim-days/s", "threads": nth,
                         "sample": f"simulated hours {W} to {W + hrs:g}, host and device both restarted "
                                   f"from the device's state at hour {W} (the start of the timed window)",
                         "wall_s": c["wall"], "gpu_same_window": grate, "gpu_over_cpu_same_window": grate / rate}
        allc = legs[nthreads]
        out["cpu_baseline"] = {"value": allc["value"], "unit": "sim-days/s", "cores": nthreads,
                               "physical_cores": physical_cores(), "kind": "port",
                               "sample": allc["sample"] + f"; oracle RHS, same integrator and driver, host "
                                                          f"vectors, OpenMP {nthreads} threads",
                               "wall_s": allc["wall_s"], "gpu_same_window": allc["gpu_same_window"],
                               "gpu_over_cpu_same_window": allc["gpu_over_cpu_same_window"],
                               "one_thread": legs[1]}
    _log("producers")
    out["producers"] = producers_run(mesh, args.seed)
    if not args.no_noah:
        _log("RHS with the Noah branches at 100,000 triangles")
        out["noah_rhs"] = noah_rhs_run(args.seed)
    if not args.no_flux:
        _log(f"Flux-PIHM: {args.flux_hours} simulated hours at 100,000 triangles")
        nth = 0 if args.no_cpu else (args.cpu_threads or len(os.sched_getaffinity(0)))
        out["flux_pihm"] = flux_pihm_run(args.seed, args.flux_hours, args.flux_cpu_hours, nth)
    if not args.no_closure:
        _log(f"closure: {args.closure_days} days at 535 triangles")
        out["closure"] = closure_run(args.closure_days)
    if not args.no_ensemble:
        _log(f"ensemble: {args.ensemble_members} members")
        out["ensemble"] = ensemble_run(args.ensemble_members, args.ensemble_hours, args.ensemble_sample, 2024)
    if not args.no_serve:
        _log(f"serving: {args.serve_seconds} s of arrivals")
        out["serving"] = serve_run(args.serve_seconds, args.serve_rate, args.serve_horizon_unit,
                                   args.serve_check, 77)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
