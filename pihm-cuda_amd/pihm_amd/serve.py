"""Continuous-batching serving of watershed solves on one device.

A ``SlotEngine`` is an ensemble over a slot context (``pihm_ctx_create_slots`` +
``pihm_ens_create_slots``): a fixed set of slots in size classes, each able to hold one
solve up to its capacity, allocated once.  ``Scheduler`` admits requests as they arrive:

* admission is by footprint: a request needs a free slot whose element, river and
  forcing-record capacity covers it, and takes the smallest such slot;
* requests wait in arrival order, and a later, smaller request that fits a free slot is
  admitted past a head that does not fit (back-fill);
* every admitted solve runs beside the ones already on the device (their vector
  operations and RHS evaluations share launches); a finished solve's result is read back
  and its slot is reused for the next request without reallocation.

Latency of a request = completion time - arrival time on the scheduler's clock.
"""
import ctypes as C
import math
from collections import deque

import numpy as np

from . import _lib

PIHM_ENS_EMPTY, PIHM_ENS_READY, PIHM_ENS_RUNNING, PIHM_ENS_DONE, PIHM_ENS_FAILED = 0, 1, 2, 3, -1

# bytes of device memory per state entry (20 integrator vectors) and, roughly, per
# element and river segment of the mesh arrays (used to report the footprint)
BYTES_PER_ENTRY = 20 * 8
BYTES_PER_ELEM_MESH = 8 * (22 + 5 * 3 + 4) + 4 * (2 * 3 + 2) + 64 + 64
BYTES_PER_SEG_MESH = 8 * (19 + 11 + 4 + 2) + 4 * 8


def slot_bytes(ne, nr):
    return BYTES_PER_ENTRY * (3 * ne + 2 * nr) + BYTES_PER_ELEM_MESH * ne + BYTES_PER_SEG_MESH * nr


class SlotEngine:
    """Device side: slots of the given capacities [(ne_cap, nr_cap, nrec_cap), ...]."""

    def __init__(self, caps, cfg, nstation_cap=1, surf_mode=2, riv_mode=2):
        self.L = _lib.lib()
        self.caps = [tuple(int(v) for v in c) for c in caps]
        n = len(self.caps)
        ne = np.array([c[0] for c in self.caps], np.int32)
        nr = np.array([c[1] for c in self.caps], np.int32)
        nrec = np.array([c[2] for c in self.caps], np.int32)
        nst = np.full(n, nstation_cap, np.int32)
        self.ctx = C.c_void_p()
        rc = self.L.pihm_ctx_create_slots(n, ne.ctypes.data_as(_lib.c_int32_p), nr.ctypes.data_as(_lib.c_int32_p),
                                          surf_mode, riv_mode, C.byref(self.ctx))
        if rc:
            raise RuntimeError(f"pihm_ctx_create_slots failed ({rc})")
        self.cfg = cfg
        self.ens = C.c_void_p()
        rc = self.L.pihm_ens_create_slots(self.ctx, nrec.ctypes.data_as(_lib.c_int32_p),
                                          nst.ctypes.data_as(_lib.c_int32_p), C.byref(cfg), C.byref(self.ens))
        if rc:
            self.L.pihm_ctx_destroy(self.ctx)
            raise RuntimeError(f"pihm_ens_create_slots failed ({rc})")
        self._keep = [None] * n
        self.nslots = n

    def fits(self, k, mesh, forcing):
        ne, nr, nrec = self.caps[k]
        return mesh.nelem <= ne and mesh.nriver <= nr and forcing.nrec <= nrec

    def load(self, k, mesh, forcing, y0, t_end):
        desc = mesh.desc()
        fs = forcing.for_mesh(mesh).struct()
        y0 = np.ascontiguousarray(y0, dtype=np.float64)
        rc = self.L.pihm_ens_load_member(self.ens, k, C.byref(desc), C.byref(fs), y0.ctypes.data_as(_lib.c_double_p),
                                         float(t_end))
        if rc:
            raise RuntimeError(f"pihm_ens_load_member({k}) failed ({rc})")
        self._keep[k] = (desc, fs, forcing, mesh)

    def advance(self, max_ticks):
        nf = C.c_int()
        r = self.L.pihm_ens_advance(self.ens, int(max_ticks), C.byref(nf))
        if r < 0:
            raise RuntimeError(f"pihm_ens_advance failed ({r})")
        return r

    def status(self, k):
        return self.L.pihm_ens_member_status(self.ens, k)

    def result(self, k):
        mesh = self._keep[k][3]
        y = np.empty(mesh.neq)
        rc = self.L.pihm_ens_get_state(self.ens, k, y.ctypes.data_as(_lib.c_double_p))
        if rc:
            raise RuntimeError(f"pihm_ens_get_state({k}) failed ({rc})")
        st = _lib.RunStats()
        self.L.pihm_ens_member_stats(self.ens, k, C.byref(st))
        return y, st

    def counters(self):
        a, b, c = C.c_long(), C.c_long(), C.c_long()
        self.L.pihm_ens_counters(self.ens, C.byref(a), C.byref(b), C.byref(c))
        return dict(ticks=a.value, op_launches=b.value, rhs_launches=c.value)

    def close(self):
        if getattr(self, "ens", None):
            self.L.pihm_ens_destroy(self.ens)
            self.ens = None
        if getattr(self, "ctx", None):
            self.L.pihm_ctx_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Request:
    """One solve: arrival time, horizon (solver seconds) and a way to get its payload
    (mesh, forcing, initial state), which may still be in preparation at arrival."""

    def __init__(self, rid, arrival, t_end, payload):
        self.rid, self.arrival, self.t_end = rid, float(arrival), float(t_end)
        self._payload = payload
        self.slot = None
        self.t_admit = self.t_done = None
        self.result = None
        self.stats = None
        self.failed = False

    def payload(self, block=False):
        """(mesh, forcing, y0), or None when not ready and block is False."""
        p = self._payload
        if callable(getattr(p, "done", None)):      # a future
            if not block and not p.done():
                return None
            return p.result()
        return p


class Scheduler:
    """Admission, back-fill and retirement on top of an engine with the SlotEngine
    interface (fits / load / advance / status / result)."""

    def __init__(self, engine, clock, ticks_per_round=8):
        self.eng = engine
        self.clock = clock
        self.ticks = ticks_per_round
        self.free = set(range(engine.nslots))
        self.running = {}
        self.done = []

    def _slot_for(self, mesh, forcing):
        best = None
        for k in self.free:
            if self.eng.fits(k, mesh, forcing):
                ne = self.eng.caps[k][0]
                if best is None or ne < self.eng.caps[best][0] or (ne == self.eng.caps[best][0] and k < best):
                    best = k
        return best

    def admit(self, queue, now):
        """Admit every queued request that fits a free slot, oldest first."""
        kept = deque()
        while queue:
            r = queue.popleft()
            p = r.payload()
            k = None if p is None else self._slot_for(p[0], p[1])
            if k is None:
                kept.append(r)
                continue
            mesh, forcing, y0 = p
            self.eng.load(k, mesh, forcing, y0, r.t_end)
            self.free.discard(k)
            r.slot = k
            r.t_admit = now
            self.running[k] = r
        queue.extend(kept)

    def step(self):
        """One round of device work; retires the members that finished."""
        if not self.running:
            return 0
        n = self.eng.advance(self.ticks)
        if n == 0:
            return 0
        now = self.clock()
        for k in list(self.running):
            s = self.eng.status(k)
            if s in (PIHM_ENS_DONE, PIHM_ENS_FAILED):
                r = self.running.pop(k)
                r.t_done = now
                r.failed = s == PIHM_ENS_FAILED
                r.result, r.stats = self.eng.result(k)
                self.free.add(k)
                self.done.append(r)
        return n

    def run(self, requests, sleep=None):
        """Serve `requests` (sorted by arrival) to completion; returns them."""
        pending = deque(requests)
        queue = deque()
        while pending or queue or self.running:
            now = self.clock()
            while pending and pending[0].arrival <= now:
                queue.append(pending.popleft())
            if queue and self.free:
                self.admit(queue, now)
            if self.running:
                self.step()
            elif pending and not queue:
                if sleep is not None:
                    sleep(max(0.0, pending[0].arrival - self.clock()))
            elif queue and not self.running and not self.free:
                raise RuntimeError("no slot can ever take the queued requests")
            elif queue and not self.running:
                # waiting for a payload still in preparation, or nothing fits
                if all(r.payload() is not None and self._slot_for(*r.payload()[:2]) is None for r in queue) and \
                        len(self.free) == self.eng.nslots:
                    raise RuntimeError("a queued request fits no slot")
                if sleep is not None:
                    sleep(0.001)
        return requests


def latency_summary(reqs):
    lat = np.array([r.t_done - r.arrival for r in reqs if r.t_done is not None])
    if len(lat) == 0:
        return {}
    return {"n": int(len(lat)), "p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
            "mean_s": float(lat.mean()), "max_s": float(lat.max())}
