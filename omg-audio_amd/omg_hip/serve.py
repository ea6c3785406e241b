"""Serving loop: resident scenes, continuous batching, HBM-budget admission.

`Server` wraps omg_server_*: a native scheduler thread batches the
requests that are waiting whenever one of its stream slots is free, so
requests arriving while others run overlap with them on the device.  A
request is one listener set's late-field trace (the reference's
`WorldLateBackend::trace_world`, crates/omg-scene/src/late.rs:81-91) plus
optionally its dome fan (`DomeProbe::sample`, crates/omg-scene/src/dome.rs:425-478).
"""
import ctypes as C

import numpy as np

from ._lib import (REQ_KEEP, REQ_RECORDS, BatchShape, Request, RequestStats, ServerConfig,
                   ServerStats, check, lib)
from . import _alloc_out, _f32, nbins
from .dome import dome_params


def _request(scene, params, sources, listeners, dome, keep, records):
    src = _f32(sources).reshape(-1, 3)
    lis = _f32(listeners).reshape(-1, 3)
    r = Request()
    r.scene = scene.h
    r.params = params
    r.sources = src.ctypes.data
    r.n_src = src.shape[0]
    r.listeners = lis.ctypes.data
    r.n_lis = lis.shape[0]
    r.dome = 1 if dome else 0
    r.flags = (REQ_KEEP if keep else 0) | (REQ_RECORDS if records else 0)
    return r, (src, lis)


class Server:
    """omg_server_*: create over resident scenes, submit, collect, replay, close."""

    def __init__(self, ctx, scenes, hbm_budget=64 << 30, slots=4, max_batch=64,
                 shape=None, dome=None, latency_rays=0):
        self.ctx = ctx
        self.scenes = list(scenes)
        cfg = ServerConfig()
        cfg.hbm_budget = int(hbm_budget)
        cfg.slots = slots
        cfg.max_batch = max_batch
        cfg.latency_rays = latency_rays
        sh = dict(lanes=1, stack_lds=12, waves_per_simd=6, regen=2, width=4)
        sh.update(shape or {})
        cfg.shape = BatchShape(sh["lanes"], sh["stack_lds"], sh["waves_per_simd"], sh["regen"],
                               sh["width"], 0.0)
        cfg.dome = dome if dome is not None else dome_params(3, n_fan=512, horizon=500.0)
        arr = (C.c_void_p * max(1, len(self.scenes)))(*[s.h.value if hasattr(s.h, "value") else s.h
                                                         for s in self.scenes])
        h = C.c_void_p()
        check(lib().omg_server_create(ctx.h, C.byref(cfg), C.cast(arr, C.c_void_p),
                                      len(self.scenes), C.byref(h)))
        self.h = h
        self.cfg = cfg
        self._specs = {}

    def submit(self, scene, params, sources, listeners, dome=True, keep=False, records=False):
        r, keepalive = _request(scene, params, sources, listeners, dome, keep, records)
        t = C.c_uint64()
        check(lib().omg_server_submit(self.h, C.byref(r), C.byref(t)))
        self._specs[t.value] = (params, keepalive[0].shape[0], keepalive[1].shape[0], dome, keep,
                                records, scene)
        return t.value

    def _collect(self, fn, ticket):
        st = RequestStats()
        if ticket not in self._specs:
            # unknown or already collected: the library reports which
            check(fn(self.h, ticket, None, None, None, None, None, C.byref(st)))
        params, ns, nl, dome, keep, records, scene = self._specs[ticket]
        out = {}
        if keep:
            bins, dirs, rec = _alloc_out(params, ns, nl, records)
            nb = params.nbands
            de = np.zeros((nl, 9, nb), np.float32) if dome else None
            dd = np.zeros((nl, 9, 3), np.float32) if dome else None
            check(fn(self.h, ticket, bins.ctypes.data_as(C.c_void_p),
                     dirs.ctypes.data_as(C.c_void_p),
                     rec.ctypes.data_as(C.c_void_p) if rec is not None else None,
                     de.ctypes.data_as(C.c_void_p) if de is not None else None,
                     dd.ctypes.data_as(C.c_void_p) if dd is not None else None, C.byref(st)))
            out.update(bins=bins, dirs=dirs, records=rec, dome_energy=de, dome_dsum=dd)
        else:
            check(fn(self.h, ticket, None, None, None, None, None, C.byref(st)))
        del self._specs[ticket]
        out.update(queue_s=st.queue_s, latency_s=st.latency_s, batch_size=st.batch_size,
                   slot=st.slot)
        return out

    def result(self, ticket):
        """Results of a finished request (raises OmgError ERR_NOT_READY while it runs)."""
        return self._collect(lib().omg_server_result, ticket)

    def wait(self, ticket):
        return self._collect(lib().omg_server_wait, ticket)

    def replay(self, requests, arrivals, queue=False):
        """Open-loop load: requests[i] = dict(scene, params, sources, listeners, dome=True,
        keep=False, records=False) submitted at arrivals[i] s.  Returns (latency_s, tickets),
        or (latency_s, queue_s, tickets) with queue=True; a refused request has latency -1
        and ticket 0."""
        n = len(requests)
        arr = (Request * max(1, n))()
        keepalive = []
        for i, q in enumerate(requests):
            r, k = _request(q["scene"], q["params"], q["sources"], q["listeners"],
                            q.get("dome", True), q.get("keep", False), q.get("records", False))
            arr[i] = r
            keepalive.append(k)
        t = np.ascontiguousarray(arrivals, np.float64)
        lat = np.zeros(n, np.float64)
        qs = np.zeros(n, np.float64)
        tickets = np.zeros(n, np.uint64)
        check(lib().omg_server_replay(self.h, C.cast(arr, C.c_void_p),
                                      t.ctypes.data_as(C.c_void_p), n,
                                      lat.ctypes.data_as(C.c_void_p),
                                      qs.ctypes.data_as(C.c_void_p),
                                      tickets.ctypes.data_as(C.c_void_p)))
        for i, q in enumerate(requests):
            if q.get("keep", False) and tickets[i]:
                self._specs[int(tickets[i])] = (q["params"], keepalive[i][0].shape[0],
                                                keepalive[i][1].shape[0], q.get("dome", True),
                                                True, q.get("records", False), q["scene"])
        return (lat, qs, tickets) if queue else (lat, tickets)

    def stats(self):
        s = ServerStats()
        check(lib().omg_server_stats_get(self.h, C.byref(s)))
        return {k: getattr(s, k) for k, _ in ServerStats._fields_}

    def close(self):
        if getattr(self, "h", None):
            lib().omg_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


__all__ = ["Server", "nbins"]
