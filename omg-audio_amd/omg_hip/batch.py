"""Batched persistent launch: many trace_world queries in one hipGraph.

A query is what one `trace_world` call takes (scene, params, sources,
listeners, overlays).  `Batch` uploads all of them once, then every
`launch()` replays one graph: clear the echogram arena, run the persistent
work-stealing tracer over all queries' rays.  Each query's echograms and
per-ray records come back laid out exactly as `trace_world` returns them;
the reference has no batch (it runs one job per `trace_world`,
crates/omg-scene/src/late.rs:81-91), so per-query results are the parity
contract.
"""
import ctypes as C
from dataclasses import dataclass

import numpy as np

from ._lib import (BATCH_COUNTERS, BATCH_RECORDS, BatchShape, Counters, Query, check, lib)
from . import TraceResult, _alloc_out, _f32, _overlays, nbins


@dataclass
class QuerySpec:
    scene: object
    params: object
    sources: np.ndarray
    listeners: np.ndarray
    overlays: tuple = ()


class Batch:
    """omg_batch_*: build once, launch many times."""

    def __init__(self, ctx, queries, records=False, counters=False):
        self.ctx = ctx
        self.specs = []
        n = len(queries)
        arr = (Query * max(1, n))()
        keep = []
        for i, q in enumerate(queries):
            if not isinstance(q, QuerySpec):
                q = QuerySpec(*q)
            src = _f32(q.sources).reshape(-1, 3)
            lis = _f32(q.listeners).reshape(-1, 3)
            ov = _overlays(list(q.overlays))
            keep += [src, lis, ov]
            arr[i].scene = q.scene.h
            arr[i].params = q.params
            arr[i].sources = src.ctypes.data
            arr[i].n_src = src.shape[0]
            arr[i].listeners = lis.ctypes.data
            arr[i].n_lis = lis.shape[0]
            arr[i].overlays = C.cast(ov, C.c_void_p).value
            arr[i].n_ovl = len(q.overlays)
            self.specs.append(QuerySpec(q.scene, q.params, src, lis, tuple(q.overlays)))
        flags = (BATCH_RECORDS if records else 0) | (BATCH_COUNTERS if counters else 0)
        h = C.c_void_p()
        check(lib().omg_batch_create(ctx.h, C.cast(arr, C.c_void_p), n, flags, C.byref(h)))
        self.h = h
        self.records = records
        self.counting = counters

    def __len__(self):
        return len(self.specs)

    def launch(self, stream=None):
        check(lib().omg_batch_launch(self.h, C.c_void_p(stream) if stream else None))

    def sync(self):
        check(lib().omg_batch_sync(self.h))

    def result(self, q):
        s = self.specs[q]
        bins, dirs, rec = _alloc_out(s.params, s.sources.shape[0], s.listeners.shape[0],
                                     self.records)
        check(lib().omg_batch_result(self.h, q, bins.ctypes.data_as(C.c_void_p),
                                     dirs.ctypes.data_as(C.c_void_p),
                                     rec.ctypes.data_as(C.c_void_p) if rec is not None else None))
        return TraceResult(bins, dirs, rec)

    def device_output(self, q):
        b, d = C.c_void_p(), C.c_void_p()
        check(lib().omg_batch_device_output(self.h, q, C.byref(b), C.byref(d)))
        return b.value, d.value

    def counters(self):
        c = Counters()
        check(lib().omg_batch_counters(self.h, C.byref(c)))
        return dict(aabb_tests=c.aabb_tests, tri_tests=c.tri_tests, bounce_rays=c.bounce_rays)

    def size(self):
        rays, items = C.c_uint64(), C.c_uint32()
        check(lib().omg_batch_size(self.h, C.byref(rays), C.byref(items)))
        return rays.value, items.value

    @property
    def shape(self):
        s = BatchShape()
        check(lib().omg_batch_get_shape(self.h, C.byref(s)))
        return dict(lanes=s.lanes, stack_lds=s.stack_lds, waves_per_simd=s.waves_per_simd,
                    regen=s.regen, width=s.width, ms=s.ms)

    def set_shape(self, lanes=1, stack_lds=16, waves_per_simd=4, regen=0, width=2):
        s = BatchShape(lanes, stack_lds, waves_per_simd, regen, width, 0.0)
        check(lib().omg_batch_set_shape(self.h, C.byref(s)))

    def tune(self, stride=16):
        """Autotune the launch shape on every `stride`-th work item; returns the choice."""
        s = BatchShape()
        check(lib().omg_batch_tune(self.h, stride, C.byref(s)))
        return dict(lanes=s.lanes, stack_lds=s.stack_lds, waves_per_simd=s.waves_per_simd,
                    regen=s.regen, width=s.width, ms=s.ms)

    def close(self):
        if getattr(self, "h", None):
            lib().omg_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


__all__ = ["Batch", "QuerySpec", "nbins"]
