"""Late-field state on the device: the Sim-side consumer of the tracer.

Mirrors the late path of ``Sim::reverb_world`` (crates/omg-scene/src/
sim.rs:349-397): ``TraceGate`` decides whether a slot is re-traced this
tick (sim.rs:38-76), a fresh echogram is folded into the slot's running
average (``Echogram::ema``, crates/omg-core/src/tracer.rs:46-57, alpha 1.0
on the first update and 0.3 after), and the average yields the reverb
parameters (``estimate_reverb``, tracer.rs:205-257), the late direction
and anisotropy (``late_direction(0.08)``, tracer.rs:62-77) and the level
scaled for the source-receiver distance change since the last trace
(sim.rs:382-395).  The averages of all slots stay resident in HBM; one
``update`` call serves any number of slots.
"""
import ctypes as C
from dataclasses import dataclass

import numpy as np

from ._lib import LateOut, LateTrace, LateUpdate, TraceGateState, check, lib

# TraceGate ages by quality tier (crates/omg-scene/src/quality.rs:133-139)
GATE_MAX_AGE = {"high": 8, "med": 12, "low": 16}


def _v3(a):
    a = np.asarray(a, dtype=np.float32).reshape(3)
    return (C.c_float * 3)(*[float(x) for x in a])


def _ptr(x):
    """A device address from an int or a tensor with data_ptr()."""
    if x is None:
        return None
    return C.c_void_p(int(x.data_ptr()) if hasattr(x, "data_ptr") else int(x))


class TraceGate:
    """``TraceGate`` (sim.rs:38-76): host logic, no device needed."""

    def __init__(self, phase=0):
        self.state = TraceGateState()
        lib().omg_trace_gate_init(C.byref(self.state), phase)

    def should_trace(self, src, lis, energy, max_age=GATE_MAX_AGE["high"]):
        e = np.ascontiguousarray(np.asarray(energy, dtype=np.float32))
        fire = C.c_int()
        check(lib().omg_trace_gate_step(C.byref(self.state), _v3(src), _v3(lis),
                                        e.ctypes.data_as(C.c_void_p), e.size, max_age,
                                        C.byref(fire)))
        return bool(fire.value)


@dataclass
class LateParams:
    rt60: np.ndarray      # ReverbParams.rt60 (params.rs:19-29)
    level: np.ndarray     # ReverbParams.level after the distance compensation
    dir: np.ndarray       # late_direction(0.08)
    aniso: float
    version: int
    traced: bool


class LateField:
    """n_slots late fields of [nbins][nbands] bins + [nbins][3] dirs on the device."""

    def __init__(self, ctx, n_slots, nbands=3, bin_dt=0.001, max_time=2.0):
        from . import nbins as _nbins
        self.ctx = ctx
        self.n_slots, self.nbands = n_slots, nbands
        self.bin_dt, self.max_time = float(np.float32(bin_dt)), float(np.float32(max_time))
        self.nbins = _nbins(max_time, bin_dt)
        h = C.c_void_p()
        check(lib().omg_late_field_create(ctx.h, n_slots, self.nbins, nbands, self.bin_dt,
                                          C.byref(h)))
        self.h = h
        # per-slot trace seeds, advanced as the reference's GPU backends do (lib.rs:869)
        self._seed = [(0x5EED0000 + s) & 0xFFFFFFFF for s in range(n_slots)]

    def gate(self, slot, src, lis, energy=None, max_age=GATE_MAX_AGE["high"]):
        e = np.ones(self.nbands, np.float32) if energy is None else \
            np.ascontiguousarray(np.asarray(energy, dtype=np.float32))
        fire = C.c_int()
        check(lib().omg_late_field_gate(self.h, slot, _v3(src), _v3(lis),
                                        e.ctypes.data_as(C.c_void_p), max_age, C.byref(fire)))
        return bool(fire.value)

    def _outs(self, out, n):
        nb = self.nbands
        return [LateParams(np.array(o.rt60[:nb], np.float32), np.array(o.level[:nb], np.float32),
                           np.array(o.dir[:], np.float32), np.float32(o.aniso), int(o.version),
                           bool(o.traced)) for o in out[:n]]

    def update(self, updates, stream=None):
        """updates: dicts with slot, traced, bins / dirs (device addresses or
        tensors, complete on `stream`), src, receiver.  Returns LateParams per update."""
        ups = list(updates)
        arr = (LateUpdate * max(1, len(ups)))()
        for i, u in enumerate(ups):
            arr[i].slot = u["slot"]
            arr[i].traced = 1 if u.get("traced") else 0
            arr[i].cur_bins = _ptr(u.get("bins")) if u.get("traced") else None
            arr[i].cur_dirs = _ptr(u.get("dirs")) if u.get("traced") else None
            arr[i].src = _v3(u["src"])
            arr[i].receiver = _v3(u["receiver"])
        out = (LateOut * max(1, len(ups)))()
        check(lib().omg_late_field_update(self.h, arr, len(ups), out,
                                          C.c_void_p(stream) if stream else None))
        return self._outs(out, len(ups))

    def trace_world(self, scene, jobs, overlays=()):
        """jobs: dicts with slot, params (TraceParams of this field's echogram
        shape), src, receiver.  Traces each on the device into the field's
        staging and folds it in (traced updates).  Returns LateParams per job."""
        from . import _overlays
        jobs = list(jobs)
        arr = (LateTrace * max(1, len(jobs)))()
        for i, j in enumerate(jobs):
            arr[i].slot = j["slot"]
            arr[i].params = j["params"]
            arr[i].src = _v3(j["src"])
            arr[i].receiver = _v3(j["receiver"])
        ov = _overlays(list(overlays))
        out = (LateOut * max(1, len(jobs)))()
        check(lib().omg_late_field_trace_world(self.h, scene.h, arr, len(jobs),
                                               C.cast(ov, C.c_void_p), len(overlays), out))
        return self._outs(out, len(jobs))

    def read(self, slot):
        bins = np.zeros((self.nbins, self.nbands), np.float32)
        dirs = np.zeros((self.nbins, 3), np.float32)
        check(lib().omg_late_field_read(self.h, slot, bins.ctypes.data_as(C.c_void_p),
                                        dirs.ctypes.data_as(C.c_void_p)))
        return bins, dirs

    def next_seed(self, slot):
        self._seed[slot] = (self._seed[slot] * 747796405 + 2891336453) & 0xFFFFFFFF
        return self._seed[slot]

    def reverb_world(self, slot, scene, src, receiver, rays, panels=(), budget=None,
                     max_age=GATE_MAX_AGE["high"]):
        """``Sim::reverb_world`` (sim.rs:349-397) for one slot: gate within the
        caller's per-tick budget (a one-element list, decremented per trace),
        trace the world on the device with the slot's next seed, fold in.
        Returns (LateParams, the seed traced with or None)."""
        from . import trace_params
        if (budget is None or budget[0] > 0) and self.gate(slot, src, receiver, max_age=max_age):
            if budget is not None:
                budget[0] -= 1
            seed = self.next_seed(slot)
            p = trace_params(rays, self.nbands, seed=seed, bin_dt=self.bin_dt,
                             max_time=self.max_time)
            out = self.trace_world(scene, [dict(slot=slot, params=p, src=src, receiver=receiver)],
                                   overlays=panels)
            return out[0], seed
        out = self.update([dict(slot=slot, traced=False, src=src, receiver=receiver)])
        return out[0], None

    def close(self):
        if getattr(self, "h", None):
            lib().omg_late_field_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
