"""Flat wire formats of the reference's web bridge and ParamBlock.

The reference queues traces as flat f32 jobs for its WebGPU driver and
takes echograms back as flat f32 buffers (crates/omg-web/src/lib.rs:151-161,
:367-378; web/gpu.js:395-440), and moves ParamBlock between threads as a
flat f32 layout (crates/omg-core/src/params.rs:62-143).  These helpers run
such job buffers on the device and read / write those layouts.
"""
import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from ._lib import (WIRE_BOX_JOB_F32S, WIRE_BOX_JOB_VERSION, WIRE_ECHO_BINS, WIRE_ECHO_F32S,
                   WIRE_PARAM_HEADER, WIRE_PARAM_PER_TAP, WIRE_WORLD_JOB_F32S,
                   WIRE_WORLD_JOB_VERSION, ParamBlock, Tap, WireWorldJob, check, lib)


def _f32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float32)).reshape(-1)


def _split(echo, n):
    """Injection layout -> (bins [n][300][3], dirs [n][300][3])."""
    e = echo[:n * WIRE_ECHO_F32S].reshape(n, 2, WIRE_ECHO_BINS, 3)
    return e[:, 0].copy(), e[:, 1].copy()


def trace_box_jobs(ctx, jobs, version=WIRE_BOX_JOB_VERSION):
    """Flat box jobs (57 f32 each) -> (ids, flat echograms [n][1800])."""
    j = _f32(jobs)
    n = j.size // WIRE_BOX_JOB_F32S
    echo = np.zeros(max(1, n) * WIRE_ECHO_F32S, np.float32)
    ids = np.zeros(max(1, n), np.uint32)
    got = C.c_uint32()
    check(lib().omg_wire_trace_box_jobs(ctx.h, j.ctypes.data_as(C.c_void_p), j.size, version,
                                        echo.ctypes.data_as(C.c_void_p),
                                        ids.ctypes.data_as(C.c_void_p), C.byref(got)))
    return ids[:got.value], echo[:got.value * WIRE_ECHO_F32S].reshape(-1, WIRE_ECHO_F32S)


def trace_world_jobs(scene, jobs, version=WIRE_WORLD_JOB_VERSION):
    """Flat world jobs (1034 f32 each) -> (ids, flat echograms [n][1800])."""
    j = _f32(jobs)
    n = j.size // WIRE_WORLD_JOB_F32S
    echo = np.zeros(max(1, n) * WIRE_ECHO_F32S, np.float32)
    ids = np.zeros(max(1, n), np.uint32)
    got = C.c_uint32()
    check(lib().omg_wire_trace_world_jobs(scene.h, j.ctypes.data_as(C.c_void_p), j.size, version,
                                          echo.ctypes.data_as(C.c_void_p),
                                          ids.ctypes.data_as(C.c_void_p), C.byref(got)))
    return ids[:got.value], echo[:got.value * WIRE_ECHO_F32S].reshape(-1, WIRE_ECHO_F32S)


def decode_world_job(job, version=WIRE_WORLD_JOB_VERSION):
    j = _f32(job)
    out = WireWorldJob()
    check(lib().omg_wire_decode_world_job(j.ctypes.data_as(C.c_void_p), version, C.byref(out)))
    return out


def split_echo(flat):
    """(bins [300][3], dirs [300][3]) of one injection-layout echogram."""
    b, d = _split(_f32(flat), 1)
    return b[0], d[0]


@dataclass
class Params:
    """ParamBlock (params.rs:62-70) with taps as (key, delay_s, dir[3], gains[3])."""
    version: int = 0
    rt60: tuple = (0.5, 0.5, 0.5)
    level: tuple = (0.0, 0.0, 0.0)
    remote: dict = None
    taps: list = field(default_factory=list)


def param_block_write_flat(p):
    pb = ParamBlock()
    pb.version = int(p.version)
    pb.rt60[:] = [float(x) for x in p.rt60]
    pb.level[:] = [float(x) for x in p.level]
    if p.remote:
        pb.has_remote = 1
        pb.remote_rt60[:] = [float(x) for x in p.remote["rt60"]]
        pb.remote_send[:] = [float(x) for x in p.remote["send"]]
        pb.remote_sh[:] = [float(x) for x in p.remote["sh"]]
    taps = (Tap * max(1, len(p.taps)))()
    for i, (key, delay, d, g) in enumerate(p.taps):
        taps[i].key = key
        taps[i].delay_s = delay
        taps[i].dir[:] = [float(x) for x in d]
        taps[i].gains[:] = [float(x) for x in g]
    n = WIRE_PARAM_HEADER + WIRE_PARAM_PER_TAP * len(p.taps)
    out = np.zeros(n, np.float32)
    got = C.c_uint32()
    check(lib().omg_param_block_write_flat(C.byref(pb), taps, len(p.taps),
                                           out.ctypes.data_as(C.c_void_p), n, C.byref(got)))
    return out


def param_block_read_flat(data):
    d = _f32(data)
    pb = ParamBlock()
    nt = C.c_uint32()
    cap = max(0, (d.size - WIRE_PARAM_HEADER) // WIRE_PARAM_PER_TAP)
    taps = (Tap * max(1, cap))()
    check(lib().omg_param_block_read_flat(d.ctypes.data_as(C.c_void_p), d.size, C.byref(pb), taps,
                                          cap, C.byref(nt)))
    remote = None
    if pb.has_remote:
        remote = dict(rt60=list(pb.remote_rt60), send=list(pb.remote_send), sh=list(pb.remote_sh))
    return Params(int(pb.version), tuple(pb.rt60), tuple(pb.level), remote,
                  [(t.key, t.delay_s, tuple(t.dir), tuple(t.gains)) for t in taps[:nt.value]])
