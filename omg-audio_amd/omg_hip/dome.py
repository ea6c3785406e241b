"""Ambient dome: DomeProbe (crates/omg-scene/src/dome.rs:293-478) on the GPU."""
import ctypes as C

import numpy as np

from ._lib import DOME_BINS, Counters, DomePanel, DomeParams, DomeRay, check, lib

# walkthrough.rs:310-326: amplitude transmission of panes and closed leaves
GLASS_TRANSMISSION = [0.50, 0.32, 0.20]
HEAVY_GLASS_TRANSMISSION = [0.18, 0.05, 0.012]
DOOR_PANEL_TRANSMISSION = [0.28, 0.12, 0.05]
HEAVY_PANEL_TRANSMISSION = [0.10, 0.03, 0.01]


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def dome_params(nbands=3, n_fan=384, n_rays=None, max_events=6, horizon=80.0, through_len=50.0):
    """Defaults are dome.rs:34-36 and :392; the sky-sphere radius is `horizon`."""
    return DomeParams(n_fan, n_fan if n_rays is None else n_rays, max_events, nbands, horizon,
                      through_len)


def make_panels(panels):
    """panels: dicts with axis, plane, lat (lo, hi), z (lo, hi), glass, trans[nb]."""
    arr = (DomePanel * max(1, len(panels)))()
    for i, p in enumerate(panels):
        arr[i].axis = p["axis"]
        arr[i].plane = p["plane"]
        arr[i].lat0, arr[i].lat1 = p["lat"]
        arr[i].z0, arr[i].z1 = p["z"]
        arr[i].glass = int(p["glass"])
        for b, v in enumerate(p["trans"]):
            arr[i].trans[b] = v
    return arr


def dome_sample(scene, eyes, params=None, panels=(), rays=False, counters=False):
    """One DomeProbe::sample pass without the EMA for every eye position.

    Returns dict(energy [n][9][nb], dsum [n][9][3], and per-ray results)."""
    params = params or dome_params(scene.nbands)
    e = np.ascontiguousarray(np.asarray(eyes, np.float32).reshape(-1, 3))
    n = e.shape[0]
    nb = params.nbands
    energy = np.zeros((n, DOME_BINS, nb), np.float32)
    dsum = np.zeros((n, DOME_BINS, 3), np.float32)
    pa = make_panels(list(panels))
    ray_arr = (DomeRay * (n * params.n_rays))() if rays else None
    cnt = Counters()
    check(lib().omg_dome_sample(scene.h, C.byref(params), _p(e), n, C.cast(pa, C.c_void_p),
                                len(panels), _p(energy), _p(dsum),
                                C.cast(ray_arr, C.c_void_p) if rays else None,
                                C.byref(cnt) if counters else None))
    out = dict(energy=energy, dsum=dsum)
    if rays:
        raw = np.frombuffer(ray_arr, dtype=np.uint32).reshape(n * params.n_rays, -1)
        out["escaped"] = raw[:, 0].reshape(n, -1)
        out["events"] = raw[:, 1].reshape(n, -1)
        out["through"] = raw[:, 2].reshape(n, -1)
        out["bin"] = raw[:, 3].reshape(n, -1)
        f = raw.view(np.float32)
        out["tp"] = f[:, 4:4 + nb].reshape(n, -1, nb)
        out["dir"] = f[:, 12:15].reshape(n, -1, 3)
    if counters:
        out["counters"] = dict(aabb_tests=cnt.aabb_tests, tri_tests=cnt.tri_tests)
    return out


def dome_sample_async(scene, d_eyes, n_eyes, d_energy, d_dsum, d_rays, params=None, panels=(),
                      stream=None):
    """Device-resident dome pass: pointers are device addresses (ints), d_rays
    is scratch for n_eyes * n_rays ray records; enqueued on `stream`."""
    params = params or dome_params(scene.nbands)
    pa = make_panels(list(panels))
    check(lib().omg_dome_sample_async(scene.h, C.byref(params), C.c_void_p(d_eyes), n_eyes,
                                      C.cast(pa, C.c_void_p), len(panels), C.c_void_p(d_energy),
                                      C.c_void_p(d_dsum), C.c_void_p(d_rays),
                                      C.c_void_p(stream) if stream else None))


def dome_status(scene):
    """Raises OmgError(ERR_OVERFLOW) if an async dome pass overflowed a crossing list."""
    check(lib().omg_dome_status(scene.h))


class DomeProbe:
    """Stateful probe: the 0.35 EMA across calls (dome.rs:460-477)."""

    def __init__(self, scene, params=None):
        self.params = params or dome_params(scene.nbands)
        self.nb = self.params.nbands
        h = C.c_void_p()
        check(lib().omg_dome_probe_create(scene.h, C.byref(self.params), C.byref(h)))
        self.h = h
        self.scene = scene

    def sample(self, eye, leaves=()):
        """-> [9] rows of (dir[3], energy[nb])"""
        out = np.zeros((DOME_BINS, 3 + self.nb), np.float32)
        pa = make_panels(list(leaves))
        e = np.asarray(eye, np.float32)
        check(lib().omg_dome_probe_sample(self.h, _p(e), C.cast(pa, C.c_void_p), len(leaves),
                                          _p(out)))
        return out

    def close(self):
        if getattr(self, "h", None):
            lib().omg_dome_probe_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
