"""Image sources on the BVH: WorldDiscovery / WorldSolve
(crates/omg-scene/src/early_world.rs:20-62) over the C ABI."""
import ctypes as C

import numpy as np

from ._lib import Blocker, MeshRecord, Surface, check, lib

AIR_ABSORPTION_PER_M = [1.0e-4, 5.0e-4, 3.0e-3]   # material.rs:75
NO_SURF = 0xFFFF                                   # pt_mesh.rs:25


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def surface_table(scene, cap=1 << 16):
    """SurfaceTable::build (pt_mesh.rs:51-80) -> ctypes array, count."""
    arr = (Surface * cap)()
    n = C.c_uint32()
    check(lib().omg_surface_table(scene.h, C.cast(arr, C.c_void_p), cap, C.byref(n)))
    return arr, n.value


def append_box(table, n_surf, mn, mx, material, nbands=3):
    """SurfaceTable::append_box (pt_mesh.rs:87-106): writes the box's six faces
    at table[n_surf:n_surf + 6]; returns the new count (the faces' base id is
    n_surf).  material: a MaterialDef."""
    from . import _materials
    if n_surf + 6 > len(table):
        raise ValueError("surface table too small")
    mat = _materials([material])
    check(lib().omg_surface_append_box(np.asarray(mn, np.float32).ctypes.data_as(C.c_void_p),
                                       np.asarray(mx, np.float32).ctypes.data_as(C.c_void_p),
                                       C.cast(mat, C.c_void_p), nbands,
                                       C.cast(C.byref(table, n_surf * C.sizeof(Surface)),
                                              C.c_void_p)))
    return n_surf + 6


class GpuWorldDiscovery:
    """WorldDiscovery::discover (early_world.rs:24-26): rotating golden fan."""

    def __init__(self, scene, n_rays=4096):
        self.scene = scene
        self.n_rays = n_rays

    def discover(self, listener, rot, cap=1 << 16, boxes=(), base=0):
        """Chains found by the fan from `listener`; boxes: overlay (min, max)
        pairs whose faces reflect too (ids base + box * 6 + face)."""
        ch = np.zeros((cap, 3), np.uint16)
        od = np.zeros(cap, np.uint8)
        n = C.c_uint32()
        lis = np.asarray(listener, np.float32)
        bx = (np.ascontiguousarray([list(mn) + list(mx) for mn, mx in boxes], np.float32)
              if len(boxes) else None)
        check(lib().omg_discover(self.scene.h, _p(bx), len(boxes), base, _p(lis), self.n_rays, rot,
                                 _p(ch), _p(od), cap, C.byref(n)))
        return ch[:n.value].copy(), od[:n.value].copy()


class GpuWorldSolve:
    """WorldSolve::solve_batch (early_world.rs:38-52): out[si * C + ci]."""

    def __init__(self, scene, surfaces=None, n_surf=None, air=None):
        self.scene = scene
        if surfaces is None:
            surfaces, n_surf = surface_table(scene)
        self.surfaces = surfaces
        self.n_surf = n_surf
        air = AIR_ABSORPTION_PER_M if air is None else air
        nb = scene.nbands
        self.air = np.asarray([air[min(b, len(air) - 1)] for b in range(nb)], np.float32)

    def solve_batch(self, sources, chains, orders, listener, extras=()):
        src = np.ascontiguousarray(np.asarray(sources, np.float32).reshape(-1, 3))
        ch = np.ascontiguousarray(np.asarray(chains, np.uint16).reshape(-1, 3))
        od = np.ascontiguousarray(np.asarray(orders, np.uint8))
        ex = (Blocker * max(1, len(extras)))()
        for i, (mn, mx, tr) in enumerate(extras):
            for k in range(3):
                ex[i].min[k] = mn[k]
                ex[i].max[k] = mx[k]
            for k, v in enumerate(tr):
                ex[i].trans[k] = v
        n = src.shape[0] * ch.shape[0]
        recs = (MeshRecord * max(1, n))()
        lis = np.asarray(listener, np.float32)
        check(lib().omg_solve_batch(self.scene.h, C.cast(self.surfaces, C.c_void_p), self.n_surf,
                                    _p(src), src.shape[0], _p(ch), _p(od), ch.shape[0], _p(lis),
                                    C.cast(ex, C.c_void_p), len(extras), _p(self.air),
                                    C.cast(recs, C.c_void_p)))
        raw = np.frombuffer(recs, dtype=np.uint32).reshape(max(1, n), -1)[:n]
        f = raw.view(np.float32)
        nb = self.scene.nbands
        return dict(valid=raw[:, 0].astype(bool), delay=f[:, 1].copy(), dir=f[:, 2:5].copy(),
                    gains=f[:, 5:5 + nb].copy())
