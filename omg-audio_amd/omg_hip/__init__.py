"""omg_hip: MI355X stochastic energy path tracer behind omg-audio's tracer seams.

The host-side mirror of the reference interface over the C ABI
(include/omg_hip.h).  Names and argument meanings follow the reference:

- ``GpuLateBackend.trace``  = ``LateBackend::trace`` (crates/omg-scene/src/late.rs:28-40)
- ``GpuWorldLateBackend.trace_world`` = ``WorldLateBackend::trace_world`` (late.rs:81-91)
- ``DomeProbe.sample`` = ``DomeProbe::sample`` (crates/omg-scene/src/dome.rs:425-478)
- ``GpuWorldDiscovery.discover`` / ``GpuWorldSolve.solve_batch`` =
  ``WorldDiscovery`` / ``WorldSolve`` (crates/omg-scene/src/early_world.rs:24-52)
- ``estimate_reverb`` / ``late_direction`` (crates/omg-core/src/tracer.rs:205-257, :62-77)
- ``LateField.reverb_world`` / ``TraceGate`` = the late path of ``Sim::reverb_world``
  (crates/omg-scene/src/sim.rs:38-76, :349-397), averages resident on the device
- ``wire``: the web bridge's flat job / echogram layouts and ParamBlock::write_flat
  (crates/omg-web/src/lib.rs:151-161, :367-378; crates/omg-core/src/params.rs:62-143)

Like the reference backends, a refused call raises ``OmgError`` (the
reference returns ``false``/``None``); construction raises when no GPU is
usable.  Nothing here computes on the CPU in place of the GPU.
"""
import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from ._lib import (DOME_BINS, LAYOUT_VERSION, MAX_BANDS, MAX_LISTENERS, MAX_OVERLAYS, Blocker,
                   Counters, DomePanel, DomeParams, DomeRay, MeshRecord, OmgError, Overlay,
                   RayRecord, SceneInfo, SceneTree, Surface, TraceParams, check, lib)
from . import _lib

__all__ = ["Context", "Scene", "MaterialDef", "trace_params", "trace_world", "trace_box",
           "estimate_reverb", "late_direction", "GpuLateBackend", "GpuWorldLateBackend",
           "ism_box", "OmgError", "Echogram", "nbins"]


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def dedup_soup(tri_verts):
    """MeshBuilder's 0.1 mm vertex dedup (omg_dedup_soup): positions [nv][3], indices [nt][3]."""
    tv = np.ascontiguousarray(np.asarray(tri_verts, np.float32).reshape(-1, 9))
    nt = tv.shape[0]
    pos = np.zeros((max(nt, 1) * 3, 3), np.float32)
    idx = np.zeros((nt, 3), np.uint32)
    nv = C.c_uint32()
    check(lib().omg_dedup_soup(_p(tv), nt, _p(pos), _p(idx), C.byref(nv)))
    return pos[:nv.value].copy(), idx


def _f32(a, shape=None):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float32))
    return a.reshape(shape) if shape is not None else a


@dataclass
class MaterialDef:
    """Per-band surface acoustics (crates/omg-core/src/material.rs:7-17)."""
    absorption: list
    scattering: float
    transmission: list

    def to_c(self):
        m = _lib.Material()
        for i, v in enumerate(self.absorption):
            m.absorption[i] = float(v)
        m.scattering = float(self.scattering)
        for i, v in enumerate(self.transmission):
            m.transmission[i] = float(v)
        return m


# material.rs:19-54
CONCRETE = MaterialDef([0.01, 0.02, 0.04], 0.10, [0.10, 0.014, 0.0015])
DRYWALL = MaterialDef([0.20, 0.08, 0.05], 0.10, [0.20, 0.045, 0.0150])
CARPET = MaterialDef([0.05, 0.30, 0.60], 0.40, [0.30, 0.100, 0.0300])
WOOD_PANEL = MaterialDef([0.20, 0.10, 0.08], 0.15, [0.22, 0.060, 0.0200])
ACOUSTIC_TILE = MaterialDef([0.30, 0.70, 0.80], 0.30, [0.30, 0.100, 0.0300])
BRICK = MaterialDef([0.02, 0.03, 0.04], 0.12, [0.045, 0.008, 0.0012])
GRASS = MaterialDef([0.11, 0.26, 0.60], 0.50, [0.10, 0.020, 0.0050])


def _materials(mats):
    arr = (_lib.Material * max(1, len(mats)))()
    for i, m in enumerate(mats):
        arr[i] = m.to_c() if isinstance(m, MaterialDef) else MaterialDef(*m).to_c()
    return arr


def nbins(max_time, bin_dt):
    """tracer.rs:14: (MAX_TIME / BIN_DT) as usize, in f32."""
    q = np.float32(max_time) / np.float32(bin_dt)
    return int(q) if q > 0 else 0


class Context:
    """A device context (GpuTracer::new, omg-gpu/src/lib.rs:33-47)."""

    def __init__(self, device=0):
        h = C.c_void_p()
        L = lib()
        _lib.check_single_runtime()     # torch may have been imported since the load
        check(L.omg_init(device, LAYOUT_VERSION, C.byref(h)))
        self.h = h
        self.device = device
        buf = C.create_string_buffer(128)
        check(lib().omg_device_name(self.h, buf, 128))
        self.device_name = buf.value.decode()

    def set_route_rays(self, min_rays=1 << 20):
        """Plain launches of at least min_rays rays run on the persistent
        regenerating kernel (0 = never; omg_set_route_rays)."""
        check(lib().omg_set_route_rays(self.h, min_rays))

    def set_handoff(self, bounce=4, node_cache16=32, waves_per_simd=6):
        """Quad-per-ray plain launches continue the rays still alive after
        `bounce` bounces on a 16-lane row each (0 = never), in a persistent
        kernel with node_cache16 top 16-wide nodes in LDS per block, sized
        for waves_per_simd resident waves per SIMD (omg_set_handoff)."""
        check(lib().omg_set_handoff(self.h, bounce, node_cache16, waves_per_simd))

    def set_handoff_lanes(self, lanes=0):
        """Lanes per ray of the hand-off kernel: 16 (one 16-wide node slot
        per lane), 8 (two slots per lane) or 0 (8 from 262144 rays, else
        16; omg_set_handoff_lanes)."""
        check(lib().omg_set_handoff_lanes(self.h, lanes))

    def set_launch_config(self, node_cache=512, stack_lds=16, lanes=0, width=4):
        """Tracer launch shape: top BVH nodes staged in LDS (64 B units), LDS
        stack depth, lanes per ray (1, 4, or 0 = chosen by launch size), BVH
        width (2 = the reference's binary tree, 4 = its 4-wide collapse)."""
        check(lib().omg_set_launch_config(self.h, node_cache, stack_lds, lanes, width))

    def close(self):
        if getattr(self, "h", None):
            lib().omg_shutdown(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Scene:
    """A mesh uploaded once with its BVH (GpuMeshTracer::new, omg-gpu/src/lib.rs:274-287)."""

    def __init__(self, ctx, h, nbands, n_materials):
        self.ctx = ctx
        self.h = h
        self.nbands = nbands
        self.n_materials = n_materials

    @classmethod
    def from_mesh(cls, ctx, positions, indices, tri_material, materials, nbands=3,
                  tri_surface=None, tri_max_edge=None, device=False):
        """device=True builds the tessellation and both trees on the GPU
        (omg_scene_create_device): the same scene, and refittable."""
        pos = _f32(positions).reshape(-1, 3)
        idx = np.ascontiguousarray(np.asarray(indices, np.uint32).reshape(-1, 3))
        nt = idx.shape[0]
        tm = np.ascontiguousarray(np.asarray(tri_material, np.uint16))
        ts = None if tri_surface is None else np.ascontiguousarray(np.asarray(tri_surface, np.uint16))
        te = None if tri_max_edge is None else _f32(tri_max_edge)
        mats = _materials(materials)
        h = C.c_void_p()
        fn = lib().omg_scene_create_device if device else lib().omg_scene_create
        check(fn(ctx.h, _p(pos), pos.shape[0], _p(idx), nt, _p(tm), _p(ts), _p(te),
                 C.cast(mats, C.c_void_p), len(materials), nbands, C.byref(h)))
        return cls(ctx, h, nbands, len(materials))

    @classmethod
    def from_soup(cls, ctx, tri_verts, tri_material, materials, nbands=3, tri_surface=None,
                  tri_max_edge=None, device=False):
        tv = _f32(tri_verts).reshape(-1, 9)
        nt = tv.shape[0]
        if device:
            pos, idx = dedup_soup(tv)
            return cls.from_mesh(ctx, pos, idx, tri_material, materials, nbands, tri_surface,
                                 tri_max_edge, device=True)
        tm = np.ascontiguousarray(np.asarray(tri_material, np.uint16))
        ts = None if tri_surface is None else np.ascontiguousarray(np.asarray(tri_surface, np.uint16))
        te = None if tri_max_edge is None else _f32(tri_max_edge)
        mats = _materials(materials)
        h = C.c_void_p()
        check(lib().omg_scene_create_soup(ctx.h, _p(tv), nt, _p(tm), _p(ts), _p(te),
                                          C.cast(mats, C.c_void_p), len(materials), nbands,
                                          C.byref(h)))
        return cls(ctx, h, nbands, len(materials))

    @classmethod
    def from_builder(cls, ctx, mb, nbands=3, device=False):
        soup = mb.soup()
        return cls.from_soup(ctx, soup["tri_verts"], soup["tri_material"], soup["materials"],
                             nbands, soup["tri_surface"], soup["tri_max_edge"], device=device)

    def export(self):
        """The resident trees: prims [n][12] and nodes [n][16] as f32 rows
        (omg_bvh_build_host's layout), nodes4 [n][32], and the tree summary."""
        t = SceneTree()
        check(lib().omg_scene_export(self.h, None, 0, None, 0, None, 0, C.byref(t)))
        prims = np.zeros((t.n_prims, 12), np.float32)
        nodes = np.zeros((t.n_nodes, 16), np.float32)
        nodes4 = np.zeros((t.n_nodes4, 32), np.float32)
        check(lib().omg_scene_export(self.h, _p(prims), t.n_prims, _p(nodes), t.n_nodes,
                                     _p(nodes4), t.n_nodes4, C.byref(t)))
        summary = {k: getattr(t, k) for k, _ in SceneTree._fields_ if k not in ("pad", "pad16")}
        summary["root_min"] = list(t.root_min)
        summary["root_max"] = list(t.root_max)
        return prims, nodes, nodes4, summary

    def refit(self, positions):
        """Moves the vertices of a scene built with device=True (omg_scene_refit)."""
        pos = _f32(positions).reshape(-1, 3)
        check(lib().omg_scene_refit(self.h, _p(pos), pos.shape[0]))

    def info(self):
        i = SceneInfo()
        check(lib().omg_scene_info_get(self.h, C.byref(i)))
        return dict(n_tris=i.n_tris, n_prims=i.n_prims, n_nodes=i.n_nodes,
                    max_depth=i.max_depth, device_bytes=i.device_bytes)

    def close(self):
        if getattr(self, "h", None):
            lib().omg_scene_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def trace_params(n_rays, nbands=3, max_bounces=64, seed=0, dir_band=1, bin_dt=0.010,
                 max_time=3.0, floor_rel=1e-7, radius=0.5, energy=None):
    """Job parameters; defaults are the reference's constants (tracer.rs:12-16, :103, :163)."""
    p = TraceParams()
    p.n_rays = n_rays
    p.nbands = nbands
    p.max_bounces = max_bounces
    p.dir_band = dir_band
    p.seed = seed
    p.bin_dt = bin_dt
    p.max_time = max_time
    p.floor_rel = floor_rel
    p.radius = radius
    for b, v in enumerate([1.0] * nbands if energy is None else energy):
        p.energy[b] = float(v)
    return p


def _overlays(panels):
    arr = (Overlay * max(1, len(panels)))()
    for i, (mn, mx, m) in enumerate(panels):
        for k in range(3):
            arr[i].min[k] = mn[k]
            arr[i].max[k] = mx[k]
        arr[i].mat = m.to_c() if isinstance(m, MaterialDef) else MaterialDef(*m).to_c()
    return arr


@dataclass
class TraceResult:
    bins: np.ndarray      # [n_src][n_lis][nbins][nbands]
    dirs: np.ndarray      # [n_src][n_lis][nbins][3]
    records: np.ndarray = None   # [n_src * n_rays][4] u32: iters, term, hit_hash, flags
    counters: dict = field(default_factory=dict)


def _alloc_out(params, n_src, n_lis, want_records):
    n = nbins(params.max_time, params.bin_dt)
    bins = np.zeros((n_src, n_lis, n, params.nbands), np.float32)
    dirs = np.zeros((n_src, n_lis, n, 3), np.float32)
    rec = np.zeros((n_src * params.n_rays, 4), np.uint32) if want_records else None
    return bins, dirs, rec


def trace_world(scene, params, sources, listeners, overlays=(), records=False, counters=False):
    """omg_trace_world: every source's rays scored against every listener."""
    src = _f32(sources).reshape(-1, 3)
    lis = _f32(listeners).reshape(-1, 3)
    bins, dirs, rec = _alloc_out(params, src.shape[0], lis.shape[0], records)
    ov = _overlays(list(overlays))
    cnt = Counters()
    check(lib().omg_trace_world(scene.h, C.byref(params), _p(src), src.shape[0], _p(lis),
                                lis.shape[0], C.cast(ov, C.c_void_p), len(overlays), _p(bins),
                                _p(dirs), _p(rec), C.byref(cnt) if counters else None))
    c = dict(aabb_tests=cnt.aabb_tests, tri_tests=cnt.tri_tests,
             bounce_rays=cnt.bounce_rays) if counters else {}
    return TraceResult(bins, dirs, rec, c)


def trace_world_hits(scene, params, sources, listeners, overlays=()):
    """omg_trace_world_hits: also returns every ray's per-bounce hit ids and distances."""
    src = _f32(sources).reshape(-1, 3)
    lis = _f32(listeners).reshape(-1, 3)
    bins, dirs, rec = _alloc_out(params, src.shape[0], lis.shape[0], True)
    n = src.shape[0] * params.n_rays
    ids = np.zeros((n, params.max_bounces), np.uint32)
    ts = np.zeros((n, params.max_bounces), np.float32)
    ov = _overlays(list(overlays))
    check(lib().omg_trace_world_hits(scene.h, C.byref(params), _p(src), src.shape[0], _p(lis),
                                     lis.shape[0], C.cast(ov, C.c_void_p), len(overlays),
                                     _p(bins), _p(dirs), _p(rec), _p(ids), _p(ts)))
    return TraceResult(bins, dirs, rec), ids, ts


def trace_box(ctx, size, walls, params, sources, listeners, records=False, counters=False):
    """omg_trace_box: the analytic Shoebox (crates/omg-core/src/scene.rs:25-82)."""
    src = _f32(sources).reshape(-1, 3)
    lis = _f32(listeners).reshape(-1, 3)
    bins, dirs, rec = _alloc_out(params, src.shape[0], lis.shape[0], records)
    w = _materials(walls)
    sz = _f32(size)
    cnt = Counters()
    check(lib().omg_trace_box(ctx.h, _p(sz), C.cast(w, C.c_void_p), params.nbands,
                              C.byref(params), _p(src), src.shape[0], _p(lis), lis.shape[0],
                              _p(bins), _p(dirs), _p(rec), C.byref(cnt) if counters else None))
    c = dict(aabb_tests=cnt.aabb_tests, tri_tests=cnt.tri_tests,
             bounce_rays=cnt.bounce_rays) if counters else {}
    return TraceResult(bins, dirs, rec, c)


def trace_world_async(scene, params, d_sources, n_src, d_listeners, n_lis, d_bins, d_dirs,
                      stream=None, overlays=(), d_records=None):
    """Device-resident launch: pointers are device addresses (ints), stream a hipStream_t."""
    ov = _overlays(list(overlays))
    check(lib().omg_trace_world_async(scene.h, C.byref(params), C.c_void_p(d_sources), n_src,
                                      C.c_void_p(d_listeners), n_lis, C.cast(ov, C.c_void_p),
                                      len(overlays), C.c_void_p(d_bins), C.c_void_p(d_dirs),
                                      C.c_void_p(d_records) if d_records else None,
                                      C.c_void_p(stream) if stream else None))


def trace_world_async_noalloc(scene, params, d_sources, n_src, d_listeners, n_lis, d_bins,
                              d_dirs, stream=None, overlays=(), d_records=None):
    """trace_world_async with no allocation in the call (a fixed device
    array holds its scratch): at most 262144 rays and 4096 materials."""
    ov = _overlays(list(overlays))
    check(lib().omg_trace_world_async_noalloc(scene.h, C.byref(params), C.c_void_p(d_sources),
                                              n_src, C.c_void_p(d_listeners), n_lis,
                                              C.cast(ov, C.c_void_p), len(overlays),
                                              C.c_void_p(d_bins), C.c_void_p(d_dirs),
                                              C.c_void_p(d_records) if d_records else None,
                                              C.c_void_p(stream) if stream else None))


def estimate_reverb(bins, bin_dt):
    """tracer.rs:205-257 -> (rt60[nb], level[nb])."""
    b = _f32(bins)
    n, nb = b.shape
    rt = np.zeros(nb, np.float32)
    lv = np.zeros(nb, np.float32)
    check(lib().omg_estimate_reverb(_p(b), n, nb, float(np.float32(bin_dt)), _p(rt), _p(lv)))
    return rt, lv


def late_direction(bins, dirs, bin_dt, from_s):
    """tracer.rs:62-77 -> (unit direction, anisotropy)."""
    b = _f32(bins)
    d = _f32(dirs)
    n, nb = b.shape
    out = np.zeros(3, np.float32)
    a = C.c_float()
    check(lib().omg_late_direction(_p(b), _p(d), n, nb, float(np.float32(bin_dt)),
                                   float(np.float32(from_s)), _p(out), C.byref(a)))
    return out, np.float32(a.value)


class Echogram:
    """The reference Echogram (tracer.rs:18-78): bins [nbins][nb], dirs [nbins][3]."""

    def __init__(self, n_bins=300, nbands=3, bin_dt=0.010):
        self.bins = np.zeros((n_bins, nbands), np.float32)
        self.dirs = np.zeros((n_bins, 3), np.float32)
        self.bin_dt = bin_dt

    def ema(self, other, alpha):
        a = np.float32(alpha)
        self.bins += a * (other.bins - self.bins)
        self.dirs += a * (other.dirs - self.dirs)

    def late_direction(self, from_s):
        return late_direction(self.bins, self.dirs, self.bin_dt, from_s)


class GpuLateBackend:
    """LateBackend (late.rs:28-49) over the analytic Shoebox.

    Seeds belong to the backend (omg-gpu/src/lib.rs:1060-1064): the rng
    argument of the reference trait is accepted and ignored.
    """

    def __init__(self, ctx, nbands=3, bin_dt=0.010, max_time=3.0, seed=0xD15EA5E):
        self.ctx = ctx
        self.nbands = nbands
        self.bin_dt = bin_dt
        self.max_time = max_time
        self.seed = seed

    def trace(self, id, room_size, walls, src, lis, n_rays, energy, rng, out):
        self.seed = (self.seed * 747796405 + 2891336453) & 0xFFFFFFFF
        p = trace_params(n_rays, self.nbands, seed=self.seed, bin_dt=self.bin_dt,
                         max_time=self.max_time, energy=energy)
        r = trace_box(self.ctx, room_size, walls, p, [src], [lis])
        out.bins[:] = r.bins[0, 0]
        out.dirs[:] = r.dirs[0, 0]
        return True

    def poll_into(self, id, out):
        return False


class GpuWorldLateBackend:
    """WorldLateBackend (late.rs:81-91) over a resident Scene (omg-gpu/src/lib.rs:845-873).

    Panels are the WithPanels overlays (min, max, material); more than
    MAX_OVERLAYS is refused (the reference truncates silently, lib.rs:691).
    """

    def __init__(self, scene, ray_mult=1, ray_cap=None, seed=0x5EEDC6D1, bin_dt=0.010,
                 max_time=3.0):
        self.scene = scene
        self.ray_mult = ray_mult
        self.ray_cap = ray_cap
        self.seed = seed
        self.bin_dt = bin_dt
        self.max_time = max_time

    def trace_world(self, id, src, lis, rays, panels, out):
        self.seed = (self.seed * 747796405 + 2891336453) & 0xFFFFFFFF
        n = rays * self.ray_mult
        if self.ray_cap:
            n = min(n, self.ray_cap)
        p = trace_params(n, self.scene.nbands, seed=self.seed, bin_dt=self.bin_dt,
                         max_time=self.max_time)
        r = trace_world(self.scene, p, [src], [lis], overlays=panels)
        out.bins[:] = r.bins[0, 0]
        out.dirs[:] = r.dirs[0, 0]
        return True


def ism_box(size, walls, src, lis, max_order=3, nbands=3):
    """ism::image_source_taps (crates/omg-core/src/ism.rs:15-84) -> rows (key, delay, dir, gains)."""
    w = _materials(walls)
    cap = 4096
    out = np.zeros((cap, 5 + nbands), np.float32)
    n = C.c_uint32()
    check(lib().omg_ism_box(_p(_f32(size)), C.cast(w, C.c_void_p), nbands, _p(_f32(src)),
                            _p(_f32(lis)), max_order, _p(out), cap, C.byref(n)))
    return out[:n.value].copy()



from .dome import DomeProbe, dome_params, dome_sample, dome_sample_async, dome_status  # noqa: E402
from .early import GpuWorldDiscovery, GpuWorldSolve, append_box, surface_table  # noqa: E402

__all__ += ["DomeProbe", "dome_params", "dome_sample", "dome_sample_async", "dome_status", "GpuWorldDiscovery", "GpuWorldSolve", "append_box",
            "surface_table"]
from .batch import Batch, QuerySpec  # noqa: E402

__all__ += ["Batch", "QuerySpec"]
from .serve import Server  # noqa: E402
from .late import GATE_MAX_AGE, LateField, LateParams, TraceGate  # noqa: E402
from . import wire  # noqa: E402,F401

__all__ += ["Server", "LateField", "LateParams", "TraceGate", "GATE_MAX_AGE", "wire"]
from .paths import AutoPaths, PathBudget, extract_edges, knife_edge_amp, bend_point  # noqa: E402

__all__ += ["AutoPaths", "PathBudget", "extract_edges", "knife_edge_amp", "bend_point"]
