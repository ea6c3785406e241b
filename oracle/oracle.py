"""ctypes front-end of the C oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/ (as the checker) and by
bench.py's cpu_baseline leg (as the timed CPU port of the reference
tracer).  Product code never imports this module.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

MAXB = 8
RNG_XORSHIFT = 0
RNG_COUNTER = 1
TERM_NAMES = ["escape", "time", "dead_transmit", "dead_reflect", "max_bounces"]


class Material(C.Structure):
    _fields_ = [("absorption", C.c_float * MAXB), ("scattering", C.c_float),
                ("transmission", C.c_float * MAXB)]


class Geom(C.Structure):
    _fields_ = [("kind", C.c_int), ("size", C.c_float * 3), ("walls", Material * 6),
                ("mesh", C.c_void_p), ("n_panels", C.c_int), ("panels", C.c_void_p),
                ("panel_mats", C.c_void_p)]


class Hit(C.Structure):
    _fields_ = [("t", C.c_float), ("normal", C.c_float * 3), ("material_ref", C.c_int),
                ("tri", C.c_uint32), ("flags", C.c_uint32)]


class TraceParams(C.Structure):
    _fields_ = [("n_rays", C.c_uint32), ("nbands", C.c_uint32), ("max_bounces", C.c_uint32),
                ("rng_mode", C.c_uint32), ("seed", C.c_uint64), ("src_index", C.c_uint32),
                ("dir_band", C.c_uint32), ("bin_dt", C.c_float), ("max_time", C.c_float),
                ("floor_rel", C.c_float), ("radius", C.c_float), ("energy", C.c_float * MAXB)]


class RayRec(C.Structure):
    _fields_ = [("iters", C.c_uint32), ("term", C.c_uint32), ("hit_hash", C.c_uint32),
                ("flags", C.c_uint32)]


class TraceJob(C.Structure):
    _fields_ = [("g", C.c_void_p), ("src", C.c_void_p), ("lis", C.c_void_p),
                ("n_lis", C.c_uint32), ("p", TraceParams), ("bins", C.c_void_p),
                ("dirs", C.c_void_p), ("stats", C.c_uint64 * 3)]


class Panel(C.Structure):
    _fields_ = [("axis", C.c_int), ("plane", C.c_float), ("lat0", C.c_float),
                ("lat1", C.c_float), ("z0", C.c_float), ("z1", C.c_float), ("glass", C.c_int),
                ("trans", C.c_float * MAXB)]


class DomeParams(C.Structure):
    _fields_ = [("n_fan", C.c_uint32), ("n_rays", C.c_uint32), ("max_events", C.c_uint32),
                ("nbands", C.c_uint32), ("horizon", C.c_float), ("through_len", C.c_float),
                ("use_project_trig", C.c_int)]


class DomeRay(C.Structure):
    _fields_ = [("escaped", C.c_uint32), ("events", C.c_uint32), ("used_through", C.c_uint32),
                ("bin", C.c_uint32), ("tp", C.c_float * MAXB)]


class Surface(C.Structure):
    _fields_ = [("n", C.c_float * 3), ("d", C.c_float), ("refl", C.c_float * MAXB),
                ("trans", C.c_float * MAXB), ("has_rect", C.c_int), ("rmin", C.c_float * 3),
                ("rmax", C.c_float * 3)]


class Aabb(C.Structure):
    _fields_ = [("min", C.c_float * 3), ("max", C.c_float * 3), ("trans", C.c_float * MAXB)]


class Record(C.Structure):
    _fields_ = [("delay", C.c_float), ("dir", C.c_float * 3), ("gains", C.c_float * MAXB)]


def build():
    """Compile liboracle.so from the C sources (make)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        vp, u32, f32 = C.c_void_p, C.c_uint32, C.c_float
        L.or_mesh_build.restype = vp
        L.or_mesh_build.argtypes = [vp, u32, vp, u32, vp, vp, vp, vp, u32, u32]
        L.or_mesh_free.argtypes = [vp]
        L.or_mesh_sizes.argtypes = [vp, vp, vp]
        L.or_mesh_export.argtypes = [vp, vp, vp]
        L.or_dedup_tris.restype = u32
        L.or_dedup_tris.argtypes = [vp, u32, vp, vp]
        L.or_mesh_raycast.argtypes = [vp, vp, vp, vp, vp, vp, vp]
        L.or_mesh_segment_hits.argtypes = [vp, vp, vp, vp, vp, vp, C.c_int, vp]
        L.or_mesh_tri_normal.argtypes = [vp, u32, vp]
        L.or_raycast_hit.restype = C.c_int
        L.or_raycast_hit.argtypes = [vp, vp, vp, vp, vp]
        L.or_nbins.restype = u32
        L.or_nbins.argtypes = [f32, f32]
        L.or_trace.argtypes = [vp, vp, vp, u32, vp, vp, vp, vp, vp, vp, vp, vp, vp]
        L.or_trace_jobs.argtypes = [vp, u32, u32]
        L.or_ema.argtypes = [vp, vp, vp, vp, u32, u32, f32]
        L.or_late_direction.argtypes = [vp, vp, u32, u32, f32, f32, vp, vp]
        L.or_estimate_reverb.argtypes = [vp, u32, u32, f32, vp, vp]
        L.or_image_source_taps.argtypes = [vp, vp, u32, vp, vp, u32, vp, vp, C.c_int]
        L.or_gate_init.argtypes = [vp, u32]
        L.or_gate_step.restype = C.c_int
        L.or_gate_step.argtypes = [vp, vp, vp, vp, u32, u32]
        L.or_late_slot_init.argtypes = [vp, vp, vp]
        L.or_late_step.argtypes = [vp, u32, u32, f32, C.c_int, vp, vp, vp, vp, vp, vp, vp, vp]
        L.or_xorshift_next.restype = C.c_uint64
        L.or_xorshift_next.argtypes = [vp]
        L.or_xorshift_f32.restype = f32
        L.or_xorshift_f32.argtypes = [vp]
        L.or_philox4x32_10.argtypes = [vp, vp, vp]
        L.or_sincosf.argtypes = [f32, vp, vp]
        L.or_dome_fan.argtypes = [u32, C.c_int, vp]
        L.or_dome_bin_of.restype = u32
        L.or_dome_bin_of.argtypes = [vp]
        L.or_dome_sample.argtypes = [vp, vp, vp, vp, C.c_int, vp, vp, vp, vp]
        L.or_surface_table.argtypes = [vp, vp, C.c_int]
        L.or_surface_append_box.argtypes = [vp, vp, vp, vp, u32]
        L.or_mesh_record.argtypes = [vp, vp, C.c_int, vp, C.c_int, vp, vp, vp, C.c_int, f32, f32,
                                     f32, vp]
        L.or_mesh_chains.argtypes = [vp, vp, C.c_int, C.c_uint16, vp, u32, u32, C.c_int, vp, vp,
                                     C.c_int]
        L.or_paths_edges.restype = C.c_int
        L.or_paths_edges.argtypes = [vp, f32, u32, vp, C.c_int]
        L.or_paths_transmission.argtypes = [vp, vp, vp, vp]
        L.or_knife_edge_amp.restype = f32
        L.or_knife_edge_amp.argtypes = [f32, f32]
        L.or_knife_edge_bands.argtypes = [f32, vp]
        L.or_bend_point.argtypes = [vp, vp, vp, vp, vp]
        L.or_paths_find.restype = C.c_int
        L.or_paths_find.argtypes = [vp, vp, C.c_int, vp, vp, vp, vp, C.c_int]
        _LIB = L
    return _LIB


def _p(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def _f32(a, shape=None):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float32))
    if shape is not None:
        a = a.reshape(shape)
    return a


def make_material(absorption, scattering, transmission):
    m = Material()
    for i, v in enumerate(absorption):
        m.absorption[i] = float(v)
    m.scattering = float(scattering)
    for i, v in enumerate(transmission):
        m.transmission[i] = float(v)
    return m


def materials_array(mats):
    """mats: list of (absorption[nb], scattering, transmission[nb])"""
    arr = (Material * max(1, len(mats)))()
    for i, (a, s, t) in enumerate(mats):
        arr[i] = make_material(a, s, t)
    return arr


# ------------------------------------------------------------------ numerics

def xorshift_stream(seed, n):
    st = C.c_uint64(max(1, seed))
    return np.array([lib().or_xorshift_next(C.byref(st)) for _ in range(n)], dtype=np.uint64)


def xorshift_f32_stream(seed, n):
    st = C.c_uint64(max(1, seed))
    return np.array([lib().or_xorshift_f32(C.byref(st)) for _ in range(n)], dtype=np.float32)


def philox(ctr, key):
    c = np.asarray(ctr, dtype=np.uint32)
    k = np.asarray(key, dtype=np.uint32)
    out = np.zeros(4, dtype=np.uint32)
    lib().or_philox4x32_10(_p(c), _p(k), _p(out))
    return out


def sincos(x):
    s, c = C.c_float(), C.c_float()
    lib().or_sincosf(float(np.float32(x)), C.byref(s), C.byref(c))
    return np.float32(s.value), np.float32(c.value)


# ------------------------------------------------------------------ mesh

def dedup_tris(tri_verts):
    tv = _f32(tri_verts).reshape(-1, 9)
    nt = tv.shape[0]
    pos = np.zeros((nt * 3, 3), dtype=np.float32)
    idx = np.zeros((nt, 3), dtype=np.uint32)
    nv = lib().or_dedup_tris(_p(tv), nt, _p(pos), _p(idx))
    return pos[:nv].copy(), idx


class Mesh:
    """The reference Mesh (mesh.rs:19-38) built by the oracle."""

    def __init__(self, positions, indices, tri_material, materials, nbands=3, tri_surface=None,
                 tri_max_edge=None):
        self.positions = _f32(positions).reshape(-1, 3)
        self.indices = np.ascontiguousarray(np.asarray(indices, dtype=np.uint32).reshape(-1, 3))
        nt = self.indices.shape[0]
        self.tri_material = np.ascontiguousarray(np.asarray(tri_material, dtype=np.uint16))
        self.tri_surface = (np.zeros(nt, np.uint16) if tri_surface is None else
                            np.ascontiguousarray(np.asarray(tri_surface, dtype=np.uint16)))
        self.tri_max_edge = (np.full(nt, 4.0, np.float32) if tri_max_edge is None else
                             _f32(tri_max_edge))
        self.materials = materials_array(materials)
        self.nbands = nbands
        self.n_materials = len(materials)
        self.ptr = lib().or_mesh_build(_p(self.positions), self.positions.shape[0],
                                       _p(self.indices), nt, _p(self.tri_material),
                                       _p(self.tri_surface), _p(self.tri_max_edge),
                                       C.cast(self.materials, C.c_void_p), len(materials), nbands)

    def __del__(self):
        if getattr(self, "ptr", None) and _LIB is not None:
            _LIB.or_mesh_free(self.ptr)
            self.ptr = None

    def sizes(self):
        n, p = C.c_uint32(), C.c_uint32()
        lib().or_mesh_sizes(self.ptr, C.byref(n), C.byref(p))
        return n.value, p.value

    def export(self):
        nn, npr = self.sizes()
        nodes = np.zeros((nn, 8), dtype=np.float32)
        prims = np.zeros((npr, 11), dtype=np.float32)
        lib().or_mesh_export(self.ptr, _p(nodes), _p(prims))
        return nodes, prims

    def raycast(self, o, d, stats=None):
        o = _f32(o)
        d = _f32(d)
        t, tri, fl = C.c_float(), C.c_uint32(), C.c_uint32()
        st = np.zeros(2, np.uint64) if stats is None else stats
        ok = lib().or_mesh_raycast(self.ptr, _p(o), _p(d), C.byref(t), C.byref(tri), _p(st),
                                   C.byref(fl))
        if not ok:
            return None
        return np.float32(t.value), tri.value, fl.value

    def segment_hits(self, a, b, cap=4096):
        a = _f32(a)
        b = _f32(b)
        ts = np.zeros(cap, np.float32)
        tris = np.zeros(cap, np.uint32)
        mats = np.zeros(cap, np.uint16)
        n = lib().or_mesh_segment_hits(self.ptr, _p(a), _p(b), _p(ts), _p(tris), _p(mats), cap,
                                       None)
        return ts[:n], tris[:n], mats[:n]

    def tri_normal(self, tri):
        n = np.zeros(3, np.float32)
        lib().or_mesh_tri_normal(self.ptr, tri, _p(n))
        return n

    def path_edges(self, budget, min_dihedral_deg=40.0):
        """AutoPaths::new's edge graph (mesh.rs:639-682): [n][7] rows a, b, importance."""
        cap = max(1, 3 * self.indices.shape[0])
        out = np.zeros((cap, 7), np.float32)
        n = lib().or_paths_edges(self.ptr, min_dihedral_deg, budget, _p(out), cap)
        return out[:n].copy()

    def path_transmission(self, a, b):
        g = np.zeros(3, np.float32)
        lib().or_paths_transmission(self.ptr, _p(_f32(a)), _p(_f32(b)), _p(g))
        return g

    def find_paths(self, edges, src, lis, edge_candidates=24, pair_edges=8, max_paths=4,
                   taps=False):
        """AutoPaths::find (paths.rs:117-332): list of dicts (key, points,
        length, gains), the direct path first."""
        e = np.ascontiguousarray(edges, np.float32).reshape(-1, 7)
        bud = PathBudget(edge_candidates, pair_edges, max_paths, 1 if taps else 0)
        cap = 1 + max_paths
        out = (FoundPath * cap)()
        n = lib().or_paths_find(self.ptr, _p(e), e.shape[0], _p(_f32(src)), _p(_f32(lis)),
                                C.byref(bud), out, cap)
        return [path_dict(out[i]) for i in range(n)]


class PathBudget(C.Structure):
    _fields_ = [("edge_candidates", C.c_uint32), ("pair_edges", C.c_uint32),
                ("max_paths", C.c_uint32), ("taps", C.c_uint32)]


class FoundPath(C.Structure):
    _fields_ = [("key", C.c_uint32), ("n_points", C.c_uint32), ("length", C.c_float),
                ("gains", C.c_float * 3), ("points", C.c_float * 75)]


def path_dict(p):
    pts = np.ctypeslib.as_array(p.points)[:3 * p.n_points].reshape(-1, 3).copy()
    return dict(key=p.key, points=pts, length=np.float32(p.length),
                gains=np.array(p.gains[:], np.float32))


def knife_edge_amp(detour, freq):
    return lib().or_knife_edge_amp(detour, freq)


def knife_edge_bands(detour):
    out = np.zeros(3, np.float32)
    lib().or_knife_edge_bands(detour, _p(out))
    return out


def bend_point(a, b, src, lis):
    p = np.zeros(3, np.float32)
    lib().or_bend_point(_p(_f32(a)), _p(_f32(b)), _p(_f32(src)), _p(_f32(lis)), _p(p))
    return p


# ------------------------------------------------------------------ geometry + tracer

def box_geom(size, walls):
    """Shoebox geometry; walls: 6 x (absorption, scattering, transmission)."""
    g = Geom()
    g.kind = 0
    for i in range(3):
        g.size[i] = float(size[i])
    for i, w in enumerate(walls):
        g.walls[i] = make_material(*w)
    g._keep = []
    return g


def mesh_geom(mesh, panels=None, panel_mats=None):
    g = Geom()
    g.kind = 1
    g.mesh = mesh.ptr
    g._keep = [mesh]
    if panels is not None and len(panels):
        pa = _f32(panels).reshape(-1, 6)
        pm = materials_array(panel_mats)
        g.n_panels = pa.shape[0]
        g.panels = _p(pa)
        g.panel_mats = C.cast(pm, C.c_void_p)
        g._keep += [pa, pm]
    return g


def nbins(max_time, bin_dt):
    return lib().or_nbins(float(np.float32(max_time)), float(np.float32(bin_dt)))


def trace_params(n_rays, nbands=3, max_bounces=64, rng_mode=RNG_COUNTER, seed=0, src_index=0,
                 dir_band=1, bin_dt=0.010, max_time=3.0, floor_rel=1e-7, radius=0.5,
                 energy=None):
    p = TraceParams()
    p.n_rays = n_rays
    p.nbands = nbands
    p.max_bounces = max_bounces
    p.rng_mode = rng_mode
    p.seed = seed
    p.src_index = src_index
    p.dir_band = dir_band
    p.bin_dt = bin_dt
    p.max_time = max_time
    p.floor_rel = floor_rel
    p.radius = radius
    e = [1.0] * nbands if energy is None else energy
    for i, v in enumerate(e):
        p.energy[i] = float(v)
    return p


def raycast_hit(geom, p, d):
    """AcousticGeometry::raycast_hit of a geometry (Shoebox, Mesh, WithPanels;
    scene.rs:41-138): dict(t, normal, panel (index or None), tri, flags) or
    None on a miss."""
    h = Hit()
    if not lib().or_raycast_hit(C.byref(geom), _p(_f32(p)), _p(_f32(d)), C.byref(h), None):
        return None
    panel = -h.material_ref - 1 if h.material_ref < 0 else None
    return dict(t=np.float32(h.t), normal=np.array(h.normal[:], np.float32), panel=panel,
                tri=h.tri, flags=h.flags)


def trace(geom, src, lis, params, rng_state=None, want_rays=False, want_hits=False,
          want_f64=False):
    """Run tracer::trace; returns dict with bins, dirs (per listener), stats, ..."""
    lis = _f32(lis).reshape(-1, 3)
    nl = lis.shape[0]
    nb = params.nbands
    n = nbins(params.max_time, params.bin_dt)
    bins = np.zeros((nl, n, nb), np.float32)
    dirs = np.zeros((nl, n, 3), np.float32)
    b64 = np.zeros((nl, n, nb), np.float64) if want_f64 else None
    d64 = np.zeros((nl, n, 3), np.float64) if want_f64 else None
    recs = np.zeros((params.n_rays, 4), np.uint32) if want_rays else None
    hits = np.zeros((params.n_rays, params.max_bounces), np.uint32) if want_hits else None
    stats = np.zeros(3, np.uint64)
    st = C.c_uint64(max(1, params.seed) if rng_state is None else rng_state)
    src = _f32(src)
    rc = lib().or_trace(C.byref(geom), _p(src), _p(lis), nl, C.byref(params), C.byref(st),
                        _p(bins), _p(dirs), _p(b64), _p(d64), _p(recs), _p(hits), _p(stats))
    assert rc == 0
    return dict(bins=bins, dirs=dirs, bins64=b64, dirs64=d64, recs=recs, hits=hits,
                stats=stats, rng_state=st.value)


def trace_jobs_threaded(geom, srcs, lis, params_list, n_threads):
    """CPU baseline: independent traces spread over a pthread pool."""
    srcs = _f32(srcs).reshape(-1, 3)
    lis = _f32(lis).reshape(-1, 3)
    nj = len(params_list)
    jobs = (TraceJob * nj)()
    keep = []
    for i, p in enumerate(params_list):
        n = nbins(p.max_time, p.bin_dt)
        b = np.zeros((lis.shape[0], n, p.nbands), np.float32)
        d = np.zeros((lis.shape[0], n, 3), np.float32)
        keep.append((b, d))
        jobs[i].g = C.cast(C.pointer(geom), C.c_void_p)
        jobs[i].src = srcs[i % srcs.shape[0]].ctypes.data
        jobs[i].lis = _p(lis)
        jobs[i].n_lis = lis.shape[0]
        jobs[i].p = p
        jobs[i].bins = _p(b)
        jobs[i].dirs = _p(d)
    lib().or_trace_jobs(C.cast(jobs, C.c_void_p), nj, n_threads)
    stats = np.array([[jobs[i].stats[k] for k in range(3)] for i in range(nj)], np.uint64)
    return keep, stats


def estimate_reverb(bins, bin_dt):
    bins = _f32(bins)
    n, nb = bins.shape
    rt = np.zeros(nb, np.float32)
    lv = np.zeros(nb, np.float32)
    lib().or_estimate_reverb(_p(bins), n, nb, float(np.float32(bin_dt)), _p(rt), _p(lv))
    return rt, lv


def late_direction(bins, dirs, bin_dt, from_s):
    bins = _f32(bins)
    dirs = _f32(dirs)
    n, nb = bins.shape
    d = np.zeros(3, np.float32)
    a = C.c_float()
    lib().or_late_direction(_p(bins), _p(dirs), n, nb, float(np.float32(bin_dt)),
                            float(np.float32(from_s)), _p(d), C.byref(a))
    return d, np.float32(a.value)


def ema(bins, dirs, obins, odirs, alpha):
    n, nb = bins.shape
    lib().or_ema(_p(bins), _p(dirs), _p(_f32(obins)), _p(_f32(odirs)), n, nb,
                 float(np.float32(alpha)))


class _Gate(C.Structure):
    _fields_ = [("last_src", C.c_float * 3), ("last_lis", C.c_float * 3),
                ("last_energy", C.c_float * MAXB), ("age", C.c_uint32)]


class TraceGate:
    """TraceGate (crates/omg-scene/src/sim.rs:38-76), or_gate_* in or_trace.c."""

    def __init__(self, phase=0):
        self.g = _Gate()
        lib().or_gate_init(C.byref(self.g), phase)

    def should_trace(self, src, lis, energy, max_age):
        e = _f32(energy)
        return bool(lib().or_gate_step(C.byref(self.g), _p(_f32(src)), _p(_f32(lis)), _p(e),
                                       e.size, max_age))


class _LateSlot(C.Structure):
    _fields_ = [("bins", C.c_void_p), ("dirs", C.c_void_p), ("version", C.c_uint32),
                ("rp_dist", C.c_float)]


class LateSlot:
    """One Sim's late state (echo_avg, version, rp_dist, sim.rs:79-112) and
    the late path of Sim::reverb_world after the trace decision
    (sim.rs:374-397), or_late_step in or_trace.c."""

    def __init__(self, nbins, nbands, bin_dt):
        self.bins = np.zeros((nbins, nbands), np.float32)
        self.dirs = np.zeros((nbins, 3), np.float32)
        self.nbins, self.nbands, self.bin_dt = nbins, nbands, float(np.float32(bin_dt))
        self.s = _LateSlot()
        lib().or_late_slot_init(C.byref(self.s), _p(self.bins), _p(self.dirs))

    def step(self, traced, cur_bins, cur_dirs, src, receiver):
        rt = np.zeros(self.nbands, np.float32)
        lv = np.zeros(self.nbands, np.float32)
        d = np.zeros(3, np.float32)
        a = C.c_float()
        cb = _f32(cur_bins) if traced else None
        cd = _f32(cur_dirs) if traced else None
        lib().or_late_step(C.byref(self.s), self.nbins, self.nbands, self.bin_dt, int(traced),
                           _p(cb), _p(cd), _p(_f32(src)), _p(_f32(receiver)), _p(rt), _p(lv),
                           _p(d), C.byref(a))
        return dict(rt60=rt, level=lv, dir=d, aniso=np.float32(a.value),
                    version=int(self.s.version))


AIR_ABSORPTION_PER_M = [1.0e-4, 5.0e-4, 3.0e-3]   # material.rs:75


def image_source_taps(size, walls, src, lis, max_order, nbands=3, air=None):
    walls_arr = (Material * 6)(*[make_material(*w) for w in walls])
    air = _f32(AIR_ABSORPTION_PER_M if air is None else air)
    cap = 4096
    out = np.zeros((cap, 5 + nbands), np.float32)
    n = lib().or_image_source_taps(_p(_f32(size)), C.cast(walls_arr, C.c_void_p), nbands,
                                   _p(_f32(src)), _p(_f32(lis)), max_order, _p(air), _p(out), cap)
    return out[:n].copy()


# ------------------------------------------------------------------ dome

def dome_fan(n_fan, project_trig=True):
    d = np.zeros((n_fan, 3), np.float32)
    lib().or_dome_fan(n_fan, int(project_trig), _p(d))
    return d


def dome_bin_of(d):
    return lib().or_dome_bin_of(_p(_f32(d)))


def make_panels(panels, nbands=3):
    """panels: list of dict(axis, plane, lat, z, glass, trans)"""
    arr = (Panel * max(1, len(panels)))()
    for i, p in enumerate(panels):
        arr[i].axis = p["axis"]
        arr[i].plane = p["plane"]
        arr[i].lat0, arr[i].lat1 = p["lat"]
        arr[i].z0, arr[i].z1 = p["z"]
        arr[i].glass = int(p["glass"])
        for b, v in enumerate(p["trans"]):
            arr[i].trans[b] = v
    return arr


def dome_sample(mesh, eye, n_fan=384, n_rays=None, max_events=6, horizon=80.0, through_len=50.0,
                panels=(), project_trig=True, want_rays=False):
    nb = mesh.nbands
    P = DomeParams(n_fan, n_fan if n_rays is None else n_rays, max_events, nb, horizon,
                   through_len, int(project_trig))
    pa = make_panels(list(panels), nb)
    e = np.zeros((9, nb), np.float32)
    ds = np.zeros((9, 3), np.float32)
    rays = (DomeRay * P.n_rays)() if want_rays else None
    stats = np.zeros(2, np.uint64)
    lib().or_dome_sample(mesh.ptr, C.byref(P), _p(_f32(eye)), C.cast(pa, C.c_void_p),
                         len(panels), _p(e), _p(ds),
                         C.cast(rays, C.c_void_p) if rays is not None else None, _p(stats))
    out = dict(e=e, dsum=ds, stats=stats)
    if want_rays:
        out["escaped"] = np.array([r.escaped for r in rays], np.uint32)
        out["events"] = np.array([r.events for r in rays], np.uint32)
        out["through"] = np.array([r.used_through for r in rays], np.uint32)
        out["bin"] = np.array([r.bin for r in rays], np.uint32)
        out["tp"] = np.array([[r.tp[b] for b in range(nb)] for r in rays], np.float32)
    return out


# ------------------------------------------------------------------ pt_mesh

def surface_table(mesh, cap=65536):
    arr = (Surface * cap)()
    n = lib().or_surface_table(mesh.ptr, C.cast(arr, C.c_void_p), cap)
    return arr, n


def mesh_record(mesh, table, n_surf, chain, order, src, lis, extras=(), air=None):
    air = AIR_ABSORPTION_PER_M if air is None else air
    ch = np.asarray(list(chain) + [0xFFFF] * (3 - len(chain)), np.uint16)
    ex = (Aabb * max(1, len(extras)))()
    for i, (mn, mx, tr) in enumerate(extras):
        for k in range(3):
            ex[i].min[k] = mn[k]
            ex[i].max[k] = mx[k]
        for k, v in enumerate(tr):
            ex[i].trans[k] = v
    rec = Record()
    ok = lib().or_mesh_record(mesh.ptr, C.cast(table, C.c_void_p), n_surf, _p(ch), order,
                              _p(_f32(src)), _p(_f32(lis)), C.cast(ex, C.c_void_p), len(extras),
                              air[0], air[1], air[2], C.byref(rec))
    if not ok:
        return None
    return dict(delay=np.float32(rec.delay), dir=np.array(rec.dir[:], np.float32),
                gains=np.array(rec.gains[:mesh.nbands], np.float32))


def mesh_chains(mesh, lis, n_rays, rot, project_trig=True, cap=1 << 16, boxes=(), base=0):
    """mesh_chains (pt_mesh.rs:375-457); boxes: overlay (min, max) pairs whose
    faces reflect too, ids base + box * 6 + face."""
    ch = np.zeros((cap, 3), np.uint16)
    od = np.zeros(cap, np.uint8)
    bx = _f32([list(mn) + list(mx) for mn, mx in boxes]) if len(boxes) else None
    n = lib().or_mesh_chains(mesh.ptr, _p(bx), len(boxes), base, _p(_f32(lis)), n_rays, rot,
                             int(project_trig), _p(ch), _p(od), cap)
    return ch[:n].copy(), od[:n].copy()
