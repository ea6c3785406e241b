"""Seeded procedural scenes for the tracer workloads.

``MeshBuilder`` mirrors the reference's authoring API (material, surface,
tri, quad, abox, coarse; crates/omg-core/src/mesh.rs:70-165) and emits a
triangle soup; the library applies the 0.1 mm vertex dedup and builds the
BVH.  All randomness is SplitMix64 from an explicit seed, so every scene
is reproducible bit for bit.
"""
import math

import numpy as np

from . import MaterialDef

MASK64 = (1 << 64) - 1


class SplitMix64:
    """SplitMix64 (Steele, Lea, Flood 2014): the scene-synthesis stream."""

    def __init__(self, seed):
        self.state = seed & MASK64

    def next_u64(self):
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def uniform(self, lo=0.0, hi=1.0):
        return lo + (hi - lo) * ((self.next_u64() >> 11) * (1.0 / (1 << 53)))

    def log_uniform(self, lo, hi):
        return math.exp(self.uniform(math.log(lo), math.log(hi)))

    def randint(self, lo, hi):
        """Uniform integer in [lo, hi]."""
        return lo + self.next_u64() % (hi - lo + 1)


class Xorshift64Star:
    """Rng (crates/omg-core/src/rng.rs:2-22): xorshift64* (Vigna 2014) with
    next_f32 = top 24 bits / 2^24, an exact f32."""

    def __init__(self, seed):
        self.state = max(seed & MASK64, 1)

    def next_u64(self):
        x = self.state
        x ^= x >> 12
        x ^= (x << 25) & MASK64
        x ^= x >> 27
        self.state = x
        return (x * 0x2545F4914F6CDD1D) & MASK64

    def next_f32(self):
        return np.float32(self.next_u64() >> 40) / np.float32(1 << 24)


class MeshBuilder:
    """Triangle-soup builder with the reference's authoring calls (mesh.rs:88-153)."""

    def __init__(self):
        self.verts = []
        self.tri_material = []
        self.tri_surface = []
        self.tri_max_edge = []
        self.materials = []
        self.next_surface = 0
        self.cur_max_edge = 0.0

    def material(self, m):
        self.materials.append(m)
        return len(self.materials) - 1

    def begin_surface(self):
        self.next_surface += 1
        return self.next_surface - 1

    def coarse(self, max_edge):
        self.cur_max_edge = max_edge

    def tri(self, a, b, c, material):
        self.verts.extend((a, b, c))
        self.tri_material.append(material)
        self.tri_surface.append(max(self.next_surface - 1, 0))
        self.tri_max_edge.append(self.cur_max_edge if self.cur_max_edge > 0 else 4.0)

    def quad(self, a, b, c, d, material):
        self.tri(a, b, c, material)
        self.tri(a, c, d, material)

    def abox(self, mn, mx, material):
        """mesh.rs:144-153: six outward-wound faces."""
        (ax, ay, az), (bx, by, bz) = mn, mx
        self.quad((ax, ay, az), (ax, by, az), (ax, by, bz), (ax, ay, bz), material)
        self.quad((bx, ay, az), (bx, ay, bz), (bx, by, bz), (bx, by, az), material)
        self.quad((ax, ay, az), (ax, ay, bz), (bx, ay, bz), (bx, ay, az), material)
        self.quad((ax, by, az), (bx, by, az), (bx, by, bz), (ax, by, bz), material)
        self.quad((ax, ay, az), (bx, ay, az), (bx, by, az), (ax, by, az), material)
        self.quad((ax, ay, bz), (ax, by, bz), (bx, by, bz), (bx, ay, bz), material)

    @property
    def n_tris(self):
        return len(self.tri_material)

    def soup(self):
        return dict(tri_verts=np.asarray(self.verts, np.float32).reshape(-1, 9),
                    tri_material=np.asarray(self.tri_material, np.uint16),
                    tri_surface=np.asarray(self.tri_surface, np.uint16),
                    tri_max_edge=np.asarray(self.tri_max_edge, np.float32),
                    materials=list(self.materials))


def _rand_material(rng, nbands, scattering):
    absorption = [rng.uniform(0.02, 0.4) for _ in range(nbands)]
    return MaterialDef(absorption, scattering, [0.0] * nbands)


def shoebox(seed=1, size=(12.0, 8.0, 4.0), nbands=3, scattering=0.3, clearance=1.0):
    """Config: one shoebox room as 12 triangles with per-surface random absorption.

    Faces follow MeshBuilder::abox's winding (mesh.rs:147-152); each face is
    its own surface with absorption U[0.02, 0.4] per band, scattering 0.3,
    transmission 0.  Returns (builder, source, listener, wall materials in
    Shoebox order x-, x+, y-, y+, z-, z+).
    """
    rng = SplitMix64(seed)
    mats = [_rand_material(rng, nbands, scattering) for _ in range(6)]
    mb = MeshBuilder()
    ids = [mb.material(m) for m in mats]
    (ax, ay, az), (bx, by, bz) = (0.0, 0.0, 0.0), size
    faces = [
        ((ax, ay, az), (ax, by, az), (ax, by, bz), (ax, ay, bz)),
        ((bx, ay, az), (bx, ay, bz), (bx, by, bz), (bx, by, az)),
        ((ax, ay, az), (ax, ay, bz), (bx, ay, bz), (bx, ay, az)),
        ((ax, by, az), (bx, by, az), (bx, by, bz), (ax, by, bz)),
        ((ax, ay, az), (bx, ay, az), (bx, by, az), (ax, by, az)),
        ((ax, ay, bz), (ax, by, bz), (bx, by, bz), (bx, ay, bz)),
    ]
    for f, m in zip(faces, ids):
        mb.begin_surface()
        mb.quad(*f, m)

    def inside():
        return tuple(rng.uniform(clearance, s - clearance) for s in size)

    src = inside()
    lis = inside()
    return mb, np.asarray(src, np.float32), np.asarray(lis, np.float32), mats


# four material classes of the city block: (absorption range, scattering, transmission)
_CLASSES = [
    ((0.02, 0.06), 0.12, [0.045, 0.008, 0.0012]),   # brick / masonry
    ((0.01, 0.05), 0.10, [0.10, 0.014, 0.0015]),    # concrete
    ((0.05, 0.25), 0.15, [0.22, 0.060, 0.0200]),    # wood / panel
    ((0.10, 0.40), 0.40, [0.30, 0.100, 0.0300]),    # soft / planted
]


def _class_materials(rng, nbands):
    mats = []
    for (lo, hi), scat, trans in _CLASSES:
        absorption = [rng.uniform(lo, hi) for _ in range(nbands)]
        tr = [trans[min(b * 3 // nbands, 2)] for b in range(nbands)]
        mats.append(MaterialDef(absorption, scat, tr))
    return mats


def _wall_cells(mb, mat, plane_axis, plane, lo, hi, z0, z1, cell, holes):
    """A vertical wall as a grid of quads with rectangular holes removed.

    plane_axis 0: wall in plane x = plane spanning y in [lo, hi];
    plane_axis 1: plane y = plane spanning x.  holes: (u0, u1, v0, v1).
    """
    nu = max(1, int(math.ceil((hi - lo) / cell)))
    nv = max(1, int(math.ceil((z1 - z0) / cell)))
    du = (hi - lo) / nu
    dv = (z1 - z0) / nv

    def P(u, v):
        return (plane, u, v) if plane_axis == 0 else (u, plane, v)

    for i in range(nu):
        u0, u1 = lo + i * du, lo + (i + 1) * du
        um = 0.5 * (u0 + u1)
        for j in range(nv):
            v0, v1 = z0 + j * dv, z0 + (j + 1) * dv
            vm = 0.5 * (v0 + v1)
            if any(h[0] <= um <= h[1] and h[2] <= vm <= h[3] for h in holes):
                continue
            mb.quad(P(u0, v0), P(u1, v0), P(u1, v1), P(u0, v1), mat)


def city_block(seed=2, n_buildings=40, target_tris=50_000, nbands=3, extent=None,
               side=(3.0, 15.0), height=(3.0, 12.0)):
    """Config: procedural city block of extruded boxes with door/window holes.

    Buildings: sides U[3, 15] m, heights U[3, 12] m, 1-3 holes per wall
    (door 1 x 2.1 m or window 1.2 x 1.2 m), walls and roofs gridded so the
    block totals about target_tris triangles; 4 material classes with
    seeded absorption; a ground plane (tessellated to 4 m by the BVH build).
    Returns (builder, bounds (min, max), building boxes).
    """
    rng = SplitMix64(seed)
    mats = _class_materials(rng, nbands)
    if extent is None:
        extent = max(60.0, 12.0 * math.sqrt(n_buildings))
    boxes = []
    for _ in range(n_buildings):
        sx, sy = rng.uniform(*side), rng.uniform(*side)
        h = rng.uniform(*height)
        x0 = rng.uniform(0.0, extent - sx)
        y0 = rng.uniform(0.0, extent - sy)
        boxes.append(((x0, y0, 0.0), (x0 + sx, y0 + sy, h), rng.randint(0, 3)))
    # total authored area -> grid cell size hitting the triangle target
    area = sum(2 * ((b[1][0] - b[0][0]) + (b[1][1] - b[0][1])) * b[1][2] +
               (b[1][0] - b[0][0]) * (b[1][1] - b[0][1]) for b in boxes)
    cell = max(0.25, math.sqrt(2.0 * area / max(1, target_tris - 2)))
    mb = MeshBuilder()
    ids = [mb.material(m) for m in mats]
    mb.begin_surface()
    g = extent * 0.25
    mb.quad((-g, -g, 0.0), (extent + g, -g, 0.0), (extent + g, extent + g, 0.0),
            (-g, extent + g, 0.0), ids[1])
    for (x0, y0, _), (x1, y1, h), mi in boxes:
        m = ids[mi]
        walls = [(0, x0, y0, y1), (0, x1, y0, y1), (1, y0, x0, x1), (1, y1, x0, x1)]
        for axis, plane, lo, hi in walls:
            holes = []
            for _ in range(rng.randint(1, 3)):
                door = rng.uniform() < 0.4
                w, hh = (1.0, 2.1) if door else (1.2, 1.2)
                if hi - lo < w + 0.4 or h < hh + 0.4:
                    continue
                u0 = rng.uniform(lo + 0.2, hi - 0.2 - w)
                v0 = 0.0 if door else rng.uniform(0.6, min(h - hh - 0.2, 0.6 + 2.5 * h))
                holes.append((u0, u0 + w, v0, v0 + hh))
            mb.begin_surface()
            _wall_cells(mb, m, axis, plane, lo, hi, 0.0, h, cell, holes)
        mb.begin_surface()
        nx = max(1, int(math.ceil((x1 - x0) / cell)))
        ny = max(1, int(math.ceil((y1 - y0) / cell)))
        for i in range(nx):
            for j in range(ny):
                ax, bx = x0 + (x1 - x0) * i / nx, x0 + (x1 - x0) * (i + 1) / nx
                ay, by = y0 + (y1 - y0) * j / ny, y0 + (y1 - y0) * (j + 1) / ny
                mb.quad((ax, ay, h), (bx, ay, h), (bx, by, h), (ax, by, h), m)
    bounds = (np.array([0.0, 0.0, 0.0], np.float32), np.array([extent, extent, 12.0], np.float32))
    return mb, bounds, [(b[0], b[1]) for b in boxes]


def indoor_points(seed, n, boxes, z=(1.2, 1.8), margin=0.5):
    """n points at ear height inside seeded building boxes (rejecting points
    inside any other building's footprint margin band)."""
    rng = SplitMix64(seed)
    out = []
    while len(out) < n:
        (x0, y0, _), (x1, y1, h) = boxes[rng.randint(0, len(boxes) - 1)]
        if h < z[1] + 0.5:
            continue
        p = (rng.uniform(x0 + margin, x1 - margin), rng.uniform(y0 + margin, y1 - margin),
             rng.uniform(*z))
        near_wall = any(
            (b0[0] - margin < p[0] < b1[0] + margin and b0[1] - margin < p[1] < b1[1] + margin) and
            not (b0[0] + margin < p[0] < b1[0] - margin and b0[1] + margin < p[1] < b1[1] - margin)
            for b0, b1 in boxes)
        if not near_wall:
            out.append(p)
    return np.asarray(out, np.float32)


def c2_workload(seed=2):
    """Plain single-launch workload: the ~50k-triangle city block, 16 sources
    (half inside buildings, half in the streets), 8 listeners (likewise)."""
    mb, bounds, boxes = city_block(seed=seed)
    src = np.vstack([indoor_points(seed * 100 + 1, 8, boxes),
                     outdoor_points(seed * 100 + 2, 8, bounds, boxes)])
    lis = np.vstack([indoor_points(seed * 100 + 3, 4, boxes),
                     outdoor_points(seed * 100 + 4, 4, bounds, boxes)])
    return mb, src, lis


def outdoor_points(seed, n, bounds, boxes, z=(1.2, 1.8), margin=0.6):
    """n points at ear height inside bounds and outside every building box."""
    rng = SplitMix64(seed)
    out = []
    lo, hi = bounds
    while len(out) < n:
        p = (rng.uniform(lo[0] + 2, hi[0] - 2), rng.uniform(lo[1] + 2, hi[1] - 2),
             rng.uniform(*z))
        if any(b0[0] - margin < p[0] < b1[0] + margin and b0[1] - margin < p[1] < b1[1] + margin
               for b0, b1 in boxes):
            continue
        out.append(p)
    return np.asarray(out, np.float32)


def c3_workload(seed=3, n_queries=256, n_scenes=32, tris=(1_000, 200_000), sources=(1, 32),
                rays=(1024, 65536), listeners=(1, 4), bands=(3, 6), seconds=(1, 4)):
    """Batched mixed-size workload: n_queries queries over a pool of
    n_scenes procedural blocks.

    Scene i: a city block with triangle target log-U[tris], 3 or 6 bands.
    Query: a pool scene, sources round(log-U[sources]) and rays per source
    round(log-U[rays]), listeners U{listeners}, half the points inside
    buildings and half in the streets, echogram 1 ms x U{seconds} s, its own
    Philox seed.  Returns (scene_specs, queries): scene_specs[i] =
    (builder, nbands), queries[j] = dict(scene, n_src, n_rays, nbands,
    max_time, seed, sources, listeners).
    """
    rng = SplitMix64(seed)
    specs, geo = [], []
    for i in range(n_scenes):
        t = int(round(rng.log_uniform(*tris)))
        nb = bands[rng.randint(0, len(bands) - 1)]
        nbld = min(160, max(4, int(round(t / 1250))))
        mb, bounds, boxes = city_block(seed=seed * 1000 + i, n_buildings=nbld, target_tris=t,
                                       nbands=nb)
        specs.append((mb, nb))
        geo.append((bounds, boxes))
    queries = []
    for j in range(n_queries):
        si = rng.randint(0, n_scenes - 1)
        n_src = int(round(rng.log_uniform(*sources)))
        n_rays = int(round(rng.log_uniform(*rays)))
        n_lis = rng.randint(*listeners)
        bounds, boxes = geo[si]
        base = seed * 100000 + j * 10
        k = n_src + n_lis
        mix = np.empty((2 * k, 3), np.float32)
        mix[0::2] = indoor_points(base + 1, k, boxes)
        mix[1::2] = outdoor_points(base + 2, k, bounds, boxes)
        src, lis = mix[:n_src], mix[n_src:k]
        queries.append(dict(scene=si, n_src=n_src, n_rays=n_rays, nbands=specs[si][1],
                            max_time=float(rng.randint(*seconds)), seed=base,
                            sources=np.ascontiguousarray(src, np.float32),
                            listeners=np.ascontiguousarray(lis, np.float32)))
    return specs, queries


def c5_workload(seed=5, n_scenes=8, tris=(5_000, 500_000), n_walkers=1000, qps=2000.0,
                seconds=2.0, sources=(1, 32), rays=(512, 16384), speed=1.4, nbands=3):
    """Serving workload: n_scenes resident city blocks (triangle targets
    log-U[tris]) and a Poisson request stream at `qps` for `seconds`.

    Listeners are n_walkers walkers, walker w living in scene w % n_scenes,
    starting at a street point and walking at `speed` m/s with a heading
    that turns by up to +-0.5 rad per request, reflected at the block's
    bounds.  A request takes a uniformly drawn walker at its arrival time,
    round(log-U[sources]) sources drawn from a per-scene pool of indoor and
    street points, and round(log-U[rays]) rays per source.  Returns
    (scene_specs [(builder, bounds, boxes)], requests [dict(scene, t,
    n_src, n_rays, sources, listener, seed)]).
    """
    rng = SplitMix64(seed)
    specs = []
    for i in range(n_scenes):
        t = int(round(rng.log_uniform(*tris)))
        nbld = min(400, max(4, int(round(t / 1250))))
        mb, bounds, boxes = city_block(seed=seed * 1000 + i, n_buildings=nbld, target_tris=t,
                                       nbands=nbands)
        specs.append((mb, bounds, boxes))
    pools = []
    for i, (_, bounds, boxes) in enumerate(specs):
        pts = np.empty((512, 3), np.float32)
        pts[0::2] = indoor_points(seed * 7000 + i, 256, boxes)
        pts[1::2] = outdoor_points(seed * 7100 + i, 256, bounds, boxes)
        pools.append(pts)
    walkers = []
    for w in range(n_walkers):
        si = w % n_scenes
        _, bounds, boxes = specs[si]
        p = outdoor_points(seed * 100000 + w, 1, bounds, boxes)[0].astype(np.float64)
        walkers.append([si, p, rng.uniform(0.0, 2.0 * math.pi), 0.0])
    reqs = []
    t = 0.0
    while True:
        t += -math.log(max(rng.uniform(), 1e-300)) / qps
        if t >= seconds:
            break
        w = walkers[rng.randint(0, n_walkers - 1)]
        si, p, heading, t_last = w
        lo, hi = specs[si][1]
        step = speed * (t - t_last)
        p[0] += step * math.cos(heading)
        p[1] += step * math.sin(heading)
        for k in range(2):
            a, b = float(lo[k]) + 1.0, float(hi[k]) - 1.0
            if p[k] < a:
                p[k] = 2 * a - p[k]
            if p[k] > b:
                p[k] = 2 * b - p[k]
        w[2] = heading + rng.uniform(-0.5, 0.5)
        w[3] = t
        n_src = int(round(rng.log_uniform(*sources)))
        n_rays = int(round(rng.log_uniform(*rays)))
        pool = pools[si]
        idx = [rng.randint(0, len(pool) - 1) for _ in range(n_src)]
        reqs.append(dict(scene=si, t=t, n_src=n_src, n_rays=n_rays,
                         sources=np.ascontiguousarray(pool[idx]),
                         listener=p.astype(np.float32).reshape(1, 3), seed=seed * 1000003 + len(reqs)))
    return specs, reqs


def _slab_cells(mb, mat, x0, x1, y0, y1, z, cell, holes):
    """A horizontal slab at height z as a grid of quads with (x0, x1, y0, y1) holes removed."""
    nx = max(1, int(math.ceil((x1 - x0) / cell)))
    ny = max(1, int(math.ceil((y1 - y0) / cell)))
    for i in range(nx):
        ax, bx = x0 + (x1 - x0) * i / nx, x0 + (x1 - x0) * (i + 1) / nx
        xm = 0.5 * (ax + bx)
        for j in range(ny):
            ay, by = y0 + (y1 - y0) * j / ny, y0 + (y1 - y0) * (j + 1) / ny
            ym = 0.5 * (ay + by)
            if any(h[0] <= xm <= h[1] and h[2] <= ym <= h[3] for h in holes):
                continue
            mb.quad((ax, ay, z), (bx, ay, z), (bx, by, z), (ax, by, z), mat)


def town(seed=4, n_buildings=400, target_tris=2_000_000, nbands=6, side=(6.0, 20.0),
         height=(6.0, 15.0), storey=3.0):
    """Large-scene stress config: a procedural town of about target_tris
    triangles.  n_buildings seeded boxes (sides U[side], heights U[height])
    with interiors: a floor slab per storey with a stair opening, two
    crossing partition walls per storey with a door each, exterior walls
    with a door on the ground storey and windows above, a roof; 4 material
    classes; a ground plane.  Surfaces are gridded so the town totals about
    target_tris triangles.  Returns (builder, bounds (min, max), boxes)."""
    rng = SplitMix64(seed)
    mats = _class_materials(rng, nbands)
    extent = max(100.0, 14.0 * math.sqrt(n_buildings))
    boxes = []
    for _ in range(n_buildings):
        sx, sy = rng.uniform(*side), rng.uniform(*side)
        h = rng.uniform(*height)
        x0 = rng.uniform(0.0, extent - sx)
        y0 = rng.uniform(0.0, extent - sy)
        boxes.append(((x0, y0, 0.0), (x0 + sx, y0 + sy, h), rng.randint(0, 3)))
    area = 0.0
    for (x0, y0, _), (x1, y1, h), _ in boxes:
        w, d = x1 - x0, y1 - y0
        storeys = max(1, int(h / storey))
        area += 2 * (w + d) * h + w * d            # exterior walls, roof
        area += (storeys - 1) * w * d              # slabs
        area += (w + d) * h                        # two partitions
    cell = max(0.25, math.sqrt(2.0 * area / max(1, target_tris - 2)))
    mb = MeshBuilder()
    ids = [mb.material(m) for m in mats]
    mb.begin_surface()
    g = extent * 0.1
    mb.quad((-g, -g, 0.0), (extent + g, -g, 0.0), (extent + g, extent + g, 0.0),
            (-g, extent + g, 0.0), ids[1])
    for (x0, y0, _), (x1, y1, h), mi in boxes:
        m = ids[mi]
        inner = ids[(mi + 2) % len(ids)]
        storeys = max(1, int(h / storey))
        # exterior walls: a door on the ground storey, windows above
        for axis, plane, lo, hi in [(0, x0, y0, y1), (0, x1, y0, y1), (1, y0, x0, x1),
                                    (1, y1, x0, x1)]:
            holes = []
            if hi - lo > 1.6:
                u0 = rng.uniform(lo + 0.3, hi - 1.3)
                holes.append((u0, u0 + 1.0, 0.0, 2.1))
            for k in range(1, storeys):
                if hi - lo > 1.8 and rng.uniform() < 0.7:
                    u0 = rng.uniform(lo + 0.3, hi - 1.5)
                    holes.append((u0, u0 + 1.2, k * storey + 0.9, k * storey + 2.1))
            mb.begin_surface()
            _wall_cells(mb, m, axis, plane, lo, hi, 0.0, h, cell, holes)
        mb.begin_surface()
        _slab_cells(mb, m, x0, x1, y0, y1, h, cell, [])
        # slabs with a stair opening
        for k in range(1, storeys):
            z = k * storey
            if z >= h - 0.5:
                break
            sx0 = rng.uniform(x0 + 0.5, max(x0 + 0.6, x1 - 1.7))
            sy0 = rng.uniform(y0 + 0.5, max(y0 + 0.6, y1 - 3.0))
            mb.begin_surface()
            _slab_cells(mb, inner, x0, x1, y0, y1, z, cell, [(sx0, sx0 + 1.2, sy0, sy0 + 2.5)])
        # two crossing partitions, a door per storey in each
        xm, ym = 0.5 * (x0 + x1), 0.5 * (y0 + y1)
        for axis, plane, lo, hi in [(1, ym, x0, x1), (0, xm, y0, y1)]:
            holes = []
            for k in range(storeys):
                if hi - lo > 1.6:
                    u0 = rng.uniform(lo + 0.3, hi - 1.3)
                    holes.append((u0, u0 + 0.9, k * storey, k * storey + 2.1))
            mb.begin_surface()
            _wall_cells(mb, inner, axis, plane, lo, hi, 0.0, h, cell, holes)
    bounds = (np.array([0.0, 0.0, 0.0], np.float32),
              np.array([extent, extent, height[1]], np.float32))
    return mb, bounds, [(b[0], b[1]) for b in boxes]


def c4_workload(seed=4):
    """Large-scene stress: the ~2M-triangle town, 64 sources (half inside
    buildings, half in the streets) and 4 listeners."""
    mb, bounds, boxes = town(seed=seed)
    src = np.vstack([indoor_points(seed * 100 + 1, 32, boxes),
                     outdoor_points(seed * 100 + 2, 32, bounds, boxes)])
    lis = np.vstack([indoor_points(seed * 100 + 3, 2, boxes),
                     outdoor_points(seed * 100 + 4, 2, bounds, boxes)])
    return mb, src, lis


def mesh_bench(n_queries=1000):
    """The reference's mesh benchmark scene and path queries
    (crates/omg-app/examples/mesh_bench.rs:15-34, :81-92): nine 8 x 6 x 3 m
    brick blocks on a 15 m grid, 200 clutter boxes drawn from Rng::new(9),
    an enclosing 55 x 55 x 12.1 m shell; one source at (4, 3, 1.5) inside the
    first block, listeners at (20 + i % 7, 20 + i % 5, 1.6).  Returns the
    builder, sources [n][3] and listeners [n][3]."""
    brick = MaterialDef([0.02, 0.03, 0.04], 0.12, [0.045, 0.008, 0.0012])   # material.rs:45-49
    mb = MeshBuilder()
    m = mb.material(brick)
    rng = Xorshift64Star(9)
    f = np.float32
    for i in range(9):
        x = f((i % 3) * 15.0)
        y = f((i // 3) * 15.0)
        mb.abox((x, y, f(0.0)), (x + f(8.0), y + f(6.0), f(3.0)), m)
    for _ in range(200):
        p = (rng.next_f32() * f(40.0), rng.next_f32() * f(40.0), f(0.0))
        sx = f(0.3) + rng.next_f32()
        sy = f(0.3) + rng.next_f32()
        sz = f(0.5) + rng.next_f32() * f(1.5)
        mb.abox(p, (p[0] + sx, p[1] + sy, p[2] + sz), m)
    mb.abox((-5.0, -5.0, -0.1), (50.0, 50.0, 12.0), m)
    lis = np.array([(20.0 + i % 7, 20.0 + i % 5, 1.6) for i in range(n_queries)], np.float32)
    src = np.tile(np.float32([4.0, 3.0, 1.5]), (n_queries, 1))
    return mb, src, lis


def path_queries(seed, n, bounds, boxes):
    """n (source, listener) pairs for the path search on a city block: ear-
    height points in the streets and inside buildings, paired at random."""
    pts = np.vstack([outdoor_points(seed * 100 + 1, 64, bounds, boxes),
                     indoor_points(seed * 100 + 2, 64, boxes)])
    rng = SplitMix64(seed)
    src, lis = [], []
    while len(src) < n:
        i = rng.randint(0, pts.shape[0] - 1)
        j = rng.randint(0, pts.shape[0] - 1)
        if i != j:
            src.append(pts[i])
            lis.append(pts[j])
    return np.asarray(src, np.float32), np.asarray(lis, np.float32)

