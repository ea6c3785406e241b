"""Pins of the CPU oracle: known-answer vectors, closed forms, brute force and
the reference's own statistical goldens and behaviour tests (CPU only)."""
import json
import math
import os

import numpy as np
import pytest

import oracle as O
from omg_hip import scenes
from helpers import (CONCRETE, DRYWALL, GOLDEN_AVG, GOLDEN_CONFIGS, GOLDEN_RAYS, TILE,
                     golden_violations, mat_tuple, oracle_mesh_from_builder)

GOLDEN_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


# ------------------------------------------------------------------ RNG / numerics

def test_philox_known_answers():
    """Random123 known-answer vectors for philox4x32-10 (Salmon et al. 2011)."""
    kat = json.load(open(os.path.join(GOLDEN_DIR, "philox4x32_10_kat.json")))
    for case in kat:
        got = O.philox(case["ctr"], case["key"])
        assert [int(x) for x in got] == case["out"]


def test_xorshift_stream_matches_fixture():
    """rng.rs:10-21 stream against a fixture computed with Python integers."""
    fx = json.load(open(os.path.join(GOLDEN_DIR, "xorshift64star.json")))
    for case in fx:
        got = O.xorshift_stream(case["seed"], len(case["u64"]))
        assert [int(x) for x in got] == case["u64"]
        f = O.xorshift_f32_stream(case["seed"], len(case["u64"]))
        assert np.array_equal(f, np.array(case["f32"], np.float32))


def test_project_sincos_accuracy():
    xs = np.float32(np.concatenate([np.linspace(0, 2 * np.pi, 20001),
                                    np.linspace(0, 1300, 20001)]))
    for x in xs[::3]:
        s, c = O.sincos(x)
        for v, r in ((s, math.sin(float(x))), (c, math.cos(float(x)))):
            ulp = np.spacing(np.float32(max(abs(r), 1e-30)))
            assert abs(float(v) - r) <= 2.0 * ulp + 1e-12, (x, v, r)


# ------------------------------------------------------------------ BVH

def _soup(seed, n=200, spread=20.0, size=2.0):
    rng = O.xorshift_f32_stream(seed, n * 9 + 10)
    mb = scenes.MeshBuilder()
    m = mb.material(CONCRETE)
    k = 0
    for _ in range(n):
        base = rng[k:k + 3] * spread
        k += 3
        a = rng[k:k + 3] * 2 - 1
        b = rng[k + 3:k + 6] * 2 - 1
        k += 6
        mb.tri(tuple(base), tuple(base + a * size), tuple(base + b * size), m)
    return mb


def _mt_f64(p0, p1, p2, o, d):
    """Textbook Moller-Trumbore in float64 (independent of the oracle)."""
    e1, e2 = p1 - p0, p2 - p0
    p = np.cross(d, e2)
    det = e1 @ p
    if abs(det) < 1e-12:
        return None
    s = o - p0
    u = (s @ p) / det
    if u < -1e-6 or u > 1 + 1e-6:
        return None
    q = np.cross(s, e1)
    v = (d @ q) / det
    if v < -1e-6 or u + v > 1 + 1e-6:
        return None
    t = (e2 @ q) / det
    return t if t > 1e-5 else None


def test_bvh_raycast_matches_brute_force():
    """mesh.rs:697-735 restated: |t_bvh - t_brute| < 1e-4 on a random soup."""
    mb = _soup(7)
    om = oracle_mesh_from_builder(mb)
    pos, idx = O.dedup_tris(mb.soup()["tri_verts"])
    P = pos.astype(np.float64)
    rng = np.random.default_rng(1)
    for _ in range(500):
        o = rng.uniform(0, 20, 3)
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        o32, d32 = o.astype(np.float32), d.astype(np.float32)
        bvh = om.raycast(o32, d32)
        ts = [_mt_f64(P[i], P[j], P[k], o32.astype(np.float64), d32.astype(np.float64))
              for i, j, k in idx]
        ts = [t for t in ts if t is not None]
        if bvh is None:
            assert not ts or min(ts) < 1.5e-5
        else:
            assert ts and abs(bvh[0] - min(ts)) < 1e-4


def test_bvh_invariants():
    """Every prim in exactly one leaf, leaves <= 4, child boxes inside their parent,
    patches <= max edge and coplanar with their original triangle (mesh.rs:241-405)."""
    mb, bounds, boxes = scenes.city_block(seed=4, n_buildings=5, target_tris=3000)
    om = oracle_mesh_from_builder(mb)
    nodes, prims = om.export()
    nu = nodes.view(np.uint32)
    seen = np.zeros(prims.shape[0], np.int32)
    stack = [0]
    while stack:
        i = stack.pop()
        a, b = int(nu[i, 6]), int(nu[i, 7])
        if a & 0x80000000:
            s = a & 0x7FFFFFFF
            assert 1 <= b <= 4
            seen[s:s + b] += 1
            v = np.stack([prims[s:s + b, 0:3], prims[s:s + b, 0:3] + prims[s:s + b, 3:6],
                          prims[s:s + b, 0:3] + prims[s:s + b, 6:9]]).reshape(-1, 3)
            assert np.all(v >= nodes[i, 0:3] - 1e-5) and np.all(v <= nodes[i, 3:6] + 1e-5)
        else:
            for c in (a, b):
                assert np.all(nodes[c, 0:3] >= nodes[i, 0:3]) and np.all(nodes[c, 3:6] <= nodes[i, 3:6])
            stack += [a, b]
    assert np.all(seen == 1)
    soup = mb.soup()
    tv = soup["tri_verts"].reshape(-1, 3, 3).astype(np.float64)
    tri = prims[:, 9].view(np.uint32)
    for k in range(0, prims.shape[0], 7):
        a, e1, e2 = prims[k, 0:3], prims[k, 3:6], prims[k, 6:9]
        edges = [np.linalg.norm(e1), np.linalg.norm(e2), np.linalg.norm(e2 - e1)]
        assert max(edges) <= 4.0 + 1e-4
        t = tv[tri[k]]
        n = np.cross(t[1] - t[0], t[2] - t[0])
        n /= np.linalg.norm(n)
        for p in (a, a + e1, a + e2):
            assert abs((p - t[0]) @ n) < 1e-4


def test_segment_hits_counts_wall_crossings():
    """mesh.rs:771-784"""
    mb = scenes.MeshBuilder()
    m = mb.material(CONCRETE)
    mb.abox((0, 0, 0), (4, 4, 3), m)
    om = oracle_mesh_from_builder(mb)
    t, _, _ = om.segment_hits((-2, 2, 1.5), (6, 2, 1.5))
    assert len(t) == 2 and t[0] < t[1]
    assert len(om.segment_hits((-2, 2, 1.5), (2, 2, 1.5))[0]) == 1
    assert len(om.segment_hits((1, 2, 1.5), (3, 2, 1.5))[0]) == 0


# ------------------------------------------------------------------ tracer

def _run_box(cfg, seed, mode):
    name, size, walls, src, lis, expect = cfg
    g = O.box_geom(size, [mat_tuple(w) for w in walls])
    n = O.nbins(3.0, 0.010)
    ab = np.zeros((n, 3), np.float32)
    ad = np.zeros((n, 3), np.float32)
    for k in range(GOLDEN_AVG):
        s = (seed + k * 0x9E3779B97F4A7C15) & ((1 << 64) - 1)
        r = O.trace(g, src, [lis], O.trace_params(GOLDEN_RAYS, rng_mode=mode, seed=s))
        O.ema(ab, ad, r["bins"][0], r["dirs"][0], np.float32(1.0 / (k + 1)))
    rt, lv = O.estimate_reverb(ab, 0.010)
    _, an = O.late_direction(ab, ad, 0.010, 0.05)
    return list(rt) + [20 * math.log10(max(float(v), 1e-9)) for v in lv] + [float(an)]


@pytest.mark.parametrize("cfg", GOLDEN_CONFIGS, ids=[c[0] for c in GOLDEN_CONFIGS])
def test_trace_goldens_both_rng_modes(cfg):
    """trace_golden.rs:198-210 with the reference's xorshift stream, and the
    counter-based stream the GPU uses, against the same goldens."""
    for seed in (0xC0FFEE, 0x5EED5EED):
        bad = golden_violations(_run_box(cfg, seed, O.RNG_XORSHIFT), cfg[5])
        assert not bad, f"{cfg[0]} seed {seed:#x}: {bad}"
    # The counter stream is held to the same goldens on the mean of six
    # independent 4-trace measurements: the goldens sit ~1.5 sigma of a single
    # measurement wide (trace_golden.rs:8-13), so any single seed of any
    # stream fails them now and then.
    m = np.mean([_run_box(cfg, 0x1234 + 977 * i, O.RNG_COUNTER) for i in range(6)], axis=0)
    bad = golden_violations(m, cfg[5])
    assert not bad, f"{cfg[0]} counter stream: {bad}"


def test_single_ray_receiver_bin():
    """A ray aimed through the listener from the source: the first bin it scores
    is floor(((d + s)/343)/dt) for its own path (tracer.rs:110-118)."""
    g = O.box_geom((40, 40, 40), [([0.999] * 3, 0.0, [0.0] * 3)] * 6)
    src, lis = (20.0, 20.0, 20.0), (23.0, 24.0, 20.0)
    r = O.trace(g, src, [lis], O.trace_params(20000, seed=5, max_bounces=1, bin_dt=0.001,
                                               max_time=2.0), want_f64=True)
    nz = np.nonzero(r["bins"][0][:, 0])[0]
    # every direct arrival lies within the 0.5 m sphere around 5 m
    assert nz.size > 0
    assert np.all((nz >= int(4.5 / 343 / 0.001)) & (nz <= int(5.5 / 343 / 0.001)))


def test_box_mesh_traces_like_shoebox():
    """mesh.rs:737-768"""
    size = (8.0, 6.0, 3.0)
    walls = [mat_tuple(CONCRETE)] * 6
    mb = scenes.MeshBuilder()
    m = mb.material(CONCRETE)
    mb.abox((0, 0, 0), size, m)
    om = oracle_mesh_from_builder(mb)
    src, lis = (2.0, 3.0, 1.5), (6.0, 2.0, 1.6)
    p = O.trace_params(8192, rng_mode=O.RNG_XORSHIFT, seed=42)
    eb = O.trace(O.box_geom(size, walls), src, [lis], p)
    em = O.trace(O.mesh_geom(om), src, [lis], p)
    rb, lb = O.estimate_reverb(eb["bins"][0], 0.01)
    rm, lm = O.estimate_reverb(em["bins"][0], 0.01)
    assert np.all((rm / rb > 0.8) & (rm / rb < 1.25)), (rm, rb)
    assert np.all((lm / np.maximum(lb, 1e-9) > 0.7) & (lm / np.maximum(lb, 1e-9) < 1.4))


def test_hole_direction_and_closed_room_diffuse():
    """tracer.rs:271-311"""
    mb = scenes.MeshBuilder()
    m = mb.material(CONCRETE)
    v = lambda x, y, z: (x, y, z)
    mb.quad(v(0, 0, 0), v(0, 6, 0), v(0, 6, 3), v(0, 0, 3), m)
    mb.quad(v(0, 0, 0), v(8, 0, 0), v(8, 0, 3), v(0, 0, 3), m)
    mb.quad(v(0, 6, 0), v(8, 6, 0), v(8, 6, 3), v(0, 6, 3), m)
    mb.quad(v(0, 0, 0), v(8, 0, 0), v(8, 6, 0), v(0, 6, 0), m)
    mb.quad(v(0, 0, 3), v(8, 0, 3), v(8, 6, 3), v(0, 6, 3), m)
    mb.quad(v(8, 0, 0), v(8, 2.5, 0), v(8, 2.5, 3), v(8, 0, 3), m)
    mb.quad(v(8, 3.5, 0), v(8, 6, 0), v(8, 6, 3), v(8, 3.5, 3), m)
    mb.quad(v(8, 2.5, 2.1), v(8, 3.5, 2.1), v(8, 3.5, 3), v(8, 2.5, 3), m)
    om = oracle_mesh_from_builder(mb)
    lis = np.array([12.0, 3.0, 1.6], np.float32)
    r = O.trace(O.mesh_geom(om), (2.0, 3.0, 1.5), [lis],
                O.trace_params(16384, rng_mode=O.RNG_XORSHIFT, seed=21))
    d, an = O.late_direction(r["bins"][0], r["dirs"][0], 0.01, 0.05)
    to_hole = np.array([8.0, 3.0, 1.05]) - lis
    to_hole /= np.linalg.norm(to_hole)
    assert d @ to_hole > 0.85 and an > 0.5
    g = O.box_geom((8, 6, 3), [mat_tuple(DRYWALL)] * 6)
    r = O.trace(g, (2, 3, 1.5), [(6, 2, 1.6)], O.trace_params(16384, rng_mode=O.RNG_XORSHIFT,
                                                               seed=5))
    _, an = O.late_direction(r["bins"][0], r["dirs"][0], 0.01, 0.15)
    assert an < 0.35


def test_rt60_tracks_absorption_and_level_is_distance_independent():
    """crates/omg-core/src/lib.rs:58-101"""
    st = [7]

    def rt(mat):
        g = O.box_geom((8, 6, 3), [mat_tuple(mat)] * 6)
        r = O.trace(g, (2, 2, 1.5), [(5, 4, 1.5)], O.trace_params(20000, rng_mode=O.RNG_XORSHIFT,
                                                                   seed=st[0]))
        st[0] += 1
        return O.estimate_reverb(r["bins"][0], 0.01)[0][1]

    live, dead, dry = rt(CONCRETE), rt(TILE), rt(DRYWALL)
    assert live > 2.0 * dead and 0.8 < dry < 3.0
    g = O.box_geom((10, 8, 4), [mat_tuple(DRYWALL)] * 6)
    lv = [O.estimate_reverb(O.trace(g, (2, 4, 1.5), [lis], O.trace_params(
        40000, rng_mode=O.RNG_XORSHIFT, seed=11 + i))["bins"][0], 0.01)[1][1]
        for i, lis in enumerate([(3, 4, 1.5), (8, 4, 1.5)])]
    assert 0.5 < lv[0] / max(lv[1], 1e-6) < 2.0


def test_estimate_reverb_closed_form():
    """An exponential echogram of known T60 returns it within one bin quantum;
    empty bands keep 0.5 s; late level sqrt(sum after 80 ms) capped at 0.7."""
    dt = 0.01
    t = (np.arange(300) + 0.5) * dt
    for t60 in (0.3, 0.8, 1.7):
        e = np.exp(-t * 6.0 * math.log(10) / t60)
        bins = np.stack([e, e, np.zeros_like(e)], 1).astype(np.float32) * 1e-3
        rt, lv = O.estimate_reverb(bins, dt)
        assert abs(rt[0] - t60) < 0.06 * t60 + dt and rt[2] == np.float32(0.5)
        late = math.sqrt(float(bins[8:, 0].astype(np.float32).sum(dtype=np.float32)))
        assert abs(lv[0] - min(late, 0.7)) < 1e-6


def test_ism_counts_and_direct_delay():
    """crates/omg-core/src/lib.rs:38-55"""
    walls = [mat_tuple(DRYWALL)] * 6
    taps = O.image_source_taps((8, 6, 3), walls, (2, 2, 1.5), (5, 4, 1.5), 3)
    assert taps.shape[0] == 63
    d = O.image_source_taps((8, 6, 3), walls, (2, 2, 1.5), (5, 4, 1.5), 0)
    assert d.shape[0] == 1
    assert abs(d[0, 1] - math.sqrt(13.0) / 343.0) < 1e-6


def test_mesh_record_box_equivalence():
    """pt_mesh_gate.rs:25-126 against the shoebox image-source lattice."""
    from test_dome_ism_gpu import live_room_mesh
    mb, walls = live_room_mesh()
    om = oracle_mesh_from_builder(mb)
    tab, n = O.surface_table(om)
    src, lis = (1.5, 1.0, 1.5), (4.5, 3.0, 1.6)
    chains = set()
    for rot in range(8):
        ch, od = O.mesh_chains(om, lis, 4096, rot)
        chains |= {(tuple(c[:o]), int(o)) for c, o in zip(ch, od)}
    recs = [O.mesh_record(om, tab, n, [], 0, src, lis)]
    recs += [O.mesh_record(om, tab, n, list(c), o, src, lis) for c, o in sorted(chains)]
    recs = [r for r in recs if r is not None]
    taps = O.image_source_taps((6, 4, 3), [mat_tuple(w) for w in walls], src, lis, 3)
    db = lambda a: 20 * np.log10(np.maximum(a, 1e-9))
    for r in recs:
        assert np.any((np.abs(taps[:, 1] - r["delay"]) * 1000 <= 0.5) &
                      np.all(np.abs(db(taps[:, 5:8]) - db(r["gains"])) <= 1.0, axis=1))
    assert len(recs) >= 40
    assert abs(recs[0]["delay"] - math.dist(src, lis) / 343.0) < 1e-5


def test_dome_open_field_closed_form_oracle():
    """Only a triangle beyond the sky sphere: an eye above grade sees every fan
    ray escape with throughput 1, so bin k holds exactly count_k / N and the
    bin direction sum is the plain sum of its fan directions (dome.rs:433-458)."""
    mb = scenes.MeshBuilder()
    m = mb.material(CONCRETE)
    mb.tri((5000, 0, 0), (5000, 1, 0), (5000, 0, 1), m)
    om = oracle_mesh_from_builder(mb)
    o = O.dome_sample(om, (0, 0, 1.6), n_fan=512, horizon=500.0, want_rays=True)
    fan = O.dome_fan(512)
    assert np.all(o["escaped"] == 1) and np.all(o["tp"] == 1.0) and not np.any(o["through"])
    bins = np.array([O.dome_bin_of(d) for d in fan])
    for k in range(9):
        assert o["e"][k, 0] == np.float32(np.sum(bins == k) / 512.0)
    assert np.all(bins[fan[:, 2] > 0.87] == 8) and set(bins[fan[:, 2] <= 0.87]) == set(range(8))
