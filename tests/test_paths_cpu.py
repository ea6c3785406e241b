"""The oracle's knife-edge path search (oracle/or_paths.c) pinned against
closed forms and the reference's own tests (diffraction.rs:52-92,
paths.rs:455-537)."""
import math

import numpy as np
import pytest

from omg_hip import scenes
from helpers import oracle_mesh_from_builder

BRICK = scenes.MaterialDef([0.02, 0.03, 0.04], 0.12, [0.045, 0.008, 0.0012])   # material.rs:45-49


def ka_f64(detour, freq):
    """Kurze-Anderson in float64 from its definition (diffraction.rs:18-26)."""
    n = 2.0 * detour * freq / 343.0
    if n <= -0.1916:
        return 1.0
    x = math.sqrt(2.0 * math.pi * abs(n))
    ratio = 1.0 if x < 1e-3 else (x / math.tanh(x) if n >= 0 else x / math.tan(x))
    return 10.0 ** (-(5.0 + 20.0 * math.log10(ratio)) / 20.0)


def test_knife_edge_matches_its_closed_form(oracle):
    for d in [-0.06, -0.03, -0.01, -1e-3, 1e-9, 1e-4, 0.01, 0.1, 0.5, 2.0, 10.0]:
        for f in (125.0, 790.0, 5600.0):
            assert abs(oracle.knife_edge_amp(d, f) - ka_f64(d, f)) <= 2e-6 * max(ka_f64(d, f), 1e-3)


def test_knife_edge_reference_properties(oracle):
    """diffraction.rs tests: grazing is -5 dB, lows bend more than highs,
    monotonic in the detour, the lit side recovers to unity."""
    a = oracle.knife_edge_amp(1e-9, 1000.0)
    assert abs(20.0 * math.log10(a) + 5.0) < 0.1
    b = oracle.knife_edge_bands(0.5)
    assert b[0] > b[1] > b[2]
    assert 20.0 * math.log10(b[0]) > -12.0 and 20.0 * math.log10(b[2]) < -20.0
    prev = 1.0
    for i in range(1, 40):
        v = oracle.knife_edge_amp(i * 0.1, 790.0)
        assert v < prev
        prev = v
    assert abs(oracle.knife_edge_amp(-0.05, 790.0) - 1.0) < 1e-6
    mid = oracle.knife_edge_amp(-0.02, 790.0)
    assert 0.56 < mid < 1.0


def test_bend_point_minimises_the_detour(oracle):
    rng = np.random.default_rng(3)
    for _ in range(200):
        a, b, s, l = (rng.uniform(-10, 10, 3).astype(np.float32) for _ in range(4))
        p = oracle.bend_point(a, b, s, l)
        ts = np.linspace(0.0, 1.0, 20001)
        pts = a[None].astype(np.float64) + (b - a)[None].astype(np.float64) * ts[:, None]
        cost = np.linalg.norm(pts - s, axis=1) + np.linalg.norm(l - pts, axis=1)
        got = np.linalg.norm(p - s) + np.linalg.norm(l - p)
        assert got <= cost.min() + 1e-4 * max(1.0, cost.min())
        # on the edge segment
        t = np.dot(p - a, b - a) / max(np.dot(b - a, b - a), 1e-12)
        assert -1e-5 <= t <= 1 + 1e-5


def _building(mb, lo, hi, t, m):
    mb.abox(lo, hi, m)
    mb.abox(tuple(x + t for x in lo), tuple(x - t for x in hi), m)


def test_transmission_wall_thickness_is_emergent(oracle):
    """paths.rs:455-484: a 0.24 m brick wall prices like the authored mass
    law, two walls square it; one slab exactly (hits at both faces)."""
    mb = scenes.MeshBuilder()
    m = mb.material(BRICK)
    _building(mb, (0.0, 0.0, 0.0), (8.0, 6.0, 3.0), 0.24, m)
    om = oracle_mesh_from_builder(mb)
    want = np.minimum(np.array(BRICK.transmission) * min(0.15 / 0.24, 2.0), 0.9)
    got = om.path_transmission((-2.0, 3.0, 1.5), (2.0, 3.0, 1.5))
    assert np.all(np.abs(got - want) < 0.02)
    got2 = om.path_transmission((-2.0, 3.0, 1.5), (10.0, 3.0, 1.5))
    assert np.all(np.abs(got2 - want * want) < 0.02)
    # exact: both faces hit, thickness from the segment parameters
    a, b = np.float32([-2.0, 3.0, 1.5]), np.float32([2.0, 3.0, 1.5])
    ts, _, _ = om.segment_hits(a, b)
    assert len(ts) == 2
    th = max((ts[1] - ts[0]) * np.float32(np.linalg.norm(b - a)), np.float32(0.01))
    k = min(np.float32(0.15) / max(th, np.float32(0.02)), np.float32(2.0))
    exact = np.minimum(np.float32(BRICK.transmission) * k, np.float32(0.9)).astype(np.float32)
    assert np.array_equal(got, exact)


def test_edge_extraction_closed_forms(oracle):
    # closed box: 12 creases at 90 degrees (sharpness 0.5), face diagonals coplanar (dropped)
    mb = scenes.MeshBuilder()
    m = mb.material(BRICK)
    mb.abox((0.0, 0.0, 0.0), (4.0, 3.0, 2.0), m)
    om = oracle_mesh_from_builder(mb)
    e = om.path_edges(64)
    assert e.shape[0] == 12
    lens = np.linalg.norm(e[:, 3:6] - e[:, 0:3], axis=1)
    assert np.allclose(e[:, 6], lens * 0.5, rtol=1e-6)
    assert sorted(np.round(lens, 4).tolist()) == sorted([4.0] * 4 + [3.0] * 4 + [2.0] * 4)
    assert np.all(np.diff(e[:, 6]) <= 0)                       # importance order
    assert om.path_edges(5).shape[0] == 5                      # the budget keeps the top
    # one open triangle: three boundary edges at sharpness 1
    mb = scenes.MeshBuilder()
    m = mb.material(BRICK)
    mb.tri((0.0, 0.0, 0.0), (3.0, 0.0, 0.0), (0.0, 4.0, 0.0), m)
    e = oracle_mesh_from_builder(mb).path_edges(64)
    assert sorted(np.round(e[:, 6], 4).tolist()) == [3.0, 4.0, 5.0]


def test_bent_paths_behind_a_building(oracle):
    """paths.rs:486-525: a solid block on a ground slab between source and
    listener: the direct path is blocked, bent paths exist, they favour the
    lows, one goes over the roof and one around the side, and they are
    longer than the straight line."""
    mb = scenes.MeshBuilder()
    m = mb.material(BRICK)
    mb.abox((10.0, -4.0, 0.0), (20.0, 4.0, 4.0), m)
    mb.abox((-40.0, -40.0, -0.5), (60.0, 40.0, 0.0), m)
    om = oracle_mesh_from_builder(mb)
    edges = om.path_edges(64)
    src, lis = np.float32([5.0, 0.0, 1.5]), np.float32([25.0, 0.0, 1.6])
    paths = om.find_paths(edges, src, lis, max_paths=6)
    assert paths[0]["key"] == 0 and paths[0]["gains"][1] < 0.1
    assert len(paths) > 1
    best = paths[1]
    assert best["gains"][0] > 2.0 * best["gains"][2]
    assert any(p["points"][1][2] > 3.5 for p in paths[1:])
    assert any(abs(p["points"][1][1]) > 3.5 and p["points"][1][2] < 3.5 for p in paths[1:])
    assert best["length"] > np.linalg.norm(lis - src) + 0.1
    g1 = [p["gains"][1] for p in paths[1:]]
    assert g1 == sorted(g1, reverse=True)


def test_clear_line_returns_only_direct(oracle):
    """paths.rs:527-536."""
    mb = scenes.MeshBuilder()
    m = mb.material(BRICK)
    mb.abox((0.0, 10.0, 0.0), (4.0, 12.0, 3.0), m)
    om = oracle_mesh_from_builder(mb)
    paths = om.find_paths(om.path_edges(64), (0.0, 0.0, 1.5), (10.0, 0.0, 1.5))
    assert len(paths) == 1 and abs(paths[0]["gains"][1] - 1.0) < 1e-6


# ---- the library's host side (no device): the same numbers as the oracle

def _lib():
    import omg_hip as H
    return H


def test_library_knife_edge_and_bend_point_match_oracle(oracle):
    """The shared numerics the device runs, host-evaluated: bit-identical to
    the oracle's (both round f64 tanh / tan / log10 / pow once to f32)."""
    H = _lib()
    rng = np.random.default_rng(9)
    detours = np.concatenate([rng.uniform(-0.1, 0.1, 3000), rng.uniform(0.0, 30.0, 3000),
                              [-0.0917, 0.0, 1e-4, 1e-9]]).astype(np.float32)
    for d in detours:
        for f in (125.0, 790.0, 5600.0):
            assert np.float32(H.knife_edge_amp(float(d), f)) == np.float32(oracle.knife_edge_amp(float(d), f))
    for _ in range(1000):
        a, b, s, l = (rng.uniform(-30, 30, 3).astype(np.float32) for _ in range(4))
        assert np.array_equal(H.bend_point(a, b, s, l), oracle.bend_point(a, b, s, l))


def test_library_edge_extraction_matches_oracle(oracle):
    """omg_paths_extract_edges (host) against the oracle on the reference's
    test meshes and the city block, budget cuts included."""
    import oracle as O
    H = _lib()
    mb, _, _ = scenes.city_block(seed=2)
    meshes = [mb]
    for build in (lambda m, mb: mb.abox((0.0, 0.0, 0.0), (4.0, 3.0, 2.0), m),
                  lambda m, mb: _building(mb, (0.0, 0.0, 0.0), (8.0, 6.0, 3.0), 0.24, m)):
        b = scenes.MeshBuilder()
        build(b.material(BRICK), b)
        meshes.append(b)
    for b in meshes:
        soup = b.soup()
        pos, idx = O.dedup_tris(soup["tri_verts"])
        om = oracle_mesh_from_builder(b)
        for budget in (0, 5, 64, 512, 3000):
            assert np.array_equal(H.extract_edges(pos, idx, budget), om.path_edges(budget))


def test_library_edge_extraction_refusals():
    """A vertex index out of range is refused; a short output buffer gets
    the top of the graph and OMG_ERR_CAP with the full count."""
    import ctypes as C
    from omg_hip import _lib
    from omg_hip.paths import EDGE_DTYPE
    L = _lib.lib()
    pos = np.float32([[0, 0, 0], [3, 0, 0], [0, 4, 0]])
    bad = np.uint32([[0, 1, 7]])
    n = C.c_uint32()
    out = np.zeros(8, EDGE_DTYPE)
    rc = L.omg_paths_extract_edges(pos.ctypes.data, 3, bad.ctypes.data, 1, 40.0, 64,
                                   out.ctypes.data, 8, C.byref(n))
    assert rc == _lib.ERR_INVALID
    good = np.uint32([[0, 1, 2]])
    rc = L.omg_paths_extract_edges(pos.ctypes.data, 3, good.ctypes.data, 1, 40.0, 64,
                                   out.ctypes.data, 2, C.byref(n))
    assert rc == _lib.ERR_CAP and n.value == 3
    assert sorted(np.round(out[:2]["importance"], 4).tolist()) == [4.0, 5.0]
