"""Scenes built on the device (omg_scene_create_device) against the host
build (omg_scene_create, itself pinned to the oracle's tree in
test_build_cpu.py): the same packed primitives bit for bit, the same
paired and 4-wide nodes (box coordinates equal as values, references
exact), the same tree summary; traces over either are bit-identical.
Refit (omg_scene_refit) keeps the tree and recomputes every box: with
unchanged vertices it reproduces the build, with moved vertices its hits
equal the oracle's on the moved mesh except rays flagged as near-edge or
exact ties (the refit keeps the original tree, whose visiting order can
differ from a fresh build's among hits at equal distance), and it
refuses moves that change the tessellation."""
import sys

import numpy as np
import pytest

import oracle as O
import omg_hip as H
from omg_hip import scenes
from helpers import assert_tree_matches_oracle, oracle_mesh_from_builder

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _mesh(mb):
    soup = mb.soup()
    pos, idx = H.dedup_soup(soup["tri_verts"])
    return dict(positions=pos, indices=idx, tri_material=soup["tri_material"],
                materials=soup["materials"], tri_surface=soup["tri_surface"],
                tri_max_edge=soup["tri_max_edge"])


def _both(ctx, m, nbands=3):
    host = H.Scene.from_mesh(ctx, nbands=nbands, **m)
    dev = H.Scene.from_mesh(ctx, nbands=nbands, device=True, **m)
    return host, dev


def _same_tree(a, b):
    pa, na, n4a, sa = a
    pb, nb, n4b, sb = b
    assert sa == sb, (sa, sb)
    assert np.array_equal(pa.view(np.uint32), pb.view(np.uint32)), "primitives differ"
    # nodes: boxes equal as values, references and padding exact
    for x, y in ((na, nb), (n4a, n4b)):
        assert x.shape == y.shape
        xv, yv = x.view(np.uint32), y.view(np.uint32)
        words = x.shape[1]
        ref_cols = [3, 7, 11, 15] if words == 16 else [c for c in range(32) if c % 8 in (3, 7)]
        box_cols = [c for c in range(words) if c not in ref_cols]
        assert np.array_equal(xv[:, ref_cols], yv[:, ref_cols]), "node references differ"
        assert np.array_equal(x[:, box_cols], y[:, box_cols]), "node boxes differ"


def _soup_builder(seed, n, spread, size, max_edge=None):
    rng = scenes.SplitMix64(seed)
    mb = scenes.MeshBuilder()
    m = mb.material(scenes.MaterialDef([0.1] * 3, 0.1, [0.0] * 3))
    if max_edge is not None:
        mb.coarse(max_edge)
    for _ in range(n):
        b = [rng.uniform(0, spread) for _ in range(3)]
        mb.tri(tuple(b), tuple(x + rng.uniform(-size, size) for x in b),
               tuple(x + rng.uniform(-size, size) for x in b), m)
    return mb


def _case(name):
    if name == "shoebox":
        return scenes.shoebox(seed=1)[0]
    if name == "city":
        return scenes.city_block(seed=2, n_buildings=8, target_tris=6000)[0]
    if name == "soup_tess":
        return _soup_builder(7, 300, 20.0, 6.0)
    if name == "coarse":
        return _soup_builder(9, 40, 30.0, 25.0, max_edge=1.5)
    if name == "stacked":
        # 64 copies of one triangle: every centroid equal, median splits only
        mb = scenes.MeshBuilder()
        m = mb.material(scenes.MaterialDef([0.2] * 3, 0.0, [0.0] * 3))
        for k in range(64):
            mb.tri((0.0, 0.0, 0.0), (1.0, 0.0, 0.0), (0.0, 1.0, 0.0), m)
        return mb
    if name.startswith("tiny"):
        return _soup_builder(11, int(name[4:]), 5.0, 1.0)
    raise KeyError(name)


@pytest.mark.parametrize("name", ["shoebox", "city", "soup_tess", "coarse", "stacked", "tiny1",
                                  "tiny4", "tiny5", "tiny9"])
def test_device_build_equals_host_build(ctx, oracle, name):
    mb = _case(name)
    m = _mesh(mb)
    _say(f"case device build {name}: {len(m['indices'])} tris")
    host, dev = _both(ctx, m)
    try:
        a, b = host.export(), dev.export()
        _same_tree(a, b)
        # and the device tree against the oracle's own build, node for node
        pb, nb, _, tb = b
        assert_tree_matches_oracle(pb, nb, tb["root_ref"], oracle_mesh_from_builder(mb))
        _say(f"  {a[3]['n_prims']} prims, {a[3]['n_nodes']} nodes, {a[3]['n_nodes4']} 4-wide, "
             f"depth {a[3]['max_depth']}: identical to the host build and the oracle")
    finally:
        host.close()
        dev.close()


def test_device_build_empty_scene(ctx):
    mat = scenes.MaterialDef([0.1] * 3, 0.1, [0.0] * 3)
    m = dict(positions=np.zeros((0, 3), np.float32), indices=np.zeros((0, 3), np.uint32),
             tri_material=np.zeros(0, np.uint16), materials=[mat])
    host, dev = _both(ctx, m)
    try:
        _same_tree(host.export(), dev.export())
    finally:
        host.close()
        dev.close()


def test_device_build_large_city(ctx):
    mb = scenes.city_block(seed=3, n_buildings=60, target_tris=200_000)[0]
    m = _mesh(mb)
    _say(f"case device build large city: {len(m['indices'])} tris")
    host, dev = _both(ctx, m)
    try:
        _same_tree(host.export(), dev.export())
    finally:
        host.close()
        dev.close()


def test_trace_over_device_build_is_identical(ctx):
    mb, bounds, boxes = scenes.city_block(seed=5, n_buildings=8, target_tris=6000)
    src = np.vstack([scenes.indoor_points(51, 2, boxes), scenes.outdoor_points(52, 1, bounds, boxes)])
    lis = scenes.outdoor_points(54, 2, bounds, boxes)
    _say("case trace over device build: city 6k, 3 sources x 2048 rays, widths 2 and 4")
    host, dev = _both(ctx, _mesh(mb))
    kw = dict(n_rays=2048, nbands=3, seed=77, bin_dt=0.001, max_time=2.0)
    try:
        for width in (2, 4):
            try:
                ctx.set_launch_config(512, 16, 0, width)
                a = H.trace_world(host, H.trace_params(**kw), src, lis, records=True)
                b = H.trace_world(dev, H.trace_params(**kw), src, lis, records=True)
            finally:
                ctx.set_launch_config()
            assert np.array_equal(a.records, b.records)
            assert np.array_equal(a.bins, b.bins) and np.array_equal(a.dirs, b.dirs)
    finally:
        host.close()
        dev.close()


def test_refit_unchanged_vertices_reproduces_the_build(ctx):
    mb = _case("soup_tess")
    m = _mesh(mb)
    dev = H.Scene.from_mesh(ctx, device=True, **m)
    try:
        before = dev.export()
        dev.refit(m["positions"])
        _same_tree(before, dev.export())
    finally:
        dev.close()


def test_refit_moved_vertices_matches_the_oracle(ctx, oracle):
    mb, bounds, boxes = scenes.city_block(seed=5, n_buildings=8, target_tris=6000)
    m = _mesh(mb)
    dev = H.Scene.from_mesh(ctx, device=True, **m)
    moved = m["positions"] + np.array([0.75, -0.5, 0.25], np.float32)
    mats = [(list(x.absorption), x.scattering, list(x.transmission)) for x in m["materials"]]
    om = O.Mesh(moved, m["indices"], m["tri_material"], mats, 3, m["tri_surface"],
                m["tri_max_edge"])
    om0 = oracle_mesh_from_builder(mb)
    src = np.vstack([scenes.indoor_points(51, 2, boxes), scenes.outdoor_points(52, 1, bounds, boxes)])
    src = src + np.array([0.75, -0.5, 0.25])
    lis = scenes.outdoor_points(54, 2, bounds, boxes) + np.array([0.75, -0.5, 0.25])
    try:
        if om.sizes()[1] != om0.sizes()[1]:
            with pytest.raises(H.OmgError):
                dev.refit(moved)
            _say("case refit moved city: the move changes the tessellation; refused")
            return
        dev.refit(moved)
        _, _, _, t = dev.export()
        kw = dict(n_rays=2000, nbands=3, seed=91, bin_dt=0.001, max_time=2.0)
        gpu = H.trace_world(dev, H.trace_params(**kw), src, lis, records=True)
        orec = []
        for si, s in enumerate(src):
            p = O.trace_params(rng_mode=O.RNG_COUNTER, src_index=si, **kw)
            orec.append(O.trace(O.mesh_geom(om), s, lis, p, want_rays=True)["recs"])
        orec = np.vstack(orec)
        flagged = (orec[:, 3] & 3) != 0
        diff = np.any(gpu.records != orec, axis=1)
        _say(f"case refit moved city: {diff.sum()} of {flagged.sum()} flagged rays differ, "
             f"{(diff & ~flagged).sum()} unflagged")
        assert not np.any(diff & ~flagged)
        # the refit root box is the moved mesh's: the oracle's root node
        onodes, _ = om.export()
        assert np.array_equal(np.float32(t["root_min"]), onodes[0, 0:3])
        assert np.array_equal(np.float32(t["root_max"]), onodes[0, 3:6])
    finally:
        dev.close()


def test_refit_refusals(ctx):
    mb = _case("soup_tess")
    m = _mesh(mb)
    dev = H.Scene.from_mesh(ctx, device=True, **m)
    host = H.Scene.from_mesh(ctx, **m)
    try:
        before = dev.export()
        with pytest.raises(H.OmgError):
            dev.refit(m["positions"] * 3.0)   # longer edges: more patches
        _same_tree(before, dev.export())
        with pytest.raises(H.OmgError):
            dev.refit(m["positions"][:-1])    # another vertex count
        with pytest.raises(H.OmgError):
            host.refit(m["positions"])        # built on the host
    finally:
        dev.close()
        host.close()


def test_device_build_rejects_bad_input(ctx):
    m = _mesh(_case("tiny5"))
    bad = dict(m)
    bad["indices"] = m["indices"].copy()
    bad["indices"][2, 1] = len(m["positions"])
    with pytest.raises(H.OmgError):
        H.Scene.from_mesh(ctx, device=True, **bad)
    bad = dict(m)
    bad["tri_material"] = m["tri_material"].copy()
    bad["tri_material"][1] = len(m["materials"])
    with pytest.raises(H.OmgError):
        H.Scene.from_mesh(ctx, device=True, **bad)
    bad = dict(m)
    bad["positions"] = m["positions"].copy()
    bad["positions"][0, 0] = np.nan
    with pytest.raises(H.OmgError):
        H.Scene.from_mesh(ctx, device=True, **bad)
