"""GPU parity at the edges of the input space: an empty mesh, one ray,
ragged ray counts, the listener and overlay caps, 1 and 8 bands, a long
bounce limit with no energy floor, and a batch whose queries are as
uneven as they get.  Every case is checked ray by ray against the oracle."""
import sys

import numpy as np
import pytest

import oracle as O
import omg_hip as H
from omg_hip import scenes
from helpers import echogram_rel_errors, mat_tuple, oracle_mesh_from_builder

pytestmark = pytest.mark.gpu

ECHO_TOL = 1e-4


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _check(gpu, geom, sources, listeners, n_rays, **kw):
    rec = gpu.records.reshape(len(sources), n_rays, 4)
    for si, s in enumerate(sources):
        p = O.trace_params(n_rays, rng_mode=O.RNG_COUNTER, src_index=si, **kw)
        o = O.trace(geom, s, listeners, p, want_rays=True, want_f64=True)
        assert np.array_equal(rec[si], o["recs"]), f"source {si}"
        eb, ed = echogram_rel_errors(gpu.bins[si], gpu.dirs[si], o["bins64"], o["dirs64"],
                                     kw.get("dir_band", 1))
        assert eb <= ECHO_TOL and ed <= ECHO_TOL, (si, eb, ed)


def test_empty_mesh_every_ray_escapes(ctx, oracle):
    mb = scenes.MeshBuilder()
    mb.material(H.MaterialDef([0.1] * 3, 0.1, [0.0] * 3))
    scene = H.Scene.from_builder(ctx, mb)
    assert scene.info()["n_tris"] == 0
    kw = dict(nbands=3, seed=3, bin_dt=0.001, max_time=1.0)
    src, lis = np.array([[0.0, 0.0, 0.0]]), np.array([[3.0, 0.0, 0.0], [0.0, -2.0, 1.0]])
    _say("case empty mesh")
    gpu = H.trace_world(scene, H.trace_params(500, **kw), src, lis, records=True)
    assert np.all(gpu.records[:, 0] == 1) and np.all(gpu.records[:, 1] == 0)
    _check(gpu, O.mesh_geom(oracle_mesh_from_builder(mb)), src, lis, 500, **kw)
    scene.close()


@pytest.mark.parametrize("n_rays,n_src", [(1, 1), (1, 5), (3, 2), (257, 3), (1023, 1)])
def test_ragged_ray_counts(ctx, oracle, n_rays, n_src):
    mb, bounds, boxes = scenes.city_block(seed=61, n_buildings=5, target_tris=2500)
    src = scenes.outdoor_points(611, n_src, bounds, boxes)
    lis = scenes.outdoor_points(612, 2, bounds, boxes)
    kw = dict(nbands=3, seed=17, bin_dt=0.001, max_time=2.0)
    _say(f"case ragged: {n_src} sources x {n_rays} rays")
    scene = H.Scene.from_builder(ctx, mb)
    for width in (2, 4):
        try:
            ctx.set_launch_config(width=width)
            gpu = H.trace_world(scene, H.trace_params(n_rays, **kw), src, lis, records=True)
        finally:
            ctx.set_launch_config()
        _check(gpu, O.mesh_geom(oracle_mesh_from_builder(mb)), src, lis, n_rays, **kw)
    scene.close()


def test_listener_and_overlay_caps_with_eight_bands(ctx, oracle):
    """64 listeners and 64 overlay boxes (the caps), 8 bands."""
    mb, src, _, _ = scenes.shoebox(seed=7, nbands=8)
    rng = np.random.default_rng(5)
    lis = np.column_stack([rng.uniform(1, 11, 64), rng.uniform(1, 7, 64), rng.uniform(1, 3, 64)])
    mat = H.MaterialDef([0.05 * (b + 1) for b in range(8)], 0.25, [0.3 / (b + 1) for b in range(8)])
    panels = []
    for i in range(64):
        x, y = 1.0 + (i % 8) * 1.3, 1.0 + (i // 8) * 0.8
        panels.append(((x, y, 0.5), (x + 0.3, y + 0.3, 1.5), mat))
    kw = dict(nbands=8, seed=23, dir_band=5, bin_dt=0.001, max_time=1.0)
    _say("case caps: 64 listeners, 64 overlays, 8 bands")
    scene = H.Scene.from_builder(ctx, mb, nbands=8)
    gpu = H.trace_world(scene, H.trace_params(400, **kw), [src], lis, overlays=panels,
                        records=True)
    om = oracle_mesh_from_builder(mb, nbands=8)
    g = O.mesh_geom(om, [list(p[0]) + list(p[1]) for p in panels], [mat_tuple(p[2]) for p in panels])
    _check(gpu, g, [src], lis, 400, **kw)
    scene.close()


def test_one_band_and_long_bounce_limit(ctx, oracle):
    """1 band; 1000 bounces with no energy floor: rays end on the time limit."""
    mb, src, lis, _ = scenes.shoebox(seed=8, nbands=1)
    kw = dict(nbands=1, seed=29, dir_band=0, bin_dt=0.001, max_time=0.5, max_bounces=1000,
              floor_rel=0.0)
    _say("case 1 band, 1000 bounces, no floor")
    scene = H.Scene.from_builder(ctx, mb, nbands=1)
    gpu = H.trace_world(scene, H.trace_params(300, **kw), [src], [lis], records=True)
    assert np.any(gpu.records[:, 1] == 1)       # some rays stop on the time limit
    _check(gpu, O.mesh_geom(oracle_mesh_from_builder(mb, nbands=1)), [src], [lis], 300, **kw)
    scene.close()


def test_batch_of_extremes(ctx, oracle):
    """One batch mixing a 1-ray query, a 1-source 8-band box, a query on an
    empty mesh and a 20k-ray query: each equals its plain launch."""
    empty = scenes.MeshBuilder()
    empty.material(H.MaterialDef([0.1] * 3, 0.1, [0.0] * 3))
    box8, s8, l8, _ = scenes.shoebox(seed=9, nbands=8)
    city, bounds, boxes = scenes.city_block(seed=62, n_buildings=5, target_tris=2500)
    se = H.Scene.from_builder(ctx, empty)
    s8s = H.Scene.from_builder(ctx, box8, nbands=8)
    sc = H.Scene.from_builder(ctx, city)
    pts = scenes.outdoor_points(621, 4, bounds, boxes)
    qs = [H.QuerySpec(sc, H.trace_params(1, 3, seed=1, bin_dt=0.001, max_time=1.0), pts[:1], pts[1:2]),
          H.QuerySpec(s8s, H.trace_params(700, 8, seed=2, bin_dt=0.001, max_time=1.0), [s8], [l8]),
          H.QuerySpec(se, H.trace_params(50, 3, seed=3, bin_dt=0.001, max_time=1.0),
                      [[0.0, 0.0, 0.0]], [[2.0, 0.0, 0.0]]),
          H.QuerySpec(sc, H.trace_params(20000, 3, seed=4, bin_dt=0.001, max_time=2.0), pts[:2],
                      pts[2:4])]
    _say("case batch of extremes")
    b = H.Batch(ctx, qs, records=True)
    assert b.shape["width"] == 4
    refs = [H.trace_world(q.scene, q.params, q.sources, q.listeners, records=True, counters=True)
            for q in qs]
    # the binary walk regenerating between bounces, and the default 4-wide
    # walk regenerating at walk steps (band capacity 8, an empty mesh whose
    # walks end before they start, a one-ray query)
    for shape in [(1, 8, 8, 1, 2), (1, 12, 6, 2, 4)]:
        b.set_shape(*shape)
        b.launch()
        for i, ref in enumerate(refs):
            got = b.result(i)
            assert np.array_equal(got.records, ref.records), (shape, i)
            eb, ed = echogram_rel_errors(got.bins, got.dirs, ref.bins.astype(np.float64),
                                         ref.dirs.astype(np.float64))
            assert eb <= ECHO_TOL and ed <= ECHO_TOL, (shape, i, eb, ed)
    b.close()
    for s in (se, s8s, sc):
        s.close()
