"""GPU parity of the batched persistent launch.

The batch has no reference counterpart of its own: each query of a batch
must give what the same query gives through the single-job entry point
(omg_trace_world), which is itself checked against the oracle.  Per-ray
records must be bit-identical, echograms within 1e-4 relative (the two
sides sum the same f32 contributions in different atomic orders), the
traversal counters exactly equal.  One query is also checked against the
oracle directly.  Every launch shape, graph replays and the autotuner
are covered.
"""
import sys

import numpy as np
import pytest

import oracle as O
import omg_hip as H
from omg_hip import scenes
from helpers import echogram_rel_errors, oracle_mesh_from_builder

pytestmark = pytest.mark.gpu

ECHO_TOL = 1e-4


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def pool(ctx):
    """Three scenes: a 12-triangle box (3 bands), small city blocks of 3 and 6 bands."""
    box_mb, box_src, box_lis, _ = scenes.shoebox(seed=1)
    c3, b3, bx3 = scenes.city_block(seed=31, n_buildings=6, target_tris=3000, nbands=3)
    c6, b6, bx6 = scenes.city_block(seed=32, n_buildings=8, target_tris=6000, nbands=6)
    s_box = H.Scene.from_builder(ctx, box_mb, nbands=3)
    s3 = H.Scene.from_builder(ctx, c3, nbands=3)
    s6 = H.Scene.from_builder(ctx, c6, nbands=6)
    pts3 = np.vstack([scenes.indoor_points(301, 4, bx3), scenes.outdoor_points(302, 4, b3, bx3)])
    pts6 = np.vstack([scenes.indoor_points(601, 4, bx6), scenes.outdoor_points(602, 4, b6, bx6)])
    yield dict(box=(s_box, box_mb, np.array([box_src]), np.array([box_lis])),
               c3=(s3, c3, pts3[:3], pts3[3:6]), c6=(s6, c6, pts6[:4], pts6[4:8]))
    for s in (s_box, s3, s6):
        s.close()


def _queries(pool):
    """Uneven ray counts (not multiples of 64), 1-4 sources, 1-4 listeners,
    3 and 6 bands, one query with an overlay panel."""
    panel = ((2.0, 2.0, 0.0), (2.2, 6.0, 3.0),
             H.MaterialDef([0.1, 0.2, 0.3], 0.2, [0.5, 0.4, 0.3]))
    out = []
    spec = [("box", 1000, 1, 1, 2.0, ()), ("c3", 333, 3, 2, 1.0, ()), ("c6", 257, 4, 4, 3.0, ()),
            ("c3", 64, 1, 3, 2.0, (panel,)), ("box", 4096, 1, 1, 2.0, ()), ("c6", 1, 2, 1, 1.0, ()),
            ("c3", 2000, 2, 1, 4.0, ())]
    for i, (name, n_rays, ns, nl, tmax, ovl) in enumerate(spec):
        scene, _, src, lis = pool[name]
        nb = 6 if name == "c6" else 3
        p = H.trace_params(n_rays, nbands=nb, seed=0xB00 + i, dir_band=1, bin_dt=0.001,
                           max_time=tmax, floor_rel=1e-6,
                           energy=[1.0 + 0.1 * b for b in range(nb)])
        out.append(H.QuerySpec(scene, p, src[:ns], lis[:nl], ovl))
    return out


def _plain(q):
    """The query through omg_trace_world; counting makes it walk the binary tree."""
    return H.trace_world(q.scene, q.params, q.sources, q.listeners, q.overlays, records=True,
                         counters=True)


def _same(a, b, label):
    """a's rays equal b's, every one (the 4-wide walks visit the binary
    walk's leaves in its order, so flagged rays are no exception), and the
    echograms agree within the f32 summation-order tolerance."""
    diff = np.any(a.records != b.records, axis=1)
    mism = np.nonzero(diff)[0]
    assert mism.size == 0, f"{label}: {mism.size} rays differ, first {mism[:5]}"
    eb, ed = echogram_rel_errors(a.bins, a.dirs, b.bins.astype(np.float64),
                                 b.dirs.astype(np.float64))
    assert eb <= ECHO_TOL and ed <= ECHO_TOL, (label, eb, ed)
    return eb, ed


def test_batch_matches_plain_per_query(ctx, pool):
    qs = _queries(pool)
    _say(f"case batch vs plain: {len(qs)} queries")
    plain = [_plain(q) for q in qs]
    b = H.Batch(ctx, qs, records=True, counters=True)
    rays, items = b.size()
    assert rays == sum(q.params.n_rays * len(q.sources) for q in qs)
    assert items == sum(len(q.sources) * ((q.params.n_rays + 63) // 64) for q in qs)
    b.launch()
    worst = 0.0
    for i, q in enumerate(qs):
        eb, ed = _same(b.result(i), plain[i], f"query {i}")
        worst = max(worst, eb, ed)
    c = b.counters()
    for k in ("aabb_tests", "tri_tests", "bounce_rays"):
        assert c[k] == sum(p.counters[k] for p in plain), k
    _say(f"  records bit-identical, counters equal, worst echogram rel diff {worst:.2e}")
    b.close()


def test_counting_batch_on_a_quad_shape_walks_the_binary_tree(ctx, pool):
    """Counters are the reference walk's: a counting batch set to the
    whole-stack quad shape runs its binary-tree counterpart."""
    qs = _queries(pool)
    _say("case counting batch, quad regen shape")
    plain = [_plain(q) for q in qs]
    b = H.Batch(ctx, qs, records=True, counters=True)
    b.set_shape(4, 40, 6, 1, 4)
    b.launch()
    for i in range(len(qs)):
        _same(b.result(i), plain[i], f"query {i}")
    c = b.counters()
    for k in ("aabb_tests", "tri_tests", "bounce_rays"):
        assert c[k] == sum(p.counters[k] for p in plain), k
    b.close()


def test_batch_query_against_oracle(ctx, pool):
    """The 6-band query straight against the oracle's f64 re-accumulation."""
    qs = _queries(pool)
    q = qs[2]
    _say("case batch query 2 (6 bands, 4 sources, 4 listeners) vs oracle")
    b = H.Batch(ctx, [qs[0], q, qs[1]], records=True)
    b.launch()
    got = b.result(1)
    b.close()
    _, mb, _, _ = pool["c6"]
    g = O.mesh_geom(oracle_mesh_from_builder(mb, nbands=6))
    rec = got.records.reshape(len(q.sources), q.params.n_rays, 4)
    for si, s in enumerate(q.sources):
        p = O.trace_params(q.params.n_rays, nbands=6, rng_mode=O.RNG_COUNTER, seed=q.params.seed,
                           src_index=si, dir_band=1, bin_dt=0.001, max_time=q.params.max_time,
                           floor_rel=1e-6, energy=[1.0 + 0.1 * b for b in range(6)])
        o = O.trace(g, s, q.listeners, p, want_rays=True, want_f64=True)
        # the default batch walks the 4-wide tree: every ray is the oracle's
        diff = np.any(rec[si] != o["recs"], axis=1)
        assert not diff.any(), f"source {si}: {int(diff.sum())} rays differ"
        eb, ed = echogram_rel_errors(got.bins[si], got.dirs[si], o["bins64"], o["dirs64"], 1)
        assert eb <= ECHO_TOL and ed <= ECHO_TOL, (si, eb, ed)


@pytest.mark.parametrize("lanes,stack,wps,regen,width",
                         [(1, 16, 2, 0, 2), (1, 16, 8, 0, 2), (4, 16, 4, 0, 2), (1, 32, 4, 0, 2),
                          (4, 32, 6, 0, 2), (1, 16, 4, 1, 2), (1, 16, 8, 1, 2), (1, 8, 2, 1, 2),
                          (1, 12, 6, 1, 2), (1, 8, 8, 1, 4), (1, 16, 4, 1, 4), (4, 16, 4, 0, 4),
                          (1, 16, 6, 0, 4), (4, 16, 6, 1, 2), (4, 40, 7, 1, 4), (4, 40, 4, 0, 4),
                          (1, 12, 6, 2, 4), (1, 16, 4, 2, 4), (1, 12, 1, 2, 4)])
def test_batch_shapes_and_replays(ctx, pool, lanes, stack, wps, regen, width):
    qs = _queries(pool)
    _say(f"case batch shape lanes={lanes} stack_lds={stack} waves/simd={wps} regen={regen} "
         f"width={width}")
    ref = [_plain(q) for q in qs]
    b = H.Batch(ctx, qs, records=True)
    b.set_shape(lanes, stack, wps, regen, width)
    assert b.shape["lanes"] == lanes and b.shape["stack_lds"] == stack
    assert b.shape["regen"] == regen and b.shape["width"] == width
    for rep in range(2):
        b.launch()
        for i in range(len(qs)):
            _same(b.result(i), ref[i], f"replay {rep} query {i}")
    b.close()


def test_batch_tune(ctx, pool):
    qs = _queries(pool)
    _say("case batch autotune")
    b = H.Batch(ctx, qs, records=True)
    shape = b.tune(stride=2)
    assert shape["lanes"] in (1, 4) and shape["stack_lds"] in (8, 12, 16, 32, 40)
    assert shape["width"] in (2, 4)
    assert 1 <= shape["waves_per_simd"] <= 8 and shape["ms"] > 0
    assert shape["regen"] in (0, 1, 2)
    b.launch()
    for i, q in enumerate(qs):
        _same(b.result(i), _plain(q), f"tuned query {i}")
    # a second batch of the same class starts from the cached choice
    b2 = H.Batch(ctx, qs, records=True)
    s2 = b2.shape
    assert (s2["lanes"], s2["stack_lds"], s2["waves_per_simd"]) == \
        (shape["lanes"], shape["stack_lds"], shape["waves_per_simd"])
    b2.close()
    b.close()


def test_batch_refusals(ctx, pool):
    scene = pool["c3"][0]
    good = H.trace_params(100, nbands=3)
    src = pool["c3"][2][:1]
    lis = pool["c3"][3][:1]
    with pytest.raises(H.OmgError) as e:
        H.Batch(ctx, [H.QuerySpec(scene, H.trace_params(100, nbands=6), src, lis)])
    assert e.value.code == H._lib.ERR_INVALID
    with pytest.raises(H.OmgError) as e:
        H.Batch(ctx, [H.QuerySpec(scene, good, src, np.zeros((65, 3), np.float32))])
    assert e.value.code == H._lib.ERR_CAP
    b = H.Batch(ctx, [H.QuerySpec(scene, good, src, lis)])
    with pytest.raises(H.OmgError):
        b.counters()          # created without counters
    with pytest.raises(H.OmgError):
        b.set_shape(2, 16, 4)
    with pytest.raises(H.OmgError):
        b.set_shape(4, 8, 4, regen=1)    # a regenerating quad keeps 16 entries (binary tree)
    with pytest.raises(H.OmgError):
        b.set_shape(4, 40, 4, regen=1, width=2)   # ... or 40 (4-wide tree only)
    with pytest.raises(H.OmgError):
        b.set_shape(1, 40, 4, regen=1, width=4)   # the whole-stack walk runs on a quad
    with pytest.raises(H.OmgError):
        b.set_shape(1, 32, 4, regen=1)   # and keeps at most 16 stack entries in LDS
    with pytest.raises(H.OmgError):
        b.set_shape(1, 32, 4, regen=0, width=4)   # the 4-wide item walk keeps 16 entries
    with pytest.raises(H.OmgError):
        b.set_shape(4, 40, 4, regen=2, width=4)   # rays replaced at walk steps: one lane per ray,
    with pytest.raises(H.OmgError):
        b.set_shape(1, 12, 4, regen=2, width=2)   # over the 4-wide tree,
    with pytest.raises(H.OmgError):
        b.set_shape(1, 8, 4, regen=2, width=4)    # with 12 or 16 stack entries in LDS
    with pytest.raises(H.OmgError):
        b.set_shape(1, 12, 4, regen=3, width=4)
    b.close()


def test_plain_launch_routed_to_the_persistent_kernel(ctx, pool):
    """omg_trace_world of a job at or above the routing threshold runs on
    the persistent regenerating kernel: every query of the uneven mix gives
    the single-job kernel's rays, counters and echograms (binary walk), and
    so does the 4-wide walk, routed or not."""
    qs = _queries(pool)
    _say(f"case plain routed vs single-job kernel: {len(qs)} queries")
    ctx.set_route_rays(0)
    single = [_plain(q) for q in qs]
    wide1 = [H.trace_world(q.scene, q.params, q.sources, q.listeners, q.overlays, records=True)
             for q in qs]
    ctx.set_route_rays(1)
    try:
        routed = [_plain(q) for q in qs]
        wide2 = [H.trace_world(q.scene, q.params, q.sources, q.listeners, q.overlays,
                               records=True) for q in qs]
    finally:
        ctx.set_route_rays()
    for i in range(len(qs)):
        _same(routed[i], single[i], f"query {i}")
        for k in ("aabb_tests", "tri_tests", "bounce_rays"):
            assert routed[i].counters[k] == single[i].counters[k], (i, k)
        _same(wide2[i], single[i], f"query {i} 4-wide")
        _same(wide1[i], single[i], f"query {i} 4-wide single-job")
