"""GPU parity of the late-field tracer against the CPU oracle (counter RNG mode).

Per ray: bounce count, termination reason, the FNV hash of the hit
triangle ids and the near-edge/tie flags must be bit-identical.  Echogram
bins must agree with the oracle's f64 re-accumulation within 1e-4
relative (the f32 error of both sides is reported), RT60 within 1 %.
"""
import sys

import numpy as np
import pytest

import oracle as O
import omg_hip as H
from omg_hip import scenes
from helpers import (GOLDEN_AVG, GOLDEN_CONFIGS, GOLDEN_RAYS, echogram_rel_errors,
                     golden_violations, mat_tuple, oracle_mesh_from_builder)

pytestmark = pytest.mark.gpu

ECHO_TOL = 1e-4
RT60_TOL = 0.01


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def oracle_trace_sources(geom, sources, listeners, **kw):
    """Oracle traces of every source (src_index = source position)."""
    outs = []
    for si, s in enumerate(sources):
        p = O.trace_params(rng_mode=O.RNG_COUNTER, src_index=si, **kw)
        outs.append(O.trace(geom, s, listeners, p, want_rays=True, want_f64=True))
    return outs


def compare(gpu, outs, n_rays, dir_band=1, label=""):
    rec = gpu.records.reshape(len(outs), n_rays, 4)
    worst_b = worst_d = 0.0
    worst_b32 = 0.0
    for si, o in enumerate(outs):
        ref = o["recs"]
        mism = np.nonzero(np.any(rec[si] != ref, axis=1))[0]
        assert mism.size == 0, (f"{label} source {si}: {mism.size} rays differ, first {mism[:5]}: "
                                f"gpu {rec[si][mism[:3]].tolist()} oracle {ref[mism[:3]].tolist()}")
        eb, ed = echogram_rel_errors(gpu.bins[si], gpu.dirs[si], o["bins64"], o["dirs64"], dir_band)
        eb32, _ = echogram_rel_errors(o["bins"], o["dirs"], o["bins64"], o["dirs64"], dir_band)
        worst_b, worst_d, worst_b32 = max(worst_b, eb), max(worst_d, ed), max(worst_b32, eb32)
    _say(f"{label}: rays bit-identical; echogram rel err vs f64: gpu bins {worst_b:.2e} "
         f"dirs {worst_d:.2e} (oracle f32 bins {worst_b32:.2e})")
    assert worst_b <= ECHO_TOL and worst_d <= ECHO_TOL


def test_c1_shoebox_mesh_parity(ctx, oracle):
    mb, src, lis, _ = scenes.shoebox(seed=1)
    _say("case c1 shoebox mesh: 12 tris, 4096 rays, 3 bands, 1 ms x 2 s, floor 1e-6")
    kw = dict(n_rays=4096, nbands=3, max_bounces=64, seed=0xC1, bin_dt=0.001, max_time=2.0,
              floor_rel=1e-6)
    scene = H.Scene.from_builder(ctx, mb)
    gpu = H.trace_world(scene, H.trace_params(**kw), [src], [lis], records=True, counters=True)
    om = oracle_mesh_from_builder(mb)
    g = O.mesh_geom(om)
    outs = oracle_trace_sources(g, [src], [lis], **kw)
    compare(gpu, outs, 4096, label="c1")
    assert gpu.counters["bounce_rays"] == int(outs[0]["stats"][2])
    assert gpu.counters["aabb_tests"] == int(outs[0]["stats"][0])
    assert gpu.counters["tri_tests"] == int(outs[0]["stats"][1])
    rt_g, _ = H.estimate_reverb(gpu.bins[0, 0], 0.001)
    rt_o, _ = O.estimate_reverb(outs[0]["bins64"][0].astype(np.float32), 0.001)
    assert np.all(np.abs(rt_g / rt_o - 1.0) <= RT60_TOL), (rt_g, rt_o)


@pytest.mark.parametrize("cfg", GOLDEN_CONFIGS, ids=[c[0] for c in GOLDEN_CONFIGS])
def test_box_parity_and_goldens(ctx, oracle, cfg):
    name, size, walls, src, lis, expect = cfg
    _say(f"case box {name}")
    wt = [mat_tuple(w) for w in walls]
    g = O.box_geom(size, wt)
    n = H.nbins(3.0, 0.010)
    metrics = []
    for rep in range(6):
        # one golden measurement = GOLDEN_AVG averaged traces (trace_golden.rs:125-143)
        avg_b = np.zeros((n, 3), np.float32)
        avg_d = np.zeros((n, 3), np.float32)
        for k in range(GOLDEN_AVG):
            kw = dict(n_rays=GOLDEN_RAYS, nbands=3, seed=0x5EED0000 + 16 * rep + k)
            gpu = H.trace_box(ctx, size, walls, H.trace_params(**kw), [src], [lis], records=True)
            if rep == 0:
                outs = oracle_trace_sources(g, [src], [lis], **kw)
                compare(gpu, outs, GOLDEN_RAYS, label=f"box {name} k={k}")
            O.ema(avg_b, avg_d, gpu.bins[0, 0], gpu.dirs[0, 0], np.float32(1.0 / (k + 1)))
        rt, lv = H.estimate_reverb(avg_b, 0.010)
        _, an = H.late_direction(avg_b, avg_d, 0.010, 0.05)
        metrics.append(list(rt) + [20 * np.log10(max(v, 1e-9)) for v in lv] + [an])
    # the goldens are ~1.5 sigma of one measurement wide: hold the mean of six
    m = np.mean(metrics, axis=0)
    bad = golden_violations(m, expect)
    assert not bad, f"{name}: {bad}"


def test_city_block_multi_listener_parity(ctx, oracle):
    mb, bounds, boxes = scenes.city_block(seed=2, n_buildings=6, target_tris=4000)
    lis = scenes.outdoor_points(21, 4, bounds, boxes)
    src = scenes.outdoor_points(22, 2, bounds, boxes)
    _say(f"case city block: {mb.n_tris} tris, 2 sources x 1024 rays x 4 listeners")
    kw = dict(n_rays=1024, nbands=3, seed=77, bin_dt=0.001, max_time=2.0)
    scene = H.Scene.from_builder(ctx, mb)
    gpu = H.trace_world(scene, H.trace_params(**kw), src, lis, records=True)
    om = oracle_mesh_from_builder(mb)
    outs = oracle_trace_sources(O.mesh_geom(om), src, lis, **kw)
    compare(gpu, outs, 1024, label="city block")


@pytest.mark.parametrize("shape", [(0, 16, 1), (512, 32, 1), (0, 16, 4), (2048, 16, 4),
                                   (512, 32, 4)])
def test_launch_shapes_give_identical_rays(ctx, oracle, shape):
    """Every launch shape (LDS node prefix, LDS stack depth, lanes per ray)
    reproduces the oracle ray for ray: results do not depend on tuning."""
    nc, sk, lanes = shape
    mb, bounds, boxes = scenes.city_block(seed=5, n_buildings=8, target_tris=6000)
    src = np.vstack([scenes.indoor_points(51, 2, boxes), scenes.outdoor_points(52, 1, bounds, boxes)])
    lis = np.vstack([scenes.indoor_points(53, 3, boxes), scenes.outdoor_points(54, 2, bounds, boxes)])
    _say(f"case launch shape node_cache={nc} stack_lds={sk} lanes={lanes}")
    kw = dict(n_rays=1000, nbands=3, seed=31, bin_dt=0.001, max_time=2.0)
    scene = H.Scene.from_builder(ctx, mb)
    try:
        ctx.set_launch_config(nc, sk, lanes, 2)
        gpu = H.trace_world(scene, H.trace_params(**kw), src, lis, records=True, counters=True)
    finally:
        ctx.set_launch_config()
    om = oracle_mesh_from_builder(mb)
    outs = oracle_trace_sources(O.mesh_geom(om), src, lis, **kw)
    compare(gpu, outs, 1000, label=f"shape {shape}")
    assert gpu.counters["aabb_tests"] == sum(int(o["stats"][0]) for o in outs)
    assert gpu.counters["tri_tests"] == sum(int(o["stats"][1]) for o in outs)
    assert gpu.counters["bounce_rays"] == sum(int(o["stats"][2]) for o in outs)


@pytest.mark.parametrize("scene_name", ["city", "c1"])
@pytest.mark.parametrize("shape", [(0, 16, 1, 4, 0, 0), (256, 16, 1, 4, 0, 0), (0, 16, 4, 4, 0, 0),
                                   (256, 16, 4, 4, 0, 0), (256, 32, 4, 4, 0, 0),
                                   (256, 16, 4, 4, 1, 16), (256, 16, 4, 4, 8, 16),
                                   (256, 16, 4, 4, 1, 8), (256, 16, 4, 4, 8, 8),
                                   (0, 16, 16, 16, 0, 0), (512, 16, 16, 16, 0, 0)])
def test_width4_walk_gives_the_binary_walks_hits(ctx, oracle, scene_name, shape):
    """The 4-wide and 16-wide nodes are visited in the binary walk's order,
    so every ray - flagged near-edge or exact-tie hits included - has the
    oracle's record, and every bounce the binary walk's hit id and
    distance.  Shapes (LDS node prefix, stack, lanes, width, hand-off
    bounce, hand-off lanes): one lane per ray; a quad with the whole stack
    in LDS (16 requested); a quad with 16 LDS entries and the rest in
    scratch (32 requested); a quad handing its rays to 16-lane rows (one
    node slot per lane) or to 8-lane rows (two slots per lane) after 1 or
    8 bounces; a 16-lane row per ray from the start."""
    nc, sk, lanes, width, handoff, hlanes = shape
    if scene_name == "city":
        mb, bounds, boxes = scenes.city_block(seed=5, n_buildings=8, target_tris=6000)
        src = np.vstack([scenes.indoor_points(51, 2, boxes),
                         scenes.outdoor_points(52, 1, bounds, boxes)])
        lis = scenes.outdoor_points(54, 2, bounds, boxes)
    else:
        mb, s1, l1, _ = scenes.shoebox(seed=1)
        src, lis = np.array([s1]), np.array([l1])
    _say(f"case wide walk {scene_name} node_cache={nc} stack_lds={sk} lanes={lanes} width={width} "
         f"handoff={handoff} handoff_lanes={hlanes}")
    kw = dict(n_rays=1000, nbands=3, seed=41, bin_dt=0.001, max_time=2.0)
    scene = H.Scene.from_builder(ctx, mb)
    tree = scene.export()[3]
    assert tree["has4"] and tree["has16"]
    try:
        ctx.set_launch_config(nc, sk, lanes, width)
        ctx.set_handoff(handoff)
        ctx.set_handoff_lanes(hlanes)
        gpu, ids, ts = H.trace_world_hits(scene, H.trace_params(**kw), src, lis)
    finally:
        ctx.set_launch_config()
        ctx.set_handoff()
        ctx.set_handoff_lanes()
    try:
        ctx.set_launch_config(512, 16, 0, 2)
        ref, ids2, ts2 = H.trace_world_hits(scene, H.trace_params(**kw), src, lis)
    finally:
        ctx.set_launch_config()
    om = oracle_mesh_from_builder(mb)
    outs = oracle_trace_sources(O.mesh_geom(om), src, lis, **kw)
    orec = np.vstack([o["recs"] for o in outs])
    assert np.array_equal(ref.records, orec)
    flagged = int(np.sum((orec[:, 3] & 3) != 0))
    diff = np.any(gpu.records != orec, axis=1)
    assert not diff.any(), f"{int(diff.sum())} rays differ ({flagged} flagged in the oracle)"
    assert np.array_equal(ids, ids2) and np.array_equal(ts, ts2)
    _say(f"  every ray identical hit by hit ({flagged} carry a near-edge or tie flag)")
    compare(gpu, outs, 1000, label=f"width {width} {scene_name}")


def test_overlays_six_bands(ctx, oracle):
    mb, src, lis, _ = scenes.shoebox(seed=3, nbands=6)
    panel_mat = H.MaterialDef([0.1, 0.2, 0.3, 0.4, 0.5, 0.6], 0.2,
                              [0.3, 0.2, 0.1, 0.05, 0.02, 0.01])
    panels = [((5.0, 2.0, 0.0), (5.1, 6.0, 2.5), panel_mat),
              ((2.0, 3.0, 1.0), (3.0, 4.0, 1.8), panel_mat)]
    _say("case overlays + 6 bands")
    kw = dict(n_rays=2048, nbands=6, seed=9, dir_band=2, bin_dt=0.001, max_time=2.0)
    scene = H.Scene.from_builder(ctx, mb, nbands=6)
    gpu = H.trace_world(scene, H.trace_params(**kw), [src], [lis], overlays=panels, records=True)
    om = oracle_mesh_from_builder(mb, nbands=6)
    g = O.mesh_geom(om, [list(p[0]) + list(p[1]) for p in panels],
                    [mat_tuple(p[2]) for p in panels])
    outs = oracle_trace_sources(g, [src], [lis], **kw)
    compare(gpu, outs, 2048, dir_band=2, label="overlays 6 bands")


def test_refusals(ctx):
    mb, src, lis, _ = scenes.shoebox(seed=1)
    scene = H.Scene.from_builder(ctx, mb)
    too_many = [((0, 0, 0), (1, 1, 1), H.CONCRETE)] * 65
    with pytest.raises(H.OmgError) as e:
        H.trace_world(scene, H.trace_params(64), [src], [lis], overlays=too_many)
    assert e.value.code == H._lib.ERR_CAP
    with pytest.raises(H.OmgError) as e:
        H.trace_world(scene, H.trace_params(64, nbands=4), [src], [lis])
    assert e.value.code == H._lib.ERR_INVALID
