"""GPU parity of the ambient dome and the image-source kernels against the oracle."""
import sys

import numpy as np
import pytest

import oracle as O
import omg_hip as H
from omg_hip import scenes
from helpers import CONCRETE, WOOD, mat_tuple, oracle_mesh_from_builder

pytestmark = pytest.mark.gpu

GLASS = [0.50, 0.32, 0.20]          # walkthrough.rs:310
DOOR_PANEL = [0.28, 0.12, 0.05]     # walkthrough.rs:323


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def house():
    """A 10 x 8 x 3 m brick room with a door hole (x = 10 wall) and a glazed
    window (y = 0 wall) on a ground plane, every wall its own surface."""
    mb = scenes.MeshBuilder()
    brick = mb.material(H.BRICK)
    ground = mb.material(H.GRASS)
    mb.begin_surface()
    mb.quad((-60, -60, 0), (70, -60, 0), (70, 70, 0), (-60, 70, 0), ground)
    door = (3.5, 4.5, 0.0, 2.1)      # (u0, u1, z0, z1) on the x = 10 wall, u = y
    window = (4.0, 6.0, 1.0, 2.2)    # on the y = 0 wall, u = x
    walls = [(0, 0.0, 0.0, 8.0, []), (0, 10.0, 0.0, 8.0, [door]), (1, 0.0, 0.0, 10.0, [window]),
             (1, 8.0, 0.0, 10.0, [])]
    for axis, plane, lo, hi, holes in walls:
        mb.begin_surface()
        scenes._wall_cells(mb, brick, axis, plane, lo, hi, 0.0, 3.0, 0.5, holes)
    mb.begin_surface()
    mb.quad((0, 0, 3), (10, 0, 3), (10, 8, 3), (0, 8, 3), brick)
    glass = dict(axis=1, plane=0.0, lat=(4.0, 6.0), z=(1.0, 2.2), glass=True, trans=GLASS)
    leaf = dict(axis=0, plane=10.0, lat=(3.5, 4.5), z=(0.0, 2.1), glass=False, trans=DOOR_PANEL)
    return mb, glass, leaf


def compare_dome(scene, om, eyes, params, panels, label):
    g = H.dome_sample(scene, eyes, params, panels, rays=True)
    for ei, eye in enumerate(eyes):
        o = O.dome_sample(om, eye, n_fan=params.n_fan, n_rays=params.n_rays,
                          max_events=params.max_events, horizon=params.horizon,
                          through_len=params.through_len, panels=panels, want_rays=True)
        for key in ("escaped", "events", "through", "bin"):
            bad = np.nonzero(g[key][ei] != o[key])[0]
            assert bad.size == 0, f"{label} eye {ei}: {key} differs at rays {bad[:5]}"
        assert np.array_equal(g["tp"][ei], o["tp"]), f"{label} eye {ei}: throughput bits differ"
        assert np.array_equal(g["energy"][ei], o["e"]), f"{label} eye {ei}: bin energies differ"
        assert np.array_equal(g["dsum"][ei], o["dsum"]), f"{label} eye {ei}: bin directions differ"
    _say(f"{label}: {len(eyes)} eyes x {params.n_rays} rays bit-identical")
    return g


def test_dome_house_parity_and_behaviour(ctx, oracle):
    mb, glass, leaf = house()
    scene = H.Scene.from_builder(ctx, mb)
    om = oracle_mesh_from_builder(mb)
    params = H.dome_params(3, n_fan=384)
    eyes = [(5.0, 4.0, 1.6), (30.0, 30.0, 1.6), (8.5, 4.0, 1.6), (5.0, 1.0, 1.6)]
    _say("case dome house: open + closed leaf")
    g_open = compare_dome(scene, om, eyes, params, [glass], "dome open door")
    g_shut = compare_dome(scene, om, eyes, params, [glass, leaf], "dome closed leaf")
    tot = lambda g, i, b=1: float(g["energy"][i][:, b].sum())
    # open field hears most of the dome, indoors only a fraction (dome.rs:497-505)
    assert 0.4 < tot(g_open, 1) < 1.0
    assert 1e-4 < tot(g_open, 0) < 0.25 * tot(g_open, 1)
    # closing the leaf never raises the inflow (dome.rs:541-556)
    for i in (0, 2, 3):
        assert tot(g_shut, i) <= tot(g_open, i) + 1e-6
    # glass favours the lows (dome.rs:638-653), behind the window
    lo, hi = tot(g_shut, 3, 0), tot(g_shut, 3, 2)
    assert lo > 2.0 * hi


def test_dome_open_field_closed_form(ctx, oracle):
    """Only a triangle beyond the sky sphere: every fan ray escapes with
    throughput 1, so bin k holds exactly count_k / N (dome.rs:433-458)."""
    mb = scenes.MeshBuilder()
    m = mb.material(H.CONCRETE)
    mb.tri((5000, 0, 0), (5000, 1, 0), (5000, 0, 1), m)
    scene = H.Scene.from_builder(ctx, mb)
    params = H.dome_params(3, n_fan=512, horizon=500.0)
    g = H.dome_sample(scene, [(0.0, 0.0, 1.6)], params, rays=True)
    fan = O.dome_fan(512)
    assert np.all(g["escaped"][0] == 1) and np.all(g["tp"][0] == 1.0)
    assert np.array_equal(g["dir"][0], fan)
    bins = np.array([O.dome_bin_of(d) for d in fan])
    for k in range(9):
        assert g["energy"][0][k, 0] == np.float32(np.sum(bins == k) / 512.0)


def test_dome_city_block_500m_sky(ctx, oracle):
    mb, bounds, boxes = scenes.city_block(seed=2, n_buildings=6, target_tris=4000)
    scene = H.Scene.from_builder(ctx, mb)
    om = oracle_mesh_from_builder(mb)
    eyes = scenes.outdoor_points(5, 3, bounds, boxes)
    compare_dome(scene, om, [tuple(e) for e in eyes], H.dome_params(3, n_fan=512, horizon=500.0),
                 [], "dome city block")


def live_room_mesh():
    """The golden live-room built as a mesh, one surface per wall (pt_mesh_gate.rs:27-50)."""
    walls = [CONCRETE, CONCRETE, CONCRETE, CONCRETE, WOOD, CONCRETE]
    quads = [(((0, 0, 0), (0, 4, 0), (0, 4, 3), (0, 0, 3)), 0),
             (((6, 0, 0), (6, 0, 3), (6, 4, 3), (6, 4, 0)), 1),
             (((0, 0, 0), (0, 0, 3), (6, 0, 3), (6, 0, 0)), 2),
             (((0, 4, 0), (6, 4, 0), (6, 4, 3), (0, 4, 3)), 3),
             (((0, 0, 0), (6, 0, 0), (6, 4, 0), (0, 4, 0)), 4),
             (((0, 0, 3), (0, 4, 3), (6, 4, 3), (6, 0, 3)), 5)]
    mb = scenes.MeshBuilder()
    for q, wi in quads:
        m = mb.material(walls[wi])
        mb.begin_surface()
        mb.quad(*q, m)
    return mb, walls


def test_discovery_and_solve_parity(ctx, oracle):
    mb, walls = live_room_mesh()
    scene = H.Scene.from_builder(ctx, mb)
    om = oracle_mesh_from_builder(mb)
    src, lis = (1.5, 1.0, 1.5), (4.5, 3.0, 1.6)
    disc = H.GpuWorldDiscovery(scene, n_rays=4096)
    for rot in (0, 3):
        gc, go = disc.discover(lis, rot)
        oc, oo = O.mesh_chains(om, lis, 4096, rot)
        assert np.array_equal(gc, oc) and np.array_equal(go, oo), f"rot {rot}: chains differ"
    _say(f"discovery: {len(gc)} chains identical")
    tab, n_surf = H.surface_table(scene)
    otab, on = O.surface_table(om)
    assert n_surf == on
    solver = H.GpuWorldSolve(scene)
    chains = np.vstack([np.full((1, 3), 0xFFFF, np.uint16), gc])
    orders = np.concatenate([[0], go]).astype(np.uint8)
    srcs = [src, (2.0, 2.5, 1.2)]
    rec = solver.solve_batch(srcs, chains, orders, lis)
    n = len(orders)
    n_valid = 0
    for si, s in enumerate(srcs):
        for ci in range(n):
            r = O.mesh_record(om, otab, on, chains[ci][:orders[ci]], int(orders[ci]), s, lis)
            k = si * n + ci
            assert rec["valid"][k] == (r is not None), f"source {si} chain {ci}: validity differs"
            if r is None:
                continue
            n_valid += 1
            assert rec["delay"][k] == r["delay"] and np.array_equal(rec["dir"][k], r["dir"])
            np.testing.assert_allclose(rec["gains"][k], r["gains"], rtol=2e-6)
    _say(f"solve: {n_valid} records match the oracle")
    # box equivalence gate (pt_mesh_gate.rs:85-125) against the shoebox lattice
    taps = O.image_source_taps((6, 4, 3), [mat_tuple(w) for w in walls], src, lis, 3)
    db = lambda a: 20 * np.log10(np.maximum(a, 1e-9))
    valid = np.nonzero(rec["valid"][:n])[0]
    for k in valid:
        ok = np.any((np.abs(taps[:, 1] - rec["delay"][k]) * 1000 <= 0.5) &
                    np.all(np.abs(db(taps[:, 5:8]) - db(rec["gains"][k])) <= 1.0, axis=1))
        assert ok, f"mesh path {k} (order {orders[k]}) matches no image source"
    direct = valid[orders[valid] == 0]
    assert direct.size == 1
    true_delay = np.float32(np.linalg.norm(np.subtract(src, lis)) / 343.0)
    assert abs(rec["delay"][direct[0]] - true_delay) < 1e-5


def _solve_vs_oracle(solver, om, otab, on, srcs, chains, orders, lis, extras=()):
    rec = solver.solve_batch(srcs, chains, orders, lis, extras=extras)
    n = len(orders)
    n_valid = 0
    for si, s in enumerate(srcs):
        for ci in range(n):
            r = O.mesh_record(om, otab, on, chains[ci][:orders[ci]], int(orders[ci]), s, lis,
                              extras=extras)
            k = si * n + ci
            assert rec["valid"][k] == (r is not None), f"source {si} chain {ci}: validity differs"
            if r is None:
                continue
            n_valid += 1
            assert rec["delay"][k] == r["delay"] and np.array_equal(rec["dir"][k], r["dir"])
            np.testing.assert_allclose(rec["gains"][k], r["gains"], rtol=2e-6)
    return rec, n_valid


def test_overlay_faces_boxes_and_extras_match_the_oracle(ctx, oracle):
    """Furniture through the C ABI: the table-top room of pt_mesh_gate.rs:
    180-240 - discovery over the box (omg_discover with boxes) gives the
    oracle's chains and finds the top face; the solve of chains that reflect
    off the box faces (omg_surface_append_box rows, rect-validated) and of
    legs crossing blockers (extras, Aabb::clip) gives the oracle's records,
    with the exact table-top delay and the off-rect refusal."""
    from test_ism_overlays_cpu import OPAQUE, golden_box, oracle_table_with_box
    mb = golden_box((6.0, 4.0, 3.0))
    scene = H.Scene.from_builder(ctx, mb)
    om = oracle_mesh_from_builder(mb)
    mn, mx = (2.0, 1.5, 0.0), (3.0, 2.5, 0.8)
    tab, n = H.surface_table(scene)
    otab, on = oracle_table_with_box(om, mn, mx, H.WOOD_PANEL)
    assert n == on
    assert H.append_box(tab, n, mn, mx, H.WOOD_PANEL) == n + 6
    top = n + 5
    src, lis = (1.0, 2.0, 1.5), (4.0, 2.0, 1.6)
    disc = H.GpuWorldDiscovery(scene, n_rays=4096)
    found = set()
    for rot in range(4):
        gc, go = disc.discover(lis, rot, boxes=[(mn, mx)], base=n)
        oc, oo = O.mesh_chains(om, lis, 4096, rot, boxes=[(mn, mx)], base=n)
        assert np.array_equal(gc, oc) and np.array_equal(go, oo), f"rot {rot}: chains differ"
        found |= {(tuple(int(x) for x in c[:o]), int(o)) for c, o in zip(gc, go)}
    assert any(top in c for c, _ in found), "discovery must bounce off the table top"
    _say(f"discovery over the box: {len(found)} chains identical, top face found")
    solver = H.GpuWorldSolve(scene, tab, n + 6)
    lst = sorted(found)
    chains = np.array([list(c) + [0xFFFF] * (3 - o) for c, o in [((), 0)] + lst], np.uint16)
    orders = np.array([0] + [o for _, o in lst], np.uint8)
    srcs = [src, (5.0, 3.0, 1.2)]
    blockers = [((2.6, 0.5, 0.0), (2.8, 3.5, 1.2), OPAQUE), ((4.5, 0.2, 1.0), (5.5, 1.0, 2.8),
                                                              (0.5, 0.25, 0.125))]
    for extras in ((), blockers):
        rec, nv = _solve_vs_oracle(solver, om, otab, n + 6, srcs, chains, orders, lis, extras)
        _say(f"solve with {len(extras)} blockers: {nv} records match the oracle")
    # the table-top chain alone: exact mirror delay, refused off the rect
    one = np.array([[top, 0xFFFF, 0xFFFF]], np.uint16)
    rec = solver.solve_batch([src], one, np.array([1], np.uint8), lis)
    img = np.array([1.0, 2.0, 2.0 * 0.8 - 1.5])
    assert rec["valid"][0]
    assert abs(float(rec["delay"][0]) - np.linalg.norm(img - np.array(lis)) / 343.0) < 1e-5
    off = solver.solve_batch([src], one, np.array([1], np.uint8), (1.2, 0.3, 1.6))
    assert not off["valid"][0]
    scene.close()


def _probe_restated(passes, nb):
    """DomeProbe::sample's EMA and bin directions (dome.rs:460-477) restated
    over the oracle's per-pass bin sums: alpha 1 on the first call, 0.35
    after; dir = normalize(dsum) (times the reciprocal length), or, for a
    silent bin, its sector centre (zenith (0, 0, 1))."""
    f = np.float32
    ema = np.zeros((9, nb), np.float32)
    outs = []
    for k, (e, dsum) in enumerate(passes):
        alpha = f(1.0) if k == 0 else f(0.35)
        ema = (ema + alpha * (e - ema)).astype(np.float32)
        out = np.zeros((9, 3 + nb), np.float32)
        for b in range(9):
            x, y, z = (f(v) for v in dsum[b])
            ln = np.sqrt(f(f(f(x * x) + f(y * y)) + f(z * z)))
            if ln > f(1e-6):
                inv = f(1.0) / ln
                out[b, :3] = [x * inv, y * inv, z * inv]
            elif b == 8:
                out[b, :3] = [0.0, 0.0, 1.0]
            else:
                az = f(f(f(b) + f(0.5)) * f(2.0 * np.pi) / f(8.0)) - f(np.pi)
                out[b, :3] = [np.cos(az), np.sin(az), 0.0]
            out[b, 3:] = ema[b]
        outs.append(out)
    return outs


@pytest.mark.parametrize("n_rays", [384, 256, 160])
def test_dome_probe_drop_in(ctx, oracle, n_rays):
    """omg_dome_probe_sample, the DomeProbe::sample drop-in (dome.rs:425-478):
    three calls from three eyes (the EMA primes on the first, then blends at
    0.35), at the full fan and the 256- and 160-ray tiers (even stride
    j * 384 / n, dome.rs:431-434), in a sealed room whose walls pass
    nothing and a house with a door and a window: every output row equals
    the restatement over the oracle's bin sums; in the sealed room every
    bin is silent and carries its placeholder direction."""
    mb, glass, leaf = house()
    scene = H.Scene.from_builder(ctx, mb)
    om = oracle_mesh_from_builder(mb)
    params = H.dome_params(3, n_fan=384, n_rays=n_rays)
    eyes = [(5.0, 4.0, 1.6), (8.5, 4.0, 1.6), (12.0, 4.0, 1.6)]
    probe = H.DomeProbe(scene, params)
    got = [probe.sample(e, [glass, leaf] if k == 1 else [glass]) for k, e in enumerate(eyes)]
    probe.close()
    passes = []
    for k, e in enumerate(eyes):
        o = O.dome_sample(om, e, n_fan=384, n_rays=n_rays, panels=[glass, leaf] if k == 1 else [glass])
        passes.append((o["e"], o["dsum"]))
    want = _probe_restated(passes, 3)
    for k in range(3):
        assert np.array_equal(got[k][:, 3:], want[k][:, 3:]), f"call {k}: EMA energies differ"
        np.testing.assert_allclose(got[k][:, :3], want[k][:, :3], rtol=0, atol=2e-7,
                                   err_msg=f"call {k}: bin directions differ")
    scene.close()
    # sealed room, walls that absorb half and pass nothing: every bounce path
    # dies inside and every straight line pays zero, so all nine bins are silent
    sb = scenes.MeshBuilder()
    m = sb.material(H.MaterialDef([0.5, 0.5, 0.5], 0.0, [0.0, 0.0, 0.0]))
    sb.abox((0.0, 0.0, 0.0), (6.0, 4.0, 3.0), m)
    sealed = H.Scene.from_builder(ctx, sb)
    probe = H.DomeProbe(sealed, params)
    out = [probe.sample((3.0, 2.0, 1.5)) for _ in range(2)]
    probe.close()
    sealed.close()
    want = _probe_restated([(np.zeros((9, 3), np.float32), np.zeros((9, 3), np.float32))] * 2, 3)
    for k in range(2):
        assert np.all(out[k][:, 3:] == 0.0)
        np.testing.assert_allclose(out[k][:, :3], want[k][:, :3], rtol=0, atol=2e-7)
    _say(f"dome probe, {n_rays} rays: three calls and a sealed room match the restatement")
