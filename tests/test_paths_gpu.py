"""GPU parity of the knife-edge path search (AutoPaths::find,
crates/omg-core/src/paths.rs:117-332) against the oracle: every path's key,
points, length and per-band gains bit-identical, on the reference's own
test scenes (paths.rs:455-537) and on the city block with the world's
budget (crates/omg-scene/src/world.rs:154, :638-642)."""
import sys

import numpy as np
import pytest
import torch   # before the HIP library is loaded: torch brings its own HIP runtime

import omg_hip as H
from omg_hip import scenes
from omg_hip.paths import PathBudget
from helpers import oracle_mesh_from_builder

pytestmark = pytest.mark.gpu

BRICK = scenes.MaterialDef([0.02, 0.03, 0.04], 0.12, [0.045, 0.008, 0.0012])   # material.rs:45-49
WORLD = PathBudget(48, 32, 4, False)   # world.rs:638-642


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def assert_same_paths(gpu, ref, where=""):
    assert len(gpu) == len(ref), f"{where}: {len(gpu)} paths vs {len(ref)}"
    for k, (g, r) in enumerate(zip(gpu, ref)):
        assert g["key"] == r["key"], f"{where} path {k}: key {g['key']} vs {r['key']}"
        assert np.array_equal(g["points"], r["points"]), f"{where} path {k}: points"
        assert g["length"] == r["length"], f"{where} path {k}: length"
        assert np.array_equal(g["gains"], r["gains"]), f"{where} path {k}: gains {g['gains']} vs {r['gains']}"


def _budget_kw(b):
    return dict(edge_candidates=b.edge_candidates, pair_edges=b.pair_edges,
                max_paths=b.max_paths, taps=b.taps)


def behind_building():
    """paths.rs:486-525: a solid block on a ground slab."""
    mb = scenes.MeshBuilder()
    m = mb.material(BRICK)
    mb.abox((10.0, -4.0, 0.0), (20.0, 4.0, 4.0), m)
    mb.abox((-40.0, -40.0, -0.5), (60.0, 40.0, 0.0), m)
    return mb


@pytest.fixture(scope="module")
def city(ctx):
    mb, bounds, boxes = scenes.city_block(seed=2)
    scene = H.Scene.from_builder(ctx, mb)
    om = oracle_mesh_from_builder(mb)
    # street-level and indoor endpoints: most pairs are blocked
    pts = np.vstack([scenes.outdoor_points(71, 48, bounds, boxes),
                     scenes.indoor_points(72, 48, boxes)])
    rng = np.random.default_rng(5)
    pairs = rng.integers(0, pts.shape[0], size=(160, 2))
    pairs = pairs[pairs[:, 0] != pairs[:, 1]]
    yield scene, om, pts[pairs[:, 0]], pts[pairs[:, 1]]
    scene.close()


def test_behind_building_matches_oracle(ctx):
    mb = behind_building()
    scene = H.Scene.from_builder(ctx, mb)
    om = oracle_mesh_from_builder(mb)
    ap = H.AutoPaths(scene, 64)
    edges = om.path_edges(64)
    assert np.array_equal(ap.edges(), edges)
    src = np.float32([5.0, 0.0, 1.5])
    rng = np.random.default_rng(11)
    lis = np.vstack([[25.0, 0.0, 1.6], rng.uniform([21, -10, 0.5], [40, 10, 3.5], (40, 3))])
    srcs = np.repeat(src[None], lis.shape[0], axis=0)
    for b in (PathBudget(), PathBudget(24, 8, 6), WORLD, PathBudget(24, 8, 4, True)):
        _say(f"behind_building budget {b}")
        got = ap.find_many(srcs, lis, b)
        for q in range(lis.shape[0]):
            ref = om.find_paths(edges, srcs[q], lis[q], **_budget_kw(b))
            assert_same_paths(got[q], ref, f"query {q}")
    # the reference's assertions hold on the device result (paths.rs:506-524)
    paths = ap.find(src, [25.0, 0.0, 1.6], PathBudget(24, 8, 6))
    assert paths[0]["key"] == 0 and paths[0]["gains"][1] < 0.1 and len(paths) > 1
    assert any(p["points"][1][2] > 3.5 for p in paths[1:])
    assert any(abs(p["points"][1][1]) > 3.5 and p["points"][1][2] < 3.5 for p in paths[1:])
    ap.close()
    scene.close()


def test_clear_line_returns_only_direct(ctx):
    """paths.rs:527-536."""
    mb = scenes.MeshBuilder()
    m = mb.material(BRICK)
    mb.abox((0.0, 10.0, 0.0), (4.0, 12.0, 3.0), m)
    scene = H.Scene.from_builder(ctx, mb)
    ap = H.AutoPaths(scene, 64)
    paths = ap.find((0.0, 0.0, 1.5), (10.0, 0.0, 1.5))
    assert len(paths) == 1 and paths[0]["key"] == 0
    assert np.array_equal(paths[0]["gains"], np.ones(3, np.float32))
    ap.close()
    scene.close()


def test_wall_transmission_matches_oracle(ctx):
    """paths.rs:455-484 scene; batched segment transmission bit-identical."""
    mb = scenes.MeshBuilder()
    m = mb.material(BRICK)
    mb.abox((0.0, 0.0, 0.0), (8.0, 6.0, 3.0), m)
    mb.abox((0.24, 0.24, 0.24), (7.76, 5.76, 2.76), m)
    scene = H.Scene.from_builder(ctx, mb)
    om = oracle_mesh_from_builder(mb)
    ap = H.AutoPaths(scene, 64)
    rng = np.random.default_rng(3)
    a = rng.uniform(-3, 11, (300, 3)).astype(np.float32)
    b = rng.uniform(-3, 11, (300, 3)).astype(np.float32)
    a[0], b[0] = (-2.0, 3.0, 1.5), (2.0, 3.0, 1.5)
    got = ap.transmission(a, b)
    ref = np.array([om.path_transmission(x, y) for x, y in zip(a, b)])
    assert np.array_equal(got, ref)
    want = np.minimum(np.array(BRICK.transmission) * min(0.15 / 0.24, 2.0), 0.9)
    assert np.all(np.abs(got[0] - want) < 0.02)
    ap.close()
    scene.close()


@pytest.mark.parametrize("budget", [PathBudget(), WORLD, PathBudget(48, 32, 4, True)],
                         ids=["default", "world", "world-taps"])
def test_city_block_matches_oracle(ctx, city, budget):
    scene, om, src, lis = city
    ap = H.AutoPaths(scene, 512)
    edges = om.path_edges(512)
    assert np.array_equal(ap.edges(), edges)
    _say(f"city block {src.shape[0]} queries, budget {budget}")
    got = ap.find_many(src, lis, budget)
    blocked = 0
    for q in range(src.shape[0]):
        ref = om.find_paths(edges, src[q], lis[q], **_budget_kw(budget))
        assert_same_paths(got[q], ref, f"query {q}")
        blocked += len(ref) > 1
    assert blocked > src.shape[0] // 4
    ap.close()


def test_prefiltered_and_degenerate_graphs(ctx, city):
    """More than 1024 edges (the corridor prefilter, paths.rs:169-171), an
    empty edge graph, no pair search, no returned bent paths, and the caps."""
    scene, om, src, lis = city
    src, lis = src[:48], lis[:48]
    for budget_edges, b in [(3000, WORLD), (0, PathBudget()), (512, PathBudget(24, 0, 4)),
                            (512, PathBudget(24, 8, 0)), (512, PathBudget(0, 8, 4))]:
        _say(f"edge budget {budget_edges}, {b}")
        ap = H.AutoPaths(scene, budget_edges)
        edges = om.path_edges(budget_edges)
        assert np.array_equal(ap.edges(), edges)
        got = ap.find_many(src, lis, b)
        for q in range(src.shape[0]):
            assert_same_paths(got[q], om.find_paths(edges, src[q], lis[q], **_budget_kw(b)),
                              f"edges {budget_edges} query {q}")
        ap.close()
    ap = H.AutoPaths(scene, 512)
    with pytest.raises(H.OmgError) as e:
        ap.find(src[0], lis[0], PathBudget(257, 8, 4))
    assert e.value.code == H._lib.ERR_CAP
    with pytest.raises(H.OmgError) as e:
        ap.find(src[0], lis[0], PathBudget(24, 65, 4))
    assert e.value.code == H._lib.ERR_CAP
    assert ap.find_many(np.zeros((0, 3)), np.zeros((0, 3))) == []
    ap.close()
    with pytest.raises(H.OmgError) as e:
        H.AutoPaths(scene, 100000)
    assert e.value.code == H._lib.ERR_CAP


def test_counting_leaves_results_unchanged(ctx, city):
    """The counted pass (one lane per segment, the reference's tests) gives
    the timed passes' paths; slots past a query's path count are zero."""
    scene, om, src, lis = city
    ap = H.AutoPaths(scene, 512)
    plain, plain_n = ap.find_many(src, lis, WORLD, raw=True)
    ap.set_counting(True)
    counted, counted_n = ap.find_many(src, lis, WORLD, raw=True)
    c = ap.counters()
    assert np.array_equal(plain_n, counted_n)
    for q in range(src.shape[0]):
        assert np.array_equal(plain[q, :plain_n[q]], counted[q, :counted_n[q]]), f"query {q}"
        for raw, n in ((plain, plain_n), (counted, counted_n)):
            rest = raw[q, n[q]:]
            assert not np.any(rest["n_points"]) and not np.any(rest["points"]), f"query {q}"
    assert c["aabb"] > 0 and c["tri"] > 0
    ap.close()


def test_many_queries_split_legs_match_one_lane_walks(ctx):
    """4096 city-block queries at the world budget give the leg pass enough
    legs to split each over 2 lanes (fewer calls split over 8): the result
    equals the counted pass's one-lane walks for every query and the
    oracle's for a sample."""
    mb, bounds, boxes = scenes.city_block(seed=2)
    scene = H.Scene.from_builder(ctx, mb)
    om = oracle_mesh_from_builder(mb)
    src, lis = scenes.path_queries(7, 4096, bounds, boxes)
    ap = H.AutoPaths(scene, 512)
    _say("4096 queries, split legs")
    fast, fast_n = ap.find_many(src, lis, WORLD, raw=True)
    ap.set_counting(True)
    ref, ref_n = ap.find_many(src, lis, WORLD, raw=True)
    assert np.array_equal(fast_n, ref_n)
    for q in range(src.shape[0]):
        assert np.array_equal(fast[q, :fast_n[q]], ref[q, :ref_n[q]]), f"query {q}"
    edges = om.path_edges(512)
    got = ap.find_many(src[::16], lis[::16], WORLD)
    for k, q in enumerate(range(0, src.shape[0], 16)):
        assert_same_paths(got[k], om.find_paths(edges, src[q], lis[q], **_budget_kw(WORLD)),
                          f"query {q}")
    ap.close()
    scene.close()


def _wall_stack(n_walls):
    """n 2 x 2 m brick walls 0.02 m thick, 0.03 m apart, across x = 10..11
    m: a segment along x at y = 0.3, z = 1 crosses each twice, inside one
    triangle of each face (at y = 0 it would run through the faces'
    diagonals and hit both triangles of every face)."""
    mb = scenes.MeshBuilder()
    m = mb.material(BRICK)
    for i in range(n_walls):
        x = 10.0 + 0.03 * i
        mb.abox((x, -1.0, 0.0), (x + 0.02, 1.0, 2.0), m)
    return mb


def test_crossing_list_limits(ctx):
    """60 crossings inside one lane's interval of a split walk send the
    segment to the one-lane walk (64 kept), which matches the oracle; 80
    crossings are refused with OMG_ERR_OVERFLOW (no silent truncation)."""
    a, b = np.float32([0.0, 0.3, 1.0]), np.float32([20.0, 0.3, 1.0])
    mb = _wall_stack(30)
    scene = H.Scene.from_builder(ctx, mb)
    om = oracle_mesh_from_builder(mb)
    ap = H.AutoPaths(scene, 0)
    ts, _, _ = om.segment_hits(a, b)
    assert len(ts) == 60
    ref = om.path_transmission(a, b)
    assert np.array_equal(ap.transmission(a, b), ref)
    paths = ap.find(a, b)
    assert np.array_equal(paths[0]["gains"], ref)
    assert_same_paths(paths, om.find_paths(om.path_edges(0), a, b), "60 crossings")
    ap.close()
    scene.close()
    mb = _wall_stack(40)
    scene = H.Scene.from_builder(ctx, mb)
    ap = H.AutoPaths(scene, 0)
    for call in (lambda: ap.transmission(a, b), lambda: ap.find(a, b)):
        with pytest.raises(H.OmgError) as e:
            call()
        assert e.value.code == H._lib.ERR_OVERFLOW
    ap.close()
    scene.close()


def test_find_async_matches_find(ctx, city):
    """omg_paths_find_async on device buffers gives omg_paths_find's paths;
    omg_paths_status reports no overflow."""
    from omg_hip.paths import PATH_DTYPE
    scene, om, src, lis = city
    ap = H.AutoPaths(scene, 512)
    ref, ref_n = ap.find_many(src, lis, WORLD, raw=True)
    dev = torch.device("cuda", 0)
    n, stride = src.shape[0], 1 + WORLD.max_paths
    d_src = torch.from_numpy(np.ascontiguousarray(src)).to(dev)
    d_lis = torch.from_numpy(np.ascontiguousarray(lis)).to(dev)
    d_out = torch.zeros(n * stride * PATH_DTYPE.itemsize, dtype=torch.uint8, device=dev)
    d_n = torch.zeros(n, dtype=torch.int32, device=dev)
    d_work = torch.zeros(n, dtype=torch.int32, device=dev)
    stream = torch.cuda.Stream(device=dev)
    for _ in range(2):
        ap.find_async(d_src.data_ptr(), d_lis.data_ptr(), n, d_out.data_ptr(), d_n.data_ptr(),
                      d_work.data_ptr(), WORLD, stream.cuda_stream)
    torch.cuda.synchronize()
    ap.status()
    got = d_out.cpu().numpy().view(PATH_DTYPE).reshape(n, stride)
    got_n = d_n.cpu().numpy().astype(np.uint32)
    assert np.array_equal(got_n, ref_n)
    for q in range(n):
        assert np.array_equal(got[q], ref[q]), f"query {q}"
    ap.close()


def test_async_overflow_survives_a_blocking_call(ctx):
    """An async search that overflows, then a blocking search that does not:
    the blocking call returns OK and waits for the async one, and the next
    omg_paths_status still reports the async overflow, once."""
    from omg_hip.paths import PATH_DTYPE
    a, b = np.float32([0.0, 0.3, 1.0]), np.float32([20.0, 0.3, 1.0])
    clear_a, clear_b = np.float32([0.0, 5.0, 1.0]), np.float32([20.0, 5.0, 1.0])
    mb = _wall_stack(40)
    scene = H.Scene.from_builder(ctx, mb)
    ap = H.AutoPaths(scene, 0)
    budget = H.PathBudget()
    dev = torch.device("cuda", 0)
    stride = 1 + budget.max_paths
    d_src = torch.from_numpy(a.reshape(1, 3).copy()).to(dev)
    d_lis = torch.from_numpy(b.reshape(1, 3).copy()).to(dev)
    d_out = torch.zeros(stride * PATH_DTYPE.itemsize, dtype=torch.uint8, device=dev)
    d_n = torch.zeros(1, dtype=torch.int32, device=dev)
    d_work = torch.zeros(1, dtype=torch.int32, device=dev)
    stream = torch.cuda.Stream(device=dev)
    ap.find_async(d_src.data_ptr(), d_lis.data_ptr(), 1, d_out.data_ptr(), d_n.data_ptr(),
                  d_work.data_ptr(), budget, stream.cuda_stream)
    paths = ap.find(clear_a, clear_b, budget)
    assert len(paths) >= 1
    with pytest.raises(H.OmgError) as e:
        ap.status()
    assert e.value.code == H._lib.ERR_OVERFLOW
    ap.status()
    ap.close()
    scene.close()
