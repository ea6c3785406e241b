"""The five benchmark workloads run at their own sizes, on the GPU, against
the oracle and against the library's own single-job entry points.

- Shoebox room (12 x 8 x 4 m, 12 triangles): order-3 image sources on
  the BVH - the 4096-ray fan discovery over four rotations equals the
  oracle's chains, every (source, chain) record equals the oracle's, and
  every valid path is one of the room's analytic image sources
  (crates/omg-scene/tests/pt_mesh_gate.rs:85-125 tolerances).
- City block (~50k triangles, 16 sources x 4096 rays x 8 listeners, 3
  bands, 1 ms x 2 s): every ray's record equals the oracle's, the 128
  echograms are within 1e-4 of the oracle's f64 re-accumulation, and the
  8 listeners' 512-ray dome fans (500 m sky) are bit-identical.
- Mixed-size batch (256 queries over 32 blocks of 1k-200k triangles): each
  query of the full batch equals its plain single-job trace on the binary
  walk (records bit-identical, echograms within 1e-4), and a sample of
  queries equals the oracle.
- Town (~2M triangles, 64 sources x 1M rays x 6 bands, 128 bounces,
  floor 1e-7): the routed persistent launch equals the single-job kernel
  ray for ray on the full job, and 64 x 4096 rays equal the oracle.
- Serving over the 8 resident blocks of the serving workload: a probe
  request alone and inside a 2000 QPS mix of that workload's requests
  equals its single-job trace and dome.
"""
import sys
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import oracle as O
import omg_hip as H
from omg_hip import scenes
from helpers import echogram_rel_errors, mat_tuple, oracle_mesh_from_builder

pytestmark = pytest.mark.gpu

ECHO_TOL = 1e-4
ORACLE_THREADS = 8


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _defaults(ctx):
    """The drop-in's own launch configuration (other tests change it)."""
    ctx.set_launch_config()
    ctx.set_handoff()
    ctx.set_handoff_lanes()
    ctx.set_route_rays()


def _oracle_traces(geom, sources, listeners, n_rays, **kw):
    """Oracle trace of each source (src_index = its position), on a pool of
    host threads; the oracle's C entry releases the interpreter lock."""
    def one(si):
        p = O.trace_params(n_rays, rng_mode=O.RNG_COUNTER, src_index=si, **kw)
        return O.trace(geom, sources[si], listeners, p, want_rays=True, want_f64=True)
    with ThreadPoolExecutor(ORACLE_THREADS) as ex:
        return list(ex.map(one, range(len(sources))))


def _vs_oracle(records, bins, dirs, outs, n_rays, label, dir_band=1, sources=None):
    """records/bins/dirs of the GPU against oracle traces (one per source)."""
    rec = records.reshape(-1, n_rays, 4)
    worst_b = worst_d = 0.0
    heard = 0
    for k, o in enumerate(outs):
        si = k if sources is None else sources[k]
        diff = np.nonzero(np.any(rec[si] != o["recs"], axis=1))[0]
        assert diff.size == 0, (f"{label} source {si}: {diff.size} rays differ, first {diff[:5]}: "
                                f"gpu {rec[si][diff[:3]].tolist()} oracle {o['recs'][diff[:3]].tolist()}")
        eb, ed = echogram_rel_errors(bins[si], dirs[si], o["bins64"], o["dirs64"], dir_band)
        worst_b, worst_d = max(worst_b, eb), max(worst_d, ed)
        heard += int(np.count_nonzero(o["bins64"]))
    _say(f"  {label}: {len(outs)} x {n_rays} rays bit-identical to the oracle; echogram rel err "
         f"vs f64 bins {worst_b:.2e} dirs {worst_d:.2e} over {heard} non-zero bins")
    assert heard > 0, f"{label}: no listener hears anything"
    assert worst_b <= ECHO_TOL and worst_d <= ECHO_TOL, (label, worst_b, worst_d)


def _same(a_rec, a_bins, a_dirs, b_rec, b_bins, b_dirs, label):
    diff = np.nonzero(np.any(a_rec != b_rec, axis=1))[0]
    assert diff.size == 0, f"{label}: {diff.size} rays differ, first {diff[:5]}"
    eb, ed = echogram_rel_errors(a_bins, a_dirs, b_bins.astype(np.float64),
                                 b_dirs.astype(np.float64))
    assert eb <= ECHO_TOL and ed <= ECHO_TOL, (label, eb, ed)
    return max(eb, ed)


# ---------------------------------------------------------------- shoebox


def test_shoebox_order3_image_sources(ctx, oracle):
    mb, src, lis, walls = scenes.shoebox(seed=1)
    _say("case shoebox 12 x 8 x 4 m: order-3 image sources (4096-ray fan, 4 rotations)")
    scene = H.Scene.from_builder(ctx, mb)
    om = oracle_mesh_from_builder(mb)
    disc = H.GpuWorldDiscovery(scene, n_rays=4096)
    seen, chains, orders = set(), [], []
    for rot in range(4):
        gc, go = disc.discover(lis, rot)
        oc, oo = O.mesh_chains(om, lis, 4096, rot)
        assert np.array_equal(gc, oc) and np.array_equal(go, oo), f"rot {rot}: chains differ"
        for c, o in zip(gc, go):
            key = tuple(int(x) for x in c[:o])
            if key not in seen:
                seen.add(key)
                chains.append(c)
                orders.append(o)
    assert set(orders) == {1, 2, 3}, sorted(set(orders))
    chains = np.vstack([np.full((1, 3), 0xFFFF, np.uint16), np.array(chains, np.uint16)])
    orders = np.concatenate([[0], orders]).astype(np.uint8)
    tab, n_surf = H.surface_table(scene)
    otab, on = O.surface_table(om)
    assert n_surf == on == 6
    # the room's source and a second one, listener as the workload places them
    srcs = np.vstack([src, np.float32([3.0, 6.5, 1.2])])
    rec = H.GpuWorldSolve(scene).solve_batch(srcs, chains, orders, lis)
    n = len(orders)
    size = (12.0, 8.0, 4.0)
    wt = [mat_tuple(w) for w in walls]
    db = lambda a: 20 * np.log10(np.maximum(a, 1e-9))
    n_valid = {0: 0, 1: 0, 2: 0, 3: 0}
    for si, s in enumerate(srcs):
        taps = O.image_source_taps(size, wt, s, lis, 3)
        for ci in range(n):
            r = O.mesh_record(om, otab, on, chains[ci][:orders[ci]], int(orders[ci]), s, lis)
            k = si * n + ci
            assert rec["valid"][k] == (r is not None), f"source {si} chain {ci}: validity differs"
            if r is None:
                continue
            n_valid[int(orders[ci])] += 1
            assert rec["delay"][k] == r["delay"] and np.array_equal(rec["dir"][k], r["dir"])
            np.testing.assert_allclose(rec["gains"][k], r["gains"], rtol=2e-6)
            ok = np.any((np.abs(taps[:, 1] - rec["delay"][k]) * 1000 <= 0.5) &
                        np.all(np.abs(db(taps[:, 5:8]) - db(rec["gains"][k])) <= 1.0, axis=1))
            assert ok, f"source {si} chain {chains[ci][:orders[ci]]}: no image source matches"
    _say(f"  {n - 1} chains, valid paths per order {n_valid}: oracle-identical, "
         "each an image source")
    assert n_valid[0] == 2 and n_valid[1] >= 2 * 6 - 2 and n_valid[3] > 0
    scene.close()


# ------------------------------------------------------------- city block


def test_city_block_full_workload(ctx, oracle):
    _defaults(ctx)
    mb, src, lis = scenes.c2_workload()
    kw = dict(nbands=3, seed=0xC2, bin_dt=0.001, max_time=2.0)
    _say(f"case city block {mb.n_tris} tris: {src.shape[0]} sources x 4096 rays x "
         f"{lis.shape[0]} listeners, plus {lis.shape[0]} dome fans")
    scene = H.Scene.from_builder(ctx, mb)
    info = scene.info()
    gpu = H.trace_world(scene, H.trace_params(4096, **kw), src, lis, records=True)
    om = oracle_mesh_from_builder(mb)
    outs = _oracle_traces(O.mesh_geom(om), src, lis, 4096, **kw)
    _vs_oracle(gpu.records, gpu.bins, gpu.dirs, outs, 4096, "city block")
    assert gpu.bins.shape[:2] == (16, 8)
    dp = H.dome_params(3, n_fan=512, horizon=500.0)
    g = H.dome_sample(scene, lis, dp, rays=True)
    for ei, eye in enumerate(lis):
        o = O.dome_sample(om, eye, n_fan=512, n_rays=512, max_events=dp.max_events,
                          horizon=500.0, through_len=dp.through_len, want_rays=True)
        for key in ("escaped", "events", "through", "bin"):
            bad = np.nonzero(g[key][ei] != o[key])[0]
            assert bad.size == 0, f"dome eye {ei}: {key} differs at rays {bad[:5]}"
        assert np.array_equal(g["tp"][ei], o["tp"]), f"dome eye {ei}: throughputs differ"
        assert np.array_equal(g["energy"][ei], o["e"]), f"dome eye {ei}: bin energies differ"
        assert np.array_equal(g["dsum"][ei], o["dsum"]), f"dome eye {ei}: bin directions differ"
    _say(f"  dome: {lis.shape[0]} eyes x 512 rays bit-identical ({info['n_tris']} tris)")
    # four times the rays: the hand-off runs on 8-lane rows (two node slots
    # per lane) from 262144 rays; every ray equals the 16-lane rows', and
    # the first 4096 rays of each source are the 4096-ray job's above
    big = H.trace_params(16384, **kw)
    rows8 = H.trace_world(scene, big, src, lis, records=True)
    try:
        ctx.set_handoff_lanes(16)
        rows16 = H.trace_world(scene, big, src, lis, records=True)
    finally:
        ctx.set_handoff_lanes()
    assert np.array_equal(rows8.records, rows16.records)
    d = float(np.max(np.abs(rows8.bins - rows16.bins))) / float(np.max(np.abs(rows16.bins)))
    assert d <= 1e-5, d
    head = rows8.records.reshape(src.shape[0], 16384, 4)[:, :4096].reshape(-1, 4)
    assert np.array_equal(head, gpu.records)
    _say(f"  {src.shape[0]} x 16384 rays: 8-lane rows == 16-lane rows ray for ray "
         f"(echogram rel diff {d:.1e}), first 4096 rays per source == the oracle-checked job")
    scene.close()


# ---------------------------------------------------------- mixed batch


@pytest.fixture(scope="module")
def mixed(ctx):
    specs, qs = scenes.c3_workload()
    pool = [H.Scene.from_builder(ctx, mb, nbands=nb) for mb, nb in specs]
    Q = [H.QuerySpec(pool[q["scene"]],
                     H.trace_params(q["n_rays"], nbands=q["nbands"], seed=q["seed"], bin_dt=0.001,
                                    max_time=q["max_time"]),
                     q["sources"], q["listeners"]) for q in qs]
    yield specs, qs, Q
    for s in pool:
        s.close()


def test_mixed_batch_full_workload_vs_binary_plain(ctx, mixed):
    _defaults(ctx)
    specs, qs, Q = mixed
    rays = sum(q.params.n_rays * len(q.sources) for q in Q)
    _say(f"case mixed batch: {len(Q)} queries, {rays} rays, autotuned shape")
    b = H.Batch(ctx, Q, records=True)
    shape = b.tune(stride=16)
    b.launch()
    b.sync()
    _say(f"  tuned shape {shape}")
    worst = 0.0
    for i, q in enumerate(Q):
        got = b.result(i)
        # counters make the single-job path walk the binary tree
        ref = H.trace_world(q.scene, q.params, q.sources, q.listeners, records=True,
                            counters=True)
        worst = max(worst, _same(got.records, got.bins, got.dirs, ref.records, ref.bins,
                                 ref.dirs, f"query {i}"))
    b.close()
    _say(f"  {len(Q)} queries: records bit-identical to the binary walk, worst echogram "
         f"rel diff {worst:.2e}")


def test_mixed_batch_sample_vs_oracle(ctx, mixed):
    _defaults(ctx)
    specs, qs, Q = mixed
    b = H.Batch(ctx, Q, records=True)
    b.launch()
    b.sync()
    # a 3-band and a 6-band query, on the smallest and the largest block
    by_tris = sorted(range(len(Q)), key=lambda i: specs[qs[i]["scene"]][0].n_tris)
    pick = []
    for nb in (3, 6):
        cand = [i for i in by_tris if qs[i]["nbands"] == nb]
        pick += [cand[0], cand[-1]]
    geoms = {}
    for i in pick:
        q, spec = qs[i], specs[qs[i]["scene"]]
        n_rays = q["n_rays"]
        srcs = list(range(min(4, len(q["sources"]))))
        _say(f"case mixed batch query {i}: {spec[0].n_tris} tris, {q['nbands']} bands, "
             f"{len(srcs)}/{len(q['sources'])} sources x {n_rays} rays vs the oracle")
        if q["scene"] not in geoms:
            geoms[q["scene"]] = O.mesh_geom(oracle_mesh_from_builder(spec[0], nbands=spec[1]))
        got = b.result(i)
        outs = _oracle_traces(geoms[q["scene"]], q["sources"][srcs], q["listeners"], n_rays,
                              nbands=q["nbands"], seed=q["seed"], bin_dt=0.001,
                              max_time=q["max_time"])
        _vs_oracle(got.records, got.bins, got.dirs, outs, n_rays, f"query {i}", sources=srcs)
    b.close()


# ----------------------------------------------------------------- town


@pytest.fixture(scope="module")
def town(ctx):
    mb, src, lis = scenes.c4_workload()
    scene = H.Scene.from_builder(ctx, mb, nbands=6)
    yield mb, src, lis, scene
    scene.close()


C4_KW = dict(nbands=6, max_bounces=128, floor_rel=1e-7, seed=0xC4, bin_dt=0.001, max_time=2.0)


def test_town_routed_equals_single_job(ctx, town):
    _defaults(ctx)
    mb, src, lis, scene = town
    n_rays = 1 << 20
    params = H.trace_params(n_rays, **C4_KW)
    _say(f"case town {scene.info()['n_tris']} tris: {src.shape[0]} x {n_rays} rays x 6 bands, "
         "routed persistent launch")
    routed = H.trace_world(scene, params, src, lis, records=True)
    _say("case town: single-job kernel (routing off)")
    ctx.set_route_rays(0)
    try:
        single = H.trace_world(scene, params, src, lis, records=True)
    finally:
        ctx.set_route_rays()
    d = _same(routed.records, routed.bins, routed.dirs, single.records, single.bins, single.dirs,
              "town routed vs single job")
    _say(f"  {src.shape[0] * n_rays} rays bit-identical, echogram rel diff {d:.2e}")
    # the first 4096 rays of each source are the 4096-ray job's: the RNG is
    # per (seed, source, ray) and 1/n scales energy and floor by a power of 2
    small = H.trace_world(scene, H.trace_params(4096, **C4_KW), src, lis, records=True)
    head = routed.records.reshape(src.shape[0], n_rays, 4)[:, :4096].reshape(-1, 4)
    assert np.array_equal(head, small.records)


def test_town_sample_vs_oracle(ctx, town):
    _defaults(ctx)
    mb, src, lis, scene = town
    _say(f"case town sample: {src.shape[0]} sources x 4096 rays x 6 bands vs the oracle")
    gpu = H.trace_world(scene, H.trace_params(4096, **C4_KW), src, lis, records=True)
    geom = O.mesh_geom(oracle_mesh_from_builder(mb, nbands=6))
    outs = _oracle_traces(geom, src, lis, 4096, **C4_KW)
    _vs_oracle(gpu.records, gpu.bins, gpu.dirs, outs, 4096, "town sample")


# -------------------------------------------------------------- serving


def test_serving_probe_alone_and_in_the_workload_mix(ctx):
    _defaults(ctx)
    specs, reqs = scenes.c5_workload(seconds=2.0, qps=2000.0)
    pool = [H.Scene.from_builder(ctx, mb, nbands=3) for mb, _, _ in specs]
    R = [dict(scene=pool[r["scene"]], params=H.trace_params(r["n_rays"], 3, seed=r["seed"]),
              sources=r["sources"], listeners=r["listener"], dome=True) for r in reqs]
    arr = np.array([r["t"] for r in reqs])
    big = int(np.argmax([mb.n_tris for mb, _, _ in specs]))
    n_mix = 600
    # the probe: the first request of the mix on the largest block with at least 4 sources
    pi = next(i for i in range(n_mix) if reqs[i]["scene"] == big and reqs[i]["n_src"] >= 4)
    probe = dict(R[pi], keep=True, records=True)
    sc = probe["scene"]
    _say(f"case serving: probe request {pi} on the {specs[big][0].n_tris}-tri block "
         f"({reqs[pi]['n_src']} sources x {reqs[pi]['n_rays']} rays + dome), alone")
    srv = H.Server(ctx, pool, hbm_budget=64 << 30, slots=8, max_batch=64,
                   shape=dict(lanes=1, stack_lds=12, waves_per_simd=6, regen=1, width=4),
                   latency_rays=1 << 18)
    t = srv.submit(sc, probe["params"], probe["sources"], probe["listeners"], dome=True,
                   keep=True, records=True)
    alone = srv.wait(t)
    _say(f"case serving: probe inside the workload's first {n_mix} requests at 2000 QPS")
    mix = list(R[:n_mix])
    mix[pi] = probe
    lat, tickets = srv.replay(mix, arr[:n_mix])
    assert np.all(lat > 0), "a request of the mix was refused"
    inmix = srv.wait(int(tickets[pi]))
    st = srv.stats()
    srv.close()
    _say(f"  mix: p50 {np.percentile(lat, 50) * 1e3:.2f} ms, p99 "
         f"{np.percentile(lat, 99) * 1e3:.2f} ms, batches {st['batches']}, "
         f"mean batch {st['mean_batch']:.2f}")
    plain = H.trace_world(sc, probe["params"], probe["sources"], probe["listeners"], records=True)
    dome = H.dome_sample(sc, probe["listeners"], H.dome_params(3, n_fan=512, horizon=500.0))
    for label, got in (("alone", alone), ("in mix", inmix)):
        _same(got["records"], got["bins"], got["dirs"], plain.records, plain.bins, plain.dirs,
              f"probe {label}")
        assert np.array_equal(got["dome_energy"], dome["energy"]), label
        assert np.array_equal(got["dome_dsum"], dome["dsum"]), label
    assert np.array_equal(alone["records"], inmix["records"])
    # and the probe's single-job trace is the oracle's (first two sources)
    geom = O.mesh_geom(oracle_mesh_from_builder(specs[big][0]))
    p = probe["params"]
    outs = _oracle_traces(geom, probe["sources"][:2], probe["listeners"], p.n_rays, nbands=3,
                          seed=p.seed, bin_dt=p.bin_dt, max_time=p.max_time)
    _vs_oracle(plain.records, plain.bins, plain.dirs, outs, p.n_rays, "probe vs oracle")
    for s in pool:
        s.close()
