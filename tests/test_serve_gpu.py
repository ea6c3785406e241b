"""GPU tests of the serving loop.

The serving loop has no reference counterpart: each request must give what
the single-job entry points give for it (omg_trace_world for the late
field, omg_dome_sample for the dome), alone and in a mix of concurrent
requests.  Per-ray records bit-identical, echograms within 1e-4 relative
(atomic summation order), dome sums bit-identical (ordered reduction per
eye).  Admission by HBM budget is checked with budgets that hold one
request at a time and budgets that refuse a request outright.
"""
import sys

import numpy as np
import pytest

import omg_hip as H
from omg_hip import scenes
from helpers import echogram_rel_errors

pytestmark = pytest.mark.gpu

ECHO_TOL = 1e-4


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def world(ctx):
    a, ba, bxa = scenes.city_block(seed=51, n_buildings=6, target_tris=4000, nbands=3)
    b, bb, bxb = scenes.city_block(seed=52, n_buildings=10, target_tris=12000, nbands=3)
    sa = H.Scene.from_builder(ctx, a, nbands=3)
    sb = H.Scene.from_builder(ctx, b, nbands=3)
    pa = np.vstack([scenes.indoor_points(511, 6, bxa), scenes.outdoor_points(512, 6, ba, bxa)])
    pb = np.vstack([scenes.indoor_points(521, 6, bxb), scenes.outdoor_points(522, 6, bb, bxb)])
    yield dict(scenes=(sa, sb), points=(pa, pb))
    sa.close()
    sb.close()


def _params(n_rays, seed):
    return H.trace_params(n_rays, 3, seed=seed, bin_dt=0.001, max_time=1.0, floor_rel=1e-6)


def _mix(world, n, rng):
    sa, sb = world["scenes"]
    out = []
    for i in range(n):
        k = int(rng.integers(0, 2))
        pts = world["points"][k]
        ns = int(rng.integers(1, 5))
        out.append(dict(scene=(sa, sb)[k], params=_params(int(rng.integers(100, 3000)), 9000 + i),
                        sources=pts[rng.integers(0, len(pts), ns)], listeners=pts[[i % len(pts)]],
                        dome=bool(i % 2)))
    return out


@pytest.mark.parametrize("latency_rays,shape", [(0, None), (1 << 20, None), (0, "quad")])
def test_request_alone_and_in_mix_match_plain(ctx, world, latency_rays, shape):
    """latency_rays = 0: every batch runs the regenerating walk (one lane per
    ray, or a quad per ray with the whole stack in LDS); 1 << 20: these
    small batches run a DPP quad per ray over 16-ray items."""
    sa, sb = world["scenes"]
    pts = world["points"][1]
    probe = dict(scene=sb, params=_params(2500, 4242), sources=pts[[0, 3, 7]],
                 listeners=pts[[5, 9]], dome=True, keep=True, records=True)
    dp = H.dome_params(3, n_fan=512, horizon=500.0)
    sh = dict(lanes=4, stack_lds=40, waves_per_simd=7, regen=1, width=4) if shape else None
    srv = H.Server(ctx, [sa, sb], hbm_budget=1 << 34, slots=3, max_batch=16, dome=dp,
                   latency_rays=latency_rays, shape=sh)
    _say(f"case serve (latency_rays={latency_rays}, shape {shape}): probe request alone")
    t = srv.submit(probe["scene"], probe["params"], probe["sources"], probe["listeners"],
                   dome=True, keep=True, records=True)
    alone = srv.wait(t)
    _say("case serve: probe inside a 60-request mix at 2000 QPS")
    rng = np.random.default_rng(7)
    mix = _mix(world, 60, rng)
    mix.insert(30, probe)
    arrivals = np.cumsum(rng.exponential(1.0 / 2000.0, len(mix)))
    lat, tickets = srv.replay(mix, arrivals)
    assert np.all(lat > 0), lat
    inmix = srv.wait(int(tickets[30]))
    st = srv.stats()
    srv.close()
    _say(f"  mix: p50 {np.percentile(lat, 50) * 1e3:.2f} ms, p99 {np.percentile(lat, 99) * 1e3:.2f} ms, "
         f"batches {st['batches']}, mean batch {st['mean_batch']:.2f}, in flight <= {st['max_in_flight']}")
    assert st["completed"] == st["submitted"] == 62
    plain = H.trace_world(sb, probe["params"], probe["sources"], probe["listeners"], records=True)
    dome = H.dome_sample(sb, probe["listeners"], dp)
    for label, got in (("alone", alone), ("in mix", inmix)):
        assert np.array_equal(got["records"], plain.records), label
        eb, ed = echogram_rel_errors(got["bins"], got["dirs"], plain.bins.astype(np.float64),
                                     plain.dirs.astype(np.float64))
        assert eb <= ECHO_TOL and ed <= ECHO_TOL, (label, eb, ed)
        assert np.array_equal(got["dome_energy"], dome["energy"]), label
        assert np.array_equal(got["dome_dsum"], dome["dsum"]), label
    assert np.array_equal(alone["records"], inmix["records"])


def test_budget_admission(ctx, world):
    sa, sb = world["scenes"]
    resident = sa.info()["device_bytes"] + sb.info()["device_bytes"]
    pts = world["points"][0]
    one = dict(scene=sa, params=_params(3000, 77), sources=pts[:4], listeners=pts[[6]], dome=True)
    # a budget with room for about one such request beside the scenes
    srv = H.Server(ctx, [sa, sb], hbm_budget=resident + 200_000, slots=4, max_batch=8)
    _say("case serve: budget holding one request at a time")
    ts = [srv.submit(**one) for _ in range(5)]
    for t in ts:
        srv.wait(t)
    st = srv.stats()
    assert st["completed"] == 5
    assert st["peak_committed_bytes"] <= resident + 200_000, st
    # a request that can never fit is refused
    big = dict(scene=sa, params=H.trace_params(4096, 3, bin_dt=0.001, max_time=4.0),
               sources=pts[:12], listeners=pts[:4])
    with pytest.raises(H.OmgError) as e:
        srv.submit(**big)
    assert e.value.code == H._lib.ERR_BUDGET
    assert srv.stats()["refused"] == 1
    srv.close()
    # scenes alone above the budget: the server is refused
    with pytest.raises(H.OmgError) as e:
        H.Server(ctx, [sa, sb], hbm_budget=resident - 1)
    assert e.value.code == H._lib.ERR_BUDGET


def test_results_protocol(ctx, world):
    sa, _ = world["scenes"]
    pts = world["points"][0]
    srv = H.Server(ctx, [sa], slots=1, max_batch=1)
    t = srv.submit(sa, _params(500, 5), pts[:2], pts[[3]], dome=False, keep=True)
    r = srv.wait(t)
    assert r["bins"].shape == (2, 1, H.nbins(1.0, 0.001), 3)
    assert r["latency_s"] > 0 and r["batch_size"] == 1
    with pytest.raises(H.OmgError):
        srv.result(t)                       # collected already
    _, other = world["scenes"]
    with pytest.raises(H.OmgError):
        srv.submit(other, _params(100, 1), pts[:1], pts[[1]])   # not resident here
    srv.close()
