"""GPU parity of the late-field state (the Sim-side consumer of the
tracer).  The slots' averages live on the device; every update must give
exactly what the oracle's restatement of the reference's late path gives
(crates/omg-scene/src/sim.rs:374-397: Echogram::ema, late_direction,
estimate_reverb, the distance compensation): averages bit-identical,
RT60 / level / direction / anisotropy bit-identical, versions equal."""
import sys

import numpy as np
import pytest
import torch   # before the HIP library is loaded: torch brings its own HIP runtime

import oracle as O
import omg_hip as H
from omg_hip import scenes
from helpers import echogram_rel_errors, oracle_mesh_from_builder

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _decays(rng, nbins, nb, dt):
    """Echograms as the tracer makes them: sparse early arrivals on an
    exponential tail with a random T60 per band, random late directions."""
    t = (np.arange(nbins) + 0.5) * dt
    bins = np.zeros((nbins, nb), np.float32)
    for b in range(nb):
        t60 = rng.uniform(0.2, 3.0)
        tail = 10.0 ** (-6.0 * t / t60) * (rng.random(nbins) < 0.6) * rng.uniform(0.5, 1.5, nbins)
        bins[:, b] = (0.05 * tail).astype(np.float32)
    k = rng.integers(0, nbins // 10, 5)
    bins[k] += rng.uniform(0.1, 1.0, (5, nb)).astype(np.float32)
    dirs = (rng.normal(0, 1, (nbins, 3)) * bins[:, 1:2]).astype(np.float32)
    return bins, dirs


def _same(got, exp, label):
    assert np.array_equal(got.rt60, exp["rt60"]), (label, got.rt60, exp["rt60"])
    assert np.array_equal(got.level, exp["level"]), (label, got.level, exp["level"])
    assert np.array_equal(got.dir, exp["dir"]), (label, got.dir, exp["dir"])
    assert got.aniso == exp["aniso"], (label, got.aniso, exp["aniso"])
    assert got.version == exp["version"], label


@pytest.mark.parametrize("nb,dt,tmax", [(3, 0.001, 2.0), (6, 0.0025, 4.0), (2, 0.01, 3.0)])
def test_late_field_matches_oracle_over_many_ticks(ctx, oracle, nb, dt, tmax):
    dev = torch.device("cuda", 0)
    rng = np.random.default_rng(nb)
    n_slots = 40
    lf = H.LateField(ctx, n_slots, nbands=nb, bin_dt=dt, max_time=tmax)
    n = lf.nbins
    ref = [O.LateSlot(n, nb, dt) for _ in range(n_slots)]
    _say(f"case late field: {n_slots} slots, {n} bins x {nb} bands, 12 ticks")
    for tick in range(12):
        slots = rng.choice(n_slots, int(rng.integers(1, n_slots + 1)), replace=False)
        ups, exp, keep = [], [], []
        for s in slots:
            traced = bool(rng.random() < 0.6)
            src = rng.uniform(-10, 10, 3).astype(np.float32)
            rec = rng.uniform(-10, 10, 3).astype(np.float32)
            u = dict(slot=int(s), traced=traced, src=src, receiver=rec)
            if traced:
                b, d = _decays(rng, n, nb, dt)
                if rng.random() < 0.1:
                    b[:] = 0.0                      # a silent trace: default RT60
                tb, td = torch.from_numpy(b).to(dev), torch.from_numpy(d).to(dev)
                keep += [tb, td]
                u.update(bins=tb, dirs=td)
                exp.append(ref[s].step(True, b, d, src, rec))
            else:
                exp.append(ref[s].step(False, None, None, src, rec))
            ups.append(u)
        torch.cuda.synchronize()
        got = lf.update(ups)
        for u, g, e in zip(ups, got, exp):
            _same(g, e, f"tick {tick} slot {u['slot']}")
    for s in range(n_slots):
        b, d = lf.read(s)
        assert np.array_equal(b, ref[s].bins) and np.array_equal(d, ref[s].dirs), s
    lf.close()


def test_late_field_reverb_world_follows_the_gate(ctx, oracle):
    """Sim::reverb_world on the device: gate within a per-tick budget, trace
    the world into the field's staging, fold in.  The gate decisions equal
    the oracle's; each tick's outputs equal the oracle's late step on the
    average the device holds (whose folding is checked bit for bit above);
    a slot's first trace (alpha 1.0) leaves the traced echogram itself as
    the average, which matches the oracle's trace of the same seed."""
    mb, bounds, boxes = scenes.city_block(seed=71, n_buildings=6, target_tris=3000)
    scene = H.Scene.from_builder(ctx, mb)
    pts = scenes.outdoor_points(711, 8, bounds, boxes)
    lf = H.LateField(ctx, 4, nbands=3, bin_dt=0.001, max_time=1.0)
    ref = [O.LateSlot(lf.nbins, 3, 0.001) for _ in range(4)]
    gates = [O.TraceGate(s % 8) for s in range(4)]
    g = O.mesh_geom(oracle_mesh_from_builder(mb))
    _say("case late field reverb_world: 4 slots x 20 ticks, budget 2 per tick after the first")
    rng = np.random.default_rng(3)
    pos = [(pts[s].copy(), pts[s + 4].copy()) for s in range(4)]
    n_traced, first_checked = 0, 0
    for tick in range(20):
        budget = [4 if tick == 0 else 2]   # every slot traces on its first tick
        for s in range(4):
            src, rec = pos[s]
            if rng.random() < 0.3:
                src = (src + rng.normal(0, 0.2, 3)).astype(np.float32)
                pos[s] = (src, rec)
            fire = budget[0] > 0 and gates[s].should_trace(src, rec, np.ones(3, np.float32), 8)
            before, version0 = budget[0], ref[s].s.version
            got, seed = lf.reverb_world(s, scene, src, rec, rays=512, budget=budget)
            assert (seed is not None) == fire and budget[0] == before - int(fire), (tick, s)
            b, d = lf.read(s)
            ref[s].bins[:] = b            # the oracle steps on the device's average:
            ref[s].dirs[:] = d            # folding the average into itself leaves it as is
            exp = ref[s].step(fire, b, d, src, rec)
            _same(got, exp, f"tick {tick} slot {s}")
            n_traced += int(fire)
            if fire and version0 == 0:
                p = O.trace_params(512, nbands=3, rng_mode=O.RNG_COUNTER, seed=seed, src_index=0,
                                   bin_dt=0.001, max_time=1.0)
                o = O.trace(g, src, [rec], p, want_f64=True)
                eb, ed = echogram_rel_errors(b[None], d[None], o["bins64"], o["dirs64"], 1)
                assert eb <= 1e-4 and ed <= 1e-4, (s, eb, ed)
                first_checked += 1
    assert n_traced >= 8 and first_checked == 4
    lf.close()
    scene.close()


def test_late_field_refusals(ctx):
    lf = H.LateField(ctx, 3, nbands=3, bin_dt=0.001, max_time=0.5)
    z = np.zeros(3, np.float32)
    with pytest.raises(H.OmgError):
        lf.update([dict(slot=3, traced=False, src=z, receiver=z)])
    with pytest.raises(H.OmgError):
        lf.update([dict(slot=1, traced=False, src=z, receiver=z),
                   dict(slot=1, traced=False, src=z, receiver=z)])
    with pytest.raises(H.OmgError):
        lf.update([dict(slot=0, traced=True, bins=None, dirs=None, src=z, receiver=z)])
    with pytest.raises(H.OmgError):
        H.LateField(ctx, 2, nbands=1)
    lf.close()
