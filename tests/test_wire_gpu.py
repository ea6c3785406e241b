"""GPU runs of the web bridge's flat job queues: box jobs (57 f32 each,
crates/omg-web/src/lib.rs:151-185) on the analytic Shoebox and world jobs
(1034 f32 each, lib.rs:367-410) on a mesh scene with panels, each
echogram written in the injection layout (lib.rs:159-160).  Every
echogram is checked against the oracle's trace of the decoded job (same
counter RNG and seed) within 1e-4 of its f64 re-accumulation."""
import sys

import numpy as np
import pytest

import oracle as O
import omg_hip as H
from omg_hip import scenes, wire
from helpers import echogram_rel_errors, oracle_mesh_from_builder
from test_wire_cpu import world_job

pytestmark = pytest.mark.gpu

TOL = 1e-4


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def box_job(sid, n_rays, seed, size, src, lis, energy, walls):
    """The writer of lib.rs:178-189."""
    j = [float(sid), float(n_rays), float(seed)] + list(size) + list(src) + list(lis) + list(energy)
    for ab, sc, tr in walls:
        j += list(ab) + [sc] + list(tr)
    assert len(j) == 57
    return np.array(j, np.float32)


def test_box_jobs_match_the_oracle(ctx, oracle):
    rng = np.random.default_rng(21)
    jobs, meta = [], []
    for k in range(4):
        size = rng.uniform([6, 4, 2.5], [20, 12, 5])
        src = rng.uniform([1, 1, 1], size - 1)
        lis = rng.uniform([1, 1, 1], size - 1)
        walls = [(rng.uniform(0.02, 0.4, 3), float(rng.uniform(0, 0.5)), rng.uniform(0, 0.1, 3) * (w == 1))
                 for w in range(6)]
        energy = rng.uniform(0.5, 2.0, 3)
        seed = int(rng.integers(0, 1 << 31))
        jobs.append(box_job(40 + k, 2048 + 512 * k, seed, size, src, lis, energy, walls))
        meta.append((seed, size, src, lis, energy, walls))
    _say("case wire box jobs: 4 jobs")
    ids, echo = wire.trace_box_jobs(ctx, np.concatenate(jobs))
    assert list(ids) == [40, 41, 42, 43]
    for k, (seed, size, src, lis, energy, walls) in enumerate(meta):
        j = jobs[k]
        b, d = wire.split_echo(echo[k])
        g = O.box_geom(j[3:6], [(j[15 + 7 * w:18 + 7 * w], j[18 + 7 * w], j[19 + 7 * w:22 + 7 * w])
                                for w in range(6)])
        p = O.trace_params(int(j[1]), rng_mode=O.RNG_COUNTER, seed=int(np.float64(j[2])),
                           energy=list(j[12:15]))
        o = O.trace(g, j[6:9], [j[9:12]], p, want_f64=True)
        eb, ed = echogram_rel_errors(b[None], d[None], o["bins64"], o["dirs64"], 1)
        assert eb <= TOL and ed <= TOL, (k, eb, ed)
        assert b.sum() > 0
    with pytest.raises(H.OmgError) as e:
        wire.trace_box_jobs(ctx, jobs[0], version=1)
    assert e.value.code == H._lib.ERR_VERSION
    with pytest.raises(H.OmgError):
        wire.trace_box_jobs(ctx, jobs[0][:50])


def test_world_jobs_match_the_oracle(ctx, oracle):
    mb, bounds, boxes = scenes.city_block(seed=81, n_buildings=6, target_tris=3000)
    scene = H.Scene.from_builder(ctx, mb)
    pts = scenes.outdoor_points(811, 6, bounds, boxes)
    rng = np.random.default_rng(22)
    jobs, seeds = [], []
    for k in range(3):
        panels = [(pts[k] + rng.uniform(-3, -1, 3), pts[k] + rng.uniform(1, 3, 3),
                   (rng.uniform(0, 0.4, 3), float(rng.uniform(0, 0.5)), rng.uniform(0, 0.3, 3)))
                  for _ in range(k * 2)]
        seed = (0xC6D2 + (k + 1) * 0x9E3779B9) & 0xFFFFFFFF
        jobs.append(world_job(7 + k, 256, seed, pts[k], pts[k + 3], panels))
    _say("case wire world jobs: 3 jobs with 0, 2, 4 panels")
    ids, echo = wire.trace_world_jobs(scene, np.concatenate(jobs))
    assert list(ids) == [7, 8, 9]
    om = oracle_mesh_from_builder(mb)
    for k, j in enumerate(jobs):
        dj = wire.decode_world_job(j)
        pan = [list(dj.panels[i].min) + list(dj.panels[i].max) for i in range(dj.n_panels)]
        mats = [(list(dj.panels[i].mat.absorption[:3]), dj.panels[i].mat.scattering,
                 list(dj.panels[i].mat.transmission[:3])) for i in range(dj.n_panels)]
        g = O.mesh_geom(om, pan, mats) if pan else O.mesh_geom(om)
        p = O.trace_params(dj.n_rays, rng_mode=O.RNG_COUNTER, seed=dj.seed)
        o = O.trace(g, np.array(dj.src), [np.array(dj.lis)], p, want_f64=True)
        b, d = wire.split_echo(echo[k])
        eb, ed = echogram_rel_errors(b[None], d[None], o["bins64"], o["dirs64"], 1)
        assert eb <= TOL and ed <= TOL, (k, eb, ed)
    scene.close()
