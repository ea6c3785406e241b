"""Closed forms that pin the oracle's dome probe (DomeProbe::trace_escape,
trace_through and sample; crates/omg-scene/src/dome.rs:330-459) ray by ray.

- One reflection off a ground plane: a fan ray that meets the floor
  within the horizon leaves with throughput (1 - absorption) per band
  exactly, after two events (the floor, then the open sky); a ray that
  meets it beyond the horizon, or never, leaves with 1 (dome.rs:350-379).
- Straight through one wall: the through ray pays t^2 per crossing
  (dome.rs:398-406), and the per-band max with the bounce throughput is
  taken when the bounce path is weaker than 0.5 in band 0 (dome.rs:
  439-450).
- Glass panes multiply a bounce ray by g^2 and let it go on; a door leaf
  ends it, and the through ray then pays the leaf's t^2 (dome.rs:336-349,
  :407-420).
A plausible slip - t instead of t^2, absorption instead of 1 - absorption,
no max, the horizon rule - fails one of them.
"""
import numpy as np

import oracle as O
from omg_hip import MaterialDef, scenes
from helpers import oracle_mesh_from_builder

EYE = (0.0, 0.0, 1.6)


def _mesh(quads, mat):
    mb = scenes.MeshBuilder()
    m = mb.material(mat)
    mb.coarse(1000.0)
    for q in quads:
        mb.quad(*q, m)
    return oracle_mesh_from_builder(mb)


def test_one_ground_reflection_costs_one_minus_absorption():
    alpha = (0.2, 0.3, 0.45)
    om = _mesh([((-2000, -2000, 0), (2000, -2000, 0), (2000, 2000, 0), (-2000, 2000, 0))],
               MaterialDef(list(alpha), 0.0, [0.0, 0.0, 0.0]))
    o = O.dome_sample(om, EYE, n_fan=384, horizon=80.0, want_rays=True)
    fan = O.dome_fan(384)
    keep = np.float32(1.0) - np.float32(alpha)
    n_refl = 0
    for j in range(384):
        d = fan[j]
        assert o["escaped"][j] == 1 and o["through"][j] == 0
        if d[2] >= 0:
            assert np.all(o["tp"][j] == 1.0) and o["events"][j] == 1
            continue
        t = 1.6 / -float(d[2])
        if abs(t - 80.0) < 1.0:
            continue                              # at the horizon: either rule
        if t > 80.0:
            assert np.all(o["tp"][j] == 1.0) and o["events"][j] == 1
        else:
            n_refl += 1
            assert np.array_equal(o["tp"][j], keep), (j, o["tp"][j])
            assert o["events"][j] == 2
    assert n_refl > 150


def test_straight_through_one_wall_pays_t_squared():
    """A wall at x = 5 that absorbs 90 % and transmits t: a ray into it
    bounces back at 0.1 (< 0.5, so the through ray is traced) and the
    through ray crosses once at t^2; the ray keeps the per-band max."""
    t = np.float32([0.5, 0.3, 0.1])
    om = _mesh([((5, -300, -300), (5, 300, -300), (5, 300, 300), (5, -300, 300))],
               MaterialDef([0.9, 0.9, 0.9], 0.0, list(t)))
    o = O.dome_sample(om, EYE, n_fan=384, horizon=80.0, want_rays=True)
    fan = O.dome_fan(384)
    bounce = np.float32(1.0) - np.float32(0.9)
    n_through = 0
    for j in range(384):
        d = fan[j]
        if d[0] <= 0.12 or d[2] < -0.03:
            continue                              # misses the wall, or near the ground rule
        n_through += 1
        assert o["through"][j] == 1
        bp = bounce if o["escaped"][j] else np.float32(0.0)
        want = np.maximum(t * t, np.float32(bp))
        assert np.array_equal(o["tp"][j], want), (j, o["tp"][j], want)
    assert n_through > 30
    # rays away from the wall meet nothing
    away = fan[:, 0] < -0.01
    assert np.all(o["tp"][away] == 1.0)


def test_glass_and_leaf_panels():
    """Panels in the plane x = 5 (|y| < 10, 0 < z < 10), no mesh in the way:
    a glass pane passes a bounce ray at g^2 (dome.rs:341-347); a leaf blocks
    it, and the straight ray then pays the leaf's t^2 (dome.rs:407-420)."""
    far = _mesh([((5000, 0, 0), (5000, 1, 0), (5000, 1, 1), (5000, 0, 1))],
                MaterialDef([0.1] * 3, 0.0, [0.0] * 3))
    g = np.float32([0.9, 0.8, 0.6])
    fan = O.dome_fan(384)
    # the rays that cross the rectangle, with a margin
    tt = 5.0 / np.maximum(fan[:, 0], 1e-9)
    y, z = tt * fan[:, 1], EYE[2] + tt * fan[:, 2]
    inside = (fan[:, 0] > 0.05) & (np.abs(y) < 9.5) & (z > 0.5) & (z < 9.5)
    assert inside.sum() > 20
    for glass in (1, 0):
        pan = dict(axis=0, plane=5.0, lat=(-10.0, 10.0), z=(0.0, 10.0), glass=glass, trans=list(g))
        o = O.dome_sample(far, EYE, n_fan=384, horizon=80.0, panels=[pan], want_rays=True)
        for j in np.nonzero(inside)[0]:
            if glass:
                # g^2 >= 0.5 in band 0: the bounce path alone
                assert o["escaped"][j] == 1 and o["through"][j] == 0
                assert np.array_equal(o["tp"][j], g * g), (j, o["tp"][j])
            else:
                assert o["escaped"][j] == 0 and o["through"][j] == 1
                # the 50 m straight line must stay above grade (dome.rs:393-395)
                want = g * g if EYE[2] + 50.0 * fan[j, 2] >= -0.05 else np.zeros(3, np.float32)
                assert np.array_equal(o["tp"][j], want), (j, o["tp"][j])
        assert np.all(o["tp"][fan[:, 0] < 0] == 1.0)


def test_tier_stride_takes_every_kth_fan_ray():
    """dome.rs:431-434: n_rays < N_RAYS traces fan ray j * N / n: bins and
    throughputs of a subsampled pass equal those rays of the full pass."""
    om = _mesh([((-2000, -2000, 0), (2000, -2000, 0), (2000, 2000, 0), (-2000, 2000, 0))],
               MaterialDef([0.2, 0.3, 0.45], 0.0, [0.0] * 3))
    full = O.dome_sample(om, EYE, n_fan=384, horizon=80.0, want_rays=True)
    for n in (256, 160):
        sub = O.dome_sample(om, EYE, n_fan=384, n_rays=n, horizon=80.0, want_rays=True)
        idx = [(j * 384) // n for j in range(n)]
        assert np.array_equal(sub["tp"], full["tp"][idx])
        assert np.array_equal(sub["bin"], full["bin"][idx])
