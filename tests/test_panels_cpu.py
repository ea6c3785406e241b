"""Overlay boxes in the tracer (WithPanels::raycast_hit, crates/omg-core/src/
scene.rs:90-138) pinned on the oracle: closed-form entry-face hits and the
reference's own gate that a door leaf over a hole seals it
(crates/omg-gpu/tests/mesh_parity.rs:87-119, on the CPU tracer here).

Closed forms: a ray entering a box through the face of axis a at
t0 = (plane - p_a) / d_a is hit at t0 with the normal -sign(d_a) on that
axis; a ray that starts inside (entry at or before 1e-4) passes; where a
mesh wall is nearer, or exactly as near, the mesh wins (the box must be
strictly nearer).
"""
import numpy as np

import oracle as O
from omg_hip import CONCRETE, WOOD_PANEL, scenes
from helpers import mat_tuple, oracle_mesh_from_builder


def holed_room():
    """mesh_parity.rs:16-29: an 8 x 6 x 3 m room with a 1 x 2.1 m hole in
    the x = 8 wall (y 2.5-3.5)."""
    mb = scenes.MeshBuilder()
    m = mb.material(CONCRETE)
    for q in (((0, 0, 0), (0, 6, 0), (0, 6, 3), (0, 0, 3)),
              ((0, 0, 0), (8, 0, 0), (8, 0, 3), (0, 0, 3)),
              ((0, 6, 0), (8, 6, 0), (8, 6, 3), (0, 6, 3)),
              ((0, 0, 0), (8, 0, 0), (8, 6, 0), (0, 6, 0)),
              ((0, 0, 3), (8, 0, 3), (8, 6, 3), (0, 6, 3)),
              ((8, 0, 0), (8, 2.5, 0), (8, 2.5, 3), (8, 0, 3)),
              ((8, 3.5, 0), (8, 6, 0), (8, 6, 3), (8, 3.5, 3)),
              ((8, 2.5, 2.1), (8, 3.5, 2.1), (8, 3.5, 3), (8, 2.5, 3))):
        mb.quad(*q, m)
    return mb


LEAF = ((7.96, 2.5, 0.0), (8.04, 3.5, 2.1))


def _geom(om, boxes):
    return O.mesh_geom(om, [list(mn) + list(mx) for mn, mx in boxes],
                       [mat_tuple(WOOD_PANEL)] * len(boxes))


def test_entry_face_hits_in_closed_form():
    om = oracle_mesh_from_builder(holed_room())
    bare = O.mesh_geom(om)
    g = _geom(om, [LEAF, ((2.0, 1.0, 0.0), (4.0, 2.0, 1.0))])
    f = np.float32
    # through the hole: no mesh hit; the leaf's x-min face at (7.96 - 2) / 1
    assert O.raycast_hit(bare, (2, 3, 1.5), (1, 0, 0)) is None
    h = O.raycast_hit(g, (2, 3, 1.5), (1, 0, 0))
    assert h["panel"] == 0 and h["t"] == f(f(7.96) - f(2.0))
    assert np.array_equal(h["normal"], [-1, 0, 0])
    # the y-min face of the second box, from below in y
    h = O.raycast_hit(g, (3, 0.5, 0.5), (0, 1, 0))
    assert h["panel"] == 1 and h["t"] == f(0.5) and np.array_equal(h["normal"], [0, -1, 0])
    # its top face from above, oblique: entry through z = 1 at t = (1 - 2.5) / dz
    d = np.float32([0.0, 0.6, -0.8])
    h = O.raycast_hit(g, (3, 0.2, 2.5), d)
    assert h["panel"] == 1 and h["t"] == f(f(f(1.0) - f(2.5)) / d[2])
    assert np.array_equal(h["normal"], [0, 0, 1])
    # leaving from inside the leaf: its entry is behind the origin, not taken;
    # the ray goes on through the hole and meets nothing
    assert O.raycast_hit(g, (8.0, 3.0, 1.5), (1, 0, 0)) is None
    # a mesh wall nearer than the box wins: along -x from inside the room the
    # x = 0 wall is at 2, the box (behind it) would be at 3
    g2 = _geom(om, [((-1.5, 2.0, 1.0), (-1.0, 4.0, 2.0))])
    h = O.raycast_hit(g2, (2, 3, 1.5), (-1, 0, 0))
    assert h["panel"] is None and h["t"] == f(2.0)
    # a box face on the mesh wall's plane ties with it: the mesh keeps it
    g3 = _geom(om, [((-1.0, 2.0, 1.0), (0.0, 4.0, 2.0))])
    h = O.raycast_hit(g3, (2, 3, 1.5), (-1, 0, 0))
    assert h["panel"] is None and h["t"] == f(2.0)
    # ... and a box face in front of it wins
    g4 = _geom(om, [((0.5, 2.0, 1.0), (0.7, 4.0, 2.0))])
    h = O.raycast_hit(g4, (2, 3, 1.5), (-1, 0, 0))
    assert h["panel"] == 0 and h["t"] == f(f(0.7) - f(2.0)) / f(-1.0)
    assert np.array_equal(h["normal"], [1, 0, 0])


def test_panel_over_the_hole_seals_it():
    """mesh_parity.rs:87-119: the outside listener's mid-band energy with the
    leaf in the doorway is below 15 % of the open doorway's."""
    om = oracle_mesh_from_builder(holed_room())
    src, lis = (2.0, 3.0, 1.5), (11.0, 3.0, 1.6)
    p = O.trace_params(16384, nbands=3, rng_mode=O.RNG_COUNTER, seed=7)
    open_e = O.trace(O.mesh_geom(om), src, [lis], p)["bins"][0][:, 1].sum()
    closed_e = O.trace(_geom(om, [LEAF]), src, [lis], p)["bins"][0][:, 1].sum()
    assert open_e > 0
    assert closed_e < 0.15 * open_e, (closed_e, open_e)
