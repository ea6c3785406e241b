"""Image sources with furniture: overlay-box faces as reflectors and
transient blockers (`extras`) on the legs, on the oracle and in the
library's host helpers.

Pins, on the oracle (the CPU restatement the GPU tests compare against):
- the reference's table-top gate (crates/omg-scene/tests/pt_mesh_gate.rs:
  180-240): the chain [table top] solves to the exact mirror image across
  the top plane (delay within 1e-5 of |image - listener| / 343), the
  reflection point off the rect is refused, and discovery over the box
  finds the top face;
- Aabb::clip on the legs (crates/omg-core/src/pt.rs:42-69, pt_mesh.rs:
  145-181): a leg that crosses a blocker once carries exactly its
  transmission per band (gains = clear gains x transmission), a leg that
  clears it carries nothing (gains unchanged), the situations of
  crates/omg-core/tests/pt_occluder.rs:77-117 (pillar between source and
  listener: direct and floor bounce through it, ceiling bounce over it);
- the library's host SurfaceTable::append_box equals the oracle's bit for
  bit.
"""
import numpy as np

import oracle as O
from omg_hip import CONCRETE, WOOD_PANEL, append_box, scenes, _lib
from helpers import mat_tuple, oracle_mesh_from_builder

OPAQUE = (0.02, 0.002, 0.0)          # pt.rs:39


def golden_box(size):
    """The reference test's room: six quads, one surface each
    (x = 0, x = X, y = 0, y = Y, z = 0 floor, z = Z ceiling)."""
    X, Y, Z = size
    mb = scenes.MeshBuilder()
    m = mb.material(CONCRETE)
    quads = [((0, 0, 0), (0, Y, 0), (0, Y, Z), (0, 0, Z)),
             ((X, 0, 0), (X, 0, Z), (X, Y, Z), (X, Y, 0)),
             ((0, 0, 0), (0, 0, Z), (X, 0, Z), (X, 0, 0)),
             ((0, Y, 0), (X, Y, 0), (X, Y, Z), (0, Y, Z)),
             ((0, 0, 0), (X, 0, 0), (X, Y, 0), (0, Y, 0)),
             ((0, 0, Z), (0, Y, Z), (X, Y, Z), (X, 0, Z))]
    for q in quads:
        mb.begin_surface()
        mb.quad(*q, m)
    return mb


def oracle_table_with_box(om, mn, mx, mat):
    tab, n = O.surface_table(om)
    box = (O.Surface * 6)()
    O.lib().or_surface_append_box(O.C.cast(box, O.C.c_void_p), O._p(O._f32(mn)),
                                  O._p(O._f32(mx)),
                                  O.C.byref(O.make_material(*mat_tuple(mat))), 3)
    for k in range(6):
        tab[n + k] = box[k]
    return tab, n


def test_table_top_reflects_exactly_on_the_oracle():
    om = oracle_mesh_from_builder(golden_box((6.0, 4.0, 3.0)))
    mn, mx = (2.0, 1.5, 0.0), (3.0, 2.5, 0.8)
    tab, base = oracle_table_with_box(om, mn, mx, WOOD_PANEL)
    top = base + 5                      # z max face
    src, lis = (1.0, 2.0, 1.5), (4.0, 2.0, 1.6)
    r = O.mesh_record(om, tab, base + 6, [top], 1, src, lis)
    assert r is not None, "table-top bounce must solve"
    img = np.array([1.0, 2.0, 2.0 * 0.8 - 1.5])
    want = np.linalg.norm(img - np.array(lis)) / 343.0
    assert abs(float(r["delay"]) - want) < 1e-5
    assert r["gains"][1] > 1e-4 and r["dir"][2] < -0.1
    # a reflection point off the rect does not validate
    assert O.mesh_record(om, tab, base + 6, [top], 1, src, (1.2, 0.3, 1.6)) is None
    # discovery over the box finds the top face
    found = False
    for rot in range(4):
        ch, od = O.mesh_chains(om, lis, 4096, rot, boxes=[(mn, mx)], base=base)
        found |= any(top in c[:o] for c, o in zip(ch, od))
    assert found, "discovery must bounce off the table top"
    # and without the box it cannot
    ch, od = O.mesh_chains(om, lis, 4096, 0)
    assert not any(c[k] >= base for c, o in zip(ch, od) for k in range(o))


def test_box_faces_in_discovery_are_numbered_like_append_box():
    """A ray straight down from above the table meets its top face first:
    face id base + 5 (z max); straight up from under it (a box lifted off
    the floor), base + 4 (z min); along +x from its left, base + 0."""
    om = oracle_mesh_from_builder(golden_box((6.0, 4.0, 3.0)))
    mn, mx = (2.0, 1.5, 0.5), (3.0, 2.5, 0.8)
    base = 6
    for lis, want in (((2.5, 2.0, 2.0), base + 5), ((2.5, 2.0, 0.2), base + 4),
                      ((1.0, 2.0, 0.65), base + 0), ((4.0, 2.0, 0.65), base + 1)):
        ch, od = O.mesh_chains(om, lis, 4096, 0, boxes=[(mn, mx)], base=base)
        firsts = {int(c[0]) for c, o in zip(ch, od) if o == 1}
        assert want in firsts, (lis, want, sorted(firsts))


def _pillar_room():
    om = oracle_mesh_from_builder(golden_box((8.0, 6.0, 3.0)))
    tab, n = O.surface_table(om)
    pillar = ((3.6, 2.2, 0.0), (4.4, 3.8, 2.2), OPAQUE)
    return om, tab, n, pillar


def test_blocker_on_a_leg_carries_its_transmission():
    """pt_occluder.rs:77-117 on the mesh solve: the pillar between source and
    listener scales the direct path by its transmission per band, and the
    floor bounce - whose reflection point (3.92, 3, 0) lies under the
    pillar, so both legs cross it - by the transmission squared; the
    ceiling bounce passes over it unchanged."""
    om, tab, n, pillar = _pillar_room()
    src, lis = (1.5, 3.0, 1.5), (6.5, 3.0, 1.6)
    floor, ceiling = 4, 5
    for chain, crossings in (([], 1), ([floor], 2), ([ceiling], 0)):
        clear = O.mesh_record(om, tab, n, chain, len(chain), src, lis)
        blocked = O.mesh_record(om, tab, n, chain, len(chain), src, lis, extras=[pillar])
        assert clear is not None
        if not crossings:
            assert blocked is not None and np.array_equal(blocked["gains"], clear["gains"])
            assert blocked["delay"] == clear["delay"]
            continue
        want = clear["gains"] * np.float32(OPAQUE) ** crossings
        if blocked is None:
            # dropped only below the audible floor (pt_mesh.rs:27, :311)
            assert np.max(want) < 2e-5
            continue
        assert blocked["delay"] == clear["delay"] and np.array_equal(blocked["dir"], clear["dir"])
        np.testing.assert_allclose(blocked["gains"], want, rtol=1e-6, atol=0.0)
    # a blocker the legs miss changes nothing; two crossings multiply
    far = ((7.0, 0.2, 0.0), (7.5, 0.6, 0.5), OPAQUE)
    thin = ((3.0, 2.0, 0.0), (3.1, 4.0, 3.0), (0.5, 0.25, 0.125))
    clear = O.mesh_record(om, tab, n, [], 0, src, lis)
    assert np.array_equal(O.mesh_record(om, tab, n, [], 0, src, lis, extras=[far])["gains"],
                          clear["gains"])
    two = O.mesh_record(om, tab, n, [], 0, src, lis, extras=[thin, pillar])
    np.testing.assert_allclose(two["gains"],
                               clear["gains"] * np.float32(thin[2]) * np.float32(OPAQUE),
                               rtol=2e-6, atol=0.0)


def test_host_append_box_equals_the_oracle():
    mn, mx = (2.0, 1.5, 0.0), (3.0, 2.5, 0.8)
    tab = (_lib.Surface * 8)()
    assert append_box(tab, 2, mn, mx, WOOD_PANEL) == 8
    ob = (O.Surface * 6)()
    O.lib().or_surface_append_box(O.C.cast(ob, O.C.c_void_p), O._p(O._f32(mn)), O._p(O._f32(mx)),
                                  O.C.byref(O.make_material(*mat_tuple(WOOD_PANEL))), 3)
    for k in range(6):
        a, b = tab[2 + k], ob[k]
        assert list(a.n) == list(b.n) and a.d == b.d and a.has_rect == b.has_rect == 1
        assert list(a.rmin) == list(b.rmin) and list(a.rmax) == list(b.rmax)
        assert list(a.refl[:3]) == list(b.refl[:3]) and list(a.trans[:3]) == list(b.trans[:3])
    # face k: axis k // 2, min side for even k; its plane is the box's bound there
    for k in range(6):
        axis, side = divmod(k, 2)
        want_n = [0.0, 0.0, 0.0]
        want_n[axis] = 1.0 if side else -1.0
        assert list(tab[2 + k].n) == want_n
        plane = np.float32((mx if side else mn)[axis])
        assert tab[2 + k].rmin[axis] == tab[2 + k].rmax[axis] == plane
        assert tab[2 + k].d == np.float32(want_n[axis]) * plane
