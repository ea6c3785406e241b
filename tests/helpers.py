"""Shared test scaffolding: one scene, two builds (oracle and product)."""
import numpy as np

import oracle as O
from omg_hip import MaterialDef, scenes


def mat_tuple(m):
    return (list(m.absorption), m.scattering, list(m.transmission))


def oracle_mesh_from_builder(mb, nbands=3):
    """Build the oracle's Mesh from a MeshBuilder soup (oracle dedup, oracle BVH)."""
    soup = mb.soup()
    pos, idx = O.dedup_tris(soup["tri_verts"])
    mats = [mat_tuple(m) for m in soup["materials"]]
    return O.Mesh(pos, idx, soup["tri_material"], mats, nbands, soup["tri_surface"],
                  soup["tri_max_edge"])


def rel_err(a, ref, scale=None):
    """max |a - ref| / (|ref| or scale), over entries where the scale is non-zero."""
    a = np.asarray(a, np.float64)
    ref = np.asarray(ref, np.float64)
    s = np.abs(ref) if scale is None else np.asarray(scale, np.float64)
    diff = np.abs(a - ref)
    nz = s > 0
    if np.any(diff[~nz] > 0):
        return np.inf
    if not np.any(nz):
        return 0.0
    return float(np.max(diff[nz] / s[nz]))


def echogram_rel_errors(bins, dirs, bins64, dirs64, dir_band=1):
    """Per-bin relative error of f32 echograms against f64 re-accumulations.

    Energy bins are compared relative to the f64 value; direction sums
    (signed, may cancel) relative to the bin's dir-band energy, which
    bounds their magnitude."""
    eb = rel_err(bins, bins64)
    scale = np.repeat(np.abs(bins64[..., dir_band:dir_band + 1]), 3, axis=-1)
    ed = rel_err(dirs, dirs64, scale)
    return eb, ed


CONCRETE = MaterialDef([0.01, 0.02, 0.04], 0.10, [0.10, 0.014, 0.0015])
DRYWALL = MaterialDef([0.20, 0.08, 0.05], 0.10, [0.20, 0.045, 0.0150])
CARPET = MaterialDef([0.05, 0.30, 0.60], 0.40, [0.30, 0.100, 0.0300])
WOOD = MaterialDef([0.20, 0.10, 0.08], 0.15, [0.22, 0.060, 0.0200])
TILE = MaterialDef([0.30, 0.70, 0.80], 0.30, [0.30, 0.100, 0.0300])

# crates/omg-core/tests/trace_golden.rs:51-109
GOLDEN_CONFIGS = [
    ("live-room", (6.0, 4.0, 3.0), [CONCRETE] * 4 + [WOOD, CONCRETE], (1.5, 1.0, 1.5),
     (4.5, 3.0, 1.6), [1.19, 1.05, 1.05, -6.8, -3.4, -4.7, 0.010]),
    ("dead-room", (12.0, 9.0, 4.0), [TILE] * 4 + [CARPET, TILE], (3.0, 2.0, 1.5),
     (9.0, 7.0, 1.6), [0.49, 0.20, 0.16, -26.6, -42.9, -50.5, 0.187]),
    ("hall", (16.0, 3.5, 3.0), [DRYWALL] * 4 + [WOOD, CONCRETE], (1.5, 1.75, 1.5),
     (14.0, 1.75, 1.6), [0.75, 1.32, 1.46, -19.8, -10.3, -8.4, 0.014]),
]
GOLDEN_RAYS = 4096          # trace_golden.rs:22
GOLDEN_AVG = 4              # trace_golden.rs:28
RT60_REL_TOL = 0.12         # trace_golden.rs:33
LEVEL_DB_TOL = 1.0          # trace_golden.rs:34
ANISO_TOL = 0.06            # trace_golden.rs:35
ANISO_LEVEL_FLOOR_DB = -40.0  # trace_golden.rs:40


def golden_violations(got, expect):
    """assert_golden (trace_golden.rs:145-179) as a list of violations."""
    bad = []
    for b in range(3):
        rel = abs(got[b] - expect[b]) / expect[b]
        if rel > RT60_REL_TOL:
            bad.append(f"rt60[{b}] {got[b]:.3f} vs {expect[b]}")
        tol = 4.0 if expect[3 + b] < ANISO_LEVEL_FLOOR_DB else LEVEL_DB_TOL
        if abs(got[3 + b] - expect[3 + b]) > tol:
            bad.append(f"level[{b}] {got[3 + b]:.2f} vs {expect[3 + b]}")
    if expect[4] > ANISO_LEVEL_FLOOR_DB and abs(got[6] - expect[6]) > ANISO_TOL:
        bad.append(f"aniso {got[6]:.3f} vs {expect[6]}")
    return bad


def c1_scene():
    """Config 1: 12 x 8 x 4 m shoebox as 12 triangles, seed 1."""
    return scenes.shoebox(seed=1)


def assert_tree_matches_oracle(prims, nodes, root, om):
    """The product's packed primitives [n][12] and paired nodes [n][16]
    (omg_bvh_build_host / omg_scene_export rows, root reference `root`)
    are the oracle Mesh's tree: primitives bit for bit in the same order,
    and walking both trees together, the same child boxes and leaf ranges
    at every node (boxes compared as values: a zero may carry either sign)."""
    onodes, oprims = om.export()
    pv = prims.view(np.uint32)
    ov = oprims.view(np.uint32)
    assert pv.shape[0] == ov.shape[0]
    for a, b in [((0, 1, 2), (0, 1, 2)), ((4, 5, 6), (3, 4, 5)), ((8, 9, 10), (6, 7, 8)),
                 ((3,), (9,)), ((7,), (10,))]:
        assert np.array_equal(pv[:, list(a)], ov[:, list(b)])
    on_u = onodes.view(np.uint32)

    def oref(i):
        a, b = on_u[i, 6], on_u[i, 7]
        if a & 0x80000000:
            return ("leaf", int(a & 0x7FFFFFFF), int(b))
        return ("node", i)

    def pref(r):
        if r & 0x80000000:
            return ("leaf", int(r & 0x0FFFFFFF), int((r >> 28) & 7))
        return ("node", int(r))

    stack = [(0, root)]
    seen = 0
    while stack:
        oi, r = stack.pop()
        o = oref(oi)
        p = pref(r)
        assert o[0] == p[0]
        if o[0] == "leaf":
            assert o[1:] == p[1:]
            continue
        seen += 1
        n = nodes[p[1]]
        nu = n.view(np.uint32)
        la, lb = int(on_u[oi, 6]), int(on_u[oi, 7])
        assert np.array_equal(n[0:3], onodes[la, 0:3]) and np.array_equal(n[4:7], onodes[la, 3:6])
        assert np.array_equal(n[8:11], onodes[lb, 0:3]) and np.array_equal(n[12:15], onodes[lb, 3:6])
        stack.append((la, int(nu[3])))
        stack.append((lb, int(nu[7])))
    assert seen == nodes.shape[0]
