"""CPU tests of the library's host side: scene build equals the oracle's,
and the C ABI exports every declared symbol."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import oracle as O
from omg_hip import _lib, scenes
from helpers import assert_tree_matches_oracle, oracle_mesh_from_builder

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "omg_hip.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|void|const char \*)\s*\*?(omg_\w+)\s*\(", text,
                                 re.M)))


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    syms = declared_symbols()
    assert len(syms) >= 20
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing
    bound = {s[0] for s in _lib.SIGNATURES}
    assert not [s for s in syms if s not in bound], "binding lacks declared entry points"


_RUNTIME_PROBE = """
import sys
sys.path.insert(0, {pkg!r})
from omg_hip import _lib
if {torch_first}:
    import torch
_lib.lib()
if not {torch_first}:
    import torch
try:
    _lib.check_single_runtime()
    print("one", len(_lib.hip_runtimes()))
except RuntimeError as e:
    print("refused", e)
"""


@pytest.mark.parametrize("torch_first", [True, False])
def test_a_second_hip_runtime_is_refused(torch_first):
    """torch imported first: the library binds to torch's HIP runtime (one
    mapped); torch imported after the library: two runtimes, refused."""
    import subprocess
    import sys
    pkg = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "omg-audio_amd")
    out = subprocess.run([sys.executable, "-c", _RUNTIME_PROBE.format(pkg=pkg, torch_first=torch_first)],
                         capture_output=True, text=True, timeout=600)
    assert out.returncode == 0, out.stderr[-2000:]
    last = out.stdout.strip().splitlines()[-1]
    if torch_first:
        assert last == "one 1", last
    else:
        assert last.startswith("refused") and "import torch before omg_hip" in last, last


def _host_build(mb):
    L = _lib.lib()
    soup = mb.soup()
    tv = np.ascontiguousarray(soup["tri_verts"])
    nt = tv.shape[0]
    pos = np.zeros((nt * 3, 3), np.float32)
    idx = np.zeros((nt, 3), np.uint32)
    nv = C.c_uint32()
    _lib.check(L.omg_dedup_soup(tv.ctypes.data, nt, pos.ctypes.data, idx.ctypes.data,
                                C.byref(nv)))
    pos = pos[:nv.value].copy()
    npr, nn, root, depth = C.c_uint32(), C.c_uint32(), C.c_uint32(), C.c_uint32()
    args = [pos.ctypes.data, nv.value, idx.ctypes.data, nt, soup["tri_material"].ctypes.data,
            soup["tri_surface"].ctypes.data, soup["tri_max_edge"].ctypes.data]
    _lib.check(L.omg_bvh_build_host(*args, None, 0, C.byref(npr), None, 0, C.byref(nn),
                                    C.byref(root), C.byref(depth)))
    prims = np.zeros((npr.value, 12), np.float32)
    nodes = np.zeros((max(1, nn.value), 16), np.float32)
    _lib.check(L.omg_bvh_build_host(*args, prims.ctypes.data, npr.value, C.byref(npr),
                                    nodes.ctypes.data, nn.value, C.byref(nn), C.byref(root),
                                    C.byref(depth)))
    return pos, idx, prims, nodes[:nn.value], root.value, depth.value


@pytest.mark.parametrize("which", ["shoebox", "city", "soup"])
def test_scene_build_matches_oracle(which, oracle):
    if which == "shoebox":
        mb = scenes.shoebox(seed=1)[0]
    elif which == "city":
        mb = scenes.city_block(seed=2, n_buildings=8, target_tris=6000)[0]
    else:
        # random triangle soup as in mesh.rs:697-715 (tri edges up to 4 m: tessellation kicks in)
        rng = scenes.SplitMix64(7)
        mb = scenes.MeshBuilder()
        m = mb.material(scenes.MaterialDef([0.1] * 3, 0.1, [0.0] * 3))
        for _ in range(200):
            b = [rng.uniform(0, 20) for _ in range(3)]
            mb.tri(tuple(b), tuple(x + rng.uniform(-3, 3) for x in b),
                   tuple(x + rng.uniform(-3, 3) for x in b), m)
    pos, idx, prims, nodes, root, depth = _host_build(mb)
    om = oracle_mesh_from_builder(mb)
    opos, oidx = O.dedup_tris(mb.soup()["tri_verts"])
    assert np.array_equal(pos, opos) and np.array_equal(idx, oidx)
    assert_tree_matches_oracle(prims, nodes, root, om)


def test_host_postprocessing_matches_oracle(oracle):
    """estimate_reverb / late_direction / ism_box run on the host: same bits as the oracle."""
    import omg_hip as H
    rng = np.random.default_rng(3)
    t = np.arange(300) * 0.01
    bins = (np.exp(-t / 0.2)[:, None] * rng.uniform(0.5, 1.5, (300, 3))).astype(np.float32)
    dirs = rng.normal(size=(300, 3)).astype(np.float32) * bins[:, 1:2]
    assert all(np.array_equal(a, b) for a, b in zip(H.estimate_reverb(bins, 0.01),
                                                    O.estimate_reverb(bins, 0.01)))
    d1, a1 = H.late_direction(bins, dirs, 0.01, 0.05)
    d2, a2 = O.late_direction(bins, dirs, 0.01, 0.05)
    assert np.array_equal(d1, d2) and a1 == a2
    walls = [H.DRYWALL] * 6
    taps = H.ism_box((8, 6, 3), walls, (2, 2, 1.5), (5, 4, 1.5), 3)
    ref = O.image_source_taps((8, 6, 3), [(m.absorption, m.scattering, m.transmission)
                                          for m in walls], (2, 2, 1.5), (5, 4, 1.5), 3)
    assert taps.shape == (63, 8) and np.array_equal(taps, ref)


def _host_build4(mb):
    L = _lib.lib()
    soup = mb.soup()
    pos, idx = O.dedup_tris(soup["tri_verts"])
    nt = idx.shape[0]
    args = [pos.ctypes.data, pos.shape[0], idx.ctypes.data, nt, soup["tri_material"].ctypes.data,
            soup["tri_surface"].ctypes.data, soup["tri_max_edge"].ctypes.data]
    n4, root4, stack4, usable = C.c_uint32(), C.c_uint32(), C.c_uint32(), C.c_uint32()
    _lib.check(L.omg_bvh4_build_host(*args, None, 0, C.byref(n4), C.byref(root4),
                                     C.byref(stack4), C.byref(usable)))
    nodes4 = np.zeros((max(1, n4.value), 32), np.float32)
    _lib.check(L.omg_bvh4_build_host(*args, nodes4.ctypes.data, n4.value, C.byref(n4),
                                     C.byref(root4), C.byref(stack4), C.byref(usable)))
    return nodes4[:n4.value], root4.value, stack4.value, usable.value


@pytest.mark.parametrize("which", ["shoebox", "city", "soup"])
def test_bvh4_collapse_covers_the_binary_tree(which):
    """The 4-wide tree is the binary tree two levels at a time: node i of
    binary node X holds in slots 0-1 the children of X's left child (or
    that child itself in slot 0 when it is a leaf, slot 1 unused) and in
    slots 2-3 those of its right child, with the binary boxes and leaf
    references; 4-wide nodes are numbered breadth-first; the stack bound
    is the largest sum of (children - 1) along a root-to-leaf path."""
    if which == "shoebox":
        mb = scenes.shoebox(seed=1)[0]
    elif which == "city":
        mb = scenes.city_block(seed=2, n_buildings=8, target_tris=6000)[0]
    else:
        rng = scenes.SplitMix64(7)
        mb = scenes.MeshBuilder()
        m = mb.material(scenes.MaterialDef([0.1] * 3, 0.1, [0.0] * 3))
        for _ in range(200):
            b = [rng.uniform(0, 20) for _ in range(3)]
            mb.tri(tuple(b), tuple(x + rng.uniform(-3, 3) for x in b),
                   tuple(x + rng.uniform(-3, 3) for x in b), m)
    _, _, _, nodes, root, _ = _host_build(mb)
    n4, root4, stack4, usable = _host_build4(mb)
    nu = nodes.view(np.uint32)
    if root & 0x80000000:
        assert root4 == root and n4.shape[0] == 0
        return
    assert root4 == 0
    u4 = n4.view(np.uint32)
    NONE = 0xFFFFFFFF

    def child(b, side):
        """(ref, min, max) of binary node b's left (0) or right (1) child."""
        o = 8 * side
        return int(nu[b, 3 + 4 * side]), nodes[b, o:o + 3], nodes[b, o + 4:o + 7]

    seen, need = [], {}
    order = [(0, root)]          # breadth-first: (4-wide node, its binary node)
    for i, b in order:
        assert i == len(need)
        kids = 0
        for h in range(2):
            ref, mn, mx = child(b, h)
            if ref & 0x80000000:
                want = [(ref, mn, mx), (NONE, None, None)]
            else:
                want = [child(ref, 0), child(ref, 1)]
            for k, (wref, wmn, wmx) in enumerate(want):
                j = 2 * h + k
                got = int(u4[i, 8 * j + 3])
                if wref == NONE:
                    assert got == NONE
                    continue
                kids += 1
                assert np.array_equal(n4[i, 8 * j:8 * j + 3], wmn)
                assert np.array_equal(n4[i, 8 * j + 4:8 * j + 7], wmx)
                if wref & 0x80000000:
                    assert got == wref
                    seen.append(wref)
                else:
                    assert got == len(order)
                    order.append((got, wref))
        need[i] = kids - 1
    for i, _ in reversed(order):
        deeper = [need[int(u4[i, 8 * j + 3])] for j in range(4)
                  if not (u4[i, 8 * j + 3] & 0x80000000)]
        need[i] += max(deeper, default=0)
    assert need[0] == stack4
    assert len(order) == n4.shape[0]
    leaves = []
    stack = [root]
    while stack:
        r = stack.pop()
        for side in range(2):
            ref = child(r, side)[0]
            (leaves if ref & 0x80000000 else stack).append(ref)
    assert sorted(seen) == sorted(leaves)
    assert usable == (1 if stack4 + 1 <= 64 else 0)


def _two_pointer(left):
    """The host partition (mesh.rs:384-395): scan from the front, swap each
    right-side element with the back."""
    order = list(range(len(left)))
    i, j = 0, len(order)
    while i < j:
        if left[order[i]]:
            i += 1
        else:
            j -= 1
            order[i], order[j] = order[j], order[i]
    return order


def _closed_form(left):
    """Destination of every element as the device build computes it from
    prefix counts (build_kernels.hip, k_ptable / k_permute)."""
    n = len(left)
    L = np.asarray(left, bool)
    nl = int(L.sum())
    lb = np.concatenate([[0], np.cumsum(L)[:-1]]).astype(int)
    c = nl - int(L[:nl].sum())
    a_r = np.zeros(n, int)
    b_l = np.zeros(n, int)
    for p in range(n):
        if not L[p] and p < nl:
            a_r[(p - lb[p])] = p
        if L[p] and p >= nl:
            b_l[nl - 1 - lb[p]] = p
    out = [None] * n
    for p in range(n):
        if L[p]:
            d = p if p < nl else a_r[nl - 1 - lb[p]]
        else:
            rb = p - lb[p]
            bprev = n if rb == 0 else b_l[rb - 1]
            bc = n if c == 0 else b_l[c - 1]
            d = bprev - 1 if rb < c else (bc - 1 if rb == c and p < bc else p - 1)
        assert out[d] is None
        out[d] = p
    return out


def test_partition_closed_form_matches_two_pointer():
    rng = np.random.default_rng(5)
    for _ in range(3000):
        n = int(rng.integers(1, 48))
        left = list(rng.random(n) < rng.random())
        assert _closed_form(left) == _two_pointer(left), left
