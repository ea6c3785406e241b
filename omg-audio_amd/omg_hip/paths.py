"""Knife-edge path search over the C ABI: AutoPaths
(crates/omg-core/src/paths.rs:62-332) with the reference's method names.

`AutoPaths(scene, edge_budget)` extracts the budgeted edge graph once
(paths.rs:70-77); `find(src, lis, budget)` returns the direct path first,
then the strongest bent paths (paths.rs:117-332); `find_many` runs many
(source, listener) queries in one launch; `transmission(a, b)` prices a
straight segment (paths.rs:84-113).
"""
import ctypes as C
from dataclasses import dataclass

import numpy as np

from ._lib import (PATH_MAX_POINTS, Counters, Path, PathBudget as _Budget, PathEdge, check,
                   lib)


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _f32(a):
    return np.ascontiguousarray(np.asarray(a, np.float32))


@dataclass
class PathBudget:
    """PathBudget (paths.rs:20-45); the defaults are the reference's."""
    edge_candidates: int = 24
    pair_edges: int = 8
    max_paths: int = 4
    taps: bool = False

    def c(self):
        return _Budget(self.edge_candidates, self.pair_edges, self.max_paths,
                       1 if self.taps else 0)


def _path_dicts(raw, counts):
    """omg_path records [n][stride] -> per query a list of dicts (key,
    points, length, gains), the direct path first (FoundPath, paths.rs:48-60)."""
    out = []
    for q in range(raw.shape[0]):
        paths = []
        for k in range(int(counts[q])):
            r = raw[q, k]
            n = int(r["n_points"])
            paths.append(dict(key=int(r["key"]), points=r["points"][:n].copy(),
                              length=np.float32(r["length"]), gains=r["gains"].copy()))
        out.append(paths)
    return out


PATH_DTYPE = np.dtype([("key", np.uint32), ("n_points", np.uint32), ("length", np.float32),
                       ("gains", np.float32, 3), ("points", np.float32, (PATH_MAX_POINTS, 3))])
assert PATH_DTYPE.itemsize == C.sizeof(Path)
EDGE_DTYPE = np.dtype([("a", np.float32, 3), ("importance", np.float32), ("b", np.float32, 3),
                       ("reserved", np.uint32)])
assert EDGE_DTYPE.itemsize == C.sizeof(PathEdge)


def _edge_rows(e):
    """[n][7] rows a, b, importance (the oracle's layout)."""
    return np.hstack([e["a"], e["b"], e["importance"][:, None]]).astype(np.float32)


def extract_edges(positions, indices, budget, min_dihedral_deg=40.0):
    """extract_edges + budget_edges (mesh.rs:639-682) on the host: [n][7]
    rows a, b, importance, importance-descending."""
    pos = _f32(positions).reshape(-1, 3)
    idx = np.ascontiguousarray(np.asarray(indices, np.uint32).reshape(-1, 3))
    n = C.c_uint32()
    cap = max(1, 3 * idx.shape[0])
    out = np.zeros(cap, EDGE_DTYPE)
    check(lib().omg_paths_extract_edges(_p(pos), pos.shape[0], _p(idx), idx.shape[0],
                                        min_dihedral_deg, budget, _p(out), cap, C.byref(n)))
    return _edge_rows(out[:n.value])


def knife_edge_amp(detour_m, freq_hz):
    """knife_edge_amp (diffraction.rs:27-42), host evaluation of the shared code."""
    return lib().omg_knife_edge_amp(detour_m, freq_hz)


def bend_point(a, b, src, lis):
    """bend_point (paths.rs:411-436), host evaluation of the shared code."""
    p = np.zeros(3, np.float32)
    lib().omg_bend_point(_p(_f32(a)), _p(_f32(b)), _p(_f32(src)), _p(_f32(lis)), _p(p))
    return p


class AutoPaths:
    """AutoPaths (paths.rs:62-77) on a 3-band scene's device BVH."""

    def __init__(self, scene, edge_budget=64):
        self.scene = scene
        h = C.c_void_p()
        check(lib().omg_paths_create(scene.h, edge_budget, C.byref(h)))
        self.h = h

    def edge_count(self):
        n = C.c_uint32()
        check(lib().omg_paths_edges(self.h, None, 0, C.byref(n)))
        return n.value

    def edges(self):
        """The device edge graph as [n][7] rows a, b, importance."""
        n = self.edge_count()
        out = np.zeros(max(1, n), EDGE_DTYPE)
        m = C.c_uint32()
        check(lib().omg_paths_edges(self.h, _p(out), max(1, n), C.byref(m)))
        return _edge_rows(out[:n])

    def transmission(self, a, b):
        """AutoPaths::transmission (paths.rs:84-113); a, b: [3] or [n][3]."""
        a2 = _f32(a).reshape(-1, 3)
        b2 = _f32(b).reshape(-1, 3)
        g = np.zeros((a2.shape[0], 3), np.float32)
        check(lib().omg_paths_transmission(self.h, _p(a2), _p(b2), a2.shape[0], _p(g)))
        return g[0] if np.asarray(a).ndim == 1 else g

    def find_many(self, src, lis, budget=None, raw=False):
        """AutoPaths::find for every row of src / lis ([n][3]): per query a
        list of paths (raw=True: the omg_path records and counts)."""
        budget = budget or PathBudget()
        s = _f32(src).reshape(-1, 3)
        l = _f32(lis).reshape(-1, 3)
        n = s.shape[0]
        stride = 1 + budget.max_paths
        out = np.zeros((max(1, n), stride), PATH_DTYPE)
        cnt = np.zeros(max(1, n), np.uint32)
        b = budget.c()
        check(lib().omg_paths_find(self.h, C.byref(b), _p(s), _p(l), n, _p(out), _p(cnt)))
        if raw:
            return out[:n], cnt[:n]
        return _path_dicts(out[:n], cnt[:n])

    def find_async(self, d_src, d_lis, n, d_out, d_n_paths, d_work, budget=None, stream=None):
        """omg_paths_find_async: device pointers (ints), enqueued on `stream`
        (a HIP stream handle, None = the context's) without waiting."""
        b = (budget or PathBudget()).c()
        check(lib().omg_paths_find_async(self.h, C.byref(b), d_src, d_lis, n, d_out, d_n_paths,
                                         d_work, stream))

    def status(self):
        """Waits for queued work; raises OmgError(ERR_OVERFLOW) after an overflow."""
        check(lib().omg_paths_status(self.h))

    def find(self, src, lis, budget=None):
        """AutoPaths::find (paths.rs:117-332) for one query."""
        return self.find_many([src], [lis], budget)[0]

    def set_counting(self, on=True):
        check(lib().omg_paths_set_counting(self.h, 1 if on else 0))

    def counters(self):
        c = Counters()
        check(lib().omg_paths_counters(self.h, C.byref(c)))
        return dict(aabb=c.aabb_tests, tri=c.tri_tests)

    def close(self):
        if self.h:
            lib().omg_paths_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
