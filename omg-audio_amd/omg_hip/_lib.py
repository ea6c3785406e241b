"""ctypes binding of libomg_hip.so (the C ABI declared in include/omg_hip.h).

This is the binding a host language adds over the C ABI; the Rust
equivalent is shown in INTEGRATION.md.  Loading fails loudly when the
library has not been built: there is no fallback path.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
LIB_PATH = os.path.join(PKG, "libomg_hip.so")

LAYOUT_VERSION = 1
MAX_BANDS = 8
MAX_OVERLAYS = 64
MAX_LISTENERS = 64
DOME_BINS = 9

OK = 0
ERR_NO_DEVICE = -1
ERR_INVALID = -2
ERR_CAP = -3
ERR_VERSION = -4
ERR_DEVICE = -5
ERR_BUDGET = -6
ERR_NOT_READY = -7
ERR_OVERFLOW = -8

F8 = C.c_float * MAX_BANDS


class Material(C.Structure):
    _fields_ = [("absorption", F8), ("scattering", C.c_float), ("transmission", F8)]


class Overlay(C.Structure):
    _fields_ = [("min", C.c_float * 3), ("max", C.c_float * 3), ("mat", Material)]


class TraceParams(C.Structure):
    _fields_ = [("n_rays", C.c_uint32), ("nbands", C.c_uint32), ("max_bounces", C.c_uint32),
                ("dir_band", C.c_uint32), ("seed", C.c_uint64), ("bin_dt", C.c_float),
                ("max_time", C.c_float), ("floor_rel", C.c_float), ("radius", C.c_float),
                ("energy", F8)]


class RayRecord(C.Structure):
    _fields_ = [("iters", C.c_uint32), ("term", C.c_uint32), ("hit_hash", C.c_uint32),
                ("flags", C.c_uint32)]


class Counters(C.Structure):
    _fields_ = [("aabb_tests", C.c_uint64), ("tri_tests", C.c_uint64),
                ("bounce_rays", C.c_uint64)]


PATH_MAX_POINTS = 25
PATH_MAX_EDGES = 4096
PATH_MAX_CANDIDATES = 256
PATH_MAX_PAIR_EDGES = 64


class PathEdge(C.Structure):
    _fields_ = [("a", C.c_float * 3), ("importance", C.c_float), ("b", C.c_float * 3),
                ("reserved", C.c_uint32)]


class PathBudget(C.Structure):
    _fields_ = [("edge_candidates", C.c_uint32), ("pair_edges", C.c_uint32),
                ("max_paths", C.c_uint32), ("taps", C.c_uint32)]


class Path(C.Structure):
    _fields_ = [("key", C.c_uint32), ("n_points", C.c_uint32), ("length", C.c_float),
                ("gains", C.c_float * 3), ("points", C.c_float * (3 * PATH_MAX_POINTS))]


class SceneInfo(C.Structure):
    _fields_ = [("n_tris", C.c_uint32), ("n_prims", C.c_uint32), ("n_nodes", C.c_uint32),
                ("max_depth", C.c_uint32), ("device_bytes", C.c_uint64)]


class SceneTree(C.Structure):
    _fields_ = [("n_prims", C.c_uint32), ("n_nodes", C.c_uint32), ("n_nodes4", C.c_uint32),
                ("root_ref", C.c_uint32), ("root4", C.c_uint32), ("stack4", C.c_uint32),
                ("has4", C.c_uint32), ("max_depth", C.c_uint32), ("empty", C.c_uint32),
                ("pad", C.c_uint32), ("root_min", C.c_float * 3), ("root_max", C.c_float * 3),
                ("n_nodes16", C.c_uint32), ("stack16", C.c_uint32), ("has16", C.c_uint32),
                ("pad16", C.c_uint32)]


class DomePanel(C.Structure):
    _fields_ = [("axis", C.c_uint32), ("plane", C.c_float), ("lat0", C.c_float),
                ("lat1", C.c_float), ("z0", C.c_float), ("z1", C.c_float),
                ("glass", C.c_uint32), ("trans", F8)]


class DomeParams(C.Structure):
    _fields_ = [("n_fan", C.c_uint32), ("n_rays", C.c_uint32), ("max_events", C.c_uint32),
                ("nbands", C.c_uint32), ("horizon", C.c_float), ("through_len", C.c_float)]


class DomeRay(C.Structure):
    _fields_ = [("escaped", C.c_uint32), ("events", C.c_uint32), ("used_through", C.c_uint32),
                ("bin", C.c_uint32), ("tp", F8), ("dir", C.c_float * 3)]


class Surface(C.Structure):
    _fields_ = [("n", C.c_float * 3), ("d", C.c_float), ("refl", F8), ("trans", F8),
                ("has_rect", C.c_uint32), ("rmin", C.c_float * 3), ("rmax", C.c_float * 3)]


class Blocker(C.Structure):
    _fields_ = [("min", C.c_float * 3), ("max", C.c_float * 3), ("trans", F8)]


class MeshRecord(C.Structure):
    _fields_ = [("valid", C.c_uint32), ("delay", C.c_float), ("dir", C.c_float * 3),
                ("gains", F8)]


class Query(C.Structure):
    _fields_ = [("scene", C.c_void_p), ("params", TraceParams), ("sources", C.c_void_p),
                ("n_src", C.c_uint32), ("listeners", C.c_void_p), ("n_lis", C.c_uint32),
                ("overlays", C.c_void_p), ("n_ovl", C.c_uint32)]


class BatchShape(C.Structure):
    _fields_ = [("lanes", C.c_uint32), ("stack_lds", C.c_uint32), ("waves_per_simd", C.c_uint32),
                ("regen", C.c_uint32), ("width", C.c_uint32), ("ms", C.c_float)]


BATCH_RECORDS = 1
BATCH_COUNTERS = 2


class Request(C.Structure):
    _fields_ = [("scene", C.c_void_p), ("params", TraceParams), ("sources", C.c_void_p),
                ("n_src", C.c_uint32), ("listeners", C.c_void_p), ("n_lis", C.c_uint32),
                ("dome", C.c_uint32), ("flags", C.c_uint32)]


REQ_KEEP = 1
REQ_RECORDS = 2


class ServerConfig(C.Structure):
    _fields_ = [("hbm_budget", C.c_uint64), ("slots", C.c_uint32), ("max_batch", C.c_uint32),
                ("latency_rays", C.c_uint32), ("shape", BatchShape), ("dome", DomeParams)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_uint64), ("completed", C.c_uint64), ("refused", C.c_uint64),
                ("batches", C.c_uint64), ("committed_bytes", C.c_uint64),
                ("peak_committed_bytes", C.c_uint64), ("rays", C.c_uint64),
                ("mean_batch", C.c_double), ("max_in_flight", C.c_uint32)]


class RequestStats(C.Structure):
    _fields_ = [("queue_s", C.c_double), ("latency_s", C.c_double), ("batch_size", C.c_uint32),
                ("slot", C.c_uint32)]


class TraceGateState(C.Structure):
    _fields_ = [("last_src", C.c_float * 3), ("last_lis", C.c_float * 3),
                ("last_energy", C.c_float * MAX_BANDS), ("age", C.c_uint32)]


class LateUpdate(C.Structure):
    _fields_ = [("slot", C.c_uint32), ("traced", C.c_uint32), ("cur_bins", C.c_void_p),
                ("cur_dirs", C.c_void_p), ("src", C.c_float * 3), ("receiver", C.c_float * 3)]


class LateTrace(C.Structure):
    _fields_ = [("slot", C.c_uint32), ("params", TraceParams), ("src", C.c_float * 3),
                ("receiver", C.c_float * 3)]


class LateOut(C.Structure):
    _fields_ = [("rt60", C.c_float * MAX_BANDS), ("level", C.c_float * MAX_BANDS),
                ("dir", C.c_float * 3), ("aniso", C.c_float), ("version", C.c_uint32),
                ("traced", C.c_uint32)]


WIRE_BOX_JOB_F32S = 57
WIRE_BOX_JOB_VERSION = 2
WIRE_WORLD_MAX_PANELS = 64
WIRE_WORLD_JOB_F32S = 10 + WIRE_WORLD_MAX_PANELS * 16
WIRE_WORLD_JOB_VERSION = 3
WIRE_ECHO_BINS = 300
WIRE_ECHO_F32S = WIRE_ECHO_BINS * 6
WIRE_PARAM_HEADER = 24
WIRE_PARAM_PER_TAP = 9


class WireWorldJob(C.Structure):
    _fields_ = [("id", C.c_uint32), ("n_rays", C.c_uint32), ("seed", C.c_uint32),
                ("src", C.c_float * 3), ("lis", C.c_float * 3), ("n_panels", C.c_uint32),
                ("panels", Overlay * WIRE_WORLD_MAX_PANELS)]


class Tap(C.Structure):
    _fields_ = [("key", C.c_uint32), ("delay_s", C.c_float), ("dir", C.c_float * 3),
                ("gains", C.c_float * 3)]


class ParamBlock(C.Structure):
    _fields_ = [("version", C.c_uint64), ("rt60", C.c_float * 3), ("level", C.c_float * 3),
                ("has_remote", C.c_uint32), ("remote_rt60", C.c_float * 3),
                ("remote_send", C.c_float * 3), ("remote_sh", C.c_float * 9)]


# (name, restype, argtypes) of every exported entry point
_vp, _u32, _f32, _i = C.c_void_p, C.c_uint32, C.c_float, C.c_int
SIGNATURES = [
    ("omg_init", _i, [_i, _u32, C.POINTER(_vp)]),
    ("omg_shutdown", None, [_vp]),
    ("omg_last_error", C.c_char_p, []),
    ("omg_device_name", _i, [_vp, C.c_char_p, _u32]),
    ("omg_set_launch_config", _i, [_vp, _u32, _u32, _u32, _u32]),
    ("omg_scene_create", _i, [_vp, _vp, _u32, _vp, _u32, _vp, _vp, _vp, _vp, _u32, _u32,
                              C.POINTER(_vp)]),
    ("omg_scene_create_soup", _i, [_vp, _vp, _u32, _vp, _vp, _vp, _vp, _u32, _u32,
                                   C.POINTER(_vp)]),
    ("omg_scene_info_get", _i, [_vp, _vp]),
    ("omg_dedup_soup", _i, [_vp, _u32, _vp, _vp, C.POINTER(_u32)]),
    ("omg_bvh_build_host", _i, [_vp, _u32, _vp, _u32, _vp, _vp, _vp, _vp, _u32,
                                C.POINTER(_u32), _vp, _u32, C.POINTER(_u32), C.POINTER(_u32),
                                C.POINTER(_u32)]),
    ("omg_bvh4_build_host", _i, [_vp, _u32, _vp, _u32, _vp, _vp, _vp, _vp, _u32,
                                 C.POINTER(_u32), C.POINTER(_u32), C.POINTER(_u32),
                                 C.POINTER(_u32)]),
    ("omg_scene_create_device", _i, [_vp, _vp, _u32, _vp, _u32, _vp, _vp, _vp, _vp, _u32, _u32,
                                     C.POINTER(_vp)]),
    ("omg_scene_export", _i, [_vp, _vp, _u32, _vp, _u32, _vp, _u32, _vp]),
    ("omg_scene_refit", _i, [_vp, _vp, _u32]),
    ("omg_scene_destroy", None, [_vp]),
    ("omg_trace_world", _i, [_vp, _vp, _vp, _u32, _vp, _u32, _vp, _u32, _vp, _vp, _vp, _vp]),
    ("omg_trace_world_hits", _i, [_vp, _vp, _vp, _u32, _vp, _u32, _vp, _u32, _vp, _vp, _vp, _vp,
                                  _vp]),
    ("omg_trace_box", _i, [_vp, _vp, _vp, _u32, _vp, _vp, _u32, _vp, _u32, _vp, _vp, _vp, _vp]),
    ("omg_trace_world_async", _i, [_vp, _vp, _vp, _u32, _vp, _u32, _vp, _u32, _vp, _vp, _vp,
                                   _vp]),
    ("omg_trace_world_async_noalloc", _i, [_vp, _vp, _vp, _u32, _vp, _u32, _vp, _u32, _vp, _vp,
                                           _vp, _vp]),
    ("omg_batch_create", _i, [_vp, _vp, _u32, _u32, C.POINTER(_vp)]),
    ("omg_batch_launch", _i, [_vp, _vp]),
    ("omg_batch_sync", _i, [_vp]),
    ("omg_batch_result", _i, [_vp, _u32, _vp, _vp, _vp]),
    ("omg_batch_device_output", _i, [_vp, _u32, C.POINTER(_vp), C.POINTER(_vp)]),
    ("omg_batch_counters", _i, [_vp, _vp]),
    ("omg_batch_size", _i, [_vp, C.POINTER(C.c_uint64), C.POINTER(_u32)]),
    ("omg_batch_get_shape", _i, [_vp, _vp]),
    ("omg_batch_set_shape", _i, [_vp, _vp]),
    ("omg_batch_tune", _i, [_vp, _u32, _vp]),
    ("omg_batch_destroy", None, [_vp]),
    ("omg_server_create", _i, [_vp, _vp, _vp, _u32, C.POINTER(_vp)]),
    ("omg_server_submit", _i, [_vp, _vp, C.POINTER(C.c_uint64)]),
    ("omg_server_result", _i, [_vp, C.c_uint64, _vp, _vp, _vp, _vp, _vp, _vp]),
    ("omg_server_wait", _i, [_vp, C.c_uint64, _vp, _vp, _vp, _vp, _vp, _vp]),
    ("omg_server_replay", _i, [_vp, _vp, _vp, _u32, _vp, _vp, _vp]),
    ("omg_server_stats_get", _i, [_vp, _vp]),
    ("omg_server_destroy", None, [_vp]),
    ("omg_trace_gate_init", None, [_vp, _u32]),
    ("omg_trace_gate_step", _i, [_vp, _vp, _vp, _vp, _u32, _u32, C.POINTER(_i)]),
    ("omg_late_field_create", _i, [_vp, _u32, _u32, _u32, _f32, C.POINTER(_vp)]),
    ("omg_late_field_gate", _i, [_vp, _u32, _vp, _vp, _vp, _u32, C.POINTER(_i)]),
    ("omg_late_field_update", _i, [_vp, _vp, _u32, _vp, _vp]),
    ("omg_late_field_trace_world", _i, [_vp, _vp, _vp, _u32, _vp, _u32, _vp]),
    ("omg_late_field_read", _i, [_vp, _u32, _vp, _vp]),
    ("omg_late_field_destroy", None, [_vp]),
    ("omg_wire_decode_world_job", _i, [_vp, _u32, _vp]),
    ("omg_wire_trace_box_jobs", _i, [_vp, _vp, _u32, _u32, _vp, _vp, C.POINTER(_u32)]),
    ("omg_wire_trace_world_jobs", _i, [_vp, _vp, _u32, _u32, _vp, _vp, C.POINTER(_u32)]),
    ("omg_param_block_write_flat", _i, [_vp, _vp, _u32, _vp, _u32, C.POINTER(_u32)]),
    ("omg_param_block_read_flat", _i, [_vp, _u32, _vp, _vp, _u32, C.POINTER(_u32)]),
    ("omg_estimate_reverb", _i, [_vp, _u32, _u32, _f32, _vp, _vp]),
    ("omg_late_direction", _i, [_vp, _vp, _u32, _u32, _f32, _f32, _vp, _vp]),
    ("omg_dome_sample", _i, [_vp, _vp, _vp, _u32, _vp, _u32, _vp, _vp, _vp, _vp]),
    ("omg_dome_sample_async", _i, [_vp, _vp, _vp, _u32, _vp, _u32, _vp, _vp, _vp, _vp]),
    ("omg_dome_status", _i, [_vp]),
    ("omg_dome_probe_create", _i, [_vp, _vp, C.POINTER(_vp)]),
    ("omg_dome_probe_sample", _i, [_vp, _vp, _vp, _u32, _vp]),
    ("omg_dome_probe_destroy", None, [_vp]),
    ("omg_ism_box", _i, [_vp, _vp, _u32, _vp, _vp, _u32, _vp, _u32, C.POINTER(_u32)]),
    ("omg_surface_table", _i, [_vp, _vp, _u32, C.POINTER(_u32)]),
    ("omg_surface_append_box", _i, [_vp, _vp, _vp, _u32, _vp]),
    ("omg_discover", _i, [_vp, _vp, _u32, C.c_uint16, _vp, _u32, _u32, _vp, _vp, _u32,
                          C.POINTER(_u32)]),
    ("omg_solve_batch", _i, [_vp, _vp, _u32, _vp, _u32, _vp, _vp, _u32, _vp, _vp, _u32, _vp,
                             _vp]),
    ("omg_set_route_rays", _i, [_vp, C.c_uint64]),
    ("omg_set_handoff", _i, [_vp, C.c_uint32, C.c_uint32, C.c_uint32]),
    ("omg_set_handoff_lanes", _i, [_vp, C.c_uint32]),
    ("omg_paths_extract_edges", _i, [_vp, _u32, _vp, _u32, _f32, _u32, _vp, _u32,
                                     C.POINTER(_u32)]),
    ("omg_paths_create", _i, [_vp, _u32, C.POINTER(_vp)]),
    ("omg_paths_edges", _i, [_vp, _vp, _u32, C.POINTER(_u32)]),
    ("omg_paths_transmission", _i, [_vp, _vp, _vp, _u32, _vp]),
    ("omg_paths_find", _i, [_vp, _vp, _vp, _vp, _u32, _vp, _vp]),
    ("omg_paths_find_async", _i, [_vp, _vp, _vp, _vp, _u32, _vp, _vp, _vp, _vp]),
    ("omg_paths_status", _i, [_vp]),
    ("omg_paths_set_counting", _i, [_vp, _u32]),
    ("omg_paths_counters", _i, [_vp, _vp]),
    ("omg_knife_edge_amp", _f32, [_f32, _f32]),
    ("omg_bend_point", None, [_vp, _vp, _vp, _vp, _vp]),
    ("omg_paths_destroy", None, [_vp]),
]

_LIB = None


class OmgError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"omg_hip error {code}: {msg}")
        self.code = code


def hip_runtimes():
    """The distinct HIP runtime libraries (libamdhip64) mapped into this process."""
    try:
        with open("/proc/self/maps") as f:
            paths = {line.split()[-1] for line in f if "libamdhip64" in line}
    except OSError:
        return []
    return sorted({os.path.realpath(p) for p in paths})


def check_single_runtime():
    """Raises when two HIP runtimes are mapped.  PyTorch wheels carry their
    own libamdhip64: imported after this library has loaded ROCm's, torch
    maps a second runtime, whose device pointers and streams mean nothing
    to this library's.  Import torch before the first use of omg_hip (the
    library then binds to torch's runtime), or do not import torch."""
    rts = hip_runtimes()
    if len(rts) > 1:
        raise RuntimeError("two HIP runtimes are mapped in this process (" + ", ".join(rts) +
                           "): import torch before omg_hip loads libomg_hip.so")


def lib():
    """Load libomg_hip.so; raises if it has not been built, or if loading it
    would leave two HIP runtimes in the process."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built: run omg-audio_amd/build.py")
        L = C.CDLL(LIB_PATH)
        check_single_runtime()
        for name, res, args in SIGNATURES:
            if not hasattr(L, name):
                continue
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def check(rc):
    if rc != OK:
        msg = lib().omg_last_error()
        raise OmgError(rc, msg.decode() if msg else "")
    return rc
