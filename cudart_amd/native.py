"""ctypes binding of the C ABI in include/rk.h and include/rk_scenes.h.

The library is loaded from cudart_amd/lib/librk.so (built in-tree by
cudart_amd/build.py). There is no fallback: if the library is missing the
import raises, so nothing silently runs on the CPU instead of the GPU.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB_DIR = Path(__file__).resolve().parent / "lib"
LIB_PATH = LIB_DIR / "librk.so"

RK_SUCCESS = 0
RK_MORTON_REF32, RK_MORTON_30, RK_MORTON_63 = 0, 1, 2
RK_HANDOFF_AUTO = 0xFFFFFFFE  # rk_raytrace_set_handoff: the limit chosen per call
RK_INFO_NODE_COUNT, RK_INFO_REAL_DEPTH, RK_INFO_FLAGS, RK_INFO_SPLITS, RK_INFO_LEAVES = 0, 1, 2, 3, 4
RK_INFO_WORDS = 8
RK_FLAG_WALK_PAST_END, RK_FLAG_WALK_PAST_DEPTH, RK_FLAG_MERGE2X_OOB, RK_FLAG_NODES_GT_N = 1, 2, 4, 8
RK_FLAG_SAH_NO_CUT, RK_FLAG_SPIN_LIMIT = 16, 64
RK_TRACE_RESET_DEPTH = 1


class RkError(RuntimeError):
    pass


class Float4(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("z", C.c_float), ("w", C.c_float)]


class Float2(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float)]


class AABox(C.Structure):
    _fields_ = [("min", Float4), ("max", Float4)]


class TraceOptions(C.Structure):
    _fields_ = [("det_epsilon", C.c_float), ("flags", C.c_uint32), ("config", C.c_uint32), ("reserved", C.c_uint32)]


class TraceTuning(C.Structure):
    _fields_ = [("config", C.c_uint32), ("workgroup", C.c_uint32), ("ray_run", C.c_uint32), ("bands", C.c_uint32),
                ("lds_stack", C.c_uint32), ("xcd_strips", C.c_uint32), ("ms", C.c_float),
                ("steal_oldest", C.c_uint32), ("leaf_batch", C.c_uint32), ("donate_min", C.c_uint32),
                ("bound_every", C.c_uint32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


DET_EPSILON_REFERENCE = 0.000001


def trace_options(det_epsilon: float | None, config: int = 0, reset_depth: bool = False) -> TraceOptions | None:
    """None = the reference threshold (1e-6); 0.0 = reject only parallel rays.
    config: a traversal variant (rk_fast_autotune); 0 = the default.
    reset_depth: RK_TRACE_RESET_DEPTH (every ray starts at FLT_MAX, every depth written)."""
    if det_epsilon is None and config == 0 and not reset_depth:
        return None
    eps = DET_EPSILON_REFERENCE if det_epsilon is None else float(det_epsilon)
    return TraceOptions(eps, RK_TRACE_RESET_DEPTH if reset_depth else 0, int(config), 0)


class BatchMesh(C.Structure):
    _fields_ = [("face_offset", C.c_uint32), ("face_count", C.c_uint32), ("vertex_offset", C.c_uint32),
                ("vertex_count", C.c_uint32), ("ray_offset", C.c_uint32), ("ray_count", C.c_uint32),
                ("reserved", C.c_uint32 * 2)]


BATCH_MESH_DTYPE = np.dtype([("face_offset", "<u4"), ("face_count", "<u4"), ("vertex_offset", "<u4"),
                             ("vertex_count", "<u4"), ("ray_offset", "<u4"), ("ray_count", "<u4"),
                             ("reserved", "<u4", 2)])


class BatchConfig(C.Structure):
    _fields_ = [("leaf_max", C.c_uint32), ("trace_mode", C.c_uint32), ("trace_wg", C.c_uint32),
                ("stack_depth", C.c_uint32), ("ray_order", C.c_uint32), ("donate_min", C.c_uint32),
                ("reserved", C.c_uint32 * 2)]

    def as_dict(self):
        return {"leaf_max": self.leaf_max, "trace_mode": self.trace_mode, "trace_wg": self.trace_wg,
                "stack_depth": self.stack_depth, "ray_order": self.ray_order, "donate_min": self.donate_min}


class CameraRays(C.Structure):
    _fields_ = [("origin", Float4), ("direction", Float4), ("step_x", Float4), ("step_y", Float4)]


class ServerConfig(C.Structure):
    _fields_ = [("slots", C.c_uint32), ("leaf_max", C.c_uint32), ("rebuild_every", C.c_uint32),
                ("max_faces", C.c_uint32), ("max_pixels", C.c_uint32), ("det_epsilon", C.c_float),
                ("amplitude_frac", C.c_float), ("omega", C.c_float), ("hbm_budget", C.c_uint64),
                ("plain", C.c_uint32), ("trace_config", C.c_uint32), ("reserved", C.c_uint32 * 2)]


class Request(C.Structure):
    _fields_ = [("scene", C.c_uint32), ("width", C.c_uint32), ("height", C.c_uint32), ("flags", C.c_uint32),
                ("time", C.c_float), ("reserved_f", C.c_float), ("camera", CameraRays), ("arrival", C.c_double),
                ("faces_id", C.c_void_p), ("coords", C.c_void_p), ("depths", C.c_void_p)]


class Result(C.Structure):
    _fields_ = [("checksum", C.c_uint64), ("hits", C.c_uint64), ("arrival", C.c_double), ("start", C.c_double),
                ("done", C.c_double), ("scene", C.c_uint32), ("pixels", C.c_uint32), ("rebuilt", C.c_uint32),
                ("uploaded", C.c_uint32), ("evictions", C.c_uint32), ("slot", C.c_uint32),
                ("build_flags", C.c_uint32), ("status", C.c_int32)]


class ServerStats(C.Structure):
    _fields_ = [("completed", C.c_uint64), ("rebuilds", C.c_uint64), ("refits", C.c_uint64),
                ("uploads", C.c_uint64), ("evictions", C.c_uint64), ("rays", C.c_uint64),
                ("first_arrival", C.c_double), ("last_done", C.c_double), ("p50_ms", C.c_double),
                ("p99_ms", C.c_double), ("mean_ms", C.c_double), ("max_ms", C.c_double),
                ("mrays_per_s", C.c_double), ("peak_resident_bytes", C.c_uint64), ("failed", C.c_uint64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class FastBVH(C.Structure):
    _fields_ = [("nodes", C.c_void_p), ("tris", C.c_void_p), ("indices", C.c_void_p), ("boxes", C.c_void_p),
                ("face_count", C.c_uint32), ("reserved", C.c_uint32)]


_lib = None

_P = C.c_void_p
_U32 = C.c_uint32
_I = C.c_int
_SZ = C.c_size_t

_SIGS = {
    "rk_status_string": (C.c_char_p, [_I]),
    "rk_device_count": (_I, [C.POINTER(_I)]),
    "rk_malloc": (_I, [C.POINTER(_P), _SZ]),
    "rk_free": (_I, [_P]),
    "rk_malloc_host": (_I, [C.POINTER(_P), _SZ]),
    "rk_free_host": (_I, [_P]),
    "rk_memcpy": (_I, [_P, _P, _SZ, _I, _P]),
    "rk_stream_create": (_I, [C.POINTER(_P)]),
    "rk_stream_destroy": (_I, [_P]),
    "rk_stream_synchronize": (_I, [_P]),
    "rk_fill_u32": (_I, [_P, _P, _SZ, _U32]),
    "rk_fill_f32": (_I, [_P, _P, _SZ, C.c_float]),
    "rk_fill_f2": (_I, [_P, _P, _SZ, Float2]),
    "rk_fill_f4": (_I, [_P, _P, _SZ, Float4]),
    "rk_calc_primary_rays": (_I, [_P, _I, _I, Float4, Float4, Float4, Float4, _P, _P]),
    "rk_shade": (_I, [_P, _P, _P, _I, _I, Float4, Float4, _P, _I]),
    "rk_calc_mesh_boxes": (_I, [_P, _P, _P, _U32, _P, _P]),
    "rk_reorder_workspace_size": (_SZ, [_U32]),
    "rk_reorder_mesh_faces": (_I, [_P, _P, _P, _U32, _P, _SZ]),
    "rk_sort_workspace_size": (_SZ, [_U32, _U32]),
    "rk_radix_sort_general": (_I, [_P, _P, _P, _U32, _U32, _I, _P, _SZ]),
    "rk_radix_sort": (_I, [_P, _P, _U32, _U32, _P, _SZ]),
    "rk_radix_sort_pairs_u64": (_I, [_P, _P, _P, _U32, _U32, _P, _SZ]),
    "rk_bitonic_sort": (_I, [_P, _P, _U32, _P, _SZ]),
    "rk_bitonic_sort_general": (_I, [_P, _P, _P, _U32, _P, _SZ]),
    "rk_bubble_sort512": (_I, [_P, _P, _P, _U32, _I]),
    "rk_workspace_flags": (_I, [_P, _P, C.POINTER(_U32)]),
    "rk_calc_morton_codes": (_I, [_P, _P, _U32, C.POINTER(AABox), _I, _P]),
    "rk_build_lbvh_workspace_size": (_SZ, [_U32]),
    "rk_build_lbvh": (_I, [_P, _P, _P, _U32, C.POINTER(AABox), _I, _U32, _P, _P, _P, _U32, _P, _P, _SZ]),
    "rk_build_sah_workspace_size": (_SZ, [_U32, _U32]),
    "rk_build_sah": (_I, [_P, _P, _P, _U32, C.POINTER(AABox), _P, _P, _P, _U32, _U32, _P, _P, _SZ]),
    "rk_raytrace_mesh_tree": (_I, [_P, _P, _U32, _P, _U32, _P, _P, _P, _I, _I, _P, _P, _P, _I]),
    "rk_raytrace_set_handoff": (_I, [_U32, _U32, _U32]),
    "rk_raytrace_get_handoff": (_I, [C.POINTER(_U32), C.POINTER(_U32), C.POINTER(_U32)]),
    "rk_fast_build_workspace_size": (_SZ, [_U32, _I]),
    "rk_fast_build": (_I, [_P, _P, _P, _U32, C.POINTER(AABox), _I, _U32, C.POINTER(FastBVH), _P, _P, _SZ]),
    "rk_fast_build_noalloc": (_I, [_P, _P, _P, _U32, C.POINTER(AABox), _I, _U32, C.POINTER(FastBVH), _P]),
    "rk_fast_trace": (_I, [_P, C.POINTER(FastBVH), _P, _P, _U32, _P, _P, _P, _P]),
    "rk_fast_trace_counted": (_I, [_P, C.POINTER(FastBVH), _P, _P, _U32, _P, _P, _P, _P, _P]),
    "rk_fast_trace_sorted_workspace_size": (_SZ, [_U32]),
    "rk_fast_trace_sorted": (_I, [_P, C.POINTER(FastBVH), _P, _P, _U32, _P, _P, _P, _P, _P, _SZ]),
    "rk_fast_build_keyed": (_I, [_P, _P, _P, _U32, _P, _I, _U32, C.POINTER(FastBVH), _P, _P, _P, _SZ]),
    "rk_fast_refit": (_I, [_P, _P, _P, _U32, _I, _U32, C.POINTER(FastBVH), _P, _P, _SZ]),
    "rk_fast_trace_camera": (_I, [_P, C.POINTER(FastBVH), C.POINTER(CameraRays), _I, _I, _P, _P, _P, _P]),
    "rk_fast_render_camera": (_I, [_P, C.POINTER(FastBVH), C.POINTER(CameraRays), _I, _I, _P, _P, _P, Float4, Float4,
                                   _P, _I, _P, _P, _P, _P]),
    "rk_fast_trace_variant": (_I, [_U32, _I, C.POINTER(TraceTuning)]),
    "rk_fast_autotune_workspace_size": (_SZ, [_U32]),
    "rk_fast_autotune": (_I, [_P, C.POINTER(FastBVH), _P, _P, _U32, _P, _P, _P, _P, _I, _P, _SZ,
                              C.POINTER(TraceTuning)]),
    "rk_fast_autotune_camera": (_I, [_P, C.POINTER(FastBVH), C.POINTER(CameraRays), _I, _I, _P, _P, _P, _P, _I,
                                     C.POINTER(TraceTuning)]),
    "rk_deform_vertices": (_I, [_P, _P, _P, _U32, C.c_float, C.c_float, C.c_float, C.c_uint64, _I]),
    "rk_frame_checksum": (_I, [_P, _P, _P, _P, _U32, _P, _U32, _P]),
    # serving
    "rk_server_config_default": (ServerConfig, []),
    "rk_server_create": (_I, [C.POINTER(ServerConfig), C.c_uint64, C.POINTER(_P)]),
    "rk_server_destroy": (_I, [_P]),
    "rk_server_scene_bytes": (_SZ, [_U32, _U32]),
    "rk_server_add_scene": (_I, [_P, _P, _U32, _P, _U32, _I, C.POINTER(_U32)]),
    "rk_server_now": (C.c_double, [_P]),
    "rk_server_submit": (_I, [_P, C.POINTER(Request), C.POINTER(C.c_uint64)]),
    "rk_server_progress": (_I, [_P, C.POINTER(_U32)]),
    "rk_server_wait_all": (_I, [_P]),
    "rk_server_result": (_I, [_P, C.c_uint64, C.POINTER(Result)]),
    "rk_server_stats_get": (_I, [_P, C.c_uint64, C.c_uint64, C.POINTER(ServerStats)]),
    "rk_server_run_poisson": (_I, [_P, C.c_double, _U32, _U32, _U32, C.c_uint64, C.POINTER(ServerStats),
                                   C.POINTER(C.c_uint64)]),
    "rk_interpolate_data": (_I, [_P, _P, _P, _I, _I, _P, _U32, _P, _P, _P, _P, _P, _P, _P, _P, _P]),
    # batches, graphs, autotuning
    "rk_batch_config_default": (BatchConfig, []),
    "rk_batch_build_workspace_size": (_SZ, [_U32, _U32]),
    "rk_batch_build": (_I, [_P, _P, _U32, _P, _P, _U32, _P, C.POINTER(BatchConfig), C.POINTER(FastBVH), _P, _P,
                            _SZ]),
    "rk_batch_trace_workspace_size": (_SZ, [_U32, _U32]),
    "rk_batch_trace": (_I, [_P, _P, _U32, C.POINTER(FastBVH), _P, _P, _U32, _P, _P, _P, _P,
                            C.POINTER(BatchConfig), _P, _SZ]),
    "rk_batch_autotune": (_I, [_P, _P, _U32, _P, _P, _U32, _P, C.POINTER(FastBVH), _P, _P, _P, _U32, _P, _P, _P,
                               _P, _U32, C.POINTER(BatchConfig), C.POINTER(C.c_float), _P, _SZ]),
    "rk_graph_capture_begin": (_I, [_P]),
    "rk_graph_capture_end": (_I, [_P, C.POINTER(_P)]),
    "rk_graph_launch": (_I, [_P, _P]),
    "rk_graph_destroy": (_I, [_P]),
    # scenes
    "rk_scene_icosphere": (_I, [_U32, _P, _P, C.POINTER(_U32), C.POINTER(_U32)]),
    "rk_scene_heightfield": (_I, [_U32, C.c_uint64, C.c_float, _P, _P, C.POINTER(_U32), C.POINTER(_U32)]),
    "rk_scene_soup": (_I, [_U32, C.c_uint64, C.c_float, C.c_float, C.c_float, C.c_float, _P, _P,
                           C.POINTER(_U32), C.POINTER(_U32)]),
    "rk_scene_hitbox": (_I, [_P, _U32, C.POINTER(AABox)]),
    "rk_camera_lookat": (_I, [Float4, Float4, C.c_float, C.c_float, _I, _I, _P]),
    "rk_camera_lookat_up": (_I, [Float4, Float4, Float4, C.c_float, C.c_float, _I, _I, _P]),
    "rk_primary_rays_host": (_I, [_P, _I, _I, _P, _P]),
    "rk_rays_random": (_I, [_U32, C.c_uint64, C.c_float, C.c_float, _P, _P]),
}


def lib() -> C.CDLL:
    """Load librk.so once and attach the signatures. Raises if it is not built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RkError(f"{LIB_PATH} is not built; run `python -m cudart_amd.build` (or __graft_entry__.build())")
        L = C.CDLL(str(LIB_PATH))
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def declared_symbols() -> list[str]:
    return list(_SIGS)


def check(status: int, what: str = "") -> None:
    if status != RK_SUCCESS:
        msg = lib().rk_status_string(status).decode()
        raise RkError(f"{what}: {msg} ({status})")


def ptr(a) -> int:
    """Device or host address of a torch tensor or numpy array."""
    if isinstance(a, np.ndarray):
        return a.ctypes.data
    return a.data_ptr()


def f4(v) -> Float4:
    v = list(v) + [0.0] * (4 - len(v))
    return Float4(*[float(x) for x in v[:4]])


def aabox(b: np.ndarray) -> AABox:
    """np.float32 array (2,4) -> AABox."""
    b = np.asarray(b, dtype=np.float32).reshape(2, 4)
    return AABox(f4(b[0]), f4(b[1]))


# ---- host scene generators (numpy) ----------------------------------------

def icosphere(levels: int):
    L = lib()
    nv, nf = _U32(), _U32()
    check(L.rk_scene_icosphere(levels, None, None, C.byref(nv), C.byref(nf)), "icosphere")
    V = np.zeros((nv.value, 4), np.float32)
    F = np.zeros((nf.value, 4), np.uint32)
    check(L.rk_scene_icosphere(levels, V.ctypes.data, F.ctypes.data, C.byref(nv), C.byref(nf)), "icosphere")
    return V, F


def heightfield(grid: int, seed: int = 0xC2, amplitude: float = 0.2):
    L = lib()
    nv, nf = _U32(), _U32()
    check(L.rk_scene_heightfield(grid, seed, amplitude, None, None, C.byref(nv), C.byref(nf)), "heightfield")
    V = np.zeros((nv.value, 4), np.float32)
    F = np.zeros((nf.value, 4), np.uint32)
    check(L.rk_scene_heightfield(grid, seed, amplitude, V.ctypes.data, F.ctypes.data, C.byref(nv), C.byref(nf)),
          "heightfield")
    return V, F


def soup(n: int, seed: int = 0xC3, mu: float = float(np.log(0.002)), sigma: float = 0.5, lo: float = 0.0,
         hi: float = 1.0):
    L = lib()
    V = np.zeros((3 * n, 4), np.float32)
    F = np.zeros((n, 4), np.uint32)
    nv, nf = _U32(), _U32()
    check(L.rk_scene_soup(n, seed, mu, sigma, lo, hi, V.ctypes.data, F.ctypes.data, C.byref(nv), C.byref(nf)),
          "soup")
    return V, F


def hitbox(V: np.ndarray) -> np.ndarray:
    b = AABox()
    check(lib().rk_scene_hitbox(V.ctypes.data, V.shape[0], C.byref(b)), "hitbox")
    return np.array([[b.min.x, b.min.y, b.min.z, b.min.w], [b.max.x, b.max.y, b.max.z, b.max.w]], np.float32)


def camera_lookat(eye, target, fov: float, width: int, height: int, up=None) -> np.ndarray:
    out = np.zeros((4, 4), np.float32)
    ratio = float(width) / float(height)
    if up is None:
        check(lib().rk_camera_lookat(f4(eye), f4(target), fov, ratio, width, height, out.ctypes.data), "camera")
    else:
        check(lib().rk_camera_lookat_up(f4(eye), f4(target), f4(up), fov, ratio, width, height, out.ctypes.data),
              "camera")
    return out


def primary_rays_host(cam: np.ndarray, width: int, height: int):
    O = np.zeros((width * height, 4), np.float32)
    D = np.zeros((width * height, 4), np.float32)
    cam = np.ascontiguousarray(cam, np.float32)
    check(lib().rk_primary_rays_host(cam.ctypes.data, width, height, O.ctypes.data, D.ctypes.data), "rays")
    return O, D


def random_rays(n: int, seed: int = 0xC3, lo: float = 0.0, hi: float = 1.0):
    O = np.zeros((n, 4), np.float32)
    D = np.zeros((n, 4), np.float32)
    check(lib().rk_rays_random(n, seed, lo, hi, O.ctypes.data, D.ctypes.data), "random rays")
    return O, D
