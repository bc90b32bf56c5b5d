"""ORACLE — test infrastructure only (see oracle.cpp).

numpy wrappers around oracle/_build/liboracle.so, the CPU restatement of the
reference build and trace. Imported only by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker or the timed CPU baseline.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
SO = HERE / "_build" / "liboracle.so"

FLAG_WALK_PAST_END, FLAG_WALK_PAST_DEPTH, FLAG_MERGE2X_OOB, FLAG_NODES_GT_N = 1, 2, 4, 8

_lib = None
_P = C.c_void_p
_U = C.c_uint32
_I = C.c_int

_SIGS = {
    "ora_calc_boxes": (None, [_P, _P, _U, _P, _P]),
    "ora_morton32": (None, [_P, _U, _P, _P, _P]),
    "ora_morton_wide": (None, [_P, _U, _P, _P, _I, _P]),
    "ora_sort_pairs_ref": (_I, [_P, _P, _U, _U, _I]),
    "ora_sort_pairs_u64": (None, [_P, _P, _U]),
    "ora_bitonic_sort_general": (_I, [_P, _P, _U]),
    "ora_bitonic_sort": (_I, [_P, _U]),
    "ora_bubble_sort512": (_I, [_P, _P, _U, _I]),
    "ora_find_splits": (_U, [_P, _U, _P, _P]),
    "ora_build_lbvh_ref": (_I, [_P, _P, _U, _P, _U, _I, _P, _P, _P, _P, _P, _P, _P]),
    "ora_refit_recursive": (None, [_P]),
    "ora_tri_ray": (_I, [_P, _P, _P, _P, _P, _P]),
    "ora_box_ray": (_I, [_P, _P, _P, _P, _P]),
    "ora_trace_ref_ray": (_U, [_P, _P, _P, _P, _P, _U, _P, _P, _P, _P]),
    "ora_trace_ref_packet": (_U, [_P, _P, _P, _P, _P, _U, _U, _P, _P, _P]),
    "ora_trace_bruteforce": (None, [_P, _P, _U, _P, _P, _U, _P, _P, _P, C.c_float]),
    "ora_trace_closest": (None, [_P, _P, _P, _P, _P, _P, _U, _P, _P, _P, C.c_float]),
    "ora_hit_t_f64": (None, [_P, _P, _P, _P, _P, _U, _P]),
    "ora_build_sah": (_I, [_P, _P, _U, _P, _P, _U, _U, _P, _P]),
    "ora_reorder_faces": (None, [_P, _P, _U]),
    "ora_num_threads": (_I, []),
    "ora_interpolate": (None, [_P, _P, _U, _P, _U, _P, _P, _P, _P, _P, _P, _P, _P, _P]),
    "ora_camera_look_at": (None, [_P, _P, _P, _P]),
    "ora_camera_set_angle": (None, [_P, _P]),
    "ora_primary_ray_constants": (None, [_P, _P, C.c_float, C.c_float, _I, _I, _P]),
    "ora_primary_rays": (None, [_P, _I, _I, _P, _P]),
    "ora_shade": (None, [_P, _P, _I, _I, _P, _P, _P, _I]),
}

NODE_DTYPE = np.dtype([("w0", "<u4"), ("w1", "<u4"), ("bmin", "<f4", 3), ("bmax", "<f4", 3)])


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not SO.exists():
            subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
        L = C.CDLL(str(SO))
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _p(a: np.ndarray) -> int:
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data


def calc_boxes(F: np.ndarray, V: np.ndarray):
    n = F.shape[0]
    boxes = np.zeros((n, 2, 4), np.float32)
    idx = np.zeros(n, np.uint32)
    lib().ora_calc_boxes(_p(F), _p(V), n, _p(boxes), _p(idx))
    return boxes, idx


def morton32(boxes: np.ndarray, root: np.ndarray) -> np.ndarray:
    root = np.ascontiguousarray(root, np.float32)
    out = np.zeros(boxes.shape[0], np.uint32)
    lib().ora_morton32(_p(boxes), boxes.shape[0], _p(root[0]), _p(root[1]), _p(out))
    return out


def morton_wide(boxes: np.ndarray, root: np.ndarray, bits: int) -> np.ndarray:
    root = np.ascontiguousarray(root, np.float32)
    out = np.zeros(boxes.shape[0], np.uint64)
    lib().ora_morton_wide(_p(boxes), boxes.shape[0], _p(root[0]), _p(root[1]), bits, _p(out))
    return out


def sort_pairs_ref(keys: np.ndarray, vals: np.ndarray, bits: int = 32, init: bool = False):
    k = np.ascontiguousarray(keys, np.uint32).copy()
    v = np.ascontiguousarray(vals, np.uint32).copy()
    oob = lib().ora_sort_pairs_ref(_p(k), _p(v), k.shape[0], bits, int(init))
    return k, v, bool(oob)


def sort_pairs_u64(keys: np.ndarray, vals: np.ndarray):
    k = np.ascontiguousarray(keys, np.uint64).copy()
    v = np.ascontiguousarray(vals, np.uint32).copy()
    lib().ora_sort_pairs_u64(_p(k), _p(v), k.shape[0])
    return k, v


def bitonic_sort_general(keys: np.ndarray, vals: np.ndarray):
    """CUDA_bitonicSortGeneral<uint, uint> at any length; returns keys, values
    and whether the reference's Merge2x would write past the arrays."""
    k = np.ascontiguousarray(keys, np.uint32).copy()
    v = np.ascontiguousarray(vals, np.uint32).copy()
    oob = lib().ora_bitonic_sort_general(_p(k), _p(v), k.shape[0])
    return k, v, bool(oob)


def bitonic_sort(values: np.ndarray):
    v = np.ascontiguousarray(values, np.uint32).copy()
    lib().ora_bitonic_sort(_p(v), v.shape[0])
    return v


def bubble_sort512(keys: np.ndarray, indices: np.ndarray, init: bool = False):
    k = np.ascontiguousarray(keys, np.uint32).copy()
    v = np.ascontiguousarray(indices, np.uint32).copy()
    if lib().ora_bubble_sort512(_p(k), _p(v), k.shape[0], int(init)) != 0:
        raise ValueError("CUDA_bubbleSort512 takes at most 512 elements")
    return k, v


def find_splits(codes: np.ndarray):
    codes = np.ascontiguousarray(codes, np.uint32)
    A = np.zeros(max(1, codes.shape[0]), np.uint32)
    cnt = np.zeros(32, np.uint32)
    S = lib().ora_find_splits(_p(codes), codes.shape[0], _p(A), _p(cnt))
    return A[:S], cnt


class LBVH:
    """Result of the reference LBVH build."""

    def __init__(self, **kw):
        self.__dict__.update(kw)


def build_lbvh_ref(F: np.ndarray, V: np.ndarray, hitbox: np.ndarray, tree_depth: int = 12,
                   complete: bool = True) -> LBVH:
    F = np.ascontiguousarray(F, np.uint32).copy()
    V = np.ascontiguousarray(V, np.float32)
    hb = np.ascontiguousarray(hitbox, np.float32).reshape(2, 4)
    n = F.shape[0]
    boxes = np.zeros((n, 2, 4), np.float32)
    idx = np.zeros(n, np.uint32)
    nodes = np.zeros(2 * n, NODE_DTYPE)
    M, rd, fl = _U(), _U(), _U()
    codes = np.zeros(n, np.uint32)
    r = lib().ora_build_lbvh_ref(_p(F), _p(V), n, _p(hb), tree_depth, int(complete), _p(boxes), _p(idx),
                                 nodes.ctypes.data, C.byref(M), C.byref(rd), C.byref(fl), _p(codes))
    if r != 0:
        raise ValueError("oracle build failed")
    return LBVH(faces=F, boxes=boxes, indices=idx, nodes=nodes[:M.value].copy(), node_count=M.value,
                real_depth=rd.value, flags=fl.value, codes=codes)


def refit_recursive(nodes: np.ndarray) -> np.ndarray:
    out = np.ascontiguousarray(nodes).copy()
    lib().ora_refit_recursive(out.ctypes.data)
    return out


def tri_ray(o, d, v0, e1, e2):
    a = [np.ascontiguousarray(np.asarray(x, np.float32).reshape(4)) for x in (o, d, v0, e1, e2)]
    tuv = np.zeros(3, np.float32)
    h = lib().ora_tri_ray(*[_p(x) for x in a], _p(tuv))
    return bool(h), tuv


def box_ray(bmin, bmax, o, d):
    a = [np.ascontiguousarray(np.asarray(x, np.float32).reshape(4)) for x in (bmin, bmax, o, d)]
    t = np.zeros(1, np.float32)
    h = lib().ora_box_ray(*[_p(x) for x in a], _p(t))
    return bool(h), float(t[0])


def _outputs(n: int, face_id=None, coords=None, depths=None):
    fid = np.zeros(n, np.uint32) if face_id is None else np.ascontiguousarray(face_id, np.uint32).copy()
    uv = np.zeros((n, 2), np.float32) if coords is None else np.ascontiguousarray(coords, np.float32).copy()
    dep = (np.full(n, np.finfo(np.float32).max, np.float32) if depths is None
           else np.ascontiguousarray(depths, np.float32).copy())
    return fid, uv, dep


def trace_ref_ray(nodes, F, V, O, D, face_id=None, coords=None, depths=None, counters=False):
    n = O.shape[0]
    fid, uv, dep = _outputs(n, face_id, coords, depths)
    cnt = np.zeros((n, 2), np.uint32) if counters else None
    ovf = lib().ora_trace_ref_ray(nodes.ctypes.data, _p(F), _p(V), _p(O), _p(D), n, _p(fid), _p(uv), _p(dep),
                                  _p(cnt) if counters else None)
    if counters:
        return fid, uv, dep, ovf, cnt
    return fid, uv, dep, ovf


def trace_ref_packet(nodes, F, V, O, D, width, height, face_id=None, coords=None, depths=None):
    n = O.shape[0]
    assert n == width * height
    fid, uv, dep = _outputs(n, face_id, coords, depths)
    ovf = lib().ora_trace_ref_packet(nodes.ctypes.data, _p(F), _p(V), _p(O), _p(D), width, height, _p(fid),
                                     _p(uv), _p(dep))
    return fid, uv, dep, ovf


DET_EPSILON_REFERENCE = 0.000001


def trace_bruteforce(F, V, O, D, face_id=None, coords=None, depths=None, det_epsilon=DET_EPSILON_REFERENCE):
    """Closest hit over every face, ties to the smallest face index."""
    n = O.shape[0]
    fid, uv, dep = _outputs(n, face_id, coords, depths)
    lib().ora_trace_bruteforce(_p(F), _p(V), F.shape[0], _p(O), _p(D), n, _p(fid), _p(uv), _p(dep), det_epsilon)
    return fid, uv, dep


def trace_closest(tree: "LBVH", V, O, D, face_id=None, coords=None, depths=None,
                  det_epsilon=DET_EPSILON_REFERENCE):
    """The trace_bruteforce result found through a reference-layout tree
    (faces reordered, indices mapping back); face ids are original ids."""
    n = O.shape[0]
    fid, uv, dep = _outputs(n, face_id, coords, depths)
    lib().ora_trace_closest(tree.nodes.ctypes.data, _p(tree.faces), _p(tree.indices), _p(V), _p(O), _p(D), n,
                            _p(fid), _p(uv), _p(dep), det_epsilon)
    return fid, uv, dep


def hit_t_f64(F, V, O, D, face_id) -> np.ndarray:
    face_id = np.ascontiguousarray(face_id, np.uint32)
    out = np.zeros(face_id.shape[0], np.float64)
    lib().ora_hit_t_f64(_p(F), _p(V), _p(O), _p(D), _p(face_id), face_id.shape[0], _p(out))
    return out


def build_sah(boxes, indices, root, nodes=None, n_existing=0):
    n = indices.shape[0]
    idx = np.ascontiguousarray(indices, np.uint32).copy()
    cap = 4 * n + 8
    out = np.zeros(cap, NODE_DTYPE)
    if nodes is not None:
        out[: nodes.shape[0]] = nodes
    M, fl = _U(), _U()
    rt = np.ascontiguousarray(root, np.float32).reshape(2, 4)
    lib().ora_build_sah(_p(np.ascontiguousarray(boxes, np.float32)), _p(idx), n, _p(rt), out.ctypes.data, cap,
                        n_existing, C.byref(M), C.byref(fl))
    return out[: M.value].copy(), idx, fl.value


def reorder_faces(F, idx):
    F2 = np.ascontiguousarray(F, np.uint32).copy()
    lib().ora_reorder_faces(_p(F2), _p(np.ascontiguousarray(idx, np.uint32)), F2.shape[0])
    return F2


def interpolate(face_id, coords, F, V, normals=None, colors=None, texcoords=None, init=None):
    """Per-pixel position, normal, colour, texcoords and material of the hit
    face (CUDA_interpolateData.cu:45-141); 'init' (p, n, c, t, m) are the
    buffers' contents before the call (pixels without a face keep them)."""
    n = face_id.shape[0]
    fid = np.ascontiguousarray(face_id, np.uint32)
    uv = np.ascontiguousarray(coords, np.float32)
    p, nn, c, t, m = init if init is not None else (np.zeros((n, 4), np.float32), np.zeros((n, 4), np.float32),
                                                    np.zeros((n, 4), np.float32), np.zeros((n, 2), np.float32),
                                                    np.zeros(n, np.uint32))
    p, nn, c, t, m = (np.ascontiguousarray(x).copy() for x in (p, nn, c, t, m))
    opt = lambda a, dt: None if a is None else _p(np.ascontiguousarray(a, dt))
    keep = [np.ascontiguousarray(a, np.float32) if a is not None else None for a in (normals, colors, texcoords)]
    lib().ora_interpolate(_p(fid), _p(uv), n, _p(np.ascontiguousarray(F, np.uint32)), F.shape[0],
                          _p(np.ascontiguousarray(V, np.float32)), None if keep[0] is None else _p(keep[0]),
                          None if keep[1] is None else _p(keep[1]), None if keep[2] is None else _p(keep[2]),
                          _p(p), _p(nn), _p(c), _p(t), _p(m))
    return p, nn, c, t, m


def shade(depths, normals, width, height, back_color, cam_dir, pitch=None):
    """kernel_shade: the packed pixels (uint32, height x pitch) of a frame."""
    pitch = width if pitch is None else pitch
    d = np.ascontiguousarray(depths, np.float32)
    nrm = np.ascontiguousarray(normals, np.float32)
    back = np.ascontiguousarray(back_color, np.float32)
    cam = np.zeros(4, np.float32)
    cam[:len(cam_dir)] = cam_dir
    out = np.zeros(height * pitch, np.uint32)
    lib().ora_shade(_p(d), _p(nrm), width, height, _p(back), _p(cam), _p(out), pitch)
    return out


class Camera:
    """The reference camera's state (CUDACamera.cpp:29-50 defaults) driven
    through the oracle's restatement of setAngle, lookAt and the host part of
    CUDA_calcPrimaryRays. Values are float32, as the reference keeps them."""

    def __init__(self):
        self.center = np.zeros(3, np.float32)
        self.angle = np.array([np.float32(3.14159265359) * np.float32(0.5), 0, 0], np.float32)
        self.dir = np.array([0, 0, 1], np.float32)
        self.fov = np.float32(np.float32(3.14159265359) * np.float32(0.5))
        self.ratio = np.float32(np.float32(4.0) / np.float32(3.0))

    def set_center(self, x, y, z):
        self.center = np.array([x, y, z], np.float32)

    def set_angle(self, x, y, z):
        self.angle = np.array([x, y, z], np.float32)
        self.dir = np.zeros(3, np.float32)
        lib().ora_camera_set_angle(_p(self.angle), _p(self.dir))

    def look_at(self, x, y, z):
        pt = np.array([x, y, z], np.float32)
        self.dir = np.zeros(3, np.float32)
        lib().ora_camera_look_at(_p(self.center), _p(pt), _p(self.angle), _p(self.dir))

    def set_fov(self, fov):
        self.fov = np.float32(fov)

    def set_ratio(self, width, height=None):
        self.ratio = np.float32(width) if height is None else np.float32(np.float32(width) / np.float32(height))

    def ray_constants(self, width: int, height: int) -> np.ndarray:
        out = np.zeros((4, 4), np.float32)
        lib().ora_primary_ray_constants(_p(self.center), _p(self.angle), float(self.fov), float(self.ratio), width,
                                        height, _p(out))
        return out

    def primary_rays(self, width: int, height: int):
        cst = self.ray_constants(width, height)
        O = np.zeros((width * height, 4), np.float32)
        D = np.zeros((width * height, 4), np.float32)
        lib().ora_primary_rays(_p(cst), width, height, _p(O), _p(D))
        return O, D


def num_threads() -> int:
    return lib().ora_num_threads()
