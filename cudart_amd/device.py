"""Device-side convenience layer over the C ABI for torch tensors.

torch provides device memory and the stream; every computation goes through
librk.so (include/rk.h). There is no CPU path: without the library or a GPU
these functions raise.
"""
from __future__ import annotations

import contextlib
import ctypes as C
from dataclasses import dataclass

import numpy as np
import torch

from . import native as N

_U32 = torch.int32  # uint32 payloads are carried in int32 tensors


def _dev(a, dtype=None) -> torch.Tensor:
    if isinstance(a, torch.Tensor):
        return a.to("cuda") if a.device.type != "cuda" else a
    t = torch.from_numpy(np.ascontiguousarray(a))
    return t.to("cuda")


def _stream_ptr(stream) -> int | None:
    """The HIP stream of 'stream', or of torch's current stream when None."""
    s = torch.cuda.current_stream() if stream is None else stream
    return int(s.cuda_stream) or None


def _on(stream):
    """Allocations made inside belong to 'stream' (torch's caching allocator
    reuses a freed block only in order on the stream it was allocated on), so a
    temporary freed on return is not handed out while the kernels that use it
    still run on 'stream'."""
    return torch.cuda.stream(stream) if stream is not None else contextlib.nullcontext()


def u32_view(t: torch.Tensor) -> np.ndarray:
    return t.cpu().numpy().view(np.uint32)


def to_u32_tensor(a: np.ndarray) -> torch.Tensor:
    return torch.from_numpy(np.ascontiguousarray(a, np.uint32).view(np.int32)).to("cuda")


def _check(s, what):
    N.check(s, what)


def calc_mesh_boxes(F: torch.Tensor, V: torch.Tensor, stream=None):
    with _on(stream):
        n = F.shape[0]
        boxes = torch.empty((n, 2, 4), dtype=torch.float32, device="cuda")
        idx = torch.empty(n, dtype=_U32, device="cuda")
        _check(N.lib().rk_calc_mesh_boxes(_stream_ptr(stream), F.data_ptr(), V.data_ptr(), n, boxes.data_ptr(),
                                          idx.data_ptr()), "calc_mesh_boxes")
        return boxes, idx


def morton(boxes: torch.Tensor, root: np.ndarray, mode: int, stream=None) -> torch.Tensor:
    with _on(stream):
        n = boxes.shape[0]
        dt = torch.int64 if mode == N.RK_MORTON_63 else torch.int32
        out = torch.empty(n, dtype=dt, device="cuda")
        rb = N.aabox(root)
        _check(N.lib().rk_calc_morton_codes(_stream_ptr(stream), boxes.data_ptr(), n, C.byref(rb), mode, out.data_ptr()),
               "morton")
        return out


def workspace(nbytes: int) -> torch.Tensor:
    return torch.empty(max(1, nbytes), dtype=torch.uint8, device="cuda")


def sort_general(keys: torch.Tensor, vals: torch.Tensor, bits: int = 32, init: bool = False, stream=None):
    with _on(stream):
        n = keys.shape[0]
        ws = workspace(N.lib().rk_sort_workspace_size(n, 4))
        _check(N.lib().rk_radix_sort_general(_stream_ptr(stream), keys.data_ptr(), vals.data_ptr(), n, bits, int(init),
                                             ws.data_ptr(), ws.numel()), "radix_sort_general")
        return keys, vals


def sort_values(values: torch.Tensor, bits: int = 32, stream=None):
    with _on(stream):
        n = values.shape[0]
        ws = workspace(N.lib().rk_sort_workspace_size(n, 4))
        _check(N.lib().rk_radix_sort(_stream_ptr(stream), values.data_ptr(), n, bits, ws.data_ptr(), ws.numel()),
               "radix_sort")
        return values


def sort_pairs_u64(keys: torch.Tensor, vals: torch.Tensor, bits: int = 64, stream=None):
    with _on(stream):
        n = keys.shape[0]
        ws = workspace(N.lib().rk_sort_workspace_size(n, 8))
        _check(N.lib().rk_radix_sort_pairs_u64(_stream_ptr(stream), keys.data_ptr(), vals.data_ptr(), n, bits,
                                               ws.data_ptr(), ws.numel()), "radix_sort_pairs_u64")
        return keys, vals


def bitonic_sort(values: torch.Tensor, stream=None, ws: torch.Tensor | None = None):
    """rk_bitonic_sort (CUDA_bitonicSort<uint>): values ascending, in place."""
    with _on(stream):
        n = values.shape[0]
        need = N.lib().rk_sort_workspace_size(n, 4)
        if ws is None or ws.numel() < need:
            ws = workspace(need)
        _check(N.lib().rk_bitonic_sort(_stream_ptr(stream), values.data_ptr(), n, ws.data_ptr(), ws.numel()),
               "bitonic_sort")
        return values, ws


def bitonic_sort_general(keys: torch.Tensor, vals: torch.Tensor, stream=None, ws: torch.Tensor | None = None):
    """rk_bitonic_sort_general (CUDA_bitonicSortGeneral<uint, uint>), in place."""
    with _on(stream):
        n = keys.shape[0]
        need = N.lib().rk_sort_workspace_size(n, 4)
        if ws is None or ws.numel() < need:
            ws = workspace(need)
        _check(N.lib().rk_bitonic_sort_general(_stream_ptr(stream), keys.data_ptr(), vals.data_ptr(), n,
                                               ws.data_ptr(), ws.numel()), "bitonic_sort_general")
        return keys, vals, ws


def bubble_sort512(keys: torch.Tensor, indices: torch.Tensor, init: bool = False, stream=None):
    """rk_bubble_sort512 (CUDA_bubbleSort512), in place."""
    _check(N.lib().rk_bubble_sort512(_stream_ptr(stream), keys.data_ptr(), indices.data_ptr(), keys.shape[0],
                                     int(init)), "bubble_sort512")
    return keys, indices


def workspace_flags(ws: torch.Tensor, stream=None) -> int:
    """The RK_FLAG_* word a sort left at the start of its workspace."""
    f = N._U32()
    _check(N.lib().rk_workspace_flags(_stream_ptr(stream), ws.data_ptr(), C.byref(f)), "workspace_flags")
    return f.value


@dataclass
class RefTree:
    faces: torch.Tensor      # (n,4) int32, reordered when complete
    boxes: torch.Tensor      # (n,2,4) float32
    indices: torch.Tensor    # (n,) int32
    nodes: torch.Tensor      # (2n, 8) int32 (32-B bvnodes), first node_count valid
    info: torch.Tensor       # (8,) int32
    node_count: int = -1

    def host_info(self) -> np.ndarray:
        return u32_view(self.info)


class LBVHBuilder:
    """Reusable buffers for rk_build_lbvh on meshes of up to n faces."""

    def __init__(self, n: int):
        self.n = n
        self.ws = workspace(N.lib().rk_build_lbvh_workspace_size(n))
        self.boxes = torch.empty((n, 2, 4), dtype=torch.float32, device="cuda")
        self.indices = torch.empty(n, dtype=_U32, device="cuda")
        self.nodes = torch.empty((2 * n, 8), dtype=_U32, device="cuda")
        self.info = torch.zeros(N.RK_INFO_WORDS, dtype=_U32, device="cuda")

    def build(self, F: torch.Tensor, V: torch.Tensor, hitbox: np.ndarray, tree_depth: int = 12,
              complete: bool = True, stream=None) -> RefTree:
        n = F.shape[0]
        assert n <= self.n
        hb = N.aabox(hitbox)
        _check(N.lib().rk_build_lbvh(_stream_ptr(stream), F.data_ptr(), V.data_ptr(), n, C.byref(hb), int(complete),
                                     tree_depth, self.boxes.data_ptr(), self.indices.data_ptr(),
                                     self.nodes.data_ptr(), 2 * self.n, self.info.data_ptr(), self.ws.data_ptr(),
                                     self.ws.numel()), "build_lbvh")
        return RefTree(F, self.boxes[:n], self.indices[:n], self.nodes, self.info)


def build_lbvh(F: torch.Tensor, V: torch.Tensor, hitbox: np.ndarray, tree_depth: int = 12, complete: bool = True,
               stream=None) -> RefTree:
    t = LBVHBuilder(F.shape[0]).build(F, V, hitbox, tree_depth, complete, stream)
    torch.cuda.synchronize()
    t.node_count = int(t.host_info()[N.RK_INFO_NODE_COUNT])
    return t


def new_outputs(n: int, tmax: float | None = None):
    fid = torch.zeros(n, dtype=_U32, device="cuda")
    uv = torch.zeros((n, 2), dtype=torch.float32, device="cuda")
    dep = torch.full((n,), float(np.finfo(np.float32).max) if tmax is None else tmax, dtype=torch.float32,
                     device="cuda")
    return fid, uv, dep


def trace_ref(tree: RefTree, V: torch.Tensor, O: torch.Tensor, D: torch.Tensor, width: int, height: int,
              coherency: bool = True, out=None, stream=None):
    with _on(stream):
        n = width * height
        fid, uv, dep = out if out is not None else new_outputs(n)
        node_count = tree.node_count if tree.node_count > 0 else tree.nodes.shape[0]
        _check(N.lib().rk_raytrace_mesh_tree(_stream_ptr(stream), tree.nodes.data_ptr(), node_count, tree.faces.data_ptr(),
                                             tree.faces.shape[0], V.data_ptr(), O.data_ptr(), D.data_ptr(), width, height, fid.data_ptr(),
                                             uv.data_ptr(), dep.data_ptr(), int(coherency)), "raytrace_mesh_tree")
        return fid, uv, dep


def set_ref_handoff(ray_faces: int = N.RK_HANDOFF_AUTO, packet_faces: int = N.RK_HANDOFF_AUTO,
                    list_cap: int = 65536) -> None:
    """Hand-off limits of trace_ref's second passes (process-wide; 0xFFFFFFFF turns one off)."""
    _check(N.lib().rk_raytrace_set_handoff(ray_faces, packet_faces, list_cap), "raytrace_set_handoff")


def get_ref_handoff() -> tuple[int, int, int]:
    a, b, c = N._U32(), N._U32(), N._U32()
    _check(N.lib().rk_raytrace_get_handoff(C.byref(a), C.byref(b), C.byref(c)), "raytrace_get_handoff")
    return a.value, b.value, c.value


@dataclass
class FastTree:
    nodes: torch.Tensor
    tris: torch.Tensor
    indices: torch.Tensor
    boxes: torch.Tensor
    info: torch.Tensor
    desc: N.FastBVH


class FastBuilder:
    """Reusable buffers for rk_fast_build on meshes of up to n faces."""

    def __init__(self, n: int, morton_mode: int = N.RK_MORTON_30):
        self.n = n
        self.mode = morton_mode
        self.ws = workspace(N.lib().rk_fast_build_workspace_size(n, morton_mode))
        self.nodes = torch.empty((max(1, n), 16), dtype=_U32, device="cuda")
        self.tris = torch.empty((n, 12), dtype=torch.float32, device="cuda")
        self.indices = torch.empty(n, dtype=_U32, device="cuda")
        self.boxes = torch.empty((n, 2, 4), dtype=torch.float32, device="cuda")
        self.info = torch.zeros(N.RK_INFO_WORDS, dtype=_U32, device="cuda")

    def build(self, F: torch.Tensor, V: torch.Tensor, hitbox: np.ndarray, leaf_max: int = 1,
              stream=None) -> FastTree:
        n = F.shape[0]
        assert n <= self.n
        desc = N.FastBVH(self.nodes.data_ptr(), self.tris.data_ptr(), self.indices.data_ptr(), self.boxes.data_ptr(),
                         n, 0)
        hb = N.aabox(hitbox)
        _check(N.lib().rk_fast_build(_stream_ptr(stream), F.data_ptr(), V.data_ptr(), n, C.byref(hb), self.mode,
                                     leaf_max, C.byref(desc), self.info.data_ptr(), self.ws.data_ptr(),
                                     self.ws.numel()), "fast_build")
        return FastTree(self.nodes, self.tris, self.indices[:n], self.boxes[:n], self.info, desc)


def _opt(det_epsilon, config=0, reset_depth=False):
    o = N.trace_options(det_epsilon, config, reset_depth)
    return None if o is None else C.byref(o)


def fast_trace(tree: FastTree, O: torch.Tensor, D: torch.Tensor, out=None, stream=None, det_epsilon=None, config=0,
               reset_depth=False):
    """det_epsilon None = the reference threshold (1e-6); 0.0 = only parallel rays rejected.
    config: a traversal variant from autotune() (0 = the default). reset_depth:
    every ray starts at tmax FLT_MAX and every depth is written (no fill needed)."""
    with _on(stream):
        n = O.shape[0]
        fid, uv, dep = out if out is not None else new_outputs(n)
        _check(N.lib().rk_fast_trace(_stream_ptr(stream), C.byref(tree.desc), O.data_ptr(), D.data_ptr(), n,
                                     fid.data_ptr(), uv.data_ptr(), dep.data_ptr(),
                                     _opt(det_epsilon, config, reset_depth)),
               "fast_trace")
        return fid, uv, dep


def autotune(tree: FastTree, O: torch.Tensor, D: torch.Tensor, out, det_epsilon=None, reps=3, stream=None,
             reset_depth=False):
    """rk_fast_autotune: times every traversal variant on these rays (out[2]
    holds tmax and is restored before each call) and returns the fastest
    (N.TraceTuning); out then holds the frame it traced."""
    with _on(stream):
        n = O.shape[0]
        fid, uv, dep = out
        ws = workspace(N.lib().rk_fast_autotune_workspace_size(n))
        best = N.TraceTuning()
        _check(N.lib().rk_fast_autotune(_stream_ptr(stream), C.byref(tree.desc), O.data_ptr(), D.data_ptr(), n,
                                        fid.data_ptr(), uv.data_ptr(), dep.data_ptr(),
                                        _opt(det_epsilon, 0, reset_depth), reps,
                                        ws.data_ptr(), ws.numel(), C.byref(best)), "fast_autotune")
        return best


def autotune_camera(tree: FastTree, cam: np.ndarray, width: int, height: int, out, det_epsilon=None, reps=3,
                    stream=None):
    """rk_fast_autotune_camera: the fastest camera-kernel variant for this frame."""
    with _on(stream):
        fid, uv, dep = out
        best = N.TraceTuning()
        c = camera_rays(cam)
        _check(N.lib().rk_fast_autotune_camera(_stream_ptr(stream), C.byref(tree.desc), C.byref(c), width, height,
                                               fid.data_ptr(), uv.data_ptr(), dep.data_ptr(), _opt(det_epsilon),
                                               reps, C.byref(best)), "fast_autotune_camera")
        return best


def trace_variant(config: int, camera: bool = False) -> N.TraceTuning:
    t = N.TraceTuning()
    _check(N.lib().rk_fast_trace_variant(config, int(camera), C.byref(t)), "fast_trace_variant")
    return t


def fast_trace_counted(tree: FastTree, O: torch.Tensor, D: torch.Tensor, out=None, stream=None, det_epsilon=None):
    """fast_trace plus the total child-box and triangle tests performed."""
    with _on(stream):
        n = O.shape[0]
        fid, uv, dep = out if out is not None else new_outputs(n)
        cnt = torch.zeros(2, dtype=torch.int64, device="cuda")
        _check(N.lib().rk_fast_trace_counted(_stream_ptr(stream), C.byref(tree.desc), O.data_ptr(), D.data_ptr(), n,
                                             fid.data_ptr(), uv.data_ptr(), dep.data_ptr(), _opt(det_epsilon),
                                             cnt.data_ptr()), "fast_trace_counted")
        return fid, uv, dep, cnt


class SAHBuilder:
    """Reusable buffers for rk_build_sah on meshes of up to n faces, on top of
    an LBVH of up to 'existing' nodes (0: pure SAH trees only)."""

    def __init__(self, n: int, existing: int = 0):
        self.n, self.existing = n, existing
        self.cap = existing + 2 * n + 1
        self.ws = workspace(N.lib().rk_build_sah_workspace_size(n, existing))
        self.boxes = torch.empty((n, 2, 4), dtype=torch.float32, device="cuda")
        self.indices = torch.empty(n, dtype=_U32, device="cuda")
        self.nodes = torch.zeros((self.cap, 8), dtype=_U32, device="cuda")
        self.info = torch.zeros(N.RK_INFO_WORDS, dtype=_U32, device="cuda")

    def build(self, F: torch.Tensor, V: torch.Tensor, hitbox: np.ndarray, lbvh: RefTree | None = None,
              stream=None) -> RefTree:
        """A pure SAH tree (lbvh None) or the Hybrid refinement of an LBVH built
        with complete=False (its boxes and indices are used in place)."""
        n = F.shape[0]
        assert n <= self.n
        existing = 0 if lbvh is None else int(lbvh.node_count)
        assert existing <= self.existing
        if lbvh is None:
            boxes, indices = self.boxes[:n], self.indices[:n]
        else:
            boxes, indices = lbvh.boxes, lbvh.indices
            self.nodes[:existing] = lbvh.nodes[:existing]
        hb = N.aabox(hitbox)
        _check(N.lib().rk_build_sah(_stream_ptr(stream), F.data_ptr(), V.data_ptr(), n, C.byref(hb),
                                    boxes.data_ptr(), indices.data_ptr(), self.nodes.data_ptr(), existing,
                                    existing + 2 * n + 1, self.info.data_ptr(), self.ws.data_ptr(),
                                    self.ws.numel()), "build_sah")
        return RefTree(F, boxes, indices, self.nodes, self.info)


def build_sah(F: torch.Tensor, V: torch.Tensor, hitbox: np.ndarray, lbvh: RefTree | None = None, stream=None):
    """rk_build_sah: a pure SAH tree (lbvh None) or the Hybrid refinement of an
    LBVH built with complete=False (its boxes, indices and nodes are reused)."""
    existing = 0 if lbvh is None else int(lbvh.node_count)
    t = SAHBuilder(F.shape[0], existing).build(F, V, hitbox, lbvh, stream)
    torch.cuda.synchronize()
    t.node_count = int(t.host_info()[N.RK_INFO_NODE_COUNT])
    return t


# ---- batches -------------------------------------------------------------------

def batch_config(leaf_max=1, trace_mode=1, trace_wg=256, stack_depth=64, ray_order=0, donate_min=0) -> N.BatchConfig:
    c = N.BatchConfig()
    c.leaf_max, c.trace_mode, c.trace_wg, c.stack_depth = leaf_max, trace_mode, trace_wg, stack_depth
    c.ray_order, c.donate_min = ray_order, donate_min
    return c


class Batch:
    """Many meshes built and traced together (rk_batch_build / rk_batch_trace).
    F holds each mesh's faces with mesh-local vertex indices, V the meshes'
    vertices back to back, O/D their rays back to back; 'table' is from
    batch_table, 'hitboxes' one (2, 4) box per mesh."""

    def __init__(self, table: np.ndarray, F: torch.Tensor, V: torch.Tensor, hitboxes: np.ndarray,
                 O: torch.Tensor, D: torch.Tensor):
        self.m = int(table.shape[0])
        self.nf, self.nr = int(F.shape[0]), int(O.shape[0])
        self.table = table
        self.meshes = torch.from_numpy(table.view(np.uint8).copy()).cuda()
        self.hitboxes = torch.from_numpy(np.ascontiguousarray(hitboxes, np.float32).reshape(self.m, 2, 4)).cuda()
        self.F, self.V, self.O, self.D = F, V, O, D
        self.nodes = torch.empty((self.nf, 16), dtype=_U32, device="cuda")
        self.tris = torch.empty((self.nf, 12), dtype=torch.float32, device="cuda")
        self.indices = torch.empty(self.nf, dtype=_U32, device="cuda")
        self.info = torch.zeros(N.RK_INFO_WORDS, dtype=_U32, device="cuda")
        self.bvh = N.FastBVH(self.nodes.data_ptr(), self.tris.data_ptr(), self.indices.data_ptr(), None, self.nf, 0)
        self.build_bytes = N.lib().rk_batch_build_workspace_size(self.m, self.nf)
        self.trace_bytes = N.lib().rk_batch_trace_workspace_size(self.m, self.nr)
        self.ws = workspace(self.build_bytes + self.trace_bytes)
        self.config = N.lib().rk_batch_config_default()

    def build(self, config: N.BatchConfig | None = None, stream=None):
        c = config or self.config
        _check(N.lib().rk_batch_build(_stream_ptr(stream), self.meshes.data_ptr(), self.m, self.F.data_ptr(),
                                      self.V.data_ptr(), self.nf, self.hitboxes.data_ptr(), C.byref(c),
                                      C.byref(self.bvh), self.info.data_ptr(), self.ws.data_ptr(), self.build_bytes),
               "batch_build")

    def trace(self, out, config: N.BatchConfig | None = None, det_epsilon=None, stream=None):
        c = config or self.config
        fid, uv, dep = out
        _check(N.lib().rk_batch_trace(_stream_ptr(stream), self.meshes.data_ptr(), self.m, C.byref(self.bvh),
                                      self.O.data_ptr(), self.D.data_ptr(), self.nr, fid.data_ptr(), uv.data_ptr(),
                                      dep.data_ptr(), _opt(det_epsilon), C.byref(c),
                                      self.ws.data_ptr() + self.build_bytes, self.trace_bytes), "batch_trace")
        return out

    def autotune(self, out, reps: int = 3, det_epsilon=None, stream=None):
        fid, uv, dep = out
        best, ms = N.BatchConfig(), C.c_float(0.0)
        _check(N.lib().rk_batch_autotune(_stream_ptr(stream), self.meshes.data_ptr(), self.m, self.F.data_ptr(),
                                         self.V.data_ptr(), self.nf, self.hitboxes.data_ptr(), C.byref(self.bvh),
                                         self.info.data_ptr(), self.O.data_ptr(), self.D.data_ptr(), self.nr,
                                         fid.data_ptr(), uv.data_ptr(), dep.data_ptr(), _opt(det_epsilon), reps,
                                         C.byref(best), C.byref(ms), self.ws.data_ptr(), self.ws.numel()),
               "batch_autotune")
        self.config = best
        return best, ms.value


class Graph:
    """rk_* calls issued on 'stream' (a torch.cuda.Stream) inside 'with
    g.capture():' are recorded once and replayed by g.launch()."""

    def __init__(self, stream):
        self.stream = stream
        self.handle = C.c_void_p()

    def capture(self):
        g = self

        class _Cap:
            def __enter__(self_):
                _check(N.lib().rk_graph_capture_begin(_stream_ptr(g.stream)), "graph_capture_begin")
                return g

            def __exit__(self_, *exc):
                _check(N.lib().rk_graph_capture_end(_stream_ptr(g.stream), C.byref(g.handle)), "graph_capture_end")
                return False

        return _Cap()

    def launch(self):
        _check(N.lib().rk_graph_launch(self.handle, _stream_ptr(self.stream)), "graph_launch")

    def __del__(self):
        if self.handle:
            N.lib().rk_graph_destroy(self.handle)
            self.handle = C.c_void_p()


# ---- camera frames, deformation, serving ------------------------------------------

def camera_rays(cam: np.ndarray) -> N.CameraRays:
    """rk_camera_rays from the (4, 4) array of native.camera_lookat."""
    c = np.ascontiguousarray(cam, np.float32)
    return N.CameraRays(*(N.Float4(*map(float, c[k])) for k in range(4)))


def fast_trace_camera(tree: FastTree, cam: np.ndarray, width: int, height: int, out=None, det_epsilon=None,
                      stream=None, config=0):
    with _on(stream):
        n = width * height
        fid, uv, dep = out if out is not None else new_outputs(n)
        c = camera_rays(cam)
        _check(N.lib().rk_fast_trace_camera(_stream_ptr(stream), C.byref(tree.desc), C.byref(c), width, height,
                                            fid.data_ptr(), uv.data_ptr(), dep.data_ptr(), _opt(det_epsilon, config)),
               "fast_trace_camera")
        return fid, uv, dep


def deform(base: torch.Tensor, amplitude: float, omega: float, time: float, seed: int, along_z: bool,
           out: torch.Tensor | None = None, stream=None) -> torch.Tensor:
    with _on(stream):
        out = torch.empty_like(base) if out is None else out
        _check(N.lib().rk_deform_vertices(_stream_ptr(stream), base.data_ptr(), out.data_ptr(), base.shape[0], amplitude,
                                          omega, time, seed, int(along_z)), "deform")
        return out


class Server:
    """rk_server_*: a request loop over deforming scenes (see include/rk.h)."""

    def __init__(self, arena_bytes: int, **cfg):
        c = N.lib().rk_server_config_default()
        for k, v in cfg.items():
            setattr(c, k, v)
        self.cfg = c
        self.h = C.c_void_p()
        _check(N.lib().rk_server_create(C.byref(c), int(arena_bytes), C.byref(self.h)), "server_create")
        self.scenes = []

    def add_scene(self, V: np.ndarray, F: np.ndarray, along_z: bool) -> int:
        V = np.ascontiguousarray(V, np.float32)
        F = np.ascontiguousarray(F, np.uint32)
        sid = N._U32()
        _check(N.lib().rk_server_add_scene(self.h, V.ctypes.data, V.shape[0], F.ctypes.data, F.shape[0],
                                           int(along_z), C.byref(sid)), "server_add_scene")
        self.scenes.append((V, F, along_z))
        return sid.value

    def submit(self, scene: int, time: float, cam: np.ndarray, width: int, height: int, flags: int = 0,
               out=None) -> int:
        r = N.Request()
        r.scene, r.width, r.height, r.flags, r.time = scene, width, height, flags, time
        r.camera = camera_rays(cam)
        r.arrival = -1.0
        if out is not None:
            r.faces_id, r.coords, r.depths = (x.data_ptr() for x in out)
        t = C.c_uint64()
        _check(N.lib().rk_server_submit(self.h, C.byref(r), C.byref(t)), "server_submit")
        return t.value

    def wait_all(self):
        _check(N.lib().rk_server_wait_all(self.h), "server_wait_all")

    def result(self, ticket: int) -> N.Result:
        r = N.Result()
        _check(N.lib().rk_server_result(self.h, ticket, C.byref(r)), "server_result")
        return r

    def run_poisson(self, rate: float, count: int, side_lo: int, side_hi: int, seed: int):
        st, first = N.ServerStats(), C.c_uint64()
        _check(N.lib().rk_server_run_poisson(self.h, rate, count, side_lo, side_hi, seed, C.byref(st),
                                             C.byref(first)), "server_run_poisson")
        return st, first.value

    def close(self):
        if self.h:
            N.lib().rk_server_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        self.close()


def interpolate(faces_id: torch.Tensor, coords: torch.Tensor, width: int, height: int, F: torch.Tensor,
                V: torch.Tensor, normals=None, colors=None, texcoords=None, outs=None, stream=None):
    """rk_interpolate_data; outs = (points, normals, colors, texcoords, materials)
    tensors (any may be None), allocated zeroed when not given."""
    with _on(stream):
        n = width * height
        if outs is None:
            outs = (torch.zeros((n, 4), dtype=torch.float32, device="cuda"),
                    torch.zeros((n, 4), dtype=torch.float32, device="cuda"),
                    torch.zeros((n, 4), dtype=torch.float32, device="cuda"),
                    torch.zeros((n, 2), dtype=torch.float32, device="cuda"),
                    torch.zeros(n, dtype=_U32, device="cuda"))
        ptr = lambda t: None if t is None else t.data_ptr()
        _check(N.lib().rk_interpolate_data(_stream_ptr(stream), faces_id.data_ptr(), coords.data_ptr(), width, height,
                                           F.data_ptr(), F.shape[0], V.data_ptr(), ptr(normals), ptr(colors),
                                           ptr(texcoords), *(ptr(t) for t in outs)), "interpolate_data")
        return outs


def shade(depths: torch.Tensor, normals: torch.Tensor, width: int, height: int, back_color, cam_dir,
          pixels: torch.Tensor | None = None, pitch: int | None = None, stream=None):
    """rk_shade: one packed A,R,G,B word per pixel, rows of 'pitch' words."""
    with _on(stream):
        pitch = width if pitch is None else pitch
        if pixels is None:
            pixels = torch.zeros(pitch * height, dtype=_U32, device="cuda")
        _check(N.lib().rk_shade(_stream_ptr(stream), depths.data_ptr(), normals.data_ptr(), width, height,
                                N.Float4(*map(float, back_color)), N.Float4(*map(float, cam_dir)),
                                pixels.data_ptr(), pitch), "shade")
        return pixels


def render_camera(tree: FastTree, cam: np.ndarray, width: int, height: int, faces: torch.Tensor, V: torch.Tensor,
                  back_color, cam_dir, normals: torch.Tensor | None = None, pixels: torch.Tensor | None = None,
                  pitch: int | None = None, out=None, det_epsilon=None, stream=None):
    """rk_fast_render_camera: rays, closest hit, hit normal and shaded pixel in
    one launch. faces: the mesh's faces in the tree's leaf order. out: None or
    (faces_id, coords, depths) to receive the trace's outputs too."""
    with _on(stream):
        pitch = width if pitch is None else pitch
        if pixels is None:
            pixels = torch.zeros(pitch * height, dtype=_U32, device="cuda")
        c = camera_rays(cam)
        o = (None, None, None) if out is None else tuple(t.data_ptr() for t in out)
        _check(N.lib().rk_fast_render_camera(_stream_ptr(stream), C.byref(tree.desc), C.byref(c), width, height,
                                             faces.data_ptr(), V.data_ptr(),
                                             None if normals is None else normals.data_ptr(),
                                             N.Float4(*map(float, back_color)), N.Float4(*map(float, cam_dir)),
                                             pixels.data_ptr(), pitch, *o, _opt(det_epsilon)), "fast_render_camera")
        return pixels


def fast_trace_sorted(tree: FastTree, O: torch.Tensor, D: torch.Tensor, out=None, stream=None, det_epsilon=None,
                      ws: torch.Tensor | None = None, config=0, reset_depth=False):
    """rk_fast_trace_sorted: the same results as fast_trace, rays traced in
    origin-Morton order (for incoherent rays); config: an rk_fast_trace variant."""
    with _on(stream):
        n = O.shape[0]
        fid, uv, dep = out if out is not None else new_outputs(n)
        need = N.lib().rk_fast_trace_sorted_workspace_size(n)
        if ws is None or ws.numel() < need:
            ws = workspace(need)
        _check(N.lib().rk_fast_trace_sorted(_stream_ptr(stream), C.byref(tree.desc), O.data_ptr(), D.data_ptr(), n,
                                            fid.data_ptr(), uv.data_ptr(), dep.data_ptr(),
                                            _opt(det_epsilon, config, reset_depth),
                                            ws.data_ptr(), ws.numel()), "fast_trace_sorted")
        return fid, uv, dep
