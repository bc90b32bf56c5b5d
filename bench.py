"""Benchmark of the mesh-tree build + closest-hit trace path on one MI355X.

Headline workload: the 708x708 fBm heightfield (999,698 FP32 triangles) and
1920x1080 primary rays from a camera at 30 degrees elevation, 45 degrees
azimuth, distance 1.5 from (0.5, 0.5, 0.1), 60 degree fov.

A step is one frame with the tree resident in HBM: a fill kernel resets tmax
and rk_fast_trace finds the closest hit of all 2,073,600 rays. `value` is
Mrays/s over the timed frames. The triangle test uses the reference's
absolute parallel-ray threshold 1e-6, so the hits are the reference's hits.
That threshold rejects the grazing hits on this mesh, whose edges are ~1.4e-3
so that |det| ~ 1e-6; the frame with threshold 0 (only exactly parallel rays
rejected) is reported beside it (mrays_s_det_epsilon_0,
hit_fraction_det_epsilon_0). The same frame with the rays formed inside
the trace kernel is camera_fused_mrays_s. lbvh_mtri_s times rk_fast_build
(30-bit Morton, one face per leaf). The reference tree (rk_build_lbvh, depth
12) and both reference traversals are timed once on the same frame, and so
are the reference's 32-bin SAH build (rk_build_sah), its per-ray traversal of
the SAH tree, and the Hybrid build (depth-12 LBVH, each leaf refined by SAH).

Further legs, each skippable except the first:
- icosphere: the reference's CPU-runnable case (20,480 faces, 256x256 rays):
  reference build and both reference traversals, the fast tree, the oracle.
- soup: 10M random faces, 63-bit keys, 2^24 incoherent rays; build and
  frames with the rays as given and sorted by origin (rk_fast_trace_sorted).
- batch: 4096 soups built and traced together (autotuned, one captured graph
  per step) against one rk_fast_build + rk_fast_trace per mesh.
- serve: the rk_server request loop over 32 deforming scenes, Poisson
  arrivals; p50/p99 latency, and saturated throughput against the same loop
  on one slot and against the plain per-request path (rebuild, ray arrays,
  trace, checksum, one request at a time).
- cpu_baseline: the oracle, i.e. the reference algorithm (its LBVH build and
  per-ray traversal, reference threshold) on the host's cores, on a bounded
  sample of the frame; each leg carries its own cpu_baseline of the same kind.

python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu] [--no-ref]
                [--no-soup] [--no-batch] [--no-serve] [--profile-only | --profile-soup]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

# Each serving slot is a HIP stream; streams only run concurrently on distinct
# hardware queues, and HIP opens 4 per process unless told otherwise. Raised
# to 16 (one per serving slot) before the first HIP call.
if int(os.environ.get("GPU_MAX_HW_QUEUES", "4")) < 16:
    os.environ["GPU_MAX_HW_QUEUES"] = "16"

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "Mrays/s closest-hit + LBVH Mtri/s, batched vs plain on 1 GPU; p99 latency"
FP32_PEAK_TFLOPS = 157.3
OPS_PER_BOX_TEST = 24   # algorithmic FP32 ops of one slab test
OPS_PER_TRI_TEST = 53   # algorithmic FP32 ops of one Moller-Trumbore test
HEADLINE_DET_EPSILON = 1e-6  # the reference's absolute parallel-ray threshold (CUDA_raytraceMeshTree.cu:36)
# HBM bytes per launch of the headline trace kernel (the default traversal
# variant, tmax reset in the kernel), from two rocprofv3 --pmc runs of
# `bench.py --profile-only` (profiles/fast_build_trace_heightfield708.md):
# FETCH_SIZE 107,351 KiB, scaled x2 by a 256 MiB copy, plus WRITE_SIZE 10,576 KiB
# (means of the two warm frames at the 1e-6 threshold).
TRACE_TRAFFIC_BYTES = (107_351 * 2 + 10_576) * 1024


# The fast build's kernels, per launch, measured with rocprofv3 on the headline
# heightfield (`bench.py --profile-only`) and on the 10M soup (`--profile-soup`):
# average duration (us) from --kernel-trace, HBM bytes read and written (KiB)
# from --pmc FETCH_SIZE (x2, the factor a 256 MiB copy gives) and WRITE_SIZE
# (profiles/fast_build_trace_heightfield708.md, profiles/soup10m_build_trace.md).
# The heightfield's climbs through the splits that cross chunks run inside
# k_fast_hierarchy (up to 2^21 faces); the soup's in k_fast_hierarchy_top.
BUILD_KERNELS = {  # kernel: (us, HBM read KiB, HBM write KiB, launches per build)
    "heightfield": {
        "k_fast_prep": (15.12, 32_289, 42_960, 1),
        "k_rs_hist": (12.20, 3_992, 1_285, 1),
        "k_rs_scatter": (16.46, 10_937, 10_366, 4),
        "k_fast_hierarchy": (109.70, 51_675, 119_632, 1),
    },
    "soup": {
        "k_fast_prep": (211.5, 625_591, 468_775, 1),
        "k_rs_hist": (66.08, 78_298, 18_713, 1),
        "k_rs_scatter": (60.33, 132_323, 134_816, 8),
        "k_fast_hierarchy": (854.0, 2_928_171, 1_192_198, 1),
        "k_fast_hierarchy_top": (126.2, 51_644, 38_306, 1),
    },
}
HBM_PEAK_GBS = 8000.0


def build_algorithmic_bytes(n_faces, n_verts, key_bytes):
    """Compulsory HBM bytes of one rk_fast_build with one face per leaf: the
    key kernel reads faces (16 B) and vertices and writes boxes (32), keys and
    indices (4); the sort reads the keys once for its counts and each of its
    passes reads and writes key and index; the hierarchy reads keys, indices,
    faces and vertices and writes a 64-B record per split and a 48-B triangle
    record per face. Vertices are counted once per kernel that reads them."""
    passes = key_bytes
    per_face = (16 + 32 + key_bytes + 4) + key_bytes + passes * 2 * (key_bytes + 4) + (key_bytes + 4 + 16 + 64 + 48)
    return per_face * n_faces + 32 * n_verts


def build_roofline(workload, n_faces, n_verts, key_bytes, build_ms):
    alg = build_algorithmic_bytes(n_faces, n_verts, key_bytes)
    achieved = alg / (build_ms * 1e-3) / 1e9
    kernels = BUILD_KERNELS.get(workload, {})
    per = {}
    traffic = 0
    for k, (us, rd_kib, wr_kib, launches) in kernels.items():
        b = (rd_kib + wr_kib) * 1024
        traffic += b * launches
        per[k] = {"us": us, "launches": launches, "hbm_mb": round(b / 1e6, 1),
                  "hbm_frac": round(b / (us * 1e-6) / 1e9 / HBM_PEAK_GBS, 3)}
    return {"bound": "hbm", "kernel": "rk_fast_build (all its kernels)", "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic or None, "algorithmic_bytes": alg, "per_kernel": per,
            "note": "achieved = compulsory bytes of the build (build_algorithmic_bytes) / its time here; "
                    "per_kernel = measured HBM bytes per launch / average launch time / peak, from the "
                    "rocprofv3 runs cited at BUILD_KERNELS"}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def heightfield_scene(native, width=1920, height=1080, grid=708):
    V, F = native.heightfield(grid, seed=0xC2, amplitude=0.2)
    hb = native.hitbox(V)
    el, az, dist = np.radians(30.0), np.radians(45.0), 1.5
    tgt = np.array([0.5, 0.5, 0.1])
    eye = tgt + dist * np.array([np.cos(el) * np.cos(az), np.cos(el) * np.sin(az), np.sin(el)])
    cam = native.camera_lookat(eye, tgt, 60 * 0.01745329252, width, height, up=[0, 0, 1])
    O, D = native.primary_rays_host(cam, width, height)
    return V, F, hb, O, D, cam


class Timer:
    """HIP events on torch's current stream, the stream every rk_* call here
    is issued on (the default stream)."""

    def __init__(self, torch):
        self.torch = torch

    def time(self, fn, reps):
        t = self.torch
        s, e = t.cuda.Event(enable_timing=True), t.cuda.Event(enable_timing=True)
        t.cuda.synchronize()
        s.record()
        for _ in range(reps):
            fn()
        e.record()
        t.cuda.synchronize()
        return s.elapsed_time(e) / reps  # ms


def cpu_trace_sampled(oracle, ref, V, O, D, budget_s):
    """Per-ray reference traversal (the oracle) of every stride-th ray, the
    stride chosen so the timed pass takes about budget_s / 2 seconds."""
    n = O.shape[0]
    stride = max(1, n // 1024)
    t0 = time.perf_counter()
    oracle.trace_ref_ray(ref.nodes, ref.faces, V, np.ascontiguousarray(O[::stride]), np.ascontiguousarray(D[::stride]))
    dt = time.perf_counter() - t0
    if dt < budget_s / 4 and stride > 1:
        stride = max(1, int(stride * dt / (budget_s / 2)))
    O2, D2 = np.ascontiguousarray(O[::stride]), np.ascontiguousarray(D[::stride])
    t0 = time.perf_counter()
    oracle.trace_ref_ray(ref.nodes, ref.faces, V, O2, D2)
    return O2.shape[0], time.perf_counter() - t0, stride


def cpu_baseline(V, F, hb, O, D, budget_s=20.0):
    """The reference algorithm on the host (the oracle): LBVH build (depth 12)
    and per-ray traversal with the reference threshold, on a sample of rays."""
    import oracle
    threads = oracle.num_threads()
    t0 = time.perf_counter()
    ref = oracle.build_lbvh_ref(F, V, hb, 12, True)
    build_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    oracle.build_lbvh_ref(F, V, hb, 31, True)  # the deepest tree the reference builds (tree_depth 31)
    build31_s = time.perf_counter() - t0
    n = O.shape[0]
    stride = 64
    t0 = time.perf_counter()
    oracle.trace_ref_ray(ref.nodes, ref.faces, V, np.ascontiguousarray(O[::stride]), np.ascontiguousarray(D[::stride]))
    dt = time.perf_counter() - t0
    rays = (n + stride - 1) // stride
    if dt < budget_s / 4:
        stride2 = max(1, int(stride * dt / (budget_s / 2)))
        O2, D2 = np.ascontiguousarray(O[::stride2]), np.ascontiguousarray(D[::stride2])
        t0 = time.perf_counter()
        oracle.trace_ref_ray(ref.nodes, ref.faces, V, O2, D2)
        dt = time.perf_counter() - t0
        rays, stride = O2.shape[0], stride2
    return {
        "value": round(rays / dt / 1e6, 4),
        "unit": "Mrays/s",
        "cores": int(threads),
        "kind": "port",
        "sample": (f"every {stride}th ray of the 1920x1080 frame ({rays} rays)" if stride > 1 else
                   f"all {rays} rays of the 1920x1080 frame") + ": reference LBVH (depth 12) and "
                  f"per-ray traversal (CUDA_raytraceMeshTree.cu:111-219) with the reference threshold; the "
                  f"reference build of all {F.shape[0]} faces took {build_s:.3f} s",
        "lbvh_mtri_s": round(F.shape[0] / build_s / 1e6, 4),
        "lbvh_d31_mtri_s": round(F.shape[0] / build31_s / 1e6, 4),
        "host_cores_visible": os.cpu_count(),
    }


def icosphere_leg(torch, dev, native, timer, steps, warmup, with_cpu=False):
    """The reference's own CPU-runnable case: icosphere (5 subdivisions,
    20,480 faces, unit radius), the reference LBVH (depth 12, 32-bit Morton),
    256 x 256 primary rays from (0, 0, 3) at the origin with a 60 degree fov,
    closest hit at the reference threshold. A step is one frame: clear the
    outputs, trace. Timed: the reference build and both reference traversals,
    the fast build and trace, and the oracle on the host."""
    V, F = native.icosphere(5)
    hb = native.hitbox(V)
    W = H = 256
    cam = native.camera_lookat([0, 0, 3], [0, 0, 0], 60 * 0.01745329252, W, H)
    O, D = native.primary_rays_host(cam, W, H)
    n, nr = F.shape[0], O.shape[0]
    Ft, Vt = dev.to_u32_tensor(F), torch.from_numpy(V).cuda()
    Ot, Dt = torch.from_numpy(O).cuda(), torch.from_numpy(D).cuda()
    f32max = float(np.finfo(np.float32).max)
    rb = dev.LBVHBuilder(n)
    Fr = Ft.clone()
    rtree = rb.build(Fr, Vt, hb, 12, True)
    torch.cuda.synchronize()
    rtree.node_count = int(rtree.host_info()[native.RK_INFO_NODE_COUNT])
    ref_build = lambda: rb.build(Fr, Vt, hb, 12, True)
    out = dev.new_outputs(nr)

    def ref_frame(coherency):
        native.check(native.lib().rk_fill_f32(None, out[2].data_ptr(), nr, f32max), "fill")
        dev.trace_ref(rtree, Vt, Ot, Dt, W, H, coherency, out=out)

    for _ in range(warmup):
        ref_build()
    ref_build_ms = timer.time(ref_build, steps)
    rtree = rb.build(Ft.clone(), Vt, hb, 12, True)  # a build of the faces in mesh order, as traced below
    torch.cuda.synchronize()
    rtree.node_count = int(rtree.host_info()[native.RK_INFO_NODE_COUNT])
    for _ in range(warmup):
        ref_frame(True)
    pkt_ms = timer.time(lambda: ref_frame(True), steps)
    ref_dep = out[2].clone()
    ray_ms = timer.time(lambda: ref_frame(False), steps)
    builder = dev.FastBuilder(n, native.RK_MORTON_30)
    fbuild = lambda: builder.build(Ft, Vt, hb, 1)
    tree = fbuild()
    fout = dev.new_outputs(nr)

    def fast_frame(config=0):
        dev.fast_trace(tree, Ot, Dt, out=fout, det_epsilon=HEADLINE_DET_EPSILON, reset_depth=True, config=config)

    for _ in range(warmup):
        fbuild()
    fbuild_ms = timer.time(fbuild, steps)
    tree = fbuild()
    for _ in range(warmup):
        fast_frame()
    fast_default_ms = timer.time(fast_frame, steps)
    # the traversal variant for this frame, as for the headline frame
    tune = dev.autotune(tree, Ot, Dt, fout, det_epsilon=HEADLINE_DET_EPSILON, reps=25, reset_depth=True)
    for _ in range(warmup):
        fast_frame(tune.config)
    fast_ms = timer.time(lambda: fast_frame(tune.config), steps)
    res = {
        "workload": "icosphere 5 subdivisions (20,480 faces), 256x256 primary rays, closest hit, threshold 1e-6",
        "faces": n, "rays": nr,
        "reference_lbvh_mtri_s": round(n / (ref_build_ms * 1e-3) / 1e6, 2),
        "reference_packet_mrays_s": round(nr / (pkt_ms * 1e-3) / 1e6, 2),
        "reference_per_ray_mrays_s": round(nr / (ray_ms * 1e-3) / 1e6, 2),
        "fast_lbvh_mtri_s": round(n / (fbuild_ms * 1e-3) / 1e6, 2),
        "fast_mrays_s": round(nr / (fast_ms * 1e-3) / 1e6, 2),
        "fast_frame_ms": round(fast_ms, 4),
        "fast_autotuned": {"config": tune.config, "workgroup": tune.workgroup, "ray_run": tune.ray_run,
                           "donate_min": tune.donate_min, "bound_every": tune.bound_every},
        "fast_mrays_s_default_variant": round(nr / (fast_default_ms * 1e-3) / 1e6, 2),
        "same_depths": bool(torch.equal(fout[2], ref_dep)),
        "hit_fraction": round(int((ref_dep < f32max).sum().item()) / nr, 4),
        "cpu_baseline": None,
    }
    if with_cpu:
        import oracle
        bt, tt = [], []
        for _ in range(5):
            t0 = time.perf_counter()
            ref = oracle.build_lbvh_ref(F, V, hb, 12, True)
            bt.append(time.perf_counter() - t0)
            t0 = time.perf_counter()
            oracle.trace_ref_ray(ref.nodes, ref.faces, V, O, D)
            tt.append(time.perf_counter() - t0)
        b, t = float(np.median(bt)), float(np.median(tt))
        res["cpu_baseline"] = {
            "value": round(nr / t / 1e6, 4), "unit": "Mrays/s", "cores": int(oracle.num_threads()), "kind": "port",
            "lbvh_mtri_s": round(n / b / 1e6, 3),
            "sample": "the whole workload, median of 5: reference LBVH build (depth 12) and per-ray traversal of "
                      "all 65,536 rays"}
    return res


def soup_leg(torch, dev, native, timer, steps, warmup, with_cpu=False):
    """Random soup of 10M faces (centroids uniform in the unit cube, lognormal
    edges around 0.002), 63-bit Morton keys sorted as uint64 pairs, and 2^24
    incoherent rays (origins uniform in the cube, directions uniform on the
    sphere); one closest-hit frame of all rays per step."""
    n, nr = 10_000_000, 1 << 24
    V, F = native.soup(n, seed=0xC3)
    hb = native.hitbox(V)
    O, D = native.random_rays(nr, seed=0xC3)
    Ft, Vt = dev.to_u32_tensor(F), torch.from_numpy(V).cuda()
    Ot, Dt = torch.from_numpy(O).cuda(), torch.from_numpy(D).cuda()
    cpu = None
    if with_cpu:
        import oracle
        t0 = time.perf_counter()
        ref = oracle.build_lbvh_ref(F, V, hb, 12, True)
        cpu_build_s = time.perf_counter() - t0
        rays, dt, stride = cpu_trace_sampled(oracle, ref, V, O, D, 10.0)
        cpu = {"value": round(rays / dt / 1e6, 4), "unit": "Mrays/s", "cores": int(oracle.num_threads()),
               "kind": "port", "lbvh_mtri_s": round(n / cpu_build_s / 1e6, 3),
               "sample": f"reference LBVH (depth 12, 32-bit Morton) of all {n} faces and per-ray traversal of "
                         f"every {stride}th ray ({rays} rays), reference threshold"}
        del ref
    del V, F, O, D
    builder = dev.FastBuilder(n, native.RK_MORTON_63)
    build = lambda: builder.build(Ft, Vt, hb, 1)
    tree = build()
    fid, uv, dep = dev.new_outputs(nr)
    f32max = float(np.finfo(np.float32).max)

    tuned = {"config": 0}

    def frame():
        dev.fast_trace(tree, Ot, Dt, out=(fid, uv, dep), det_epsilon=HEADLINE_DET_EPSILON, config=tuned["config"],
                       reset_depth=True)

    for _ in range(warmup):
        build()
    build_ms = timer.time(build, steps)
    for _ in range(warmup):
        frame()
    default_ms = timer.time(frame, steps)
    tuning = dev.autotune(tree, Ot, Dt, (fid, uv, dep), det_epsilon=HEADLINE_DET_EPSILON, reps=1, reset_depth=True)
    tuned["config"] = int(tuning.config)
    for _ in range(warmup):
        frame()
    frame_ms = timer.time(frame, steps)
    hits = int((dep < f32max).sum().item())
    ref = (fid.clone(), dep.clone())
    # the same frame with the rays traced in origin-Morton order (key kernel +
    # 30-bit radix sort + permuted trace, all inside the timed step)
    sws = dev.workspace(native.lib().rk_fast_trace_sorted_workspace_size(nr))

    def frame_sorted():
        dev.fast_trace_sorted(tree, Ot, Dt, out=(fid, uv, dep), det_epsilon=HEADLINE_DET_EPSILON, ws=sws,
                              config=tuned["config"], reset_depth=True)

    for _ in range(warmup):
        frame_sorted()
    sorted_ms = timer.time(frame_sorted, steps)
    same = bool(torch.equal(fid, ref[0]) and torch.equal(dep, ref[1]))
    return {
        "workload": "soup 10M faces, 63-bit Morton, one face per leaf, 2^24 incoherent rays, closest hit",
        "lbvh63_mtri_s": round(n / (build_ms * 1e-3) / 1e6, 2),
        "build_ms": round(build_ms, 3),
        "build_roofline": build_roofline("soup", n, 3 * n, 8, build_ms),
        "mrays_s": round(nr / (frame_ms * 1e-3) / 1e6, 2),
        "frame_ms": round(frame_ms, 3),
        "trace_autotuned": tuning.as_dict(),
        "mrays_s_default_variant": round(nr / (default_ms * 1e-3) / 1e6, 2),
        "hit_fraction": round(hits / nr, 4),
        "sorted_rays_mrays_s": round(nr / (sorted_ms * 1e-3) / 1e6, 2),
        "sorted_rays_frame_ms": round(sorted_ms, 3),
        "sorted_rays_same_results": same,
        "cpu_baseline": cpu,
    }


def frame_loop_leg(torch, dev, native, timer, steps, warmup, tree, cam, W, H, F, V, Vt, cam_config, rtree, Ot, Dt,
                   with_cpu=False):
    """The frame of samples/02_how_to_use_the_raytracer.cpp:67-85 on the
    heightfield: primary rays, closest hit, the hit face's normal
    (rk_interpolate_data) and the shaded pixel (rk_shade). Fast tree: the camera
    trace kernel, then the two kernels, or all of it in rk_fast_render_camera;
    reference tree: rk_calc_primary_rays, a fill, the packet traversal, the two
    kernels. Face normals (the heightfield has no vertex normals)."""
    log("frame loop: rays, trace, normals, shading")
    nr = W * H
    order = dev.u32_view(tree.indices)
    Fs = dev.to_u32_tensor(np.ascontiguousarray(F[order]))  # faces in leaf order, as the builds leave the mesh
    c = np.asarray(cam, np.float64)
    look = c[1, :3] - c[2, :3] * (W / 2.0) - c[3, :3] * (H / 2.0)  # the centre pixel's ray
    look = (look / np.linalg.norm(look)).astype(np.float32)
    back = (58 / 255.0, 110 / 255.0, 165 / 255.0, 1.0)
    fid, uv, dep = dev.new_outputs(nr)
    nrm = torch.zeros((nr, 4), dtype=torch.float32, device="cuda")
    px = torch.zeros(nr, dtype=torch.int32, device="cuda")
    px2 = torch.zeros(nr, dtype=torch.int32, device="cuda")
    outs = (None, nrm, None, None, None)
    trace = lambda: dev.fast_trace_camera(tree, cam, W, H, out=(fid, uv, dep), det_epsilon=HEADLINE_DET_EPSILON,
                                          config=cam_config)
    interp = lambda: dev.interpolate(fid, uv, W, H, Fs, Vt, outs=outs)
    shade = lambda: dev.shade(dep, nrm, W, H, back, look, pixels=px)

    def sequence():
        trace()
        interp()
        shade()

    fused = lambda: dev.render_camera(tree, cam, W, H, Fs, Vt, back, look, pixels=px2,
                                      det_epsilon=HEADLINE_DET_EPSILON)
    for _ in range(warmup):
        sequence()
        fused()
    seq_ms = timer.time(sequence, steps)
    fused_ms = timer.time(fused, steps)
    trace_ms = timer.time(trace, steps)
    interp_ms = timer.time(interp, steps)
    shade_ms = timer.time(shade, steps)
    sequence()
    fused()
    same = bool(torch.equal(px, px2))
    # the reference tree's frame
    O, D = torch.empty_like(Ot), torch.empty_like(Dt)
    cr = dev.camera_rays(cam)
    Fr = rtree.faces
    rfid, ruv, rdep = dev.new_outputs(nr)
    rnrm = torch.zeros((nr, 4), dtype=torch.float32, device="cuda")
    rpx = torch.zeros(nr, dtype=torch.int32, device="cuda")
    f32max = float(np.finfo(np.float32).max)

    def ref_frame():
        native.check(native.lib().rk_calc_primary_rays(None, W, H, cr.origin, cr.direction, cr.step_x, cr.step_y,
                                                       O.data_ptr(), D.data_ptr()), "primary rays")
        native.check(native.lib().rk_fill_f32(None, rdep.data_ptr(), nr, f32max), "fill")
        dev.trace_ref(rtree, Vt, O, D, W, H, True, out=(rfid, ruv, rdep))
        dev.interpolate(rfid, ruv, W, H, Fr, Vt, outs=(None, rnrm, None, None, None))
        dev.shade(rdep, rnrm, W, H, back, look, pixels=rpx)

    ref_frame()
    ref_ms = timer.time(ref_frame, 2)
    res = {
        "workload": "heightfield 708x708, 1920x1080: primary rays + closest hit + hit normal + shaded pixel "
                    "(samples/02 frame loop), threshold 1e-6",
        "fast_sequence_ms": round(seq_ms, 4),
        "fast_sequence_frames_s": round(1e3 / seq_ms, 1),
        "fused_ms": round(fused_ms, 4),
        "fused_frames_s": round(1e3 / fused_ms, 1),
        "fused_same_pixels": same,
        "stage_ms": {"camera_trace": round(trace_ms, 4), "interpolate_normals": round(interp_ms, 4),
                     "shade": round(shade_ms, 4)},
        "reference_packet_frame_ms": round(ref_ms, 3),
    }
    if with_cpu:
        import oracle
        fh, uh, dh = (x.cpu().numpy() for x in (fid, uv, dep))
        fh = fh.view(np.uint32)
        Fsh = np.ascontiguousarray(F[order])
        t0 = time.perf_counter()
        on = oracle.interpolate(fh, uh, Fsh, V)[1]
        opx = oracle.shade(dh, on, W, H, np.array(back, np.float32), look)
        dt = time.perf_counter() - t0
        res["cpu_baseline"] = {"value": round(dt * 1e3, 2), "unit": "ms per frame (normals + shading only)",
                               "cores": 1, "kind": "port",
                               "sample": "the whole frame's hit normals and pixels from the GPU trace's outputs",
                               "same_pixels": bool(np.array_equal(opx, px.cpu().numpy().view(np.uint32)))}
    return res


def batch_leg(torch, dev, native, timer, steps, warmup, with_cpu=False):
    """4096 soups (log-uniform 64..65536 faces, 256..16384 rays each): one
    step builds every mesh and traces every ray. Batched = the autotuned
    rk_batch_build + rk_batch_trace captured in one graph; plain = one
    rk_fast_build + rk_fast_trace per mesh, in a loop on one stream."""
    from cudart_amd import scenes
    table, F, V, hb, O, D = scenes.soup_batch()
    nf, nr, m = F.shape[0], O.shape[0], table.shape[0]
    Ft, Vt = dev.to_u32_tensor(F), torch.from_numpy(V).cuda()
    Ot, Dt = torch.from_numpy(O).cuda(), torch.from_numpy(D).cuda()
    b = dev.Batch(table, Ft, Vt, hb, Ot, Dt)
    out = dev.new_outputs(nr)
    f32max = float(np.finfo(np.float32).max)
    config, tuned_ms = b.autotune(out, reps=2, det_epsilon=HEADLINE_DET_EPSILON)
    s = torch.cuda.Stream()
    g = dev.Graph(s)
    torch.cuda.synchronize()
    with torch.cuda.stream(s):
        with g.capture():
            native.check(native.lib().rk_fill_f32(dev._stream_ptr(s), out[2].data_ptr(), nr, f32max), "fill")
            b.build(stream=s)
            b.trace(out, det_epsilon=HEADLINE_DET_EPSILON, stream=s)

    def batched():
        g.launch()

    ev = (torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))

    def time_on(stream, fn, reps):
        torch.cuda.synchronize()
        ev[0].record(stream)
        for _ in range(reps):
            fn()
        ev[1].record(stream)
        torch.cuda.synchronize()
        return ev[0].elapsed_time(ev[1]) / reps

    for _ in range(warmup):
        batched()
    batch_ms = time_on(s, batched, steps)
    hits = int((out[2] < f32max).sum().item())

    builder = dev.FastBuilder(int(table["face_count"].max()), native.RK_MORTON_30)
    pout = dev.new_outputs(nr)
    views = []
    for k in range(m):
        f0, fc = int(table["face_offset"][k]), int(table["face_count"][k])
        v0, vc = int(table["vertex_offset"][k]), int(table["vertex_count"][k])
        r0, rc = int(table["ray_offset"][k]), int(table["ray_count"][k])
        views.append((Ft[f0:f0 + fc], Vt[v0:v0 + vc], hb[k], Ot[r0:r0 + rc], Dt[r0:r0 + rc],
                      tuple(x[r0:r0 + rc] for x in pout), rc))

    def plain():
        for Fm, Vm, hbm, Om, Dm, outm, rc in views:
            tree = builder.build(Fm, Vm, hbm, 4)
            native.check(native.lib().rk_fill_f32(None, outm[2].data_ptr(), rc, f32max), "fill")
            dev.fast_trace(tree, Om, Dm, out=outm, det_epsilon=HEADLINE_DET_EPSILON)

    plain()
    plain_ms = timer.time(plain, 2)
    same = bool(torch.equal(pout[2], out[2]))
    cpu = None
    if with_cpu:
        # the reference per mesh on the host: LBVH build and per-ray traversal of
        # every ray of every 64th mesh
        import oracle
        rays = faces = 0
        t0 = time.perf_counter()
        for k in range(0, m, 64):
            f0, fc = int(table["face_offset"][k]), int(table["face_count"][k])
            v0, vc = int(table["vertex_offset"][k]), int(table["vertex_count"][k])
            r0, rc = int(table["ray_offset"][k]), int(table["ray_count"][k])
            Fm = np.ascontiguousarray(F[f0:f0 + fc])
            Vm = np.ascontiguousarray(V[v0:v0 + vc])
            ref = oracle.build_lbvh_ref(Fm, Vm, hb[k], 12, True)
            oracle.trace_ref_ray(ref.nodes, ref.faces, Vm, np.ascontiguousarray(O[r0:r0 + rc]),
                                 np.ascontiguousarray(D[r0:r0 + rc]))
            rays += rc
            faces += fc
        dt = time.perf_counter() - t0
        cpu = {"value": round(rays / dt / 1e6, 4), "unit": "Mrays/s", "cores": int(oracle.num_threads()),
               "kind": "port", "mtri_s": round(faces / dt / 1e6, 3),
               "sample": f"every 64th mesh ({len(range(0, m, 64))} meshes, {faces} faces, {rays} rays): reference "
                         f"LBVH build (depth 12) and per-ray traversal, build and trace together"}
    return {
        "workload": "4096 soups, log-uniform 64..65536 faces and 256..16384 rays per mesh; build all + trace all",
        "meshes": m, "faces": nf, "rays": nr,
        "batched_ms": round(batch_ms, 3),
        "batched_mrays_s": round(nr / (batch_ms * 1e-3) / 1e6, 2),
        "batched_mtri_s": round(nf / (batch_ms * 1e-3) / 1e6, 2),
        "plain_ms": round(plain_ms, 3),
        "speedup_vs_plain": round(plain_ms / batch_ms, 2),
        "same_depths_as_plain": same,
        "hit_fraction": round(hits / nr, 4),
        "autotuned": config.as_dict(),
        "autotune_best_ms": round(tuned_ms, 3),
        "graph": "fill + rk_batch_build + rk_batch_trace captured once, replayed per step",
        "cpu_baseline": cpu,
    }


def serve_leg(dev, native, budget=200 << 30, rate=2000.0, count=4000, with_cpu=False):
    """Serving: 32 deforming scenes (log-uniform 1K..2M faces, alternately a
    terrain and a soup) behind rk_server; Poisson arrivals at 'rate' per
    second, square frames of log-uniform side 64..1024, refit or rebuild per
    request. Latency p50/p99 is arrival to observed completion. Saturation:
    all requests queued at once, 16 slots (continuous batching) against the
    same loop with 1 slot (one request at a time), and against the plain
    per-request path: each request through the plain entry points (rebuild
    from scratch, primary rays into arrays, trace, checksum), one at a time."""
    import math
    from cudart_amd import scenes
    rng = scenes.splitmix64(0xC5)
    counts = scenes.log_uniform_counts(32, 1000, 2_000_000, rng)
    pool = []
    for k, n in enumerate(counts):
        if k % 2 == 0:
            V, F = native.heightfield(int(math.sqrt(int(n) / 2)) + 1, seed=0xC5 + k)
            pool.append((V, F, True))
        else:
            V, F = native.soup(int(n), seed=0xC5 + k, mu=float(math.log(0.002 * math.sqrt(1e7 / int(n)))))
            pool.append((V, F, False))
    need = sum(native.lib().rk_server_scene_bytes(V.shape[0], F.shape[0]) for V, F, _ in pool)
    arena = min(budget, int(need * 1.1) + (256 << 20))

    def server(slots):
        srv = dev.Server(arena, slots=slots, hbm_budget=budget, det_epsilon=HEADLINE_DET_EPSILON)
        for V, F, z in pool:
            srv.add_scene(V, F, z)
        srv.run_poisson(0.0, 64, 64, 1024, 1)  # upload and build every scene once
        return srv

    srv = server(16)
    st, _ = srv.run_poisson(rate, count, 64, 1024, 0xC5)
    sat, _ = srv.run_poisson(0.0, 1000, 64, 1024, 2)
    srv.close()
    # the same stream under an HBM budget of half the scenes: requests of
    # evicted scenes upload them again from pinned host memory and rebuild
    sizes = [native.lib().rk_server_scene_bytes(V.shape[0], F.shape[0]) for V, F, _ in pool]
    tight = max(max(sizes), need // 2)
    srvt = dev.Server(tight, slots=16, hbm_budget=tight, det_epsilon=HEADLINE_DET_EPSILON)
    for V, F, z in pool:
        srvt.add_scene(V, F, z)
    tst, _ = srvt.run_poisson(rate, count // 2, 64, 1024, 0xC6)
    srvt.close()
    one = server(1)
    single, _ = one.run_poisson(0.0, 1000, 64, 1024, 2)
    one.close()
    # the same 1000 requests through the plain entry points, one at a time
    # (rebuild from scratch, ray arrays, trace, checksum, wait)
    srvp = dev.Server(arena, slots=1, hbm_budget=budget, det_epsilon=HEADLINE_DET_EPSILON, plain=1)
    for V, F, z in pool:
        srvp.add_scene(V, F, z)
    srvp.run_poisson(0.0, 64, 64, 1024, 1)
    plain, _ = srvp.run_poisson(0.0, 1000, 64, 1024, 2)
    srvp.close()
    cpu = None
    if with_cpu:
        # the reference per request on the host: rebuild the scene's LBVH
        # (depth 12) and trace its camera frame per ray, for 12 requests drawn
        # like the stream's (scene uniform, side log-uniform in 64..1024)
        import oracle
        r = np.random.default_rng(0xC5)
        rays = 0
        t0 = time.perf_counter()
        for _ in range(12):
            V, F, z = pool[int(r.integers(0, len(pool)))]
            side = int(round(math.exp(r.uniform(math.log(64), math.log(1024)))))
            hbs = native.hitbox(V)
            ctr = 0.5 * (hbs[0, :3] + hbs[1, :3])
            diag = float(np.linalg.norm(hbs[1, :3] - hbs[0, :3]))
            el, az = math.radians(30.0), r.uniform(0, 2 * math.pi)
            eye = ctr + 1.2 * diag * np.array([math.cos(el) * math.cos(az), math.cos(el) * math.sin(az), math.sin(el)])
            Oc, Dc = native.primary_rays_host(native.camera_lookat(eye, ctr, math.radians(60.0), side, side,
                                                                   up=[0, 0, 1]), side, side)
            ref = oracle.build_lbvh_ref(F, V, hbs, 12, True)
            oracle.trace_ref_ray(ref.nodes, ref.faces, V, Oc, Dc)
            rays += side * side
        dt = time.perf_counter() - t0
        cpu = {"value": round(rays / dt / 1e6, 4), "unit": "Mrays/s", "cores": int(oracle.num_threads()),
               "kind": "port", "ms_per_request": round(dt / 12 * 1e3, 2),
               "sample": f"12 requests drawn like the stream's ({rays} rays): per request, the reference LBVH "
                         f"rebuild (depth 12) and per-ray traversal of the frame, one request at a time"}
    return {
        "workload": "32 deforming scenes (1K..2M faces, terrain/soup), Poisson arrivals, frames 64^2..1024^2",
        "faces_total": int(sum(F.shape[0] for _, F, _ in pool)),
        "rate_req_s": rate, "requests": int(st.completed),
        "p50_ms": round(st.p50_ms, 3), "p99_ms": round(st.p99_ms, 3), "mean_ms": round(st.mean_ms, 3),
        "max_ms": round(st.max_ms, 3),
        "mrays_s": round(st.mrays_per_s, 2),
        "rebuilds": int(st.rebuilds), "refits": int(st.refits),
        "uploads": int(st.uploads), "evictions": int(st.evictions),
        "resident_gb": round(st.peak_resident_bytes / 2 ** 30, 3), "budget_gb": budget / 2 ** 30,
        "slots": 16, "hw_queues": int(os.environ.get("GPU_MAX_HW_QUEUES", "4")),
        "evicting_budget": {"budget_gb": round(tight / 2 ** 30, 3), "requests": int(tst.completed),
                            "p50_ms": round(tst.p50_ms, 3), "p99_ms": round(tst.p99_ms, 3),
                            "max_ms": round(tst.max_ms, 3), "mrays_s": round(tst.mrays_per_s, 2),
                            "uploads": int(tst.uploads), "evictions": int(tst.evictions),
                            "rebuilds": int(tst.rebuilds), "refits": int(tst.refits),
                            "resident_gb": round(tst.peak_resident_bytes / 2 ** 30, 3)},
        "saturated_mrays_s": round(sat.mrays_per_s, 2),
        "one_slot_mrays_s": round(single.mrays_per_s, 2),
        "served_vs_one_at_a_time": round(sat.mrays_per_s / single.mrays_per_s, 2) if single.mrays_per_s else None,
        "plain_per_request_mrays_s": round(plain.mrays_per_s, 2),
        "served_vs_plain": round(sat.mrays_per_s / plain.mrays_per_s, 2) if plain.mrays_per_s else None,
        "failed_requests": int(st.failed + sat.failed),
        "cpu_baseline": cpu,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-ref", action="store_true")
    ap.add_argument("--no-soup", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--leaf", type=int, default=1)
    ap.add_argument("--profile-only", action="store_true", help="a few builds and frames, no JSON (for rocprofv3)")
    ap.add_argument("--profile-soup", action="store_true",
                    help="the 10M-face soup: two builds and two frames, no JSON (for rocprofv3)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (each GPU would trace its own "
              "share of rays or meshes with no exchange); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    from cudart_amd import device as dev
    from cudart_amd import native

    torch.cuda.set_device(0)
    timer = Timer(torch)
    if args.profile_soup:
        soup_leg(torch, dev, native, timer, 2, 0)
        torch.cuda.synchronize()
        log("profile run done")
        return
    W, H = 1920, 1080
    log("generating scene")
    V, F, hb, O, D, cam = heightfield_scene(native, W, H)
    n, nr = F.shape[0], O.shape[0]
    Ft, Vt = dev.to_u32_tensor(F), torch.from_numpy(V).cuda()
    Ot, Dt = torch.from_numpy(O).cuda(), torch.from_numpy(D).cuda()
    f32max = float(np.finfo(np.float32).max)

    builder = dev.FastBuilder(n, native.RK_MORTON_30)
    build = lambda: builder.build(Ft, Vt, hb, args.leaf)
    tree = build()
    fid, uv, dep = dev.new_outputs(nr)
    tuned = {"config": 0}

    def frame(eps=HEADLINE_DET_EPSILON, config=None):
        # tmax reset to FLT_MAX and trace in one launch (RK_TRACE_RESET_DEPTH)
        dev.fast_trace(tree, Ot, Dt, out=(fid, uv, dep), det_epsilon=eps,
                       config=tuned["config"] if config is None else config, reset_depth=True)

    def frame_fill_then_trace():
        native.check(native.lib().rk_fill_f32(None, dep.data_ptr(), nr, f32max), "fill")
        dev.fast_trace(tree, Ot, Dt, out=(fid, uv, dep), det_epsilon=HEADLINE_DET_EPSILON, config=tuned["config"])

    torch.cuda.synchronize()
    if args.profile_only:
        for _ in range(3):
            build()
        for _ in range(3):  # the default traversal variant (the tuned ones are within 1-2% of it on this frame)
            frame()
        frame(0.0)
        rb = dev.LBVHBuilder(n)
        rtree = rb.build(Ft.clone(), Vt, hb, 12, True)
        torch.cuda.synchronize()
        rtree.node_count = int(rtree.host_info()[native.RK_INFO_NODE_COUNT])
        dev.trace_ref(rtree, Vt, Ot, Dt, W, H, False)
        dev.trace_ref(rtree, Vt, Ot, Dt, W, H, True)  # the packet traversal too
        a = torch.ones(64 << 20, dtype=torch.float32, device="cuda")  # 256 MiB copy: scales FETCH/WRITE_SIZE
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        log("profile run done")
        return

    log("timing builds")
    for _ in range(args.warmup):
        build()
    build_ms = timer.time(build, args.steps)
    flags = int(dev.u32_view(tree.info)[native.RK_INFO_FLAGS])

    log("autotuning the traversal on the frame")
    tuning = dev.autotune(tree, Ot, Dt, (fid, uv, dep), det_epsilon=HEADLINE_DET_EPSILON, reps=25,
                          reset_depth=True)
    tuned["config"] = tuning.config

    log("timing frames")
    for _ in range(args.warmup):
        frame()
    frame_ms = timer.time(frame, args.steps)
    default_ms = timer.time(lambda: frame(config=0), args.steps)
    hits = int((dep < f32max).sum().item())
    trace_ms = frame_ms  # one launch per frame: the trace kernel resets tmax itself
    fill_then_trace_ms = timer.time(frame_fill_then_trace, args.steps)
    eps0_frame_ms = timer.time(lambda: frame(0.0), args.steps)
    eps0_hits = int((dep < f32max).sum().item())
    # the same frame with the rays formed inside the trace kernel (its own tuned variant)
    cout = dev.new_outputs(nr)
    cam_tuning = dev.autotune_camera(tree, cam, W, H, cout, det_epsilon=HEADLINE_DET_EPSILON, reps=25)
    cam_frame = lambda: dev.fast_trace_camera(tree, cam, W, H, out=cout, det_epsilon=HEADLINE_DET_EPSILON,
                                              config=cam_tuning.config)
    for _ in range(args.warmup):
        cam_frame()
    cam_ms = timer.time(cam_frame, args.steps)
    frame()
    cam_same = bool(torch.equal(cout[2], dep))
    dep.fill_(f32max)
    _, _, _, cnt = dev.fast_trace_counted(tree, Ot, Dt, out=(fid, uv, dep), det_epsilon=HEADLINE_DET_EPSILON)
    torch.cuda.synchronize()
    n_box, n_tri = (int(x) for x in cnt.cpu().numpy())
    flops = OPS_PER_BOX_TEST * n_box + OPS_PER_TRI_TEST * n_tri
    achieved_tflops = flops / (trace_ms * 1e-3) / 1e12

    result = {
        "metric": METRIC,
        "value": round(nr / (frame_ms * 1e-3) / 1e6, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(frame_ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "heightfield 708x708 fBm (999,698 FP32 triangles), 1920x1080 primary rays, closest hit",
            "faces": n, "rays": nr, "tree": "LBVH 30-bit Morton (Karras), " + ("one face per leaf" if args.leaf == 1 else f"leaf <= {args.leaf} faces"),
            "step": "one frame: reset tmax + closest-hit trace of every ray, tree resident in HBM (one launch, "
                    "RK_TRACE_RESET_DEPTH)",
            "det_epsilon": HEADLINE_DET_EPSILON,
        },
        "hit_fraction": round(hits / nr, 4),
        "trace_kernel_ms": round(trace_ms, 4),
        "trace_autotuned": tuning.as_dict(),
        "mrays_s_default_variant": round(nr / (default_ms * 1e-3) / 1e6, 2),
        "mrays_s_fill_then_trace": round(nr / (fill_then_trace_ms * 1e-3) / 1e6, 2),
        "camera_autotuned": cam_tuning.as_dict(),
        "mrays_s_det_epsilon_0": round(nr / (eps0_frame_ms * 1e-3) / 1e6, 2),
        "hit_fraction_det_epsilon_0": round(eps0_hits / nr, 4),
        "camera_fused_mrays_s": round(nr / (cam_ms * 1e-3) / 1e6, 2),
        "camera_fused_same_depths": cam_same,
        "lbvh_mtri_s": round(n / (build_ms * 1e-3) / 1e6, 2),
        "lbvh_build_ms": round(build_ms, 4),
        "build_roofline": build_roofline("heightfield", n, V.shape[0], 4, build_ms),
        "box_tests_per_ray": round(n_box / nr, 2),
        "tri_tests_per_ray": round(n_tri / nr, 2),
        "build_flags": flags,
        "roofline": {
            "bound": "valu",
            "kernel": "k_fast_trace",
            "achieved": round(achieved_tflops, 4),
            "peak": FP32_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": round(achieved_tflops / FP32_PEAK_TFLOPS, 5),
            "traffic": TRACE_TRAFFIC_BYTES,
            "algorithmic_bytes": 36 * nr + 12 * hits,
            "note": "algorithmic FP32 ops = 24 per slab test + 53 per triangle test, counted on this frame "
                    "(default variant); achieved over the tuned variant's kernel time (the whole frame: one "
                    "launch that also resets tmax); 32 B/ray read (origin, direction) + 4 B depth written per "
                    "ray + 12 B id and coords per hit of compulsory HBM traffic makes this kernel "
                    "VALU/latency-bound, not HBM-bound; "
                    "traffic = HBM bytes per launch from rocprofv3 FETCH_SIZE/WRITE_SIZE runs (rays, outputs "
                    "and node/triangle records fetched past the L2)",
        },
    }

    if not args.no_ref:
        log("reference tree and traversals")
        rb = dev.LBVHBuilder(n)
        rtree = rb.build(Ft.clone(), Vt, hb, 12, True)
        torch.cuda.synchronize()
        Fr = rtree.faces
        ref_build_ms = timer.time(lambda: rb.build(Fr, Vt, hb, 12, True), 5)
        rtree = rb.build(Ft.clone(), Vt, hb, 12, True)
        torch.cuda.synchronize()
        rtree.node_count = int(rtree.host_info()[native.RK_INFO_NODE_COUNT])
        rf, ruv, rdep = dev.new_outputs(nr)
        ref_ray_ms = timer.time(lambda: dev.trace_ref(rtree, Vt, Ot, Dt, W, H, False, out=(rf, ruv, rdep)), 1)
        rf, ruv, rdep = dev.new_outputs(nr)
        ref_pkt_ms = timer.time(lambda: dev.trace_ref(rtree, Vt, Ot, Dt, W, H, True, out=(rf, ruv, rdep)), 1)
        result["reference_lbvh_mtri_s"] = round(n / (ref_build_ms * 1e-3) / 1e6, 2)
        rb31 = dev.LBVHBuilder(n)
        Fr31 = Ft.clone()
        rb31.build(Fr31, Vt, hb, 31, True)
        torch.cuda.synchronize()
        result["reference_lbvh_d31_mtri_s"] = round(n / (timer.time(lambda: rb31.build(Fr31, Vt, hb, 31, True), 5)
                                                         * 1e-3) / 1e6, 2)
        del rb31, Fr31
        result["reference_per_ray_mrays_s"] = round(nr / (ref_ray_ms * 1e-3) / 1e6, 3)
        result["reference_packet_mrays_s"] = round(nr / (ref_pkt_ms * 1e-3) / 1e6, 3)
        result["reference_node_count"] = rtree.node_count
        result["frame_loop"] = frame_loop_leg(torch, dev, native, timer, args.steps, args.warmup, tree, cam, W, H, F, V,
                                              Vt, cam_tuning.config, rtree, Ot, Dt, with_cpu=not args.no_cpu)

        log("reference 32-bin SAH and Hybrid builds")
        Fs = Ft.clone()
        sb = dev.SAHBuilder(n)
        sah = lambda: sb.build(Fs, Vt, hb)  # rk_build_sah waits for each level, so it returns built
        sah()
        sah_ms = timer.time(sah, 5)
        stree = sb.build(Ft.clone(), Vt, hb)
        torch.cuda.synchronize()
        stree.node_count = int(stree.host_info()[native.RK_INFO_NODE_COUNT])
        sf, suv, sdep = dev.new_outputs(nr)
        sah_ray_ms = timer.time(lambda: dev.trace_ref(stree, Vt, Ot, Dt, W, H, False, out=(sf, suv, sdep)), 1)
        hl = dev.LBVHBuilder(n)
        hs = dev.SAHBuilder(n, 1 << 13)  # a depth-12 LBVH has at most 2^13 - 1 nodes
        Fh = Ft.clone()

        def hybrid():
            lb = hl.build(Fh, Vt, hb, 12, False)
            torch.cuda.synchronize()
            lb.node_count = int(lb.host_info()[native.RK_INFO_NODE_COUNT])
            return hs.build(Fh, Vt, hb, lbvh=lb)

        hybrid()
        hybrid_ms = timer.time(hybrid, 5)
        result["reference_sah_mtri_s"] = round(n / (sah_ms * 1e-3) / 1e6, 2)
        result["reference_sah_build_ms"] = round(sah_ms, 3)
        result["reference_sah_node_count"] = stree.node_count
        result["reference_sah_per_ray_mrays_s"] = round(nr / (sah_ray_ms * 1e-3) / 1e6, 3)
        result["reference_sah_same_depths"] = bool(torch.equal(sdep, rdep))
        result["reference_hybrid_mtri_s"] = round(n / (hybrid_ms * 1e-3) / 1e6, 2)
        result["reference_hybrid_build_ms"] = round(hybrid_ms, 3)
        if not args.no_cpu:
            import oracle
            ob, oi = oracle.calc_boxes(F, V)
            t0 = time.perf_counter()
            oracle.build_sah(ob, oi, hb)
            result["reference_sah_cpu_mtri_s"] = round(n / (time.perf_counter() - t0) / 1e6, 3)

    log("icosphere 20,480 faces, 256x256 rays")
    result["icosphere"] = icosphere_leg(torch, dev, native, timer, args.steps, args.warmup, not args.no_cpu)

    if not args.no_soup:
        log("soup 10M faces, 63-bit keys, 2^24 incoherent rays")
        result["soup"] = soup_leg(torch, dev, native, timer, min(args.steps, 10), min(args.warmup, 2), not args.no_cpu)

    if not args.no_batch:
        log("batch of 4096 meshes: autotune, graph, batched vs plain")
        result["batch"] = batch_leg(torch, dev, native, timer, min(args.steps, 10), min(args.warmup, 2),
                                    not args.no_cpu)

    if not args.no_serve:
        log("serving: Poisson 2000 req/s over 32 deforming scenes")
        result["serve"] = serve_leg(dev, native, with_cpu=not args.no_cpu)

    if not args.no_cpu:
        log("cpu baseline")
        result["cpu_baseline"] = cpu_baseline(V, F, hb, O, D)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
