"""Synthetic multi-mesh workloads: batches of random triangle soups with rays.

Sizes come from a SplitMix64 stream (seeded per batch) so every machine
regenerates the same batch; each mesh's soup and rays come from the library's
own generators (rk_scene_soup, rk_rays_random).
"""
from __future__ import annotations

import math
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from . import native as N

_M64 = (1 << 64) - 1


def splitmix64(state: int):
    """Yields the SplitMix64 sequence (Steele, Lea, Flood 2014) from 'state'."""
    while True:
        state = (state + 0x9E3779B97F4A7C15) & _M64
        z = state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
        yield z ^ (z >> 31)


def log_uniform_counts(count: int, lo: int, hi: int, rng) -> np.ndarray:
    """count integers, log-uniform in [lo, hi]."""
    out = np.empty(count, np.uint32)
    for i in range(count):
        u = (next(rng) >> 11) * (1.0 / (1 << 53))
        out[i] = min(hi, int(math.floor(lo * (hi / lo) ** u)))
    return out


def batch_table(face_counts, vertex_counts, ray_counts) -> np.ndarray:
    """rk_batch_mesh rows for meshes stored back to back (faces, vertices and
    rays each in mesh order with no gaps)."""
    fc = np.asarray(face_counts, np.uint64)
    vc = np.asarray(vertex_counts, np.uint64)
    rc = np.asarray(ray_counts, np.uint64)
    t = np.zeros(fc.shape[0], N.BATCH_MESH_DTYPE)
    t["face_count"], t["vertex_count"], t["ray_count"] = fc, vc, rc
    t["face_offset"] = np.concatenate([[0], np.cumsum(fc)[:-1]])
    t["vertex_offset"] = np.concatenate([[0], np.cumsum(vc)[:-1]])
    t["ray_offset"] = np.concatenate([[0], np.cumsum(rc)[:-1]])
    return t


def soup_batch(mesh_count: int = 4096, seed: int = 0xC4, faces=(64, 65536), rays=(256, 16384),
               reference_faces: int = 10_000_000):
    """A batch of soups in the unit cube. Mesh m has a log-uniform face count
    and ray count; its edges are lognormal (sigma 0.5) around
    0.002 * sqrt(reference_faces / faces_m), so every mesh has the total
    triangle area, and hence the hit rate, of a 10M-face soup with 0.002
    edges. Rays start uniform in the cube with directions uniform on the
    sphere.

    Returns (table, F, V, hitboxes, O, D): F holds mesh-local vertex indices."""
    rng = splitmix64(seed)
    fc = log_uniform_counts(mesh_count, faces[0], faces[1], rng)
    rc = log_uniform_counts(mesh_count, rays[0], rays[1], rng)
    seeds = [next(rng) for _ in range(2 * mesh_count)]
    nf, nr = int(fc.sum()), int(rc.sum())
    F = np.zeros((nf, 4), np.uint32)
    V = np.zeros((3 * nf, 4), np.float32)
    O = np.zeros((nr, 4), np.float32)
    D = np.zeros((nr, 4), np.float32)
    hb = np.zeros((mesh_count, 2, 4), np.float32)
    L = N.lib()
    fo = np.concatenate([[0], np.cumsum(fc.astype(np.int64))[:-1]])
    ro = np.concatenate([[0], np.cumsum(rc.astype(np.int64))[:-1]])

    def make(m):  # ctypes releases the GIL, so meshes generate in parallel
        n, r, f0, r0 = int(fc[m]), int(rc[m]), int(fo[m]), int(ro[m])
        mu = math.log(0.002 * math.sqrt(reference_faces / n))
        Fm, Vm = F[f0:f0 + n], V[3 * f0:3 * (f0 + n)]
        nv, nff = N._U32(), N._U32()
        N.check(L.rk_scene_soup(n, seeds[2 * m], mu, 0.5, 0.0, 1.0, Vm.ctypes.data, Fm.ctypes.data,
                                N.C.byref(nv), N.C.byref(nff)), "soup")
        hb[m] = N.hitbox(Vm)
        N.check(L.rk_rays_random(r, seeds[2 * m + 1], 0.0, 1.0, O[r0:r0 + r].ctypes.data, D[r0:r0 + r].ctypes.data),
                "random rays")

    with ThreadPoolExecutor(max_workers=min(16, os.cpu_count() or 1)) as ex:
        list(ex.map(make, range(mesh_count)))
    table = batch_table(fc, 3 * fc.astype(np.uint64), rc)
    return table, F, V, hb, O, D
