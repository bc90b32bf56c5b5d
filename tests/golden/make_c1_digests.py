"""Writes c1_digests.json: SHA-256 digests of the C1 workload's outputs as
the oracle computes them. C1 is the icosphere (subdivision 5, 20,480 faces)
with 256x256 primary rays from (0, 0, 3) at the origin, 60-degree fov.

The digests cover:
- the mesh generator's vertex and face bytes;
- the reference LBVH at depths 12 and 31: node words and boxes, indices, and
  the reordered faces;
- the per-ray and packet traversals of the frame: face ids, (u, v) and depths.

The reference ships no golden vectors, and it cannot be built or run here
(it needs nvcc, the legacy texture API and an un-vendored header). These
digests therefore pin the oracle against silent changes, and let the GPU
tests compare with committed data without running the oracle.

Run from the repository root: python tests/golden/make_c1_digests.py"""
import hashlib
import json
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))


def digest(*arrays) -> str:
    """SHA-256 over the arrays' bytes, float arrays with -0 made +0 (min/max of
    signed zeros may pick either; the trees compare them as equal floats)."""
    h = hashlib.sha256()
    for a in arrays:
        a = np.ascontiguousarray(a)
        if a.dtype.kind == "f":
            a = a + a.dtype.type(0)
        h.update(a.tobytes())
    return h.hexdigest()


def node_fields(nodes):
    """(words, box min, box max) of bvnode records (oracle.NODE_DTYPE)."""
    return np.stack([nodes["w0"], nodes["w1"]], 1), nodes["bmin"], nodes["bmax"]


def c1_outputs(rk, build, trace):
    """Digests of one implementation's C1 outputs. build(F, V, hitbox, depth)
    returns (node records, indices, reordered faces, handle); trace(handle,
    V, O, D, coherency) returns (face ids, uv, depths)."""
    V, F = rk.icosphere(5)
    hb = rk.hitbox(V)
    cam = rk.camera_lookat([0, 0, 3], [0, 0, 0], 60 * 0.01745329252, 256, 256)
    O, D = rk.primary_rays_host(cam, 256, 256)
    out = {"mesh": digest(V, F), "rays": digest(O, D)}
    for depth in (12, 31):
        nodes, idx, faces, handle = build(F, V, hb, depth)
        out[f"lbvh_d{depth}"] = digest(*node_fields(nodes), idx, faces)
        if depth == 12:
            for coh in (False, True):
                out[f"trace_d12_{'packet' if coh else 'ray'}"] = digest(*trace(handle, V, O, D, coh))
    return out


def oracle_outputs():
    from cudart_amd import native as rk
    import oracle

    def build(F, V, hb, depth):
        t = oracle.build_lbvh_ref(F, V, hb, depth, True)
        return t.nodes, t.indices, t.faces, t

    def trace(t, V, O, D, coh):
        if coh:
            f, uv, d, _ = oracle.trace_ref_packet(t.nodes, t.faces, V, O, D, 256, 256)
        else:
            f, uv, d, _ = oracle.trace_ref_ray(t.nodes, t.faces, V, O, D)
        return f, uv, d

    return c1_outputs(rk, build, trace)


if __name__ == "__main__":
    out = oracle_outputs()
    path = Path(__file__).with_name("c1_digests.json")
    path.write_text(json.dumps(out, indent=1) + "\n")
    print(path, out)
