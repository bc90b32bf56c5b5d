"""The C++ drop-in classes (include/renderkit/) driven like the reference's
sample 02: the tree, the primary rays and both traversals match the oracle
bit for bit; the .cdm round trip and the fast tree give the same hits."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parent.parent
EXE = ROOT / "tests" / "cpp" / "_build" / "sample02_headless"
F32MAX = np.finfo(np.float32).max


def _read_chunks(path):
    data = path.read_bytes()
    hdr = np.frombuffer(data[:16], np.uint32)
    hitbox = np.frombuffer(data[16:48], np.float32).reshape(2, 4)
    off, chunks = 48, []
    while off < len(data):
        n = int(np.frombuffer(data[off:off + 4], np.uint32)[0])
        chunks.append(data[off + 4:off + 4 + n])
        off += 4 + n
    return hdr, hitbox, chunks


@pytest.mark.parametrize("sub,w,h", [(5, 256, 256), (3, 100, 70)])
def test_sample02_dropin_matches_oracle(rk, ora, tmp_path, sub, w, h):
    assert EXE.exists(), "build() compiles tests/cpp first"
    out = tmp_path / "out.bin"
    r = subprocess.run([str(EXE), str(out), str(sub), str(w), str(h), str(tmp_path)], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr
    hdr, hitbox, ch = _read_chunks(out)
    nv, nf, M, real_depth = (int(x) for x in hdr)
    V, F = rk.icosphere(sub)
    assert (nv, nf) == V.shape[0:1] + F.shape[0:1]
    np.testing.assert_array_equal(hitbox, rk.hitbox(V))
    ref = ora.build_lbvh_ref(F, V, hitbox, 12, True)
    assert M == ref.node_count and real_depth == ref.real_depth
    nodes = np.frombuffer(ch[0], ora.NODE_DTYPE)
    for k in ("w0", "w1", "bmin", "bmax"):
        assert np.array_equal(nodes[k], ref.nodes[k]), k
    np.testing.assert_array_equal(np.frombuffer(ch[1], np.uint32), ref.indices)
    np.testing.assert_array_equal(np.frombuffer(ch[2], np.uint32).reshape(-1, 4), ref.faces)
    # frame 1 (packet), then the rays it traced, then frame 2 (per-ray)
    pf, puv, pd = (np.frombuffer(ch[i], t) for i, t in ((3, np.uint32), (4, np.float32), (5, np.float32)))
    O = np.frombuffer(ch[6], np.float32).reshape(-1, 4).copy()
    D = np.frombuffer(ch[7], np.float32).reshape(-1, 4).copy()
    # the drop-in camera and CUDA_calcPrimaryRays give the oracle's rays bit for bit
    # (setCenter(0, 0, 3), lookAt(0, 0, 0), setFOV(60 * DEG2RAD), setRatio(w, h))
    cam = ora.Camera()
    cam.set_center(0, 0, 3)
    cam.look_at(0, 0, 0)
    cam.set_fov(np.float32(60.0) * np.float32(0.01745329252))
    cam.set_ratio(w, h)
    Oo, Do = cam.primary_rays(w, h)
    np.testing.assert_array_equal(O, Oo)
    np.testing.assert_array_equal(D, Do)
    of, ouv, od, _ = ora.trace_ref_packet(ref.nodes, ref.faces, V, O, D, w, h)
    np.testing.assert_array_equal(pf, of)
    np.testing.assert_array_equal(puv.reshape(-1, 2), ouv)
    np.testing.assert_array_equal(pd, od)
    rf, ruv, rd = (np.frombuffer(ch[i], t) for i, t in ((8, np.uint32), (9, np.float32), (10, np.float32)))
    of, ouv, od, _ = ora.trace_ref_ray(ref.nodes, ref.faces, V, O, D)
    np.testing.assert_array_equal(rf, of)
    np.testing.assert_array_equal(ruv.reshape(-1, 2), ouv)
    np.testing.assert_array_equal(rd, od)
    # .cdm round trip: identical packet frame
    for a, b in ((11, 3), (12, 4), (13, 5)):
        assert ch[a] == ch[b]
    # fast tree: same depths; hits identical in original face ids except exact t ties
    fd = np.frombuffer(ch[16], np.float32)
    np.testing.assert_array_equal(fd, od)
    hit = od < F32MAX
    fast_ids = np.frombuffer(ch[17], np.uint32)[np.frombuffer(ch[14], np.uint32)[hit]]
    ref_ids = ref.indices[of[hit]]
    assert (fast_ids == ref_ids).mean() > 0.99


LOOP = ROOT / "tests" / "cpp" / "_build" / "sample02_loop"


def _loop_chunks(path):
    data = path.read_bytes()
    hdr = np.frombuffer(data[:16], np.uint32)
    off, chunks = 16, []
    while off < len(data):
        n = int(np.frombuffer(data[off:off + 4], np.uint32)[0])
        chunks.append(data[off + 4:off + 4 + n])
        off += 4 + n
    return hdr, chunks


@pytest.mark.parametrize("scene,frames", [("cube", 3), ("ico", 2)])
def test_sample02_frame_loop(rk, ora, tmp_path, scene, frames):
    """Sample 02's loop (clear, primary rays, trace, interpolate points and
    normals, updateRenderSurface, renderer and engine update) until the headless
    window closes. The sample's own scene is a createCube model: its hitbox
    stays empty, but its 12 faces make a single SAH leaf, whose triangles are
    tested without a box test (CUDA_raytraceMeshTree.cu:145-178), so the cube
    shows. Rays, depths, normals and the shaded pixels equal the oracle's."""
    assert LOOP.exists(), "build() compiles tests/cpp first"
    out = tmp_path / "loop.bin"
    r = subprocess.run([str(LOOP), str(out), scene, str(frames)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    hdr, ch = _loop_chunks(out)
    w, h, rframes, eframes = (int(x) for x in hdr)
    assert (w, h, rframes, eframes) == (400, 300, frames, frames)
    V = np.frombuffer(ch[0], np.float32).reshape(-1, 4)
    F = np.frombuffer(ch[1], np.uint32).reshape(-1, 4)
    O = np.frombuffer(ch[2], np.float32).reshape(-1, 4)
    D = np.frombuffer(ch[3], np.float32).reshape(-1, 4)
    fid = np.frombuffer(ch[4], np.uint32)
    uv = np.frombuffer(ch[5], np.float32).reshape(-1, 2)
    dep = np.frombuffer(ch[6], np.float32)
    pts = np.frombuffer(ch[7], np.float32).reshape(-1, 4)
    nrm = np.frombuffer(ch[8], np.float32).reshape(-1, 4)
    px = np.frombuffer(ch[9], np.uint32)
    extra = np.frombuffer(ch[10], np.float32)
    cam_dir, back = extra[:3], extra[4:8]
    np.testing.assert_array_equal(back, np.array([0.227451, 0.431373, 0.647059, 1.0], np.float32))
    cam = ora.Camera()
    cam.set_ratio(400, 300)
    cam.set_center(-10, 20, 30)
    cam.look_at(0, 0, 0)
    np.testing.assert_array_equal(cam_dir, cam.dir)
    Oo, Do = cam.primary_rays(w, h)
    np.testing.assert_array_equal(O, Oo)
    np.testing.assert_array_equal(D, Do)
    of, ouv, od = ora.trace_bruteforce(F, V, O, D)
    np.testing.assert_array_equal(dep, od)
    hit = od < F32MAX
    assert 0.01 < hit.mean() < 0.9  # the icosphere (radius 10) covers 4.5% of the 90-degree view
    assert (fid[hit] == of[hit]).mean() > 0.99  # the packet traversal keeps the first of exact t ties
    same = hit & (fid == of)
    np.testing.assert_array_equal(uv[same], ouv[same])
    assert np.all(px[~hit] == 0xFF3A6EA5) and np.all(px[hit] != 0xFF3A6EA5)
    p0, n0, _, _, _ = ora.interpolate(fid, uv, F, V)
    np.testing.assert_array_equal(pts, p0)
    np.testing.assert_array_equal(nrm, n0)
    np.testing.assert_array_equal(px, ora.shade(dep, nrm, w, h, back, cam_dir))
