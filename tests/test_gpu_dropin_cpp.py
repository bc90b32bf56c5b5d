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
