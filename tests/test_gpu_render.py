"""The sample-02 frame after the trace: rk_shade against the oracle's
kernel_shade, and rk_fast_render_camera (rays, closest hit, hit normal and
pixel in one launch) against the three-kernel sequence it replaces
(rk_fast_trace_camera, rk_interpolate_data, rk_shade) and against the oracle's
interpolation and shading of the same hits."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

F32MAX = np.finfo(np.float32).max
BACK = (58 / 255.0, 110 / 255.0, 165 / 255.0, 1.0)


@pytest.fixture(scope="module")
def dev():
    import torch
    from cudart_amd import device
    torch.cuda.init()
    return device


def _t(a):
    import torch
    a = np.ascontiguousarray(a)
    return torch.from_numpy(a.view(np.int32) if a.dtype == np.uint32 else a).to("cuda")


def _h(t, dt=None):
    import torch
    torch.cuda.synchronize()
    a = t.cpu().numpy()
    return a.view(dt) if dt is not None else a


def test_shade_equals_oracle(dev, ora):
    w, h, pitch = 97, 61, 104
    n = w * h
    rng = np.random.default_rng(4)
    dep = np.where(rng.random(n) < 0.6, rng.random(n) * 10, F32MAX).astype(np.float32)
    dep[::17] = np.nextafter(np.float32(F32MAX), np.float32(0))  # one ulp below FLT_MAX: a hit
    nrm = rng.normal(size=(n, 4)).astype(np.float32)
    nrm[:, :3] /= np.linalg.norm(nrm[:, :3], axis=1, keepdims=True)
    cam = np.array([0.3, -0.5, -0.81], np.float32)
    cam /= np.linalg.norm(cam)
    got = _h(dev.shade(_t(dep), _t(nrm), w, h, BACK, cam, pitch=pitch), np.uint32)
    want = ora.shade(dep, nrm, w, h, np.array(BACK, np.float32), cam, pitch)
    for y in range(h):  # the pitch padding is not part of the frame
        np.testing.assert_array_equal(got[y * pitch:y * pitch + w], want[y * pitch:y * pitch + w])


def _scene(rk, name, w, h):
    if name == "hf":
        V, F = rk.heightfield(96, seed=5)
        el, az = np.radians(25.0), np.radians(40.0)
        tgt = np.array([0.5, 0.5, 0.1])
        eye = tgt + 1.4 * np.array([np.cos(el) * np.cos(az), np.cos(el) * np.sin(az), np.sin(el)])
        cam = rk.camera_lookat(eye, tgt, np.radians(60.0), w, h, up=[0, 0, 1])
    else:
        V, F = rk.icosphere(4)
        eye, tgt = np.array([0.4, 0.3, 2.6]), np.zeros(3)
        cam = rk.camera_lookat(eye, tgt, np.radians(50.0), w, h)
    look = (tgt - eye) / np.linalg.norm(tgt - eye)
    return V, F, cam, look.astype(np.float32)


@pytest.mark.parametrize("name,w,h,pitch,vnormals", [("hf", 131, 77, 131, False), ("hf", 200, 120, 208, True),
                                                     ("ico", 160, 100, 160, True), ("ico", 64, 64, 70, False)])
def test_render_camera_equals_the_three_kernels(dev, rk, ora, name, w, h, pitch, vnormals):
    V, F, cam, look = _scene(rk, name, w, h)
    tree = dev.FastBuilder(F.shape[0]).build(_t(F), _t(V), rk.hitbox(V), 1)
    order = _h(tree.indices, np.uint32)
    Fs = np.ascontiguousarray(F[order])  # the faces in leaf order
    N = None
    if vnormals:
        rng = np.random.default_rng(7)
        N = rng.normal(size=(V.shape[0], 4)).astype(np.float32)
    # the sequence: camera trace, normals, shading
    fid, uv, dep = dev.fast_trace_camera(tree, cam, w, h)
    nrm = dev.interpolate(fid, uv, w, h, _t(Fs), _t(V), normals=None if N is None else _t(N))[1]
    seq = _h(dev.shade(dep, nrm, w, h, BACK, look, pitch=pitch), np.uint32)
    # one launch
    import torch
    out = (torch.zeros(w * h, dtype=torch.int32, device="cuda"), torch.zeros((w * h, 2), device="cuda"),
           torch.zeros(w * h, device="cuda"))
    px = _h(dev.render_camera(tree, cam, w, h, _t(Fs), _t(V), BACK, look, normals=None if N is None else _t(N),
                              pitch=pitch, out=out), np.uint32)
    for y in range(h):
        np.testing.assert_array_equal(px[y * pitch:y * pitch + w], seq[y * pitch:y * pitch + w], err_msg=f"row {y}")
    for a, b in zip(out, (fid, uv, dep)):
        np.testing.assert_array_equal(_h(a), _h(b))
    # the oracle's interpolation and shading of the same hits
    fh, uh, dh = _h(fid, np.uint32), _h(uv), _h(dep)
    hit = dh < F32MAX
    assert 0.05 < hit.mean() < 0.99
    onrm = ora.interpolate(fh, uh, Fs, V, normals=N)[1]
    want = ora.shade(dh, onrm, w, h, np.array(BACK, np.float32), look, pitch)
    for y in range(h):
        np.testing.assert_array_equal(px[y * pitch:y * pitch + w], want[y * pitch:y * pitch + w], err_msg=f"row {y}")
