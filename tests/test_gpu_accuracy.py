"""Error of the hit distance t against FP64. The GPU computes t with the
reference's FP32 Moller-Trumbore arithmetic (contraction off); the oracle
recomputes t in double for the same face (oracle.hit_t_f64). The figures are
printed and, when RK_ACCURACY_JSON names a file, written there as JSON."""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

F32MAX = np.finfo(np.float32).max
RESULTS = {}


@pytest.fixture(scope="module")
def dev():
    import torch
    from cudart_amd import device
    torch.cuda.init()
    return device


def _t(a):
    import torch
    a = np.ascontiguousarray(a)
    if a.dtype == np.uint32:
        a = a.view(np.int32)
    return torch.from_numpy(a).to("cuda")


def _h(t, dtype=None):
    import torch
    torch.cuda.synchronize()
    a = t.cpu().numpy()
    return a.view(dtype) if dtype is not None else a


def _stats(name, ora, F, V, O, D, order, fid, dep):
    hit = dep < F32MAX
    orig = order[fid.view(np.uint32)[hit]]
    t64 = ora.hit_t_f64(F, V, np.ascontiguousarray(O[hit]), np.ascontiguousarray(D[hit]), orig)
    t32 = dep[hit].astype(np.float64)
    ok = np.isfinite(t64) & (t64 > 0)
    err = np.abs(t32[ok] - t64[ok])
    rel = err / t64[ok]
    ulp = err / np.spacing(dep[hit][ok]).astype(np.float64)
    s = {"hits": int(hit.sum()), "max_rel": float(rel.max()), "mean_rel": float(rel.mean()),
         "max_ulp": float(ulp.max()), "mean_ulp": float(ulp.mean()), "p999_ulp": float(np.quantile(ulp, 0.999))}
    RESULTS[name] = s
    print(name, s)
    out = os.environ.get("RK_ACCURACY_JSON")
    if out:
        with open(out, "w") as f:
            json.dump(RESULTS, f, indent=1)
    return s


def test_t_error_icosphere(dev, rk, ora):
    V, F = rk.icosphere(5)
    cam = rk.camera_lookat([0, 0, 3], [0, 0, 0], 60 * 0.01745329252, 256, 256)
    O, D = rk.primary_rays_host(cam, 256, 256)
    tree = dev.FastBuilder(F.shape[0], rk.RK_MORTON_30).build(_t(F), _t(V), rk.hitbox(V), 4)
    fid, uv, dep = dev.fast_trace(tree, _t(O), _t(D))
    s = _stats("C1 icosphere 256x256", ora, F, V, O, D, _h(tree.indices, np.uint32), _h(fid), _h(dep))
    assert s["hits"] > 10000 and s["mean_rel"] < 1e-6 and s["max_rel"] < 1e-4


def test_t_error_heightfield(dev, rk, ora):
    V, F = rk.heightfield(708)
    el, az, dist = np.radians(30.0), np.radians(45.0), 1.5
    tgt = np.array([0.5, 0.5, 0.1])
    eye = tgt + dist * np.array([np.cos(el) * np.cos(az), np.cos(el) * np.sin(az), np.sin(el)])
    cam = rk.camera_lookat(eye, tgt, 60 * 0.01745329252, 1920, 1080, up=[0, 0, 1])
    O, D = rk.primary_rays_host(cam, 1920, 1080)
    tree = dev.FastBuilder(F.shape[0], rk.RK_MORTON_30).build(_t(F), _t(V), rk.hitbox(V), 4)
    fid, uv, dep = dev.fast_trace(tree, _t(O), _t(D), det_epsilon=0.0)
    s = _stats("C2 heightfield 1920x1080", ora, F, V, O, D, _h(tree.indices, np.uint32), _h(fid), _h(dep))
    assert s["hits"] > 100000 and s["mean_rel"] < 1e-5


def test_t_error_soup(dev, rk, ora):
    n = 10_000_000
    V, F = rk.soup(n, seed=0xC3)
    O, D = rk.random_rays(1 << 20, seed=0xC3)
    tree = dev.FastBuilder(n, rk.RK_MORTON_63).build(_t(F), _t(V), rk.hitbox(V), 4)
    fid, uv, dep = dev.fast_trace(tree, _t(O), _t(D))
    s = _stats("C3 soup 10M, 2^20 rays", ora, F, V, O, D, _h(tree.indices, np.uint32), _h(fid), _h(dep))
    assert s["hits"] > 500000 and s["mean_rel"] < 1e-5
