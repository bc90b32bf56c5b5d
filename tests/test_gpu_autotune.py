"""Traversal variants and their autotuner: every variant of rk_fast_trace and
of rk_fast_trace_camera returns the same frame, equal to the oracle's closest
hits; the tuners return a valid variant and leave the frame it traced."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

F32MAX = np.finfo(np.float32).max
N_TRACE_VARIANTS = 27
N_CAMERA_VARIANTS = 20


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


def _h(t):
    import torch
    torch.cuda.synchronize()
    return t.cpu().numpy()


def _scene(rk, w, h):
    V, F = rk.heightfield(96, seed=5)
    el, az = np.radians(25.0), np.radians(40.0)
    tgt = np.array([0.5, 0.5, 0.1])
    eye = tgt + 1.4 * np.array([np.cos(el) * np.cos(az), np.cos(el) * np.sin(az), np.sin(el)])
    cam = rk.camera_lookat(eye, tgt, np.radians(60.0), w, h, up=[0, 0, 1])
    return V, F, cam


def test_variant_table(dev, rk):
    key = lambda v: (v.workgroup, v.ray_run, v.lds_stack, v.steal_oldest, v.leaf_batch, v.donate_min, v.bound_every)
    shapes = {key(dev.trace_variant(k)) for k in range(1, N_TRACE_VARIANTS + 1)}
    assert len(shapes) == N_TRACE_VARIANTS
    d = dev.trace_variant(0)
    assert (d.workgroup, d.ray_run, d.lds_stack) == (64, 8, 8)
    ckey = lambda v: (v.workgroup, v.bands, v.xcd_strips, v.steal_oldest, v.leaf_batch, v.donate_min)
    cams = {ckey(dev.trace_variant(k, True)) for k in range(1, N_CAMERA_VARIANTS + 1)}
    assert len(cams) == N_CAMERA_VARIANTS
    with pytest.raises(rk.RkError):
        dev.trace_variant(N_TRACE_VARIANTS + 1)
    with pytest.raises(rk.RkError):
        dev.trace_variant(N_CAMERA_VARIANTS + 1, True)


# 131 x 77 leaves the last band, the last run of each row and the last XCD strip part empty;
# leaves of up to 4 faces exercise the postponed leaves' face loop
@pytest.mark.parametrize("w,h,leaf", [(131, 77, 1), (256, 160, 1), (256, 160, 4)])
def test_every_variant_same_frame(dev, rk, ora, w, h, leaf):
    V, F, cam = _scene(rk, w, h)
    tree = dev.FastBuilder(F.shape[0]).build(_t(F), _t(V), rk.hitbox(V), leaf)
    O, D = rk.primary_rays_host(cam, w, h)
    Ot, Dt = _t(O), _t(D)
    of, ouv, od = ora.trace_bruteforce(F, V, O, D, det_epsilon=0.0)
    hit = od < F32MAX
    assert hit.mean() > 0.05
    order = _h(tree.indices).view(np.uint32)
    for k in range(0, N_TRACE_VARIANTS + 1):
        fid, uv, dep = (_h(x) for x in dev.fast_trace(tree, Ot, Dt, det_epsilon=0.0, config=k))
        np.testing.assert_array_equal(dep, od, err_msg=f"trace variant {k}")
        np.testing.assert_array_equal(order[fid.view(np.uint32)[hit]], of[hit], err_msg=f"trace variant {k}")
        np.testing.assert_array_equal(uv[hit], ouv[hit], err_msg=f"trace variant {k}")
    # a finite tmax on every third ray (the compacting variants test it at the root)
    base = dev.new_outputs(w * h)
    base[2][::3] = 0.9
    want = [_h(x) for x in dev.fast_trace(tree, Ot, Dt, out=tuple(t.clone() for t in base), det_epsilon=0.0)]
    assert not np.array_equal(want[2], od)
    for k in range(1, N_TRACE_VARIANTS + 1):
        got = [_h(x) for x in dev.fast_trace(tree, Ot, Dt, out=tuple(t.clone() for t in base), det_epsilon=0.0,
                                             config=k)]
        for a, b in zip(got, want):
            np.testing.assert_array_equal(a, b, err_msg=f"trace variant {k}, finite tmax")
    for k in range(0, N_CAMERA_VARIANTS + 1):
        fid, uv, dep = (_h(x) for x in dev.fast_trace_camera(tree, cam, w, h, det_epsilon=0.0, config=k))
        np.testing.assert_array_equal(dep, od, err_msg=f"camera variant {k}")
        np.testing.assert_array_equal(order[fid.view(np.uint32)[hit]], of[hit], err_msg=f"camera variant {k}")
        assert np.all(fid.view(np.uint32)[~hit] == 0xFFFFFFFF), f"camera variant {k}"


def test_autotuners_choose_and_leave_the_frame(dev, rk, ora):
    w, h = 320, 200
    V, F, cam = _scene(rk, w, h)
    tree = dev.FastBuilder(F.shape[0]).build(_t(F), _t(V), rk.hitbox(V), 1)
    O, D = rk.primary_rays_host(cam, w, h)
    Ot, Dt = _t(O), _t(D)
    ref = [_h(x) for x in dev.fast_trace(tree, Ot, Dt)]
    out = dev.new_outputs(w * h)
    out[2][::3] = 1.25  # a finite tmax on some rays: restored before each timed call
    want = [_h(x) for x in dev.fast_trace(tree, Ot, Dt, out=tuple(t.clone() for t in out))]
    best = dev.autotune(tree, Ot, Dt, out, reps=2)
    assert 1 <= best.config <= N_TRACE_VARIANTS and best.ms > 0
    np.testing.assert_array_equal(_h(out[2]), want[2])
    hit = want[2] < F32MAX
    np.testing.assert_array_equal(_h(out[0])[hit], want[0][hit])
    assert not np.array_equal(want[2], ref[2])  # the finite tmax mattered
    cout = dev.new_outputs(w * h)
    cbest = dev.autotune_camera(tree, cam, w, h, cout, reps=2)
    assert 1 <= cbest.config <= N_CAMERA_VARIANTS and cbest.ms > 0
    np.testing.assert_array_equal(_h(cout[2]), ref[2])


def test_reset_depth_equals_fill_then_trace(dev, rk, ora):
    """RK_TRACE_RESET_DEPTH leaves what filling depths with FLT_MAX and tracing
    leaves: every depth written (FLT_MAX on a miss), ids and coords of misses
    untouched; for every rk_fast_trace variant and for the sorted trace."""
    w, h = 200, 120
    V, F, cam = _scene(rk, w, h)
    tree = dev.FastBuilder(F.shape[0]).build(_t(F), _t(V), rk.hitbox(V), 1)
    O, D = rk.primary_rays_host(cam, w, h)
    Ot, Dt = _t(O), _t(D)
    n = w * h
    rng = np.random.default_rng(9)
    fid0 = rng.integers(0, 1 << 20, n, dtype=np.uint32)
    uv0 = rng.random((n, 2), dtype=np.float32)
    dep0 = rng.random(n, dtype=np.float32) * 0.01  # tiny depths: a trace reading them would hit nothing
    filled = (_t(fid0), _t(uv0), _t(np.full(n, F32MAX, np.float32)))
    want = [_h(x) for x in dev.fast_trace(tree, Ot, Dt, out=filled)]
    hit = want[2] < F32MAX
    assert 0.05 < hit.mean() < 0.95
    np.testing.assert_array_equal(want[0][~hit], fid0[~hit])
    for k in range(0, N_TRACE_VARIANTS + 1):
        out = (_t(fid0), _t(uv0), _t(dep0))
        got = [_h(x) for x in dev.fast_trace(tree, Ot, Dt, out=out, config=k, reset_depth=True)]
        for a, b in zip(got, want):
            np.testing.assert_array_equal(a, b, err_msg=f"trace variant {k}")
    out = (_t(fid0), _t(uv0), _t(dep0))
    got = [_h(x) for x in dev.fast_trace_sorted(tree, Ot, Dt, out=out, reset_depth=True)]
    for a, b in zip(got, want):
        np.testing.assert_array_equal(a, b, err_msg="sorted trace")
