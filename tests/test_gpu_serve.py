"""Camera frames, deformation + refit, and the serving loop, through the C ABI.
A served frame must equal the oracle's closest hits for the deformed scene, be
the same alone and in a mix of concurrent requests (with evictions), and not
depend on refit vs rebuild."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

F32MAX = np.finfo(np.float32).max
M64 = (1 << 64) - 1


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


def _mix64(z):
    z = np.asarray(z, np.uint64)
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def _checksum(orig_ids, uv, dep):
    """Host statement of the frame checksum (wrapping sum of per-ray hashes)."""
    n = dep.shape[0]
    i = np.arange(n, dtype=np.uint64)
    o = np.asarray(orig_ids, np.uint32).astype(np.uint64)
    d = dep.view(np.uint32).astype(np.uint64)
    u = uv[:, 0].copy().view(np.uint32).astype(np.uint64)
    v = uv[:, 1].copy().view(np.uint32).astype(np.uint64)
    with np.errstate(over="ignore"):
        inner = _mix64((d << np.uint64(32)) ^ (u * np.uint64(0x100000001B3)) ^ v)
        h = _mix64(((i << np.uint64(32)) | o) ^ inner)
        total = int(h.sum(dtype=np.uint64))
    hits = int((np.asarray(orig_ids, np.uint32) != 0xFFFFFFFF).sum())
    return total & M64, hits


def _camera(rk, V, side, az=0.7, height=None):
    hb = rk.hitbox(V)
    c = 0.5 * (hb[0, :3] + hb[1, :3])
    diag = float(np.linalg.norm(hb[1, :3] - hb[0, :3]))
    el = np.radians(30.0)
    eye = c + 1.2 * diag * np.array([np.cos(el) * np.cos(az), np.cos(el) * np.sin(az), np.sin(el)])
    return rk.camera_lookat(eye, c, np.radians(60.0), side, height or side, up=[0, 0, 1])


# 96x96 fills every band and run of the camera kernel's wave mapping; the
# others leave the last band and the last run of each row partly empty.
@pytest.mark.parametrize("w,h", [(96, 96), (101, 37), (13, 130), (1, 1)])
def test_camera_frame_equals_ray_arrays(dev, rk, w, h):
    V, F = rk.heightfield(128)
    tree = dev.FastBuilder(F.shape[0], rk.RK_MORTON_30).build(_t(F), _t(V), rk.hitbox(V), 4)
    cam = _camera(rk, V, w, height=h)
    O, D = rk.primary_rays_host(cam, w, h)
    fid, uv, dep = (_h(x) for x in dev.fast_trace(tree, _t(O), _t(D), det_epsilon=0.0))
    cf, cuv, cdep = (_h(x) for x in dev.fast_trace_camera(tree, cam, w, h, det_epsilon=0.0))
    assert cdep.shape == dep.shape == (w * h,)
    np.testing.assert_array_equal(cdep, dep)
    hit = dep < F32MAX
    if w * h > 1000:
        assert hit.mean() > 0.05  # 0.14 of the 96x96 rays hit the terrain
    np.testing.assert_array_equal(cf[hit], fid[hit])
    np.testing.assert_array_equal(cuv[hit], uv[hit])
    assert np.all(cf.view(np.uint32)[~hit] == 0xFFFFFFFF) and np.all(cuv[~hit] == 0)


@pytest.mark.parametrize("kind", ["soup", "terrain"])
def test_refit_after_deform_equals_rebuild_and_oracle(dev, rk, ora, kind):
    import torch
    if kind == "soup":
        V, F = rk.soup(3000, seed=4, mu=float(np.log(0.05)))
    else:
        V, F = rk.heightfield(60)
    n = F.shape[0]
    hb = rk.hitbox(V)
    amp = 0.05 * float((hb[1, :3] - hb[0, :3]).max())
    grow = np.array([[-amp] * 3 + [0], [amp] * 3 + [0]], np.float32)
    hbg = hb + grow
    Vt, Ft = _t(V), _t(F)
    ws = dev.workspace(rk.lib().rk_fast_build_workspace_size(n, rk.RK_MORTON_30))
    b = dev.FastBuilder(n, rk.RK_MORTON_30)
    keys = torch.empty(n, dtype=torch.int32, device="cuda")
    info = torch.zeros(rk.RK_INFO_WORDS, dtype=torch.int32, device="cuda")
    desc = rk.FastBVH(b.nodes.data_ptr(), b.tris.data_ptr(), b.indices.data_ptr(), None, n, 0)
    rk.check(rk.lib().rk_fast_build_keyed(None, Ft.data_ptr(), Vt.data_ptr(), n, rk.C.byref(rk.aabox(hbg)),
                                          rk.RK_MORTON_30, 4, rk.C.byref(desc), info.data_ptr(), keys.data_ptr(),
                                          ws.data_ptr(), ws.numel()), "build_keyed")
    Vd = dev.deform(Vt, amp, 6.2831853, 0.37, 99, kind == "terrain")
    rk.check(rk.lib().rk_fast_refit(None, Ft.data_ptr(), Vd.data_ptr(), n, rk.RK_MORTON_30, 4, rk.C.byref(desc),
                                    keys.data_ptr(), ws.data_ptr(), ws.numel()), "refit")
    refit = dev.FastTree(b.nodes, b.tris, b.indices, b.boxes, b.info, desc)
    cam = _camera(rk, V, 64)
    rf, ruv, rdep = (_h(x) for x in dev.fast_trace_camera(refit, cam, 64, 64, det_epsilon=0.0))
    rorder = _h(b.indices, np.uint32).copy()
    fresh = dev.FastBuilder(n, rk.RK_MORTON_30).build(Ft, Vd, hbg, 4)
    ff, fuv, fdep = (_h(x) for x in dev.fast_trace_camera(fresh, cam, 64, 64, det_epsilon=0.0))
    forder = _h(fresh.indices, np.uint32)
    np.testing.assert_array_equal(rdep, fdep)
    hit = fdep < F32MAX
    assert hit.mean() > 0.05
    np.testing.assert_array_equal(rorder[rf.view(np.uint32)[hit]], forder[ff.view(np.uint32)[hit]])
    np.testing.assert_array_equal(ruv[hit], fuv[hit])
    Vh = _h(Vd)
    O, D = rk.primary_rays_host(cam, 64, 64)
    of, ouv, od = ora.trace_bruteforce(F, Vh, O, D, det_epsilon=0.0)
    np.testing.assert_array_equal(fdep, od)
    np.testing.assert_array_equal(forder[ff.view(np.uint32)[hit]], of[hit])


def _scenes(rk, count, seed):
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        if k % 2:
            V, F = rk.heightfield(int(rng.integers(20, 120)), seed=seed + k)
            out.append((V, F, True))
        else:
            V, F = rk.soup(int(rng.integers(1000, 30000)), seed=seed + k, mu=float(np.log(0.03)))
            out.append((V, F, False))
    return out


def test_served_frames_match_oracle(dev, rk, ora):
    import torch
    scenes = _scenes(rk, 3, 11)
    srv = dev.Server(1 << 30, slots=2, rebuild_every=2, det_epsilon=0.0)
    ids = [srv.add_scene(V, F, z) for V, F, z in scenes]
    jobs = []
    for k, t in enumerate([0.0, 0.21, 0.45, 0.8, 1.3]):
        s = ids[k % len(ids)]
        V, F, z = scenes[s]
        side = 48
        cam = _camera(rk, V, side, az=0.3 + k)
        out = dev.new_outputs(side * side)
        jobs.append((srv.submit(s, t, cam, side, side, out=out), s, t, cam, side, out))
    srv.wait_all()
    for ticket, s, t, cam, side, out in jobs:
        V, F, z = scenes[s]
        res = srv.result(ticket)
        hb = rk.hitbox(V)
        # the server's float32 product amplitude_frac * extent
        amp = float(np.float32(srv.cfg.amplitude_frac) * np.float32((hb[1, :3] - hb[0, :3]).max()))
        Vd = _h(dev.deform(_t(V), amp, float(srv.cfg.omega), t, 0x5EED0000 + s, z))
        O, D = rk.primary_rays_host(cam, side, side)
        of, ouv, od = ora.trace_bruteforce(F, Vd, O, D, det_epsilon=0.0)
        np.testing.assert_array_equal(_h(out[2]), od)
        orig = np.where(od < F32MAX, of, 0xFFFFFFFF).astype(np.uint32)
        ouv = np.where((od < F32MAX)[:, None], ouv, 0).astype(np.float32)
        assert (res.checksum, res.hits) == _checksum(orig, ouv, od)
        torch.cuda.synchronize()
    srv.close()


def test_served_frames_same_alone_and_in_mix(dev, rk):
    """60 requests over 8 scenes on 4 slots with an arena too small for all
    scenes (evictions), then each request alone on a 1-slot server, half of
    them forced to refit and half to rebuild: every checksum is identical."""
    scenes = _scenes(rk, 8, 29)
    sizes = [rk.lib().rk_server_scene_bytes(V.shape[0], F.shape[0]) for V, F, _ in scenes]
    arena = max(max(sizes), int(0.5 * sum(sizes)))
    mix = dev.Server(arena, slots=4, rebuild_every=3, det_epsilon=0.0)
    alone = dev.Server(2 * sum(sizes), slots=1, rebuild_every=3, det_epsilon=0.0)
    for V, F, z in scenes:
        mix.add_scene(V, F, z)
        alone.add_scene(V, F, z)
    rng = np.random.default_rng(3)
    reqs = []
    for k in range(60):
        s = int(rng.integers(0, len(scenes)))
        side = int(rng.integers(16, 96))
        reqs.append((s, float(rng.uniform(0, 2)), _camera(rk, scenes[s][0], side, az=float(rng.uniform(0, 6.3))),
                     side))
    tickets = [mix.submit(s, t, cam, side, side) for s, t, cam, side in reqs]
    mix.wait_all()
    got = [mix.result(t) for t in tickets]
    assert sum(r.evictions for r in got) > 0
    assert len({r.slot for r in got}) > 1
    for k, (s, t, cam, side) in enumerate(reqs):
        flags = 2 if k % 2 else 1  # force refit / force rebuild
        a = alone.submit(s, t, cam, side, side, flags=flags)
        alone.wait_all()
        r = alone.result(a)
        assert (r.checksum, r.hits) == (got[k].checksum, got[k].hits), k
    mix.close()
    alone.close()


def test_plain_path_same_frames(dev, rk):
    """The plain per-request path (rebuild from scratch, ray arrays, separate
    checksum, one request at a time) gives every request of a concurrent mix
    the same checksum, frame by frame."""
    scenes = _scenes(rk, 5, 53)
    sizes = [rk.lib().rk_server_scene_bytes(V.shape[0], F.shape[0]) for V, F, _ in scenes]
    mix = dev.Server(2 * sum(sizes), slots=4, rebuild_every=3, det_epsilon=0.0)
    plain = dev.Server(2 * sum(sizes), slots=1, det_epsilon=0.0, plain=1)
    for V, F, z in scenes:
        mix.add_scene(V, F, z)
        plain.add_scene(V, F, z)
    rng = np.random.default_rng(5)
    reqs = []
    for k in range(30):
        s = int(rng.integers(0, len(scenes)))
        side = int(rng.integers(16, 96))
        reqs.append((s, float(rng.uniform(0, 2)), _camera(rk, scenes[s][0], side, az=float(rng.uniform(0, 6.3))),
                     side))
    tickets = [mix.submit(s, t, cam, side, side) for s, t, cam, side in reqs]
    mix.wait_all()
    ptickets = [plain.submit(s, t, cam, side, side) for s, t, cam, side in reqs]
    plain.wait_all()
    for k, (a, b) in enumerate(zip(tickets, ptickets)):
        ra, rb = mix.result(a), plain.result(b)
        assert rb.rebuilt == 1
        assert (ra.checksum, ra.hits) == (rb.checksum, rb.hits), k
    mix.close()
    plain.close()


def test_poisson_driver_statistics(dev, rk):
    scenes = _scenes(rk, 6, 41)
    sizes = [rk.lib().rk_server_scene_bytes(V.shape[0], F.shape[0]) for V, F, _ in scenes]
    srv = dev.Server(max(max(sizes), int(0.6 * sum(sizes))), slots=4, max_pixels=128 * 128)
    for V, F, z in scenes:
        srv.add_scene(V, F, z)
    st, first = srv.run_poisson(4000.0, 200, 16, 128, 7)
    assert st.completed == 200 and st.rebuilds + st.refits == 200
    assert 0 < st.p50_ms <= st.p99_ms <= st.max_ms
    assert st.rays > 200 * 16 * 16 and st.mrays_per_s > 0
    assert st.uploads >= len(scenes) and st.evictions > 0
    res = [srv.result(first + k) for k in range(200)]
    assert all(r.done >= r.start >= r.arrival for r in res)
    srv.close()


def test_server_rejects_what_cannot_fit(dev, rk):
    V, F = rk.soup(5000, seed=1)
    need = rk.lib().rk_server_scene_bytes(V.shape[0], F.shape[0])
    srv = dev.Server(need // 2, slots=1, max_pixels=64 * 64)
    with pytest.raises(rk.RkError):
        srv.add_scene(V, F, False)
    V2, F2 = rk.soup(100, seed=2)
    s = srv.add_scene(V2, F2, False)
    with pytest.raises(rk.RkError):
        srv.submit(s, 0.0, _camera(rk, V2, 65), 65, 65)
    srv.close()


def _server_hitbox(V, amp):
    """The server's Morton box of a scene: the vertex hitbox grown by the
    deformation amplitude on every side (rk_server_add_scene)."""
    hb = np.zeros((2, 4), np.float32)
    hb[0, :3] = V[:, :3].min(0) - np.float32(amp)
    hb[1, :3] = V[:, :3].max(0) + np.float32(amp)
    return hb


def test_served_frames_at_full_scale(dev, rk, ora):
    """The serving workload's largest sizes: a 2,000,000-face terrain and a
    2,000,000-face soup at 1024 x 1024, under an HBM budget that holds one of
    them, so the third request (terrain again) evicts the soup, uploads and
    rebuilds. Every 37th pixel of each frame equals the oracle's closest hit on
    the deformed scene; each result reports a clean build."""
    terrain = rk.heightfield(1001, seed=0xC5)
    soup = rk.soup(2_000_000, seed=0xC5)
    scenes = [(terrain[0], terrain[1], True), (soup[0], soup[1], False)]
    sizes = [rk.lib().rk_server_scene_bytes(V.shape[0], F.shape[0]) for V, F, _ in scenes]
    srv = dev.Server(int(1.2 * max(sizes)), slots=2, det_epsilon=0.0)
    ids = [srv.add_scene(V, F, z) for V, F, z in scenes]
    side = 1024
    jobs = []
    for k, (s, t) in enumerate([(0, 0.13), (1, 0.29), (0, 0.61)]):
        V, F, z = scenes[s]
        cam = _camera(rk, V, side, az=0.4 + k)
        out = dev.new_outputs(side * side)
        print(f"full-scale request {k}: scene {s}, {F.shape[0]} faces, {side}x{side}", flush=True)
        jobs.append((srv.submit(ids[s], t, cam, side, side, out=out), s, t, cam, out))
    srv.wait_all()
    res = [srv.result(j[0]) for j in jobs]
    assert [r.status for r in res] == [0, 0, 0] and [r.build_flags for r in res] == [0, 0, 0]
    assert res[2].evictions >= 1 and res[2].uploaded == 1 and res[2].rebuilt == 1
    sel = np.arange(0, side * side, 37)
    for (ticket, s, t, cam, out), r in zip(jobs, res):
        V, F, z = scenes[s]
        hb = rk.hitbox(V)
        amp = float(np.float32(srv.cfg.amplitude_frac) * np.float32((hb[1, :3] - hb[0, :3]).max()))
        Vd = _h(dev.deform(_t(V), amp, float(srv.cfg.omega), t, 0x5EED0000 + s, z))
        O, D = rk.primary_rays_host(cam, side, side)
        Os, Ds = np.ascontiguousarray(O[sel]), np.ascontiguousarray(D[sel])
        otree = ora.build_lbvh_ref(F, Vd, rk.hitbox(Vd), 31, True)
        of, ouv, od = ora.trace_closest(otree, Vd, Os, Ds, det_epsilon=0.0)
        fid, uv, dep = _h(out[0], np.uint32), _h(out[1]), _h(out[2])
        np.testing.assert_array_equal(dep[sel], od)
        hit = od < F32MAX
        assert hit.sum() > 1000
        np.testing.assert_array_equal(uv[sel][hit], ouv[hit])
        # every request here rebuilds, so the frame's sorted positions are those
        # of a fresh build over the deformed scene in the server's Morton box
        tree = dev.FastBuilder(F.shape[0], rk.RK_MORTON_30).build(_t(F), _t(Vd), _server_hitbox(V, amp), 1)
        order = _h(tree.indices, np.uint32)
        orig_sel = order[fid[sel][hit]]
        np.testing.assert_array_equal(orig_sel, of[hit])
        orig = np.where(dep < F32MAX, order[np.minimum(fid, F.shape[0] - 1)], 0xFFFFFFFF).astype(np.uint32)
        assert (r.checksum, r.hits) == _checksum(orig, np.where((dep < F32MAX)[:, None], uv, 0).astype(np.float32),
                                                 dep)
    srv.close()
