"""Batches of meshes (rk_batch_build / rk_batch_trace), graph capture and the
autotuner, through the C ABI: every ray's closest hit must equal the
oracle's for its own mesh, bit for bit (t, u, v, original face id)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

F32MAX = np.finfo(np.float32).max


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


def _small_batch(rk, seed=5):
    """Meshes of 1 to 5000 faces (soups, a heightfield, an icosphere), each
    with its own rays, some with none."""
    from cudart_amd import scenes
    rng = np.random.default_rng(seed)
    meshes = []
    for i, n in enumerate([1, 2, 3, 17, 64, 100, 1000, 3000, 5000]):
        V, F = rk.soup(n, seed=seed * 100 + i, mu=float(np.log(0.08)))
        meshes.append((V, F))
    meshes.append(rk.heightfield(40))
    meshes.append(rk.icosphere(3))
    ray_counts = [int(x) for x in rng.integers(0, 700, len(meshes))]
    ray_counts[3] = 0
    Vs, Fs, Os, Ds, hbs = [], [], [], [], []
    for k, (V, F) in enumerate(meshes):
        hb = rk.hitbox(V)
        lo, hi = hb[0, :3].min(), hb[1, :3].max()
        O, D = rk.random_rays(ray_counts[k], seed=seed * 1000 + k, lo=float(lo), hi=float(hi))
        Vs.append(V), Fs.append(F), Os.append(O), Ds.append(D), hbs.append(hb)
    table = scenes.batch_table([f.shape[0] for f in Fs], [v.shape[0] for v in Vs], ray_counts)
    cat = lambda xs, w: np.concatenate(xs) if sum(x.shape[0] for x in xs) else np.zeros((0, w), xs[0].dtype)
    return table, meshes, np.concatenate(Fs), np.concatenate(Vs), np.stack(hbs), cat(Os, 4), cat(Ds, 4)


def _oracle_per_mesh(ora, table, meshes, O, D, det_epsilon):
    n = O.shape[0]
    fid = np.zeros(n, np.uint32)
    uv = np.zeros((n, 2), np.float32)
    dep = np.full(n, F32MAX, np.float32)
    for m, (V, F) in enumerate(meshes):
        r0, rc = int(table["ray_offset"][m]), int(table["ray_count"][m])
        if rc == 0:
            continue
        f, u, d = ora.trace_bruteforce(F, V, np.ascontiguousarray(O[r0:r0 + rc]), np.ascontiguousarray(D[r0:r0 + rc]),
                                       det_epsilon=det_epsilon)
        fid[r0:r0 + rc], uv[r0:r0 + rc], dep[r0:r0 + rc] = f + table["face_offset"][m], u, d
    return fid, uv, dep


def _check(batch, out, ref):
    fid, uv, dep = (_h(x) for x in out)
    rf, ruv, rdep = ref
    np.testing.assert_array_equal(dep, rdep)
    hit = rdep < F32MAX
    order = _h(batch.indices, np.uint32)
    np.testing.assert_array_equal(order[fid.view(np.uint32)[hit]], rf[hit])
    np.testing.assert_array_equal(uv[hit], ruv[hit])
    return hit.mean()


@pytest.mark.parametrize("leaf_max", [1, 4, 16])
def test_batch_every_trace_config_matches_oracle(dev, rk, ora, leaf_max):
    table, meshes, F, V, hb, O, D = _small_batch(rk)
    b = dev.Batch(table, _t(F), _t(V), hb, _t(O), _t(D))
    ref = _oracle_per_mesh(ora, table, meshes, O, D, 0.0)
    for ray_order in (0, 1):
        for mode in (0, 1):
            for wg in (64, 128, 256):
                for stack in (64, 96):
                    for donate in (0, 32):
                        c = dev.batch_config(leaf_max, mode, wg, stack, ray_order, donate)
                        b.build(c)
                        out = dev.new_outputs(O.shape[0])
                        b.trace(out, c, det_epsilon=0.0)
                        # 0.209 of these rays hit
                        assert _check(b, out, ref) > 0.15, (ray_order, mode, wg, stack, donate)


def test_batch_equals_plain_per_mesh(dev, rk):
    """Batched results equal rk_fast_build + rk_fast_trace per mesh."""
    import torch
    from cudart_amd import scenes
    table, F, V, hb, O, D = scenes.soup_batch(48, seed=77, faces=(64, 8192), rays=(64, 2048))
    b = dev.Batch(table, _t(F), _t(V), hb, _t(O), _t(D))
    b.build()
    out = b.trace(dev.new_outputs(O.shape[0]), det_epsilon=0.0)
    bf, buv, bdep = (_h(x) for x in out)
    border = _h(b.indices, np.uint32)
    builder = dev.FastBuilder(int(table["face_count"].max()), rk.RK_MORTON_30)
    for m in range(table.shape[0]):
        f0, fc = int(table["face_offset"][m]), int(table["face_count"][m])
        v0, vc = int(table["vertex_offset"][m]), int(table["vertex_count"][m])
        r0, rc = int(table["ray_offset"][m]), int(table["ray_count"][m])
        tree = builder.build(_t(F[f0:f0 + fc]), _t(V[v0:v0 + vc]), hb[m], 4)
        pf, puv, pdep = (_h(x) for x in dev.fast_trace(tree, _t(O[r0:r0 + rc]), _t(D[r0:r0 + rc]), det_epsilon=0.0))
        np.testing.assert_array_equal(bdep[r0:r0 + rc], pdep)
        hit = pdep < F32MAX
        porder = _h(tree.indices, np.uint32)
        np.testing.assert_array_equal(border[bf.view(np.uint32)[r0:r0 + rc][hit]] - f0, porder[pf.view(np.uint32)[hit]])
        np.testing.assert_array_equal(buv[r0:r0 + rc][hit], puv[hit])
        torch.cuda.synchronize()


@pytest.mark.parametrize("ray_order", [0, 1])
def test_batch_graph_replays_match_eager(dev, rk, ray_order):
    """Build + trace captured once; each replay, with eager work and
    synchronisation in between, reproduces the eager results (with
    ray_order 1 the ray sort is captured too)."""
    import torch
    from cudart_amd import native, scenes
    table, F, V, hb, O, D = scenes.soup_batch(64, seed=91, faces=(64, 4096), rays=(64, 1024))
    b = dev.Batch(table, _t(F), _t(V), hb, _t(O), _t(D))
    b.config = dev.batch_config(ray_order=ray_order)
    nr = O.shape[0]
    b.build()
    eager = b.trace(dev.new_outputs(nr), det_epsilon=0.0)
    eager = [_h(x) for x in eager]
    out = dev.new_outputs(nr)
    s = torch.cuda.Stream()
    g = dev.Graph(s)
    torch.cuda.synchronize()
    with torch.cuda.stream(s):
        with g.capture():
            native.check(native.lib().rk_fill_f32(dev._stream_ptr(s), out[2].data_ptr(), nr, float(F32MAX)), "fill")
            native.check(native.lib().rk_fill_u32(dev._stream_ptr(s), out[0].data_ptr(), nr, 0), "fill")
            b.build(stream=s)
            b.trace(out, det_epsilon=0.0, stream=s)
    for rep in range(3):
        out[0].fill_(7)
        out[1].fill_(-1.0)
        out[2].fill_(-1.0)
        torch.cuda.synchronize()
        b.indices.fill_(0)  # the replay rebuilds the tree too
        torch.cuda.synchronize()
        g.launch()
        s.synchronize()
        got = [_h(x) for x in out]
        hit = eager[2] < F32MAX
        np.testing.assert_array_equal(got[2], eager[2])
        np.testing.assert_array_equal(got[0][hit], eager[0][hit])
        np.testing.assert_array_equal(got[1][hit], eager[1][hit])
        np.testing.assert_array_equal(got[0][~hit], 0)  # the captured fill ran


def test_batch_autotune(dev, rk):
    from cudart_amd import scenes
    table, F, V, hb, O, D = scenes.soup_batch(32, seed=23, faces=(64, 4096), rays=(256, 2048))
    b = dev.Batch(table, _t(F), _t(V), hb, _t(O), _t(D))
    b.build()
    ref = [_h(x) for x in b.trace(dev.new_outputs(O.shape[0]), det_epsilon=0.0)]
    out = dev.new_outputs(O.shape[0])
    best, ms = b.autotune(out, reps=2, det_epsilon=0.0)
    assert best.leaf_max in (1, 2, 4, 8) and best.trace_mode in (0, 1) and best.trace_wg in (64, 128, 256)
    assert best.stack_depth in (64, 96) and best.ray_order in (0, 1) and best.donate_min in (16, 32) and ms > 0
    got = [_h(x) for x in out]
    np.testing.assert_array_equal(got[2], ref[2])  # depths do not depend on the tree


def test_batch_4096_meshes_sampled(dev, rk, ora):
    """The full 4096-mesh batch (log-uniform 64..65536 faces and 256..16384
    rays per mesh): every ray of every 128th mesh against the oracle through
    its own tree of that mesh."""
    from cudart_amd import scenes
    table, F, V, hb, O, D = scenes.soup_batch()
    b = dev.Batch(table, _t(F), _t(V), hb, _t(O), _t(D))
    b.build()
    fid, uv, dep = (_h(x) for x in b.trace(dev.new_outputs(O.shape[0])))
    # rays sorted by origin within each mesh: the same results
    sf, suv, sdep = (_h(x) for x in b.trace(dev.new_outputs(O.shape[0]), dev.batch_config(ray_order=1)))
    np.testing.assert_array_equal(sf, fid)
    np.testing.assert_array_equal(suv, uv)
    np.testing.assert_array_equal(sdep, dep)
    del sf, suv, sdep
    order = _h(b.indices, np.uint32)
    hits = 0
    for m in range(0, table.shape[0], 128):
        f0, fc = int(table["face_offset"][m]), int(table["face_count"][m])
        v0, vc = int(table["vertex_offset"][m]), int(table["vertex_count"][m])
        r0, rc = int(table["ray_offset"][m]), int(table["ray_count"][m])
        Fm, Vm = np.ascontiguousarray(F[f0:f0 + fc]), np.ascontiguousarray(V[v0:v0 + vc])
        otree = ora.build_lbvh_ref(Fm, Vm, hb[m], 31, True)
        of, ouv, od = ora.trace_closest(otree, Vm, np.ascontiguousarray(O[r0:r0 + rc]),
                                        np.ascontiguousarray(D[r0:r0 + rc]))
        np.testing.assert_array_equal(dep[r0:r0 + rc], od)
        hit = od < F32MAX
        np.testing.assert_array_equal(order[fid.view(np.uint32)[r0:r0 + rc][hit]] - f0, of[hit])
        np.testing.assert_array_equal(uv[r0:r0 + rc][hit], ouv[hit])
        hits += int(hit.sum())
    assert hits > 10000
