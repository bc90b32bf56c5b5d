"""Parity of the gfx950 path (through the C ABI) with the oracle, bit-exact:
boxes, Morton codes, sort permutations with the reference tie order, the
reference LBVH (topology, boxes, indices, reordered faces, real_depth), both
reference traversals, and the fast tree's closest hits."""
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
    elif a.dtype == np.uint64:
        a = a.view(np.int64)
    return torch.from_numpy(a).to("cuda")


def _h(t, dtype=None):
    import torch
    torch.cuda.synchronize()
    a = t.cpu().numpy()
    if dtype is not None:
        a = a.view(dtype)
    return a


def _scene(rk, name):
    if name == "ico3":
        V, F = rk.icosphere(3)
    elif name == "ico5":
        V, F = rk.icosphere(5)
    elif name == "hf128":
        V, F = rk.heightfield(128)
    elif name == "hf32":
        V, F = rk.heightfield(32)
    elif name == "hf160":
        V, F = rk.heightfield(160)
    elif name == "hf708":
        V, F = rk.heightfield(708)
    elif name == "soup50k":
        V, F = rk.soup(50000, seed=11)
    elif name == "tiny":
        V, F = rk.soup(37, seed=2)
    else:
        raise KeyError(name)
    return V, F


@pytest.mark.parametrize("name", ["ico3", "hf128", "soup50k"])
def test_boxes_and_morton(dev, rk, ora, name):
    V, F = _scene(rk, name)
    hb = rk.hitbox(V)
    boxes, idx = dev.calc_mesh_boxes(_t(F), _t(V))
    ob, oi = ora.calc_boxes(F, V)
    np.testing.assert_array_equal(_h(boxes), ob)
    np.testing.assert_array_equal(_h(idx, np.uint32), oi)
    m32 = dev.morton(boxes, hb, rk.RK_MORTON_REF32)
    np.testing.assert_array_equal(_h(m32, np.uint32), ora.morton32(ob, hb))
    m30 = dev.morton(boxes, hb, rk.RK_MORTON_30)
    np.testing.assert_array_equal(_h(m30, np.uint32).astype(np.uint64), ora.morton_wide(ob, hb, 10))
    m63 = dev.morton(boxes, hb, rk.RK_MORTON_63)
    np.testing.assert_array_equal(_h(m63, np.uint64), ora.morton_wide(ob, hb, 21))


@pytest.mark.parametrize("n", [1, 2, 3, 100, 511, 512, 513, 700, 4096, 20480, 32768, 32769, 100000, 1 << 20])
@pytest.mark.parametrize("dist", ["dups", "wide"])
def test_sort_general_reference_tie_order(dev, ora, n, dist):
    rng = np.random.default_rng(n * 7 + len(dist))
    if dist == "dups":
        keys = rng.integers(0, 5, n, dtype=np.uint32)
    else:
        keys = rng.integers(0, 2 ** 32, n, dtype=np.uint64).astype(np.uint32)
        keys[::17] = keys[0]
    ok, ov, _ = ora.sort_pairs_ref(keys, np.zeros(n, np.uint32), 32, init=True)
    k, v = _t(keys), _t(np.zeros(n, np.uint32))
    dev.sort_general(k, v, 32, init=True)
    np.testing.assert_array_equal(_h(k, np.uint32), ok)
    np.testing.assert_array_equal(_h(v, np.uint32), ov)


@pytest.mark.parametrize("n,bits", [(9000, 8), (70000, 8), (50000, 21), (300000, 16)])
def test_sort_general_low_byte_classes(dev, ora, n, bits):
    rng = np.random.default_rng(n + bits)
    keys = rng.integers(0, 1 << bits, n, dtype=np.uint32)
    vals = rng.integers(0, 1 << 31, n, dtype=np.uint32)
    ok, ov, _ = ora.sort_pairs_ref(keys, vals, bits, init=False)
    k, v = _t(keys), _t(vals)
    dev.sort_general(k, v, bits, init=False)
    np.testing.assert_array_equal(_h(k, np.uint32), ok)
    np.testing.assert_array_equal(_h(v, np.uint32), ov)


@pytest.mark.parametrize("n", [5, 1000, 40000, 500000])
def test_sort_values(dev, n):
    rng = np.random.default_rng(n)
    vals = rng.permutation(n * 3)[:n].astype(np.uint32)
    t = _t(vals)
    dev.sort_values(t)
    np.testing.assert_array_equal(_h(t, np.uint32), np.sort(vals))


@pytest.mark.parametrize("n", [2, 777, 65537, 1 << 20])
def test_sort_pairs_u64_stable(dev, ora, n):
    rng = np.random.default_rng(n)
    keys = rng.integers(0, 2 ** 63, n, dtype=np.uint64)
    keys[::5] = keys[1]
    ok, ov = ora.sort_pairs_u64(keys, np.arange(n, dtype=np.uint32))
    k, v = _t(keys), _t(np.arange(n, dtype=np.uint32))
    dev.sort_pairs_u64(k, v)
    np.testing.assert_array_equal(_h(k, np.uint64), ok)
    np.testing.assert_array_equal(_h(v, np.uint32), ov)


def _compare_tree(dev, rk, ora, V, F, depth, complete):
    hb = rk.hitbox(V)
    ref = ora.build_lbvh_ref(F, V, hb, depth, complete)
    Ft = _t(F)
    tree = dev.build_lbvh(Ft, _t(V), hb, depth, complete)
    info = tree.host_info()
    M = int(info[rk.RK_INFO_NODE_COUNT])
    assert M == ref.node_count
    assert int(info[rk.RK_INFO_REAL_DEPTH]) == ref.real_depth
    assert int(info[rk.RK_INFO_FLAGS]) & 0xFF == ref.flags
    np.testing.assert_array_equal(_h(tree.indices, np.uint32), ref.indices)
    np.testing.assert_array_equal(_h(tree.boxes), ref.boxes)
    np.testing.assert_array_equal(_h(Ft, np.uint32), ref.faces)
    nodes = _h(tree.nodes, np.uint32)[:M].copy().view(ora.NODE_DTYPE).reshape(M)
    np.testing.assert_array_equal(nodes["w0"], ref.nodes["w0"])
    np.testing.assert_array_equal(nodes["w1"], ref.nodes["w1"])
    # boxes compare as floats: -0 == +0 (min/max of signed zeros may differ)
    assert np.array_equal(nodes["bmin"], ref.nodes["bmin"])
    assert np.array_equal(nodes["bmax"], ref.nodes["bmax"])
    return tree, ref


@pytest.mark.parametrize("name", ["tiny", "ico3", "ico5", "hf128", "soup50k"])
@pytest.mark.parametrize("depth", [0, 4, 12, 31])
def test_reference_lbvh_bit_exact(dev, rk, ora, name, depth):
    V, F = _scene(rk, name)
    _compare_tree(dev, rk, ora, V, F, depth, True)


@pytest.mark.parametrize("name", ["ico5", "soup50k"])
def test_reference_lbvh_incomplete(dev, rk, ora, name):
    V, F = _scene(rk, name)
    _compare_tree(dev, rk, ora, V, F, 12, False)


def test_reference_lbvh_heightfield_full_size(dev, rk, ora):
    V, F = _scene(rk, "hf708")
    assert F.shape[0] == 999698
    _compare_tree(dev, rk, ora, V, F, 12, True)


def _camera_rays(rk, name, w, h):
    if name.startswith("ico"):
        cam = rk.camera_lookat([0, 0, 3], [0, 0, 0], 60 * 0.01745329252, w, h)
    else:
        el, az, dist = np.radians(30.0), np.radians(45.0), 1.5
        tgt = np.array([0.5, 0.5, 0.1])
        eye = tgt + dist * np.array([np.cos(el) * np.cos(az), np.cos(el) * np.sin(az), np.sin(el)])
        cam = rk.camera_lookat(eye, tgt, 60 * 0.01745329252, w, h, up=[0, 0, 1])
    return rk.primary_rays_host(cam, w, h)


@pytest.fixture(params=["default", "all", "all_cap7"])
def handoff(request, dev):
    """The reference traversals' second passes: at their default limits; with
    every ray or packet that reaches a non-empty leaf handed off; and the same
    with a hand-off list of 7 entries, so that the rest finish in the first pass."""
    if request.param == "all":
        dev.set_ref_handoff(0, 0, 65536)
    elif request.param == "all_cap7":
        dev.set_ref_handoff(0, 0, 7)
    yield request.param
    dev.set_ref_handoff()


def test_reference_handoff_limits(dev):
    assert dev.get_ref_handoff() == (0xFFFFFFFE, 0xFFFFFFFE, 65536)
    dev.set_ref_handoff(1, 2, 3)
    assert dev.get_ref_handoff() == (1, 2, 3)
    with pytest.raises(Exception):
        dev.set_ref_handoff(1, 2, 0)
    with pytest.raises(Exception):
        dev.set_ref_handoff(1, 2, 65537)
    dev.set_ref_handoff()
    assert dev.get_ref_handoff() == (0xFFFFFFFE, 0xFFFFFFFE, 65536)


@pytest.mark.parametrize("name,w,h", [("ico5", 256, 256), ("hf128", 200, 150), ("ico3", 33, 29),
                                      ("hf708", 1920, 1080)])
@pytest.mark.parametrize("coherency", [False, True])
def test_reference_trace_bit_exact(dev, rk, ora, name, w, h, coherency, handoff):
    """Both reference traversals (per-ray and packet) against the oracle's
    restatement of each, every ray, up to the full 1920 x 1080 heightfield frame,
    with and without their second passes."""
    V, F = _scene(rk, name)
    tree, ref = _compare_tree(dev, rk, ora, V, F, 12, True)
    O, D = _camera_rays(rk, name, w, h)
    n = w * h
    rng = np.random.default_rng(3)
    fid0 = rng.integers(0, 1000, n, dtype=np.uint32)
    uv0 = rng.random((n, 2), dtype=np.float32)
    dep0 = np.full(n, F32MAX, np.float32)
    dep0[::7] = 2.5  # some rays start with a finite tmax
    out = (_t(fid0), _t(uv0), _t(dep0))
    dev.trace_ref(tree, _t(V), _t(O), _t(D), w, h, coherency, out=out)
    if coherency:
        of, ouv, od, ovf = ora.trace_ref_packet(ref.nodes, ref.faces, V, O, D, w, h, fid0, uv0, dep0)
    else:
        of, ouv, od, ovf = ora.trace_ref_ray(ref.nodes, ref.faces, V, O, D, fid0, uv0, dep0)
    assert ovf == 0
    np.testing.assert_array_equal(_h(out[0], np.uint32), of)
    np.testing.assert_array_equal(_h(out[1]), ouv)
    np.testing.assert_array_equal(_h(out[2]), od)
    assert (od < dep0).sum() > n // 20


@pytest.mark.parametrize("depth", [4, 12])
@pytest.mark.parametrize("coherency", [False, True])
def test_reference_trace_duplicate_faces(dev, rk, ora, depth, coherency, handoff):
    """Every face three times: identical triangles hit at identical t, so the
    reference's strict t < depth keeps the first copy in leaf order. Depth 4
    gives leaves of about 960 faces (many rounds of 64 in the per-ray kernel,
    copies on both sides of a round's boundary and of the parts of a chunk in
    the packet traversal's second pass), depth 12 small ones."""
    V, F0 = rk.icosphere(4)
    F = np.ascontiguousarray(np.repeat(F0, 3, axis=0))
    tree, ref = _compare_tree(dev, rk, ora, V, F, depth, True)
    w, h = 96, 80
    cam = rk.camera_lookat([0.3, 0.2, 3], [0, 0, 0], 60 * 0.01745329252, w, h)
    O, D = rk.primary_rays_host(cam, w, h)
    n = w * h
    fid0 = np.zeros(n, np.uint32)
    uv0 = np.zeros((n, 2), np.float32)
    dep0 = np.full(n, F32MAX, np.float32)
    out = (_t(fid0), _t(uv0), _t(dep0))
    dev.trace_ref(tree, _t(V), _t(O), _t(D), w, h, coherency, out=out)
    if coherency:
        of, ouv, od, ovf = ora.trace_ref_packet(ref.nodes, ref.faces, V, O, D, w, h, fid0, uv0, dep0)
    else:
        of, ouv, od, ovf = ora.trace_ref_ray(ref.nodes, ref.faces, V, O, D, fid0, uv0, dep0)
    assert ovf == 0
    hit = od < F32MAX
    assert hit.mean() > 0.15
    np.testing.assert_array_equal(_h(out[0], np.uint32), of)
    np.testing.assert_array_equal(_h(out[1]), ouv)
    np.testing.assert_array_equal(_h(out[2]), od)


@pytest.mark.parametrize("name,w,h,depth,coherency", [("hf128", 200, 150, 12, False), ("ico5", 128, 96, 12, False),
                                                      ("hf128", 200, 150, 31, False), ("hf160", 64, 48, 12, True),
                                                      ("hf160", 64, 48, 31, True), ("ico5", 64, 48, 12, True)])
def test_reference_trace_rays_crossing_every_box(dev, rk, ora, name, w, h, depth, coherency):
    """Rays with a direction component of exactly 0: the reference's normalised
    inverse direction is then (0, 0, NaN) up to order, every box test passes
    at t = 0 and the per-ray traversal tests every face of the mesh (the packet
    traversal: every ray of a packet holding one). The library hands such a ray
    or packet to its second pass at the first leaf, and other rays past their
    face budget; the results stay bit-identical to the oracle's."""
    V, F = _scene(rk, name)
    assert F.shape[0] > 8192
    tree, ref = _compare_tree(dev, rk, ora, V, F, depth, True)
    O, D = _camera_rays(rk, name, w, h)
    O, D = O.copy(), D.copy()
    n = w * h
    D[::5, 2] = 0.0
    D[1::11, 0] = 0.0
    # some of them start at a vertex: a hit at t = 0 on a face there ends their
    # walk's forcing of the packet's, and freezes their own result
    sel = np.arange(0, n, 10)
    O[sel, :3] = V[F[(sel * 7919) % F.shape[0], 0], :3]
    rng = np.random.default_rng(5)
    fid0 = rng.integers(0, 1000, n, dtype=np.uint32)
    uv0 = rng.random((n, 2), dtype=np.float32)
    dep0 = np.full(n, F32MAX, np.float32)
    dep0[::7] = 2.5
    dep0[3::13] = 0.0
    out = (_t(fid0), _t(uv0), _t(dep0))
    dev.trace_ref(tree, _t(V), _t(O), _t(D), w, h, coherency, out=out)
    _, _, _, _, cnt = ora.trace_ref_ray(ref.nodes, ref.faces, V, O, D, fid0, uv0, dep0, counters=True)
    assert (cnt[:, 1] > 8192).sum() > n // 10  # the second pass has work
    if coherency:
        of, ouv, od, ovf = ora.trace_ref_packet(ref.nodes, ref.faces, V, O, D, w, h, fid0, uv0, dep0)
    else:
        of, ouv, od, ovf = ora.trace_ref_ray(ref.nodes, ref.faces, V, O, D, fid0, uv0, dep0)
    assert ovf == 0
    assert (od < dep0).sum() > n // 20
    assert ((od == 0) & (dep0 > 0)).sum() > 0  # hits at t = 0
    np.testing.assert_array_equal(_h(out[0], np.uint32), of)
    np.testing.assert_array_equal(_h(out[1]), ouv)
    np.testing.assert_array_equal(_h(out[2]), od)


def test_reference_trace_two_streams(dev, rk, ora):
    """Both traversals on two streams at once, three times over: each call has
    its own hand-off buffers and triangle records (stream-ordered), so both
    equal the oracle."""
    import torch
    V, F = _scene(rk, "hf128")
    tree, ref = _compare_tree(dev, rk, ora, V, F, 12, True)
    w, h = 200, 150
    O, D = _camera_rays(rk, "hf128", w, h)
    D = D.copy()
    D[::5, 2] = 0.0  # rays that pass every box test: the split path
    n = w * h
    Vt, Ot, Dt = _t(V), _t(O), _t(D)
    streams = (torch.cuda.Stream(), torch.cuda.Stream())
    outs = [dev.new_outputs(n) for _ in range(6)]
    torch.cuda.synchronize()
    for k in range(6):
        dev.trace_ref(tree, Vt, Ot, Dt, w, h, k % 2 == 1, out=outs[k], stream=streams[k % 2])
    torch.cuda.synchronize()
    expect = {False: ora.trace_ref_ray(ref.nodes, ref.faces, V, O, D), True: ora.trace_ref_packet(ref.nodes, ref.faces,
                                                                                                  V, O, D, w, h)}
    for k in range(6):
        of, ouv, od, _ = expect[k % 2 == 1]
        np.testing.assert_array_equal(_h(outs[k][0], np.uint32), of)
        np.testing.assert_array_equal(_h(outs[k][1]), ouv)
        np.testing.assert_array_equal(_h(outs[k][2]), od)


def _chain_tree(V, levels, per_leaf):
    """A tree of `levels` inner nodes in a chain (split axis x, the mesh's box):
    each has a leaf of per_leaf faces and the next inner node (a leaf after the
    last) as its children, the leaf first at even levels and second at odd
    ones. A ray then takes the chain first at every other level, pushing the
    leaf, and tests the leaf first at the others: its stack grows by one entry
    every two levels, and overflows deep enough down."""
    hb = np.asarray(V[:, :3], np.float32)
    lo, hi = hb.min(0), hb.max(0)
    nodes = np.zeros((2 * levels + 1, 8), np.uint32)
    inner, nxt = 0, 1
    for d in range(levels):
        leaf, chain = (nxt, nxt + 1) if d % 2 == 0 else (nxt + 1, nxt)
        nodes[inner, 0] = 0 << 1  # axis x
        nodes[inner, 1] = nxt
        nodes[inner, 2:5] = lo.view(np.uint32)
        nodes[inner, 5:8] = hi.view(np.uint32)
        nodes[leaf, 0] = (per_leaf << 1) | 1
        nodes[leaf, 1] = d * per_leaf
        inner, nxt = chain, nxt + 2
    nodes[inner, 0] = (per_leaf << 1) | 1
    nodes[inner, 1] = levels * per_leaf
    return nodes


@pytest.mark.parametrize("coherency", [False, True])
def test_reference_trace_stack_overflow(dev, rk, ora, coherency, handoff):
    """A tree deeper than the reference's stacks (96 entries per ray, 128 per
    packet): the rays that take the chain first stop where the reference's
    stack would overflow, in the first pass, the second and the split of the
    rays that pass every box test (whose subtree walks start with the
    reference's stack below them)."""
    V, F = rk.heightfield(32)
    levels, per_leaf = 270, 6
    assert (levels + 1) * per_leaf <= F.shape[0]
    nodes = _chain_tree(V, levels, per_leaf)
    w, h = 48, 36
    O, D = _camera_rays(rk, "hf32", w, h)
    O, D = O.copy(), D.copy()
    n = w * h
    D[1::3, 0] = np.abs(D[1::3, 0])
    D[::5, 2] = 0.0  # every box test passes
    D[2::7, 1] = 0.0
    fid0 = np.zeros(n, np.uint32)
    uv0 = np.zeros((n, 2), np.float32)
    dep0 = np.full(n, F32MAX, np.float32)
    from cudart_amd.device import RefTree
    tree = RefTree(_t(F), None, None, _t(nodes), None, node_count=nodes.shape[0])
    out = (_t(fid0), _t(uv0), _t(dep0))
    dev.trace_ref(tree, _t(V), _t(O), _t(D), w, h, coherency, out=out)
    if coherency:
        of, ouv, od, ovf = ora.trace_ref_packet(nodes, F, V, O, D, w, h, fid0, uv0, dep0)
    else:
        of, ouv, od, ovf = ora.trace_ref_ray(nodes, F, V, O, D, fid0, uv0, dep0)
    assert ovf > 0
    assert (od < F32MAX).sum() > 10  # hits in the leaves tested before the stacks overflow
    np.testing.assert_array_equal(_h(out[0], np.uint32), of)
    np.testing.assert_array_equal(_h(out[1]), ouv)
    np.testing.assert_array_equal(_h(out[2]), od)


def test_reference_trace_handoff_list_full(dev, rk, ora):
    """More rays crossing every box than the hand-off list holds (65536 per
    call): the rays past its end finish in the first pass, with the same
    results."""
    V, F = rk.heightfield(66)
    assert F.shape[0] > 8192
    tree, ref = _compare_tree(dev, rk, ora, V, F, 12, True)
    w, h = 480, 180
    O, D = _camera_rays(rk, "hf66", w, h)
    O, D = O.copy(), D.copy()
    D[:, 2] = 0.0
    D[::9, 2] = -0.25
    O[1::2, 2] = np.linspace(0.04, 0.16, O[1::2].shape[0], dtype=np.float32)  # through the terrain (z in [0, 0.2])
    n = w * h
    fid0 = np.zeros(n, np.uint32)
    uv0 = np.zeros((n, 2), np.float32)
    dep0 = np.full(n, F32MAX, np.float32)
    out = (_t(fid0), _t(uv0), _t(dep0))
    dev.trace_ref(tree, _t(V), _t(O), _t(D), w, h, False, out=out)
    of, ouv, od, ovf, cnt = ora.trace_ref_ray(ref.nodes, ref.faces, V, O, D, fid0, uv0, dep0, counters=True)
    assert ovf == 0
    assert (cnt[:, 1] > 8192).sum() > 65536
    assert (od < F32MAX).sum() > n // 20
    np.testing.assert_array_equal(_h(out[0], np.uint32), of)
    np.testing.assert_array_equal(_h(out[1]), ouv)
    np.testing.assert_array_equal(_h(out[2]), od)


def _check_fast_boxes(tree, F, V, leaf_max):
    """Every child slot of every reachable node holds the padded union of the
    boxes of the faces below it, the right face count, and covers them."""
    nodes = _h(tree.nodes, np.uint32)
    order = _h(tree.indices, np.uint32)
    fb = np.stack([V[F[:, 0], :3], V[F[:, 1], :3], V[F[:, 2], :3]], 1)
    lo_f, hi_f = fb.min(1)[order], fb.max(1)[order]
    nf = nodes.view(np.float32)
    n = F.shape[0]
    stack, seen = [0], 0
    while stack:
        i = stack.pop()
        seen += 1
        for side in (0, 1):
            ref = int(nodes[i, 12 + side])
            cnt = int(nodes[i, 14 + side])
            box = nf[i, 4 * side:4 * side + 4]
            lo = np.array([box[0], box[2], nf[i, 8 + 2 * side]])
            hi = np.array([box[1], box[3], nf[i, 9 + 2 * side]])
            if ref & 0x80000000:
                first, c = (ref & 0x7FFFFFFF) >> 4, (ref & 15) + 1
                assert c == cnt and c <= leaf_max and first + c <= n
                rng = slice(first, first + c)
            else:
                stack.append(ref)
                rng = None
            if rng is not None:
                assert np.all(lo <= lo_f[rng].min(0)) and np.all(hi >= hi_f[rng].max(0))
    assert seen >= max(1, (n - 1) // leaf_max // 2)


def _fast_vs_bruteforce(dev, rk, ora, V, F, O, D, mode, leaf_max, tmax=None, det_epsilon=1e-6, sorted_rays=False):
    hb = rk.hitbox(V)
    b = dev.FastBuilder(F.shape[0], mode)
    tree = b.build(_t(F), _t(V), hb, leaf_max)
    n = O.shape[0]
    fid, uv, dep = dev.new_outputs(n)
    if tmax is not None:
        dep.fill_(tmax)
    trace = dev.fast_trace_sorted if sorted_rays else dev.fast_trace
    trace(tree, _t(O), _t(D), out=(fid, uv, dep), det_epsilon=det_epsilon)
    bf, buv, bdep = ora.trace_bruteforce(F, V, O, D, depths=None if tmax is None else np.full(n, tmax, np.float32),
                                         det_epsilon=det_epsilon)
    g_dep = _h(dep)
    hit = bdep < (F32MAX if tmax is None else tmax)
    np.testing.assert_array_equal(g_dep, bdep)
    g_pos = _h(fid, np.uint32)
    order = _h(tree.indices, np.uint32)
    np.testing.assert_array_equal(order[g_pos[hit]], bf[hit])
    np.testing.assert_array_equal(_h(uv)[hit], buv[hit])
    assert int(_h(tree.info, np.uint32)[rk.RK_INFO_FLAGS]) == 0
    if F.shape[0] > 1:
        _check_fast_boxes(tree, F, V, leaf_max)
    return hit.mean()


@pytest.mark.parametrize("mode", [1, 2])
def test_fast_tree_det_epsilon_zero(dev, rk, ora, mode):
    V, F = _scene(rk, "hf128")
    O, D = _camera_rays(rk, "hf128", 80, 60)
    assert _fast_vs_bruteforce(dev, rk, ora, V, F, O, D, mode, 4, det_epsilon=0.0) > 0.05


@pytest.mark.parametrize("leaf_max", [1, 4])
def test_fast_tree_full_size_heightfield(dev, rk, ora, leaf_max):
    """1M faces, 1920x1080 camera: every 37th ray of the frame against the
    oracle's closest hit through its own depth-31 tree, det_epsilon 0. Leaf 1
    is the benchmark's tree."""
    V, F = _scene(rk, "hf708")
    hb = rk.hitbox(V)
    O, D = _camera_rays(rk, "hf708", 1920, 1080)
    tree = dev.FastBuilder(F.shape[0], 1).build(_t(F), _t(V), hb, leaf_max)
    fid, uv, dep = dev.fast_trace(tree, _t(O), _t(D), det_epsilon=0.0)
    sel = np.arange(0, O.shape[0], 37)
    Os, Ds = np.ascontiguousarray(O[sel]), np.ascontiguousarray(D[sel])
    otree = ora.build_lbvh_ref(F, V, hb, 31, True)
    of, ouv, od = ora.trace_closest(otree, V, Os, Ds, det_epsilon=0.0)
    g_dep = _h(dep)[sel]
    np.testing.assert_array_equal(g_dep, od)
    hit = od < F32MAX
    assert hit.sum() > 1000
    g_orig = _h(tree.indices, np.uint32)[_h(fid, np.uint32)[sel][hit]]
    np.testing.assert_array_equal(g_orig, of[hit])
    np.testing.assert_array_equal(_h(uv)[sel][hit], ouv[hit])


@pytest.mark.parametrize("leaf_max", [1, 4])
def test_fast_tree_soup_10m_63bit(dev, rk, ora, leaf_max):
    """10M-face soup (lognormal edges around 0.002), 63-bit Morton keys sorted
    as uint64 pairs, 16.7M incoherent rays: the sort order equals the
    oracle's stable sort of its 21/21/21 codes, and every 4099th ray's
    closest hit equals the oracle's through its own tree. Leaf 1 is the
    benchmark's tree."""
    n, nr = 10_000_000, 1 << 24
    V, F = rk.soup(n, seed=0xC3)
    hb = rk.hitbox(V)
    O, D = rk.random_rays(nr, seed=0xC3)
    tree = dev.FastBuilder(n, rk.RK_MORTON_63).build(_t(F), _t(V), hb, leaf_max)
    fid, uv, dep = dev.fast_trace(tree, _t(O), _t(D))
    # the same rays traced in origin-Morton order give the same results
    sf, suv, sdep = dev.fast_trace_sorted(tree, _t(O), _t(D))
    np.testing.assert_array_equal(_h(sf), _h(fid))
    np.testing.assert_array_equal(_h(suv), _h(uv))
    np.testing.assert_array_equal(_h(sdep), _h(dep))
    del sf, suv, sdep
    order = _h(tree.indices, np.uint32)
    ob, _ = ora.calc_boxes(F, V)
    codes = ora.morton_wide(ob, hb, 21)
    del ob
    _, oorder = ora.sort_pairs_u64(codes, np.arange(n, dtype=np.uint32))
    np.testing.assert_array_equal(order, oorder)
    del codes, oorder
    sel = np.arange(0, nr, 4099)
    Os, Ds = np.ascontiguousarray(O[sel]), np.ascontiguousarray(D[sel])
    otree = ora.build_lbvh_ref(F, V, hb, 31, True)
    of, ouv, od = ora.trace_closest(otree, V, Os, Ds)
    g_dep = _h(dep)[sel]
    np.testing.assert_array_equal(g_dep, od)
    hit = od < F32MAX
    assert hit.mean() > 0.5
    np.testing.assert_array_equal(order[_h(fid, np.uint32)[sel][hit]], of[hit])
    np.testing.assert_array_equal(_h(uv)[sel][hit], ouv[hit])


@pytest.mark.parametrize("mode", [1, 2])
@pytest.mark.parametrize("leaf_max", [1, 4, 8])
def test_fast_tree_closest_hit_icosphere(dev, rk, ora, mode, leaf_max):
    V, F = _scene(rk, "ico3")
    O, D = _camera_rays(rk, "ico3", 64, 64)
    assert _fast_vs_bruteforce(dev, rk, ora, V, F, O, D, mode, leaf_max) > 0.2


@pytest.mark.parametrize("mode", [1, 2])
def test_fast_tree_closest_hit_soup(dev, rk, ora, mode):
    V, F = rk.soup(3000, seed=9, mu=float(np.log(0.05)))
    O, D = rk.random_rays(4000, seed=10)
    assert _fast_vs_bruteforce(dev, rk, ora, V, F, O, D, mode, 4) > 0.05


def _flat_grid(g):
    """A flat (g-1)x(g-1) grid in z = 0 on [0, 1]^2 with spacing 1/(g-1), two
    faces per cell; with g - 1 a power of two every vertex is exact."""
    xs = np.arange(g, dtype=np.float32) / np.float32(g - 1)
    X, Y = np.meshgrid(xs, xs, indexing="xy")
    V = np.stack([X.ravel(), Y.ravel(), np.zeros(g * g, np.float32), np.zeros(g * g, np.float32)], 1)
    F = []
    for j in range(g - 1):
        for i in range(g - 1):
            a, b, c, d = j * g + i, j * g + i + 1, (j + 1) * g + i, (j + 1) * g + i + 1
            F.append((a, b, d, 0))
            F.append((a, d, c, 0))
    return np.ascontiguousarray(V, np.float32), np.asarray(F, np.uint32)


@pytest.mark.parametrize("leaf_max", [1, 2, 4])
def test_fast_trace_exact_ties_in_shared_waves(dev, rk, ora, leaf_max):
    """Rays through grid vertices hit up to six faces at exactly the same t;
    grazing rays cross the whole grid and keep their lanes busy while the
    others finish and help them. Within each wave the three kinds are mixed,
    so subtrees move between lanes and ties are resolved across lanes: every
    hit must equal brute force's, ties to the smallest original face id."""
    g = 65
    V, F = _flat_grid(g)
    rng = np.random.default_rng(17)
    n = 4096
    O = np.zeros((n, 4), np.float32)
    D = np.zeros((n, 4), np.float32)
    vi = rng.integers(0, g, size=(n, 2)).astype(np.float32) / np.float32(g - 1)
    kind = np.arange(n) % 4
    # straight down through a vertex: t = 1 exactly for every face at it
    O[:, 0], O[:, 1], O[:, 2] = vi[:, 0], vi[:, 1], 1.0
    D[:, 2] = -1.0
    # tilted through a vertex
    tilt = kind == 1
    O[tilt, 0] = vi[tilt, 0] + 0.25
    O[tilt, 1] = vi[tilt, 1] - 0.125
    D[tilt, 0], D[tilt, 1], D[tilt, 2] = -0.25, 0.125, -1.0
    # grazing: across the grid just above it, descending slowly
    graze = kind >= 2
    O[graze, 0] = -0.25
    O[graze, 1] = rng.random(graze.sum()).astype(np.float32)
    O[graze, 2] = 0.0005
    D[graze, 0], D[graze, 1], D[graze, 2] = 1.0, 0.0, -0.0005
    hit = _fast_vs_bruteforce(dev, rk, ora, V, F, O, D, 1, leaf_max, det_epsilon=0.0)
    assert hit > 0.9


@pytest.mark.parametrize("dist", [1e2, 1e3, 1e4, pytest.param(1e6, marks=pytest.mark.xfail(
    strict=True, reason="documented limit: at 1e6 times the scene's size the float32 origin carries ~0.06 units of "
    "rounding, more than the boxes' padding, and hits at box boundaries are lost"))])
def test_fast_trace_far_origins(dev, rk, ora, dist):
    """Rays from far away (dist times the scene's size) aimed at points of a
    small soup: the padded box test stays conservative while the rounding of
    the origin term (|o| * 2^-24) is small against the padding, so every hit
    equals brute force's. At 1e6 times the scene's size it is not (half the
    hits are lost at box boundaries): the documented limit of the fast trace."""
    V, F = rk.soup(2000, seed=21, mu=float(np.log(0.05)))
    rng = np.random.default_rng(int(dist))
    n = 2048
    target = rng.random((n, 3)).astype(np.float64)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    O = np.zeros((n, 4), np.float32)
    D = np.zeros((n, 4), np.float32)
    O[:, :3] = (target - d * dist).astype(np.float32)
    D[:, :3] = d.astype(np.float32)
    hit = _fast_vs_bruteforce(dev, rk, ora, V, F, O, D, 2, 1, det_epsilon=0.0)
    assert hit > 0.02


@pytest.mark.parametrize("nr", [1, 255, 257, 4000])
def test_fast_trace_sorted_closest_hit_soup(dev, rk, ora, nr):
    """Rays traced in origin-Morton order (a key kernel, a radix sort, then
    the trace through the permutation) hit what brute force hits, at ragged
    ray counts."""
    V, F = rk.soup(3000, seed=9, mu=float(np.log(0.05)))
    O, D = rk.random_rays(nr, seed=10)
    _fast_vs_bruteforce(dev, rk, ora, V, F, O, D, 2, 4, sorted_rays=True)


def test_fast_trace_sorted_tmax_and_single_face(dev, rk, ora):
    V, F = _scene(rk, "hf128")
    O, D = _camera_rays(rk, "hf128", 80, 60)
    _fast_vs_bruteforce(dev, rk, ora, V, F, O, D, 1, 4, tmax=1.2, sorted_rays=True)
    V = np.array([[0, 0, 0, 0], [1, 0, 0, 0], [0, 1, 0, 0]], np.float32)
    F = np.array([[0, 1, 2, 0]], np.uint32)
    O = np.array([[0.25, 0.25, 1, 0], [2, 2, 1, 0], [0.1, 0.1, -1, 0]], np.float32)
    D = np.array([[0, 0, -1, 0], [0, 0, -1, 0], [0, 0, 1, 0]], np.float32)
    _fast_vs_bruteforce(dev, rk, ora, V, F, O, D, 1, 4, sorted_rays=True)


def test_fast_trace_sorted_workspace_too_small(dev, rk):
    import torch
    V, F = rk.soup(100, seed=3)
    tree = dev.FastBuilder(100, 1).build(_t(F), _t(V), rk.hitbox(V), 4)
    O, D = rk.random_rays(1000, seed=4)
    ws = torch.empty(64, dtype=torch.uint8, device="cuda")
    Ot, Dt = _t(O), _t(D)
    fid, uv, dep = dev.new_outputs(1000)
    st = dev.N.lib().rk_fast_trace_sorted(None, dev.C.byref(tree.desc), Ot.data_ptr(), Dt.data_ptr(), 1000,
                                          fid.data_ptr(), uv.data_ptr(), dep.data_ptr(), None, ws.data_ptr(),
                                          ws.numel())
    assert st == 4  # RK_ERROR_WORKSPACE


def test_fast_tree_heightfield_and_tmax(dev, rk, ora):
    V, F = _scene(rk, "hf128")
    O, D = _camera_rays(rk, "hf128", 80, 60)
    # the reference's absolute det threshold (1e-6) rejects grazing hits on
    # this 128x128 terrain, so most rays pass through
    assert _fast_vs_bruteforce(dev, rk, ora, V, F, O, D, 1, 4) > 0.05
    _fast_vs_bruteforce(dev, rk, ora, V, F, O, D, 1, 4, tmax=1.2)


def test_fast_tree_single_face(dev, rk, ora):
    V = np.array([[0, 0, 0, 0], [1, 0, 0, 0], [0, 1, 0, 0]], np.float32)
    F = np.array([[0, 1, 2, 0]], np.uint32)
    O = np.array([[0.25, 0.25, 1, 0], [2, 2, 1, 0]], np.float32)
    D = np.array([[0, 0, -1, 0], [0, 0, -1, 0]], np.float32)
    _fast_vs_bruteforce(dev, rk, ora, V, F, O, D, 1, 4)


def test_fast_vs_reference_trace_c1(dev, rk, ora):
    """The fast tree and the reference tree give the same closest hits on the
    256x256 icosphere view, in original face ids, except where two faces are
    hit at exactly the same t (the reference keeps the first one it visits)."""
    V, F = _scene(rk, "ico5")
    O, D = _camera_rays(rk, "ico5", 256, 256)
    hb = rk.hitbox(V)
    ref = ora.build_lbvh_ref(F, V, hb, 12, True)
    rf, ruv, rd, _ = ora.trace_ref_ray(ref.nodes, ref.faces, V, O, D)
    b = dev.FastBuilder(F.shape[0], 1)
    tree = b.build(_t(F), _t(V), hb, 4)
    fid, uv, dep = dev.fast_trace(tree, _t(O), _t(D))
    g_dep = _h(dep)
    np.testing.assert_array_equal(g_dep, rd)
    hit = rd < F32MAX
    g_orig = _h(tree.indices, np.uint32)[_h(fid, np.uint32)[hit]]
    r_orig = ref.indices[rf[hit]]
    diff = g_orig != r_orig
    assert diff.mean() < 0.01
    # every difference is an exact tie
    gt = ora.hit_t_f64(F, V, O[hit][diff], D[hit][diff], g_orig[diff])
    rt = ora.hit_t_f64(F, V, O[hit][diff], D[hit][diff], r_orig[diff])
    assert np.allclose(gt, rt, rtol=1e-6)


def _compare_sah(rk, ora, F, Ft, tree, onodes, oidx, oflags):
    M = tree.node_count
    assert M == onodes.shape[0]
    assert int(tree.host_info()[rk.RK_INFO_FLAGS]) == oflags
    np.testing.assert_array_equal(_h(tree.indices, np.uint32), oidx)
    np.testing.assert_array_equal(_h(Ft, np.uint32), ora.reorder_faces(F, oidx))
    nodes = _h(tree.nodes, np.uint32)[:M].copy().view(ora.NODE_DTYPE).reshape(M)
    np.testing.assert_array_equal(nodes["w0"], onodes["w0"])
    np.testing.assert_array_equal(nodes["w1"], onodes["w1"])
    assert np.array_equal(nodes["bmin"], onodes["bmin"]) and np.array_equal(nodes["bmax"], onodes["bmax"])


@pytest.mark.parametrize("name", ["tiny", "ico3", "ico5", "hf128", "soup50k", "hf708"])
def test_sah_bit_exact(dev, rk, ora, name):
    """hf708 is the full-size heightfield (999,698 faces) of the benchmark."""
    V, F = _scene(rk, name)
    hb = rk.hitbox(V)
    boxes, idx = ora.calc_boxes(F, V)
    onodes, oidx, oflags = ora.build_sah(boxes, idx, hb)
    Ft = _t(F)
    tree = dev.build_sah(Ft, _t(V), hb)
    _compare_sah(rk, ora, F, Ft, tree, onodes, oidx, oflags)


@pytest.mark.parametrize("name,depth", [("ico5", 12), ("hf128", 6), ("soup50k", 12), ("hf708", 12)])
def test_hybrid_bit_exact(dev, rk, ora, name, depth):
    V, F = _scene(rk, name)
    hb = rk.hitbox(V)
    ref = ora.build_lbvh_ref(F, V, hb, depth, False)
    onodes, oidx, oflags = ora.build_sah(ref.boxes, ref.indices, hb, ref.nodes, ref.node_count)
    Ft = _t(F)
    lb = dev.build_lbvh(Ft, _t(V), hb, depth, False)
    assert lb.node_count == ref.node_count
    tree = dev.build_sah(Ft, _t(V), hb, lbvh=lb)
    _compare_sah(rk, ora, F, Ft, tree, onodes, oidx, oflags)


def _degenerate_scene(rk, case):
    """Faces that share one centroid, where the reference's sweep finds no
    finite cost (a NaN for every split): alone, in two clusters, and mixed
    into a soup. Sizes put such items on both sides of 1024 faces."""
    V, F = rk.soup(4000, seed=7)
    if case == "one300":
        F = np.repeat(F[:1], 300, axis=0)
    elif case == "one5000":
        F = np.repeat(F[:1], 5000, axis=0)
    elif case == "two_clusters":
        F = np.concatenate([np.repeat(F[:1], 2000, axis=0), np.repeat(F[1:2], 3000, axis=0)])
    elif case == "soup_plus_dups":
        F = np.concatenate([F, np.repeat(F[5:6], 1500, axis=0), np.repeat(F[9:10], 230, axis=0)])
    return V, np.ascontiguousarray(F)


@pytest.mark.parametrize("case", ["one300", "one5000", "two_clusters", "soup_plus_dups"])
def test_sah_degenerate_no_cut(dev, rk, ora, case):
    V, F = _degenerate_scene(rk, case)
    hb = rk.hitbox(V)
    boxes, idx = ora.calc_boxes(F, V)
    onodes, oidx, oflags = ora.build_sah(boxes, idx, hb)
    assert oflags & rk.RK_FLAG_SAH_NO_CUT
    Ft = _t(F)
    tree = dev.build_sah(Ft, _t(V), hb)
    _compare_sah(rk, ora, F, Ft, tree, onodes, oidx, oflags)


def test_sah_builder_reuse(dev, rk, ora):
    """One SAHBuilder's buffers reused: a larger mesh, then a smaller one, then
    a Hybrid refinement; each tree equals the oracle's."""
    sb = dev.SAHBuilder(50_000, 1 << 13)
    for name in ["soup50k", "ico5"]:
        V, F = _scene(rk, name)
        hb = rk.hitbox(V)
        boxes, idx = ora.calc_boxes(F, V)
        onodes, oidx, oflags = ora.build_sah(boxes, idx, hb)
        Ft = _t(F)
        tree = sb.build(Ft, _t(V), hb)
        tree.node_count = int(tree.host_info()[rk.RK_INFO_NODE_COUNT])
        _compare_sah(rk, ora, F, Ft, tree, onodes, oidx, oflags)
    ref = ora.build_lbvh_ref(F, V, hb, 12, False)
    onodes, oidx, oflags = ora.build_sah(ref.boxes, ref.indices, hb, ref.nodes, ref.node_count)
    Ft = _t(F)
    lb = dev.build_lbvh(Ft, _t(V), hb, 12, False)
    tree = sb.build(Ft, _t(V), hb, lbvh=lb)
    tree.node_count = int(tree.host_info()[rk.RK_INFO_NODE_COUNT])
    _compare_sah(rk, ora, F, Ft, tree, onodes, oidx, oflags)


@pytest.mark.parametrize("coherency", [False, True])
def test_sah_tree_trace_bit_exact(dev, rk, ora, coherency):
    V, F = _scene(rk, "ico5")
    hb = rk.hitbox(V)
    boxes, idx = ora.calc_boxes(F, V)
    onodes, oidx, _ = ora.build_sah(boxes, idx, hb)
    OF = ora.reorder_faces(F, oidx)
    tree = dev.build_sah(_t(F), _t(V), hb)
    O, D = _camera_rays(rk, "ico5", 128, 128)
    fid, uv, dep = dev.trace_ref(tree, _t(V), _t(O), _t(D), 128, 128, coherency)
    if coherency:
        of, ouv, od, _ = ora.trace_ref_packet(onodes, OF, V, O, D, 128, 128)
    else:
        of, ouv, od, _ = ora.trace_ref_ray(onodes, OF, V, O, D)
    np.testing.assert_array_equal(_h(fid, np.uint32), of)
    np.testing.assert_array_equal(_h(uv), ouv)
    np.testing.assert_array_equal(_h(dep), od)


@pytest.mark.parametrize("attributes", [False, True])
def test_interpolate_matches_oracle(dev, rk, ora, attributes):
    """Positions, normals, colours, texcoords and materials at the hits of a
    reference trace (reordered faces), bit-exact with the oracle; pixels that
    missed keep face 0 from the cleared buffer, as in the reference frame loop."""
    V, F = _scene(rk, "ico3")
    F = F.copy()
    F[:, 3] = np.arange(F.shape[0], dtype=np.uint32) * 7 % 13
    hb = rk.hitbox(V)
    tree = dev.build_lbvh(_t(F), _t(V), hb, 12, True)
    O, D = _camera_rays(rk, "ico3", 64, 48)
    fid, uv, dep = dev.trace_ref(tree, _t(V), _t(O), _t(D), 64, 48, False)
    rng = np.random.default_rng(8)
    attrs = [None, None, None]
    if attributes:
        attrs = [rng.normal(size=(V.shape[0], 4)).astype(np.float32),
                 rng.uniform(size=(V.shape[0], 4)).astype(np.float32),
                 rng.uniform(size=(V.shape[0], 2)).astype(np.float32)]
    outs = dev.interpolate(fid, uv, 64, 48, tree.faces, _t(V), *(None if a is None else _t(a) for a in attrs))
    Fr = _h(tree.faces, np.uint32)
    ref = ora.interpolate(_h(fid, np.uint32), _h(uv), Fr, V, *attrs)
    for g, r in zip(outs, ref):
        np.testing.assert_array_equal(_h(g).view(r.dtype), r)
    assert (_h(dep) < F32MAX).mean() > 0.2
