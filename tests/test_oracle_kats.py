"""Pins of the oracle (CPU, no GPU): closed forms, known answers, hand-drawn
trees, invariants and independent formulations, chosen so that a dropped term,
a swapped axis or an off-by-one in the oracle fails one of them."""
import itertools

import numpy as np
import pytest

F32MAX = np.finfo(np.float32).max


def box_of_center(c):
    c = np.asarray(c, np.float32)
    b = np.zeros((1, 2, 4), np.float32)
    b[0, 0, :3] = c
    b[0, 1, :3] = c
    return b


UNIT = np.array([[0, 0, 0, 0], [1, 1, 1, 0]], np.float32)


# ---- Morton -----------------------------------------------------------------

@pytest.mark.parametrize("center,code", [
    ((0, 0, 0), 0x00000000),
    ((1, 1, 1), 0xFFFFFFFF),
    ((1, 0, 0), 0x92492492),   # x bits at 3k+1
    ((0, 1, 0), 0x49249249),   # y bits at 3k
    ((0, 0, 1), 0x24924924),   # z bits at 3k+2 (10 bits)
])
def test_morton32_known_answers(ora, center, code):
    assert ora.morton32(box_of_center(center), UNIT)[0] == code


def _bitloop32(q):
    qx, qy, qz = q
    code = 0
    for k in range(11):
        code |= ((qy >> k) & 1) << (3 * k)
        code |= ((qx >> k) & 1) << (3 * k + 1)
        if k < 10:
            code |= ((qz >> k) & 1) << (3 * k + 2)
    return code & 0xFFFFFFFF


def test_morton32_matches_bit_loop(ora):
    rng = np.random.default_rng(1)
    root = np.array([[-1.5, -0.5, 2.0, 0], [2.5, 3.0, 2.75, 0]], np.float32)
    n = 4000
    boxes = np.zeros((n, 2, 4), np.float32)
    lo = root[0, :3] + rng.random((n, 3), dtype=np.float32) * (root[1, :3] - root[0, :3])
    ext = rng.random((n, 3), dtype=np.float32) * 0.01
    boxes[:, 0, :3] = lo
    boxes[:, 1, :3] = np.minimum(lo + ext, root[1, :3])
    codes = ora.morton32(boxes, root)
    length = root[1] - root[0]
    s = np.array([np.float32(2047) / length[0], np.float32(2047) / length[1], np.float32(1023) / length[2]],
                 np.float32)
    c = (boxes[:, 0] + boxes[:, 1]) * np.float32(0.5)
    q = ((c[:, :3] - root[0, :3]) * s).astype(np.float32)
    qi = np.floor(np.maximum(q, 0)).astype(np.uint64)
    for i in range(0, n, 37):
        assert codes[i] == _bitloop32(tuple(int(x) for x in qi[i])), i


@pytest.mark.parametrize("bits", [10, 21])
def test_morton_wide_matches_bit_loop(ora, bits):
    rng = np.random.default_rng(bits)
    n = 2000
    boxes = np.zeros((n, 2, 4), np.float32)
    c = rng.random((n, 3), dtype=np.float32)
    c[0] = 1.0  # the top cell clamps to 2^b - 1
    c[1] = 0.0
    boxes[:, 0, :3] = c
    boxes[:, 1, :3] = c
    codes = ora.morton_wide(boxes, UNIT, bits)
    cells = np.float32(1 << bits)
    q = np.minimum(np.floor(c * (cells / np.float32(1.0))).astype(np.int64), (1 << bits) - 1)
    for i in range(0, n, 13):
        ref = 0
        for k in range(bits):
            ref |= ((int(q[i, 0]) >> k) & 1) << (3 * k + 2)
            ref |= ((int(q[i, 1]) >> k) & 1) << (3 * k + 1)
            ref |= ((int(q[i, 2]) >> k) & 1) << (3 * k)
        assert int(codes[i]) == ref, i
    assert int(codes[0]) == (1 << (3 * bits)) - 1


# ---- boxes ------------------------------------------------------------------

def test_boxes_closed_form(ora):
    V = np.array([[0, 1, 2, 5], [3, -1, 2, -7], [1, 1, -4, 0.5], [9, 9, 9, 9]], np.float32)
    F = np.array([[0, 1, 2, 0], [3, 3, 3, 7], [2, 0, 1, 1]], np.uint32)
    boxes, idx = ora.calc_boxes(F, V)
    assert list(idx) == [0, 1, 2]
    np.testing.assert_array_equal(boxes[0, 0], [0, -1, -4, -7])   # w is merged too
    np.testing.assert_array_equal(boxes[0, 1], [3, 1, 2, 5])
    np.testing.assert_array_equal(boxes[1, 0], V[3])
    np.testing.assert_array_equal(boxes[1, 1], V[3])
    np.testing.assert_array_equal(boxes[2], boxes[0])


# ---- sorts ------------------------------------------------------------------

def test_sort_small_all_permutations_stable(ora):
    for keys in ([3, 1, 2, 1, 3, 0, 1], [5, 5, 5, 5], [2, 1, 2, 1, 2, 1, 2, 1]):
        for perm in set(itertools.permutations(range(len(keys)))):
            k = np.array([keys[p] for p in perm], np.uint32)
            sk, sv, _ = ora.sort_pairs_ref(k, np.arange(len(k), dtype=np.uint32), 32, init=True)
            expect = sorted(range(len(k)), key=lambda i: (k[i], i))
            assert list(sv) == expect


def _batcher_bitonic(k, v):
    """Bitonic sorting network on 128 elements (Batcher 1968): in stage (i, j),
    element n is compare-exchanged with n^j, ascending when n & i == 0."""
    k, v = k.copy(), v.copy()
    n = np.arange(128)
    i = 2
    while i <= 128:
        j = i // 2
        while j >= 1:
            partner = n ^ j
            lo = n[partner > n]
            hi = lo ^ j
            up = (lo & i) == 0
            a, b = k[lo], k[hi]
            sw = np.where(up, a > b, a < b)
            kl, kh = k[lo].copy(), k[hi].copy()
            vl, vh = v[lo].copy(), v[hi].copy()
            k[lo] = np.where(sw, kh, kl)
            k[hi] = np.where(sw, kl, kh)
            v[lo] = np.where(sw, vh, vl)
            v[hi] = np.where(sw, vl, vh)
            j //= 2
        i *= 2
    return k, v


@pytest.mark.parametrize("n", [513, 640, 1000, 4096, 20480, 32768])
def test_sort_bitonic_class_tie_order(ora, n):
    rng = np.random.default_rng(n)
    keys = rng.integers(0, 7, n, dtype=np.uint32)  # heavy duplicates
    sk, sv, _ = ora.sort_pairs_ref(keys, np.zeros(n, np.uint32), 32, init=True)
    # independent formulation: network per full 128-chunk, stable for the tail
    k2, v2 = keys.copy(), np.arange(n, dtype=np.uint32)
    for o in range(0, n, 128):
        m = min(128, n - o)
        if m == 128:
            k2[o:o + 128], v2[o:o + 128] = _batcher_bitonic(k2[o:o + 128], v2[o:o + 128])
        else:
            order = np.argsort(k2[o:o + m], kind="stable")
            k2[o:o + m], v2[o:o + m] = k2[o:o + m][order], v2[o:o + m][order]
    order = np.argsort(k2, kind="stable")
    np.testing.assert_array_equal(sk, k2[order])
    np.testing.assert_array_equal(sv, v2[order])
    assert np.all(np.diff(sk.astype(np.int64)) >= 0)
    assert sorted(sv.tolist()) == list(range(n))


@pytest.mark.parametrize("n,bits", [(32769, 32), (40000, 21), (9000, 8), (70000, 8)])
def test_sort_radix_class_stable_on_low_bytes(ora, n, bits):
    rng = np.random.default_rng(n + bits)
    keys = rng.integers(0, 1 << min(bits, 31), n, dtype=np.uint32)
    keys[::3] = keys[0]
    sk, sv, _ = ora.sort_pairs_ref(keys, np.zeros(n, np.uint32), bits, init=True)
    byte_depth = 1 + (bits - 1) // 8
    if n <= byte_depth * 8192:
        pytest.skip("bitonic class")
    mask = np.uint32((1 << (8 * byte_depth)) - 1) if byte_depth < 4 else np.uint32(0xFFFFFFFF)
    order = np.argsort(keys & mask, kind="stable")
    np.testing.assert_array_equal(sv, order.astype(np.uint32))


def test_merge2x_oob_trigger(ora):
    # 2 bins with a partial second bin: the reference writes past n (BSG.cu:254)
    _, _, oob = ora.sort_pairs_ref(np.arange(700, dtype=np.uint32)[::-1].copy(), np.zeros(700, np.uint32))
    assert oob  # 700 -> 6 bins, 700 % 128 != 0, last pair (4,5) partial
    _, _, oob = ora.sort_pairs_ref(np.arange(640, dtype=np.uint32), np.zeros(640, np.uint32))
    assert not oob  # 640 = 5 full bins


# ---- splits / topology ------------------------------------------------------

def test_find_splits_levels(ora):
    codes = np.array([0, 0, 0x40000000, 0x80000000, 0x80000001, 0xA0000000], np.uint32)
    A, cnt = ora.find_splits(codes)
    # splits at 2 (level 1), 3 (level 0), 4 (level 31), 5 (level 2)
    assert cnt[0] == 1 and cnt[1] == 1 and cnt[2] == 1 and cnt[31] == 1 and cnt.sum() == 4
    assert list(A) == [3, 2, 5, 4]


def _mesh_from_codes(codes):
    """A mesh whose i-th face (one point triangle) has Morton code codes[i] in the
    unit root box, listed in code order so the sort is the identity."""
    codes = np.asarray(codes, np.uint32)
    n = codes.shape[0]
    q = np.zeros((n, 3), np.int64)
    for k in range(11):
        q[:, 1] |= ((codes >> (3 * k)) & 1).astype(np.int64) << k
        q[:, 0] |= ((codes >> (3 * k + 1)) & 1).astype(np.int64) << k
        if k < 10:
            q[:, 2] |= ((codes >> (3 * k + 2)) & 1).astype(np.int64) << k
    # cell centres: (q + 0.5) / cells, exact in float32
    c = np.stack([(q[:, 0] + 0.5) / 2047.0, (q[:, 1] + 0.5) / 2047.0, (q[:, 2] + 0.5) / 1023.0], 1)
    c = c.astype(np.float32)
    V = np.zeros((3 * n, 4), np.float32)
    for m in range(3):
        V[m::3, :3] = c
    V[1::3, 2] += np.float32(1e-6)  # non-zero z extent for each leaf
    F = np.stack([np.arange(0, 3 * n, 3), np.arange(1, 3 * n, 3), np.arange(2, 3 * n, 3), np.zeros(n)], 1)
    return F.astype(np.uint32), V


def _semantic(nodes):
    out = []
    for nd in nodes:
        if nd["w0"] & 1:
            out.append(("L", int(nd["w1"]), int(nd["w0"] >> 1)))
        else:
            out.append(("N", int((nd["w0"] >> 1) & 3), int(nd["w1"])))
    return out


def test_topology_walk_into_next_level(ora):
    # Splits: 16 (level 1), 32 (level 0), 48 (level 2). The right level-1 range
    # [32,64) has no level-1 split; its split pointer walks on into the level-2
    # list and splits at 48 (CUDA_buildMeshTreeLBVH.cu:381-392).
    codes = [0] * 16 + [0x40000000] * 16 + [0x80000000] * 16 + [0xA0000000] * 16
    F, V = _mesh_from_codes(codes)
    r = ora.build_lbvh_ref(F, V, UNIT, tree_depth=12, complete=True)
    assert list(r.codes) == codes
    assert _semantic(r.nodes) == [
        ("N", 0, 1),             # 0: root [0,64) split 32, axis 0
        ("N", 1, 3),             # 1: [0,32) L1 split 16, axis 1
        ("N", 1, 5),             # 2: [32,64) L1 split 48 (a level-2 split), axis 1
        ("L", 0, 16), ("L", 16, 16),
        ("L", 32, 16), ("L", 48, 16),
    ]
    assert r.flags & ora.FLAG_WALK_PAST_END  # [48,64) at L2 walks past the last split
    assert r.real_depth == 3


def test_topology_leaf_rules(ora):
    # all codes equal: no split at all; 300 > 224 so the root is median-split
    F, V = _mesh_from_codes([7] * 300)
    r = ora.build_lbvh_ref(F, V, UNIT, 12, True)
    assert _semantic(r.nodes) == [("N", 0, 1), ("L", 0, 150), ("L", 150, 150)]
    # incomplete tree keeps the 300-face leaf
    r = ora.build_lbvh_ref(F, V, UNIT, 12, False)
    assert _semantic(r.nodes) == [("L", 0, 300)]
    # exactly 224 faces: a leaf (createLeaf splits only > 224)
    F, V = _mesh_from_codes([7] * 224)
    assert _semantic(ora.build_lbvh_ref(F, V, UNIT, 12, True).nodes) == [("L", 0, 224)]
    # 448 faces: 224/224 from createLeaf, then each 224 splits again (median rule '< 224')
    F, V = _mesh_from_codes([7] * 448)
    assert _semantic(ora.build_lbvh_ref(F, V, UNIT, 12, True).nodes) == [
        ("N", 0, 1), ("N", 1, 3), ("N", 1, 5), ("L", 0, 112), ("L", 112, 112), ("L", 224, 112), ("L", 336, 112)]
    # size < 16 is a leaf even with a split inside
    F, V = _mesh_from_codes([0] * 7 + [0x80000000] * 8)
    assert _semantic(ora.build_lbvh_ref(F, V, UNIT, 12, True).nodes) == [("L", 0, 15)]
    # 16 faces with a level-0 split: splits
    F, V = _mesh_from_codes([0] * 8 + [0x80000000] * 8)
    assert _semantic(ora.build_lbvh_ref(F, V, UNIT, 12, True).nodes) == [("N", 0, 1), ("L", 0, 8), ("L", 8, 8)]
    # tree_depth 0: the root is a leaf
    F, V = _mesh_from_codes([0] * 8 + [0x80000000] * 8)
    assert _semantic(ora.build_lbvh_ref(F, V, UNIT, 0, True).nodes) == [("L", 0, 16)]


def _check_tree_invariants(r, n):
    nodes = r.nodes
    M = nodes.shape[0]
    assert M % 2 == 1
    leaves = []
    rank = 0
    seen_children = set()

    def walk(i, lo, hi, depth):
        nonlocal rank
        nd = nodes[i]
        if nd["w0"] & 1:
            a, c = int(nd["w1"]), int(nd["w0"] >> 1)
            leaves.append((a, a + c))
            return depth
        c = int(nd["w1"])
        assert c == 2 * rank + 1, "children of the internal node of preorder rank r are 2r+1, 2r+2"
        rank += 1
        seen_children.update([c, c + 1])
        return max(walk(c, lo, hi, depth + 1), walk(c + 1, lo, hi, depth + 1))

    depth = walk(0, 0, n, 0)
    assert sorted(seen_children) == list(range(1, M))
    leaves.sort()
    assert leaves[0][0] == 0 and leaves[-1][1] == n
    for (a0, b0), (a1, b1) in zip(leaves, leaves[1:]):
        assert b0 == a1 and a0 < b0
    assert r.real_depth == depth + 1


@pytest.mark.parametrize("scene", ["ico3", "hf64", "soup"])
@pytest.mark.parametrize("depth", [4, 12, 31])
def test_topology_invariants_and_refit(ora, rk, scene, depth):
    if scene == "ico3":
        V, F = rk.icosphere(3)
    elif scene == "hf64":
        V, F = rk.heightfield(64)
    else:
        V, F = rk.soup(3000, seed=7)
    hb = rk.hitbox(V)
    r = ora.build_lbvh_ref(F, V, hb, depth, True)
    _check_tree_invariants(r, F.shape[0])
    assert np.all(np.diff(r.codes.astype(np.int64)) >= 0)
    assert sorted(r.indices.tolist()) == list(range(F.shape[0]))
    # the pass emulation of the refit equals the recursive rule
    rec = ora.refit_recursive(r.nodes)
    for f in ("bmin", "bmax"):
        np.testing.assert_array_equal(rec[f], r.nodes[f])


def test_refit_flat_leaf_empties_ancestors(ora):
    # two leaves; the right leaf is flat in z, so the root keeps the empty box
    codes = [0] * 16 + [0x80000000] * 16
    F, V = _mesh_from_codes(codes)
    V[3 * 16 + 1::3, 2] = V[3 * 16::3, 2][: 16]  # right leaf: remove the z offset
    r = ora.build_lbvh_ref(F, V, UNIT, 12, True)
    assert r.nodes[2]["bmax"][2] - r.nodes[2]["bmin"][2] == 0.0
    np.testing.assert_array_equal(r.nodes[0]["bmin"], [F32MAX] * 3)
    np.testing.assert_array_equal(r.nodes[0]["bmax"], [-F32MAX] * 3)
    assert r.nodes[1]["bmax"][2] > r.nodes[1]["bmin"][2]


# ---- intersection -----------------------------------------------------------

def test_moller_trumbore_known_answer(ora):
    hit, tuv = ora.tri_ray([0.25, 0.25, 1, 0], [0, 0, -1, 0], [0, 0, 0, 0], [1, 0, 0, 0], [0, 1, 0, 0])
    assert hit and tuv[0] == 1.0 and tuv[1] == 0.25 and tuv[2] == 0.25


@pytest.mark.parametrize("b,expect", [(0.999e-3, False), (1.001e-3, True), (-0.999e-3, False), (-1.001e-3, True)])
def test_moller_trumbore_det_threshold(ora, b, expect):
    # det = e1.x * e2.y for d = (0,0,-1): the reference rejects |det| < 1e-6 (absolute)
    o = [0.0001, 0.0001 * np.sign(b), 1, 0]
    hit, _ = ora.tri_ray(o, [0, 0, -1, 0], [0, 0, 0, 0], [1e-3, 0, 0, 0], [0, b, 0, 0])
    assert hit == expect


def test_box_zero_direction_component_always_enters(ora):
    # 1/0 = inf, normalize gives NaN there and the slab test passes (RT.cu:128)
    hit, _ = ora.box_ray([5, 5, 5, 0], [6, 6, 6, 0], [0, 0, 0, 0], [1, 0, 0, 0])
    assert hit
    hit, _ = ora.box_ray([5, 5, 5, 0], [6, 6, 6, 0], [0, 0, 0, 0], [1, 0.01, 0.02, 0])
    assert not hit
    # an inverted (empty) box is entered by every ray
    hit, _ = ora.box_ray([F32MAX] * 3 + [0], [-F32MAX] * 3 + [0], [0, 0, 0, 0], [1, 0.3, 0.2, 0])
    assert hit


def test_trace_ref_matches_bruteforce_on_icosphere(ora, rk):
    V, F = rk.icosphere(3)
    hb = rk.hitbox(V)
    r = ora.build_lbvh_ref(F, V, hb, 12, True)
    cam = rk.camera_lookat([0, 0, 3], [0, 0, 0], 60 * 0.01745329252, 48, 48)
    O, D = rk.primary_rays_host(cam, 48, 48)
    fid, uv, dep, ovf = ora.trace_ref_ray(r.nodes, r.faces, V, O, D)
    assert ovf == 0
    bf, buv, bdep = ora.trace_bruteforce(F, V, O, D)
    hit = dep < F32MAX
    assert hit.sum() > 0.2 * hit.size  # the unit sphere covers ~29% of a 60 degree view from z = 3
    np.testing.assert_array_equal(dep, bdep)
    same = r.indices[fid[hit]] == bf[hit]
    # ties on t (a ray through a shared edge or vertex) may pick either face
    assert same.mean() > 0.99
    np.testing.assert_array_equal(uv[hit][same], buv[hit][same])
    pf, puv, pdep, povf = ora.trace_ref_packet(r.nodes, r.faces, V, O, D, 48, 48)
    assert povf == 0
    np.testing.assert_array_equal(pdep, dep)
    t64 = ora.hit_t_f64(r.faces, V, O, D, fid)
    err = np.abs(dep[hit].astype(np.float64) - t64[hit]) / t64[hit]
    assert err.max() < 1e-5


@pytest.mark.parametrize("eps", [1e-6, 0.0])
def test_trace_closest_equals_bruteforce(ora, rk, eps):
    # the tree-based closest-hit oracle agrees with brute force (all faces) on
    # t, u, v and the tie rule, with the reference threshold and with eps = 0
    V, F = rk.heightfield(40, seed=5)
    hb = rk.hitbox(V)
    el, az = np.radians(20.0), np.radians(30.0)
    tgt = np.array([0.5, 0.5, 0.1])
    eye = tgt + 1.5 * np.array([np.cos(el) * np.cos(az), np.cos(el) * np.sin(az), np.sin(el)])
    cam = rk.camera_lookat(eye, tgt, 60 * 0.01745329252, 64, 48, up=[0, 0, 1])
    O, D = rk.primary_rays_host(cam, 64, 48)
    bf, buv, bd = ora.trace_bruteforce(F, V, O, D, det_epsilon=eps)
    for depth in (6, 31):
        tree = ora.build_lbvh_ref(F, V, hb, depth, True)
        cf, cuv, cd = ora.trace_closest(tree, V, O, D, det_epsilon=eps)
        np.testing.assert_array_equal(cd, bd)
        hit = bd < F32MAX
        np.testing.assert_array_equal(cf[hit], bf[hit])
        np.testing.assert_array_equal(cuv[hit], buv[hit])
    assert hit.mean() > 0.05  # the terrain fills about a tenth of this view


def test_trace_miss_leaves_outputs_unchanged(ora, rk):
    V, F = rk.icosphere(1)
    r = ora.build_lbvh_ref(F, V, rk.hitbox(V), 12, True)
    O = np.array([[0, 0, 5, 0], [10, 10, 10, 0]], np.float32)
    D = np.array([[0, 0, -1, 0], [1, 0, 0, 0]], np.float32)
    fid, uv, dep, _ = ora.trace_ref_ray(r.nodes, r.faces, V, O, D, face_id=np.array([77, 88], np.uint32),
                                        coords=np.full((2, 2), 0.5, np.float32),
                                        depths=np.array([F32MAX, 9.0], np.float32))
    assert dep[0] < 5 and fid[0] != 77
    assert fid[1] == 88 and dep[1] == 9.0 and list(uv[1]) == [0.5, 0.5]


def test_trace_depth_is_tmax(ora, rk):
    V, F = rk.icosphere(2)
    r = ora.build_lbvh_ref(F, V, rk.hitbox(V), 12, True)
    O = np.array([[0, 0, 3, 0]], np.float32)
    D = np.array([[0.01, 0.02, -1, 0]], np.float32)
    _, _, d0, _ = ora.trace_ref_ray(r.nodes, r.faces, V, O, D)
    _, _, d1, _ = ora.trace_ref_ray(r.nodes, r.faces, V, O, D, depths=np.array([d0[0]], np.float32))
    assert d1[0] == d0[0]  # strict t < depth: an equal hit is not taken, depth unchanged


def test_sah_two_clusters(ora):
    # two well-separated clusters of 20 unit boxes: the root cut separates them
    n = 40
    boxes = np.zeros((n, 2, 4), np.float32)
    xs = np.concatenate([np.linspace(0, 1, 20), np.linspace(100, 101, 20)]).astype(np.float32)
    rng = np.random.default_rng(0)
    perm = rng.permutation(n)
    boxes[:, 0, 0] = xs[perm]
    boxes[:, 1, 0] = xs[perm] + 0.01
    boxes[:, 1, 1:3] = 0.01
    root = np.array([[0, 0, 0, 0], [101.01, 0.01, 0.01, 0]], np.float32)
    nodes, idx, fl = ora.build_sah(boxes, np.arange(n, dtype=np.uint32), root)
    assert fl == 0
    assert not (nodes[0]["w0"] & 1) and ((nodes[0]["w0"] >> 1) & 3) == 0
    c = nodes[0]["w1"]
    left, right = nodes[c], nodes[c + 1]
    assert left["w0"] & 1 and right["w0"] & 1
    assert (left["w0"] >> 1) == 20 and (right["w0"] >> 1) == 20
    assert np.all(boxes[idx[:20], 0, 0] < 50) and np.all(boxes[idx[20:], 0, 0] > 50)
    assert left["bmax"][0] == np.float32(1.01) and right["bmin"][0] == np.float32(100)


# ---- attribute interpolation (CUDA_interpolateData.cu:45-141) ----------------------

def test_interpolate_known_answers(ora):
    V = np.array([[0, 0, 0, 0], [4, 0, 0, 0], [0, 8, 0, 0]], np.float32)
    F = np.array([[0, 1, 2, 7], [0, 2, 1, 9]], np.uint32)  # counter-clockwise, clockwise
    fid = np.array([0, 0, 0, 0, 1, 5], np.uint32)
    uv = np.array([[0, 0], [1, 0], [0, 1], [0.25, 0.5], [0.25, 0.5], [0.1, 0.1]], np.float32)
    init = (np.full((6, 4), 7, np.float32), np.full((6, 4), 7, np.float32), np.full((6, 4), 7, np.float32),
            np.full((6, 2), 7, np.float32), np.full(6, 3, np.uint32))
    p, n, c, t, m = ora.interpolate(fid, uv, F, V, init=init)
    np.testing.assert_array_equal(p[:4], [V[0], V[1], V[2], [1, 4, 0, 0]])  # vertices and an exact interior point
    np.testing.assert_array_equal(p[4], [2, 2, 0, 0])  # face 1 swaps v1, v2: 0.25*(0,8) + 0.5*(4,0)
    np.testing.assert_array_equal(n[:4], np.tile([0, 0, 1, 0], (4, 1)))  # face normal of a CCW xy triangle
    np.testing.assert_array_equal(n[4], [0, 0, -1, 0])
    np.testing.assert_array_equal(c[:5], np.ones((5, 4)))  # no colours: white
    np.testing.assert_array_equal(t[:5], uv[:5])  # no texture coordinates: (u, v)
    np.testing.assert_array_equal(m[:5], [7, 7, 7, 7, 9])
    for a, b in zip((p, n, c, t, m), init):  # face id out of range: untouched
        np.testing.assert_array_equal(a[5], b[5])


def test_interpolate_attributes(ora):
    V = np.array([[0, 0, 0, 0], [1, 0, 0, 0], [0, 1, 0, 0]], np.float32)
    F = np.array([[0, 1, 2, 0]], np.uint32)
    normals = np.array([[0, 0, 2, 0], [0, 0, 2, 0], [0, 0, 2, 0]], np.float32)
    colors = np.array([[1, 0, 0, 1], [0, 1, 0, 1], [0, 0, 1, 1]], np.float32)
    tex = np.array([[0, 0], [1, 0], [0, 1]], np.float32)
    fid = np.zeros(2, np.uint32)
    uv = np.array([[0.25, 0.25], [0.5, 0.0]], np.float32)
    p, n, c, t, m = ora.interpolate(fid, uv, F, V, normals, colors, tex)
    np.testing.assert_array_equal(n, [[0, 0, 1, 0], [0, 0, 1, 0]])  # equal normals normalise to themselves
    np.testing.assert_array_equal(c, [[0.5, 0.25, 0.25, 1], [0.5, 0.5, 0, 1]])  # barycentric weights of the corners
    np.testing.assert_array_equal(t, uv)  # texcoords equal to positions: (u, v) again
    np.testing.assert_array_equal(p[:, :2], uv)


# ---- SAH rules, hand-computed (CUDA_buildMeshTreeSAH.cu:224-645, CUDABin.h:44-111) ----

def _sah_node(nd):
    if nd["w0"] & 1:
        return ("L", int(nd["w1"]), int(nd["w0"] >> 1))
    return ("N", int((nd["w0"] >> 1) & 3), int(nd["w1"]))


def _boxes(lo, hi):
    n = len(lo)
    b = np.zeros((n, 2, 4), np.float32)
    b[:, 0, :3] = np.asarray(lo, np.float32)
    b[:, 1, :3] = np.asarray(hi, np.float32)
    return b


def _points(centres):
    return _boxes(centres, centres)


def test_sah_line_bins_and_tie_to_larger_split():
    """64 boxes [i-0.5, i+0.5] x [0,1] x [0,1], i = 0..63, fed in shuffled order.

    Centroid extent 63 along x only: axis 0, scale = 31/63, bin(i) =
    (int)(i * fl(31/63)) = 0 for i = 0,1,2, then two per bin, 63 alone in bin 31
    (SAH.cu:312-325, 373-385). The cut after bin 14 leaves 31 boxes on the left
    and 33 on the right; the cut after bin 15 leaves 33 and 31. The group of k
    adjacent boxes has area 2(2k+1), exact in float, so the two costs are
    2 + 0.125 * (a + b) and 2 + 0.125 * (b + a) with a = fl(fl(126/258) * 31),
    b = fl(fl(134/258) * 33): an exact tie, and no other cut is within 0.7%.
    The right sweep keeps the first strict minimum from bin 31 down
    (SAH.cu:452-472), so the cut after bin 15 wins: 33 left, 31 right. A scale
    of 32/63 would make a 32/32 cut available; '<=' would pick 31/33."""
    import oracle as ora
    i = np.arange(64, dtype=np.float32)
    boxes = _boxes(np.stack([i - 0.5, 0 * i, 0 * i], 1), np.stack([i + 0.5, 0 * i + 1, 0 * i + 1], 1))
    perm = np.random.default_rng(3).permutation(64).astype(np.uint32)
    root = np.array([[-0.5, 0, 0, 0], [63.5, 1, 1, 0]], np.float32)
    nodes, idx, fl = ora.build_sah(boxes, perm, root)
    assert fl == 0
    assert _sah_node(nodes[0]) == ("N", 0, 1)
    assert sorted(idx[:33].tolist()) == list(range(33))  # left subtree: boxes 0..32
    assert sorted(idx[33:].tolist()) == list(range(33, 64))
    # children boxes are the bin unions (SAH.cu:501-523); the node keeps the item box
    np.testing.assert_array_equal(nodes[1]["bmin"], [-0.5, 0, 0])
    np.testing.assert_array_equal(nodes[1]["bmax"], [32.5, 1, 1])
    np.testing.assert_array_equal(nodes[2]["bmin"], [32.5, 0, 0])
    np.testing.assert_array_equal(nodes[2]["bmax"], [63.5, 1, 1])
    np.testing.assert_array_equal(nodes[0]["bmin"], root[0, :3])
    np.testing.assert_array_equal(nodes[0]["bmax"], root[1, :3])


@pytest.mark.parametrize("a,b,axis", [
    ((0, 0, 0), (8, 8, 0), 1),    # x == y > z: 'x > y' fails, 'y > z' holds -> 1
    ((0, 0, 0), (8, 3, 8), 2),    # x > y, x == z: 'x > z' fails -> 2
    ((0, 0, 0), (3, 8, 8), 2),    # y > x, y == z: 'y > z' fails -> 2
    ((0, 0, 0), (8, 8, 8), 2),    # all equal -> 2
    ((0, 0, 0), (8, 7, 7), 0),
    ((0, 0, 0), (7, 8, 7), 1),
])
def test_sah_axis_ties(a, b, axis):
    """Two clusters of 10 point boxes (SAH.cu:312-325): ties between centroid
    extents go to the later axis. Point boxes have area 0, so every cut between
    the clusters costs exactly 2 and the 20 boxes split 10/10 into two leaves."""
    import oracle as ora
    boxes = _points([a] * 10 + [b] * 10)
    idx = np.array([k // 2 + (k % 2) * 10 for k in range(20)], np.uint32)  # interleaved
    root = np.array([[-1, -1, -1, 0], [9, 9, 9, 0]], np.float32)
    nodes, out, fl = ora.build_sah(boxes, idx, root)
    assert [_sah_node(nd) for nd in nodes] == [("N", axis, 1), ("L", 0, 10), ("L", 10, 10)]
    assert out.tolist() == list(range(20))  # stable partition in index order
    np.testing.assert_array_equal(nodes[1]["bmin"], a)
    np.testing.assert_array_equal(nodes[2]["bmax"], b)


@pytest.mark.parametrize("n,leaf", [(16, True), (17, False)])
def test_sah_min_leaf_size(n, leaf):
    """At most 16 objects is a leaf before any binning (SAH.cu:239-255)."""
    import oracle as ora
    boxes = _points([(0, 0, 0)] * (n // 2) + [(5, 0, 0)] * (n - n // 2))
    root = np.array([[-1, -1, -1, 0], [6, 1, 1, 0]], np.float32)
    nodes, _, _ = ora.build_sah(boxes, np.arange(n, dtype=np.uint32), root)
    if leaf:
        assert [_sah_node(nd) for nd in nodes] == [("L", 0, n)]
    else:
        assert [_sah_node(nd) for nd in nodes] == [("N", 0, 1), ("L", 0, n // 2), ("L", n // 2, n - n // 2)]


@pytest.mark.parametrize("root_x,leaf", [(6.0, False), (4.0, True), (5.5, False)])
def test_sah_leaf_cost_rule(root_x, leaf):
    """10 unit boxes [0,1]^3 and 10 boxes [9,10]x[0,1]^2 under an item box
    [0,root_x] x [0,1]^2 (area 4*root_x + 2). Each child has area 6, so the best
    cost is T = 2 + 0.125 * 20 * 6 / (4*root_x + 2) (CUDABin.h:104-111): 2.577
    for root_x = 6 and 2.833 for root_x = 4. The node is a leaf when it has fewer
    than 224 objects and T > (20 + 1) * 0.125 = 2.625 (SAH.cu:479-496), so the
    first splits and the second does not (a threshold of 20 * 0.125 would make
    both leaves, a box cost of 1 instead of 2 both splits). With root_x = 5.5 the
    area is exactly 24 and T = 2 + 0.125 * (0.25 * 10 + 0.25 * 10) = 2.625, equal
    to the threshold in float: the strict '>' splits."""
    import oracle as ora
    boxes = _boxes([(0, 0, 0)] * 10 + [(9, 0, 0)] * 10, [(1, 1, 1)] * 10 + [(10, 1, 1)] * 10)
    root = np.array([[0, 0, 0, 0], [root_x, 1, 1, 0]], np.float32)
    nodes, _, _ = ora.build_sah(boxes, np.arange(20, dtype=np.uint32), root)
    if leaf:
        assert [_sah_node(nd) for nd in nodes] == [("L", 0, 20)]
    else:
        assert [_sah_node(nd) for nd in nodes] == [("N", 0, 1), ("L", 0, 10), ("L", 10, 10)]
        np.testing.assert_array_equal(nodes[1]["bmax"], [1, 1, 1])
        np.testing.assert_array_equal(nodes[2]["bmin"], [9, 0, 0])


@pytest.mark.parametrize("n,leaf", [(223, True), (224, False)])
def test_sah_max_leaf_size(n, leaf):
    """n boxes of side 200 whose centres step by 1/1024 along x: every child box
    is almost the parent box, so every cut costs about 2 + 0.125 * n and the
    cost test alone would make a leaf. A leaf needs fewer than 224 objects
    (SAH.cu:479-496), so 223 is a leaf and 224 splits."""
    import oracle as ora
    c = np.arange(n, dtype=np.float32) / np.float32(1024)
    lo = np.stack([c - 100, np.full(n, -100, np.float32), np.full(n, -100, np.float32)], 1)
    hi = np.stack([c + 100, np.full(n, 100, np.float32), np.full(n, 100, np.float32)], 1)
    boxes = _boxes(lo, hi)
    root = np.array([[-100, -100, -100, 0], [100 + c[-1], 100, 100, 0]], np.float32)
    nodes, _, fl = ora.build_sah(boxes, np.arange(n, dtype=np.uint32), root)
    assert fl == 0
    if leaf:
        assert [_sah_node(nd) for nd in nodes] == [("L", 0, n)]
    else:
        assert _sah_node(nodes[0])[0] == "N" and _sah_node(nodes[0])[1] == 0


def test_sah_degenerate_centroids():
    """All centroids equal: the scale is 31/0 = inf, (c - min) * inf is NaN and
    converts to bin 0; every cut has an empty side whose area term is
    inf * 0 = NaN, which never wins, so m_min stays -1. Below 224 objects the
    cost test (T_min = FLT_MAX) makes a leaf; at 224 or more the reference would
    push the same item forever, and the oracle makes a leaf and sets bit 0x10."""
    import oracle as ora
    for n, flag in ((40, 0), (300, 0x10)):
        boxes = _boxes([(0, 0, 0)] * n, [(1, 1, 1)] * n)
        root = np.array([[0, 0, 0, 0], [1, 1, 1, 0]], np.float32)
        nodes, _, fl = ora.build_sah(boxes, np.arange(n, dtype=np.uint32), root)
        assert [_sah_node(nd) for nd in nodes] == [("L", 0, n)]
        assert fl == flag


def test_hybrid_refines_only_large_leaves():
    """Hybrid (CUDAMeshTree.cpp:98-103, SAH.cu:711-735, kernel_initStack
    SAH.cu:647-666): every leaf of an existing tree is pushed with its own box as
    the item box, and new node ids continue from the existing count. A hand-built
    LBVH: root 0 over leaves 1 (objects 0..15, two separable clusters) and 2
    (objects 16..47: clusters at x = 0 and x = 10, interleaved). Leaf 1 has 16
    objects and stays a leaf; leaf 2 becomes an internal node with its original
    (deliberately loose) box and children 3 and 4 of 16 objects each."""
    import oracle as ora
    cent = [(0, 0, 0)] * 8 + [(5, 5, 5)] * 8 + [(0, 0, 0), (10, 0, 0)] * 16
    boxes = _points(cent)
    nodes = np.zeros(3, ora.NODE_DTYPE)
    nodes[0]["w0"], nodes[0]["w1"] = 0 | (2 << 1), 1
    nodes[0]["bmin"], nodes[0]["bmax"] = (-1, -2, -3), (11, 12, 13)
    nodes[1]["w0"], nodes[1]["w1"] = 1 | (16 << 1), 0
    nodes[1]["bmin"], nodes[1]["bmax"] = (0, 0, 0), (5, 5, 5)
    nodes[2]["w0"], nodes[2]["w1"] = 1 | (32 << 1), 16
    nodes[2]["bmin"], nodes[2]["bmax"] = (-100, -100, -100), (100, 100, 100)
    root = np.array([[-1, -2, -3, 0], [11, 12, 13, 0]], np.float32)
    out, idx, fl = ora.build_sah(boxes, np.arange(48, dtype=np.uint32), root, nodes=nodes, n_existing=3)
    assert fl == 0
    assert [_sah_node(nd) for nd in out] == [("N", 2, 1), ("L", 0, 16), ("N", 0, 3), ("L", 16, 16), ("L", 32, 16)]
    for k in (0, 1, 2):  # boxes of the existing nodes are kept
        np.testing.assert_array_equal(out[k]["bmin"], nodes[k]["bmin"])
        np.testing.assert_array_equal(out[k]["bmax"], nodes[k]["bmax"])
    np.testing.assert_array_equal(out[3]["bmax"], [0, 0, 0])
    np.testing.assert_array_equal(out[4]["bmin"], [10, 0, 0])
    assert idx[:16].tolist() == list(range(16))
    assert idx[16:32].tolist() == list(range(16, 48, 2))  # stable partition
    assert idx[32:].tolist() == list(range(17, 48, 2))


def test_hybrid_lbvh_pass_keeps_face_order(ora):
    """The LBVH pass of Hybrid (tree_complete = false) does not reorder the
    mesh's faces (CUDA_buildMeshTreeLBVH.cu:669-676 runs only when complete)."""
    F, V = _mesh_from_codes([0x80000000] * 20 + [0] * 20)
    r = ora.build_lbvh_ref(F, V, UNIT, 12, False)
    np.testing.assert_array_equal(r.faces, F)
    assert r.indices.tolist() == list(range(20, 40)) + list(range(20))
    r = ora.build_lbvh_ref(F, V, UNIT, 12, True)
    np.testing.assert_array_equal(r.faces, F[r.indices])


def test_refit_flat_left_leaf_empties_ancestors(ora):
    # the left leaf is flat in z: the root keeps the empty box (LBVH.cu:282)
    codes = [0] * 16 + [0x80000000] * 16
    F, V = _mesh_from_codes(codes)
    V[1:3 * 16:3, 2] = V[0:3 * 16:3, 2]  # left leaf: remove the z offset
    r = ora.build_lbvh_ref(F, V, UNIT, 12, True)
    assert r.nodes[1]["bmax"][2] - r.nodes[1]["bmin"][2] == 0.0
    assert r.nodes[2]["bmax"][2] > r.nodes[2]["bmin"][2]
    np.testing.assert_array_equal(r.nodes[0]["bmin"], [F32MAX] * 3)
    np.testing.assert_array_equal(r.nodes[0]["bmax"], [-F32MAX] * 3)
    # both leaves with z extent: the root is their union
    F, V = _mesh_from_codes(codes)
    r = ora.build_lbvh_ref(F, V, UNIT, 12, True)
    np.testing.assert_array_equal(r.nodes[0]["bmin"], np.minimum(r.nodes[1]["bmin"], r.nodes[2]["bmin"]))
    np.testing.assert_array_equal(r.nodes[0]["bmax"], np.maximum(r.nodes[1]["bmax"], r.nodes[2]["bmax"]))


@pytest.mark.parametrize("o,hit,tuv", [
    ((1.0, 0.0, 1.0), True, (1.0, 1.0, 0.0)),      # u == 1 at vertex v1: accepted ('u > 1' rejects)
    ((0.5, 0.5, 1.0), True, (1.0, 0.5, 0.5)),      # u + v == 1 on edge v1-v2: accepted
    ((0.0, 0.5, 1.0), True, (1.0, 0.0, 0.5)),      # u == 0: accepted ('u < 0' rejects)
    ((0.5, 0.0, 1.0), True, (1.0, 0.5, 0.0)),      # v == 0: accepted
    ((0.25, 0.25, 0.0), True, (0.0, 0.25, 0.25)),  # t == 0: accepted ('t >= 0')
    ((0.25, 0.25, -0.5), False, None),             # t < 0
    ((1.0078125, 0.0, 1.0), False, None),          # u = 1 + 2^-7
    ((0.5078125, 0.5, 1.0), False, None),          # u + v = 1 + 2^-7
    ((-0.0078125, 0.5, 1.0), False, None),         # u = -2^-7
    ((0.5, -0.0078125, 1.0), False, None),         # v = -2^-7
])
def test_moller_trumbore_boundaries(ora, o, hit, tuv):
    """Triangle (0,0,0), (1,0,0), (0,1,0), ray straight down: det = 1, u = o.x,
    v = o.y, t = o.z, all exact (RT.cu:57-89). The bounds are inclusive."""
    h, got = ora.tri_ray([*o, 0], [0, 0, -1, 0], [0, 0, 0, 0], [1, 0, 0, 0], [0, 1, 0, 0])
    assert h == hit
    if hit:
        np.testing.assert_array_equal(got, tuv)


# ---- the standalone sorts (samples/09_how_to_use_a_sorter.cpp:36-44) ---------------

def _batcher_rows(k, v):
    """The Batcher bitonic network of _batcher_bitonic applied to every row of
    (rows, 128) arrays at once."""
    k, v = k.copy(), v.copy()
    n = np.arange(128)
    i = 2
    while i <= 128:
        j = i // 2
        while j >= 1:
            lo = n[(n ^ j) > n]
            hi = lo ^ j
            up = (lo & i) == 0
            a, b = k[:, lo], k[:, hi]
            sw = np.where(up, a > b, a < b)
            kl, kh, vl, vh = a.copy(), b.copy(), v[:, lo].copy(), v[:, hi].copy()
            k[:, lo], k[:, hi] = np.where(sw, kh, kl), np.where(sw, kl, kh)
            v[:, lo], v[:, hi] = np.where(sw, vh, vl), np.where(sw, vl, vh)
            j //= 2
        i *= 2
    return k, v


@pytest.mark.parametrize("n", [40000, 300001, 1397121])
def test_bitonic_general_any_length(ora, n):
    """CUDA_bitonicSortGeneral called directly has no 32768 limit: the bucket
    merges (CUDA_bitonicSortGeneral.cu:584-640) run until one bucket is left.
    Independent formulation: network per full 128-chunk, stable partial chunk,
    then a stable sort by key; the values follow the keys."""
    rng = np.random.default_rng(n)
    keys = rng.integers(0, 11, n, dtype=np.uint32)
    vals = np.arange(n, dtype=np.uint32)
    sk, sv, _ = ora.bitonic_sort_general(keys, vals)
    full = n // 128 * 128
    k2, v2 = keys.copy(), vals.copy()
    rk_, rv_ = _batcher_rows(k2[:full].reshape(-1, 128), v2[:full].reshape(-1, 128))
    k2[:full], v2[:full] = rk_.reshape(-1), rv_.reshape(-1)
    tail = np.argsort(k2[full:], kind="stable")
    k2[full:], v2[full:] = k2[full:][tail], v2[full:][tail]
    order = np.argsort(k2, kind="stable")
    np.testing.assert_array_equal(sk, k2[order])
    np.testing.assert_array_equal(sv, v2[order])
    np.testing.assert_array_equal(ora.bitonic_sort(keys), np.sort(keys))


def test_bubble_sort512_stable_and_bounded(ora):
    rng = np.random.default_rng(5)
    for n in (0, 1, 2, 77, 511, 512):
        keys = rng.integers(0, 4, n, dtype=np.uint32)
        idx = rng.integers(0, 1 << 31, n, dtype=np.uint32)
        k, i = ora.bubble_sort512(keys, idx, False)
        order = np.argsort(keys, kind="stable")
        np.testing.assert_array_equal(k, keys[order])
        np.testing.assert_array_equal(i, idx[order])
        k, i = ora.bubble_sort512(keys, idx, True)
        np.testing.assert_array_equal(i, order.astype(np.uint32))
    with pytest.raises(ValueError):
        ora.bubble_sort512(np.zeros(513, np.uint32), np.zeros(513, np.uint32))


# ---- camera and primary rays (CUDACamera.cpp:29-142, CUDA_calcPrimaryRays.cu:20-55) ----

PI32 = np.float32(3.14159265359)


def test_camera_axis_frame_exact(ora):
    """angle (0, 0, 0): camDir = (1, 0, 0), camUp = (0, 1, 0), so right =
    (0, 0, 1) and up = (0, 1, 0) exactly. fov = PI/2 gives tan(fov/2) = 1.0f, so
    with ratio 2 on a 4 x 2 surface: rayDir = up + 2 right + camDir = (1, 1, 2),
    rayX = 2*2*right/4 = (0, 0, 1), rayY = 2*1*up/2 = (0, 1, 0); pixel (x, y)
    looks along (1, 1 - y, 2 - x)."""
    c = ora.Camera()
    c.set_center(0.5, -2.0, 7.0)
    c.set_angle(0, 0, 0)
    np.testing.assert_array_equal(c.dir, [1, 0, 0])
    c.set_fov(PI32 * np.float32(0.5))
    c.set_ratio(2.0)
    cst = c.ray_constants(4, 2)
    np.testing.assert_array_equal(cst, [[0.5, -2, 7, 0], [1, 1, 2, 0], [0, 0, 1, 0], [0, 1, 0, 0]])
    O, D = c.primary_rays(4, 2)
    want = np.array([[1, 1 - y, 2 - x, 0] for y in range(2) for x in range(4)], np.float32)
    np.testing.assert_array_equal(D, want)
    np.testing.assert_array_equal(O, np.tile([0.5, -2, 7, 0], (8, 1)).astype(np.float32))


def test_camera_roll_pi_flips_the_frame(ora):
    """roll = PI: camUp = (cos 0 * sin PI, cos PI, 0) = (-8.7e-8, -1, 0), so
    right = (0, 0, -1) and up = (0, -1, 0) exactly: rayDir = (1, -1, -2),
    rayX = (0, 0, -1), rayY = (0, -1, 0) on the same 4 x 2 surface."""
    c = ora.Camera()
    c.set_angle(0, 0, PI32)
    c.set_fov(PI32 * np.float32(0.5))
    c.set_ratio(4, 2)
    np.testing.assert_array_equal(c.ray_constants(4, 2), [[0, 0, 0, 0], [1, -1, -2, 0], [0, 0, -1, 0],
                                                          [0, -1, 0, 0]])


def test_camera_defaults(ora):
    """CUDACamera.cpp:29-50: fov PI/2, ratio 4/3, center 0, angle (PI/2, 0, 0),
    dir (0, 0, 1)."""
    c = ora.Camera()
    assert c.fov == PI32 * np.float32(0.5) and c.ratio == np.float32(4) / np.float32(3)
    np.testing.assert_array_equal(c.angle, [PI32 * np.float32(0.5), 0, 0])
    np.testing.assert_array_equal(c.dir, [0, 0, 1])


@pytest.mark.parametrize("point,angle_x,dir_", [
    ((2, 0, 0), np.float32(0), (1, 0, 0)),                              # dir.x > 0: asin(z / cos y)
    ((0, 0, 5), np.float32(np.arccos(np.float32(0))), (0, 0, 1)),        # dir.z > 0: acos(x / cos y)
    ((-3, 0, 0), PI32, (-1, 0, 0)),                                      # else: acos(-x / cos -y) + PI
    ((0, 0, -1), np.float32(np.float32(np.arccos(np.float32(0))) + PI32), (0, 0, -1)),
])
def test_camera_look_at_branches(ora, point, angle_x, dir_):
    """angleFromDir (CUDACamera.cpp:70-88): the three branches of the azimuth;
    the roll (angle.z) is kept."""
    c = ora.Camera()
    c.set_angle(0, 0, 0.25)
    c.look_at(*point)
    assert c.angle[0] == angle_x and c.angle[1] == 0 and c.angle[2] == np.float32(0.25)
    np.testing.assert_array_equal(c.dir, dir_)


def test_camera_look_at_pole_clamps_yaw(ora):
    """Looking straight up: yaw = asin(1) is clamped to PI/2 - 0.001 (in float),
    and dir.x = dir.z = 0 takes the third branch: acos(-0) + PI."""
    c = ora.Camera()
    c.look_at(0, 4, 0)
    assert c.angle[1] == PI32 * np.float32(0.5) - np.float32(0.001)
    assert c.angle[0] == np.float32(np.float32(np.arccos(np.float32(0))) + PI32)
    c.look_at(0, -4, 0)
    assert c.angle[1] == -PI32 * np.float32(0.5) + np.float32(0.001)


def test_camera_frame_follows_look_at(ora):
    """For any look direction off the poles, the camera's forward vector camDir
    (rays of a 2 x 2 surface: rayDir - rayX - rayY) and setAngle(angle) both point
    along normalize(point - center) within 5e-5, and right and up are
    orthogonal to it. A wrong branch, a swapped pitch/yaw or a sign error turns
    a direction around and fails this."""
    rng = np.random.default_rng(7)
    for _ in range(300):
        ctr = rng.uniform(-5, 5, 3).astype(np.float32)
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        if abs(d[1]) > 0.99:
            continue
        pt = (ctr + rng.uniform(0.5, 4) * d).astype(np.float32)
        c = ora.Camera()
        c.set_center(*ctr)
        c.look_at(*pt)
        want = (pt.astype(np.float64) - ctr) / np.linalg.norm(pt.astype(np.float64) - ctr)
        np.testing.assert_allclose(c.dir, want, atol=2e-6)
        cst = c.ray_constants(2, 2).astype(np.float64)
        fwd = cst[1, :3] - cst[2, :3] - cst[3, :3]
        np.testing.assert_allclose(fwd, want, atol=5e-5)  # float asin/acos of the azimuth amplify rounding
        assert abs(cst[2, :3] @ want) < 1e-4 and abs(cst[3, :3] @ want) < 1e-4
        # pixel (0, 0) is up and to the right of centre: +up, +right (y * up + x * right + camDir)
        assert cst[3, 1] > 0
        a = c.angle.copy()
        c2 = ora.Camera()
        c2.set_angle(*a)
        np.testing.assert_allclose(c2.dir, want, atol=5e-5)


def test_camera_c1_frame(ora):
    """The icosphere camera (centre (0, 0, 3), looking at the origin, 60 degrees,
    256 x 256): azimuth 3 PI / 2 in float, right = (1, 0, ~0), up = (0, 1, 0),
    and the steps are 2 tan(30 deg) / 256 apart."""
    c = ora.Camera()
    c.set_center(0, 0, 3)
    c.look_at(0, 0, 0)
    c.set_fov(np.float32(60) * np.float32(0.01745329252))
    c.set_ratio(256, 256)
    assert c.angle[0] == np.float32(np.float32(np.arccos(np.float32(0))) + PI32)
    cst = c.ray_constants(256, 256)
    t = np.tan(np.float32(np.float32(60) * np.float32(0.01745329252)) * np.float32(0.5))
    np.testing.assert_allclose(cst[1], [t, t, -1, 0], rtol=1e-6, atol=1e-7)
    np.testing.assert_allclose(cst[2], [2 * t / 256, 0, 0, 0], rtol=1e-6, atol=1e-9)
    np.testing.assert_allclose(cst[3], [0, 2 * t / 256, 0, 0], rtol=1e-6, atol=1e-9)
    O, D = c.primary_rays(256, 256)
    assert D[128 * 256 + 5, 1] == 0  # row 128 has dir.y == 0 exactly


# ---- shading (CUDAShaderKernel.cu:14-43, make_pixel in kernels/CUDAShaderCommon.cuh) ----

BACK = np.array([0.227451, 0.431373, 0.647059, 1.0], np.float32)  # CUDAShader.cpp:28-33


def test_shade_known_pixels(ora):
    """Hand-computed pixels: the default back colour is (58, 110, 165, 255)
    (0.227451 * 255 = 58.000005 and so on), a normal facing the camera is
    white, c = 0.5 gives 127 in every channel (truncation of 127.5), c = 0.999
    gives 254 (not rounded), a normal facing away clamps to 0. w of the normal
    (even infinite) and of the camera direction are ignored."""
    H = np.finfo(np.float32).max
    depths = np.array([H, 1.0, 1.0, 1.0, 1.0, np.nextafter(H, np.float32(0))], np.float32)
    normals = np.array([[0, 0, 1, 0], [0, 0, 1, 7], [0, 0, 0.5, np.inf], [0, 0, 0.999, 0], [0, 0, -1, 0],
                        [0, 0, 1, 0]], np.float32)
    px = ora.shade(depths, normals, 6, 1, BACK, [0, 0, -1, 3])
    assert [hex(p) for p in px] == ["0xff3a6ea5", "0xffffffff", "0x7f7f7f7f", "0xfefefefe", "0x0", "0xffffffff"]


def test_shade_channel_order_and_pitch(ora):
    """A << 24 | R << 16 | G << 8 | B; rows start at y * pitch and the pixels
    past the width of each row are left alone."""
    H = np.finfo(np.float32).max
    d = np.full(4, H, np.float32)
    n = np.zeros((4, 4), np.float32)
    for back, want in (([1, 0, 0, 0], 0x00FF0000), ([0, 1, 0, 0], 0x0000FF00), ([0, 0, 1, 0], 0x000000FF),
                       ([0, 0, 0, 1], 0xFF000000), ([2, -1, 0.5, 0], 0x00FF007F)):
        px = ora.shade(d, n, 2, 2, np.array(back, np.float32), [0, 0, 1], pitch=3).reshape(2, 3)
        assert np.all(px[:, :2] == want) and np.all(px[:, 2] == 0)


def test_shade_obliquity(ora):
    """c = -dot(cam, n) over random unit normals equals the clamp of the cosine
    between the view direction and the normal, to a step of 1/255."""
    rng = np.random.default_rng(9)
    n = rng.normal(size=(500, 4)).astype(np.float32)
    n[:, :3] /= np.linalg.norm(n[:, :3], axis=1, keepdims=True)
    cam = np.array([0.6, 0.0, -0.8], np.float32)
    px = ora.shade(np.ones(500, np.float32), n, 500, 1, BACK, cam)
    c = np.clip(-(n[:, :3].astype(np.float64) @ cam), 0, 1)
    got = px & 0xFF
    assert np.all(np.abs(got - np.floor(c * 255)) <= 1)
    assert np.all((px >> 24) == got) and np.all(((px >> 16) & 0xFF) == got)
