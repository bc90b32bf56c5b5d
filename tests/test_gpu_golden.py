"""The GPU reference path reproduces the committed C1 digests without running
the oracle: reference LBVH at depths 12 and 31 (node words and boxes, indices,
reordered faces) and both traversals of the 256x256 frame."""
import json
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
GOLDEN = Path(__file__).resolve().parent / "golden"


def test_gpu_matches_committed_c1_digests(rk):
    import torch
    from cudart_amd import device as dev
    from test_golden import load_maker
    maker = load_maker()
    node_dtype = np.dtype([("w0", "<u4"), ("w1", "<u4"), ("bmin", "<f4", 3), ("bmax", "<f4", 3)])

    def t(a):
        a = np.ascontiguousarray(a)
        return torch.from_numpy(a.view(np.int32) if a.dtype == np.uint32 else a).cuda()

    def h(x):
        torch.cuda.synchronize()
        return x.cpu().numpy()

    def build(F, V, hb, depth):
        tree = dev.build_lbvh(t(F), t(V), hb, depth, True)
        tree.V = t(V)
        M = tree.node_count
        nodes = h(tree.nodes)[:M].copy().view(node_dtype).reshape(M)
        return nodes, h(tree.indices).view(np.uint32), h(tree.faces).view(np.uint32), tree

    def trace(tree, V, O, D, coh):
        fid, uv, dep = dev.trace_ref(tree, tree.V, t(O), t(D), 256, 256, coh)
        return h(fid).view(np.uint32), h(uv), h(dep)

    want = json.loads((GOLDEN / "c1_digests.json").read_text())
    assert maker.c1_outputs(rk, build, trace) == want
