"""rk_fast_build_noalloc (the build with its workspace in the library's own
device array) against rk_fast_build with a caller workspace, through the C ABI:
every output word equal, on consecutive calls whose second is smaller (the
array keeps the first call's contents), for both key widths; sizes beyond the
array are refused before any launch."""
import ctypes as C

import numpy as np
import pytest
import torch

from cudart_amd import device as dev
from cudart_amd import native as N

pytestmark = pytest.mark.gpu

RK_ERROR_WORKSPACE = 4
LIMIT = {N.RK_MORTON_30: 1_944_805, N.RK_MORTON_63: 1_720_320}


def _buffers(n):
    z = lambda *shape, dt=torch.int32: torch.zeros(shape, dtype=dt, device="cuda")
    return dict(nodes=z(max(1, n), 16), tris=z(n, 12, dt=torch.float32), indices=z(n),
                boxes=z(n, 2, 4, dt=torch.float32), info=z(N.RK_INFO_WORDS))


def _desc(b, n):
    return N.FastBVH(b["nodes"].data_ptr(), b["tris"].data_ptr(), b["indices"].data_ptr(), b["boxes"].data_ptr(), n, 0)


def _build(F, V, hb, mode, leaf_max, noalloc):
    n = F.shape[0]
    b = _buffers(n)
    desc = _desc(b, n)
    box = N.aabox(hb)
    L = N.lib()
    if noalloc:
        s = L.rk_fast_build_noalloc(None, F.data_ptr(), V.data_ptr(), n, C.byref(box), mode, leaf_max, C.byref(desc),
                                    b["info"].data_ptr())
    else:
        ws = dev.workspace(L.rk_fast_build_workspace_size(n, mode))
        s = L.rk_fast_build(None, F.data_ptr(), V.data_ptr(), n, C.byref(box), mode, leaf_max, C.byref(desc),
                            b["info"].data_ptr(), ws.data_ptr(), ws.numel())
    N.check(s, "noalloc" if noalloc else "fast_build")
    torch.cuda.synchronize()
    return b, desc


def _trace(desc, O, D):
    n = O.shape[0]
    fid, uv, dep = dev.new_outputs(n)
    N.check(N.lib().rk_fast_trace(None, C.byref(desc), O.data_ptr(), D.data_ptr(), n, fid.data_ptr(), uv.data_ptr(),
                                  dep.data_ptr(), None), "fast_trace")
    torch.cuda.synchronize()
    return fid.cpu(), uv.cpu(), dep.cpu()


@pytest.mark.parametrize("mode,leaf_max", [(N.RK_MORTON_30, 1), (N.RK_MORTON_63, 4)])
def test_noalloc_build_equals_workspace_build(mode, leaf_max):
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a).view(np.int32) if a.dtype == np.uint32 else a).cuda()
    # a 300,000-face soup (climbs across chunks in a second pass), then a smaller
    # 120 x 120 terrain (28,322 faces, climbs inline) into the same array
    cases = [N.soup(300_000, seed=0x5A), N.heightfield(120, seed=7)]
    for V, F in cases:
        hb = N.hitbox(V)
        Ft, Vt = t(F), t(V)
        print(f"noalloc case: mode={mode} leaf_max={leaf_max} faces={F.shape[0]}", flush=True)
        ref, rdesc = _build(Ft, Vt, hb, mode, leaf_max, noalloc=False)
        got, gdesc = _build(Ft, Vt, hb, mode, leaf_max, noalloc=True)
        for k in ("nodes", "tris", "indices", "boxes", "info"):
            assert torch.equal(ref[k], got[k]), k
        assert int(got["info"][N.RK_INFO_FLAGS]) == 0
        O, D = N.random_rays(20_000, seed=F.shape[0])
        a = _trace(rdesc, t(O), t(D))
        b = _trace(gdesc, t(O), t(D))
        for x, y in zip(a, b):
            assert torch.equal(x, y)
        assert (a[2] < np.finfo(np.float32).max).sum() > 100


def test_noalloc_rejects_sizes_beyond_the_array():
    L = N.lib()
    for mode, lim in LIMIT.items():
        assert L.rk_fast_build_workspace_size(lim, mode) <= (128 << 20) < L.rk_fast_build_workspace_size(lim + 1, mode)
    b = _buffers(16)
    desc = _desc(b, 16)
    box = N.aabox(np.array([[0, 0, 0, 0], [1, 1, 1, 0]], np.float32))
    F = torch.zeros((16, 4), dtype=torch.int32, device="cuda")
    V = torch.zeros((16, 4), dtype=torch.float32, device="cuda")
    for mode, lim in LIMIT.items():
        # refused on the size alone, before any kernel reads the (too small) arrays
        s = L.rk_fast_build_noalloc(None, F.data_ptr(), V.data_ptr(), lim + 1, C.byref(box), mode, 1, C.byref(desc),
                                    b["info"].data_ptr())
        assert s == RK_ERROR_WORKSPACE
