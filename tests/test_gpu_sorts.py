"""The reference's standalone sorts through the C ABI, against the oracle:
CUDA_bitonicSort<uint> (rk_bitonic_sort), CUDA_bitonicSortGeneral<uint, uint>
(rk_bitonic_sort_general, with the reference tie order at any length) and
CUDA_bubbleSort512 (rk_bubble_sort512); then the C++ free functions driven as
samples/09_how_to_use_a_sorter.cpp drives them, on its 1,397,121 keys."""
import subprocess
from pathlib import Path

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parent.parent
EXE = ROOT / "tests" / "cpp" / "_build" / "sample09_sorts"
SIZES = [0, 1, 2, 127, 128, 129, 255, 256, 257, 512, 513, 700, 32768, 32769, 300001, 1397121]


def _keys(n, seed, distinct):
    rng = np.random.default_rng(seed)
    return rng.integers(0, distinct, n, dtype=np.uint32)


def _t(a):
    return torch.from_numpy(np.ascontiguousarray(a, np.uint32).view(np.int32)).cuda()


def _h(t):
    return t.cpu().numpy().view(np.uint32)


@pytest.mark.parametrize("n", SIZES)
@pytest.mark.parametrize("distinct", [7, 2097152])
def test_bitonic_sort_general_tie_order(rk, ora, n, distinct):
    from cudart_amd import device as dev
    keys = _keys(n, n + distinct, distinct)
    vals = np.arange(n, dtype=np.uint32)
    k, v, ws = dev.bitonic_sort_general(_t(keys), _t(vals))
    ek, ev, _ = ora.bitonic_sort_general(keys, vals)
    np.testing.assert_array_equal(_h(k), ek)
    np.testing.assert_array_equal(_h(v), ev)
    assert dev.workspace_flags(ws) == 0


@pytest.mark.parametrize("n", SIZES)
def test_bitonic_sort_values(rk, ora, n):
    from cudart_amd import device as dev
    vals = _keys(n, 3 * n + 1, 2097152)
    vals[::5] = 0xFFFFFFFF  # full 32-bit values
    v, ws = dev.bitonic_sort(_t(vals))
    np.testing.assert_array_equal(_h(v), ora.bitonic_sort(vals))
    np.testing.assert_array_equal(_h(v), np.sort(vals))
    assert dev.workspace_flags(ws) == 0


@pytest.mark.parametrize("n", [0, 1, 2, 3, 64, 100, 511, 512])
@pytest.mark.parametrize("init", [False, True])
def test_bubble_sort512(rk, ora, n, init):
    from cudart_amd import device as dev
    keys = _keys(n, n, 5)
    idx = np.arange(n, dtype=np.uint32)[::-1].copy() * 3 + 11
    k, i = dev.bubble_sort512(_t(keys), _t(idx), init)
    ek, ei = ora.bubble_sort512(keys, idx, init)
    np.testing.assert_array_equal(_h(k), ek)
    np.testing.assert_array_equal(_h(i), ei)


def test_bubble_sort512_rejects_more_than_512(rk):
    t = torch.zeros(513, dtype=torch.int32, device="cuda")
    assert rk.lib().rk_bubble_sort512(None, t.data_ptr(), t.data_ptr(), 513, 0) == 1  # RK_ERROR_INVALID_VALUE


def test_radix_sorts_leave_zero_flags(rk, ora):
    from cudart_amd import device as dev
    n = 200003
    keys = _keys(n, 9, 1000)
    k, v = _t(keys), torch.zeros(n, dtype=torch.int32, device="cuda")
    ws = dev.workspace(rk.lib().rk_sort_workspace_size(n, 4))
    ws.fill_(0xAB)  # the call clears the word itself
    assert rk.lib().rk_radix_sort_general(None, k.data_ptr(), v.data_ptr(), n, 32, 1, ws.data_ptr(), ws.numel()) == 0
    assert dev.workspace_flags(ws) == 0
    ek, ev, _ = ora.sort_pairs_ref(keys, np.zeros(n, np.uint32), 32, init=True)
    np.testing.assert_array_equal(_h(k), ek)
    np.testing.assert_array_equal(_h(v), ev)


@pytest.mark.parametrize("n", [1397121, 300])
def test_sample09_free_functions(rk, ora, tmp_path, n):
    assert EXE.exists(), "build() compiles tests/cpp first"
    out = tmp_path / "sorts.bin"
    r = subprocess.run([str(EXE), str(out), str(n), "9"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    data = out.read_bytes()
    assert int(np.frombuffer(data[:4], np.uint32)[0]) == n
    off, ch = 4, []
    while off < len(data):
        b = int(np.frombuffer(data[off:off + 4], np.uint32)[0])
        ch.append(np.frombuffer(data[off + 4:off + 4 + b], np.uint32))
        off += 4 + b
    inp = ch[0]
    assert inp.max() < 2097152
    np.testing.assert_array_equal(ch[1], np.sort(inp))                  # CUDA_radixSort, 21 bits
    np.testing.assert_array_equal(ch[2], ora.bitonic_sort(inp))         # CUDA_bitonicSort<uint>
    ek, ev, _ = ora.sort_pairs_ref(inp, np.arange(n, dtype=np.uint32), 21)  # CUDA_radixSortGeneral, 21 bits
    np.testing.assert_array_equal(ch[3], ek)
    np.testing.assert_array_equal(ch[4], ev)
    ek, ev, _ = ora.bitonic_sort_general(inp, np.arange(n, dtype=np.uint32))  # CUDA_bitonicSortGeneral<uint,uint>
    np.testing.assert_array_equal(ch[5], ek)
    np.testing.assert_array_equal(ch[6], ev)
    m = min(n, 512)
    ek, ei = ora.bubble_sort512(inp[:m], np.zeros(m, np.uint32), True)  # CUDA_bubbleSort512, init
    np.testing.assert_array_equal(ch[7], ek)
    np.testing.assert_array_equal(ch[8], ei)
