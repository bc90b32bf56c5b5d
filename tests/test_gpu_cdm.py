"""The .cdm format against bytes assembled by hand from the reference's reader
and writer (CUDAMesh.cpp:134-400, CUDAMeshTree.cpp:114-289): chunks of int
magic, int length (the whole chunk, header included) and a payload; the hitbox
payload is one aabox, every other payload a uint count and the raw array.
Loaders read known chunks in sequence and skip unknown ones by their length, so
one file holds a mesh and its tree (samples 05 and 06)."""
import struct
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
EXE = ROOT / "tests" / "cpp" / "_build" / "cdm_format"

HITBOX, VERTICES, NORMALS, COLORS, TEXCOORDS, FACES = 0x3027AD79, 0xBA5C1CC7, 0x1DD14C48, 0x757B31AB, 0x2B77D260, 0x8E934163
NODES, INDICES, BOXES = 0x9077086B, 0x2CC8BD7D, 0xE54B7842


def _chunk(magic, payload):
    return struct.pack("<II", magic, 8 + len(payload)) + payload


def _array(magic, a):
    a = np.ascontiguousarray(a)
    return _chunk(magic, struct.pack("<I", a.shape[0]) + a.tobytes())


def _scene():
    V = np.array([[0, 0, 0, 0], [1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 1, 0], [1, 1, 0.5, 0]], np.float32)
    Nn = np.tile(np.array([0, 0, 1, 0], np.float32), (5, 1))
    Cc = np.arange(20, dtype=np.float32).reshape(5, 4) / 20
    T = np.arange(10, dtype=np.float32).reshape(5, 2) / 10
    F = np.array([[0, 1, 2, 7], [1, 4, 2, 8], [0, 2, 3, 9]], np.uint32)
    hb = np.array([[0, 0, 0, 0], [1, 1, 1, 0]], np.float32)
    nodes = np.zeros(3, [("w0", "<u4"), ("w1", "<u4"), ("bmin", "<f4", 3), ("bmax", "<f4", 3)])
    nodes[0] = (0 | (2 << 1), 1, (0, 0, 0), (1, 1, 1))
    nodes[1] = (1 | (2 << 1), 0, (0, 0, 0), (1, 1, 0.5))
    nodes[2] = (1 | (1 << 1), 2, (0, 0, 0), (0, 1, 1))
    idx = np.array([2, 0, 1], np.uint32)
    boxes = np.zeros((3, 2, 4), np.float32)
    for k, f in enumerate(F):
        boxes[k, 0] = V[f[:3]].min(0)
        boxes[k, 1] = V[f[:3]].max(0)
    return V, Nn, Cc, T, F, hb, nodes, idx, boxes


def _file(unknown=False):
    V, Nn, Cc, T, F, hb, nodes, idx, boxes = _scene()
    mesh = (_chunk(HITBOX, hb.tobytes()) + _array(VERTICES, V) + _array(NORMALS, Nn) + _array(COLORS, Cc)
            + _array(TEXCOORDS, T) + _array(FACES, F))
    tree = _array(NODES, nodes) + _array(INDICES, idx) + _array(BOXES, boxes)
    if unknown:  # a chunk neither loader knows, between the two halves and at the start
        junk = _chunk(0x12345678, b"\xAB" * 13)
        return junk + mesh + junk + tree
    return mesh + tree


def _run(tmp_path, data):
    src, out, dmp = tmp_path / "in.cdm", tmp_path / "out.cdm", tmp_path / "dump.bin"
    src.write_bytes(data)
    r = subprocess.run([str(EXE), str(src), str(out), str(dmp)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    raw, off, arrays = dmp.read_bytes(), 0, []
    while off < len(raw):
        n = int(np.frombuffer(raw[off:off + 4], np.uint32)[0])
        arrays.append(raw[off + 4:off + 4 + n])
        off += 4 + n
    return out.read_bytes(), arrays


def test_hand_assembled_layout_chunks():
    data = _file()
    # the hitbox chunk: magic, length 40, the box; no count word
    assert struct.unpack("<II", data[:8]) == (HITBOX, 40)
    assert struct.unpack("<II", data[40:48]) == (VERTICES, 12 + 5 * 16)
    assert struct.unpack("<I", data[48:52]) == (5,)


@pytest.mark.gpu
@pytest.mark.parametrize("unknown", [False, True])
def test_cdm_load_matches_hand_assembled_bytes(tmp_path, unknown):
    assert EXE.exists(), "build() compiles tests/cpp first"
    V, Nn, Cc, T, F, hb, nodes, idx, boxes = _scene()
    saved, arrays = _run(tmp_path, _file(unknown))
    expect = [hb, V, Nn, Cc, T, F, nodes, idx, boxes]
    for got, exp in zip(arrays, expect):
        assert got == np.ascontiguousarray(exp).tobytes()
    # saving the loaded mesh (overwrite) and tree (append) gives the reference layout byte for byte
    assert saved == _file(False)
