"""The C-ABI library loads and exports every function the public headers
declare; host-only entry points (scene generators) run without a GPU."""
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent


def _declared():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = h.read_text()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(rk_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol(rk):
    lib_path = rk.LIB_PATH
    out = subprocess.run(["nm", "-D", "--defined-only", str(lib_path)], capture_output=True, text=True, check=True)
    exported = {l.split()[-1] for l in out.stdout.splitlines() if " T " in l}
    declared = _declared()
    assert declared, "no declarations parsed"
    missing = sorted(declared - exported)
    assert not missing, missing
    for name in declared:
        getattr(rk.lib(), name)


def test_bindings_cover_every_declaration(rk):
    assert _declared() <= set(rk.declared_symbols())


def test_status_strings_and_errors(rk):
    L = rk.lib()
    assert L.rk_status_string(0) == b"success"
    assert L.rk_status_string(4) == b"workspace too small"
    # invalid arguments are reported, never exit
    assert L.rk_calc_mesh_boxes(None, None, None, 5, None, None) == 1
    assert L.rk_build_lbvh(None, None, None, 0, None, 1, 12, None, None, None, 0, None, None, 0) == 1
    assert L.rk_radix_sort_general(None, None, None, 10, 0, 0, None, 0) == 1
    assert L.rk_fast_build_workspace_size(1000, 1) > 0
    # lengths of 2^30 or more are refused before anything is launched
    big, fake = 1 << 30, 4096
    assert L.rk_radix_sort(None, fake, big, 32, fake, 16) == 6
    assert L.rk_radix_sort_general(None, fake, fake, big, 32, 1, fake, 16) == 6
    assert L.rk_radix_sort_pairs_u64(None, fake, fake, big, 64, fake, 16) == 6
    assert L.rk_bitonic_sort(None, fake, big, fake, 16) == 6
    assert L.rk_bitonic_sort_general(None, fake, fake, big, fake, 16) == 6
    assert L.rk_bubble_sort512(None, fake, fake, 513, 0) == 1


def test_reference_handoff_limits(rk):
    """rk_raytrace_set_handoff / get_handoff: defaults, bounds of the list size,
    NULL outputs; host-only."""
    import ctypes as C
    L = rk.lib()
    a, b, c = C.c_uint32(), C.c_uint32(), C.c_uint32()
    get = lambda: (L.rk_raytrace_get_handoff(C.byref(a), C.byref(b), C.byref(c)), a.value, b.value, c.value)
    assert get() == (0, 0xFFFFFFFE, 0xFFFFFFFE, 65536)
    assert L.rk_raytrace_set_handoff(1, 2, 0) == 1
    assert L.rk_raytrace_set_handoff(1, 2, 65537) == 1
    assert get() == (0, 0xFFFFFFFE, 0xFFFFFFFE, 65536)  # a refused call changes nothing
    assert L.rk_raytrace_get_handoff(None, C.byref(b), C.byref(c)) == 1
    try:
        assert L.rk_raytrace_set_handoff(0xFFFFFFFF, 7, 1) == 0
        assert get() == (0, 0xFFFFFFFF, 7, 1)
    finally:
        assert L.rk_raytrace_set_handoff(0xFFFFFFFE, 0xFFFFFFFE, 65536) == 0


def test_renderkit_exports_the_reference_sort_prototypes(rk):
    """The C++ free functions as the reference's callers declare them
    (samples/09_how_to_use_a_sorter.cpp:36-44, CUDA_radixSortGeneral.cu:27-30),
    and CUDA_bitonicSortGeneral also in the form of its definition."""
    out = subprocess.run(["nm", "-D", "--defined-only", str(rk.LIB_DIR / "librenderkit.so")], capture_output=True,
                         text=True, check=True).stdout
    for sym in ("_Z14CUDA_radixSortPjjj", "_Z21CUDA_radixSortGeneralPjS_jjb", "_Z16CUDA_bitonicSortIjEvPT_j",
                "_Z23CUDA_bitonicSortGeneralIjjEvPT0_S1_j", "_Z23CUDA_bitonicSortGeneralIjjEvPT_PT0_j",
                "_Z18CUDA_bubbleSort512PjS_jb"):
        assert sym in out, sym


def test_scene_generators_are_deterministic(rk):
    V, F = rk.icosphere(5)
    assert V.shape == (10242, 4) and F.shape == (20480, 4)
    assert np.allclose(np.linalg.norm(V[:, :3], axis=1), 1.0, atol=1e-6)
    hb = rk.hitbox(V)
    assert np.all(hb[0, :3] < -0.99) and np.all(hb[1, :3] > 0.99) and hb[0, 3] == 0 and hb[1, 3] == 0
    V2, F2 = rk.icosphere(5)
    assert np.array_equal(V, V2) and np.array_equal(F, F2)
    V, F = rk.heightfield(64, seed=0xC2)
    assert F.shape[0] == 2 * 63 * 63 and V[:, 2].min() == 0.0 and np.isclose(V[:, 2].max(), 0.2)
    V, F = rk.soup(1000, seed=3)
    assert V.shape == (3000, 4)
    Vb, _ = rk.soup(1000, seed=3)
    assert np.array_equal(V, Vb)
    O, D = rk.random_rays(100, seed=5)
    assert np.allclose(np.linalg.norm(D[:, :3], axis=1), 1.0, atol=1e-6)


def test_camera_matches_reference_convention(rk):
    # setCenter(0,0,3); lookAt(0,0,0): angle.x = 3*pi/2, y-up frame
    cam = rk.camera_lookat([0, 0, 3], [0, 0, 0], 60 * 0.01745329252, 256, 256)
    np.testing.assert_allclose(cam[0, :3], [0, 0, 3])
    O, D = rk.primary_rays_host(cam, 256, 256)
    c = D[128 * 256 + 128, :3]
    assert c[2] < -0.99 and abs(c[0]) < 1e-2 and abs(c[1]) < 1e-2
    assert D[128 * 256, 1] == 0.0  # row 128: y * rayY cancels the up term exactly


@pytest.mark.parametrize("eye,tgt,w,h", [((0, 0, 3), (0, 0, 0), 256, 256), ((1.2, 0.7, -2.5), (0.1, -0.3, 0.4), 100, 70),
                                         ((-4, 1, 0.5), (0, 0, 0), 33, 17)])
def test_host_camera_equals_oracle(rk, ora, eye, tgt, w, h):
    """The library's look-at camera and host ray generator give the oracle's
    restatement of CUDACamera::lookAt + CUDA_calcPrimaryRays bit for bit."""
    fov = np.float32(60.0) * np.float32(0.01745329252)
    cam = rk.camera_lookat(eye, tgt, float(fov), w, h)
    c = ora.Camera()
    c.set_center(*eye)
    c.look_at(*tgt)
    c.set_fov(fov)
    c.set_ratio(w, h)
    np.testing.assert_array_equal(cam, c.ray_constants(w, h))
    O, D = rk.primary_rays_host(cam, w, h)
    Oo, Do = c.primary_rays(w, h)
    np.testing.assert_array_equal(O, Oo)
    np.testing.assert_array_equal(D, Do)


def test_camera_api_host_program():
    """CUDACamera movement, accessors and matrices (tests/cpp/camera_api.cpp, no GPU call)."""
    exe = ROOT / "tests" / "cpp" / "_build" / "camera_api"
    assert exe.exists(), "build() compiles tests/cpp first"
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stdout + r.stderr
