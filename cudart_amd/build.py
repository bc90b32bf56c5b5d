"""Build the native libraries in-tree.

  cudart_amd/lib/librk.so       C ABI (include/rk.h, include/rk_scenes.h): gfx950
                                kernels built with hipcc, host generators with g++.
  cudart_amd/lib/librenderkit.so  the C++ drop-in classes (include/renderkit/),
                                host-only g++, linked against librk.so.
  oracle/_build/liboracle.so    the CPU restatement used by tests (oracle/Makefile).

Every HIP source is compiled with -ffp-contract=off (the reference arithmetic is
reproduced operation by operation) and with packed-FP32 instructions disabled.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
PKG = ROOT / "cudart_amd"
CSRC = PKG / "csrc"
CPP = PKG / "cpp"
LIB = PKG / "lib"
OBJ = PKG / "_build"
ORACLE = ROOT / "oracle"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("RK_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-Wno-unused-command-line-argument",
    f"-I{ROOT / 'include'}",
]
HOST_FLAGS = ["-O2", "-fPIC", "-std=c++17", "-ffp-contract=off", "-pthread", f"-I{ROOT / 'include'}"]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {cmd[0]} {cmd[-1]}")


def _compile(src: Path) -> Path:
    out = OBJ / (src.stem + (".hip.o" if src.suffix == ".hip" else ".o"))
    if src.suffix == ".hip":
        _run([HIPCC, *HIP_FLAGS, "-c", str(src), "-o", str(out)])
    else:
        _run(["g++", *HOST_FLAGS, "-c", str(src), "-o", str(out)])
    return out


def build_rk(jobs: int = 8) -> Path:
    OBJ.mkdir(exist_ok=True)
    LIB.mkdir(exist_ok=True)
    srcs = sorted(CSRC.glob("*.hip")) + sorted(CSRC.glob("*.cpp"))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    out = LIB / "librk.so"
    tmp = LIB / "librk.so.tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-pthread", *map(str, objs), "-o", str(tmp)])
    os.replace(tmp, out)
    return out


def build_renderkit() -> Path | None:
    srcs = sorted(CPP.glob("*.cpp"))
    if not srcs:
        return None
    out = LIB / "librenderkit.so"
    tmp = LIB / "librenderkit.so.tmp"
    _run(["g++", *HOST_FLAGS, "-shared", *map(str, srcs), f"-L{LIB}", "-lrk", "-Wl,-rpath,$ORIGIN", "-o", str(tmp)])
    os.replace(tmp, out)
    return out


def build_oracle() -> Path:
    _run(["make", "-s", "-C", str(ORACLE)])
    return ORACLE / "_build" / "liboracle.so"


def build_cpp_tests() -> None:
    """The C++ programs the GPU tests run (drop-in usage of include/renderkit/)."""
    src_dir = ROOT / "tests" / "cpp"
    out_dir = src_dir / "_build"
    out_dir.mkdir(exist_ok=True)
    for src in sorted(src_dir.glob("*.cpp")):
        exe = out_dir / src.stem
        _run(["g++", *HOST_FLAGS, str(src), f"-L{LIB}", "-lrenderkit", "-lrk", f"-Wl,-rpath,{LIB}", "-o", str(exe)])


def build_all(jobs: int = 8) -> None:
    build_rk(jobs)
    build_renderkit()
    build_oracle()
    build_cpp_tests()


if __name__ == "__main__":
    build_all()
    print("built", LIB / "librk.so")
