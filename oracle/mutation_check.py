"""Mutation check of the oracle (test infrastructure).

Applies each one-line mutation below to a scratch copy of oracle.cpp, rebuilds
the oracle there and runs the CPU test suite against it. Every mutation must
make at least one test fail; a mutation that survives names a rule of the
reference that no known answer pins.

    python oracle/mutation_check.py [name ...]
"""
import shutil
import subprocess
import sys
import tempfile
from pathlib import Path
REPO = Path(__file__).resolve().parent.parent
MUT = [
    ("scale 31->32", "scale = 31.0f / ext.x", "scale = 32.0f / ext.x"),
    ("sweep < -> <=", "if (T < T_min) { T_min = T; m_min = m - 1; }", "if (T <= T_min) { T_min = T; m_min = m - 1; }"),
    ("axis x>y -> >=", "if (ext.x > ext.y) axis", "if (ext.x >= ext.y) axis"),
    ("axis x>z -> >=", "axis = (ext.x > ext.z) ? 0 : 2;", "axis = (ext.x >= ext.z) ? 0 : 2;"),
    ("axis y>z -> >=", "else axis = (ext.y > ext.z) ? 1 : 2;", "else axis = (ext.y >= ext.z) ? 1 : 2;"),
    ("leaf size<224 -> <=", "if (size < 224 && T_min", "if (size <= 224 && T_min"),
    ("leaf (size+1) -> size", "T_min > (float)(size + 1) * 0.125f", "T_min > (float)(size) * 0.125f"),
    ("leaf T> -> >=", "T_min > (float)(size + 1) * 0.125f", "T_min >= (float)(size + 1) * 0.125f"),
    ("box cost 2 -> 1", "float T = 1.0f * 2.f", "float T = 1.0f * 1.f"),
    ("min leaf <=16 -> <16", "if (size <= 16) { leaf_node(); continue; }", "if (size < 16) { leaf_node(); continue; }"),
    ("partition <= -> <", "if (bin_of(boxes[id]) <= m_min)", "if (bin_of(boxes[id]) < m_min)"),
    ("hybrid root box", "queue.push_back({b, k, nd.w1", "queue.push_back({*root, k, nd.w1"),
    ("hybrid ids", "    count = n_existing;\n", "    count = n_existing + 2;\n"),
    ("refit left >0 -> >=", "if (l.bmax[2] - l.bmin[2] > 0.0f) {", "if (l.bmax[2] - l.bmin[2] >= 0.0f) {"),
    ("refit right >0 -> >=", "if (r.bmax[2] - r.bmin[2] > 0.0f) {", "if (r.bmax[2] - r.bmin[2] >= 0.0f) {"),
    ("tri u>1 -> >=", "if (u < 0.0f || u > 1.0f) return false;", "if (u < 0.0f || u >= 1.0f) return false;"),
    ("tri u<0 -> <=", "if (u < 0.0f || u > 1.0f) return false;", "if (u <= 0.0f || u > 1.0f) return false;"),
    ("tri v<0 -> <=", "if (v < 0.0f || u + v > 1.0f) return false;", "if (v <= 0.0f || u + v > 1.0f) return false;"),
    ("tri u+v>1 -> >=", "if (v < 0.0f || u + v > 1.0f) return false;", "if (v < 0.0f || u + v >= 1.0f) return false;"),
    ("tri t>=0 -> >", "return t >= 0.0f;\n}\n\n// intersectBoxRay", "return t > 0.0f;\n}\n\n// intersectBoxRay"),
    ("cam ray x step sign", "sub4(sub4(cst[1], muls(cst[2], (float)x))", "add4(sub4(cst[1], muls(cst[2], (float)x))"),
    ("cam ray y step sign", "muls(cst[3], (float)y));", "muls(cst[3], -(float)y));"),
    ("cam ratio", "float x = y * ratio;", "float x = y / ratio;"),
    ("cam 2x step", "muls(muls(right, 2.0f * x)", "muls(muls(right, x)"),
    ("cam pitch/yaw swap", "F4 camDir = {cosf(pitch) * cosf(yaw), sinf(yaw), sinf(pitch) * cosf(yaw), 0.0f};",
     "F4 camDir = {cosf(yaw) * cosf(pitch), sinf(pitch), sinf(yaw) * cosf(pitch), 0.0f};"),
    ("cam camUp roll", "F4 camUp = {cosf(pitch) * sinf(roll), cosf(roll), 0.0f, 0.0f};",
     "F4 camUp = {0.0f, 1.0f, 0.0f, 0.0f};"),
    ("cam up order", "F4 up = normalize4(cross4(right, camDir));", "F4 up = normalize4(cross4(camDir, right));"),
    ("cam branch3 +PI", "else angle.x = acosf(-dir.x / cosf(-angle.y)) + CAM_PI;",
     "else angle.x = acosf(-dir.x / cosf(-angle.y));"),
    ("cam yaw clamp", "if (angle.y > CAM_PI * 0.5f - 0.001f) angle.y = CAM_PI * 0.5f - 0.001f;\n  if (dir.x",
     "if (angle.y > CAM_PI * 0.5f - 0.01f) angle.y = CAM_PI * 0.5f - 0.01f;\n  if (dir.x"),
    ("cam dirFromAngle z", "dir.z = sinf(angle.x) * cosf(angle.y);", "dir.z = cosf(angle.x) * sinf(angle.y);"),
    ("shade R/B swap", "(((uint32_t)v.x & 0xFFu) << 16) | (((uint32_t)v.y & 0xFFu) << 8) | ((uint32_t)v.z & 0xFFu) |",
     "(((uint32_t)v.z & 0xFFu) << 16) | (((uint32_t)v.y & 0xFFu) << 8) | ((uint32_t)v.x & 0xFFu) |"),
    ("shade round", "v = muls(v, 255.0f);", "v = add4(muls(v, 255.0f), F4{0.5f, 0.5f, 0.5f, 0.5f});"),
    ("shade depth <=", "if (depths[k] < HUGE_F) {", "if (depths[k] <= HUGE_F) {"),
    ("shade sign", "float c = -dot4(cam, n);", "float c = dot4(cam, n);"),
    ("shade normal w", "        n.w = 0.0f;\n        float c", "        float c"),
]
only = sys.argv[1:]
base = Path(tempfile.mkdtemp(prefix="oracle_mut_"))
shutil.rmtree(base)
def _ignored_top_dirs():
    """Top-level directories that git ignores (working files, run outputs, caches)."""
    r = subprocess.run(["git", "-C", str(REPO), "ls-files", "--others", "--ignored", "--exclude-standard",
                        "--directory"], capture_output=True, text=True, check=True)
    return {l.rstrip("/") for l in r.stdout.splitlines() if l.endswith("/") and "/" not in l.rstrip("/")}


_TOP_SKIP = _ignored_top_dirs() | {".git"}


def _skip(d, names):  # the oracle's own build is redone per mutation; the rest is copied as built
    skip = {"__pycache__", ".pytest_cache"}
    if Path(d).resolve() == REPO:
        skip |= _TOP_SKIP
    if Path(d).resolve() == (REPO / "oracle").resolve():
        skip.add("_build")
    return [n for n in names if n in skip]


shutil.copytree(REPO, base, ignore=_skip)
src = (REPO / "oracle/oracle.cpp").read_text()
res = []
for name, a, b in MUT:
    if only and name not in only:
        continue
    n = src.count(a)
    assert n >= 1, (name, "pattern not found")
    (base / "oracle/oracle.cpp").write_text(src.replace(a, b))
    subprocess.run(["make", "-s", "-C", str(base / "oracle")], check=True)
    r = subprocess.run([sys.executable, "-m", "pytest", "-x", "-q", "-m", "not gpu", "-p", "no:cacheprovider",
                        "tests"], cwd=base, capture_output=True, text=True)
    last = [l for l in r.stdout.splitlines() if l.strip()][-1]
    failed = [l for l in r.stdout.splitlines() if l.startswith("FAILED")]
    res.append((name, r.returncode != 0, failed[:1], last))
    print(f"{name:28s} caught={r.returncode != 0}  {failed[:1]} {last}", flush=True)
(base / "oracle/oracle.cpp").write_text(src)
shutil.rmtree(base)
ok = all(c for _, c, _, _ in res)
print("ALL CAUGHT" if ok else "SOME SURVIVED")
sys.exit(0 if ok else 1)
