// ORACLE — test infrastructure only.
//
// A CPU restatement of cudart's mesh-tree build and closest-hit trace path.
// Each function cites the reference file:line it follows. It is loaded only by
// tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, and
// only as the checker or the timed CPU baseline. The product never links it.
//
// Compiled with g++ -O2 -ffp-contract=off (no fast-math, SSE float), so every
// float operation below is one IEEE-754 single-precision operation, in the
// order written.
//
// Where the reference is nondeterministic or undefined, the oracle uses a
// fixed definition. Each case is marked "DEFINED:" at the line it affects:
//   - the radix scatter's lane serialisation (RSG.cu:183-231): stable;
//   - the kernel_findSplits append order (LBVH.cu:227): ascending position;
//   - the Merge2x out-of-bounds write (BSG.cu:254, BS.cu:196): ignored;
//     the intended sort is computed and the trigger condition is reported;
//   - the split-pointer walk past the end of the split list
//     (LBVH.cu:381-392): leaf, reported;
//   - face_id / coords on a miss (RT.cu:116-118,215-216): left unchanged;
//   - out-of-image threads in the packet kernel (RT.cu:235-242): they do
//     not take part in the block-wide hit vote;
//   - SAH node numbering (SAH.cu:625): breadth-first, children allocated in
//     queue order;
//   - normalize() uses 1.0f/sqrtf, not the device rsqrtf (types/_cuda_float4.h:181-186);
//   - the camera's trigonometry is the host C library's sinf/cosf/tanf/asinf/acosf.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>
#include <chrono>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORA_API extern "C" __attribute__((visibility("default")))

namespace {

struct F4 { float x, y, z, w; };
struct F3 { float x, y, z; };
struct U4 { uint32_t x, y, z, w; };
struct Box { F4 mn, mx; };
struct Node { uint32_t w0, w1; float bmin[3], bmax[3]; };
static_assert(sizeof(Node) == 32, "bvnode is 32 bytes");
static_assert(sizeof(Box) == 32, "aabox is 32 bytes");

const float HUGE_F = 3.402823466e+38f;  // CUDART_NORM_HUGE_F (FLT_MAX), CUDABox.h:50-54

// Device fminf/fmaxf are IEEE minNum/maxNum: a NaN operand yields the other.
inline float fmn(float a, float b) { if (a != a) return b; if (b != b) return a; return a < b ? a : b; }
inline float fmx(float a, float b) { if (a != a) return b; if (b != b) return a; return a > b ? a : b; }

inline F4 sub4(F4 a, F4 b) { return {a.x - b.x, a.y - b.y, a.z - b.z, a.w - b.w}; }
inline F4 add4(F4 a, F4 b) { return {a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w}; }
inline F4 mul4(F4 a, F4 b) { return {a.x * b.x, a.y * b.y, a.z * b.z, a.w * b.w}; }
inline F4 muls(F4 a, float s) { return {a.x * s, a.y * s, a.z * s, a.w * s}; }
inline F4 min4(F4 a, F4 b) { return {fmn(a.x, b.x), fmn(a.y, b.y), fmn(a.z, b.z), fmn(a.w, b.w)}; }
inline F4 max4(F4 a, F4 b) { return {fmx(a.x, b.x), fmx(a.y, b.y), fmx(a.z, b.z), fmx(a.w, b.w)}; }
// types/_cuda_float4.h:171-174: ((x*x + y*y) + z*z) + w*w, left to right.
inline float dot4(F4 a, F4 b) { return a.x * b.x + a.y * b.y + a.z * b.z + a.w * b.w; }
// types/_cuda_float4.h:198-208: w := 0.
inline F4 cross4(F4 a, F4 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x, 0.0f};
}
// types/_cuda_float4.h:181-186 with rsqrtf(x) read as 1.0f/sqrtf(x) (the
// reference's own host definition, types/_cuda_float.h:21-24).
inline F4 normalize4(F4 v) { float inv_len = 1.0f / sqrtf(dot4(v, v)); return muls(v, inv_len); }

// cvt.rzi.u32.f32 semantics: truncate, saturate, NaN -> 0.
inline uint32_t f2u_sat(float f) {
  if (!(f > 0.0f)) return 0u;
  if (f >= 4294967296.0f) return 0xFFFFFFFFu;
  return (uint32_t)f;
}
// cvt.rzi.s32.f32 semantics: truncate, saturate, NaN -> 0.
inline int f2i_sat(float f) {
  if (f != f) return 0;
  if (f >= 2147483648.0f) return 2147483647;
  if (f <= -2147483648.0f) return (-2147483647 - 1);
  return (int)f;
}

inline int nthreads() {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

}  // namespace

// ---------------------------------------------------------------------------
// A6: per-face boxes and index init — CUDA_calcMeshBoxes.cu:30-36 (initIndices),
// :38-59 (calcBoxes: reset(v0), merge(v1), merge(v2) as float4 incl. .w).
ORA_API void ora_calc_boxes(const U4* faces, const F4* verts, uint32_t n, Box* boxes, uint32_t* indices) {
#pragma omp parallel for schedule(static)
  for (long k = 0; k < (long)n; ++k) {
    U4 f = faces[k];
    F4 v0 = verts[f.x], v1 = verts[f.y], v2 = verts[f.z];
    Box b;
    b.mn = v0; b.mx = v0;                      // CUDABox.h:86-90 reset(v)
    b.mn = min4(b.mn, v1); b.mx = max4(b.mx, v1);  // CUDABox.h:112-116 merge(v)
    b.mn = min4(b.mn, v2); b.mx = max4(b.mx, v2);
    boxes[k] = b;
    indices[k] = (uint32_t)k;
  }
}

// ---------------------------------------------------------------------------
// A7: Morton codes. Mode 0 = the reference 32-bit code, LBVH.cu:115-173:
//   sx = 2047/len.x, sy = 2047/len.y, sz = 1023/len.z; c = 0.5f*(min+max);
//   q = (uint)((c - root.min) * s); bit k of y -> 3k, x -> 3k+1, z -> 3k+2.
// Mode 1 = 30-bit 10/10/10 code, mode 2 = 63-bit 21/21/21 code (new modes
// with no reference counterpart: q = min((uint)((c-min)*(2^b/len)), 2^b-1),
// x in the highest bit of each triple).
static inline uint32_t morton_ref32(const Box& b, F4 rmin, F4 rmax) {
  F4 len = sub4(rmax, rmin);
  float sx = 2047.0f / len.x, sy = 2047.0f / len.y, sz = 1023.0f / len.z;
  F4 c = muls(add4(b.mn, b.mx), 0.5f);
  uint32_t cx = f2u_sat((c.x - rmin.x) * sx);
  uint32_t cy = f2u_sat((c.y - rmin.y) * sy);
  uint32_t cz = f2u_sat((c.z - rmin.z) * sz);
  uint32_t code = 0;
  code |= (cy & 0x0001) << 0;  code |= (cx & 0x0001) << 1;  code |= (cz & 0x0001) << 2;
  code |= (cy & 0x0002) << 2;  code |= (cx & 0x0002) << 3;  code |= (cz & 0x0002) << 4;
  code |= (cy & 0x0004) << 4;  code |= (cx & 0x0004) << 5;  code |= (cz & 0x0004) << 6;
  code |= (cy & 0x0008) << 6;  code |= (cx & 0x0008) << 7;  code |= (cz & 0x0008) << 8;
  code |= (cy & 0x0010) << 8;  code |= (cx & 0x0010) << 9;  code |= (cz & 0x0010) << 10;
  code |= (cy & 0x0020) << 10; code |= (cx & 0x0020) << 11; code |= (cz & 0x0020) << 12;
  code |= (cy & 0x0040) << 12; code |= (cx & 0x0040) << 13; code |= (cz & 0x0040) << 14;
  code |= (cy & 0x0080) << 14; code |= (cx & 0x0080) << 15; code |= (cz & 0x0080) << 16;
  code |= (cy & 0x0100) << 16; code |= (cx & 0x0100) << 17; code |= (cz & 0x0100) << 18;
  code |= (cy & 0x0200) << 18; code |= (cx & 0x0200) << 19; code |= (cz & 0x0200) << 20;
  code |= (cy & 0x0400) << 20; code |= (cx & 0x0400) << 21;
  return code;
}

static inline uint64_t morton_wide(const Box& b, F4 rmin, F4 rmax, int bits) {
  F4 len = sub4(rmax, rmin);
  float cells = (float)(1u << bits);
  uint32_t qmax = (1u << bits) - 1u;
  float sx = cells / len.x, sy = cells / len.y, sz = cells / len.z;
  F4 c = muls(add4(b.mn, b.mx), 0.5f);
  uint32_t q[3] = {f2u_sat((c.x - rmin.x) * sx), f2u_sat((c.y - rmin.y) * sy), f2u_sat((c.z - rmin.z) * sz)};
  uint64_t code = 0;
  for (int a = 0; a < 3; ++a) if (q[a] > qmax) q[a] = qmax;
  for (int k = 0; k < bits; ++k) {
    code |= (uint64_t)((q[0] >> k) & 1u) << (3 * k + 2);
    code |= (uint64_t)((q[1] >> k) & 1u) << (3 * k + 1);
    code |= (uint64_t)((q[2] >> k) & 1u) << (3 * k + 0);
  }
  return code;
}

ORA_API void ora_morton32(const Box* boxes, uint32_t n, const F4* root_min, const F4* root_max, uint32_t* codes) {
  F4 rmin = *root_min, rmax = *root_max;
#pragma omp parallel for schedule(static)
  for (long k = 0; k < (long)n; ++k) codes[k] = morton_ref32(boxes[k], rmin, rmax);
}

ORA_API void ora_morton_wide(const Box* boxes, uint32_t n, const F4* root_min, const F4* root_max, int bits, uint64_t* codes) {
  F4 rmin = *root_min, rmax = *root_max;
#pragma omp parallel for schedule(static)
  for (long k = 0; k < (long)n; ++k) codes[k] = morton_wide(boxes[k], rmin, rmax, bits);
}

// ---------------------------------------------------------------------------
// A8-A10: sorts with the reference's size-class tie rules.

// CUDA_bubbleSort512.cu:42-96 — odd-even transposition over 2n phases, swap on '>'.
static void odd_even_sort(uint32_t* k, uint32_t* v, uint32_t n) {
  for (uint32_t i = 0; i < 2 * n; ++i) {
    uint32_t start = (i & 1u) ? 1u : 0u;
    for (uint32_t a = start; a + 1 < n; a += 2) {
      if (k[a] > k[a + 1]) { std::swap(k[a], k[a + 1]); if (v) std::swap(v[a], v[a + 1]); }
    }
  }
}

// BSG.cu:62-105 — bitonic network over one full 128-element bin.
static void bitonic128(uint32_t* k, uint32_t* idx) {
  const uint32_t C = 128;
  for (uint32_t i = 2; i <= C; i <<= 1) {
    for (uint32_t j = i >> 1; j > 0; j >>= 1) {
      for (uint32_t n = 0; n < C; ++n) {
        uint32_t ixj = n ^ j;
        if (ixj > n) {
          uint32_t ka = k[n], kb = k[ixj];
          bool sw = ((n & i) == 0) ? (ka > kb) : (ka < kb);
          if (sw) { std::swap(k[n], k[ixj]); if (idx) std::swap(idx[n], idx[ixj]); }
        }
      }
    }
  }
}

// Stable merge of [a,p) and [p,q) with '<=' (BSG.cu:214-244, :292-328, :415-436).
static void stable_merge(const uint32_t* k, const uint32_t* v, uint32_t a, uint32_t p, uint32_t q,
                         uint32_t* ko, uint32_t* vo) {
  uint32_t i = a, j = p, c = a;
  while (i < p && j < q) {
    if (k[i] <= k[j]) { ko[c] = k[i]; if (v) vo[c] = v[i]; ++i; }
    else { ko[c] = k[j]; if (v) vo[c] = v[j]; ++j; }
    ++c;
  }
  while (i < p) { ko[c] = k[i]; if (v) vo[c] = v[i]; ++i; ++c; }
  while (j < q) { ko[c] = k[j]; if (v) vo[c] = v[j]; ++j; ++c; }
}

// CUDA_bitonicSortGeneral BSG.cu:504-690: the permutation is carried in a local
// index array and applied to the values at the end (kernel_bitonicSortFinal).
// Returns 1 when kernel_bitonicMerge2x would write past 'n' (BSG.cu:254).
static int bitonic_general(uint32_t* keys, uint32_t* vals, uint32_t n) {
  if (n <= 1) return 0;
  std::vector<uint32_t> idx(n);
  for (uint32_t i = 0; i < n; ++i) idx[i] = i;
  const uint32_t BIN = 128, BUCKET = 256;
  uint32_t bins = (n + BIN - 1) / BIN;
  for (uint32_t b = 0; b < bins; ++b) {
    uint32_t o = b * BIN, size = std::min(BIN, n - o);
    if (size <= 1) continue;
    if (size == BIN) bitonic128(keys + o, idx.data() + o);
    else odd_even_sort(keys + o, idx.data() + o, size);  // BSG.cu:107-162
  }
  int oob = 0;
  if (n > BIN) {
    std::vector<uint32_t> k2(keys, keys + n), i2(idx.begin(), idx.end());
    for (uint32_t b = 0; b + 1 < bins; b += 2) {  // kernel_bitonicMerge2x, bins/2 blocks
      uint32_t o = b * BIN, p = o + BIN, q = std::min(p + BIN, n);
      if (q - o < BUCKET) oob = 1;
      stable_merge(keys, idx.data(), o, p, q, k2.data(), i2.data());
    }
    std::copy(k2.begin(), k2.begin() + n, keys);
    std::copy(i2.begin(), i2.begin() + n, idx.begin());
    if (n > BUCKET) {  // BSG.cu:584-640: bucket merges until bucket_size >= n
      for (uint32_t bs = BUCKET; bs < n; bs *= 2) {
        for (uint32_t a = 0; a < n; a += 2 * bs) {
          uint32_t p = std::min(n, a + bs), q = std::min(n, a + 2 * bs);
          stable_merge(keys, idx.data(), a, p, q, k2.data(), i2.data());
        }
        std::copy(k2.begin(), k2.begin() + n, keys);
        std::copy(i2.begin(), i2.begin() + n, idx.begin());
      }
    }
  }
  if (vals) {  // kernel_bitonicSortFinal BSG.cu:484-490
    std::vector<uint32_t> old(vals, vals + n);
    for (uint32_t i = 0; i < n; ++i) vals[i] = old[idx[i]];
  }
  return oob;
}

// CUDA_radixSortGeneral RSG.cu:397-482 (keys with values).
// DEFINED: the radix scatter (RSG.cu:183-231) is stable.
// Returns 1 when the Merge2x out-of-bounds write would fire.
ORA_API int ora_sort_pairs_ref(uint32_t* keys, uint32_t* vals, uint32_t n, uint32_t bits, int init_indices) {
  uint32_t byte_depth = 1 + (bits - 1) / 8;
  if (n <= 512) {  // RSG.cu:401-406 -> CUDA_bubbleSort512.cu:114-167
    if (init_indices) for (uint32_t i = 0; i < n; ++i) vals[i] = i;
    if (n <= 1) return 0;
    odd_even_sort(keys, vals, n);
    return 0;
  }
  if (init_indices) for (uint32_t i = 0; i < n; ++i) vals[i] = i;  // RSG.cu:409-422
  if (n <= byte_depth * 256u * 32u) return bitonic_general(keys, vals, n);  // RSG.cu:424-429
  uint32_t mask = byte_depth >= 4 ? 0xFFFFFFFFu : ((1u << (8 * byte_depth)) - 1u);
  std::vector<uint32_t> perm(n);
  for (uint32_t i = 0; i < n; ++i) perm[i] = i;
  std::stable_sort(perm.begin(), perm.end(), [&](uint32_t a, uint32_t b) { return (keys[a] & mask) < (keys[b] & mask); });
  std::vector<uint32_t> k2(n), v2(n);
  for (uint32_t i = 0; i < n; ++i) { k2[i] = keys[perm[i]]; v2[i] = vals[perm[i]]; }
  std::copy(k2.begin(), k2.end(), keys);
  std::copy(v2.begin(), v2.end(), vals);
  return 0;
}

// CUDA_bitonicSortGeneral<uint, uint> (BSG.cu:504-690) called directly, at any
// length. Returns 1 when kernel_bitonicMerge2x would write past 'n'.
ORA_API int ora_bitonic_sort_general(uint32_t* keys, uint32_t* vals, uint32_t n) {
  return bitonic_general(keys, vals, n);
}

// CUDA_bitonicSort<uint> (BS.cu:402-518): the same networks and merges on the
// values alone (BS.cu:35-136 is BSG.cu's chunk kernel without the indices).
ORA_API int ora_bitonic_sort(uint32_t* values, uint32_t n) { return bitonic_general(values, nullptr, n); }

// CUDA_bubbleSort512 (CUDA_bubbleSort512.cu:114-167): n <= 512, indices set to
// 0..n-1 first when init_indices. Returns -1 for n > 512 (the reference asserts).
ORA_API int ora_bubble_sort512(uint32_t* keys, uint32_t* indices, uint32_t n, int init_indices) {
  if (n > 512) return -1;
  if (init_indices) for (uint32_t i = 0; i < n; ++i) indices[i] = i;
  if (n > 1) odd_even_sort(keys, indices, n);
  return 0;
}

// Stable 64-bit key sort (the 63-bit Morton mode; no reference counterpart).
ORA_API void ora_sort_pairs_u64(uint64_t* keys, uint32_t* vals, uint32_t n) {
  std::vector<uint32_t> perm(n);
  for (uint32_t i = 0; i < n; ++i) perm[i] = i;
  std::stable_sort(perm.begin(), perm.end(), [&](uint32_t a, uint32_t b) { return keys[a] < keys[b]; });
  std::vector<uint64_t> k2(n);
  std::vector<uint32_t> v2(n);
  for (uint32_t i = 0; i < n; ++i) { k2[i] = keys[perm[i]]; v2[i] = vals[perm[i]]; }
  std::copy(k2.begin(), k2.end(), keys);
  std::copy(v2.begin(), v2.end(), vals);
}

// ---------------------------------------------------------------------------
// A11-A16: the reference LBVH build (CUDA_buildMeshTreeLBVH, LBVH.cu:457-681).

namespace {

inline Node make_node(uint32_t axis, uint32_t child) {  // LBVH.cu:85-96
  Node n;
  n.w0 = (axis & 3u) << 1;  // type 0; bits 3-31 are left uninitialised by the reference
  n.w1 = child;
  for (int a = 0; a < 3; ++a) { n.bmin[a] = +HUGE_F; n.bmax[a] = -HUGE_F; }
  return n;
}
inline Node make_leaf(uint32_t obj_id, uint32_t count) {  // LBVH.cu:98-109
  Node n;
  n.w0 = 1u | (count << 1);
  n.w1 = obj_id;
  for (int a = 0; a < 3; ++a) { n.bmin[a] = +HUGE_F; n.bmax[a] = -HUGE_F; }
  return n;
}
inline bool is_leaf(const Node& n) { return (n.w0 & 1u) != 0; }
inline uint32_t leaf_count(const Node& n) { return n.w0 >> 1; }
inline uint32_t node_axis(const Node& n) { return (n.w0 >> 1) & 3u; }

struct Topo {
  const uint32_t* A;       // split positions grouped by level, each level ascending
  uint32_t S;              // number of splits
  uint32_t idx[34];        // idx[k] = first split of level k (idx[32] = S)
  uint32_t ptr[32];        // host_level_splits_ptr (LBVH.cu:421)
  uint32_t d;              // tree_depth
  bool complete;           // tree_complete
  std::vector<Node>* nodes;
  uint32_t count;          // node_count (LBVH.cu:596)
  uint32_t real_depth;     // LBVH.cu:598
  uint32_t flags;          // diagnostics, see ora_build_lbvh_ref
};

enum { FLAG_WALK_PAST_END = 1, FLAG_WALK_PAST_DEPTH = 2, FLAG_MERGE2X_OOB = 4, FLAG_NODES_GT_N = 8 };

void set_node(Topo& t, uint32_t id, const Node& n) {
  if (t.nodes->size() <= id) t.nodes->resize(id + 1);
  (*t.nodes)[id] = n;
}

uint32_t createLeaf(Topo& t, uint32_t a, uint32_t b, uint32_t L, uint32_t id);

// createNode(median) LBVH.cu:314-339
uint32_t createNodeMedian(Topo& t, uint32_t a, uint32_t b, uint32_t L, uint32_t id) {
  if (b - a < 224) return createLeaf(t, a, b, L, id);
  uint32_t m = (a + b) >> 1;
  uint32_t left = ++t.count;
  uint32_t right = ++t.count;
  set_node(t, id, make_node(L % 3, left));
  createNodeMedian(t, a, m, L + 1, left);
  createNodeMedian(t, m, b, L + 1, right);
  return id;
}

// createLeaf LBVH.cu:341-358
uint32_t createLeaf(Topo& t, uint32_t a, uint32_t b, uint32_t L, uint32_t id) {
  if (t.complete && (b - a > 224)) return createNodeMedian(t, a, b, L, id);
  if (L > t.real_depth) t.real_depth = L;
  set_node(t, id, make_leaf(a, b - a));
  return id;
}

// createNode(split) LBVH.cu:360-411. The walk reads the concatenated
// split array exactly as the reference's pointer does, including past the
// end of the node's own level (into the lists of deeper levels).
uint32_t createNodeSplit(Topo& t, uint32_t a, uint32_t b, uint32_t L, uint32_t id) {
  if (b - a < 16) return createLeaf(t, a, b, L, id);
  if (L == t.d) return createLeaf(t, a, b, L, id);
  if (t.ptr[L] == t.idx[L + 1]) return createLeaf(t, a, b, L, id);
  uint32_t s;
  while (true) {
    if (t.ptr[L] >= t.S) {  // DEFINED: the reference reads past its host array here
      t.flags |= FLAG_WALK_PAST_END;
      return createLeaf(t, a, b, L, id);
    }
    s = t.A[t.ptr[L]];
    if (s >= b) {
      if (t.ptr[L] >= t.idx[t.d + 1]) t.flags |= FLAG_WALK_PAST_DEPTH;
      return createLeaf(t, a, b, L, id);
    }
    if (s > a) break;
    t.ptr[L]++;
  }
  if (t.ptr[L] >= t.idx[t.d + 1]) t.flags |= FLAG_WALK_PAST_DEPTH;
  uint32_t left = ++t.count;
  uint32_t right = ++t.count;
  set_node(t, id, make_node(L % 3, left));
  createNodeSplit(t, a, s, L + 1, left);
  createNodeSplit(t, s, b, L + 1, right);
  return id;
}

}  // namespace

// Splits and level grouping, LBVH.cu:175-233 (kernel_findSplits) and :565-594.
// DEFINED: splits are appended in ascending position (the reference appends in
// atomic order), so after the stable level grouping every level is ascending;
// the reference re-sorts levels 1..d, so only levels > d differ, and those are
// read only when a walk passes level d (FLAG_WALK_PAST_DEPTH).
// Out: A[] (capacity n), level_count[32]. Returns S.
ORA_API uint32_t ora_find_splits(const uint32_t* codes, uint32_t n, uint32_t* A, uint32_t* level_count) {
  for (int k = 0; k < 32; ++k) level_count[k] = 0;
  std::vector<uint32_t> lev;
  std::vector<uint32_t> pos;
  for (uint32_t i = 0; i + 1 < n; ++i) {
    uint32_t diff = codes[i] ^ codes[i + 1];
    if (diff != 0) {
      uint32_t level = 0;
      while (!(diff & (0x80000000u >> level))) ++level;  // LBVH.cu:190-221
      level_count[level]++;
      pos.push_back(i + 1);
      lev.push_back(level);
    }
  }
  uint32_t S = (uint32_t)pos.size();
  uint32_t start[33];
  start[0] = 0;
  for (int k = 0; k < 32; ++k) start[k + 1] = start[k] + level_count[k];
  for (uint32_t i = 0; i < S; ++i) A[start[lev[i]]++] = pos[i];
  return S;
}

// Full reference LBVH build on sorted-by-Morton data.
// In:  faces (modified in place when tree_complete, A16), verts, n, root box,
//      tree_depth (<= 31), tree_complete.
// Out: boxes[n] (original face order), indices[n] (sorted permutation),
//      nodes (capacity 2n), *n_nodes, *real_depth, *flags, codes_out[n]
//      (sorted Morton codes, may be NULL).
ORA_API int ora_build_lbvh_ref(U4* faces, const F4* verts, uint32_t n, const Box* root, uint32_t tree_depth,
                               int tree_complete, Box* boxes, uint32_t* indices, Node* nodes_out,
                               uint32_t* n_nodes, uint32_t* real_depth_out, uint32_t* flags_out,
                               uint32_t* codes_out) {
  if (n == 0 || tree_depth > 31) return 1;
  ora_calc_boxes(faces, verts, n, boxes, indices);                               // LBVH.cu:474-482
  std::vector<uint32_t> codes(n);
  ora_morton32(boxes, n, &root->mn, &root->mx, codes.data());                    // LBVH.cu:485-513
  uint32_t flags = 0;
  if (ora_sort_pairs_ref(codes.data(), indices, n, 32, 1)) flags |= FLAG_MERGE2X_OOB;  // LBVH.cu:516-529
  std::vector<uint32_t> A(n > 0 ? n : 1);
  uint32_t level_count[32];
  uint32_t S = ora_find_splits(codes.data(), n, A.data(), level_count);          // LBVH.cu:531-563
  // LBVH.cu:579-594: the per-level value sorts of levels 1..d can trigger the
  // Merge2x out-of-bounds write into the next level (BS.cu:196).
  for (uint32_t k = 1; k <= tree_depth; ++k) {
    uint32_t c = level_count[k];
    uint32_t bins = (c + 127) / 128;
    if (c > 128 && c <= 32768 && (bins % 2) == 0 && (c % 128) != 0) flags |= FLAG_MERGE2X_OOB;
  }
  Topo t;
  t.A = A.data(); t.S = S; t.d = tree_depth; t.complete = tree_complete != 0;
  t.idx[0] = 0;
  for (int k = 1; k <= 33; ++k) t.idx[k] = t.idx[k - 1] + (k - 1 < 32 ? level_count[k - 1] : 0);
  for (int k = 0; k < 32; ++k) t.ptr[k] = t.idx[k];
  std::vector<Node> nodes;
  nodes.reserve(64);
  t.nodes = &nodes; t.count = 0; t.real_depth = 0; t.flags = 0;
  createNodeSplit(t, 0, n, 0, 0);                                                // LBVH.cu:444-448
  uint32_t M = t.count + 1;                                                      // LBVH.cu:450
  uint32_t rd = t.real_depth + 1;                                                // LBVH.cu:452
  flags |= t.flags;
  if (M > n) flags |= FLAG_NODES_GT_N;
  // kernel_calcBoundingBoxes LBVH.cu:235-303, launched real_depth times
  // (LBVH.cu:635-649). Emulated in ascending node order within a pass.
  for (uint32_t pass = 0; pass < rd; ++pass) {
    for (uint32_t k = 0; k < M; ++k) {
      Node& nd = nodes[k];
      if (pass == 0 && is_leaf(nd)) {
        Box nb;
        nb.mn = {+HUGE_F, +HUGE_F, +HUGE_F, +HUGE_F};
        nb.mx = {-HUGE_F, -HUGE_F, -HUGE_F, -HUGE_F};
        uint32_t lo = nd.w1, hi = nd.w1 + leaf_count(nd);
        for (uint32_t i = lo; i < hi; ++i) {
          const Box& tb = boxes[indices[i]];
          nb.mn = min4(nb.mn, tb.mn);
          nb.mx = max4(nb.mx, tb.mx);
        }
        nd.bmin[0] = nb.mn.x; nd.bmin[1] = nb.mn.y; nd.bmin[2] = nb.mn.z;
        nd.bmax[0] = nb.mx.x; nd.bmax[1] = nb.mx.y; nd.bmax[2] = nb.mx.z;
      } else if (pass != 0 && !is_leaf(nd)) {
        if (nd.bmax[2] - nd.bmin[2] < 0.0f) {
          const Node& l = nodes[nd.w1];
          if (l.bmax[2] - l.bmin[2] > 0.0f) {
            const Node& r = nodes[nd.w1 + 1];
            if (r.bmax[2] - r.bmin[2] > 0.0f) {
              for (int a = 0; a < 3; ++a) {
                nd.bmin[a] = fmn(l.bmin[a], r.bmin[a]);
                nd.bmax[a] = fmx(l.bmax[a], r.bmax[a]);
              }
            }
          }
        }
      }
    }
  }
  for (uint32_t k = 0; k < M; ++k) nodes_out[k] = nodes[k];
  *n_nodes = M;
  *real_depth_out = rd;
  *flags_out = flags;
  if (codes_out) std::copy(codes.begin(), codes.end(), codes_out);
  if (tree_complete) {  // CUDA_reorderMeshFaces, CUDA_reorderMeshFaces.cu:26-38, LBVH.cu:669-676
    std::vector<U4> old(faces, faces + n);
    for (uint32_t i = 0; i < n; ++i) faces[i] = old[indices[i]];
  }
  return 0;
}

// The recursive form of the refit rule, used to check the pass emulation
// above: an internal node's box is the union of its children's boxes if both
// children are valid (leaf: z-extent > 0; internal: valid children), else
// EMPTY. Writes boxes into nodes (topology untouched). Leaf boxes are read
// from the input.
static bool refit_rec(Node* nodes, uint32_t k) {
  Node& nd = nodes[k];
  if (is_leaf(nd)) return nd.bmax[2] - nd.bmin[2] > 0.0f;
  bool lv = refit_rec(nodes, nd.w1);
  bool rv = refit_rec(nodes, nd.w1 + 1);
  if (lv && rv) {
    const Node& l = nodes[nd.w1];
    const Node& r = nodes[nd.w1 + 1];
    for (int a = 0; a < 3; ++a) { nd.bmin[a] = fmn(l.bmin[a], r.bmin[a]); nd.bmax[a] = fmx(l.bmax[a], r.bmax[a]); }
    return true;
  }
  for (int a = 0; a < 3; ++a) { nd.bmin[a] = +HUGE_F; nd.bmax[a] = -HUGE_F; }
  return false;
}
ORA_API void ora_refit_recursive(Node* nodes) { refit_rec(nodes, 0); }

// ---------------------------------------------------------------------------
// A22/A23: intersection primitives.

namespace {

const float EPS = 0.000001f;  // RT.cu:36

// intersectTriRay RT.cu:57-89 (vert0/edge1/edge2 have w = 0, RT.cu:174-176).
// eps is the reference's EPSILON unless a caller asks for another threshold.
inline bool tri_ray(F4 orig, F4 dir, F4 v0, F4 e1, F4 e2, float& t, float& u, float& v, float eps = EPS) {
  F4 pvec = cross4(dir, e2);
  float det = dot4(e1, pvec);
  if (det > -eps && det < eps) return false;
  float inv_det = 1.0f / det;
  F4 tvec = sub4(orig, v0);
  u = dot4(tvec, pvec) * inv_det;
  if (u < 0.0f || u > 1.0f) return false;
  F4 qvec = cross4(tvec, e1);
  v = dot4(dir, qvec) * inv_det;
  if (v < 0.0f || u + v > 1.0f) return false;
  t = dot4(e2, qvec) * inv_det;
  return t >= 0.0f;
}

// intersectBoxRay RT.cu:91-105 (bmin/bmax are make_float4(float3), w = 0).
inline bool box_ray(F4 bmin, F4 bmax, F4 orig, F4 invdir, float& t) {
  F4 v1 = mul4(sub4(bmin, orig), invdir);
  F4 v2 = mul4(sub4(bmax, orig), invdir);
  F4 vmin = min4(v1, v2);
  F4 vmax = max4(v1, v2);
  float tmin = fmx(vmin.z, fmx(vmin.x, vmin.y));
  float tmax = fmn(vmax.z, fmn(vmax.x, vmax.y));
  t = tmin;
  return (tmax >= tmin) && (tmax >= 0.0f);
}

inline void tri_setup(const U4* faces, const F4* verts, uint32_t f, F4& v0, F4& e1, F4& e2) {
  U4 face = faces[f];
  v0 = verts[face.x];
  e1 = sub4(verts[face.y], v0);
  e2 = sub4(verts[face.z], v0);
  v0.w = 0.0f; e1.w = 0.0f; e2.w = 0.0f;
}

inline F4 invdir_of(F4 dir) {  // RT.cu:128
  return normalize4(F4{1.0f / dir.x, 1.0f / dir.y, 1.0f / dir.z, 0.0f});
}

}  // namespace

ORA_API int ora_tri_ray(const F4* orig, const F4* dir, const F4* v0, const F4* e1, const F4* e2, float* tuv) {
  float t = 0, u = 0, v = 0;
  bool h = tri_ray(*orig, *dir, *v0, *e1, *e2, t, u, v);
  tuv[0] = t; tuv[1] = u; tuv[2] = v;
  return h ? 1 : 0;
}

ORA_API int ora_box_ray(const F4* bmin, const F4* bmax, const F4* orig, const F4* dir, float* t) {
  return box_ray(*bmin, *bmax, *orig, invdir_of(*dir), *t) ? 1 : 0;
}

// kernel_raytraceMeshTree RT.cu:111-219: per-ray DFS, stack of 96.
// In: face_id/coords/depth hold the buffer's prior contents (depth = tmax).
// DEFINED: a miss leaves face_id and coords unchanged.
// Counters (optional, per ray): [0] box tests, [1] triangle tests.
// Returns the number of rays whose stack would overflow 96 entries.
ORA_API uint32_t ora_trace_ref_ray(const Node* nodes, const U4* faces, const F4* verts, const F4* origs,
                                   const F4* dirs, uint32_t n_rays, uint32_t* face_id, float* coords,
                                   float* depths, uint32_t* counters) {
  uint32_t overflow = 0;
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : overflow)
  for (long l = 0; l < (long)n_rays; ++l) {
    F4 orig = origs[l], dir = dirs[l];
    F4 invdir = invdir_of(dir);
    float depth = depths[l];
    uint32_t fid = face_id[l];
    float cu = coords[2 * l], cv = coords[2 * l + 1];
    uint32_t stack[96];
    int sp = 0;
    stack[sp++] = 0;
    uint32_t nbox = 0, ntri = 0;
    bool ovf = false;
    while (sp > 0) {
      uint32_t nid = stack[--sp];
      const Node& nd = nodes[nid];
      if (is_leaf(nd)) {
        uint32_t tri_id = nd.w1, cnt = leaf_count(nd);
        for (uint32_t k = 0; k < cnt; ++k) {
          F4 v0, e1, e2;
          tri_setup(faces, verts, tri_id + k, v0, e1, e2);
          float t, u, v;
          ++ntri;
          if (tri_ray(orig, dir, v0, e1, e2, t, u, v)) {
            if (t < depth) { fid = tri_id + k; cu = u; cv = v; depth = t; }
          }
        }
      } else {
        float t;
        F4 bmin{nd.bmin[0], nd.bmin[1], nd.bmin[2], 0.0f};
        F4 bmax{nd.bmax[0], nd.bmax[1], nd.bmax[2], 0.0f};
        ++nbox;
        if (box_ray(bmin, bmax, orig, invdir, t)) {
          if (t < depth) {
            uint32_t c = nd.w1;
            bool neg;  // CUDANode.h:80-127
            uint32_t ax = node_axis(nd);
            if (ax == 0) neg = dir.x < 0.0f; else if (ax == 1) neg = dir.y < 0.0f; else neg = dir.z < 0.0f;
            uint32_t nearc = neg ? c + 1 : c, farc = neg ? c : c + 1;
            if (sp + 2 > 96) { ovf = true; break; }
            stack[sp++] = farc;
            stack[sp++] = nearc;
          }
        }
      }
    }
    if (ovf) ++overflow;
    face_id[l] = fid; coords[2 * l] = cu; coords[2 * l + 1] = cv; depths[l] = depth;
    if (counters) { counters[2 * l] = nbox; counters[2 * l + 1] = ntri; }
  }
  return overflow;
}

// kernel_raytraceMeshTreePacket RT.cu:222-377: one shared DFS per 16x12 block.
// A node's children are pushed if any ray of the block hits its box with
// t < that ray's depth; near/far order comes from thread 0's ray; every ray
// tests every triangle of every visited leaf.
// DEFINED: threads outside the image do not vote (the reference reads their
// uninitialised registers).
ORA_API uint32_t ora_trace_ref_packet(const Node* nodes, const U4* faces, const F4* verts, const F4* origs,
                                      const F4* dirs, uint32_t width, uint32_t height, uint32_t* face_id,
                                      float* coords, float* depths) {
  const uint32_t BW = 16, BH = 12;
  uint32_t gx = (width + BW - 1) / BW, gy = (height + BH - 1) / BH;
  uint32_t overflow = 0;
#pragma omp parallel for schedule(dynamic, 4) reduction(+ : overflow)
  for (long blk = 0; blk < (long)gx * gy; ++blk) {
    uint32_t bx = (uint32_t)(blk % gx), by = (uint32_t)(blk / gx);
    uint32_t ids[BW * BH];
    int nr = 0;
    F4 orig[BW * BH], dir[BW * BH], inv[BW * BH];
    float depth[BW * BH], cu[BW * BH], cv[BW * BH];
    uint32_t fid[BW * BH];
    for (uint32_t ty = 0; ty < BH; ++ty)
      for (uint32_t tx = 0; tx < BW; ++tx) {
        uint32_t x = bx * BW + tx, y = by * BH + ty;
        if (x < width && y < height) {
          uint32_t l = y * width + x;
          ids[nr] = l; orig[nr] = origs[l]; dir[nr] = dirs[l]; inv[nr] = invdir_of(dir[nr]);
          depth[nr] = depths[l]; fid[nr] = face_id[l]; cu[nr] = coords[2 * l]; cv[nr] = coords[2 * l + 1];
          ++nr;
        }
      }
    // Thread 0 (tx = ty = 0) is always inside the image; it is ids[0].
    uint32_t stack[128];
    int sp = 0;
    stack[sp++] = 0;
    bool ovf = false;
    while (sp > 0) {
      uint32_t nid = stack[--sp];
      const Node& nd = nodes[nid];
      if (is_leaf(nd)) {
        uint32_t tri_id = nd.w1, cnt = leaf_count(nd);
        for (uint32_t k = 0; k < cnt; ++k) {
          F4 v0, e1, e2;
          tri_setup(faces, verts, tri_id + k, v0, e1, e2);
          for (int r = 0; r < nr; ++r) {
            float t, u, v;
            if (tri_ray(orig[r], dir[r], v0, e1, e2, t, u, v)) {
              if (t < depth[r]) { fid[r] = tri_id + k; cu[r] = u; cv[r] = v; depth[r] = t; }
            }
          }
        }
      } else {
        F4 bmin{nd.bmin[0], nd.bmin[1], nd.bmin[2], 0.0f};
        F4 bmax{nd.bmax[0], nd.bmax[1], nd.bmax[2], 0.0f};
        bool hit = false;
        for (int r = 0; r < nr && !hit; ++r) {
          float t;
          if (box_ray(bmin, bmax, orig[r], inv[r], t) && t < depth[r]) hit = true;
        }
        if (hit) {
          uint32_t c = nd.w1;
          uint32_t ax = node_axis(nd);
          bool neg;
          if (ax == 0) neg = dir[0].x < 0.0f; else if (ax == 1) neg = dir[0].y < 0.0f; else neg = dir[0].z < 0.0f;
          uint32_t nearc = neg ? c + 1 : c, farc = neg ? c : c + 1;
          if (sp + 2 > 128) { ovf = true; break; }
          stack[sp++] = farc;
          stack[sp++] = nearc;
        }
      }
    }
    if (ovf) ++overflow;
    for (int r = 0; r < nr; ++r) {
      uint32_t l = ids[r];
      face_id[l] = fid[r]; coords[2 * l] = cu[r]; coords[2 * l + 1] = cv[r]; depths[l] = depth[r];
    }
  }
  return overflow;
}

// Closest hit by brute force over every face, with the reference arithmetic
// (RT.cu:57-89) and the reference acceptance (t < depth). Ties on t go to the
// smallest face index (the order-independent rule of the wide-BVH trace).
// face ids are indices into 'faces' as given.
ORA_API void ora_trace_bruteforce(const U4* faces, const F4* verts, uint32_t n_faces, const F4* origs,
                                  const F4* dirs, uint32_t n_rays, uint32_t* face_id, float* coords,
                                  float* depths, float eps) {
#pragma omp parallel for schedule(dynamic, 16)
  for (long l = 0; l < (long)n_rays; ++l) {
    F4 orig = origs[l], dir = dirs[l];
    float depth = depths[l];
    uint32_t fid = face_id[l];
    float cu = coords[2 * l], cv = coords[2 * l + 1];
    bool any = false;
    for (uint32_t f = 0; f < n_faces; ++f) {
      F4 v0, e1, e2;
      tri_setup(faces, verts, f, v0, e1, e2);
      float t, u, v;
      if (tri_ray(orig, dir, v0, e1, e2, t, u, v, eps)) {
        if (t < depth || (any && t == depth && f < fid)) { fid = f; cu = u; cv = v; depth = t; any = true; }
      }
    }
    face_id[l] = fid; coords[2 * l] = cu; coords[2 * l + 1] = cv; depths[l] = depth;
  }
}

// The same closest hit as ora_trace_bruteforce, found through a tree in the
// reference layout (any of the builds above) with a conservative box test in
// double precision (boxes widened by 1e-6 relative), so it scales to large
// meshes. 'faces' are the tree's (reordered) faces and indices[] maps a
// position to the original face id used for ties and returned in face_id.
ORA_API void ora_trace_closest(const Node* nodes, const U4* faces, const uint32_t* indices, const F4* verts,
                               const F4* origs, const F4* dirs, uint32_t n_rays, uint32_t* face_id, float* coords,
                               float* depths, float eps) {
#pragma omp parallel for schedule(dynamic, 64)
  for (long l = 0; l < (long)n_rays; ++l) {
    F4 orig = origs[l], dir = dirs[l];
    float depth = depths[l];
    uint32_t fid = face_id[l];
    float cu = coords[2 * l], cv = coords[2 * l + 1];
    bool any = false;
    const double o[3] = {orig.x, orig.y, orig.z};
    double inv[3];
    const float dd[3] = {dir.x, dir.y, dir.z};
    for (int a = 0; a < 3; ++a) inv[a] = dd[a] != 0.0f ? 1.0 / (double)dd[a] : 1e300;
    std::vector<uint32_t> stack;
    stack.reserve(128);
    stack.push_back(0);
    while (!stack.empty()) {
      const Node& nd = nodes[stack.back()];
      stack.pop_back();
      if (is_leaf(nd)) {
        for (uint32_t k = 0; k < leaf_count(nd); ++k) {
          uint32_t pos = nd.w1 + k;
          F4 v0, e1, e2;
          tri_setup(faces, verts, pos, v0, e1, e2);
          float t, u, v;
          if (tri_ray(orig, dir, v0, e1, e2, t, u, v, eps)) {
            uint32_t id = indices[pos];
            if (t < depth || (any && t == depth && id < fid)) { fid = id; cu = u; cv = v; depth = t; any = true; }
          }
        }
        continue;
      }
      double t0 = 0.0, t1 = (double)depth * (1.0 + 1e-6) + 1e-30;
      bool empty = nd.bmin[0] > nd.bmax[0] || nd.bmin[1] > nd.bmax[1] || nd.bmin[2] > nd.bmax[2];
      bool hit = true;
      if (!empty) {
        for (int a = 0; a < 3 && hit; ++a) {
          double lo = nd.bmin[a], hi = nd.bmax[a], pad = 1e-6 * (std::fabs(lo) + std::fabs(hi) + (hi - lo)) + 1e-30;
          lo -= pad; hi += pad;
          if (dd[a] == 0.0f) { hit = o[a] >= lo && o[a] <= hi; continue; }
          double ta = (lo - o[a]) * inv[a], tb = (hi - o[a]) * inv[a];
          if (ta > tb) std::swap(ta, tb);
          t0 = std::max(t0, ta);
          t1 = std::min(t1, tb);
          hit = t0 <= t1;
        }
      }
      if (hit) { stack.push_back(nd.w1); stack.push_back(nd.w1 + 1); }
    }
    face_id[l] = fid; coords[2 * l] = cu; coords[2 * l + 1] = cv; depths[l] = depth;
  }
}

// t of a hit recomputed in double precision with the same formula
// (the FP64 reference for the error statement); NaN if not a hit in double.
ORA_API void ora_hit_t_f64(const U4* faces, const F4* verts, const F4* origs, const F4* dirs, const uint32_t* face_id,
                           uint32_t n, double* t_out) {
#pragma omp parallel for schedule(static)
  for (long l = 0; l < (long)n; ++l) {
    U4 f = faces[face_id[l]];
    const F4 &a = verts[f.x], &b = verts[f.y], &c = verts[f.z];
    double o[3] = {origs[l].x, origs[l].y, origs[l].z}, d[3] = {dirs[l].x, dirs[l].y, dirs[l].z};
    double v0[3] = {a.x, a.y, a.z};
    double e1[3] = {(double)b.x - a.x, (double)b.y - a.y, (double)b.z - a.z};
    double e2[3] = {(double)c.x - a.x, (double)c.y - a.y, (double)c.z - a.z};
    double p[3] = {d[1] * e2[2] - d[2] * e2[1], d[2] * e2[0] - d[0] * e2[2], d[0] * e2[1] - d[1] * e2[0]};
    double det = e1[0] * p[0] + e1[1] * p[1] + e1[2] * p[2];
    double s[3] = {o[0] - v0[0], o[1] - v0[1], o[2] - v0[2]};
    double q[3] = {s[1] * e1[2] - s[2] * e1[1], s[2] * e1[0] - s[0] * e1[2], s[0] * e1[1] - s[1] * e1[0]};
    t_out[l] = det != 0.0 ? (e2[0] * q[0] + e2[1] * q[1] + e2[2] * q[2]) / det : NAN;
  }
}

// ---------------------------------------------------------------------------
// Attribute interpolation, kernel_interpolateData CUDA_interpolateData.cu:45-141:
// a = (1 - u) - v; value = (a*x0 + u*x1) + v*x2 componentwise (float4/float2
// operators, types/_cuda_float4.h:80-111, types/_cuda_float2.h:42-76); normals
// normalize(...) or, without vertex normals, normalize(cross(v1 - v0, v2 - v0));
// colours (1,1,1,1) and texcoords (u, v) when the mesh has none; material face.w.
// DEFINED: normalize is v * (1/sqrtf(dot)) (the device code's rsqrtf is an
// approximation); pixels with face id >= face_count are left unchanged (the
// reference reads past its faces).
ORA_API void ora_interpolate(const uint32_t* faces_id, const float* coords, uint32_t count, const U4* faces,
                             uint32_t face_count, const F4* verts, const F4* norms, const F4* cols, const float* texs,
                             F4* out_p, F4* out_n, F4* out_c, float* out_t, uint32_t* out_m) {
  for (uint32_t l = 0; l < count; ++l) {
    const uint32_t fid = faces_id[l];
    if (fid >= face_count) continue;
    const float u = coords[2 * l], v = coords[2 * l + 1];
    const U4 f = faces[fid];
    const F4 v0 = verts[f.x], v1 = verts[f.y], v2 = verts[f.z];
    const float a = (1.0f - u) - v;
    auto bary = [&](F4 x0, F4 x1, F4 x2) { return add4(add4(muls(x0, a), muls(x1, u)), muls(x2, v)); };
    if (out_p) out_p[l] = bary(v0, v1, v2);
    if (out_n) out_n[l] = norms ? normalize4(bary(norms[f.x], norms[f.y], norms[f.z]))
                                : normalize4(cross4(sub4(v1, v0), sub4(v2, v0)));
    if (out_c) out_c[l] = cols ? bary(cols[f.x], cols[f.y], cols[f.z]) : F4{1.0f, 1.0f, 1.0f, 1.0f};
    if (out_t) {
      if (texs) {
        out_t[2 * l] = (texs[2 * f.x] * a + texs[2 * f.y] * u) + texs[2 * f.z] * v;
        out_t[2 * l + 1] = (texs[2 * f.x + 1] * a + texs[2 * f.y + 1] * u) + texs[2 * f.z + 1] * v;
      } else {
        out_t[2 * l] = u;
        out_t[2 * l + 1] = v;
      }
    }
    if (out_m) out_m[l] = f.w;
  }
}

// ---------------------------------------------------------------------------
// A17/A18: binned SAH build, kernel_buildMeshTree SAH.cu:224-645 with the host
// driver SAH.cu:672-789 and the Hybrid entry kernel_initStack SAH.cu:647-666.
// DEFINED: items are processed breadth-first in queue order and each internal
// node takes children (count, count+1) in that order (the reference uses
// atomicAdd order, SAH.cu:625).
// DEFINED: when no cut has a finite cost (m_min == -1) and the node has >= 224
// objects, the reference pushes the same range again forever; here it becomes
// a leaf (FLAG bit 16 reported).

namespace {

struct Bin {  // CUDABin.h:15-113
  F3 bmin, bmax;
  int count, lcount, rcount;
  float larea, rarea;
};

inline float box_area(F4 mn, F4 mx) {  // CUDABox.h:35-44
  F4 v = sub4(mx, mn);
  return 2.f * (v.x * v.y + v.x * v.z + v.y * v.z);
}

struct Item { Box box; uint32_t node_id, lo, hi; };

}  // namespace

ORA_API int ora_build_sah(const Box* boxes, uint32_t* indices, uint32_t n, const Box* root,
                          Node* nodes, uint32_t nodes_cap, uint32_t n_existing, uint32_t* n_nodes_out,
                          uint32_t* flags_out) {
  const int NB = 32;
  uint32_t count;
  std::vector<Item> queue;
  uint32_t flags = 0;
  if (n_existing == 0) {  // SAH.cu:703-710
    queue.push_back({*root, 0, 0, n});
    count = 1;
  } else {  // SAH.cu:711-735 + kernel_initStack SAH.cu:647-666
    count = n_existing;
    for (uint32_t k = 0; k < n_existing; ++k) {
      const Node& nd = nodes[k];
      if (is_leaf(nd)) {
        Box b;
        b.mn = {nd.bmin[0], nd.bmin[1], nd.bmin[2], 0.0f};
        b.mx = {nd.bmax[0], nd.bmax[1], nd.bmax[2], 0.0f};
        queue.push_back({b, k, nd.w1, nd.w1 + leaf_count(nd)});
      }
    }
  }
  std::vector<uint32_t> buf(n);
  size_t head = 0;
  while (head < queue.size()) {
    Item it = queue[head++];
    uint32_t size = it.hi - it.lo;
    auto put = [&](uint32_t id, const Node& nd) { if (id < nodes_cap) nodes[id] = nd; else flags |= 0x20; };
    auto leaf_node = [&](void) {
      Node nd = make_leaf(it.lo, size);
      nd.bmin[0] = it.box.mn.x; nd.bmin[1] = it.box.mn.y; nd.bmin[2] = it.box.mn.z;
      nd.bmax[0] = it.box.mx.x; nd.bmax[1] = it.box.mx.y; nd.bmax[2] = it.box.mx.z;
      put(it.node_id, nd);
    };
    if (size <= 16) { leaf_node(); continue; }  // SAH.cu:239-255
    // Centroid bounds SAH.cu:275-345 (float4 centroids incl. w).
    F4 cmn = {+HUGE_F, +HUGE_F, +HUGE_F, +HUGE_F}, cmx = {-HUGE_F, -HUGE_F, -HUGE_F, -HUGE_F};
    for (uint32_t l = it.lo; l < it.hi; ++l) {
      const Box& b = boxes[indices[l]];
      F4 c = muls(add4(b.mn, b.mx), 0.5f);
      cmn = min4(cmn, c); cmx = max4(cmx, c);
    }
    F4 ext = sub4(cmx, cmn);
    int axis;
    if (ext.x > ext.y) axis = (ext.x > ext.z) ? 0 : 2;
    else axis = (ext.y > ext.z) ? 1 : 2;
    float scale, bmin0;
    if (axis == 0) { scale = 31.0f / ext.x; bmin0 = cmn.x; }
    else if (axis == 1) { scale = 31.0f / ext.y; bmin0 = cmn.y; }
    else { scale = 31.0f / ext.z; bmin0 = cmn.z; }
    auto bin_of = [&](const Box& b) {  // SAH.cu:373-385
      F4 c = muls(add4(b.mn, b.mx), 0.5f);
      float cc = axis == 0 ? c.x : (axis == 1 ? c.y : c.z);
      int m = f2i_sat((cc - bmin0) * scale);
      if (m < 0) m = 0; else if (m >= NB - 1) m = NB - 1;
      return m;
    };
    Bin bins[NB];
    for (int m = 0; m < NB; ++m) {
      bins[m].bmin = {+HUGE_F, +HUGE_F, +HUGE_F}; bins[m].bmax = {-HUGE_F, -HUGE_F, -HUGE_F};
      bins[m].count = 0; bins[m].lcount = 0; bins[m].rcount = 0; bins[m].larea = 0.f; bins[m].rarea = 0.f;
    }
    for (uint32_t l = it.lo; l < it.hi; ++l) {  // insertObject CUDABin.h:63-69
      const Box& b = boxes[indices[l]];
      Bin& bn = bins[bin_of(b)];
      bn.bmin = {fmn(bn.bmin.x, b.mn.x), fmn(bn.bmin.y, b.mn.y), fmn(bn.bmin.z, b.mn.z)};
      bn.bmax = {fmx(bn.bmax.x, b.mx.x), fmx(bn.bmax.y, b.mx.y), fmx(bn.bmax.z, b.mx.z)};
      bn.count++;
    }
    // Sweeps SAH.cu:424-474.
    F4 lmn = {+HUGE_F, +HUGE_F, +HUGE_F, +HUGE_F}, lmx = {-HUGE_F, -HUGE_F, -HUGE_F, -HUGE_F};
    int lc = 0;
    for (int m = 0; m < NB - 1; ++m) {
      lc += bins[m].count;
      lmn = {fmn(lmn.x, bins[m].bmin.x), fmn(lmn.y, bins[m].bmin.y), fmn(lmn.z, bins[m].bmin.z), lmn.w};
      lmx = {fmx(lmx.x, bins[m].bmax.x), fmx(lmx.y, bins[m].bmax.y), fmx(lmx.z, bins[m].bmax.z), lmx.w};
      bins[m].larea += box_area(lmn, lmx);
      bins[m].lcount += lc;
    }
    float T_min = +HUGE_F;
    int m_min = -1;
    F4 rmn = {+HUGE_F, +HUGE_F, +HUGE_F, +HUGE_F}, rmx = {-HUGE_F, -HUGE_F, -HUGE_F, -HUGE_F};
    int rc = 0;
    float pa = box_area(it.box.mn, it.box.mx);
    for (int m = NB - 1; m > 0; --m) {
      rc += bins[m].count;
      rmn = {fmn(rmn.x, bins[m].bmin.x), fmn(rmn.y, bins[m].bmin.y), fmn(rmn.z, bins[m].bmin.z), rmn.w};
      rmx = {fmx(rmx.x, bins[m].bmax.x), fmx(rmx.y, bins[m].bmax.y), fmx(rmx.z, bins[m].bmax.z), rmx.w};
      Bin& cb = bins[m - 1];
      cb.rarea += box_area(rmn, rmx);
      cb.rcount += rc;
      // getSAH CUDABin.h:104-111 with box_cost 1.0, obj_cost 0.125
      float T = 1.0f * 2.f + 0.125f * ((cb.larea / pa) * (float)cb.lcount + (cb.rarea / pa) * (float)cb.rcount);
      if (T < T_min) { T_min = T; m_min = m - 1; }
    }
    if (size < 224 && T_min > (float)(size + 1) * 0.125f) { leaf_node(); continue; }  // SAH.cu:479-496
    if (m_min < 0) { flags |= 0x10; leaf_node(); continue; }
    // Children boxes SAH.cu:501-530.
    F4 Lmn = {+HUGE_F, +HUGE_F, +HUGE_F, +HUGE_F}, Lmx = {-HUGE_F, -HUGE_F, -HUGE_F, -HUGE_F};
    F4 Rmn = Lmn, Rmx = Lmx;
    int lcnt = 0, rcnt = 0;
    for (int m = 0; m <= m_min; ++m) {
      lcnt += bins[m].count;
      Lmn = {fmn(Lmn.x, bins[m].bmin.x), fmn(Lmn.y, bins[m].bmin.y), fmn(Lmn.z, bins[m].bmin.z), Lmn.w};
      Lmx = {fmx(Lmx.x, bins[m].bmax.x), fmx(Lmx.y, bins[m].bmax.y), fmx(Lmx.z, bins[m].bmax.z), Lmx.w};
    }
    for (int m = m_min + 1; m < NB; ++m) {
      rcnt += bins[m].count;
      Rmn = {fmn(Rmn.x, bins[m].bmin.x), fmn(Rmn.y, bins[m].bmin.y), fmn(Rmn.z, bins[m].bmin.z), Rmn.w};
      Rmx = {fmx(Rmx.x, bins[m].bmax.x), fmx(Rmx.y, bins[m].bmax.y), fmx(Rmx.z, bins[m].bmax.z), Rmx.w};
    }
    // Stable partition SAH.cu:534-610.
    uint32_t lo_off = it.lo, ro_off = it.lo + (uint32_t)lcnt, li = 0, ri = 0;
    for (uint32_t l = it.lo; l < it.hi; ++l) {
      uint32_t id = indices[l];
      if (bin_of(boxes[id]) <= m_min) buf[lo_off + li++] = id; else buf[ro_off + ri++] = id;
    }
    for (uint32_t l = it.lo; l < it.hi; ++l) indices[l] = buf[l];
    uint32_t c = count; count += 2;  // SAH.cu:621-639
    Node nd = make_node((uint32_t)axis, c);
    nd.bmin[0] = it.box.mn.x; nd.bmin[1] = it.box.mn.y; nd.bmin[2] = it.box.mn.z;
    nd.bmax[0] = it.box.mx.x; nd.bmax[1] = it.box.mx.y; nd.bmax[2] = it.box.mx.z;
    put(it.node_id, nd);
    Box lb{Lmn, Lmx}, rb{Rmn, Rmx};
    queue.push_back({lb, c, lo_off, lo_off + li});
    queue.push_back({rb, c + 1, ro_off, ro_off + ri});
  }
  *n_nodes_out = count;
  *flags_out = flags;
  return 0;
}

// Face reorder, CUDA_reorderMeshFaces.cu:26-38.
ORA_API void ora_reorder_faces(U4* faces, const uint32_t* indices, uint32_t n) {
  std::vector<U4> old(faces, faces + n);
  for (uint32_t i = 0; i < n; ++i) faces[i] = old[indices[i]];
}

ORA_API int ora_num_threads(void) { return nthreads(); }

// ---- camera and primary rays --------------------------------------------------
// CUDACamera state is (center, angle, dir, fov, ratio); angle.x = pitch (the
// azimuth in the xz plane), angle.y = yaw (the elevation towards +y), angle.z =
// roll (CUDACamera.h:235-247). PI is the reference's 3.14159265359f
// (CUDACommon.h:108). CUDACamera.cpp is host code, so its normalize() is
// v * (1.0f / sqrtf(dot)) (types/_cuda_float.h:21-24, types/_cuda_float3.h:203-208)
// with dot = x*x + y*y + z*z in order.
// DEFINED: sinf/cosf/tanf/asinf/acosf are the host C library's (the reference
// calls its compiler's host math library; no result of the reference pins them).
namespace {
const float CAM_PI = 3.14159265359f;

inline F3 normalize3(F3 v) {
  float inv_len = 1.0f / sqrtf(v.x * v.x + v.y * v.y + v.z * v.z);
  return {v.x * inv_len, v.y * inv_len, v.z * inv_len};
}

// dirFromAngle, CUDACamera.cpp:56-68 (clamps angle in place, then dir).
void cam_dir_from_angle(F3& angle, F3& dir) {
  if (angle.x < 0.0f) angle.x += CAM_PI * 2.0f;
  if (angle.x > CAM_PI * 2.0f) angle.x -= CAM_PI * 2.0f;
  if (angle.y < -CAM_PI * 0.5f + 0.001f) angle.y = -CAM_PI * 0.5f + 0.001f;
  if (angle.y > CAM_PI * 0.5f - 0.001f) angle.y = CAM_PI * 0.5f - 0.001f;
  dir.x = cosf(angle.x) * cosf(angle.y);
  dir.y = sinf(angle.y);
  dir.z = sinf(angle.x) * cosf(angle.y);
  dir = normalize3(dir);
}

// angleFromDir, CUDACamera.cpp:70-88.
void cam_angle_from_dir(F3& angle, F3& dir) {
  dir = normalize3(dir);
  angle.y = asinf(dir.y);
  if (angle.y < -CAM_PI * 0.5f + 0.001f) angle.y = -CAM_PI * 0.5f + 0.001f;
  if (angle.y > CAM_PI * 0.5f - 0.001f) angle.y = CAM_PI * 0.5f - 0.001f;
  if (dir.x > 0) angle.x = asinf(dir.z / cosf(angle.y));
  else if (dir.z > 0) angle.x = acosf(dir.x / cosf(angle.y));
  else angle.x = acosf(-dir.x / cosf(-angle.y)) + CAM_PI;
  if (angle.x < 0.0f) angle.x += CAM_PI * 2.0f;
  if (angle.x > CAM_PI * 2.0f) angle.x -= CAM_PI * 2.0f;
}
}  // namespace

// CUDACamera::lookAt, CUDACamera.cpp:119-135: dir = point - center, then
// angleFromDir. io_angle holds the camera's angle before the call (angle.z,
// the roll, is kept). Writes the new angle and dir.
ORA_API void ora_camera_look_at(const float* center, const float* point, float* io_angle, float* dir_out) {
  F3 angle = {io_angle[0], io_angle[1], io_angle[2]};
  F3 dir = {point[0] - center[0], point[1] - center[1], point[2] - center[2]};
  cam_angle_from_dir(angle, dir);
  io_angle[0] = angle.x; io_angle[1] = angle.y; io_angle[2] = angle.z;
  dir_out[0] = dir.x; dir_out[1] = dir.y; dir_out[2] = dir.z;
}

// CUDACamera::setAngle, CUDACamera.cpp:101-117: angle, then dirFromAngle.
ORA_API void ora_camera_set_angle(float* io_angle, float* dir_out) {
  F3 angle = {io_angle[0], io_angle[1], io_angle[2]};
  F3 dir = {0.0f, 0.0f, 0.0f};
  cam_dir_from_angle(angle, dir);
  io_angle[0] = angle.x; io_angle[1] = angle.y; io_angle[2] = angle.z;
  dir_out[0] = dir.x; dir_out[1] = dir.y; dir_out[2] = dir.z;
}

// CUDA_calcPrimaryRays host part, CUDA_calcPrimaryRays.cu:43-55, on float4
// (w = 0 throughout: make_float4(float3) and cross() set w = 0). The float4
// normalize is v * (1.0f / sqrtf(dot)) (DEFINED, as above); float4 / float is
// a * (1.0f / s) (types/_cuda_float4.h:123-128); scalar * float4 multiplies each
// component by the scalar (:109-112). out = {rayOrig, rayDir, rayX, rayY}.
ORA_API void ora_primary_ray_constants(const float* center, const float* angle, float fov, float ratio, int width,
                                       int height, F4* out) {
  float pitch = angle[0], yaw = angle[1], roll = angle[2];
  float y = tanf(fov * 0.5f);
  float x = y * ratio;
  F4 camDir = {cosf(pitch) * cosf(yaw), sinf(yaw), sinf(pitch) * cosf(yaw), 0.0f};
  F4 camUp = {cosf(pitch) * sinf(roll), cosf(roll), 0.0f, 0.0f};
  F4 right = normalize4(cross4(camDir, camUp));
  F4 up = normalize4(cross4(right, camDir));
  out[0] = {center[0], center[1], center[2], 0.0f};
  out[1] = add4(add4(muls(up, y), muls(right, x)), camDir);
  out[2] = muls(muls(right, 2.0f * x), 1.0f / (float)width);
  out[3] = muls(muls(up, 2.0f * y), 1.0f / (float)height);
}

// kernel_calcPrimaryRays, CUDA_calcPrimaryRays.cu:20-32: pixel (x, y) at
// n = y * width + x gets origin rayOrig and direction
// rayDir - (float)x * rayX - (float)y * rayY, evaluated left to right.
ORA_API void ora_primary_rays(const F4* cst, int width, int height, F4* origins, F4* directions) {
  for (int y = 0; y < height; ++y)
    for (int x = 0; x < width; ++x) {
      size_t n = (size_t)y * width + x;
      origins[n] = cst[0];
      directions[n] = sub4(sub4(cst[1], muls(cst[2], (float)x)), muls(cst[3], (float)y));
    }
}

// ---- shading -----------------------------------------------------------------
// make_pixel, kernels/CUDAShaderCommon.cuh: v * 255, fminf(v, 255), fmaxf(v, 0)
// (device minNum/maxNum), truncate to uint, pack A << 24 | R << 16 | G << 8 | B.
namespace {
inline uint32_t make_pixel(F4 v) {
  v = muls(v, 255.0f);
  v = {fmx(fmn(v.x, 255.0f), 0.0f), fmx(fmn(v.y, 255.0f), 0.0f), fmx(fmn(v.z, 255.0f), 0.0f),
       fmx(fmn(v.w, 255.0f), 0.0f)};
  return (((uint32_t)v.x & 0xFFu) << 16) | (((uint32_t)v.y & 0xFFu) << 8) | ((uint32_t)v.z & 0xFFu) |
         (((uint32_t)v.w & 0xFFu) << 24);
}
}  // namespace

// kernel_shade, CUDAShaderKernel.cu:14-43: pixel k = y * width + x reads
// depths[k]; a hit (depth < CUDART_NORM_HUGE_F) is make_pixel(make_float4(
// -dot(cam_dir, normal))) with cam_dir.w = normal.w = 0; a miss is
// make_pixel(back_color). The result goes to pixels[y * pitch + x].
ORA_API void ora_shade(const float* depths, const F4* normals, int width, int height, const F4* back_color,
                       const F4* cam_dir, uint32_t* pixels, int pitch) {
  F4 cam = *cam_dir;
  cam.w = 0.0f;
  for (int y = 0; y < height; ++y)
    for (int x = 0; x < width; ++x) {
      size_t k = (size_t)y * width + x;
      uint32_t px;
      if (depths[k] < HUGE_F) {
        F4 n = normals[k];
        n.w = 0.0f;
        float c = -dot4(cam, n);
        px = make_pixel(F4{c, c, c, c});
      } else {
        px = make_pixel(*back_color);
      }
      pixels[(size_t)y * pitch + x] = px;
    }
}
