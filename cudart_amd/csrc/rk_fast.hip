// The throughput tree: a Morton LBVH built entirely on the GPU and a
// closest-hit traversal tuned for gfx950. No reference counterpart (the
// reference trees stop at tree_depth with leaves of up to 224 triangles); the
// triangle test is the reference's (CUDA_raytraceMeshTree.cu:57-89), so a
// hit's t, u, v are bit-identical to the reference's for the same triangle.
//
// Build:
//   k_fast_prep     per-face box (API array) and 30/63-bit Morton key
//   radix sort      (key, face) pairs
//   k_fast_hierarchy  topology, node boxes and leaf-ordered triangle records
//                     (v0, v1-v0, v2-v0; 48 B) bottom-up (agglomerative
//                     LBVH, Apetrei 2014) for the splits inside each
//                     512-position chunk; subtrees of <= leaf_max faces
//                     become leaves; up to 2^21 faces it also climbs the
//                     splits that cross chunks
//   k_fast_hierarchy_top  above that, the climbs through the splits that cross
//                     chunks
//
// Node record (64 B): float4 (c0.lo.x, c0.hi.x, c0.lo.y, c0.hi.y),
//                     float4 (c1.lo.x, c1.hi.x, c1.lo.y, c1.hi.y),
//                     float4 (c0.lo.z, c0.hi.z, c1.lo.z, c1.hi.z),
//                     uint4  (ref0, ref1, count0, count1)
// ref: node index, or kLeafBit | first << 4 | (count - 1). The root is
// record 0 (a batch mesh's root: its first position); a node covering sorted
// positions [l, r] is record r when it is its parent's left child and record
// l when it is the right child, so the children of the split between p and
// p + 1 are records p and p + 1. n records for n faces.
#include <mutex>

#include "rk_device.h"

template <typename K>
rk_status rk_radix_sort_impl(hipStream_t st, K* keys, uint32_t* vals, uint32_t n, int first_pass, int passes,
                             void* ws, size_t ws_bytes, uint32_t* err, bool head_cleared = false,
                             const uint32_t** digit_counts = nullptr);

namespace {

using rk::Box;

constexpr uint32_t kLeafBit = 0x80000000u;
constexpr uint32_t kMaxLeaf = 16;

struct FNode { float4 a, b, c; uint4 d; };
static_assert(sizeof(FNode) == 64, "fast node is 64 B");

__device__ __forceinline__ uint32_t spread10(uint32_t v) {
  v &= 0x3FFu;
  v = (v | (v << 16)) & 0x030000FFu;
  v = (v | (v << 8)) & 0x0300F00Fu;
  v = (v | (v << 4)) & 0x030C30C3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}
__device__ __forceinline__ uint64_t spread21(uint64_t v) {
  v &= 0x1FFFFFull;
  v = (v | (v << 32)) & 0x1F00000000FFFFull;
  v = (v | (v << 16)) & 0x1F0000FF0000FFull;
  v = (v | (v << 8)) & 0x100F00F00F00F00Full;
  v = (v | (v << 4)) & 0x10C30C30C30C30C3ull;
  v = (v | (v << 2)) & 0x1249249249249249ull;
  return v;
}

// Also clears the build's info words, the hierarchy's split counters
// (flags, n words) and the sort's error word and digit counts (sort_head,
// head_words), which saves three fill launches per build.
template <typename K>
__global__ __launch_bounds__(256) void k_fast_prep(const uint4* __restrict__ faces, const float4* __restrict__ verts,
                                                   uint32_t n, float4 rmin, float4 rmax, Box* __restrict__ boxes,
                                                   K* __restrict__ keys, uint32_t* __restrict__ vals,
                                                   uint32_t* __restrict__ info, uint32_t* __restrict__ flags,
                                                   uint32_t* __restrict__ sort_head, uint32_t head_words) {
  uint32_t k = blockIdx.x * 256 + threadIdx.x;
  if (k < RK_INFO_WORDS) info[k] = 0u;
  if (k < head_words) sort_head[k] = 0u;
  if (k >= n) return;
  flags[k] = 0u;
  uint4 f = faces[k];
  float4 v0 = verts[f.x], v1 = verts[f.y], v2 = verts[f.z];
  Box b;
  b.mn = rk::min4(rk::min4(v0, v1), v2);
  b.mx = rk::max4(rk::max4(v0, v1), v2);
  if (boxes) {  // the API's per-face boxes, when the caller keeps them
    float4* o = reinterpret_cast<float4*>(boxes + k);
    o[0] = b.mn;
    o[1] = b.mx;
  }
  float4 len = rk::sub4(rmax, rmin);
  float4 c = rk::muls4(rk::add4(b.mn, b.mx), 0.5f);
  if (sizeof(K) == 4) {
    float sx = 1024.0f / len.x, sy = 1024.0f / len.y, sz = 1024.0f / len.z;
    uint32_t qx = min(rk::f2u_sat((c.x - rmin.x) * sx), 1023u);
    uint32_t qy = min(rk::f2u_sat((c.y - rmin.y) * sy), 1023u);
    uint32_t qz = min(rk::f2u_sat((c.z - rmin.z) * sz), 1023u);
    keys[k] = (K)((spread10(qx) << 2) | (spread10(qy) << 1) | spread10(qz));
  } else {
    float sx = 2097152.0f / len.x, sy = 2097152.0f / len.y, sz = 2097152.0f / len.z;
    uint32_t qx = min(rk::f2u_sat((c.x - rmin.x) * sx), 2097151u);
    uint32_t qy = min(rk::f2u_sat((c.y - rmin.y) * sy), 2097151u);
    uint32_t qz = min(rk::f2u_sat((c.z - rmin.z) * sz), 2097151u);
    keys[k] = (K)((spread21(qx) << 2) | (spread21(qy) << 1) | spread21(qz));
  }
  vals[k] = k;
}

// Split key between sorted positions i and i+1 (Apetrei 2014, "Fast and
// simple agglomerative LBVH construction"): the XOR of the two Morton keys,
// with the top bit set so that any difference in the keys outranks the index
// tie-break used for equal keys (the XOR of the two positions). Larger means
// the two positions separate higher in the tree. Two splits that bound the
// same node never compare equal.
template <typename K>
__device__ __forceinline__ uint64_t split_key(const K* __restrict__ keys, uint32_t i) {
  const uint64_t x = (uint64_t)(keys[i] ^ keys[i + 1]);
  return x ? (x | 0x8000000000000000ull) : (uint64_t)(i ^ (i + 1));
}

// Outward padding: a relative 2^-18 of the coordinate magnitude plus the same
// fraction of the extent, so every box contains what the triangle test can
// accept inside it despite rounding on either side.
__device__ __forceinline__ void pad_box(float& lo, float& hi) {
  float m = rk::fmx(fabsf(lo), fabsf(hi));
  float e = hi - lo;
  float p = (m + e) * 3.814697265625e-06f + 1e-30f;
  lo = lo - p;
  hi = hi + p;
}

// A finished subtree as its parent sees it: box (unpadded), reference, and
// the far end of its leaf range (the end the sibling does not know).
struct Part {
  float lox, loy, loz, hix, hiy, hiz;
  uint32_t ref, end;
};
static_assert(sizeof(Part) == 32, "part is 32 B");
// A part exchanged through global memory at a split that crosses chunks, with
// the split key at its outer end (left end for a left child, right end for a
// right child), which its parent's side test needs: carried with the part, so
// the climb does not load keys after each exchange.
struct GPart {
  Part p;
  uint64_t outer;
  uint64_t pad;
};
static_assert(sizeof(GPart) == 48, "global part is 48 B");

__device__ __forceinline__ void write_node(FNode* __restrict__ nodes, uint32_t idx, const Part& c0, const Part& c1,
                                           uint32_t n0, uint32_t n1) {
  float a0 = c0.lox, a1 = c0.hix, a2 = c0.loy, a3 = c0.hiy, a4 = c0.loz, a5 = c0.hiz;
  float b0 = c1.lox, b1 = c1.hix, b2 = c1.loy, b3 = c1.hiy, b4 = c1.loz, b5 = c1.hiz;
  pad_box(a0, a1); pad_box(a2, a3); pad_box(a4, a5);
  pad_box(b0, b1); pad_box(b2, b3); pad_box(b4, b5);
  FNode* p = nodes + idx;
  p->a = make_float4(a0, a1, a2, a3);
  p->b = make_float4(b0, b1, b2, b3);
  p->c = make_float4(a4, a5, b4, b5);
  p->d = make_uint4(c0.ref, c1.ref, n0, n1);
}

// Max-scan over one workgroup of kLocalThreads threads (waves of 64). The
// hierarchy pass's chunks are one position per thread: 512 measured best on
// the 1M-face heightfield and the 10M-face soup (builds of 224 and 1718 us,
// against 236 and 1816 with 1024 and 222 and 1810 with 256; the same trees).
#ifndef RK_HIER_THREADS
#define RK_HIER_THREADS 512
#endif
constexpr uint32_t kLocalThreads = RK_HIER_THREADS;
constexpr uint32_t kLocalWaves = kLocalThreads / 64;

__device__ __forceinline__ uint64_t shfl_up64(uint64_t v, int d) {
  uint32_t lo = __shfl_up((uint32_t)v, d, 64), hi = __shfl_up((uint32_t)(v >> 32), d, 64);
  return ((uint64_t)hi << 32) | lo;
}
__device__ __forceinline__ uint64_t shfl_down64(uint64_t v, int d) {
  uint32_t lo = __shfl_down((uint32_t)v, d, 64), hi = __shfl_down((uint32_t)(v >> 32), d, 64);
  return ((uint64_t)hi << 32) | lo;
}
// before = max of v over threads < t (0 for t = 0), after = max over threads
// > t (0 for the last thread), t = threadIdx.x.
__device__ __forceinline__ void block_excl_max(uint64_t v, uint64_t* s_wave, uint64_t& before, uint64_t& after) {
  const uint32_t t = threadIdx.x, lane = t & 63, w = t >> 6;
  uint64_t up = v, dn = v;
  for (int d = 1; d < 64; d <<= 1) {
    uint64_t u = shfl_up64(up, d), q = shfl_down64(dn, d);
    if (lane >= (uint32_t)d) up = up > u ? up : u;
    if (lane + d < 64) dn = dn > q ? dn : q;
  }
  if (lane == 63) s_wave[w] = up;  // wave max
  __syncthreads();
  uint64_t b = 0, a = 0;
  for (uint32_t k = 0; k < kLocalWaves; ++k) {
    if (k < w) b = b > s_wave[k] ? b : s_wave[k];
    if (k > w) a = a > s_wave[k] ? a : s_wave[k];
  }
  uint64_t ub = shfl_up64(up, 1), da = shfl_down64(dn, 1);
  if (lane > 0) b = b > ub ? b : ub;
  if (lane < 63) a = a > da ? a : da;
  before = b;
  after = a;
}

constexpr uint32_t kChunk = kLocalThreads;  // sorted positions per workgroup, one per thread
constexpr uint32_t kPendLeft = 0x80000000u;  // a handed-on climb whose part is its split's left child

__device__ __forceinline__ Part merge_parts(const Part& lc, const Part& rc) {
  Part up;
  up.lox = rk::fmn(lc.lox, rc.lox);
  up.loy = rk::fmn(lc.loy, rc.loy);
  up.loz = rk::fmn(lc.loz, rc.loz);
  up.hix = rk::fmx(lc.hix, rc.hix);
  up.hiy = rk::fmx(lc.hiy, rc.hiy);
  up.hiz = rk::fmx(lc.hiz, rc.hiz);
  return up;
}

// Parent of the subtree [l, r]: the split at r when split_key(r) <
// split_key(l - 1) (the subtree is its left child), else the split at l - 1.
__device__ __forceinline__ bool left_of_parent(bool has_l, bool has_r, uint64_t kl, uint64_t kr) {
  return !has_l || (has_r && kr < kl);
}

// Completes split p from its two children: writes its record unless it
// collapses into a leaf, and returns the merged part (end not set).
// The tree's positions are [lo, hi]; its root record goes to lo. Any other
// node [nl, nr] goes to nr when it is its parent's left child (next_left)
// and to nl when it is the right child, so the two children of a split p sit
// in records p and p + 1, next to each other (the numbering of Karras 2012).
__device__ __forceinline__ Part finish_split(FNode* __restrict__ nodes, uint32_t lo, uint32_t hi, uint32_t leaf_max,
                                             uint32_t p, uint32_t nl, uint32_t nr, const Part& lc, const Part& rc,
                                             bool next_left) {
  const uint32_t cnt = nr - nl + 1;
  const bool root = nl == lo && nr == hi;
  const uint32_t idx = root ? lo : (next_left ? nr : nl);
  if (root || cnt > leaf_max) write_node(nodes, idx, lc, rc, p - nl + 1, nr - p);
  Part up = merge_parts(lc, rc);
  up.ref = (!root && cnt <= leaf_max) ? (kLeafBit | (nl << 4) | (cnt - 1)) : idx;
  return up;
}

// 16-B loads and stores that are coherent across XCDs, as relaxed agent-scope
// atomics are (the same cache policy bit), but one instruction per 16 B.
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ void store_part_coherent(GPart* p, const Part& v, uint64_t outer) {
  const u32x4 a = {__float_as_uint(v.lox), __float_as_uint(v.loy), __float_as_uint(v.loz), __float_as_uint(v.hix)};
  const u32x4 b = {__float_as_uint(v.hiy), __float_as_uint(v.hiz), v.ref, v.end};
  const u32x4 c = {(uint32_t)outer, (uint32_t)(outer >> 32), 0u, 0u};
  asm volatile(
      "global_store_dwordx4 %0, %1, off sc1\n\tglobal_store_dwordx4 %0, %2, off offset:16 sc1\n\t"
      "global_store_dwordx4 %0, %3, off offset:32 sc1"
      :
      : "v"(p), "v"(a), "v"(b), "v"(c)
      : "memory");
}
__device__ __forceinline__ Part load_part_coherent(const GPart* p, uint64_t& outer) {
  u32x4 a, b, c;
  asm volatile(
      "global_load_dwordx4 %0, %3, off sc1\n\tglobal_load_dwordx4 %1, %3, off offset:16 sc1\n\t"
      "global_load_dwordx4 %2, %3, off offset:32 sc1\n\ts_waitcnt vmcnt(0)"
      : "=&v"(a), "=&v"(b), "=&v"(c)
      : "v"(p)
      : "memory");
  Part v;
  v.lox = __uint_as_float(a.x);
  v.loy = __uint_as_float(a.y);
  v.loz = __uint_as_float(a.z);
  v.hix = __uint_as_float(a.w);
  v.hiy = __uint_as_float(b.x);
  v.hiz = __uint_as_float(b.y);
  v.ref = b.z;
  v.end = b.w;
  outer = (uint64_t)c.x | ((uint64_t)c.y << 32);
  return v;
}

// Climb of the subtree [l, r] (box and ref in me) through splits that cross
// chunks, starting at its parent on side as_left: publish the part in its
// slot with XCD-coherent stores, waited for before the split's counter moves
// (published: already in its slot); the first to arrive stops, the second
// merges both and climbs on. Every split above a crossing split crosses
// chunks too, so the whole climb exchanges through global memory.
// outer: the split key at the outer end of [l, r] (~0 past the array's end).
__device__ __forceinline__ void climb_crossing(uint32_t n, uint32_t leaf_max, uint32_t* __restrict__ gflags,
                                               GPart* __restrict__ gparts, FNode* __restrict__ nodes, uint32_t lo,
                                               uint32_t hi, Part me, uint64_t outer, uint32_t l, uint32_t r,
                                               bool as_left, bool published) {
  for (uint32_t guard = 0; guard < n; ++guard) {
    const uint32_t p = as_left ? r : l - 1;
    const uint32_t mine = as_left ? p : p + 1, other = as_left ? p + 1 : p;
    if (!published) {
      me.end = as_left ? l : r;
      store_part_coherent(gparts + mine, me, outer);
      __asm__ volatile("s_waitcnt vmcnt(0)" ::: "memory");
    }
    published = false;
    const uint32_t old = __hip_atomic_fetch_add(gflags + p, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (old == 0) return;
    uint64_t sib_outer;
    const Part sib = load_part_coherent(gparts + other, sib_outer);
    const uint32_t nl = as_left ? l : sib.end, nr = as_left ? sib.end : r;
    const bool root = nl == lo && nr == hi;  // the (mesh) root
    bool next_left = false;
    // the merged range's end keys: the left child's outer key and the right child's
    const uint64_t kl = as_left ? outer : sib_outer, kr = as_left ? sib_outer : outer;
    if (!root) next_left = left_of_parent(nl > 0, nr + 1 < n, kl, kr);
    outer = next_left ? kl : kr;
    me = as_left ? finish_split(nodes, lo, hi, leaf_max, p, nl, nr, me, sib, next_left)
                 : finish_split(nodes, lo, hi, leaf_max, p, nl, nr, sib, me, next_left);
    if (root) return;
    l = nl;
    r = nr;
    as_left = next_left;
  }
}

// The hierarchy in one bottom-up pass (agglomerative LBVH, Apetrei 2014): a
// subtree covering sorted positions [l, r] is a child of the split at r or
// at l - 1, whichever has the smaller split key; of the two subtrees that
// reach a split, the second to arrive merges both and climbs on. This gives
// the binary radix tree of Karras 2012 without per-node binary searches. Each
// thread starts at its leaf, whose triangle record it also writes.
//
// Exchange at a split: both children publish their part before either knows
// it is second, each in its own slot, the position next to the split (p for
// the left child, p + 1 for the right one); then a counter decides. A slot is
// next written by an ancestor of the split, which completes only after the
// split's reader is done.
//
// A workgroup owns kChunk consecutive positions, one per thread. A split
// whose range lies inside them is reached only by this workgroup, and its
// exchange goes through LDS. The range of split p ends, on each side, at the
// nearest split with a larger key, so it lies inside [c0, c1) when such a
// split exists both in [c0 - 1, p - 1] and in [p + 1, c1 - 1] (prefix and
// suffix maxima of the chunk's split keys). A climb that reaches a split
// crossing chunks goes on through global memory (climb_crossing) when
// inline_top is set; otherwise it leaves its part in the split's slot and its
// position in the chunk's list of handed-on climbs (pend, pend_count), and
// ends, and k_fast_hierarchy_top continues it. The chain of crossing splits
// above a chunk takes a few memory round trips per level; on a large tree,
// waiting for it here would hold the workgroup's LDS while other chunks queue.
// With kBatch the positions hold many meshes, sorted by (mesh, Morton code):
// keys are mesh << 30 | 30-bit code, so every split between two meshes ranks
// above every split inside one, and a climb stops at its mesh's root, whose
// record goes to the mesh's first position. Mesh m's nodes and triangle
// records then occupy [face_offset, face_offset + face_count).
template <typename K, bool kBatch>
__global__ __launch_bounds__(kLocalThreads) void k_fast_hierarchy(const K* __restrict__ keys,
                                                                  const uint32_t* __restrict__ order,
                                                                  const uint4* __restrict__ faces,
                                                                  const float4* __restrict__ verts, uint32_t n,
                                                                  uint32_t leaf_max, uint32_t* __restrict__ gflags,
                                                                  GPart* __restrict__ gparts,
                                                                  FNode* __restrict__ nodes,
                                                                  float4* __restrict__ tris,
                                                                  const rk_batch_mesh* __restrict__ meshes,
                                                                  uint32_t* __restrict__ pend,
                                                                  uint32_t* __restrict__ pend_count,
                                                                  uint32_t inline_top) {
  __shared__ Part s_part[kChunk];
  __shared__ uint32_t s_npend;
  __shared__ uint64_t s_key[kChunk + 1];  // s_key[k] = split_key(c0 - 1 + k), ~0 past either end
  __shared__ uint32_t s_flag[kChunk];
  __shared__ uint8_t s_local[kChunk];     // s_local[k]: split c0 + k lies inside the chunk
  __shared__ uint64_t s_wave[kLocalWaves];
  const uint32_t t = threadIdx.x;
  const uint32_t c0 = blockIdx.x * kChunk;
  const uint32_t c1 = min(c0 + kChunk, n);
  const uint32_t i = c0 + t;
  // leaf loads first: their latency overlaps the split-key scan
  uint32_t fid = 0;
  uint32_t lo = 0, hi = n - 1, voff = 0;  // this leaf's mesh: positions [lo, hi]
  float4 v0 = make_float4(0.0f, 0.0f, 0.0f, 0.0f), v1 = v0, v2 = v0;
  if (i < c1) {
    fid = order[i];
    if (kBatch) {
      const rk_batch_mesh m = meshes[(uint32_t)((uint64_t)keys[i] >> 30)];
      lo = m.face_offset;
      hi = m.face_offset + m.face_count - 1;
      voff = m.vertex_offset;
    }
    const uint4 f = faces[fid];
    v0 = verts[f.x + voff];
    v1 = verts[f.y + voff];
    v2 = verts[f.z + voff];
  }
  for (uint32_t k = t; k <= kChunk; k += kLocalThreads) {
    const int64_t x = (int64_t)c0 - 1 + (int64_t)k;
    s_key[k] = (x >= 0 && x + 1 < (int64_t)n) ? split_key(keys, (uint32_t)x) : ~0ull;
  }
  s_flag[t] = 0u;
  if (t == 0) s_npend = 0u;
  __syncthreads();
  {
    // inside test of split p = c0 - 1 + k for k = t (k >= 1)
    const uint64_t d = s_key[t];
    uint64_t before, after;
    block_excl_max(d, s_wave, before, after);
    const uint64_t last = s_key[kChunk];
    after = after > last ? after : last;
    const bool inside = t >= 1 && c0 + t < c1 && before > d && after > d;
    if (t >= 1) s_local[t - 1] = inside ? 1 : 0;
    if (t == kLocalThreads - 1) s_local[kChunk - 1] = 0;  // split c0 + kChunk - 1 always crosses
  }
  __syncthreads();
  // Climbs from this thread's leaf through the chunk's splits; stops as the
  // first to arrive at a split, at the (mesh) root, or at a split that
  // crosses chunks, whose climb k_fast_hierarchy_top continues.
  auto climb = [&]() {
    if (i >= c1) return;
    const float4 e1 = rk::sub4(v1, v0), e2 = rk::sub4(v2, v0);
    tris[3 * (size_t)i + 0] = make_float4(v0.x, v0.y, v0.z, __uint_as_float(fid));
    tris[3 * (size_t)i + 1] = make_float4(e1.x, e1.y, e1.z, 0.0f);
    tris[3 * (size_t)i + 2] = make_float4(e2.x, e2.y, e2.z, 0.0f);
    Part me;
    me.lox = rk::fmn(rk::fmn(v0.x, v1.x), v2.x);
    me.loy = rk::fmn(rk::fmn(v0.y, v1.y), v2.y);
    me.loz = rk::fmn(rk::fmn(v0.z, v1.z), v2.z);
    me.hix = rk::fmx(rk::fmx(v0.x, v1.x), v2.x);
    me.hiy = rk::fmx(rk::fmx(v0.y, v1.y), v2.y);
    me.hiz = rk::fmx(rk::fmx(v0.z, v1.z), v2.z);
    me.ref = kLeafBit | (i << 4);  // count - 1 = 0
    if (lo == hi) {  // a single face: root with one leaf child and one empty child
      Part empty;
      empty.lox = empty.loy = empty.loz = RK_HUGE;
      empty.hix = empty.hiy = empty.hiz = -RK_HUGE;
      empty.ref = kLeafBit;
      write_node(nodes, lo, me, empty, 1, 0);
      return;
    }
    // the side of [l, r] under its parent, from the split keys at its ends
    // (and the split key at the outer end of [l, r] in outer)
    auto side = [&](uint32_t l, uint32_t r, uint64_t& outer) {
      const bool has_l = l > 0, has_r = r + 1 < n;
      uint64_t kl = ~0ull, kr = ~0ull;
      if (has_l) kl = (l >= c0 && l - c0 <= kChunk) ? s_key[l - c0] : split_key(keys, l - 1);
      if (has_r) kr = (r + 1 >= c0 && r + 1 - c0 <= kChunk) ? s_key[r + 1 - c0] : split_key(keys, r);
      const bool left = left_of_parent(has_l, has_r, kl, kr);
      outer = left ? kl : kr;
      return left;
    };
    uint32_t l = i, r = i;
    uint64_t outer = ~0ull;
    bool as_left = side(l, r, outer);
    // Each step widens [l, r] by at least one position: n steps bound it.
    for (uint32_t guard = 0; guard < n; ++guard) {
      const uint32_t p = as_left ? r : l - 1;
      const uint32_t mine = as_left ? p : p + 1, other = as_left ? p + 1 : p;
      me.end = as_left ? l : r;
      if (!(p >= c0 && p - c0 < kChunk && s_local[p - c0])) {
        if (inline_top) {  // small trees: the chain above the chunks is short
          climb_crossing(n, leaf_max, gflags, gparts, nodes, lo, hi, me, outer, l, r, as_left, false);
          return;
        }
        // crossing split: publish the part in its slot and hand the climb on
        gparts[mine].p = me;
        gparts[mine].outer = outer;
        const uint32_t k = atomicAdd(&s_npend, 1u);
        pend[c0 + k] = mine | (as_left ? kPendLeft : 0u);
        return;
      }
      s_part[mine - c0] = me;  // a local split has p, p + 1 in [c0, c1)
      const uint32_t old =
          __hip_atomic_fetch_add(s_flag + (p - c0), 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP);
      if (old == 0) return;
      const Part sib = s_part[other - c0];
      const uint32_t nl = as_left ? l : sib.end, nr = as_left ? sib.end : r;
      const bool root = nl == lo && nr == hi;  // the (mesh) root
      const bool next_left = root ? false : side(nl, nr, outer);
      me = as_left ? finish_split(nodes, lo, hi, leaf_max, p, nl, nr, me, sib, next_left)
                   : finish_split(nodes, lo, hi, leaf_max, p, nl, nr, sib, me, next_left);
      if (root) return;
      l = nl;
      r = nr;
      as_left = next_left;
    }
  };
  climb();
  __syncthreads();
  if (t == 0) pend_count[blockIdx.x] = s_npend;
}

// The splits that cross chunks, above the chunks' own: thread k of workgroup
// c continues the k-th climb that chunk c handed on, whose part is in its
// slot already (written by k_fast_hierarchy).
template <typename K, bool kBatch>
__global__ __launch_bounds__(64) void k_fast_hierarchy_top(const K* __restrict__ keys, uint32_t n, uint32_t leaf_max,
                                                           uint32_t* __restrict__ gflags, GPart* __restrict__ gparts,
                                                           FNode* __restrict__ nodes,
                                                           const rk_batch_mesh* __restrict__ meshes,
                                                           const uint32_t* __restrict__ pend,
                                                           const uint32_t* __restrict__ pend_count) {
  const uint32_t c0 = blockIdx.x * kChunk;
  const uint32_t count = pend_count[blockIdx.x];
  for (uint32_t k = threadIdx.x; k < count; k += 64) {
    const uint32_t e = pend[c0 + k];
    const uint32_t mine = e & ~kPendLeft;
    const bool as_left = (e & kPendLeft) != 0;
    uint64_t outer;
    const Part me = load_part_coherent(gparts + mine, outer);
    const uint32_t l = as_left ? me.end : mine, r = as_left ? mine : me.end;
    uint32_t lo = 0, hi = n - 1;
    if (kBatch) {
      const rk_batch_mesh m = meshes[(uint32_t)((uint64_t)keys[mine] >> 30)];
      lo = m.face_offset;
      hi = m.face_offset + m.face_count - 1;
    }
    climb_crossing(n, leaf_max, gflags, gparts, nodes, lo, hi, me, outer, l, r, as_left, true);
  }
}

// ---- trace -------------------------------------------------------------------

// The traversal stack holds at most one entry per tree level, and a radix
// tree over b-bit keys with position tie-breaks (< 2^27 positions) is at most
// b + 27 levels deep: 57 for 30-bit keys, 90 for 63-bit keys.
#ifndef RK_TRACE_STACK
#define RK_TRACE_STACK 96
#endif
constexpr int kStack = RK_TRACE_STACK;
#ifndef RK_LDS_STACK
#define RK_LDS_STACK 8
#endif
// Stack entries per lane kept in LDS: 8 x 8 B x 64 lanes x 32 waves = 128 KB,
// so the traces still run 32 waves per CU.
constexpr int kLdsStack = RK_LDS_STACK;

// intersectTriRay, CUDA_raytraceMeshTree.cu:57-89, with vert0.w = edge1.w =
// edge2.w = 0 (:174-176) and the reference's four-term dot products. eps is
// the reference's EPSILON (1e-6, :36) unless the caller chose another.
__device__ __forceinline__ bool tri_ray(float4 orig, float4 dir, float4 v0, float4 e1, float4 e2, float eps, float& t,
                                        float& u, float& v) {
  float4 pvec = rk::cross4(dir, e2);
  float det = rk::dot4(e1, pvec);
  if (det > -eps && det < eps) return false;
  float inv_det = 1.0f / det;
  float4 tvec = rk::sub4(orig, v0);
  u = rk::dot4(tvec, pvec) * inv_det;
  if (u < 0.0f || u > 1.0f) return false;
  float4 qvec = rk::cross4(tvec, e1);
  v = rk::dot4(dir, qvec) * inv_det;
  if (v < 0.0f || u + v > 1.0f) return false;
  t = rk::dot4(e2, qvec) * inv_det;
  return t >= 0.0f;
}

__device__ __forceinline__ float safe_inv(float d) {
  float a = fabsf(d) > 1e-20f ? d : (d < 0.0f ? -1e-20f : 1e-20f);
  return 1.0f / a;
}

// The slab test of a node record's two child boxes against a ray given by its
// inverse direction (ix, iy, iz) and origin terms (ox = o.x * ix, ...), over
// [0, bound]: entry distances and whether each box is hit.
struct Slab2 {
  float tmin0, tmin1;
  bool h0, h1;
};
__device__ __forceinline__ Slab2 slab2(const FNode& nd, float ix, float iy, float iz, float ox, float oy, float oz,
                                       float bound) {
  const float a0 = __builtin_fmaf(nd.a.x, ix, -ox), a1 = __builtin_fmaf(nd.a.y, ix, -ox);
  const float a2 = __builtin_fmaf(nd.a.z, iy, -oy), a3 = __builtin_fmaf(nd.a.w, iy, -oy);
  const float a4 = __builtin_fmaf(nd.c.x, iz, -oz), a5 = __builtin_fmaf(nd.c.y, iz, -oz);
  const float b0 = __builtin_fmaf(nd.b.x, ix, -ox), b1 = __builtin_fmaf(nd.b.y, ix, -ox);
  const float b2 = __builtin_fmaf(nd.b.z, iy, -oy), b3 = __builtin_fmaf(nd.b.w, iy, -oy);
  const float b4 = __builtin_fmaf(nd.c.z, iz, -oz), b5 = __builtin_fmaf(nd.c.w, iz, -oz);
  const float tmin0 = fmaxf(fmaxf(fminf(a0, a1), fminf(a2, a3)), fmaxf(fminf(a4, a5), 0.0f));
  const float tmax0 = fminf(fminf(fmaxf(a0, a1), fmaxf(a2, a3)), fminf(fmaxf(a4, a5), bound));
  const float tmin1 = fmaxf(fmaxf(fminf(b0, b1), fminf(b2, b3)), fmaxf(fminf(b4, b5), 0.0f));
  const float tmax1 = fminf(fminf(fmaxf(b0, b1), fmaxf(b2, b3)), fminf(fmaxf(b4, b5), bound));
  return Slab2{tmin0, tmin1, tmin0 <= tmax0, tmin1 <= tmax1};
}

// A traversal's result: the sorted position of the closest hit (~0: none),
// its t (the tmax it started with when there is no hit), u and v.
struct RayHit {
  uint32_t pos;
  float t, u, v;
};

// Closest hit of each lane's ray in the tree rooted at 'root' (positions and
// node records share one index space of n_tris). Depth-first, nearer child
// first; a pending child is skipped once the hit found is nearer than its
// entry point. Ties on t go to the smallest original face id, so the result
// does not depend on the tree or on the order of the visits.
//
// The wave shares the work. Every lane starts on its own ray (has_ray); a
// lane with nothing left to do takes the newest pending subtree of a lane
// that still has some (one per donor and step) and traverses it for that
// lane's ray, culling against the owner's best hit so far. A helper's best
// hit is merged into its owner's when the subtree is done, by (t, original
// id). A frame then waits less on its few long rays (grazing rays near a
// horizon), which would otherwise keep a wave running with one active lane.
// Every lane of the wave must call it; xchg is this wave's kXchgWords words of
// LDS.
//
// The stack: its first kLds entries live in LDS (lds: this workgroup's
// kLds x kWG entries, entry e of thread x at e * kWG + x), the rest in
// per-lane scratch. Pending entries rarely go deeper than a dozen or two, and
// an LDS access takes a fraction of the time of a scratch access that misses
// the L1 (the scratch of thousands of resident lanes does not fit in it).
// LDS words per wave for the exchange of pending subtrees: reference, entry
// distance, owner lane and donor lane of each of up to 64 donations.
constexpr int kXchgWords = 4 * 64;

// kOldest: a donor hands out its oldest pending entry (the bottom of its
// stack: the subtree nearest the root, the largest) instead of its newest, as
// a work-stealing deque does, so a helper gets more work and soon has pending
// entries of its own to hand on. The stack is then the range [sb, sp); it is
// reset to empty at [0, 0) whenever it drains, and an index still stands for
// one level of the current path, so sp stays within the depth bound.
//
// kLeafBatch > 0 postpones leaves: a lane that reaches a leaf keeps it as its
// one pending leaf and goes on to its next pending subtree, and the wave
// tests the pending leaves together once kLeafBatch lanes hold one (or a lane
// cannot go on without its leaf: it reached a second leaf, or its subtree is
// done). With one face per leaf some lane of a busy wave is at a leaf on
// almost every step, so testing leaves at once issues the triangle test on
// almost every step. A postponed leaf leaves the culling bound looser for a
// few steps, so a lane may visit a node it would have culled; the closest hit
// is the same. kDonateMin: the exchange runs only on steps where at least this
// many lanes are idle.
// The faces of leaf 'ref' against the job's ray (jo, jd). own: the ray is
// this lane's, and a hit goes to its best (o_*); otherwise to this helper's
// best (j_*), kept only if it can beat the owner's best as of this step (dmax,
// d_hit). Ties on t go to the smaller original id. A macro, not a function: the
// best-hit variables stay in registers only if no reference to them is taken.
#define RK_LEAF_HITS(ref)                                                                              \
  do {                                                                                                 \
    const uint32_t first_ = ((ref) & ~kLeafBit) >> 4;                                                 \
    const uint32_t cnt_ = first_ < n_tris ? min(((ref) & 15u) + 1, n_tris - first_) : 0u;             \
    for (uint32_t k = 0; k < cnt_; ++k) {                                                              \
      const float4 r0 = tris[3 * (size_t)(first_ + k) + 0];                                            \
      float4 v0 = r0;                                                                                  \
      v0.w = 0.0f;                                                                                     \
      const float4 e1 = tris[3 * (size_t)(first_ + k) + 1], e2 = tris[3 * (size_t)(first_ + k) + 2];   \
      float t, u, v;                                                                                   \
      if (kCount) ++n_tri;                                                                             \
      if (tri_ray(jo, jd, v0, e1, e2, eps, t, u, v)) {                                                 \
        const uint32_t oid = __float_as_uint(r0.w);                                                    \
        if (own) {                                                                                     \
          if (t < o_t || (o_pos != 0xFFFFFFFFu && t == o_t && oid < o_oid)) {                          \
            o_t = t; o_pos = first_ + k; o_oid = oid; o_u = u; o_v = v;                                \
          }                                                                                            \
        } else if ((t < dmax || (d_hit && t == dmax)) &&                                               \
                   (j_pos == 0xFFFFFFFFu || t < j_t || (t == j_t && oid < j_oid))) {                   \
          j_t = t; j_pos = first_ + k; j_oid = oid; j_u = u; j_v = v;                                  \
        }                                                                                              \
      }                                                                                                \
    }                                                                                                  \
  } while (0)

//
// kBoundEvery: helpers read their owner's best so far on every kBoundEvery-th
// step only (1: every step); in between they cull against the value they last
// read, which can only be larger, so they may visit more nodes but find the
// same closest hit.
template <bool kCount, int kStackN, int kLds, int kWG, bool kOldest = false, int kLeafBatch = 0,
          int kDonateMin = 1, int kBoundEvery = 1>
__device__ __forceinline__ RayHit trace_ray_shared(const FNode* __restrict__ nodes, const float4* __restrict__ tris,
                                                   uint32_t n_tris, uint32_t root, bool has_ray, float4 orig,
                                                   float4 dir, float depth0, float eps, uint32_t& n_box,
                                                   uint32_t& n_tri, uint2* lds, uint32_t* xchg) {
  const uint32_t lane = __lane_id();
  const uint32_t max_visits = 4 * n_tris + 64;  // a subtree visit touches each node once
  // this lane's own ray: its best hit
  uint32_t o_pos = 0xFFFFFFFFu, o_oid = 0xFFFFFFFFu;
  float o_t = depth0, o_u = 0.0f, o_v = 0.0f;
  // the job: a subtree (cur + stack) of the owner's ray, and a helper's own best
  int owner = has_ray ? (int)lane : -1;
  float4 jo = orig, jd = dir;
  float ix = safe_inv(jd.x), iy = safe_inv(jd.y), iz = safe_inv(jd.z);
  float ox = jo.x * ix, oy = jo.y * iy, oz = jo.z * iz;
  uint32_t j_pos = 0xFFFFFFFFu, j_oid = 0xFFFFFFFFu;
  float j_t = 0.0f, j_u = 0.0f, j_v = 0.0f;
  uint32_t cur = root, visits = 0;
  uint32_t pend = 0u;  // a postponed leaf (a leaf reference always has kLeafBit set; 0: none)
  uint32_t step = 0;   // loop steps so far (the same in every lane)
  float h_t = depth0;  // kBoundEvery > 1: the owner's best as this helper last read it
  bool h_hit = false;
  static_assert(kStackN > kLds, "scratch part of the stack");
  uint32_t stack_ref[kStackN - kLds];
  float stack_t[kStackN - kLds];
  // An LDS-typed pointer (an entry is ref | t << 32, the layout of a uint2):
  // with a generic one the compiler merges the two branches of st_get into one
  // flat load through a selected address, which costs a flat access (and a
  // wait on every memory counter) per pop.
  typedef __attribute__((address_space(3))) unsigned long long lds_u64;
  lds_u64* my_lds = (lds_u64*)(lds + threadIdx.x);
  int sp = 0, sb = 0;
  auto st_get = [&](int i, uint32_t& r, float& t) {
    if (i < kLds) {
      const unsigned long long e = my_lds[i * kWG];
      r = (uint32_t)e;
      t = __uint_as_float((uint32_t)(e >> 32));
    } else {
      r = stack_ref[i - kLds];
      t = stack_t[i - kLds];
    }
  };
  auto st_set = [&](int i, uint32_t r, float t) {
    if (i < kLds) {
      my_lds[i * kWG] = (unsigned long long)r | ((unsigned long long)__float_as_uint(t) << 32);
    } else {
      stack_ref[i - kLds] = r;
      stack_t[i - kLds] = t;
    }
  };
  for (;;) {
    const bool working = owner >= 0;
    const uint64_t idle = __ballot(!working);
    if (idle == ~0ull) break;
    bool fresh = false;  // this lane took a new job this step
    // hand the newest pending subtree of up to popc(idle) lanes to idle lanes
    const uint64_t donors = __ballot(working && sp > sb);
    if (idle != 0ull && donors != 0ull && (kDonateMin <= 1 || __popcll(idle) >= kDonateMin)) {
      const uint32_t nd = min((uint32_t)__popcll(idle), (uint32_t)__popcll(donors));
      const uint64_t lt = rk::lanemask_lt();
      const uint32_t rd = (uint32_t)__popcll(donors & lt), ri = (uint32_t)__popcll(idle & lt);
      if (working && sp > sb && rd < nd) {
        // the newest entry: taking it is a pop, so the stack stays one
        // entry per level of the current path at most (kOldest: the oldest)
        uint32_t r0;
        float t0;
        if (kOldest) {
          st_get(sb, r0, t0);
          if (++sb == sp) sb = sp = 0;
        } else {
          --sp;
          st_get(sp, r0, t0);
        }
        __hip_atomic_store(xchg + rd, r0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
        __hip_atomic_store(xchg + 64 + rd, __float_as_uint(t0), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
        __hip_atomic_store(xchg + 128 + rd, (uint32_t)owner, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
        __hip_atomic_store(xchg + 192 + rd, lane, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
      }
      __builtin_amdgcn_wave_barrier();
      const bool take = !working && ri < nd;
      uint32_t g_ref = 0, g_t = 0, g_own = lane, g_don = lane;
      if (take) {
        g_ref = __hip_atomic_load(xchg + ri, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
        g_t = __hip_atomic_load(xchg + 64 + ri, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
        g_own = __hip_atomic_load(xchg + 128 + ri, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
        g_don = __hip_atomic_load(xchg + 192 + ri, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
      }
      __builtin_amdgcn_wave_barrier();
      // Every lane takes part in the exchange. The donor's job is a subtree of
      // the owner's ray, so the donor holds that ray and its slab terms (the
      // same bits the owner computed): copying them saves three divisions.
      const int src = (int)g_own, don = (int)g_don;
      const float4 no = make_float4(__shfl(jo.x, don, 64), __shfl(jo.y, don, 64), __shfl(jo.z, don, 64),
                                    __shfl(jo.w, don, 64));
      const float4 ndr = make_float4(__shfl(jd.x, don, 64), __shfl(jd.y, don, 64), __shfl(jd.z, don, 64),
                                     __shfl(jd.w, don, 64));
      const float nix = __shfl(ix, don, 64), niy = __shfl(iy, don, 64), niz = __shfl(iz, don, 64);
      const float nox = __shfl(ox, don, 64), noy = __shfl(oy, don, 64), noz = __shfl(oz, don, 64);
      const float od = __shfl(o_t, src, 64);
      if (take && __uint_as_float(g_t) <= od) {
        fresh = true;
        owner = src;
        jo = no;
        jd = ndr;
        ix = nix;
        iy = niy;
        iz = niz;
        ox = nox;
        oy = noy;
        oz = noz;
        cur = g_ref;
        sp = 0;
        sb = 0;
        visits = 0;
        j_pos = 0xFFFFFFFFu;
        j_oid = 0xFFFFFFFFu;
      }
    }
    const bool own = owner == (int)lane;
    // the owner's best so far bounds a helper's subtree
    float dmax = o_t;
    bool d_hit = o_pos != 0xFFFFFFFFu;
    // a lane that just took a job reads its new owner's bound at once
    if (kBoundEvery <= 1 || (step % (uint32_t)kBoundEvery) == 0u || __ballot(fresh) != 0ull) {
      if (__ballot(owner >= 0 && !own) != 0ull) {
        const int src = owner >= 0 ? owner : (int)lane;
        const float od = __shfl(o_t, src, 64);
        const uint32_t op = __shfl(o_pos, src, 64);
        if (!own) {
          dmax = od;
          d_hit = op != 0xFFFFFFFFu;
        }
      }
      if (kBoundEvery > 1) {
        h_t = dmax;
        h_hit = d_hit;
      }
    } else if (!own) {
      dmax = h_t;
      d_hit = h_hit;
    }
    ++step;
    bool job_end = false, flush_me = false, end_after = false;
    if (owner >= 0) {
      float bound = own ? o_t : (j_pos != 0xFFFFFFFFu ? fminf(dmax, j_t) : dmax);
      bool need_pop = false;
      if (++visits > max_visits) {
        job_end = true;
        pend = 0u;
      } else if (!(cur & kLeafBit)) {
        if (cur >= n_tris) {
          job_end = true;  // never taken for a tree built here
        } else {
          const FNode nd = nodes[cur];
          const Slab2 sl = slab2(nd, ix, iy, iz, ox, oy, oz, bound);
          const float tmin0 = sl.tmin0, tmin1 = sl.tmin1;
          const bool h0 = sl.h0, h1 = sl.h1;
          if (kCount) n_box += 2;
          if (h0 && h1) {
            const bool first0 = tmin0 <= tmin1;
            if (sp < kStackN) {
              st_set(sp, first0 ? nd.d.y : nd.d.x, first0 ? tmin1 : tmin0);
              ++sp;
            }
            cur = first0 ? nd.d.x : nd.d.y;
          } else if (h0) {
            cur = nd.d.x;
          } else if (h1) {
            cur = nd.d.y;
          } else {
            need_pop = true;
          }
        }
      } else if (kLeafBatch == 0) {
        RK_LEAF_HITS(cur);
        need_pop = true;
      } else if (pend == 0u) {
        pend = cur;  // postponed: go on to the next pending subtree
        need_pop = true;
      } else {
        flush_me = true;  // a second leaf: test the pending one first, keep cur
      }
      if (need_pop) {
        // the nearest pending entry that can still hold a closer hit
        bound = own ? o_t : (j_pos != 0xFFFFFFFFu ? fminf(dmax, j_t) : dmax);
        bool found = false;
        while (sp > sb) {
          --sp;
          uint32_t r;
          float t;
          st_get(sp, r, t);
          if (t <= bound) {
            cur = r;
            found = true;
            break;
          }
        }
        if (!found) {
          if (kLeafBatch > 0 && pend != 0u) {
            flush_me = true;  // the subtree is done but for its pending leaf
            end_after = true;
          } else {
            job_end = true;
          }
        }
      }
    }
    if (kLeafBatch > 0) {
      const uint64_t pm = __ballot(pend != 0u);
      if (pm != 0ull && (__ballot(flush_me) != 0ull || __popcll(pm) >= kLeafBatch)) {
        if (pend != 0u) {
          RK_LEAF_HITS(pend);
          pend = 0u;
        }
        if (end_after) job_end = true;
      }
    }
    // merge finished helpers' hits into their owners, one at a time
    uint64_t m = __ballot(job_end && !own && j_pos != 0xFFFFFFFFu);
    while (m != 0ull) {
      const int e = __builtin_ctzll(m);
      m &= m - 1;
      const int ow = __shfl(owner, e, 64);
      const float ct = __shfl(j_t, e, 64);
      const uint32_t coid = __shfl(j_oid, e, 64), cpos = __shfl(j_pos, e, 64);
      const float cu = __shfl(j_u, e, 64), cv = __shfl(j_v, e, 64);
      if ((int)lane == ow && (ct < o_t || (o_pos != 0xFFFFFFFFu && ct == o_t && coid < o_oid))) {
        o_t = ct; o_pos = cpos; o_oid = coid; o_u = cu; o_v = cv;
      }
    }
    if (job_end) owner = -1;
  }
  return RayHit{o_pos, o_t, o_u, o_v};
}
#undef RK_LEAF_HITS

// Sort key of a ray: the 30-bit Morton code of its origin in the tree's root
// box (the union of the root record's two child boxes), clamped to the box.
__global__ __launch_bounds__(256) void k_ray_keys(const FNode* __restrict__ nodes, const float4* __restrict__ origins,
                                                  uint32_t n, uint32_t* __restrict__ keys,
                                                  uint32_t* __restrict__ vals) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const FNode r = nodes[0];
  const float lx = rk::fmn(r.a.x, r.b.x), hx = rk::fmx(r.a.y, r.b.y);
  const float ly = rk::fmn(r.a.z, r.b.z), hy = rk::fmx(r.a.w, r.b.w);
  const float lz = rk::fmn(r.c.x, r.c.z), hz = rk::fmx(r.c.y, r.c.w);
  const float4 o = origins[i];
  const uint32_t qx = min(rk::f2u_sat((o.x - lx) * (1024.0f / (hx - lx))), 1023u);
  const uint32_t qy = min(rk::f2u_sat((o.y - ly) * (1024.0f / (hy - ly))), 1023u);
  const uint32_t qz = min(rk::f2u_sat((o.z - lz) * (1024.0f / (hz - lz))), 1023u);
  keys[i] = (spread10(qx) << 2) | (spread10(qy) << 1) | spread10(qz);
  vals[i] = i;
}

// Rays of one frame differ in cost by orders of magnitude: on a terrain seen
// at a low angle, the rays that graze the horizon make hundreds of node
// visits and the rays that leave the root box make one, and the grazing rays
// fill a few rows. A wave lasts as long as its costliest rays; its idle lanes
// help (trace_ray_shared), but a wave of 64 grazing rays has none idle. So a
// wave does not take 64 consecutive rays: it takes runs of kSpreadRun
// consecutive rays (a 128-B line of origins and one of directions) from
// 64 / kSpreadRun bands of the ray array, the run at the same place in each
// band. Runs of 8 measured best on the 1920x1080 heightfield frame (0.224 ms
// against 0.257 ms for 64 consecutive rays; 0.276 ms for runs of 4, 0.231
// for runs of 16). 0 traces the rays in array order. Sorted traces keep
// their order.
#ifndef RK_TRACE_SPREAD
#define RK_TRACE_SPREAD 8
#endif
constexpr uint32_t kSpreadRun = RK_TRACE_SPREAD;

// The ray of thread t (n_rays when none).
__device__ __forceinline__ uint32_t spread_ray(uint32_t t, uint32_t n_rays) {
  return rk::band_ray<kSpreadRun>(t, n_rays);
}

// order (may be NULL): thread t traces ray order[t], so rays that start near
// each other run in the same wave; results still go to the ray's own slot.
// Threads per workgroup of k_fast_trace and k_fast_trace_camera. One wave per
// workgroup measured best on the 1920x1080 heightfield frame (rays 0.211 ms,
// camera 0.199 ms; 128: 0.218 / 0.200; 256: 0.228 / 0.209; 512: 0.251 /
// 0.230) and within 1% on the 10M soup's incoherent rays: a finished wave
// frees its slot at once, where a larger workgroup waits for its slowest wave.
#ifndef RK_TRACE_WG
#define RK_TRACE_WG 64
#endif
constexpr int kTraceWG = RK_TRACE_WG;
// Pending subtrees are handed to idle lanes only on steps where at least this
// many lanes are idle (trace_ray_shared's kDonateMin): each exchange costs
// every lane of the wave a dozen shuffles and two LDS round trips, and a lane
// that waits a few steps for work costs less. 16 measured best on the
// 1920x1080 heightfield frame (0.183 ms against 0.189 with 1 and 0.215 with
// 32) and better than 1 on the 10M soup's incoherent rays (45.9 against 48.4
// ms; 42.5 with 32, which the autotuner picks there).
#ifndef RK_TRACE_DONATE_MIN
#define RK_TRACE_DONATE_MIN 16
#endif
constexpr int kTraceDonateMin = RK_TRACE_DONATE_MIN;
static_assert(kTraceWG % 64 == 0 && kTraceWG <= 1024, "whole waves");

// kWG threads per workgroup, runs of kRun rays per band (spread_ray; 0: 64
// consecutive rays per wave), kLds stack entries per lane in LDS. The
// defaults above are one variant; rk_fast_autotune chooses among several.
template <bool kCount, int kWG, uint32_t kRun, int kLds, bool kOldest = false, int kLeafBatch = 0, int kDonateMin = 1,
          int kBoundEvery = 1>
__global__ __launch_bounds__(kWG) void k_fast_trace_v(const FNode* __restrict__ nodes, const float4* __restrict__ tris,
                                                      uint32_t n_tris, const float4* __restrict__ origins,
                                                      const float4* __restrict__ dirs, uint32_t n_rays,
                                                      uint32_t* __restrict__ faces_id, float2* __restrict__ coords,
                                                      float* __restrict__ depths, float eps,
                                                      unsigned long long* __restrict__ counters,
                                                      const uint32_t* __restrict__ order, uint32_t flags) {
  const uint32_t t = blockIdx.x * kWG + threadIdx.x;
  const bool reset = (flags & RK_TRACE_RESET_DEPTH) != 0u;
  const uint32_t l = order ? (t < n_rays ? order[t] : n_rays) : rk::band_ray<kRun>(t, n_rays);
  const bool has_ray = l < n_rays;
  __shared__ uint2 s_stack[kLds * kWG];
  __shared__ uint32_t s_xchg[kWG / 64][kXchgWords];
  uint32_t n_box = 0, n_tri = 0;
  const float4 zero4 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  const RayHit h = trace_ray_shared<kCount, kStack, kLds, kWG, kOldest, kLeafBatch, kDonateMin, kBoundEvery>(
      nodes, tris, n_tris, 0u, has_ray, has_ray ? origins[l] : zero4, has_ray ? dirs[l] : zero4,
      has_ray ? (reset ? RK_HUGE : depths[l]) : 0.0f, eps, n_box, n_tri, s_stack, s_xchg[threadIdx.x >> 6]);
  if (kCount) {
    atomicAdd(counters, (unsigned long long)n_box);
    atomicAdd(counters + 1, (unsigned long long)n_tri);
  }
  if (has_ray) {
    if (h.pos != 0xFFFFFFFFu) {
      faces_id[l] = h.pos;
      coords[l] = make_float2(h.u, h.v);
      depths[l] = h.t;
    } else if (reset) {
      depths[l] = RK_HUGE;  // the reset a fill kernel would have done; the miss's id and coords stay
    }
  }
}

inline unsigned blocks_for(uint32_t n, uint32_t per) { return (n + per - 1) / per; }

// Up to this many faces the chunk pass also climbs the splits that cross
// chunks, in the thread that reaches them: the chain above a few thousand
// chunks is short, and a second launch would cost more than it saves (the 1M
// heightfield: 224 us against 229 us with k_fast_hierarchy_top). Above it, a
// workgroup waiting on that chain would hold its LDS while thousands of
// chunks queue, so the climbs go to k_fast_hierarchy_top.
#ifndef RK_INLINE_TOP_FACES
#define RK_INLINE_TOP_FACES (1u << 21)
#endif
constexpr uint32_t kInlineTopFaces = RK_INLINE_TOP_FACES;

// Workspace of the hierarchy passes over n positions.
struct HierWs {
  uint32_t* flags;       // counters of splits that cross chunks (zeroed per build)
  GPart* parts;          // parts exchanged at splits that cross chunks
  uint32_t* pend;        // per chunk, the climbs handed on to the top pass
  uint32_t* pend_count;  // per chunk, how many
};
inline size_t hier_ws_bytes(size_t n) {
  return rk_align_up(n * 4) + rk_align_up(n * sizeof(GPart)) + rk_align_up(n * 4) +
         rk_align_up(((n + kChunk - 1) / kChunk) * 4);
}
inline bool take_hier_ws(RkArena& ar, uint32_t n, HierWs& w) {
  w.flags = ar.take<uint32_t>(n);
  w.parts = ar.take<GPart>(n);
  w.pend = ar.take<uint32_t>(n);
  w.pend_count = ar.take<uint32_t>(blocks_for(n, kChunk));
  return w.flags && w.parts && w.pend && w.pend_count;
}
// Clears the counters and runs both hierarchy passes.
template <typename K, bool kBatch>
rk_status launch_hierarchy(hipStream_t st, const K* keys, const uint32_t* order, const uint4* f, const float4* v,
                           uint32_t n, uint32_t leaf_max, const HierWs& w, FNode* nodes, float4* tris,
                           const rk_batch_mesh* meshes, bool flags_clear = false) {
  if (!flags_clear) {
    rk_status s = rk_fill_u32(reinterpret_cast<rk_stream>(st), w.flags, n, 0u);
    if (s != RK_SUCCESS) return s;
  }
  const unsigned chunks = blocks_for(n, kChunk);
  const uint32_t inline_top = n <= kInlineTopFaces ? 1u : 0u;
  hipLaunchKernelGGL((k_fast_hierarchy<K, kBatch>), dim3(chunks), dim3(kLocalThreads), 0, st, keys, order, f, v, n,
                     leaf_max, w.flags, w.parts, nodes, tris, meshes, w.pend, w.pend_count, inline_top);
  RK_CHECK_LAUNCH();
  if (!inline_top) {
    hipLaunchKernelGGL((k_fast_hierarchy_top<K, kBatch>), dim3(chunks), dim3(64), 0, st, keys, n, leaf_max, w.flags,
                       w.parts, nodes, meshes, w.pend, w.pend_count);
    RK_CHECK_LAUNCH();
  }
  return RK_SUCCESS;
}

}  // namespace

extern "C" size_t rk_fast_build_workspace_size(uint32_t face_count, int morton_mode) {
  size_t n = face_count ? face_count : 1;
  size_t kb = morton_mode == RK_MORTON_63 ? 8 : 4;
  size_t b = 0;
  b += rk_align_up(n * kb);               // keys
  b += hier_ws_bytes(n);
  b += rk_sort_workspace_size((uint32_t)n, (uint32_t)kb);
  return b + 4096;
}

template <typename K>
static rk_status fast_build_keys(hipStream_t st, const uint4* f, const float4* v, uint32_t n, float4 rmin,
                                 float4 rmax, const rk_fast_bvh* bvh, uint32_t leaf_max, K* keys, const HierWs& hw,
                                 void* sort_ws, size_t sort_bytes, uint32_t* info) {
  uint32_t* order = bvh->indices;
  // the sort's workspace starts with its error word and digit counts
  const uint32_t head_words = 1u + (uint32_t)sizeof(K) * 256u;
  const uint32_t grid_n = std::max(n, std::max((uint32_t)RK_INFO_WORDS, head_words));
  hipLaunchKernelGGL(k_fast_prep<K>, dim3(blocks_for(grid_n, 256)), dim3(256), 0, st, f, v, n, rmin, rmax,
                     reinterpret_cast<Box*>(bvh->boxes), keys, order, info, hw.flags, static_cast<uint32_t*>(sort_ws),
                     head_words);
  RK_CHECK_LAUNCH();
  rk_status s = rk_radix_sort_impl<K>(st, keys, order, n, 0, (int)sizeof(K), sort_ws, sort_bytes,
                                      info + RK_INFO_FLAGS, true);
  if (s != RK_SUCCESS) return s;
  return launch_hierarchy<K, false>(st, keys, order, f, v, n, leaf_max, hw, static_cast<FNode*>(bvh->nodes),
                                    static_cast<float4*>(bvh->tris), (const rk_batch_mesh*)nullptr, true);
}

static rk_status fast_build_common(rk_stream stream, const rk_uint4* faces, const rk_float4* vertices,
                                   uint32_t face_count, const rk_aabox* hitbox, int morton_mode, uint32_t leaf_max,
                                   const rk_fast_bvh* bvh, uint32_t* info, void* sorted_keys, void* ws,
                                   size_t ws_bytes) {
  const uint32_t n = face_count;
  if (n == 0 || !faces || !vertices || !hitbox || !bvh || !info) return RK_ERROR_INVALID_VALUE;
  if (!bvh->nodes || !bvh->tris || !bvh->indices) return RK_ERROR_INVALID_VALUE;
  if (morton_mode != RK_MORTON_30 && morton_mode != RK_MORTON_63) return RK_ERROR_INVALID_VALUE;
  if (leaf_max < 1 || leaf_max > kMaxLeaf) return RK_ERROR_INVALID_VALUE;
  if (n >= (1u << 27)) return RK_ERROR_UNSUPPORTED;
  hipStream_t st = rk_hip_stream(stream);
  const bool wide = morton_mode == RK_MORTON_63;
  RkArena ar(ws, ws_bytes);
  void* keys = wide ? (void*)ar.take<uint64_t>(n) : (void*)ar.take<uint32_t>(n);
  HierWs hw;
  const bool hw_ok = take_hier_ws(ar, n, hw);
  size_t sort_bytes = rk_sort_workspace_size(n, wide ? 8 : 4);
  void* sort_ws = ar.take<char>(sort_bytes);
  if (!keys || !hw_ok || !sort_ws) return RK_ERROR_WORKSPACE;
  if (sorted_keys) keys = sorted_keys;
  float4 rmin = make_float4(hitbox->min.x, hitbox->min.y, hitbox->min.z, hitbox->min.w);
  float4 rmax = make_float4(hitbox->max.x, hitbox->max.y, hitbox->max.z, hitbox->max.w);
  const uint4* f = reinterpret_cast<const uint4*>(faces);
  const float4* v = reinterpret_cast<const float4*>(vertices);
  if (wide)
    return fast_build_keys<uint64_t>(st, f, v, n, rmin, rmax, bvh, leaf_max, static_cast<uint64_t*>(keys), hw,
                                     sort_ws, sort_bytes, info);
  return fast_build_keys<uint32_t>(st, f, v, n, rmin, rmax, bvh, leaf_max, static_cast<uint32_t*>(keys), hw,
                                   sort_ws, sort_bytes, info);
}

extern "C" rk_status rk_fast_build(rk_stream stream, const rk_uint4* faces, const rk_float4* vertices,
                                   uint32_t face_count, const rk_aabox* hitbox, int morton_mode, uint32_t leaf_max,
                                   const rk_fast_bvh* bvh, uint32_t* info, void* ws, size_t ws_bytes) {
  return fast_build_common(stream, faces, vertices, face_count, hitbox, morton_mode, leaf_max, bvh, info, nullptr,
                           ws, ws_bytes);
}

extern "C" rk_status rk_fast_build_keyed(rk_stream stream, const rk_uint4* faces, const rk_float4* vertices,
                                         uint32_t face_count, const rk_aabox* hitbox, int morton_mode,
                                         uint32_t leaf_max, const rk_fast_bvh* bvh, uint32_t* info, void* sorted_keys,
                                         void* ws, size_t ws_bytes) {
  if (!sorted_keys) return RK_ERROR_INVALID_VALUE;
  return fast_build_common(stream, faces, vertices, face_count, hitbox, morton_mode, leaf_max, bvh, info,
                           sorted_keys, ws, ws_bytes);
}

// ---- rk_fast_build without a caller workspace ------------------------------------
// The workspace of rk_fast_build (keys, hierarchy slots and counters, sort
// buffers and look-back status) lives in one file-scope device array. Every
// device the library runs on holds its own copy until the process ends. Host
// code never passes the array's name to a kernel: the device address comes
// from hipGetSymbolAddress, once per device, on that device's first call.
namespace {
constexpr size_t kNoAllocScratchBytes = RK_NOALLOC_SCRATCH_BYTES;
__device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

rk_status scratch_for_current_device(void** out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return RK_ERROR_DEVICE;
  if (dev < 0 || dev >= kMaxDevices) return RK_ERROR_UNSUPPORTED;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return RK_ERROR_DEVICE;
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return RK_SUCCESS;
}
}  // namespace

extern "C" rk_status rk_fast_build_noalloc(rk_stream stream, const rk_uint4* faces, const rk_float4* vertices,
                                           uint32_t face_count, const rk_aabox* hitbox, int morton_mode,
                                           uint32_t leaf_max, const rk_fast_bvh* bvh, uint32_t* info) {
  if (morton_mode != RK_MORTON_30 && morton_mode != RK_MORTON_63) return RK_ERROR_INVALID_VALUE;
  if (face_count >= (1u << 27) || rk_fast_build_workspace_size(face_count, morton_mode) > kNoAllocScratchBytes)
    return RK_ERROR_WORKSPACE;
  void* ws = nullptr;
  rk_status s = scratch_for_current_device(&ws);
  if (s != RK_SUCCESS) return s;
  return fast_build_common(stream, faces, vertices, face_count, hitbox, morton_mode, leaf_max, bvh, info, nullptr, ws,
                           kNoAllocScratchBytes);
}

static rk_status fast_refit(rk_stream stream, const rk_uint4* faces, const rk_float4* vertices, uint32_t face_count,
                            int morton_mode, uint32_t leaf_max, const rk_fast_bvh* bvh, const void* sorted_keys,
                            void* ws, size_t ws_bytes, bool counters_clear);

extern "C" rk_status rk_fast_refit(rk_stream stream, const rk_uint4* faces, const rk_float4* vertices,
                                   uint32_t face_count, int morton_mode, uint32_t leaf_max, const rk_fast_bvh* bvh,
                                   const void* sorted_keys, void* ws, size_t ws_bytes) {
  return fast_refit(stream, faces, vertices, face_count, morton_mode, leaf_max, bvh, sorted_keys, ws, ws_bytes,
                    false);
}

// For the request loop: where a refit over n faces keeps its split counters
// in ws (n words), and the refit that takes them as already zero.
uint32_t* rk_fast_refit_counters(void* ws, size_t ws_bytes, uint32_t n) {
  RkArena ar(ws, ws_bytes);
  HierWs hw;
  return take_hier_ws(ar, n, hw) ? hw.flags : nullptr;
}
rk_status rk_fast_refit_precleared(rk_stream stream, const rk_uint4* faces, const rk_float4* vertices,
                                   uint32_t face_count, int morton_mode, uint32_t leaf_max, const rk_fast_bvh* bvh,
                                   const void* sorted_keys, void* ws, size_t ws_bytes) {
  return fast_refit(stream, faces, vertices, face_count, morton_mode, leaf_max, bvh, sorted_keys, ws, ws_bytes, true);
}

static rk_status fast_refit(rk_stream stream, const rk_uint4* faces, const rk_float4* vertices, uint32_t face_count,
                            int morton_mode, uint32_t leaf_max, const rk_fast_bvh* bvh, const void* sorted_keys,
                            void* ws, size_t ws_bytes, bool counters_clear) {
  const uint32_t n = face_count;
  if (n == 0 || !faces || !vertices || !bvh || !sorted_keys) return RK_ERROR_INVALID_VALUE;
  if (!bvh->nodes || !bvh->tris || !bvh->indices) return RK_ERROR_INVALID_VALUE;
  if (morton_mode != RK_MORTON_30 && morton_mode != RK_MORTON_63) return RK_ERROR_INVALID_VALUE;
  if (leaf_max < 1 || leaf_max > kMaxLeaf) return RK_ERROR_INVALID_VALUE;
  if (n >= (1u << 27)) return RK_ERROR_UNSUPPORTED;
  hipStream_t st = rk_hip_stream(stream);
  RkArena ar(ws, ws_bytes);
  HierWs hw;
  if (!take_hier_ws(ar, n, hw)) return RK_ERROR_WORKSPACE;
  const uint4* f = reinterpret_cast<const uint4*>(faces);
  const float4* v = reinterpret_cast<const float4*>(vertices);
  FNode* nodes = static_cast<FNode*>(bvh->nodes);
  float4* tris = static_cast<float4*>(bvh->tris);
  if (morton_mode == RK_MORTON_63)
    return launch_hierarchy<uint64_t, false>(st, static_cast<const uint64_t*>(sorted_keys), bvh->indices, f, v, n,
                                             leaf_max, hw, nodes, tris, (const rk_batch_mesh*)nullptr,
                                             counters_clear);
  return launch_hierarchy<uint32_t, false>(st, static_cast<const uint32_t*>(sorted_keys), bvh->indices, f, v, n,
                                           leaf_max, hw, nodes, tris, (const rk_batch_mesh*)nullptr, counters_clear);
}


// ---- batches ------------------------------------------------------------------

namespace {

// The mesh whose [offset, offset + count) holds g, for offsets in increasing
// order (mesh_count >= 1): the last mesh with offset <= g.
__device__ __forceinline__ uint32_t mesh_of_face(const rk_batch_mesh* __restrict__ meshes, uint32_t mesh_count,
                                                 uint32_t g) {
  uint32_t lo = 0, hi = mesh_count - 1;
  while (lo < hi) {
    const uint32_t mid = (lo + hi + 1) >> 1;
    if (meshes[mid].face_offset <= g) lo = mid; else hi = mid - 1;
  }
  return lo;
}
__device__ __forceinline__ uint32_t mesh_of_ray(const rk_batch_mesh* __restrict__ meshes, uint32_t mesh_count,
                                                uint32_t r) {
  uint32_t lo = 0, hi = mesh_count - 1;
  while (lo < hi) {
    const uint32_t mid = (lo + hi + 1) >> 1;
    if (meshes[mid].ray_offset <= r) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// Per-face box centroid quantised to 10 bits per axis in its mesh's hitbox,
// keyed (mesh << 30 | code) so one sort orders every mesh at once.
__global__ __launch_bounds__(256) void k_batch_prep(const rk_batch_mesh* __restrict__ meshes, uint32_t mesh_count,
                                                    const uint4* __restrict__ faces, const float4* __restrict__ verts,
                                                    uint32_t n, const rk_aabox* __restrict__ hitboxes,
                                                    uint64_t* __restrict__ keys, uint32_t* __restrict__ vals) {
  const uint32_t g = blockIdx.x * 256 + threadIdx.x;
  if (g >= n) return;
  const uint32_t m = mesh_of_face(meshes, mesh_count, g);
  const uint32_t voff = meshes[m].vertex_offset;
  const uint4 f = faces[g];
  const float4 v0 = verts[f.x + voff], v1 = verts[f.y + voff], v2 = verts[f.z + voff];
  const float4 mn = rk::min4(rk::min4(v0, v1), v2), mx = rk::max4(rk::max4(v0, v1), v2);
  const rk_aabox hb = hitboxes[m];
  const float4 rmin = make_float4(hb.min.x, hb.min.y, hb.min.z, 0.0f);
  const float4 len = make_float4(hb.max.x - hb.min.x, hb.max.y - hb.min.y, hb.max.z - hb.min.z, 0.0f);
  const float4 c = rk::muls4(rk::add4(mn, mx), 0.5f);
  const float sx = 1024.0f / len.x, sy = 1024.0f / len.y, sz = 1024.0f / len.z;
  const uint32_t qx = min(rk::f2u_sat((c.x - rmin.x) * sx), 1023u);
  const uint32_t qy = min(rk::f2u_sat((c.y - rmin.y) * sy), 1023u);
  const uint32_t qz = min(rk::f2u_sat((c.z - rmin.z) * sz), 1023u);
  keys[g] = ((uint64_t)m << 30) | (uint64_t)((spread10(qx) << 2) | (spread10(qy) << 1) | spread10(qz));
  vals[g] = g;
}

// Sort key of ray r: (mesh << 30 | 30-bit Morton code of its origin in the
// mesh's root box), so one sort orders the rays of every mesh by where they
// start and keeps each mesh's rays together. A mesh of one face has no root
// record and keys all its rays alike.
__global__ __launch_bounds__(256) void k_batch_ray_keys(const rk_batch_mesh* __restrict__ meshes,
                                                        uint32_t mesh_count, const FNode* __restrict__ nodes,
                                                        const float4* __restrict__ origins, uint32_t n,
                                                        uint64_t* __restrict__ keys, uint32_t* __restrict__ vals) {
  const uint32_t r = blockIdx.x * 256 + threadIdx.x;
  if (r >= n) return;
  const uint32_t m = mesh_of_ray(meshes, mesh_count, r);
  uint32_t code = 0;
  if (meshes[m].face_count >= 2) {
    const FNode root = nodes[meshes[m].face_offset];
    const float lx = rk::fmn(root.a.x, root.b.x), hx = rk::fmx(root.a.y, root.b.y);
    const float ly = rk::fmn(root.a.z, root.b.z), hy = rk::fmx(root.a.w, root.b.w);
    const float lz = rk::fmn(root.c.x, root.c.z), hz = rk::fmx(root.c.y, root.c.w);
    const float4 o = origins[r];
    const uint32_t qx = min(rk::f2u_sat((o.x - lx) * (1024.0f / (hx - lx))), 1023u);
    const uint32_t qy = min(rk::f2u_sat((o.y - ly) * (1024.0f / (hy - ly))), 1023u);
    const uint32_t qz = min(rk::f2u_sat((o.z - lz) * (1024.0f / (hz - lz))), 1023u);
    code = (spread10(qx) << 2) | (spread10(qy) << 1) | spread10(qz);
  }
  keys[r] = ((uint64_t)m << 30) | code;
  vals[r] = r;
}

// Closest hit of ray r (r >= n_rays: none, the lane only helps the others of
// its wave), in its mesh's tree. Every lane of the wave calls it.
template <int kStackN, int kWG, int kDM>
__device__ __forceinline__ void batch_trace_one(const rk_batch_mesh* __restrict__ meshes, uint32_t mesh_count,
                                                const FNode* __restrict__ nodes, const float4* __restrict__ tris,
                                                uint32_t n_tris, const float4* __restrict__ origins,
                                                const float4* __restrict__ dirs, uint32_t r, uint32_t n_rays,
                                                uint32_t* __restrict__ faces_id, float2* __restrict__ coords,
                                                float* __restrict__ depths, float eps, uint2* lds_stack,
                                                uint32_t* xchg) {
  const bool has_ray = r < n_rays;
  const uint32_t root = has_ray ? meshes[mesh_of_ray(meshes, mesh_count, r)].face_offset : 0u;
  const float4 zero4 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  uint32_t nb = 0, nt = 0;
  const RayHit h = trace_ray_shared<false, kStackN, kLdsStack, kWG, false, 0, kDM>(
      nodes, tris, n_tris, root, has_ray, has_ray ? origins[r] : zero4, has_ray ? dirs[r] : zero4,
      has_ray ? depths[r] : 0.0f, eps, nb, nt, lds_stack, xchg);
  if (has_ray && h.pos != 0xFFFFFFFFu) {
    faces_id[r] = h.pos;
    coords[r] = make_float2(h.u, h.v);
    depths[r] = h.t;
  }
}

// One ray per thread, kWG rays per workgroup.
template <int kWG, int kStackN, int kDM>
__global__ __launch_bounds__(kWG) void k_batch_trace_grid(const rk_batch_mesh* __restrict__ meshes,
                                                          uint32_t mesh_count, const FNode* __restrict__ nodes,
                                                          const float4* __restrict__ tris, uint32_t n_tris,
                                                          const float4* __restrict__ origins,
                                                          const float4* __restrict__ dirs, uint32_t n_rays,
                                                          uint32_t* __restrict__ faces_id, float2* __restrict__ coords,
                                                          float* __restrict__ depths, float eps,
                                                          const uint32_t* __restrict__ order) {
  const uint32_t t = blockIdx.x * kWG + threadIdx.x;
  const uint32_t r = t < n_rays ? (order ? order[t] : t) : n_rays;
  __shared__ uint2 s_stack[kLdsStack * kWG];
  __shared__ uint32_t s_xchg[kWG / 64][kXchgWords];
  batch_trace_one<kStackN, kWG, kDM>(meshes, mesh_count, nodes, tris, n_tris, origins, dirs, r, n_rays, faces_id, coords,
                                depths, eps, s_stack, s_xchg[threadIdx.x >> 6]);
}

// Persistent grid: every wave takes 64-ray tiles from its XCD's queue
// (workgroups are dealt to the eight XCDs in turn, so blockIdx % 8 is the
// XCD) and moves to the next queue once one runs dry; a queue that ran dry
// stays dry, so after eight dry queues in a row the wave has seen all work.
constexpr uint32_t kQueues = 8, kQueueStride = 32;  // counters 128 B apart

template <int kWG, int kStackN, int kDM>
__global__ __launch_bounds__(kWG) void k_batch_trace_queue(const rk_batch_mesh* __restrict__ meshes,
                                                           uint32_t mesh_count, const FNode* __restrict__ nodes,
                                                           const float4* __restrict__ tris, uint32_t n_tris,
                                                           const float4* __restrict__ origins,
                                                           const float4* __restrict__ dirs, uint32_t n_rays,
                                                           uint32_t* __restrict__ faces_id,
                                                           float2* __restrict__ coords, float* __restrict__ depths,
                                                           float eps, uint32_t* __restrict__ queues,
                                                           const uint32_t* __restrict__ order) {
  __shared__ uint2 s_stack[kLdsStack * kWG];
  __shared__ uint32_t s_xchg[kWG / 64][kXchgWords];
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t n_tiles = (n_rays + 63) / 64;
  const uint32_t per = (n_tiles + kQueues - 1) / kQueues;
  uint32_t q = blockIdx.x % kQueues;
  for (uint32_t dry = 0; dry < kQueues;) {
    uint32_t t = 0;
    if (lane == 0) t = atomicAdd(queues + q * kQueueStride, 1u);
    t = __shfl(t, 0, 64);
    const uint32_t tile = q * per + t;
    if (t >= per || tile >= n_tiles) {
      q = (q + 1) % kQueues;
      ++dry;
      continue;
    }
    const uint32_t p = tile * 64 + lane;
    const uint32_t r = p < n_rays ? (order ? order[p] : p) : n_rays;
    batch_trace_one<kStackN, kWG, kDM>(meshes, mesh_count, nodes, tris, n_tris, origins, dirs, r, n_rays, faces_id, coords,
                                  depths, eps, s_stack, s_xchg[threadIdx.x >> 6]);
  }
}

uint32_t key_passes_for(uint32_t mesh_count) {
  uint32_t bits = 30;
  while (bits < 64 && (1ull << (bits - 30)) < mesh_count) ++bits;
  return (bits + 7) / 8;
}

bool config_ok(const rk_batch_config* c) {
  return c && c->leaf_max >= 1 && c->leaf_max <= kMaxLeaf && c->trace_mode <= 1 &&
         (c->trace_wg == 64 || c->trace_wg == 128 || c->trace_wg == 256) &&
         (c->stack_depth == 64 || c->stack_depth == 96) && c->ray_order <= 1 &&
         (c->donate_min == 0 || c->donate_min == (uint32_t)kTraceDonateMin || c->donate_min == 32);
}

}  // namespace

extern "C" rk_batch_config rk_batch_config_default(void) {
  rk_batch_config c = {};
  c.leaf_max = 1;
  c.trace_mode = 1;
  c.trace_wg = 256;
  c.stack_depth = 64;
  return c;
}

extern "C" size_t rk_batch_build_workspace_size(uint32_t mesh_count, uint32_t face_count) {
  (void)mesh_count;
  size_t n = face_count ? face_count : 1;
  size_t b = 0;
  b += rk_align_up(n * 8);  // (mesh, code) keys
  b += hier_ws_bytes(n);
  b += rk_sort_workspace_size((uint32_t)n, 8);
  return b + 4096;
}

extern "C" rk_status rk_batch_build(rk_stream stream, const rk_batch_mesh* meshes, uint32_t mesh_count,
                                    const rk_uint4* faces, const rk_float4* vertices, uint32_t face_count,
                                    const rk_aabox* hitboxes, const rk_batch_config* config, const rk_fast_bvh* bvh,
                                    uint32_t* info, void* ws, size_t ws_bytes) {
  const uint32_t n = face_count;
  if (!meshes || mesh_count == 0 || mesh_count > RK_BATCH_MAX_MESHES || n == 0 || !faces || !vertices ||
      !hitboxes || !bvh || !info || !config_ok(config))
    return RK_ERROR_INVALID_VALUE;
  if (!bvh->nodes || !bvh->tris || !bvh->indices) return RK_ERROR_INVALID_VALUE;
  if (n >= (1u << 27) || n < mesh_count) return RK_ERROR_UNSUPPORTED;
  hipStream_t st = rk_hip_stream(stream);
  RkArena ar(ws, ws_bytes);
  uint64_t* keys = ar.take<uint64_t>(n);
  HierWs hw;
  const bool hw_ok = take_hier_ws(ar, n, hw);
  const size_t sort_bytes = rk_sort_workspace_size(n, 8);
  void* sort_ws = ar.take<char>(sort_bytes);
  if (!keys || !hw_ok || !sort_ws) return RK_ERROR_WORKSPACE;
  rk_status s;
  if ((s = rk_fill_u32(stream, info, RK_INFO_WORDS, 0u)) != RK_SUCCESS) return s;
  const uint4* f = reinterpret_cast<const uint4*>(faces);
  const float4* v = reinterpret_cast<const float4*>(vertices);
  uint32_t* order = bvh->indices;
  hipLaunchKernelGGL(k_batch_prep, dim3(blocks_for(n, 256)), dim3(256), 0, st, meshes, mesh_count, f, v, n, hitboxes,
                     keys, order);
  RK_CHECK_LAUNCH();
  s = rk_radix_sort_impl<uint64_t>(st, keys, order, n, 0, (int)key_passes_for(mesh_count), sort_ws, sort_bytes,
                                   info + RK_INFO_FLAGS);
  if (s != RK_SUCCESS) return s;
  return launch_hierarchy<uint64_t, true>(st, keys, order, f, v, n, config->leaf_max, hw,
                                          static_cast<FNode*>(bvh->nodes), static_cast<float4*>(bvh->tris), meshes);
}

extern "C" size_t rk_batch_trace_workspace_size(uint32_t mesh_count, uint32_t ray_count) {
  (void)mesh_count;
  const size_t n = ray_count ? ray_count : 1;
  // the flags word, the queue counters, then (ray_order 1) the ray keys, the ray
  // order and the sort's own space
  return rk_align_up(4) + rk_align_up(kQueues * kQueueStride * 4) + rk_align_up(n * 8) + rk_align_up(n * 4) +
         rk_sort_workspace_size((uint32_t)n, 8) + 1024;
}

template <int kWG, int kStackN, int kDM>
static void launch_batch_trace(hipStream_t st, uint32_t mode, uint32_t grid_persistent, const rk_batch_mesh* meshes,
                               uint32_t mesh_count, const FNode* nodes, const float4* tris, uint32_t n_tris,
                               const float4* o, const float4* d, uint32_t n_rays, uint32_t* fid, float2* uv,
                               float* dep, float eps, uint32_t* queues, const uint32_t* order) {
  if (mode == 0)
    hipLaunchKernelGGL((k_batch_trace_grid<kWG, kStackN, kDM>), dim3(blocks_for(n_rays, kWG)), dim3(kWG), 0, st, meshes,
                       mesh_count, nodes, tris, n_tris, o, d, n_rays, fid, uv, dep, eps, order);
  else
    hipLaunchKernelGGL((k_batch_trace_queue<kWG, kStackN, kDM>), dim3(grid_persistent), dim3(kWG), 0, st, meshes,
                       mesh_count, nodes, tris, n_tris, o, d, n_rays, fid, uv, dep, eps, queues, order);
}

static inline float batch_eps_of(const rk_trace_options* o) {
  return o ? o->det_epsilon : RK_DET_EPSILON_REFERENCE;
}

extern "C" rk_status rk_batch_trace(rk_stream stream, const rk_batch_mesh* meshes, uint32_t mesh_count,
                                    const rk_fast_bvh* bvh, const rk_float4* origins, const rk_float4* directions,
                                    uint32_t ray_count, uint32_t* faces_id, rk_float2* coords, float* depths,
                                    const rk_trace_options* options, const rk_batch_config* config, void* ws,
                                    size_t ws_bytes) {
  if (!meshes || mesh_count == 0 || mesh_count > RK_BATCH_MAX_MESHES || !bvh || !bvh->nodes || !bvh->tris ||
      bvh->face_count == 0 || !config_ok(config))
    return RK_ERROR_INVALID_VALUE;
  if (ray_count == 0) return RK_SUCCESS;
  if (!origins || !directions || !faces_id || !coords || !depths) return RK_ERROR_INVALID_VALUE;
  hipStream_t st = rk_hip_stream(stream);
  RkArena ar(ws, ws_bytes);
  uint32_t* flags = ar.take<uint32_t>(1);  // word 0 of ws: RK_FLAG_* bits of this call
  uint32_t* queues = ar.take<uint32_t>(kQueues * kQueueStride);
  if (!flags || !queues) return RK_ERROR_WORKSPACE;
  {
    rk_status s = rk_fill_u32(stream, flags, 1, 0u);
    if (s != RK_SUCCESS) return s;
  }
  uint32_t grid_persistent = 0;
  if (config->trace_mode == 1) {
    rk_status s = rk_fill_u32(stream, queues, kQueues * kQueueStride, 0u);
    if (s != RK_SUCCESS) return s;
    int dev = 0, cus = 0;
    RK_HIP(hipGetDevice(&dev));
    RK_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    // 16 waves per CU, a multiple of the eight XCDs
    grid_persistent = (uint32_t)cus * 16u * 64u / config->trace_wg;
    grid_persistent = (grid_persistent + kQueues - 1) / kQueues * kQueues;
  }
  const FNode* nodes = static_cast<const FNode*>(bvh->nodes);
  const float4* tris = static_cast<const float4*>(bvh->tris);
  const float4* o = reinterpret_cast<const float4*>(origins);
  const float4* d = reinterpret_cast<const float4*>(directions);
  float2* uv = reinterpret_cast<float2*>(coords);
  const float eps = batch_eps_of(options);
  const uint32_t nt = bvh->face_count, mode = config->trace_mode;
  uint32_t* order = nullptr;
  if (config->ray_order == 1) {
    uint64_t* keys = ar.take<uint64_t>(ray_count);
    order = ar.take<uint32_t>(ray_count);
    const size_t sort_bytes = rk_sort_workspace_size(ray_count, 8);
    void* sort_ws = ar.take<char>(sort_bytes);
    if (!keys || !order || !sort_ws) return RK_ERROR_WORKSPACE;
    hipLaunchKernelGGL(k_batch_ray_keys, dim3(blocks_for(ray_count, 256)), dim3(256), 0, st, meshes, mesh_count,
                       nodes, o, ray_count, keys, order);
    RK_CHECK_LAUNCH();
    rk_status s = rk_radix_sort_impl<uint64_t>(st, keys, order, ray_count, 0, (int)key_passes_for(mesh_count),
                                               sort_ws, sort_bytes, flags);
    if (s != RK_SUCCESS) return s;
  }
#define RK_BATCH_TRACE(WG, SD, DM)                                                                                 \
  launch_batch_trace<WG, SD, DM>(st, mode, grid_persistent, meshes, mesh_count, nodes, tris, nt, o, d, ray_count,  \
                                 faces_id, uv, depths, eps, queues, order)
#define RK_BATCH_TRACE_DM(WG, SD)                        \
  do {                                                   \
    if (dm32) RK_BATCH_TRACE(WG, SD, 32);                \
    else RK_BATCH_TRACE(WG, SD, kTraceDonateMin);        \
  } while (0)
  const bool deep = config->stack_depth == 96, dm32 = config->donate_min == 32;
  switch (config->trace_wg) {
    case 64: if (deep) RK_BATCH_TRACE_DM(64, 96); else RK_BATCH_TRACE_DM(64, 64); break;
    case 128: if (deep) RK_BATCH_TRACE_DM(128, 96); else RK_BATCH_TRACE_DM(128, 64); break;
    default: if (deep) RK_BATCH_TRACE_DM(256, 96); else RK_BATCH_TRACE_DM(256, 64); break;
  }
#undef RK_BATCH_TRACE_DM
#undef RK_BATCH_TRACE
  RK_CHECK_LAUNCH();
  return RK_SUCCESS;
}

static inline float det_eps_of(const rk_trace_options* o) {
  return o ? o->det_epsilon : RK_DET_EPSILON_REFERENCE;
}
static inline uint32_t trace_flags_of(const rk_trace_options* o) { return o ? (o->flags & RK_TRACE_RESET_DEPTH) : 0u; }


extern "C" rk_status rk_fast_trace_counted(rk_stream stream, const rk_fast_bvh* bvh, const rk_float4* origins,
                                           const rk_float4* directions, uint32_t n_rays, uint32_t* faces_id,
                                           rk_float2* coords, float* depths, const rk_trace_options* options,
                                           uint64_t* counters) {
  if (!bvh || !bvh->nodes || !bvh->tris || bvh->face_count == 0 || !counters) return RK_ERROR_INVALID_VALUE;
  if (n_rays == 0) return RK_SUCCESS;
  if (!origins || !directions || !faces_id || !coords || !depths) return RK_ERROR_INVALID_VALUE;
  hipLaunchKernelGGL((k_fast_trace_v<true, kTraceWG, kSpreadRun, kLdsStack, false, 0, kTraceDonateMin>), dim3(blocks_for(n_rays, kTraceWG)), dim3(kTraceWG), 0, rk_hip_stream(stream),
                     static_cast<const FNode*>(bvh->nodes), static_cast<const float4*>(bvh->tris), bvh->face_count,
                     reinterpret_cast<const float4*>(origins), reinterpret_cast<const float4*>(directions), n_rays,
                     faces_id, reinterpret_cast<float2*>(coords), depths, det_eps_of(options),
                     reinterpret_cast<unsigned long long*>(counters), (const uint32_t*)nullptr, trace_flags_of(options));
  RK_CHECK_LAUNCH();
  return RK_SUCCESS;
}

// ---- camera frames, deformation, frame checksums --------------------------------

namespace {

// Bands of the camera kernel: 8 bands of 8-pixel runs measured best on the
// 1920x1080 heightfield frame with the 64-VGPR traversal (0.181 ms; 16 bands
// of 4-pixel runs 0.184, 4 bands 0.192). At 79 VGPRs 16 bands had been best
// (0.212 ms; 8 bands 0.216, 32 bands 0.218, 64 bands of single pixels 0.231,
// and 8x8 tiles per wave 0.249).
#ifndef RK_CAM_BANDS
#define RK_CAM_BANDS 8
#endif
constexpr int kCamBands = RK_CAM_BANDS, kCamRun = 64 / RK_CAM_BANDS;
static_assert(kCamBands >= 1 && 64 % kCamBands == 0, "runs tile a wave");

__device__ __forceinline__ uint64_t mix64(uint64_t z) {  // SplitMix64 finaliser
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
__device__ __forceinline__ float u01_of(uint64_t h) { return (float)(h >> 40) * (1.0f / 16777216.0f); }

// v = base + amplitude * sin(omega * time + phase_v) * axis_v; phase and axis
// hashed from the vertex index (axis uniform on the sphere, or +z).
// Threads below clear_n also zero clear[i] (the request loop clears a
// refit's split counters here instead of in a launch of its own).
__global__ __launch_bounds__(256) void k_deform(const float4* __restrict__ base, float4* __restrict__ out,
                                                uint32_t count, float amplitude, float omega, float time,
                                                uint64_t seed, int along_z, uint32_t* __restrict__ clear,
                                                uint32_t clear_n) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i < clear_n) clear[i] = 0u;
  if (i >= count) return;
  const uint64_t h0 = mix64(seed ^ (0x9E3779B97F4A7C15ull * ((uint64_t)i + 1)));
  const uint64_t h1 = mix64(h0 + 0x632BE59BD9B4E019ull), h2 = mix64(h1 + 0x632BE59BD9B4E019ull);
  const float phase = 6.2831853f * u01_of(h0);
  float ax = 0.0f, ay = 0.0f, az = 1.0f;
  if (!along_z) {
    const float z = 2.0f * u01_of(h1) - 1.0f, phi = 6.2831853f * u01_of(h2);
    const float r = sqrtf(fmaxf(0.0f, 1.0f - z * z));
    ax = r * cosf(phi);
    ay = r * sinf(phi);
    az = z;
  }
  const float s = amplitude * sinf(omega * time + phase);
  const float4 b = base[i];
  out[i] = make_float4(b.x + s * ax, b.y + s * ay, b.z + s * az, b.w);
}

// Order-independent frame checksum: the wrapping sum over rays of a hash of
// (ray, original face id, depth, u, v bits), and the hit count. Original ids
// make it independent of the tree the frame was traced with. A fixed grid of
// at most kSumBlocks workgroups walks the frame, reduces in registers, waves
// and LDS, and adds once per workgroup: same-address atomics from every wave
// would serialise across XCDs.
constexpr uint32_t kSumBlocks = 256;
static_assert(kSumBlocks <= 256, "k_checksum_to_host's last workgroup reads one share per thread");

// One ray's term of the frame checksum.
__device__ __forceinline__ uint64_t ray_hash(uint32_t i, uint32_t orig, float depth, float2 c) {
  return mix64(((uint64_t)i << 32 | orig) ^ mix64(((uint64_t)__float_as_uint(depth) << 32) ^
                                                  ((uint64_t)__float_as_uint(c.x) * 0x100000001B3ull) ^
                                                  __float_as_uint(c.y)));
}

// Wave and workgroup sums of (h, n), valid in thread 0 of the workgroup.
__device__ __forceinline__ void block_sum2(uint64_t& h, uint64_t& n) {
  __shared__ uint64_t s_h[4], s_n[4];
  for (int d = 32; d >= 1; d >>= 1) {  // lane 0 ends with the wave's sums
    h += shfl_down64(h, d);
    n += shfl_down64(n, d);
  }
  const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) { s_h[w] = h; s_n[w] = n; }
  __syncthreads();
  h = s_h[0] + s_h[1] + s_h[2] + s_h[3];
  n = s_n[0] + s_n[1] + s_n[2] + s_n[3];
}

// This workgroup's share of the frame checksum: (hash sum, hits), valid in thread 0.
__device__ __forceinline__ void block_checksum(const uint32_t* __restrict__ faces_id,
                                               const float2* __restrict__ coords, const float* __restrict__ depths,
                                               uint32_t count, const uint32_t* __restrict__ indices,
                                               uint32_t n_faces, uint64_t& bh, uint64_t& bn) {
  uint64_t h = 0, hit = 0;
  for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < count; i += gridDim.x * 256) {
    const uint32_t f = faces_id[i];
    const bool is_hit = f != 0xFFFFFFFFu && f < n_faces;
    h += ray_hash(i, is_hit ? (indices ? indices[f] : f) : 0xFFFFFFFFu, depths[i], coords[i]);
    hit += is_hit ? 1 : 0;
  }
  block_sum2(h, hit);
  bh = h;
  bn = hit;
}

__global__ __launch_bounds__(256) void k_checksum(const uint32_t* __restrict__ faces_id,
                                                  const float2* __restrict__ coords,
                                                  const float* __restrict__ depths, uint32_t count,
                                                  const uint32_t* __restrict__ indices, uint32_t n_faces,
                                                  unsigned long long* __restrict__ out) {
  uint64_t bh, bn;
  block_checksum(faces_id, coords, depths, count, indices, n_faces, bh, bn);
  if (threadIdx.x == 0) {
    atomicAdd(out, (unsigned long long)bh);
    atomicAdd(out + 1, (unsigned long long)bn);
  }
}

// The end of a one-launch checksum, called by every thread of every
// workgroup with the workgroup's share in thread 0 (bh, bn): each workgroup
// stores its share in partials (2 words each), and the last one to count
// itself in adds the shares, writes the total to out[0..1] (host memory
// mapped for the device), sets the counter back to zero for the next launch,
// and then writes done_value to out[2]: the host sees the frame finished
// without a runtime query.
// out[0] = hash sum, out[1] = hits; with build_flags, out[3] = *build_flags;
// out[2] = done_value last.
__device__ __forceinline__ void finish_checksum(uint64_t bh, uint64_t bn, uint64_t* __restrict__ partials,
                                                uint32_t* __restrict__ counter, uint64_t* out, uint64_t done_value,
                                                const uint32_t* build_flags = nullptr) {
  __shared__ uint32_t s_last;
  if (threadIdx.x == 0) {
    partials[2 * blockIdx.x] = bh;
    partials[2 * blockIdx.x + 1] = bn;
    __threadfence();
    s_last = atomicAdd(counter, 1u) == gridDim.x - 1 ? 1u : 0u;
  }
  __syncthreads();
  if (!s_last) return;  // workgroup-uniform
  __threadfence();
  uint64_t h = 0, n = 0;
  for (uint32_t b = threadIdx.x; b < gridDim.x; b += blockDim.x) {
    h += __hip_atomic_load(partials + 2 * b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    n += __hip_atomic_load(partials + 2 * b + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  for (int d = 32; d >= 1; d >>= 1) {
    h += shfl_down64(h, d);
    n += shfl_down64(n, d);
  }
  __shared__ uint64_t s_h[4], s_n[4];
  const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) { s_h[w] = h; s_n[w] = n; }
  __syncthreads();
  if (threadIdx.x == 0) {
    out[0] = s_h[0] + s_h[1] + s_h[2] + s_h[3];
    out[1] = s_n[0] + s_n[1] + s_n[2] + s_n[3];
    if (build_flags) out[3] = __hip_atomic_load(build_flags, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    *counter = 0u;
    __threadfence_system();  // the total lands before the done word
    __hip_atomic_store(out + 2, done_value, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}

// The checksum of a frame in memory in one launch (finish_checksum).
__global__ __launch_bounds__(256) void k_checksum_to_host(const uint32_t* __restrict__ faces_id,
                                                          const float2* __restrict__ coords,
                                                          const float* __restrict__ depths, uint32_t count,
                                                          const uint32_t* __restrict__ indices, uint32_t n_faces,
                                                          uint64_t* __restrict__ partials,
                                                          uint32_t* __restrict__ counter, uint64_t* out,
                                                          uint64_t done_value, const uint32_t* build_flags) {
  uint64_t bh, bn;
  block_checksum(faces_id, coords, depths, count, indices, n_faces, bh, bn);
  finish_checksum(bh, bn, partials, counter, out, done_value, build_flags);
}

// Which pixel a camera-kernel lane traces. Wave g takes a row segment of
// 64 / kBands pixels in each of kBands horizontal bands of the image, at the
// same place in every band, for the reason given at spread_ray.
//
// With kXcd, workgroups are dealt to the eight XCDs round-robin (block b runs
// on XCD b mod 8), and XCD x takes the x-th of eight vertical strips of the
// image: its L2 then holds the nodes of one eighth of the view, where each L2
// would otherwise fetch the nodes of the whole view. The waves of a strip are
// taken row segment by row segment; the horizon, which crosses every strip,
// spreads over all eight. Blocks past a strip's waves have no pixels.
template <int kWG, int kBands, bool kXcd>
__device__ __forceinline__ bool camera_pixel(int width, int height, int& x, int& y) {
  constexpr int kRun = 64 / kBands;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int segs = (width + kRun - 1) / kRun, band_h = (height + kBands - 1) / kBands;
  int seg, row_in_band;
  if (kXcd) {
    const int sw = (segs + 7) / 8;  // segments per strip
    const int xcd = blockIdx.x & 7;
    const int w = (int)(blockIdx.x >> 3) * (kWG / 64) + wave;
    seg = xcd * sw + w % sw;
    row_in_band = seg < segs ? w / sw : band_h;  // the last strip can be narrower
  } else {
    const int g = (int)blockIdx.x * (kWG / 64) + wave;
    seg = g % segs;
    row_in_band = g / segs;
  }
  x = seg * kRun + lane % kRun;
  y = (lane / kRun) * band_h + row_in_band;
  return row_in_band < band_h && x < width && y < height;
}

// Primary ray of pixel (x, y) generated in the trace kernel itself (the
// reference's CUDA_calcPrimaryRays.cu:43-55 formula: origin, and direction
// (dir - x * step_x) - y * step_y), so a frame reads no ray arrays. Every pixel is written:
// a miss gives face ~0, coords (0, 0), depth FLT_MAX.
template <int kWG, int kBands, bool kXcd, bool kOldest = false, int kLeafBatch = 0, int kDonateMin = 1>
__global__ __launch_bounds__(kWG) void k_fast_trace_camera_v(const FNode* __restrict__ nodes,
                                                             const float4* __restrict__ tris, uint32_t n_tris,
                                                             float4 ro, float4 rd, float4 rx, float4 ry, int width,
                                                             int height, uint32_t* __restrict__ faces_id,
                                                             float2* __restrict__ coords, float* __restrict__ depths,
                                                             float eps) {
  int x, y;
  const bool has_ray = camera_pixel<kWG, kBands, kXcd>(width, height, x, y);
  const float4 d = rk::sub4(rk::sub4(rd, rk::muls4(rx, (float)x)), rk::muls4(ry, (float)y));
  uint32_t nb = 0, nt = 0;
  __shared__ uint2 s_stack[kLdsStack * kWG];
  __shared__ uint32_t s_xchg[kWG / 64][kXchgWords];
  const RayHit h = trace_ray_shared<false, kStack, kLdsStack, kWG, kOldest, kLeafBatch, kDonateMin>(
      nodes, tris, n_tris, 0u, has_ray, ro, d,
                                                                            RK_HUGE, eps, nb, nt, s_stack,
                                                                            s_xchg[threadIdx.x >> 6]);
  if (!has_ray) return;
  const size_t p = (size_t)y * (size_t)width + (size_t)x;
  const bool hit = h.pos != 0xFFFFFFFFu;
  faces_id[p] = h.pos;
  coords[p] = hit ? make_float2(h.u, h.v) : make_float2(0.0f, 0.0f);
  depths[p] = hit ? h.t : RK_HUGE;
}

// The sample-02 frame in one launch: each pixel's primary ray, its closest hit
// (as k_fast_trace_camera_v), the hit face's normal at (u, v) (as
// rk_interpolate_data) and the pixel (as rk_shade), with the same arithmetic in
// the same order as those three kernels, so the pixels are bit-identical to
// theirs. faces: the mesh's faces in the tree's leaf order (faces[position]),
// as the reference's builds leave them. faces_id, coords and depths, when not
// NULL, receive what k_fast_trace_camera_v writes.
template <int kWG, int kBands, int kDM>
__global__ __launch_bounds__(kWG) void k_fast_render_camera(
    const FNode* __restrict__ nodes, const float4* __restrict__ tris, uint32_t n_tris, float4 ro, float4 rd,
    float4 rx, float4 ry, int width, int height, const uint4* __restrict__ faces, const float4* __restrict__ verts,
    const float4* __restrict__ norms, float4 back, float4 cam, uint32_t* __restrict__ pixels, int pitch,
    uint32_t* __restrict__ faces_id, float2* __restrict__ coords, float* __restrict__ depths, float eps) {
  int x, y;
  const bool has_ray = camera_pixel<kWG, kBands, false>(width, height, x, y);
  const float4 d = rk::sub4(rk::sub4(rd, rk::muls4(rx, (float)x)), rk::muls4(ry, (float)y));
  uint32_t nb = 0, nt = 0;
  __shared__ uint2 s_stack[kLdsStack * kWG];
  __shared__ uint32_t s_xchg[kWG / 64][kXchgWords];
  const RayHit h = trace_ray_shared<false, kStack, kLdsStack, kWG, false, 0, kDM>(
      nodes, tris, n_tris, 0u, has_ray, ro, d, RK_HUGE, eps, nb, nt, s_stack, s_xchg[threadIdx.x >> 6]);
  if (!has_ray) return;
  const bool hit = h.pos != 0xFFFFFFFFu;  // then t < FLT_MAX: the shader's depth test
  if (faces_id) {
    const size_t p = (size_t)y * (size_t)width + (size_t)x;
    faces_id[p] = h.pos;
    coords[p] = hit ? make_float2(h.u, h.v) : make_float2(0.0f, 0.0f);
    depths[p] = hit ? h.t : RK_HUGE;
  }
  pixels[(size_t)y * (size_t)pitch + (size_t)x] =
      hit ? rk::shade_hit(cam, rk::face_normal_at(faces[h.pos], h.u, h.v, verts, norms)) : rk::make_pixel(back);
}

// Workgroups of a camera-kernel variant for a width x height frame.
uint32_t camera_blocks(int width, int height, int wg, int bands, bool xcd) {
  const int run = 64 / bands;
  const uint64_t segs = (uint64_t)((width + run - 1) / run), band_h = (uint64_t)((height + bands - 1) / bands);
  const uint64_t wpb = (uint64_t)(wg / 64);
  if (xcd) {
    const uint64_t sw = (segs + 7) / 8;
    return (uint32_t)(8 * ((sw * band_h + wpb - 1) / wpb));
  }
  return (uint32_t)((segs * band_h + wpb - 1) / wpb);
}

}  // namespace

// ---- traversal variants and the autotuner ------------------------------------------
namespace {

typedef void (*TraceKernel)(const FNode*, const float4*, uint32_t, const float4*, const float4*, uint32_t, uint32_t*,
                            float2*, float*, float, unsigned long long*, const uint32_t*, uint32_t);
typedef void (*CameraKernel)(const FNode*, const float4*, uint32_t, float4, float4, float4, float4, int, int,
                             uint32_t*, float2*, float*, float);
struct TraceVariant {
  uint32_t wg, run, lds, oldest, leaf_batch, donate_min, bound_every;
  TraceKernel kernel;
};
struct CameraVariant {
  uint32_t wg, bands, xcd, oldest, leaf_batch, donate_min;
  CameraKernel kernel;
};

// rk_trace_options.config k selects entry k - 1 (0: the first, the default).
#define RK_TRACE_VARIANT(wg, run, lds, old, lb, dm) \
  {wg, run, lds, old, lb, dm, 1, k_fast_trace_v<false, wg, run, lds, old != 0, lb, dm>}
#define RK_TRACE_VARIANT_B(dm, be) {64, 8, 8, 0, 0, dm, be, k_fast_trace_v<false, 64, 8, 8, false, 0, dm, be>}

const TraceVariant kTraceVariants[] = {
    RK_TRACE_VARIANT(kTraceWG, kSpreadRun, kLdsStack, 0, 0, kTraceDonateMin),
    RK_TRACE_VARIANT(64, 4, 8, 0, 0, 1),   RK_TRACE_VARIANT(64, 16, 8, 0, 0, 1), RK_TRACE_VARIANT(64, 0, 8, 0, 0, 1),
    RK_TRACE_VARIANT(64, 8, 4, 0, 0, 1),   RK_TRACE_VARIANT(64, 8, 12, 0, 0, 1), RK_TRACE_VARIANT(128, 8, 8, 0, 0, 1),
    RK_TRACE_VARIANT(128, 16, 8, 0, 0, 1), RK_TRACE_VARIANT(256, 8, 8, 0, 0, 1), RK_TRACE_VARIANT(64, 8, 8, 1, 0, 1),
    RK_TRACE_VARIANT(64, 4, 8, 1, 0, 1),   RK_TRACE_VARIANT(64, 0, 8, 1, 0, 1),  RK_TRACE_VARIANT(128, 8, 8, 1, 0, 1),
    RK_TRACE_VARIANT(64, 8, 8, 0, 8, 1),   RK_TRACE_VARIANT(64, 8, 8, 0, 16, 1), RK_TRACE_VARIANT(64, 8, 8, 0, 32, 1),
    RK_TRACE_VARIANT(64, 8, 8, 0, 0, 8),   RK_TRACE_VARIANT(64, 8, 8, 0, 0, 1),  RK_TRACE_VARIANT(64, 8, 8, 0, 16, 8),
    RK_TRACE_VARIANT(64, 8, 8, 0, 16, 16), RK_TRACE_VARIANT(64, 8, 8, 0, 32, 16), RK_TRACE_VARIANT(64, 8, 8, 0, 0, 32),
    RK_TRACE_VARIANT(64, 8, 8, 1, 0, 16),
    RK_TRACE_VARIANT_B(16, 2),             RK_TRACE_VARIANT_B(16, 4),            RK_TRACE_VARIANT_B(32, 2),
    RK_TRACE_VARIANT_B(32, 4),
};
#undef RK_TRACE_VARIANT
#undef RK_TRACE_VARIANT_B
#define RK_CAMERA_VARIANT(wg, bands, xcd, old, lb, dm) \
  {wg, bands, xcd, old, lb, dm, k_fast_trace_camera_v<wg, bands, xcd != 0, old != 0, lb, dm>}
const CameraVariant kCameraVariants[] = {
    RK_CAMERA_VARIANT(kTraceWG, kCamBands, 0, 0, 0, kTraceDonateMin),
    RK_CAMERA_VARIANT(64, 8, 1, 0, 0, 1),   RK_CAMERA_VARIANT(64, 4, 0, 0, 0, 1),   RK_CAMERA_VARIANT(64, 4, 1, 0, 0, 1),
    RK_CAMERA_VARIANT(64, 16, 0, 0, 0, 1),  RK_CAMERA_VARIANT(64, 16, 1, 0, 0, 1),  RK_CAMERA_VARIANT(128, 8, 0, 0, 0, 1),
    RK_CAMERA_VARIANT(128, 8, 1, 0, 0, 1),  RK_CAMERA_VARIANT(128, 16, 0, 0, 0, 1), RK_CAMERA_VARIANT(128, 16, 1, 0, 0, 1),
    RK_CAMERA_VARIANT(64, 8, 0, 1, 0, 1),   RK_CAMERA_VARIANT(64, 4, 0, 1, 0, 1),   RK_CAMERA_VARIANT(128, 8, 0, 1, 0, 1),
    RK_CAMERA_VARIANT(64, 8, 1, 1, 0, 1),   RK_CAMERA_VARIANT(64, 8, 0, 0, 16, 1),  RK_CAMERA_VARIANT(64, 8, 0, 0, 32, 1),
    RK_CAMERA_VARIANT(64, 8, 0, 0, 16, 16), RK_CAMERA_VARIANT(64, 8, 0, 0, 0, 1),   RK_CAMERA_VARIANT(64, 8, 0, 0, 0, 32),
    RK_CAMERA_VARIANT(128, 8, 0, 1, 0, 16),
};
#undef RK_CAMERA_VARIANT
constexpr uint32_t kNumTraceVariants = sizeof(kTraceVariants) / sizeof(kTraceVariants[0]);
constexpr uint32_t kNumCameraVariants = sizeof(kCameraVariants) / sizeof(kCameraVariants[0]);

const TraceVariant* trace_variant(uint32_t config) {
  return config == 0 ? &kTraceVariants[0] : (config <= kNumTraceVariants ? &kTraceVariants[config - 1] : nullptr);
}
const CameraVariant* camera_variant(uint32_t config) {
  return config == 0 ? &kCameraVariants[0] : (config <= kNumCameraVariants ? &kCameraVariants[config - 1] : nullptr);
}

rk_status launch_trace_variant(hipStream_t st, const TraceVariant& tv, const rk_fast_bvh* bvh,
                               const rk_float4* origins, const rk_float4* directions, uint32_t n_rays,
                               uint32_t* faces_id, rk_float2* coords, float* depths, float eps, uint32_t flags) {
  hipLaunchKernelGGL(tv.kernel, dim3(blocks_for(n_rays, tv.wg)), dim3(tv.wg), 0, st,
                     static_cast<const FNode*>(bvh->nodes), static_cast<const float4*>(bvh->tris), bvh->face_count,
                     reinterpret_cast<const float4*>(origins), reinterpret_cast<const float4*>(directions), n_rays,
                     faces_id, reinterpret_cast<float2*>(coords), depths, eps, (unsigned long long*)nullptr,
                     (const uint32_t*)nullptr, flags);
  RK_CHECK_LAUNCH();
  return RK_SUCCESS;
}

void describe_trace(uint32_t config, rk_trace_tuning* out) {
  const TraceVariant& v = *trace_variant(config);
  *out = rk_trace_tuning{};
  out->config = config;
  out->workgroup = v.wg;
  out->ray_run = v.run;
  out->lds_stack = v.lds;
  out->steal_oldest = v.oldest;
  out->leaf_batch = v.leaf_batch;
  out->donate_min = v.donate_min;
  out->bound_every = v.bound_every;
}
void describe_camera(uint32_t config, rk_trace_tuning* out) {
  const CameraVariant& v = *camera_variant(config);
  *out = rk_trace_tuning{};
  out->config = config;
  out->workgroup = v.wg;
  out->ray_run = 64 / v.bands;
  out->bands = v.bands;
  out->lds_stack = kLdsStack;
  out->xcd_strips = v.xcd;
  out->steal_oldest = v.oldest;
  out->leaf_batch = v.leaf_batch;
  out->donate_min = v.donate_min;
  out->bound_every = 1;
}

// Times fn() reps times after one untimed call, each call preceded by
// prep() outside the timed region; returns the mean in ms.
template <typename Prep, typename Fn>
rk_status time_calls(hipStream_t st, hipEvent_t e0, hipEvent_t e1, int reps, Prep prep, Fn fn, float* ms) {
  float total = 0.0f;
  for (int r = 0; r <= reps; ++r) {
    rk_status s = prep();
    if (s != RK_SUCCESS) return s;
    RK_HIP(hipEventRecord(e0, st));
    s = fn();
    if (s != RK_SUCCESS) return s;
    RK_HIP(hipEventRecord(e1, st));
    RK_HIP(hipEventSynchronize(e1));
    float t = 0.0f;
    RK_HIP(hipEventElapsedTime(&t, e0, e1));
    if (r > 0) total += t;
  }
  *ms = total / (float)reps;
  return RK_SUCCESS;
}

struct EventPair {
  hipEvent_t e0 = nullptr, e1 = nullptr;
  ~EventPair() {
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
  }
};

}  // namespace

extern "C" rk_status rk_fast_trace_variant(uint32_t config, int camera, rk_trace_tuning* out) {
  if (!out) return RK_ERROR_INVALID_VALUE;
  if (camera ? !camera_variant(config) : !trace_variant(config)) return RK_ERROR_INVALID_VALUE;
  if (camera)
    describe_camera(config, out);
  else
    describe_trace(config, out);
  return RK_SUCCESS;
}

extern "C" size_t rk_fast_autotune_workspace_size(uint32_t n_rays) {
  return rk_align_up((size_t)(n_rays ? n_rays : 1) * 4) + 256;
}

extern "C" rk_status rk_fast_autotune(rk_stream stream, const rk_fast_bvh* bvh, const rk_float4* origins,
                                      const rk_float4* directions, uint32_t n_rays, uint32_t* faces_id,
                                      rk_float2* coords, float* depths, const rk_trace_options* options, int reps,
                                      void* ws, size_t ws_bytes, rk_trace_tuning* best) {
  if (!bvh || !bvh->nodes || !bvh->tris || bvh->face_count == 0 || !best) return RK_ERROR_INVALID_VALUE;
  if (n_rays == 0 || !origins || !directions || !faces_id || !coords || !depths) return RK_ERROR_INVALID_VALUE;
  if (reps < 1 || reps > 100) return RK_ERROR_INVALID_VALUE;
  if (!ws || ws_bytes < rk_fast_autotune_workspace_size(n_rays)) return RK_ERROR_WORKSPACE;
  hipStream_t st = rk_hip_stream(stream);
  const float eps = det_eps_of(options);
  float* tmax = static_cast<float*>(ws);
  RK_HIP(hipMemcpyAsync(tmax, depths, 4ull * n_rays, hipMemcpyDeviceToDevice, st));
  EventPair ev;
  RK_HIP(hipEventCreate(&ev.e0));
  RK_HIP(hipEventCreate(&ev.e1));
  // with RK_TRACE_RESET_DEPTH the calls do not read depths: nothing to restore
  const bool reset = (trace_flags_of(options) & RK_TRACE_RESET_DEPTH) != 0u;
  auto restore = [&]() -> rk_status {
    if (!reset) RK_HIP(hipMemcpyAsync(depths, tmax, 4ull * n_rays, hipMemcpyDeviceToDevice, st));
    return RK_SUCCESS;
  };
  uint32_t best_cfg = 0;
  float best_ms = 0.0f;
  for (uint32_t k = 0; k < kNumTraceVariants; ++k) {
    float ms = 0.0f;
    const TraceVariant& tv = kTraceVariants[k];
    rk_status s = time_calls(st, ev.e0, ev.e1, reps, restore, [&]() {
      return launch_trace_variant(st, tv, bvh, origins, directions, n_rays, faces_id, coords, depths, eps,
                                  trace_flags_of(options));
    }, &ms);
    if (s != RK_SUCCESS) return s;
    if (k == 0 || ms < best_ms) {
      best_ms = ms;
      best_cfg = k + 1;
    }
  }
  // leave the caller's frame traced by the chosen variant
  rk_status s = restore();
  if (s != RK_SUCCESS) return s;
  s = launch_trace_variant(st, *trace_variant(best_cfg), bvh, origins, directions, n_rays, faces_id, coords, depths,
                           eps, trace_flags_of(options));
  if (s != RK_SUCCESS) return s;
  describe_trace(best_cfg, best);
  best->ms = best_ms;
  return RK_SUCCESS;
}

extern "C" rk_status rk_fast_autotune_camera(rk_stream stream, const rk_fast_bvh* bvh, const rk_camera_rays* camera,
                                             int width, int height, uint32_t* faces_id, rk_float2* coords,
                                             float* depths, const rk_trace_options* options, int reps,
                                             rk_trace_tuning* best) {
  if (!bvh || !bvh->nodes || !bvh->tris || bvh->face_count == 0 || !camera || !best) return RK_ERROR_INVALID_VALUE;
  if (width <= 0 || height <= 0 || !faces_id || !coords || !depths) return RK_ERROR_INVALID_VALUE;
  if ((uint64_t)width * (uint64_t)height > 0xFFFFFFFFull) return RK_ERROR_UNSUPPORTED;
  if (reps < 1 || reps > 100) return RK_ERROR_INVALID_VALUE;
  hipStream_t st = rk_hip_stream(stream);
  EventPair ev;
  RK_HIP(hipEventCreate(&ev.e0));
  RK_HIP(hipEventCreate(&ev.e1));
  rk_trace_options opt = options ? *options : rk_trace_options{RK_DET_EPSILON_REFERENCE, 0u, 0u, 0u};
  uint32_t best_cfg = 0;
  float best_ms = 0.0f;
  for (uint32_t k = 0; k < kNumCameraVariants; ++k) {
    float ms = 0.0f;
    opt.config = k + 1;
    rk_status s = time_calls(st, ev.e0, ev.e1, reps, []() { return (rk_status)RK_SUCCESS; }, [&]() {
      return rk_fast_trace_camera(stream, bvh, camera, width, height, faces_id, coords, depths, &opt);
    }, &ms);
    if (s != RK_SUCCESS) return s;
    if (k == 0 || ms < best_ms) {
      best_ms = ms;
      best_cfg = k + 1;
    }
  }
  opt.config = best_cfg;
  rk_status s = rk_fast_trace_camera(stream, bvh, camera, width, height, faces_id, coords, depths, &opt);
  if (s != RK_SUCCESS) return s;
  describe_camera(best_cfg, best);
  best->ms = best_ms;
  return RK_SUCCESS;
}

extern "C" rk_status rk_fast_trace(rk_stream stream, const rk_fast_bvh* bvh, const rk_float4* origins,
                                   const rk_float4* directions, uint32_t n_rays, uint32_t* faces_id, rk_float2* coords,
                                   float* depths, const rk_trace_options* options) {
  if (!bvh || !bvh->nodes || !bvh->tris || bvh->face_count == 0) return RK_ERROR_INVALID_VALUE;
  if (n_rays == 0) return RK_SUCCESS;
  if (!origins || !directions || !faces_id || !coords || !depths) return RK_ERROR_INVALID_VALUE;
  const TraceVariant* tv = trace_variant(options ? options->config : 0u);
  if (!tv) return RK_ERROR_INVALID_VALUE;
  return launch_trace_variant(rk_hip_stream(stream), *tv, bvh, origins, directions, n_rays, faces_id, coords, depths,
                              det_eps_of(options), trace_flags_of(options));
}

extern "C" rk_status rk_fast_trace_camera(rk_stream stream, const rk_fast_bvh* bvh, const rk_camera_rays* camera,
                                          int width, int height, uint32_t* faces_id, rk_float2* coords, float* depths,
                                          const rk_trace_options* options) {
  if (!bvh || !bvh->nodes || !bvh->tris || bvh->face_count == 0 || !camera) return RK_ERROR_INVALID_VALUE;
  if (width <= 0 || height <= 0 || !faces_id || !coords || !depths) return RK_ERROR_INVALID_VALUE;
  if ((uint64_t)width * (uint64_t)height > 0xFFFFFFFFull) return RK_ERROR_UNSUPPORTED;
  auto f4 = [](const rk_float4& v) { return make_float4(v.x, v.y, v.z, v.w); };
  const CameraVariant* cv = camera_variant(options ? options->config : 0u);
  if (!cv) return RK_ERROR_INVALID_VALUE;
  dim3 grid(camera_blocks(width, height, (int)cv->wg, (int)cv->bands, cv->xcd != 0));
  hipLaunchKernelGGL(cv->kernel, grid, dim3(cv->wg), 0, rk_hip_stream(stream), static_cast<const FNode*>(bvh->nodes),
                     static_cast<const float4*>(bvh->tris), bvh->face_count, f4(camera->origin),
                     f4(camera->direction), f4(camera->step_x), f4(camera->step_y), width, height, faces_id,
                     reinterpret_cast<float2*>(coords), depths, det_eps_of(options));
  RK_CHECK_LAUNCH();
  return RK_SUCCESS;
}

extern "C" rk_status rk_fast_render_camera(rk_stream stream, const rk_fast_bvh* bvh, const rk_camera_rays* camera,
                                           int width, int height, const rk_uint4* faces, const rk_float4* vertices,
                                           const rk_float4* normals, rk_float4 back_color, rk_float4 cam_dir,
                                           uint32_t* pixels, int pitch, uint32_t* faces_id, rk_float2* coords,
                                           float* depths, const rk_trace_options* options) {
  if (!bvh || !bvh->nodes || !bvh->tris || bvh->face_count == 0 || !camera || !faces || !vertices || !pixels)
    return RK_ERROR_INVALID_VALUE;
  if (width <= 0 || height <= 0 || pitch < width) return RK_ERROR_INVALID_VALUE;
  if ((faces_id || coords || depths) && !(faces_id && coords && depths)) return RK_ERROR_INVALID_VALUE;
  if ((uint64_t)width * (uint64_t)height > 0xFFFFFFFFull || (uint64_t)pitch * (uint64_t)height > 0xFFFFFFFFull)
    return RK_ERROR_UNSUPPORTED;
  auto f4 = [](const rk_float4& v) { return make_float4(v.x, v.y, v.z, v.w); };
  dim3 grid(camera_blocks(width, height, kTraceWG, kCamBands, false));
  hipLaunchKernelGGL((k_fast_render_camera<kTraceWG, kCamBands, kTraceDonateMin>), grid, dim3(kTraceWG), 0,
                     rk_hip_stream(stream), static_cast<const FNode*>(bvh->nodes),
                     static_cast<const float4*>(bvh->tris), bvh->face_count, f4(camera->origin),
                     f4(camera->direction), f4(camera->step_x), f4(camera->step_y), width, height,
                     reinterpret_cast<const uint4*>(faces), reinterpret_cast<const float4*>(vertices),
                     reinterpret_cast<const float4*>(normals), f4(back_color),
                     make_float4(cam_dir.x, cam_dir.y, cam_dir.z, 0.0f), pixels, pitch, faces_id,
                     reinterpret_cast<float2*>(coords), depths, det_eps_of(options));
  RK_CHECK_LAUNCH();
  return RK_SUCCESS;
}

extern "C" rk_status rk_deform_vertices(rk_stream stream, const rk_float4* base, rk_float4* out, uint32_t count,
                                        float amplitude, float omega, float time, uint64_t seed, int along_z) {
  if (count == 0) return RK_SUCCESS;
  if (!base || !out) return RK_ERROR_INVALID_VALUE;
  hipLaunchKernelGGL(k_deform, dim3(blocks_for(count, 256)), dim3(256), 0, rk_hip_stream(stream),
                     reinterpret_cast<const float4*>(base), reinterpret_cast<float4*>(out), count, amplitude, omega,
                     time, seed, along_z, (uint32_t*)nullptr, 0u);
  RK_CHECK_LAUNCH();
  return RK_SUCCESS;
}

// For the request loop (rk_serve.hip): rk_deform_vertices that also zeroes
// clear[0, clear_n).
rk_status rk_deform_and_clear(hipStream_t st, const rk_float4* base, rk_float4* out, uint32_t count, float amplitude,
                              float omega, float time, uint64_t seed, int along_z, uint32_t* clear,
                              uint32_t clear_n) {
  if (!base || !out || (clear_n && !clear)) return RK_ERROR_INVALID_VALUE;
  const uint32_t n = std::max(count, clear_n);
  if (n == 0) return RK_SUCCESS;
  hipLaunchKernelGGL(k_deform, dim3(blocks_for(n, 256)), dim3(256), 0, st, reinterpret_cast<const float4*>(base),
                     reinterpret_cast<float4*>(out), count, amplitude, omega, time, seed, along_z, clear, clear_n);
  RK_CHECK_LAUNCH();
  return RK_SUCCESS;
}


extern "C" rk_status rk_frame_checksum(rk_stream stream, const uint32_t* faces_id, const rk_float2* coords,
                                       const float* depths, uint32_t count, const uint32_t* indices,
                                       uint32_t face_count, uint64_t* out) {
  if (!out) return RK_ERROR_INVALID_VALUE;
  rk_status s = rk_fill_u32(stream, reinterpret_cast<uint32_t*>(out), 4, 0u);
  if (s != RK_SUCCESS || count == 0) return s;
  if (!faces_id || !coords || !depths) return RK_ERROR_INVALID_VALUE;
  hipLaunchKernelGGL(k_checksum, dim3(std::min(blocks_for(count, 256), kSumBlocks)), dim3(256), 0,
                     rk_hip_stream(stream), faces_id,
                     reinterpret_cast<const float2*>(coords), depths, count, indices, face_count,
                     reinterpret_cast<unsigned long long*>(out));
  RK_CHECK_LAUNCH();
  return RK_SUCCESS;
}

// The frame checksum for the request loop: one launch, the total written to
// out_host[0..1], a device-writable mapping of pinned coherent host memory,
// then done_value to out_host[2]. partials holds 2 * kSumBlocks words and
// counter one word that is zero between launches (zero it once before the
// first).
rk_status rk_frame_checksum_to_host(hipStream_t st, const uint32_t* faces_id, const rk_float2* coords,
                                    const float* depths, uint32_t count, const uint32_t* indices,
                                    uint32_t face_count, uint64_t* partials, uint32_t* counter, uint64_t* out_host,
                                    uint64_t done_value, const uint32_t* build_flags) {
  if (!faces_id || !coords || !depths || !partials || !counter || !out_host) return RK_ERROR_INVALID_VALUE;
  hipLaunchKernelGGL(k_checksum_to_host, dim3(std::max(1u, std::min(blocks_for(count, 256), kSumBlocks))), dim3(256),
                     0, st, faces_id, reinterpret_cast<const float2*>(coords), depths, count, indices, face_count,
                     partials, counter, out_host, done_value, build_flags);
  RK_CHECK_LAUNCH();
  return RK_SUCCESS;
}

extern "C" size_t rk_fast_trace_sorted_workspace_size(uint32_t n_rays) {
  const size_t n = n_rays ? n_rays : 1;
  return rk_align_up(n * 4) * 2 + rk_sort_workspace_size((uint32_t)n, 4) + 1024;
}

extern "C" rk_status rk_fast_trace_sorted(rk_stream stream, const rk_fast_bvh* bvh, const rk_float4* origins,
                                          const rk_float4* directions, uint32_t n_rays, uint32_t* faces_id,
                                          rk_float2* coords, float* depths, const rk_trace_options* options,
                                          void* ws, size_t ws_bytes) {
  if (!bvh || !bvh->nodes || !bvh->tris || bvh->face_count == 0) return RK_ERROR_INVALID_VALUE;
  if (n_rays == 0) return RK_SUCCESS;
  if (!origins || !directions || !faces_id || !coords || !depths) return RK_ERROR_INVALID_VALUE;
  hipStream_t st = rk_hip_stream(stream);
  RkArena ar(ws, ws_bytes);
  const size_t sort_bytes = rk_sort_workspace_size(n_rays, 4);
  void* sort_ws = ar.take<char>(sort_bytes);  // first: its word 0 is the sort's error word
  uint32_t* keys = ar.take<uint32_t>(n_rays);
  uint32_t* order = ar.take<uint32_t>(n_rays);
  if (!keys || !order || !sort_ws) return RK_ERROR_WORKSPACE;
  const FNode* nodes = static_cast<const FNode*>(bvh->nodes);
  const float4* o = reinterpret_cast<const float4*>(origins);
  hipLaunchKernelGGL(k_ray_keys, dim3(blocks_for(n_rays, 256)), dim3(256), 0, st, nodes, o, n_rays, keys, order);
  RK_CHECK_LAUNCH();
  const TraceVariant* tv = trace_variant(options ? options->config : 0u);
  if (!tv) return RK_ERROR_INVALID_VALUE;
  rk_status s = rk_radix_sort_impl<uint32_t>(st, keys, order, n_rays, 0, 4, sort_ws, sort_bytes, nullptr);
  if (s != RK_SUCCESS) return s;
  // the variant's workgroup, stack and exchange settings; the order replaces its ray runs
  hipLaunchKernelGGL(tv->kernel, dim3(blocks_for(n_rays, tv->wg)), dim3(tv->wg), 0, st, nodes,
                     static_cast<const float4*>(bvh->tris), bvh->face_count, o,
                     reinterpret_cast<const float4*>(directions), n_rays, faces_id, reinterpret_cast<float2*>(coords),
                     depths, det_eps_of(options), (unsigned long long*)nullptr, (const uint32_t*)order,
                     trace_flags_of(options));
  RK_CHECK_LAUNCH();
  return RK_SUCCESS;
}
