// The throughput tree: a Morton LBVH built entirely on the GPU and a
// closest-hit traversal tuned for gfx950. No reference counterpart (the
// reference trees stop at tree_depth with leaves of up to 224 triangles); the
// triangle test is the reference's (CUDA_raytraceMeshTree.cu:57-89), so a
// hit's t, u, v are bit-identical to the reference's for the same triangle.
//
// Build:
//   k_fast_prep     per-face box (API array) and 30/63-bit Morton key
//   radix sort      (key, face) pairs
//   k_karras        internal node i of n-1 (Karras 2012): range, split,
//                   children, parent links
//   k_fast_refit    one thread per leaf walks to the root; the second thread
//                   to reach a node finishes it (atomic counter), writing each
//                   child's padded box, count and reference into the parent's
//                   64-B record; subtrees of <= leaf_max faces become leaves
//   k_fast_tris     leaf-ordered triangle records (v0, v1-v0, v2-v0), 48 B
//
// Node record (64 B): float4 (c0.lo.x, c0.hi.x, c0.lo.y, c0.hi.y),
//                     float4 (c1.lo.x, c1.hi.x, c1.lo.y, c1.hi.y),
//                     float4 (c0.lo.z, c0.hi.z, c1.lo.z, c1.hi.z),
//                     uint4  (ref0, ref1, count0, count1)
// ref: internal node index, or kLeafBit | first << 4 | (count - 1).
#include "rk_device.h"

template <typename K>
rk_status rk_radix_sort_impl(hipStream_t st, K* keys, uint32_t* vals, uint32_t n, int first_pass, int passes,
                             void* ws, size_t ws_bytes, uint32_t* err);

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

template <typename K>
__global__ __launch_bounds__(256) void k_fast_prep(const uint4* __restrict__ faces, const float4* __restrict__ verts,
                                                   uint32_t n, float4 rmin, float4 rmax, Box* __restrict__ boxes,
                                                   K* __restrict__ keys, uint32_t* __restrict__ vals) {
  uint32_t k = blockIdx.x * 256 + threadIdx.x;
  if (k >= n) return;
  uint4 f = faces[k];
  float4 v0 = verts[f.x], v1 = verts[f.y], v2 = verts[f.z];
  Box b;
  b.mn = rk::min4(rk::min4(v0, v1), v2);
  b.mx = rk::max4(rk::max4(v0, v1), v2);
  float4* o = reinterpret_cast<float4*>(boxes + k);
  o[0] = b.mn;
  o[1] = b.mx;
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

template <typename K>
__device__ __forceinline__ int kdelta(const K* __restrict__ keys, uint32_t n, int i, int j) {
  if (j < 0 || j >= (int)n) return -1;
  K a = keys[i], b = keys[j];
  const int kb = (int)sizeof(K) * 8;
  if (a == b) return kb + __clz((int)((uint32_t)i ^ (uint32_t)j));
  if (sizeof(K) == 4) return __clz((int)(uint32_t)(a ^ b));
  return __clzll((long long)(uint64_t)(a ^ b));
}

// Internal node i (Karras 2012, "Maximizing parallelism in the construction of
// BVHs, octrees and k-d trees"). Children are leaf (sorted position) or
// internal index; parents are recorded for the bottom-up pass.
template <typename K>
__global__ __launch_bounds__(256) void k_karras(const K* __restrict__ keys, uint32_t n, uint32_t* __restrict__ child,
                                                uint32_t* __restrict__ range_first, uint32_t* __restrict__ parent_node,
                                                uint32_t* __restrict__ parent_leaf) {
  const int i = (int)(blockIdx.x * 256 + threadIdx.x);
  if (i >= (int)n - 1) return;
  const int d = (kdelta(keys, n, i, i + 1) - kdelta(keys, n, i, i - 1)) >= 0 ? 1 : -1;
  const int dmin = kdelta(keys, n, i, i - d);
  int lmax = 2;
  while (kdelta(keys, n, i, i + lmax * d) > dmin) lmax <<= 1;
  int l = 0;
  for (int t = lmax >> 1; t >= 1; t >>= 1)
    if (kdelta(keys, n, i, i + (l + t) * d) > dmin) l += t;
  const int j = i + l * d;
  const int dnode = kdelta(keys, n, i, j);
  int s = 0;
  int t = l;
  do {
    t = (t + 1) >> 1;
    if (kdelta(keys, n, i, i + (s + t) * d) > dnode) s += t;
  } while (t > 1);
  const int gamma = i + s * d + (d < 0 ? -1 : 0);
  const int lo = i < j ? i : j, hi = i < j ? j : i;
  uint32_t left, right;
  if (lo == gamma) { left = kLeafBit | (uint32_t)gamma; parent_leaf[gamma] = (uint32_t)i; }
  else { left = (uint32_t)gamma; parent_node[gamma] = (uint32_t)i; }
  if (hi == gamma + 1) { right = kLeafBit | (uint32_t)(gamma + 1); parent_leaf[gamma + 1] = (uint32_t)i; }
  else { right = (uint32_t)(gamma + 1); parent_node[gamma + 1] = (uint32_t)i; }
  child[2 * i] = left;
  child[2 * i + 1] = right;
  range_first[i] = (uint32_t)lo;
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

__device__ __forceinline__ void write_slot(FNode* __restrict__ nodes, uint32_t parent, uint32_t side, float lox,
                                           float loy, float loz, float hix, float hiy, float hiz, uint32_t ref,
                                           uint32_t count) {
  pad_box(lox, hix);
  pad_box(loy, hiy);
  pad_box(loz, hiz);
  FNode* p = nodes + parent;
  float* ab = side ? reinterpret_cast<float*>(&p->b) : reinterpret_cast<float*>(&p->a);
  ab[0] = lox; ab[1] = hix; ab[2] = loy; ab[3] = hiy;
  float* cz = reinterpret_cast<float*>(&p->c) + 2 * side;
  cz[0] = loz; cz[1] = hiz;
  uint32_t* dd = reinterpret_cast<uint32_t*>(&p->d);
  dd[side] = ref;
  dd[2 + side] = count;
}

__global__ __launch_bounds__(256) void k_fast_refit(const Box* __restrict__ boxes, const uint32_t* __restrict__ order,
                                                    uint32_t n, const uint32_t* __restrict__ child,
                                                    const uint32_t* __restrict__ range_first,
                                                    const uint32_t* __restrict__ parent_node,
                                                    const uint32_t* __restrict__ parent_leaf,
                                                    uint32_t* __restrict__ flags, float* __restrict__ nbox,
                                                    uint32_t* __restrict__ ncount, FNode* __restrict__ nodes,
                                                    uint32_t leaf_max) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  Box b = rk::load_box(boxes + order[i]);
  float lox = b.mn.x, loy = b.mn.y, loz = b.mn.z, hix = b.mx.x, hiy = b.mx.y, hiz = b.mx.z;
  if (n == 1) {  // a single face: root with one leaf child and one empty child
    write_slot(nodes, 0, 0, lox, loy, loz, hix, hiy, hiz, kLeafBit | 0u, 1);
    FNode* p = nodes;
    p->b = make_float4(RK_HUGE, -RK_HUGE, RK_HUGE, -RK_HUGE);
    reinterpret_cast<float*>(&p->c)[2] = RK_HUGE;
    reinterpret_cast<float*>(&p->c)[3] = -RK_HUGE;
    reinterpret_cast<uint32_t*>(&p->d)[1] = kLeafBit | 0u;
    reinterpret_cast<uint32_t*>(&p->d)[3] = 0;
    return;
  }
  uint32_t count = 1;
  uint32_t ref = kLeafBit | (i << 4);  // count - 1 = 0
  uint32_t node = parent_leaf[i];
  uint32_t self = kLeafBit | i;
  // A Karras tree over n keys with index tie-breaks is at most 64 + 32 deep;
  // the cap only bounds the loop should the parent links ever be corrupt.
  // Sibling exchange without release/acquire: on a multi-XCD device an
  // agent-scope release writes back the whole local L2 and an acquire
  // invalidates it. Instead the exchanged words go through relaxed
  // agent-scope atomic stores and loads (coherent across XCDs), and the
  // counter increment waits until those stores are acknowledged.
  for (uint32_t guard = 0; guard < 4096u && node < n - 1; ++guard) {
    const uint32_t side = (child[2 * node] == self) ? 0u : 1u;
    write_slot(nodes, node, side, lox, loy, loz, hix, hiy, hiz, ref, count);
    uint32_t* mine = reinterpret_cast<uint32_t*>(nbox) + (size_t)(2 * node + side) * 6;
    rk::atomic_store_relaxed(mine + 0, __float_as_uint(lox));
    rk::atomic_store_relaxed(mine + 1, __float_as_uint(loy));
    rk::atomic_store_relaxed(mine + 2, __float_as_uint(loz));
    rk::atomic_store_relaxed(mine + 3, __float_as_uint(hix));
    rk::atomic_store_relaxed(mine + 4, __float_as_uint(hiy));
    rk::atomic_store_relaxed(mine + 5, __float_as_uint(hiz));
    rk::atomic_store_relaxed(ncount + 2 * node + side, count);
    __asm__ volatile("s_waitcnt vmcnt(0)" ::: "memory");
    const uint32_t old = __hip_atomic_fetch_add(flags + node, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (old == 0) return;  // the sibling finishes this node
    const uint32_t* other = reinterpret_cast<const uint32_t*>(nbox) + (size_t)(2 * node + (side ^ 1u)) * 6;
    lox = rk::fmn(lox, __uint_as_float(rk::atomic_load_relaxed(other + 0)));
    loy = rk::fmn(loy, __uint_as_float(rk::atomic_load_relaxed(other + 1)));
    loz = rk::fmn(loz, __uint_as_float(rk::atomic_load_relaxed(other + 2)));
    hix = rk::fmx(hix, __uint_as_float(rk::atomic_load_relaxed(other + 3)));
    hiy = rk::fmx(hiy, __uint_as_float(rk::atomic_load_relaxed(other + 4)));
    hiz = rk::fmx(hiz, __uint_as_float(rk::atomic_load_relaxed(other + 5)));
    count += rk::atomic_load_relaxed(ncount + 2 * node + (side ^ 1u));
    if (node == 0) return;
    ref = (count <= leaf_max) ? (kLeafBit | (range_first[node] << 4) | (count - 1)) : node;
    self = node;
    node = parent_node[node];
  }
}

__global__ __launch_bounds__(256) void k_fast_tris(const uint4* __restrict__ faces, const float4* __restrict__ verts,
                                                   const uint32_t* __restrict__ order, uint32_t n,
                                                   float4* __restrict__ tris) {
  uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  uint32_t fid = order[i];
  uint4 f = faces[fid];
  float4 v0 = verts[f.x];
  float4 e1 = rk::sub4(verts[f.y], v0);
  float4 e2 = rk::sub4(verts[f.z], v0);
  tris[3 * i + 0] = make_float4(v0.x, v0.y, v0.z, __uint_as_float(fid));
  tris[3 * i + 1] = make_float4(e1.x, e1.y, e1.z, 0.0f);
  tris[3 * i + 2] = make_float4(e2.x, e2.y, e2.z, 0.0f);
}

// ---- trace -------------------------------------------------------------------

constexpr int kStack = 64;

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

template <bool kCount>
__global__ __launch_bounds__(256) void k_fast_trace(const FNode* __restrict__ nodes, const float4* __restrict__ tris,
                                                    uint32_t n_tris, const float4* __restrict__ origins,
                                                    const float4* __restrict__ dirs, uint32_t n_rays,
                                                    uint32_t* __restrict__ faces_id, float2* __restrict__ coords,
                                                    float* __restrict__ depths, float eps,
                                                    unsigned long long* __restrict__ counters) {
  uint32_t n_box = 0, n_tri = 0;
  const uint32_t n_nodes = n_tris > 1 ? n_tris - 1 : 1;
  const uint32_t max_visits = 4 * n_tris + 64;  // a tree visit touches each node once
  const uint32_t l = blockIdx.x * 256 + threadIdx.x;
  if (l >= n_rays) return;
  const float4 orig = origins[l], dir = dirs[l];
  float depth = depths[l];
  uint32_t best = 0xFFFFFFFFu, best_oid = 0xFFFFFFFFu;
  float bu = 0.0f, bv = 0.0f;
  const float ix = safe_inv(dir.x), iy = safe_inv(dir.y), iz = safe_inv(dir.z);
  const float ox = orig.x * ix, oy = orig.y * iy, oz = orig.z * iz;
  uint32_t stack_ref[kStack];
  float stack_t[kStack];
  int sp = 0;
  uint32_t cur = 0;  // root
  uint32_t visits = 0;
  while (++visits <= max_visits) {
    if (!(cur & kLeafBit)) {
      if (cur >= n_nodes) break;  // never taken for a tree built by rk_fast_build
      const FNode nd = nodes[cur];
      const float a0 = __builtin_fmaf(nd.a.x, ix, -ox), a1 = __builtin_fmaf(nd.a.y, ix, -ox);
      const float a2 = __builtin_fmaf(nd.a.z, iy, -oy), a3 = __builtin_fmaf(nd.a.w, iy, -oy);
      const float a4 = __builtin_fmaf(nd.c.x, iz, -oz), a5 = __builtin_fmaf(nd.c.y, iz, -oz);
      const float b0 = __builtin_fmaf(nd.b.x, ix, -ox), b1 = __builtin_fmaf(nd.b.y, ix, -ox);
      const float b2 = __builtin_fmaf(nd.b.z, iy, -oy), b3 = __builtin_fmaf(nd.b.w, iy, -oy);
      const float b4 = __builtin_fmaf(nd.c.z, iz, -oz), b5 = __builtin_fmaf(nd.c.w, iz, -oz);
      const float tmin0 = fmaxf(fmaxf(fminf(a0, a1), fminf(a2, a3)), fmaxf(fminf(a4, a5), 0.0f));
      const float tmax0 = fminf(fminf(fmaxf(a0, a1), fmaxf(a2, a3)), fminf(fmaxf(a4, a5), depth));
      const float tmin1 = fmaxf(fmaxf(fminf(b0, b1), fminf(b2, b3)), fmaxf(fminf(b4, b5), 0.0f));
      const float tmax1 = fminf(fminf(fmaxf(b0, b1), fmaxf(b2, b3)), fminf(fmaxf(b4, b5), depth));
      const bool h0 = tmin0 <= tmax0, h1 = tmin1 <= tmax1;
      if (kCount) n_box += 2;
      if (h0 && h1) {
        const bool first0 = tmin0 <= tmin1;
        const uint32_t nearr = first0 ? nd.d.x : nd.d.y, farr = first0 ? nd.d.y : nd.d.x;
        const float fart = first0 ? tmin1 : tmin0;
        if (sp < kStack) { stack_ref[sp] = farr; stack_t[sp] = fart; ++sp; }
        cur = nearr;
        continue;
      }
      if (h0) { cur = nd.d.x; continue; }
      if (h1) { cur = nd.d.y; continue; }
    } else {
      const uint32_t first = (cur & ~kLeafBit) >> 4;
      const uint32_t cnt = first < n_tris ? min((cur & 15u) + 1, n_tris - first) : 0u;
      for (uint32_t k = 0; k < cnt; ++k) {
        const float4 r0 = tris[3 * (first + k) + 0];
        float4 v0 = r0;
        v0.w = 0.0f;
        const float4 e1 = tris[3 * (first + k) + 1], e2 = tris[3 * (first + k) + 2];
        float t, u, v;
        if (kCount) ++n_tri;
        if (tri_ray(orig, dir, v0, e1, e2, eps, t, u, v)) {
          const uint32_t oid = __float_as_uint(r0.w);
          if (t < depth || (best != 0xFFFFFFFFu && t == depth && oid < best_oid)) {
            depth = t; best = first + k; best_oid = oid; bu = u; bv = v;
          }
        }
      }
    }
    // pop the nearest pending entry that can still hold a closer hit
    bool found = false;
    while (sp > 0) {
      --sp;
      if (stack_t[sp] <= depth) { cur = stack_ref[sp]; found = true; break; }
    }
    if (!found) break;
  }
  if (kCount) {
    atomicAdd(counters, (unsigned long long)n_box);
    atomicAdd(counters + 1, (unsigned long long)n_tri);
  }
  if (best != 0xFFFFFFFFu) {
    faces_id[l] = best;
    coords[l] = make_float2(bu, bv);
    depths[l] = depth;
  }
}

inline unsigned blocks_for(uint32_t n, uint32_t per) { return (n + per - 1) / per; }

}  // namespace

extern "C" size_t rk_fast_build_workspace_size(uint32_t face_count, int morton_mode) {
  size_t n = face_count ? face_count : 1;
  size_t kb = morton_mode == RK_MORTON_63 ? 8 : 4;
  size_t b = 0;
  b += rk_align_up(n * kb);        // keys
  b += rk_align_up(2 * n * 4);     // child
  b += rk_align_up(n * 4) * 3;     // range_first, parent_node, parent_leaf
  b += rk_align_up(n * 4);         // flags
  b += rk_align_up(2 * n * 6 * 4); // child boxes
  b += rk_align_up(2 * n * 4);     // child counts
  b += rk_sort_workspace_size((uint32_t)n, (uint32_t)kb);
  return b + 4096;
}

extern "C" rk_status rk_fast_build(rk_stream stream, const rk_uint4* faces, const rk_float4* vertices,
                                   uint32_t face_count, const rk_aabox* hitbox, int morton_mode, uint32_t leaf_max,
                                   const rk_fast_bvh* bvh, uint32_t* info, void* ws, size_t ws_bytes) {
  const uint32_t n = face_count;
  if (n == 0 || !faces || !vertices || !hitbox || !bvh || !info) return RK_ERROR_INVALID_VALUE;
  if (!bvh->nodes || !bvh->tris || !bvh->indices || !bvh->boxes) return RK_ERROR_INVALID_VALUE;
  if (morton_mode != RK_MORTON_30 && morton_mode != RK_MORTON_63) return RK_ERROR_INVALID_VALUE;
  if (leaf_max < 1 || leaf_max > kMaxLeaf) return RK_ERROR_INVALID_VALUE;
  if (n >= (1u << 27)) return RK_ERROR_UNSUPPORTED;
  hipStream_t st = rk_hip_stream(stream);
  const bool wide = morton_mode == RK_MORTON_63;
  RkArena ar(ws, ws_bytes);
  void* keys = wide ? (void*)ar.take<uint64_t>(n) : (void*)ar.take<uint32_t>(n);
  uint32_t* child = ar.take<uint32_t>(2 * (size_t)n);
  uint32_t* range_first = ar.take<uint32_t>(n);
  uint32_t* parent_node = ar.take<uint32_t>(n);
  uint32_t* parent_leaf = ar.take<uint32_t>(n);
  uint32_t* flags = ar.take<uint32_t>(n);
  float* nbox = ar.take<float>(2 * (size_t)n * 6);
  uint32_t* ncount = ar.take<uint32_t>(2 * (size_t)n);
  size_t sort_bytes = rk_sort_workspace_size(n, wide ? 8 : 4);
  void* sort_ws = ar.take<char>(sort_bytes);
  if (!keys || !child || !range_first || !parent_node || !parent_leaf || !flags || !nbox || !ncount || !sort_ws)
    return RK_ERROR_WORKSPACE;
  rk_status s;
  if ((s = rk_fill_u32(stream, info, RK_INFO_WORDS, 0u)) != RK_SUCCESS) return s;
  if ((s = rk_fill_u32(stream, flags, n, 0u)) != RK_SUCCESS) return s;
  float4 rmin = make_float4(hitbox->min.x, hitbox->min.y, hitbox->min.z, hitbox->min.w);
  float4 rmax = make_float4(hitbox->max.x, hitbox->max.y, hitbox->max.z, hitbox->max.w);
  const uint4* f = reinterpret_cast<const uint4*>(faces);
  const float4* v = reinterpret_cast<const float4*>(vertices);
  uint32_t* order = bvh->indices;
  if (wide) {
    hipLaunchKernelGGL(k_fast_prep<uint64_t>, dim3(blocks_for(n, 256)), dim3(256), 0, st, f, v, n, rmin, rmax,
                       reinterpret_cast<Box*>(bvh->boxes), static_cast<uint64_t*>(keys), order);
    RK_CHECK_LAUNCH();
    s = rk_radix_sort_impl<uint64_t>(st, static_cast<uint64_t*>(keys), order, n, 0, 8, sort_ws, sort_bytes,
                                     info + RK_INFO_FLAGS);
    if (s != RK_SUCCESS) return s;
    if (n > 1) {
      hipLaunchKernelGGL(k_karras<uint64_t>, dim3(blocks_for(n - 1, 256)), dim3(256), 0, st,
                         static_cast<const uint64_t*>(keys), n, child, range_first, parent_node, parent_leaf);
      RK_CHECK_LAUNCH();
    }
  } else {
    hipLaunchKernelGGL(k_fast_prep<uint32_t>, dim3(blocks_for(n, 256)), dim3(256), 0, st, f, v, n, rmin, rmax,
                       reinterpret_cast<Box*>(bvh->boxes), static_cast<uint32_t*>(keys), order);
    RK_CHECK_LAUNCH();
    s = rk_radix_sort_impl<uint32_t>(st, static_cast<uint32_t*>(keys), order, n, 0, 4, sort_ws, sort_bytes,
                                     info + RK_INFO_FLAGS);
    if (s != RK_SUCCESS) return s;
    if (n > 1) {
      hipLaunchKernelGGL(k_karras<uint32_t>, dim3(blocks_for(n - 1, 256)), dim3(256), 0, st,
                         static_cast<const uint32_t*>(keys), n, child, range_first, parent_node, parent_leaf);
      RK_CHECK_LAUNCH();
    }
  }
  hipLaunchKernelGGL(k_fast_refit, dim3(blocks_for(n, 256)), dim3(256), 0, st,
                     reinterpret_cast<const Box*>(bvh->boxes), order, n, child, range_first, parent_node, parent_leaf,
                     flags, nbox, ncount, static_cast<FNode*>(bvh->nodes), leaf_max);
  RK_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_fast_tris, dim3(blocks_for(n, 256)), dim3(256), 0, st, f, v, order, n,
                     static_cast<float4*>(bvh->tris));
  RK_CHECK_LAUNCH();
  return RK_SUCCESS;
}

static inline float det_eps_of(const rk_trace_options* o) {
  return o ? o->det_epsilon : RK_DET_EPSILON_REFERENCE;
}

extern "C" rk_status rk_fast_trace(rk_stream stream, const rk_fast_bvh* bvh, const rk_float4* origins,
                                   const rk_float4* directions, uint32_t n_rays, uint32_t* faces_id, rk_float2* coords,
                                   float* depths, const rk_trace_options* options) {
  if (!bvh || !bvh->nodes || !bvh->tris || bvh->face_count == 0) return RK_ERROR_INVALID_VALUE;
  if (n_rays == 0) return RK_SUCCESS;
  if (!origins || !directions || !faces_id || !coords || !depths) return RK_ERROR_INVALID_VALUE;
  hipLaunchKernelGGL(k_fast_trace<false>, dim3(blocks_for(n_rays, 256)), dim3(256), 0, rk_hip_stream(stream),
                     static_cast<const FNode*>(bvh->nodes), static_cast<const float4*>(bvh->tris), bvh->face_count,
                     reinterpret_cast<const float4*>(origins), reinterpret_cast<const float4*>(directions), n_rays,
                     faces_id, reinterpret_cast<float2*>(coords), depths, det_eps_of(options),
                     (unsigned long long*)nullptr);
  RK_CHECK_LAUNCH();
  return RK_SUCCESS;
}

extern "C" rk_status rk_fast_trace_counted(rk_stream stream, const rk_fast_bvh* bvh, const rk_float4* origins,
                                           const rk_float4* directions, uint32_t n_rays, uint32_t* faces_id,
                                           rk_float2* coords, float* depths, const rk_trace_options* options,
                                           uint64_t* counters) {
  if (!bvh || !bvh->nodes || !bvh->tris || bvh->face_count == 0 || !counters) return RK_ERROR_INVALID_VALUE;
  if (n_rays == 0) return RK_SUCCESS;
  if (!origins || !directions || !faces_id || !coords || !depths) return RK_ERROR_INVALID_VALUE;
  hipLaunchKernelGGL(k_fast_trace<true>, dim3(blocks_for(n_rays, 256)), dim3(256), 0, rk_hip_stream(stream),
                     static_cast<const FNode*>(bvh->nodes), static_cast<const float4*>(bvh->tris), bvh->face_count,
                     reinterpret_cast<const float4*>(origins), reinterpret_cast<const float4*>(directions), n_rays,
                     faces_id, reinterpret_cast<float2*>(coords), depths, det_eps_of(options),
                     reinterpret_cast<unsigned long long*>(counters));
  RK_CHECK_LAUNCH();
  return RK_SUCCESS;
}
