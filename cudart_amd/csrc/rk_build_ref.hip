// The reference LBVH build (CUDA_buildMeshTreeLBVH, CUDA_buildMeshTreeLBVH.cu:457-681),
// rebuilt as a device-resident pipeline for gfx950:
//   k_boxes_morton   per-face box + reference 32-bit Morton code (A6, A7)
//   radix sort       (code, face) pairs with the reference tie order (A8-A10)
//   k_split_levels   level of each adjacent-code split (A11)
//   radix sort       splits grouped by level, ascending within a level (A12)
//   k_topo_ref       the host recursion of :309-455, level-synchronous (A13, A14)
//   k_leaf_boxes     leaf boxes, one wave per leaf (A15, pass 0)
//   k_refit_ref      internal boxes with the z-extent rule (A15, passes 1..)
//   reorder          faces'[i] = faces[indices[i]] (A16)
// No step goes through the host: the reference's device->host copies
// (:428, :651) and its host recursion are gone.
#include <algorithm>

#include "rk_device.h"

template <typename K>
rk_status rk_radix_sort_impl(hipStream_t st, K* keys, uint32_t* vals, uint32_t n, int first_pass, int passes,
                             void* ws, size_t ws_bytes, uint32_t* err, bool head_cleared = false,
                             const uint32_t** digit_counts = nullptr);
rk_status rk_sort_general_impl(hipStream_t st, uint32_t* keys, uint32_t* indices, uint32_t length,
                               uint32_t bit_depth, int init_indices, void* ws, size_t ws_bytes, uint32_t* err);

namespace {

using rk::Box;
using rk::Node;

__device__ __forceinline__ Box face_box(const uint4* __restrict__ faces, const float4* __restrict__ verts, uint32_t k) {
  uint4 f = faces[k];
  float4 v0 = verts[f.x], v1 = verts[f.y], v2 = verts[f.z];
  Box b;  // CUDA_calcMeshBoxes.cu:53-55: reset(v0), merge(v1), merge(v2)
  b.mn = rk::min4(rk::min4(v0, v1), v2);
  b.mx = rk::max4(rk::max4(v0, v1), v2);
  return b;
}

__device__ __forceinline__ uint32_t spread10(uint32_t v) {  // bit k -> bit 3k, k < 10
  v &= 0x3FFu;
  v = (v | (v << 16)) & 0x030000FFu;
  v = (v | (v << 8)) & 0x0300F00Fu;
  v = (v | (v << 4)) & 0x030C30C3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}
__device__ __forceinline__ uint64_t spread21(uint64_t v) {  // bit k -> bit 3k, k < 21
  v &= 0x1FFFFFull;
  v = (v | (v << 32)) & 0x1F00000000FFFFull;
  v = (v | (v << 16)) & 0x1F0000FF0000FFull;
  v = (v | (v << 8)) & 0x100F00F00F00F00Full;
  v = (v | (v << 4)) & 0x10C30C30C30C30C3ull;
  v = (v | (v << 2)) & 0x1249249249249249ull;
  return v;
}

// kernel_calcMortonCodes, CUDA_buildMeshTreeLBVH.cu:115-173.
__device__ __forceinline__ uint32_t morton_ref32(const Box& b, float4 rmin, float4 rmax) {
  float4 len = rk::sub4(rmax, rmin);
  float sx = 2047.0f / len.x, sy = 2047.0f / len.y, sz = 1023.0f / len.z;
  float4 c = rk::muls4(rk::add4(b.mn, b.mx), 0.5f);
  uint32_t cx = rk::f2u_sat((c.x - rmin.x) * sx);
  uint32_t cy = rk::f2u_sat((c.y - rmin.y) * sy);
  uint32_t cz = rk::f2u_sat((c.z - rmin.z) * sz);
  // y bit k -> 3k, x bit k -> 3k+1 (k <= 10), z bit k -> 3k+2 (k <= 9); bit 10
  // of y and x land on bits 30 and 31.
  return spread10(cy) | (spread10(cx) << 1) | (spread10(cz) << 2) | (((cy >> 10) & 1u) << 30) |
         (((cx >> 10) & 1u) << 31);
}
// 30- and 63-bit codes (no reference counterpart): q = min((uint)((c-min)*(2^b/len)), 2^b-1),
// x in the highest bit of each triple.
__device__ __forceinline__ uint32_t morton30(const Box& b, float4 rmin, float4 rmax) {
  float4 len = rk::sub4(rmax, rmin);
  float sx = 1024.0f / len.x, sy = 1024.0f / len.y, sz = 1024.0f / len.z;
  float4 c = rk::muls4(rk::add4(b.mn, b.mx), 0.5f);
  uint32_t qx = min(rk::f2u_sat((c.x - rmin.x) * sx), 1023u);
  uint32_t qy = min(rk::f2u_sat((c.y - rmin.y) * sy), 1023u);
  uint32_t qz = min(rk::f2u_sat((c.z - rmin.z) * sz), 1023u);
  return (spread10(qx) << 2) | (spread10(qy) << 1) | spread10(qz);
}
__device__ __forceinline__ uint64_t morton63(const Box& b, float4 rmin, float4 rmax) {
  float4 len = rk::sub4(rmax, rmin);
  float sx = 2097152.0f / len.x, sy = 2097152.0f / len.y, sz = 2097152.0f / len.z;
  float4 c = rk::muls4(rk::add4(b.mn, b.mx), 0.5f);
  uint32_t qx = min(rk::f2u_sat((c.x - rmin.x) * sx), 2097151u);
  uint32_t qy = min(rk::f2u_sat((c.y - rmin.y) * sy), 2097151u);
  uint32_t qz = min(rk::f2u_sat((c.z - rmin.z) * sz), 2097151u);
  return (spread21(qx) << 2) | (spread21(qy) << 1) | spread21(qz);
}

__global__ __launch_bounds__(256) void k_calc_boxes(const uint4* __restrict__ faces, const float4* __restrict__ verts,
                                                    uint32_t n, Box* __restrict__ boxes, uint32_t* __restrict__ indices) {
  uint32_t k = blockIdx.x * 256 + threadIdx.x;
  if (k >= n) return;
  Box b = face_box(faces, verts, k);
  float4* o = reinterpret_cast<float4*>(boxes + k);
  o[0] = b.mn;
  o[1] = b.mx;
  indices[k] = k;
}

__global__ __launch_bounds__(256) void k_boxes_morton32(const uint4* __restrict__ faces,
                                                        const float4* __restrict__ verts, uint32_t n, float4 rmin,
                                                        float4 rmax, Box* __restrict__ boxes,
                                                        uint32_t* __restrict__ indices, uint32_t* __restrict__ codes) {
  uint32_t k = blockIdx.x * 256 + threadIdx.x;
  if (k >= n) return;
  Box b = face_box(faces, verts, k);
  float4* o = reinterpret_cast<float4*>(boxes + k);
  o[0] = b.mn;
  o[1] = b.mx;
  indices[k] = k;
  codes[k] = morton_ref32(b, rmin, rmax);
}

__global__ __launch_bounds__(256) void k_morton(const Box* __restrict__ boxes, uint32_t n, float4 rmin, float4 rmax,
                                                int mode, void* __restrict__ codes) {
  uint32_t k = blockIdx.x * 256 + threadIdx.x;
  if (k >= n) return;
  Box b = rk::load_box(boxes + k);
  if (mode == RK_MORTON_REF32) static_cast<uint32_t*>(codes)[k] = morton_ref32(b, rmin, rmax);
  else if (mode == RK_MORTON_30) static_cast<uint32_t*>(codes)[k] = morton30(b, rmin, rmax);
  else static_cast<uint64_t*>(codes)[k] = morton63(b, rmin, rmax);
}

// kernel_findSplits, CUDA_buildMeshTreeLBVH.cu:175-233: for adjacent sorted
// codes with code[s]^code[s+1] != 0, level = clz (0 = bit 31). Here every
// position gets a key (32 = no split) so one stable 8-bit pass groups the
// splits by level, ascending within a level.
// The per-level counts come from the digit counts of that pass (its first
// 33 digits), except for a single split, which no sort runs on: then this
// kernel counts it (level_count, may be NULL otherwise).
__global__ __launch_bounds__(256) void k_split_levels(const uint32_t* __restrict__ codes, uint32_t n,
                                                      uint32_t* __restrict__ lev, uint32_t* __restrict__ pos,
                                                      uint32_t* __restrict__ level_count) {
  uint32_t s = blockIdx.x * 256 + threadIdx.x;
  if (s + 1 < n) {
    uint32_t diff = codes[s] ^ codes[s + 1];
    uint32_t level = diff ? (uint32_t)__clz((int)diff) : 32u;
    lev[s] = level;
    pos[s] = s + 1;
    if (level_count) atomicAdd(&level_count[level], 1u);
  }
}

constexpr uint32_t kTopoThreads = 1024;
constexpr uint32_t kMaxLevels = 96;
constexpr uint32_t kKindMedian = 2u;  // rec flag bit 1: median-split subtree
constexpr uint32_t kInternal = 1u;    // rec flag bit 0: internal node

struct TopoWs {
  uint32_t* ra;      // range start (sorted position)
  uint32_t* rb;      // range end
  uint32_t* rflag;   // kInternal | kKindMedian | level << 8
  uint32_t* rchild;  // record index of the left child (right = +1)
  uint32_t* rcnt;    // internal nodes in the subtree
  uint32_t* rrank;   // preorder rank among internal nodes
  uint32_t* rid;     // node id
  uint32_t* lev_begin;  // kMaxLevels + 1 record offsets
  uint32_t* leaf_ids;   // node ids of leaves
  uint32_t cap;
};

__device__ __forceinline__ uint32_t upper_bound(const uint32_t* __restrict__ A, uint32_t lo, uint32_t hi, uint32_t a) {
  while (lo < hi) {
    uint32_t mid = (lo + hi) >> 1;
    if (A[mid] <= a) lo = mid + 1; else hi = mid;
  }
  return lo;
}

// Where the reference's level-L split pointer stops when every level-L split
// is <= a: it walks on through the (ascending) lists of levels L+1, L+2, ...
// and stops at the first entry > a (CUDA_buildMeshTreeLBVH.cu:381-392).
__device__ uint32_t chain_walk(const uint32_t* __restrict__ A, const uint32_t* s_idx, uint32_t L, uint32_t a,
                               uint32_t S) {
  for (uint32_t k = L + 1; k < 32; ++k) {
    uint32_t lo = s_idx[k], hi = s_idx[k + 1];
    if (hi > lo && A[hi - 1] > a) return upper_bound(A, lo, hi, a);
  }
  return S;
}

__device__ __forceinline__ uint32_t block_excl_scan1024(uint32_t v, uint32_t* s_tmp /*16*/, uint32_t* total) {
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint32_t incl = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    uint32_t y = __shfl_up(incl, off, 64);
    if (lane >= (uint32_t)off) incl += y;
  }
  if (lane == 63) s_tmp[wave] = incl;
  __syncthreads();
  uint32_t wprefix = 0, tot = 0;
  for (uint32_t w = 0; w < kTopoThreads / 64; ++w) {
    uint32_t c = s_tmp[w];
    wprefix += (w < wave) ? c : 0u;
    tot += c;
  }
  __syncthreads();
  *total = tot;
  return wprefix + incl - v;
}

// The reference topology (createNode/createLeaf, CUDA_buildMeshTreeLBVH.cu:309-455),
// one workgroup, level by level. Nodes of a level are kept in ascending range
// order, which is the order the reference's depth-first recursion visits them,
// so the state of each level's split pointer is a function of the range start:
//  - a node with a level-L split > a (prefix of the level) stops at the first
//    one; it splits there if that split is < b;
//  - for the remaining nodes (suffix) the pointer has left level L: the first
//    such node walks into the deeper lists (chain_walk); if it stops exactly at
//    the start of level L+1's list, every later suffix node meets the entry test
//    ptr == idx[L+1] (:375) and is a leaf; otherwise each later node stops at
//    chain_walk(a).
// Preorder ranks follow from per-subtree internal-node counts; an internal node
// of rank r has children 2r+1, 2r+2 (:325-326, :394-395).
__global__ __launch_bounds__(1024) void k_topo_ref(const uint32_t* __restrict__ A,
                                                   const uint32_t* __restrict__ level_count, uint32_t n, uint32_t d,
                                                   int complete, TopoWs w, Node* __restrict__ nodes,
                                                   uint32_t* __restrict__ info) {
  __shared__ uint32_t s_idx[34];
  __shared__ uint32_t s_lev[kMaxLevels + 2];
  __shared__ uint32_t s_tmp[16];
  __shared__ uint32_t s_j0, s_q0, s_flags, s_maxleaf, s_nleaves;
  const uint32_t tid = threadIdx.x;
  if (tid == 0) {
    s_idx[0] = 0;
    for (int k = 0; k < 33; ++k) s_idx[k + 1] = s_idx[k] + level_count[k];
    s_flags = 0;
    s_maxleaf = 0;
    s_nleaves = 0;
    w.ra[0] = 0;
    w.rb[0] = n;
    w.rflag[0] = 0;
    s_lev[0] = 0;
  }
  __syncthreads();
  const uint32_t S = s_idx[32];
  const uint32_t past_depth = s_idx[d + 1 < 33 ? d + 1 : 32];
  uint32_t cur = 0, end = 1, L = 0;
  while (cur < end) {
    if (L >= kMaxLevels) {
      if (tid == 0) s_flags |= 0x80000000u;
      break;
    }
    const uint32_t cntL = L < 32 ? level_count[L] : 0u;
    const uint32_t EL = L < 32 ? s_idx[L + 1] : S;
    const uint32_t lastA = cntL ? A[EL - 1] : 0u;
    if (tid == 0) s_j0 = 0xFFFFFFFFu;
    __syncthreads();
    for (uint32_t r = cur + tid; r < end; r += kTopoThreads) {
      uint32_t a = w.ra[r], b = w.rb[r], f = w.rflag[r];
      bool walker = !(f & kKindMedian) && (b - a) >= 16 && L < d;
      if (walker && (cntL == 0 || a >= lastA)) atomicMin(&s_j0, r);
    }
    __syncthreads();
    if (tid == 0) {
      uint32_t q0 = S;
      if (s_j0 != 0xFFFFFFFFu && cntL != 0) q0 = chain_walk(A, s_idx, L, w.ra[s_j0], S);
      s_q0 = q0;
    }
    __syncthreads();
    const uint32_t j0 = s_j0, q0 = s_q0;
    uint32_t out = end;
    for (uint32_t chunk = cur; chunk < end; chunk += kTopoThreads) {
      const uint32_t r = chunk + tid;
      const bool in = r < end;
      uint32_t internal = 0, split = 0, ckind = 0, a = 0, b = 0, f = 0;
      if (in) {
        a = w.ra[r]; b = w.rb[r]; f = w.rflag[r];
        const uint32_t size = b - a;
        if (f & kKindMedian) {  // createNode(median), :314-339
          if (size >= 224) { internal = 1; split = (a + b) >> 1; ckind = kKindMedian; }
        } else {                 // createNode(split), :360-411
          bool leafish = true;
          if (size >= 16 && L < d) {
            uint32_t q = S;
            bool entry_leaf = false;
            if (cntL == 0) entry_leaf = true;
            else if (a < lastA) q = upper_bound(A, s_idx[L], EL, a);
            else if (r == j0) q = q0;
            else if (q0 == EL) entry_leaf = true;
            else q = chain_walk(A, s_idx, L, a, S);
            if (!entry_leaf) {
              if (q >= S) {
                atomicOr(&s_flags, (uint32_t)RK_FLAG_WALK_PAST_END);
              } else {
                if (q >= past_depth) atomicOr(&s_flags, (uint32_t)RK_FLAG_WALK_PAST_DEPTH);
                uint32_t sp = A[q];
                if (sp < b) { leafish = false; internal = 1; split = sp; ckind = 0; }
              }
            }
          }
          // createLeaf, :341-358: a complete tree splits oversized leaves at the median.
          if (leafish && complete && size > 224) { internal = 1; split = (a + b) >> 1; ckind = kKindMedian; }
        }
        if (!internal) atomicMax(&s_maxleaf, L);
      }
      uint32_t total;
      uint32_t excl = block_excl_scan1024(internal * 2u, s_tmp, &total);
      if (in) {
        w.rflag[r] = (f & kKindMedian) | internal | (L << 8);
        if (internal) {
          uint32_t pos = out + excl;
          if (pos + 2 <= w.cap) {
            w.rchild[r] = pos;
            w.ra[pos] = a; w.rb[pos] = split; w.rflag[pos] = ckind;
            w.ra[pos + 1] = split; w.rb[pos + 1] = b; w.rflag[pos + 1] = ckind;
          } else {
            atomicOr(&s_flags, 0x40000000u);
          }
        }
      }
      out += total;
      __syncthreads();
    }
    cur = end;
    end = out < w.cap ? out : w.cap;
    ++L;
    if (tid == 0) s_lev[L] = cur;
    __syncthreads();
  }
  const uint32_t nlev = L;
  const uint32_t R = end;
  if (tid == 0) s_lev[nlev] = R;
  __syncthreads();
  // Internal-node counts per subtree, bottom-up.
  for (int lv = (int)nlev - 1; lv >= 0; --lv) {
    for (uint32_t r = s_lev[lv] + tid; r < s_lev[lv + 1]; r += kTopoThreads) {
      uint32_t c = 0;
      if (w.rflag[r] & kInternal) {
        uint32_t lc = w.rchild[r];
        c = 1 + ((w.rflag[lc] & kInternal) ? w.rcnt[lc] : 0u) + ((w.rflag[lc + 1] & kInternal) ? w.rcnt[lc + 1] : 0u);
      }
      w.rcnt[r] = c;
    }
    __syncthreads();
  }
  // Preorder ranks and ids, top-down.
  if (tid == 0) { w.rrank[0] = 0; w.rid[0] = 0; }
  __syncthreads();
  for (uint32_t lv = 0; lv < nlev; ++lv) {
    for (uint32_t r = s_lev[lv] + tid; r < s_lev[lv + 1]; r += kTopoThreads) {
      if (w.rflag[r] & kInternal) {
        uint32_t R0 = w.rrank[r], lc = w.rchild[r];
        w.rid[lc] = 2 * R0 + 1;
        w.rid[lc + 1] = 2 * R0 + 2;
        if (w.rflag[lc] & kInternal) w.rrank[lc] = R0 + 1;
        if (w.rflag[lc + 1] & kInternal) w.rrank[lc + 1] = R0 + 1 + ((w.rflag[lc] & kInternal) ? w.rcnt[lc] : 0u);
      }
    }
    __syncthreads();
  }
  // Emit bvnodes (make_node / make_leaf, :85-109) with empty boxes.
  for (uint32_t r = tid; r < R; r += kTopoThreads) {
    uint32_t f = w.rflag[r], id = w.rid[r];
    Node nd;
    if (f & kInternal) {
      nd.w0 = (((f >> 8) % 3u) & 3u) << 1;
      nd.w1 = 2 * w.rrank[r] + 1;
    } else {
      uint32_t a = w.ra[r], b = w.rb[r];
      nd.w0 = 1u | ((b - a) << 1);
      nd.w1 = a;
      uint32_t li = atomicAdd(&s_nleaves, 1u);
      w.leaf_ids[li] = id;
    }
    nd.bmin[0] = nd.bmin[1] = nd.bmin[2] = RK_HUGE;
    nd.bmax[0] = nd.bmax[1] = nd.bmax[2] = -RK_HUGE;
    nodes[id] = nd;
  }
  for (uint32_t k = tid; k <= nlev; k += kTopoThreads) w.lev_begin[k] = s_lev[k];
  __syncthreads();
  if (tid == 0) {
    uint32_t flags = s_flags & 0xFFFFu;
    if (R > n) flags |= RK_FLAG_NODES_GT_N;
    // The reference's kernel_bitonicMerge2x writes past the end of its input
    // (CUDA_bitonicSortGeneral.cu:254, CUDA_bitonicSort.cu:196) when the last
    // pair of 128-bins is partial: in the Morton sort (512 < n <= 32768) and
    // in the per-level value sorts of levels 1..d (128 < count <= 32768).
    auto merge2x_oob = [](uint32_t c) { return ((c + 127) / 128) % 2 == 0 && (c % 128) != 0; };
    if (n > 512 && n <= 32768 && merge2x_oob(n)) flags |= RK_FLAG_MERGE2X_OOB;
    for (uint32_t k = 1; k <= d; ++k) {
      uint32_t c = level_count[k];
      if (c > 128 && c <= 32768 && merge2x_oob(c)) flags |= RK_FLAG_MERGE2X_OOB;
    }
    info[RK_INFO_NODE_COUNT] = R;
    info[RK_INFO_REAL_DEPTH] = s_maxleaf + 1;
    atomicOr(&info[RK_INFO_FLAGS], flags | (s_flags & 0xC0000000u));
    info[RK_INFO_SPLITS] = S;
    info[RK_INFO_LEAVES] = s_nleaves;
    info[5] = nlev;
  }
}

// Leaf boxes (kernel_calcBoundingBoxes pass 0, CUDA_buildMeshTreeLBVH.cu:246-268):
// merge of the leaf's triangle boxes via indices. One wave per leaf.
__global__ __launch_bounds__(256) void k_leaf_boxes(const uint32_t* __restrict__ leaf_ids, uint32_t n,
                                                    const uint32_t* __restrict__ info,
                                                    const uint32_t* __restrict__ indices, const Box* __restrict__ boxes,
                                                    Node* __restrict__ nodes) {
  const uint32_t nleaves = info[RK_INFO_LEAVES];
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t wave_global = (blockIdx.x * 256 + threadIdx.x) >> 6;
  const uint32_t nwaves = (gridDim.x * 256) >> 6;
  for (uint32_t li = wave_global; li < nleaves; li += nwaves) {
    const uint32_t id = leaf_ids[li];
    const uint32_t w0 = nodes[id].w0, first = nodes[id].w1;
    uint32_t cnt = rk::leaf_count(w0);
    if (first >= n) cnt = 0;
    else if (cnt > n - first) cnt = n - first;
    float mnx = RK_HUGE, mny = RK_HUGE, mnz = RK_HUGE, mxx = -RK_HUGE, mxy = -RK_HUGE, mxz = -RK_HUGE;
    // Faces lane, lane + 64, lane + 128, ... in that order, kLeafFetch at a
    // time with all their loads issued first: a face costs two dependent loads
    // (its index, then its box), and the leaves of an incomplete tree (the
    // Hybrid's first pass) hold thousands of faces.
    constexpr uint32_t kLeafFetch = 8;
    for (uint32_t i0 = lane; i0 < cnt; i0 += 64 * kLeafFetch) {
      uint32_t ix[kLeafFetch];
#pragma unroll
      for (uint32_t k = 0; k < kLeafFetch; ++k) ix[k] = i0 + 64 * k < cnt ? indices[first + i0 + 64 * k] : 0u;
#pragma unroll
      for (uint32_t k = 0; k < kLeafFetch; ++k) {
        if (i0 + 64 * k < cnt) {
          const Box b = rk::load_box(boxes + ix[k]);
          mnx = rk::fmn(mnx, b.mn.x); mny = rk::fmn(mny, b.mn.y); mnz = rk::fmn(mnz, b.mn.z);
          mxx = rk::fmx(mxx, b.mx.x); mxy = rk::fmx(mxy, b.mx.y); mxz = rk::fmx(mxz, b.mx.z);
        }
      }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      mnx = rk::fmn(mnx, __shfl_xor(mnx, off, 64)); mny = rk::fmn(mny, __shfl_xor(mny, off, 64));
      mnz = rk::fmn(mnz, __shfl_xor(mnz, off, 64)); mxx = rk::fmx(mxx, __shfl_xor(mxx, off, 64));
      mxy = rk::fmx(mxy, __shfl_xor(mxy, off, 64)); mxz = rk::fmx(mxz, __shfl_xor(mxz, off, 64));
    }
    if (lane == 0) {
      Node* nd = nodes + id;
      nd->bmin[0] = mnx; nd->bmin[1] = mny; nd->bmin[2] = mnz;
      nd->bmax[0] = mxx; nd->bmax[1] = mxy; nd->bmax[2] = mxz;
    }
  }
}

// Internal boxes, bottom-up by level in one workgroup. The fixpoint of the
// reference's real_depth racy passes (CUDA_buildMeshTreeLBVH.cu:269-301): a node
// takes the union of its children only if both have z-extent > 0, else it
// stays EMPTY (and so does every ancestor).
__global__ __launch_bounds__(1024) void k_refit_ref(const uint32_t* __restrict__ rflag,
                                                    const uint32_t* __restrict__ rid,
                                                    const uint32_t* __restrict__ lev_begin,
                                                    const uint32_t* __restrict__ info, Node* __restrict__ nodes) {
  const uint32_t nlev = info[5];
  const uint32_t M = info[RK_INFO_NODE_COUNT];
  for (int lv = (int)nlev - 1; lv >= 0; --lv) {
    const uint32_t b0 = lev_begin[lv], b1 = lev_begin[lv + 1];
    for (uint32_t r = b0 + threadIdx.x; r < b1; r += 1024) {
      if (!(rflag[r] & kInternal)) continue;
      Node* nd = nodes + rid[r];
      if (nd->w1 + 1 >= M) continue;
      const Node* l = nodes + nd->w1;
      const Node* q = nodes + nd->w1 + 1;
      if (l->bmax[2] - l->bmin[2] > 0.0f && q->bmax[2] - q->bmin[2] > 0.0f) {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
          nd->bmin[a] = rk::fmn(l->bmin[a], q->bmin[a]);
          nd->bmax[a] = rk::fmx(l->bmax[a], q->bmax[a]);
        }
      }
    }
    __syncthreads();
  }
}

__global__ __launch_bounds__(256) void k_gather_faces(const uint4* __restrict__ src, const uint32_t* __restrict__ idx,
                                                      uint4* __restrict__ dst, uint32_t n) {
  uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) dst[i] = src[idx[i]];
}
__global__ __launch_bounds__(256) void k_copy_faces(const uint4* __restrict__ src, uint4* __restrict__ dst, uint32_t n) {
  uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

inline unsigned blocks_for(uint32_t n, uint32_t per) { return (n + per - 1) / per; }

}  // namespace

extern "C" rk_status rk_calc_mesh_boxes(rk_stream stream, const rk_uint4* faces, const rk_float4* vertices,
                                        uint32_t count, rk_aabox* boxes, uint32_t* indices) {
  if (count == 0) return RK_SUCCESS;
  if (!faces || !vertices || !boxes || !indices) return RK_ERROR_INVALID_VALUE;
  hipLaunchKernelGGL(k_calc_boxes, dim3(blocks_for(count, 256)), dim3(256), 0, rk_hip_stream(stream),
                     reinterpret_cast<const uint4*>(faces), reinterpret_cast<const float4*>(vertices), count,
                     reinterpret_cast<Box*>(boxes), indices);
  RK_CHECK_LAUNCH();
  return RK_SUCCESS;
}

extern "C" rk_status rk_calc_morton_codes(rk_stream stream, const rk_aabox* boxes, uint32_t count,
                                          const rk_aabox* root_box, int mode, void* codes) {
  if (count == 0) return RK_SUCCESS;
  if (!boxes || !root_box || !codes || mode < 0 || mode > 2) return RK_ERROR_INVALID_VALUE;
  float4 rmin = make_float4(root_box->min.x, root_box->min.y, root_box->min.z, root_box->min.w);
  float4 rmax = make_float4(root_box->max.x, root_box->max.y, root_box->max.z, root_box->max.w);
  hipLaunchKernelGGL(k_morton, dim3(blocks_for(count, 256)), dim3(256), 0, rk_hip_stream(stream),
                     reinterpret_cast<const Box*>(boxes), count, rmin, rmax, mode, codes);
  RK_CHECK_LAUNCH();
  return RK_SUCCESS;
}

extern "C" size_t rk_reorder_workspace_size(uint32_t count) { return rk_align_up((size_t)(count ? count : 1) * 16) + 256; }

extern "C" rk_status rk_reorder_mesh_faces(rk_stream stream, rk_uint4* faces, const uint32_t* indices, uint32_t count,
                                           void* ws, size_t ws_bytes) {
  if (count == 0) return RK_SUCCESS;
  if (!faces || !indices) return RK_ERROR_INVALID_VALUE;
  RkArena ar(ws, ws_bytes);
  uint4* tmp = ar.take<uint4>(count);
  if (!tmp) return RK_ERROR_WORKSPACE;
  hipStream_t st = rk_hip_stream(stream);
  uint4* f = reinterpret_cast<uint4*>(faces);
  hipLaunchKernelGGL(k_gather_faces, dim3(blocks_for(count, 256)), dim3(256), 0, st, f, indices, tmp, count);
  RK_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_copy_faces, dim3(blocks_for(count, 256)), dim3(256), 0, st, tmp, f, count);
  RK_CHECK_LAUNCH();
  return RK_SUCCESS;
}

extern "C" size_t rk_build_lbvh_workspace_size(uint32_t face_count) {
  size_t n = face_count ? face_count : 1;
  size_t cap = 2 * n;
  size_t b = 0;
  b += rk_align_up(n * 4);            // codes
  b += rk_align_up(n * 4) * 2;        // level keys, positions
  b += rk_align_up(64 * 4);           // level counts
  b += rk_align_up(cap * 4) * 7;      // topology records
  b += rk_align_up((kMaxLevels + 2) * 4);
  b += rk_align_up(n * 4);            // leaf ids
  b += rk_align_up(n * 16);           // face reorder temp
  b += rk_sort_workspace_size((uint32_t)n, 4);
  return b + 4096;
}

extern "C" rk_status rk_build_lbvh(rk_stream stream, rk_uint4* faces, const rk_float4* vertices, uint32_t face_count,
                                   const rk_aabox* hitbox, int tree_complete, uint32_t tree_depth, rk_aabox* boxes,
                                   uint32_t* indices, rk_bvnode* nodes, uint32_t nodes_capacity, uint32_t* info,
                                   void* ws, size_t ws_bytes) {
  const uint32_t n = face_count;
  if (n == 0 || !faces || !vertices || !hitbox || !boxes || !indices || !nodes || !info) return RK_ERROR_INVALID_VALUE;
  if (tree_depth > 31) return RK_ERROR_INVALID_VALUE;
  if (n >= (1u << 30)) return RK_ERROR_UNSUPPORTED;
  if ((size_t)nodes_capacity < 2 * (size_t)n - 1) return RK_ERROR_INVALID_VALUE;
  hipStream_t st = rk_hip_stream(stream);
  const size_t cap = 2 * (size_t)n;
  RkArena ar(ws, ws_bytes);
  uint32_t* codes = ar.take<uint32_t>(n);
  uint32_t* lev = ar.take<uint32_t>(n);
  uint32_t* pos = ar.take<uint32_t>(n);
  uint32_t* level_count = ar.take<uint32_t>(64);
  TopoWs tw;
  tw.ra = ar.take<uint32_t>(cap); tw.rb = ar.take<uint32_t>(cap); tw.rflag = ar.take<uint32_t>(cap);
  tw.rchild = ar.take<uint32_t>(cap); tw.rcnt = ar.take<uint32_t>(cap); tw.rrank = ar.take<uint32_t>(cap);
  tw.rid = ar.take<uint32_t>(cap);
  tw.lev_begin = ar.take<uint32_t>(kMaxLevels + 2);
  tw.leaf_ids = ar.take<uint32_t>(n);
  tw.cap = (uint32_t)cap;
  uint4* ftmp = ar.take<uint4>(n);
  size_t sort_bytes = rk_sort_workspace_size(n, 4);
  void* sort_ws = ar.take<char>(sort_bytes);
  if (!codes || !lev || !pos || !level_count || !tw.ra || !tw.rb || !tw.rflag || !tw.rchild || !tw.rcnt ||
      !tw.rrank || !tw.rid || !tw.lev_begin || !tw.leaf_ids || !ftmp || !sort_ws)
    return RK_ERROR_WORKSPACE;

  rk_status s;
  if ((s = rk_fill_u32(stream, info, RK_INFO_WORDS, 0u)) != RK_SUCCESS) return s;
  if ((s = rk_fill_u32(stream, level_count, 64, 0u)) != RK_SUCCESS) return s;
  float4 rmin = make_float4(hitbox->min.x, hitbox->min.y, hitbox->min.z, hitbox->min.w);
  float4 rmax = make_float4(hitbox->max.x, hitbox->max.y, hitbox->max.z, hitbox->max.w);
  uint4* f = reinterpret_cast<uint4*>(faces);
  hipLaunchKernelGGL(k_boxes_morton32, dim3(blocks_for(n, 256)), dim3(256), 0, st, f,
                     reinterpret_cast<const float4*>(vertices), n, rmin, rmax, reinterpret_cast<Box*>(boxes), indices,
                     codes);
  RK_CHECK_LAUNCH();
  // CUDA_radixSortGeneral(codes, indices, N, 32, true), CUDA_buildMeshTreeLBVH.cu:522-528.
  s = rk_sort_general_impl(st, codes, indices, n, 32, 0, sort_ws, sort_bytes, info + RK_INFO_FLAGS);
  if (s != RK_SUCCESS) return s;
  const uint32_t* counts = level_count;  // zero when there is no split
  if (n > 1) {
    // one split is not sorted, so its level is counted here; more are counted
    // by the sort that groups them by level
    hipLaunchKernelGGL(k_split_levels, dim3(blocks_for(n - 1, 256)), dim3(256), 0, st, codes, n, lev, pos,
                       n == 2 ? level_count : (uint32_t*)nullptr);
    RK_CHECK_LAUNCH();
    const uint32_t* digit_counts = nullptr;
    s = rk_radix_sort_impl<uint32_t>(st, lev, pos, n - 1, 0, 1, sort_ws, sort_bytes, info + RK_INFO_FLAGS, false,
                                     &digit_counts);
    if (s != RK_SUCCESS) return s;
    if (digit_counts) counts = digit_counts;
  }
  hipLaunchKernelGGL(k_topo_ref, dim3(1), dim3(kTopoThreads), 0, st, pos, counts, n, tree_depth, tree_complete,
                     tw, reinterpret_cast<Node*>(nodes), info);
  RK_CHECK_LAUNCH();
  unsigned lb = blocks_for(n, 64 * 4);
  if (lb > 2048) lb = 2048;
  hipLaunchKernelGGL(k_leaf_boxes, dim3(lb), dim3(256), 0, st, tw.leaf_ids, n, info, indices,
                     reinterpret_cast<const Box*>(boxes), reinterpret_cast<Node*>(nodes));
  RK_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_refit_ref, dim3(1), dim3(1024), 0, st, tw.rflag, tw.rid, tw.lev_begin, info,
                     reinterpret_cast<Node*>(nodes));
  RK_CHECK_LAUNCH();
  if (tree_complete) {  // CUDA_reorderMeshFaces, CUDA_buildMeshTreeLBVH.cu:669-676
    hipLaunchKernelGGL(k_gather_faces, dim3(blocks_for(n, 256)), dim3(256), 0, st, f, indices, ftmp, n);
    RK_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_copy_faces, dim3(blocks_for(n, 256)), dim3(256), 0, st, ftmp, f, n);
    RK_CHECK_LAUNCH();
  }
  return RK_SUCCESS;
}
