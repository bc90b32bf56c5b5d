// Closest-hit traversal of the reference tree layout (bvnode, 32 B) with the
// reference arithmetic (CUDA_raytraceMeshTree.cu), for gfx950:
//   k_trace_ray     per-ray DFS (kernel_raytraceMeshTree, :111-219), one lane
//                   per ray for the nodes; a leaf's faces tested by its lane,
//                   or by the whole wave when few lanes hold a leaf; stack in
//                   LDS and scratch; a ray that passes every box test is
//                   handed off to
//   k_units_split, k_units_walk, k_units_merge  its walk cut into subtrees,
//                   walked for groups of rays of one direction octant, and
//                   merged in walk order; another ray past a face budget to
//   k_trace_ray_wave  the same walk, one wave per ray, leaves tested by the
//                   whole wave;
//   k_trace_packet  the 16x12 packet traversal (kernel_raytraceMeshTreePacket,
//                   :222-377), one 192-thread workgroup (3 waves) per tile,
//                   block-uniform stack in LDS, one barrier vote per node; a
//                   packet past a face budget is handed off to
//   k_trace_packet_split  the same walk with five threads per ray, each
//                   testing a fifth of every chunk of faces, or, for a packet
//                   holding a ray that passes every box test, to
//   k_packet_forced  the packet's walk by parts of 16 rays on different CUs.
// The results of both passes are bit-identical to the reference's.
// Built with -ffp-contract=off, so the triangle and box tests are the same
// float operations in the same order as the reference source.
#include <algorithm>
#include <atomic>
#include <mutex>

#include "rk_device.h"

namespace {

constexpr float kEps = 0.000001f;  // CUDA_raytraceMeshTree.cu:36
constexpr float kUMargin = 0.001f;
#ifndef RK_REF_U_PRETEST
#define RK_REF_U_PRETEST 1
#endif
#ifndef RK_REF_SHARED_TVEC
#define RK_REF_SHARED_TVEC 1
#endif

// intersectTriRay, CUDA_raytraceMeshTree.cu:57-89.
__device__ __forceinline__ bool tri_ray(float4 orig, float4 dir, float4 v0, float4 e1, float4 e2, float& t, float& u,
                                        float& v) {
  float4 pvec = rk::cross4(dir, e2);
  float det = rk::dot4(e1, pvec);
  if (det > -kEps && det < kEps) return false;
  float4 tvec = rk::sub4(orig, v0);
  const float un = rk::dot4(tvec, pvec);
  // Most faces fail on u. Estimated with the hardware reciprocal (within
  // 2^-21 of 1/det), a u below -kUMargin or above 1 + kUMargin means a u
  // computed with the correctly rounded 1/det is outside [0, 1] too (with the
  // same sign, also past the float range); NaN and every other case go on to
  // the exact division. The result is the reference's in every case.
#if RK_REF_U_PRETEST
  const float ua = un * __builtin_amdgcn_rcpf(det);
  if (ua < -kUMargin || ua > 1.0f + kUMargin) return false;
#endif
  float inv_det = 1.0f / det;
  u = un * inv_det;
  if (u < 0.0f || u > 1.0f) return false;
  float4 qvec = rk::cross4(tvec, e1);
  v = rk::dot4(dir, qvec) * inv_det;
  if (v < 0.0f || u + v > 1.0f) return false;
  t = rk::dot4(e2, qvec) * inv_det;
  return t >= 0.0f;
}

// tri_ray with tvec = orig - v0 given: for rays that share an origin (bit for
// bit, w included) it is computed once per face, with tri_ray's operation.
// The same results as tri_ray.
__device__ __forceinline__ bool tri_ray_tv(float4 dir, float4 tvec, float4 e1, float4 e2, float& t, float& u,
                                           float& v) {
  float4 pvec = rk::cross4(dir, e2);
  float det = rk::dot4(e1, pvec);
  if (det > -kEps && det < kEps) return false;
  const float un = rk::dot4(tvec, pvec);
#if RK_REF_U_PRETEST
  const float ua = un * __builtin_amdgcn_rcpf(det);
  if (ua < -kUMargin || ua > 1.0f + kUMargin) return false;
#endif
  float inv_det = 1.0f / det;
  u = un * inv_det;
  if (u < 0.0f || u > 1.0f) return false;
  float4 qvec = rk::cross4(tvec, e1);
  v = rk::dot4(dir, qvec) * inv_det;
  if (v < 0.0f || u + v > 1.0f) return false;
  t = rk::dot4(e2, qvec) * inv_det;
  return t >= 0.0f;
}

__device__ __forceinline__ bool same_bits(float4 a, float4 b) {
  return __float_as_uint(a.x) == __float_as_uint(b.x) && __float_as_uint(a.y) == __float_as_uint(b.y) &&
         __float_as_uint(a.z) == __float_as_uint(b.z) && __float_as_uint(a.w) == __float_as_uint(b.w);
}

// intersectBoxRay, CUDA_raytraceMeshTree.cu:91-105.
__device__ __forceinline__ bool box_ray(float4 bmin, float4 bmax, float4 orig, float4 invdir, float& t) {
  float4 v1 = rk::mul4(rk::sub4(bmin, orig), invdir);
  float4 v2 = rk::mul4(rk::sub4(bmax, orig), invdir);
  float4 vmin = rk::min4(v1, v2);
  float4 vmax = rk::max4(v1, v2);
  float tmin = rk::fmx(vmin.z, rk::fmx(vmin.x, vmin.y));
  float tmax = rk::fmn(vmax.z, rk::fmn(vmax.x, vmax.y));
  t = tmin;
  return (tmax >= tmin) && (tmax >= 0.0f);
}

// invdir = normalize(1/dir), :128, with normalize read as v * (1.0f / sqrtf(dot(v, v))).
__device__ __forceinline__ float4 invdir_of(float4 dir) {
  float4 v = make_float4(1.0f / dir.x, 1.0f / dir.y, 1.0f / dir.z, 0.0f);
  float inv_len = 1.0f / __builtin_sqrtf(rk::dot4(v, v));
  return rk::muls4(v, inv_len);
}

// Whether box_ray passes every box for this invdir while t < depth allows:
// a direction component of 0 makes 1/dir infinite, the normalisation factor
// 0 and that component NaN; the other components are then +-0, and
// box_ray's tmin and tmax are both 0.
__device__ __forceinline__ bool crosses_every_box(float4 invdir) {
  return __builtin_isnan(invdir.x) || __builtin_isnan(invdir.y) || __builtin_isnan(invdir.z);
}

__device__ __forceinline__ void tri_setup(const uint4* __restrict__ faces, const float4* __restrict__ verts, uint32_t f,
                                          float4& v0, float4& e1, float4& e2) {
  uint4 face = faces[f];
  v0 = verts[face.x];
  e1 = rk::sub4(verts[face.y], v0);
  e2 = rk::sub4(verts[face.z], v0);
  v0.w = 0.0f;
  e1.w = 0.0f;
  e2.w = 0.0f;
}

// A face's tri_setup terms, v0, e1 = v1 - v0 and e2 = v2 - v0 (w 0),
// precomputed per call in face order (nine floats, 36 B per face): one
// contiguous load per face test instead of the face's indices, then its
// vertices.
constexpr size_t kRecFloats = 9;
__global__ __launch_bounds__(256) void k_tri_records(const uint4* __restrict__ faces,
                                                     const float4* __restrict__ verts, uint32_t face_count,
                                                     float* __restrict__ tris) {
  const uint32_t f = blockIdx.x * 256 + threadIdx.x;
  if (f >= face_count) return;
  float4 v0, e1, e2;
  tri_setup(faces, verts, f, v0, e1, e2);
  float3* r = reinterpret_cast<float3*>(tris + kRecFloats * (size_t)f);
  r[0] = make_float3(v0.x, v0.y, v0.z);
  r[1] = make_float3(e1.x, e1.y, e1.z);
  r[2] = make_float3(e2.x, e2.y, e2.z);
}

// One of a record's three vectors, w 0 as tri_setup leaves it.
__device__ __forceinline__ float4 rec_vec(const float* __restrict__ r) {
  const float3 v = *reinterpret_cast<const float3*>(r);
  return make_float4(v.x, v.y, v.z, 0.0f);
}

// tri_setup from the records when the call has them, else from the mesh.
__device__ __forceinline__ void tri_get(const uint4* __restrict__ faces, const float4* __restrict__ verts,
                                        const float* __restrict__ tris, uint32_t f, float4& v0, float4& e1,
                                        float4& e2) {
  if (tris) {
    const float* r = tris + kRecFloats * (size_t)f;
    v0 = rec_vec(r);
    e1 = rec_vec(r + 3);
    e2 = rec_vec(r + 6);
  } else {
    tri_setup(faces, verts, f, v0, e1, e2);
  }
}

struct NodeRegs { uint32_t w0, w1; float4 bmin, bmax; };

__device__ __forceinline__ NodeRegs load_node(const rk::Node* __restrict__ nodes, uint32_t id) {
  const uint4* p = reinterpret_cast<const uint4*>(nodes + id);
  uint4 a = p[0], b = p[1];
  NodeRegs r;
  r.w0 = a.x; r.w1 = a.y;
  r.bmin = make_float4(__uint_as_float(a.z), __uint_as_float(a.w), __uint_as_float(b.x), 0.0f);
  r.bmax = make_float4(__uint_as_float(b.y), __uint_as_float(b.z), __uint_as_float(b.w), 0.0f);
  return r;
}

// getChildren, CUDANode.h:113-127: a negative direction on the split axis visits child_id+1 first.
__device__ __forceinline__ void children_order(uint32_t w0, uint32_t w1, float4 dir, uint32_t& nearc, uint32_t& farc) {
  uint32_t ax = rk::node_axis(w0);
  bool neg = ax == 0 ? dir.x < 0.0f : (ax == 1 ? dir.y < 0.0f : dir.z < 0.0f);
  nearc = neg ? w1 + 1 : w1;
  farc = neg ? w1 : w1 + 1;
}

constexpr int kStack = 96;  // STACK_SIZE, :45
#ifndef RK_REF_LDS_STACK
#define RK_REF_LDS_STACK 32
#endif
constexpr int kLdsStack = RK_REF_LDS_STACK;  // entries per lane in LDS (8 KB per wave); the rest in scratch

__device__ __forceinline__ uint32_t readlane_u32(uint32_t v, int l) {
  return (uint32_t)__builtin_amdgcn_readlane((int)v, l);
}
__device__ __forceinline__ float readlane_f32(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ __forceinline__ float4 readlane_f4(float4 v, int l) {
  return make_float4(readlane_f32(v.x, l), readlane_f32(v.y, l), readlane_f32(v.z, l), readlane_f32(v.w, l));
}

// The hand-off list of the per-ray traversal: word 0 counts the rays handed
// off, the ray indices start at word kHandoffHead.
constexpr uint32_t kHandoffHead = 64;
constexpr uint32_t kHandoffCap = 1u << 16;  // rays per call (256 KB)
constexpr uint32_t kWaveGrid = 2048;     // one-wave workgroups merging the split rays
constexpr uint32_t kRayWaveGrid = 8192;  // one-wave workgroups walking the other rays handed off
#ifndef RK_REF_WAVE_PER
#define RK_REF_WAVE_PER 4
#endif
constexpr int kWavePer = RK_REF_WAVE_PER;  // faces per lane per round of the second pass's leaf tests

struct LeafBest { float t, u, v; uint32_t k; bool found; };

// The whole wave tests faces [first, first + cnt) against one ray (o, d, dep:
// the same in every lane) and returns the first face of smallest t < dep, the
// face a sequential pass keeping hits with t < depth would end on: its depth
// after such a pass is the smallest t below dep, and a later face with the
// same t fails t < depth. A round takes up to 64 * kPer faces, lane x the
// consecutive faces [x * per, x * per + per), so that lane order is face
// order: each lane keeps its first face of smallest t below dep, and the
// lowest lane holding the wave's smallest t has the first such face. The
// next round compares with that t.
template <int kPer>
__device__ __forceinline__ LeafBest wave_leaf(const uint4* __restrict__ faces, const float4* __restrict__ verts,
                                              const float* __restrict__ tris, float4 o, float4 d, uint32_t first,
                                              uint32_t cnt, float dep) {
  const uint32_t lane = threadIdx.x & 63u;
  LeafBest r;
  r.t = dep;
  r.u = 0.0f;
  r.v = 0.0f;
  r.k = 0;
  r.found = false;
  const float inf = __int_as_float(0x7f800000);
  for (uint32_t c0 = 0; c0 < cnt; c0 += 64u * kPer) {
    const uint32_t m = min(cnt - c0, 64u * kPer);
    const uint32_t per = (m + 63u) >> 6;
    const uint32_t k0 = lane * per;
    float4 p0[kPer], p1[kPer], p2[kPer];
    if (tris) {  // v0, e1, e2
#pragma unroll
      for (int j = 0; j < kPer; ++j) {
        if ((uint32_t)j < per && k0 + j < m) {
          const float* q = tris + kRecFloats * (size_t)(first + c0 + k0 + j);
          p0[j] = rec_vec(q);
          p1[j] = rec_vec(q + 3);
          p2[j] = rec_vec(q + 6);
        }
      }
    } else {  // the vertices
      uint4 f[kPer];
#pragma unroll
      for (int j = 0; j < kPer; ++j)
        f[j] = ((uint32_t)j < per && k0 + j < m) ? faces[first + c0 + k0 + j] : make_uint4(0, 0, 0, 0);
#pragma unroll
      for (int j = 0; j < kPer; ++j) {
        if ((uint32_t)j < per && k0 + j < m) {
          p0[j] = verts[f[j].x];
          p1[j] = verts[f[j].y];
          p2[j] = verts[f[j].z];
        }
      }
    }
    float bt = inf, bu = 0.0f, bv = 0.0f;
    uint32_t bk = 0;
    bool any = false;
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
      if ((uint32_t)j < per && k0 + j < m) {
        float4 v0 = p0[j], e1 = p1[j], e2 = p2[j];
        if (!tris) {  // tri_setup's arithmetic: v0, v1 - v0, v2 - v0, w cleared
          e1 = rk::sub4(p1[j], v0);
          e2 = rk::sub4(p2[j], v0);
          v0.w = 0.0f;
          e1.w = 0.0f;
          e2.w = 0.0f;
        }
        float t, u, v;
        if (tri_ray(o, d, v0, e1, e2, t, u, v) && t < r.t && t < bt) {
          bt = t;  // t < dep <= FLT_MAX or +inf: t is finite, below the sentinel
          bu = u;
          bv = v;
          bk = c0 + k0 + (uint32_t)j;
          any = true;
        }
      }
    }
    if (__ballot(any) == 0ull) continue;
    float mn = any ? bt : inf;
    for (int off = 32; off > 0; off >>= 1) mn = fminf(mn, __shfl_xor(mn, off, 64));
    const int w = __builtin_ctzll(__ballot(any && bt == mn));
    r.t = __shfl(bt, w, 64);  // the winner's own bits (bt == mn holds for -0 against +0 too)
    r.u = __shfl(bu, w, 64);
    r.v = __shfl(bv, w, 64);
    r.k = (uint32_t)__shfl((int)bk, w, 64);
    r.found = true;
  }
  return r;
}

// Rays that pass every box test (crosses_every_box) are split by subtree. The
// walk of such a ray visits every node below one that passes its box test,
// in an order fixed by the direction, as long as its depth stays above 0
// (its box tests give t = +-0); at depth 0 no later face can be kept (t < 0
// never holds). So the subtrees at tree level kUnitLevel (and the leaves
// above it), each walked on its own from the ray's starting depth, give the
// walk's result when merged in walk order, first smallest t winning. A
// subtree's walk differs from the reference's only after the reference's
// depth has reached 0, when its result can no longer change. The stack the
// reference holds when it pops a subtree's root is known from the walk above
// the cut, so a stack overflow stops both at the same node; the visit limit
// is checked on the sum (past it, the ray is walked again by one wave). The
// subtrees are walked for groups of up to 64 rays of one direction octant at
// once (k_units_walk). On the 1920x1080 heightfield (1,949 such rays), cuts
// at levels 6, 8, 9 and 10 gave 15.7, 8.5, 9.9 and 9.4 ms per frame on the
// depth-12 LBVH and 8.8, 5.9, 6.2 and 6.8 ms on the SAH tree.
#ifndef RK_REF_UNIT_LEVEL
#define RK_REF_UNIT_LEVEL 8
#endif
constexpr int kUnitLevel = RK_REF_UNIT_LEVEL;
constexpr uint32_t kUnits = 1u << kUnitLevel;  // subtrees per ray at most: the cut has <= 2^kUnitLevel nodes
#ifndef RK_REF_CROSS_CAP
#define RK_REF_CROSS_CAP 4096
#endif
constexpr uint32_t kCrossCap = RK_REF_CROSS_CAP;          // rays split per call
constexpr uint32_t kUnitGrid = 8192;          // workgroups walking the subtrees
// A subtree walk takes two waves, each testing half of every batch of faces,
// when the tree's leaves hold 64 faces or more on average (node_count / 2
// leaves): the heightfield's depth-12 LBVH (155 faces per leaf) 7.7 -> 6.8 ms
// per frame; with small leaves the second wave only repeats the walk
// (depth-31 LBVH, 42 faces per leaf: 9.5 against 9.2 ms; SAH tree, 21 faces
// per leaf: 5.6 against 6.7 ms), and four waves were slower on all three.
#ifndef RK_REF_UNIT_WAVES
#define RK_REF_UNIT_WAVES 2
#endif
constexpr uint32_t kUnitWaveFaces = 64;  // faces per leaf from which a walk takes RK_REF_UNIT_WAVES waves
constexpr uint32_t kNoFace = 0xffffffffu;
constexpr uint32_t kUnitStop = 1u;            // the walk stopped here (stack overflow, child out of range)

struct UnitResult { float t, u, v; uint32_t face, pops, flags, pad0, pad1; };

// The split rays' state in the hand-off buffer.
struct UnitArgs {
  uint32_t* list;        // [cap] the rays, in hand-off order
  uint32_t* nunits;      // [cap] subtrees of each
  uint32_t* top_pops;    // [cap] pops of the walk above the cut
  uint32_t* diverged;    // [cap] 1: the ray's walk left its group's; walked again by one wave
  uint32_t* bucket;      // [8 * cap] the slots of the rays of each direction octant
  uint32_t* unit_node;   // [cap * kUnits] each subtree's root, in walk order
  uint32_t* unit_sp;     // [cap * kUnits] the reference's stack size after popping that root
  UnitResult* res;       // [cap * kUnits]
  uint32_t cap;
};
constexpr uint32_t kBucketHead = 8;  // handoff[8 + o]: rays of octant o

// The per-ray traversal (kernel_raytraceMeshTree, :111-219), one lane per ray:
// pop a node, test its box against the depth so far, push the far child, then
// the near one; at a leaf, test its faces in order, keeping a hit only when
// t < depth. Each ray's result is bit-identical to the reference's.
//
// The lanes of a wave walk their nodes until each has reached a leaf or its
// end; then the leaves are tested, and the walk goes on. A frame's cost sits
// in a few rays that graze the scene and cross dozens of leaves of a few
// hundred faces each: on the 1920x1080 heightfield the waves that hold them
// ran alone for most of the frame (0.5 waves per CU on average). So when at
// most kCoopMax lanes hold a leaf, the whole wave tests each of those leaves
// in turn, 64 faces at a time, one face per lane: the reference keeps the
// first face of smallest t below the depth the leaf was entered with, so a
// round keeps the smallest t below the current depth, ties to the smallest
// face index, and the next round compares with the depth that leaves (the
// same depth after every leaf as the reference's, and so the same later
// pops). When more lanes hold a leaf, each lane tests its own, its faces
// fetched kLeafBatch at a time with all their loads issued before the first
// test (a face costs two dependent loads: the face, then its vertices).
//
// The stack keeps the reference's 96 entries and its overflow rule; the first
// kLdsStack entries of each lane live in LDS (entry e of lane x at
// e * 64 + x), deeper ones in scratch (trees of depth <= 31 need at most 33),
// so a CU holds 20 waves of this kernel instead of 6.
#ifndef RK_REF_LEAF_BATCH
#define RK_REF_LEAF_BATCH 2
#endif
constexpr int kLeafBatch = RK_REF_LEAF_BATCH;

template <int kCoopMax>
__global__ __launch_bounds__(64) void k_trace_ray(const rk::Node* __restrict__ nodes,
                                                  const uint4* __restrict__ faces,
                                                  const float4* __restrict__ verts,
                                                  const float4* __restrict__ origins,
                                                  const float4* __restrict__ dirs, uint32_t n_rays,
                                                  uint32_t node_count, uint32_t face_count,
                                                  uint32_t* __restrict__ faces_id, float2* __restrict__ coords,
                                                  float* __restrict__ depths, uint32_t* __restrict__ handoff,
                                                  uint32_t handoff_cap, uint32_t handoff_faces, UnitArgs ua,
                                                  const float* __restrict__ tris) {
  __shared__ uint32_t s_stack[kLdsStack * 64];
  const int lane = (int)threadIdx.x;
  const uint32_t l = blockIdx.x * 64 + (uint32_t)lane;
  const bool has_ray = l < n_rays;
  const float4 zero4 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  const float4 orig = has_ray ? origins[l] : zero4, dir = has_ray ? dirs[l] : zero4;
  const float4 invdir = has_ray ? invdir_of(dir) : zero4;
  float depth = has_ray ? depths[l] : 0.0f;
  uint32_t fid = has_ray ? faces_id[l] : 0u;
  float2 coord = has_ray ? coords[l] : make_float2(0.0f, 0.0f);
  uint32_t spill[kStack - kLdsStack];
  typedef __attribute__((address_space(3))) uint32_t lds_u32;
  lds_u32* my_lds = (lds_u32*)(s_stack + lane);
  int sp = 0;
  if (has_ray) {
    my_lds[0] = 0u;
    sp = 1;
  }
  bool active = has_ray;
  bool may_hand_off = handoff_cap != 0u, handed_off = false;
  uint32_t tested = 0;  // faces of the leaves reached, while may_hand_off
  // A direction with a component of 0 leaves a NaN in invdir: every box test
  // then passes (at t = 0, while depth > 0) and the ray walks the whole tree,
  // so it goes to the split at its first leaf (else to the second pass).
  const bool crossing = crosses_every_box(invdir);
  const uint32_t limit = crossing ? 0u : handoff_faces;
  uint32_t visits = 0;
  const uint32_t max_visits = 2 * node_count + 2;
  for (;;) {
    // each lane walks on to its next leaf, or to its end
    bool at_leaf = false;
    uint32_t leaf_first = 0, leaf_cnt = 0;
    while (active && !at_leaf) {
      if (sp == 0 || ++visits > max_visits) {
        active = false;
        break;
      }
      --sp;
      const uint32_t nid = sp < kLdsStack ? my_lds[sp * 64] : spill[sp - kLdsStack];
      if (nid >= node_count) {  // never taken for a tree built by this library
        active = false;
        break;
      }
      const NodeRegs nd = load_node(nodes, nid);
      if (rk::node_is_leaf(nd.w0)) {
        leaf_first = nd.w1;
        leaf_cnt = leaf_first < face_count ? min(rk::leaf_count(nd.w0), face_count - leaf_first) : 0u;
        if (may_hand_off) {
          if (leaf_cnt > limit - tested) {
            may_hand_off = false;
            if (crossing && ua.cap != 0u) {
              const uint32_t slot = atomicAdd(handoff + 1, 1u);
              if (slot < ua.cap) {
                ua.list[slot] = l;
                handed_off = true;
                active = false;
                break;
              }
            }
            const uint32_t slot = atomicAdd(handoff, 1u);
            if (slot < handoff_cap) {
              handoff[kHandoffHead + slot] = l;
              handed_off = true;
              active = false;
              break;
            }
          } else {
            tested += leaf_cnt;
          }
        }
        at_leaf = true;
      } else {
        float t;
        if (box_ray(nd.bmin, nd.bmax, orig, invdir, t) && t < depth) {
          uint32_t nearc, farc;
          children_order(nd.w0, nd.w1, dir, nearc, farc);
          if (sp + 2 > kStack) {  // depth > 94: cannot occur for trees of depth <= 94
            active = false;
            break;
          }
          if (sp < kLdsStack) my_lds[sp * 64] = farc;
          else spill[sp - kLdsStack] = farc;
          ++sp;
          if (sp < kLdsStack) my_lds[sp * 64] = nearc;
          else spill[sp - kLdsStack] = nearc;
          ++sp;
        }
      }
    }
    uint64_t pending = __ballot(at_leaf);
    if (pending == 0ull) break;  // every lane is done
    if (__popcll(pending) > kCoopMax) {
      // each lane its own leaf, in face order
      if (at_leaf) {
        for (uint32_t k0 = 0; k0 < leaf_cnt; k0 += kLeafBatch) {
          float4 p0[kLeafBatch], p1[kLeafBatch], p2[kLeafBatch];
          if (tris) {  // v0, e1, e2
#pragma unroll
            for (int b = 0; b < kLeafBatch; ++b) {
              if (k0 + b < leaf_cnt) {
                const float* q = tris + kRecFloats * (size_t)(leaf_first + k0 + b);
                p0[b] = rec_vec(q);
                p1[b] = rec_vec(q + 3);
                p2[b] = rec_vec(q + 6);
              }
            }
          } else {  // the vertices
            uint4 f[kLeafBatch];
#pragma unroll
            for (int b = 0; b < kLeafBatch; ++b)
              f[b] = k0 + b < leaf_cnt ? faces[leaf_first + k0 + b] : make_uint4(0, 0, 0, 0);
#pragma unroll
            for (int b = 0; b < kLeafBatch; ++b) {
              if (k0 + b < leaf_cnt) {
                p0[b] = verts[f[b].x];
                p1[b] = verts[f[b].y];
                p2[b] = verts[f[b].z];
              }
            }
          }
#pragma unroll
          for (int b = 0; b < kLeafBatch; ++b) {
            if (k0 + b < leaf_cnt) {
              float4 v0 = p0[b], e1 = p1[b], e2 = p2[b];
              if (!tris) {  // tri_setup's arithmetic: v0, v1 - v0, v2 - v0, w cleared
                e1 = rk::sub4(p1[b], v0);
                e2 = rk::sub4(p2[b], v0);
                v0.w = 0.0f;
                e1.w = 0.0f;
                e2.w = 0.0f;
              }
              float t, u, v;
              if (tri_ray(orig, dir, v0, e1, e2, t, u, v)) {
                if (t < depth) { fid = leaf_first + k0 + b; coord = make_float2(u, v); depth = t; }
              }
            }
          }
        }
      }
    } else {
      // the whole wave on each pending leaf in turn
      while (pending != 0ull) {
        const int L = __builtin_ctzll(pending);
        pending &= pending - 1;
        const uint32_t first = readlane_u32(leaf_first, L);
        const LeafBest b = wave_leaf<1>(faces, verts, tris, readlane_f4(orig, L), readlane_f4(dir, L), first,
                                     readlane_u32(leaf_cnt, L), readlane_f32(depth, L));
        if (lane == L && b.found) {
          depth = b.t;
          fid = first + b.k;
          coord = make_float2(b.u, b.v);
        }
      }
    }
  }
  if (has_ray && !handed_off) {
    faces_id[l] = fid;
    coords[l] = coord;
    depths[l] = depth;
  }
}

__global__ __launch_bounds__(64) void k_handoff_clear(uint32_t* __restrict__ handoff) {
  if (threadIdx.x < kHandoffHead) handoff[threadIdx.x] = 0u;
}

// The walk of a split ray above the cut, one lane per ray: the reference's
// pops of the nodes above kUnitLevel (box tests at the starting depth), each
// node at the cut, leaf above it or child out of range recorded as a subtree
// root with the stack size after its pop. The stack holds at most
// kUnitLevel + 1 entries here.
__global__ __launch_bounds__(64) void k_units_split(const rk::Node* __restrict__ nodes,
                                                    const float4* __restrict__ origins,
                                                    const float4* __restrict__ dirs,
                                                    const float* __restrict__ depths, uint32_t node_count,
                                                    uint32_t* __restrict__ handoff, UnitArgs ua) {
  const uint32_t i = blockIdx.x * 64 + threadIdx.x;
  if (i >= min(handoff[1], ua.cap)) return;
  const uint32_t l = ua.list[i];
  const float4 orig = origins[l], dir = dirs[l];
  const float4 invdir = invdir_of(dir);
  const float depth = depths[l];
  uint32_t st[2 * kUnitLevel + 2], lv[2 * kUnitLevel + 2];
  int sp = 1;
  st[0] = 0u;
  lv[0] = 0u;
  uint32_t nu = 0, pops = 0;
  uint32_t* unode = ua.unit_node + (size_t)i * kUnits;
  uint32_t* usp = ua.unit_sp + (size_t)i * kUnits;
  while (sp > 0 && nu < kUnits) {
    --sp;
    const uint32_t nid = st[sp], level = lv[sp];
    bool unit = level >= (uint32_t)kUnitLevel || nid >= node_count;
    NodeRegs nd;
    if (!unit) {
      nd = load_node(nodes, nid);
      unit = rk::node_is_leaf(nd.w0);
    }
    if (unit) {  // its walk counts its pop and stops at a child out of range
      unode[nu] = nid;
      usp[nu] = (uint32_t)sp;
      ++nu;
      if (nid >= node_count) break;
      continue;
    }
    ++pops;
    float t;
    if (box_ray(nd.bmin, nd.bmax, orig, invdir, t) && t < depth) {
      uint32_t nearc, farc;
      children_order(nd.w0, nd.w1, dir, nearc, farc);
      if (sp + 2 > 2 * kUnitLevel + 2) break;  // cannot occur: one entry per level, two at the deepest
      st[sp] = farc;
      lv[sp] = level + 1u;
      ++sp;
      st[sp] = nearc;
      lv[sp] = level + 1u;
      ++sp;
    }
  }
  ua.nunits[i] = nu;
  ua.top_pops[i] = pops;
  ua.diverged[i] = 0u;
  // children_order's choice at every node: rays of one octant walk alike
  const uint32_t o = (dir.x < 0.0f ? 1u : 0u) | (dir.y < 0.0f ? 2u : 0u) | (dir.z < 0.0f ? 4u : 0u);
  const uint32_t slot = atomicAdd(handoff + kBucketHead + o, 1u);
  ua.bucket[o * ua.cap + slot] = i;  // slot < cap: at most cap rays in all
}

// The walk of ray l by the whole wave with the reference's rules (the same
// pops, box tests, stack overflow and visit limits), each leaf tested by the
// whole wave (wave_leaf), from node root with the reference's stack holding
// sp_base entries below it (0 and 0: the ray's walk), from depth, fid and
// coord. The stack's entry e is lane e % 64 of st0 (e < 64) or st1.
struct WalkResult { float depth; uint32_t fid; float2 coord; uint32_t pops; bool stopped; };

__device__ __forceinline__ WalkResult wave_walk(uint32_t l, uint32_t root, uint32_t sp_base, float depth,
                                                uint32_t fid, float2 coord, const rk::Node* __restrict__ nodes,
                                                const uint4* __restrict__ faces, const float4* __restrict__ verts,
                                                const float4* __restrict__ origins,
                                                const float4* __restrict__ dirs, uint32_t node_count,
                                                uint32_t face_count, const float* __restrict__ tris) {
  const int lane = (int)(threadIdx.x & 63u);
  const uint32_t max_visits = 2 * node_count + 2;
  const float4 orig = origins[l], dir = dirs[l];
  const float4 invdir = invdir_of(dir);
  uint32_t st0 = lane == 0 ? root : 0u, st1 = 0u;
  int sp = 1;
  uint32_t visits = 0;
  bool stopped = false;
  while (sp > 0 && ++visits <= max_visits) {
    --sp;
    const uint32_t nid = readlane_u32(sp < 64 ? st0 : st1, sp & 63);
    if (nid >= node_count) {  // never taken for a tree built by this library
      stopped = true;
      break;
    }
    const NodeRegs nd = load_node(nodes, nid);
    if (rk::node_is_leaf(nd.w0)) {
      const uint32_t first = nd.w1;
      const uint32_t cnt = first < face_count ? min(rk::leaf_count(nd.w0), face_count - first) : 0u;
      const LeafBest b = wave_leaf<kWavePer>(faces, verts, tris, orig, dir, first, cnt, depth);
      if (b.found) {
        depth = b.t;
        fid = first + b.k;
        coord = make_float2(b.u, b.v);
      }
    } else {
      float t;
      if (box_ray(nd.bmin, nd.bmax, orig, invdir, t) && t < depth) {
        uint32_t nearc, farc;
        children_order(nd.w0, nd.w1, dir, nearc, farc);
        if ((uint32_t)sp + sp_base + 2u > (uint32_t)kStack) {  // cannot occur for trees of depth <= 94
          stopped = true;
          break;
        }
        if (lane == (sp & 63)) {
          if (sp < 64) st0 = farc;
          else st1 = farc;
        }
        ++sp;
        if (lane == (sp & 63)) {
          if (sp < 64) st0 = nearc;
          else st1 = nearc;
        }
        ++sp;
      }
    }
  }
  WalkResult r;
  r.depth = depth;
  r.fid = fid;
  r.coord = coord;
  r.pops = visits;
  r.stopped = stopped;
  return r;
}

// The second pass of the per-ray traversal: one wave per ray handed off, its
// walk again from the root.
__global__ __launch_bounds__(64) void k_trace_ray_wave(const rk::Node* __restrict__ nodes,
                                                       const uint4* __restrict__ faces,
                                                       const float4* __restrict__ verts,
                                                       const float4* __restrict__ origins,
                                                       const float4* __restrict__ dirs, uint32_t node_count,
                                                       uint32_t face_count, const uint32_t* __restrict__ handoff,
                                                       uint32_t handoff_cap, uint32_t* __restrict__ faces_id,
                                                       float2* __restrict__ coords, float* __restrict__ depths,
                                                       const float* __restrict__ tris) {
  const uint32_t count = min(handoff[0], handoff_cap);
  for (uint32_t i = blockIdx.x; i < count; i += gridDim.x) {
    const uint32_t l = handoff[kHandoffHead + i];
    const WalkResult r = wave_walk(l, 0u, 0u, depths[l], faces_id[l], coords[l], nodes, faces, verts, origins,
                                   dirs, node_count, face_count, tris);
    if (threadIdx.x == 0) {
      faces_id[l] = r.fid;
      coords[l] = r.coord;
      depths[l] = r.depth;
    }
  }
}

// The subtrees of the split rays, by groups of up to 64 rays of one octant:
// one wave per group and subtree, lane r on the group's ray r. The rays of a
// group walk alike (children_order depends on the octant only, and their box
// tests pass at t = +-0 while their depth is above 0), so the wave walks the
// subtree once, from the root its first ray (the leader) recorded: a node is
// visited when a ray still walking (not diverged, depth above 0) passes its
// box test, and a ray still walking that fails a visited node's test has left
// the group's walk (diverged: the ray is walked again by one wave). A leaf's
// faces go to LDS 64 at a time, and each ray tests them in face order against
// its own depth, as the reference does. A ray whose subtree list differs from
// the leader's is marked diverged too. Each ray gets its UnitResult.
template <int kUW>
__global__ __launch_bounds__(64 * kUW) void k_units_walk(const rk::Node* __restrict__ nodes,
                                                   const uint4* __restrict__ faces,
                                                   const float4* __restrict__ verts,
                                                   const float4* __restrict__ origins,
                                                   const float4* __restrict__ dirs, const float* __restrict__ depths,
                                                   uint32_t node_count, uint32_t face_count,
                                                   const uint32_t* __restrict__ handoff, UnitArgs ua,
                                                   const float* __restrict__ tris) {
  __shared__ float4 s_v0[64 * kUW], s_e1[64 * kUW], s_e2[64 * kUW];
  __shared__ float4 s_part[kUW > 1 ? 64 * kUW : 1];  // a wave's first smallest (t, u, v, face) per ray
  const int lane = (int)threadIdx.x & 63, wv = (int)threadIdx.x >> 6;
  uint32_t bcount[8], groups = 0;
#pragma unroll
  for (int o = 0; o < 8; ++o) {
    bcount[o] = handoff[kBucketHead + o];
    groups += (bcount[o] + 63u) / 64u;
  }
  const uint32_t jobs = groups * kUnits;
  const uint32_t max_visits = 2 * node_count + 2;
  for (uint32_t jj = blockIdx.x; jj < jobs; jj += gridDim.x) {
    // subtree k of every group before subtree k + 1 of any: the waves running
    // together walk the same part of the tree and share it in the L2
    const uint32_t k = jj / groups;
    uint32_t g = jj % groups, o = 0;
    while (g >= (bcount[o] + 63u) / 64u) {
      g -= (bcount[o] + 63u) / 64u;
      ++o;
    }
    const uint32_t n_in = min(64u, bcount[o] - g * 64u);
    const bool has_ray = (uint32_t)lane < n_in;
    const uint32_t i = has_ray ? ua.bucket[o * ua.cap + g * 64u + (uint32_t)lane] : 0u;
    const uint32_t lead = readlane_u32(i, 0);
    const uint32_t nu_lead = ua.nunits[lead];
    if (k == 0 && has_ray && wv == 0 && ua.nunits[i] != nu_lead) ua.diverged[i] = 1u;
    if (k >= nu_lead) continue;
    const uint32_t root = ua.unit_node[lead * kUnits + k], sp_base = ua.unit_sp[lead * kUnits + k];
    bool walking = has_ray && ua.nunits[i] == nu_lead && ua.unit_node[i * kUnits + k] == root &&
                   ua.unit_sp[i * kUnits + k] == sp_base;
    bool diverged = has_ray && !walking;
    const uint32_t l = has_ray ? ua.list[i] : 0u;
    const float4 zero4 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    const float4 orig = has_ray ? origins[l] : zero4, dir = has_ray ? dirs[l] : zero4;
    const float4 invdir = has_ray ? invdir_of(dir) : zero4;
    float depth = has_ray ? depths[l] : 0.0f;
    const float4 dir0 = readlane_f4(dir, 0);  // the octant's: the leader's direction
    const float4 orig0 = readlane_f4(orig, 0);
    // rays from one camera share their origin: tvec once per face
    const bool shared = RK_REF_SHARED_TVEC && __ballot(has_ray && !same_bits(orig, orig0)) == 0ull;
    uint32_t fid = kNoFace;
    float2 coord = make_float2(0.0f, 0.0f);
    uint32_t st0 = lane == 0 ? root : 0u, st1 = 0u;
    int sp = 1;
    uint32_t visits = 0;
    bool stopped = false;
    while (sp > 0 && ++visits <= max_visits) {
      if (__ballot(walking && depth > 0.0f) == 0ull) break;  // every ray's result here is final
      --sp;
      const uint32_t nid = readlane_u32(sp < 64 ? st0 : st1, sp & 63);
      if (nid >= node_count) {  // never taken for a tree built by this library
        stopped = true;
        break;
      }
      const NodeRegs nd = load_node(nodes, nid);
      if (rk::node_is_leaf(nd.w0)) {
        const uint32_t first = nd.w1;
        const uint32_t cnt = first < face_count ? min(rk::leaf_count(nd.w0), face_count - first) : 0u;
        // Batches of 64 * kUW faces; wave w tests [w * 64, w * 64 + 64) of each,
        // keeping its first smallest t below the depth at the batch's start,
        // and the waves' results are merged in wave order: a sequential pass
        // keeping t < depth ends on the same face (the first of smallest t).
        for (uint32_t c0 = 0; c0 < cnt; c0 += 64u * kUW) {
          const uint32_t cc = min(64u * kUW, cnt - c0);
          if (threadIdx.x < cc) {
            float4 v0, e1, e2;
            tri_get(faces, verts, tris, first + c0 + threadIdx.x, v0, e1, e2);
            s_v0[threadIdx.x] = shared ? rk::sub4(orig0, v0) : v0;  // tvec, or v0
            s_e1[threadIdx.x] = e1;
            s_e2[threadIdx.x] = e2;
          }
          __syncthreads();
          const uint32_t q0 = (uint32_t)wv * 64u, q1 = min(cc, q0 + 64u);
          float pt = depth, pu = 0.0f, pv = 0.0f;
          uint32_t pq = 0xFFFFFFFFu;  // none
          if (walking) {
            if (shared) {
              for (uint32_t q = q0; q < q1; ++q) {
                float t, u, v;
                if (tri_ray_tv(dir, s_v0[q], s_e1[q], s_e2[q], t, u, v)) {
                  if (t < pt) { pq = q; pu = u; pv = v; pt = t; }
                }
              }
            } else {
              for (uint32_t q = q0; q < q1; ++q) {
                float t, u, v;
                if (tri_ray(orig, dir, s_v0[q], s_e1[q], s_e2[q], t, u, v)) {
                  if (t < pt) { pq = q; pu = u; pv = v; pt = t; }
                }
              }
            }
          }
          if (kUW == 1) {
            if (pq != 0xFFFFFFFFu) { fid = first + c0 + pq; coord = make_float2(pu, pv); depth = pt; }
          } else {
            s_part[threadIdx.x] = make_float4(pt, pu, pv, __uint_as_float(pq));
            __syncthreads();
            if (walking) {
#pragma unroll
              for (int w = 0; w < kUW; ++w) {
                const float4 r = s_part[w * 64 + lane];
                const uint32_t rq = __float_as_uint(r.w);
                if (rq != 0xFFFFFFFFu && r.x < depth) { fid = first + c0 + rq; coord = make_float2(r.y, r.z); depth = r.x; }
              }
            }
          }
          __syncthreads();
        }
      } else {
        float t;
        const bool pass = box_ray(nd.bmin, nd.bmax, orig, invdir, t) && t < depth;
        const bool live = walking && depth > 0.0f;  // at depth 0 a ray's result is final
        if (__ballot(live && pass) != 0ull) {
          if (live && !pass) {
            walking = false;
            diverged = true;
          }
          uint32_t nearc, farc;
          children_order(nd.w0, nd.w1, dir0, nearc, farc);
          if ((uint32_t)sp + sp_base + 2u > (uint32_t)kStack) {  // cannot occur for trees of depth <= 94
            stopped = true;
            break;
          }
          if (lane == (sp & 63)) {
            if (sp < 64) st0 = farc;
            else st1 = farc;
          }
          ++sp;
          if (lane == (sp & 63)) {
            if (sp < 64) st0 = nearc;
            else st1 = nearc;
          }
          ++sp;
        }
      }
    }
    if (has_ray && wv == 0) {
      if (diverged) ua.diverged[i] = 1u;
      UnitResult r;
      r.t = depth;
      r.u = coord.x;
      r.v = coord.y;
      r.face = fid;
      r.pops = visits;
      r.flags = stopped ? kUnitStop : 0u;
      r.pad0 = 0u;
      r.pad1 = 0u;
      ua.res[i * kUnits + k] = r;
    }
  }
}

// The split rays' results: the subtrees' results merged in walk order (the
// first smallest t below the depth so far wins; a subtree that stopped ends
// the walk), when the pops add up to no more than the visit limit; else the
// ray's walk by the whole wave.
__global__ __launch_bounds__(64) void k_units_merge(const rk::Node* __restrict__ nodes,
                                                    const uint4* __restrict__ faces,
                                                    const float4* __restrict__ verts,
                                                    const float4* __restrict__ origins,
                                                    const float4* __restrict__ dirs, uint32_t node_count,
                                                    uint32_t face_count, const uint32_t* __restrict__ handoff,
                                                    UnitArgs ua, uint32_t* __restrict__ faces_id,
                                                    float2* __restrict__ coords, float* __restrict__ depths,
                                                    const float* __restrict__ tris) {
  const uint32_t count = min(handoff[1], ua.cap);
  const uint64_t max_visits = 2 * node_count + 2;
  for (uint32_t i = blockIdx.x; i < count; i += gridDim.x) {
    const uint32_t l = ua.list[i];
    float depth = depths[l];
    uint32_t fid = faces_id[l];
    float2 coord = coords[l];
    uint64_t pops = ua.top_pops[i];
    const uint32_t nu = ua.diverged[i] != 0u ? 0u : ua.nunits[i];  // a diverged ray's results may be missing
    for (uint32_t k = 0; k < nu; ++k) {
      const UnitResult r = ua.res[(size_t)i * kUnits + k];
      pops += r.pops;
      if (r.face != kNoFace && r.t < depth) {
        depth = r.t;
        fid = r.face;
        coord = make_float2(r.u, r.v);
      }
      if (r.flags & kUnitStop) break;
    }
    if (pops > max_visits || ua.diverged[i] != 0u) {
      const WalkResult w = wave_walk(l, 0u, 0u, depths[l], faces_id[l], coords[l], nodes, faces, verts, origins,
                                     dirs, node_count, face_count, tris);
      depth = w.depth;
      fid = w.fid;
      coord = w.coord;
    }
    if (threadIdx.x == 0) {
      faces_id[l] = fid;
      coords[l] = coord;
      depths[l] = depth;
    }
  }
}

#ifndef RK_REF_HANDOFF_FACES
#define RK_REF_HANDOFF_FACES 4096
#endif

#ifndef RK_REF_COOP_MAX
#define RK_REF_COOP_MAX 16
#endif

constexpr int kPW = 16, kPH = 12, kPT = kPW * kPH;  // PACKET_BLOCK_*, :47-49
constexpr int kPStack = 128;                       // PACKET_STACK_SIZE, :51
constexpr int kPCache = 224;                       // TRIANGLES_MAX, :39

// One packet of the packet traversal (kernel_raytraceMeshTreePacket,
// :222-377): the block walks the tree once for its 16x12 rays, visiting a
// node when any of them passes its box test with t < its depth, in the order
// the first ray's direction gives; at a leaf every ray tests every face
// against its own depth, in face order. The block holds kS threads per ray:
// thread x serves ray x % 192, and with kS > 1 tests part x / 192 of each
// chunk of faces (consecutive faces, parts in face order). Each part keeps its
// first face of smallest t below the ray's depth at the chunk's start, and the
// first part with the smallest t wins: the face a sequential pass keeping hits
// with t < depth ends on, with the same depth after the chunk.
//
// A packet whose leaves pass handoff_faces faces (while handoff_cap != 0) is
// put on the hand-off list and left where it stands: its rays' hits so far in
// the outputs, and in its record of the state area its stack with the leaf
// back on top, sp and the visits before that leaf's pop (so that the second
// pass, given the record as resume, makes the same pops from there). Returns
// whether it was.
// A packet's record: sp, visits, the forcing ray (the first ray of the
// packet that passes every box test, or kNone), its forced slot (or kNone;
// kDone once the forced walk's results are in the outputs), the stack.
constexpr uint32_t kPStateWords = kPStack + 4;
constexpr uint32_t kNone = 0xffffffffu, kDone = 0xfffffffeu;

template <int kS>
__device__ __forceinline__ bool trace_packet_job(
    const rk::Node* __restrict__ nodes, const uint4* __restrict__ faces, const float4* __restrict__ verts,
    const float4* __restrict__ origins, const float4* __restrict__ dirs, int width, int height, uint32_t node_count,
    uint32_t face_count, uint32_t* __restrict__ faces_id, float2* __restrict__ coords, float* __restrict__ depths,
    int bx, int by, uint32_t* __restrict__ handoff, uint32_t handoff_cap, uint32_t handoff_faces,
    uint32_t* __restrict__ state, uint32_t* __restrict__ forced_list, uint32_t forced_cap,
    const uint32_t* __restrict__ resume, const float* __restrict__ tris, uint32_t* s_stack, float4* s_v0,
    float4* s_e1, float4* s_e2, float4* s_dir0, uint32_t* s_slot, float* s_pt, float* s_pu, float* s_pv,
    uint32_t* s_pk) {
  const int ray = (int)threadIdx.x % kPT, part = (int)threadIdx.x / kPT;
  const int tx = ray % kPW, ty = ray / kPW;
  const int x = bx * kPW + tx, y = by * kPH + ty;
  const bool in_image = x < width && y < height;
  const uint32_t l = in_image ? (uint32_t)(y * width + x) : 0u;
  float4 orig = make_float4(0.f, 0.f, 0.f, 0.f), dir = orig, invdir = orig;
  float depth = 0.0f;
  uint32_t fid = 0;
  float2 coord = make_float2(0.f, 0.f);
  if (in_image) {
    orig = origins[l];
    dir = dirs[l];
    invdir = invdir_of(dir);
    depth = depths[l];
    fid = faces_id[l];
    coord = coords[l];
  }
  int sp = 1;
  uint32_t visits = 0;
  if (resume) {
    sp = (int)resume[0];
    visits = resume[1];
    for (int k = (int)threadIdx.x; k < sp; k += kPT * kS) s_stack[k] = resume[4 + k];
  }
  if (threadIdx.x == 0) {  // thread (0,0) of a block is always inside the image
    *s_dir0 = dir;
    if (!resume) s_stack[0] = 0;
  }
  bool may_hand_off = handoff_cap != 0u;
  uint32_t limit = handoff_faces, forcer = kNone;
  if (may_hand_off) {  // a packet holding a ray that passes every box test visits every node
    if (threadIdx.x == 0) *s_slot = kNone;
    __syncthreads();
    if (in_image && part == 0 && crosses_every_box(invdir)) atomicMin(s_slot, (uint32_t)ray);
    __syncthreads();
    forcer = *s_slot;
    if (forcer != kNone) limit = 0u;
  }
  __syncthreads();
  const float4 dir0 = *s_dir0;
  uint32_t tested = 0;  // block-uniform: faces of the leaves visited, while may_hand_off
  const uint32_t max_visits = 2 * node_count + 2;
  while (sp > 0 && ++visits <= max_visits) {
    const uint32_t nid = s_stack[--sp];
    if (nid >= node_count) break;  // block-uniform; never taken for a tree built by this library
    const NodeRegs nd = load_node(nodes, nid);  // block-uniform address
    if (rk::node_is_leaf(nd.w0)) {
      const uint32_t tri_id = nd.w1;
      const uint32_t cnt = tri_id < face_count ? min(rk::leaf_count(nd.w0), face_count - tri_id) : 0u;
      if (may_hand_off) {
        if (cnt > limit - tested) {
          may_hand_off = false;
          if (threadIdx.x == 0) {
            const uint32_t slot = atomicAdd(handoff, 1u);
            if (slot < handoff_cap) handoff[kHandoffHead + slot] = (uint32_t)by * gridDim.x + (uint32_t)bx;
            *s_slot = slot;
          }
          __syncthreads();
          const uint32_t slot = *s_slot;
          __syncthreads();
          if (slot < handoff_cap) {
            uint32_t* rec = state + (size_t)slot * kPStateWords;
            for (int k = (int)threadIdx.x; k < sp; k += kPT * kS) rec[4 + k] = s_stack[k];
            if (threadIdx.x == 0) {
              rec[0] = (uint32_t)sp + 1u;  // sp < kPStack after the pop: the leaf fits
              rec[1] = visits - 1u;
              rec[4 + sp] = nid;
              uint32_t f = kNone;
              if (forcer != kNone) {
                f = atomicAdd(handoff + 2, 1u);
                if (f < forced_cap) forced_list[f] = slot;
                else f = kNone;
              }
              rec[2] = f != kNone ? forcer : kNone;
              rec[3] = f;
            }
            if (in_image && part == 0) {
              faces_id[l] = fid;
              coords[l] = coord;
              depths[l] = depth;
            }
            return true;
          }
        } else {
          tested += cnt;
        }
      }
      for (uint32_t c0 = 0; c0 < cnt; c0 += kPCache) {
        const uint32_t cc = (cnt - c0) < (uint32_t)kPCache ? (cnt - c0) : (uint32_t)kPCache;
        __syncthreads();
        for (uint32_t k = threadIdx.x; k < cc; k += kPT * kS) {
          float4 v0, e1, e2;
          tri_get(faces, verts, tris, tri_id + c0 + k, v0, e1, e2);
          s_v0[k] = v0; s_e1[k] = e1; s_e2[k] = e2;
        }
        __syncthreads();
        if (kS == 1) {
          if (in_image) {
            for (uint32_t k = 0; k < cc; ++k) {
              float t, u, v;
              if (tri_ray(orig, dir, s_v0[k], s_e1[k], s_e2[k], t, u, v)) {
                if (t < depth) { fid = tri_id + c0 + k; coord = make_float2(u, v); depth = t; }
              }
            }
          }
        } else {
          const uint32_t q = (cc + kS - 1) / kS;
          const uint32_t k0 = (uint32_t)part * q, k1 = min(cc, k0 + q);
          float pt = depth, pu = 0.0f, pv = 0.0f;
          uint32_t pk = 0xffffffffu;  // none
          if (in_image) {
            for (uint32_t k = k0; k < k1; ++k) {
              float t, u, v;
              if (tri_ray(orig, dir, s_v0[k], s_e1[k], s_e2[k], t, u, v)) {
                if (t < pt) { pk = k; pu = u; pv = v; pt = t; }
              }
            }
          }
          s_pt[threadIdx.x] = pt;
          s_pu[threadIdx.x] = pu;
          s_pv[threadIdx.x] = pv;
          s_pk[threadIdx.x] = pk;
          __syncthreads();
          if (in_image) {
#pragma unroll
            for (int p = 0; p < kS; ++p) {
              const int o = p * kPT + ray;
              const uint32_t k = s_pk[o];
              if (k != 0xffffffffu && s_pt[o] < depth) {
                depth = s_pt[o];
                fid = tri_id + c0 + k;
                coord = make_float2(s_pu[o], s_pv[o]);
              }
            }
          }
        }
      }
      __syncthreads();
    } else {
      int hit = 0;
      if (in_image) {
        float t;
        if (box_ray(nd.bmin, nd.bmax, orig, invdir, t) && t < depth) hit = 1;
      }
      const int any = __syncthreads_or(hit);
      if (any) {
        uint32_t nearc, farc;
        children_order(nd.w0, nd.w1, dir0, nearc, farc);
        if (sp + 2 > kPStack) break;  // block-uniform; cannot occur for trees of depth <= 126
        if (threadIdx.x == 0) { s_stack[sp] = farc; s_stack[sp + 1] = nearc; }
        sp += 2;
      }
      __syncthreads();
    }
  }
  if (in_image && part == 0) {
    faces_id[l] = fid;
    coords[l] = coord;
    depths[l] = depth;
  }
  return false;
}

// The packet traversal's first pass: one 192-thread block per packet.
__global__ __launch_bounds__(kPT) void k_trace_packet(const rk::Node* __restrict__ nodes,
                                                      const uint4* __restrict__ faces,
                                                      const float4* __restrict__ verts,
                                                      const float4* __restrict__ origins,
                                                      const float4* __restrict__ dirs, int width, int height,
                                                      uint32_t node_count, uint32_t face_count,
                                                      uint32_t* __restrict__ faces_id, float2* __restrict__ coords,
                                                      float* __restrict__ depths, uint32_t* __restrict__ handoff,
                                                      uint32_t handoff_cap, uint32_t handoff_faces,
                                                      uint32_t* __restrict__ forced_list, uint32_t forced_cap,
                                                      const float* __restrict__ tris) {
  __shared__ uint32_t s_stack[kPStack];
  __shared__ float4 s_v0[kPCache], s_e1[kPCache], s_e2[kPCache];
  __shared__ float4 s_dir0;
  __shared__ uint32_t s_slot;
  trace_packet_job<1>(nodes, faces, verts, origins, dirs, width, height, node_count, face_count, faces_id, coords,
                      depths, (int)blockIdx.x, (int)blockIdx.y, handoff, handoff_cap, handoff_faces,
                      handoff + kHandoffHead + handoff_cap, forced_list, forced_cap, nullptr, tris, s_stack, s_v0,
                      s_e1, s_e2, &s_dir0, &s_slot, nullptr, nullptr, nullptr, nullptr);
}

// The second pass: the packets handed off, each resumed from its record,
// kPacketSplit threads per ray.
#ifndef RK_REF_PACKET_SPLIT
#define RK_REF_PACKET_SPLIT 5
#endif
constexpr int kPacketSplit = RK_REF_PACKET_SPLIT;

__global__ __launch_bounds__(kPT * kPacketSplit) void k_trace_packet_split(
    const rk::Node* __restrict__ nodes, const uint4* __restrict__ faces, const float4* __restrict__ verts,
    const float4* __restrict__ origins, const float4* __restrict__ dirs, int width, int height, int packets_x,
    uint32_t node_count, uint32_t face_count, const uint32_t* __restrict__ handoff, uint32_t handoff_cap,
    uint32_t* __restrict__ faces_id, float2* __restrict__ coords, float* __restrict__ depths,
    const float* __restrict__ tris) {
  __shared__ uint32_t s_stack[kPStack];
  __shared__ float4 s_v0[kPCache], s_e1[kPCache], s_e2[kPCache];
  __shared__ float4 s_dir0;
  __shared__ uint32_t s_slot;
  __shared__ float s_pt[kPT * kPacketSplit], s_pu[kPT * kPacketSplit], s_pv[kPT * kPacketSplit];
  __shared__ uint32_t s_pk[kPT * kPacketSplit];
  const uint32_t count = min(handoff[0], handoff_cap);
  for (uint32_t i = blockIdx.x; i < count; i += gridDim.x) {
    const uint32_t pk = handoff[kHandoffHead + i];
    const uint32_t* rec = handoff + kHandoffHead + handoff_cap + (size_t)i * kPStateWords;
    if (rec[3] == kDone) continue;  // traced by the forced walk
    trace_packet_job<kPacketSplit>(nodes, faces, verts, origins, dirs, width, height, node_count, face_count,
                                   faces_id, coords, depths, (int)(pk % (uint32_t)packets_x),
                                   (int)(pk / (uint32_t)packets_x), nullptr, 0u, 0u, nullptr, nullptr, 0u, rec,
                                   tris, s_stack, s_v0, s_e1, s_e2, &s_dir0, &s_slot, s_pt, s_pu, s_pv, s_pk);
    __syncthreads();  // the next packet rewrites the stack and s_dir0
  }
}

// The packets holding a ray that passes every box test (the forcing ray), by
// parts of kFRays rays on different CUs. While the forcing ray's depth is
// above 0 it passes every box that box_ray passes (at t = +-0), so the packet
// visits every such node, in the order its first ray gives, whatever its
// other rays hold: each part walks that order from the packet's record on its
// own, checking the forcing ray's box test at every node (a failure ends the
// part: diverged), and each of its rays tests every face of every leaf in face
// order against its own depth, kFParts threads per ray (consecutive faces of
// each chunk each, merged in part order). The results go to temp;
// k_packet_commit puts them in the outputs when no part diverged and the
// forcing ray's final depth is above 0 (then it was above 0 all along), and
// otherwise leaves the packet to k_trace_packet_split, from its record. On the
// 1920x1080 heightfield (149 such packets), parts of 48, 24, 16 and 12 rays
// gave 68.6, 47.4, 44.9 and 59.8 ms per frame (67.1 ms with all on one CU).
#ifndef RK_REF_FORCED_RAYS
#define RK_REF_FORCED_RAYS 16
#endif
constexpr int kFRays = RK_REF_FORCED_RAYS, kFThreads = kPT, kFParts = kFThreads / kFRays;
constexpr int kFPackets = kPT / kFRays;  // parts of a packet
constexpr uint32_t kForcedCap = 1024;    // forced packets per call
constexpr uint32_t kForcedGrid = 2048;

__global__ __launch_bounds__(kFThreads) void k_packet_forced(
    const rk::Node* __restrict__ nodes, const uint4* __restrict__ faces, const float4* __restrict__ verts,
    const float4* __restrict__ origins, const float4* __restrict__ dirs, int width, int height, int packets_x,
    uint32_t node_count, uint32_t face_count, const uint32_t* __restrict__ handoff, uint32_t handoff_cap,
    const uint32_t* __restrict__ forced_list, uint32_t forced_cap, float4* __restrict__ temp,
    uint32_t* __restrict__ fflags, const uint32_t* __restrict__ faces_id, const float2* __restrict__ coords,
    const float* __restrict__ depths, const float* __restrict__ tris) {
  __shared__ uint32_t s_stack[kPStack];
  __shared__ float4 s_v0[kPCache], s_e1[kPCache], s_e2[kPCache];
  __shared__ float s_pt[kFThreads], s_pu[kFThreads], s_pv[kFThreads];
  __shared__ uint32_t s_pk[kFThreads];
  const uint32_t nf = min(handoff[2], forced_cap);
  const uint32_t jobs = nf * (uint32_t)kFPackets;
  const uint32_t* state = handoff + kHandoffHead + handoff_cap;
  const uint32_t max_visits = 2 * node_count + 2;
  const int r = (int)threadIdx.x % kFRays, fp = (int)threadIdx.x / kFRays;
  for (uint32_t job = blockIdx.x; job < jobs; job += gridDim.x) {
    const uint32_t f = job % nf, part = job / nf;  // part k of every packet before part k + 1 of any
    const uint32_t slot = forced_list[f];
    const uint32_t* rec = state + (size_t)slot * kPStateWords;
    const uint32_t pk = handoff[kHandoffHead + slot];
    const int bx = (int)(pk % (uint32_t)packets_x), by = (int)(pk / (uint32_t)packets_x);
    const int q = (int)part * kFRays + r;  // the ray's place in the packet
    const int x = bx * kPW + q % kPW, y = by * kPH + q / kPW;
    const bool in_image = x < width && y < height;
    const uint32_t l = in_image ? (uint32_t)(y * width + x) : 0u;
    float4 orig = make_float4(0.f, 0.f, 0.f, 0.f), dir = orig;
    float depth = 0.0f;
    uint32_t fid = 0;
    float2 coord = make_float2(0.f, 0.f);
    if (in_image) {  // the hits so far, left in the outputs at the hand-off
      orig = origins[l];
      dir = dirs[l];
      depth = depths[l];
      fid = faces_id[l];
      coord = coords[l];
    }
    const uint32_t R = rec[2];
    const uint32_t lr = (uint32_t)((by * kPH + (int)R / kPW) * width + bx * kPW + (int)R % kPW);
    const float4 orig_r = origins[lr], invdir_r = invdir_of(dirs[lr]);
    const float4 dir0 = dirs[(uint32_t)(by * kPH) * (uint32_t)width + (uint32_t)(bx * kPW)];
    int sp = (int)rec[0];
    uint32_t visits = rec[1];
    for (int k = (int)threadIdx.x; k < sp; k += kFThreads) s_stack[k] = rec[4 + k];
    __syncthreads();
    bool diverged = false;
    while (sp > 0 && ++visits <= max_visits) {
      const uint32_t nid = s_stack[--sp];
      if (nid >= node_count) break;  // block-uniform; never taken for a tree built by this library
      const NodeRegs nd = load_node(nodes, nid);
      if (rk::node_is_leaf(nd.w0)) {
        const uint32_t tri_id = nd.w1;
        const uint32_t cnt = tri_id < face_count ? min(rk::leaf_count(nd.w0), face_count - tri_id) : 0u;
        for (uint32_t c0 = 0; c0 < cnt; c0 += kPCache) {
          const uint32_t cc = (cnt - c0) < (uint32_t)kPCache ? (cnt - c0) : (uint32_t)kPCache;
          __syncthreads();
          for (uint32_t k = threadIdx.x; k < cc; k += kFThreads) {
            float4 v0, e1, e2;
            tri_get(faces, verts, tris, tri_id + c0 + k, v0, e1, e2);
            s_v0[k] = v0;
            s_e1[k] = e1;
            s_e2[k] = e2;
          }
          __syncthreads();
          const uint32_t qn = (cc + kFParts - 1) / kFParts;
          const uint32_t k0 = (uint32_t)fp * qn, k1 = min(cc, k0 + qn);
          float pt = depth, pu = 0.0f, pv = 0.0f;
          uint32_t pkf = kNone;
          if (in_image) {
            for (uint32_t k = k0; k < k1; ++k) {
              float t, u, v;
              if (tri_ray(orig, dir, s_v0[k], s_e1[k], s_e2[k], t, u, v)) {
                if (t < pt) { pkf = k; pu = u; pv = v; pt = t; }
              }
            }
          }
          s_pt[threadIdx.x] = pt;
          s_pu[threadIdx.x] = pu;
          s_pv[threadIdx.x] = pv;
          s_pk[threadIdx.x] = pkf;
          __syncthreads();
          if (in_image) {
#pragma unroll
            for (int p2 = 0; p2 < kFParts; ++p2) {
              const int o = p2 * kFRays + r;
              const uint32_t k = s_pk[o];
              if (k != kNone && s_pt[o] < depth) {
                depth = s_pt[o];
                fid = tri_id + c0 + k;
                coord = make_float2(s_pu[o], s_pv[o]);
              }
            }
          }
        }
        __syncthreads();
      } else {
        float t;
        if (!box_ray(nd.bmin, nd.bmax, orig_r, invdir_r, t)) {  // block-uniform
          diverged = true;
          break;
        }
        uint32_t nearc, farc;
        children_order(nd.w0, nd.w1, dir0, nearc, farc);
        if (sp + 2 > kPStack) break;  // block-uniform; cannot occur for trees of depth <= 126
        __syncthreads();  // every thread has read the entry popped above
        if (threadIdx.x == 0) { s_stack[sp] = farc; s_stack[sp + 1] = nearc; }
        sp += 2;
        __syncthreads();
      }
    }
    if (in_image && fp == 0) temp[(size_t)f * kPT + (uint32_t)q] = make_float4(depth, coord.x, coord.y, __uint_as_float(fid));
    if (threadIdx.x == 0) fflags[f * kFPackets + part] = diverged ? 1u : 0u;
    __syncthreads();  // the next job rewrites the stack
  }
}

// The forced walks' results, per packet: into the outputs (and the packet's
// record marked kDone) when no part diverged and the forcing ray's final
// depth is above 0.
__global__ __launch_bounds__(kPT) void k_packet_commit(
    int width, int height, int packets_x, uint32_t* __restrict__ handoff, uint32_t handoff_cap,
    const uint32_t* __restrict__ forced_list, uint32_t forced_cap, const float4* __restrict__ temp,
    const uint32_t* __restrict__ fflags, uint32_t* __restrict__ faces_id, float2* __restrict__ coords,
    float* __restrict__ depths) {
  const uint32_t nf = min(handoff[2], forced_cap);
  uint32_t* state = handoff + kHandoffHead + handoff_cap;
  for (uint32_t f = blockIdx.x; f < nf; f += gridDim.x) {
    const uint32_t slot = forced_list[f];
    uint32_t* rec = state + (size_t)slot * kPStateWords;
    bool ok = temp[(size_t)f * kPT + rec[2]].x > 0.0f;
    for (int p = 0; p < kFPackets; ++p) ok = ok && fflags[f * kFPackets + p] == 0u;
    if (!ok) continue;  // block-uniform
    const uint32_t pk = handoff[kHandoffHead + slot];
    const int bx = (int)(pk % (uint32_t)packets_x), by = (int)(pk / (uint32_t)packets_x);
    const int q = (int)threadIdx.x;
    const int x = bx * kPW + q % kPW, y = by * kPH + q / kPW;
    if (x < width && y < height) {
      const uint32_t l = (uint32_t)(y * width + x);
      const float4 v = temp[(size_t)f * kPT + (uint32_t)q];
      depths[l] = v.x;
      coords[l] = make_float2(v.y, v.z);
      faces_id[l] = __float_as_uint(v.w);
    }
    __syncthreads();  // every thread has read the record before it is marked
    if (threadIdx.x == 0) rec[3] = kDone;
  }
}

}  // namespace

#ifndef RK_REF_PACKET_HANDOFF_FACES
#define RK_REF_PACKET_HANDOFF_FACES 4096
#endif
constexpr uint32_t kPacketGrid = 1024;  // blocks of the packet traversal's second pass

static std::atomic<uint32_t> g_ray_faces{RK_HANDOFF_AUTO};
static std::atomic<uint32_t> g_packet_faces{RK_HANDOFF_AUTO};
// RK_HANDOFF_AUTO: below 2^19 rays the first pass holds fewer waves than the
// GPU runs at once (about 5,000), and rays handed off early to a wave of their
// own finish sooner. On the 256x256 icosphere (65,536 rays) the per-ray frame
// took 0.85 ms past 256 faces against 1.69 ms past 4096, the packet frame
// 2.05 ms past 512 against 2.29; on the 1920x1080 heightfield 4096 is best
// (per-ray 8.1 ms against 11.1-11.5 ms past 256-2048; packets flat).
constexpr uint32_t kAutoRaysBelow = 1u << 19;
constexpr uint32_t kAutoSmallRayFaces = 256, kAutoSmallPacketFaces = 512;
static uint32_t limit_of(uint32_t set, uint32_t n_rays, uint32_t small_limit, uint32_t large_limit) {
  if (set != RK_HANDOFF_AUTO) return set;
  return n_rays < kAutoRaysBelow ? small_limit : large_limit;
}
static std::atomic<uint32_t> g_list_cap{kHandoffCap};
constexpr uint32_t kNoHandoff = 0xffffffffu;

// A memory pool per device for the hand-off buffers, which keeps what calls
// free (the default pool returns it at every synchronisation, and the next
// call would allocate again: 0.3-1 ms per call on the 1920x1080 frame).
static std::mutex g_pool_mutex;
static hipMemPool_t g_pools[64];
static bool g_pool_tried[64];

static hipMemPool_t handoff_pool() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return nullptr;
  std::lock_guard<std::mutex> lock(g_pool_mutex);
  if (!g_pool_tried[dev]) {
    g_pool_tried[dev] = true;
    hipMemPoolProps props{};
    props.allocType = hipMemAllocationTypePinned;
    props.handleTypes = hipMemHandleTypeNone;
    props.location.type = hipMemLocationTypeDevice;
    props.location.id = dev;
    hipMemPool_t pool = nullptr;
    if (hipMemPoolCreate(&pool, &props) == hipSuccess) {
      uint64_t keep = UINT64_MAX;
      if (hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep) == hipSuccess) {
        g_pools[dev] = pool;
      } else {
        (void)hipMemPoolDestroy(pool);
      }
    }
    (void)hipGetLastError();
  }
  return g_pools[dev];
}

// The hand-off buffer of one call (words 32-bit words), its counters cleared:
// a stream-ordered allocation, so that calls on other streams have their own.
// Without it (nullptr) every ray or packet stays in the first pass, with the
// same results.
static uint32_t* handoff_alloc(hipStream_t st, size_t words, uint32_t faces) {
  uint32_t* ho = nullptr;
  if (faces == kNoHandoff || words <= kHandoffHead) return nullptr;
  hipMemPool_t pool = handoff_pool();
  const size_t bytes = words * sizeof(uint32_t);
  const hipError_t e = pool ? hipMallocFromPoolAsync(reinterpret_cast<void**>(&ho), bytes, pool, st)
                            : hipMallocAsync(reinterpret_cast<void**>(&ho), bytes, st);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return nullptr;
  }
  hipLaunchKernelGGL(k_handoff_clear, dim3(1), dim3(64), 0, st, ho);
  return ho;
}

#ifndef RK_REF_TRI_RECORDS
#define RK_REF_TRI_RECORDS 1
#endif

// The call's triangle records (k_tri_records), from the same pool, or nullptr:
// the kernels then read the faces and vertices.
static float* tri_records_alloc(hipStream_t st, const uint4* faces, const float4* verts, uint32_t face_count) {
  if (!RK_REF_TRI_RECORDS) return nullptr;
  float* tris = nullptr;
  hipMemPool_t pool = handoff_pool();
  const size_t bytes = (size_t)face_count * kRecFloats * sizeof(float);
  const hipError_t e = pool ? hipMallocFromPoolAsync(reinterpret_cast<void**>(&tris), bytes, pool, st)
                            : hipMallocAsync(reinterpret_cast<void**>(&tris), bytes, st);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return nullptr;
  }
  hipLaunchKernelGGL(k_tri_records, dim3((face_count + 255) / 256), dim3(256), 0, st, faces, verts, face_count, tris);
  return tris;
}

extern "C" rk_status rk_raytrace_mesh_tree(rk_stream stream, const rk_bvnode* nodes, uint32_t node_count,
                                           const rk_uint4* faces, uint32_t face_count, const rk_float4* vertices,
                                           const rk_float4* origins,
                                           const rk_float4* directions, int width, int height, uint32_t* faces_id,
                                           rk_float2* coords, float* depths, int coherency) {
  if (width <= 0 || height <= 0) return RK_ERROR_INVALID_VALUE;
  if (!nodes || node_count == 0 || !faces || face_count == 0 || !vertices || !origins || !directions || !faces_id || !coords || !depths)
    return RK_ERROR_INVALID_VALUE;
  hipStream_t st = rk_hip_stream(stream);
  const uint32_t n = (uint32_t)width * (uint32_t)height;
  if (coherency) {
    // A packet whose leaves pass g_packet_faces faces, or one holding a ray
    // that passes every box test (the packet then visits every leaf, and all
    // its rays test every face; at its first leaf), is traced on by the second
    // pass, kPacketSplit threads per ray, from where it was left.
    const uint32_t gx = (uint32_t)(width + kPW - 1) / kPW, gy = (uint32_t)(height + kPH - 1) / kPH;
    const uint32_t faces_limit = limit_of(g_packet_faces.load(std::memory_order_relaxed), n, kAutoSmallPacketFaces,
                                          RK_REF_PACKET_HANDOFF_FACES);
    const uint32_t cap = std::min<uint32_t>(gx * gy, g_list_cap.load(std::memory_order_relaxed));
    const uint32_t fcap = std::min<uint32_t>(cap, kForcedCap);
    // [head | list | records | forced list | forced flags (4 each) | forced results (192 float4 each)]
    const size_t flist_off = kHandoffHead + (size_t)cap * (1u + kPStateWords);
    const size_t temp_off = (flist_off + (1ull + kFPackets) * fcap + 3) & ~size_t(3);
    uint32_t* ho = handoff_alloc(st, temp_off + 4ull * kPT * fcap, faces_limit);
    uint32_t* flist = ho ? ho + flist_off : nullptr;
    uint32_t* fflags = ho ? ho + flist_off + fcap : nullptr;
    float4* temp = ho ? reinterpret_cast<float4*>(ho + temp_off) : nullptr;
    float* tris = tri_records_alloc(st, reinterpret_cast<const uint4*>(faces),
                                     reinterpret_cast<const float4*>(vertices), face_count);
    hipLaunchKernelGGL(k_trace_packet, dim3(gx, gy), dim3(kPT), 0, st, reinterpret_cast<const rk::Node*>(nodes),
                       reinterpret_cast<const uint4*>(faces), reinterpret_cast<const float4*>(vertices),
                       reinterpret_cast<const float4*>(origins), reinterpret_cast<const float4*>(directions), width,
                       height, node_count, face_count, faces_id, reinterpret_cast<float2*>(coords), depths, ho,
                       ho ? cap : 0u, faces_limit, flist, fcap, tris);
    if (ho) {
      hipLaunchKernelGGL(k_packet_forced, dim3(std::min<uint32_t>(fcap * kFPackets, kForcedGrid)), dim3(kFThreads), 0,
                         st, reinterpret_cast<const rk::Node*>(nodes), reinterpret_cast<const uint4*>(faces),
                         reinterpret_cast<const float4*>(vertices), reinterpret_cast<const float4*>(origins),
                         reinterpret_cast<const float4*>(directions), width, height, (int)gx, node_count, face_count,
                         ho, cap, flist, fcap, temp, fflags, faces_id, reinterpret_cast<const float2*>(coords),
                         depths, tris);
      hipLaunchKernelGGL(k_packet_commit, dim3(fcap), dim3(kPT), 0, st, width, height, (int)gx, ho, cap, flist, fcap,
                         temp, fflags, faces_id, reinterpret_cast<float2*>(coords), depths);
      hipLaunchKernelGGL(k_trace_packet_split, dim3(std::min<uint32_t>(cap, kPacketGrid)), dim3(kPT * kPacketSplit),
                         0, st, reinterpret_cast<const rk::Node*>(nodes), reinterpret_cast<const uint4*>(faces),
                         reinterpret_cast<const float4*>(vertices), reinterpret_cast<const float4*>(origins),
                         reinterpret_cast<const float4*>(directions), width, height, (int)gx, node_count, face_count,
                         ho, cap, faces_id, reinterpret_cast<float2*>(coords), depths, tris);
      if (hipFreeAsync(ho, st) != hipSuccess) return RK_ERROR_DEVICE;
    }
    if (tris && hipFreeAsync(tris, st) != hipSuccess) return RK_ERROR_DEVICE;
  } else {
    // A ray whose leaves pass g_ray_faces faces, or a ray that passes every
    // box test (at its first leaf), is traced again by the second pass, one
    // wave per ray.
    const uint32_t faces_limit = limit_of(g_ray_faces.load(std::memory_order_relaxed), n, kAutoSmallRayFaces,
                                          RK_REF_HANDOFF_FACES);
    const uint32_t cap = std::min<uint32_t>(n, g_list_cap.load(std::memory_order_relaxed));
    const uint32_t ucap = std::min<uint32_t>(cap, kCrossCap);
    // [head | list | split rays: list, nunits, top_pops, diverged, buckets (8) | unit_node, unit_sp |
    //  results (8 words each)]
    const size_t res_off = (kHandoffHead + cap + 12ull * ucap + 2ull * ucap * kUnits + 7) & ~size_t(7);
    uint32_t* ho = handoff_alloc(st, res_off + 8ull * ucap * kUnits, faces_limit);
    UnitArgs ua{};
    if (ho) {
      uint32_t* p = ho + kHandoffHead + cap;
      ua.list = p;
      ua.nunits = p + ucap;
      ua.top_pops = p + 2 * ucap;
      ua.diverged = p + 3 * ucap;
      ua.bucket = p + 4 * ucap;
      ua.unit_node = p + 12 * ucap;
      ua.unit_sp = p + 12 * ucap + (size_t)ucap * kUnits;
      ua.res = reinterpret_cast<UnitResult*>(ho + res_off);
      ua.cap = ucap;
    }
    const rk::Node* nd = reinterpret_cast<const rk::Node*>(nodes);
    const uint4* fc = reinterpret_cast<const uint4*>(faces);
    const float4* vt = reinterpret_cast<const float4*>(vertices);
    const float4* org = reinterpret_cast<const float4*>(origins);
    const float4* drs = reinterpret_cast<const float4*>(directions);
    float2* uv = reinterpret_cast<float2*>(coords);
    float* tris = tri_records_alloc(st, fc, vt, face_count);
    hipLaunchKernelGGL(k_trace_ray<RK_REF_COOP_MAX>, dim3((n + 63) / 64), dim3(64), 0, st, nd, fc, vt, org,
                       drs, n, node_count, face_count, faces_id, uv, depths, ho, ho ? cap : 0u, faces_limit, ua,
                       tris);
    if (ho && ucap != 0u) {
      hipLaunchKernelGGL(k_units_split, dim3((ucap + 63) / 64), dim3(64), 0, st, nd, org, drs, depths, node_count,
                         ho, ua);
      const dim3 ugrid(std::min<uint32_t>(ucap * kUnits, kUnitGrid));
      if ((uint64_t)face_count >= (uint64_t)kUnitWaveFaces * ((node_count + 1) / 2))
        hipLaunchKernelGGL(k_units_walk<RK_REF_UNIT_WAVES>, ugrid, dim3(64 * RK_REF_UNIT_WAVES), 0, st, nd, fc, vt, org,
                           drs, depths, node_count, face_count, ho, ua, tris);
      else
        hipLaunchKernelGGL(k_units_walk<1>, ugrid, dim3(64), 0, st, nd, fc, vt, org, drs, depths, node_count,
                           face_count, ho, ua, tris);
      hipLaunchKernelGGL(k_units_merge, dim3(std::min<uint32_t>(ucap, kWaveGrid)), dim3(64), 0, st, nd, fc, vt, org,
                         drs, node_count, face_count, ho, ua, faces_id, uv, depths, tris);
    }
    if (ho) {
      hipLaunchKernelGGL(k_trace_ray_wave, dim3(std::min<uint32_t>(cap, kRayWaveGrid)), dim3(64), 0, st, nd, fc, vt,
                         org, drs, node_count, face_count, ho, cap, faces_id, uv, depths, tris);
      if (hipFreeAsync(ho, st) != hipSuccess) return RK_ERROR_DEVICE;
    }
    if (tris && hipFreeAsync(tris, st) != hipSuccess) return RK_ERROR_DEVICE;
  }
  RK_CHECK_LAUNCH();
  return RK_SUCCESS;
}

extern "C" rk_status rk_raytrace_set_handoff(uint32_t ray_faces, uint32_t packet_faces, uint32_t list_cap) {
  if (list_cap == 0 || list_cap > kHandoffCap) return RK_ERROR_INVALID_VALUE;
  g_ray_faces.store(ray_faces, std::memory_order_relaxed);
  g_packet_faces.store(packet_faces, std::memory_order_relaxed);
  g_list_cap.store(list_cap, std::memory_order_relaxed);
  return RK_SUCCESS;
}

extern "C" rk_status rk_raytrace_get_handoff(uint32_t* ray_faces, uint32_t* packet_faces, uint32_t* list_cap) {
  if (!ray_faces || !packet_faces || !list_cap) return RK_ERROR_INVALID_VALUE;
  *ray_faces = g_ray_faces.load(std::memory_order_relaxed);
  *packet_faces = g_packet_faces.load(std::memory_order_relaxed);
  *list_cap = g_list_cap.load(std::memory_order_relaxed);
  return RK_SUCCESS;
}
