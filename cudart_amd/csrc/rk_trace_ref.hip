// Closest-hit traversal of the reference tree layout (bvnode, 32 B) with the
// reference arithmetic (CUDA_raytraceMeshTree.cu), for gfx950:
//   k_trace_ray     per-ray DFS (kernel_raytraceMeshTree, :111-219), one lane
//                   per ray, a leaf's faces fetched two at a time, stack in
//                   LDS and scratch;
//   k_trace_packet  the 16x12 packet traversal (kernel_raytraceMeshTreePacket,
//                   :222-377), one 192-thread workgroup (3 waves) per tile,
//                   block-uniform stack in LDS, one barrier vote per node.
// Built with -ffp-contract=off, so the triangle and box tests are the same
// float operations in the same order as the reference source.
#include "rk_device.h"

namespace {

constexpr float kEps = 0.000001f;  // CUDA_raytraceMeshTree.cu:36

// intersectTriRay, CUDA_raytraceMeshTree.cu:57-89.
__device__ __forceinline__ bool tri_ray(float4 orig, float4 dir, float4 v0, float4 e1, float4 e2, float& t, float& u,
                                        float& v) {
  float4 pvec = rk::cross4(dir, e2);
  float det = rk::dot4(e1, pvec);
  if (det > -kEps && det < kEps) return false;
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
constexpr int kLdsStack = 32;  // entries per lane in LDS (8 KB per wave); the rest in scratch

// The per-ray traversal (kernel_raytraceMeshTree, :111-219), one lane per ray:
// pop a node, test its box against the depth so far, push the far child, then
// the near one; at a leaf, test its faces in order, keeping a hit only when
// t < depth. The same operations in the same order as the reference, so the
// results are bit-identical.
//
// A leaf's faces are fetched kLeafBatch at a time, all their loads issued
// before the first test: a face costs two dependent loads (the face, then
// its vertices), and a leaf holds up to a few hundred faces, so one face at a
// time leaves each lane waiting on memory. The tests still run in face order.
// On the 1920x1080 heightfield frame (depth-12 tree): 291 ms per frame with
// batches of 2, 315 with 4, 294 with 8, against 395 ms one face at a time
// with the whole stack in LDS (6 waves per CU). Testing each leaf with the
// whole wave, 64 faces at a time, was slower (458 ms): on terrain rows every
// lane reaches its leaves together, and the wave then walks them one by one.
//
// The stack keeps the reference's 96 entries and its overflow rule; the first
// kLdsStack entries of each lane live in LDS (entry e of lane x at
// e * 64 + x), deeper ones in scratch (trees of depth <= 31 need at most 33),
// so a CU holds 20 waves of this kernel instead of 6.
#ifndef RK_REF_LEAF_BATCH
#define RK_REF_LEAF_BATCH 2
#endif
constexpr int kLeafBatch = RK_REF_LEAF_BATCH;

__global__ __launch_bounds__(64) void k_trace_ray(const rk::Node* __restrict__ nodes,
                                                  const uint4* __restrict__ faces,
                                                  const float4* __restrict__ verts,
                                                  const float4* __restrict__ origins,
                                                  const float4* __restrict__ dirs, uint32_t n_rays,
                                                  uint32_t node_count, uint32_t face_count,
                                                  uint32_t* __restrict__ faces_id, float2* __restrict__ coords,
                                                  float* __restrict__ depths) {
  __shared__ uint32_t s_stack[kLdsStack * 64];
  const uint32_t l = blockIdx.x * 64 + threadIdx.x;
  if (l >= n_rays) return;
  const float4 orig = origins[l], dir = dirs[l];
  const float4 invdir = invdir_of(dir);
  float depth = depths[l];
  uint32_t fid = faces_id[l];
  float2 coord = coords[l];
  uint32_t spill[kStack - kLdsStack];
  typedef __attribute__((address_space(3))) uint32_t lds_u32;
  lds_u32* my_lds = (lds_u32*)(s_stack + threadIdx.x);
  int sp = 0;
  my_lds[0] = 0u;
  sp = 1;
  uint32_t visits = 0;
  const uint32_t max_visits = 2 * node_count + 2;
  while (sp > 0 && ++visits <= max_visits) {
    --sp;
    const uint32_t nid = sp < kLdsStack ? my_lds[sp * 64] : spill[sp - kLdsStack];
    if (nid >= node_count) break;  // never taken for a tree built by this library
    const NodeRegs nd = load_node(nodes, nid);
    if (rk::node_is_leaf(nd.w0)) {
      const uint32_t tri_id = nd.w1;
      const uint32_t cnt = tri_id < face_count ? min(rk::leaf_count(nd.w0), face_count - tri_id) : 0u;
      for (uint32_t k0 = 0; k0 < cnt; k0 += kLeafBatch) {
        uint4 f[kLeafBatch];
#pragma unroll
        for (int b = 0; b < kLeafBatch; ++b) f[b] = k0 + b < cnt ? faces[tri_id + k0 + b] : make_uint4(0, 0, 0, 0);
        float4 p0[kLeafBatch], p1[kLeafBatch], p2[kLeafBatch];
#pragma unroll
        for (int b = 0; b < kLeafBatch; ++b) {
          if (k0 + b < cnt) {
            p0[b] = verts[f[b].x];
            p1[b] = verts[f[b].y];
            p2[b] = verts[f[b].z];
          }
        }
#pragma unroll
        for (int b = 0; b < kLeafBatch; ++b) {
          if (k0 + b < cnt) {
            // tri_setup's arithmetic: v0, v1 - v0, v2 - v0, w cleared
            float4 v0 = p0[b];
            float4 e1 = rk::sub4(p1[b], v0), e2 = rk::sub4(p2[b], v0);
            v0.w = 0.0f;
            e1.w = 0.0f;
            e2.w = 0.0f;
            float t, u, v;
            if (tri_ray(orig, dir, v0, e1, e2, t, u, v)) {
              if (t < depth) { fid = tri_id + k0 + b; coord = make_float2(u, v); depth = t; }
            }
          }
        }
      }
    } else {
      float t;
      if (box_ray(nd.bmin, nd.bmax, orig, invdir, t)) {
        if (t < depth) {
          uint32_t nearc, farc;
          children_order(nd.w0, nd.w1, dir, nearc, farc);
          if (sp + 2 > kStack) break;  // depth > 94: cannot occur for trees of depth <= 94
          if (sp < kLdsStack) my_lds[sp * 64] = farc;
          else spill[sp - kLdsStack] = farc;
          ++sp;
          if (sp < kLdsStack) my_lds[sp * 64] = nearc;
          else spill[sp - kLdsStack] = nearc;
          ++sp;
        }
      }
    }
  }
  faces_id[l] = fid;
  coords[l] = coord;
  depths[l] = depth;
}

constexpr int kPW = 16, kPH = 12, kPT = kPW * kPH;  // PACKET_BLOCK_*, :47-49
constexpr int kPStack = 128;                       // PACKET_STACK_SIZE, :51
constexpr int kPCache = 224;                       // TRIANGLES_MAX, :39

__global__ __launch_bounds__(kPT) void k_trace_packet(const rk::Node* __restrict__ nodes,
                                                      const uint4* __restrict__ faces,
                                                      const float4* __restrict__ verts,
                                                      const float4* __restrict__ origins,
                                                      const float4* __restrict__ dirs, int width, int height,
                                                      uint32_t node_count, uint32_t face_count,
                                                      uint32_t* __restrict__ faces_id, float2* __restrict__ coords,
                                                      float* __restrict__ depths) {
  __shared__ uint32_t s_stack[kPStack];
  __shared__ float4 s_v0[kPCache], s_e1[kPCache], s_e2[kPCache];
  __shared__ float4 s_dir0;
  const int tx = threadIdx.x % kPW, ty = threadIdx.x / kPW;
  const int x = blockIdx.x * kPW + tx, y = blockIdx.y * kPH + ty;
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
  if (threadIdx.x == 0) {  // thread (0,0) of a block is always inside the image
    s_dir0 = dir;
    s_stack[0] = 0;
  }
  int sp = 1;
  __syncthreads();
  const float4 dir0 = s_dir0;
  uint32_t visits = 0;
  const uint32_t max_visits = 2 * node_count + 2;
  while (sp > 0 && ++visits <= max_visits) {
    const uint32_t nid = s_stack[--sp];
    if (nid >= node_count) break;  // block-uniform; never taken for a tree built by this library
    const NodeRegs nd = load_node(nodes, nid);  // block-uniform address
    if (rk::node_is_leaf(nd.w0)) {
      const uint32_t tri_id = nd.w1;
      const uint32_t cnt = tri_id < face_count ? min(rk::leaf_count(nd.w0), face_count - tri_id) : 0u;
      for (uint32_t c0 = 0; c0 < cnt; c0 += kPCache) {
        const uint32_t cc = (cnt - c0) < (uint32_t)kPCache ? (cnt - c0) : (uint32_t)kPCache;
        __syncthreads();
        for (uint32_t k = threadIdx.x; k < cc; k += kPT) {
          float4 v0, e1, e2;
          tri_setup(faces, verts, tri_id + c0 + k, v0, e1, e2);
          s_v0[k] = v0; s_e1[k] = e1; s_e2[k] = e2;
        }
        __syncthreads();
        if (in_image) {
          for (uint32_t k = 0; k < cc; ++k) {
            float t, u, v;
            if (tri_ray(orig, dir, s_v0[k], s_e1[k], s_e2[k], t, u, v)) {
              if (t < depth) { fid = tri_id + c0 + k; coord = make_float2(u, v); depth = t; }
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
  if (in_image) {
    faces_id[l] = fid;
    coords[l] = coord;
    depths[l] = depth;
  }
}

}  // namespace

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
    dim3 grid((width + kPW - 1) / kPW, (height + kPH - 1) / kPH);
    hipLaunchKernelGGL(k_trace_packet, grid, dim3(kPT), 0, st, reinterpret_cast<const rk::Node*>(nodes),
                       reinterpret_cast<const uint4*>(faces), reinterpret_cast<const float4*>(vertices),
                       reinterpret_cast<const float4*>(origins), reinterpret_cast<const float4*>(directions), width,
                       height, node_count, face_count, faces_id, reinterpret_cast<float2*>(coords), depths);
  } else {
    hipLaunchKernelGGL(k_trace_ray, dim3((n + 63) / 64), dim3(64), 0, st,
                       reinterpret_cast<const rk::Node*>(nodes), reinterpret_cast<const uint4*>(faces),
                       reinterpret_cast<const float4*>(vertices), reinterpret_cast<const float4*>(origins),
                       reinterpret_cast<const float4*>(directions), n, node_count, face_count, faces_id, reinterpret_cast<float2*>(coords),
                       depths);
  }
  RK_CHECK_LAUNCH();
  return RK_SUCCESS;
}
