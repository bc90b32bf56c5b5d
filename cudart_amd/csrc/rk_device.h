// Device-side helpers shared by the gfx950 kernels. Every kernel file is built
// with -ffp-contract=off, so each float expression below is evaluated as the
// IEEE single-precision operations it spells, in order; where a kernel wants
// an FMA it calls __builtin_fmaf explicitly.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/rk.h"

#define RK_HUGE 3.402823466e+38f  // CUDART_NORM_HUGE_F (FLT_MAX), CUDABox.h:50-54

namespace rk {

struct Box { float4 mn, mx; };  // aabox, CUDABox.h:17-148

struct Node {  // bvnode, CUDANode.h:20-48
  uint32_t w0, w1;
  float bmin[3], bmax[3];
};

static_assert(sizeof(Box) == 32, "aabox size");
static_assert(sizeof(Node) == 32, "bvnode size");

// IEEE minNum/maxNum (v_min_f32 / v_max_f32): a NaN operand yields the other.
__device__ __forceinline__ float fmn(float a, float b) { return __builtin_fminf(a, b); }
__device__ __forceinline__ float fmx(float a, float b) { return __builtin_fmaxf(a, b); }

__device__ __forceinline__ float4 sub4(float4 a, float4 b) { return make_float4(a.x - b.x, a.y - b.y, a.z - b.z, a.w - b.w); }
__device__ __forceinline__ float4 add4(float4 a, float4 b) { return make_float4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w); }
__device__ __forceinline__ float4 mul4(float4 a, float4 b) { return make_float4(a.x * b.x, a.y * b.y, a.z * b.z, a.w * b.w); }
__device__ __forceinline__ float4 muls4(float4 a, float s) { return make_float4(a.x * s, a.y * s, a.z * s, a.w * s); }
__device__ __forceinline__ float4 min4(float4 a, float4 b) { return make_float4(fmn(a.x, b.x), fmn(a.y, b.y), fmn(a.z, b.z), fmn(a.w, b.w)); }
__device__ __forceinline__ float4 max4(float4 a, float4 b) { return make_float4(fmx(a.x, b.x), fmx(a.y, b.y), fmx(a.z, b.z), fmx(a.w, b.w)); }
// dot: ((x*x + y*y) + z*z) + w*w, types/_cuda_float4.h:171-174.
__device__ __forceinline__ float dot4(float4 a, float4 b) { return a.x * b.x + a.y * b.y + a.z * b.z + a.w * b.w; }
// cross with w = 0, types/_cuda_float4.h:198-208.
__device__ __forceinline__ float4 cross4(float4 a, float4 b) {
  return make_float4(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x, 0.0f);
}

// Truncating float->uint with saturation and NaN -> 0 (the semantics of the
// reference's (uint) cast on its GPU), written out so no out-of-range
// conversion reaches the optimiser.
__device__ __forceinline__ uint32_t f2u_sat(float f) {
  if (!(f > 0.0f)) return 0u;
  if (f >= 4294967296.0f) return 0xFFFFFFFFu;
  return (uint32_t)f;
}
__device__ __forceinline__ int f2i_sat(float f) {
  if (f != f) return 0;
  if (f >= 2147483648.0f) return 2147483647;
  if (f <= -2147483648.0f) return (-2147483647 - 1);
  return (int)f;
}

__device__ __forceinline__ Box load_box(const Box* p) {
  const float4* q = reinterpret_cast<const float4*>(p);
  Box b; b.mn = q[0]; b.mx = q[1];
  return b;
}

__device__ __forceinline__ uint32_t node_is_leaf(uint32_t w0) { return w0 & 1u; }
__device__ __forceinline__ uint32_t leaf_count(uint32_t w0) { return w0 >> 1; }
__device__ __forceinline__ uint32_t node_axis(uint32_t w0) { return (w0 >> 1) & 3u; }

// Wave64 helpers.
__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }
__device__ __forceinline__ uint64_t lanemask_lt() {
  uint32_t l = __lane_id();
  return l == 0 ? 0ull : (~0ull >> (64 - l));
}

// Relaxed device-scope atomics used for look-back status words.
__device__ __forceinline__ uint32_t atomic_load_relaxed(const uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void atomic_store_relaxed(uint32_t* p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Ray of thread t when each wave takes runs of kRun consecutive rays from
// 64 / kRun equal bands of an array of n_rays, the run at the same place in
// each band (n_rays when thread t has none; kRun 0: ray t). A bijection of
// [0, n_rays) onto the threads below 64 * ceil(n_rays / 64).
template <uint32_t kRun>
__device__ __forceinline__ uint32_t band_ray(uint32_t t, uint32_t n_rays) {
  static_assert(kRun == 0 || 64 % kRun == 0, "runs tile a wave");
  if (kRun == 0) return t < n_rays ? t : n_rays;
  const uint32_t waves = (n_rays + 63) / 64, w = t >> 6, q = t & 63;
  const uint32_t r = ((q / kRun) * waves + w) * kRun + q % kRun;  // < 64 * waves
  return (w < waves && r < n_rays) ? r : n_rays;
}

// Attribute interpolation at barycentrics (u, v) with a = (1 - u) - v:
// (a * x0 + u * x1) + v * x2 per component, the reference's float4 operator
// order (types/_cuda_float4.h:80-111).
__device__ __forceinline__ float4 bary4(float a, float u, float v, float4 x0, float4 x1, float4 x2) {
  return add4(add4(muls4(x0, a), muls4(x1, u)), muls4(x2, v));
}
// v * (1 / sqrtf(dot(v, v))), the reference's normalize.
__device__ __forceinline__ float4 normalize4(float4 v) {
  const float inv = 1.0f / sqrtf(dot4(v, v));
  return muls4(v, inv);
}
// The normal of face f at (u, v): interpolated vertex normals, normalised, or
// without them the face normal normalize(cross(v1 - v0, v2 - v0))
// (kernel_interpolateData, CUDA_interpolateData.cu:45-141).
__device__ __forceinline__ float4 face_normal_at(const uint4 f, float u, float v, const float4* __restrict__ verts,
                                                 const float4* __restrict__ norms) {
  if (norms) return normalize4(bary4((1.0f - u) - v, u, v, norms[f.x], norms[f.y], norms[f.z]));
  const float4 v0 = verts[f.x], v1 = verts[f.y], v2 = verts[f.z];
  return normalize4(cross4(sub4(v1, v0), sub4(v2, v0)));
}
// make_pixel (kernels/CUDAShaderCommon.cuh): scale by 255, clamp to [0, 255]
// with fminf/fmaxf, truncate, pack as A << 24 | R << 16 | G << 8 | B.
__device__ __forceinline__ uint32_t make_pixel(float4 v) {
  v = muls4(v, 255.0f);
  v = make_float4(fmaxf(fminf(v.x, 255.0f), 0.0f), fmaxf(fminf(v.y, 255.0f), 0.0f), fmaxf(fminf(v.z, 255.0f), 0.0f),
                  fmaxf(fminf(v.w, 255.0f), 0.0f));
  return (((uint32_t)v.x & 0xFFu) << 16) | (((uint32_t)v.y & 0xFFu) << 8) | ((uint32_t)v.z & 0xFFu) |
         (((uint32_t)v.w & 0xFFu) << 24);
}
// kernel_shade's pixel (CUDAShaderKernel.cu:14-43) for a hit with normal n
// (w ignored) seen along cam (w = 0): grey -dot(cam, n).
__device__ __forceinline__ uint32_t shade_hit(float4 cam, float4 n) {
  n.w = 0.0f;
  const float c = -dot4(cam, n);
  return make_pixel(make_float4(c, c, c, c));
}

}  // namespace rk

// Host-side launch check: every entry point returns RK_ERROR_LAUNCH when a
// launch reports an error.
#define RK_CHECK_LAUNCH()                                   \
  do {                                                      \
    hipError_t _e = hipGetLastError();                      \
    if (_e != hipSuccess) return RK_ERROR_LAUNCH;           \
  } while (0)

#define RK_HIP(call)                                        \
  do {                                                      \
    hipError_t _e = (call);                                 \
    if (_e != hipSuccess) return RK_ERROR_DEVICE;           \
  } while (0)

static inline hipStream_t rk_hip_stream(rk_stream s) { return reinterpret_cast<hipStream_t>(s); }

// Workspace carving: aligned sub-buffers from one caller-provided block.
struct RkArena {
  char* base;
  size_t size;
  size_t used;
  explicit RkArena(void* p, size_t n) : base(static_cast<char*>(p)), size(n), used(0) {}
  template <typename T>
  T* take(size_t count) {
    size_t off = (used + 255) & ~size_t(255);
    size_t bytes = count * sizeof(T);
    if (off + bytes > size) return nullptr;
    used = off + bytes;
    return reinterpret_cast<T*>(base + off);
  }
};
static inline size_t rk_align_up(size_t x) { return (x + 255) & ~size_t(255); }
