// Runtime entry points: device memory, streams, surface fills, primary rays.
#include <new>

#include "rk_device.h"

extern "C" const char* rk_status_string(rk_status s) {
  switch (s) {
    case RK_SUCCESS: return "success";
    case RK_ERROR_INVALID_VALUE: return "invalid value";
    case RK_ERROR_OUT_OF_MEMORY: return "out of device memory";
    case RK_ERROR_LAUNCH: return "kernel launch failed";
    case RK_ERROR_WORKSPACE: return "workspace too small";
    case RK_ERROR_DEVICE: return "device runtime error";
    case RK_ERROR_UNSUPPORTED: return "unsupported";
    default: return "unknown status";
  }
}

extern "C" rk_status rk_device_count(int* count) {
  if (!count) return RK_ERROR_INVALID_VALUE;
  int c = 0;
  if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
  *count = c;
  return RK_SUCCESS;
}

extern "C" rk_status rk_malloc(void** ptr, size_t bytes) {
  if (!ptr || bytes == 0) return RK_ERROR_INVALID_VALUE;
  if (hipMalloc(ptr, bytes) != hipSuccess) { *ptr = nullptr; return RK_ERROR_OUT_OF_MEMORY; }
  return RK_SUCCESS;
}

extern "C" rk_status rk_free(void* ptr) {
  if (!ptr) return RK_SUCCESS;
  RK_HIP(hipFree(ptr));
  return RK_SUCCESS;
}

extern "C" rk_status rk_malloc_host(void** ptr, size_t bytes) {
  if (!ptr || bytes == 0) return RK_ERROR_INVALID_VALUE;
  if (hipHostMalloc(ptr, bytes, hipHostMallocDefault) != hipSuccess) { *ptr = nullptr; return RK_ERROR_OUT_OF_MEMORY; }
  return RK_SUCCESS;
}

extern "C" rk_status rk_free_host(void* ptr) {
  if (!ptr) return RK_SUCCESS;
  RK_HIP(hipHostFree(ptr));
  return RK_SUCCESS;
}

extern "C" rk_status rk_memcpy(void* dst, const void* src, size_t bytes, int kind, rk_stream stream) {
  if (bytes == 0) return RK_SUCCESS;
  if (!dst || !src) return RK_ERROR_INVALID_VALUE;
  hipMemcpyKind k;
  switch (kind) {
    case 0: k = hipMemcpyHostToDevice; break;
    case 1: k = hipMemcpyDeviceToHost; break;
    case 2: k = hipMemcpyDeviceToDevice; break;
    default: return RK_ERROR_INVALID_VALUE;
  }
  if (stream == nullptr) {
    RK_HIP(hipMemcpy(dst, src, bytes, k));
  } else {
    RK_HIP(hipMemcpyAsync(dst, src, bytes, k, rk_hip_stream(stream)));
  }
  return RK_SUCCESS;
}

extern "C" rk_status rk_stream_create(rk_stream* stream) {
  if (!stream) return RK_ERROR_INVALID_VALUE;
  hipStream_t s;
  RK_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  *stream = reinterpret_cast<rk_stream>(s);
  return RK_SUCCESS;
}

extern "C" rk_status rk_stream_destroy(rk_stream stream) {
  if (!stream) return RK_SUCCESS;
  RK_HIP(hipStreamDestroy(rk_hip_stream(stream)));
  return RK_SUCCESS;
}

extern "C" rk_status rk_stream_synchronize(rk_stream stream) {
  RK_HIP(hipStreamSynchronize(rk_hip_stream(stream)));
  return RK_SUCCESS;
}

// ---- fills (kernel_fillSurface, CUDA_fillSurface.cu:17) ----------------------
template <typename T>
__global__ __launch_bounds__(256) void k_fill(T* __restrict__ dst, size_t count, T value) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < count; i += stride) dst[i] = value;
}

template <typename T>
static rk_status launch_fill(rk_stream stream, T* dst, size_t count, T value) {
  if (count == 0) return RK_SUCCESS;
  if (!dst) return RK_ERROR_INVALID_VALUE;
  size_t blocks = (count + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(k_fill<T>, dim3((unsigned)blocks), dim3(256), 0, rk_hip_stream(stream), dst, count, value);
  RK_CHECK_LAUNCH();
  return RK_SUCCESS;
}

extern "C" rk_status rk_fill_u32(rk_stream stream, uint32_t* dst, size_t count, uint32_t value) {
  return launch_fill<uint32_t>(stream, dst, count, value);
}
extern "C" rk_status rk_fill_f32(rk_stream stream, float* dst, size_t count, float value) {
  return launch_fill<float>(stream, dst, count, value);
}
extern "C" rk_status rk_fill_f2(rk_stream stream, rk_float2* dst, size_t count, rk_float2 value) {
  return launch_fill<float2>(stream, reinterpret_cast<float2*>(dst), count, make_float2(value.x, value.y));
}
extern "C" rk_status rk_fill_f4(rk_stream stream, rk_float4* dst, size_t count, rk_float4 value) {
  return launch_fill<float4>(stream, reinterpret_cast<float4*>(dst), count,
                             make_float4(value.x, value.y, value.z, value.w));
}

// ---- primary rays (kernel_calcPrimaryRays, CUDA_calcPrimaryRays.cu:20-32) ----
// dir = rayDir - (float)x * rayX - (float)y * rayY, evaluated left to right.
__global__ __launch_bounds__(256) void k_primary_rays(float4* __restrict__ origins, float4* __restrict__ dirs,
                                                      int width, int height, float4 ro, float4 rd, float4 rx,
                                                      float4 ry) {
  int x = blockIdx.x * 16 + (threadIdx.x & 15);
  int y = blockIdx.y * 16 + (threadIdx.x >> 4);
  if (x >= width || y >= height) return;
  int n = y * width + x;
  float fx = (float)x, fy = (float)y;
  float4 d = rk::sub4(rk::sub4(rd, rk::muls4(rx, fx)), rk::muls4(ry, fy));
  origins[n] = ro;
  dirs[n] = d;
}

extern "C" rk_status rk_calc_primary_rays(rk_stream stream, int width, int height, rk_float4 ray_orig,
                                          rk_float4 ray_dir, rk_float4 ray_x, rk_float4 ray_y, rk_float4* origins,
                                          rk_float4* directions) {
  if (width <= 0 || height <= 0 || !origins || !directions) return RK_ERROR_INVALID_VALUE;
  dim3 grid((width + 15) / 16, (height + 15) / 16);
  hipLaunchKernelGGL(k_primary_rays, grid, dim3(256), 0, rk_hip_stream(stream), reinterpret_cast<float4*>(origins),
                     reinterpret_cast<float4*>(directions), width, height,
                     make_float4(ray_orig.x, ray_orig.y, ray_orig.z, ray_orig.w),
                     make_float4(ray_dir.x, ray_dir.y, ray_dir.z, ray_dir.w),
                     make_float4(ray_x.x, ray_x.y, ray_x.z, ray_x.w), make_float4(ray_y.x, ray_y.y, ray_y.z, ray_y.w));
  RK_CHECK_LAUNCH();
  return RK_SUCCESS;
}

// ---- graphs ------------------------------------------------------------------
// A captured sequence of rk_* calls, replayed with one launch.
struct rk_graph_st {
  hipGraph_t graph;
  hipGraphExec_t exec;
};

extern "C" rk_status rk_graph_capture_begin(rk_stream stream) {
  if (!stream) return RK_ERROR_INVALID_VALUE;  // the default stream cannot be captured
  RK_HIP(hipStreamBeginCapture(rk_hip_stream(stream), hipStreamCaptureModeThreadLocal));
  return RK_SUCCESS;
}

extern "C" rk_status rk_graph_capture_end(rk_stream stream, rk_graph* graph) {
  if (!stream || !graph) return RK_ERROR_INVALID_VALUE;
  hipGraph_t g = nullptr;
  RK_HIP(hipStreamEndCapture(rk_hip_stream(stream), &g));
  hipGraphExec_t e = nullptr;
  hipError_t err = hipGraphInstantiate(&e, g, nullptr, nullptr, 0);
  if (err != hipSuccess) {
    (void)hipGraphDestroy(g);
    return RK_ERROR_LAUNCH;
  }
  rk_graph_st* h = new (std::nothrow) rk_graph_st{g, e};
  if (!h) {
    (void)hipGraphExecDestroy(e);
    (void)hipGraphDestroy(g);
    return RK_ERROR_OUT_OF_MEMORY;
  }
  *graph = h;
  return RK_SUCCESS;
}

extern "C" rk_status rk_graph_launch(rk_graph graph, rk_stream stream) {
  if (!graph) return RK_ERROR_INVALID_VALUE;
  RK_HIP(hipGraphLaunch(graph->exec, rk_hip_stream(stream)));
  return RK_SUCCESS;
}

extern "C" rk_status rk_graph_destroy(rk_graph graph) {
  if (!graph) return RK_SUCCESS;
  (void)hipGraphExecDestroy(graph->exec);
  (void)hipGraphDestroy(graph->graph);
  delete graph;
  return RK_SUCCESS;
}

// ---- autotuning -----------------------------------------------------------------
// Candidates: leaf_max 1/2/4/8 for the build; for the trace, rays as given or
// sorted, both modes, three workgroup sizes, both exact stack depths and
// donation at 16 or 32 idle lanes. Each candidate is timed on the
// caller's batch with HIP events on the caller's stream (median of reps).
namespace {
float median_of(float* v, uint32_t n) {
  for (uint32_t i = 1; i < n; ++i)
    for (uint32_t j = i; j > 0 && v[j - 1] > v[j]; --j) {
      float t = v[j]; v[j] = v[j - 1]; v[j - 1] = t;
    }
  return v[n / 2];
}
}  // namespace

extern "C" rk_status rk_batch_autotune(rk_stream stream, const rk_batch_mesh* meshes, uint32_t mesh_count,
                                       const rk_uint4* faces, const rk_float4* vertices, uint32_t face_count,
                                       const rk_aabox* hitboxes, const rk_fast_bvh* bvh, uint32_t* info,
                                       const rk_float4* origins, const rk_float4* directions, uint32_t ray_count,
                                       uint32_t* faces_id, rk_float2* coords, float* depths,
                                       const rk_trace_options* options, uint32_t reps, rk_batch_config* best,
                                       float* best_ms, void* ws, size_t ws_bytes) {
  if (!best || reps == 0 || reps > 64) return RK_ERROR_INVALID_VALUE;
  const size_t build_bytes = rk_batch_build_workspace_size(mesh_count, face_count);
  if (ws_bytes < build_bytes + rk_batch_trace_workspace_size(mesh_count, ray_count)) return RK_ERROR_WORKSPACE;
  void* trace_ws = static_cast<char*>(ws) + build_bytes;
  const size_t trace_bytes = ws_bytes - build_bytes;
  hipStream_t st = rk_hip_stream(stream);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  RK_HIP(hipEventCreate(&e0));
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    return RK_ERROR_DEVICE;
  }
  static const uint32_t leaves[] = {1, 2, 4, 8};
  static const uint32_t wgs[] = {64, 128, 256};
  static const uint32_t stacks[] = {64, 96};
  static const uint32_t donates[] = {16, 32};
  float samples[64];
  float best_total = 3.4e38f, best_build = 0.0f;
  rk_batch_config pick = rk_batch_config_default();
  rk_status s = RK_SUCCESS;
  auto timed = [&](auto&& call, float* out_ms) -> rk_status {
    for (uint32_t r = 0; r < reps; ++r) {
      if (hipEventRecord(e0, st) != hipSuccess) return RK_ERROR_DEVICE;
      rk_status cs = call();
      if (cs != RK_SUCCESS) return cs;
      if (hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) return RK_ERROR_DEVICE;
      if (hipEventElapsedTime(&samples[r], e0, e1) != hipSuccess) return RK_ERROR_DEVICE;
    }
    *out_ms = median_of(samples, reps);
    return RK_SUCCESS;
  };
  for (uint32_t li = 0; li < 4 && s == RK_SUCCESS; ++li) {
    rk_batch_config c = rk_batch_config_default();
    c.leaf_max = leaves[li];
    float build_ms = 0.0f;
    s = timed([&] { return rk_batch_build(stream, meshes, mesh_count, faces, vertices, face_count, hitboxes, &c, bvh,
                                          info, ws, build_bytes); },
              &build_ms);
    for (uint32_t oi = 0; oi < 2 && s == RK_SUCCESS; ++oi)
    for (uint32_t mode = 0; mode < 2 && s == RK_SUCCESS; ++mode)
      for (uint32_t wi = 0; wi < 3 && s == RK_SUCCESS; ++wi)
        for (uint32_t si = 0; si < 2 && s == RK_SUCCESS; ++si)
        for (uint32_t di = 0; di < 2 && s == RK_SUCCESS; ++di) {
          c.ray_order = oi;
          c.trace_mode = mode;
          c.trace_wg = wgs[wi];
          c.stack_depth = stacks[si];
          c.donate_min = donates[di];
          float trace_ms = 0.0f;
          s = timed([&] {
                rk_status fs = rk_fill_f32(stream, depths, ray_count, 3.402823466e38f);
                if (fs != RK_SUCCESS) return fs;
                return rk_batch_trace(stream, meshes, mesh_count, bvh, origins, directions, ray_count, faces_id,
                                      coords, depths, options, &c, trace_ws, trace_bytes);
              },
              &trace_ms);
          if (s == RK_SUCCESS && build_ms + trace_ms < best_total) {
            best_total = build_ms + trace_ms;
            best_build = build_ms;
            pick = c;
          }
        }
  }
  (void)best_build;
  // leave the outputs as the chosen configuration produces them
  if (s == RK_SUCCESS)
    s = rk_batch_build(stream, meshes, mesh_count, faces, vertices, face_count, hitboxes, &pick, bvh, info, ws,
                       build_bytes);
  if (s == RK_SUCCESS) s = rk_fill_f32(stream, depths, ray_count, 3.402823466e38f);
  if (s == RK_SUCCESS)
    s = rk_batch_trace(stream, meshes, mesh_count, bvh, origins, directions, ray_count, faces_id, coords, depths,
                       options, &pick, trace_ws, trace_bytes);
  if (s == RK_SUCCESS && hipStreamSynchronize(st) != hipSuccess) s = RK_ERROR_DEVICE;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (s == RK_SUCCESS) {
    *best = pick;
    if (best_ms) *best_ms = best_total;
  }
  return s;
}

// ---- attribute interpolation (kernel_interpolateData, CUDA_interpolateData.cu:45-141) ----
// For each pixel, the hit face's attributes at barycentrics (u, v):
// a = (1 - u) - v; attribute = (a * x0 + u * x1) + v * x2 per component, the
// reference's float4 operator order (types/_cuda_float4.h:80-111). Normals are
// normalised as v * (1 / sqrtf(dot(v, v))) (the reference's device normalize uses
// the approximate rsqrtf); without vertex normals the face normal
// normalize(cross(v1 - v0, v2 - v0)) is used. Without colours, (1, 1, 1, 1);
// without texture coordinates, (u, v); materials are face.w. A pixel whose face
// id is not a face of the mesh (>= face_count) is left unchanged.
namespace {

using rk::bary4;
using rk::normalize4;

__global__ __launch_bounds__(256) void k_interpolate(const uint32_t* __restrict__ faces_id,
                                                     const float2* __restrict__ coords, int width, int height,
                                                     const uint4* __restrict__ faces, uint32_t face_count,
                                                     const float4* __restrict__ verts, const float4* __restrict__ norms,
                                                     const float4* __restrict__ cols, const float2* __restrict__ texs,
                                                     float4* __restrict__ out_p, float4* __restrict__ out_n,
                                                     float4* __restrict__ out_c, float2* __restrict__ out_t,
                                                     uint32_t* __restrict__ out_m) {
  const int x = blockIdx.x * 16 + (threadIdx.x & 15);
  const int y = blockIdx.y * 16 + (threadIdx.x >> 4);
  if (x >= width || y >= height) return;
  const size_t l = (size_t)y * (size_t)width + (size_t)x;
  const uint32_t fid = faces_id[l];
  if (fid >= face_count) return;
  const float2 c = coords[l];
  const uint4 f = faces[fid];
  const float4 v0 = verts[f.x], v1 = verts[f.y], v2 = verts[f.z];
  const float a = (1.0f - c.x) - c.y;
  if (out_p) out_p[l] = bary4(a, c.x, c.y, v0, v1, v2);
  if (out_n) out_n[l] = rk::face_normal_at(f, c.x, c.y, verts, norms);
  if (out_c) out_c[l] = cols ? bary4(a, c.x, c.y, cols[f.x], cols[f.y], cols[f.z]) : make_float4(1.0f, 1.0f, 1.0f, 1.0f);
  if (out_t) {
    if (texs) {
      const float2 t0 = texs[f.x], t1 = texs[f.y], t2 = texs[f.z];
      out_t[l] = make_float2((t0.x * a + t1.x * c.x) + t2.x * c.y, (t0.y * a + t1.y * c.x) + t2.y * c.y);
    } else {
      out_t[l] = c;
    }
  }
  if (out_m) out_m[l] = f.w;
}

}  // namespace

extern "C" rk_status rk_interpolate_data(rk_stream stream, const uint32_t* faces_id, const rk_float2* coords,
                                         int width, int height, const rk_uint4* faces, uint32_t face_count,
                                         const rk_float4* vertices, const rk_float4* normals, const rk_float4* colors,
                                         const rk_float2* texcoords, rk_float4* out_points, rk_float4* out_normals,
                                         rk_float4* out_colors, rk_float2* out_texcoords, uint32_t* out_materials) {
  if (width <= 0 || height <= 0 || !faces_id || !coords || !faces || !vertices) return RK_ERROR_INVALID_VALUE;
  dim3 grid((unsigned)((width + 15) / 16), (unsigned)((height + 15) / 16));
  hipLaunchKernelGGL(k_interpolate, grid, dim3(256), 0, rk_hip_stream(stream), faces_id,
                     reinterpret_cast<const float2*>(coords), width, height, reinterpret_cast<const uint4*>(faces),
                     face_count, reinterpret_cast<const float4*>(vertices), reinterpret_cast<const float4*>(normals),
                     reinterpret_cast<const float4*>(colors), reinterpret_cast<const float2*>(texcoords),
                     reinterpret_cast<float4*>(out_points), reinterpret_cast<float4*>(out_normals),
                     reinterpret_cast<float4*>(out_colors), reinterpret_cast<float2*>(out_texcoords), out_materials);
  RK_CHECK_LAUNCH();
  return RK_SUCCESS;
}

// ---- shading (kernel_shade, CUDAShaderKernel.cu:14-43) ---------------------------
namespace {

// One thread per pixel, in row order over the input surfaces (k = y * width +
// x); the output row of pixel k starts at y * pitch.
__global__ __launch_bounds__(256) void k_shade(const float* __restrict__ depths, const float4* __restrict__ normals,
                                               int width, uint32_t count, int pitch, float4 back, float4 cam,
                                               uint32_t* __restrict__ pixels) {
  const uint32_t k = blockIdx.x * 256u + threadIdx.x;
  if (k >= count) return;
  const uint32_t y = k / (uint32_t)width, x = k - y * (uint32_t)width;
  const uint32_t px = depths[k] < RK_HUGE ? rk::shade_hit(cam, normals[k]) : rk::make_pixel(back);
  pixels[(size_t)y * (uint32_t)pitch + x] = px;
}

}  // namespace

extern "C" rk_status rk_shade(rk_stream stream, const float* depths, const rk_float4* normals, int width, int height,
                              rk_float4 back_color, rk_float4 cam_dir, uint32_t* pixels, int pitch) {
  if (width <= 0 || height <= 0 || pitch < width || !depths || !normals || !pixels) return RK_ERROR_INVALID_VALUE;
  const uint64_t count = (uint64_t)width * (uint64_t)height;
  if (count > 0xFFFFFFFFull) return RK_ERROR_UNSUPPORTED;
  hipLaunchKernelGGL(k_shade, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, rk_hip_stream(stream), depths,
                     reinterpret_cast<const float4*>(normals), width, (uint32_t)count, pitch,
                     make_float4(back_color.x, back_color.y, back_color.z, back_color.w),
                     make_float4(cam_dir.x, cam_dir.y, cam_dir.z, 0.0f), pixels);
  RK_CHECK_LAUNCH();
  return RK_SUCCESS;
}
