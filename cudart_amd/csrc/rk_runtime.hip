// Runtime entry points: device memory, streams, surface fills, primary rays.
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
