// Seeded scene and ray generators (host). See include/rk_scenes.h.
#include "../../include/rk_scenes.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <map>
#include <thread>
#include <utility>
#include <vector>

namespace {

inline uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}
// Uniform double in [0,1) from stream 'stream' element 'k' of seed 'seed'.
inline double u01(uint64_t seed, uint64_t stream, uint64_t k) {
  uint64_t h = splitmix64(seed ^ splitmix64(stream * 0x632BE59BD9B4E019ull + k));
  return (double)(h >> 11) * (1.0 / 9007199254740992.0);
}

template <typename F>
void parallel_for(uint64_t n, F f) {
  unsigned hw = std::thread::hardware_concurrency();
  unsigned T = hw ? std::min(hw, 32u) : 4u;
  if (n < 65536 || T <= 1) { for (uint64_t i = 0; i < n; ++i) f(i); return; }
  std::vector<std::thread> th;
  uint64_t chunk = (n + T - 1) / T;
  for (unsigned t = 0; t < T; ++t) {
    uint64_t a = t * chunk, b = std::min(n, a + chunk);
    if (a >= b) break;
    th.emplace_back([=, &f]() { for (uint64_t i = a; i < b; ++i) f(i); });
  }
  for (auto& x : th) x.join();
}

struct V3 { float x, y, z; };
inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 muls(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline V3 normalize(V3 v) { float inv = 1.0f / sqrtf(dot(v, v)); return muls(v, inv); }

}  // namespace

extern "C" rk_status rk_scene_icosphere(uint32_t levels, rk_float4* vertices, rk_uint4* faces,
                                        uint32_t* vertex_count, uint32_t* face_count) {
  if (levels > 10) return RK_ERROR_INVALID_VALUE;
  uint64_t nf = 20ull << (2 * levels), nv = 10ull * (1ull << (2 * levels)) + 2;
  if (vertex_count) *vertex_count = (uint32_t)nv;
  if (face_count) *face_count = (uint32_t)nf;
  if (!vertices || !faces) return RK_SUCCESS;
  const double t = (1.0 + std::sqrt(5.0)) / 2.0;
  std::vector<std::array<double, 3>> P = {
      {-1, t, 0}, {1, t, 0}, {-1, -t, 0}, {1, -t, 0}, {0, -1, t}, {0, 1, t},
      {0, -1, -t}, {0, 1, -t}, {t, 0, -1}, {t, 0, 1}, {-t, 0, -1}, {-t, 0, 1}};
  for (auto& p : P) {
    double l = std::sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
    p = {p[0] / l, p[1] / l, p[2] / l};
  }
  std::vector<std::array<uint32_t, 3>> F = {
      {0, 11, 5}, {0, 5, 1}, {0, 1, 7}, {0, 7, 10}, {0, 10, 11}, {1, 5, 9}, {5, 11, 4},
      {11, 10, 2}, {10, 7, 6}, {7, 1, 8}, {3, 9, 4}, {3, 4, 2}, {3, 2, 6}, {3, 6, 8},
      {3, 8, 9}, {4, 9, 5}, {2, 4, 11}, {6, 2, 10}, {8, 6, 7}, {9, 8, 1}};
  for (uint32_t lv = 0; lv < levels; ++lv) {
    std::map<std::pair<uint32_t, uint32_t>, uint32_t> cache;
    auto mid = [&](uint32_t a, uint32_t b) {
      auto key = std::make_pair(std::min(a, b), std::max(a, b));
      auto it = cache.find(key);
      if (it != cache.end()) return it->second;
      std::array<double, 3> m = {(P[a][0] + P[b][0]) * 0.5, (P[a][1] + P[b][1]) * 0.5, (P[a][2] + P[b][2]) * 0.5};
      double l = std::sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
      P.push_back({m[0] / l, m[1] / l, m[2] / l});
      uint32_t id = (uint32_t)P.size() - 1;
      cache[key] = id;
      return id;
    };
    std::vector<std::array<uint32_t, 3>> G;
    G.reserve(F.size() * 4);
    for (auto& f : F) {
      uint32_t ab = mid(f[0], f[1]), bc = mid(f[1], f[2]), ca = mid(f[2], f[0]);
      G.push_back({f[0], ab, ca});
      G.push_back({f[1], bc, ab});
      G.push_back({f[2], ca, bc});
      G.push_back({ab, bc, ca});
    }
    F.swap(G);
  }
  for (size_t i = 0; i < P.size(); ++i) vertices[i] = {(float)P[i][0], (float)P[i][1], (float)P[i][2], 0.0f};
  for (size_t i = 0; i < F.size(); ++i) faces[i] = {F[i][0], F[i][1], F[i][2], 0u};
  return RK_SUCCESS;
}

namespace {
inline double lattice(uint64_t seed, int octave, int64_t i, int64_t j) {
  uint64_t k = ((uint64_t)(uint32_t)i << 32) ^ (uint64_t)(uint32_t)j;
  return u01(seed, 1000 + (uint64_t)octave, k);
}
inline double smooth(double f) { return f * f * (3.0 - 2.0 * f); }
double value_noise(uint64_t seed, int octave, double x, double y) {
  double fx = std::floor(x), fy = std::floor(y);
  int64_t ix = (int64_t)fx, iy = (int64_t)fy;
  double tx = smooth(x - fx), ty = smooth(y - fy);
  double a = lattice(seed, octave, ix, iy), b = lattice(seed, octave, ix + 1, iy);
  double c = lattice(seed, octave, ix, iy + 1), d = lattice(seed, octave, ix + 1, iy + 1);
  double ab = a + (b - a) * tx, cd = c + (d - c) * tx;
  return ab + (cd - ab) * ty;
}
double fbm(uint64_t seed, double x, double y) {
  double sum = 0.0, amp = 1.0, freq = 4.0;
  for (int o = 0; o < 6; ++o) {
    sum += amp * value_noise(seed, o, x * freq, y * freq);
    amp *= 0.5;
    freq *= 2.0;
  }
  return sum;
}
}  // namespace

extern "C" rk_status rk_scene_heightfield(uint32_t grid, uint64_t seed, float amplitude, rk_float4* vertices,
                                          rk_uint4* faces, uint32_t* vertex_count, uint32_t* face_count) {
  if (grid < 2 || grid > 46341) return RK_ERROR_INVALID_VALUE;
  uint64_t nv = (uint64_t)grid * grid, nf = 2ull * (grid - 1) * (grid - 1);
  if (vertex_count) *vertex_count = (uint32_t)nv;
  if (face_count) *face_count = (uint32_t)nf;
  if (!vertices || !faces) return RK_SUCCESS;
  std::vector<double> h(nv);
  const double inv = 1.0 / (double)(grid - 1);
  parallel_for(nv, [&](uint64_t v) {
    uint64_t i = v % grid, j = v / grid;
    h[v] = fbm(seed, (double)i * inv, (double)j * inv);
  });
  double lo = *std::min_element(h.begin(), h.end()), hi = *std::max_element(h.begin(), h.end());
  double span = hi > lo ? hi - lo : 1.0;
  parallel_for(nv, [&](uint64_t v) {
    uint64_t i = v % grid, j = v / grid;
    vertices[v] = {(float)((double)i * inv), (float)((double)j * inv), (float)((double)amplitude * (h[v] - lo) / span),
                   0.0f};
  });
  const uint64_t cells = (uint64_t)(grid - 1) * (grid - 1);
  parallel_for(cells, [&](uint64_t c) {
    uint32_t i = (uint32_t)(c % (grid - 1)), j = (uint32_t)(c / (grid - 1));
    uint32_t v00 = j * grid + i, v10 = v00 + 1, v01 = v00 + grid, v11 = v01 + 1;
    faces[2 * c] = {v00, v10, v11, 0u};
    faces[2 * c + 1] = {v00, v11, v01, 0u};
  });
  return RK_SUCCESS;
}

extern "C" rk_status rk_scene_soup(uint32_t n, uint64_t seed, float mu, float sigma, float lo, float hi,
                                   rk_float4* vertices, rk_uint4* faces, uint32_t* vertex_count, uint32_t* face_count) {
  if (n == 0 || n > 400000000u) return RK_ERROR_INVALID_VALUE;
  if (vertex_count) *vertex_count = 3 * n;
  if (face_count) *face_count = n;
  if (!vertices || !faces) return RK_SUCCESS;
  const double two_pi = 6.283185307179586;
  parallel_for(n, [&](uint64_t k) {
    double c[3];
    for (int a = 0; a < 3; ++a) c[a] = lo + (hi - lo) * u01(seed, a, k);
    double e[2][3];
    for (int m = 0; m < 2; ++m) {
      double z = 2.0 * u01(seed, 3 + 4 * m, k) - 1.0;
      double phi = two_pi * u01(seed, 4 + 4 * m, k);
      double r = std::sqrt(std::max(0.0, 1.0 - z * z));
      double g1 = u01(seed, 5 + 4 * m, k), g2 = u01(seed, 6 + 4 * m, k);
      double gauss = std::sqrt(-2.0 * std::log(1.0 - g1)) * std::cos(two_pi * g2);
      double len = std::exp((double)mu + (double)sigma * gauss);
      e[m][0] = r * std::cos(phi) * len;
      e[m][1] = r * std::sin(phi) * len;
      e[m][2] = z * len;
    }
    float v0[3], v1[3], v2[3];
    for (int a = 0; a < 3; ++a) {
      double b0 = c[a] - (e[0][a] + e[1][a]) / 3.0;
      v0[a] = (float)b0;
      v1[a] = (float)(b0 + e[0][a]);
      v2[a] = (float)(b0 + e[1][a]);
    }
    vertices[3 * k + 0] = {v0[0], v0[1], v0[2], 0.0f};
    vertices[3 * k + 1] = {v1[0], v1[1], v1[2], 0.0f};
    vertices[3 * k + 2] = {v2[0], v2[1], v2[2], 0.0f};
    faces[k] = {(uint32_t)(3 * k), (uint32_t)(3 * k + 1), (uint32_t)(3 * k + 2), 0u};
  });
  return RK_SUCCESS;
}

extern "C" rk_status rk_scene_hitbox(const rk_float4* vertices, uint32_t vertex_count, rk_aabox* hitbox) {
  if (!hitbox || (vertex_count && !vertices)) return RK_ERROR_INVALID_VALUE;
  const float H = 3.402823466e+38f;
  rk_aabox b = {{H, H, H, H}, {-H, -H, -H, -H}};
  for (uint32_t i = 0; i < vertex_count; ++i) {
    const rk_float4& v = vertices[i];
    float w = 0.0f;  // setVertex(uint, const float*) stores w = 0 (CUDAModel.h:178-183)
    b.min.x = fminf(b.min.x, v.x); b.min.y = fminf(b.min.y, v.y); b.min.z = fminf(b.min.z, v.z); b.min.w = fminf(b.min.w, w);
    b.max.x = fmaxf(b.max.x, v.x); b.max.y = fmaxf(b.max.y, v.y); b.max.z = fmaxf(b.max.z, v.z); b.max.w = fmaxf(b.max.w, w);
  }
  *hitbox = b;
  return RK_SUCCESS;
}

namespace {
// CUDA_calcPrimaryRays.cu:43-55, host part, from an orthonormal camera frame.
void frame_to_rays(V3 center, V3 camDir, V3 right, V3 up, float fov, float ratio, int width, int height,
                   rk_float4 out[4]) {
  float y = tanf(fov * 0.5f);
  float x = y * ratio;
  V3 rd = add(add(muls(up, y), muls(right, x)), camDir);
  float iw = 1.0f / (float)width, ih = 1.0f / (float)height;  // float4 / float is a * (1/s)
  V3 rx = muls(muls(right, 2.0f * x), iw);
  V3 ry = muls(muls(up, 2.0f * y), ih);
  out[0] = {center.x, center.y, center.z, 0.0f};
  out[1] = {rd.x, rd.y, rd.z, 0.0f};
  out[2] = {rx.x, rx.y, rx.z, 0.0f};
  out[3] = {ry.x, ry.y, ry.z, 0.0f};
}
}  // namespace

extern "C" rk_status rk_camera_lookat(rk_float4 eye, rk_float4 target, float fov, float ratio, int width, int height,
                                      rk_float4 out[4]) {
  if (!out || width <= 0 || height <= 0) return RK_ERROR_INVALID_VALUE;
  const float PI = 3.14159265359f;
  // angleFromDir, CUDACamera.cpp:70-88.
  V3 dir = normalize(V3{target.x - eye.x, target.y - eye.y, target.z - eye.z});
  float ay = asinf(dir.y);
  if (ay < -PI * 0.5f + 0.001f) ay = -PI * 0.5f + 0.001f;
  if (ay > PI * 0.5f - 0.001f) ay = PI * 0.5f - 0.001f;
  float ax;
  if (dir.x > 0) ax = asinf(dir.z / cosf(ay));
  else if (dir.z > 0) ax = acosf(dir.x / cosf(ay));
  else ax = acosf(-dir.x / cosf(-ay)) + PI;
  if (ax < 0.0f) ax += PI * 2.0f;
  if (ax > PI * 2.0f) ax -= PI * 2.0f;
  float pitch = ax, yaw = ay, roll = 0.0f;
  // CUDA_calcPrimaryRays.cu:46-50.
  V3 camDir = {cosf(pitch) * cosf(yaw), sinf(yaw), sinf(pitch) * cosf(yaw)};
  V3 camUp = {cosf(pitch) * sinf(roll), cosf(roll), 0.0f};
  V3 right = normalize(cross(camDir, camUp));
  V3 up = normalize(cross(right, camDir));
  frame_to_rays(V3{eye.x, eye.y, eye.z}, camDir, right, up, fov, ratio, width, height, out);
  return RK_SUCCESS;
}

extern "C" rk_status rk_camera_lookat_up(rk_float4 eye, rk_float4 target, rk_float4 upv, float fov, float ratio,
                                         int width, int height, rk_float4 out[4]) {
  if (!out || width <= 0 || height <= 0) return RK_ERROR_INVALID_VALUE;
  V3 camDir = normalize(V3{target.x - eye.x, target.y - eye.y, target.z - eye.z});
  V3 right = normalize(cross(camDir, V3{upv.x, upv.y, upv.z}));
  V3 up = normalize(cross(right, camDir));
  frame_to_rays(V3{eye.x, eye.y, eye.z}, camDir, right, up, fov, ratio, width, height, out);
  return RK_SUCCESS;
}

extern "C" rk_status rk_primary_rays_host(const rk_float4 cam[4], int width, int height, rk_float4* origins,
                                          rk_float4* directions) {
  if (!cam || !origins || !directions || width <= 0 || height <= 0) return RK_ERROR_INVALID_VALUE;
  const rk_float4 ro = cam[0], rd = cam[1], rx = cam[2], ry = cam[3];
  parallel_for((uint64_t)width * height, [&](uint64_t n) {
    float fx = (float)(n % width), fy = (float)(n / width);
    origins[n] = ro;
    directions[n] = {(rd.x - fx * rx.x) - fy * ry.x, (rd.y - fx * rx.y) - fy * ry.y, (rd.z - fx * rx.z) - fy * ry.z,
                     (rd.w - fx * rx.w) - fy * ry.w};
  });
  return RK_SUCCESS;
}

extern "C" rk_status rk_rays_random(uint32_t n, uint64_t seed, float lo, float hi, rk_float4* origins,
                                    rk_float4* directions) {
  if (!origins || !directions) return RK_ERROR_INVALID_VALUE;
  const double two_pi = 6.283185307179586;
  parallel_for(n, [&](uint64_t k) {
    float o[3];
    for (int a = 0; a < 3; ++a) o[a] = (float)(lo + (hi - lo) * u01(seed, 20 + a, k));
    double z = 2.0 * u01(seed, 23, k) - 1.0, phi = two_pi * u01(seed, 24, k);
    double r = std::sqrt(std::max(0.0, 1.0 - z * z));
    origins[k] = {o[0], o[1], o[2], 0.0f};
    directions[k] = {(float)(r * std::cos(phi)), (float)(r * std::sin(phi)), (float)z, 0.0f};
  });
  return RK_SUCCESS;
}
