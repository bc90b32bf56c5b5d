// renderkit drop-in: common typedefs, vector types and error handling.
// Mirrors CUDACommon.h (typedefs :21-30, PI/DEG2RAD :108-110, _assert :123-133)
// and the float2/float3/float4/uint4 layouts of the reference's types/ headers,
// without any GPU runtime header: device work goes through include/rk.h.
#ifndef RENDERKIT_CUDACOMMON_H
#define RENDERKIT_CUDACOMMON_H

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <string>

#include "../rk.h"
#include "../rk_scenes.h"

typedef unsigned long long qword;
typedef unsigned int dword;
typedef unsigned short word;
typedef unsigned int uint;
typedef unsigned short ushort;
typedef unsigned char uchar;
typedef unsigned char byte;

#ifndef PI
#define PI (3.14159265359f)
#endif
#define DEG2RAD (0.01745329252f)
#define RAD2DEG (57.2957795131f)

struct float2 { float x, y; };
struct float3 { float x, y, z; };
struct alignas(16) float4 { float x, y, z, w; };
struct alignas(16) uint4 { uint x, y, z, w; };
struct uint2 { uint x, y; };
// A 4 x 4 float matrix, row-major (m[row][column]).
struct float4x4 { float m[4][4]; };

inline float2 make_float2(float x, float y) { return {x, y}; }
inline float2 make_float2(float s) { return {s, s}; }
inline float3 make_float3(float x, float y, float z) { return {x, y, z}; }
inline float3 make_float3(float s) { return {s, s, s}; }
inline float3 make_float3(const float4& a) { return {a.x, a.y, a.z}; }
inline float4 make_float4(float x, float y, float z, float w) { return {x, y, z, w}; }
inline float4 make_float4(float s) { return {s, s, s, s}; }
inline float4 make_float4(const float3& a) { return {a.x, a.y, a.z, 0.0f}; }
inline float4 make_float4(const float3& a, float w) { return {a.x, a.y, a.z, w}; }
inline uint4 make_uint4(uint x, uint y, uint z, uint w) { return {x, y, z, w}; }

inline float3 operator-(const float3& a, const float3& b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline float3 operator+(const float3& a, const float3& b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline float3 operator*(const float3& a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline float4 operator-(const float4& a, const float4& b) { return {a.x - b.x, a.y - b.y, a.z - b.z, a.w - b.w}; }
inline float4 operator+(const float4& a, const float4& b) { return {a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w}; }
inline float4 operator*(const float4& a, float s) { return {a.x * s, a.y * s, a.z * s, a.w * s}; }
inline float4 operator*(float s, const float4& a) { return {a.x * s, a.y * s, a.z * s, a.w * s}; }
inline float dot(const float3& a, const float3& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline float3 cross(const float3& a, const float3& b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline float3 normalize(const float3& v) { return v * (1.0f / sqrtf(dot(v, v))); }

static_assert(sizeof(float4) == sizeof(rk_float4), "float4 layout");
static_assert(sizeof(uint4) == sizeof(rk_uint4), "uint4 layout");
static_assert(sizeof(float2) == sizeof(rk_float2), "float2 layout");

namespace renderkit {

// The reference's _assert prints the location and exits the process
// (CUDACommon.h:123-133). Kept as the default; a caller that prefers an
// exception calls renderkit::setThrowOnError(true).
bool& throwOnErrorFlag();
inline void setThrowOnError(bool on) { throwOnErrorFlag() = on; }

struct Error : std::runtime_error {
  using std::runtime_error::runtime_error;
};

inline void _assert(bool predicate, const char* file, int line, const char* message) {
  if (predicate) return;
  char buf[1024];
  std::snprintf(buf, sizeof(buf), "Error in %s, line %d : %s", file, line, message);
  if (throwOnErrorFlag()) throw Error(buf);
  std::fprintf(stderr, "%s\n\n", buf);
  std::exit(-1);
}

inline void rkCheck(rk_status s, const char* file, int line, const char* what) {
  if (s == RK_SUCCESS) return;
  std::string m = std::string(what) + " : " + rk_status_string(s);
  _assert(false, file, line, m.c_str());
}

}  // namespace renderkit

#define RK_CALL(expr) ::renderkit::rkCheck((expr), __FILE__, __LINE__, #expr)

#endif
