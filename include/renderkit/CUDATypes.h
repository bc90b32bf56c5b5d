// renderkit drop-in: the POD tree types with the reference layouts.
//   aabox  (CUDABox.h:17-148)  float4 min, max; 32 B
//   bvnode (CUDANode.h:20-48)  type/axis|count bits, child_id|obj_id, float3 box_min, box_max; 32 B
//   bvbin  (CUDABin.h:15-113)  SAH bin; 48 B
#ifndef RENDERKIT_CUDATYPES_H
#define RENDERKIT_CUDATYPES_H

#include <cfloat>

#include "CUDACommon.h"

namespace renderkit {

struct alignas(16) aabox {
  float4 min;
  float4 max;

  float getArea() const {  // CUDABox.h:35-44
    float4 v = max - min;
    return 2.f * (v.x * v.y + v.x * v.z + v.y * v.z);
  }
  void reset() {  // CUDABox.h:50-54
    min = make_float4(+FLT_MAX);
    max = make_float4(-FLT_MAX);
  }
  void reset(const float4& v) { min = v; max = v; }
  void reset(const float4& _min, const float4& _max) { min = _min; max = _max; }
  void reset(const float3& _min, const float3& _max) { min = make_float4(_min); max = make_float4(_max); }
  void merge(const float4& v) {  // CUDABox.h:112-116 (device fminf/fmaxf semantics)
    min = {fminf(min.x, v.x), fminf(min.y, v.y), fminf(min.z, v.z), fminf(min.w, v.w)};
    max = {fmaxf(max.x, v.x), fmaxf(max.y, v.y), fmaxf(max.z, v.z), fmaxf(max.w, v.w)};
  }
  void merge(const aabox& b) {
    min = {fminf(min.x, b.min.x), fminf(min.y, b.min.y), fminf(min.z, b.min.z), fminf(min.w, b.min.w)};
    max = {fmaxf(max.x, b.max.x), fmaxf(max.y, b.max.y), fmaxf(max.z, b.max.z), fmaxf(max.w, b.max.w)};
  }
};
typedef aabox CUDABox;

struct alignas(16) bvnode {
  uint word0;  // bit 0 type (0 node, 1 leaf); node: bits 1-2 axis; leaf: bits 1-31 obj_count
  uint word1;  // node: child_id (children child_id, child_id + 1); leaf: obj_id
  float3 box_min;
  float3 box_max;

  bool isLeaf() const { return (word0 & 1u) != 0; }                 // CUDANode.h:60-66
  bool isNode() const { return (word0 & 1u) == 0; }                 // :68-74
  int getAxis() const { return (int)((word0 >> 1) & 3u); }          // :76-87
  uint getChildId() const { return word1; }
  uint getObjectId() const { return word1; }
  uint getObjectCount() const { return word0 >> 1; }
  // getChildren, CUDANode.h:113-127: a negative direction on the split axis
  // visits child_id + 1 first.
  void getChildren(const float4& dir, int& near_child_id, int& far_child_id) const {
    int axis = getAxis();
    bool neg = axis == 0 ? dir.x < 0.f : (axis == 1 ? dir.y < 0.f : dir.z < 0.f);
    near_child_id = (int)(neg ? word1 + 1 : word1);
    far_child_id = (int)(neg ? word1 : word1 + 1);
  }
};
typedef bvnode CUDANode;

struct alignas(16) bvbin {
  float3 box_min;
  float3 box_max;
  int obj_count;
  int left_count;
  int right_count;
  float left_area;
  float right_area;
  int padding;
};
typedef bvbin CUDABin;

static_assert(sizeof(aabox) == 32 && sizeof(aabox) == sizeof(rk_aabox), "aabox is 32 B");
static_assert(sizeof(bvnode) == 32 && sizeof(bvnode) == sizeof(rk_bvnode), "bvnode is 32 B");
static_assert(sizeof(bvbin) == 48, "bvbin is 48 B");

}  // namespace renderkit

#endif
