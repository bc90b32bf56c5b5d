// renderkit drop-in: CUDAModel (CUDAModel.h), the host-side mesh a CUDAMesh is
// built from. Vertices are float4 with w = 0; faces are uint4 (x, y, z vertex
// ids, w material). The hitbox follows the reference rule: only
// setVertex(uint, const float*) merges into it (CUDAModel.h:178-183).
#ifndef RENDERKIT_CUDAMODEL_H
#define RENDERKIT_CUDAMODEL_H

#include <vector>

#include "CUDATypes.h"

namespace renderkit {

class CUDAModel {
 public:
  bool hasVertices = false, hasNormals = false, hasColors = false, hasTexcoords = false;

  CUDAModel() { box.reset(); }
  CUDAModel(uint vertex_count, uint face_count) { box.reset(); allocate(vertex_count, face_count); }

  void allocate(uint vertex_count, uint face_count) {
    vertices.assign(vertex_count, make_float4(0.0f));
    faces.assign(face_count, make_uint4(0, 0, 0, 0));
    hasVertices = vertex_count > 0;
  }

  // CUDAModel.cpp:74-101: resize, keeping the first entries.
  void setVertexCount(uint count) {
    vertices.resize(count, make_float4(0.0f));
    if (hasNormals) normals.resize(count, make_float4(0.0f));
    if (hasColors) colors.resize(count, make_float4(0.0f));
    if (hasTexcoords) texcoords.resize(count, make_float2(0.0f));
  }
  void setFaceCount(uint count) { faces.resize(count, make_uint4(0, 0, 0, 0)); }
  // CUDAModel.cpp:318-349: an axis-aligned box centred at the origin, 8
  // vertices and 12 faces. Its vertices are set through setVertex(uint, float,
  // float, float), which does not merge into the hitbox: the hitbox stays empty,
  // as in the reference.
  void createCube(float width, float height, float depth) {
    hasVertices = true;
    setVertexCount(8);
    const float hx = 0.5f * width, hy = 0.5f * height, hz = 0.5f * depth;
    for (uint i = 0; i < 8; ++i) setVertex(i, (i & 4) ? hx : -hx, (i & 2) ? hy : -hy, (i & 1) ? hz : -hz);
    static const uint tri[12][3] = {{0, 1, 2}, {1, 2, 3}, {4, 5, 6}, {5, 6, 7}, {0, 1, 4}, {1, 4, 5},
                                    {2, 3, 6}, {3, 6, 7}, {0, 2, 4}, {2, 4, 6}, {1, 3, 5}, {3, 5, 7}};
    setFaceCount(12);
    for (uint f = 0; f < 12; ++f) setFace(f, tri[f][0], tri[f][1], tri[f][2], 0);
  }

  uint getVertexCount() const { return (uint)vertices.size(); }
  uint getFaceCount() const { return (uint)faces.size(); }
  CUDABox getHitbox() const { return box; }

  float4* getVertexPointer() { return vertices.data(); }
  uint4* getFacePointer() { return faces.data(); }
  float4* getNormalPointer() { return normals.data(); }
  float4* getColorPointer() { return colors.data(); }
  float2* getTexcoordPointer() { return texcoords.data(); }

  void setVertex(uint id, float x, float y, float z) { vertices[id] = make_float4(x, y, z, 0.0f); }
  void setVertex(uint id, float3 v) { vertices[id] = make_float4(v); }
  void setVertex(uint id, float4 v) { vertices[id] = v; }
  void setVertex(uint id, const float* xyz) {  // the only setter that updates the hitbox
    vertices[id] = make_float4(xyz[0], xyz[1], xyz[2], 0.0f);
    box.merge(vertices[id]);
  }
  float3 getVertex(uint id) const { return make_float3(vertices[id]); }
  void setFace(uint id, uint a, uint b, uint c, uint material = 0) { faces[id] = make_uint4(a, b, c, material); }

 private:
  CUDABox box;
  std::vector<float4> vertices, normals, colors;
  std::vector<float2> texcoords;
  std::vector<uint4> faces;
};

}  // namespace renderkit

#endif
