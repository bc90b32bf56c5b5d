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
