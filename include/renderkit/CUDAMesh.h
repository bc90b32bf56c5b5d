// renderkit drop-in: CUDAMesh (CUDAMesh.h:19-157). Device arrays float4
// vertices[V], uint4 faces[N] (+ optional normals, colors, texcoords) and the
// model's hitbox. Builds reorder the faces in place (CUDA_reorderMeshFaces).
#ifndef RENDERKIT_CUDAMESH_H
#define RENDERKIT_CUDAMESH_H

#include "CUDAArray.h"
#include "CUDAModel.h"

namespace renderkit {

class CUDAMesh {
 public:
  CUDAMesh();
  ~CUDAMesh();

  CUDAModel* getModel() const { return model; }
  CUDAArray<float4>* getVerticesArray() const { return vertices; }
  CUDAArray<float4>* getNormalsArray() const { return normals; }
  CUDAArray<float4>* getColorsArray() const { return colors; }
  CUDAArray<float2>* getTexcoordsArray() const { return texcoords; }
  CUDAArray<uint4>* getFacesArray() const { return faces; }
  int getVertexCount() const { return vertices ? vertices->getUnitCount() : 0; }
  int getFaceCount() const { return faces ? faces->getUnitCount() : 0; }
  CUDABox getHitbox() const { return hitbox; }

  // CUDAMesh.cpp:67-132
  void buildFromModel(CUDAModel* _model);
  // Direct construction from host arrays (w of vertices as given; hitbox = the
  // merge of every vertex as float4 with w = 0, as CUDAModel accumulates it).
  void buildFromArrays(const float4* host_vertices, uint vertex_count, const uint4* host_faces, uint face_count);
  // Chunked .cdm file, CUDAMesh.cpp:134-400: HITBOX (int magic, int length, aabox),
  // then VERTICES, NORMALS, COLORS, TEXCOORDS, FACES (int magic, int length, uint
  // count, raw array); unknown chunks are skipped by their length.
  void quickLoad(const char* path);
  void quickSave(const char* path, bool overwrite = true);

 private:
  void release();
  CUDABox hitbox;
  CUDAArray<float4>* vertices = nullptr;
  CUDAArray<float4>* normals = nullptr;
  CUDAArray<float4>* colors = nullptr;
  CUDAArray<float2>* texcoords = nullptr;
  CUDAArray<uint4>* faces = nullptr;
  CUDAModel* model = nullptr;
};

}  // namespace renderkit

#endif
