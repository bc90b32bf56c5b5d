// renderkit drop-in: CUDABuffer (CUDABuffer.h:19-211). Ray/hit surfaces of
// width x height: in_origins/in_directions (float4), out_faces_id (uint),
// out_coords (float2, barycentric u, v), out_depths (float; read as tmax by
// the trace, written with the hit distance).
#ifndef RENDERKIT_CUDABUFFER_H
#define RENDERKIT_CUDABUFFER_H

#include "CUDAArray.h"

namespace renderkit {

class CUDABuffer {
  friend class CUDARaytracer;

 public:
  CUDABuffer(int _width, int _height);  // CUDABuffer.cpp:43-63
  ~CUDABuffer();

  int getWidth() { return width; }
  int getHeight() { return height; }
  CUDASurface<float4>* getInputOriginSurface() const { return in_origins; }
  CUDASurface<float4>* getInputDirectionSurface() const { return in_directions; }
  CUDASurface<float4>* getOutputPointSurface() const { return out_points; }
  CUDASurface<float4>* getOutputNormalSurface() const { return out_normals; }
  CUDASurface<float4>* getOutputColorSurface() const { return out_colors; }
  CUDASurface<float2>* getOutputTexcoordSurface() const { return out_texcoords; }
  CUDASurface<uint>* getOutputMaterialSurface() const { return out_materials; }
  CUDASurface<uint>* getOutputFaceIdSurface() const { return out_faces_id; }
  CUDASurface<float2>* getOutputCoordSurface() const { return out_coords; }
  CUDASurface<float>* getOutputDepthSurface() const { return out_depths; }

  // CUDABuffer.cpp:102-107: origins and directions := 0.
  void clearInputSurfaces();
  // CUDABuffer.cpp:109-131: faces_id := 0, coords := 0, depths := FLT_MAX (the
  // initial tmax), and the optional surfaces := 0.
  void clearOutputSurfaces();

 private:
  int width, height;
  CUDASurface<uint>* out_faces_id = nullptr;
  CUDASurface<float2>* out_coords = nullptr;
  CUDASurface<float>* out_depths = nullptr;
  CUDASurface<float4>* out_points = nullptr;
  CUDASurface<float4>* out_normals = nullptr;
  CUDASurface<float4>* out_colors = nullptr;
  CUDASurface<float2>* out_texcoords = nullptr;
  CUDASurface<uint>* out_materials = nullptr;
  CUDASurface<float4>* in_origins = nullptr;
  CUDASurface<float4>* in_directions = nullptr;
};

}  // namespace renderkit

#endif
