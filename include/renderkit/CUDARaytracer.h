// renderkit drop-in: CUDARaytracer (CUDARaytracer.h:52-214). Owns an internal
// camera and buffer, sized from the engine's renderer surface at initialize().
// raytraceMeshTree(mesh, tree, coherency) traces the buffer's rays: the
// reference tree with the packet (coherency = true, default) or per-ray
// traversal, or the fast tree when the tree was built with buildMeshTreeFast().
#ifndef RENDERKIT_CUDARAYTRACER_H
#define RENDERKIT_CUDARAYTRACER_H

#include "CUDABuffer.h"
#include "CUDACamera.h"
#include "CUDAMeshTree.h"
#include "CUDARenderer.h"

namespace renderkit {

class CUDARaytracer {
 public:
  explicit CUDARaytracer(CUDAEngine* _engine);  // CUDARaytracer.cpp:34-42
  virtual ~CUDARaytracer();

  CUDAEngine* getEngine() { return engine; }
  CUDACamera* getCamera() { return current_camera; }
  void setCamera(CUDACamera* c) { current_camera = c; }
  CUDABuffer* getBuffer() { return current_buffer; }
  void setBuffer(CUDABuffer* b) { current_buffer = b; }

  virtual void initialize();  // CUDARaytracer.cpp:53-83
  virtual void finalize();    // :85-92

  // CUDA_calcPrimaryRays (CUDARaytracer.cpp:125-128).
  void calcPrimaryRays();
  // CUDA_raytraceMeshTree (CUDARaytracer.cpp:130-135): returns early for a mesh
  // without vertices.
  void raytraceMeshTree(CUDAMesh* mesh, CUDAMeshTree* tree, bool coherency = true);
  // Stream-ordered variant for callers that manage streams (rk_stream).
  void raytraceMeshTreeAsync(CUDAMesh* mesh, CUDAMeshTree* tree, bool coherency, rk_stream stream);

  // CUDARaytracer.cpp:144-161: allocates the requested output surfaces of the
  // current buffer when missing, then interpolates the mesh's attributes at
  // the traced hits (rk_interpolate_data). Pixels whose face id is not a face
  // of the mesh keep their previous contents.
  void interpolateOutputSurfaces(CUDAMesh* mesh, bool points = true, bool normals = true, bool colors = false,
                                 bool texcoords = false, bool materials = false);

  // Parallel-ray threshold of the triangle test for fast trees (reference
  // trees always use the reference's 1e-6, CUDA_raytraceMeshTree.cu:36).
  void setDetEpsilon(float eps) { det_epsilon = eps; }
  float getDetEpsilon() const { return det_epsilon; }

 protected:
  float det_epsilon = RK_DET_EPSILON_REFERENCE;
  CUDAEngine* engine;
  CUDACamera* internal_camera = nullptr;
  CUDACamera* current_camera = nullptr;
  CUDABuffer* internal_buffer = nullptr;
  CUDABuffer* current_buffer = nullptr;
};

}  // namespace renderkit

#endif
