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
#include "CUDAShader.h"
#include "CUDAWindow.h"

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
  CUDAShader* getShader() { return current_shader; }
  void setShader(CUDAShader* s) { current_shader = s; }

  // CUDARaytracer.cpp:53-83: a buffer the size of the renderer's surface and a
  // base shader.
  virtual void initialize();
  virtual void finalize();  // :85-92

  // CUDARaytracer.cpp:110-121: runs the current shader on the current buffer
  // into the renderer's surface.
  void updateRenderSurface();
  // CUDARaytracer.cpp:94-108: the reference saves the shaded surface as a JPEG
  // through its Direct3D renderer and asserts for any other renderer; there is
  // no Direct3D renderer here, so this asserts as it does for those.
  void saveRenderSurface(const char* path);

  // CUDA_calcPrimaryRays (CUDARaytracer.cpp:125-128).
  void calcPrimaryRays();
  // CUDA_raytraceMeshTree (CUDARaytracer.cpp:130-135): returns early for a mesh
  // without vertices.
  void raytraceMeshTree(CUDAMesh* mesh, CUDAMeshTree* tree, bool coherency = true);
  // Stream-ordered variant for callers that manage streams (rk_stream).
  void raytraceMeshTreeAsync(CUDAMesh* mesh, CUDAMeshTree* tree, bool coherency, rk_stream stream);
  // CUDARaytracer.cpp:137-142. The reference rasterises the faces of the mesh
  // (CUDA_rasterizeMeshTree, a Direct3D-era experimental path) into the same
  // output surfaces as a primary-ray trace; this library has no rasteriser and
  // fills them with the coherent traversal of the buffer's rays instead: the
  // closest face, (u, v) and distance of each pixel's ray.
  void rasterizeMeshTree(CUDAMesh* mesh, CUDAMeshTree* tree);

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
  CUDAShader* internal_shader = nullptr;
  CUDAShader* current_shader = nullptr;
};

}  // namespace renderkit

#endif
