// The reference's sample 02 frame loop (samples/02_how_to_use_the_raytracer.cpp:67-85),
// headless, through the drop-in classes only: CUDAModel -> CUDAMesh ->
// CUDAMeshTree::buildMeshTreeLBVH(12) -> CUDARaytracer::calcPrimaryRays ->
// raytraceMeshTree (packet and per-ray) -> outputs; then the .cdm quickSave /
// quickLoad round trip and the fast tree. Writes everything to one binary file
// for the test to compare against the oracle.
//
//   sample02_headless <out.bin> <subdivisions> <width> <height> <tmpdir>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/renderkit/CUDARaytracer.h"

using namespace renderkit;

template <typename T>
static void dump(FILE* f, CUDAArray<T>* a) {
  std::vector<T> h(a->getUnitCount());
  a->copyDataToHost(h.data());
  uint32_t bytes = (uint32_t)(h.size() * sizeof(T));
  std::fwrite(&bytes, 4, 1, f);
  std::fwrite(h.data(), 1, bytes, f);
}

static void frame(CUDARaytracer& rt, CUDAMesh& mesh, CUDAMeshTree& tree, bool coherency, FILE* f) {
  rt.getBuffer()->clearOutputSurfaces();
  rt.calcPrimaryRays();
  rt.raytraceMeshTree(&mesh, &tree, coherency);
  CUDABuffer* b = rt.getBuffer();
  dump(f, b->getOutputFaceIdSurface());
  dump(f, b->getOutputCoordSurface());
  dump(f, b->getOutputDepthSurface());
}

int main(int argc, char** argv) {
  if (argc < 6) {
    std::fprintf(stderr, "usage: %s out.bin subdivisions width height tmpdir\n", argv[0]);
    return 2;
  }
  setThrowOnError(true);
  const char* out = argv[1];
  uint32_t sub = (uint32_t)std::atoi(argv[2]);
  int W = std::atoi(argv[3]), H = std::atoi(argv[4]);
  std::string tmp = argv[5];

  uint32_t nv = 0, nf = 0;
  RK_CALL(rk_scene_icosphere(sub, nullptr, nullptr, &nv, &nf));
  std::vector<rk_float4> v(nv);
  std::vector<rk_uint4> fc(nf);
  RK_CALL(rk_scene_icosphere(sub, v.data(), fc.data(), &nv, &nf));

  CUDAModel model(nv, nf);
  for (uint32_t i = 0; i < nv; ++i) {
    float xyz[3] = {v[i].x, v[i].y, v[i].z};
    model.setVertex(i, xyz);  // the setter that accumulates the hitbox
  }
  for (uint32_t i = 0; i < nf; ++i) model.setFace(i, fc[i].x, fc[i].y, fc[i].z, fc[i].w);

  CUDAEngine engine;
  CUDAHeadlessRenderer renderer(&engine, W, H);
  renderer.initialize();
  CUDARaytracer rt(&engine);
  rt.initialize();
  CUDACamera* cam = rt.getCamera();
  cam->setCenter(0.0f, 0.0f, 3.0f);
  cam->lookAt(0.0f, 0.0f, 0.0f);
  cam->setFOV(60.0f * DEG2RAD);
  cam->setRatio(W, H);

  CUDAMesh mesh;
  mesh.buildFromModel(&model);
  CUDAMeshTree tree(&mesh);
  tree.buildMeshTreeLBVH(12);

  FILE* f = std::fopen(out, "wb");
  if (!f) return 3;
  uint32_t hdr[4] = {nv, nf, tree.getNodeCount(), tree.getRealDepth()};
  std::fwrite(hdr, 4, 4, f);
  CUDABox hb = mesh.getHitbox();
  std::fwrite(&hb, sizeof(hb), 1, f);
  dump(f, tree.getNodesArray());
  dump(f, tree.getIndicesArray());
  dump(f, mesh.getFacesArray());
  frame(rt, mesh, tree, true, f);   // packet traversal (the reference default)
  dump(f, rt.getBuffer()->getInputOriginSurface());
  dump(f, rt.getBuffer()->getInputDirectionSurface());
  frame(rt, mesh, tree, false, f);  // per-ray traversal

  // .cdm round trip: save the mesh and tree, load them into fresh objects, trace again
  std::string path = tmp + "/scene.cdm";
  mesh.quickSave(path.c_str(), true);
  tree.quickSave(path.c_str(), false);  // appended to the same file
  CUDAMesh mesh2;
  mesh2.quickLoad(path.c_str());
  CUDAMeshTree tree2(&mesh2);
  tree2.quickLoad(path.c_str());
  frame(rt, mesh2, tree2, true, f);

  // the fast tree on a copy of the mesh
  CUDAMesh mesh3;
  mesh3.buildFromModel(&model);
  CUDAMeshTree tree3(&mesh3);
  tree3.buildMeshTreeFast(4);
  frame(rt, mesh3, tree3, true, f);
  dump(f, tree3.getIndicesArray());
  std::fclose(f);
  rt.finalize();
  renderer.finalize();
  std::printf("sample02 ok: %u faces, %u nodes\n", nf, hdr[2]);
  return 0;
}
