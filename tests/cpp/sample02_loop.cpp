// The frame loop of the reference's sample 02 (samples/02_how_to_use_the_raytracer.cpp:67-85)
// run through the drop-in classes, with the set-up of its initEngine, initWindow,
// initRenderer, initRaytracer and createScene (:104-196), headless: the window
// is a CUDAHeadlessWindow that closes after a number of frames and the
// renderer a CUDAHeadlessRenderer of the sample's 400 x 300 window.
//
//   sample02_loop <out.bin> <cube|ico> <frames>
//
// "cube" is the sample's scene (a 20-unit cube built with the SAH tree);
// "ico" is a subdivided icosphere of radius 10 whose vertices go through
// setVertex(uint, const float*) and so have a hitbox. After the loop the
// program writes the mesh, the buffer's surfaces, the shaded render surface,
// the camera direction and the shader's back colour for the test to check.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/renderkit/CUDARaytracer.h"

using namespace renderkit;

static CUDAEngine* engine = nullptr;
static CUDAWindow* window = nullptr;
static CUDARenderer* renderer = nullptr;
static CUDARaytracer* raytracer = nullptr;
static CUDACamera* camera = nullptr;
static CUDAModel* model = nullptr;
static CUDAMesh* mesh = nullptr;
static CUDAMeshTree* tree = nullptr;

static void mainLoop() {
  while (window->isActive()) {
    raytracer->getBuffer()->clearOutputSurfaces();
    raytracer->calcPrimaryRays();
    raytracer->raytraceMeshTree(mesh, tree);
    raytracer->interpolateOutputSurfaces(mesh, true, true, false, false, false);
    raytracer->updateRenderSurface();
    renderer->update();
    engine->update();
  }
}

template <typename T>
static void dump(FILE* f, CUDAArray<T>* a) {
  std::vector<T> h(a->getUnitCount());
  a->copyDataToHost(h.data());
  uint32_t bytes = (uint32_t)(h.size() * sizeof(T));
  std::fwrite(&bytes, 4, 1, f);
  std::fwrite(h.data(), 1, bytes, f);
}

static void dumpHost(FILE* f, const void* p, uint32_t bytes) {
  std::fwrite(&bytes, 4, 1, f);
  std::fwrite(p, 1, bytes, f);
}

int main(int argc, char** argv) {
  if (argc < 4) {
    std::fprintf(stderr, "usage: %s out.bin cube|ico frames\n", argv[0]);
    return 2;
  }
  setThrowOnError(true);
  const bool cube = std::strcmp(argv[2], "cube") == 0;
  const unsigned long long frames = std::strtoull(argv[3], nullptr, 10);

  engine = new CUDAEngine();
  window = new CUDAHeadlessWindow(engine, frames);
  window->setLeft(120);
  window->setTop(80);
  window->setWidth(400);
  window->setHeight(300);
  window->show();

  renderer = new CUDAHeadlessRenderer(engine, window->getWidth(), window->getHeight());
  renderer->disableFullscreen();
  renderer->disableVSync();
  renderer->useHardware();
  renderer->initialize();

  raytracer = new CUDARaytracer(engine);
  camera = raytracer->getCamera();
  camera->setRatio(renderer->getSurface()->getWidth(), renderer->getSurface()->getHeight());
  camera->setCenter(-10.0f, 20.0f, 30.0f);
  camera->lookAt(0.0f, 0.0f, 0.0f);
  camera->update();
  raytracer->initialize();

  model = new CUDAModel();
  if (cube) {
    model->createCube(20.0f, 20.0f, 20.0f);
  } else {
    uint32_t nv = 0, nf = 0;
    RK_CALL(rk_scene_icosphere(4, nullptr, nullptr, &nv, &nf));
    std::vector<rk_float4> v(nv);
    std::vector<rk_uint4> fc(nf);
    RK_CALL(rk_scene_icosphere(4, v.data(), fc.data(), &nv, &nf));
    model->allocate(nv, nf);
    for (uint32_t i = 0; i < nv; ++i) {
      float xyz[3] = {10.0f * v[i].x, 10.0f * v[i].y, 10.0f * v[i].z};
      model->setVertex(i, xyz);
    }
    for (uint32_t i = 0; i < nf; ++i) model->setFace(i, fc[i].x, fc[i].y, fc[i].z, fc[i].w);
  }
  mesh = new CUDAMesh();
  mesh->buildFromModel(model);
  tree = new CUDAMeshTree(mesh);
  tree->buildMeshTreeSAH();

  mainLoop();

  FILE* f = std::fopen(argv[1], "wb");
  if (!f) return 3;
  CUDABuffer* b = raytracer->getBuffer();
  uint32_t hdr[4] = {(uint32_t)b->getWidth(), (uint32_t)b->getHeight(),
                     (uint32_t)static_cast<CUDAHeadlessRenderer*>(renderer)->getFrameCount(),
                     (uint32_t)engine->getFrameCount()};
  std::fwrite(hdr, 4, 4, f);
  dump(f, mesh->getVerticesArray());
  dump(f, mesh->getFacesArray());
  dump(f, b->getInputOriginSurface());
  dump(f, b->getInputDirectionSurface());
  dump(f, b->getOutputFaceIdSurface());
  dump(f, b->getOutputCoordSurface());
  dump(f, b->getOutputDepthSurface());
  dump(f, b->getOutputPointSurface());
  dump(f, b->getOutputNormalSurface());
  dump(f, renderer->getSurface());
  const float3 d = camera->getDir();
  const float4 back = raytracer->getShader()->getBackColor();
  const float extra[8] = {d.x, d.y, d.z, 0.0f, back.x, back.y, back.z, back.w};
  dumpHost(f, extra, sizeof(extra));
  std::fclose(f);

  delete tree;
  delete mesh;
  delete model;
  raytracer->finalize();
  delete raytracer;
  renderer->finalize();
  delete renderer;
  window->hide();
  delete window;
  delete engine;
  return 0;
}
