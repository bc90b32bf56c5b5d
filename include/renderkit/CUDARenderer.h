// renderkit drop-in: CUDARenderer (CUDARenderer.h:20-193), the abstract owner
// of the render surface whose size sets the ray buffer (CUDARaytracer.cpp:73-78),
// and CUDAHeadlessRenderer, a concrete renderer with a device surface and no
// display (the reference's D3D/OpenGL renderers have no counterpart on a
// headless MI355X node). RAW surface files follow CUDARenderer.cpp:128-211.
#ifndef RENDERKIT_CUDARENDERER_H
#define RENDERKIT_CUDARENDERER_H

#include <string>
#include <vector>

#include "CUDAArray.h"

namespace renderkit {

class CUDAEngine;

const int CUDA_FORMAT_BMP = 0;
const int CUDA_FORMAT_JPG = 1;
const int CUDA_FORMAT_TGA = 2;
const int CUDA_FORMAT_PNG = 3;
const int CUDA_FORMAT_DDS = 4;
const int CUDA_FORMAT_PPM = 5;
const int CUDA_FORMAT_DIB = 6;
const int CUDA_FORMAT_HDR = 7;
const int CUDA_FORMAT_PFM = 8;
const int CUDA_FORMAT_RAW = 9;

class CUDARenderer {
 public:
  struct CUDALabel {
    std::string text;
    int x, y;
    float r, g, b, a;
  };

  explicit CUDARenderer(CUDAEngine* _engine);
  virtual ~CUDARenderer();

  CUDAEngine* getEngine() { return engine; }
  virtual void lockSurface() = 0;
  virtual void unlockSurface() = 0;
  CUDARenderSurface<uint>* getSurface() { return surface; }
  void enableFullscreen(int w, int h) { fullscreen = true; (void)w; (void)h; }
  void disableFullscreen() { fullscreen = false; }
  void enableVSync() { verticalSync = true; }
  void disableVSync() { verticalSync = false; }
  void useHardware() { hardware = true; }
  void useSoftware() { hardware = false; }
  void insertLabel(const std::string& text, int x, int y, float r = 1.f, float g = 1.f, float b = 1.f, float a = 1.f) {
    labels.push_back({text, x, y, r, g, b, a});
  }
  void clearForeground() { labels.clear(); }

  virtual void initialize() = 0;
  virtual void finalize() = 0;
  virtual void update() = 0;
  // RAW format only: int header 0x194B7EA2, unit_count, unit_size, size, width,
  // height, pitch, then the bytes.
  virtual void saveSurface(const char* path, const int format, CUDASurface<byte>* surface);
  virtual CUDASurface<byte>* loadSurface(const char* path);

 protected:
  CUDAEngine* engine;
  std::vector<CUDALabel> labels;
  bool fullscreen = false, verticalSync = false, hardware = true;
  CUDARenderSurface<uint>* surface = nullptr;
};

// A renderer with a device-resident ARGB surface of a fixed size and no display.
class CUDAHeadlessRenderer : public CUDARenderer {
 public:
  CUDAHeadlessRenderer(CUDAEngine* _engine, int width, int height);
  ~CUDAHeadlessRenderer() override;
  void lockSurface() override {}
  void unlockSurface() override {}
  void initialize() override;
  void finalize() override;
  void update() override { ++frames; }
  unsigned long long getFrameCount() const { return frames; }

 private:
  int width, height;
  CUDAArray<uint>* storage = nullptr;
  unsigned long long frames = 0;
};

// CUDAEngine (CUDAEngine.h:19-130): links the window, renderer and raytracer
// of an application and keeps its frame clock. update() ends a frame:
// getDelta() is the time since the previous update(); the current and average
// frame rates are refreshed once a second (CUDAEngine.cpp:113-132).
class CUDARaytracer;
class CUDAWindow;
class CUDAEngine {
 public:
  CUDAEngine();
  CUDAWindow* getWindow() { return window; }
  CUDARenderer* getRenderer() { return renderer; }
  void setRenderer(CUDARenderer* r) { renderer = r; }
  CUDARaytracer* getRaytracer() { return raytracer; }
  // Seconds on a monotonic clock.
  double getTime();
  double getDelta() { return frame_delta; }
  double getCurrentFramerate() { return current_framerate; }
  double getAverageFramerate() { return average_framerate; }
  void update();
  // Frames ended by update() so far (no counterpart in the reference).
  unsigned long long getFrameCount() const { return frames; }
  double getFrameTime() const { return frame_delta; }

  CUDAWindow* window = nullptr;
  CUDARenderer* renderer = nullptr;
  CUDARaytracer* raytracer = nullptr;

 private:
  double frame_time = 0.0, cycle_time = 0.0, total_time = 0.0, frame_delta = 0.0;
  double current_framerate = 0.0, average_framerate = 0.0;
  unsigned long long cycle_count = 0, total_count = 0, frames = 0;
};

}  // namespace renderkit

#endif
