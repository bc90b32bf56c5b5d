// renderkit drop-in: CUDAWindow (CUDAWindow.h:21-300) without a display. An
// MI355X node has no window system, so the reference's Win32/X11 windows have
// no counterpart; CUDAHeadlessWindow keeps the frame loop of a windowed
// program running (while (window->isActive()) { ... engine->update(); }) for a
// set number of frames, then reports itself inactive, as a closed window does.
#ifndef RENDERKIT_CUDAWINDOW_H
#define RENDERKIT_CUDAWINDOW_H

#include "CUDARenderer.h"

namespace renderkit {

class CUDAWindow {
 public:
  explicit CUDAWindow(CUDAEngine* _engine);  // links itself to the engine
  virtual ~CUDAWindow();

  CUDAEngine* getEngine() { return engine; }
  void setLeft(int v = 120) { left = v; }
  void setTop(int v = 80) { top = v; }
  int getLeft() { return left; }
  int getTop() { return top; }
  void setWidth(int v = 800) { width = v; }
  void setHeight(int v = 600) { height = v; }
  int getWidth() { return width; }
  int getHeight() { return height; }
  int getInnerWidth() { return width; }
  int getInnerHeight() { return height; }
  virtual void show() { visible = true; }
  virtual void hide() { visible = false; }
  bool isVisible() { return visible; }
  // False once the window was closed.
  virtual bool isActive() { return !closed; }
  // Closes the window: isActive() turns false and the close callback runs.
  void close();

  typedef void (*LPCLOSECALLBACK)();
  void registerCloseCallback(LPCLOSECALLBACK cb) { on_close = cb; }
  void unregisterCloseCallback(LPCLOSECALLBACK cb) { if (on_close == cb) on_close = nullptr; }

 protected:
  CUDAEngine* engine;
  int left = 120, top = 80, width = 800, height = 600;
  bool visible = false, closed = false;
  LPCLOSECALLBACK on_close = nullptr;
};

// Active for the first 'frames' frames of the engine (counted by
// CUDAEngine::update()), then closed.
class CUDAHeadlessWindow : public CUDAWindow {
 public:
  CUDAHeadlessWindow(CUDAEngine* _engine, unsigned long long frames) : CUDAWindow(_engine), frame_limit(frames) {}
  bool isActive() override;

 private:
  unsigned long long frame_limit;
};

}  // namespace renderkit

#endif
