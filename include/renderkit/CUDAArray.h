// renderkit drop-in: CUDAArray<T> (CUDAArray.h:19-213) and CUDASurface<T>
// (CUDASurface.h), RAII device arrays over rk_malloc/rk_free/rk_memcpy.
#ifndef RENDERKIT_CUDAARRAY_H
#define RENDERKIT_CUDAARRAY_H

#include "CUDACommon.h"

namespace renderkit {

template <typename unit_type>
class CUDAArray {
 protected:
  int unit_count;
  int unit_size;
  int size;
  unit_type* data;

 public:
  CUDAArray() : unit_count(0), unit_size(sizeof(unit_type)), size(0), data(nullptr) {}
  // CUDAArray.h:60-70
  explicit CUDAArray(int _unit_count) : CUDAArray(_unit_count, (int)sizeof(unit_type)) {}
  CUDAArray(int _unit_count, int _unit_size) : data(nullptr) {
    _assert(_unit_count > 0, __FILE__, __LINE__,
            "CUDAArray::CUDAArray() : Invalid parameter '_unit_count', must be positive.");
    _assert(_unit_size > 0, __FILE__, __LINE__,
            "CUDAArray::CUDAArray() : Invalid parameter '_unit_size', must be positive.");
    unit_count = _unit_count;
    unit_size = _unit_size;
    size = unit_count * unit_size;
    initialize();
  }
  virtual ~CUDAArray() { finalize(); }
  CUDAArray(const CUDAArray&) = delete;
  CUDAArray& operator=(const CUDAArray&) = delete;

  int getUnitCount() { return unit_count; }
  int getUnitSize() { return unit_size; }
  int getSize() { return size; }
  unit_type* getPointer() { return data; }

  void copyDataFromHost(const unit_type* host_ptr) { RK_CALL(rk_memcpy(data, host_ptr, size, 0, nullptr)); }
  void copyDataToHost(unit_type* host_ptr) { RK_CALL(rk_memcpy(host_ptr, data, size, 1, nullptr)); }
  void copyDataFromDevice(const unit_type* device_ptr) { RK_CALL(rk_memcpy(data, device_ptr, size, 2, nullptr)); }
  void copyDataToDevice(unit_type* device_ptr) { RK_CALL(rk_memcpy(device_ptr, data, size, 2, nullptr)); }

 protected:
  void initialize() {
    void* p = nullptr;
    _assert(rk_malloc(&p, (size_t)size) == RK_SUCCESS, __FILE__, __LINE__,
            "CUDAArray::initialize() : Unable to allocate enough video memory.");
    data = static_cast<unit_type*>(p);
  }
  void finalize() {
    if (data) rk_free(data);
    data = nullptr;
  }
};

template <typename unit_type>
class CUDASurface : public CUDAArray<unit_type> {
  friend class CUDARenderer;

 protected:
  int width = 0;
  int height = 0;
  int pitch = 0;

 public:
  CUDASurface() {}
  CUDASurface(int _width, int _height) : CUDASurface(_width, _height, (int)sizeof(unit_type)) {}
  CUDASurface(int _width, int _height, int _unit_size) {
    _assert(_width > 0, __FILE__, __LINE__, "CUDASurface::CUDASurface() : Invalid parameter '_width', must be positive.");
    _assert(_height > 0, __FILE__, __LINE__,
            "CUDASurface::CUDASurface() : Invalid parameter '_height', must be positive.");
    width = _width;
    height = _height;
    this->unit_count = width * height;
    this->unit_size = _unit_size;
    this->size = this->unit_count * this->unit_size;
    pitch = width * this->unit_size;
    this->initialize();
  }
  int getWidth() { return width; }
  int getHeight() { return height; }
  int getPitch() { return pitch; }

  // Used by CUDARenderer::loadSurface (CUDARenderer.cpp:160-200).
  void assign(int w, int h, int unit_size, int pitch_) {
    this->finalize();
    width = w;
    height = h;
    this->unit_size = unit_size;
    this->unit_count = w * h;
    this->size = this->unit_count * unit_size;
    pitch = pitch_;
    this->initialize();
  }
};

// A surface that does not own its memory (CUDARenderSurface.h:88-91).
template <typename unit_type>
class CUDARenderSurface : public CUDASurface<unit_type> {
 public:
  CUDARenderSurface(int _width, int _height) {
    _assert(_width > 0 && _height > 0, __FILE__, __LINE__, "CUDARenderSurface::CUDARenderSurface() : Invalid size.");
    this->width = _width;
    this->height = _height;
    this->unit_count = _width * _height;
    this->unit_size = sizeof(unit_type);
    this->size = this->unit_count * this->unit_size;
    this->pitch = this->width * this->unit_size;
  }
  void setPointer(unit_type* p) { this->data = p; }
  ~CUDARenderSurface() override { this->data = nullptr; }
};

}  // namespace renderkit

#endif
