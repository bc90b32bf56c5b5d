// renderkit drop-in: CUDACamera (CUDACamera.h, CUDACamera.cpp:29-267). A y-up
// pinhole camera: center, angles (x = pitch, the azimuth; y = yaw, the
// elevation; z = roll), fov (radians) and aspect ratio.
#ifndef RENDERKIT_CUDACAMERA_H
#define RENDERKIT_CUDACAMERA_H

#include "CUDACommon.h"

namespace renderkit {

class CUDACamera {
 public:
  CUDACamera();

  void setCenter(float x, float y, float z);
  void setCenter(float3 c);
  float3 getCenter() const { return center; }
  void setAngle(float x, float y, float z);
  void setAngle(float3 a);
  float3 getAngle() const { return angle; }
  void lookAt(float x, float y, float z);
  void lookAt(float3 p);
  void translate(float x, float y, float z);
  void rotate(float x, float y, float z);
  void setFOV(float _fov) { fov = _fov; dirty = true; }
  float getFOV() const { return fov; }
  void setRatio(float _ratio) { ratio = _ratio; dirty = true; }
  void setRatio(int width, int height) { ratio = (float)width / (float)height; dirty = true; }
  float getRatio() const { return ratio; }
  float3 getDirection() const { return dir; }
  float pitch() const { return angle.x; }
  float yaw() const { return angle.y; }
  float roll() const { return angle.z; }

  // The per-frame constants of CUDA_calcPrimaryRays (CUDA_calcPrimaryRays.cu:43-55):
  // out[0] ray origin, out[1] ray direction of pixel (0, 0), out[2] step per x,
  // out[3] step per y, for a width x height surface.
  void primaryRayConstants(int width, int height, rk_float4 out[4]) const;

 private:
  void dirFromAngle();
  void angleFromDir();
  float fov, ratio, zNearPlane, zFarPlane;
  float3 center, angle, dir;
  bool dirty;
};

}  // namespace renderkit

#endif
