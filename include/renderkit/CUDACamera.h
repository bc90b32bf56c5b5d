// renderkit drop-in: CUDACamera (CUDACamera.h:20-325, CUDACamera.cpp:29-267). A y-up
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
  // Moves along the view direction, or sideways in the xz plane
  // (CUDACamera.cpp:163-195).
  void moveForward(float k);
  void moveBackward(float k);
  void moveLeft(float k);
  void moveRight(float k);
  void setFOV(float _fov) { fov = _fov; dirty = true; }
  float getFOV() const { return fov; }
  void setRatio(float _ratio) { ratio = _ratio; dirty = true; }
  void setRatio(int width, int height) { ratio = (float)width / (float)height; dirty = true; }
  float getRatio() const { return ratio; }
  float3 getDirection() const { return dir; }
  float3 getDir() const { return dir; }
  void setZNearPlane(float z) { zNearPlane = z; dirty = true; }
  float getZNearPlane() const { return zNearPlane; }
  void setZFarPlane(float z) { zFarPlane = z; dirty = true; }
  float getZFarPlane() const { return zFarPlane; }
  float pitch() const { return angle.x; }
  float yaw() const { return angle.y; }
  float roll() const { return angle.z; }
  float x() const { return center.x; }
  float y() const { return center.y; }
  float z() const { return center.z; }
  float dx() const { return dir.x; }
  float dy() const { return dir.y; }
  float dz() const { return dir.z; }

  // View and projection matrices of the camera (CUDACamera.cpp:208-263), for
  // rasterising clients; the ray tracer does not use them. Row vectors
  // (p' = p * M), left-handed: the view matrix takes the centre to the origin and
  // the view direction to +z; the projection maps z = zNearPlane to 0 and
  // z = zFarPlane to 1 (Direct3D conventions; the reference's matrix helpers,
  // ../../common/mat44.h, are not part of its sources). update() recomputes them
  // after a change; the getters call it when needed.
  float4x4 getViewMatrix();
  float4x4 getProjMatrix();
  float4x4 getInvViewMatrix();
  float4x4 getInvProjMatrix();
  void update();

  // The per-frame constants of CUDA_calcPrimaryRays (CUDA_calcPrimaryRays.cu:43-55):
  // out[0] ray origin, out[1] ray direction of pixel (0, 0), out[2] step per x,
  // out[3] step per y, for a width x height surface.
  void primaryRayConstants(int width, int height, rk_float4 out[4]) const;

 private:
  void dirFromAngle();
  void angleFromDir();
  float fov, ratio, zNearPlane, zFarPlane;
  float3 center, angle, dir;
  float4x4 matView, matProj, matInvView, matInvProj;
  bool dirty;
};

}  // namespace renderkit

#endif
