// renderkit drop-in: CUDAShader (CUDAShader.h:21-85, CUDAShader.cpp). The
// reference's base shader: a hit pixel is grey by the obliquity of its
// interpolated normal to the camera's direction, a miss takes the back colour
// (rk_shade).
#ifndef RENDERKIT_CUDASHADER_H
#define RENDERKIT_CUDASHADER_H

#include "CUDABuffer.h"
#include "CUDACamera.h"

namespace renderkit {

class CUDAShader {
 public:
  CUDAShader();  // back colour (0.227451, 0.431373, 0.647059, 1), CUDAShader.cpp:26-34
  virtual ~CUDAShader() {}

  float4 getBackColor() { return backColor; }
  void setBackColor(float _r = 0.0f, float _g = 0.0f, float _b = 0.0f, float _a = 0.0f) {
    backColor = make_float4(_r, _g, _b, _a);
  }

  // Shades 'input' (its depth and normal surfaces; interpolateOutputSurfaces
  // with normals fills the latter) into 'output' as seen by 'camera'.
  virtual void run(CUDARenderSurface<uint>* output, CUDABuffer* input, CUDACamera* camera);

 protected:
  float4 backColor;
};

}  // namespace renderkit

#endif
