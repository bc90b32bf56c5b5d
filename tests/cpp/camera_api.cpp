// The CUDACamera surface beyond the ray constants, host only (no GPU call):
// movement (CUDACamera.cpp:163-195), accessors, and the view/projection
// matrices. Prints "ok" and exits 0 when every check holds.
#include <cmath>
#include <cstdio>

#include "../../include/renderkit/CUDACamera.h"

using namespace renderkit;

static int failures = 0;
static void check(bool ok, const char* what) {
  if (!ok) {
    std::printf("FAILED: %s\n", what);
    ++failures;
  }
}
static bool near(float a, float b, float tol = 1e-4f) { return std::fabs(a - b) <= tol * (1.0f + std::fabs(b)); }
static float3 xform(const float4x4& m, float3 p, float* w_out = nullptr) {
  float r[4];
  for (int j = 0; j < 4; ++j) r[j] = p.x * m.m[0][j] + p.y * m.m[1][j] + p.z * m.m[2][j] + m.m[3][j];
  if (w_out) *w_out = r[3];
  return make_float3(r[0], r[1], r[2]);
}

int main() {
  CUDACamera c;
  check(c.getZNearPlane() == 1.0f && c.getZFarPlane() == 1999.999f, "default planes");
  c.setCenter(1.0f, 2.0f, 3.0f);
  c.setAngle(0.0f, 0.0f, 0.0f);  // dir (1, 0, 0)
  check(c.dx() == 1.0f && c.dy() == 0.0f && c.dz() == 0.0f, "dir of angle 0");
  c.moveForward(2.0f);
  check(c.x() == 3.0f && c.y() == 2.0f && c.z() == 3.0f, "moveForward");
  c.moveBackward(1.0f);
  check(c.x() == 2.0f, "moveBackward");
  c.moveLeft(1.0f);  // x -= k dir.z, z += k dir.x
  check(c.x() == 2.0f && c.z() == 4.0f, "moveLeft");
  c.moveRight(2.0f);  // x += k dir.z, z -= k dir.x
  check(c.x() == 2.0f && c.z() == 2.0f, "moveRight");

  // view: the centre to the origin, a point ahead on +z, up (0, 1, 0) on +y
  c.setCenter(-10.0f, 20.0f, 30.0f);
  c.lookAt(0.0f, 0.0f, 0.0f);
  c.setZNearPlane(0.5f);
  c.setZFarPlane(100.0f);
  c.setFOV(PI * 0.5f);
  c.setRatio(2.0f);
  float4x4 v = c.getViewMatrix();
  float3 e = xform(v, c.getCenter());
  check(near(e.x, 0) && near(e.y, 0) && near(e.z, 0, 1e-3f), "view: centre at origin");
  float3 d = c.getDir();
  float3 ahead = xform(v, make_float3(-10.0f + 5.0f * d.x, 20.0f + 5.0f * d.y, 30.0f + 5.0f * d.z));
  check(near(ahead.x, 0, 1e-3f) && near(ahead.y, 0, 1e-3f) && near(ahead.z, 5.0f), "view: dir to +z");
  float3 upw = xform(v, make_float3(-10.0f, 21.0f, 30.0f));
  check(upw.y > 0.0f, "view: world up is up");
  // projection: z = near -> 0, z = far -> 1 after the divide by w = z
  float4x4 p = c.getProjMatrix();
  float w = 0;
  float3 pn = xform(p, make_float3(0, 0, 0.5f), &w);
  check(near(pn.z / w, 0.0f, 1e-5f), "proj: near plane at 0");
  float3 pf = xform(p, make_float3(0, 0, 100.0f), &w);
  check(near(pf.z / w, 1.0f, 1e-5f), "proj: far plane at 1");
  float3 edge = xform(p, make_float3(2.0f, 1.0f, 1.0f), &w);  // tan(fov/2) = 1, ratio 2: the frustum corner
  check(near(edge.x / w, 1.0f) && near(edge.y / w, 1.0f), "proj: fov and ratio");
  // inverses
  float4x4 iv = c.getInvViewMatrix(), ip = c.getInvProjMatrix();
  float3 q = xform(iv, xform(v, make_float3(3.0f, -4.0f, 7.0f)));
  check(near(q.x, 3.0f) && near(q.y, -4.0f) && near(q.z, 7.0f), "inverse view");
  float w1 = 0;
  float3 r = xform(p, make_float3(0.3f, -0.2f, 5.0f), &w1);
  float r4[4] = {r.x, r.y, r.z, w1}, back[4] = {0, 0, 0, 0};
  for (int j = 0; j < 4; ++j)
    for (int i = 0; i < 4; ++i) back[j] += r4[i] * ip.m[i][j];
  check(near(back[0], 0.3f) && near(back[1], -0.2f) && near(back[2], 5.0f) && near(back[3], 1.0f), "inverse proj");
  // a changed camera recomputes its matrices
  c.translate(1.0f, 0.0f, 0.0f);
  float3 e2 = xform(c.getViewMatrix(), c.getCenter());
  check(near(e2.x, 0, 1e-3f) && near(e2.z, 0, 1e-3f), "update after translate");
  if (failures == 0) std::printf("ok\n");
  return failures == 0 ? 0 : 1;
}
