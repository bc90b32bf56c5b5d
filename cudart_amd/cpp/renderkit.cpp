// renderkit drop-in classes (include/renderkit/) over the C ABI (include/rk.h).
// Host-only: no GPU runtime header is included; device work is rk_* calls.
#include <chrono>
#include <cstring>
#include <vector>

#include "../../include/renderkit/CUDARaytracer.h"

namespace renderkit {

bool& throwOnErrorFlag() {
  static bool flag = false;
  return flag;
}

namespace {

// Scratch device memory for one call, freed on scope exit (the reference
// allocates and frees its temporaries inside every CUDA_* call).
struct Workspace {
  void* p = nullptr;
  size_t bytes = 0;
  explicit Workspace(size_t n) : bytes(n) {
    if (n) _assert(rk_malloc(&p, n) == RK_SUCCESS, __FILE__, __LINE__, "Unable to allocate a build workspace.");
  }
  ~Workspace() { if (p) rk_free(p); }
};

rk_aabox toRk(const CUDABox& b) {
  rk_aabox r;
  std::memcpy(&r, &b, sizeof(r));
  return r;
}

// .cdm chunks (CUDAMesh.cpp:134-400, CUDAMeshTree.cpp:114-289): int magic,
// int length (of the whole chunk, header included), then the payload. An array
// chunk's payload is uint count and the raw array; the hitbox chunk's payload is
// the aabox alone (CUDAMesh.cpp:294-296). Readers skip unknown chunks by their
// length and read known ones in sequence, as the reference does.
void writeChunk(FILE* f, int magic, const void* data, uint count, size_t unit) {
  int length = (int)(sizeof(int) * 2 + sizeof(uint) + (size_t)count * unit);
  std::fwrite(&magic, sizeof(int), 1, f);
  std::fwrite(&length, sizeof(int), 1, f);
  std::fwrite(&count, sizeof(uint), 1, f);
  if (count) std::fwrite(data, unit, count, f);
}

void writeHitboxChunk(FILE* f, int magic, const CUDABox& box) {
  int length = (int)(sizeof(int) * 2 + sizeof(box));
  std::fwrite(&magic, sizeof(int), 1, f);
  std::fwrite(&length, sizeof(int), 1, f);
  std::fwrite(&box, sizeof(box), 1, f);
}

const int HITBOX_HEADER = 0x3027AD79;
const int VERTICES_HEADER = (int)0xBA5C1CC7;
const int NORMALS_HEADER = 0x1DD14C48;
const int COLORS_HEADER = 0x757B31AB;
const int TEXCOORDS_HEADER = 0x2B77D260;
const int FACES_HEADER = (int)0x8E934163;
const int NODES_HEADER = (int)0x9077086B;
const int INDICES_HEADER = 0x2CC8BD7D;
const int BOXES_HEADER = (int)0xE54B7842;

template <typename T>
void saveArray(FILE* f, int magic, CUDAArray<T>* a) {
  if (!a) return;
  std::vector<T> host(a->getUnitCount());
  a->copyDataToHost(host.data());
  writeChunk(f, magic, host.data(), (uint)host.size(), sizeof(T));
}

template <typename T>
CUDAArray<T>* loadArray(FILE* f, uint count) {
  std::vector<T> host(count);
  _assert(std::fread(host.data(), sizeof(T), count, f) == count, __FILE__, __LINE__, "quickLoad() : Truncated chunk.");
  auto* a = new CUDAArray<T>((int)count);
  a->copyDataFromHost(host.data());
  return a;
}

}  // namespace

// ---- CUDAMesh -------------------------------------------------------------------

CUDAMesh::CUDAMesh() { hitbox.reset(); }
CUDAMesh::~CUDAMesh() { release(); }

void CUDAMesh::release() {
  delete faces; faces = nullptr;
  delete texcoords; texcoords = nullptr;
  delete colors; colors = nullptr;
  delete normals; normals = nullptr;
  delete vertices; vertices = nullptr;
}

void CUDAMesh::buildFromModel(CUDAModel* _model) {
  release();
  model = _model;
  hitbox = model->getHitbox();
  uint nv = model->getVertexCount(), nf = model->getFaceCount();
  if (nv > 0) {
    if (model->hasVertices) { vertices = new CUDAArray<float4>((int)nv); vertices->copyDataFromHost(model->getVertexPointer()); }
    if (model->hasNormals) { normals = new CUDAArray<float4>((int)nv); normals->copyDataFromHost(model->getNormalPointer()); }
    if (model->hasColors) { colors = new CUDAArray<float4>((int)nv); colors->copyDataFromHost(model->getColorPointer()); }
    if (model->hasTexcoords) { texcoords = new CUDAArray<float2>((int)nv); texcoords->copyDataFromHost(model->getTexcoordPointer()); }
  }
  if (nf > 0) { faces = new CUDAArray<uint4>((int)nf); faces->copyDataFromHost(model->getFacePointer()); }
}

void CUDAMesh::buildFromArrays(const float4* hv, uint nv, const uint4* hf, uint nf) {
  release();
  model = nullptr;
  rk_aabox hb;
  RK_CALL(rk_scene_hitbox(reinterpret_cast<const rk_float4*>(hv), nv, &hb));
  std::memcpy(&hitbox, &hb, sizeof(hb));
  if (nv) { vertices = new CUDAArray<float4>((int)nv); vertices->copyDataFromHost(hv); }
  if (nf) { faces = new CUDAArray<uint4>((int)nf); faces->copyDataFromHost(hf); }
}

void CUDAMesh::quickSave(const char* path, bool overwrite) {
  FILE* f = std::fopen(path, overwrite ? "wb" : "ab");
  _assert(f != nullptr, __FILE__, __LINE__, "CUDAMesh::quickSave() : Invalid file.");
  writeHitboxChunk(f, HITBOX_HEADER, hitbox);
  saveArray(f, VERTICES_HEADER, vertices);
  saveArray(f, NORMALS_HEADER, normals);
  saveArray(f, COLORS_HEADER, colors);
  saveArray(f, TEXCOORDS_HEADER, texcoords);
  saveArray(f, FACES_HEADER, faces);
  std::fclose(f);
}

void CUDAMesh::quickLoad(const char* path) {
  FILE* f = std::fopen(path, "rb");
  _assert(f != nullptr, __FILE__, __LINE__, "CUDAMesh::quickLoad() : Invalid file.");
  release();
  model = nullptr;
  std::fseek(f, 0, SEEK_END);
  long file_length = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  long off;
  while ((off = std::ftell(f)) < file_length) {
    int header = 0, length = 0;
    uint count = 0;
    if (std::fread(&header, sizeof(int), 1, f) != 1 || std::fread(&length, sizeof(int), 1, f) != 1) break;
    bool known = header == HITBOX_HEADER || header == VERTICES_HEADER || header == NORMALS_HEADER ||
                 header == COLORS_HEADER || header == TEXCOORDS_HEADER || header == FACES_HEADER;
    if (!known) { std::fseek(f, off + length, SEEK_SET); continue; }
    if (header == HITBOX_HEADER) {
      _assert(std::fread(&hitbox, sizeof(hitbox), 1, f) == 1, __FILE__, __LINE__, "CUDAMesh::quickLoad() : Truncated file.");
      continue;
    }
    _assert(std::fread(&count, sizeof(uint), 1, f) == 1, __FILE__, __LINE__, "CUDAMesh::quickLoad() : Truncated file.");
    if (header == VERTICES_HEADER) { delete vertices; vertices = loadArray<float4>(f, count); }
    else if (header == NORMALS_HEADER) { delete normals; normals = loadArray<float4>(f, count); }
    else if (header == COLORS_HEADER) { delete colors; colors = loadArray<float4>(f, count); }
    else if (header == TEXCOORDS_HEADER) { delete texcoords; texcoords = loadArray<float2>(f, count); }
    else { delete faces; faces = loadArray<uint4>(f, count); }
  }
  std::fclose(f);
}

// ---- CUDAMeshTree -----------------------------------------------------------------

CUDAMeshTree::CUDAMeshTree(CUDAMesh* _mesh) : mesh(_mesh) {  // CUDAMeshTree.cpp:51-70
  int n = mesh->getFaceCount();
  if (n > 0) {
    boxes = new CUDAArray<aabox>(n);
    indices = new CUDAArray<uint>(n);
  }
}

CUDAMeshTree::~CUDAMeshTree() {
  delete nodes;
  delete boxes;
  delete indices;
  releaseFast();
}

void CUDAMeshTree::releaseFast() {
  delete fast_nodes; fast_nodes = nullptr;
  delete fast_tris; fast_tris = nullptr;
}

void CUDAMeshTree::updateNodes(bvnode* nodes_buffer, uint node_count) {
  delete nodes;
  nodes = new CUDAArray<bvnode>((int)node_count);
  nodes->copyDataFromDevice(nodes_buffer);
}

void CUDAMeshTree::buildMeshTreeLBVH(uint tree_depth) {
  uint n = (uint)mesh->getFaceCount();
  _assert(n > 0 && boxes && indices, __FILE__, __LINE__, "CUDAMeshTree::buildMeshTreeLBVH() : Empty mesh.");
  releaseFast();
  Workspace ws(rk_build_lbvh_workspace_size(n));
  CUDAArray<bvnode> buffer((int)(2 * n));
  CUDAArray<uint> info(RK_INFO_WORDS);
  CUDABox hb = mesh->getHitbox();
  rk_aabox rhb = toRk(hb);
  RK_CALL(rk_build_lbvh(nullptr, reinterpret_cast<rk_uint4*>(mesh->getFacesArray()->getPointer()),
                        reinterpret_cast<const rk_float4*>(mesh->getVerticesArray()->getPointer()), n, &rhb, 1,
                        tree_depth, reinterpret_cast<rk_aabox*>(boxes->getPointer()), indices->getPointer(),
                        reinterpret_cast<rk_bvnode*>(buffer.getPointer()), 2 * n, info.getPointer(), ws.p, ws.bytes));
  info.copyDataToHost(info_host);
  updateNodes(buffer.getPointer(), info_host[RK_INFO_NODE_COUNT]);
}

void CUDAMeshTree::buildMeshTreeSAH() {
  uint n = (uint)mesh->getFaceCount();
  _assert(n > 0 && boxes && indices, __FILE__, __LINE__, "CUDAMeshTree::buildMeshTreeSAH() : Empty mesh.");
  releaseFast();
  uint existing = getNodeCount();
  uint cap = existing + 2 * n + 1;
  CUDAArray<bvnode> buffer((int)cap);
  if (existing) buffer.copyDataFromDevice(nodes->getPointer());
  CUDAArray<uint> info(RK_INFO_WORDS);
  Workspace ws(rk_build_sah_workspace_size(n, existing));
  CUDABox hb = mesh->getHitbox();
  rk_aabox rhb = toRk(hb);
  RK_CALL(rk_build_sah(nullptr, reinterpret_cast<rk_uint4*>(mesh->getFacesArray()->getPointer()),
                       reinterpret_cast<const rk_float4*>(mesh->getVerticesArray()->getPointer()), n, &rhb,
                       reinterpret_cast<rk_aabox*>(boxes->getPointer()), indices->getPointer(),
                       reinterpret_cast<rk_bvnode*>(buffer.getPointer()), existing, cap, info.getPointer(), ws.p,
                       ws.bytes));
  info.copyDataToHost(info_host);
  updateNodes(buffer.getPointer(), info_host[RK_INFO_NODE_COUNT]);
}

void CUDAMeshTree::buildMeshTreeHybrid(uint lbvh_tree_depth) {  // CUDAMeshTree.cpp:98-103
  uint n = (uint)mesh->getFaceCount();
  _assert(n > 0 && boxes && indices, __FILE__, __LINE__, "CUDAMeshTree::buildMeshTreeHybrid() : Empty mesh.");
  releaseFast();
  {
    Workspace ws(rk_build_lbvh_workspace_size(n));
    CUDAArray<bvnode> buffer((int)(2 * n));
    CUDAArray<uint> info(RK_INFO_WORDS);
    rk_aabox rhb = toRk(mesh->getHitbox());
    RK_CALL(rk_build_lbvh(nullptr, reinterpret_cast<rk_uint4*>(mesh->getFacesArray()->getPointer()),
                          reinterpret_cast<const rk_float4*>(mesh->getVerticesArray()->getPointer()), n, &rhb, 0,
                          lbvh_tree_depth, reinterpret_cast<rk_aabox*>(boxes->getPointer()), indices->getPointer(),
                          reinterpret_cast<rk_bvnode*>(buffer.getPointer()), 2 * n, info.getPointer(), ws.p, ws.bytes));
    info.copyDataToHost(info_host);
    updateNodes(buffer.getPointer(), info_host[RK_INFO_NODE_COUNT]);
  }
  buildMeshTreeSAH();
}

void CUDAMeshTree::buildMeshTreeFast(uint leaf_max, bool morton63) {
  uint n = (uint)mesh->getFaceCount();
  _assert(n > 0 && boxes && indices, __FILE__, __LINE__, "CUDAMeshTree::buildMeshTreeFast() : Empty mesh.");
  releaseFast();
  delete nodes;
  nodes = nullptr;
  int mode = morton63 ? RK_MORTON_63 : RK_MORTON_30;
  fast_nodes = new CUDAArray<float4>((int)(4 * n));
  fast_tris = new CUDAArray<float4>((int)(3 * n));
  Workspace ws(rk_fast_build_workspace_size(n, mode));
  CUDAArray<uint> info(RK_INFO_WORDS);
  rk_fast_bvh d = {fast_nodes->getPointer(), fast_tris->getPointer(), indices->getPointer(),
                   reinterpret_cast<rk_aabox*>(boxes->getPointer()), n, 0};
  rk_aabox rhb = toRk(mesh->getHitbox());
  RK_CALL(rk_fast_build(nullptr, reinterpret_cast<const rk_uint4*>(mesh->getFacesArray()->getPointer()),
                        reinterpret_cast<const rk_float4*>(mesh->getVerticesArray()->getPointer()), n, &rhb, mode,
                        leaf_max, &d, info.getPointer(), ws.p, ws.bytes));
  // faces in leaf order, as the reference builds leave the mesh (CUDA_reorderMeshFaces)
  Workspace rws(rk_reorder_workspace_size(n));
  RK_CALL(rk_reorder_mesh_faces(nullptr, reinterpret_cast<rk_uint4*>(mesh->getFacesArray()->getPointer()),
                                indices->getPointer(), n, rws.p, rws.bytes));
  info.copyDataToHost(info_host);
  RK_CALL(rk_stream_synchronize(nullptr));
}

void CUDAMeshTree::quickSave(const char* path, bool overwrite) {
  // The reference opens append mode and then back-patches chunk lengths at
  // positions that, in append mode, land at the end of the file
  // (CUDAMeshTree.cpp:201,222-228); here each chunk's length is known before
  // it is written, so both modes produce well-formed chunks.
  FILE* f = std::fopen(path, overwrite ? "wb" : "ab");
  _assert(f != nullptr, __FILE__, __LINE__, "CUDAMeshTree::quickSave() : Invalid file.");
  saveArray(f, NODES_HEADER, nodes);
  saveArray(f, INDICES_HEADER, indices);
  saveArray(f, BOXES_HEADER, boxes);
  std::fclose(f);
}

void CUDAMeshTree::quickLoad(const char* path) {
  FILE* f = std::fopen(path, "rb");
  _assert(f != nullptr, __FILE__, __LINE__, "CUDAMeshTree::quickLoad() : Invalid file.");
  std::fseek(f, 0, SEEK_END);
  long file_length = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  long off;
  while ((off = std::ftell(f)) < file_length) {
    int header = 0, length = 0;
    uint count = 0;
    if (std::fread(&header, sizeof(int), 1, f) != 1 || std::fread(&length, sizeof(int), 1, f) != 1) break;
    if (header != NODES_HEADER && header != INDICES_HEADER && header != BOXES_HEADER) {
      std::fseek(f, off + length, SEEK_SET);
      continue;
    }
    _assert(std::fread(&count, sizeof(uint), 1, f) == 1, __FILE__, __LINE__, "CUDAMeshTree::quickLoad() : Truncated file.");
    if (header == NODES_HEADER) {
      delete nodes;
      nodes = loadArray<bvnode>(f, count);
      releaseFast();
    } else if (header == INDICES_HEADER) {
      // the reference asserts '!=' here (CUDAMeshTree.cpp:166), rejecting valid files
      _assert(indices && (uint)indices->getUnitCount() == count, __FILE__, __LINE__,
              "CUDAMeshTree::quickLoad() : Invalid index count.");
      std::vector<uint> host(count);
      _assert(std::fread(host.data(), sizeof(uint), count, f) == count, __FILE__, __LINE__, "quickLoad() : Truncated chunk.");
      indices->copyDataFromHost(host.data());
    } else {
      _assert(boxes && (uint)boxes->getUnitCount() == count, __FILE__, __LINE__,  // :182, inverted in the reference
              "CUDAMeshTree::quickLoad() : Invalid box count.");
      std::vector<aabox> host(count);
      _assert(std::fread(host.data(), sizeof(aabox), count, f) == count, __FILE__, __LINE__, "quickLoad() : Truncated chunk.");
      boxes->copyDataFromHost(host.data());
    }
  }
  std::fclose(f);
}

// ---- CUDABuffer -------------------------------------------------------------------

CUDABuffer::CUDABuffer(int _width, int _height) {
  _assert(_width > 0, __FILE__, __LINE__, "CUDABuffer::CUDABuffer() : Invalid parameter '_width', must be positive.");
  _assert(_height > 0, __FILE__, __LINE__, "CUDABuffer::CUDABuffer() : Invalid parameter '_height', must be positive.");
  width = _width;
  height = _height;
  in_origins = new CUDASurface<float4>(width, height);
  in_directions = new CUDASurface<float4>(width, height);
  out_faces_id = new CUDASurface<uint>(width, height);
  out_coords = new CUDASurface<float2>(width, height);
  out_depths = new CUDASurface<float>(width, height);
}

CUDABuffer::~CUDABuffer() {
  delete out_depths; delete out_coords; delete out_faces_id; delete out_materials; delete out_texcoords;
  delete out_colors; delete out_normals; delete out_points; delete in_directions; delete in_origins;
}

void CUDABuffer::clearInputSurfaces() {
  size_t n = (size_t)width * height;
  rk_float4 z = {0.f, 0.f, 0.f, 0.f};
  RK_CALL(rk_fill_f4(nullptr, reinterpret_cast<rk_float4*>(in_origins->getPointer()), n, z));
  RK_CALL(rk_fill_f4(nullptr, reinterpret_cast<rk_float4*>(in_directions->getPointer()), n, z));
}

void CUDABuffer::clearOutputSurfaces() {
  size_t n = (size_t)width * height;
  rk_float2 z2 = {0.f, 0.f};
  rk_float4 z4 = {0.f, 0.f, 0.f, 0.f};
  RK_CALL(rk_fill_u32(nullptr, out_faces_id->getPointer(), n, 0u));
  RK_CALL(rk_fill_f2(nullptr, reinterpret_cast<rk_float2*>(out_coords->getPointer()), n, z2));
  RK_CALL(rk_fill_f32(nullptr, out_depths->getPointer(), n, 3.402823466e+38f));
  if (out_materials) RK_CALL(rk_fill_u32(nullptr, out_materials->getPointer(), n, 0u));
  if (out_texcoords) RK_CALL(rk_fill_f2(nullptr, reinterpret_cast<rk_float2*>(out_texcoords->getPointer()), n, z2));
  if (out_colors) RK_CALL(rk_fill_f4(nullptr, reinterpret_cast<rk_float4*>(out_colors->getPointer()), n, z4));
  if (out_normals) RK_CALL(rk_fill_f4(nullptr, reinterpret_cast<rk_float4*>(out_normals->getPointer()), n, z4));
  if (out_points) RK_CALL(rk_fill_f4(nullptr, reinterpret_cast<rk_float4*>(out_points->getPointer()), n, z4));
}

// ---- CUDACamera -------------------------------------------------------------------

CUDACamera::CUDACamera() {  // CUDACamera.cpp:29-50
  fov = PI * 0.5f;
  ratio = 4.0f / 3.0f;
  zNearPlane = 1.000f;
  zFarPlane = 1999.999f;
  center = make_float3(0.0f, 0.0f, 0.0f);
  angle = make_float3(PI * 0.5f, 0.0f, 0.0f);
  dir = make_float3(0.0f, 0.0f, 1.0f);
  dirty = true;
}

void CUDACamera::dirFromAngle() {  // CUDACamera.cpp:56-68
  if (angle.x < 0.0f) angle.x += PI * 2.0f;
  if (angle.x > PI * 2.0f) angle.x -= PI * 2.0f;
  if (angle.y < -PI * 0.5f + 0.001f) angle.y = -PI * 0.5f + 0.001f;
  if (angle.y > PI * 0.5f - 0.001f) angle.y = PI * 0.5f - 0.001f;
  dir = normalize(make_float3(cosf(angle.x) * cosf(angle.y), sinf(angle.y), sinf(angle.x) * cosf(angle.y)));
}

void CUDACamera::angleFromDir() {  // CUDACamera.cpp:70-88
  dir = normalize(dir);
  angle.y = asinf(dir.y);
  if (angle.y < -PI * 0.5f + 0.001f) angle.y = -PI * 0.5f + 0.001f;
  if (angle.y > PI * 0.5f - 0.001f) angle.y = PI * 0.5f - 0.001f;
  if (dir.x > 0) angle.x = asinf(dir.z / cosf(angle.y));
  else if (dir.z > 0) angle.x = acosf(dir.x / cosf(angle.y));
  else angle.x = acosf(-dir.x / cosf(-angle.y)) + PI;
  if (angle.x < 0.0f) angle.x += PI * 2.0f;
  if (angle.x > PI * 2.0f) angle.x -= PI * 2.0f;
}

void CUDACamera::setCenter(float x, float y, float z) { center = make_float3(x, y, z); dirty = true; }
void CUDACamera::setCenter(float3 c) { center = c; dirty = true; }
void CUDACamera::setAngle(float x, float y, float z) { angle = make_float3(x, y, z); dirFromAngle(); dirty = true; }
void CUDACamera::setAngle(float3 a) { angle = a; dirFromAngle(); dirty = true; }
void CUDACamera::lookAt(float x, float y, float z) { dir = make_float3(x, y, z) - center; angleFromDir(); dirty = true; }
void CUDACamera::lookAt(float3 p) { dir = p - center; angleFromDir(); dirty = true; }
void CUDACamera::translate(float x, float y, float z) { center.x += x; center.y += y; center.z += z; dirty = true; }
void CUDACamera::rotate(float x, float y, float z) {
  angle.x -= x;
  angle.y -= y;
  angle.z -= z;
  dirFromAngle();
  dirty = true;
}

void CUDACamera::moveForward(float k) {  // CUDACamera.cpp:163-170
  center.x += k * dir.x;
  center.y += k * dir.y;
  center.z += k * dir.z;
  dirty = true;
}
void CUDACamera::moveBackward(float k) {  // :172-179
  center.x -= k * dir.x;
  center.y -= k * dir.y;
  center.z -= k * dir.z;
  dirty = true;
}
void CUDACamera::moveLeft(float k) {  // :181-187
  center.x -= k * dir.z;
  center.z += k * dir.x;
  dirty = true;
}
void CUDACamera::moveRight(float k) {  // :189-195
  center.x += k * dir.z;
  center.z -= k * dir.x;
  dirty = true;
}

namespace {
float4x4 identity4() {
  float4x4 m = {};
  for (int i = 0; i < 4; ++i) m.m[i][i] = 1.0f;
  return m;
}
// Rotation by 'angle' about 'axis' for row vectors (p' = p * M); identity for
// a zero angle or a zero axis.
float4x4 rotation_axis(float3 axis, float angle) {
  const float len2 = dot(axis, axis);
  if (angle == 0.0f || !(len2 > 0.0f)) return identity4();
  const float3 a = normalize(axis);
  const float c = cosf(angle), s = sinf(angle), t = 1.0f - c;
  float4x4 m = identity4();
  m.m[0][0] = t * a.x * a.x + c;       m.m[0][1] = t * a.x * a.y + s * a.z; m.m[0][2] = t * a.x * a.z - s * a.y;
  m.m[1][0] = t * a.x * a.y - s * a.z; m.m[1][1] = t * a.y * a.y + c;       m.m[1][2] = t * a.y * a.z + s * a.x;
  m.m[2][0] = t * a.x * a.z + s * a.y; m.m[2][1] = t * a.y * a.z - s * a.x; m.m[2][2] = t * a.z * a.z + c;
  return m;
}
float3 transform_point(const float4x4& m, float3 p) {  // (p, 1) * M, w dropped
  return make_float3(p.x * m.m[0][0] + p.y * m.m[1][0] + p.z * m.m[2][0] + m.m[3][0],
                     p.x * m.m[0][1] + p.y * m.m[1][1] + p.z * m.m[2][1] + m.m[3][1],
                     p.x * m.m[0][2] + p.y * m.m[1][2] + p.z * m.m[2][2] + m.m[3][2]);
}
// Left-handed look-at: the eye goes to the origin, at - eye to +z.
float4x4 look_at_lh(float3 eye, float3 at, float3 up) {
  const float3 z = normalize(at - eye);
  const float3 x = normalize(cross(up, z));
  const float3 y = cross(z, x);
  float4x4 m = {};
  m.m[0][0] = x.x; m.m[0][1] = y.x; m.m[0][2] = z.x;
  m.m[1][0] = x.y; m.m[1][1] = y.y; m.m[1][2] = z.y;
  m.m[2][0] = x.z; m.m[2][1] = y.z; m.m[2][2] = z.z;
  m.m[3][0] = -dot(x, eye); m.m[3][1] = -dot(y, eye); m.m[3][2] = -dot(z, eye);
  m.m[3][3] = 1.0f;
  return m;
}
// Left-handed perspective: vertical fov, aspect = width / height; view z in
// [zn, zf] maps to depth [0, 1] after the divide by w = z.
float4x4 perspective_fov_lh(float fovy, float aspect, float zn, float zf) {
  const float ys = 1.0f / tanf(fovy * 0.5f), xs = ys / aspect;
  float4x4 m = {};
  m.m[0][0] = xs;
  m.m[1][1] = ys;
  m.m[2][2] = zf / (zf - zn);
  m.m[2][3] = 1.0f;
  m.m[3][2] = -zn * zf / (zf - zn);
  return m;
}
// Gauss-Jordan inverse with partial pivoting, in double; a singular matrix
// gives all-NaN.
float4x4 inverse4(const float4x4& m) {
  double a[4][8];
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) {
      a[i][j] = m.m[i][j];
      a[i][j + 4] = i == j ? 1.0 : 0.0;
    }
  for (int c = 0; c < 4; ++c) {
    int p = c;
    for (int r = c + 1; r < 4; ++r)
      if (std::fabs(a[r][c]) > std::fabs(a[p][c])) p = r;
    if (a[p][c] == 0.0) {
      float4x4 bad;
      for (auto& row : bad.m)
        for (float& v : row) v = NAN;
      return bad;
    }
    if (p != c)
      for (int j = 0; j < 8; ++j) std::swap(a[p][j], a[c][j]);
    const double inv = 1.0 / a[c][c];
    for (int j = 0; j < 8; ++j) a[c][j] *= inv;
    for (int r = 0; r < 4; ++r) {
      if (r == c) continue;
      const double f = a[r][c];
      if (f != 0.0)
        for (int j = 0; j < 8; ++j) a[r][j] -= f * a[c][j];
    }
  }
  float4x4 out;
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) out.m[i][j] = (float)a[i][j + 4];
  return out;
}
}  // namespace

// CUDACamera.cpp:244-263: the view looks from the centre along dir (at a
// point 4999.999 away), with the up vector (0, 1, 0) turned by the roll about
// the axis (dir.x, 0, dir.y), as the reference writes it.
void CUDACamera::update() {
  const float3 target = center + dir * 4999.999f;
  const float3 axis = make_float3(dir.x, 0.0f, dir.y);
  const float4x4 up_rot = rotation_axis(axis, angle.z);
  const float3 up = transform_point(up_rot, make_float3(0.0f, 1.0f, 0.0f));
  matView = look_at_lh(center, target, up);
  matInvView = inverse4(matView);
  matProj = perspective_fov_lh(fov, ratio, zNearPlane, zFarPlane);
  matInvProj = inverse4(matProj);
  dirty = false;
}
float4x4 CUDACamera::getViewMatrix() { if (dirty) update(); return matView; }
float4x4 CUDACamera::getProjMatrix() { if (dirty) update(); return matProj; }
float4x4 CUDACamera::getInvViewMatrix() { if (dirty) update(); return matInvView; }
float4x4 CUDACamera::getInvProjMatrix() { if (dirty) update(); return matInvProj; }

void CUDACamera::primaryRayConstants(int width, int height, rk_float4 out[4]) const {
  // CUDA_calcPrimaryRays.cu:43-55
  float y = tanf(fov * 0.5f);
  float x = y * ratio;
  float3 camDir = make_float3(cosf(pitch()) * cosf(yaw()), sinf(yaw()), sinf(pitch()) * cosf(yaw()));
  float3 camUp = make_float3(cosf(pitch()) * sinf(roll()), cosf(roll()), 0.0f);
  float3 right = normalize(cross(camDir, camUp));
  float3 up = normalize(cross(right, camDir));
  float3 rd = (up * y + right * x) + camDir;
  float iw = 1.0f / (float)width, ih = 1.0f / (float)height;
  float3 rx = (right * (2.0f * x)) * iw;
  float3 ry = (up * (2.0f * y)) * ih;
  out[0] = {center.x, center.y, center.z, 0.0f};
  out[1] = {rd.x, rd.y, rd.z, 0.0f};
  out[2] = {rx.x, rx.y, rx.z, 0.0f};
  out[3] = {ry.x, ry.y, ry.z, 0.0f};
}

// ---- CUDARenderer / CUDAEngine ----------------------------------------------------

CUDARenderer::CUDARenderer(CUDAEngine* _engine) : engine(_engine) { if (engine) engine->renderer = this; }
CUDARenderer::~CUDARenderer() {}

void CUDARenderer::saveSurface(const char* path, const int format, CUDASurface<byte>* s) {
  _assert(format == CUDA_FORMAT_RAW, __FILE__, __LINE__, "CUDARenderer::save() : Unsupported file format.");
  FILE* f = std::fopen(path, "wb");
  _assert(f != nullptr, __FILE__, __LINE__, "CUDASurface::save() : Invalid file.");
  int header = 0x194B7EA2;
  int v[6] = {s->getUnitCount(), s->getUnitSize(), s->getSize(), s->getWidth(), s->getHeight(), s->getPitch()};
  std::fwrite(&header, sizeof(int), 1, f);
  std::fwrite(v, sizeof(int), 6, f);
  std::vector<byte> host(s->getSize());
  s->copyDataToHost(host.data());
  std::fwrite(host.data(), 1, host.size(), f);
  std::fclose(f);
}

CUDASurface<byte>* CUDARenderer::loadSurface(const char* path) {
  FILE* f = std::fopen(path, "rb");
  _assert(f != nullptr, __FILE__, __LINE__, "CUDASurface::load() : Invalid file.");
  int header = 0, v[6] = {0};
  bool ok = std::fread(&header, sizeof(int), 1, f) == 1 && header == 0x194B7EA2 && std::fread(v, sizeof(int), 6, f) == 6;
  if (!ok) {
    std::fclose(f);
    _assert(false, __FILE__, __LINE__, "CUDARenderer::load() : Unsupported file format.");
    return nullptr;
  }
  auto* s = new CUDASurface<byte>();
  s->assign(v[3], v[4], v[1], v[5]);
  std::vector<byte> host(v[2]);
  _assert(std::fread(host.data(), 1, host.size(), f) == host.size(), __FILE__, __LINE__, "CUDARenderer::load() : Truncated file.");
  s->copyDataFromHost(host.data());
  std::fclose(f);
  return s;
}

CUDAHeadlessRenderer::CUDAHeadlessRenderer(CUDAEngine* e, int w, int h) : CUDARenderer(e), width(w), height(h) {
  surface = new CUDARenderSurface<uint>(w, h);
}
CUDAHeadlessRenderer::~CUDAHeadlessRenderer() { finalize(); delete surface; surface = nullptr; }
void CUDAHeadlessRenderer::initialize() {
  if (!storage) storage = new CUDAArray<uint>(width * height);
  surface->setPointer(storage->getPointer());
}
void CUDAHeadlessRenderer::finalize() {
  if (surface) surface->setPointer(nullptr);
  delete storage;
  storage = nullptr;
}

CUDAEngine::CUDAEngine() {  // CUDAEngine.cpp:30-54
  const double t = getTime();
  frame_time = cycle_time = total_time = t;
}

double CUDAEngine::getTime() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

void CUDAEngine::update() {  // CUDAEngine.cpp:113-132
  ++cycle_count;
  ++frames;
  const double t = getTime();
  frame_delta = t - frame_time;
  frame_time = t;
  if (frame_time - cycle_time >= 1.0) {
    total_count += cycle_count;
    average_framerate = (double)total_count / (frame_time - total_time);
    current_framerate = (double)cycle_count / (frame_time - cycle_time);
    cycle_time = frame_time;
    cycle_count = 0;
  }
}

// ---- CUDAWindow ----------------------------------------------------------------------

CUDAWindow::CUDAWindow(CUDAEngine* _engine) : engine(_engine) { if (engine) engine->window = this; }
CUDAWindow::~CUDAWindow() {
  if (engine && engine->window == this) engine->window = nullptr;
}
void CUDAWindow::close() {
  if (closed) return;
  closed = true;
  if (on_close) on_close();
}
bool CUDAHeadlessWindow::isActive() {
  if (!closed && engine && engine->getFrameCount() >= frame_limit) close();
  return !closed;
}

// ---- CUDAShader ------------------------------------------------------------------------

CUDAShader::CUDAShader() { backColor = make_float4(0.227451f, 0.431373f, 0.647059f, 1.0f); }

void CUDAShader::run(CUDARenderSurface<uint>* output, CUDABuffer* input, CUDACamera* camera) {
  _assert(output && input && camera, __FILE__, __LINE__, "CUDAShader::run() : Invalid parameter.");
  _assert(input->getOutputNormalSurface() != nullptr, __FILE__, __LINE__,
          "CUDAShader::run() : The buffer has no normal surface (interpolateOutputSurfaces with normals).");
  _assert(output->getWidth() == input->getWidth() && output->getHeight() == input->getHeight(), __FILE__, __LINE__,
          "CUDAShader::run() : Output and buffer sizes differ.");
  const float3 d = camera->getDir();
  const rk_float4 back = {backColor.x, backColor.y, backColor.z, backColor.w};
  RK_CALL(rk_shade(nullptr, input->getOutputDepthSurface()->getPointer(),
                   reinterpret_cast<const rk_float4*>(input->getOutputNormalSurface()->getPointer()),
                   input->getWidth(), input->getHeight(), back, rk_float4{d.x, d.y, d.z, 0.0f},
                   output->getPointer(), output->getPitch() / (int)sizeof(uint)));
  RK_CALL(rk_stream_synchronize(nullptr));
}

// ---- CUDARaytracer ------------------------------------------------------------------

CUDARaytracer::CUDARaytracer(CUDAEngine* _engine) : engine(_engine) {
  if (engine) engine->raytracer = this;
  internal_camera = new CUDACamera();
  current_camera = internal_camera;
}

CUDARaytracer::~CUDARaytracer() {
  finalize();
  delete internal_camera;
}

void CUDARaytracer::updateRenderSurface() {  // CUDARaytracer.cpp:110-121
  CUDARenderer* renderer = engine->getRenderer();
  renderer->lockSurface();
  current_shader->run(renderer->getSurface(), current_buffer, current_camera);
  renderer->unlockSurface();
}

void CUDARaytracer::saveRenderSurface(const char* path) {  // CUDARaytracer.cpp:94-108
  (void)path;
  _assert(false, __FILE__, __LINE__,
          "CUDARaytracer::saveRenderSurface() : Function not available, invalid renderer type, CUDAD3DRenderer "
          "needed.");
}

void CUDARaytracer::rasterizeMeshTree(CUDAMesh* mesh, CUDAMeshTree* tree) {
  if (mesh->getVertexCount() == 0) return;  // CUDARaytracer.cpp:139
  raytraceMeshTree(mesh, tree, true);
}

void CUDARaytracer::initialize() {  // CUDARaytracer.cpp:53-83
  int count = 0;
  rk_device_count(&count);
  _assert(count != 0, __FILE__, __LINE__, "Unable to find any GPU.\n");
  _assert(engine && engine->getRenderer(), __FILE__, __LINE__, "CUDARaytracer::initialize() : No renderer.");
  CUDARenderer* renderer = engine->getRenderer();
  int width = renderer->getSurface()->getWidth();
  int height = renderer->getSurface()->getHeight();
  delete internal_buffer;
  internal_buffer = new CUDABuffer(width, height);
  current_buffer = internal_buffer;
  if (!internal_shader) internal_shader = new CUDAShader();
  current_shader = internal_shader;
}

void CUDARaytracer::finalize() {
  if (current_buffer == internal_buffer) current_buffer = nullptr;
  delete internal_buffer;
  internal_buffer = nullptr;
  if (current_shader == internal_shader) current_shader = nullptr;
  delete internal_shader;
  internal_shader = nullptr;
}

void CUDARaytracer::calcPrimaryRays() {
  CUDABuffer* b = current_buffer;
  rk_float4 c[4];
  current_camera->primaryRayConstants(b->getWidth(), b->getHeight(), c);
  RK_CALL(rk_calc_primary_rays(nullptr, b->getWidth(), b->getHeight(), c[0], c[1], c[2], c[3],
                               reinterpret_cast<rk_float4*>(b->getInputOriginSurface()->getPointer()),
                               reinterpret_cast<rk_float4*>(b->getInputDirectionSurface()->getPointer())));
}

void CUDARaytracer::raytraceMeshTree(CUDAMesh* mesh, CUDAMeshTree* tree, bool coherency) {
  raytraceMeshTreeAsync(mesh, tree, coherency, nullptr);
  RK_CALL(rk_stream_synchronize(nullptr));
}

void CUDARaytracer::raytraceMeshTreeAsync(CUDAMesh* mesh, CUDAMeshTree* tree, bool coherency, rk_stream stream) {
  if (mesh->getVertexCount() == 0) return;  // CUDARaytracer.cpp:132
  CUDABuffer* b = current_buffer;
  int w = b->getWidth(), h = b->getHeight();
  const rk_float4* o = reinterpret_cast<const rk_float4*>(b->getInputOriginSurface()->getPointer());
  const rk_float4* d = reinterpret_cast<const rk_float4*>(b->getInputDirectionSurface()->getPointer());
  uint* fid = b->getOutputFaceIdSurface()->getPointer();
  rk_float2* uv = reinterpret_cast<rk_float2*>(b->getOutputCoordSurface()->getPointer());
  float* dep = b->getOutputDepthSurface()->getPointer();
  if (tree->isFast()) {
    rk_fast_bvh desc = {tree->getFastNodes(), tree->getFastTriangles(), tree->getIndicesArray()->getPointer(),
                        reinterpret_cast<rk_aabox*>(tree->getBoxesArray()->getPointer()),
                        (uint32_t)mesh->getFaceCount(), 0};
    rk_trace_options opt = {det_epsilon, 0u, 0u, 0u};
    RK_CALL(rk_fast_trace(stream, &desc, o, d, (uint32_t)w * (uint32_t)h, fid, uv, dep, &opt));
    return;
  }
  _assert(tree->getNodeCount() > 0, __FILE__, __LINE__, "CUDARaytracer::raytraceMeshTree() : Tree not built.");
  RK_CALL(rk_raytrace_mesh_tree(stream, reinterpret_cast<const rk_bvnode*>(tree->getNodesArray()->getPointer()),
                                tree->getNodeCount(), reinterpret_cast<const rk_uint4*>(mesh->getFacesArray()->getPointer()),
                                (uint32_t)mesh->getFaceCount(),
                                reinterpret_cast<const rk_float4*>(mesh->getVerticesArray()->getPointer()), o, d, w, h,
                                fid, uv, dep, coherency ? 1 : 0));
}

void CUDARaytracer::interpolateOutputSurfaces(CUDAMesh* mesh, bool points, bool normals, bool colors,
                                              bool texcoords, bool materials) {
  CUDABuffer* b = current_buffer;
  const int w = b->getWidth(), h = b->getHeight();
  if (points && !b->out_points) b->out_points = new CUDASurface<float4>(w, h);
  if (normals && !b->out_normals) b->out_normals = new CUDASurface<float4>(w, h);
  if (colors && !b->out_colors) b->out_colors = new CUDASurface<float4>(w, h);
  if (texcoords && !b->out_texcoords) b->out_texcoords = new CUDASurface<float2>(w, h);
  if (materials && !b->out_materials) b->out_materials = new CUDASurface<uint>(w, h);
  auto p4 = [](CUDASurface<float4>* s) { return s ? reinterpret_cast<rk_float4*>(s->getPointer()) : nullptr; };
  auto a4 = [](CUDAArray<float4>* a) { return a ? reinterpret_cast<const rk_float4*>(a->getPointer()) : nullptr; };
  RK_CALL(rk_interpolate_data(
      nullptr, b->getOutputFaceIdSurface()->getPointer(),
      reinterpret_cast<const rk_float2*>(b->getOutputCoordSurface()->getPointer()), w, h,
      reinterpret_cast<const rk_uint4*>(mesh->getFacesArray()->getPointer()), (uint32_t)mesh->getFaceCount(),
      a4(mesh->getVerticesArray()), a4(mesh->getNormalsArray()), a4(mesh->getColorsArray()),
      mesh->getTexcoordsArray() ? reinterpret_cast<const rk_float2*>(mesh->getTexcoordsArray()->getPointer()) : nullptr,
      p4(b->out_points), p4(b->out_normals), p4(b->out_colors),
      b->out_texcoords ? reinterpret_cast<rk_float2*>(b->out_texcoords->getPointer()) : nullptr,
      b->out_materials ? b->out_materials->getPointer() : nullptr));
  RK_CALL(rk_stream_synchronize(nullptr));  // the reference returns after the kernel (CUDA_interpolateData.cu:250)
}

}  // namespace renderkit
