// .cdm files through the drop-in classes, as samples 05 and 06 use them: one file
// holding a mesh's chunks and then its tree's (samples/05_how_to_quick_save_a_mesh.cpp:129-134,
// samples/06_how_to_quick_load_a_mesh.cpp:175-181).
//
//   cdm_format <in.cdm> <out.cdm> <dump.bin>
// loads the mesh and then the tree from in.cdm, writes every loaded array to
// dump.bin ([uint32 bytes][bytes] per array: hitbox, vertices, normals, colors,
// texcoords, faces, nodes, indices, boxes), then saves the mesh (overwrite) and
// the tree (append) to out.cdm.
#include <cstdio>
#include <vector>

#include "../../include/renderkit/CUDAMeshTree.h"

using namespace renderkit;

template <typename T>
static void dump(FILE* f, CUDAArray<T>* a) {
  uint32_t bytes = 0;
  if (!a) { std::fwrite(&bytes, 4, 1, f); return; }
  std::vector<T> h(a->getUnitCount());
  a->copyDataToHost(h.data());
  bytes = (uint32_t)(h.size() * sizeof(T));
  std::fwrite(&bytes, 4, 1, f);
  std::fwrite(h.data(), 1, bytes, f);
}

int main(int argc, char** argv) {
  if (argc < 4) { std::fprintf(stderr, "usage: %s in.cdm out.cdm dump.bin\n", argv[0]); return 2; }
  setThrowOnError(true);
  CUDAMesh mesh;
  mesh.quickLoad(argv[1]);
  CUDAMeshTree tree(&mesh);
  tree.quickLoad(argv[1]);
  FILE* f = std::fopen(argv[3], "wb");
  if (!f) return 1;
  CUDABox hb = mesh.getHitbox();
  uint32_t bytes = sizeof(hb);
  std::fwrite(&bytes, 4, 1, f);
  std::fwrite(&hb, 1, bytes, f);
  dump(f, mesh.getVerticesArray());
  dump(f, mesh.getNormalsArray());
  dump(f, mesh.getColorsArray());
  dump(f, mesh.getTexcoordsArray());
  dump(f, mesh.getFacesArray());
  dump(f, tree.getNodesArray());
  dump(f, tree.getIndicesArray());
  dump(f, tree.getBoxesArray());
  std::fclose(f);
  mesh.quickSave(argv[2], true);
  tree.quickSave(argv[2], false);
  return 0;
}
