// renderkit drop-in: CUDAMeshTree (CUDAMeshTree.h:22-164). Owns the nodes
// (made by a build), per-face boxes and the sorted face indices. The builds run
// on the GPU through include/rk.h:
//   buildMeshTreeLBVH(d)   rk_build_lbvh(tree_complete = true)  (CUDAMeshTree.cpp:88-91)
//   buildMeshTreeSAH()     rk_build_sah                          (:93-96)
//   buildMeshTreeHybrid(d) rk_build_lbvh(false) + rk_build_sah  (:98-103)
//   buildMeshTreeFast()    the throughput tree (rk_fast_build); no reference counterpart
#ifndef RENDERKIT_CUDAMESHTREE_H
#define RENDERKIT_CUDAMESHTREE_H

#include "CUDAMesh.h"

namespace renderkit {

class CUDAMeshTree {
 public:
  CUDAArray<bvnode>* nodes = nullptr;
  CUDAArray<aabox>* boxes = nullptr;
  CUDAArray<uint>* indices = nullptr;
  CUDAMesh* mesh = nullptr;

  explicit CUDAMeshTree(CUDAMesh* mesh);
  ~CUDAMeshTree();

  CUDAMesh* getMesh() const { return mesh; }
  CUDABox getHitbox() const { return mesh->getHitbox(); }
  CUDAArray<aabox>* getBoxesArray() const { return boxes; }
  CUDAArray<uint>* getIndicesArray() const { return indices; }
  CUDAArray<bvnode>* getNodesArray() const { return nodes; }
  uint getNodeCount() const { return nodes ? nodes->getUnitCount() : 0; }

  void buildMeshTreeLBVH(uint tree_depth = 12);
  void buildMeshTreeSAH();
  void buildMeshTreeHybrid(uint lbvh_tree_depth = 12);
  // The throughput tree: Karras LBVH with 30-bit (or 63-bit) Morton keys and
  // leaves of <= leaf_max faces. Faces are reordered in leaf order, like the
  // reference builds; isFast() trees are traced by rk_fast_trace.
  void buildMeshTreeFast(uint leaf_max = 1, bool morton63 = false);
  bool isFast() const { return fast_nodes != nullptr; }
  void* getFastNodes() const { return fast_nodes ? fast_nodes->getPointer() : nullptr; }
  void* getFastTriangles() const { return fast_tris ? fast_tris->getPointer() : nullptr; }

  // Diagnostics of the last build (RK_INFO_* words).
  uint getRealDepth() const { return info_host[RK_INFO_REAL_DEPTH]; }
  uint getBuildFlags() const { return info_host[RK_INFO_FLAGS]; }

  void updateNodes(bvnode* nodes_buffer, uint node_count);  // CUDAMeshTree.cpp:105-112
  // Chunked .cdm file, CUDAMeshTree.cpp:114-289 (NODES, INDICES, BOXES chunks).
  // Fixed against the reference: count checks test equality (the reference's
  // asserts at :166,182 are inverted) and append mode writes correct lengths.
  void quickLoad(const char* path);
  void quickSave(const char* path, bool overwrite = false);

 private:
  void releaseFast();
  CUDAArray<float4>* fast_nodes = nullptr;  // 64-B records, 4 float4 each
  CUDAArray<float4>* fast_tris = nullptr;   // 48-B records, 3 float4 each
  uint info_host[RK_INFO_WORDS] = {0};
};

}  // namespace renderkit

#endif
