/*
 * rk.h — C ABI of the MI355X (gfx950) mesh-tree build and closest-hit trace
 * library. This is the drop-in boundary: the C++ host classes in
 * include/renderkit/ (CUDAMeshTree, CUDAMesh, CUDABuffer, CUDARaytracer, ...)
 * call only these functions. The reference's own boundary is a set of C++
 * free functions declared as prototypes by their callers
 * (CUDAMeshTree.cpp:26-38, CUDARaytracer.h:24-36, CUDA_buildMeshTreeLBVH.cu:23-29);
 * each declaration below names the one it replaces.
 *
 * Conventions
 *  - Plain pointers and sizes only. Pointers marked "device" address GPU
 *    memory (from rk_malloc or any HIP/torch allocation on the current device).
 *  - Every call is stream-ordered on 'stream' (NULL = the default stream) and
 *    returns without waiting for the GPU, except where noted.
 *  - Temporaries come from a caller-provided workspace ('ws', 'ws_bytes'), sized
 *    by the matching *_workspace_size function. The library keeps no global
 *    device state (one exception: rk_fast_build_noalloc's own array), so calls
 *    on different streams with different workspaces are independent and may run
 *    concurrently.
 *  - Errors are returned as rk_status; nothing exits the process.
 */
#ifndef RK_H
#define RK_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t rk_status;
#define RK_SUCCESS 0
#define RK_ERROR_INVALID_VALUE 1
#define RK_ERROR_OUT_OF_MEMORY 2
#define RK_ERROR_LAUNCH 3
#define RK_ERROR_WORKSPACE 4
#define RK_ERROR_DEVICE 5
#define RK_ERROR_UNSUPPORTED 6

/* POD types with the reference's layouts (16-byte aligned where the reference
 * uses __builtin_align__(16)). */
typedef struct rk_float2 { float x, y; } rk_float2;
typedef struct rk_float4 { float x, y, z, w; } rk_float4;
typedef struct rk_uint4 { uint32_t x, y, z, w; } rk_uint4;
/* aabox, CUDABox.h:17-148: float4 min, float4 max (32 B). */
typedef struct rk_aabox { rk_float4 min, max; } rk_aabox;
/* bvnode, CUDANode.h:20-48 (32 B). word0 bit 0 = type (0 node, 1 leaf);
 * node: bits 1-2 = split axis, word1 = child_id (children child_id, child_id+1);
 * leaf: bits 1-31 = obj_count, word1 = obj_id. Then float3 box_min, box_max. */
typedef struct rk_bvnode { uint32_t word0, word1; float box_min[3], box_max[3]; } rk_bvnode;

/* An opaque stream handle (a hipStream_t). NULL is the default stream. */
typedef struct rk_stream_st* rk_stream;

/* Build diagnostics written to the device array 'info' by the builders. */
#define RK_INFO_NODE_COUNT 0 /* number of nodes M */
#define RK_INFO_REAL_DEPTH 1 /* max leaf level + 1 (LBVH.cu:452) */
#define RK_INFO_FLAGS 2      /* RK_FLAG_* bits */
#define RK_INFO_SPLITS 3     /* number of Morton splits S (LBVH.cu:569) */
#define RK_INFO_LEAVES 4     /* number of leaves */
#define RK_INFO_WORDS 8
#define RK_FLAG_WALK_PAST_END 1   /* split walk ran past the last split (reference reads past its array) */
#define RK_FLAG_WALK_PAST_DEPTH 2 /* split walk reached levels > tree_depth (reference order there is nondeterministic) */
#define RK_FLAG_MERGE2X_OOB 4     /* reference bitonic Merge2x would write out of bounds (BSG.cu:254, BS.cu:196) */
#define RK_FLAG_NODES_GT_N 8      /* M > n: the reference's N-node buffers would overflow (LBVH.cu:466-470) */
#define RK_FLAG_SAH_NO_CUT 16     /* SAH node with >= 224 objects and no finite-cost cut (reference loops forever) */
#define RK_FLAG_SPIN_LIMIT 64     /* a look-back wait hit its iteration limit; results invalid */

/* ---- runtime ------------------------------------------------------------- */
const char* rk_status_string(rk_status s);
/* Returns the number of visible GPUs in *count (0 if none). */
rk_status rk_device_count(int* count);
/* CUDAArray.h:144-148 cudaMalloc / :150-154 cudaFree. */
rk_status rk_malloc(void** ptr, size_t bytes);
rk_status rk_free(void* ptr);
/* Pinned host memory (cudaMallocHost in the reference's quickLoad/quickSave). */
rk_status rk_malloc_host(void** ptr, size_t bytes);
rk_status rk_free_host(void* ptr);
/* kind: 0 host->device, 1 device->host, 2 device->device. Synchronous when
 * stream is NULL (CUDAArray.h:169-208 copyData*), else stream-ordered. */
rk_status rk_memcpy(void* dst, const void* src, size_t bytes, int kind, rk_stream stream);
rk_status rk_stream_create(rk_stream* stream);
rk_status rk_stream_destroy(rk_stream stream);
rk_status rk_stream_synchronize(rk_stream stream);

/* ---- surfaces (CUDA_fillSurface.cu:17-61) ---------------------------------- */
/* template<T> CUDA_fillSurface(CUDASurface<T>*, T), CUDABuffer.cpp:27-32. */
rk_status rk_fill_u32(rk_stream stream, uint32_t* dst, size_t count, uint32_t value);
rk_status rk_fill_f32(rk_stream stream, float* dst, size_t count, float value);
rk_status rk_fill_f2(rk_stream stream, rk_float2* dst, size_t count, rk_float2 value);
rk_status rk_fill_f4(rk_stream stream, rk_float4* dst, size_t count, rk_float4 value);

/* ---- primary rays (CUDA_calcPrimaryRays.cu:20-32, kernel part) ------------- */
/* origins[y*w+x] = ray_orig; directions[y*w+x] = (ray_dir - x*ray_x) - y*ray_y. */
rk_status rk_calc_primary_rays(rk_stream stream, int width, int height, rk_float4 ray_orig, rk_float4 ray_dir,
                               rk_float4 ray_x, rk_float4 ray_y, rk_float4* origins, rk_float4* directions);

/* ---- attribute interpolation --------------------------------------------------- */
/* void CUDA_interpolateData(CUDABuffer*, CUDAMesh*), CUDARaytracer.h:46 /
 * kernels/CUDA_interpolateData.cu:217. Per pixel of a traced frame, the hit
 * face's position, normal, colour, texture coordinates and material at its
 * barycentrics: a = (1 - u) - v, value = (a*x0 + u*x1) + v*x2 in the reference's
 * order (kernels/CUDA_interpolateData.cu:45-141). normals/colors/texcoords may be
 * NULL (mesh without them): the face normal normalize(cross(v1-v0, v2-v0)),
 * (1,1,1,1) and (u, v) are written instead. Any out_* may be NULL (not wanted).
 * normalize is v * (1/sqrtf(dot(v, v))), where the reference's device code uses
 * the approximate rsqrtf. A pixel whose face id is >= face_count (a miss that
 * kept an out-of-range id) is left unchanged; the reference reads past its faces. */
rk_status rk_interpolate_data(rk_stream stream, const uint32_t* faces_id, const rk_float2* coords, int width,
                              int height, const rk_uint4* faces, uint32_t face_count, const rk_float4* vertices,
                              const rk_float4* normals, const rk_float4* colors, const rk_float2* texcoords,
                              rk_float4* out_points, rk_float4* out_normals, rk_float4* out_colors,
                              rk_float2* out_texcoords, uint32_t* out_materials);

/* ---- shading ------------------------------------------------------------------ */
/* void runCUDAShader(CUDARenderSurface<uint>*, CUDABuffer*, CUDACamera*, CUDAShader*),
 * CUDAShader.cpp:8-15 / CUDAShaderKernel.cu:14-76. Per pixel k = y*width + x of a
 * traced and interpolated frame: when depths[k] < FLT_MAX, c = -dot(cam_dir,
 * normals[k]) with both w taken as 0, and the pixel is make_pixel(c, c, c, c);
 * otherwise make_pixel(back_color). make_pixel (kernels/CUDAShaderCommon.cuh)
 * scales by 255, clamps to [0, 255] with fminf/fmaxf, truncates and packs
 * A << 24 | R << 16 | G << 8 | B. Pixel (x, y) goes to pixels[y*pitch + x]
 * (pitch in pixels, >= width). cam_dir is the camera's direction. */
rk_status rk_shade(rk_stream stream, const float* depths, const rk_float4* normals, int width, int height,
                   rk_float4 back_color, rk_float4 cam_dir, uint32_t* pixels, int pitch);

/* ---- build pieces ----------------------------------------------------------- */
/* void CUDA_calcMeshBoxes(aabox*, uint*, uint4*, float4*, uint), CUDA_calcMeshBoxes.cu:65. */
rk_status rk_calc_mesh_boxes(rk_stream stream, const rk_uint4* faces, const rk_float4* vertices, uint32_t count,
                             rk_aabox* boxes, uint32_t* indices);
/* void CUDA_reorderMeshFaces(uint4*, uint*, uint), CUDA_reorderMeshFaces.cu:44. */
size_t rk_reorder_workspace_size(uint32_t count);
rk_status rk_reorder_mesh_faces(rk_stream stream, rk_uint4* faces, const uint32_t* indices, uint32_t count, void* ws,
                                size_t ws_bytes);

/* ---- sorts ------------------------------------------------------------------ */
/* Workspace for any sort below of 'length' elements with keys of key_bytes (4 or 8).
 * Every sort here accepts length < 2^30 (RK_ERROR_UNSUPPORTED beyond). The first 4
 * bytes of the workspace receive the call's RK_FLAG_* bits (RK_FLAG_SPIN_LIMIT: a
 * look-back wait of the radix pass hit its iteration limit and the output is not
 * sorted); each call clears them first. Read them with rk_workspace_flags. */
size_t rk_sort_workspace_size(uint32_t length, uint32_t key_bytes);
/* Copies the RK_FLAG_* word at the start of 'ws' to *flags after waiting for
 * 'stream' (no reference counterpart: the reference's sorts cannot fail silently
 * because they cannot fail at all). */
rk_status rk_workspace_flags(rk_stream stream, const void* ws, uint32_t* flags);
/* void CUDA_radixSortGeneral(uint* keys, uint* indices, uint length, uint bit_depth, bool init_indices),
 * CUDA_buildMeshTreeLBVH.cu:29 / CUDA_radixSortGeneral.cu:397. Same tie order by size class:
 * length <= 512 stable; <= byte_depth*8192 bitonic order within 128-chunks then stable by key;
 * else stable by the low 8*byte_depth bits. */
rk_status rk_radix_sort_general(rk_stream stream, uint32_t* keys, uint32_t* indices, uint32_t length,
                                uint32_t bit_depth, int init_indices, void* ws, size_t ws_bytes);
/* void CUDA_radixSort(uint* values, uint length, uint bit_depth), CUDA_buildMeshTreeLBVH.cu:27 / CUDA_radixSort.cu:357. */
rk_status rk_radix_sort(rk_stream stream, uint32_t* values, uint32_t length, uint32_t bit_depth, void* ws,
                        size_t ws_bytes);
/* Stable (u64 key, u32 value) sort on the low 'bit_depth' bits (63-bit Morton keys; no reference counterpart). */
rk_status rk_radix_sort_pairs_u64(rk_stream stream, uint64_t* keys, uint32_t* values, uint32_t length,
                                  uint32_t bit_depth, void* ws, size_t ws_bytes);
/* template <typename T> void CUDA_bitonicSort(T* values, uint length),
 * samples/09_how_to_use_a_sorter.cpp:39-40 / kernels/CUDA_bitonicSort.cu:402-518, instantiated
 * for unsigned int only (:523-537). Sorts values ascending in place, any length. */
rk_status rk_bitonic_sort(rk_stream stream, uint32_t* values, uint32_t length, void* ws, size_t ws_bytes);
/* template <typename K, typename V> void CUDA_bitonicSortGeneral(K* keys, V* values, uint length),
 * samples/09_how_to_use_a_sorter.cpp:42-43 / kernels/CUDA_bitonicSortGeneral.cu:504-690,
 * instantiated for <uint, uint> only (:694-697). Sorts (key, value) pairs by key, any length,
 * in the reference's tie order: a bitonic network inside each full 128-element chunk (on keys
 * only), the partial last chunk stably, then stable merges. Where the reference's Merge2x
 * writes past 'length' (BSG.cu:254) this writes nothing outside the arrays. */
rk_status rk_bitonic_sort_general(rk_stream stream, uint32_t* keys, uint32_t* values, uint32_t length, void* ws,
                                  size_t ws_bytes);
/* void CUDA_bubbleSort512(uint* keys, uint* indices, uint length, bool init_indices = false),
 * kernels/CUDA_radixSortGeneral.cu:27 / kernels/CUDA_bubbleSort512.cu:114-167. A stable sort of
 * at most 512 (key, index) pairs in one workgroup; init_indices first sets indices[i] = i. Needs
 * no workspace. length > 512: RK_ERROR_INVALID_VALUE (the reference asserts, :116). */
rk_status rk_bubble_sort512(rk_stream stream, uint32_t* keys, uint32_t* indices, uint32_t length, int init_indices);

/* ---- Morton codes (kernel_calcMortonCodes, CUDA_buildMeshTreeLBVH.cu:115-173) ---- */
#define RK_MORTON_REF32 0 /* the reference 11/11/10-bit code */
#define RK_MORTON_30 1    /* 10/10/10 bits */
#define RK_MORTON_63 2    /* 21/21/21 bits (written as uint64) */
rk_status rk_calc_morton_codes(rk_stream stream, const rk_aabox* boxes, uint32_t count, const rk_aabox* root_box,
                               int mode, void* codes);

/* ---- reference LBVH build --------------------------------------------------- */
/* void CUDA_buildMeshTreeLBVH(CUDAMesh*, CUDAMeshTree*, bool tree_complete, uint tree_depth),
 * CUDAMeshTree.cpp:26-32 / CUDA_buildMeshTreeLBVH.cu:457. Inputs are the mesh arrays and its
 * hitbox (host value, the Morton root box). Writes boxes[n], indices[n], nodes[M] (M <= 2n-1,
 * nodes_capacity >= 2n-1), info[RK_INFO_WORDS] (device) and, when tree_complete, reorders
 * faces in place. tree_depth <= 31. */
size_t rk_build_lbvh_workspace_size(uint32_t face_count);
rk_status rk_build_lbvh(rk_stream stream, rk_uint4* faces, const rk_float4* vertices, uint32_t face_count,
                        const rk_aabox* hitbox, int tree_complete, uint32_t tree_depth, rk_aabox* boxes,
                        uint32_t* indices, rk_bvnode* nodes, uint32_t nodes_capacity, uint32_t* info, void* ws,
                        size_t ws_bytes);

/* ---- reference binned SAH build ------------------------------------------------ */
/* void CUDA_buildMeshTreeSAH(CUDAMesh*, CUDAMeshTree*), CUDAMeshTree.cpp:34-38 /
 * CUDA_buildMeshTreeSAH.cu:672. 32-bin SAH with the reference's costs, leaf rules and
 * stable partition. existing_nodes == 0: boxes/indices are computed and the root box is
 * the hitbox; existing_nodes > 0 (Hybrid, CUDAMeshTree.cpp:98-103): the first
 * existing_nodes entries of 'nodes' are an LBVH built with tree_complete = 0 (boxes and
 * indices from that build) and each of its leaves is refined. Node ids are breadth-first
 * (the reference's are in atomicAdd order). nodes_capacity >= existing_nodes + 2n - 1.
 * Faces are reordered in place. Synchronises 'stream' once per tree level, except on the first
 * levels (while a level can hold at most min(4096, n/256) nodes); the info words and the face
 * reorder are still queued on 'stream' when it returns. */
size_t rk_build_sah_workspace_size(uint32_t face_count, uint32_t existing_nodes);
rk_status rk_build_sah(rk_stream stream, rk_uint4* faces, const rk_float4* vertices, uint32_t face_count,
                       const rk_aabox* hitbox, rk_aabox* boxes, uint32_t* indices, rk_bvnode* nodes,
                       uint32_t existing_nodes, uint32_t nodes_capacity, uint32_t* info, void* ws, size_t ws_bytes);

/* ---- reference trace -------------------------------------------------------- */
/* void CUDA_raytraceMeshTree(CUDABuffer*, CUDAMesh*, CUDAMeshTree*, bool coherency),
 * CUDARaytracer.h:30-36 / CUDA_raytraceMeshTree.cu:434. faces_id/coords/depths are the
 * buffer's output surfaces; depths is read as tmax. On a miss the three outputs are left
 * unchanged. coherency != 0 selects the 16x12 packet traversal (the reference default). */
rk_status rk_raytrace_mesh_tree(rk_stream stream, const rk_bvnode* nodes, uint32_t node_count,
                                const rk_uint4* faces, uint32_t face_count, const rk_float4* vertices,
                                const rk_float4* origins,
                                const rk_float4* directions, int width, int height, uint32_t* faces_id,
                                rk_float2* coords, float* depths, int coherency);

/* Second passes of rk_raytrace_mesh_tree (no reference counterpart). A ray of the per-ray
 * traversal whose leaves pass ray_faces faces is traced again from the start by a wave of
 * its own; a packet whose leaves pass packet_faces faces is resumed where it stands with
 * five threads per ray. A ray with a direction component of 0 (every box test passes for
 * it, so it walks the whole tree), or a packet holding one, goes at its first leaf, to
 * walks split by subtree or by parts of the packet on several CUs. The
 * results are the same. What is handed off goes to a per-call list of list_cap entries
 * (1 .. 65536, device memory from the stream's pool); past its end, rays and packets finish
 * in the first pass. UINT32_MAX turns a second pass off. RK_HANDOFF_AUTO (the default for
 * both limits) takes 256 faces per ray and 512 per packet when the call has fewer than 2^19
 * rays (its first pass then holds too few waves to fill the GPU), else 4096 for both. The
 * default list size is 65536. Process-wide, read by each call. */
#define RK_HANDOFF_AUTO 0xFFFFFFFEu
rk_status rk_raytrace_set_handoff(uint32_t ray_faces, uint32_t packet_faces, uint32_t list_cap);
rk_status rk_raytrace_get_handoff(uint32_t* ray_faces, uint32_t* packet_faces, uint32_t* list_cap);

/* ---- wide-BVH fast path (no reference counterpart) --------------------------- */
/* A second tree for throughput: Morton LBVH (binary radix tree built bottom-up in one
 * pass with its boxes, leaf collapse), stored as 64-B two-child nodes and leaf-ordered triangle records
 * (v0, v1-v0, v2-v0; 48 B). The trace uses the reference triangle arithmetic, so a
 * hit's t/u/v are bit-identical to the reference's for the same triangle; ties on t
 * go to the smallest original face id.
 *   nodes: capacity n x 64 B (root at 0); tris: n x 48 B; indices: n (sorted
 *   position -> original face id); boxes: n per-face boxes in original order
 *   (CUDAMeshTree's boxes array), or NULL to skip them.
 * morton_mode is RK_MORTON_30 or RK_MORTON_63; leaf_max in [1, 16]. */
typedef struct rk_fast_bvh {
  void* nodes;
  void* tris;
  uint32_t* indices;
  rk_aabox* boxes;
  uint32_t face_count;
  uint32_t reserved;
} rk_fast_bvh;
size_t rk_fast_build_workspace_size(uint32_t face_count, int morton_mode);
rk_status rk_fast_build(rk_stream stream, const rk_uint4* faces, const rk_float4* vertices, uint32_t face_count,
                        const rk_aabox* hitbox, int morton_mode, uint32_t leaf_max, const rk_fast_bvh* bvh,
                        uint32_t* info, void* ws, size_t ws_bytes);
/* Trace options. det_epsilon is the triangle test's parallel-ray threshold:
 * the reference rejects |det| < 1e-6 (CUDA_raytraceMeshTree.cu:36,65), an
 * absolute bound that rejects every hit once edges are ~1e-3 long; 0 rejects
 * only exactly parallel rays. NULL options = the reference threshold. */
#define RK_DET_EPSILON_REFERENCE 0.000001f
/* rk_trace_options.flags. RK_TRACE_RESET_DEPTH: every ray starts with tmax
 * FLT_MAX (depths is not read) and every ray's depth is written, FLT_MAX for a
 * miss, whose faces_id and coords are left as they are: the state that
 * filling depths with FLT_MAX and tracing leaves, in one launch. For
 * rk_fast_trace, rk_fast_trace_sorted, rk_fast_trace_counted and rk_fast_autotune. */
#define RK_TRACE_RESET_DEPTH 1u
typedef struct rk_trace_options {
  float det_epsilon;
  uint32_t flags;  /* RK_TRACE_* */
  uint32_t config;   /* traversal variant (rk_fast_autotune, rk_fast_trace_variant); 0: the default */
  uint32_t reserved;
} rk_trace_options;
/* Closest hit for n_rays rays. faces_id receives the sorted position (index into
 * bvh->tris; original face id = bvh->indices[position]); depths is read as tmax. */
rk_status rk_fast_trace(rk_stream stream, const rk_fast_bvh* bvh, const rk_float4* origins,
                        const rk_float4* directions, uint32_t n_rays, uint32_t* faces_id, rk_float2* coords,
                        float* depths, const rk_trace_options* options);
/* rk_fast_trace for incoherent rays: the rays are first ordered by the Morton code
 * of their origin in the tree's root box (a key kernel and a radix sort in ws),
 * then traced in that order, so rays of a wave start near each other and share
 * node and triangle fetches. Results are written to each ray's own slot and are
 * the same as rk_fast_trace's. Pays off when rays start scattered through the
 * scene (secondary or random rays); camera rays are coherent already. The first
 * 4 bytes of ws receive the sort's RK_FLAG_* bits, as for the sorts above.
 * options->config selects an rk_fast_trace variant (its ray runs do not apply:
 * the order replaces them). */
size_t rk_fast_trace_sorted_workspace_size(uint32_t n_rays);
rk_status rk_fast_trace_sorted(rk_stream stream, const rk_fast_bvh* bvh, const rk_float4* origins,
                               const rk_float4* directions, uint32_t n_rays, uint32_t* faces_id, rk_float2* coords,
                               float* depths, const rk_trace_options* options, void* ws, size_t ws_bytes);
/* rk_fast_trace that also adds the number of child-box tests and triangle tests
 * it performs to counters[0] and counters[1] (device, uint64). Same results. */
rk_status rk_fast_trace_counted(rk_stream stream, const rk_fast_bvh* bvh, const rk_float4* origins,
                                const rk_float4* directions, uint32_t n_rays, uint32_t* faces_id, rk_float2* coords,
                                float* depths, const rk_trace_options* options, uint64_t* counters);

/* rk_fast_build that also leaves the sorted Morton keys in sorted_keys (device,
 * face_count keys of 4 B, or 8 B for RK_MORTON_63), for rk_fast_refit. */
rk_status rk_fast_build_keyed(rk_stream stream, const rk_uint4* faces, const rk_float4* vertices, uint32_t face_count,
                              const rk_aabox* hitbox, int morton_mode, uint32_t leaf_max, const rk_fast_bvh* bvh,
                              uint32_t* info, void* sorted_keys, void* ws, size_t ws_bytes);
/* rk_fast_build with no workspace argument, for callers that cannot allocate
 * (code under HIP graph capture, tight serving loops, hosts that forbid
 * hipMalloc). No counterpart in the reference. It calls no allocator: the
 * workspace is a file-scope device array of RK_NOALLOC_SCRATCH_BYTES (128 MiB),
 * one per device, held until the process ends. face_count is limited to what
 * fits it, rk_fast_build_workspace_size(face_count, morton_mode) <=
 * RK_NOALLOC_SCRATCH_BYTES: 1,944,805 faces with RK_MORTON_30 and 1,720,320
 * with RK_MORTON_63; larger counts return RK_ERROR_WORKSPACE. Everything the
 * build reads from the array is cleared by the build itself. All calls on one
 * device share that array, so calls that overlap on one device (concurrent
 * streams, threads) race: order them on one stream. Same results as
 * rk_fast_build. */
#define RK_NOALLOC_SCRATCH_BYTES (128ull << 20)
rk_status rk_fast_build_noalloc(rk_stream stream, const rk_uint4* faces, const rk_float4* vertices,
                                uint32_t face_count, const rk_aabox* hitbox, int morton_mode, uint32_t leaf_max,
                                const rk_fast_bvh* bvh, uint32_t* info);
/* Refit after the vertices moved: keeps the topology of the rk_fast_build_keyed
 * that produced sorted_keys and bvh->indices (same faces, same morton_mode and
 * leaf_max) and recomputes every node box and triangle record from 'vertices'.
 * Workspace: rk_fast_build_workspace_size. Runs only the hierarchy passes
 * (no key computation, no sort). */
rk_status rk_fast_refit(rk_stream stream, const rk_uint4* faces, const rk_float4* vertices, uint32_t face_count,
                        int morton_mode, uint32_t leaf_max, const rk_fast_bvh* bvh, const void* sorted_keys, void* ws,
                        size_t ws_bytes);

/* A pinhole camera's primary rays, as CUDA_calcPrimaryRays.cu:43-55 forms them:
 * pixel (x, y) has origin 'origin' and direction (direction - x*step_x) - y*step_y. */
typedef struct rk_camera_rays {
  rk_float4 origin, direction, step_x, step_y;
} rk_camera_rays;
/* A camera frame traced without ray arrays: each pixel's primary ray is formed
 * in the trace kernel (CUDA_calcPrimaryRays.cu:20-55 fused into the trace), tmax
 * is FLT_MAX, and every pixel is written: a hit gives its sorted position, (u, v)
 * and t; a miss gives 0xFFFFFFFF, (0, 0) and FLT_MAX. Outputs are row-major
 * width x height. */
rk_status rk_fast_trace_camera(rk_stream stream, const rk_fast_bvh* bvh, const rk_camera_rays* camera, int width,
                               int height, uint32_t* faces_id, rk_float2* coords, float* depths,
                               const rk_trace_options* options);
/* The frame loop of samples/02_how_to_use_the_raytracer.cpp:67-85 in one launch:
 * CUDA_calcPrimaryRays.cu:20-55, the closest hit, CUDA_interpolateData.cu:45-141
 * (the hit face's normal only) and CUDAShaderKernel.cu:14-43, i.e. what
 * rk_fast_trace_camera, then rk_interpolate_data (out_normals) and rk_shade
 * leave in pixels, bit for bit, without the intermediate surfaces. No single
 * reference counterpart: it replaces that sequence of four calls for callers of
 * the fast tree. faces: the mesh's faces in the tree's leaf order (the order the
 * reference's builds leave the mesh in); normals: per-vertex normals or NULL
 * (face normals); pixels: height rows of pitch words (pitch >= width);
 * faces_id, coords, depths: all NULL, or all set to receive what
 * rk_fast_trace_camera writes. options->config is not used (the default camera
 * variant). */
rk_status rk_fast_render_camera(rk_stream stream, const rk_fast_bvh* bvh, const rk_camera_rays* camera, int width,
                                int height, const rk_uint4* faces, const rk_float4* vertices, const rk_float4* normals,
                                rk_float4 back_color, rk_float4 cam_dir, uint32_t* pixels, int pitch,
                                uint32_t* faces_id, rk_float2* coords, float* depths,
                                const rk_trace_options* options);

/* Traversal variants and their autotuner (no counterpart in the reference, whose
 * launch shapes are fixed: 16x12 packets, 96-entry stacks, RT.cu:434-480). The
 * results of every variant are identical; they differ in speed. A variant of
 * rk_fast_trace sets threads per workgroup (workgroup), how many consecutive rays
 * a wave takes from each band of the ray array (ray_run; 0: 64 consecutive rays)
 * and how many stack entries per lane live in LDS (lds_stack); a variant of
 * rk_fast_trace_camera sets the workgroup, the bands of rows a wave draws its
 * pixels from (bands, of ray_run = 64 / bands pixels each) and whether each of
 * the eight XCDs takes one vertical strip of the image (xcd_strips); in either
 * kind, lanes that are done may take the oldest pending subtree of a lane that
 * is not, rather than its newest (steal_oldest). The variant
 * is rk_trace_options.config; rk_fast_trace_variant describes one (camera != 0:
 * a camera variant). */
typedef struct rk_trace_tuning {
  uint32_t config;
  uint32_t workgroup, ray_run, bands, lds_stack, xcd_strips;
  float ms;              /* mean time per call of the chosen variant in the tuning run */
  uint32_t steal_oldest; /* idle lanes take a busy lane's oldest pending subtree, not its newest */
  uint32_t leaf_batch;   /* leaves are postponed until this many lanes hold one (0: tested at once) */
  uint32_t donate_min;   /* pending subtrees are handed out only when this many lanes are idle */
  uint32_t bound_every;  /* lanes helping another lane's ray read its best hit every this many steps */
} rk_trace_tuning;
rk_status rk_fast_trace_variant(uint32_t config, int camera, rk_trace_tuning* out);
/* Times every rk_fast_trace variant on the caller's rays (reps timed calls each
 * after one untimed call, tmax restored from depths' contents at entry before
 * each call), writes the fastest to *best, and leaves the frame traced by it in
 * faces_id/coords/depths. ws: rk_fast_autotune_workspace_size(n_rays) bytes
 * (a copy of the entry tmax). Synchronises 'stream'. */
size_t rk_fast_autotune_workspace_size(uint32_t n_rays);
rk_status rk_fast_autotune(rk_stream stream, const rk_fast_bvh* bvh, const rk_float4* origins,
                           const rk_float4* directions, uint32_t n_rays, uint32_t* faces_id, rk_float2* coords,
                           float* depths, const rk_trace_options* options, int reps, void* ws, size_t ws_bytes,
                           rk_trace_tuning* best);
/* The same for rk_fast_trace_camera variants on the caller's camera frame
 * (options->config is ignored). Synchronises 'stream'. */
rk_status rk_fast_autotune_camera(rk_stream stream, const rk_fast_bvh* bvh, const rk_camera_rays* camera, int width,
                                  int height, uint32_t* faces_id, rk_float2* coords, float* depths,
                                  const rk_trace_options* options, int reps, rk_trace_tuning* best);

/* Deforming meshes: out[v] = base[v] + amplitude * sin(omega * time + phase_v) * axis_v,
 * with phase_v and axis_v (uniform on the sphere, or +z when along_z != 0) hashed
 * from seed and v. out.w = base.w. */
rk_status rk_deform_vertices(rk_stream stream, const rk_float4* base, rk_float4* out, uint32_t count, float amplitude,
                             float omega, float time, uint64_t seed, int along_z);
/* out[0] = wrapping sum over rays of a hash of (ray index, original face id,
 * depth, u, v bits), out[1] = hits (faces_id != 0xFFFFFFFF). The original id is
 * indices[faces_id] when indices is not NULL, so the sum does not depend on the
 * tree that produced the frame. out: 2 device uint64. */
rk_status rk_frame_checksum(rk_stream stream, const uint32_t* faces_id, const rk_float2* coords, const float* depths,
                            uint32_t count, const uint32_t* indices, uint32_t face_count, uint64_t* out);

/* ---- batches: many meshes built and traced together (no reference counterpart) ----
 * The reference builds and traces one mesh per call (CUDAMeshTree.cpp:88-103,
 * CUDARaytracer.cpp:130-135); a caller with many meshes loops over them. A batch
 * holds all meshes in shared arrays: faces of mesh m are faces[face_offset ..
 * face_offset + face_count) and index the mesh's own vertices, which start at
 * vertices[vertex_offset]; its rays are origins/directions[ray_offset .. +ray_count).
 * Meshes are listed in order of face_offset and of ray_offset, with no gaps
 * (face_offset[m+1] = face_offset[m] + face_count[m], the same for rays), and
 * face_count >= 1. The batch tree is one rk_fast_bvh over all faces: mesh m's
 * root is node face_offset[m], its nodes and triangle records lie in
 * [face_offset, face_offset + face_count), and indices[] maps a sorted position
 * to the batch face index. At most 4096 meshes and 2^27 faces per batch. */
typedef struct rk_batch_mesh {
  uint32_t face_offset, face_count;
  uint32_t vertex_offset, vertex_count;
  uint32_t ray_offset, ray_count;
  uint32_t reserved[2];
} rk_batch_mesh;
#define RK_BATCH_MAX_MESHES 4096u

/* Tuning of a batch. leaf_max in [1, 16]; trace_mode 0 = one workgroup per
 * trace_wg rays, 1 = a persistent grid that takes 64-ray tiles from eight
 * work queues (one per XCD, stealing when its own runs dry); trace_wg is 64,
 * 128 or 256; stack_depth is 64 or 96 (the per-ray traversal stack); ray_order
 * 0 traces the rays as given, 1 first sorts them (within their mesh) by the
 * Morton code of their origin in the mesh's root box, for incoherent rays;
 * donate_min: a wave hands its pending subtrees to idle lanes only on steps
 * where at least this many of its lanes are idle, 16 or 32 (0: 16). The
 * results are the same whatever the configuration. */
typedef struct rk_batch_config {
  uint32_t leaf_max;
  uint32_t trace_mode;
  uint32_t trace_wg;
  uint32_t stack_depth;
  uint32_t ray_order;
  uint32_t donate_min;
  uint32_t reserved[2];
} rk_batch_config;
/* leaf_max 1, persistent queues, 256-thread workgroups, stack 64, rays as given. */
rk_batch_config rk_batch_config_default(void);

/* Builds every mesh of the batch: per-mesh 30-bit Morton keys in each mesh's
 * hitbox (hitboxes: device, one per mesh), one radix sort of (mesh, code)
 * keys for the whole batch, one hierarchy pass. bvh->nodes and bvh->tris hold
 * face_count records, bvh->indices face_count entries; bvh->boxes is unused.
 * meshes: device array of mesh_count entries. */
size_t rk_batch_build_workspace_size(uint32_t mesh_count, uint32_t face_count);
rk_status rk_batch_build(rk_stream stream, const rk_batch_mesh* meshes, uint32_t mesh_count, const rk_uint4* faces,
                         const rk_float4* vertices, uint32_t face_count, const rk_aabox* hitboxes,
                         const rk_batch_config* config, const rk_fast_bvh* bvh, uint32_t* info, void* ws,
                         size_t ws_bytes);
/* Closest hit of every ray of every mesh in that mesh's tree. faces_id receives
 * the sorted position (original batch face = bvh->indices[position]); depths is
 * read as tmax; a miss leaves the outputs unchanged. The first 4 bytes of ws
 * receive the call's RK_FLAG_* bits (the ray sort of ray_order 1), as for the sorts. */
size_t rk_batch_trace_workspace_size(uint32_t mesh_count, uint32_t ray_count);
rk_status rk_batch_trace(rk_stream stream, const rk_batch_mesh* meshes, uint32_t mesh_count, const rk_fast_bvh* bvh,
                         const rk_float4* origins, const rk_float4* directions, uint32_t ray_count, uint32_t* faces_id,
                         rk_float2* coords, float* depths, const rk_trace_options* options,
                         const rk_batch_config* config, void* ws, size_t ws_bytes);

/* ---- graphs ---------------------------------------------------------------------
 * Any sequence of rk_* calls issued on one non-NULL stream between begin and
 * end is recorded into a graph (hipStreamBeginCapture) and replayed by
 * rk_graph_launch with the same pointers and sizes. The library records no
 * memset or host copy in the calls that can be captured (the fast and batch
 * builds and traces, fills, sorts); rk_build_sah and the reference LBVH
 * (which read results back) cannot be captured. */
typedef struct rk_graph_st* rk_graph;
rk_status rk_graph_capture_begin(rk_stream stream);
rk_status rk_graph_capture_end(rk_stream stream, rk_graph* graph);
rk_status rk_graph_launch(rk_graph graph, rk_stream stream);
rk_status rk_graph_destroy(rk_graph graph);

/* ---- autotuning ------------------------------------------------------------------
 * Times candidate configurations on the caller's own batch: a build for each
 * leaf_max in {1, 2, 4, 8}, then each trace variant (rays as given or sorted,
 * both modes, workgroups of 64, 128 and 256, stacks of 64 and 96), each the median of 'reps' runs (1..64)
 * timed with HIP events on 'stream', and writes the fastest build + trace pair
 * to *best (its time to *best_ms when not NULL). Each trace starts from
 * depths = FLT_MAX. On return the tree and outputs are those of the chosen
 * configuration. ws_bytes >= rk_batch_build_workspace_size +
 * rk_batch_trace_workspace_size. Synchronises 'stream'. */
rk_status rk_batch_autotune(rk_stream stream, const rk_batch_mesh* meshes, uint32_t mesh_count,
                            const rk_uint4* faces, const rk_float4* vertices, uint32_t face_count,
                            const rk_aabox* hitboxes, const rk_fast_bvh* bvh, uint32_t* info,
                            const rk_float4* origins, const rk_float4* directions, uint32_t ray_count,
                            uint32_t* faces_id, rk_float2* coords, float* depths, const rk_trace_options* options,
                            uint32_t reps, rk_batch_config* best, float* best_ms, void* ws, size_t ws_bytes);

/* ---- serving (no reference counterpart) --------------------------------------------
 * A request loop over deforming scenes. Scenes are added once (host arrays, kept
 * in pinned memory) and live in one device arena of arena_bytes (at most
 * cfg.hbm_budget); a scene is made resident when a request needs it, evicting
 * scenes no running request uses, least recently used first. A queued request is
 * admitted, in arrival order, when one of cfg.slots slots (a stream with its own
 * workspace and frame buffers) is free and its scene fits; admitted requests run
 * concurrently. Per request: upload if just made resident, deform the scene's
 * vertices to the request's time (amplitude cfg.amplitude_frac of the scene's
 * extent, as rk_deform_vertices), rebuild (rk_fast_build_keyed) when the scene has
 * no tree or has been refitted cfg.rebuild_every times since its last build (0:
 * always), else refit (as rk_fast_refit), trace the camera frame
 * (rk_fast_trace_camera) and checksum it (as rk_frame_checksum, the result written
 * to pinned host memory with the request's completion). A frame depends only on
 * the scene, time, camera and size, not on what else runs or on refit vs rebuild.
 * The host calls are not thread-safe: one thread drives a server. */
#define RK_REQUEST_FORCE_REBUILD 1u
#define RK_REQUEST_FORCE_REFIT 2u  /* when the scene has a tree */
typedef struct rk_server_config {
  uint32_t slots;          /* requests on the GPU at once, 1..64 */
  uint32_t leaf_max;       /* 1..16 */
  uint32_t rebuild_every;  /* refits between rebuilds of a scene */
  uint32_t max_faces;      /* largest scene that can be added */
  uint32_t max_pixels;     /* largest frame rendered into the slot buffers */
  float det_epsilon;       /* triangle test threshold */
  float amplitude_frac;    /* deformation amplitude / scene extent */
  float omega;             /* deformation angular rate, radians per unit of request time */
  uint64_t hbm_budget;     /* cap on arena_bytes */
  uint32_t plain;          /* != 0: the plain per-request path, a baseline (below) */
  uint32_t trace_config;   /* rk_fast_trace_camera variant (rk_fast_autotune_camera); 0: the default */
  uint32_t reserved[2];
} rk_server_config;
/* 16 slots, leaf 1, rebuild every 8 refits, 2^21 faces, 1024^2 pixels, threshold 0,
 * amplitude 1% of extent, omega 2 pi, budget 200 GiB, plain 0. Slots are streams: they
 * overlap only on distinct hardware queues (GPU_MAX_HW_QUEUES, 4 by default).
 * plain != 0 runs every request the way a caller of the plain entry points would,
 * as a baseline for the request loop: deform, rk_fast_build from scratch, clear the
 * frame, rk_calc_primary_rays into ray arrays, rk_fast_trace, rk_frame_checksum and a
 * copy of it to the host. Frames and checksums equal the request loop's. */
rk_server_config rk_server_config_default(void);
typedef struct rk_server_st* rk_server;
rk_status rk_server_create(const rk_server_config* cfg, uint64_t arena_bytes, rk_server* server);
rk_status rk_server_destroy(rk_server server);
/* Device bytes a scene occupies while resident. */
size_t rk_server_scene_bytes(uint32_t vertex_count, uint32_t face_count);
/* Host arrays; deform_along_z != 0 moves vertices along +z (terrains), else along
 * per-vertex random directions. Faces must index vertex_count vertices. */
rk_status rk_server_add_scene(rk_server server, const rk_float4* vertices, uint32_t vertex_count,
                              const rk_uint4* faces, uint32_t face_count, int deform_along_z, uint32_t* scene_id);
typedef struct rk_request {
  uint32_t scene;
  uint32_t width, height;
  uint32_t flags;           /* RK_REQUEST_* */
  float time;               /* scene time: the deformation phase */
  float reserved_f;
  rk_camera_rays camera;
  double arrival;           /* server clock (rk_server_now); < 0 = at submission */
  uint32_t* faces_id;       /* optional device outputs of width*height, all three or */
  rk_float2* coords;        /* none (then the slot's buffers; only the checksum */
  float* depths;            /* is returned) */
} rk_request;
typedef struct rk_result {
  uint64_t checksum, hits;  /* rk_frame_checksum of the frame */
  double arrival, start, done;  /* server clock, seconds: done is when completion was observed */
  uint32_t scene, pixels;
  uint32_t rebuilt;         /* 1 rebuild, 0 refit */
  uint32_t uploaded;        /* the scene was uploaded for this request */
  uint32_t evictions;       /* scenes evicted to admit it */
  uint32_t slot;
  uint32_t build_flags;     /* RK_FLAG_* bits of the scene's latest build, read after this frame */
  rk_status status;         /* RK_SUCCESS, or RK_ERROR_DEVICE when build_flags has RK_FLAG_SPIN_LIMIT:
                               the tree and this frame are invalid (the scene's next request rebuilds) */
} rk_result;
double rk_server_now(rk_server server);
/* Queues a request; *ticket identifies it. Nothing runs until progress. */
rk_status rk_server_submit(rk_server server, const rk_request* request, uint64_t* ticket);
/* Non-blocking: records finished requests, admits queued ones, launches their work.
 * When a launch fails, that error is returned by this call and every later one:
 * requests already running still finish and are recorded, queued ones are not
 * admitted. A request whose build reports RK_FLAG_SPIN_LIMIT finishes with
 * rk_result.status = RK_ERROR_DEVICE; the server goes on. */
rk_status rk_server_progress(rk_server server, uint32_t* completed);
/* Progress until nothing is queued or running, or a launch fails (its error). */
rk_status rk_server_wait_all(rk_server server);
/* RK_ERROR_UNSUPPORTED while the request has not finished. */
rk_status rk_server_result(rk_server server, uint64_t ticket, rk_result* result);
typedef struct rk_server_stats {
  uint64_t completed, rebuilds, refits, uploads, evictions, rays;
  double first_arrival, last_done;
  double p50_ms, p99_ms, mean_ms, max_ms;  /* latency = done - arrival */
  double mrays_per_s;                      /* rays / (last_done - first_arrival) */
  uint64_t peak_resident_bytes;
  uint64_t failed;                         /* finished requests whose rk_result.status is not RK_SUCCESS */
} rk_server_stats;
/* Statistics over tickets [first_ticket, first_ticket + count). */
rk_status rk_server_stats_get(rk_server server, uint64_t first_ticket, uint64_t count, rk_server_stats* stats);
/* Open-loop driver: count requests with Poisson arrivals at 'rate' per second
 * (rate 0: all arrive at once), each for a scene drawn uniformly, at time =
 * its arrival, with a square frame of side log-uniform in [side_lo, side_hi] and
 * a 60-degree camera at 30 degrees elevation on an orbit around the scene. Runs
 * until all are done and returns their statistics. */
rk_status rk_server_run_poisson(rk_server server, double rate, uint32_t count, uint32_t side_lo, uint32_t side_hi,
                                uint64_t seed, rk_server_stats* stats, uint64_t* first_ticket);

#ifdef __cplusplus
}
#endif

#endif /* RK_H */
