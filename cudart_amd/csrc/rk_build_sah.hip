// Binned SAH build (CUDA_buildMeshTreeSAH, CUDA_buildMeshTreeSAH.cu:224-789)
// and the Hybrid refinement of LBVH leaves (kernel_initStack, :647-666), for
// gfx950. The reference processes one node per 32-thread block with thread 0
// doing the binning, sweep and partition serially, and a host round trip per
// level through a 4096-entry global stack. Here each level's nodes ("items")
// are cut into 1024-face chunks, one workgroup per chunk:
//   k_sah_prep       chunk offsets per item (one workgroup scan), accumulator reset
//   k_sah_bounds     centroid bounds per item (ordered-int atomics)
//   k_sah_bins       32 bins per item: count + primitive-box bounds (LDS, then atomics)
//   k_sah_decide     one half-wave per item, one lane per bin: the reference
//                    sweep (:424-474), leaf test (:479-496), children boxes
//                    (:501-530)
//   k_sah_small      items of 17 to 1024 faces, one wave each: all of the
//                    above and the partition, in LDS; the chunk kernels take
//                    the items of more than 1024 faces
//   k_sah_emit       1024 items per workgroup: scan of the internal items,
//                    node records, next-level items; child ids in
//                    breadth-first order
//   k_sah_count      per-chunk left counts
//   k_sah_offsets    per-item exclusive scan over chunks, one wave per item
//   k_sah_scatter    stable partition (bin <= m_min goes left, :572-594)
//   k_sah_copyback   partitioned ranges back into indices
// Min/max are exact in any order, so the atomics and the sweep's scans
// reproduce the reference's serial merges; each cost is the reference's
// expression on the same values. Node numbering: items are processed level by
// level in queue order, and the k-th internal item of a level takes children
// (count + 2k, count + 2k + 1), the breadth-first equivalent of the
// reference's atomicAdd(node_count, 2).
#include "rk_device.h"

namespace {

using rk::Box;
using rk::Node;

constexpr uint32_t kChunk = 1024;
// threads per workgroup of the per-chunk kernels (bounds, bins, count,
// scatter, copyback): one chunk each
#ifndef RK_SAH_CHUNK_THREADS
#define RK_SAH_CHUNK_THREADS 128
#endif
constexpr int kCT = RK_SAH_CHUNK_THREADS;
static_assert(kCT % 64 == 0 && kCT <= 256, "whole waves, at most 4");
constexpr uint32_t kEmitGroup = 1024;  // items per workgroup of k_sah_emit
constexpr int kBins = 32;
constexpr uint32_t kLeafMin = 16;    // MIN_LEAF_SIZE, :55
constexpr uint32_t kLeafMax = 224;   // MAX_LEAF_SIZE, :56

struct Item {
  uint32_t node, lo, hi, pad;
  float bmin[3], bmax[3];
  float pad2[2];
};
static_assert(sizeof(Item) == 48, "item is 48 B");

// Accumulators of an item of more than kChunk faces: its centroid bounds
// (6 ordered ints) in kBoundSlots copies, which its chunks update in turn so
// that fewer atomics meet on one address, then kBinSlots copies of 32 bins
// of (count, 3 min, 3 max ordered ints).
constexpr int kBoundSlots = 8;
constexpr int kBinSlots = 4;  // copies of the bins, likewise
constexpr int kBinBase = 6 * kBoundSlots;
constexpr int kBinWords = kBins * 7;
constexpr int kAccWords = kBinBase + kBinSlots * kBinWords;

// Per-item decision written by k_sah_decide.
struct Decision {
  int axis;
  float scale, bmin0;
  int m_min;        // -1: leaf
  uint32_t lcount;
  uint32_t pad[3];
  float child[12];  // internal items: left min xyz, left max xyz, right min xyz, right max xyz
};
static_assert(sizeof(Decision) == 80, "decision is 80 B");

__device__ __forceinline__ int f2o(float f) {  // order-preserving float -> int
  int i = __float_as_int(f);
  return i >= 0 ? i : (i ^ 0x7FFFFFFF);
}
__device__ __forceinline__ float o2f(int i) { return __int_as_float(i >= 0 ? i : (i ^ 0x7FFFFFFF)); }

__device__ __forceinline__ float box_area(float mnx, float mny, float mnz, float mxx, float mxy, float mxz) {
  float vx = mxx - mnx, vy = mxy - mny, vz = mxz - mnz;  // CUDABox.h:35-44
  return 2.f * (vx * vy + vx * vz + vy * vz);
}

__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* s_tmp, uint32_t* total, int nthreads) {
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint32_t incl = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    uint32_t y = __shfl_up(incl, off, 64);
    if (lane >= (uint32_t)off) incl += y;
  }
  if (lane == 63) s_tmp[wave] = incl;
  __syncthreads();
  uint32_t wprefix = 0, tot = 0;
  for (int w = 0; w < nthreads / 64; ++w) {
    uint32_t c = s_tmp[w];
    wprefix += ((uint32_t)w < wave) ? c : 0u;
    tot += c;
  }
  __syncthreads();
  *total = tot;
  return wprefix + incl - v;
}

// ctl words: [0] flags; per level parity p (level & 1), at kCtlLevel(p):
// +0 items (m), +1 items of more than kChunk faces, +2 chunks (T), +3 slots
// taken, +4 node count before the level. A level of parity p reads its own
// words and writes the next level's m, big-item count and node count; its
// k_sah_prep clears the next level's counters, which nothing of this level
// uses.
constexpr int kCtlFlags = 0, kCtlWords = 32;
__device__ __host__ constexpr int kCtlLevel(uint32_t p) { return 8 + 8 * (int)p; }
constexpr int kLvM = 0, kLvBig = 1, kLvT = 2, kLvSlot = 3, kLvCount = 4;

// Items of more than kChunk faces get a slot of accumulators (kept in the
// item's pad word) and a run of chunks, in any order among items; chunk_item
// maps a chunk back to its item. Smaller items are k_sah_small's. One wave
// per item, launched over the level's items (with a bound, kPrepItems per
// workgroup), and the grid's threads also clear the emit groups' counters.
constexpr uint32_t kPrepItems = 4;
__global__ __launch_bounds__(64 * kPrepItems) void k_sah_prep(Item* __restrict__ items, uint32_t* __restrict__ ctl,
                                                              uint32_t par, uint32_t* __restrict__ chunk_base,
                                                              uint32_t* __restrict__ chunk_item,
                                                              int* __restrict__ acc,
                                                              uint32_t* __restrict__ group_internal) {
  uint32_t* lv = ctl + kCtlLevel(par);
  const uint32_t m = lv[kLvM];
  const uint32_t t = blockIdx.x * (64 * kPrepItems) + threadIdx.x, lane = threadIdx.x & 63;
  if (t == 0) {
    uint32_t* nx = ctl + kCtlLevel(par ^ 1u);
    nx[kLvBig] = 0u;
    nx[kLvT] = 0u;
    nx[kLvSlot] = 0u;
  }
  if (t < (m + kEmitGroup - 1) / kEmitGroup) group_internal[t] = 0u;
  const uint32_t i = t >> 6;
  if (i >= m) return;  // wave-uniform
  const uint32_t size = items[i].hi - items[i].lo;
  if (size <= kChunk) return;
  const uint32_t nch = (size + kChunk - 1) / kChunk;
  uint32_t slot = 0, base = 0;
  if (lane == 0) {
    slot = atomicAdd(lv + kLvSlot, 1u);
    base = atomicAdd(lv + kLvT, nch);
    items[i].pad = slot;
    chunk_base[i] = base;
  }
  slot = __shfl(slot, 0, 64);
  base = __shfl(base, 0, 64);
  for (uint32_t j = lane; j < nch; j += 64) chunk_item[base + j] = i;
  int* a = acc + (size_t)slot * kAccWords;  // the bins are reset by k_sah_bounds
  if (lane < 6 * kBoundSlots) a[lane] = (lane % 6) < 3 ? f2o(RK_HUGE) : f2o(-RK_HUGE);
}

struct ChunkRange { uint32_t item, slot, lo, hi; bool first, ok; };

__device__ __forceinline__ ChunkRange chunk_range(const Item* __restrict__ items, const uint32_t* __restrict__ ctl,
                                                  uint32_t par, const uint32_t* __restrict__ chunk_base,
                                                  const uint32_t* __restrict__ chunk_item, uint32_t c) {
  ChunkRange r;
  r.ok = c < ctl[kCtlLevel(par) + kLvT];
  if (!r.ok) return r;
  r.item = chunk_item[c];
  const Item it = items[r.item];
  const uint32_t j = c - chunk_base[r.item];
  r.slot = it.pad;
  r.first = j == 0;
  r.lo = it.lo + j * kChunk;
  r.hi = min(it.hi, r.lo + kChunk);
  return r;
}

// The boxes of a chunk's faces, kPer per thread, with every gather in flight
// at once: the face ids first, then the boxes.
constexpr int kPer = kChunk / kCT;
__device__ __forceinline__ void chunk_boxes(const ChunkRange& r, const uint32_t* __restrict__ indices,
                                            const Box* __restrict__ boxes, Box (&b)[kPer], bool (&in)[kPer]) {
  uint32_t id[kPer];
#pragma unroll
  for (int u = 0; u < kPer; ++u) {
    const uint32_t l = r.lo + threadIdx.x + (uint32_t)u * kCT;
    in[u] = l < r.hi;
    id[u] = in[u] ? indices[l] : 0u;
  }
#pragma unroll
  for (int u = 0; u < kPer; ++u)
    if (in[u]) b[u] = rk::load_box(boxes + id[u]);
}

__device__ __forceinline__ float center_axis(const Box& b, int a) {  // 0.5f * (min + max), :292
  return a == 0 ? (b.mn.x + b.mx.x) * 0.5f : (a == 1 ? (b.mn.y + b.mx.y) * 0.5f : (b.mn.z + b.mx.z) * 0.5f);
}

__global__ __launch_bounds__(kCT) void k_sah_bounds(const Item* __restrict__ items, const uint32_t* __restrict__ ctl,
                                                    uint32_t par, const uint32_t* __restrict__ chunk_base,
                                                    const uint32_t* __restrict__ chunk_item,
                                                    const uint32_t* __restrict__ indices, const Box* __restrict__ boxes,
                                                    int* __restrict__ acc) {
  const ChunkRange r = chunk_range(items, ctl, par, chunk_base, chunk_item, blockIdx.x);
  if (!r.ok) return;
  int* a = acc + (size_t)r.slot * kAccWords;
  if (r.first) {  // the item's first chunk resets its bins for k_sah_bins
    int* bins = a + kBinBase;
    for (int w = threadIdx.x; w < kBinSlots * kBinWords; w += kCT) {
      int k = (w % kBinWords) % 7;
      bins[w] = k == 0 ? 0 : (k < 4 ? f2o(RK_HUGE) : f2o(-RK_HUGE));
    }
  }
  // ordered ints, as the atomics compare them
  int mn[3] = {f2o(RK_HUGE), f2o(RK_HUGE), f2o(RK_HUGE)}, mx[3] = {f2o(-RK_HUGE), f2o(-RK_HUGE), f2o(-RK_HUGE)};
  Box bb[kPer];
  bool in[kPer];
  chunk_boxes(r, indices, boxes, bb, in);
#pragma unroll
  for (int u = 0; u < kPer; ++u) {
    if (!in[u]) continue;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const int c = f2o(center_axis(bb[u], q));
      mn[q] = min(mn[q], c);
      mx[q] = max(mx[q], c);
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1)
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      mn[q] = min(mn[q], __shfl_xor(mn[q], off, 64));
      mx[q] = max(mx[q], __shfl_xor(mx[q], off, 64));
    }
  // one set of atomics per workgroup, into one of the item's bound slots
  __shared__ int s_red[kCT / 64][6];
  const uint32_t w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0)
    for (int q = 0; q < 3; ++q) { s_red[w][q] = mn[q]; s_red[w][3 + q] = mx[q]; }
  __syncthreads();
  if (threadIdx.x == 0) {
    int* bs = a + 6 * (blockIdx.x % kBoundSlots);
    for (int q = 0; q < 3; ++q) {
      int lo = s_red[0][q], hi = s_red[0][3 + q];
      for (int v = 1; v < kCT / 64; ++v) { lo = min(lo, s_red[v][q]); hi = max(hi, s_red[v][3 + q]); }
      atomicMin(bs + q, lo);
      atomicMax(bs + 3 + q, hi);
    }
  }
}

// An item's centroid bounds from its bound slots.
__device__ __forceinline__ void acc_bounds(const int* __restrict__ a, int (&cb)[6]) {
  for (int q = 0; q < 3; ++q) { cb[q] = a[q]; cb[3 + q] = a[3 + q]; }
  for (int v = 1; v < kBoundSlots; ++v)
    for (int q = 0; q < 3; ++q) { cb[q] = min(cb[q], a[6 * v + q]); cb[3 + q] = max(cb[3 + q], a[6 * v + 3 + q]); }
}

// Axis and bin scale of an item from its centroid bounds (:302-342).
__device__ __forceinline__ void item_axis(const int (&a)[6], int& axis, float& scale, float& bmin0) {
  float ex = o2f(a[3]) - o2f(a[0]), ey = o2f(a[4]) - o2f(a[1]), ez = o2f(a[5]) - o2f(a[2]);
  if (ex > ey) axis = (ex > ez) ? 0 : 2;
  else axis = (ey > ez) ? 1 : 2;
  float e = axis == 0 ? ex : (axis == 1 ? ey : ez);
  scale = (float)(kBins - 1) / e;
  bmin0 = o2f(a[axis]);
}

__device__ __forceinline__ int bin_of(const Box& b, int axis, float scale, float bmin0) {  // :375-385
  int m = rk::f2i_sat((center_axis(b, axis) - bmin0) * scale);
  if (m < 0) m = 0;
  else if (m >= kBins - 1) m = kBins - 1;
  return m;
}

__global__ __launch_bounds__(kCT) void k_sah_bins(const Item* __restrict__ items, const uint32_t* __restrict__ ctl,
                                                  uint32_t par, const uint32_t* __restrict__ chunk_base,
                                                  const uint32_t* __restrict__ chunk_item,
                                                  const uint32_t* __restrict__ indices, const Box* __restrict__ boxes,
                                                  int* __restrict__ acc) {
  __shared__ int s_bin[kBins * 7];
  const ChunkRange r = chunk_range(items, ctl, par, chunk_base, chunk_item, blockIdx.x);
  if (!r.ok) return;
  for (int w = threadIdx.x; w < kBins * 7; w += kCT) {
    int k = w % 7;
    s_bin[w] = k == 0 ? 0 : (k < 4 ? f2o(RK_HUGE) : f2o(-RK_HUGE));
  }
  __syncthreads();
  int* a = acc + (size_t)r.slot * kAccWords;
  int cb[6];
  acc_bounds(a, cb);
  int axis;
  float scale, bmin0;
  item_axis(cb, axis, scale, bmin0);
  Box bb[kPer];
  bool in[kPer];
  chunk_boxes(r, indices, boxes, bb, in);
#pragma unroll
  for (int u = 0; u < kPer; ++u) {
    if (!in[u]) continue;
    const Box& b = bb[u];
    int m = bin_of(b, axis, scale, bmin0);
    int* s = s_bin + m * 7;
    atomicAdd(s, 1);
    atomicMin(s + 1, f2o(b.mn.x)); atomicMin(s + 2, f2o(b.mn.y)); atomicMin(s + 3, f2o(b.mn.z));
    atomicMax(s + 4, f2o(b.mx.x)); atomicMax(s + 5, f2o(b.mx.y)); atomicMax(s + 6, f2o(b.mx.z));
  }
  __syncthreads();
  int* slot_bins = a + kBinBase + (blockIdx.x % kBinSlots) * kBinWords;
  for (int w = threadIdx.x; w < kBins * 7; w += kCT) {
    int k = w % 7, v = s_bin[w];
    int* g = slot_bins + w;
    if (k == 0) { if (v) atomicAdd(g, v); }
    else if (k < 4) { if (v != f2o(RK_HUGE)) atomicMin(g, v); }
    else { if (v != f2o(-RK_HUGE)) atomicMax(g, v); }
  }
}

__device__ __forceinline__ Node make_node(uint32_t axis, uint32_t child, const float* bmin, const float* bmax) {
  Node n;
  n.w0 = (axis & 3u) << 1;
  n.w1 = child;
  for (int a = 0; a < 3; ++a) { n.bmin[a] = bmin[a]; n.bmax[a] = bmax[a]; }
  return n;
}
__device__ __forceinline__ Node make_leaf(uint32_t first, uint32_t count, const float* bmin, const float* bmax) {
  Node n;
  n.w0 = 1u | (count << 1);
  n.w1 = first;
  for (int a = 0; a < 3; ++a) { n.bmin[a] = bmin[a]; n.bmax[a] = bmax[a]; }
  return n;
}

// Items per workgroup of k_sah_decide (two per wave, one lane per bin) and
// per workgroup of k_sah_emit.
constexpr uint32_t kDecideItems = 8;
constexpr uint32_t kEmitItems = kEmitGroup;
static_assert(kEmitItems % kDecideItems == 0, "a decide workgroup's items lie in one emit group");

// The reference's sweep over one item's bins (:424-474) and children boxes
// (:501-530), one lane per bin: lane k of each 32-lane half holds bin k
// (count, ordered-int box) of its half's item, and every lane of the wave
// calls this. The left sweep's running box and count after bin k are the
// inclusive prefix at lane k, and the right sweep's after bin k the inclusive
// suffix; min, max and integer sums give the same values in any order, and
// each lane then evaluates the reference's area and cost expressions on those
// values. The reference keeps the first strict minimum of its sweep from bin
// 31 down, i.e. the largest split index among the minimal costs. The
// children's boxes are the prefix at the chosen split and the suffix after it.
struct Sweep {
  int m_min;  // -1: no finite cost below +huge
  float T_min;
  int lcount;
  float child[12];  // left min xyz, left max xyz, right min xyz, right max xyz
};

__device__ __forceinline__ Sweep sweep_bins(int cnt, const float (&bmn)[3], const float (&bmx)[3], float pa,
                                            uint32_t k) {
  int lc = cnt, rc = cnt;
  float lmn[3], lmx[3], rmn[3], rmx[3];
  for (int q = 0; q < 3; ++q) { lmn[q] = rmn[q] = bmn[q]; lmx[q] = rmx[q] = bmx[q]; }
#pragma unroll
  for (int off = 1; off < 32; off <<= 1) {
    const int ylc = __shfl_up(lc, off, 32), yrc = __shfl_down(rc, off, 32);
    float ylmn[3], ylmx[3], yrmn[3], yrmx[3];
    for (int q = 0; q < 3; ++q) {
      ylmn[q] = __shfl_up(lmn[q], off, 32); ylmx[q] = __shfl_up(lmx[q], off, 32);
      yrmn[q] = __shfl_down(rmn[q], off, 32); yrmx[q] = __shfl_down(rmx[q], off, 32);
    }
    if (k >= (uint32_t)off) {
      lc += ylc;
      for (int q = 0; q < 3; ++q) { lmn[q] = rk::fmn(lmn[q], ylmn[q]); lmx[q] = rk::fmx(lmx[q], ylmx[q]); }
    }
    if (k + off < 32) {
      rc += yrc;
      for (int q = 0; q < 3; ++q) { rmn[q] = rk::fmn(rmn[q], yrmn[q]); rmx[q] = rk::fmx(rmx[q], yrmx[q]); }
    }
  }
  // cost of the split after bin k (k <= 30): prefix at k, suffix at k + 1
  const float rarea = 0.f + box_area(rmn[0], rmn[1], rmn[2], rmx[0], rmx[1], rmx[2]);
  const float rarea_next = __shfl_down(rarea, 1, 32);
  const int rc_next = __shfl_down(rc, 1, 32);
  const float larea = 0.f + box_area(lmn[0], lmn[1], lmn[2], lmx[0], lmx[1], lmx[2]);
  const int lcount = 0 + lc;
  const float T = 1.0f * 2.f + 0.125f * ((larea / pa) * (float)lcount + (rarea_next / pa) * (float)(0 + rc_next));
  float bt = T;
  int bj = (k < 31 && T < RK_HUGE) ? (int)k : -1;  // NaN and +huge never win a strict <
#pragma unroll
  for (int off = 16; off > 0; off >>= 1) {
    const float ot = __shfl_xor(bt, off, 32);
    const int oj = __shfl_xor(bj, off, 32);
    if (oj >= 0 && (bj < 0 || ot < bt || (ot == bt && oj > bj))) { bt = ot; bj = oj; }
  }
  Sweep r;
  r.m_min = bj;
  r.T_min = bj >= 0 ? bt : RK_HUGE;
  const int src_l = bj >= 0 ? bj : 0, src_r = bj >= 0 ? bj + 1 : 0;
  for (int q = 0; q < 3; ++q) {
    r.child[q] = __shfl(lmn[q], src_l, 32);
    r.child[3 + q] = __shfl(lmx[q], src_l, 32);
    r.child[6 + q] = __shfl(rmn[q], src_r, 32);
    r.child[9 + q] = __shfl(rmx[q], src_r, 32);
  }
  r.lcount = __shfl(lc, src_l, 32);
  return r;
}

// The leaf test (:479-496) on a sweep's result; an item with no finite cut
// that is too large for a leaf becomes one anyway and is flagged in the flags word
// (when ctl is given).
__device__ __forceinline__ Decision make_decision(uint32_t size, int axis, float scale, float bmin0, const Sweep& sw,
                                                  uint32_t* __restrict__ ctl) {
  Decision d;
  d.axis = axis; d.scale = scale; d.bmin0 = bmin0; d.m_min = -1; d.lcount = 0;
  d.pad[0] = d.pad[1] = d.pad[2] = 0;
  for (int q = 0; q < 12; ++q) d.child[q] = 0.f;
  bool leaf = size < kLeafMax && sw.T_min > (float)(size + 1) * 0.125f;
  if (!leaf && sw.m_min < 0) {
    leaf = true;
    if (ctl) atomicOr(ctl + kCtlFlags, (uint32_t)RK_FLAG_SAH_NO_CUT);
  }
  if (!leaf) {
    d.m_min = sw.m_min;
    d.lcount = (uint32_t)sw.lcount;
    for (int q = 0; q < 12; ++q) d.child[q] = sw.child[q];
  }
  return d;
}

__device__ __forceinline__ Decision leaf_decision() {
  Decision d;
  d.axis = 0; d.scale = 0.f; d.bmin0 = 0.f; d.m_min = -1; d.lcount = 0;
  d.pad[0] = d.pad[1] = d.pad[2] = 0;
  for (int q = 0; q < 12; ++q) d.child[q] = 0.f;
  return d;
}

// Items of more than kChunk faces, decided from their chunks' accumulators;
// k_sah_small decides the rest. One 32-lane half-wave per item.
__global__ __launch_bounds__(256) void k_sah_decide(const Item* __restrict__ items, uint32_t* __restrict__ ctl,
                                                    uint32_t par, const int* __restrict__ acc,
                                                    Decision* __restrict__ dec,
                                                    uint32_t* __restrict__ group_internal) {
  __shared__ uint32_t s_internal;
  const uint32_t m = ctl[kCtlLevel(par) + kLvM];
  const uint32_t k = threadIdx.x & 31;
  const uint32_t i = blockIdx.x * kDecideItems + (threadIdx.x >> 5);
  if (threadIdx.x == 0) s_internal = 0;
  __syncthreads();
  Item it;
  uint32_t size = 0;
  if (i < m) {
    it = items[i];
    size = it.hi - it.lo;
  }
  const bool split_test = i < m && size > kChunk;
  int axis = 0, cnt = 0;
  float scale = 0.f, bmin0 = 0.f, pa = 0.f;
  float bmn[3] = {RK_HUGE, RK_HUGE, RK_HUGE}, bmx[3] = {-RK_HUGE, -RK_HUGE, -RK_HUGE};
  if (split_test) {
    const int* a = acc + (size_t)it.pad * kAccWords;
    int cb[6];
    acc_bounds(a, cb);
    item_axis(cb, axis, scale, bmin0);
    int bi[7];
    for (int q = 0; q < 7; ++q) bi[q] = a[kBinBase + k * 7 + q];
    for (int v = 1; v < kBinSlots; ++v) {
      const int* b = a + kBinBase + v * kBinWords + k * 7;
      bi[0] += b[0];
      for (int q = 1; q < 4; ++q) bi[q] = min(bi[q], b[q]);
      for (int q = 4; q < 7; ++q) bi[q] = max(bi[q], b[q]);
    }
    cnt = bi[0];
    for (int q = 0; q < 3; ++q) { bmn[q] = o2f(bi[1 + q]); bmx[q] = o2f(bi[4 + q]); }
    pa = box_area(it.bmin[0], it.bmin[1], it.bmin[2], it.bmax[0], it.bmax[1], it.bmax[2]);
  }
  const Sweep sw = sweep_bins(cnt, bmn, bmx, pa, k);
  if (k == 0 && split_test) {
    const Decision d = make_decision(size, axis, scale, bmin0, sw, ctl);
    if (d.m_min >= 0) atomicAdd(&s_internal, 1u);
    dec[i] = d;
  }
  __syncthreads();
  // kEmitItems is a multiple of kDecideItems: the workgroup's items share a group
  if (threadIdx.x == 0 && s_internal)
    atomicAdd(group_internal + blockIdx.x * kDecideItems / kEmitItems, s_internal);
}

// Items of kLeafMin + 1 to kChunk faces, one wave each, start to end: centroid
// bounds, bins, the sweep and the stable partition, with the item's face ids
// held in LDS so that the partition writes back in place. Every step reduces
// or scans the same values as the chunk kernels, so the decision is the same.
// Items of at most kLeafMin faces are leaves (:239-255).
constexpr uint32_t kSmallWaves = 4;
constexpr int kSmallPer = 4;  // faces per lane in flight
static_assert(kEmitItems % kSmallWaves == 0, "a small-item workgroup's items lie in one emit group");
__global__ __launch_bounds__(64 * kSmallWaves) void k_sah_small(const Item* __restrict__ items,
                                                                uint32_t* __restrict__ ctl, uint32_t par,
                                                                uint32_t* __restrict__ indices,
                                                                const Box* __restrict__ boxes,
                                                                Decision* __restrict__ dec,
                                                                uint32_t* __restrict__ group_internal) {
  __shared__ uint32_t s_ids[kSmallWaves][kChunk];
  __shared__ uint8_t s_binof[kSmallWaves][kChunk];
  __shared__ int s_bins[kSmallWaves][kBins * 7];
  __shared__ uint32_t s_internal;
  const uint32_t m = ctl[kCtlLevel(par) + kLvM];
  const uint32_t w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const uint32_t i = blockIdx.x * kSmallWaves + w;
  if (threadIdx.x == 0) s_internal = 0;
  __syncthreads();
  Item it;
  uint32_t size = 0;
  if (i < m) {
    it = items[i];
    size = it.hi - it.lo;
  }
  if (i < m && size <= kLeafMin) {  // wave-uniform
    if (lane == 0) dec[i] = leaf_decision();
  } else if (size > kLeafMin && size <= kChunk) {  // wave-uniform
    uint32_t* ids = s_ids[w];
    uint8_t* binof = s_binof[w];
    int* bins = s_bins[w];
    // centroid bounds, ordered ints as in k_sah_bounds' atomics. Faces go
    // kSmallPer per lane at a time, every gather in flight at once; the first
    // 64 * kSmallPer boxes stay in registers for the binning pass.
    int cmn[3] = {f2o(RK_HUGE), f2o(RK_HUGE), f2o(RK_HUGE)}, cmx[3] = {f2o(-RK_HUGE), f2o(-RK_HUGE), f2o(-RK_HUGE)};
    Box keep[kSmallPer];
    for (uint32_t l0 = 0; l0 < size; l0 += 64 * kSmallPer) {
      uint32_t id[kSmallPer];
      Box b[kSmallPer];
#pragma unroll
      for (int u = 0; u < kSmallPer; ++u) {
        const uint32_t l = l0 + lane + 64u * u;
        id[u] = l < size ? indices[it.lo + l] : 0u;
      }
#pragma unroll
      for (int u = 0; u < kSmallPer; ++u)
        if (l0 + lane + 64u * u < size) b[u] = rk::load_box(boxes + id[u]);
#pragma unroll
      for (int u = 0; u < kSmallPer; ++u) {
        const uint32_t l = l0 + lane + 64u * u;
        if (l >= size) continue;
        ids[l] = id[u];
        if (l0 == 0) keep[u] = b[u];
        for (int a = 0; a < 3; ++a) {
          const int c = f2o(center_axis(b[u], a));
          cmn[a] = min(cmn[a], c);
          cmx[a] = max(cmx[a], c);
        }
      }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
      for (int a = 0; a < 3; ++a) {
        cmn[a] = min(cmn[a], __shfl_xor(cmn[a], off, 64));
        cmx[a] = max(cmx[a], __shfl_xor(cmx[a], off, 64));
      }
    int cb[6] = {cmn[0], cmn[1], cmn[2], cmx[0], cmx[1], cmx[2]};
    int axis;
    float scale, bmin0;
    item_axis(cb, axis, scale, bmin0);
    for (int q = lane; q < kBins * 7; q += 64) {
      const int kq = q % 7;
      bins[q] = kq == 0 ? 0 : (kq < 4 ? f2o(RK_HUGE) : f2o(-RK_HUGE));
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    for (uint32_t l0 = 0; l0 < size; l0 += 64 * kSmallPer) {
      Box b[kSmallPer];
#pragma unroll
      for (int u = 0; u < kSmallPer; ++u) {
        const uint32_t l = l0 + lane + 64u * u;
        if (l < size) b[u] = l0 == 0 ? keep[u] : rk::load_box(boxes + ids[l]);
      }
#pragma unroll
      for (int u = 0; u < kSmallPer; ++u) {
        const uint32_t l = l0 + lane + 64u * u;
        if (l >= size) continue;
        const int bn = bin_of(b[u], axis, scale, bmin0);
        binof[l] = (uint8_t)bn;
        int* sb = bins + bn * 7;
        atomicAdd(sb, 1);
        atomicMin(sb + 1, f2o(b[u].mn.x)); atomicMin(sb + 2, f2o(b[u].mn.y)); atomicMin(sb + 3, f2o(b[u].mn.z));
        atomicMax(sb + 4, f2o(b[u].mx.x)); atomicMax(sb + 5, f2o(b[u].mx.y)); atomicMax(sb + 6, f2o(b[u].mx.z));
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    const uint32_t k = lane & 31;
    const int* bk = bins + k * 7;
    const int cnt = bk[0];
    float bmn[3], bmx[3];
    for (int q = 0; q < 3; ++q) { bmn[q] = o2f(bk[1 + q]); bmx[q] = o2f(bk[4 + q]); }
    const float pa = box_area(it.bmin[0], it.bmin[1], it.bmin[2], it.bmax[0], it.bmax[1], it.bmax[2]);
    const Sweep sw = sweep_bins(cnt, bmn, bmx, pa, k);
    const Decision d = make_decision(size, axis, scale, bmin0, sw, lane == 0 ? ctl : nullptr);
    if (d.m_min >= 0) {
      // stable partition of the item's ids, bin <= m_min first (:534-610)
      uint32_t left_before = 0, right_before = 0;
      for (uint32_t l0 = 0; l0 < size; l0 += 64) {
        const uint32_t l = l0 + lane;
        uint32_t id = 0;
        bool is_left = false;
        if (l < size) {
          id = ids[l];
          is_left = (int)binof[l] <= d.m_min;
        }
        const uint64_t lm = __ballot(l < size && is_left), rm = __ballot(l < size && !is_left);
        const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
        if (l < size) {
          const uint32_t out = is_left ? it.lo + left_before + (uint32_t)__popcll(lm & below)
                                       : it.lo + d.lcount + right_before + (uint32_t)__popcll(rm & below);
          indices[out] = id;
        }
        left_before += (uint32_t)__popcll(lm);
        right_before += (uint32_t)__popcll(rm);
      }
      if (lane == 0) atomicAdd(&s_internal, 1u);
    }
    if (lane == 0) dec[i] = d;
  }
  __syncthreads();
  // kEmitItems is a multiple of kSmallWaves: the workgroup's items share a group
  if (threadIdx.x == 0 && s_internal)
    atomicAdd(group_internal + blockIdx.x * kSmallWaves / kEmitItems, s_internal);
}

// Node records and the next level's items, in item order: the k-th internal
// item of the level takes children (count + 2k, count + 2k + 1). Workgroup g
// handles items [g, g + 1) * kEmitItems and starts after the internal items
// counted in the groups before it. The grid may hold more workgroups than the
// level has groups; the one with the last group (workgroup 0 when the level
// is empty) writes the next level's item and node counts.
__global__ __launch_bounds__(1024) void k_sah_emit(const Item* __restrict__ items, uint32_t* __restrict__ ctl,
                                                   uint32_t par, const Decision* __restrict__ dec,
                                                   const uint32_t* __restrict__ group_internal,
                                                   Item* __restrict__ next, Node* __restrict__ nodes,
                                                   uint32_t nodes_cap) {
  __shared__ uint32_t s_tmp[16];
  const uint32_t m = ctl[kCtlLevel(par) + kLvM];
  const uint32_t count0 = ctl[kCtlLevel(par) + kLvCount];
  const uint32_t g = blockIdx.x;
  uint32_t before = 0, before_total;
  for (uint32_t h = threadIdx.x; h < g; h += 1024) before += group_internal[h];
  block_excl_scan(before, s_tmp, &before_total, 1024);
  const uint32_t i = g * kEmitItems + threadIdx.x;
  uint32_t internal = 0;
  Item it;
  Decision d;
  if (i < m) {
    it = items[i];
    d = dec[i];
    internal = d.m_min >= 0 ? 1u : 0u;
  }
  uint32_t total;
  const uint32_t ex = block_excl_scan(internal, s_tmp, &total, 1024);
  // children of more than kChunk faces, for the next level's launches
  uint32_t big = 0, big_total;
  if (internal) big = (d.lcount > kChunk ? 1u : 0u) + (it.hi - it.lo - d.lcount > kChunk ? 1u : 0u);
  block_excl_scan(big, s_tmp, &big_total, 1024);
  uint32_t* nx = ctl + kCtlLevel(par ^ 1u);
  if (threadIdx.x == 0 && big_total) atomicAdd(nx + kLvBig, big_total);
  if (i < m) {
    if (internal) {
      const uint32_t k = before_total + ex;
      const uint32_t c = count0 + 2 * k;
      if (it.node < nodes_cap) nodes[it.node] = make_node((uint32_t)d.axis, c, it.bmin, it.bmax);
      Item li, ri;
      li.node = c; li.lo = it.lo; li.hi = it.lo + d.lcount; li.pad = 0;
      ri.node = c + 1; ri.lo = it.lo + d.lcount; ri.hi = it.hi; ri.pad = 0;
      for (int q = 0; q < 3; ++q) {
        li.bmin[q] = d.child[q]; li.bmax[q] = d.child[3 + q];
        ri.bmin[q] = d.child[6 + q]; ri.bmax[q] = d.child[9 + q];
      }
      li.pad2[0] = li.pad2[1] = ri.pad2[0] = ri.pad2[1] = 0.f;
      next[2 * k] = li;
      next[2 * k + 1] = ri;
    } else if (it.node < nodes_cap) {
      nodes[it.node] = make_leaf(it.lo, it.hi - it.lo, it.bmin, it.bmax);
    }
  }
  if (threadIdx.x == 0 && g == (m ? (m - 1) / kEmitItems : 0u)) {
    const uint32_t all = before_total + total;
    nx[kLvM] = 2 * all;
    nx[kLvCount] = count0 + 2 * all;
    if (count0 + 2 * all > nodes_cap) atomicOr(ctl + kCtlFlags, 0x20u);
  }
}

__global__ __launch_bounds__(kCT) void k_sah_count(const Item* __restrict__ items, const uint32_t* __restrict__ ctl,
                                                   uint32_t par, const uint32_t* __restrict__ chunk_item,
                                                   const uint32_t* __restrict__ chunk_base,
                                                   const uint32_t* __restrict__ indices, const Box* __restrict__ boxes,
                                                   const Decision* __restrict__ dec, uint32_t* __restrict__ chunk_left) {
  __shared__ uint32_t s_cnt;
  const ChunkRange r = chunk_range(items, ctl, par, chunk_base, chunk_item, blockIdx.x);
  if (!r.ok) return;
  const Decision d = dec[r.item];
  if (threadIdx.x == 0) s_cnt = 0;
  __syncthreads();
  if (d.m_min >= 0) {
    Box bb[kPer];
    bool in[kPer];
    chunk_boxes(r, indices, boxes, bb, in);
    uint32_t c = 0;
#pragma unroll
    for (int u = 0; u < kPer; ++u)
      if (in[u]) c += bin_of(bb[u], d.axis, d.scale, d.bmin0) <= d.m_min ? 1u : 0u;
    atomicAdd(&s_cnt, c);
  }
  __syncthreads();
  if (threadIdx.x == 0) chunk_left[blockIdx.x] = s_cnt;
}

// Exclusive scan of chunk_left within each item (an item's chunks are
// contiguous), one wave per item.
__global__ __launch_bounds__(256) void k_sah_offsets(const Item* __restrict__ items, const uint32_t* __restrict__ ctl,
                                                     uint32_t par, const uint32_t* __restrict__ chunk_base,
                                                     const uint32_t* __restrict__ chunk_left,
                                                     uint32_t* __restrict__ chunk_loff) {
  const uint32_t m = ctl[kCtlLevel(par) + kLvM];
  const uint32_t i = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (i >= m) return;  // wave-uniform
  const uint32_t size = items[i].hi - items[i].lo;
  if (size <= kChunk) return;  // no chunks
  const uint32_t b = chunk_base[i], nch = (size + kChunk - 1) / kChunk;
  uint32_t run = 0;
  for (uint32_t c0 = 0; c0 < nch; c0 += 64) {
    const uint32_t c = c0 + lane;
    const uint32_t v = c < nch ? chunk_left[b + c] : 0u;
    uint32_t incl = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const uint32_t y = __shfl_up(incl, off, 64);
      if (lane >= (uint32_t)off) incl += y;
    }
    if (c < nch) chunk_loff[b + c] = run + incl - v;
    run += __shfl(incl, 63, 64);
  }
}

__global__ __launch_bounds__(kCT) void k_sah_scatter(const Item* __restrict__ items, const uint32_t* __restrict__ ctl,
                                                     uint32_t par, const uint32_t* __restrict__ chunk_item,
                                                     const uint32_t* __restrict__ chunk_base,
                                                     const uint32_t* __restrict__ indices,
                                                     const Box* __restrict__ boxes, const Decision* __restrict__ dec,
                                                     const uint32_t* __restrict__ chunk_loff,
                                                     uint32_t* __restrict__ tmp) {
  __shared__ uint32_t s_tmp[8];
  const ChunkRange r = chunk_range(items, ctl, par, chunk_base, chunk_item, blockIdx.x);
  if (!r.ok) return;
  const Decision d = dec[r.item];
  if (d.m_min < 0) return;
  const Item it = items[r.item];
  const uint32_t loff = chunk_loff[blockIdx.x];
  const uint32_t chunk_start = r.lo - it.lo;  // elements of this item before the chunk
  uint32_t left_before = 0;                    // left elements of this chunk before the current step
  for (uint32_t s = r.lo; s < r.hi; s += kCT) {
    const uint32_t l = s + threadIdx.x;
    uint32_t id = 0, is_left = 0;
    const bool in = l < r.hi;
    if (in) {
      id = indices[l];
      is_left = bin_of(rk::load_box(boxes + id), d.axis, d.scale, d.bmin0) <= d.m_min ? 1u : 0u;
    }
    uint32_t total;
    const uint32_t ex = block_excl_scan(is_left, s_tmp, &total, kCT);
    if (in) {
      const uint32_t pos_in_item = chunk_start + (l - r.lo);
      const uint32_t lrank = loff + left_before + ex;            // left elements before this one
      const uint32_t out = is_left ? (it.lo + lrank) : (it.lo + d.lcount + (pos_in_item - lrank));
      tmp[out] = id;
    }
    left_before += total;
  }
}

__global__ __launch_bounds__(kCT) void k_sah_copyback(const Item* __restrict__ items, const uint32_t* __restrict__ ctl,
                                                      uint32_t par, const uint32_t* __restrict__ chunk_item,
                                                      const uint32_t* __restrict__ chunk_base,
                                                      const Decision* __restrict__ dec,
                                                      const uint32_t* __restrict__ tmp, uint32_t* __restrict__ indices) {
  const ChunkRange r = chunk_range(items, ctl, par, chunk_base, chunk_item, blockIdx.x);
  if (!r.ok) return;
  if (dec[r.item].m_min < 0) return;
  for (uint32_t l = r.lo + threadIdx.x; l < r.hi; l += kCT) indices[l] = tmp[l];
}

// Hybrid entry (kernel_initStack, :647-666): every leaf of the existing tree
// becomes an item, in ascending node order, with the leaf's box.
__global__ __launch_bounds__(1024) void k_sah_init_from_leaves(const Node* __restrict__ nodes, uint32_t existing,
                                                               Item* __restrict__ items, uint32_t* __restrict__ ctl) {
  __shared__ uint32_t s_tmp[16];
  __shared__ uint32_t s_carry, s_big;
  if (threadIdx.x == 0) s_carry = s_big = 0;
  __syncthreads();
  for (uint32_t c0 = 0; c0 < existing; c0 += 1024) {
    const uint32_t k = c0 + threadIdx.x;
    uint32_t leaf = 0;
    Node nd;
    if (k < existing) {
      nd = nodes[k];
      leaf = rk::node_is_leaf(nd.w0);
    }
    uint32_t total, big_total;
    const uint32_t ex = block_excl_scan(leaf, s_tmp, &total, 1024);
    block_excl_scan(leaf && rk::leaf_count(nd.w0) > kChunk ? 1u : 0u, s_tmp, &big_total, 1024);
    if (leaf) {
      Item it;
      it.node = k; it.lo = nd.w1; it.hi = nd.w1 + rk::leaf_count(nd.w0); it.pad = 0;
      for (int q = 0; q < 3; ++q) { it.bmin[q] = nd.bmin[q]; it.bmax[q] = nd.bmax[q]; }
      it.pad2[0] = it.pad2[1] = 0.f;
      items[s_carry + ex] = it;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      s_carry += total;
      s_big += big_total;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    uint32_t* lv = ctl + kCtlLevel(0);
    lv[kLvM] = s_carry;
    lv[kLvBig] = s_big;
    lv[kLvCount] = existing;
    ctl[kCtlFlags] = 0;
  }
}

__global__ void k_sah_init_root(Item* __restrict__ items, uint32_t* __restrict__ ctl, uint32_t n, float4 rmin,
                                float4 rmax) {
  Item it;
  it.node = 0; it.lo = 0; it.hi = n; it.pad = 0;
  it.bmin[0] = rmin.x; it.bmin[1] = rmin.y; it.bmin[2] = rmin.z;
  it.bmax[0] = rmax.x; it.bmax[1] = rmax.y; it.bmax[2] = rmax.z;
  it.pad2[0] = it.pad2[1] = 0.f;
  items[0] = it;
  uint32_t* lv = ctl + kCtlLevel(0);
  lv[kLvM] = 1;
  lv[kLvBig] = n > kChunk ? 1u : 0u;
  lv[kLvCount] = 1;
  ctl[kCtlFlags] = 0;
}

// par: the parity of the first level that did not run, whose node count is the total.
__global__ void k_sah_finish(const uint32_t* __restrict__ ctl, uint32_t par, uint32_t* __restrict__ info) {
  info[RK_INFO_NODE_COUNT] = ctl[kCtlLevel(par) + kLvCount];
  info[RK_INFO_FLAGS] |= ctl[kCtlFlags];
}

__global__ __launch_bounds__(256) void k_boxes_iota(const uint4* __restrict__ faces, const float4* __restrict__ verts,
                                                    uint32_t n, Box* __restrict__ boxes, uint32_t* __restrict__ idx) {
  uint32_t k = blockIdx.x * 256 + threadIdx.x;
  if (k >= n) return;
  uint4 f = faces[k];
  float4 v0 = verts[f.x], v1 = verts[f.y], v2 = verts[f.z];
  float4* o = reinterpret_cast<float4*>(boxes + k);
  o[0] = rk::min4(rk::min4(v0, v1), v2);
  o[1] = rk::max4(rk::max4(v0, v1), v2);
  idx[k] = k;
}

__global__ __launch_bounds__(256) void k_gather_faces2(const uint4* __restrict__ src, const uint32_t* __restrict__ idx,
                                                       uint4* __restrict__ dst, uint32_t n) {
  uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) dst[i] = src[idx[i]];
}
__global__ __launch_bounds__(256) void k_copy_faces2(const uint4* __restrict__ src, uint4* __restrict__ dst, uint32_t n) {
  uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

}  // namespace

extern "C" size_t rk_build_sah_workspace_size(uint32_t face_count, uint32_t existing_nodes) {
  size_t n = face_count ? face_count : 1;
  size_t m = n + existing_nodes + 1;   // items per level
  size_t big = n / kChunk + 1;         // items of more than kChunk faces per level
  size_t T = n / kChunk + big + 1;     // chunks per level: their ceil(size / kChunk)
  size_t b = 0;
  b += rk_align_up(m * sizeof(Item)) * 2;
  b += rk_align_up(m * 4);                         // chunk_base
  b += rk_align_up(big * kAccWords * 4);           // accumulators
  b += rk_align_up(m * sizeof(Decision));
  b += rk_align_up(T * 4) * 3;                     // chunk_item, chunk_left, chunk_loff
  b += rk_align_up(n * 4);                         // partition temp
  b += rk_align_up(n * 16);                        // face reorder temp
  b += rk_align_up(kCtlWords * 4);
  b += rk_align_up((m / kEmitGroup + 1) * 4);      // internal items per emit group
  return b + 4096;
}

// void CUDA_buildMeshTreeSAH(CUDAMesh*, CUDAMeshTree*), CUDAMeshTree.cpp:34-38 /
// CUDA_buildMeshTreeSAH.cu:672-789. existing_nodes == 0: boxes and indices are
// (re)computed and the root is the mesh hitbox; existing_nodes > 0 (Hybrid):
// boxes/indices come from the LBVH pass and every leaf among nodes[0,
// existing_nodes) is refined. Faces are reordered by the final indices.
// Synchronises 'stream' once per tree level once a level can hold more than
// min(4096, n/256) items (the level's counts size the next launches); the
// info words and the face reorder are queued on 'stream' when it returns.
extern "C" rk_status rk_build_sah(rk_stream stream, rk_uint4* faces, const rk_float4* vertices, uint32_t face_count,
                                  const rk_aabox* hitbox, rk_aabox* boxes, uint32_t* indices, rk_bvnode* nodes,
                                  uint32_t existing_nodes, uint32_t nodes_capacity, uint32_t* info, void* ws,
                                  size_t ws_bytes) {
  const uint32_t n = face_count;
  if (n == 0 || !faces || !vertices || !hitbox || !boxes || !indices || !nodes || !info) return RK_ERROR_INVALID_VALUE;
  if (n >= (1u << 30)) return RK_ERROR_UNSUPPORTED;
  if ((size_t)nodes_capacity < (size_t)existing_nodes + 2 * (size_t)n - 1) return RK_ERROR_INVALID_VALUE;
  hipStream_t st = rk_hip_stream(stream);
  const size_t m_cap = (size_t)n + existing_nodes + 1;
  const size_t big_cap = n / kChunk + 1;
  const size_t T_cap = n / kChunk + big_cap + 1;
  RkArena ar(ws, ws_bytes);
  Item* items_a = ar.take<Item>(m_cap);
  Item* items_b = ar.take<Item>(m_cap);
  uint32_t* chunk_base = ar.take<uint32_t>(m_cap);
  int* acc = ar.take<int>(big_cap * kAccWords);
  Decision* dec = ar.take<Decision>(m_cap);
  uint32_t* chunk_item = ar.take<uint32_t>(T_cap);
  uint32_t* chunk_left = ar.take<uint32_t>(T_cap);
  uint32_t* chunk_loff = ar.take<uint32_t>(T_cap);
  uint32_t* tmp = ar.take<uint32_t>(n);
  uint4* ftmp = ar.take<uint4>(n);
  uint32_t* ctl = ar.take<uint32_t>(kCtlWords);
  uint32_t* group_internal = ar.take<uint32_t>(m_cap / kEmitGroup + 1);
  if (!items_a || !items_b || !chunk_base || !acc || !dec || !chunk_item || !chunk_left || !chunk_loff || !tmp ||
      !ftmp || !ctl || !group_internal)
    return RK_ERROR_WORKSPACE;
  rk_status s;
  if ((s = rk_fill_u32(stream, ctl, kCtlWords, 0u)) != RK_SUCCESS) return s;
  if ((s = rk_fill_u32(stream, info, RK_INFO_WORDS, 0u)) != RK_SUCCESS) return s;
  if (existing_nodes == 0) {
    hipLaunchKernelGGL(k_boxes_iota, dim3((n + 255) / 256), dim3(256), 0, st, reinterpret_cast<const uint4*>(faces),
                       reinterpret_cast<const float4*>(vertices), n, reinterpret_cast<Box*>(boxes), indices);
    RK_CHECK_LAUNCH();
    float4 rmin = make_float4(hitbox->min.x, hitbox->min.y, hitbox->min.z, hitbox->min.w);
    float4 rmax = make_float4(hitbox->max.x, hitbox->max.y, hitbox->max.z, hitbox->max.w);
    hipLaunchKernelGGL(k_sah_init_root, dim3(1), dim3(1), 0, st, items_a, ctl, n, rmin, rmax);
  } else {
    hipLaunchKernelGGL(k_sah_init_from_leaves, dim3(1), dim3(1024), 0, st, reinterpret_cast<const Node*>(nodes),
                       existing_nodes, items_a, ctl);
  }
  RK_CHECK_LAUNCH();
  // this level's item count and items of more than kChunk faces, or upper
  // bounds of them while !exact
  uint32_t lvh[2] = {0u, 0u};
  RK_HIP(hipMemcpyAsync(lvh, ctl + kCtlLevel(0), sizeof(lvh), hipMemcpyDeviceToHost, st));
  RK_HIP(hipStreamSynchronize(st));
  uint32_t m = lvh[0], big = lvh[1];
  bool exact = true;
  // The first levels, while their item count is bounded by a few thousand,
  // run without reading it back: a level has at most twice the items of the
  // one before, and every kernel reads the exact counts on the device, so a
  // grid sized by a bound is only larger than needed. From the first level
  // beyond that bound on, the counts are read back after every level to size
  // the next level's grids (and to skip the chunk kernels when no item has
  // more than kChunk faces), and an empty level ends the build.
  const uint32_t async_items = std::min<uint32_t>(4096u, n / 256);
  const uint32_t big_bound = n / (kChunk + 1);
  bool shallow = true;
  Item* cur = items_a;
  Item* nxt = items_b;
  const Box* bx = reinterpret_cast<const Box*>(boxes);
  Node* nd = reinterpret_cast<Node*>(nodes);
  uint32_t level = 0;
  for (; !(exact && m == 0); ++level) {
    if (level > 4096) return RK_ERROR_LAUNCH;
    const uint32_t par = level & 1u;
    const unsigned grid = (unsigned)(n / kChunk + std::min(big, big_bound) + 1);  // chunks: see the workspace
    hipLaunchKernelGGL(k_sah_prep, dim3(m ? (m + kPrepItems - 1) / kPrepItems : 1u), dim3(64 * kPrepItems), 0, st,
                       cur, ctl, par, chunk_base, chunk_item, acc, group_internal);
    if (big) {
      hipLaunchKernelGGL(k_sah_bounds, dim3(grid), dim3(kCT), 0, st, cur, ctl, par, chunk_base, chunk_item, indices,
                         bx, acc);
      hipLaunchKernelGGL(k_sah_bins, dim3(grid), dim3(kCT), 0, st, cur, ctl, par, chunk_base, chunk_item, indices, bx,
                         acc);
      hipLaunchKernelGGL(k_sah_decide, dim3((m + kDecideItems - 1) / kDecideItems), dim3(256), 0, st, cur, ctl, par,
                         acc, dec, group_internal);
    }
    hipLaunchKernelGGL(k_sah_small, dim3((m + kSmallWaves - 1) / kSmallWaves), dim3(64 * kSmallWaves), 0, st, cur,
                       ctl, par, indices, bx, dec, group_internal);
    hipLaunchKernelGGL(k_sah_emit, dim3(m ? (m + kEmitItems - 1) / kEmitItems : 1u), dim3(1024), 0, st, cur, ctl,
                       par, dec, group_internal, nxt, nd, nodes_capacity);
    if (big) {
      hipLaunchKernelGGL(k_sah_count, dim3(grid), dim3(kCT), 0, st, cur, ctl, par, chunk_item, chunk_base, indices,
                         bx, dec, chunk_left);
      hipLaunchKernelGGL(k_sah_offsets, dim3((m + 3) / 4), dim3(256), 0, st, cur, ctl, par, chunk_base, chunk_left,
                         chunk_loff);
      hipLaunchKernelGGL(k_sah_scatter, dim3(grid), dim3(kCT), 0, st, cur, ctl, par, chunk_item, chunk_base, indices,
                         bx, dec, chunk_loff, tmp);
      hipLaunchKernelGGL(k_sah_copyback, dim3(grid), dim3(kCT), 0, st, cur, ctl, par, chunk_item, chunk_base, dec, tmp,
                         indices);
    }
    RK_CHECK_LAUNCH();
    const uint64_t bound = std::min<uint64_t>(2 * (uint64_t)m, m_cap);
    if (shallow && bound <= async_items) {
      m = (uint32_t)bound;
      big = (uint32_t)std::min<uint64_t>(bound, big_bound);
      exact = false;
    } else {
      shallow = false;
      RK_HIP(hipMemcpyAsync(lvh, ctl + kCtlLevel(par ^ 1u), sizeof(lvh), hipMemcpyDeviceToHost, st));
      RK_HIP(hipStreamSynchronize(st));
      m = lvh[0];
      big = lvh[1];
      exact = true;
    }
    std::swap(cur, nxt);
  }
  hipLaunchKernelGGL(k_sah_finish, dim3(1), dim3(1), 0, st, ctl, level & 1u, info);
  RK_CHECK_LAUNCH();
  // CUDA_reorderMeshFaces, :778-784
  uint4* f = reinterpret_cast<uint4*>(faces);
  hipLaunchKernelGGL(k_gather_faces2, dim3((n + 255) / 256), dim3(256), 0, st, f, indices, ftmp, n);
  hipLaunchKernelGGL(k_copy_faces2, dim3((n + 255) / 256), dim3(256), 0, st, ftmp, f, n);
  RK_CHECK_LAUNCH();
  return RK_SUCCESS;
}
