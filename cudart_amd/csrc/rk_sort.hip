// Stable LSD radix sort for gfx950, 8-bit digits, onesweep style:
//   1. one histogram kernel counts the digits of every pass in a single read;
//   2. one scatter kernel per pass. Each 256-thread workgroup takes a 4096-key
//      tile from a dynamic tile counter, ranks its keys stably in LDS with
//      wave64 ballots, publishes its per-digit counts, resolves its global
//      per-digit prefix by decoupled look-back over earlier tiles, reorders the
//      tile by digit in LDS and writes it out coalesced.
// Look-back waits are bounded (kSpinLimit); on expiry the kernel sets an error
// bit and finishes, so a fault in the protocol cannot hang the GPU.
//
// The reference's sorts (CUDA_radixSortGeneral.cu, CUDA_radixSort.cu,
// CUDA_bitonicSortGeneral.cu, CUDA_bitonicSort.cu, CUDA_bubbleSort512.cu)
// differ in tie order by size class; rk_radix_sort_general reproduces that
// order (see the size-class notes at the entry point).
#include "rk_device.h"

namespace {

constexpr int kWaves = 4;
constexpr int kThreads = 64 * kWaves;       // 256
constexpr int kItems = 16;                  // keys per lane
constexpr int kWaveTile = 64 * kItems;      // 1024
constexpr int kTile = kThreads * kItems;    // 4096
constexpr uint32_t kAgg = 1u << 30;
constexpr uint32_t kInc = 2u << 30;
constexpr uint32_t kCountMask = (1u << 30) - 1u;
constexpr uint32_t kSpinLimit = 1u << 20;
constexpr int kMaxPasses = 8;

// Exclusive scan of one value per thread over a 256-thread block.
__device__ __forceinline__ uint32_t block_excl_scan256(uint32_t v, uint32_t* s_tmp /*>= 4*/) {
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint32_t incl = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    uint32_t y = __shfl_up(incl, off, 64);
    if (lane >= (uint32_t)off) incl += y;
  }
  if (lane == 63) s_tmp[wave] = incl;
  __syncthreads();
  uint32_t wprefix = 0;
#pragma unroll
  for (int w = 0; w < kWaves; ++w) wprefix += (w < (int)wave) ? s_tmp[w] : 0u;
  __syncthreads();
  return wprefix + incl - v;
}

template <typename K>
__global__ __launch_bounds__(256) void k_rs_hist(const K* __restrict__ keys, uint32_t n, int passes, int first_pass,
                                                 uint32_t* __restrict__ hist) {
  __shared__ uint32_t s[kMaxPasses * 256];
  for (int i = threadIdx.x; i < passes * 256; i += 256) s[i] = 0;
  __syncthreads();
  size_t stride = (size_t)gridDim.x * 256;
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += stride) {
    K k = keys[i];
    for (int p = 0; p < passes; ++p) atomicAdd(&s[p * 256 + (uint32_t)((k >> (8 * (p + first_pass))) & 0xFF)], 1u);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < passes * 256; i += 256) {
    uint32_t c = s[i];
    if (c) atomicAdd(&hist[i], c);
  }
}

template <typename K, bool kVals>
__global__ __launch_bounds__(256) void k_rs_scatter(const K* __restrict__ kin, const uint32_t* __restrict__ vin,
                                                    K* __restrict__ kout, uint32_t* __restrict__ vout, uint32_t n,
                                                    int shift, const uint32_t* __restrict__ hist,
                                                    uint32_t* __restrict__ tile_ctr, uint32_t* __restrict__ status,
                                                    uint32_t* __restrict__ err) {
  __shared__ uint32_t s_tile;
  __shared__ uint32_t s_cnt[kWaves][256];
  __shared__ uint32_t s_dstart[256];
  __shared__ uint32_t s_gbase[256];
  __shared__ uint32_t s_tmp[8];
  __shared__ K s_keys[kTile];
  __shared__ uint32_t s_vals[kVals ? kTile : 1];

  const uint32_t tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  if (tid == 0) s_tile = atomicAdd(tile_ctr, 1u);
  for (int i = tid; i < kWaves * 256; i += 256) (&s_cnt[0][0])[i] = 0;
  __syncthreads();
  const uint32_t tile = s_tile;
  const size_t base = (size_t)tile * kTile;
  const size_t wbase = base + (size_t)wave * kWaveTile;

  K key[kItems];
  uint32_t val[kItems];
  uint32_t rank[kItems];
  uint32_t dig[kItems];
#pragma unroll
  for (int j = 0; j < kItems; ++j) {
    size_t idx = wbase + (size_t)j * 64 + lane;
    if (idx < n) {
      key[j] = kin[idx];
      if (kVals) val[j] = vin[idx];
    } else {
      key[j] = 0;
      if (kVals) val[j] = 0;
    }
  }
  // Stable in-tile ranking: iteration j of wave w covers keys
  // [wbase + 64j, wbase + 64j + 64) in lane order, so (w, j, lane) is memory order.
#pragma unroll
  for (int j = 0; j < kItems; ++j) {
    size_t idx = wbase + (size_t)j * 64 + lane;
    bool valid = idx < n;
    uint32_t d = (uint32_t)(key[j] >> shift) & 0xFFu;
    dig[j] = d;
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      uint32_t bit = (d >> b) & 1u;
      uint64_t bb = __ballot(bit);
      peers &= bit ? bb : ~bb;
    }
    uint32_t lower = (uint32_t)__popcll(peers & rk::lanemask_lt());
    uint32_t cnt = (uint32_t)__popcll(peers);
    uint32_t old = s_cnt[wave][d];
    rank[j] = old + lower;
    __builtin_amdgcn_wave_barrier();
    if (valid && lower + 1 == cnt) s_cnt[wave][d] = old + cnt;
    __builtin_amdgcn_wave_barrier();
  }
  __syncthreads();

  const uint32_t d = tid;
  uint32_t run = 0;
#pragma unroll
  for (int w = 0; w < kWaves; ++w) {
    uint32_t c = s_cnt[w][d];
    s_cnt[w][d] = run;
    run += c;
  }
  const uint32_t total = run;
  uint32_t* my_status = status + (size_t)tile * 256 + d;
  rk::atomic_store_relaxed(my_status, (tile == 0 ? kInc : kAgg) | total);

  const uint32_t local_excl = block_excl_scan256(total, s_tmp);
  const uint32_t global_excl = block_excl_scan256(hist[d], s_tmp);

  uint32_t prefix = 0;
  if (tile > 0) {
    int32_t j = (int32_t)tile - 1;
    uint32_t spins = 0;
    while (j >= 0) {
      uint32_t s = rk::atomic_load_relaxed(status + (size_t)j * 256 + d);
      uint32_t flag = s & ~kCountMask;
      if (flag == 0) {
        if (++spins > kSpinLimit) {
          atomicOr(err, (uint32_t)RK_FLAG_SPIN_LIMIT);
          break;
        }
        __builtin_amdgcn_s_sleep(1);
        continue;
      }
      prefix += s & kCountMask;
      if (flag == kInc) break;
      --j;
    }
    rk::atomic_store_relaxed(my_status, kInc | ((prefix + total) & kCountMask));
  }
  s_dstart[d] = local_excl;
  s_gbase[d] = global_excl + prefix - local_excl;
  __syncthreads();

#pragma unroll
  for (int j = 0; j < kItems; ++j) {
    size_t idx = wbase + (size_t)j * 64 + lane;
    if (idx < n) {
      uint32_t dd = dig[j];
      uint32_t pos = s_dstart[dd] + s_cnt[wave][dd] + rank[j];
      s_keys[pos] = key[j];
      if (kVals) s_vals[pos] = val[j];
    }
  }
  __syncthreads();

  const uint32_t tile_n = (n - base) < (size_t)kTile ? (uint32_t)(n - base) : (uint32_t)kTile;
  for (uint32_t i = tid; i < tile_n; i += 256) {
    K k = s_keys[i];
    uint32_t dd = (uint32_t)(k >> shift) & 0xFFu;
    size_t o = (size_t)s_gbase[dd] + i;
    if (o < n) {  // always true unless the look-back failed (err is set then)
      kout[o] = k;
      if (kVals) vout[o] = s_vals[i];
    }
  }
}

template <typename K, bool kVals>
__global__ __launch_bounds__(256) void k_copy_back(const K* __restrict__ ks, const uint32_t* __restrict__ vs,
                                                   K* __restrict__ kd, uint32_t* __restrict__ vd, uint32_t n) {
  size_t stride = (size_t)gridDim.x * 256;
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += stride) {
    kd[i] = ks[i];
    if (kVals) vd[i] = vs[i];
  }
}

// The reference's middle size class (CUDA_bitonicSortGeneral.cu:39-169):
// each full 128-element chunk goes through the bitonic network on (key, value);
// the partial last chunk is sorted stably (odd-even transposition in the
// reference, :107-162). A stable key sort afterwards reproduces the order of
// the reference's stable '<=' merges (:172-481).
__global__ __launch_bounds__(128) void k_bitonic_chunks(uint32_t* __restrict__ keys, uint32_t* __restrict__ vals,
                                                        uint32_t n) {
  __shared__ uint32_t sk[128], sv[128];
  const uint32_t t = threadIdx.x, o = blockIdx.x * 128u;
  const uint32_t size = (n - o) < 128u ? (n - o) : 128u;
  if (size <= 1) return;
  if (t < size) { sk[t] = keys[o + t]; sv[t] = vals[o + t]; }
  __syncthreads();
  if (size == 128) {
    for (uint32_t i = 2; i <= 128; i <<= 1) {
      for (uint32_t j = i >> 1; j > 0; j >>= 1) {
        uint32_t ixj = t ^ j;
        if (ixj > t) {
          uint32_t ka = sk[t], kb = sk[ixj];
          bool sw = ((t & i) == 0) ? (ka > kb) : (ka < kb);
          if (sw) {
            uint32_t va = sv[t], vb = sv[ixj];
            sk[t] = kb; sk[ixj] = ka; sv[t] = vb; sv[ixj] = va;
          }
        }
        __syncthreads();
      }
    }
    keys[o + t] = sk[t];
    vals[o + t] = sv[t];
  } else if (t < size) {
    uint32_t k = sk[t], r = 0;
    for (uint32_t u = 0; u < size; ++u) r += (sk[u] < k || (sk[u] == k && u < t)) ? 1u : 0u;
    keys[o + r] = k;
    vals[o + r] = sv[t];
  }
}

__global__ __launch_bounds__(256) void k_iota(uint32_t* __restrict__ v, uint32_t n) {
  size_t stride = (size_t)gridDim.x * 256;
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += stride) v[i] = (uint32_t)i;
}

inline uint32_t tiles_of(uint32_t n) { return (n + kTile - 1) / kTile; }
inline unsigned grid_for(size_t n, size_t per_block, unsigned cap) {
  size_t g = (n + per_block - 1) / per_block;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (unsigned)g;
}

}  // namespace

// Size of the workspace for a sort of 'length' keys of 'key_bytes' bytes.
extern "C" size_t rk_sort_workspace_size(uint32_t length, uint32_t key_bytes) {
  size_t n = length ? length : 1;
  size_t T = tiles_of((uint32_t)n);
  size_t P = key_bytes >= 8 ? 8 : 4;
  return rk_align_up(P * 256 * 4 + P * 4 + 4 + P * T * 256 * 4) + rk_align_up(n * (key_bytes >= 8 ? 8 : 4)) +
         rk_align_up(n * 4) + 1024;
}

// Sort keys (and values) stably on digits first_pass .. first_pass+passes-1.
// 'err' (device, may be NULL) receives RK_FLAG_SPIN_LIMIT on a look-back timeout.
template <typename K>
rk_status rk_radix_sort_impl(hipStream_t st, K* keys, uint32_t* vals, uint32_t n, int first_pass, int passes,
                             void* ws, size_t ws_bytes, uint32_t* err) {
  if (n <= 1 || passes <= 0) return RK_SUCCESS;
  if (!keys || passes > kMaxPasses) return RK_ERROR_INVALID_VALUE;
  const uint32_t T = tiles_of(n);
  RkArena ar(ws, ws_bytes);
  const size_t ctl_words = (size_t)passes * 256 + passes + 1 + (size_t)passes * T * 256;
  uint32_t* ctl = ar.take<uint32_t>(ctl_words);
  K* kalt = ar.take<K>(n);
  uint32_t* valt = vals ? ar.take<uint32_t>(n) : nullptr;
  if (!ctl || !kalt || (vals && !valt)) return RK_ERROR_WORKSPACE;
  uint32_t* hist = ctl;
  uint32_t* ctr = hist + (size_t)passes * 256;
  uint32_t* werr = ctr + passes;
  uint32_t* status = werr + 1;
  if (!err) err = werr;

  rk_status s = rk_fill_u32(reinterpret_cast<rk_stream>(st), ctl, ctl_words, 0u);
  if (s != RK_SUCCESS) return s;
  hipLaunchKernelGGL(k_rs_hist<K>, dim3(grid_for(n, 256 * 16, 256)), dim3(256), 0, st, keys, n, passes, first_pass,
                     hist);
  RK_CHECK_LAUNCH();
  K* ks = keys;
  K* kd = kalt;
  uint32_t* vs = vals;
  uint32_t* vd = valt;
  for (int p = 0; p < passes; ++p) {
    int shift = 8 * (p + first_pass);
    if (vals) {
      hipLaunchKernelGGL((k_rs_scatter<K, true>), dim3(T), dim3(256), 0, st, ks, vs, kd, vd, n, shift,
                         hist + (size_t)p * 256, ctr + p, status + (size_t)p * T * 256, err);
    } else {
      hipLaunchKernelGGL((k_rs_scatter<K, false>), dim3(T), dim3(256), 0, st, ks, (const uint32_t*)nullptr, kd,
                         (uint32_t*)nullptr, n, shift, hist + (size_t)p * 256, ctr + p,
                         status + (size_t)p * T * 256, err);
    }
    RK_CHECK_LAUNCH();
    std::swap(ks, kd);
    std::swap(vs, vd);
  }
  if (passes & 1) {
    if (vals) {
      hipLaunchKernelGGL((k_copy_back<K, true>), dim3(grid_for(n, 256 * 8, 4096)), dim3(256), 0, st, ks, vs, keys,
                         vals, n);
    } else {
      hipLaunchKernelGGL((k_copy_back<K, false>), dim3(grid_for(n, 256 * 8, 4096)), dim3(256), 0, st, ks,
                         (const uint32_t*)nullptr, keys, (uint32_t*)nullptr, n);
    }
    RK_CHECK_LAUNCH();
  }
  return RK_SUCCESS;
}

template rk_status rk_radix_sort_impl<uint32_t>(hipStream_t, uint32_t*, uint32_t*, uint32_t, int, int, void*,
                                                size_t, uint32_t*);
template rk_status rk_radix_sort_impl<uint64_t>(hipStream_t, uint64_t*, uint32_t*, uint32_t, int, int, void*,
                                                size_t, uint32_t*);

// The reference tie order of CUDA_radixSortGeneral (CUDA_radixSortGeneral.cu:397-482):
//  length <= 512 (bubble, CUDA_bubbleSort512.cu): stable by key;
//  length <= byte_depth*8192 (bitonic general): network order inside each
//   128-chunk, then stable by key;
//  larger (LSD radix on the low byte_depth bytes): stable by those bytes.
rk_status rk_sort_general_impl(hipStream_t st, uint32_t* keys, uint32_t* indices, uint32_t length,
                               uint32_t bit_depth, int init_indices, void* ws, size_t ws_bytes, uint32_t* err) {
  if (bit_depth == 0 || bit_depth > 32) return RK_ERROR_INVALID_VALUE;
  if (length == 0) return RK_SUCCESS;
  if (!keys || !indices) return RK_ERROR_INVALID_VALUE;
  const uint32_t byte_depth = 1 + (bit_depth - 1) / 8;
  if (init_indices) {
    hipLaunchKernelGGL(k_iota, dim3(grid_for(length, 256 * 8, 4096)), dim3(256), 0, st, indices, length);
    RK_CHECK_LAUNCH();
  }
  if (length <= 1) return RK_SUCCESS;
  if (length <= 512) return rk_radix_sort_impl<uint32_t>(st, keys, indices, length, 0, 4, ws, ws_bytes, err);
  if (length <= byte_depth * 8192u) {
    hipLaunchKernelGGL(k_bitonic_chunks, dim3((length + 127) / 128), dim3(128), 0, st, keys, indices, length);
    RK_CHECK_LAUNCH();
    return rk_radix_sort_impl<uint32_t>(st, keys, indices, length, 0, 4, ws, ws_bytes, err);
  }
  return rk_radix_sort_impl<uint32_t>(st, keys, indices, length, 0, (int)byte_depth, ws, ws_bytes, err);
}

extern "C" rk_status rk_radix_sort_general(rk_stream stream, uint32_t* keys, uint32_t* indices, uint32_t length,
                                           uint32_t bit_depth, int init_indices, void* ws, size_t ws_bytes) {
  return rk_sort_general_impl(rk_hip_stream(stream), keys, indices, length, bit_depth, init_indices, ws, ws_bytes,
                              nullptr);
}

// CUDA_radixSort (CUDA_radixSort.cu:357-392): values only, sorted on the low
// byte_depth bytes (the bitonic class compares whole values; for values below
// 2^(8*byte_depth) both agree, and equal values are indistinguishable).
extern "C" rk_status rk_radix_sort(rk_stream stream, uint32_t* values, uint32_t length, uint32_t bit_depth, void* ws,
                                   size_t ws_bytes) {
  if (bit_depth == 0 || bit_depth > 32) return RK_ERROR_INVALID_VALUE;
  if (length <= 1) return RK_SUCCESS;
  const uint32_t byte_depth = 1 + (bit_depth - 1) / 8;
  int passes = (length <= byte_depth * 8192u) ? 4 : (int)byte_depth;
  return rk_radix_sort_impl<uint32_t>(rk_hip_stream(stream), values, nullptr, length, 0, passes, ws, ws_bytes,
                                      nullptr);
}

extern "C" rk_status rk_radix_sort_pairs_u64(rk_stream stream, uint64_t* keys, uint32_t* values, uint32_t length,
                                             uint32_t bit_depth, void* ws, size_t ws_bytes) {
  if (bit_depth == 0 || bit_depth > 64) return RK_ERROR_INVALID_VALUE;
  if (length <= 1) return RK_SUCCESS;
  int passes = (int)(1 + (bit_depth - 1) / 8);
  return rk_radix_sort_impl<uint64_t>(rk_hip_stream(stream), keys, values, length, 0, passes, ws, ws_bytes, nullptr);
}
