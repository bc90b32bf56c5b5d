// Stable LSD radix sort for gfx950, 8-bit digits, onesweep style:
//   1. one histogram kernel counts the digits of every pass in a single read;
//   2. one scatter kernel per pass. Each 512-thread workgroup sorts the
//      8192-key tile (4096 below 2^21 keys) of its block index: it ranks its keys stably with
//      per-(wave, digit) lane masks in LDS, publishes its per-digit counts,
//      resolves its global per-digit prefix by decoupled look-back over
//      earlier tiles, reorders the tile by digit in LDS and writes it out in
//      per-digit runs, keys first, then values through the same LDS.
// Tiles follow the block index rather than a counter that every workgroup
// increments (on MI355X such a counter serialises the workgroups' start
// across the eight XCDs). A workgroup only waits on lower tiles; workgroups
// are dispatched in block order on each XCD, so the lowest unfinished tile is
// always resident or next to be dispatched on its XCD, whose earlier
// workgroups have all finished. Look-back waits are bounded (kSpinLimit); on
// expiry the kernel sets an error bit and finishes, so a fault in the
// protocol cannot hang the GPU.
//
// The reference's sorts (CUDA_radixSortGeneral.cu, CUDA_radixSort.cu,
// CUDA_bitonicSortGeneral.cu, CUDA_bitonicSort.cu, CUDA_bubbleSort512.cu)
// differ in tie order by size class; rk_radix_sort_general reproduces that
// order (see the size-class notes at the entry point).
#include "rk_device.h"

namespace {

#ifndef RK_SCATTER_WAVES
#define RK_SCATTER_WAVES 8
#endif
#ifndef RK_SCATTER_ITEMS
#define RK_SCATTER_ITEMS 16
#endif
#ifndef RK_SCATTER_ITEMS_SMALL
#define RK_SCATTER_ITEMS_SMALL 8
#endif
constexpr int kWaves = RK_SCATTER_WAVES;    // waves per scatter workgroup
constexpr int kThreads = 64 * kWaves;
constexpr int kItems = RK_SCATTER_ITEMS;    // keys per lane
constexpr int kItemsSmall = RK_SCATTER_ITEMS_SMALL;  // keys per lane below kSmallTilesBelow keys
// Below 2^21 keys the tiles are of 4096 keys: tiles of 8192 would leave most
// of the 256 CUs idle (1M keys: 123 tiles). On the 1M-face heightfield build
// 214 vs 224 us; on 10M keys the larger tiles are faster (1.73 vs 1.80 ms).
constexpr size_t kSmallTilesBelow = size_t(1) << 21;
static_assert(kThreads >= 256, "one thread per digit");
constexpr uint32_t kAgg = 1u << 30;
constexpr uint32_t kInc = 2u << 30;
constexpr uint32_t kCountMask = (1u << 30) - 1u;
constexpr uint32_t kSpinLimit = 1u << 20;
constexpr int kMaxPasses = 8;
#ifndef RK_SORT_LOOK
#define RK_SORT_LOOK 8
#endif
constexpr int kLook = RK_SORT_LOOK;  // look-back status loads in flight per digit

// Exclusive scan over the first 256 threads of the block (one value per
// digit); every thread of the block calls it, threads 256.. get no result.
__device__ __forceinline__ uint32_t block_excl_scan256(uint32_t v, uint32_t* s_tmp /*>= 4*/) {
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint32_t incl = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    uint32_t y = __shfl_up(incl, off, 64);
    if (lane >= (uint32_t)off) incl += y;
  }
  if (lane == 63 && wave < 4) s_tmp[wave] = incl;
  __syncthreads();
  uint32_t wprefix = 0;
#pragma unroll
  for (int w = 0; w < 4; ++w) wprefix += (w < (int)wave) ? s_tmp[w] : 0u;
  __syncthreads();
  return wprefix + incl - v;
}

// Digit counts of every pass in one read. Block b counts keys
// [b * chunk, (b + 1) * chunk), eight loads in flight per thread; a wave whose
// 64 keys share a digit (the high digits of spatially ordered keys) adds them
// with one LDS atomic instead of 64 to the same address.
constexpr int kHistUnroll = 8;
constexpr uint32_t kHistMaxBlocks = 1024;
#ifndef RK_HIST_KEYS_PER_BLOCK
#define RK_HIST_KEYS_PER_BLOCK (256 * kHistUnroll * 2)
#endif
constexpr size_t kHistKeysPerBlock = RK_HIST_KEYS_PER_BLOCK;

// Also zeroes the look-back status words of the passes that follow (zero,
// zero_words), which saves a fill launch per sort.
template <typename K>
__global__ __launch_bounds__(256) void k_rs_hist(const K* __restrict__ keys, uint32_t n, int passes, int first_pass,
                                                 uint32_t chunk, uint32_t* __restrict__ hist,
                                                 uint32_t* __restrict__ zero, size_t zero_words) {
  __shared__ uint32_t s[kMaxPasses * 256];
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < zero_words; i += (size_t)gridDim.x * 256) zero[i] = 0u;
  for (int i = threadIdx.x; i < passes * 256; i += 256) s[i] = 0;
  __syncthreads();
  const size_t b0 = (size_t)blockIdx.x * chunk;
  const size_t b1 = b0 + chunk < (size_t)n ? b0 + chunk : (size_t)n;
  for (size_t i0 = b0; i0 < b1; i0 += 256 * kHistUnroll) {
    K k[kHistUnroll];
    bool ok[kHistUnroll];
#pragma unroll
    for (int u = 0; u < kHistUnroll; ++u) {
      const size_t i = i0 + (size_t)u * 256 + threadIdx.x;
      ok[u] = i < b1;
      k[u] = ok[u] ? keys[i] : (K)0;
    }
#pragma unroll
    for (int u = 0; u < kHistUnroll; ++u) {
      const uint64_t active = __ballot(ok[u]);
      if (!ok[u]) continue;
      for (int p = 0; p < passes; ++p) {
        const uint32_t d = (uint32_t)(k[u] >> (8 * (p + first_pass))) & 0xFFu;
        const uint32_t d0 = __builtin_amdgcn_readfirstlane(d);
        if (__ballot(d == d0) == active) {
          if (__lane_id() == (uint32_t)__builtin_ctzll(active)) atomicAdd(&s[p * 256 + d0], (uint32_t)__popcll(active));
        } else {
          atomicAdd(&s[p * 256 + d], 1u);
        }
      }
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < passes * 256; i += 256) {
    uint32_t c = s[i];
    if (c) atomicAdd(&hist[i], c);
  }
}

// One pass over one tile of kThreads * kIt keys. The tile's keys are ranked stably
// per digit and moved to their place in the tile's digit order in LDS before
// the look-back, so their registers are free while it waits; the values are
// loaded during the look-back and follow through the same LDS buffer once the
// keys are out, so a workgroup holds a tile of keys in LDS, not keys and values.
template <typename K, bool kVals, int kIt>
__global__ __launch_bounds__(kThreads) void k_rs_scatter(const K* __restrict__ kin, const uint32_t* __restrict__ vin,
                                                         K* __restrict__ kout, uint32_t* __restrict__ vout, uint32_t n,
                                                         int shift, const uint32_t* __restrict__ hist,
                                                         uint32_t* __restrict__ status, uint32_t* __restrict__ err) {
  constexpr int kTileT = kThreads * kIt, kWaveTileT = 64 * kIt;  // keys per tile and per wave
  __shared__ uint32_t s_cnt[kWaves][256];
  __shared__ uint32_t s_gbase[256];
  __shared__ uint32_t s_tmp[8];
  __shared__ K s_buf[kTileT];

  const uint32_t tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  // While ranking, s_buf holds one 64-bit lane mask per (wave, digit): the
  // lanes of the current 64 keys that carry that digit.
  static_assert(sizeof(K) * kTileT >= sizeof(uint64_t) * kWaves * 256, "digit masks fit in the tile buffer");
  uint64_t* s_mask = reinterpret_cast<uint64_t*>(s_buf);
  for (int i = tid; i < kWaves * 256; i += kThreads) {
    (&s_cnt[0][0])[i] = 0;
    s_mask[i] = 0;
  }
  __syncthreads();
  const uint32_t tile = blockIdx.x;
  const size_t base = (size_t)tile * kTileT;
  const size_t wbase = base + (size_t)wave * kWaveTileT;

  K key[kIt];
  uint32_t rank[kIt];
#pragma unroll
  for (int j = 0; j < kIt; ++j) {
    size_t idx = wbase + (size_t)j * 64 + lane;
    key[j] = idx < n ? kin[idx] : (K)0;
  }
  // Stable in-tile ranking: iteration j of wave w covers keys
  // [wbase + 64j, wbase + 64j + 64) in lane order, so (w, j, lane) is memory order.
  // Each lane ORs its bit into its digit's mask; reading the mask back gives
  // the lanes that share the digit, so the lane's rank among them is the
  // count of lower lanes. The group's last lane advances the digit's counter
  // and clears the mask for the next 64 keys.
  const uint64_t my_bit = 1ull << lane;
  const uint64_t lower_lanes = my_bit - 1;
  uint64_t* wmask = s_mask + wave * 256;
#pragma unroll
  for (int j = 0; j < kIt; ++j) {
    size_t idx = wbase + (size_t)j * 64 + lane;
    bool valid = idx < n;
    uint32_t d = (uint32_t)(key[j] >> shift) & 0xFFu;
    if (valid) __hip_atomic_fetch_or(wmask + d, my_bit, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    __builtin_amdgcn_wave_barrier();
    const uint64_t peers =
        valid ? __hip_atomic_load(wmask + d, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) : 0ull;
    const uint32_t lower = (uint32_t)__popcll(peers & lower_lanes);
    const uint32_t cnt = (uint32_t)__popcll(peers);
    const uint32_t old = s_cnt[wave][d];
    rank[j] = old + lower;
    __builtin_amdgcn_wave_barrier();
    if (valid && lower + 1 == cnt) {
      s_cnt[wave][d] = old + cnt;
      __hip_atomic_store(wmask + d, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
    __builtin_amdgcn_wave_barrier();
  }
  __syncthreads();

  // Per digit (thread d < 256): the tile's count, published at once for the
  // look-back of later tiles, and each wave's start within the digit's run.
  const uint32_t d = tid;
  uint32_t total = 0;
  if (tid < 256) {
#pragma unroll
    for (int w = 0; w < kWaves; ++w) {
      uint32_t c = s_cnt[w][d];
      s_cnt[w][d] = total;
      total += c;
    }
    rk::atomic_store_relaxed(status + (size_t)tile * 256 + d, (tile == 0 ? kInc : kAgg) | total);
  }
  const uint32_t local_excl = block_excl_scan256(total, s_tmp);
  if (tid < 256) {
#pragma unroll
    for (int w = 0; w < kWaves; ++w) s_cnt[w][d] += local_excl;
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kIt; ++j) {
    size_t idx = wbase + (size_t)j * 64 + lane;
    if (idx < n) {
      const uint32_t dd = (uint32_t)(key[j] >> shift) & 0xFFu;
      rank[j] += s_cnt[wave][dd];  // position in the tile's digit order
      s_buf[rank[j]] = key[j];
    }
  }
  uint32_t val[kIt];
  if (kVals) {
#pragma unroll
    for (int j = 0; j < kIt; ++j) {
      size_t idx = wbase + (size_t)j * 64 + lane;
      val[j] = idx < n ? vin[idx] : 0u;
    }
  }
  const uint32_t global_excl = block_excl_scan256(tid < 256 ? hist[tid & 255] : 0u, s_tmp);

  // Decoupled look-back, kLook predecessors' status words per round trip:
  // add counts from tile j down until an inclusive prefix; stop the round at
  // a tile that has published nothing yet and reload from there.
  if (tid < 256) {
    uint32_t prefix = 0;
    if (tile > 0) {
      int32_t j = (int32_t)tile - 1;
      uint32_t spins = 0;
      for (;;) {
        uint32_t sv[kLook];
#pragma unroll
        for (int w = 0; w < kLook; ++w)
          sv[w] = (j - w >= 0) ? rk::atomic_load_relaxed(status + (size_t)(j - w) * 256 + d) : kInc;
        int used = 0;
        bool inc = false, stalled = false;
#pragma unroll
        for (int w = 0; w < kLook; ++w) {
          if (inc || stalled) continue;
          const uint32_t flag = sv[w] & ~kCountMask;
          if (flag == 0) {
            stalled = true;
            continue;
          }
          prefix += sv[w] & kCountMask;
          ++used;
          if (flag == kInc) inc = true;
        }
        if (inc) break;
        j -= used;
        if (used == 0) {
          if (++spins > kSpinLimit) {
            atomicOr(err, (uint32_t)RK_FLAG_SPIN_LIMIT);
            break;
          }
          __builtin_amdgcn_s_sleep(1);
        }
      }
      rk::atomic_store_relaxed(status + (size_t)tile * 256 + d, kInc | ((prefix + total) & kCountMask));
    }
    s_gbase[d] = global_excl + prefix - local_excl;
  }
  __syncthreads();

  const uint32_t tile_n = (n - base) < (size_t)kTileT ? (uint32_t)(n - base) : (uint32_t)kTileT;
  uint32_t opos[kIt];
#pragma unroll
  for (int m = 0; m < kIt; ++m) {
    const uint32_t i = tid + (uint32_t)m * kThreads;
    opos[m] = 0xFFFFFFFFu;
    if (i < tile_n) {
      const K k = s_buf[i];
      const uint32_t dd = (uint32_t)(k >> shift) & 0xFFu;
      const size_t o = (size_t)s_gbase[dd] + i;
      if (o < n) {  // always true unless the look-back failed (err is set then)
        kout[o] = k;
        opos[m] = (uint32_t)o;
      }
    }
  }
  if (kVals) {
    uint32_t* s_val = reinterpret_cast<uint32_t*>(s_buf);
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kIt; ++j) {
      size_t idx = wbase + (size_t)j * 64 + lane;
      if (idx < n) s_val[rank[j]] = val[j];
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < kIt; ++m) {
      const uint32_t i = tid + (uint32_t)m * kThreads;
      if (i < tile_n && opos[m] != 0xFFFFFFFFu) vout[opos[m]] = s_val[i];
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

// Stable sort of at most 512 keys (and values) in one workgroup: each thread
// ranks its key by counting the smaller keys and the equal keys before it. This
// is the order of the reference's odd-even transposition sort, which swaps only
// on '>' (CUDA_bubbleSort512.cu:42-96) and is therefore stable. 'init' takes
// the values to be 0..n-1 (kernel_bubbleSortInit, CUDA_bubbleSort512.cu:98-104).
// 'clear' (may be NULL) is set to 0. Every global read happens before the
// barrier and every write after it, so the sort is in place.
constexpr uint32_t kSmallSort = 512;
__global__ __launch_bounds__(512) void k_small_sort(uint32_t* __restrict__ keys, uint32_t* __restrict__ vals,
                                                    uint32_t n, int init, uint32_t* __restrict__ clear) {
  __shared__ uint32_t sk[kSmallSort];
  const uint32_t t = threadIdx.x;
  if (clear && t == 0) *clear = 0u;
  uint32_t k = 0, v = 0;
  if (t < n) {
    k = keys[t];
    sk[t] = k;
    if (vals) v = init ? t : vals[t];
  }
  __syncthreads();
  if (t < n) {
    uint32_t r = 0;
    for (uint32_t u = 0; u < n; ++u) {
      const uint32_t x = sk[u];
      r += (x < k || (x == k && u < t)) ? 1u : 0u;
    }
    keys[r] = k;
    if (vals) vals[r] = v;
  }
}

__global__ __launch_bounds__(256) void k_iota(uint32_t* __restrict__ v, uint32_t n) {
  size_t stride = (size_t)gridDim.x * 256;
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += stride) v[i] = (uint32_t)i;
}

inline int items_of(uint32_t n) { return (size_t)n < kSmallTilesBelow ? kItemsSmall : kItems; }
inline uint32_t tiles_of(uint32_t n) {
  const uint32_t tile = (uint32_t)(kThreads * items_of(n));
  return (uint32_t)(((size_t)n + tile - 1) / tile);
}
// The error word of a sort called with this workspace: its first 4 bytes.
inline uint32_t* ws_flags_word(void* ws, size_t ws_bytes) {
  return (ws && ws_bytes >= 4) ? static_cast<uint32_t*>(ws) : nullptr;
}
inline rk_status clear_ws_flags(hipStream_t st, void* ws, size_t ws_bytes) {
  uint32_t* w = ws_flags_word(ws, ws_bytes);
  return w ? rk_fill_u32(reinterpret_cast<rk_stream>(st), w, 1, 0u) : RK_SUCCESS;
}
rk_status launch_small_sort(hipStream_t st, uint32_t* keys, uint32_t* vals, uint32_t n, int init, uint32_t* clear) {
  hipLaunchKernelGGL(k_small_sort, dim3(1), dim3(kSmallSort), 0, st, keys, vals, n, init, clear);
  RK_CHECK_LAUNCH();
  return RK_SUCCESS;
}
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
// 'err' (device, may be NULL) receives RK_FLAG_SPIN_LIMIT on a look-back
// timeout; when it is NULL the first word of 'ws' does, cleared by this call.
// The look-back status words count keys in 30 bits, so n < 2^30.
// head_cleared: the caller has zeroed the first rk_sort_head_words(passes)
// words of ws (error word and digit counts) in an earlier kernel on 'st', so
// no fill runs here.
// digit_counts (may be NULL): receives the device address of the first pass's
// 256 digit counts, valid once the sort has run on 'st' (NULL when n <= 1).
template <typename K>
rk_status rk_radix_sort_impl(hipStream_t st, K* keys, uint32_t* vals, uint32_t n, int first_pass, int passes,
                             void* ws, size_t ws_bytes, uint32_t* err, bool head_cleared = false,
                             const uint32_t** digit_counts = nullptr) {
  if (digit_counts) *digit_counts = nullptr;
  if (n >= kAgg) return RK_ERROR_UNSUPPORTED;
  if (n <= 1 || passes <= 0) return err ? RK_SUCCESS : clear_ws_flags(st, ws, ws_bytes);
  if (!keys || passes > kMaxPasses) return RK_ERROR_INVALID_VALUE;
  const uint32_t T = tiles_of(n);
  RkArena ar(ws, ws_bytes);
  const size_t ctl_words = 1 + (size_t)passes * 256 + (size_t)passes * T * 256;
  uint32_t* ctl = ar.take<uint32_t>(ctl_words);  // first in ws: ctl[0] is the error word
  K* kalt = ar.take<K>(n);
  uint32_t* valt = vals ? ar.take<uint32_t>(n) : nullptr;
  if (!ctl || !kalt || (vals && !valt)) return RK_ERROR_WORKSPACE;
  uint32_t* werr = ctl;
  uint32_t* hist = ctl + 1;
  uint32_t* status = hist + (size_t)passes * 256;
  if (!err) err = werr;
  if (digit_counts) *digit_counts = hist;

  if (!head_cleared) {
    rk_status s = rk_fill_u32(reinterpret_cast<rk_stream>(st), ctl, 1 + (size_t)passes * 256, 0u);
    if (s != RK_SUCCESS) return s;
  }
  const unsigned hist_blocks = grid_for(n, kHistKeysPerBlock, kHistMaxBlocks);
  const uint32_t chunk = (uint32_t)(((size_t)n + hist_blocks - 1) / hist_blocks);
  hipLaunchKernelGGL(k_rs_hist<K>, dim3(hist_blocks), dim3(256), 0, st, keys, n, passes, first_pass, chunk, hist,
                     status, (size_t)passes * T * 256);
  RK_CHECK_LAUNCH();
  K* ks = keys;
  K* kd = kalt;
  uint32_t* vs = vals;
  uint32_t* vd = valt;
  for (int p = 0; p < passes; ++p) {
    int shift = 8 * (p + first_pass);
    const bool small = items_of(n) == kItemsSmall;
    if (vals) {
      if (small)
        hipLaunchKernelGGL((k_rs_scatter<K, true, kItemsSmall>), dim3(T), dim3(kThreads), 0, st, ks, vs, kd, vd, n,
                           shift, hist + (size_t)p * 256, status + (size_t)p * T * 256, err);
      else
        hipLaunchKernelGGL((k_rs_scatter<K, true, kItems>), dim3(T), dim3(kThreads), 0, st, ks, vs, kd, vd, n, shift,
                           hist + (size_t)p * 256, status + (size_t)p * T * 256, err);
    } else {
      if (small)
        hipLaunchKernelGGL((k_rs_scatter<K, false, kItemsSmall>), dim3(T), dim3(kThreads), 0, st, ks,
                           (const uint32_t*)nullptr, kd, (uint32_t*)nullptr, n, shift, hist + (size_t)p * 256,
                           status + (size_t)p * T * 256, err);
      else
        hipLaunchKernelGGL((k_rs_scatter<K, false, kItems>), dim3(T), dim3(kThreads), 0, st, ks,
                           (const uint32_t*)nullptr, kd, (uint32_t*)nullptr, n, shift, hist + (size_t)p * 256,
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
                                                size_t, uint32_t*, bool, const uint32_t**);
template rk_status rk_radix_sort_impl<uint64_t>(hipStream_t, uint64_t*, uint32_t*, uint32_t, int, int, void*,
                                                size_t, uint32_t*, bool, const uint32_t**);

// The reference tie order of CUDA_radixSortGeneral (CUDA_radixSortGeneral.cu:397-482):
//  length <= 512 (bubble, CUDA_bubbleSort512.cu): stable by key;
//  length <= byte_depth*8192 (bitonic general): network order inside each
//   128-chunk, then stable by key;
//  larger (LSD radix on the low byte_depth bytes): stable by those bytes.
rk_status rk_sort_general_impl(hipStream_t st, uint32_t* keys, uint32_t* indices, uint32_t length,
                               uint32_t bit_depth, int init_indices, void* ws, size_t ws_bytes, uint32_t* err) {
  if (bit_depth == 0 || bit_depth > 32) return RK_ERROR_INVALID_VALUE;
  if (length >= kAgg) return RK_ERROR_UNSUPPORTED;
  if (length == 0) return err ? RK_SUCCESS : clear_ws_flags(st, ws, ws_bytes);
  if (!keys || !indices) return RK_ERROR_INVALID_VALUE;
  const uint32_t byte_depth = 1 + (bit_depth - 1) / 8;
  if (length <= kSmallSort)
    return launch_small_sort(st, keys, indices, length, init_indices, err ? nullptr : ws_flags_word(ws, ws_bytes));
  if (init_indices) {
    hipLaunchKernelGGL(k_iota, dim3(grid_for(length, 256 * 8, 4096)), dim3(256), 0, st, indices, length);
    RK_CHECK_LAUNCH();
  }
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
  if (length >= kAgg) return RK_ERROR_UNSUPPORTED;
  if (length > 1 && !values) return RK_ERROR_INVALID_VALUE;
  if (length <= kSmallSort)
    return launch_small_sort(rk_hip_stream(stream), values, nullptr, length, 0, ws_flags_word(ws, ws_bytes));
  const uint32_t byte_depth = 1 + (bit_depth - 1) / 8;
  int passes = (length <= byte_depth * 8192u) ? 4 : (int)byte_depth;
  return rk_radix_sort_impl<uint32_t>(rk_hip_stream(stream), values, nullptr, length, 0, passes, ws, ws_bytes,
                                      nullptr);
}

extern "C" rk_status rk_radix_sort_pairs_u64(rk_stream stream, uint64_t* keys, uint32_t* values, uint32_t length,
                                             uint32_t bit_depth, void* ws, size_t ws_bytes) {
  if (bit_depth == 0 || bit_depth > 64) return RK_ERROR_INVALID_VALUE;
  int passes = (int)(1 + (bit_depth - 1) / 8);
  return rk_radix_sort_impl<uint64_t>(rk_hip_stream(stream), keys, values, length, 0, passes, ws, ws_bytes, nullptr);
}

// template <typename T> void CUDA_bitonicSort(T* values, uint length),
// samples/09_how_to_use_a_sorter.cpp:39-40 / CUDA_bitonicSort.cu:402-518
// (instantiated for unsigned int, :523-537). The values are their own keys, so
// any correct sort gives the reference's result: at most 512 values in one
// workgroup, more by the radix passes over all 32 bits.
extern "C" rk_status rk_bitonic_sort(rk_stream stream, uint32_t* values, uint32_t length, void* ws, size_t ws_bytes) {
  if (length >= kAgg) return RK_ERROR_UNSUPPORTED;
  if (length > 1 && !values) return RK_ERROR_INVALID_VALUE;
  hipStream_t st = rk_hip_stream(stream);
  if (length <= kSmallSort) return launch_small_sort(st, values, nullptr, length, 0, ws_flags_word(ws, ws_bytes));
  return rk_radix_sort_impl<uint32_t>(st, values, nullptr, length, 0, 4, ws, ws_bytes, nullptr);
}

// template <typename K, typename V> void CUDA_bitonicSortGeneral(K* keys, V* values, uint length),
// samples/09_how_to_use_a_sorter.cpp:42-43 / CUDA_bitonicSortGeneral.cu:504-690
// (instantiated for <uint, uint>, :694-697), at any length. The reference sorts
// each full 128-element chunk with a bitonic network, the partial last chunk
// by odd-even transposition, then merges chunk pairs and buckets of doubling
// size with stable '<=' merges (:172-481) until one bucket remains, carrying a
// local index array that finally permutes the values. The same order is the
// chunk networks followed by a stable sort on the whole key.
extern "C" rk_status rk_bitonic_sort_general(rk_stream stream, uint32_t* keys, uint32_t* values, uint32_t length,
                                             void* ws, size_t ws_bytes) {
  hipStream_t st = rk_hip_stream(stream);
  if (length >= kAgg) return RK_ERROR_UNSUPPORTED;
  if (length <= 1) return clear_ws_flags(st, ws, ws_bytes);
  if (!keys || !values) return RK_ERROR_INVALID_VALUE;
  hipLaunchKernelGGL(k_bitonic_chunks, dim3((length + 127) / 128), dim3(128), 0, st, keys, values, length);
  RK_CHECK_LAUNCH();
  if (length <= 128) return clear_ws_flags(st, ws, ws_bytes);
  return rk_radix_sort_impl<uint32_t>(st, keys, values, length, 0, 4, ws, ws_bytes, nullptr);
}

// void CUDA_bubbleSort512(uint* keys, uint* indices, uint length, bool init_indices),
// CUDA_radixSortGeneral.cu:27 / CUDA_bubbleSort512.cu:114-167: a stable sort of
// at most 512 (key, index) pairs; init_indices first sets indices[i] = i. The
// reference asserts on length > 512 (:116); here that is RK_ERROR_INVALID_VALUE.
extern "C" rk_status rk_bubble_sort512(rk_stream stream, uint32_t* keys, uint32_t* indices, uint32_t length,
                                       int init_indices) {
  if (length > kSmallSort) return RK_ERROR_INVALID_VALUE;
  if (length == 0) return RK_SUCCESS;
  if (!keys || !indices) return RK_ERROR_INVALID_VALUE;
  return launch_small_sort(rk_hip_stream(stream), keys, indices, length, init_indices, nullptr);
}

// The RK_FLAG_* bits a sort left in the first word of its workspace (waits for 'stream').
extern "C" rk_status rk_workspace_flags(rk_stream stream, const void* ws, uint32_t* flags) {
  if (!ws || !flags) return RK_ERROR_INVALID_VALUE;
  hipStream_t st = rk_hip_stream(stream);
  uint32_t h = 0;
  RK_HIP(hipMemcpyAsync(&h, ws, 4, hipMemcpyDeviceToHost, st));
  RK_HIP(hipStreamSynchronize(st));
  *flags = h;
  return RK_SUCCESS;
}
