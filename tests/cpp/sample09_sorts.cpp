// The sorts of the reference's sample 09 (samples/09_how_to_use_a_sorter.cpp),
// called through prototypes declared here exactly as the sample declares them
// (:36-44), so the program links against librenderkit.so the way the sample would.
// Device memory comes from the C ABI. Values are (x * RAND_MAX + y) % 2097152 as in
// the sample, from a seeded generator; bitonicSortGeneral gets value = position so
// that its tie order is visible. Every result is checked sorted, as the sample does,
// and written to one binary file for the test to compare with the oracle.
//
//   sample09_sorts <out.bin> <length> <seed>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../include/rk.h"

typedef unsigned int uint;

void CUDA_radixSort( uint* values, uint length, uint bit_depth = 32 );

void CUDA_radixSortGeneral( uint* keys, uint* indices, uint length, uint bit_depth = 32, bool init_indices = false );

template <typename value_type>
void CUDA_bitonicSort( value_type* values, uint length );

template <typename key_type, typename value_type>
void CUDA_bitonicSortGeneral( value_type* keys, value_type* values, uint length );

void CUDA_bubbleSort512( uint* keys, uint* indices, uint length, bool init_indices = false );

static uint64_t g_state;
static uint32_t next_u32() {  // SplitMix64
  uint64_t z = (g_state += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return (uint32_t)((z ^ (z >> 31)) >> 32);
}

static void check(rk_status s, const char* what) {
  if (s != RK_SUCCESS) { std::fprintf(stderr, "%s: %s\n", what, rk_status_string(s)); std::exit(1); }
}

static void dump(FILE* f, const std::vector<uint>& h) {
  uint32_t bytes = (uint32_t)(h.size() * 4);
  std::fwrite(&bytes, 4, 1, f);
  std::fwrite(h.data(), 1, bytes, f);
}

int main(int argc, char** argv) {
  if (argc < 4) { std::fprintf(stderr, "usage: %s out.bin length seed\n", argv[0]); return 2; }
  const uint n = (uint)std::strtoul(argv[2], nullptr, 10);
  g_state = std::strtoull(argv[3], nullptr, 10);
  std::vector<uint> input(n), h(n), hidx(n);
  for (uint k = 0; k < n; ++k) input[k] = ((next_u32() & 0x7FFF) * 32767u + (next_u32() & 0x7FFF)) % 2097152u;
  uint *d_keys = nullptr, *d_vals = nullptr;
  check(rk_malloc((void**)&d_keys, (size_t)n * 4 + 4), "malloc");
  check(rk_malloc((void**)&d_vals, (size_t)n * 4 + 4), "malloc");
  auto up = [&](uint* d, const std::vector<uint>& v) { check(rk_memcpy(d, v.data(), (size_t)n * 4, 0, nullptr), "h2d"); };
  auto down = [&](std::vector<uint>& v, const uint* d) { check(rk_memcpy(v.data(), d, (size_t)n * 4, 1, nullptr), "d2h"); };
  auto sorted = [&](const std::vector<uint>& v, const char* what) {
    for (uint k = 1; k < v.size(); ++k)
      if (v[k] < v[k - 1]) { std::fprintf(stderr, "%s not sorted at %u\n", what, k); std::exit(1); }
  };
  FILE* f = std::fopen(argv[1], "wb");
  if (!f) return 1;
  std::fwrite(&n, 4, 1, f);
  dump(f, input);

  // testRadixSort (:124-156): values, 21 bits
  up(d_keys, input);
  CUDA_radixSort(d_keys, n, 21);
  down(h, d_keys); sorted(h, "radixSort"); dump(f, h);

  // testBitonicSort (:100-122): values
  up(d_keys, input);
  CUDA_bitonicSort<uint>(d_keys, n);
  down(h, d_keys); sorted(h, "bitonicSort"); dump(f, h);

  // testRadixSortGeneral (:197-232): keys with indices 0..n-1 given, 21 bits
  for (uint k = 0; k < n; ++k) hidx[k] = k;
  up(d_keys, input); up(d_vals, hidx);
  CUDA_radixSortGeneral(d_keys, d_vals, n, 21, false);
  down(h, d_keys); sorted(h, "radixSortGeneral"); dump(f, h);
  down(h, d_vals); dump(f, h);

  // testBitonicSortGeneral (:158-195): keys with values = positions
  up(d_keys, input); up(d_vals, hidx);
  CUDA_bitonicSortGeneral<uint, uint>(d_keys, d_vals, n);
  down(h, d_keys); sorted(h, "bitonicSortGeneral"); dump(f, h);
  down(h, d_vals); dump(f, h);

  // CUDA_bubbleSort512 on the first min(n, 512) keys, indices initialised
  const uint m = n < 512 ? n : 512;
  std::vector<uint> hk(input.begin(), input.begin() + m), hv(m, 0xFFFFFFFFu);
  check(rk_memcpy(d_keys, hk.data(), (size_t)m * 4, 0, nullptr), "h2d");
  check(rk_memcpy(d_vals, hv.data(), (size_t)m * 4, 0, nullptr), "h2d");
  CUDA_bubbleSort512(d_keys, d_vals, m, true);
  check(rk_memcpy(hk.data(), d_keys, (size_t)m * 4, 1, nullptr), "d2h");
  check(rk_memcpy(hv.data(), d_vals, (size_t)m * 4, 1, nullptr), "d2h");
  sorted(hk, "bubbleSort512"); dump(f, hk); dump(f, hv);

  std::fclose(f);
  rk_free(d_keys);
  rk_free(d_vals);
  return 0;
}
