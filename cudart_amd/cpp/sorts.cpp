// The reference's sort entry points (include/renderkit/CUDASort.h) over the C ABI.
// Each call allocates its workspace and frees it before returning, as the
// reference allocates and frees its temporaries inside every call
// (CUDA_radixSortGeneral.cu:437-439, CUDA_bitonicSortGeneral.cu:508,573-574).
#include "../../include/renderkit/CUDASort.h"

namespace {

struct SortWorkspace {
  void* p = nullptr;
  size_t bytes = 0;
  explicit SortWorkspace(uint length) : bytes(rk_sort_workspace_size(length, 4)) {
    renderkit::_assert(rk_malloc(&p, bytes) == RK_SUCCESS, __FILE__, __LINE__, "Unable to allocate a sort workspace.");
  }
  ~SortWorkspace() { if (p) rk_free(p); }
  // Waits for the sort and fails like the reference's _assert if a look-back
  // wait gave up (the output would not be sorted).
  void finish(const char* what) {
    uint32_t flags = 0;
    RK_CALL(rk_workspace_flags(nullptr, p, &flags));
    renderkit::_assert((flags & RK_FLAG_SPIN_LIMIT) == 0, __FILE__, __LINE__, what);
  }
};

}  // namespace

void CUDA_radixSort(uint* values, uint length, uint bit_depth) {
  SortWorkspace ws(length);
  RK_CALL(rk_radix_sort(nullptr, values, length, bit_depth, ws.p, ws.bytes));
  ws.finish("CUDA_radixSort() : sort did not complete.");
}

void CUDA_radixSortGeneral(uint* keys, uint* indices, uint length, uint bit_depth, bool init_indices) {
  SortWorkspace ws(length);
  RK_CALL(rk_radix_sort_general(nullptr, keys, indices, length, bit_depth, init_indices ? 1 : 0, ws.p, ws.bytes));
  ws.finish("CUDA_radixSortGeneral() : sort did not complete.");
}

template <>
void CUDA_bitonicSort<unsigned int>(unsigned int* values, uint length) {
  SortWorkspace ws(length);
  RK_CALL(rk_bitonic_sort(nullptr, values, length, ws.p, ws.bytes));
  ws.finish("CUDA_bitonicSort() : sort did not complete.");
}

template <>
void CUDA_bitonicSortGeneral<unsigned int, unsigned int>(unsigned int* keys, unsigned int* values, uint length) {
  SortWorkspace ws(length);
  RK_CALL(rk_bitonic_sort_general(nullptr, keys, values, length, ws.p, ws.bytes));
  ws.finish("CUDA_bitonicSortGeneral() : sort did not complete.");
}

void CUDA_bubbleSort512(uint* keys, uint* indices, uint length, bool init_indices) {
  renderkit::_assert(length <= 512, __FILE__, __LINE__, "CUDA_bubbleSort512() : Unable to execute CUDA kernel.");
  RK_CALL(rk_bubble_sort512(nullptr, keys, indices, length, init_indices ? 1 : 0));
  RK_CALL(rk_stream_synchronize(nullptr));
}
