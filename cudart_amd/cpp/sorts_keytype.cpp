// CUDA_bitonicSortGeneral in the form of its definition,
// template <typename key_type, typename value_type> (key_type* keys, value_type* values, uint length)
// (kernels/CUDA_bitonicSortGeneral.cu:504-505). Its callers declare (value_type*, value_type*),
// which is a different symbol; that form is in sorts.cpp. The two are kept in separate
// translation units so that each explicit specialization names one primary template.
#include "../../include/renderkit/CUDACommon.h"

template <typename key_type, typename value_type>
void CUDA_bitonicSortGeneral(key_type* keys, value_type* values, uint length);

template <>
void CUDA_bitonicSortGeneral<unsigned int, unsigned int>(unsigned int* keys, unsigned int* values, uint length) {
  void* ws = nullptr;
  const size_t bytes = rk_sort_workspace_size(length, 4);
  renderkit::_assert(rk_malloc(&ws, bytes) == RK_SUCCESS, __FILE__, __LINE__, "Unable to allocate a sort workspace.");
  rk_status s = rk_bitonic_sort_general(nullptr, keys, values, length, ws, bytes);
  uint32_t flags = 0;
  if (s == RK_SUCCESS) s = rk_workspace_flags(nullptr, ws, &flags);
  rk_free(ws);
  RK_CALL(s);
  renderkit::_assert((flags & RK_FLAG_SPIN_LIMIT) == 0, __FILE__, __LINE__, "CUDA_bitonicSortGeneral() : sort did not complete.");
}
