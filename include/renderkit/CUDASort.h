// renderkit drop-in: the reference's sort entry points as C++ free functions at
// global scope. The reference has no header for them: its callers declare
// prototypes (samples/09_how_to_use_a_sorter.cpp:36-44,
// kernels/CUDA_buildMeshTreeLBVH.cu:23-29, kernels/CUDA_radixSortGeneral.cu:27-30,
// kernels/CUDA_radixSort.cu:27). A caller may include this header or keep its own
// prototypes; both link against librenderkit.so. Every function takes device
// pointers, runs on the default stream, returns after the sort has finished and,
// like the reference's _assert, reports a failure and ends the process (or throws
// after renderkit::setThrowOnError(true)).
#ifndef RENDERKIT_CUDASORT_H
#define RENDERKIT_CUDASORT_H

#include "CUDACommon.h"

// kernels/CUDA_radixSort.cu:357-392: values sorted on their low 8*ceil(bit_depth/8) bits.
void CUDA_radixSort(uint* values, uint length, uint bit_depth = 32);
// kernels/CUDA_radixSortGeneral.cu:397-482: (key, index) pairs, the reference tie order by size class.
void CUDA_radixSortGeneral(uint* keys, uint* indices, uint length, uint bit_depth = 32, bool init_indices = false);
// kernels/CUDA_bitonicSort.cu:402-518 (unsigned int only, :523-537).
template <typename value_type>
void CUDA_bitonicSort(value_type* values, uint length);
// kernels/CUDA_bitonicSortGeneral.cu:504-690 (<uint, uint> only, :694-697), declared as its
// callers declare it (samples/09_how_to_use_a_sorter.cpp:42-43, CUDA_radixSortGeneral.cu:29-30);
// the (key_type* keys, value_type* values) form of its definition is exported as well.
template <typename key_type, typename value_type>
void CUDA_bitonicSortGeneral(value_type* keys, value_type* values, uint length);
// kernels/CUDA_bubbleSort512.cu:114-167: length <= 512.
void CUDA_bubbleSort512(uint* keys, uint* indices, uint length, bool init_indices = false);

#endif
