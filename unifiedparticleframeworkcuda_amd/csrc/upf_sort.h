// upf_sort.h -- device radix sort entry points.
#ifndef UPF_SORT_H_
#define UPF_SORT_H_

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace upf {

size_t radix_sort_temp_bytes(int n);
hipError_t radix_sort(const uint32_t* keys, const uint32_t* vals, uint32_t* keys_out, uint32_t* vals_out, int n,
                      int bits, void* temp, hipStream_t st);

}  // namespace upf

#endif
