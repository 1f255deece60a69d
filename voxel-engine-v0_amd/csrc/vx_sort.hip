// vx_sort.hip — the device radix sort that orders a call's ray queries before
// they are traced (rocPRIM's radix_sort_pairs; its own translation unit, so
// the library's templates compile once and apart from the pass kernels).
#include <hip/hip_runtime.h>

#include <rocprim/device/device_radix_sort.hpp>

#include "vx_launch.h"

hipError_t vx_sort_pairs_u32(void* temp, size_t* temp_bytes, const uint32_t* keys_in, uint32_t* keys_out,
                             const uint32_t* vals_in, uint32_t* vals_out, uint32_t n, uint32_t end_bit, hipStream_t st) {
  return rocprim::radix_sort_pairs(temp, *temp_bytes, keys_in, keys_out, vals_in, vals_out, n, 0u, end_bit, st);
}
