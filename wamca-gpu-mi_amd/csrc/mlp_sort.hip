// mlp_sort.hip -- device radix sort of (uint64 key, int32 value) pairs for the
// TOTAL-mode merge.  rocPRIM's LSD radix sort is stable, so entries with
// equal keys keep their input order.  Kept in its own translation unit: the
// library templates are slow to compile.
#include <hip/hip_runtime.h>

#include <rocprim/device/device_radix_sort.hpp>

#include "mlp_launch.h"

hipError_t mlp_sort_pairs_u64(void* tmp, size_t* tmp_bytes, const uint64_t* kin, uint64_t* kout,
                              const int32_t* vin, int32_t* vout, int n, int end_bit,
                              hipStream_t st) {
  return rocprim::radix_sort_pairs(tmp, *tmp_bytes, kin, kout, vin, vout, (size_t)n, 0u,
                                   (unsigned)end_bit, st);
}
