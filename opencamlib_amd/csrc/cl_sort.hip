// Stable sort of CL points by their Hilbert cell key (the low key_bits) (rocPRIM LSD
// radix sort: equal keys keep input order, so the point -> wavefront
// assignment, and with it every result, is the same on every run).
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <cstring>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/iterator/counting_iterator.hpp>

namespace ocl_amd {

// keys_in: n keys; keys_out: n scratch keys; perm: n output indices (sorted
// order).  temp == nullptr: *temp_bytes receives the scratch size.
hipError_t cl_sort_by_key(void* temp, size_t* temp_bytes, const uint32_t* keys_in, uint32_t* keys_out,
                          uint32_t* perm, uint32_t n, uint32_t key_bits, hipStream_t st) {
    rocprim::counting_iterator<uint32_t> idx(0u);
    return rocprim::radix_sort_pairs(temp, *temp_bytes, keys_in, keys_out, idx, perm, n, 0u, key_bits, st);
}

}  // namespace ocl_amd
