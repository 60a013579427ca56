// k_prims.hip -- device-wide exclusive scan and LSD radix sort from rocPRIM,
// isolated in one translation unit so its templates compile once.  Both
// follow rocPRIM's two-phase protocol: with tmp == nullptr a call only
// stores the work-area size it needs in tmp_bytes.
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include "kernels.hpp"

// out[i] = in[0] + ... + in[i-1] (out[0] = 0), m entries.
hipError_t prim_exclusive_scan_u32(const uint32_t* in, uint32_t* out, uint64_t m, void* tmp,
                                   size_t& tmp_bytes, hipStream_t s) {
    return rocprim::exclusive_scan(tmp, tmp_bytes, in, out, 0u, (size_t)m, rocprim::plus<uint32_t>(), s);
}

// Stable sort of (key, value) pairs by key bits [0, end_bit).
hipError_t prim_sort_pairs_u64_u32(const uint64_t* kin, uint64_t* kout, const uint32_t* vin,
                                   uint32_t* vout, uint64_t m, int end_bit, void* tmp,
                                   size_t& tmp_bytes, hipStream_t s) {
    return rocprim::radix_sort_pairs(tmp, tmp_bytes, kin, kout, vin, vout, (size_t)m, 0u, (unsigned)end_bit, s);
}
