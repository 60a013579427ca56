// k_prims.hip -- device-wide scan and radix sort (hipCUB), isolated in one
// translation unit.  With tmp == nullptr each call only reports tmp_bytes.
#include <hipcub/hipcub.hpp>

#include "kernels.hpp"

hipError_t prim_exclusive_scan_u32(const uint32_t* in, uint32_t* out, uint64_t m, void* tmp,
                                   size_t& tmp_bytes, hipStream_t s) {
    return hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, in, out, (int64_t)m, s);
}

hipError_t prim_sort_pairs_u64_u32(const uint64_t* kin, uint64_t* kout, const uint32_t* vin,
                                   uint32_t* vout, uint64_t m, int end_bit, void* tmp,
                                   size_t& tmp_bytes, hipStream_t s) {
    return hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, kin, kout, vin, vout, (int64_t)m, 0, end_bit, s);
}
