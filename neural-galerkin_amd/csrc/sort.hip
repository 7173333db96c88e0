// Stable radix sorts (rocPRIM) used for point bucketing and vertex de-duplication.
#include <rocprim/device/device_radix_sort.hpp>

#include "common.h"

namespace ngs {

int sort_pairs_u32(const uint32_t* kin, uint32_t* kout, const int32_t* vin, int32_t* vout, int64_t n, int bits,
                   void* tmp, size_t* tmp_bytes, hipStream_t st) {
    hipError_t e = rocprim::radix_sort_pairs(tmp, *tmp_bytes, kin, kout, vin, vout, (size_t)n, 0, bits, st);
    if (e != hipSuccess) {
        set_error(std::string("radix_sort_pairs(u32): ") + hipGetErrorString(e));
        return NGS_ERR_HIP;
    }
    return NGS_OK;
}

int sort_pairs_u64(const uint64_t* kin, uint64_t* kout, const int32_t* vin, int32_t* vout, int64_t n, int bits,
                   void* tmp, size_t* tmp_bytes, hipStream_t st) {
    hipError_t e = rocprim::radix_sort_pairs(tmp, *tmp_bytes, kin, kout, vin, vout, (size_t)n, 0, bits, st);
    if (e != hipSuccess) {
        set_error(std::string("radix_sort_pairs(u64): ") + hipGetErrorString(e));
        return NGS_ERR_HIP;
    }
    return NGS_OK;
}

}  // namespace ngs
