// rocPRIM device primitives: stable radix sorts (point bucketing, vertex de-duplication) and
// single-pass exclusive scans (decoupled look-back: one read of the input, one write of the output).
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>
#include <rocprim/iterator/counting_iterator.hpp>
#include <rocprim/iterator/transform_iterator.hpp>

#include <algorithm>

#include "common.h"

namespace ngs {

namespace {
// rocPRIM sorts up to 1M items by merge sort whatever the key width. For keys of at most 24 bits
// (three 8-bit digit passes) its onesweep radix path is faster from 256k items on gfx950 (u32 keys
// of 14-20 bits with int32 values: 300k items 115 -> 94 us, 1M items 160 -> 102 us; at 100k items
// merge sort is faster, 50 against 90 us). Both sorts are stable, so the output is the same.
using onesweep_only = rocprim::radix_sort_config<rocprim::default_config, rocprim::default_config,
                                                 rocprim::default_config, 0>;
constexpr int64_t kOnesweepMinItems = int64_t(1) << 18;
constexpr int kOnesweepMaxBits = 24;
}  // namespace

// A size query (tmp == nullptr) returns enough for either algorithm at up to `bits` key bits, so a
// caller may size its workspace for wider keys than it then sorts.
int sort_pairs_u32(const uint32_t* kin, uint32_t* kout, const int32_t* vin, int32_t* vout, int64_t n, int bits,
                   void* tmp, size_t* tmp_bytes, hipStream_t st) {
    hipError_t e;
    if (tmp == nullptr) {
        size_t a = 0, b = 0;
        e = rocprim::radix_sort_pairs(nullptr, a, kin, kout, vin, vout, (size_t)n, 0, bits, st);
        if (e == hipSuccess && n >= kOnesweepMinItems)
            e = rocprim::radix_sort_pairs<onesweep_only>(nullptr, b, kin, kout, vin, vout, (size_t)n, 0,
                                                         std::min(bits, kOnesweepMaxBits), st);
        *tmp_bytes = std::max(a, b);
    } else if (n >= kOnesweepMinItems && bits <= kOnesweepMaxBits) {
        e = rocprim::radix_sort_pairs<onesweep_only>(tmp, *tmp_bytes, kin, kout, vin, vout, (size_t)n, 0, bits, st);
    } else {
        e = rocprim::radix_sort_pairs(tmp, *tmp_bytes, kin, kout, vin, vout, (size_t)n, 0, bits, st);
    }
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

namespace {
// Element i of the (n + 1)-long sequence counts[0], ..., counts[n - 1], 0: its exclusive scan ends with the total.
struct PaddedCount {
    const int32_t* v;
    int64_t n;
    __host__ __device__ int32_t operator()(int64_t i) const { return i < n ? v[i] : 0; }
};
struct WordPopc {
    const uint32_t* w;
    __host__ __device__ int32_t operator()(int64_t i) const { return __builtin_popcount(w[i]); }
};

template <class F>
int scan_seq(F f, int64_t len, int32_t* out, void* tmp, size_t* tmp_bytes, hipStream_t st, const char* what) {
    auto in = rocprim::make_transform_iterator(rocprim::counting_iterator<int64_t>(0), f);
    hipError_t e = rocprim::exclusive_scan(tmp, *tmp_bytes, in, out, int32_t(0), (size_t)len, rocprim::plus<int32_t>(), st);
    if (e != hipSuccess) {
        set_error(std::string(what) + ": " + hipGetErrorString(e));
        return NGS_ERR_HIP;
    }
    return NGS_OK;
}
}  // namespace

int scan_counts_device(const int32_t* counts, int64_t n, int32_t* offsets, void* tmp, size_t* tmp_bytes, hipStream_t st) {
    return scan_seq(PaddedCount{counts, n}, n + 1, offsets, tmp, tmp_bytes, st, "exclusive_scan(i32)");
}

int scan_popc_device(const uint32_t* words, int64_t n, int32_t* rank, void* tmp, size_t* tmp_bytes, hipStream_t st) {
    return scan_seq(WordPopc{words}, n, rank, tmp, tmp_bytes, st, "exclusive_scan(popc)");
}

}  // namespace ngs
