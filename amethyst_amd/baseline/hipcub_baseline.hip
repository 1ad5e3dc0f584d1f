// hipcub_baseline.hip -- the GPU library baselines the reference times beside
// its own kernels, on this GPU: hipCUB (rocPRIM) over unpacked data, one u64
// per element (the reference's "baseline" format).
//   src/benchs/bench_compressstore_baseline_cub.cu:29-35  cub::DeviceSelect::Flagged
//       over u64 elements with a bit-iterator mask
//   src/benchs/bench_hop_baseline_cub.cu:32-38            cub::DeviceReduce over u64
//       elements with the HOP operator (sum)
// Measurement only: bench.py times these next to the library's own operators
// and a GPU test checks their results against them. Nothing in amethyst_amd
// calls this library.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>
#include <cstdint>

namespace {

// flag i = bit i of an LSB-first u64 mask (the library's mask layout)
struct MaskBit {
    const uint64_t* mask;
    __host__ __device__ bool operator()(uint64_t i) const { return (mask[i >> 6] >> (i & 63)) & 1; }
};

} // namespace

extern "C" {

// Compaction of n u64 elements by an LSB-first bitmask into `out`, the count
// into *count_dev. With temp == nullptr only *temp_bytes is set. 0 on success.
int hb_select_flagged_u64(const uint64_t* in, const uint64_t* mask, uint64_t* out, uint64_t* count_dev, uint64_t n,
                          void* temp, size_t* temp_bytes, hipStream_t s)
{
    hipcub::CountingInputIterator<uint64_t> idx(0);
    hipcub::TransformInputIterator<bool, MaskBit, hipcub::CountingInputIterator<uint64_t>> flags(idx, MaskBit{mask});
    // (32-bit item counts where they fit: the library's 32-bit offset path)
    hipError_t e = n < (1ull << 31)
                       ? hipcub::DeviceSelect::Flagged(temp, *temp_bytes, in, flags, out, (unsigned int*)count_dev,
                                                       (int)n, s)
                       : hipcub::DeviceSelect::Flagged(temp, *temp_bytes, in, flags, out,
                                                       (unsigned long long*)count_dev, (int64_t)n, s);
    return e == hipSuccess ? 0 : 1;
}

// The same with one flag byte per element (nonzero: selected), the form
// rocPRIM reads without an index transform.
int hb_select_flagged_u64_bytes(const uint64_t* in, const uint8_t* flags, uint64_t* out, uint64_t* count_dev,
                                uint64_t n, void* temp, size_t* temp_bytes, hipStream_t s)
{
    hipError_t e = n < (1ull << 31)
                       ? hipcub::DeviceSelect::Flagged(temp, *temp_bytes, in, flags, out, (unsigned int*)count_dev,
                                                       (int)n, s)
                       : hipcub::DeviceSelect::Flagged(temp, *temp_bytes, in, flags, out,
                                                       (unsigned long long*)count_dev, (int64_t)n, s);
    return e == hipSuccess ? 0 : 1;
}

// Wrapping sum of n u64 elements into *out_dev. With temp == nullptr only
// *temp_bytes is set. 0 on success.
int hb_reduce_sum_u64(const uint64_t* in, uint64_t* out_dev, uint64_t n, void* temp, size_t* temp_bytes, hipStream_t s)
{
    hipError_t e = n < (1ull << 31) ? hipcub::DeviceReduce::Sum(temp, *temp_bytes, in, out_dev, (int)n, s)
                                     : hipcub::DeviceReduce::Sum(temp, *temp_bytes, in, out_dev, (int64_t)n, s);
    return e == hipSuccess ? 0 : 1;
}

} // extern "C"
