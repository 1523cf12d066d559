// The device-resident point batches' sorts and scans, by hipCUB: a stable
// partition by kind (radix sort on the 2-bit kind keys carrying the point
// indices in their original order within each kind), the stable sort by
// face of a gradient batch, and the inclusive scan of its segment heads.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include "umcp_kernels.hpp"

namespace umcp {

hipError_t point_partition_by_kind(const unsigned char *keys_in, unsigned char *keys_out, const int *vals_in,
                                   int *vals_out, int n, void *temp, size_t *temp_bytes, hipStream_t st)
{
    return hipcub::DeviceRadixSort::SortPairs(temp, *temp_bytes, keys_in, keys_out, vals_in, vals_out, n, 0, 2, st);
}

hipError_t point_sort_by_face(const int *keys_in, int *keys_out, const int *vals_in, int *vals_out, int n, int end_bit,
                              void *temp, size_t *temp_bytes, hipStream_t st)
{
    return hipcub::DeviceRadixSort::SortPairs(temp, *temp_bytes, keys_in, keys_out, vals_in, vals_out, n, 0, end_bit,
                                              st);
}

hipError_t point_inclusive_sum(const int *in, int *out, int n, void *temp, size_t *temp_bytes, hipStream_t st)
{
    return hipcub::DeviceScan::InclusiveSum(temp, *temp_bytes, in, out, n, st);
}

}  // namespace umcp
