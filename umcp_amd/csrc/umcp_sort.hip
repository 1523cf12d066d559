// Stable partition of a batch of points by kind (the device-resident point
// evaluation): hipCUB's radix sort on the 2-bit kind keys carries the point
// indices along in their original order within each kind.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include "umcp_kernels.hpp"

namespace umcp {

hipError_t point_partition_by_kind(const unsigned char *keys_in, unsigned char *keys_out, const int *vals_in,
                                   int *vals_out, int n, void *temp, size_t *temp_bytes, hipStream_t st)
{
    return hipcub::DeviceRadixSort::SortPairs(temp, *temp_bytes, keys_in, keys_out, vals_in, vals_out, n, 0, 2, st);
}

}  // namespace umcp
