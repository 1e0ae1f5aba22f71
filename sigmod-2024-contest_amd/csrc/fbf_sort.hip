// fbf_sort.hip -- the sorted layouts' row order, on the device.
//
// SortDataset (hybrid_graph.cpp:505-524) orders the rows by (label,
// timestamp) and by timestamp with std::sort; this build adds the row id as
// the last key, so the order is unique.  Each order is one stable LSD radix
// sort of packed keys -- (label, timestamp) as ((label ^ 2^31) << 32) |
// f2ord(timestamp), or f2ord(timestamp) alone over the low 32 bits -- with the
// row ids 0..n-1 as values: stability keeps equal keys in id order.
#include <rocprim/device/device_radix_sort.hpp>

#include "fbf_internal.hpp"

namespace fbf {

__global__ void iota_u32_kernel(uint32_t *__restrict__ out, uint64_t n) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) out[t] = (uint32_t)t;
}

hipError_t sort_rows_by_key(const uint64_t *keys_dev, uint64_t n, int end_bit, uint32_t *ids_out_dev,
                            hipStream_t st) {
    if (n == 0) return hipSuccess;
    if (n > 0xFFFFFFFFull || (end_bit != 32 && end_bit != 64)) return hipErrorInvalidValue;
    uint64_t *keys_out = nullptr;
    uint32_t *ids_in = nullptr;
    void *temp = nullptr;
    size_t temp_bytes = 0;
    hipError_t e = hipMalloc(&keys_out, n * sizeof(uint64_t));
    if (e == hipSuccess) e = hipMalloc(&ids_in, n * sizeof(uint32_t));
    if (e == hipSuccess) {
        hipLaunchKernelGGL(iota_u32_kernel, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, st, ids_in, n);
        e = hipGetLastError();
    }
    if (e == hipSuccess)
        e = rocprim::radix_sort_pairs(nullptr, temp_bytes, keys_dev, keys_out, ids_in, ids_out_dev, (size_t)n, 0,
                                      (unsigned)end_bit, st);
    if (e == hipSuccess) e = hipMalloc(&temp, temp_bytes > 0 ? temp_bytes : 1);
    if (e == hipSuccess)
        e = rocprim::radix_sort_pairs(temp, temp_bytes, keys_dev, keys_out, ids_in, ids_out_dev, (size_t)n, 0,
                                      (unsigned)end_bit, st);
    const hipError_t es = hipStreamSynchronize(st);
    if (e == hipSuccess) e = es;
    if (temp) (void)hipFree(temp);
    if (ids_in) (void)hipFree(ids_in);
    if (keys_out) (void)hipFree(keys_out);
    return e;
}

}  // namespace fbf
