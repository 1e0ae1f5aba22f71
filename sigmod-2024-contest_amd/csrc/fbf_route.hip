// fbf_route.hip -- the router's range step and query order on the device.
//
// route_kernel: one thread per query turns its filter into a contiguous
// range [s, e) of a sorted layout (hybrid_graph.cpp:176-199: lower / upper
// bound over sorted_both by label or (label, timestamp), over
// sorted_timestamp by timestamp; type 0 the whole timestamp layout), forms
// the selectivity (e - s) / nb (:152, :183-198) and the brute-force decision
// sel < bf_thr || (type == 3 && sel < bf_thr3) (:200-201).
//
// sort_queries: the batch order of ReadSortedQuery (utils.h:514-547):
// ascending (uint32 type, int32 v, float l, float r), with the query index as
// the last key so the order is unique (std::sort leaves ties unspecified).
// Two stable LSD radix passes over packed keys: (l, r) first, then (type, v).
#include <rocprim/device/device_radix_sort.hpp>

#include "fbf_internal.hpp"

namespace fbf {

__global__ void route_kernel(const float *__restrict__ metas, uint64_t nq, const int32_t *__restrict__ lab_b,
                             const float *__restrict__ ts_b, const float *__restrict__ ts_t, uint64_t n, float nb,
                             float bf_thr, float bf_thr3, uint32_t *__restrict__ ranges,
                             uint64_t *__restrict__ key_lr, uint64_t *__restrict__ key_tv, uint32_t *__restrict__ idx) {
    const uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nq) return;
    const float tf = metas[4 * q], vf = metas[4 * q + 1], l = metas[4 * q + 2], r = metas[4 * q + 3];
    const int type = (int)tf;
    uint32_t layout = 1, s = 0, e = (uint32_t)n, bf = 0;
    if (!(tf == (float)type) || type < 0 || type > 3) {
        layout = 0xFFFFFFFFu;  // unknown filter type: no range
        e = 0;
    } else if (type == 2) {
        uint64_t lo = 0, hi = n;
        while (lo < hi) {
            const uint64_t m = (lo + hi) / 2;
            if (ts_t[m] < l) lo = m + 1; else hi = m;
        }
        s = (uint32_t)lo;
        lo = 0; hi = n;
        while (lo < hi) {
            const uint64_t m = (lo + hi) / 2;
            if (ts_t[m] <= r) lo = m + 1; else hi = m;
        }
        e = (uint32_t)(lo > s ? lo : s);
    } else if (type != 0) {
        layout = 0;
        const int32_t v = (int32_t)vf;  // labels compare as integral values (utils.h:99,109)
        uint64_t lo = 0, hi = n;
        while (lo < hi) {
            const uint64_t m = (lo + hi) / 2;
            const bool below = lab_b[m] < v || (type == 3 && lab_b[m] == v && ts_b[m] < l);
            if (below) lo = m + 1; else hi = m;
        }
        s = (uint32_t)lo;
        lo = 0; hi = n;
        while (lo < hi) {
            const uint64_t m = (lo + hi) / 2;
            const bool notabove = lab_b[m] < v || (lab_b[m] == v && (type == 1 || ts_b[m] <= r));
            if (notabove) lo = m + 1; else hi = m;
        }
        e = (uint32_t)(lo > s ? lo : s);
    }
    if (layout != 0xFFFFFFFFu) {
        // type 0 keeps sel = 1 (hybrid_graph.cpp:175)
        const float sel = type == 0 ? 1.0f : (float)(e - s) / nb;
        bf = (sel < bf_thr || (type == 3 && sel < bf_thr3)) ? 1u : 0u;
    }
    ranges[4 * q + 0] = layout;
    ranges[4 * q + 1] = s;
    ranges[4 * q + 2] = e;
    ranges[4 * q + 3] = bf;
    // Sort keys (ReadSortedQuery's comparator: uint32 type, int32 v, then
    // float l, r; -0.0 orders as +0.0, as the float compares do).
    const uint32_t tu = (uint32_t)tf;
    const int32_t vi = (int32_t)vf;
    key_tv[q] = ((uint64_t)tu << 32) | ((uint32_t)vi ^ 0x80000000u);
    key_lr[q] = ((uint64_t)f2ord(l == 0.0f ? 0.0f : l) << 32) | f2ord(r == 0.0f ? 0.0f : r);
    idx[q] = (uint32_t)q;
}

__global__ void gather_u64_kernel(const uint64_t *__restrict__ src, const uint32_t *__restrict__ idx, uint64_t n,
                                  uint64_t *__restrict__ dst) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[idx[i]];
}

// Temporary storage of the two radix passes for nq queries (bytes).
size_t route_temp_bytes(uint64_t nq) {
    size_t t = 0;
    (void)rocprim::radix_sort_pairs(nullptr, t, (const uint64_t *)nullptr, (uint64_t *)nullptr,
                                    (const uint32_t *)nullptr, (uint32_t *)nullptr, (size_t)nq, 0, 64);
    return t;
}

hipError_t launch_route(const float *metas, uint64_t nq, const int32_t *lab_b, const float *ts_b, const float *ts_t,
                        uint64_t n, float nb, float bf_thr, float bf_thr3, uint32_t *ranges, uint32_t *order,
                        uint64_t *keys, uint32_t *vals, void *temp, size_t temp_bytes, hipStream_t st) {
    if (nq == 0) return hipSuccess;
    // keys: 4 * nq u64 (lr, tv, sorted, gathered); vals: 2 * nq u32 (iota, pass-1 order)
    uint64_t *key_lr = keys, *key_tv = keys + nq, *k_sorted = keys + 2 * nq, *k_gather = keys + 3 * nq;
    uint32_t *idx = vals, *ord1 = vals + nq;
    const uint32_t blocks = (uint32_t)((nq + 255) / 256);
    hipLaunchKernelGGL(route_kernel, dim3(blocks), dim3(256), 0, st, metas, nq, lab_b, ts_b, ts_t, n, nb, bf_thr,
                       bf_thr3, ranges, key_lr, key_tv, idx);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || order == nullptr) return e;
    size_t tb = temp_bytes;
    e = rocprim::radix_sort_pairs(temp, tb, key_lr, k_sorted, idx, ord1, (size_t)nq, 0, 64, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(gather_u64_kernel, dim3(blocks), dim3(256), 0, st, key_tv, ord1, nq, k_gather);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    tb = temp_bytes;
    return rocprim::radix_sort_pairs(temp, tb, k_gather, k_sorted, ord1, order, (size_t)nq, 0, 64, st);
}

}  // namespace fbf
