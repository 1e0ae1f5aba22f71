// fbf_select.hpp -- wave-level bounded top-K on 64-bit (distance, id) keys.
//
// Every work item is one wave64 workgroup.  Candidates that beat the current
// threshold of their query are appended to that query's list; when a list
// could overflow on the next tile it is sorted (bitonic, in LDS) and cut back
// to its K smallest keys, and the threshold becomes the K-th key.  Keys are
// unique (the low word is the row id), so "key < threshold" is an exact test:
// a key at or above the K-th smallest of any subset can never enter the
// top-K of the whole set.  This replaces maxPQIFCS (bruteforce.h:57-117) and
// ResultMaxHeap (utils.h:34-78), whose per-element heap walk does not map to
// 64 lanes.
#pragma once

#include "fbf_internal.hpp"

namespace fbf {

__device__ inline uint32_t lane_id() { return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u)); }

// Number of set bits of m below this lane.
__device__ inline uint32_t lanes_below(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

__device__ inline uint64_t wave_ballot(bool p) { return __ballot(p); }

__device__ inline uint32_t uniform_u32(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ inline uint64_t uniform_u64(uint64_t v) {
    uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return ((uint64_t)hi << 32) | lo;
}

// Synchronisation policies for the list routines.  BlockSync: the workgroup
// is one wave, __syncthreads() orders its LDS and global traffic.  WaveSync:
// a wave of a larger workgroup working on data only it touches: wait for all
// its outstanding memory operations, then a wave barrier (no cross-wave wait).
struct BlockSync {
    __device__ static void sync() { __syncthreads(); }
};
struct WaveSync {
    __device__ static void sync() {
        __asm__ volatile("" ::: "memory");
        __builtin_amdgcn_s_waitcnt(0);
        __builtin_amdgcn_wave_barrier();
        __asm__ volatile("" ::: "memory");
    }
};

// Ascending bitonic sort of B keys held in LDS by one wave.
template <int B, class Sync = BlockSync>
__device__ inline void wave_sort_lds(uint64_t *s) {
    const uint32_t lane = threadIdx.x & 63;
#pragma unroll 1
    for (int size = 2; size <= B; size <<= 1) {
#pragma unroll 1
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
#pragma unroll 2
            for (int p0 = 0; p0 < B / 2; p0 += 64) {
                const int p = p0 + (int)lane;
                const int i = 2 * p - (p & (stride - 1));
                const int j = i + stride;
                const bool up = (i & size) == 0;
                uint64_t x = s[i], y = s[j];
                if ((x > y) == up) {
                    s[i] = y;
                    s[j] = x;
                }
            }
            Sync::sync();
        }
    }
}

// Ascending bitonic sort of the first n keys in LDS (n a power of two, at
// least 2) by one wave.
template <class Sync = BlockSync>
__device__ inline void wave_sort_lds_n(uint64_t *s, uint32_t n) {
    const uint32_t lane = threadIdx.x & 63;
#pragma unroll 1
    for (uint32_t size = 2; size <= n; size <<= 1) {
#pragma unroll 1
        for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
#pragma unroll 1
            for (uint32_t p0 = 0; p0 < n / 2; p0 += 64) {
                const uint32_t p = p0 + lane;
                if (p < n / 2) {
                    const uint32_t i = 2 * p - (p & (stride - 1));
                    const uint32_t j = i + stride;
                    const bool up = (i & size) == 0;
                    const uint64_t x = s[i], y = s[j];
                    if ((x > y) == up) {
                        s[i] = y;
                        s[j] = x;
                    }
                }
            }
            Sync::sync();
        }
    }
}

// The smallest power of two at or above max(c, 2).
__device__ inline uint32_t pow2_at_least(uint32_t c) {
    uint32_t n = 2;
    while (n < c) n <<= 1;
    return n;
}

// A query's candidate list in global memory (list) with its count and
// threshold.  The LDS buffer sbuf (B keys) is scratch for compaction.
template <int B, class Sync = BlockSync>
__device__ inline void compact_global(uint64_t *list, uint32_t &cnt, uint64_t &tau, uint32_t K,
                                      uint64_t *sbuf) {
    const uint32_t lane = threadIdx.x & 63;
    Sync::sync();  // list stores of this wave complete before they are read back
#pragma unroll
    for (int i0 = 0; i0 < B; i0 += 64) {
        const uint32_t i = i0 + lane;
        sbuf[i] = i < cnt ? list[i] : kKeyMax;
    }
    Sync::sync();
    wave_sort_lds<B, Sync>(sbuf);
    const uint32_t keep = cnt < K ? cnt : K;
#pragma unroll
    for (int i0 = 0; i0 < B; i0 += 64) {
        const uint32_t i = i0 + lane;
        if (i < keep) list[i] = sbuf[i];
    }
    tau = (keep == K) ? uniform_u64(sbuf[K - 1]) : kKeyMax;
    cnt = keep;
    Sync::sync();
}

// The same with the list itself in LDS (merge kernels); only the smallest
// power of two of keys holding the list is sorted.
template <int B>
__device__ inline void compact_lds(uint64_t *list, uint32_t &cnt, uint64_t &tau, uint32_t K) {
    const uint32_t lane = threadIdx.x;
    const uint32_t n = pow2_at_least(cnt);
    __syncthreads();
    for (uint32_t i = cnt + lane; i < n; i += 64) list[i] = kKeyMax;
    __syncthreads();
    wave_sort_lds_n(list, n);
    const uint32_t keep = cnt < K ? cnt : K;
    tau = (keep == K) ? uniform_u64(list[K - 1]) : kKeyMax;
    cnt = keep;
    __syncthreads();
}

// Offer one key per lane to a global list; compacts when the next tile could
// overflow the list.  cnt and tau are wave-uniform.
template <int B>
__device__ inline void offer_global(uint64_t *list, uint32_t &cnt, uint64_t &tau, uint32_t K,
                                    uint64_t *sbuf, bool pass, uint64_t key) {
    const uint64_t m = wave_ballot(pass);
    if (m == 0) return;
    if (pass) list[cnt + lanes_below(m)] = key;
    cnt += (uint32_t)__popcll(m);
    if (cnt > (uint32_t)(B - 64)) compact_global<B>(list, cnt, tau, K, sbuf);
}

template <int B>
__device__ inline void offer_lds(uint64_t *list, uint32_t &cnt, uint64_t &tau, uint32_t K, bool pass,
                                 uint64_t key) {
    const uint64_t m = wave_ballot(pass);
    if (m == 0) return;
    if (pass) list[cnt + lanes_below(m)] = key;
    cnt += (uint32_t)__popcll(m);
    if (cnt > (uint32_t)(B - 64)) compact_lds<B>(list, cnt, tau, K);
}

// Publish this item's thresholds to the per-query shared thresholds and take
// the tighter of the two (one 64-bit atomicMin per query both publishes and
// fetches).  Valid because a key at or above the K-th smallest key of ANY
// subset of a query's candidates cannot be in that query's top-K.  s_tau[j]
// is the item's threshold of its query j, iq_qi(j) the query's batch index.
template <class Sync = BlockSync, class QiFn>
__device__ inline void share_thresholds(uint64_t *gtau, uint32_t nq, uint64_t *s_tau, QiFn iq_qi) {
    if (gtau == nullptr) return;
    Sync::sync();
    for (uint32_t j = threadIdx.x & 63; j < nq; j += 64) {
        const uint64_t mine = s_tau[j];
        const uint64_t old = atomicMin((unsigned long long *)&gtau[iq_qi(j)], (unsigned long long)mine);
        s_tau[j] = old < mine ? old : mine;
    }
    Sync::sync();
}

}  // namespace fbf
