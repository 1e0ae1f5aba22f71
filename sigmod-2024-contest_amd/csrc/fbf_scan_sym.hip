// fbf_scan_sym.hip -- symmetric int8 scan (the router's brute-force kernel).
//
// Replaces bruteforce() / bruteforce_subgraph() (bruteforce.h:244-289) and
// their per-row L2SqrSQ8_sym (pyglass/glass/simd/distance.hpp:329-366): exact
// int32 sum over the 112-byte code of (x_i - q_i)^2, evaluated as
//     |x|^2 - 2 x.q + |q|^2
// with v_dot4_i32_i8 (four int8 products per instruction, int32 accumulate;
// every term is exact, so the result equals the reference bit for bit).
// One wave per work item; a 64-row tile (7 KiB) is loaded with coalesced
// 16-byte loads into LDS, then each lane owns one row.
#include "fbf_select.hpp"

namespace fbf {

constexpr int kCodeInt4 = 7;  // 112 bytes = 7 x 16

__device__ inline int dot_codes(const int4 (&x)[kCodeInt4], const int4 *q) {
    int acc = 0;
#pragma unroll
    for (int c = 0; c < kCodeInt4; ++c) {
        const int4 v = q[c];
        acc = __builtin_amdgcn_sdot4(x[c].x, v.x, acc, false);
        acc = __builtin_amdgcn_sdot4(x[c].y, v.y, acc, false);
        acc = __builtin_amdgcn_sdot4(x[c].z, v.z, acc, false);
        acc = __builtin_amdgcn_sdot4(x[c].w, v.w, acc, false);
    }
    return acc;
}

template <int B>
__device__ __attribute__((always_inline)) void scan_sym_item(const ScanSymArgs &a, const WorkItem &W, uint64_t *sbuf,
                              uint64_t *s_tau, uint32_t *s_cnt, int *s_qq, int4 *sq, int4 *stile) {
    const uint32_t lane = threadIdx.x;
    const uint32_t nq = W.q_count;
    cptr<ItemQuery> iq = as_const(a.iq) + W.q_begin;
    for (uint32_t j = 0; j < nq; ++j) {
        s_tau[j] = kKeyMax;
        s_cnt[j] = 0;
    }
    __syncthreads();
    for (uint32_t e = lane; e < nq * kCodeInt4; e += 64) {
        const uint32_t j = e / kCodeInt4, c = e % kCodeInt4;
        sq[e] = reinterpret_cast<const int4 *>(a.qcodes + (uint64_t)iq[j].qi * 112)[c];
    }
    __syncthreads();
    if (lane < nq) {
        int4 qx[kCodeInt4];
#pragma unroll
        for (int c = 0; c < kCodeInt4; ++c) qx[c] = sq[lane * kCodeInt4 + c];
        s_qq[lane] = dot_codes(qx, sq + lane * kCodeInt4);
    }
    __syncthreads();

    const int4 *src = reinterpret_cast<const int4 *>(a.codes);
#pragma unroll 1
    for (uint32_t t = W.row_begin; t < W.row_end; t += kRowTile) {
        // Coalesced tile load: int4 number (t*7 + lane + 64c) of the code array.
#pragma unroll
        for (int c = 0; c < kCodeInt4; ++c) {
            const uint32_t e = lane + 64 * c;
            const uint64_t row = t + e / kCodeInt4;
            int4 v = make_int4(0, 0, 0, 0);
            if (row < a.n_rows) v = src[(uint64_t)t * kCodeInt4 + e];
            stile[e] = v;
        }
        __syncthreads();
        int4 x[kCodeInt4];
#pragma unroll
        for (int c = 0; c < kCodeInt4; ++c) x[c] = stile[lane * kCodeInt4 + c];
        __syncthreads();
        const uint32_t row = t + lane;
        int xxs = 0;
#pragma unroll
        for (int c = 0; c < kCodeInt4; ++c) {
            xxs = __builtin_amdgcn_sdot4(x[c].x, x[c].x, xxs, false);
            xxs = __builtin_amdgcn_sdot4(x[c].y, x[c].y, xxs, false);
            xxs = __builtin_amdgcn_sdot4(x[c].z, x[c].z, xxs, false);
            xxs = __builtin_amdgcn_sdot4(x[c].w, x[c].w, xxs, false);
        }
        const bool row_ok = row < W.row_end;
        const uint32_t id = (a.ids && row_ok) ? a.ids[row] : row;
#pragma unroll 1
        for (uint32_t j = 0; j < nq; ++j) {
            const ItemQuery e = iq[j];
            if (t >= e.e || t + kRowTile <= e.s) continue;
            const int xq = dot_codes(x, sq + j * kCodeInt4);
            const int dist = xxs - 2 * xq + s_qq[j];
            const bool in = row_ok && row >= e.s && row < e.e;
            const uint64_t key = ((uint64_t)(uint32_t)dist << 32) | id;
            uint64_t tau = s_tau[j];
            uint32_t cnt = s_cnt[j];
            uint64_t *list = a.lists + list_slot(e, W) * (uint64_t)B;
            offer_global<B>(list, cnt, tau, a.K, sbuf, in && key < tau, key);
            s_tau[j] = tau;
            s_cnt[j] = cnt;
        }
    }
    for (uint32_t j = 0; j < nq; ++j) {
        uint64_t tau = s_tau[j];
        uint32_t cnt = s_cnt[j];
        const uint64_t slot = list_slot(iq[j], W);
        uint64_t *list = a.lists + slot * (uint64_t)B;
        compact_global<B>(list, cnt, tau, a.K, sbuf);
        if (lane == 0) a.counts[slot] = cnt;
    }
}

template <int B>
__global__ __launch_bounds__(64) void scan_sym_kernel(ScanSymArgs a) {
    __shared__ uint64_t sbuf[B];
    __shared__ uint64_t s_tau[kMaxGroup];
    __shared__ uint32_t s_cnt[kMaxGroup];
    __shared__ int s_qq[kMaxGroup];
    __shared__ int4 sq[kMaxGroup * kCodeInt4];
    __shared__ int4 stile[kRowTile * kCodeInt4];
    if (a.queue == nullptr) {
        const WorkItem W = as_const(a.items)[blockIdx.x];
        scan_sym_item<B>(a, W, sbuf, s_tau, s_cnt, s_qq, sq, stile);
        return;
    }
    for (;;) {
        uint32_t idx = 0;
        if (threadIdx.x == 0) idx = atomicAdd(a.queue, 1u);
        idx = __builtin_amdgcn_readfirstlane(__shfl(idx, 0));
        if (idx >= a.n_items) break;
        const WorkItem W = as_const(a.items)[idx];
        scan_sym_item<B>(a, W, sbuf, s_tau, s_cnt, s_qq, sq, stile);
    }
}

hipError_t launch_scan_sq8sym(const ScanSymArgs &a, int B, hipStream_t st) {
    if (a.n_items == 0) return hipSuccess;
    // A list holds K kept keys plus one tile of appends (offer_global).
    if ((int)a.K + kRowTile > B) return hipErrorInvalidValue;
    uint64_t grid = a.n_items;
    if (a.queue) {
        int dev = 0, cus = 0;
        (void)hipGetDevice(&dev);
        (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        grid = std::min<uint64_t>((uint64_t)cus * 8, a.n_items);
    }
    switch (B) {
    case 256: hipLaunchKernelGGL(scan_sym_kernel<256>, dim3((uint32_t)grid), dim3(64), 0, st, a); break;
    case 512: hipLaunchKernelGGL(scan_sym_kernel<512>, dim3((uint32_t)grid), dim3(64), 0, st, a); break;
    case 1024: hipLaunchKernelGGL(scan_sym_kernel<1024>, dim3((uint32_t)grid), dim3(64), 0, st, a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// ---------------------------------------------------------------- one query, one launch

template <int B>
__global__ __launch_bounds__(64) void sym_one_kernel(SymOneArgs a) {
    __shared__ uint64_t sbuf[B];
    __shared__ int4 stile[kRowTile * kCodeInt4];
    __shared__ uint32_t s_last;
    const uint32_t lane = threadIdx.x;
    int4 q[kCodeInt4];
#pragma unroll
    for (int c = 0; c < kCodeInt4; ++c) q[c] = a.q[c];
    const int qq = dot_codes(q, q);
    const uint32_t r0 = blockIdx.x * a.chunk;
    const uint32_t r1 = min(r0 + a.chunk, a.n);
    uint64_t *list = a.lists + (uint64_t)blockIdx.x * B;
    uint32_t cnt = 0;
    uint64_t tau = kKeyMax;
    const int4 *src = reinterpret_cast<const int4 *>(a.codes);
#pragma unroll 1
    for (uint32_t t = r0; t < r1; t += kRowTile) {
        // Coalesced tile load: int4 number (t*7 + lane + 64c) of the slice.
#pragma unroll
        for (int c = 0; c < kCodeInt4; ++c) {
            const uint32_t e = lane + 64 * c;
            const uint32_t row = t + e / kCodeInt4;
            stile[e] = row < r1 ? src[(uint64_t)t * kCodeInt4 + e] : make_int4(0, 0, 0, 0);
        }
        __syncthreads();
        int4 x[kCodeInt4];
#pragma unroll
        for (int c = 0; c < kCodeInt4; ++c) x[c] = stile[lane * kCodeInt4 + c];
        __syncthreads();
        const uint32_t row = t + lane;
        const int dist = dot_codes(x, x) - 2 * dot_codes(x, q) + qq;
        const uint64_t key = ((uint64_t)(uint32_t)dist << 32) | row;
        offer_global<B>(list, cnt, tau, a.K, sbuf, row < r1 && key < tau, key);
    }
    if (lane == 0) a.counts[blockIdx.x] = cnt;
    // Publish this list, then count arrivals; the last workgroup merges.
    __threadfence();
    if (lane == 0) {
        const uint32_t old = __hip_atomic_fetch_add(a.done, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        s_last = old == gridDim.x - 1 ? 1u : 0u;
    }
    __syncthreads();
    if (s_last == 0) return;
    __threadfence();
    uint32_t mc = 0;
    uint64_t mt = kKeyMax;
#pragma unroll 1
    for (uint32_t g = 0; g < gridDim.x; ++g) {
        const uint32_t c = __hip_atomic_load(&a.counts[g], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const uint64_t *L = a.lists + (uint64_t)g * B;
#pragma unroll 1
        for (uint32_t i = 0; i < c; i += 64) {
            const uint32_t ii = i + lane;
            const uint64_t key = ii < c ? __hip_atomic_load(&L[ii], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                                        : kKeyMax;
            offer_lds<B>(sbuf, mc, mt, a.K, key < mt, key);
        }
    }
    compact_lds<B>(sbuf, mc, mt, a.K);  // ascending, mc <= K
    for (uint32_t i = lane; i < a.K; i += 64) {
        const bool v = i < mc;
        a.out[i] = v ? (uint32_t)sbuf[i] : kPadId;
        a.out[a.K + i] = v ? (uint32_t)(sbuf[i] >> 32) : 0x7FFFFFFFu;
    }
    if (lane == 0) __hip_atomic_store(a.done, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ---------------------------------------------------------------- many one-slice searches, one launch

template <int B>
__global__ __launch_bounds__(64) void sym_multi_kernel(SymMultiArgs a) {
    __shared__ uint64_t sbuf[B];
    __shared__ int4 stile[kRowTile * kCodeInt4];
    __shared__ uint32_t s_last;
    const uint32_t lane = threadIdx.x;
    const uint32_t rq = blockIdx.y;
    cptr<SymReq> R = as_const(a.reqs) + rq;
    const uint32_t nblk = R->nblk;
    if (blockIdx.x >= nblk) return;  // the request needs fewer workgroups
    const uint32_t n = R->n, chunk = R->chunk, list0 = R->list0;
    int4 q[kCodeInt4];
#pragma unroll
    for (int c = 0; c < kCodeInt4; ++c) q[c] = R->q[c];
    const int qq = dot_codes(q, q);
    const uint32_t r0 = blockIdx.x * chunk;
    const uint32_t r1 = min(r0 + chunk, n);
    uint64_t *list = a.lists + (uint64_t)(list0 + blockIdx.x) * B;
    uint32_t cnt = 0;
    uint64_t tau = kKeyMax;
    const int4 *src = reinterpret_cast<const int4 *>(R->codes);
    uint64_t *gt = a.gtau + rq;
    uint32_t step = 0;
    // The next tile is loaded into registers while this one is scored.
    int4 nxt[kCodeInt4];
    auto load_tile = [&](uint32_t t) __attribute__((always_inline)) {
#pragma unroll
        for (int c = 0; c < kCodeInt4; ++c) {
            const uint32_t e = lane + 64 * c;
            const uint32_t row = t + e / kCodeInt4;
            nxt[c] = row < r1 ? src[(uint64_t)t * kCodeInt4 + e] : make_int4(0, 0, 0, 0);
        }
    };
    load_tile(r0);
#pragma unroll 1
    for (uint32_t t = r0; t < r1; t += kRowTile) {
#pragma unroll
        for (int c = 0; c < kCodeInt4; ++c) stile[lane + 64 * c] = nxt[c];
        if (t + kRowTile < r1) load_tile(t + kRowTile);
        __syncthreads();
        int4 x[kCodeInt4];
#pragma unroll
        for (int c = 0; c < kCodeInt4; ++c) x[c] = stile[lane * kCodeInt4 + c];
        __syncthreads();
        const uint32_t row = t + lane;
        const int dist = dot_codes(x, x) - 2 * dot_codes(x, q) + qq;
        const uint64_t key = ((uint64_t)(uint32_t)dist << 32) | row;
        offer_global<B>(list, cnt, tau, a.K, sbuf, row < r1 && key < tau, key);
        // Every 4 tiles: publish this list's threshold and take the tighter
        // of it and the request's (the K-th key of any of the request's lists
        // bounds its K-th key).
        if ((++step & 3u) == 0) {
            uint64_t g = 0;
            if (lane == 0) g = atomicMin((unsigned long long *)gt, (unsigned long long)tau);
            g = uniform_u64(__shfl(g, 0));
            tau = g < tau ? g : tau;
        }
    }
    if (lane == 0) {
        (void)atomicMin((unsigned long long *)gt, (unsigned long long)tau);
        a.counts[list0 + blockIdx.x] = cnt;
    }
    // Publish this list, then count the request's arrivals; its last
    // workgroup merges.
    __threadfence();
    if (lane == 0) {
        const uint32_t old = __hip_atomic_fetch_add(&a.done[rq], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        s_last = old == nblk - 1 ? 1u : 0u;
    }
    __syncthreads();
    if (s_last == 0) return;
    __threadfence();
    uint32_t mc = 0;
    uint64_t mt = kKeyMax;
    // Every list's count in one load (lane g: list g; nblk <= 64), then each
    // list's keys with all of its loads in flight at once, issued while the
    // previous list is offered: about one memory round trip per list rather
    // than one per 64 keys (C1-like router slices from 16 threads: 26.1 k ->
    // 29.1 k calls/s).
    const uint32_t my_cnt = lane < nblk
        ? __hip_atomic_load(&a.counts[list0 + lane], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0u;
    constexpr int kChunks = (B - 64) / 64;  // a list holds at most B - 64 keys
    uint64_t nx[kChunks];
    auto load_list = [&](uint32_t g) __attribute__((always_inline)) {
        const uint32_t c = __builtin_amdgcn_readlane(my_cnt, g);
        const uint64_t *L = a.lists + (uint64_t)(list0 + g) * B;
#pragma unroll
        for (int j = 0; j < kChunks; ++j) {
            const uint32_t ii = 64 * j + lane;
            nx[j] = ii < c ? __hip_atomic_load(&L[ii], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : kKeyMax;
        }
    };
    load_list(0);
#pragma unroll 1
    for (uint32_t g = 0; g < nblk; ++g) {
        const uint32_t c = __builtin_amdgcn_readlane(my_cnt, g);
        uint64_t kv[kChunks];
#pragma unroll
        for (int j = 0; j < kChunks; ++j) kv[j] = nx[j];
        if (g + 1 < nblk) load_list(g + 1);  // in flight while list g is offered
#pragma unroll
        for (int j = 0; j < kChunks; ++j)
            if ((uint32_t)(64 * j) < c) offer_lds<B>(sbuf, mc, mt, a.K, kv[j] < mt, kv[j]);
    }
    compact_lds<B>(sbuf, mc, mt, a.K);  // ascending, mc <= K
    uint32_t *out = a.out + (uint64_t)rq * 2 * a.K;
    for (uint32_t i = lane; i < a.K; i += 64) {
        const bool v = i < mc;
        out[i] = v ? (uint32_t)sbuf[i] : kPadId;
        out[a.K + i] = v ? (uint32_t)(sbuf[i] >> 32) : 0x7FFFFFFFu;
    }
    if (lane == 0) {
        __hip_atomic_store(gt, kKeyMax, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(&a.done[rq], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

hipError_t launch_sym_multi(const SymMultiArgs &a, uint32_t grid_x, int B, hipStream_t st) {
    if (a.nreq == 0) return hipSuccess;
    if (grid_x == 0 || a.nreq > 65535) return hipErrorInvalidValue;
    if ((int)a.K + kRowTile > B) return hipErrorInvalidValue;
    const dim3 grid(grid_x, a.nreq);
    switch (B) {
    case 256: hipLaunchKernelGGL(sym_multi_kernel<256>, grid, dim3(64), 0, st, a); break;
    case 512: hipLaunchKernelGGL(sym_multi_kernel<512>, grid, dim3(64), 0, st, a); break;
    case 1024: hipLaunchKernelGGL(sym_multi_kernel<1024>, grid, dim3(64), 0, st, a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_sym_one(const SymOneArgs &a, int B, hipStream_t st) {
    if (a.n == 0 || a.chunk == 0 || a.chunk % kRowTile != 0) return hipErrorInvalidValue;
    if ((int)a.K + kRowTile > B) return hipErrorInvalidValue;
    const uint32_t grid = (a.n + a.chunk - 1) / a.chunk;
    switch (B) {
    case 256: hipLaunchKernelGGL(sym_one_kernel<256>, dim3(grid), dim3(64), 0, st, a); break;
    case 512: hipLaunchKernelGGL(sym_one_kernel<512>, dim3(grid), dim3(64), 0, st, a); break;
    case 1024: hipLaunchKernelGGL(sym_one_kernel<1024>, dim3(grid), dim3(64), 0, st, a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace fbf

namespace fbf {

// ---------------------------------------------------------------- symmetric int8 encoding (device)

// Global absmax (SQ8SymmetricQuantizer::train, sq8_sym_quant.hpp:37-43):
// non-negative floats order like their bit patterns, so an unsigned atomicMax
// on the bits reduces them exactly.
__global__ void absmax_kernel(const float *__restrict__ x, uint64_t count, uint32_t *__restrict__ out_bits) {
    uint32_t m = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
         i += (uint64_t)gridDim.x * blockDim.x)
        m = max(m, __builtin_bit_cast(uint32_t, fabsf(x[i])));
    for (int off = 32; off > 0; off >>= 1) m = max(m, (uint32_t)__shfl_xor((int)m, off));
    if ((threadIdx.x & 63) == 0) atomicMax(out_bits, m);
}

// SQ8SymmetricQuantizer::encode (sq8_sym_quant.hpp:63-74): round(clamp(x /
// alpha, -1, 1) * 127), half away from zero, 112-byte codes zero padded.
// Row p of the output takes source row perm[p] (perm may be null: identity);
// rows at or beyond n are all zero.
__global__ void sym_encode_kernel(const float *__restrict__ rows, const uint32_t *__restrict__ perm,
                                  uint64_t n, uint64_t n_out, float alpha, int4 *__restrict__ out) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_out * kCodeInt4) return;
    const uint64_t p = t / kCodeInt4;
    const int c = (int)(t % kCodeInt4);
    uint32_t w[4] = {0, 0, 0, 0};
    if (p < n) {
        const float *x = rows + (uint64_t)(perm ? perm[p] : p) * kDim;
        for (int e = 0; e < 16; ++e) {
            const int k = 16 * c + e;
            int8_t code = 0;
            if (k < kDim) {
                float v = x[k] / alpha;
                if (v > 1.0f) v = 1.0f;
                if (v < -1.0f) v = -1.0f;
                code = (int8_t)roundf(v * 127.0f);
            }
            w[e / 4] |= (uint32_t)(uint8_t)code << (8 * (e % 4));
        }
    }
    out[t] = make_int4((int)w[0], (int)w[1], (int)w[2], (int)w[3]);
}

hipError_t launch_absmax(const float *x, uint64_t count, uint32_t *out_bits, hipStream_t st) {
    hipLaunchKernelGGL(absmax_kernel, dim3(1024), dim3(256), 0, st, x, count, out_bits);
    return hipGetLastError();
}

hipError_t launch_sym_encode(const float *rows, const uint32_t *perm, uint64_t n, uint64_t n_out, float alpha,
                             int8_t *out, hipStream_t st) {
    const uint64_t total = n_out * kCodeInt4;
    if (total == 0) return hipSuccess;
    hipLaunchKernelGGL(sym_encode_kernel, dim3((uint32_t)((total + 255) / 256)), dim3(256), 0, st, rows, perm, n,
                       n_out, alpha, reinterpret_cast<int4 *>(out));
    return hipGetLastError();
}

}  // namespace fbf

namespace fbf {

// ---------------------------------------------------------------- type-2 sub-interval path

// Absmax over the rows perm[0..n) of rows (one time bucket's quantiser,
// SQ8SymmetricQuantizer::train over that bucket's rows, sq8_sym_quant.hpp:32-44).
__global__ void absmax_rows_kernel(const float *__restrict__ rows, const uint32_t *__restrict__ perm, uint64_t n,
                                   uint32_t *__restrict__ out_bits) {
    uint32_t m = 0;
    const uint64_t total = n * kDim;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (uint64_t)gridDim.x * blockDim.x)
        m = max(m, __builtin_bit_cast(uint32_t, fabsf(rows[(uint64_t)perm[i / kDim] * kDim + i % kDim])));
    for (int off = 32; off > 0; off >>= 1) m = max(m, (uint32_t)__shfl_xor((int)m, off));
    if ((threadIdx.x & 63) == 0) atomicMax(out_bits, m);
}

hipError_t launch_absmax_rows(const float *rows, const uint32_t *perm, uint64_t n, uint32_t *out_bits, hipStream_t st) {
    if (n == 0) return hipSuccess;
    const uint64_t blocks = std::min<uint64_t>(1024, (n * kDim + 255) / 256);
    hipLaunchKernelGGL(absmax_rows_kernel, dim3((uint32_t)blocks), dim3(256), 0, st, rows, perm, n, out_bits);
    return hipGetLastError();
}

// Query codes of each slice: query sl[i].q encoded with its bucket's alpha
// (the bucket quantiser's encode, hybrid_graph.cpp:402-405), 112 bytes.
__global__ void sym_encode_slices_kernel(const float *__restrict__ queries, const SliceInfo *__restrict__ sl,
                                         uint64_t n, int4 *__restrict__ out) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * kCodeInt4) return;
    const uint64_t i = t / kCodeInt4;
    const int c = (int)(t % kCodeInt4);
    const SliceInfo S = sl[i];
    const float *x = queries + (uint64_t)S.q * kDim;
    uint32_t w[4] = {0, 0, 0, 0};
    for (int e = 0; e < 16; ++e) {
        const int k = 16 * c + e;
        int8_t code = 0;
        if (k < kDim) {
            float v = x[k] / S.alpha;
            if (v > 1.0f) v = 1.0f;
            if (v < -1.0f) v = -1.0f;
            code = (int8_t)roundf(v * 127.0f);
        }
        w[e / 4] |= (uint32_t)(uint8_t)code << (8 * (e % 4));
    }
    out[t] = make_int4((int)w[0], (int)w[1], (int)w[2], (int)w[3]);
}

hipError_t launch_sym_encode_slices(const float *queries, const SliceInfo *sl, uint64_t n, int8_t *out, hipStream_t st) {
    const uint64_t total = n * kCodeInt4;
    if (total == 0) return hipSuccess;
    hipLaunchKernelGGL(sym_encode_slices_kernel, dim3((uint32_t)((total + 255) / 256)), dim3(256), 0, st, queries, sl,
                       n, reinterpret_cast<int4 *>(out));
    return hipGetLastError();
}

// A pool distance as the reference forms it: the int distance times alpha
// times alpha in float (bruteforce.h:287), converted to int by the pool's
// maybe_pop_emplace(int, int) (truncation toward zero, hybrid_graph.cpp:415);
// values at or above 2^31 saturate.
__device__ inline uint32_t pool_dist_of(float f) {
    return f >= 2147483648.0f ? 0x7FFFFFFFu : (uint32_t)(int32_t)f;
}

// Pool keys of the SMALL slices: slice i's kb merged (int distance, code row)
// results become (pool distance, 1 << 31 | timestamp-layout position) keys in
// its pool slot; pads stay kKeyMax (never offered).  Slot counts are kb.
__global__ void pool_keys_kernel(const uint32_t *__restrict__ ids, const uint32_t *__restrict__ dbits,
                                 const SliceInfo *__restrict__ sl, uint64_t n, uint32_t kb, uint32_t stride,
                                 uint64_t *__restrict__ lists, uint32_t *__restrict__ counts) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * kb) return;
    const uint64_t i = t / kb;
    const uint32_t j = (uint32_t)(t % kb);
    const SliceInfo S = sl[i];
    const uint32_t id = ids[t];
    uint64_t key = kKeyMax;
    if (id != kPadId) {
        const float f = ((float)(int32_t)dbits[t] * S.alpha) * S.alpha;
        const uint32_t pos = (uint32_t)((int64_t)id + S.row_to_pos);
        key = ((uint64_t)pool_dist_of(f) << 32) | 0x80000000u | pos;
    }
    lists[(uint64_t)S.pool_slot * stride + j] = key;
    if (j == 0) counts[S.pool_slot] = kb;
}

hipError_t launch_pool_keys(const uint32_t *ids, const uint32_t *dbits, const SliceInfo *sl, uint64_t n, uint32_t kb,
                            uint32_t stride, uint64_t *lists, uint32_t *counts, hipStream_t st) {
    const uint64_t total = n * kb;
    if (total == 0) return hipSuccess;
    hipLaunchKernelGGL(pool_keys_kernel, dim3((uint32_t)((total + 255) / 256)), dim3(256), 0, st, ids, dbits, sl, n,
                       kb, stride, lists, counts);
    return hipGetLastError();
}

// Pool keys of the caller's candidates (the graph searches of FULL / MEDIUM
// buckets, hybrid_graph.cpp:421-458): query q's n_ext (position, distance)
// pairs into its ext slot, distances converted as the pool converts them;
// pad ids (0xFFFFFFFF) and positions at or beyond n_pos are not offered.
__global__ void ext_keys_kernel(const uint32_t *__restrict__ pos, const float *__restrict__ dist,
                                const uint32_t *__restrict__ ext_slot, uint64_t nq, uint32_t n_ext, uint64_t n_pos,
                                uint32_t stride, uint64_t *__restrict__ lists, uint32_t *__restrict__ counts) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nq * n_ext) return;
    const uint64_t q = t / n_ext;
    const uint32_t j = (uint32_t)(t % n_ext);
    const uint32_t p = pos[t];
    const float f = dist[t];
    uint64_t key = kKeyMax;
    if (p != kPadId && (uint64_t)p < n_pos && f >= 0.0f) key = ((uint64_t)pool_dist_of(f) << 32) | 0x80000000u | p;
    const uint32_t slot = ext_slot[q];
    lists[(uint64_t)slot * stride + j] = key;
    if (j == 0) counts[slot] = n_ext;
}

hipError_t launch_ext_keys(const uint32_t *pos, const float *dist, const uint32_t *ext_slot, uint64_t nq, uint32_t n_ext,
                           uint64_t n_pos, uint32_t stride, uint64_t *lists, uint32_t *counts, hipStream_t st) {
    const uint64_t total = nq * n_ext;
    if (total == 0) return hipSuccess;
    hipLaunchKernelGGL(ext_keys_kernel, dim3((uint32_t)((total + 255) / 256)), dim3(256), 0, st, pos, dist, ext_slot,
                       nq, n_ext, n_pos, stride, lists, counts);
    return hipGetLastError();
}

}  // namespace fbf
