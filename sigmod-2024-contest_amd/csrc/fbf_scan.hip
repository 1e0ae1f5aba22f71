// fbf_scan.hip -- exact FP32 filtered scan, merge / rerank, layout builders.
//
// Exact scan (replaces the scalar loop of Bruteforce, utils.h:105-139, and the
// per-row normal_l2 / L2SqrRef calls, distance.h:62-71, distance.hpp:165-173):
// one wave64 per work item, one candidate row per lane held in 100 VGPRs, the
// item's queries read as wave-uniform scalars, so every (query, row) pair costs
// one v_sub_f32 and one v_fma_f32 per dimension in the canonical order
//     t = x_i - q_i;  acc = fma(t, t, acc),   i = 0 .. 99.
// Rows stream from HBM once per item and are reused across the item's queries.
#include "fbf_select.hpp"

#pragma clang fp contract(off)

namespace fbf {

// ---------------------------------------------------------------- layout build

// FP32 scan copy of one sorted layout: position p lives in block p / 64 as
// 25 float4 groups, group g of lane p % 64 at float4 index (blk*25 + g)*64 + lane,
// so a wave's load of group g is one contiguous 1 KiB burst.  Rows at or
// beyond n (tile padding) are zero.
__global__ void build_soa_kernel(const float *__restrict__ rows, const uint32_t *__restrict__ perm,
                                 uint64_t n, uint64_t n_pad, float4 *__restrict__ soa) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t total = n_pad * (kDim / 4);
    if (t >= total) return;
    const uint64_t p = t % n_pad;        // consecutive threads take consecutive rows
    const uint32_t g = (uint32_t)(t / n_pad);
    float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
    if (p < n) {
        const float *src = rows + (uint64_t)perm[p] * kDim + 4 * g;
        v = make_float4(src[0], src[1], src[2], src[3]);
    }
    const uint64_t blk = p / kRowTile, lane = p % kRowTile;
    soa[(blk * (kDim / 4) + g) * kRowTile + lane] = v;
}

hipError_t launch_build_soa(const float *rows, const uint32_t *perm, uint64_t n, float *soa,
                            hipStream_t st) {
    const uint64_t n_pad = (n + kRowTile - 1) / kRowTile * kRowTile;
    const uint64_t total = n_pad * (kDim / 4);
    const uint64_t blocks = (total + 255) / 256;
    hipLaunchKernelGGL(build_soa_kernel, dim3((uint32_t)blocks), dim3(256), 0, st, rows, perm, n,
                       n_pad, reinterpret_cast<float4 *>(soa));
    return hipGetLastError();
}

// Fill kernel for counters (used instead of a memset so the search can be
// captured in a graph).
__global__ void fill_u32_kernel(uint32_t *p, uint32_t v, uint64_t n) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

hipError_t launch_fill_u32(uint32_t *p, uint32_t v, uint64_t n, hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(fill_u32_kernel, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, st, p, v, n);
    return hipGetLastError();
}

// ---------------------------------------------------------------- exact scan


__device__ inline void load_row(const float4 *__restrict__ soa, uint32_t tile_row0, uint32_t lane,
                                float (&x)[kDim]) {
    const float4 *src = soa + (uint64_t)(tile_row0 / kRowTile) * (kDim / 4) * kRowTile + lane;
#pragma unroll
    for (int g = 0; g < kDim / 4; ++g) {
        const float4 v = src[g * kRowTile];
        x[4 * g + 0] = v.x;
        x[4 * g + 1] = v.y;
        x[4 * g + 2] = v.z;
        x[4 * g + 3] = v.w;
    }
}

template <int B>
__device__ __attribute__((always_inline)) void scan_f32_item(const ScanArgs &a, const WorkItem &W, uint64_t *sbuf,
                              uint64_t *s_tau, uint32_t *s_cnt, float4 *sq) {
    const uint32_t lane = threadIdx.x;
    const uint32_t nq = W.q_count;
    const float4 *soa = reinterpret_cast<const float4 *>(a.soa);
    cptr<ItemQuery> iq = as_const(a.iq) + W.q_begin;
    // Start from the thresholds earlier items of the same queries published.
    for (uint32_t j = 0; j < nq; ++j) {
        // every lane writes the same value and reads only its own writes
        s_tau[j] = a.gtau ? uniform_u64(__hip_atomic_load(&a.gtau[iq[j].qi], __ATOMIC_RELAXED,
                                                          __HIP_MEMORY_SCOPE_AGENT))
                          : kKeyMax;
        s_cnt[j] = 0;
    }
    // The item's query vectors, staged once in LDS; the scan reads them at
    // wave-uniform addresses (LDS broadcast).
    __syncthreads();
    for (uint32_t e = lane; e < nq * (kDim / 4); e += 64) {
        const uint32_t j = e / (kDim / 4), g = e % (kDim / 4);
        const float4 *src = reinterpret_cast<const float4 *>(a.queries + (uint64_t)iq[j].qi * kDim);
        sq[j * (kDim / 4) + g] = src[g];
    }
    __syncthreads();

    uint32_t since_share = 0;
#pragma unroll 1
    for (uint32_t t = W.row_begin; t < W.row_end; t += kRowTile) {
        if (++since_share == 32) {
            since_share = 0;
            share_thresholds(a.gtau, nq, s_tau, [&](uint32_t j) { return iq[j].qi; });
        }
        const uint32_t row = t + lane;
        float x[kDim];
        load_row(soa, t, lane, x);
        const uint32_t id = a.ids[row];  // ids are padded to the tile multiple
        const bool row_ok = row < W.row_end;

#pragma unroll 1
        for (uint32_t j = 0; j < nq; ++j) {
            const ItemQuery e = iq[j];
            if (t >= e.e || t + kRowTile <= e.s) continue;  // tile outside this query's range
            // One query at a time: the compiler keeps a single FMA chain and
            // consumes each LDS broadcast read right away.
            const float4 *qv = sq + j * (kDim / 4);
            float acc = 0.f;
#pragma unroll
            for (int g = 0; g < kDim / 4; ++g) {
                const float4 v = qv[g];
                float tt;
                tt = x[4 * g + 0] - v.x; acc = __builtin_fmaf(tt, tt, acc);
                tt = x[4 * g + 1] - v.y; acc = __builtin_fmaf(tt, tt, acc);
                tt = x[4 * g + 2] - v.z; acc = __builtin_fmaf(tt, tt, acc);
                tt = x[4 * g + 3] - v.w; acc = __builtin_fmaf(tt, tt, acc);
            }
            const bool in = row_ok && row >= e.s && row < e.e;
            const uint64_t key = ((uint64_t)f2ord(acc) << 32) | id;
            uint64_t tau = s_tau[j];
            uint32_t cnt = s_cnt[j];
            uint64_t *list = a.lists + list_slot(e, W) * (uint64_t)B;
            offer_global<B>(list, cnt, tau, a.K, sbuf, in && key < tau, key);
            s_tau[j] = tau;
            s_cnt[j] = cnt;
        }
    }
    // Final cut: each list sorted and at most K long.
    for (uint32_t j = 0; j < nq; ++j) {
        uint64_t tau = s_tau[j];
        uint32_t cnt = s_cnt[j];
        const uint64_t slot = list_slot(iq[j], W);
        uint64_t *list = a.lists + slot * (uint64_t)B;
        compact_global<B>(list, cnt, tau, a.K, sbuf);
        if (lane == 0) a.counts[slot] = cnt;
        s_tau[j] = tau;
    }
    share_thresholds(a.gtau, nq, s_tau, [&](uint32_t j) { return iq[j].qi; });
}

template <int B>
__global__ __launch_bounds__(64) void scan_f32_kernel(ScanArgs a) {
    __shared__ uint64_t sbuf[B];
    __shared__ uint64_t s_tau[kMaxGroup];
    __shared__ uint32_t s_cnt[kMaxGroup];
    __shared__ float4 sq[kMaxGroup * (kDim / 4)];
    if (a.queue == nullptr) {
        const WorkItem W = as_const(a.items)[blockIdx.x];
        scan_f32_item<B>(a, W, sbuf, s_tau, s_cnt, sq);
        return;
    }
    // Persistent form: each wave pulls item indices until the queue is drained.
    for (;;) {
        uint32_t idx = 0;
        if (threadIdx.x == 0) idx = atomicAdd(a.queue, 1u);
        idx = __builtin_amdgcn_readfirstlane(__shfl(idx, 0));
        if (idx >= a.n_items) break;
        const WorkItem W = as_const(a.items)[idx];
        scan_f32_item<B>(a, W, sbuf, s_tau, s_cnt, sq);
    }
}

hipError_t launch_scan_f32(const ScanArgs &a, int B, hipStream_t st) {
    if (a.n_items == 0) return hipSuccess;
    // A list holds K kept keys plus one tile of appends (offer_global).
    if ((int)a.K + kRowTile > B) return hipErrorInvalidValue;
    uint64_t grid = a.n_items;
    if (a.queue) {
        int dev = 0, cus = 0;
        hipGetDevice(&dev);
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        grid = (uint64_t)cus * 8;
        if (grid > a.n_items) grid = a.n_items;
    }
    switch (B) {
    case 256: hipLaunchKernelGGL(scan_f32_kernel<256>, dim3((uint32_t)grid), dim3(64), 0, st, a); break;
    case 512: hipLaunchKernelGGL(scan_f32_kernel<512>, dim3((uint32_t)grid), dim3(64), 0, st, a); break;
    case 1024: hipLaunchKernelGGL(scan_f32_kernel<1024>, dim3((uint32_t)grid), dim3(64), 0, st, a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// ---------------------------------------------------------------- merge / rerank

// Canonical FP32 distance of a row-major original-order row (rerank gather).
__device__ inline float canon_dist_row(const float *__restrict__ row, const float *__restrict__ q) {
    const float4 *r4 = reinterpret_cast<const float4 *>(row);
    float acc = 0.f;
#pragma unroll
    for (int g = 0; g < kDim / 4; ++g) {
        const float4 v = r4[g];
        float t;
        t = v.x - q[4 * g + 0]; acc = __builtin_fmaf(t, t, acc);
        t = v.y - q[4 * g + 1]; acc = __builtin_fmaf(t, t, acc);
        t = v.z - q[4 * g + 2]; acc = __builtin_fmaf(t, t, acc);
        t = v.w - q[4 * g + 3]; acc = __builtin_fmaf(t, t, acc);
    }
    return acc;
}

// One wave per query: stream the query's per-item lists through a running
// LDS list, keep the kprime smallest keys; exact modes emit them, the
// approximate modes recompute each candidate's canonical FP32 distance and
// emit the k smallest (distance, id).
template <int B>
__global__ __launch_bounds__(64) void merge_kernel(MergeArgs a) {
    __shared__ uint64_t list[B];
    const uint64_t q = blockIdx.x;
    const uint32_t lane = threadIdx.x;
    if (a.pass == 2 && a.flags[q] == 0) return;
    uint32_t cnt = 0;
    uint64_t tau = kKeyMax;
    const uint64_t s0 = a.slot_off[q], s1 = a.pass == 3 ? a.pre_end[q] : a.slot_off[q + 1];
    // 64 slots at a time: lane j reads the count word of slot sb + j, a wave
    // prefix sum numbers the keys of those slots, and each batch of 64 keys
    // is gathered with one load per lane (lane l finds its slot by a binary
    // search over the prefix sums), so no load waits for another.
#pragma unroll 1
    for (uint64_t sb = s0; sb < s1; sb += 64) {
        const uint64_t si = sb + lane;
        uint32_t c0 = 0, c = 0, h = 0;
        if (si < s1) {
            const uint32_t w = a.counts[si];
            c0 = list_count_c0(w);
            c = c0 + list_count_c1(w);
            h = list_count_h(w);
        }
        uint32_t incl = c;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t y = (uint32_t)__shfl_up((int)incl, o);
            if (lane >= (uint32_t)o) incl += y;
        }
        const uint32_t total = (uint32_t)__shfl((int)incl, 63);
#pragma unroll 1
        for (uint32_t g0 = 0; g0 < total; g0 += 64) {
            const uint32_t g = g0 + lane < total ? g0 + lane : total - 1;
            uint32_t lo = 0, hi = 63;  // the first slot whose inclusive count exceeds g
#pragma unroll
            for (int it = 0; it < 6; ++it) {
                const uint32_t mid = (lo + hi) >> 1;
                const uint32_t v = (uint32_t)__shfl((int)incl, (int)mid);
                if (v > g) hi = mid;
                else lo = mid + 1;
            }
            const uint32_t e = (uint32_t)__shfl((int)(incl - c), (int)lo);
            const uint32_t cj0 = (uint32_t)__shfl((int)c0, (int)lo), hj = (uint32_t)__shfl((int)h, (int)lo);
            const uint32_t ii = g - e;
            const uint64_t *L = a.lists + (sb + lo) * (uint64_t)a.list_cap;
            const uint64_t key = g0 + lane < total ? L[ii < cj0 ? ii : hj + (ii - cj0)] : kKeyMax;
            offer_lds<B>(list, cnt, tau, a.kprime, key < tau, key);
        }
    }
    compact_lds<B>(list, cnt, tau, a.kprime);  // sorted, cnt <= kprime
    if (a.pass == 3) {
        if (lane == 0) {
            a.tau0[q] = tau;
            a.gtau[q] = tau;
        }
        return;
    }
    // tau is now the k'-th smallest merged key (kKeyMax when fewer).  The
    // merged lists hold every candidate key at or below the seed tau0, so the
    // merged top-k' is exact when its k'-th key is at or below tau0.  Otherwise
    // the query is rescanned with the k'-th merged key as its threshold (the
    // k'-th key of any subset bounds the k'-th key of the whole candidate set).
    if (a.pass == 1) {
        const uint64_t t0 = a.tau0[q];
        if (t0 != kKeyMax && tau > t0) {
            if (lane == 0) {
                a.flags[q] = 1u;
                a.gtau[q] = tau;
                atomicAdd(a.nflag, 1u);
            }
            return;
        }
    }

    if (a.rerank) {
        const float *__restrict__ qv = a.queries + q * kDim;
        const uint32_t nsort = pow2_at_least(cnt);
        for (uint32_t i = lane; i < nsort; i += 64) {
            uint64_t nk = kKeyMax;
            if (i < cnt) {
                const uint32_t low = (uint32_t)list[i];
                const uint32_t id = a.key_pos ? a.ids_lay[low >> 31][low & 0x7FFFFFFFu] : low;
                const float d = canon_dist_row(a.base_rows + (uint64_t)id * kDim, qv);
                nk = ((uint64_t)f2ord(d) << 32) | id;
            }
            list[i] = nk;
        }
        __syncthreads();
        wave_sort_lds_n(list, nsort);
        if (a.cert) {
            // Rows outside the candidates have BF16 scores at or above the
            // k'-th one, ts (keys are unique, so "outside" means "key above
            // tau").  The tile forms a row's score as |x|^2_hi + |x|^2_lo +
            // sum_i bf16(x_i) (-2 bf16(q_i)) (exact products, FP32 sums).  With
            // qb = bf16(q), xb = bf16(x): qb.xb - q.x = (qb - q).xb + q.(xb - x),
            // so |error of -2 qb.xb| <= 2 |x| (|qb - q| (1 + rho) + |q| rho),
            // rho >= |xb - x| / |x| over the rows.  The hi + lo split of the
            // FP32 norm is within 2^-16 |x|^2 of it and the FP32 norm within
            // 100 * 2^-24 of |x|^2 (2.2e-5 together); summing 102 terms in
            // FP32 adds at most 111 * 2^-24 of their absolute sum (<= 2.02 |q|
            // |x| + 1.01 |x|^2); 1.4e-5, twice that, also covers an accumulator
            // that truncates.  |x| <= cert_norm.  The canonical
            // FP32 distances and |q|^2 carry relative errors below 1e-5.  Fewer
            // than k' candidates (tau = kKeyMax) means the range held no more
            // rows.
            uint32_t ok = 1u;
            if (tau != kKeyMax && cnt >= a.k) {
                float qn = 0.f, dq2 = 0.f;
                for (int d = 0; d < kDim; ++d) {
                    const float v = qv[d], dv = bf16_round(v) - v;
                    qn = __builtin_fmaf(v, v, qn);
                    dq2 = __builtin_fmaf(dv, dv, dq2);
                }
                const float xm = a.cert_norm, rho = a.cert_rho;
                const float qa = __builtin_sqrtf(qn) * 1.0001f, dq = __builtin_sqrtf(dq2) * 1.0001f;
                const float eps = 1.01f * (2.0f * xm * (dq * (1.0f + rho) + qa * rho) +
                                           1.4e-5f * (2.02f * qa * xm + 1.01f * xm * xm) + 2.2e-5f * xm * xm);
                const float ts = ord2f((uint32_t)(tau >> 32));
                const float lower = (ts + qn * (1.0f - 2e-5f) - eps) * (1.0f - 3e-5f);
                const float dk = ord2f((uint32_t)(list[a.k - 1] >> 32));
                ok = lower > dk ? 1u : 0u;
            }
            if (lane == 0) a.cert[q] = ok;
        }
        if (cnt > a.k) cnt = a.k;
    }
    for (uint32_t i = lane; i < a.k; i += 64) {
        uint32_t id = kPadId;
        float d = __builtin_inff();
        if (i < cnt) {
            const uint64_t key = list[i];
            id = (uint32_t)key;
            d = a.key_int ? __builtin_bit_cast(float, (uint32_t)(key >> 32)) : ord2f((uint32_t)(key >> 32));
        } else if (a.key_int) {
            d = __builtin_bit_cast(float, 0xFFFFFFFFu);
        }
        a.out_ids[q * a.k + i] = id;
        a.out_dists[q * a.k + i] = d;
    }
}

__global__ void gather_rows_kernel(const float *__restrict__ src, const uint32_t *__restrict__ idx, uint64_t n,
                                   float *__restrict__ dst) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * kDim) return;
    const uint64_t i = t / kDim, d = t % kDim;
    dst[t] = src[(uint64_t)idx[i] * kDim + d];
}

__global__ void scatter_results_kernel(const uint32_t *__restrict__ ids, const float *__restrict__ dd,
                                       const uint32_t *__restrict__ idx, uint64_t n, uint32_t k,
                                       uint32_t *__restrict__ out_ids, float *__restrict__ out_d) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * k) return;
    const uint64_t i = t / k, j = t % k;
    const uint64_t o = (uint64_t)idx[i] * k + j;
    out_ids[o] = ids[t];
    out_d[o] = dd[t];
}

hipError_t launch_gather_rows(const float *src, const uint32_t *idx, uint64_t n, float *dst, hipStream_t st) {
    if (n == 0) return hipSuccess;
    const uint64_t total = n * kDim;
    hipLaunchKernelGGL(gather_rows_kernel, dim3((uint32_t)((total + 255) / 256)), dim3(256), 0, st, src, idx, n, dst);
    return hipGetLastError();
}

hipError_t launch_scatter_results(const uint32_t *ids, const float *d, const uint32_t *idx, uint64_t n, uint32_t k,
                                  uint32_t *out_ids, float *out_d, hipStream_t st) {
    if (n == 0) return hipSuccess;
    const uint64_t total = n * k;
    hipLaunchKernelGGL(scatter_results_kernel, dim3((uint32_t)((total + 255) / 256)), dim3(256), 0, st, ids, d, idx,
                       n, k, out_ids, out_d);
    return hipGetLastError();
}

hipError_t launch_merge(const MergeArgs &a, hipStream_t st) {
    if (a.nq == 0) return hipSuccess;
    const uint32_t need = a.kprime > a.k ? a.kprime : a.k;
    // The LDS list keeps need keys plus batches of 64 offers (offer_lds):
    // room for several batches between compactions (each compaction sorts
    // the list), 8 KiB per query up to 416 kept keys, 16 KiB beyond.
    const int B = need <= 416 ? 1024 : 2048;
    if ((int)need + 64 > B) return hipErrorInvalidValue;
    switch (B) {
    case 1024: hipLaunchKernelGGL(merge_kernel<1024>, dim3((uint32_t)a.nq), dim3(64), 0, st, a); break;
    case 2048: hipLaunchKernelGGL(merge_kernel<2048>, dim3((uint32_t)a.nq), dim3(64), 0, st, a); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace fbf
