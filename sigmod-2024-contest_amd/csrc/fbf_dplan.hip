// fbf_dplan.hip -- the work-item planner of the seeded MFMA search, on the device.
//
// The host planner (fbf_api.hip: make_plan, concat_plans, xcd_runs) reads the
// batch's ranges on the host; a captured HIP graph built on it is fixed to
// one batch's filters.  These kernels build the same kind of plan from ranges
// in device memory (the router's output), so one captured graph serves any
// batch of its size: the router kernel, this planner and the scans, merges
// and rerank all run inside the graph.
//
// Plan (for nq queries with ranges [s, e) of layout 0 or 1; other layouts and
// empty ranges take no part):
//  * queries sorted by (layout, s, e) -- a stable LSD radix sort of packed
//    keys, 8 bits per pass (histogram, scan, scatter; no library sort, whose
//    memset nodes a captured graph must not hold);
//  * groups: consecutive sorted queries, at most gmax, cut where the shared
//    span would exceed the longest member's range by more than 25% (the host
//    rule), walked independently in windows of gmax sorted positions;
//  * a seeded query (prefix ceil(len / pdiv) >= min_prefix rows) splits its
//    range into the prefix [s, s + pre) and the rest [s + pre, e); a group's
//    prefix span and main span are cut into chunks of `chunk` rows: one work
//    item per (group, chunk) in each pass, carrying all members of the group;
//  * query q owns nc_pre + nc_main list slots from slot_off[q] (prefix slots
//    end at pre_end[q]): exclusive scans;
//  * items of each pass ordered chunk-major by a counting sort on row bins
//    (layout, first row / chunk) and split into 8 runs of equal count, one
//    per XCD (items of one bin follow each other in any order).
// Capacities (items per pass, list slots) are fixed at graph creation; a plan
// that would exceed them sets dp.status[2], writes nothing out of bounds and
// leaves every query without lists (its outputs padded).

#include <utility>

#include "fbf_internal.hpp"

namespace fbf {

// Sort key of a query: (layout, s, e) packed into 2 * bits + 1 bits (e <= n
// < 2^bits); ranges that take no part get all ones and sort last.
__global__ void dplan_keys_kernel(const uint32_t *__restrict__ ranges, uint64_t nq, uint32_t bits,
                                  uint64_t *__restrict__ keys, uint32_t *__restrict__ vals) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nq) return;
    const uint32_t layout = ranges[4 * i], s = ranges[4 * i + 1], e = ranges[4 * i + 2];
    const bool ok = layout < 2 && e > s;
    keys[i] = ok ? ((uint64_t)layout << (2 * bits)) | ((uint64_t)s << bits) | e : (1ull << (2 * bits + 1)) - 1;
    vals[i] = (uint32_t)i;
}

__device__ inline bool dplan_valid(const uint32_t *ranges, uint32_t q) {
    return ranges[4 * q] < 2 && ranges[4 * q + 2] > ranges[4 * q + 1];
}

// One pass of the query sort over digit (key >> shift) & 255: per-block digit
// counts, digit-major (hist[d * nblocks + block]), so that their exclusive
// scan is each (digit, block)'s first output position.
__global__ __launch_bounds__(256) void dplan_rhist_kernel(const uint64_t *__restrict__ keys, uint64_t n,
                                                          uint32_t shift, uint32_t nblocks,
                                                          uint32_t *__restrict__ hist) {
    __shared__ uint32_t c[256];
    c[threadIdx.x] = 0;
    __syncthreads();
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) atomicAdd(&c[(uint32_t)(keys[i] >> shift) & 255u], 1u);
    __syncthreads();
    hist[(uint64_t)threadIdx.x * nblocks + blockIdx.x] = c[threadIdx.x];
}

// Stable scatter: an element goes after the block's earlier elements of its
// digit (earlier waves: their counts in LDS; earlier lanes of its own wave:
// the lanes whose digit matches, from eight ballots).
__global__ __launch_bounds__(256) void dplan_rscatter_kernel(const uint64_t *__restrict__ kin,
                                                             const uint32_t *__restrict__ vin, uint64_t n,
                                                             uint32_t shift, uint32_t nblocks,
                                                             const uint64_t *__restrict__ hoff,
                                                             uint64_t *__restrict__ kout,
                                                             uint32_t *__restrict__ vout) {
    __shared__ uint32_t wc[4][256];
    for (uint32_t t = threadIdx.x; t < 4 * 256; t += 256) wc[t >> 8][t & 255u] = 0;
    __syncthreads();
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    const bool ok = i < n;
    const uint64_t key = ok ? kin[i] : 0;
    const uint32_t d = (uint32_t)(key >> shift) & 255u;
    const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
    uint64_t peers = __ballot(ok);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
        const bool bit = (d >> b) & 1u;
        const uint64_t m = __ballot(bit);
        peers &= bit ? m : ~m;
    }
    const uint32_t rank = (uint32_t)__popcll(peers & ((1ull << lane) - 1));
    if (ok && rank == 0) wc[w][d] = (uint32_t)__popcll(peers);
    __syncthreads();
    if (!ok) return;
    uint32_t before = rank;
    for (uint32_t v = 0; v < w; ++v) before += wc[v][d];
    const uint64_t pos = hoff[(uint64_t)d * nblocks + blockIdx.x] + before;
    kout[pos] = key;
    vout[pos] = vin[i];
}

__device__ inline uint32_t dplan_chunks(uint32_t us, uint32_t ue, uint32_t chunk) {
    if (ue <= us) return 0;
    const uint32_t b0 = us / kRowTile * kRowTile;
    return (uint32_t)(((uint64_t)ue - b0 + chunk - 1) / chunk);
}

// The prefix length of a range (0: not seeded).
__device__ inline uint32_t dplan_prefix(uint32_t s, uint32_t e, uint32_t pdiv, uint32_t min_prefix) {
    if (pdiv <= 1) return 0;
    const uint32_t len = e - s;
    const uint32_t pre = (uint32_t)(((uint64_t)len + pdiv - 1) / pdiv);
    return pre >= min_prefix ? pre : 0u;
}

// One thread per window of gmax sorted positions.  Group descriptors go to
// gd[(w * gmax + j) * 8]: first sorted position, count, layout, prefix span,
// main span; gnc[(w * gmax + j) * 2]: item counts of the two passes (0 for
// unused slots); qnc[q * 2]: the counts of q's group.
__global__ void dplan_group_kernel(const uint32_t *__restrict__ ranges, const uint32_t *__restrict__ order, uint64_t nq, uint32_t gmax, uint32_t chunk,
                                   uint32_t pdiv, uint32_t min_prefix, uint32_t *__restrict__ gd,
                                   uint32_t *__restrict__ gnc, uint32_t *__restrict__ qnc) {
    const uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t nwin = (nq + gmax - 1) / gmax;
    if (w >= nwin) return;
    const uint64_t p0 = w * gmax, p1 = p0 + gmax < nq ? p0 + gmax : nq;
    uint32_t j = 0;
    uint64_t g0 = p0;
    uint32_t gl = 0, gs = 0, ge = 0, gml = 0;
    auto close = [&](uint64_t pe) {
        // members order[g0 .. pe)
        uint32_t ps = 0xFFFFFFFFu, pe_ = 0, ms = 0xFFFFFFFFu, me = 0;
        for (uint64_t p = g0; p < pe; ++p) {
            const uint32_t q = order[p];
            const uint32_t s = ranges[4 * q + 1], e = ranges[4 * q + 2];
            const uint32_t pre = dplan_prefix(s, e, pdiv, min_prefix);
            if (pre) {
                ps = min(ps, s);
                pe_ = max(pe_, s + pre);
            }
            ms = min(ms, s + pre);
            me = max(me, e);
        }
        const uint32_t ncp = ps < pe_ ? dplan_chunks(ps, pe_, chunk) : 0u;
        const uint32_t ncm = dplan_chunks(ms, me, chunk);
        uint32_t *d = gd + (w * gmax + j) * 8;
        d[0] = (uint32_t)g0;
        d[1] = (uint32_t)(pe - g0);
        d[2] = gl;
        d[3] = ps;
        d[4] = pe_;
        d[5] = ms;
        d[6] = me;
        d[7] = 0;
        gnc[(w * gmax + j) * 2] = ncp;
        gnc[(w * gmax + j) * 2 + 1] = ncm;
        for (uint64_t p = g0; p < pe; ++p) {
            const uint32_t q = order[p];
            qnc[2 * q] = ncp;
            qnc[2 * q + 1] = ncm;
        }
        ++j;
    };
    uint64_t p = p0;
    for (; p < p1; ++p) {
        const uint32_t q = order[p];
        if (!dplan_valid(ranges, q)) break;  // ranges that take no part sort last
        const uint32_t l = ranges[4 * q], s = ranges[4 * q + 1], e = ranges[4 * q + 2];
        bool fits = p > g0 && l == gl;
        if (fits) {
            const uint32_t ns = min(gs, s), ne = max(ge, e), ml = max(gml, e - s);
            fits = (uint64_t)(ne - ns) * 4 <= (uint64_t)ml * 5 + 4 * kRowTile;
        }
        if (!fits) {
            if (p > g0) close(p);
            g0 = p;
            gl = l;
            gs = s;
            ge = e;
            gml = 0;
        }
        gs = min(gs, s);
        ge = max(ge, e);
        gml = max(gml, e - s);
    }
    if (p > g0) close(p);
    for (uint64_t r = p; r < p1; ++r) {  // empty ranges own no slots
        const uint32_t q = order[r];
        qnc[2 * q] = 0;
        qnc[2 * q + 1] = 0;
    }
    for (uint32_t jj = j; jj < gmax; ++jj) {
        gnc[(w * gmax + jj) * 2] = 0;
        gnc[(w * gmax + jj) * 2 + 1] = 0;
    }
}

// Exclusive scans of n values (stride `stride` u32 apart, field `field`) into
// out[0..n] (u64, out[n] = total), one workgroup of 1024 threads: each thread
// sums a contiguous segment, the segment sums are scanned in LDS.
__global__ __launch_bounds__(1024) void dplan_scan_kernel(const uint32_t *__restrict__ in, uint32_t stride,
                                                         uint32_t field, uint32_t field2, uint64_t n,
                                                         uint64_t *__restrict__ out) {
    __shared__ uint64_t part[1024];
    const uint32_t t = threadIdx.x;
    const uint64_t per = (n + 1023) / 1024;
    const uint64_t b = t * per, e = b + per < n ? b + per : n;
    auto val = [&](uint64_t i) -> uint64_t {
        uint64_t v = in[i * stride + field];
        if (field2 != 0xFFFFFFFFu) v += in[i * stride + field2];
        return v;
    };
    uint64_t sum = 0;
    for (uint64_t i = b; i < e; ++i) sum += val(i);
    part[t] = sum;
    __syncthreads();
    for (uint32_t off = 1; off < 1024; off <<= 1) {
        const uint64_t v = t >= off ? part[t - off] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    uint64_t acc = part[t] - sum;  // exclusive prefix of this segment
    for (uint64_t i = b; i < e; ++i) {
        out[i] = acc;
        acc += val(i);
    }
    if (t == 1023) out[n] = part[1023];
}

// Items of both passes, one thread per group slot; each item takes the next
// place in its row bin (layout, row_begin / chunk) of its pass.
__global__ void dplan_items_kernel(const uint32_t *__restrict__ gd, const uint32_t *__restrict__ gnc,
                                   const uint64_t *__restrict__ ioff_pre, const uint64_t *__restrict__ ioff_main,
                                   uint64_t ngslot, uint32_t chunk, uint64_t max_items, uint32_t nbins,
                                   WorkItem *__restrict__ it_pre, WorkItem *__restrict__ it_main,
                                   uint32_t *__restrict__ sl_pre, uint32_t *__restrict__ sl_main,
                                   uint32_t *__restrict__ bins, uint32_t *__restrict__ status) {
    const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ngslot) return;
    const uint32_t *d = gd + g * 8;
    const uint32_t half = nbins / 2;
    for (int pass = 0; pass < 2; ++pass) {
        const uint32_t nc = gnc[g * 2 + pass];
        if (nc == 0) continue;
        const uint64_t o = pass ? ioff_main[g] : ioff_pre[g];
        if (o + nc > max_items) {
            atomicOr(&status[2], 1u);
            continue;
        }
        const uint32_t us = pass ? d[5] : d[3], ue = pass ? d[6] : d[4];
        const uint32_t b0 = us / kRowTile * kRowTile;
        WorkItem *it = pass ? it_main : it_pre;
        uint32_t *sl = pass ? sl_main : sl_pre;
        uint32_t *bn = bins + (pass ? nbins : 0);
        for (uint32_t c = 0; c < nc; ++c) {
            WorkItem W{};
            W.row_begin = b0 + c * chunk;
            W.row_end = (uint32_t)min((uint64_t)W.row_begin + chunk, (uint64_t)ue);
            W.q_begin = d[0];
            W.q_count = d[1];
            W.layout = d[2];
            W.chunk = c;
            it[o + c] = W;
            sl[o + c] = atomicAdd(&bn[d[2] * half + W.row_begin / chunk], 1u);
        }
    }
}

// Per sorted position: the ItemQuery of each pass; per query: pre_end.
__global__ void dplan_iq_kernel(const uint32_t *__restrict__ ranges, const uint32_t *__restrict__ order,
                                const uint32_t *__restrict__ qnc, const uint64_t *__restrict__ slot_off, uint64_t nq,
                                uint32_t pdiv, uint32_t min_prefix, ItemQuery *__restrict__ iq_pre,
                                ItemQuery *__restrict__ iq_main, uint64_t *__restrict__ pre_end) {
    const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nq) return;
    const uint32_t q = order[p];
    const uint64_t so = slot_off[q], pe = so + qnc[2 * q];
    pre_end[q] = pe;
    if (!dplan_valid(ranges, q)) return;
    const uint32_t s = ranges[4 * q + 1], e = ranges[4 * q + 2];
    const uint32_t pre = dplan_prefix(s, e, pdiv, min_prefix);
    iq_pre[p] = ItemQuery{q, s, s + pre, (uint32_t)so};
    iq_main[p] = ItemQuery{q, s + pre, e, (uint32_t)pe};
}

// Counts, XCD runs of equal count, overflow checks.
__global__ void dplan_finish_kernel(const uint64_t *__restrict__ ioff_pre, const uint64_t *__restrict__ ioff_main,
                                    const uint64_t *__restrict__ slot_off, uint64_t ngslot, uint64_t nq,
                                    uint64_t max_items, uint64_t max_slots, uint32_t *__restrict__ status,
                                    uint64_t *__restrict__ n_items, uint32_t *__restrict__ xoff_pre,
                                    uint32_t *__restrict__ xoff_main) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    uint64_t np = ioff_pre[ngslot], nm = ioff_main[ngslot];
    const uint64_t ns = slot_off[nq];
    const bool over = np > max_items || nm > max_items || ns > max_slots || status[2] != 0;
    if (over) {
        status[2] = 1u;
        np = nm = 0;
    }
    n_items[0] = np;
    n_items[1] = nm;
    status[0] = (uint32_t)np;
    status[1] = (uint32_t)nm;
    status[3] = (uint32_t)(ns < 0xFFFFFFFFull ? ns : 0xFFFFFFFFull);
    for (int x = 0; x <= 8; ++x) {
        xoff_pre[x] = (uint32_t)(np * x / 8);
        xoff_main[x] = (uint32_t)(nm * x / 8);
    }
}

// A plan over capacity leaves every query without lists: the merges pad.
__global__ void dplan_overflow_kernel(const uint32_t *__restrict__ status, uint64_t nq, uint64_t *__restrict__ slot_off,
                                      uint64_t *__restrict__ pre_end) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i > nq || status[2] == 0) return;
    slot_off[i] = 0;
    if (i < nq) pre_end[i] = 0;
}

// Items into chunk-major order: item i of a pass goes to its bin's first
// place plus the place it took in the bin.
__global__ void dplan_place_items_kernel(const WorkItem *__restrict__ src, const uint32_t *__restrict__ sl,
                                         const uint64_t *__restrict__ boff, const uint64_t *__restrict__ n_dev,
                                         uint32_t chunk, uint32_t nbins, WorkItem *__restrict__ dst) {
    const uint64_t n = *n_dev;
    const uint32_t half = nbins / 2;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const WorkItem W = src[i];
        dst[boff[W.layout * half + W.row_begin / chunk] + sl[i]] = W;
    }
}

__global__ void dplan_zero_kernel(uint32_t *__restrict__ p, uint64_t n) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// Bits that hold every row position 0..n.
static uint32_t dplan_row_bits(uint64_t n) {
    uint32_t b = 1;
    while (b < 31 && (n >> b) != 0) ++b;
    return b;
}

hipError_t launch_dplan(const DPlan &dp, hipStream_t st) {
    const uint64_t nq = dp.nq;
    if (nq == 0) return hipSuccess;
    if (dp.gmax == 0 || dp.gmax > 256 || dp.chunk == 0 || dp.chunk % kRowTile) return hipErrorInvalidValue;
    if (dp.n_rows >= (1ull << 31) || nq >= (1ull << 31)) return hipErrorInvalidValue;
    const uint32_t bq = (uint32_t)((nq + 255) / 256);
    const uint64_t nwin = (nq + dp.gmax - 1) / dp.gmax, ngslot = nwin * dp.gmax;
    const uint32_t nbins = dplan_bins(dp.n_rows, dp.chunk);
    const uint32_t bits = dplan_row_bits(dp.n_rows);
    hipLaunchKernelGGL(dplan_zero_kernel, dim3((2 * nbins + 255) / 256), dim3(256), 0, st, dp.bins, 2ull * nbins);
    // Query order: LSD passes ping-pong between (keys, vals) and (skeys, order).
    hipLaunchKernelGGL(dplan_keys_kernel, dim3(bq), dim3(256), 0, st, dp.ranges, nq, bits, dp.keys, dp.vals);
    uint64_t *k0 = dp.keys, *k1 = dp.skeys;
    uint32_t *v0 = dp.vals, *v1 = dp.order;
    for (uint32_t shift = 0; shift < 2 * bits + 1; shift += 8) {
        hipLaunchKernelGGL(dplan_rhist_kernel, dim3(bq), dim3(256), 0, st, k0, nq, shift, bq, dp.hist);
        hipLaunchKernelGGL(dplan_scan_kernel, dim3(1), dim3(1024), 0, st, dp.hist, 1u, 0u, 0xFFFFFFFFu,
                           256ull * bq, dp.hoff);
        hipLaunchKernelGGL(dplan_rscatter_kernel, dim3(bq), dim3(256), 0, st, k0, v0, nq, shift, bq, dp.hoff, k1,
                           v1);
        std::swap(k0, k1);
        std::swap(v0, v1);
    }
    const uint32_t *order = v0;  // sorted positions -> queries
    hipLaunchKernelGGL(dplan_group_kernel, dim3((uint32_t)((nwin + 63) / 64)), dim3(64), 0, st, dp.ranges, order, nq,
                       dp.gmax, dp.chunk, dp.pdiv, dp.min_prefix, dp.gd, dp.gnc, dp.qnc);
    hipLaunchKernelGGL(dplan_scan_kernel, dim3(1), dim3(1024), 0, st, dp.gnc, 2u, 0u, 0xFFFFFFFFu, ngslot, dp.ioff_pre);
    hipLaunchKernelGGL(dplan_scan_kernel, dim3(1), dim3(1024), 0, st, dp.gnc, 2u, 1u, 0xFFFFFFFFu, ngslot,
                       dp.ioff_main);
    hipLaunchKernelGGL(dplan_scan_kernel, dim3(1), dim3(1024), 0, st, dp.qnc, 2u, 0u, 1u, nq, dp.slot_off);
    hipLaunchKernelGGL(dplan_items_kernel, dim3((uint32_t)((ngslot + 255) / 256)), dim3(256), 0, st, dp.gd, dp.gnc,
                       dp.ioff_pre, dp.ioff_main, ngslot, dp.chunk, dp.max_items, nbins, dp.items_tmp_pre,
                       dp.items_tmp_main, dp.islot_pre, dp.islot_main, dp.bins, dp.status);
    hipLaunchKernelGGL(dplan_iq_kernel, dim3(bq), dim3(256), 0, st, dp.ranges, order, dp.qnc, dp.slot_off, nq, dp.pdiv,
                       dp.min_prefix, dp.iq_pre, dp.iq_main, dp.pre_end);
    hipLaunchKernelGGL(dplan_finish_kernel, dim3(1), dim3(64), 0, st, dp.ioff_pre, dp.ioff_main, dp.slot_off, ngslot,
                       nq, dp.max_items, dp.max_slots, dp.status, dp.n_items, dp.xoff_pre, dp.xoff_main);
    hipLaunchKernelGGL(dplan_overflow_kernel, dim3((uint32_t)((nq + 256) / 256)), dim3(256), 0, st, dp.status, nq,
                       dp.slot_off, dp.pre_end);
    for (int pass = 0; pass < 2; ++pass) {
        uint64_t *bo = dp.boff + (pass ? nbins + 1 : 0);
        hipLaunchKernelGGL(dplan_scan_kernel, dim3(1), dim3(1024), 0, st, dp.bins + (pass ? nbins : 0), 1u, 0u,
                           0xFFFFFFFFu, (uint64_t)nbins, bo);
        hipLaunchKernelGGL(dplan_place_items_kernel, dim3(1024), dim3(256), 0, st,
                           pass ? dp.items_tmp_main : dp.items_tmp_pre, pass ? dp.islot_main : dp.islot_pre, bo,
                           dp.n_items + pass, dp.chunk, nbins, pass ? dp.items_main : dp.items_pre);
    }
    return hipGetLastError();
}

}  // namespace fbf
