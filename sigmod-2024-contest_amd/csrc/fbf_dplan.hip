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
//    rule), walked independently in windows of gmax sorted positions (one
//    wave per window, the walk over LDS);
//  * a seeded query (prefix ceil(len / pdiv) >= min_prefix rows) splits its
//    range into the prefix [s, s + pre) and the rest [s + pre, e); a group's
//    main span is cut into chunks of `chunk` rows, one work item per (group,
//    chunk) carrying all members; its prefix items carry at most gpre members
//    (two halves of a group larger than gpre, each with its own prefix span:
//    prefix lists start empty and compact often, and fewer waves per tile
//    hold each other up less);
//  * query q owns nc_pre + nc_main list slots from slot_off[q] (prefix slots
//    end at pre_end[q]): exclusive scans (multi-block, three kernels);
//  * items of each pass ordered chunk-major by a counting sort on their
//    first row (layout, row_begin; items of one first row follow each other
//    in any order) and split into 8 runs of about equal cost (rows x
//    queries), one per XCD: costs are summed per bin of `chunk` rows and a
//    split falls inside a bin in proportion to its cost.
// Capacities (items per pass, list slots) are fixed at graph creation; a plan
// that would exceed them sets dp.status[2], writes nothing out of bounds and
// leaves every query without lists (its outputs padded).

#include <algorithm>
#include <cmath>
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

// One wave per window of gmax sorted positions: the window's ranges go to
// LDS, lane 0 walks them and cuts the groups, then the lanes describe one
// group each -- gd[(w * gmax + j) * kGdStride]: first sorted position, count,
// layout, prefix span of the first gpre members, main span, prefix span of
// the rest; gnc[(w * gmax + j) * kGncStride]: item counts (prefix halves,
// main; 0 for unused group slots) -- and one query each: qnc[2 q] = the
// prefix chunks of q's half, qnc[2 q + 1] = the main chunks of its group.
__global__ __launch_bounds__(64) void dplan_group_kernel(const uint32_t *__restrict__ ranges,
                                                         const uint32_t *__restrict__ order, uint64_t nq,
                                                         uint32_t gmax, uint32_t gpre, uint32_t chunk, uint32_t pdiv,
                                                         uint32_t min_prefix, uint32_t *__restrict__ gd,
                                                         uint32_t *__restrict__ gnc, uint32_t *__restrict__ qnc) {
    __shared__ uint32_t sl[256], ss[256], se[256], sq[256], gid[256], gst[257], c0[256], c1[256], cm[256];
    __shared__ uint32_t s_ng, s_nv;
    const uint64_t w = blockIdx.x;
    const uint32_t lane = threadIdx.x;
    const uint64_t p0 = w * gmax;
    const uint32_t np = (uint32_t)(nq - p0 < gmax ? nq - p0 : gmax);
    for (uint32_t i = lane; i < np; i += 64) {
        const uint32_t q = order[p0 + i];
        sq[i] = q;
        sl[i] = ranges[4 * q];
        ss[i] = ranges[4 * q + 1];
        se[i] = ranges[4 * q + 2];
    }
    __syncthreads();
    if (lane == 0) {
        uint32_t ng = 0, g0 = 0, gl = 0, gs = 0, ge = 0, gml = 0, i = 0;
        for (; i < np; ++i) {
            const uint32_t l = sl[i], s = ss[i], e = se[i];
            if (!(l < 2 && e > s)) break;  // ranges that take no part sort last
            bool fits = i > g0 && l == gl;
            if (fits) {
                const uint32_t ns = min(gs, s), ne = max(ge, e), ml = max(gml, e - s);
                fits = (uint64_t)(ne - ns) * 4 <= (uint64_t)ml * 5 + 4 * kRowTile;
            }
            if (!fits) {
                if (i > g0) gst[ng++] = g0;
                g0 = i;
                gl = l;
                gs = s;
                ge = e;
                gml = 0;
            }
            gs = min(gs, s);
            ge = max(ge, e);
            gml = max(gml, e - s);
            gid[i] = ng;
        }
        if (i > g0) gst[ng++] = g0;
        gst[ng] = i;
        s_ng = ng;
        s_nv = i;
    }
    __syncthreads();
    const uint32_t ng = s_ng, nv = s_nv;
    for (uint32_t j = lane; j < gmax; j += 64) {
        uint32_t *G = gnc + (w * gmax + j) * kGncStride;
        if (j >= ng) {
            G[0] = G[1] = G[2] = 0;
            continue;
        }
        const uint32_t a = gst[j], b = gst[j + 1];
        uint32_t ps0 = 0xFFFFFFFFu, pe0 = 0, ps1 = 0xFFFFFFFFu, pe1 = 0, ms = 0xFFFFFFFFu, me = 0;
        for (uint32_t i = a; i < b; ++i) {
            const uint32_t s = ss[i], e = se[i];
            const uint32_t pre = dplan_prefix(s, e, pdiv, min_prefix);
            if (pre && i - a < gpre) {
                ps0 = min(ps0, s);
                pe0 = max(pe0, s + pre);
            } else if (pre) {
                ps1 = min(ps1, s);
                pe1 = max(pe1, s + pre);
            }
            ms = min(ms, s + pre);
            me = max(me, e);
        }
        const uint32_t n0 = ps0 < pe0 ? dplan_chunks(ps0, pe0, chunk) : 0u;
        const uint32_t n1 = ps1 < pe1 ? dplan_chunks(ps1, pe1, chunk) : 0u;
        const uint32_t nm = dplan_chunks(ms, me, chunk);
        uint32_t *d = gd + (w * gmax + j) * kGdStride;
        d[0] = (uint32_t)(p0 + a);
        d[1] = b - a;
        d[2] = sl[a];
        d[3] = ps0;
        d[4] = pe0;
        d[5] = ms;
        d[6] = me;
        d[7] = ps1;
        d[8] = pe1;
        G[0] = n0;
        G[1] = n1;
        G[2] = nm;
        c0[j] = n0;
        c1[j] = n1;
        cm[j] = nm;
    }
    __syncthreads();
    for (uint32_t i = lane; i < np; i += 64) {
        const uint32_t q = sq[i];
        if (i >= nv) {  // no range: no slots
            qnc[2 * q] = 0;
            qnc[2 * q + 1] = 0;
            continue;
        }
        const uint32_t j = gid[i];
        qnc[2 * q] = i - gst[j] < gpre ? c0[j] : c1[j];
        qnc[2 * q + 1] = cm[j];
    }
}

// Value i of a scan's input: in[i * stride + field] (+ in[i * stride + field2]
// unless field2 is 0xFFFFFFFF).
template <class T>
__device__ inline uint64_t dplan_val(const T *in, uint32_t stride, uint32_t field, uint32_t field2, uint64_t i) {
    uint64_t v = in[i * stride + field];
    if (field2 != 0xFFFFFFFFu) v += in[i * stride + field2];
    return v;
}

// Exclusive scan of a few values into out[0..n] (out[n] = total), one
// workgroup of 1024 threads: each thread sums a contiguous segment, the
// segment sums are scanned in LDS.
template <class T>
__global__ __launch_bounds__(1024) void dplan_scan_small_kernel(const T *__restrict__ in, uint32_t stride,
                                                               uint32_t field, uint32_t field2, uint64_t n,
                                                               uint64_t *__restrict__ out) {
    __shared__ uint64_t part[1024];
    const uint32_t t = threadIdx.x;
    const uint64_t per = (n + 1023) / 1024;
    const uint64_t b = t * per, e = b + per < n ? b + per : n;
    uint64_t sum = 0;
    for (uint64_t i = b; i < e; ++i) sum += dplan_val(in, stride, field, field2, i);
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
        acc += dplan_val(in, stride, field, field2, i);
    }
    if (t == 1023) out[n] = part[1023];
}

// Larger scans: per-tile sums, a scan of those, then each tile's scan.
constexpr uint32_t kScanTile = 4096;

__global__ __launch_bounds__(256) void dplan_scan_reduce_kernel(const uint32_t *__restrict__ in, uint32_t stride,
                                                                uint32_t field, uint32_t field2, uint64_t n,
                                                                uint64_t *__restrict__ part) {
    __shared__ uint64_t red[4];
    const uint64_t b0 = (uint64_t)blockIdx.x * kScanTile, b1 = b0 + kScanTile < n ? b0 + kScanTile : n;
    unsigned long long x = 0;
    for (uint64_t i = b0 + threadIdx.x; i < b1; i += 256) x += dplan_val(in, stride, field, field2, i);
    for (int o = 32; o > 0; o >>= 1) x += __shfl_down(x, o);
    if ((threadIdx.x & 63u) == 0) red[threadIdx.x >> 6] = x;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

__global__ __launch_bounds__(256) void dplan_scan_down_kernel(const uint32_t *__restrict__ in, uint32_t stride,
                                                              uint32_t field, uint32_t field2, uint64_t n,
                                                              const uint64_t *__restrict__ poff, uint64_t nb,
                                                              uint64_t *__restrict__ out) {
    __shared__ uint64_t wsum[4];
    const uint64_t b0 = (uint64_t)blockIdx.x * kScanTile, b1 = b0 + kScanTile < n ? b0 + kScanTile : n;
    const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
    uint64_t carry = poff[blockIdx.x];
    for (uint64_t base = b0; base < b1; base += 256) {
        const uint64_t i = base + threadIdx.x;
        const unsigned long long v = i < b1 ? dplan_val(in, stride, field, field2, i) : 0ull;
        unsigned long long x = v;  // inclusive scan over the wave
        for (int o = 1; o < 64; o <<= 1) {
            const unsigned long long y = __shfl_up(x, o);
            if (lane >= (uint32_t)o) x += y;
        }
        if (lane == 63) wsum[w] = x;
        __syncthreads();
        uint64_t before = carry;
        for (uint32_t u = 0; u < w; ++u) before += wsum[u];
        if (i < b1) out[i] = before + x - v;
        carry += wsum[0] + wsum[1] + wsum[2] + wsum[3];
        __syncthreads();
    }
    if (blockIdx.x == nb - 1 && threadIdx.x == 0) out[n] = poff[nb];
}

// Items of both passes, one thread per group slot; each item takes the next
// place among its pass's items of its first row (layout, row_begin) and adds
// its cost (rows x queries) to its chunk bin's (layout, row_begin / chunk).
__global__ void dplan_items_kernel(const uint32_t *__restrict__ gd, const uint32_t *__restrict__ gnc,
                                   const uint64_t *__restrict__ ioff_pre, const uint64_t *__restrict__ ioff_main,
                                   uint64_t ngslot, uint32_t gpre, uint32_t chunk, uint64_t max_items, uint32_t nrb,
                                   uint32_t nbins,
                                   WorkItem *__restrict__ it_pre, WorkItem *__restrict__ it_main,
                                   uint32_t *__restrict__ sl_pre, uint32_t *__restrict__ sl_main,
                                   uint32_t *__restrict__ bins, unsigned long long *__restrict__ bcost,
                                   uint32_t *__restrict__ status) {
    const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ngslot) return;
    const uint32_t *d = gd + g * kGdStride;
    const uint32_t *G = gnc + g * kGncStride;
    const uint32_t half = nbins / 2, rhalf = nrb / 2;
    // (pass, first item, chunks, span, first member, members)
    auto emit = [&](int pass, uint64_t o, uint32_t nc, uint32_t us, uint32_t ue, uint32_t qb, uint32_t qc) {
        const uint32_t b0 = us / kRowTile * kRowTile;
        WorkItem *it = pass ? it_main : it_pre;
        uint32_t *sl = pass ? sl_main : sl_pre;
        for (uint32_t c = 0; c < nc; ++c) {
            WorkItem W{};
            W.row_begin = b0 + c * chunk;
            W.row_end = (uint32_t)min((uint64_t)W.row_begin + chunk, (uint64_t)ue);
            W.q_begin = qb;
            W.q_count = qc;
            W.layout = d[2];
            W.chunk = c;
            it[o + c] = W;
            sl[o + c] = atomicAdd(&bins[(pass ? nrb : 0u) + d[2] * rhalf + W.row_begin / kRowTile], 1u);
            atomicAdd(&bcost[(pass ? nbins : 0u) + d[2] * half + W.row_begin / chunk],
                      (unsigned long long)(W.row_end - W.row_begin) * qc);
        }
    };
    const uint32_t np = G[0] + G[1], nm = G[2];
    if (np) {
        const uint64_t o = ioff_pre[g];
        if (o + np > max_items) {
            atomicOr(&status[2], 1u);
        } else {
            const uint32_t q0 = min(d[1], gpre);
            if (G[0]) emit(0, o, G[0], d[3], d[4], d[0], q0);
            if (G[1]) emit(0, o + G[0], G[1], d[7], d[8], d[0] + gpre, d[1] - q0);
        }
    }
    if (nm) {
        const uint64_t o = ioff_main[g];
        if (o + nm > max_items) atomicOr(&status[2], 1u);
        else emit(1, o, nm, d[5], d[6], d[0], d[1]);
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

// Counts and overflow checks.
__global__ void dplan_finish_kernel(const uint64_t *__restrict__ ioff_pre, const uint64_t *__restrict__ ioff_main,
                                    const uint64_t *__restrict__ slot_off, uint64_t ngslot, uint64_t nq,
                                    uint64_t max_items, uint64_t max_slots, uint32_t *__restrict__ status,
                                    uint64_t *__restrict__ n_items) {
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
}

// A plan over capacity leaves every query without lists: the merges pad.
__global__ void dplan_overflow_kernel(const uint32_t *__restrict__ status, uint64_t nq, uint64_t *__restrict__ slot_off,
                                      uint64_t *__restrict__ pre_end) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i > nq || status[2] == 0) return;
    slot_off[i] = 0;
    if (i < nq) pre_end[i] = 0;
}

// Items into chunk-major order: item i of a pass goes to the first place of
// its first row plus the place it took there.
__global__ void dplan_place_items_kernel(const WorkItem *__restrict__ src, const uint32_t *__restrict__ sl,
                                         const uint64_t *__restrict__ boff, const uint64_t *__restrict__ n_dev,
                                         uint32_t nrb, WorkItem *__restrict__ dst) {
    const uint64_t n = *n_dev;
    const uint32_t rhalf = nrb / 2;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const WorkItem W = src[i];
        dst[boff[W.layout * rhalf + W.row_begin / kRowTile] + sl[i]] = W;
    }
}

// XCD runs of each pass (wave = pass): 8 runs of about equal cost, the
// boundary inside a chunk bin placed in proportion to the bin's cost.  Chunk
// bin b (layout b / (nbins / 2)) starts at the first item whose first row is
// at or after its first row.  Each lane sums a segment of bins; the lane
// whose segment reaches a boundary's cost places it.
__global__ __launch_bounds__(128) void dplan_xcd_kernel(const uint64_t *__restrict__ boff, uint32_t nrb,
                                                        const unsigned long long *__restrict__ bcost, uint32_t nbins,
                                                        uint32_t chunk, const uint64_t *__restrict__ n_items,
                                                        uint32_t *__restrict__ xoff_pre,
                                                        uint32_t *__restrict__ xoff_main) {
    const uint32_t pass = threadIdx.x >> 6, lane = threadIdx.x & 63u;
    const uint64_t *bo = boff + pass * (nrb + 1);
    const unsigned long long *bc = bcost + pass * nbins;
    const uint32_t half = nbins / 2, rhalf = nrb / 2, per = chunk / kRowTile;
    auto start = [&](uint32_t b) -> uint64_t {  // first item of chunk bin b (b <= nbins)
        const uint32_t l = b / half, cb = b % half;
        if (l >= 2) return bo[nrb];
        return bo[min((uint64_t)l * rhalf + (uint64_t)cb * per, (uint64_t)(l + 1) * rhalf)];
    };
    uint32_t *xo = pass ? xoff_main : xoff_pre;
    const uint64_t n = n_items[pass];
    if (lane == 0) {
        xo[0] = 0;
        xo[8] = (uint32_t)n;
    }
    if (n == 0) {
        if (lane >= 1 && lane < 8) xo[lane] = 0;
        return;
    }
    const uint32_t seg = (nbins + 63) / 64, b0 = min(lane * seg, nbins), b1 = min(b0 + seg, nbins);
    unsigned long long mine = 0;
    for (uint32_t b = b0; b < b1; ++b) mine += bc[b];
    unsigned long long incl = mine;
    for (int o = 1; o < 64; o <<= 1) {
        const unsigned long long y = __shfl_up(incl, o);
        if (lane >= (uint32_t)o) incl += y;
    }
    const unsigned long long excl = incl - mine;
    const double total = (double)__shfl(incl, 63);
    for (int x = 1; x < 8; ++x) {
        const double target = total * x / 8.0;
        // The first bin whose inclusive cost reaches the target lies in this
        // lane's segment.
        if (!((double)incl >= target && (double)excl < target)) continue;
        unsigned long long cum = excl;
        for (uint32_t b = b0; b < b1; ++b) {
            const double c = (double)bc[b];
            if ((double)cum + c >= target) {
                const double f = c > 0 ? (target - (double)cum) / c : 0.0;
                const uint64_t s0 = start(b), s1 = start(b + 1);
                xo[x] = (uint32_t)min((double)n, (double)s0 + floor(f * (double)(s1 - s0) + 0.5));
                break;
            }
            cum += bc[b];
        }
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

uint64_t dplan_scan_parts(uint64_t nq, uint32_t gmax, uint64_t n_rows) {
    const uint64_t ngslot = (nq + gmax - 1) / gmax * gmax;
    const uint64_t n = std::max<uint64_t>({256 * ((nq + 255) / 256), ngslot, nq, (uint64_t)dplan_row_bins(n_rows)});
    return (n + kScanTile - 1) / kScanTile;
}

// Exclusive scan of n values of in (as dplan_val reads them) into out[0..n].
static void dplan_scan(const DPlan &dp, const uint32_t *in, uint32_t stride, uint32_t field, uint32_t field2,
                       uint64_t n, uint64_t *out, hipStream_t st) {
    const uint64_t nb = (n + kScanTile - 1) / kScanTile;
    hipLaunchKernelGGL(dplan_scan_reduce_kernel, dim3((uint32_t)nb), dim3(256), 0, st, in, stride, field, field2, n,
                       dp.part);
    hipLaunchKernelGGL(dplan_scan_small_kernel<uint64_t>, dim3(1), dim3(1024), 0, st, dp.part, 1u, 0u, 0xFFFFFFFFu,
                       nb, dp.poff);
    hipLaunchKernelGGL(dplan_scan_down_kernel, dim3((uint32_t)nb), dim3(256), 0, st, in, stride, field, field2, n,
                       dp.poff, nb, out);
}

hipError_t launch_dplan(const DPlan &dp, hipStream_t st) {
    const uint64_t nq = dp.nq;
    if (nq == 0) return hipSuccess;
    if (dp.gmax == 0 || dp.gmax > 256 || dp.gpre == 0 || dp.gpre > dp.gmax || dp.gmax > 2 * dp.gpre ||
        dp.chunk == 0 || dp.chunk % kRowTile)
        return hipErrorInvalidValue;
    if (dp.n_rows >= (1ull << 31) || nq >= (1ull << 31)) return hipErrorInvalidValue;
    const uint32_t bq = (uint32_t)((nq + 255) / 256);
    const uint64_t nwin = (nq + dp.gmax - 1) / dp.gmax, ngslot = nwin * dp.gmax;
    const uint32_t nbins = dplan_bins(dp.n_rows, dp.chunk), nrb = dplan_row_bins(dp.n_rows);
    const uint32_t bits = dplan_row_bits(dp.n_rows);
    hipLaunchKernelGGL(dplan_zero_kernel, dim3((2 * nrb + 255) / 256), dim3(256), 0, st, dp.bins, 2ull * nrb);
    hipLaunchKernelGGL(dplan_zero_kernel, dim3((4 * nbins + 255) / 256), dim3(256), 0, st,
                       reinterpret_cast<uint32_t *>(dp.bcost), 4ull * nbins);
    // Query order: LSD passes ping-pong between (keys, vals) and (skeys, order).
    hipLaunchKernelGGL(dplan_keys_kernel, dim3(bq), dim3(256), 0, st, dp.ranges, nq, bits, dp.keys, dp.vals);
    uint64_t *k0 = dp.keys, *k1 = dp.skeys;
    uint32_t *v0 = dp.vals, *v1 = dp.order;
    for (uint32_t shift = 0; shift < 2 * bits + 1; shift += 8) {
        hipLaunchKernelGGL(dplan_rhist_kernel, dim3(bq), dim3(256), 0, st, k0, nq, shift, bq, dp.hist);
        dplan_scan(dp, dp.hist, 1u, 0u, 0xFFFFFFFFu, 256ull * bq, dp.hoff, st);
        hipLaunchKernelGGL(dplan_rscatter_kernel, dim3(bq), dim3(256), 0, st, k0, v0, nq, shift, bq, dp.hoff, k1,
                           v1);
        std::swap(k0, k1);
        std::swap(v0, v1);
    }
    const uint32_t *order = v0;  // sorted positions -> queries
    hipLaunchKernelGGL(dplan_group_kernel, dim3((uint32_t)nwin), dim3(64), 0, st, dp.ranges, order, nq, dp.gmax,
                       dp.gpre, dp.chunk, dp.pdiv, dp.min_prefix, dp.gd, dp.gnc, dp.qnc);
    dplan_scan(dp, dp.gnc, kGncStride, 0u, 1u, ngslot, dp.ioff_pre, st);
    dplan_scan(dp, dp.gnc, kGncStride, 2u, 0xFFFFFFFFu, ngslot, dp.ioff_main, st);
    dplan_scan(dp, dp.qnc, 2u, 0u, 1u, nq, dp.slot_off, st);
    hipLaunchKernelGGL(dplan_items_kernel, dim3((uint32_t)((ngslot + 255) / 256)), dim3(256), 0, st, dp.gd, dp.gnc,
                       dp.ioff_pre, dp.ioff_main, ngslot, dp.gpre, dp.chunk, dp.max_items, nrb, nbins,
                       dp.items_tmp_pre,
                       dp.items_tmp_main, dp.islot_pre, dp.islot_main, dp.bins, dp.bcost, dp.status);
    hipLaunchKernelGGL(dplan_iq_kernel, dim3(bq), dim3(256), 0, st, dp.ranges, order, dp.qnc, dp.slot_off, nq, dp.pdiv,
                       dp.min_prefix, dp.iq_pre, dp.iq_main, dp.pre_end);
    hipLaunchKernelGGL(dplan_finish_kernel, dim3(1), dim3(64), 0, st, dp.ioff_pre, dp.ioff_main, dp.slot_off, ngslot,
                       nq, dp.max_items, dp.max_slots, dp.status, dp.n_items);
    hipLaunchKernelGGL(dplan_overflow_kernel, dim3((uint32_t)((nq + 256) / 256)), dim3(256), 0, st, dp.status, nq,
                       dp.slot_off, dp.pre_end);
    for (int pass = 0; pass < 2; ++pass) {
        uint64_t *bo = dp.boff + (pass ? nrb + 1 : 0);
        dplan_scan(dp, dp.bins + (pass ? nrb : 0), 1u, 0u, 0xFFFFFFFFu, (uint64_t)nrb, bo, st);
        hipLaunchKernelGGL(dplan_place_items_kernel, dim3(1024), dim3(256), 0, st,
                           pass ? dp.items_tmp_main : dp.items_tmp_pre, pass ? dp.islot_main : dp.islot_pre, bo,
                           dp.n_items + pass, nrb, pass ? dp.items_main : dp.items_pre);
    }
    hipLaunchKernelGGL(dplan_xcd_kernel, dim3(1), dim3(128), 0, st, dp.boff, nrb, dp.bcost, nbins, dp.chunk,
                       dp.n_items, dp.xoff_pre, dp.xoff_main);
    return hipGetLastError();
}

}  // namespace fbf
