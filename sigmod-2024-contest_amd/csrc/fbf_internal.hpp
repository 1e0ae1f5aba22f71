// fbf_internal.hpp -- types shared by the host planner and the HIP kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fbf {

constexpr int kDim = 100;        // the reference hard-codes d = 100 (utils.h:255)
constexpr int kRowTile = 64;     // rows per wave tile: one row per lane
constexpr uint32_t kPadId = 0xFFFFFFFFu;
constexpr uint64_t kKeyMax = ~0ull;

// A work item: one workgroup scans rows [row_begin, row_end) of one sorted
// layout for the queries of one group: entries q_begin .. q_begin + q_count - 1
// of the plan's group table.  The item is chunk number `chunk` of its group's
// span; query j's candidate list lives in slot iq[q_begin + j].slot_base + chunk.
struct WorkItem {
    uint32_t row_begin;   // multiple of kRowTile
    uint32_t row_end;
    uint32_t q_begin;
    uint32_t q_count;
    uint32_t layout;      // 0 = (label, ts) order, 1 = ts order
    uint32_t chunk;
    uint32_t pad[2];
};

// Per (group, query) entry: the query's index in the batch, its candidate
// range in the group's layout, and its first list slot for this pass.
struct ItemQuery {
    uint32_t qi;
    uint32_t s;
    uint32_t e;
    uint32_t slot_base;
};

__device__ __host__ inline uint64_t list_slot(const ItemQuery &e, const WorkItem &W) {
    return (uint64_t)e.slot_base + W.chunk;
}

// Count word of one list slot: the slot holds keys [0, c0) and [h, h + c1)
// (a list kept as two halves that grew separately; h = 0 and c1 = 0 for a
// contiguous list of c0 keys).  c0, c1 < 2048; h a multiple of 64 below 2^16.
__device__ __host__ inline uint32_t list_count_word(uint32_t c0, uint32_t c1, uint32_t h) {
    return c0 | (c1 << 11) | ((h >> 6) << 22);
}
__device__ __host__ inline uint32_t list_count_c0(uint32_t w) { return w & 0x7FFu; }
__device__ __host__ inline uint32_t list_count_c1(uint32_t w) { return (w >> 11) & 0x7FFu; }
__device__ __host__ inline uint32_t list_count_h(uint32_t w) { return (w >> 22) << 6; }

constexpr int kMaxGroup = 16;    // queries per work item
// Unit of MFMA item width: an item of width nw (1, 2, 4) carries at most
// 64 nw queries (widths 1 and 2: one wave per 64 queries; width 4: two waves
// of 128).
constexpr uint32_t kMfmaWaveQueries = 64;

// (distance, id) key: distance bits in the high word.  For the non-negative
// distances of the exact scan the float bits are already ordered; the
// general map also orders negative values (approximate scores).
__host__ __device__ inline uint32_t f2ord(float f) {
    uint32_t u = __builtin_bit_cast(uint32_t, f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__host__ __device__ inline float ord2f(uint32_t o) {
    uint32_t u = (o & 0x80000000u) ? (o & 0x7FFFFFFFu) : ~o;
    return __builtin_bit_cast(float, u);
}
// f rounded to bfloat16 (nearest, ties to even), as a float.
__host__ __device__ inline float bf16_round(float f) {
    const uint32_t u = __builtin_bit_cast(uint32_t, f);
    return __builtin_bit_cast(float, ((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16) << 16);
}

// Read-only data that every lane of a wave reads at the same address (work
// items, query vectors) is accessed through the constant address space so the
// compiler issues scalar loads (s_load_*) into SGPRs; the buffers are never
// written while a kernel that reads them runs.
// (The host pass of hipcc parses device functions too; only the device pass
// knows the constant address space, so the host pass sees plain pointers.)
#if defined(__HIP_DEVICE_COMPILE__)
template <class T> using cptr = const __attribute__((address_space(4))) T *;
#else
template <class T> using cptr = const T *;
#endif
template <class T> __device__ inline cptr<T> as_const(const T *p) { return (cptr<T>)p; }

// List capacity for a kept depth K: room for K kept keys plus at least two
// full tiles of appends.
// Capacity of a candidate list that keeps K keys: at least 96 slots of slack
// above K, so a compaction frees room for several tiles of appends.
inline int list_capacity(int K) {
    if (K <= 160) return 256;
    if (K <= 416) return 512;
    return 1024;
}

// ---------------------------------------------------------------- launchers
struct ScanArgs {
    const WorkItem *items;
    uint64_t n_items;
    const ItemQuery *iq;
    const float *queries;      // [nq][kDim]
    const float *soa;          // layout FP32 copy in 64-row x 4-dim blocks
    const uint32_t *ids;       // layout position -> original id
    uint64_t *lists;           // [slots][B]
    uint32_t *counts;          // [slots]
    uint64_t *gtau;            // [nq] shared per-query threshold (key)
    uint32_t K;
    uint32_t *queue;           // persistent mode work counter (or null)
};

// Symmetric int8 scan (bruteforce.h:244-289 over L2SqrSQ8_sym): rows are
// 112-byte codes, keys are (int distance, id) with id = ids[row] when ids is
// given, else the row position itself.
struct ScanSymArgs {
    const WorkItem *items;
    uint64_t n_items;
    const ItemQuery *iq;
    const int8_t *qcodes;      // [nq][112]
    const int8_t *codes;       // [rows][112]
    const uint32_t *ids;       // may be null
    uint64_t n_rows;           // rows readable from codes (tiles are clamped to it)
    uint64_t *lists;
    uint32_t *counts;
    uint32_t K;
    uint32_t *queue;
};

// One sorted layout as seen by the MFMA scans.
struct LayoutMfma {
    const void *tiles;       // row fragments in MFMA order (variant encoding)
    const float *xnorm;      // [n_pad32] |x|^2 (FP8), |xhat|^2 (SQ8) or |codes|^2 (SYM) per position
    const uint8_t *scale;    // [n_pad32] E8M0 row scale (FP8)
    const uint32_t *ids;     // position -> original id
};

struct ScanMfmaArgs {
    const WorkItem *items;
    uint64_t n_items;
    const ItemQuery *iq;
    const uint4 *qfrag;      // [nq][kQFragU4] B-operand fragments per query (launch_qfrag)
    LayoutMfma lay[2];
    uint64_t *lists;
    uint32_t *counts;
    uint64_t *gtau;          // [nq] shared per-query threshold (key), tightened by every item
    uint32_t K;
    uint32_t list_stride;    // keys between consecutive list slots (>= the launch's list capacity B)
    // Persistent form (queue != null): the items are split into 8 runs, run x
    // = [xoff[x], xoff[x + 1]), one per XCD; workgroups take items from their
    // own XCD's run (counter queue[16 x]) and then from the others' runs.
    uint32_t *queue;
    uint32_t xoff[9];
    const uint32_t *xoff_dev;  // null, or the runs in device memory (a plan built on the device)
    const uint32_t *qmask;   // [nq] or null: when set, only queries with qmask[qi] != 0 are scanned
    const uint32_t *nmask;   // null, or the number of queries set in qmask: the launch exits when 0
};

struct MergeArgs {
    const uint64_t *slot_off;  // [nq + 1]: query q's lists are slots [slot_off[q], slot_off[q + 1])
    const uint64_t *lists;
    const uint32_t *counts;
    uint32_t list_cap;         // B of the scan lists
    uint64_t nq;
    uint32_t k;                // outputs per query
    uint32_t kprime;           // merged depth before rerank (== k for exact)
    const float *queries;      // rerank only
    const float *base_rows;    // rerank only: [n][kDim] original order
    uint32_t *out_ids;
    float *out_dists;          // float distances (key_int == 0) or raw uint32 (key_int == 1)
    int rerank;
    int key_int;               // keys carry an integer distance (symmetric int8 codes)
    int key_pos;               // key low word = (layout << 31) | layout position, not the id
    const uint32_t *ids_lay[2];// position -> id per layout (key_pos)
    // Certificate of the BF16 scan's candidates (rerank only; null: none).
    // cert[q] = 1 when no row outside the k' candidates can have a canonical
    // FP32 distance at or below the k-th reranked one: the k'-th BF16 score
    // less its rounding-error bound, plus |q|^2, lies above it.  The bound
    // uses cert_norm (the largest row norm) and cert_rho (the largest
    // |bf16(x) - x| / |x| over the rows), both rounded up.
    uint32_t *cert;
    float cert_norm;
    float cert_rho;
    // Seeded scans.  pass 0: plain merge.  pass 3 (after the prefix pass):
    // merge only the slots [slot_off[q], pre_end[q]) to their kprime-th key and
    // store it as the seed in tau0[q] and gtau[q]; nothing is emitted.  pass 1:
    // queries whose k'-th merged key lies above their seed tau0[q] are not
    // emitted; they get flags[q] = 1, gtau[q] = that k'-th key and count in
    // *nflag.  pass 2: only flagged queries are merged and emitted.
    int pass;
    const uint64_t *pre_end;
    uint64_t *tau0;
    uint32_t *flags;
    uint64_t *gtau;
    uint32_t *nflag;
};

// One SMALL (query, time bucket) slice of the type-2 sub-interval path: its
// query, its bucket's alpha, the offset from a bucket-code row to its
// timestamp-layout position, and its slot in the candidate pool.
struct SliceInfo {
    uint32_t q;
    float alpha;
    int64_t row_to_pos;
    uint32_t pool_slot;
    uint32_t pad;
};

hipError_t launch_absmax_rows(const float *rows, const uint32_t *perm, uint64_t n, uint32_t *out_bits, hipStream_t st);
hipError_t launch_sym_encode_slices(const float *queries, const SliceInfo *sl, uint64_t n, int8_t *out, hipStream_t st);
hipError_t launch_pool_keys(const uint32_t *ids, const uint32_t *dbits, const SliceInfo *sl, uint64_t n, uint32_t kb,
                            uint32_t stride, uint64_t *lists, uint32_t *counts, hipStream_t st);
hipError_t launch_ext_keys(const uint32_t *pos, const float *dist, const uint32_t *ext_slot, uint64_t nq, uint32_t n_ext,
                           uint64_t n_pos, uint32_t stride, uint64_t *lists, uint32_t *counts, hipStream_t st);

// Device router: per-query ranges [layout, s, e, bf] and (order != null) the
// batch order of ReadSortedQuery.  keys: 4 nq u64, vals: 2 nq u32, temp:
// route_temp_bytes(nq) bytes.
size_t route_temp_bytes(uint64_t nq);
hipError_t launch_route(const float *metas, uint64_t nq, const int32_t *lab_b, const float *ts_b, const float *ts_t,
                        uint64_t n, float nb, float bf_thr, float bf_thr3, uint32_t *ranges, uint32_t *order,
                        uint64_t *keys, uint32_t *vals, void *temp, size_t temp_bytes, hipStream_t st);

hipError_t launch_scan_f32(const ScanArgs &a, int B, hipStream_t st);
hipError_t launch_fill_u32(uint32_t *p, uint32_t v, uint64_t n, hipStream_t st);
hipError_t launch_merge(const MergeArgs &a, hipStream_t st);
// Row ids 0..n-1 stably sorted by keys_dev[0..n) (bits [0, end_bit), end_bit
// 32 or 64) into ids_out_dev; synchronises st.
hipError_t sort_rows_by_key(const uint64_t *keys_dev, uint64_t n, int end_bit, uint32_t *ids_out_dev, hipStream_t st);
// dst[i][:] = src[idx[i]][:] (kDim floats per row), i < n.
hipError_t launch_gather_rows(const float *src, const uint32_t *idx, uint64_t n, float *dst, hipStream_t st);
// out_ids[idx[i]][:] = ids[i][:], out_d[idx[i]][:] = d[i][:] (k per row), i < n.
hipError_t launch_scatter_results(const uint32_t *ids, const float *d, const uint32_t *idx, uint64_t n, uint32_t k,
                                  uint32_t *out_ids, float *out_d, hipStream_t st);
hipError_t launch_scan_sq8sym(const ScanSymArgs &a, int B, hipStream_t st);
// One query over one slice in a single launch (the per-call drop-in): each
// workgroup scans `chunk` rows of the slice into its own list; the last to
// finish (a device counter, zero between calls and reset by that workgroup)
// merges the lists and writes out[0..K) = slice positions (kPadId pad) and
// out[K..2K) = int distances (INT32_MAX pad), ascending (distance, position).
struct SymOneArgs {
    const int8_t *codes;   // the slice's first row (112-byte rows)
    uint32_t n;            // rows in the slice
    uint32_t chunk;        // rows per workgroup (a multiple of kRowTile)
    uint32_t K;
    int4 q[7];             // the query's 112 code bytes
    uint64_t *lists;       // [gridDim.x][B]
    uint32_t *counts;      // [gridDim.x]
    uint32_t *done;        // arrival counter
    uint32_t *out;         // [2K]
};
hipError_t launch_sym_one(const SymOneArgs &a, int B, hipStream_t st);
// Many independent one-slice searches in one launch (concurrent per-call
// searches combined): request r scans its slice with nblk workgroups, each
// into its own list (slots list0 .. list0 + nblk - 1), sharing their
// thresholds through gtau[r]; the last of them to finish merges the lists
// into out[r][2K].
struct SymReq {
    const int8_t *codes;   // the slice's first row (112-byte rows)
    uint32_t n;            // rows in the slice
    uint32_t chunk;        // rows per workgroup (a multiple of kRowTile)
    uint32_t nblk;         // workgroups: ceil(n / chunk)
    uint32_t list0;        // the request's first list slot
    int4 q[7];             // the query's 112 code bytes
};
struct SymMultiArgs {
    const SymReq *reqs;    // [nreq] (device memory)
    uint32_t nreq;
    uint32_t K;
    uint64_t *lists;       // [sum of nblk][B]
    uint32_t *counts;      // [sum of nblk]
    uint32_t *done;        // [nreq] arrival counters (zero between calls)
    uint64_t *gtau;        // [nreq] shared thresholds (all ones between calls)
    uint32_t *out;         // [nreq][2K]: positions in the slice, then int distances
};
// grid_x: the largest nblk of the requests.
hipError_t launch_sym_multi(const SymMultiArgs &a, uint32_t grid_x, int B, hipStream_t st);
// variant: 0 BF16, 1 FP8, 2 SQ8, 3 symmetric int8; B: list capacity (128,
// 256, 512, 1024; at most a.list_stride); nw: item width (1, 2, 4), an
// item carries at most 64 * nw queries.  launch_build_mfma with
// variant 3 reads the layout's symmetric codes from sq8_codes and writes the
// code norms to norms.
hipError_t launch_scan_mfma(int variant, const ScanMfmaArgs &a, int B, int nw, hipStream_t st);
// The seeded MFMA search's plan built on the device from ranges in device
// memory (fbf_dplan.hip).  All pointers are device buffers of the stated
// capacity; n_items[0] / [1] receive the item counts of the two passes,
// status = {items of the prefix pass, items of the main pass, overflow}.
constexpr uint32_t kGdStride = 12;   // u32 per group descriptor of the device planner
constexpr uint32_t kGncStride = 4;   // u32 per group's item counts (prefix halves, main)
struct DPlan {
    const uint32_t *ranges;  // [nq][4] {layout, s, e, -}
    uint64_t nq;
    uint64_t n_rows;         // rows of each layout
    uint32_t gmax, gpre, chunk, pdiv, min_prefix;  // gpre: members per prefix item (gmax <= 2 gpre)
    uint64_t max_items, max_slots;
    uint64_t *keys, *skeys;        // [nq] (ping-pong buffers of the query sort)
    uint32_t *vals, *order;        // [nq]
    uint32_t *hist;                // [256 * ceil(nq / 256)]
    uint64_t *hoff;                // [256 * ceil(nq / 256) + 1]
    uint64_t *part, *poff;         // [dplan_scan_parts], [dplan_scan_parts + 1]
    uint32_t *gd;                  // [ceil(nq / gmax) * gmax][kGdStride]
    uint32_t *gnc;                 // [ceil(nq / gmax) * gmax][kGncStride]
    uint32_t *qnc;                 // [nq][2]
    uint64_t *ioff_pre, *ioff_main;  // [ceil(nq / gmax) * gmax + 1]
    uint64_t *slot_off;            // [nq + 1]
    uint64_t *pre_end;             // [nq]
    WorkItem *items_tmp_pre, *items_tmp_main, *items_pre, *items_main;  // [max_items] each
    uint32_t *islot_pre, *islot_main;  // [max_items] each: an item's place in its row bin
    uint32_t *bins;                // [2][dplan_row_bins(n_rows)]: items per first row
    uint64_t *boff;                // [2][dplan_row_bins(n_rows) + 1]
    unsigned long long *bcost;     // [2][dplan_bins(n_rows, chunk)]: rows x queries per chunk bin
    ItemQuery *iq_pre, *iq_main;   // [nq] each (indexed by sorted position)
    uint64_t *n_items;             // [2]
    uint32_t *xoff_pre, *xoff_main;  // [9] each
    uint32_t *status;              // [4]: items per pass, overflow, list slots
};
// Bins of the item order (first rows: multiples of kRowTile) and of the XCD
// split (multiples of chunk), both layouts.
inline uint32_t dplan_row_bins(uint64_t n_rows) { return (uint32_t)(2 * (n_rows / kRowTile + 1)); }
inline uint32_t dplan_bins(uint64_t n_rows, uint32_t chunk) { return (uint32_t)(2 * (n_rows / chunk + 1)); }
// Tile sums the planner's scans hold at most.
uint64_t dplan_scan_parts(uint64_t nq, uint32_t gmax, uint64_t n_rows);
hipError_t launch_dplan(const DPlan &dp, hipStream_t st);
// B-operand fragments of a batch's queries for the MFMA scans, kQFragU4
// uint4 per query (variant 3 reads the queries' symmetric int8 codes).
constexpr int kQFragU4 = 16;
hipError_t launch_qfrag(int variant, const float *queries, const float *sq8_w, const int8_t *qcodes, uint64_t nq,
                        uint4 *out, hipStream_t st);
// Resolve (and cache) the persistent grid of a scan instantiation ahead of a
// stream capture.
void warm_scan_mfma(int variant, int B, int nw);
hipError_t launch_build_mfma(int variant, const float *rows, const uint32_t *perm, const float *norms,
                             const uint8_t *sq8_codes, uint64_t n, void *out, uint8_t *scale, hipStream_t st);
hipError_t launch_row_norms(const float *rows, uint64_t n, float *out, hipStream_t st);
// max over rows of |bf16(x) - x| / |x| (FP32, as float bits; rows with |x| = 0
// count 0) folded into *out_bits with atomicMax (*out_bits must start at 0).
hipError_t launch_bf16_err_max(const float *rows, uint64_t n, uint32_t *out_bits, hipStream_t st);
hipError_t launch_gather_f32(const float *src, const uint32_t *perm, uint64_t n, uint64_t n_pad, float *dst,
                             hipStream_t st);
hipError_t launch_absmax(const float *x, uint64_t count, uint32_t *out_bits, hipStream_t st);
hipError_t launch_sym_encode(const float *rows, const uint32_t *perm, uint64_t n, uint64_t n_out, float alpha,
                             int8_t *out, hipStream_t st);
hipError_t launch_sq8_train_encode(const float *rows, uint64_t n, float *mi, float *dif, uint8_t *codes,
                                   float *xhat_norm, hipStream_t st);
hipError_t launch_build_soa(const float *rows, const uint32_t *perm, uint64_t n,
                            float *soa, hipStream_t st);

}  // namespace fbf
