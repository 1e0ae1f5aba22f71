/*
 * fbf.h -- C-ABI of the MI355X filtered brute-force top-k library (libfbf.so).
 *
 * The library replaces the brute-force branch of the reference's query router
 * (hybrid_graph.cpp:200-228 and :399-419) and the exact filtered ground-truth
 * scan (utils.h:80-158).  Plain pointers and sizes only; every entry point
 * returns an fbf_status (0 = success).  All functions are safe to call from
 * many host threads at once on distinct streams; calls on one dataset that
 * share a stream are serialised by that stream.
 *
 * Data conventions (shared with the reference's file formats, utils.h:257-276
 * and :550-566): base rows are float32 [n][d] in original id order with a
 * float32 label (integral category id) and a float32 timestamp per row; a
 * query is float32 [d] plus a meta [type, v, l, r] with type 0 = no filter,
 * 1 = label == v, 2 = l <= timestamp <= r (inclusive), 3 = both.
 *
 * Result conventions: per query, k (id, squared-L2 distance) pairs in
 * ascending (distance, id) order; ties go to the smaller original id; when
 * fewer than k rows pass the filter the tail is padded with id FBF_PAD_ID and
 * distance +inf.  Exact modes compute each distance in the canonical FP32
 * order (t = x_i - q_i; acc = fma(t, t, acc) for i = 0..d-1).
 */
#ifndef FBF_H_
#define FBF_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FBF_PAD_ID 0xFFFFFFFFu
#define FBF_CODE_BYTES 112 /* reference SQ8 code stride (distance.hpp:271,305,332) */

typedef enum fbf_status {
    FBF_OK = 0,
    FBF_ERR_INVALID_ARG = 1,   /* null pointer, bad size, k out of range, bad filter type */
    FBF_ERR_UNSUPPORTED = 2,   /* dimension or mode not built into this library */
    FBF_ERR_UNREGISTERED = 3,  /* codes pointer not inside a registered code array */
    FBF_ERR_DEVICE = 4,        /* a HIP call failed; fbf_last_error() has the text */
    FBF_ERR_NO_REPR = 5        /* the dataset was created without the representation a mode needs */
} fbf_status;

/* Distance / selection modes of fbf_search. */
typedef enum fbf_mode {
    FBF_MODE_EXACT_F32 = 0, /* plain exact FP32 scan (canonical order), bit-exact ids */
    FBF_MODE_SQ8 = 1,       /* per-dim uint8 scan (SQ8Quantizer semantics), FP32 rerank of k' */
    FBF_MODE_BF16 = 2,      /* BF16 MFMA q.x + FP32 norms, FP32 rerank of k' */
    FBF_MODE_FP8 = 3,       /* FP8 e4m3 (per-vector scale) MFMA q.x + FP32 norms, FP32 rerank */
    FBF_MODE_SQ8SYM = 4,    /* symmetric int8 scan (router's bruteforce()), FP32 rerank of k' */
    FBF_MODE_EXACT_BF16 = 5 /* the FBF_MODE_EXACT_F32 results (same ids and distances, Bruteforce
                               utils.h:80-158) from the BF16 MFMA scan and an FP32 rerank of k'
                               (default 720; work items of 256 queries over 65536 rows) candidates: a
                               query is certified when the k'-th BF16
                               score, less a bound on its rounding error, exceeds the k-th exact
                               distance; the others are searched again by the FP32 scan.  Needs
                               FBF_REPR_BF16 | FBF_REPR_F32; waits for the device (no capture). */
} fbf_mode;

/* Representations built by fbf_dataset_create (bit flags). */
#define FBF_REPR_F32 0x1u     /* FP32 scan copies in both sorted layouts (exact mode) */
#define FBF_REPR_SQ8 0x2u     /* per-dimension uint8 codes (mode SQ8) */
#define FBF_REPR_BF16 0x4u    /* BF16 MFMA tiles (mode BF16) */
#define FBF_REPR_FP8 0x8u     /* FP8 e4m3 MFMA tiles + per-row scale (mode FP8) */
#define FBF_REPR_SQ8SYM 0x10u /* symmetric int8 codes, global absmax alpha (mode SQ8SYM) */
#define FBF_REPR_TBUCKETS 0x20u /* the ten 0.1-wide time buckets of the timestamp layout, each with
                                   its own symmetric int8 codes and alpha (fbf_search_subinterval) */
#define FBF_REPR_ALL 0x3Fu

typedef struct fbf_dataset fbf_dataset;

/* Search tuning.  Zero fields take the library defaults.  kprime is the
 * over-fetch depth of the approximate modes (ignored by FBF_MODE_EXACT_F32);
 * chunk_rows is the rows per work item (default: from the batch's total span, 16 Ki to 256 Ki
 * rows, about 32 Ki work items per batch); group_queries the
 * queries sharing one streamed tile (<= 16 for the VALU scans, default 16;
 * <= 256 for the MFMA scans, default 256). */
typedef struct fbf_search_opts {
    uint32_t mode;        /* fbf_mode */
    uint32_t kprime;      /* 0 -> 3k/2 (BF16, SQ8), 4k (SQ8SYM), 9k (FP8) or 720 (EXACT_BF16), at most 896 */
    uint32_t chunk_rows;  /* 0 -> default */
    uint32_t group_queries;
    uint32_t persistent;  /* scan launch form: 0 -> default (MFMA modes: persistent; others: one
                             workgroup per work item), 1 -> persistent (a device-filling grid
                             that pulls work items from queues; the MFMA scans keep one
                             chunk-major run of items per XCD), 2 -> one workgroup per item */
    uint32_t prepass_div; /* MFMA modes: a prefix pass over the first 1/prepass_div of each large
                             range seeds the per-query thresholds (0 -> default 16; 16 for EXACT_BF16; 1 -> off) */
    uint32_t prepass_margin; /* percent of k'/prepass_div keys the prefix pass keeps; its last
                                kept key is the seed (0 -> 300 for k' <= 160, 200 up to 500, else 150; 200 for EXACT_BF16).  Queries whose seed proves
                                too tight are rescanned, so results never depend on it. */
    uint32_t max_batch;   /* fbf_search / fbf_search_host: larger batches run as consecutive
                             sub-batches of this many queries (0 -> 131072), which bounds the
                             candidate-list memory (about 0.25 MB per query at the BF16
                             defaults on 10 M rows); results do not depend on it.  Captured
                             searches (fbf_search_graph_create) are not split. */
    uint32_t metas_on_device; /* fbf_search: 1 -> metas is DEVICE memory [nq][4]; the ranges
                                 come from the device router (fbf_route) and the host plans
                                 from them.  Results equal those of host metas.
                                 fbf_search_graph_create: 1 -> an any-batch graph that reads
                                 metas at every replay and plans on the device. */
} fbf_search_opts;

/* Create a device-resident dataset on the current HIP device from host
 * arrays: vecs [n][d] float32 (original id order), labels[n], timestamps[n].
 * Sorts the rows by (label, timestamp, id) and (timestamp, id) -- the layouts
 * of SortDataset (hybrid_graph.cpp:505-524) -- and builds the representations
 * named in repr_flags.  d must be 100 in this build. */
fbf_status fbf_dataset_create(const float *vecs, const float *labels, const float *timestamps,
                              uint64_t n, uint32_t d, uint32_t repr_flags, fbf_dataset **out);

/* As fbf_dataset_create with vecs already in device memory (labels and
 * timestamps stay on the host, they drive the planner). */
fbf_status fbf_dataset_create_device(const float *vecs_dev, const float *labels,
                                     const float *timestamps, uint64_t n, uint32_t d,
                                     uint32_t repr_flags, fbf_dataset **out);

fbf_status fbf_dataset_destroy(fbf_dataset *ds);

/* Number of rows a query's filter admits (the router's e - s,
 * hybrid_graph.cpp:176-199) and its selectivity (e - s) / n. */
fbf_status fbf_filter_count(const fbf_dataset *ds, const float *meta, uint64_t *count,
                            float *selectivity);

/* Batched filtered top-k.  queries_dev: device float32 [nq][d]; metas: HOST
 * float32 [nq][4]; out_ids_dev / out_dists_dev: device [nq][k].  stream is a
 * hipStream_t (NULL = default stream).  k must be in [1, 512]. */
fbf_status fbf_search(fbf_dataset *ds, const float *queries_dev, const float *metas,
                      uint64_t nq, uint32_t k, const fbf_search_opts *opts,
                      uint32_t *out_ids_dev, float *out_dists_dev, void *stream);

/* The router's range step on the device (hybrid_graph.cpp:168-201 and the
 * query order of ReadSortedQuery, utils.h:514-547) for a batch whose metas
 * are in device memory [nq][4]: ranges_dev[q] = {layout, s, e, bf} -- the
 * query's contiguous range [s, e) of sorted layout `layout` (0 = (label,
 * timestamp), 1 = timestamp; type 0: all of layout 1), as fbf_filter_count
 * gives it, and bf = 1 when the router sends it to brute force: selectivity
 * (e - s) / n below bf_thr, or type 3 and below bf_thr3 (0 -> 0.045 and 0.08,
 * baseline.cpp:77, hybrid_graph.h:79; type 0 has selectivity 1).  An unknown
 * filter type gives {0xFFFFFFFF, 0, 0, 0}.  order_dev (may be NULL) receives
 * the query indices in ascending (type, v, l, r, index) order -- the
 * router's batch order (sorted_ids).  Device arrays; waits for its work. */
fbf_status fbf_route(fbf_dataset *ds, const float *metas_dev, uint64_t nq, float bf_thr, float bf_thr3,
                     uint32_t *ranges_dev, uint32_t *order_dev, void *stream);

/* Host-memory convenience form of fbf_search (copies in and out, waits). */
fbf_status fbf_search_host(fbf_dataset *ds, const float *queries, const float *metas,
                           uint64_t nq, uint32_t k, const fbf_search_opts *opts,
                           uint32_t *out_ids, float *out_dists);

/* Captured batch (no counterpart in the reference; it replaces the router's
 * per-batch brute-force loop, hybrid_graph.cpp:168-228, for a batch whose
 * filters repeat).  fbf_search_graph_create plans the batch exactly as
 * fbf_search does, uploads the plan into scratch owned by the graph, and
 * captures the whole device sequence (threshold fills, prefix / main /
 * rescan scans with their persistent work queues, merges and rerank) into
 * one HIP graph.  fbf_search_graph_launch replays it on any stream: it reads
 * queries_dev and writes out_ids_dev / out_dists_dev given at creation, so a
 * caller may refill the query vectors between replays (the filters, metas,
 * are fixed by the plan).  Replays on one graph must not overlap.  Searches
 * on the dataset may run meanwhile (the graph has its own scratch).  The
 * rescan count of fbf_last_search_rescans is not recorded for replays.
 *
 * Any-batch form (opts->metas_on_device = 1, modes BF16, FP8, SQ8, SQ8SYM):
 * metas is DEVICE memory [nq][4] and is read at every replay, like the
 * queries: the graph holds the device router (fbf_route's range step), a
 * planner that groups the queries and cuts their ranges into work items on
 * the device, and the seeded search, so each replay serves a new batch of nq
 * filtered queries with no host work.  Results equal fbf_search's on the same
 * queries and metas (the plans differ; the kept keys do not).  Unknown filter
 * types and empty ranges give padded outputs (id 0xFFFFFFFF, distance
 * +inf).  Buffers are sized at creation for the worst plan of nq queries,
 * the candidate lists at most 32 GB or 3/4 of the device memory free at
 * creation, whichever is more; a replay whose plan needs more list
 * slots than that returns padded outputs for every query and sets the
 * overflow flag of fbf_search_graph_plan_info (re-run such a batch with
 * fbf_search or with larger chunk_rows). */
typedef struct fbf_search_graph fbf_search_graph;
fbf_status fbf_search_graph_create(fbf_dataset *ds, const float *queries_dev, const float *metas,
                                   uint64_t nq, uint32_t k, const fbf_search_opts *opts,
                                   uint32_t *out_ids_dev, float *out_dists_dev, fbf_search_graph **out);
fbf_status fbf_search_graph_launch(fbf_search_graph *g, void *stream);
/* Waits for the device, then frees the graph and its scratch. */
fbf_status fbf_search_graph_destroy(fbf_search_graph *g);
/* Any-batch graphs: the plan of the last completed replay (the caller waits
 * for it first): work items of the prefix and main passes and the overflow
 * flag (1: the plan exceeded the graph's capacity, outputs padded).  No
 * counterpart in the reference (its plan is the OpenMP loop over queries). */
fbf_status fbf_search_graph_plan_info(fbf_search_graph *g, uint64_t *items_prefix, uint64_t *items_main,
                                      uint32_t *overflow);

/* Type-2 sub-interval search: the SMALL-bucket branch of the router's
 * type-2 path (hybrid_graph.cpp:317-420) and the global refine (:463-494) for
 * a batch of type-2 queries (metas[q][0] == 2; other types are rejected).
 * Needs FBF_REPR_TBUCKETS: the time buckets of ReadSortedBaseTimestamp
 * (utils.h:470-478) over the timestamp layout, bucket t = [lower_bound(0.1 t),
 * upper_bound(0.1 (t + 1))), each with the symmetric int8 quantiser trained on
 * its own rows (the timestamp searchers, baseline.cpp:130-141).
 *
 * Per query: SplitInterval (hybrid_graph.cpp:91-134) classes each bucket as
 * FULL (3), MEDIUM (1) or SMALL (0) -- 255 when it does not overlap [l, r] --
 * with the SMALL threshold of the router (0.5 when the query's selectivity is
 * below 0.2, else 0.2; opts->small_thr > 0 overrides); for each SMALL bucket
 * the bucket_k smallest symmetric int8 distances of its slice, the query
 * encoded with that bucket's alpha (bruteforce_subgraph), ties to the smaller
 * position; their distances int(d * alpha * alpha) enter the candidate pool,
 * which keeps the pool_k smallest (pool distance, timestamp-layout position)
 * (candidate_pool, maxPQIFCS<int>); the pool is reranked with the canonical
 * FP32 distance and the k smallest (distance, original id) are written
 * (padded with FBF_PAD_ID / +inf).  FULL and MEDIUM buckets are the reference's
 * graph searches: the caller may add their candidates as n_ext (position in
 * the timestamp layout, distance) pairs per query (device arrays [nq][n_ext];
 * pad position 0xFFFFFFFF), which enter the pool with the same conversion.
 * bucket_types (host, nq x 10, may be NULL) receives the classes.  queries_dev,
 * out_ids_dev, out_dists_dev: device; metas: host; asynchronous on stream. */
typedef struct fbf_subinterval_opts {
    float small_thr;     /* 0 -> the router's rule above */
    uint32_t bucket_k;   /* 0 -> 140 (bf_refine_k, hybrid_graph.h:80), at most 896 */
    uint32_t pool_k;     /* 0 -> 150 (global_graph_refine_k, hybrid_graph.h:84), at most 896 */
    uint32_t reserved;
} fbf_subinterval_opts;
fbf_status fbf_search_subinterval(fbf_dataset *ds, const float *queries_dev, const float *metas, uint64_t nq,
                                  uint32_t k, const fbf_subinterval_opts *opts, const uint32_t *ext_pos_dev,
                                  const float *ext_dist_dev, uint32_t n_ext, uint32_t *out_ids_dev,
                                  float *out_dists_dev, uint8_t *bucket_types, void *stream);

/* Quantiser parameters of the dataset: alpha of the symmetric int8 codes
 * (SQ8SymmetricQuantizer::alpha, sq8_sym_quant.hpp:37-44) and the per-dim
 * min / (max - min) of the uint8 codes (SQ8Quantizer::mi/dif,
 * sq8_quant.hpp:30-47).  mi and dif receive d floats each (may be NULL). */
fbf_status fbf_dataset_quant_params(const fbf_dataset *ds, float *alpha, float *mi, float *dif);

/* Rows [row0, row0 + nrows) of one sorted layout (0 = (label, timestamp, id)
 * order, 1 = (timestamp, id) order) as the dataset holds them: the original id
 * at each position (ids, may be NULL) and, when codes is not NULL, the codes
 * of representation repr -- FBF_REPR_SQ8SYM: 112-byte symmetric int8 codes
 * (the router's codes_both / codes_time, hybrid_graph.cpp:515-522, encoded as
 * SQ8SymmetricQuantizer::encode, sq8_sym_quant.hpp:63-74); FBF_REPR_SQ8: 100
 * per-dimension uint8 codes per row (SQ8Quantizer::encode, sq8_quant.hpp:51-63).
 * Host arrays; waits for the device. */
fbf_status fbf_dataset_layout_rows(const fbf_dataset *ds, uint32_t layout, uint64_t row0, uint64_t nrows,
                                   uint32_t repr, uint32_t *ids, void *codes);

/* ---------------------------------------------------------------------
 * Reference-shaped entry points (bruteforce.h:244-249 and :267-273).
 *
 * Register a host array of 112-byte symmetric int8 codes once; the library
 * uploads it and later resolves any codes_in pointer inside [host_base,
 * host_base + bytes) to the device copy.  Unregistered pointers are uploaded
 * per call (correct, slow).
 * --------------------------------------------------------------------- */
fbf_status fbf_register_codes(const void *host_base, uint64_t bytes);
fbf_status fbf_unregister_codes(const void *host_base);

/* bruteforce (bruteforce.h:244-265): SQ8-symmetric scan of npoints codes,
 * the k smallest int distances.  closest_points receives positions relative
 * to codes_in in ascending (distance, position) order -- the same SET as the
 * reference's heap whenever the k-th distance is not tied.  dim, alpha and
 * nqueries are accepted for signature parity (the reference ignores them).
 * k > npoints pads with -1.  Calls on registered arrays from several threads
 * at once are combined: a waiting caller runs every queued call of its GPU as
 * one launch (each call still gets exactly its own results), so the router's
 * OpenMP loop keeps the GPU busy without changing its call pattern. */
fbf_status fbf_bruteforce(size_t dim, size_t k, int *closest_points, size_t npoints,
                          const char *codes_in, float alpha, size_t nqueries,
                          const int8_t *queries_in);

/* bruteforce_subgraph (bruteforce.h:267-289): as fbf_bruteforce, plus
 * distances[j] = int_distance * alpha * alpha. */
fbf_status fbf_bruteforce_subgraph(size_t dim, size_t k, int *closest_points, float *distances,
                                   size_t npoints, const char *codes_in, float alpha,
                                   size_t nqueries, const int8_t *queries_in);

/* Batched form of the router's brute-force branch over one registered code
 * array: per query i the slice [starts[i], starts[i] + counts[i]) of 112-byte
 * codes, query code qcodes[i*112 .. +112], top k positions (relative to the
 * slice start) and int distances, ascending.  All arrays are host memory. */
fbf_status fbf_bruteforce_batch(const char *codes_base, const uint64_t *starts,
                                const uint64_t *counts, const int8_t *qcodes, uint64_t nq,
                                uint32_t k, int32_t *out_pos, int32_t *out_dist);

/* No counterpart in the reference.  fbf_bruteforce_batch for callers that
 * cannot allocate device memory (serving loops and harnesses that forbid
 * hipMalloc): the same arguments and results, plus the stream to run on
 * (NULL = default stream), with every scratch buffer -- plan, candidate
 * lists, query codes, results -- in a fixed device array of
 * FBF_NOALLOC_SCRATCH_BYTES per GPU that the library holds until the process
 * ends.  The call allocates nothing on the device (codes_base must lie in a
 * registered code array, registered beforehand); a batch that needs more
 * scratch than the array holds is rejected with FBF_ERR_INVALID_ARG (about
 * 2 KiB per (query, 64 Ki-row chunk) at k <= 160).  Every call on a GPU
 * shares that GPU's array: the library runs one such call at a time per
 * process, and each waits for its own results. */
#define FBF_NOALLOC_SCRATCH_BYTES (64ull << 20)
fbf_status fbf_bruteforce_batch_noalloc(const char *codes_base, const uint64_t *starts,
                                        const uint64_t *counts, const int8_t *qcodes, uint64_t nq,
                                        uint32_t k, int32_t *out_pos, int32_t *out_dist, void *stream);

/* ---------------------------------------------------------------------
 * Measurement
 * --------------------------------------------------------------------- */

/* When enabled, every kernel launch of this dataset is bracketed by HIP
 * events on its stream; fbf_profile_collect waits for them and returns the
 * launch count and summed milliseconds of the kernel named `kernel`
 * ("scan", "merge", "rerank", "build", or "all"; "scan_main", "scan_prefix"
 * and "scan_rescan" select the MFMA scans' passes: main (or only) pass,
 * seeded prefix pass, masked rescan), then clears the entries it returns.
 * "plan" returns the host time fbf_search spent planning the work items and
 * queueing their upload (one entry per search). */
fbf_status fbf_profile_enable(fbf_dataset *ds, int enable);
fbf_status fbf_profile_collect(fbf_dataset *ds, const char *kernel, uint64_t *launches,
                               double *total_ms);

/* Work counters of the last fbf_search on this dataset: P = sum over queries
 * of the filtered candidate count, the rows actually streamed by the scan
 * (including tile padding), and the number of work items. */
fbf_status fbf_last_search_stats(const fbf_dataset *ds, uint64_t *pairs, uint64_t *pairs_computed,
                                 uint64_t *work_items);

/* Algorithmic (query, candidate) pairs of the last fbf_search's seeded
 * prefix pass and of its main pass (prefix 0 when it ran none); they sum to
 * fbf_last_search_stats' pairs.  No reference counterpart: measurement. */
fbf_status fbf_last_search_pass_pairs(const fbf_dataset *ds, uint64_t *prefix_pairs, uint64_t *main_pairs);

/* Number of queries of the last fbf_search on this dataset whose seeded
 * threshold failed verification and were rescanned over their whole range
 * (0 when the search ran no prefix pass).  Waits for that search's verify
 * step to finish. */
fbf_status fbf_last_search_rescans(fbf_dataset *ds, uint64_t *queries);

/* Number of queries of the last fbf_search on this dataset that
 * FBF_MODE_EXACT_BF16 could not certify and searched again with the FP32
 * scan (0 after searches in other modes).  No reference counterpart:
 * measurement of the certified mode. */
fbf_status fbf_last_search_fallbacks(const fbf_dataset *ds, uint64_t *queries);

/* Text of the last error on the calling thread ("" when none). */
const char *fbf_last_error(void);

/* Library build string (kernel set, target). */
const char *fbf_version(void);

#ifdef __cplusplus
}
#endif

#endif /* FBF_H_ */
