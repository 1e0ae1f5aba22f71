/*
 * fbf_oracle.c -- CPU restatement of the reference's filtered brute-force path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: tests/, the smoke entry
 * point and the bench's cpu_baseline leg may load it; the product library never
 * links it, calls it or falls back to it.
 *
 * Every function cites the reference file:line whose behaviour it restates.
 * Arithmetic conventions (decided here, mirrored bit-for-bit by the HIP kernels):
 *   - canonical FP32 squared L2: for i = 0..d-1 in index order,
 *         t = x_i - q_i            (one rounding)
 *         acc = fmaf(t, t, acc)    (one rounding)
 *     compiled with -ffp-contract=off so no other fusion happens.
 *   - top-k key: (dist, id) ordered lexicographically, ties to the smaller id;
 *     keys are packed as ((uint64)f2ord(dist) << 32) | id.
 *   - fewer than k candidates: pad with id 0xFFFFFFFF and dist +inf.
 *
 * Pinning (what checks each function, see tests/test_oracle.py):
 *   orc_l2_canon        FP64 closed form within 1e-5 rel; dims-permutation
 *                       invariance is NOT expected (order is fixed); d=1 exact.
 *   orc_l2_sq8sym       KAT test_simd_distance.cpp:199-203 -> 210929; closed form.
 *   orc_sq8sym_*        absmax encodes to +/-127; half-away rounding cases.
 *   orc_sq8_*           x=mi -> 0, x=mx -> 255, truncation cases, dif=0 -> 0.
 *   orc_topk_heap       == orc_topk_fullsort (independent algorithm) on random
 *                       tiny inputs with forced ties; == numpy lexsort.
 *   orc_query_range     == predicate scan (utils.h:110-138) on every query.
 *   orc_bf_heap_ref     restates maxPQIFCS exactly; its id SET equals the sorted
 *                       variant whenever the k-th distance is not tied.
 *   orc_recall          closed form of utils.h:233-246 on hand-made cases.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_PAD_ID 0xFFFFFFFFu
#define ORC_CODE_DIM 112 /* reference code stride: distance.hpp:271,305,332 */

/* ------------------------------------------------------------------------ */
/* keys                                                                     */
/* ------------------------------------------------------------------------ */

/* Monotone map float -> uint32: a < b (as floats, no NaN) <=> f2ord(a) < f2ord(b). */
uint32_t orc_f2ord(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

float orc_ord2f(uint32_t o) {
    uint32_t u = (o & 0x80000000u) ? (o & 0x7FFFFFFFu) : ~o;
    float f;
    memcpy(&f, &u, 4);
    return f;
}

static inline uint64_t mk_key(float d, uint32_t id) {
    return ((uint64_t)orc_f2ord(d) << 32) | id;
}

/* ------------------------------------------------------------------------ */
/* distances                                                                */
/* ------------------------------------------------------------------------ */

/* Canonical FP32 squared L2.  Follows normal_l2 (distance.h:62-71) and
 * L2SqrRef (pyglass/glass/simd/distance.hpp:165-173) in meaning; the summation
 * order and the single fused multiply-add per term are this project's fixed
 * convention, because the reference's own bits depend on its compiler. */
float orc_l2_canon(const float *x, const float *q, int d) {
    float acc = 0.0f;
    for (int i = 0; i < d; ++i) {
        float t = x[i] - q[i];
        acc = fmaf(t, t, acc);
    }
    return acc;
}

/* normal_l2 as written (distance.h:62-71): v = a-b; v *= v; r += v, in source
 * order with separate roundings (built with -ffp-contract=off). */
float orc_l2_normal(const float *a, const float *b, int d) {
    float r = 0.0f;
    for (int i = 0; i < d; ++i) {
        float v = a[i] - b[i];
        v *= v;
        r += v;
    }
    return r;
}

double orc_l2_f64(const float *x, const float *q, int d) {
    double acc = 0.0;
    for (int i = 0; i < d; ++i) {
        double t = (double)x[i] - (double)q[i];
        acc += t * t;
    }
    return acc;
}

/* L2SqrSQ8_sym (pyglass/glass/simd/distance.hpp:329-366): exact int32
 * sum over the fixed 112-byte code of (x_i - y_i)^2; the d argument of the
 * reference is ignored there (:332), so it is absent here. */
int32_t orc_l2_sq8sym(const int8_t *x, const int8_t *y) {
    int32_t s = 0;
    for (int i = 0; i < ORC_CODE_DIM; ++i) {
        int32_t t = (int32_t)x[i] - (int32_t)y[i];
        s += t * t;
    }
    return s;
}

/* L2SqrSQ8_ext (pyglass/glass/simd/distance.hpp:268-301) evaluated in FP64:
 *   sum_i (255 x_i - ((c_i + 0.5) dif_i + 255 mi_i))^2  = 255^2 |x - xhat|^2 */
double orc_l2_sq8ext_f64(const float *x, const uint8_t *c, const float *mi,
                         const float *dif, int d) {
    double s = 0.0;
    for (int i = 0; i < d; ++i) {
        double y = ((double)c[i] + 0.5) * (double)dif[i] + 255.0 * (double)mi[i];
        double t = 255.0 * (double)x[i] - y;
        s += t * t;
    }
    return s;
}

/* ------------------------------------------------------------------------ */
/* quantisers                                                               */
/* ------------------------------------------------------------------------ */

/* SQ8SymmetricQuantizer::train alpha (sq8_sym_quant.hpp:37-43): max |x| over
 * all n*d values, sequential. */
float orc_sq8sym_train(const float *x, uint64_t n, int d) {
    float a = 0.0f;
    for (uint64_t i = 0; i < n * (uint64_t)d; ++i) {
        float v = fabsf(x[i]);
        if (v > a) a = v;
    }
    return a;
}

/* SQ8SymmetricQuantizer::encode (sq8_sym_quant.hpp:63-74):
 *   code = round(clamp(x/alpha, -1, 1) * 127), round = half away from zero;
 * codes are 112 bytes with zero padding (alloc2M memsets, memory.hpp:30-36). */
void orc_sq8sym_encode(const float *x, uint64_t n, int d, float alpha, int8_t *codes) {
    for (uint64_t r = 0; r < n; ++r) {
        int8_t *o = codes + r * ORC_CODE_DIM;
        memset(o, 0, ORC_CODE_DIM);
        for (int i = 0; i < d; ++i) {
            float v = x[r * d + i] / alpha;
            if (v > 1.0f) v = 1.0f;
            if (v < -1.0f) v = -1.0f;
            o[i] = (int8_t)roundf(v * 127.0f);
        }
    }
}

/* SQ8Quantizer::train (sq8_quant.hpp:30-47): per-dimension min / max. */
void orc_sq8_train(const float *x, uint64_t n, int d, float *mi, float *dif) {
    for (int j = 0; j < d; ++j) {
        mi[j] = HUGE_VALF;
        dif[j] = -HUGE_VALF; /* holds mx during the scan */
    }
    for (uint64_t i = 0; i < n; ++i)
        for (int j = 0; j < d; ++j) {
            float v = x[i * d + j];
            if (v > dif[j]) dif[j] = v;
            if (v < mi[j]) mi[j] = v;
        }
    for (int j = 0; j < d; ++j) dif[j] = dif[j] - mi[j];
}

/* SQ8Quantizer::encode (sq8_quant.hpp:51-63): code = (uint8)(clamp((x-mi)/dif,
 * 0, 1) * 255), truncating.  The reference divides by zero when dif == 0; this
 * restatement defines code 0 for a constant dimension.  Codes are d bytes. */
void orc_sq8_encode(const float *x, uint64_t n, int d, const float *mi,
                    const float *dif, uint8_t *codes) {
    for (uint64_t r = 0; r < n; ++r)
        for (int j = 0; j < d; ++j) {
            uint8_t c = 0;
            if (dif[j] > 0.0f) {
                float v = (x[r * d + j] - mi[j]) / dif[j];
                if (v < 0.0f) v = 0.0f;
                if (v > 1.0f) v = 1.0f;
                c = (uint8_t)(v * 255.0f);
            }
            codes[r * (uint64_t)d + j] = c;
        }
}

/* ------------------------------------------------------------------------ */
/* bounded top-k (ResultMaxHeap, utils.h:34-78, on (dist,id) keys)           */
/* ------------------------------------------------------------------------ */

typedef struct {
    uint64_t *h;
    int n, k;
} kheap;

static void kh_push(kheap *H, uint64_t key) {
    if (H->n < H->k) {
        int i = H->n++;
        H->h[i] = key;
        while (i > 0) {
            int p = (i - 1) >> 1;
            if (H->h[p] >= H->h[i]) break;
            uint64_t t = H->h[p]; H->h[p] = H->h[i]; H->h[i] = t;
            i = p;
        }
        return;
    }
    if (H->k == 0 || key >= H->h[0]) return;
    H->h[0] = key;
    int i = 0;
    for (;;) {
        int l = 2 * i + 1, r = l + 1, m = i;
        if (l < H->n && H->h[l] > H->h[m]) m = l;
        if (r < H->n && H->h[r] > H->h[m]) m = r;
        if (m == i) break;
        uint64_t t = H->h[m]; H->h[m] = H->h[i]; H->h[i] = t;
        i = m;
    }
}

static int cmp_u64(const void *a, const void *b) {
    uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
    return (x > y) - (x < y);
}

/* Heap contents -> ascending (dist, id), padded. */
static void kh_emit(kheap *H, int k, uint32_t *ids, float *dists) {
    qsort(H->h, H->n, 8, cmp_u64);
    for (int j = 0; j < k; ++j) {
        if (j < H->n) {
            ids[j] = (uint32_t)(H->h[j] & 0xFFFFFFFFu);
            dists[j] = orc_ord2f((uint32_t)(H->h[j] >> 32));
        } else {
            ids[j] = ORC_PAD_ID;
            dists[j] = HUGE_VALF;
        }
    }
}

/* Generic top-k over an explicit key list: heap algorithm. */
void orc_topk_heap(const uint64_t *keys, uint64_t n, int k, uint64_t *out, int *nout) {
    kheap H;
    H.h = (uint64_t *)malloc(sizeof(uint64_t) * (k > 0 ? k : 1));
    H.n = 0; H.k = k;
    for (uint64_t i = 0; i < n; ++i) kh_push(&H, keys[i]);
    qsort(H.h, H.n, 8, cmp_u64);
    memcpy(out, H.h, 8 * (size_t)H.n);
    *nout = H.n;
    free(H.h);
}

/* Independent algorithm for the same selection: full sort, take the first k. */
void orc_topk_fullsort(const uint64_t *keys, uint64_t n, int k, uint64_t *out, int *nout) {
    uint64_t *t = (uint64_t *)malloc(8 * (n ? n : 1));
    memcpy(t, keys, 8 * n);
    qsort(t, n, 8, cmp_u64);
    int m = (uint64_t)k < n ? k : (int)n;
    memcpy(out, t, 8 * (size_t)m);
    *nout = m;
    free(t);
}

/* ------------------------------------------------------------------------ */
/* filters                                                                  */
/* ------------------------------------------------------------------------ */

/* Predicate of Bruteforce (utils.h:98-138): type 0 all rows; 1 int(label) ==
 * int(v); 2 l <= ts <= r (inclusive); 3 both.  Labels are float in the file
 * and compared after conversion to int32 (utils.h:99,109). */
int orc_match(int type, int32_t v, float l, float r, int32_t label, float ts) {
    switch (type) {
    case 0: return 1;
    case 1: return label == v;
    case 2: return ts >= l && ts <= r;
    case 3: return label == v && ts >= l && ts <= r;
    default: return 0;
    }
}

/* ------------------------------------------------------------------------ */
/* exact filtered brute force                                               */
/* ------------------------------------------------------------------------ */

/* Bruteforce (utils.h:80-158): for each query scan every base row in id
 * order, apply the predicate, keep the k smallest (dist, id).  Output is
 * ascending, padded when fewer than k rows match (the reference throws there,
 * utils.h:142-146 -> :44-47).  mode 0: canonical FP32, 1: FP64 (check time
 * only), 2: normal_l2 strict order.  labels are int32, metas are nq x 4
 * [type, v, l, r] as in the query file (utils.h:550-566). */
void orc_exact_predicate(const float *base, const int32_t *labels, const float *ts,
                         uint64_t n, int d, const float *queries, const float *metas,
                         uint64_t nq, int k, int mode, uint32_t *out_ids,
                         double *out_dists) {
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t qi = 0; qi < (int64_t)nq; ++qi) {
        const float *q = queries + qi * d;
        int type = (int)metas[qi * 4 + 0];
        int32_t v = (int32_t)metas[qi * 4 + 1];
        float l = metas[qi * 4 + 2], r = metas[qi * 4 + 3];
        /* FP64 mode needs double keys: keep (double, id) pairs in a small
         * sorted array instead of the u64 heap. */
        if (mode == 1) {
            double *bd = (double *)malloc(sizeof(double) * (k + 1));
            uint32_t *bi = (uint32_t *)malloc(sizeof(uint32_t) * (k + 1));
            int cnt = 0;
            for (uint64_t j = 0; j < n; ++j) {
                if (!orc_match(type, v, l, r, labels[j], ts[j])) continue;
                double dist = orc_l2_f64(base + j * d, q, d);
                if (cnt == k && !(dist < bd[k - 1])) continue; /* ties: smaller id kept */
                int p = cnt < k ? cnt : k - 1;
                while (p > 0 && bd[p - 1] > dist) { bd[p] = bd[p - 1]; bi[p] = bi[p - 1]; --p; }
                bd[p] = dist; bi[p] = (uint32_t)j;
                if (cnt < k) ++cnt;
            }
            for (int t = 0; t < k; ++t) {
                out_ids[qi * k + t] = t < cnt ? bi[t] : ORC_PAD_ID;
                out_dists[qi * k + t] = t < cnt ? bd[t] : HUGE_VAL;
            }
            free(bd); free(bi);
            continue;
        }
        kheap H;
        H.h = (uint64_t *)malloc(8 * (k > 0 ? k : 1));
        H.n = 0; H.k = k;
        for (uint64_t j = 0; j < n; ++j) {
            if (!orc_match(type, v, l, r, labels[j], ts[j])) continue;
            float dist = mode == 0 ? orc_l2_canon(base + j * d, q, d)
                                   : orc_l2_normal(base + j * d, q, d);
            kh_push(&H, mk_key(dist, (uint32_t)j));
        }
        float *fd = (float *)malloc(sizeof(float) * (k > 0 ? k : 1));
        kh_emit(&H, k, out_ids + qi * k, fd);
        for (int t = 0; t < k; ++t) out_dists[qi * k + t] = fd[t];
        free(fd); free(H.h);
    }
}

/* ------------------------------------------------------------------------ */
/* sorted layouts and range lookup                                          */
/* ------------------------------------------------------------------------ */

static const int32_t *g_lab;
static const float *g_ts;
static int cmp_both(const void *a, const void *b) {
    uint32_t x = *(const uint32_t *)a, y = *(const uint32_t *)b;
    if (g_lab[x] != g_lab[y]) return g_lab[x] < g_lab[y] ? -1 : 1;
    if (g_ts[x] != g_ts[y]) return g_ts[x] < g_ts[y] ? -1 : 1;
    return (x > y) - (x < y);
}
static int cmp_time(const void *a, const void *b) {
    uint32_t x = *(const uint32_t *)a, y = *(const uint32_t *)b;
    if (g_ts[x] != g_ts[y]) return g_ts[x] < g_ts[y] ? -1 : 1;
    return (x > y) - (x < y);
}

/* SortDataset (hybrid_graph.cpp:505-524) with comparators cmp_by_both /
 * cmp_by_time (hybrid_graph.cpp:14-27).  The reference sorts with the unstable
 * std::sort; this restatement breaks (label, ts) and ts ties by original id so
 * the scan order is fully defined.  Not thread-safe (qsort context). */
void orc_build_layouts(const int32_t *labels, const float *ts, uint64_t n,
                       uint32_t *perm_both, uint32_t *perm_time) {
    for (uint64_t i = 0; i < n; ++i) perm_both[i] = perm_time[i] = (uint32_t)i;
    g_lab = labels; g_ts = ts;
    qsort(perm_both, n, 4, cmp_both);
    qsort(perm_time, n, 4, cmp_time);
}

/* Range lookup (hybrid_graph.cpp:176-199): lower/upper bound over the sorted
 * metadata.  layout 0 = (label, ts) order, 1 = ts order.  Type 0 uses the ts
 * layout over [0, n).  lab_b/ts_b are the labels/timestamps in "both" order,
 * ts_t the timestamps in "time" order. */
void orc_query_range(int type, int32_t v, float l, float r, const int32_t *lab_b,
                     const float *ts_b, const float *ts_t, uint64_t n, int *layout,
                     uint64_t *s, uint64_t *e) {
    uint64_t lo, hi;
    if (type == 0) { *layout = 1; *s = 0; *e = n; return; }
    if (type == 2) {
        *layout = 1;
        lo = 0; hi = n; /* first ts >= l */
        while (lo < hi) { uint64_t m = (lo + hi) / 2; if (ts_t[m] < l) lo = m + 1; else hi = m; }
        *s = lo;
        lo = 0; hi = n; /* first ts > r */
        while (lo < hi) { uint64_t m = (lo + hi) / 2; if (ts_t[m] <= r) lo = m + 1; else hi = m; }
        *e = lo;
        if (*e < *s) *e = *s;
        return;
    }
    *layout = 0;
    /* first position whose (label, ts) is not below the lower key */
    lo = 0; hi = n;
    while (lo < hi) {
        uint64_t m = (lo + hi) / 2;
        int below = lab_b[m] < v || (type == 3 && lab_b[m] == v && ts_b[m] < l);
        if (below) lo = m + 1; else hi = m;
    }
    *s = lo;
    lo = 0; hi = n;
    while (lo < hi) {
        uint64_t m = (lo + hi) / 2;
        int notabove = lab_b[m] < v || (lab_b[m] == v && (type == 1 || ts_b[m] <= r));
        if (notabove) lo = m + 1; else hi = m;
    }
    *e = lo;
    if (*e < *s) *e = *s;
}

/* Exact canonical-FP32 top-k over contiguous ranges of permuted copies: the
 * CPU path timed beside the GPU (the router's range form of Bruteforce,
 * hybrid_graph.cpp:176-199 + utils.h:80-158).  vec_b / vec_t are the base rows
 * in "both" / "time" order, id_b / id_t map positions back to original ids.
 * Output ascending (dist, id), padded. */
void orc_exact_ranges(const float *vec_b, const uint32_t *id_b, const int32_t *lab_b,
                      const float *ts_b, const float *vec_t, const uint32_t *id_t,
                      const float *ts_t, uint64_t n, int d, const float *queries,
                      const float *metas, uint64_t nq, int k, uint32_t *out_ids,
                      float *out_dists) {
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t qi = 0; qi < (int64_t)nq; ++qi) {
        int lay; uint64_t s, e;
        orc_query_range((int)metas[qi * 4], (int32_t)metas[qi * 4 + 1], metas[qi * 4 + 2],
                        metas[qi * 4 + 3], lab_b, ts_b, ts_t, n, &lay, &s, &e);
        const float *vec = lay == 0 ? vec_b : vec_t;
        const uint32_t *idm = lay == 0 ? id_b : id_t;
        const float *q = queries + qi * d;
        kheap H;
        H.h = (uint64_t *)malloc(8 * (k > 0 ? k : 1));
        H.n = 0; H.k = k;
        for (uint64_t j = s; j < e; ++j)
            kh_push(&H, mk_key(orc_l2_canon(vec + j * d, q, d), idm[j]));
        kh_emit(&H, k, out_ids + qi * k, out_dists + qi * k);
        free(H.h);
    }
}

/* The same search in FP64 (check time only: recall and k-th distance error
 * against the exact answer): each range's k smallest (FP64 distance, id),
 * ascending, padded with id 0xFFFFFFFF and +inf.  Equal to orc_exact_predicate
 * mode 1 on every query (tests/test_oracle.py). */
void orc_exact_ranges_f64(const float *vec_b, const uint32_t *id_b, const int32_t *lab_b,
                          const float *ts_b, const float *vec_t, const uint32_t *id_t,
                          const float *ts_t, uint64_t n, int d, const float *queries,
                          const float *metas, uint64_t nq, int k, uint32_t *out_ids,
                          double *out_dists) {
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t qi = 0; qi < (int64_t)nq; ++qi) {
        int lay; uint64_t s, e;
        orc_query_range((int)metas[qi * 4], (int32_t)metas[qi * 4 + 1], metas[qi * 4 + 2],
                        metas[qi * 4 + 3], lab_b, ts_b, ts_t, n, &lay, &s, &e);
        const float *vec = lay == 0 ? vec_b : vec_t;
        const uint32_t *idm = lay == 0 ? id_b : id_t;
        const float *q = queries + qi * d;
        double *bd = (double *)malloc(sizeof(double) * (k + 1));
        uint32_t *bi = (uint32_t *)malloc(sizeof(uint32_t) * (k + 1));
        int cnt = 0;
        for (uint64_t j = s; j < e; ++j) {
            const double dist = orc_l2_f64(vec + j * d, q, d);
            const uint32_t id = idm[j];
            /* keep the k smallest by (dist, id): positions are not in id order */
            if (cnt == k && !(dist < bd[k - 1] || (dist == bd[k - 1] && id < bi[k - 1]))) continue;
            int p = cnt < k ? cnt : k - 1;
            while (p > 0 && (bd[p - 1] > dist || (bd[p - 1] == dist && bi[p - 1] > id))) {
                bd[p] = bd[p - 1]; bi[p] = bi[p - 1]; --p;
            }
            bd[p] = dist; bi[p] = id;
            if (cnt < k) ++cnt;
        }
        for (int t = 0; t < k; ++t) {
            out_ids[qi * k + t] = t < cnt ? bi[t] : ORC_PAD_ID;
            out_dists[qi * k + t] = t < cnt ? bd[t] : HUGE_VAL;
        }
        free(bd); free(bi);
    }
}

/* Canonical FP32 distance of every returned (query, id) pair: out[q*k + j] =
 * orc_l2_canon(base[ids[q*k + j]], queries[q]), +inf for the pad id.  base is
 * row-major in original id order. */
void orc_l2_canon_rows(const float *base, int d, const float *queries, const uint32_t *ids,
                       uint64_t nq, int k, float *out) {
#pragma omp parallel for schedule(static)
    for (int64_t qi = 0; qi < (int64_t)nq; ++qi)
        for (int j = 0; j < k; ++j) {
            const uint32_t id = ids[qi * k + j];
            out[qi * k + j] = id == ORC_PAD_ID ? HUGE_VALF
                                               : orc_l2_canon(base + (uint64_t)id * d, queries + qi * d, d);
        }
}

/* ------------------------------------------------------------------------ */
/* reference bruteforce() semantics (bruteforce.h:51-117, 244-289)           */
/* ------------------------------------------------------------------------ */

/* Exact restatement of maxPQIFCS<int> (bruteforce.h:57-117) driven as in
 * bruteforce() (bruteforce.h:254-264): emplace the first k, then
 * must_pop_emplace (reject d >= root), then write closest_points[k-1-l] =
 * heap[l+1].id.  Positions are relative to the code slice.  dists_out (may be
 * NULL) receives the matching int distances, as bruteforce_subgraph does
 * before its alpha^2 scaling (bruteforce.h:285-288). */
void orc_bf_heap_ref(int k, uint64_t npoints, const int8_t *codes, const int8_t *qcode,
                     int32_t *closest_points, int32_t *dists_out) {
    int32_t *hd = (int32_t *)malloc(sizeof(int32_t) * (k + 5));
    int32_t *hi = (int32_t *)malloc(sizeof(int32_t) * (k + 5));
    int cnt = 0;
    for (uint64_t p = 0; p < npoints; ++p) {
        int32_t dist = orc_l2_sq8sym(qcode, codes + p * ORC_CODE_DIM);
        if ((int)p < k) { /* emplace, bruteforce.h:106-113 */
            hd[++cnt] = dist; hi[cnt] = (int32_t)p;
            int now = cnt;
            while (now != 1 && hd[now] > hd[now >> 1]) {
                int32_t t = hd[now]; hd[now] = hd[now >> 1]; hd[now >> 1] = t;
                t = hi[now]; hi[now] = hi[now >> 1]; hi[now >> 1] = t;
                now >>= 1;
            }
            continue;
        }
        if (hd[1] <= dist) continue; /* must_pop_emplace, bruteforce.h:99-105 */
        hd[1] = dist; hi[1] = (int32_t)p;
        int x = 1; /* push_down, bruteforce.h:78-87 */
        while ((x << 1) <= cnt) {
            int t = x << 1;
            if ((t | 1) <= cnt && hd[t | 1] > hd[t]) t |= 1;
            if (hd[t] <= hd[x]) break;
            int32_t a = hd[x]; hd[x] = hd[t]; hd[t] = a;
            a = hi[x]; hi[x] = hi[t]; hi[t] = a;
            x = t;
        }
    }
    for (int l = 0; l < k; ++l) {
        closest_points[k - l - 1] = hi[l + 1];
        if (dists_out) dists_out[k - l - 1] = hd[l + 1];
    }
    free(hd); free(hi);
}

/* The same selection with this project's deterministic contract: the k
 * smallest (dist, position) pairs, ascending. */
void orc_bf_sorted(int k, uint64_t npoints, const int8_t *codes, const int8_t *qcode,
                   int32_t *ids, int32_t *dists) {
    kheap H;
    H.h = (uint64_t *)malloc(8 * (k > 0 ? k : 1));
    H.n = 0; H.k = k;
    for (uint64_t p = 0; p < npoints; ++p) {
        int32_t dist = orc_l2_sq8sym(qcode, codes + p * ORC_CODE_DIM);
        /* a non-negative int distance is already a monotone key */
        kh_push(&H, ((uint64_t)(uint32_t)dist << 32) | (uint32_t)p);
    }
    qsort(H.h, H.n, 8, cmp_u64);
    for (int j = 0; j < k; ++j) {
        ids[j] = j < H.n ? (int32_t)(H.h[j] & 0xFFFFFFFFu) : -1;
        dists[j] = j < H.n ? (int32_t)(H.h[j] >> 32) : INT32_MAX;
    }
    free(H.h);
}

/* The router's brute-force branch (hybrid_graph.cpp:203-228) over contiguous
 * ranges: SQ8-symmetric scan keeping K = min(kq, e-s) candidates, map to
 * original ids, exact FP32 rerank to k.  Used as the CPU "reference route"
 * timing; rerank uses the canonical FP32 distance and (dist,id) order. */
void orc_router_bf(const int8_t *codes_b, const int8_t *codes_t, const uint32_t *id_b,
                   const int32_t *lab_b, const float *ts_b, const uint32_t *id_t,
                   const float *ts_t, const float *base, uint64_t n, int d, float alpha,
                   const float *queries, const float *metas, uint64_t nq, int kq, int k,
                   uint32_t *out_ids, float *out_dists) {
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t qi = 0; qi < (int64_t)nq; ++qi) {
        int lay; uint64_t s, e;
        orc_query_range((int)metas[qi * 4], (int32_t)metas[qi * 4 + 1], metas[qi * 4 + 2],
                        metas[qi * 4 + 3], lab_b, ts_b, ts_t, n, &lay, &s, &e);
        const int8_t *codes = (lay == 0 ? codes_b : codes_t) + s * ORC_CODE_DIM;
        const uint32_t *idm = (lay == 0 ? id_b : id_t) + s;
        const float *q = queries + qi * d;
        int8_t qc[ORC_CODE_DIM];
        orc_sq8sym_encode(q, 1, d, alpha, qc);
        uint64_t np = e - s;
        int K = (uint64_t)kq < np ? kq : (int)np;
        int32_t *cand = (int32_t *)malloc(sizeof(int32_t) * (K > 0 ? K : 1));
        int32_t *cd = (int32_t *)malloc(sizeof(int32_t) * (K > 0 ? K : 1));
        if (K > 0) orc_bf_sorted(K, np, codes, qc, cand, cd);
        kheap H;
        H.h = (uint64_t *)malloc(8 * (k > 0 ? k : 1));
        H.n = 0; H.k = k;
        for (int j = 0; j < K; ++j) {
            uint32_t id = idm[cand[j]];
            kh_push(&H, mk_key(orc_l2_canon(base + (uint64_t)id * d, q, d), id));
        }
        kh_emit(&H, k, out_ids + qi * k, out_dists + qi * k);
        free(H.h); free(cand); free(cd);
    }
}

/* ------------------------------------------------------------------------ */
/* type-2 sub-interval path (hybrid_graph.cpp:91-134, 317-494)               */
/* ------------------------------------------------------------------------ */

/* The 0.1-wide time buckets of ReadSortedBaseTimestamp (utils.h:470-478) over
 * the timestamp-order timestamps ts_t: bucket i = [lower_bound(0.1 i),
 * upper_bound(0.1 (i + 1))), the float timestamps compared with the double
 * bounds (so a timestamp equal to a bound lands in both buckets). */
void orc_time_buckets(const float *ts_t, uint64_t n, uint64_t *start, uint64_t *count) {
    for (int i = 0; i < 10; ++i) {
        const double lo = 0.1 * i, hi = 0.1 * (i + 1);
        uint64_t a = 0, b = n;
        while (a < b) { uint64_t m = (a + b) / 2; if ((double)ts_t[m] < lo) a = m + 1; else b = m; }
        const uint64_t s = a;
        a = 0; b = n;
        while (a < b) { uint64_t m = (a + b) / 2; if (!(hi < (double)ts_t[m])) a = m + 1; else b = m; }
        start[i] = s;
        count[i] = a > s ? a - s : 0;
    }
}

/* SplitInterval (hybrid_graph.cpp:91-134) of [l, r] over the buckets, with the
 * router's threshold rule (hybrid_graph.cpp:324-333): thr = 0.5 when the
 * query's selectivity (e - s) / nb over the timestamp layout is below 0.2,
 * else 0.2 (small_thr > 0 replaces the rule).  types[t]: 255 no overlap or no
 * bucket, 0 SMALL, 1 MEDIUM, 3 FULL (the reference's RFType values);
 * loc_s / loc_n: the bucket-local start and row count of a partial bucket. */
void orc_split_interval(float l, float r, const float *ts_t, uint64_t n, uint64_t nb, const uint64_t *bstart,
                        const uint64_t *bcount, float small_thr, uint8_t *types, uint64_t *loc_s, uint64_t *loc_n) {
    uint64_t a = 0, b = n;
    while (a < b) { uint64_t m = (a + b) / 2; if (ts_t[m] < l) a = m + 1; else b = m; }
    const uint64_t gs = a;
    a = 0; b = n;
    while (a < b) { uint64_t m = (a + b) / 2; if (ts_t[m] <= r) a = m + 1; else b = m; }
    const uint64_t ge = a > gs ? a : gs;
    const float sel = (float)(ge - gs) / (float)nb;
    const float thr = small_thr > 0.0f ? small_thr : ((double)sel < 0.2 ? 0.5f : 0.2f);
    for (int t = 0; t < 10; ++t) {
        types[t] = 255; loc_s[t] = 0; loc_n[t] = 0;
        if (bcount[t] == 0) continue;
        const float ts = (float)t / 10, te = (float)(t + 1) / 10;
        if (r < ts || l > te) continue;
        if (l <= ts && r >= te) { types[t] = 3; loc_n[t] = bcount[t]; continue; }
        const float lb = ts > l ? ts : l, rb = te < r ? te : r;
        const float *bt = ts_t + bstart[t];
        const uint64_t bn = bcount[t];
        a = 0; b = bn;
        while (a < b) { uint64_t m = (a + b) / 2; if (bt[m] < lb) a = m + 1; else b = m; }
        const uint64_t s = a;
        a = 0; b = bn;
        while (a < b) { uint64_t m = (a + b) / 2; if (bt[m] <= rb) a = m + 1; else b = m; }
        const uint64_t e = a > s ? a : s;
        const float bsel = (float)(e - s) / (float)bn;
        types[t] = bsel < thr ? 0 : 1;
        loc_s[t] = s;
        loc_n[t] = e - s;
    }
}

/* An int * alpha * alpha pool distance as the reference forms it
 * (bruteforce.h:287, then the implicit float -> int conversion of
 * maxPQIFCS<int>::maybe_pop_emplace, hybrid_graph.cpp:415): truncation toward
 * zero; values at or above 2^31 (undefined there) saturate. */
static uint32_t pool_dist(int32_t d, float alpha) {
    const float f = ((float)d * alpha) * alpha;
    return f >= 2147483648.0f ? 0x7FFFFFFFu : (uint32_t)(int32_t)f;
}

/* The type-2 sub-interval route for the SMALL buckets (hybrid_graph.cpp:
 * 317-420) and the global refine (:463-494), with this project's tie rules:
 * each SMALL bucket's K = min(bucket_k, n_points) smallest (int distance,
 * position) of its slice of the bucket codes (query encoded with the
 * bucket's alpha), their distances int(d * alpha * alpha); the pool keeps
 * the pool_k smallest (pool distance, timestamp-layout position) over the
 * query's buckets; the refine reranks the pool with the canonical FP32
 * distance and emits the k smallest (distance, original id), padded.
 * base: original order; id_t / ts_t: the timestamp layout; bcodes: each
 * bucket's codes (bcount[t] rows of 112 bytes from row boff[t]); types:
 * nq x 10 as orc_split_interval; ext_pos / ext_dist: n_ext caller candidates per
 * query (timestamp-layout positions, pool distances) that join the pool.
 * Queries must be type 2. */
void orc_subinterval(const float *base, const uint32_t *id_t, const float *ts_t, uint64_t n, int d,
                     const int8_t *bcodes, const uint64_t *boff, const uint64_t *bstart, const uint64_t *bcount,
                     const float *balpha, const float *queries, const float *metas, uint64_t nq, uint64_t nb,
                     float small_thr, int bucket_k, int pool_k, int k, const uint32_t *ext_pos,
                     const float *ext_dist, int n_ext, uint32_t *out_ids, float *out_dists, uint8_t *types) {
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t qi = 0; qi < (int64_t)nq; ++qi) {
        uint64_t ls[10], ln[10];
        orc_split_interval(metas[qi * 4 + 2], metas[qi * 4 + 3], ts_t, n, nb, bstart, bcount, small_thr,
                           types + qi * 10, ls, ln);
        const float *q = queries + qi * d;
        kheap P;
        P.h = (uint64_t *)malloc(8 * (pool_k > 0 ? pool_k : 1));
        P.n = 0; P.k = pool_k;
        int32_t *pos = (int32_t *)malloc(sizeof(int32_t) * (bucket_k > 0 ? bucket_k : 1));
        int32_t *dd = (int32_t *)malloc(sizeof(int32_t) * (bucket_k > 0 ? bucket_k : 1));
        for (int t = 0; t < 10; ++t) {
            if (types[qi * 10 + t] != 0 || ln[t] == 0) continue;
            int8_t qc[ORC_CODE_DIM];
            orc_sq8sym_encode(q, 1, d, balpha[t], qc);
            const int K = (uint64_t)bucket_k < ln[t] ? bucket_k : (int)ln[t];
            orc_bf_sorted(K, ln[t], bcodes + (boff[t] + ls[t]) * ORC_CODE_DIM, qc, pos, dd);
            for (int j = 0; j < K; ++j) {
                const uint32_t p = (uint32_t)(bstart[t] + ls[t] + (uint64_t)pos[j]);
                kh_push(&P, ((uint64_t)pool_dist(dd[j], balpha[t]) << 32) | p);
            }
        }
        /* the caller's candidates (graph searches of FULL / MEDIUM buckets):
         * float distance -> int as maybe_pop_emplace(int, int) converts it */
        for (int j = 0; j < n_ext; ++j) {
            const uint32_t p = ext_pos[qi * n_ext + j];
            const float f = ext_dist[qi * n_ext + j];
            if (p == ORC_PAD_ID || p >= n || !(f >= 0.0f)) continue;
            const uint32_t pd = f >= 2147483648.0f ? 0x7FFFFFFFu : (uint32_t)(int32_t)f;
            kh_push(&P, ((uint64_t)pd << 32) | p);
        }
        kheap H;
        H.h = (uint64_t *)malloc(8 * (k > 0 ? k : 1));
        H.n = 0; H.k = k;
        for (int j = 0; j < P.n; ++j) {
            const uint32_t id = id_t[(uint32_t)P.h[j]];
            kh_push(&H, mk_key(orc_l2_canon(base + (uint64_t)id * d, q, d), id));
        }
        kh_emit(&H, k, out_ids + qi * k, out_dists + qi * k);
        free(H.h); free(P.h); free(pos); free(dd);
    }
}

/* ------------------------------------------------------------------------ */
/* evaluation                                                               */
/* ------------------------------------------------------------------------ */

/* GetKNNRecall (utils.h:224-253): sum over queries of |knn_i intersect gt_i|
 * (each gt id counted once if it appears anywhere in knn_i), / (nq * k).
 * per_query (may be NULL) receives the per-query counts (recall1.bin). */
double orc_recall(const uint32_t *knn, const uint32_t *gt, uint64_t nq, int k,
                  int32_t *per_query) {
    uint64_t tot = 0;
    for (uint64_t i = 0; i < nq; ++i) {
        int c = 0;
        for (int j = 0; j < k; ++j)
            for (int t = 0; t < k; ++t)
                if (knn[i * k + t] == gt[i * k + j]) { ++c; break; }
        if (per_query) per_query[i] = c;
        tot += c;
    }
    return nq ? (double)tot / (double)nq / (double)k : 0.0;
}

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
