// fbf_api.hip -- C-ABI entry points, device-resident dataset, batch planner.
//
// The planner is the router's range step (hybrid_graph.cpp:168-199) made
// batch-wide: each query's filter becomes a contiguous range of one of two
// sorted layouts (SortDataset, hybrid_graph.cpp:505-524); queries are ordered
// by (layout, start, end) as ReadSortedQuery orders them (utils.h:514-537),
// grouped so that up to 16 queries share each streamed tile, and every group's
// row span is cut into work items of chunk_rows rows.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <thread>
#include <unordered_map>
#include <cmath>
#include <cstdio>
#include <condition_variable>
#include <cstring>
#include <map>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/fbf.h"
#include "fbf_internal.hpp"

using namespace fbf;

// ---------------------------------------------------------------- errors
static thread_local std::string g_err;

static fbf_status fail(fbf_status s, const std::string &msg) {
    g_err = msg;
    return s;
}

#define HIP_TRY(expr)                                                                    \
    do {                                                                                 \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess)                                                            \
            return fail(FBF_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

// ---------------------------------------------------------------- device buffers
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    // Grows to at least b bytes, by 1.5x of the old size or more, so a run of
    // slightly larger requests does not reallocate each time.
    hipError_t reserve(size_t b) {
        if (b <= bytes) return hipSuccess;
        const size_t old = bytes;
        if (p) hipFree(p);
        p = nullptr;
        bytes = 0;
        const size_t nb = std::max(b, old + old / 2);
        hipError_t e = hipMalloc(&p, nb);
        if (e == hipSuccess) bytes = nb;
        return e;
    }
    void release() {
        if (p) hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    template <class T> T *as() const { return static_cast<T *>(p); }
};

struct HostPinned {
    void *p = nullptr;
    size_t bytes = 0;
    hipError_t reserve(size_t b) {
        if (b <= bytes) return hipSuccess;
        const size_t old = bytes;
        if (p) hipHostFree(p);
        p = nullptr;
        bytes = 0;
        const size_t nb = std::max(b, old + old / 2);
        hipError_t e = hipHostMalloc(&p, nb, hipHostMallocDefault);
        if (e == hipSuccess) bytes = nb;
        return e;
    }
    void release() {
        if (p) hipHostFree(p);
        p = nullptr;
        bytes = 0;
    }
};

struct LayoutDev {
    uint32_t *ids = nullptr;      // [n_pad] position -> original id (pad: kPadId)
    float *soa = nullptr;         // FP32 scan copy (exact mode)
    void *bf16 = nullptr;         // BF16 MFMA row fragments
    void *fp8 = nullptr;          // FP8 e4m3 MFMA row fragments
    uint8_t *fp8_scale = nullptr; // [n_pad] E8M0 row scales
    float *xnorm = nullptr;       // [n_pad] |x|^2 per position (FP8)
    void *sq8 = nullptr;          // uint8 codes in MFMA order
    float *sq8_norm = nullptr;    // [n_pad] |xhat|^2 per position
    int8_t *sym = nullptr;        // [n_pad][112] symmetric int8 codes
    void *sym_mfma = nullptr;     // symmetric int8 codes in MFMA order (K = 128)
    float *sym_norm = nullptr;    // [n_pad] |codes|^2 per position
    void free_all() {
        for (void *p : {(void *)ids, (void *)soa, bf16, fp8, (void *)fp8_scale, (void *)xnorm, sq8,
                        (void *)sq8_norm, (void *)sym, sym_mfma, (void *)sym_norm})
            if (p) (void)hipFree(p);
        *this = LayoutDev();
    }
};

struct ProfEvent {
    int kind;  // 0 scan, 1 merge, 2 rerank(part of merge), 3 build, 4 plan (host)
    int pass = 0;  // scans: 0 single or main pass, 1 seeded prefix pass, 2 masked rescan
    hipEvent_t a = nullptr, b = nullptr;
    double host_ms = 0.0;  // kind 4: wall time on the host
};

// Per-call scratch: the uploaded plan, the per-(item, query) candidate lists
// and the pinned staging buffer the plan is copied through.
struct Scratch {
    DevBuf d_items, d_iq, d_slot_off, d_lists, d_counts, d_queue, d_q, d_qf, d_out_ids, d_out_d, d_gtau,
        d_tau0, d_flags, d_pre_end;
    // FBF_MODE_EXACT_BF16: per-query certificates and the uncertified subset.
    DevBuf d_cert, d_sub_idx, d_sub_q, d_sub_ids, d_sub_d;
    // fbf_search_subinterval: slices, candidate pool, pool slot offsets.
    DevBuf d_ti_sl, d_ti_lists, d_ti_counts, d_ti_off, d_ti_ext;
    HostPinned h_ti;                 // staging of the slices and pool offsets
    hipEvent_t ev_ti = nullptr;      // recorded after the copies out of h_ti
    HostPinned h_stage[2], h_nflag;
    hipEvent_t ev_stage[2] = {nullptr, nullptr};  // recorded after the copies out of h_stage[i]
    int stage_next = 0;
    hipEvent_t ev_nflag = nullptr;  // recorded after the rescan count is copied to h_nflag
    // Recorded on the search's stream after the last device work of every
    // search that uses this scratch; the next search (on any stream) makes its
    // stream wait for it.  An event, not the previous stream's handle: the
    // caller may destroy that stream between searches.
    hipEvent_t ev_done = nullptr;
    bool used = false;
    void release() {
        for (DevBuf *b : {&d_items, &d_iq, &d_slot_off, &d_lists, &d_counts, &d_queue, &d_q, &d_qf, &d_out_ids,
                          &d_out_d, &d_gtau, &d_tau0, &d_flags, &d_pre_end, &d_cert, &d_sub_idx, &d_sub_q,
                          &d_sub_ids, &d_sub_d, &d_ti_sl, &d_ti_lists, &d_ti_counts, &d_ti_off, &d_ti_ext})
            b->release();
        h_ti.release();
        if (ev_ti) (void)hipEventDestroy(ev_ti);
        ev_ti = nullptr;
        for (int i = 0; i < 2; ++i) {
            h_stage[i].release();
            if (ev_stage[i]) (void)hipEventDestroy(ev_stage[i]);
            ev_stage[i] = nullptr;
        }
        h_nflag.release();
        if (ev_nflag) (void)hipEventDestroy(ev_nflag);
        ev_nflag = nullptr;
        if (ev_done) (void)hipEventDestroy(ev_done);
        ev_done = nullptr;
        used = false;
    }
};

struct fbf_dataset {
    int device = 0;
    uint64_t n = 0, n_pad = 0;
    uint32_t d = 0;
    uint32_t repr = 0;
    // host metadata for the planner
    std::vector<int32_t> lab_b;
    std::vector<float> ts_b, ts_t;
    // the same sorted metadata on the device (the device router, fbf_route)
    int32_t *d_lab_b = nullptr;
    float *d_ts_b = nullptr, *d_ts_t = nullptr;
    // the device router's buffers (one routing at a time: mu_route)
    std::mutex mu_route;
    DevBuf r_keys, r_vals, r_temp, r_ranges;
    // device data
    float *base_rows = nullptr;  // [n][d] original order
    LayoutDev lay[2];
    float *sq8_mi = nullptr, *sq8_dif = nullptr, *sq8_w = nullptr;  // [d] each
    float alpha = 0.f;           // symmetric int8 quantiser scale
    float max_norm = 0.f;        // largest row norm |x| (rounded up; BF16 builds), for FBF_MODE_EXACT_BF16
    float bf16_rho = 0.f;        // largest |bf16(x) - x| / |x| over the rows (rounded up), same use
    std::vector<float> h_mi, h_dif;
    // Time buckets of the timestamp layout (FBF_REPR_TBUCKETS): bucket t holds
    // positions [tb_start[t], tb_start[t] + tb_count[t]) of layout 1, its own
    // symmetric int8 scale tb_alpha[t], and its codes at rows tb_off[t].. of tb_codes.
    uint64_t tb_start[10] = {}, tb_count[10] = {}, tb_off[10] = {};
    float tb_alpha[10] = {};
    int8_t *tb_codes = nullptr;
    uint64_t tb_rows = 0;
    Scratch scr;
    std::mutex mu;
    std::mutex mu_cert;          // one FBF_MODE_EXACT_BF16 search at a time (its subset buffers)
    // measurement
    bool profile = false;
    std::vector<ProfEvent> events;
    uint64_t last_pairs = 0, last_pairs_computed = 0, last_items = 0;
    uint64_t last_pairs_prefix = 0;  // pairs of the seeded prefix pass (0: none)
    bool last_seeded = false;  // the last search ran the seeded prefix pass (h_nflag is live)
    uint64_t last_fallbacks = 0;  // FBF_MODE_EXACT_BF16: queries searched again by the FP32 scan
};

// ---------------------------------------------------------------- helpers
static inline int32_t label_int(float v) { return (int32_t)v; }  // utils.h:99,109

struct QRange {
    uint32_t layout, s, e, qi;
};

// Range of one query (hybrid_graph.cpp:176-199 semantics, integer labels,
// inclusive timestamp bounds).  Returns false for an unknown filter type.
static bool query_range(const fbf_dataset *ds, const float *meta, QRange &r) {
    const float tf = meta[0];
    const int type = (int)tf;
    if (!(tf == (float)type) || type < 0 || type > 3) return false;
    const int32_t v = label_int(meta[1]);
    const float l = meta[2], rr = meta[3];
    const uint64_t n = ds->n;
    if (type == 0) {
        r.layout = 1; r.s = 0; r.e = (uint32_t)n;
        return true;
    }
    if (type == 2) {
        r.layout = 1;
        const float *t = ds->ts_t.data();
        const uint64_t s = std::lower_bound(t, t + n, l) - t;
        const uint64_t e = std::upper_bound(t, t + n, rr) - t;
        r.s = (uint32_t)s;
        r.e = (uint32_t)std::max(s, e);
        return true;
    }
    r.layout = 0;
    const int32_t *lab = ds->lab_b.data();
    const float *ts = ds->ts_b.data();
    uint64_t lo = 0, hi = n;
    while (lo < hi) {
        uint64_t m = (lo + hi) / 2;
        bool below = lab[m] < v || (type == 3 && lab[m] == v && ts[m] < l);
        if (below) lo = m + 1; else hi = m;
    }
    const uint64_t s = lo;
    lo = 0; hi = n;
    while (lo < hi) {
        uint64_t m = (lo + hi) / 2;
        bool notabove = lab[m] < v || (lab[m] == v && (type == 1 || ts[m] <= rr));
        if (notabove) lo = m + 1; else hi = m;
    }
    r.s = (uint32_t)s;
    r.e = (uint32_t)std::max(s, lo);
    return true;
}

static void record_begin(fbf_dataset *ds, int kind, hipStream_t st, ProfEvent &ev) {
    if (!ds->profile) return;
    ev.kind = kind;
    hipEventCreate(&ev.a);
    hipEventCreate(&ev.b);
    hipEventRecord(ev.a, st);
}
static void record_end(fbf_dataset *ds, hipStream_t st, ProfEvent &ev) {
    if (!ds->profile) return;
    hipEventRecord(ev.b, st);
    ds->events.push_back(ev);
}

// ---------------------------------------------------------------- dataset create
static fbf_status dataset_build(fbf_dataset *ds, const float *labels, const float *timestamps,
                                hipStream_t st) {
    const uint64_t n = ds->n;
    std::vector<int32_t> lab(n);
    for (uint64_t i = 0; i < n; ++i) lab[i] = label_int(labels[i]);
    std::vector<uint32_t> pb(n), pt(n);
    {
        // SortDataset's orders (comparators hybrid_graph.cpp:14-27) with the id
        // as the final key, as stable radix sorts of packed keys on the device:
        // (label ^ 2^31) << 32 | f2ord(timestamp), and f2ord(timestamp).  -0.0
        // and +0.0 compare equal as floats, so both get the key of +0.0.
        std::vector<uint64_t> kb(n), kt(n);
        for (uint64_t i = 0; i < n; ++i) {
            const float t = timestamps[i] == 0.0f ? 0.0f : timestamps[i];
            const uint32_t o = f2ord(t);
            kb[i] = ((uint64_t)((uint32_t)lab[i] ^ 0x80000000u) << 32) | o;
            kt[i] = o;
        }
        struct Guard {
            DevBuf k, p;
            ~Guard() { k.release(); p.release(); }
        } g;
        HIP_TRY(g.k.reserve(n * sizeof(uint64_t)));
        HIP_TRY(g.p.reserve(n * sizeof(uint32_t)));
        HIP_TRY(hipMemcpy(g.k.p, kb.data(), n * sizeof(uint64_t), hipMemcpyHostToDevice));
        HIP_TRY(sort_rows_by_key(g.k.as<uint64_t>(), n, 64, g.p.as<uint32_t>(), st));
        HIP_TRY(hipMemcpy(pb.data(), g.p.p, n * sizeof(uint32_t), hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(g.k.p, kt.data(), n * sizeof(uint64_t), hipMemcpyHostToDevice));
        HIP_TRY(sort_rows_by_key(g.k.as<uint64_t>(), n, 32, g.p.as<uint32_t>(), st));
        HIP_TRY(hipMemcpy(pt.data(), g.p.p, n * sizeof(uint32_t), hipMemcpyDeviceToHost));
    }
    ds->lab_b.resize(n);
    ds->ts_b.resize(n);
    ds->ts_t.resize(n);
    for (uint64_t i = 0; i < n; ++i) {
        ds->lab_b[i] = lab[pb[i]];
        ds->ts_b[i] = timestamps[pb[i]];
        ds->ts_t[i] = timestamps[pt[i]];
    }
    HIP_TRY(hipMalloc(&ds->d_lab_b, n * sizeof(int32_t)));
    HIP_TRY(hipMalloc(&ds->d_ts_b, n * sizeof(float)));
    HIP_TRY(hipMalloc(&ds->d_ts_t, n * sizeof(float)));
    HIP_TRY(hipMemcpy(ds->d_lab_b, ds->lab_b.data(), n * sizeof(int32_t), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(ds->d_ts_b, ds->ts_b.data(), n * sizeof(float), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(ds->d_ts_t, ds->ts_t.data(), n * sizeof(float), hipMemcpyHostToDevice));
    const uint64_t n_pad = ds->n_pad;
    std::vector<uint32_t> *perms[2] = {&pb, &pt};
    for (int L = 0; L < 2; ++L) {
        std::vector<uint32_t> &p = *perms[L];
        p.resize(n_pad, kPadId);
        HIP_TRY(hipMalloc(&ds->lay[L].ids, n_pad * sizeof(uint32_t)));
        HIP_TRY(hipMemcpy(ds->lay[L].ids, p.data(), n_pad * sizeof(uint32_t), hipMemcpyHostToDevice));
    }
    ProfEvent ev;
    record_begin(ds, 3, st, ev);
    const uint32_t repr = ds->repr;
    const uint64_t nb32 = n_pad / 32;  // 32-row MFMA blocks covering the padded layout
    DevBuf norms, tmp_codes, tmp_xhat, absbits;
    if (repr & (FBF_REPR_BF16 | FBF_REPR_FP8)) {
        HIP_TRY(norms.reserve(n * sizeof(float)));
        HIP_TRY(launch_row_norms(ds->base_rows, n, norms.as<float>(), st));
        // The largest |x|^2 (its bits order like the value: norms are >= 0).
        HIP_TRY(absbits.reserve(sizeof(uint32_t)));
        HIP_TRY(hipMemsetAsync(absbits.p, 0, sizeof(uint32_t), st));
        HIP_TRY(launch_absmax(norms.as<float>(), n, absbits.as<uint32_t>(), st));
        uint32_t bits = 0;
        HIP_TRY(hipMemcpyAsync(&bits, absbits.p, 4, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        float m2 = 0.f;
        std::memcpy(&m2, &bits, 4);
        ds->max_norm = (float)(std::sqrt((double)m2) * (1.0 + 1e-5));
        HIP_TRY(hipMemsetAsync(absbits.p, 0, sizeof(uint32_t), st));
        HIP_TRY(launch_bf16_err_max(ds->base_rows, n, absbits.as<uint32_t>(), st));
        HIP_TRY(hipMemcpyAsync(&bits, absbits.p, 4, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        float rho = 0.f;
        std::memcpy(&rho, &bits, 4);
        ds->bf16_rho = (float)((double)rho * 1.001 + 1e-7);
    }
    if (repr & FBF_REPR_SQ8) {
        HIP_TRY(hipMalloc(&ds->sq8_mi, kDim * sizeof(float)));
        HIP_TRY(hipMalloc(&ds->sq8_dif, kDim * sizeof(float)));
        HIP_TRY(hipMalloc(&ds->sq8_w, kDim * sizeof(float)));
        HIP_TRY(tmp_codes.reserve(n * kDim));
        HIP_TRY(tmp_xhat.reserve(n * sizeof(float)));
        HIP_TRY(launch_sq8_train_encode(ds->base_rows, n, ds->sq8_mi, ds->sq8_dif, tmp_codes.as<uint8_t>(),
                                        tmp_xhat.as<float>(), st));
        ds->h_mi.resize(kDim);
        ds->h_dif.resize(kDim);
        HIP_TRY(hipMemcpyAsync(ds->h_mi.data(), ds->sq8_mi, kDim * 4, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipMemcpyAsync(ds->h_dif.data(), ds->sq8_dif, kDim * 4, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        std::vector<float> w(kDim);
        for (int k = 0; k < kDim; ++k) w[k] = ds->h_dif[k] / 255.0f;
        HIP_TRY(hipMemcpy(ds->sq8_w, w.data(), kDim * 4, hipMemcpyHostToDevice));
    }
    if (repr & FBF_REPR_SQ8SYM) {
        HIP_TRY(absbits.reserve(sizeof(uint32_t)));
        HIP_TRY(hipMemsetAsync(absbits.p, 0, sizeof(uint32_t), st));
        HIP_TRY(launch_absmax(ds->base_rows, n * kDim, absbits.as<uint32_t>(), st));
        uint32_t bits = 0;
        HIP_TRY(hipMemcpyAsync(&bits, absbits.p, 4, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        std::memcpy(&ds->alpha, &bits, 4);
        if (!(ds->alpha > 0.f)) return fail(FBF_ERR_INVALID_ARG, "all-zero base: symmetric int8 scale undefined");
    }
    for (int L = 0; L < 2; ++L) {
        LayoutDev &D = ds->lay[L];
        if (repr & FBF_REPR_F32) {
            HIP_TRY(hipMalloc(&D.soa, n_pad * kDim * sizeof(float)));
            HIP_TRY(launch_build_soa(ds->base_rows, D.ids, n, D.soa, st));
        }
        if (repr & FBF_REPR_BF16) {
            HIP_TRY(hipMalloc(&D.bf16, nb32 * 7 * 64 * 16));
            HIP_TRY(launch_build_mfma(0, ds->base_rows, D.ids, norms.as<float>(), nullptr, n, D.bf16, nullptr, st));
        }
        if (repr & FBF_REPR_FP8) {
            HIP_TRY(hipMalloc(&D.fp8, nb32 * 2 * 64 * 32));
            HIP_TRY(hipMalloc(&D.fp8_scale, n_pad));
            HIP_TRY(hipMalloc(&D.xnorm, n_pad * sizeof(float)));
            HIP_TRY(launch_build_mfma(1, ds->base_rows, D.ids, nullptr, nullptr, n, D.fp8, D.fp8_scale, st));
            HIP_TRY(launch_gather_f32(norms.as<float>(), D.ids, n, n_pad, D.xnorm, st));
        }
        if (repr & FBF_REPR_SQ8) {
            HIP_TRY(hipMalloc(&D.sq8, nb32 * 7 * 64 * 8));
            HIP_TRY(hipMalloc(&D.sq8_norm, n_pad * sizeof(float)));
            HIP_TRY(launch_build_mfma(2, nullptr, D.ids, tmp_xhat.as<float>(), tmp_codes.as<uint8_t>(), n, D.sq8,
                                      nullptr, st));
            HIP_TRY(launch_gather_f32(tmp_xhat.as<float>(), D.ids, n, n_pad, D.sq8_norm, st));
        }
        if (repr & FBF_REPR_SQ8SYM) {
            HIP_TRY(hipMalloc(&D.sym, n_pad * 112));
            HIP_TRY(launch_sym_encode(ds->base_rows, D.ids, n, n_pad, ds->alpha, D.sym, st));
            // The same codes for the int8 matrix cores (batched searches).
            HIP_TRY(hipMalloc(&D.sym_mfma, nb32 * 4 * 64 * 16));
            HIP_TRY(hipMalloc(&D.sym_norm, n_pad * sizeof(float)));
            HIP_TRY(launch_build_mfma(3, nullptr, D.ids, D.sym_norm, reinterpret_cast<const uint8_t *>(D.sym), n,
                                      D.sym_mfma, nullptr, st));
        }
    }
    if (repr & FBF_REPR_TBUCKETS) {
        // The timestamp sub-indexes' buckets (ReadSortedBaseTimestamp,
        // utils.h:470-478: bucket i = [lower_bound(0.1 i), upper_bound(0.1 (i +
        // 1))) of the timestamp layout, float timestamps against double bounds)
        // and their quantisers, each trained on its own rows (the timestamp
        // searchers' SetData, baseline.cpp:130-141).
        const float *t = ds->ts_t.data();
        uint64_t total = 0;
        for (int b = 0; b < 10; ++b) {
            const double lo = 0.1 * b, hi = 0.1 * (b + 1);
            const uint64_t s0 = std::lower_bound(t, t + n, lo, [](float a, double v) { return (double)a < v; }) - t;
            const uint64_t e0 = std::upper_bound(t, t + n, hi, [](double v, float a) { return v < (double)a; }) - t;
            ds->tb_start[b] = s0;
            ds->tb_count[b] = e0 > s0 ? e0 - s0 : 0;
            ds->tb_off[b] = total;
            total += ds->tb_count[b];
        }
        ds->tb_rows = total;
        HIP_TRY(hipMalloc(&ds->tb_codes, std::max<uint64_t>(1, total) * 112));
        HIP_TRY(absbits.reserve(sizeof(uint32_t)));
        for (int b = 0; b < 10; ++b) {
            const uint64_t cnt = ds->tb_count[b];
            if (cnt == 0) continue;
            const uint32_t *perm = ds->lay[1].ids + ds->tb_start[b];
            HIP_TRY(hipMemsetAsync(absbits.p, 0, sizeof(uint32_t), st));
            HIP_TRY(launch_absmax_rows(ds->base_rows, perm, cnt, absbits.as<uint32_t>(), st));
            uint32_t bits = 0;
            HIP_TRY(hipMemcpyAsync(&bits, absbits.p, 4, hipMemcpyDeviceToHost, st));
            HIP_TRY(hipStreamSynchronize(st));
            std::memcpy(&ds->tb_alpha[b], &bits, 4);
            int8_t *out = ds->tb_codes + ds->tb_off[b] * 112;
            // A bucket of all-zero rows has scale 0 (the reference divides by
            // it); its codes are defined as zero.
            if (ds->tb_alpha[b] > 0.f) HIP_TRY(launch_sym_encode(ds->base_rows, perm, cnt, cnt, ds->tb_alpha[b], out, st));
            else HIP_TRY(hipMemsetAsync(out, 0, cnt * 112, st));
        }
    }
    record_end(ds, st, ev);
    HIP_TRY(hipStreamSynchronize(st));
    norms.release();
    tmp_codes.release();
    tmp_xhat.release();
    absbits.release();
    return FBF_OK;
}

static fbf_status dataset_create_common(const float *vecs, bool vecs_on_device, const float *labels,
                                        const float *timestamps, uint64_t n, uint32_t d,
                                        uint32_t repr, fbf_dataset **out) {
    if (!out || !vecs || !labels || !timestamps) return fail(FBF_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    if (n == 0 || n >= 0xFFFFFFFFull - 64) return fail(FBF_ERR_INVALID_ARG, "n out of range");
    if (d != (uint32_t)kDim) return fail(FBF_ERR_UNSUPPORTED, "this build supports d = 100 only");
    if (repr == 0) repr = FBF_REPR_F32;
    auto *ds = new fbf_dataset();
    hipGetDevice(&ds->device);
    ds->n = n;
    ds->n_pad = (n + kRowTile - 1) / kRowTile * kRowTile;
    ds->d = d;
    ds->repr = repr;
    hipError_t e = hipMalloc(&ds->base_rows, n * d * sizeof(float));
    if (e != hipSuccess) {
        delete ds;
        return fail(FBF_ERR_DEVICE, std::string("hipMalloc base: ") + hipGetErrorString(e));
    }
    e = hipMemcpy(ds->base_rows, vecs, n * d * sizeof(float),
                  vecs_on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        fbf_dataset_destroy(ds);
        return fail(FBF_ERR_DEVICE, std::string("upload base: ") + hipGetErrorString(e));
    }
    fbf_status s = dataset_build(ds, labels, timestamps, nullptr);
    if (s != FBF_OK) {
        std::string m = g_err;
        fbf_dataset_destroy(ds);
        return fail(s, m);
    }
    *out = ds;
    return FBF_OK;
}

extern "C" fbf_status fbf_dataset_create(const float *vecs, const float *labels, const float *timestamps,
                                         uint64_t n, uint32_t d, uint32_t repr_flags, fbf_dataset **out) {
    return dataset_create_common(vecs, false, labels, timestamps, n, d, repr_flags, out);
}

extern "C" fbf_status fbf_dataset_create_device(const float *vecs_dev, const float *labels,
                                                const float *timestamps, uint64_t n, uint32_t d,
                                                uint32_t repr_flags, fbf_dataset **out) {
    return dataset_create_common(vecs_dev, true, labels, timestamps, n, d, repr_flags, out);
}

extern "C" fbf_status fbf_dataset_destroy(fbf_dataset *ds) {
    if (!ds) return FBF_OK;
    hipDeviceSynchronize();
    if (ds->base_rows) (void)hipFree(ds->base_rows);
    for (auto &L : ds->lay) L.free_all();
    for (float *p : {ds->sq8_mi, ds->sq8_dif, ds->sq8_w})
        if (p) (void)hipFree(p);
    if (ds->tb_codes) (void)hipFree(ds->tb_codes);
    for (void *p : {(void *)ds->d_lab_b, (void *)ds->d_ts_b, (void *)ds->d_ts_t})
        if (p) (void)hipFree(p);
    for (DevBuf *b : {&ds->r_keys, &ds->r_vals, &ds->r_temp, &ds->r_ranges}) b->release();
    ds->scr.release();
    for (auto &ev : ds->events) {
        if (ev.a) hipEventDestroy(ev.a);
        if (ev.b) hipEventDestroy(ev.b);
    }
    delete ds;
    return FBF_OK;
}

extern "C" fbf_status fbf_filter_count(const fbf_dataset *ds, const float *meta, uint64_t *count,
                                       float *selectivity) {
    if (!ds || !meta) return fail(FBF_ERR_INVALID_ARG, "null argument");
    QRange r;
    if (!query_range(ds, meta, r)) return fail(FBF_ERR_INVALID_ARG, "bad filter type");
    if (count) *count = r.e - r.s;
    if (selectivity) *selectivity = (float)(r.e - r.s) / (float)ds->n;
    return FBF_OK;
}

// ---------------------------------------------------------------- planner
struct Plan {
    std::vector<WorkItem> items;
    std::vector<ItemQuery> iq;
    std::vector<uint64_t> slot_off;  // [nq + 1]: query q's lists are slots [slot_off[q], slot_off[q + 1])
    std::vector<uint64_t> pre_end;   // [nq], two-pass plans: end of each query's first-pass slots
    std::vector<uint32_t> nchunk;    // [nq] chunks of the query's group (slots the query owns)
    uint64_t pairs = 0, pairs_computed = 0;
};

// Run fn(begin, end) over [0, n) in up to 16 host threads (one when n is small).
template <class Fn>
static void parallel_ranges(uint64_t n, uint64_t min_per_thread, Fn fn) {
    const uint64_t hw = std::max(1u, std::thread::hardware_concurrency());
    const uint64_t nt = std::min<uint64_t>({16, hw, (n + min_per_thread - 1) / min_per_thread});
    if (nt <= 1) {
        fn(0, n);
        return;
    }
    std::vector<std::thread> th;
    const uint64_t per = (n + nt - 1) / nt;
    for (uint64_t t = 0; t < nt; ++t) {
        const uint64_t b = t * per, e = std::min(n, b + per);
        if (b < e) th.emplace_back(fn, b, e);
    }
    for (auto &x : th) x.join();
}

// Each query's candidate range (binary searches over the sorted layouts; the
// batch is split across host threads).
static fbf_status compute_ranges(const fbf_dataset *ds, const float *metas, uint64_t nq,
                                 std::vector<QRange> &qr) {
    qr.resize(nq);
    std::atomic<uint64_t> bad{UINT64_MAX};
    parallel_ranges(nq, 4096, [&](uint64_t b, uint64_t e) {
        for (uint64_t i = b; i < e; ++i) {
            if (!query_range(ds, metas + 4 * i, qr[i])) {
                uint64_t cur = bad.load();
                while (i < cur && !bad.compare_exchange_weak(cur, i)) {
                }
            }
            qr[i].qi = (uint32_t)i;
        }
    });
    if (bad.load() != UINT64_MAX)
        return fail(FBF_ERR_INVALID_ARG, "query " + std::to_string(bad.load()) + ": bad filter type");
    return FBF_OK;
}

// Group queries and cut their spans into work items.  qr[i].qi must be i.
constexpr uint64_t kTargetItems = 32768;  // automatic chunk size aims at this many work items

// Work items from query ranges.  Queries of one layout are sorted by range
// and packed greedily into groups of at most group_max whose shared span
// exceeds the longest member's range by at most 25%; a group's span is cut
// into chunks of chunk_rows (0: chosen from the total span), one work item
// each, and every item carries all
// members of its group (a member whose range misses the chunk costs one empty
// list).  Query q owns nchunk[q] consecutive list slots, one per chunk of its
// group; finish_slots numbers them.  Host work is O(nq log nq + items).
static void make_plan(const std::vector<QRange> &qr, uint32_t chunk_rows, uint32_t group_max, Plan &P) {
    const uint64_t nq = qr.size();
    P.nchunk.assign(nq, 0);
    for (const QRange &r : qr) P.pairs += r.e - r.s;
    std::vector<uint32_t> order;
    order.reserve(nq);
    for (uint64_t i = 0; i < nq; ++i)
        if (qr[i].e > qr[i].s) order.push_back((uint32_t)i);
    std::sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
        const QRange &x = qr[a], &y = qr[b];
        if (x.layout != y.layout) return x.layout < y.layout;
        if (x.s != y.s) return x.s < y.s;
        if (x.e != y.e) return x.e < y.e;
        return a < b;
    });

    auto emit_group = [&](const uint32_t *g, uint32_t gn, uint32_t us, uint32_t ue) {
        const uint32_t b0 = us / kRowTile * kRowTile;
        const uint32_t nc = (uint32_t)((ue - b0 + (uint64_t)chunk_rows - 1) / chunk_rows);
        const uint32_t q_begin = (uint32_t)P.iq.size();
        for (uint32_t k = 0; k < gn; ++k) {
            const QRange &r = qr[g[k]];
            P.iq.push_back(ItemQuery{g[k], r.s, r.e, 0});
            P.nchunk[g[k]] = nc;
            // Rows streamed for this query: the chunks its range touches.
            const uint64_t c_lo = (r.s - b0) / chunk_rows, c_hi = (r.e - 1 - b0) / chunk_rows;
            const uint64_t lo = b0 + c_lo * chunk_rows, hi = std::min<uint64_t>(b0 + (c_hi + 1) * chunk_rows, ue);
            P.pairs_computed += (hi - lo + kRowTile - 1) / kRowTile * kRowTile;
        }
        for (uint32_t c = 0; c < nc; ++c) {
            WorkItem W{};
            W.row_begin = b0 + c * chunk_rows;
            W.row_end = (uint32_t)std::min<uint64_t>((uint64_t)W.row_begin + chunk_rows, ue);
            W.q_begin = q_begin;
            W.q_count = gn;
            W.layout = qr[g[0]].layout;
            W.chunk = c;
            P.items.push_back(W);
        }
    };

    struct Group {
        size_t g0;
        uint32_t gn, us, ue;
    };
    std::vector<Group> groups;
    size_t g0 = 0;  // current group: order[g0, i)
    uint32_t g_layout = 0, g_s = 0, g_e = 0, g_maxlen = 0;
    for (size_t i = 0; i < order.size(); ++i) {
        const QRange &r = qr[order[i]];
        bool fits = i > g0 && i - g0 < group_max && r.layout == g_layout;
        if (fits) {
            const uint32_t ns = std::min(g_s, r.s), ne = std::max(g_e, r.e);
            const uint32_t ml = std::max(g_maxlen, r.e - r.s);
            // The shared span may exceed the longest member's range by 25%.
            fits = (uint64_t)(ne - ns) * 4 <= (uint64_t)ml * 5 + 4 * kRowTile;
        }
        if (!fits) {
            if (i > g0) groups.push_back(Group{g0, (uint32_t)(i - g0), g_s, g_e});
            g0 = i;
            g_layout = r.layout;
            g_s = r.s;
            g_e = r.e;
            g_maxlen = 0;
        }
        g_s = std::min(g_s, r.s);
        g_e = std::max(g_e, r.e);
        g_maxlen = std::max(g_maxlen, r.e - r.s);
    }
    if (order.size() > g0) groups.push_back(Group{g0, (uint32_t)(order.size() - g0), g_s, g_e});
    if (chunk_rows == 0) {
        // Automatic chunk: about kTargetItems items over all groups (enough
        // independent work to fill the GPU for small batches too), 16 Ki to
        // 256 Ki rows.
        uint64_t span = 0;
        for (const Group &g : groups) span += g.ue - g.us / kRowTile * kRowTile;
        const uint64_t c = (span / kTargetItems + kRowTile - 1) / kRowTile * kRowTile;
        chunk_rows = (uint32_t)std::min<uint64_t>(std::max<uint64_t>(c, 16384), 262144);
    }
    for (const Group &g : groups) emit_group(order.data() + g.g0, g.gn, g.us, g.ue);
}

// Number the list slots of a one-pass plan: query q owns nchunk[q] slots from
// slot_off[q]; item W's list of member j is slot iq[q_begin + j].slot_base + W.chunk.
static fbf_status finish_slots(Plan &P, uint64_t nq) {
    P.slot_off.assign(nq + 1, 0);
    for (uint64_t i = 0; i < nq; ++i) P.slot_off[i + 1] = P.slot_off[i] + P.nchunk[i];
    if (P.slot_off[nq] >= 0xFFFFFFFFull) return fail(FBF_ERR_INVALID_ARG, "batch needs more than 2^32 candidate lists");
    for (ItemQuery &e : P.iq) e.slot_base = (uint32_t)P.slot_off[e.qi];
    return FBF_OK;
}

// Append plan Bp (the second pass) to plan A (the first pass) and number the
// slots: query q owns A's chunks, then B's, from slot_off[q]; pre_end[q] ends
// its first-pass slots.
static fbf_status concat_plans(Plan &A, const Plan &Bp, uint64_t nq) {
    const uint32_t iqA = (uint32_t)A.iq.size();
    for (WorkItem W : Bp.items) {
        W.q_begin += iqA;
        A.items.push_back(W);
    }
    A.iq.insert(A.iq.end(), Bp.iq.begin(), Bp.iq.end());
    A.slot_off.assign(nq + 1, 0);
    A.pre_end.assign(nq, 0);
    for (uint64_t i = 0; i < nq; ++i) {
        A.pre_end[i] = A.slot_off[i] + A.nchunk[i];
        A.slot_off[i + 1] = A.pre_end[i] + Bp.nchunk[i];
    }
    if (A.slot_off[nq] >= 0xFFFFFFFFull) return fail(FBF_ERR_INVALID_ARG, "batch needs more than 2^32 candidate lists");
    for (size_t k = 0; k < A.iq.size(); ++k) {
        ItemQuery &e = A.iq[k];
        e.slot_base = (uint32_t)(k < iqA ? A.slot_off[e.qi] : A.pre_end[e.qi]);
    }
    A.pairs += Bp.pairs;
    A.pairs_computed += Bp.pairs_computed;
    return FBF_OK;
}

// Reserve scratch for a plan with list capacity B and copy the plan to the
// device through pinned memory (asynchronous on st).
static fbf_status upload_plan(Scratch &S, const Plan &P, uint64_t nq, int B, hipStream_t st) {
    const uint64_t n_slots = P.slot_off.empty() ? 0 : P.slot_off[nq];
    HIP_TRY(S.d_items.reserve(std::max<size_t>(1, P.items.size()) * sizeof(WorkItem)));
    HIP_TRY(S.d_iq.reserve(std::max<size_t>(1, P.iq.size()) * sizeof(ItemQuery)));
    HIP_TRY(S.d_slot_off.reserve((nq + 1) * sizeof(uint64_t)));
    HIP_TRY(S.d_lists.reserve(std::max<size_t>(1, n_slots) * (size_t)B * sizeof(uint64_t)));
    HIP_TRY(S.d_counts.reserve(std::max<size_t>(1, n_slots) * sizeof(uint32_t)));
    HIP_TRY(S.d_queue.reserve(8 * 64));
    const size_t b_items = P.items.size() * sizeof(WorkItem);
    const size_t b_iq = P.iq.size() * sizeof(ItemQuery);
    const size_t b_off = (nq + 1) * sizeof(uint64_t);
    const size_t b_pe = P.pre_end.size() * sizeof(uint64_t);
    if (b_pe) HIP_TRY(S.d_pre_end.reserve(b_pe));
    // Two pinned staging buffers used in turn: a buffer is refilled only after
    // the copies that read it (queued by the search before last) are done.
    const int sb = S.stage_next;
    S.stage_next ^= 1;
    if (S.ev_stage[sb]) HIP_TRY(hipEventSynchronize(S.ev_stage[sb]));
    else HIP_TRY(hipEventCreateWithFlags(&S.ev_stage[sb], hipEventDisableTiming));
    HIP_TRY(S.h_stage[sb].reserve(b_items + b_iq + b_off + b_pe + 64));
    char *h = (char *)S.h_stage[sb].p;
    std::memcpy(h, P.items.data(), b_items);
    std::memcpy(h + b_items, P.iq.data(), b_iq);
    std::memcpy(h + b_items + b_iq, P.slot_off.data(), b_off);
    if (b_pe) std::memcpy(h + b_items + b_iq + b_off, P.pre_end.data(), b_pe);
    if (b_items) HIP_TRY(hipMemcpyAsync(S.d_items.p, h, b_items, hipMemcpyHostToDevice, st));
    if (b_iq) HIP_TRY(hipMemcpyAsync(S.d_iq.p, h + b_items, b_iq, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(S.d_slot_off.p, h + b_items + b_iq, b_off, hipMemcpyHostToDevice, st));
    if (b_pe) HIP_TRY(hipMemcpyAsync(S.d_pre_end.p, h + b_items + b_iq + b_off, b_pe, hipMemcpyHostToDevice, st));
    HIP_TRY(hipEventRecord(S.ev_stage[sb], st));
    return FBF_OK;
}

static MergeArgs merge_args(const Scratch &S, int B, uint64_t nq, uint32_t k, uint32_t kprime) {
    MergeArgs m{};
    m.slot_off = S.d_slot_off.as<uint64_t>();
    m.lists = S.d_lists.as<uint64_t>();
    m.counts = S.d_counts.as<uint32_t>();
    m.list_cap = (uint32_t)B;
    m.nq = nq;
    m.k = k;
    m.kprime = kprime;
    return m;
}

// ---------------------------------------------------------------- search
// Ranges whose prefix would be shorter than this (or than twice the keys it
// keeps) are scanned in one pass.  An unseeded range starts with an open
// threshold and compacts its lists many times while they fill -- at large
// k' (720, 896: lists of 1024 keys with room for about 128 appends between
// compactions) that dominated the scan.  C1, ids identical: the certified
// exact mode 243 -> 167 ms per 100 k queries, FP8 at k' = 896 336 -> 226 ms,
// BF16 at k' = 110 90.9 -> 88.1 ms, against a minimum of 16 Ki rows and
// eight times the kept keys.
constexpr uint64_t kMinPrefixRows = 256;
// Work items in descending cost order (largest first), grouped by layout
// when each layout is its own launch: big items start early and the small ones
// fill the tail.
static void order_items(Plan &P, bool by_layout) {
    std::stable_sort(P.items.begin(), P.items.end(), [&](const WorkItem &x, const WorkItem &y) {
        if (by_layout && x.layout != y.layout) return x.layout < y.layout;
        const uint64_t cx = (uint64_t)(x.row_end - x.row_begin) * x.q_count;
        const uint64_t cy = (uint64_t)(y.row_end - y.row_begin) * y.q_count;
        if (cx != cy) return cx > cy;
        // Equal cost: chunk-major, so the items that run at the same time on an
        // XCD (workgroups are dealt round-robin to the 8 XCDs) stream the same
        // rows through that XCD's L2 for different query groups.
        if (x.layout != y.layout) return x.layout < y.layout;
        return x.row_begin < y.row_begin;
    });
}

// Persistent MFMA scans: chunk-major order (layout, first row), so that the
// items of one XCD's run that are in flight together share their rows in L2.
static void order_items_chunk_major(Plan &P) {
    std::stable_sort(P.items.begin(), P.items.end(), [&](const WorkItem &x, const WorkItem &y) {
        if (x.layout != y.layout) return x.layout < y.layout;
        if (x.row_begin != y.row_begin) return x.row_begin < y.row_begin;
        return (uint64_t)(x.row_end - x.row_begin) * x.q_count > (uint64_t)(y.row_end - y.row_begin) * y.q_count;
    });
}

// Prefix-pass order for the persistent MFMA scans: the groups (items sharing
// q_begin) in first-row order are dealt to 8 runs of about equal cost, one per
// XCD, and each run is ordered by the items' chunk index within their group.
// A group's first chunk thus starts before its later ones, which then start
// from the thresholds it has published instead of open ones (C1 BF16 prefix
// pass 11.4 -> 11.0 ms; the prefix items of one group cover different rows,
// so chunk-major order gains little L2 sharing there).
static void order_items_prefix(Plan &P) {
    struct Grp {
        uint32_t layout, row;
        double cost;
        uint32_t run;
    };
    std::vector<Grp> gs;
    std::unordered_map<uint32_t, uint32_t> gi;  // q_begin -> group
    for (const WorkItem &w : P.items) {
        const double c = (double)(w.row_end - w.row_begin) * w.q_count;
        auto it = gi.find(w.q_begin);
        if (it == gi.end()) {
            gi.emplace(w.q_begin, (uint32_t)gs.size());
            gs.push_back({w.layout, w.row_begin, c, 0});
        } else {
            Grp &g = gs[it->second];
            g.row = std::min(g.row, w.row_begin);
            g.cost += c;
        }
    }
    std::vector<uint32_t> ord(gs.size());
    for (uint32_t i = 0; i < ord.size(); ++i) ord[i] = i;
    std::sort(ord.begin(), ord.end(), [&](uint32_t a, uint32_t b) {
        return gs[a].layout != gs[b].layout ? gs[a].layout < gs[b].layout : gs[a].row < gs[b].row;
    });
    double total = 0, acc = 0;
    for (const Grp &g : gs) total += g.cost;
    for (uint32_t i : ord) {
        gs[i].run = std::min<uint32_t>(7u, (uint32_t)((acc + 0.5 * gs[i].cost) * 8.0 / std::max(total, 1.0)));
        acc += gs[i].cost;
    }
    std::vector<uint32_t> run_of(P.items.size());
    for (size_t i = 0; i < P.items.size(); ++i) run_of[i] = gs[gi[P.items[i].q_begin]].run;
    std::vector<uint32_t> idx(P.items.size());
    for (uint32_t i = 0; i < idx.size(); ++i) idx[i] = i;
    std::stable_sort(idx.begin(), idx.end(), [&](uint32_t a, uint32_t b) {
        const WorkItem &x = P.items[a], &y = P.items[b];
        if (run_of[a] != run_of[b]) return run_of[a] < run_of[b];
        if (x.chunk != y.chunk) return x.chunk < y.chunk;
        if (x.layout != y.layout) return x.layout < y.layout;
        return x.row_begin < y.row_begin;
    });
    std::vector<WorkItem> sorted(P.items.size());
    for (size_t i = 0; i < idx.size(); ++i) sorted[i] = P.items[idx[i]];
    P.items.swap(sorted);
}

// Split items [i0, i1) into 8 contiguous runs of about equal cost (rows x
// queries; equal counts when by_cost is false), one per XCD.
static void xcd_runs(const std::vector<WorkItem> &items, uint64_t i0, uint64_t i1, bool by_cost, uint32_t xoff[9]) {
    auto cost = [&](uint64_t i) -> double {
        return by_cost ? (double)(items[i].row_end - items[i].row_begin) * items[i].q_count : 1.0;
    };
    double total = 0;
    for (uint64_t i = i0; i < i1; ++i) total += cost(i);
    xoff[0] = 0;
    double acc = 0;
    uint64_t i = i0;
    for (int x = 1; x < 8; ++x) {
        const double target = total * x / 8;
        while (i < i1 && acc + 0.5 * cost(i) < target) acc += cost(i++);
        xoff[x] = (uint32_t)(i - i0);
    }
    xoff[8] = (uint32_t)(i1 - i0);
}

static fbf_status mode_check(const fbf_dataset *ds, uint32_t mode) {
    static const uint32_t need[6] = {FBF_REPR_F32, FBF_REPR_SQ8,    FBF_REPR_BF16,
                                     FBF_REPR_FP8, FBF_REPR_SQ8SYM, FBF_REPR_BF16 | FBF_REPR_F32};
    if (mode > FBF_MODE_EXACT_BF16) return fail(FBF_ERR_INVALID_ARG, "unknown mode");
    if ((ds->repr & need[mode]) != need[mode])
        return fail(FBF_ERR_NO_REPR, "dataset was created without this mode's representation");
    return FBF_OK;
}

// A search's device work captured as one HIP graph (fbf_search_graph_create).
struct GraphOut {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

// One search.  S: the scratch its plan and lists live in (the dataset's own,
// shared by its searches, or a captured graph's).  cap: capture the device
// work (everything after the plan upload) into a graph instead of running it.
static fbf_status search_impl(fbf_dataset *ds, Scratch &S, bool shared_scratch, const float *queries_dev,
                              const float *metas, uint64_t nq, uint32_t k, const fbf_search_opts *opts,
                              uint32_t *out_ids_dev, float *out_dists_dev, hipStream_t st, GraphOut *cap,
                              uint32_t *cert = nullptr, const QRange *ranges_in = nullptr) {
    if (!ds || (!metas && !ranges_in) || (nq && (!queries_dev || !out_ids_dev || !out_dists_dev)))
        return fail(FBF_ERR_INVALID_ARG, "null argument");
    if (k == 0 || k > 512) return fail(FBF_ERR_INVALID_ARG, "k must be in [1, 512]");
    if (nq == 0) return FBF_OK;
    fbf_search_opts o{};
    if (opts) o = *opts;
    fbf_status s = mode_check(ds, o.mode);
    if (s != FBF_OK) return s;
    if (o.mode == FBF_MODE_EXACT_BF16)
        return fail(FBF_ERR_INVALID_ARG, "FBF_MODE_EXACT_BF16 waits for the device between its steps; it cannot be captured");
    const bool exact = o.mode == FBF_MODE_EXACT_F32;
    const bool mfma = o.mode == FBF_MODE_BF16 || o.mode == FBF_MODE_FP8 || o.mode == FBF_MODE_SQ8 ||
                      o.mode == FBF_MODE_SQ8SYM;
    // Kept depth per list: k for the exact scan, k' (over-fetch) otherwise.
    uint32_t K = k;
    if (!exact) {
        // Default over-fetch: 3k/2 for the BF16 and SQ8 scans (their score
        // errors are small against the gaps between neighbours), 4k for the
        // symmetric int8 codes, 9k (at most 896) for FP8: e4m3 keeps three
        // mantissa bits, and on the 10 M-row Gaussian-mixture workload its
        // score errors reach the gap between the 100th and the 400th
        // neighbour (recall@100 0.957 at k' = 400, 0.994 at 896).
        const bool fine = o.mode == FBF_MODE_BF16 || o.mode == FBF_MODE_SQ8;
        K = o.kprime ? o.kprime
                     : std::min<uint32_t>(fine ? (3 * k + 1) / 2 : o.mode == FBF_MODE_FP8 ? 9 * k : 4 * k, 896);
        if (K < k) return fail(FBF_ERR_INVALID_ARG, "kprime must be >= k");
        if (K > 896) return fail(FBF_ERR_INVALID_ARG, "kprime must be <= 896");
    }
    // Rows per work item (0: the planner picks it from the batch).
    const uint32_t chunk = o.chunk_rows ? (o.chunk_rows + kRowTile - 1) / kRowTile * kRowTile : 0u;
    // MFMA items carry up to 256 queries, 256 by default (two waves of four
    // 32-query blocks share each row tile); VALU items up to 16.
    const uint32_t gcap = mfma ? 256u : (uint32_t)kMaxGroup;
    const uint32_t gmax = o.group_queries ? std::min<uint32_t>(o.group_queries, gcap) : mfma ? 256u : gcap;
    // Prefix-pass items carry at most 128 queries (two waves per tile): the
    // prefix lists start empty, so its waves compact often, and every
    // compaction holds up the other waves that share the tile ring.
    const uint32_t gmax_pre = std::min<uint32_t>(gmax, 128u);
    // MFMA item width in units of kMfmaWaveQueries queries.
    auto waves_of = [&](uint32_t g) {
        return !mfma ? 1 : g <= kMfmaWaveQueries ? 1 : g <= 2 * kMfmaWaveQueries ? 2 : 4;
    };
    if (o.persistent > 2) return fail(FBF_ERR_INVALID_ARG, "persistent must be 0, 1 or 2");
    const bool persist = o.persistent == 1 || (o.persistent == 0 && mfma);
    // Item order: chunk-major runs per XCD for the persistent MFMA scans, else
    // largest first.
    auto order = [&](Plan &Q) {
        if (mfma && persist) order_items_chunk_major(Q);
        else order_items(Q, !mfma);
    };
    std::lock_guard<std::mutex> lock(ds->mu);
    HIP_TRY(hipSetDevice(ds->device));
    const auto t_plan0 = std::chrono::steady_clock::now();
    // The dataset's scratch is reused by every search: this search's stream
    // first waits for the device work of the previous one (whatever stream
    // that ran on), and this search records the same event when its own work
    // is queued (on every return path below, failures included).
    struct DoneRecord {
        hipEvent_t ev = nullptr;
        hipStream_t st = nullptr;
        ~DoneRecord() {
            if (ev) (void)hipEventRecord(ev, st);
        }
    } done_rec;
    if (shared_scratch) {
        if (!S.ev_done) HIP_TRY(hipEventCreateWithFlags(&S.ev_done, hipEventDisableTiming));
        else if (S.used) HIP_TRY(hipStreamWaitEvent(st, S.ev_done, 0));
        S.used = true;
        done_rec.ev = S.ev_done;
        done_rec.st = st;
    }
    // Profiling events and the rescan count are host-side records of a run:
    // neither goes into a captured graph.
    const bool prof = ds->profile && cap == nullptr;
    std::vector<QRange> qr;
    if (ranges_in) {
        qr.assign(ranges_in, ranges_in + nq);
        for (uint64_t i = 0; i < nq; ++i) qr[i].qi = (uint32_t)i;
    } else {
        s = compute_ranges(ds, metas, nq, qr);
        if (s != FBF_OK) return s;
    }
    // Approximate modes: a prefix pass over the first 1/prepass_div of every
    // large range keeps only the kpre = margin * k' / prepass_div best keys of
    // that prefix; its kpre-th key (tau0) estimates the k'-th key of the whole
    // range and seeds the threshold of the main pass over the rest, so the main
    // pass appends about margin * k' candidates per query instead of the many
    // thousands an unseeded scan appends while its lists fill.  The merge
    // verifies each seeded query (exact when its k'-th merged key is <= tau0)
    // and the few that fail are rescanned over their whole range.
    // Defaults 1/16 and a margin that shrinks as k' grows: the seed estimates
    // the (margin * k')-th key from the kpre-th key of a 1/16 sample, whose
    // relative spread falls as 1/sqrt(kpre).  C1: BF16 k' = 110 at 300%
    // 88.3 ms per step (89.9 with 1/10 and 250%; 1/24 or 1/32 let seeds
    // fail); k' = 400 at 200% 104.4 ms (300%: 111.0; 150%: seeds fail, 132.7);
    // FP8 k' = 896 at 150% 148.5 ms (200%: 173.4; 300%: 219.9).
    const uint32_t pdiv = mfma ? (o.prepass_div ? o.prepass_div : 16u) : 1u;
    const uint32_t margin = o.prepass_margin ? o.prepass_margin : K <= 160 ? 300u : K <= 500 ? 200u : 150u;
    const uint32_t kpre = (uint32_t)std::min<uint64_t>(K, std::max<uint64_t>(1, ((uint64_t)K * margin + 100ull * pdiv - 1) / (100ull * pdiv)));
    Plan P;
    uint64_t n_items_a = 0;
    bool seeded = false;
    if (pdiv > 1) {
        std::vector<QRange> qa = qr, qb = qr;  // prefixes (empty when not split), the rest
        const uint64_t min_prefix = std::max<uint64_t>(kMinPrefixRows, 2ull * kpre);
        uint64_t max_prefix = 0;
        for (uint64_t i = 0; i < nq; ++i) {
            const uint64_t len = qr[i].e - qr[i].s;
            const uint64_t pre = (len + pdiv - 1) / pdiv;
            if (pre >= min_prefix) {
                qa[i].e = qr[i].s + (uint32_t)pre;
                qb[i].s = qa[i].e;
                max_prefix = std::max<uint64_t>(max_prefix, pre);
            } else {
                qa[i].e = qa[i].s;
            }
        }
        seeded = max_prefix > 0;
        if (seeded) {
            // Prefix items are ordinary chunks: the prefix merge takes the
            // kpre-th key of the union of their lists (the exact kpre-th key
            // of the whole prefix: every key up to it is in some list).
            Plan PB;
            std::thread tb([&] {
                make_plan(qb, chunk, gmax, PB);
                order(PB);
            });
            make_plan(qa, chunk, gmax_pre, P);
            if (mfma && persist) order_items_prefix(P);
            else order(P);
            tb.join();
            n_items_a = P.items.size();
            ds->last_pairs_prefix = P.pairs;
            s = concat_plans(P, PB, nq);
            if (s != FBF_OK) return s;
        }
    }
    if (!seeded) {
        ds->last_pairs_prefix = 0;
        make_plan(qr, chunk, gmax, P);
        order(P);
        n_items_a = 0;
        s = finish_slots(P, nq);
        if (s != FBF_OK) return s;
    }
    ds->last_pairs = P.pairs;
    ds->last_pairs_computed = P.pairs_computed;
    ds->last_items = P.items.size();

    // MFMA scans: room for about one K of appends between compactions once K
    // is large (a scan whose threshold is still open then compacts rarely);
    // each of the two halves a lane appends to keeps room above ceil(K / 2) + 16.
    const int B = mfma ? (K <= 160 ? 256 : K <= 300 ? 512 : 1024) : list_capacity((int)K);
    // The prefix pass keeps only kpre keys: its lists use the capacity the
    // rule above gives for kpre (128 below 64 keys), at most their B slots,
    // so each compaction sorts fewer keys and the threshold tightens after
    // fewer appends.
    const int B_pre = !seeded ? B : std::min(B, kpre <= 64 ? 128 : kpre <= 160 ? 256 : kpre <= 300 ? 512 : 1024);
    s = upload_plan(S, P, nq, B, st);
    if (s != FBF_OK) return s;
    if (prof) {
        ProfEvent ev;
        ev.kind = 4;
        ev.host_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_plan0).count();
        ds->events.push_back(ev);
    }
    uint32_t *queue = nullptr;
    if (persist) queue = S.d_queue.as<uint32_t>();
    // Every device buffer of the device work is reserved here, before it:
    // a captured graph may not allocate.
    HIP_TRY(S.d_gtau.reserve(nq * sizeof(uint64_t)));
    uint64_t *gtau = S.d_gtau.as<uint64_t>();
    if (seeded) {
        HIP_TRY(S.d_tau0.reserve(nq * sizeof(uint64_t)));
        HIP_TRY(S.d_flags.reserve((nq + 1) * sizeof(uint32_t)));
        HIP_TRY(S.h_nflag.reserve(sizeof(uint32_t)));
        if (!S.ev_nflag) HIP_TRY(hipEventCreateWithFlags(&S.ev_nflag, hipEventDisableTiming));
    }
    if (o.mode == FBF_MODE_SQ8SYM) HIP_TRY(S.d_q.reserve(nq * 112));
    if (mfma) HIP_TRY(S.d_qf.reserve(nq * kQFragU4 * sizeof(uint4)));
    ds->last_seeded = false;
    if (cap) {
        // The plan is on the device once the stream drains; the launches
        // below go into the graph.
        if (mfma) {
            const int v = o.mode == FBF_MODE_BF16 ? 0 : o.mode == FBF_MODE_FP8 ? 1 : o.mode == FBF_MODE_SQ8 ? 2 : 3;
            warm_scan_mfma(v, B, waves_of(gmax));
            if (seeded) warm_scan_mfma(v, B_pre, waves_of(gmax_pre));
        }
        HIP_TRY(hipStreamSynchronize(st));
        HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    }
    fbf_status sw = [&]() -> fbf_status {
    fbf_status s = FBF_OK;
    // Per-query thresholds shared by all work items of a query.
    HIP_TRY(launch_fill_u32(reinterpret_cast<uint32_t *>(gtau), 0xFFFFFFFFu, 2 * nq, st));

    // Merge (and rerank) of the per-item lists into the outputs.
    auto make_merge = [&]() {
        MergeArgs m = merge_args(S, B, nq, k, K);
        m.queries = queries_dev;
        m.base_rows = ds->base_rows;
        m.out_ids = out_ids_dev;
        m.out_dists = out_dists_dev;
        m.rerank = exact ? 0 : 1;
        m.key_int = 0;
        m.key_pos = mfma ? 1 : 0;  // the MFMA scans key candidates by layout position
        m.ids_lay[0] = ds->lay[0].ids;
        m.ids_lay[1] = ds->lay[1].ids;
        m.cert = cert;
        m.cert_norm = ds->max_norm;
        m.cert_rho = ds->bf16_rho;
        return m;
    };
    auto run_merge = [&](MergeArgs &m) -> fbf_status {
        ProfEvent ev;
        if (prof) record_begin(ds, exact ? 1 : 2, st, ev);
        HIP_TRY(launch_merge(m, st));
        if (prof) record_end(ds, st, ev);
        return FBF_OK;
    };

    // Scan.  Each work item writes every list slot it owns (count and keys),
    // so no clearing is needed.
    if (mfma) {
        const int variant = o.mode == FBF_MODE_BF16 ? 0 : o.mode == FBF_MODE_FP8 ? 1 : o.mode == FBF_MODE_SQ8 ? 2 : 3;
        ScanMfmaArgs a{};
        // Query codes with the dataset's scale (SQ8SymmetricQuantizer::encode).
        if (variant == 3) HIP_TRY(launch_sym_encode(queries_dev, nullptr, nq, nq, ds->alpha, S.d_q.as<int8_t>(), st));
        HIP_TRY(launch_qfrag(variant, queries_dev, ds->sq8_w, variant == 3 ? S.d_q.as<int8_t>() : nullptr, nq,
                             S.d_qf.as<uint4>(), st));
        a.items = S.d_items.as<WorkItem>();
        a.n_items = P.items.size();
        a.iq = S.d_iq.as<ItemQuery>();
        a.qfrag = S.d_qf.as<uint4>();
        for (int L = 0; L < 2; ++L) {
            const LayoutDev &D = ds->lay[L];
            a.lay[L].ids = D.ids;
            a.lay[L].tiles = variant == 0 ? D.bf16 : variant == 1 ? D.fp8 : variant == 2 ? D.sq8 : D.sym_mfma;
            a.lay[L].xnorm = variant == 1 ? D.xnorm : variant == 2 ? D.sq8_norm : D.sym_norm;
            a.lay[L].scale = D.fp8_scale;
        }
        a.lists = S.d_lists.as<uint64_t>();
        a.counts = S.d_counts.as<uint32_t>();
        a.gtau = gtau;
        a.list_stride = (uint32_t)B;
        // Item width: enough for the largest item of the launch (a wave with
        // no queries of an item only helps load its rows).
        auto scan = [&](uint64_t i0, uint64_t i1, uint32_t kk, const uint32_t *mask) -> fbf_status {
            const bool pre = seeded && i1 <= n_items_a;
            const int nw = waves_of(pre ? gmax_pre : gmax);
            if (i1 == i0) return FBF_OK;
            a.items = S.d_items.as<WorkItem>() + i0;
            a.n_items = i1 - i0;
            a.K = kk;
            a.qmask = mask;
            a.nmask = mask ? mask + nq : nullptr;  // the verify merge counts the flags there
            a.queue = queue;
            if (queue) {
                xcd_runs(P.items, i0, i1, mask == nullptr, a.xoff);
                HIP_TRY(launch_fill_u32(queue, 0u, 8 * 16, st));
            }
            ProfEvent ev;
            ev.pass = pre ? 1 : mask ? 2 : 0;
            if (prof) record_begin(ds, 0, st, ev);
            HIP_TRY(launch_scan_mfma(variant, a, pre ? B_pre : B, nw, st));
            if (prof) record_end(ds, st, ev);
            return FBF_OK;
        };
        if (seeded) {
            // Prefix pass (kpre kept), seed snapshot, main pass (k' kept).
            HIP_TRY(S.d_tau0.reserve(nq * sizeof(uint64_t)));
            HIP_TRY(S.d_flags.reserve((nq + 1) * sizeof(uint32_t)));
            HIP_TRY(launch_fill_u32(S.d_flags.as<uint32_t>(), 0u, nq + 1, st));
            s = scan(0, n_items_a, kpre, nullptr);
            if (s != FBF_OK) return s;
            uint32_t *flags = S.d_flags.as<uint32_t>();
            MergeArgs m = make_merge();
            m.tau0 = S.d_tau0.as<uint64_t>();
            m.flags = flags;
            m.gtau = gtau;
            m.nflag = flags + nq;
            m.pre_end = S.d_pre_end.as<uint64_t>();
            // Seeds: the kpre-th key of each query's prefix lists.
            m.pass = 3;
            m.kprime = kpre;
            s = run_merge(m);
            if (s != FBF_OK) return s;
            m.kprime = K;
            s = scan(n_items_a, P.items.size(), K, nullptr);
            if (s != FBF_OK) return s;
            // Verify each seeded query; rescan the failures over their whole
            // range (prefix items now keep k' too) and merge those again.
            m.pass = 1;
            s = run_merge(m);
            if (s != FBF_OK) return s;
            if (!cap) {
                HIP_TRY(hipMemcpyAsync(S.h_nflag.p, flags + nq, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
                HIP_TRY(hipEventRecord(S.ev_nflag, st));
            }
            s = scan(0, P.items.size(), K, flags);
            if (s != FBF_OK) return s;
            m.pass = 2;
            ds->last_seeded = !cap && shared_scratch;
            return run_merge(m);
        }
        s = scan(0, P.items.size(), K, nullptr);
        if (s != FBF_OK) return s;
    } else {
        // Exact FP32 scan.  Runs of equal layout in item order (per pass: each pass is sorted by layout).
        std::vector<std::pair<uint64_t, uint64_t>> runs;
        for (uint64_t i = 0; i < P.items.size();) {
            uint64_t j = i + 1;
            while (j < P.items.size() && P.items[j].layout == P.items[i].layout && j != n_items_a) ++j;
            runs.emplace_back(i, j);
            i = j;
        }
        for (auto &run : runs) {
            const uint64_t i0 = run.first, i1 = run.second;
            const uint32_t L = P.items[i0].layout;
            if (queue) HIP_TRY(launch_fill_u32(queue, 0u, 1, st));
            ProfEvent ev;
            if (prof) record_begin(ds, 0, st, ev);
            ScanArgs a{};
            a.items = S.d_items.as<WorkItem>() + i0;
            a.n_items = i1 - i0;
            a.iq = S.d_iq.as<ItemQuery>();
            a.queries = queries_dev;
            a.soa = ds->lay[L].soa;
            a.ids = ds->lay[L].ids;
            a.lists = S.d_lists.as<uint64_t>();
            a.counts = S.d_counts.as<uint32_t>();
            a.gtau = gtau;
            a.K = K;
            a.queue = queue;
            HIP_TRY(launch_scan_f32(a, B, st));
            if (prof) record_end(ds, st, ev);
        }
    }

    MergeArgs m = make_merge();
    return run_merge(m);
    }();
    if (cap) {
        // End the capture on every path (a stream left capturing is unusable).
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(st, &g);
        if (sw != FBF_OK) {
            if (g) (void)hipGraphDestroy(g);
            return sw;
        }
        if (e != hipSuccess || g == nullptr) {
            if (g) (void)hipGraphDestroy(g);
            return fail(FBF_ERR_DEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
        }
        hipGraphExec_t x = nullptr;
        const hipError_t ei = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
        if (ei != hipSuccess) {
            (void)hipGraphDestroy(g);
            return fail(FBF_ERR_DEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
        }
        cap->graph = g;
        cap->exec = x;
    }
    return sw;
}

// FBF_MODE_EXACT_BF16: the BF16 search with a certificate per query, then the
// FP32 scan over the uncertified queries only, whose rows replace theirs.
// Certified rows equal the exact search's (same candidates, canonical FP32
// distances, (distance, id) order), so the whole output does.
static fbf_status search_exact_bf16(fbf_dataset *ds, const float *queries_dev, const float *metas, uint64_t nq,
                                    uint32_t k, const fbf_search_opts &o, uint32_t *out_ids_dev,
                                    float *out_dists_dev, hipStream_t st, const QRange *ranges_in = nullptr) {
    fbf_status s = mode_check(ds, FBF_MODE_EXACT_BF16);
    if (s != FBF_OK) return s;
    if ((!metas && !ranges_in) || (nq && (!queries_dev || !out_ids_dev || !out_dists_dev)))
        return fail(FBF_ERR_INVALID_ARG, "null argument");
    if (k == 0 || k > 512) return fail(FBF_ERR_INVALID_ARG, "k must be in [1, 512]");
    std::lock_guard<std::mutex> lk(ds->mu_cert);
    if (nq == 0) return FBF_OK;
    HIP_TRY(hipSetDevice(ds->device));
    fbf_search_opts ob = o;
    ob.mode = FBF_MODE_BF16;
    // k' = 720: on 10 M rows x 100 k queries all but ~0.1% of the queries
    // are certified; a larger k' leaves the 1024-key lists too little room
    // above k' / 2 between compactions, a smaller one certifies fewer.
    if (ob.kprime == 0) ob.kprime = 720;
    // At this k' a prefix over 1/16 of each range with a 200% margin is
    // faster than the general defaults (C1: 143 ms per step against 162 with
    // 300%, 148 with 1/24 and 250%).
    if (ob.prepass_div == 0) ob.prepass_div = 16;
    if (ob.prepass_margin == 0) ob.prepass_margin = 200;
    // Items of 256 queries over 64 Ki rows: the groups of one chunk share
    // its rows in L2 (C1: 243 ms per 100 k queries, against 258 ms with
    // 256 Ki-row items of the default group size and 277 ms with 128 Ki);
    // the lists of up to 1024 keys then take slots x 8 KiB, about 40 GB for
    // 100 k queries over 10 M rows.  Sub-batches (max_batch) bound that
    // memory at some speed (50 k queries: 333 ms).
    if (ob.chunk_rows == 0) ob.chunk_rows = 65536;
    if (ob.group_queries == 0) ob.group_queries = 256;
    Scratch &S = ds->scr;
    HIP_TRY(S.d_cert.reserve(nq * sizeof(uint32_t)));
    s = search_impl(ds, S, true, queries_dev, metas, nq, k, &ob, out_ids_dev, out_dists_dev, st, nullptr,
                    S.d_cert.as<uint32_t>(), ranges_in);
    if (s != FBF_OK) return s;
    // The search statistics stay those of the BF16 search (the fallback is
    // counted by last_fallbacks).
    uint64_t pairs = 0, pairs_c = 0, items = 0, pairs_pre = 0;
    bool seeded = false;
    {
        std::lock_guard<std::mutex> lock(ds->mu);
        pairs_pre = ds->last_pairs_prefix;
        pairs = ds->last_pairs;
        pairs_c = ds->last_pairs_computed;
        items = ds->last_items;
        seeded = ds->last_seeded;
    }
    std::vector<uint32_t> hc(nq);
    HIP_TRY(hipMemcpyAsync(hc.data(), S.d_cert.p, nq * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    std::vector<uint32_t> sub;
    for (uint64_t i = 0; i < nq; ++i)
        if (hc[i] == 0) sub.push_back((uint32_t)i);
    if (!sub.empty()) {
        const uint64_t m = sub.size();
        HIP_TRY(S.d_sub_idx.reserve(m * sizeof(uint32_t)));
        HIP_TRY(S.d_sub_q.reserve(m * kDim * sizeof(float)));
        HIP_TRY(S.d_sub_ids.reserve(m * k * sizeof(uint32_t)));
        HIP_TRY(S.d_sub_d.reserve(m * k * sizeof(float)));
        std::vector<float> msub;
        std::vector<QRange> rsub;
        if (ranges_in) {
            rsub.resize(m);
            for (uint64_t j = 0; j < m; ++j) rsub[j] = ranges_in[sub[j]];
        } else {
            msub.resize(m * 4);
            for (uint64_t j = 0; j < m; ++j) std::memcpy(&msub[4 * j], metas + 4 * (uint64_t)sub[j], 4 * sizeof(float));
        }
        HIP_TRY(hipMemcpyAsync(S.d_sub_idx.p, sub.data(), m * sizeof(uint32_t), hipMemcpyHostToDevice, st));
        HIP_TRY(launch_gather_rows(queries_dev, S.d_sub_idx.as<uint32_t>(), m, S.d_sub_q.as<float>(), st));
        fbf_search_opts oe{};
        oe.mode = FBF_MODE_EXACT_F32;
        s = search_impl(ds, S, true, S.d_sub_q.as<float>(), ranges_in ? nullptr : msub.data(), m, k, &oe,
                        S.d_sub_ids.as<uint32_t>(), S.d_sub_d.as<float>(), st, nullptr, nullptr,
                        ranges_in ? rsub.data() : nullptr);
        if (s != FBF_OK) return s;
        HIP_TRY(launch_scatter_results(S.d_sub_ids.as<uint32_t>(), S.d_sub_d.as<float>(), S.d_sub_idx.as<uint32_t>(),
                                       m, k, out_ids_dev, out_dists_dev, st));
        // The subset buffers are reused by the next call: let this one finish.
        HIP_TRY(hipStreamSynchronize(st));
    }
    std::lock_guard<std::mutex> lock(ds->mu);
    ds->last_fallbacks = sub.size();
    ds->last_pairs_prefix = pairs_pre;
    ds->last_pairs = pairs;
    ds->last_pairs_computed = pairs_c;
    ds->last_items = items;
    ds->last_seeded = seeded;
    return FBF_OK;
}

// The device router for a batch: ranges (and, when order is not null, the
// ReadSortedQuery order) on stream st into ds->r_ranges; the caller holds
// ds->mu_route.
static fbf_status route_device(fbf_dataset *ds, const float *metas_dev, uint64_t nq, float bf_thr, float bf_thr3,
                               uint32_t *order_dev, hipStream_t st) {
    HIP_TRY(ds->r_ranges.reserve(std::max<uint64_t>(1, nq) * 4 * sizeof(uint32_t)));
    HIP_TRY(ds->r_keys.reserve(std::max<uint64_t>(1, nq) * 4 * sizeof(uint64_t)));
    HIP_TRY(ds->r_vals.reserve(std::max<uint64_t>(1, nq) * 2 * sizeof(uint32_t)));
    const size_t tb = route_temp_bytes(nq);
    HIP_TRY(ds->r_temp.reserve(std::max<size_t>(1, tb)));
    HIP_TRY(launch_route(metas_dev, nq, ds->d_lab_b, ds->d_ts_b, ds->d_ts_t, ds->n, (float)ds->n, bf_thr, bf_thr3,
                         ds->r_ranges.as<uint32_t>(), order_dev, ds->r_keys.as<uint64_t>(), ds->r_vals.as<uint32_t>(),
                         ds->r_temp.p, tb, st));
    return FBF_OK;
}

// Ranges of a batch whose metas are in device memory: routed on the device,
// copied to the host for the planner.
static fbf_status device_ranges(fbf_dataset *ds, const float *metas_dev, uint64_t nq, hipStream_t st,
                                std::vector<QRange> &qr) {
    std::vector<uint32_t> h(4 * nq);
    {
        std::lock_guard<std::mutex> lk(ds->mu_route);
        HIP_TRY(hipSetDevice(ds->device));
        fbf_status s = route_device(ds, metas_dev, nq, 0.045f, 0.08f, nullptr, st);
        if (s != FBF_OK) return s;
        HIP_TRY(hipMemcpyAsync(h.data(), ds->r_ranges.p, 4 * nq * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
    }
    qr.resize(nq);
    for (uint64_t i = 0; i < nq; ++i) {
        if (h[4 * i] == 0xFFFFFFFFu) return fail(FBF_ERR_INVALID_ARG, "query " + std::to_string(i) + ": bad filter type");
        qr[i] = QRange{h[4 * i], h[4 * i + 1], h[4 * i + 2], (uint32_t)i};
    }
    return FBF_OK;
}

extern "C" fbf_status fbf_route(fbf_dataset *ds, const float *metas_dev, uint64_t nq, float bf_thr, float bf_thr3,
                                uint32_t *ranges_dev, uint32_t *order_dev, void *stream) {
    if (!ds || (nq && (!metas_dev || !ranges_dev))) return fail(FBF_ERR_INVALID_ARG, "null argument");
    if (nq >= 0xFFFFFFFFull) return fail(FBF_ERR_INVALID_ARG, "batch too large");
    if (nq == 0) return FBF_OK;
    if (bf_thr == 0.f) bf_thr = 0.045f;   // baseline.cpp:77
    if (bf_thr3 == 0.f) bf_thr3 = 0.08f;  // hybrid_graph.h:79
    hipStream_t st = (hipStream_t)stream;
    std::lock_guard<std::mutex> lk(ds->mu_route);
    HIP_TRY(hipSetDevice(ds->device));
    fbf_status s = route_device(ds, metas_dev, nq, bf_thr, bf_thr3, order_dev, st);
    if (s != FBF_OK) return s;
    HIP_TRY(hipMemcpyAsync(ranges_dev, ds->r_ranges.p, 4 * nq * sizeof(uint32_t), hipMemcpyDeviceToDevice, st));
    // the router's buffers are reused by the next call: let this one finish
    HIP_TRY(hipStreamSynchronize(st));
    return FBF_OK;
}

extern "C" fbf_status fbf_search(fbf_dataset *ds, const float *queries_dev, const float *metas,
                                 uint64_t nq, uint32_t k, const fbf_search_opts *opts,
                                 uint32_t *out_ids_dev, float *out_dists_dev, void *stream) {
    if (!ds) return fail(FBF_ERR_INVALID_ARG, "null argument");
    if (opts && opts->metas_on_device) {
        // metas in device memory: the device router gives the ranges, the
        // host plans from them (searches in sub-batches as below).
        if (!metas || (nq && (!queries_dev || !out_ids_dev || !out_dists_dev)))
            return fail(FBF_ERR_INVALID_ARG, "null argument");
        if (k == 0 || k > 512) return fail(FBF_ERR_INVALID_ARG, "k must be in [1, 512]");
        if (nq == 0) return FBF_OK;
        std::vector<QRange> qr;
        fbf_status s = device_ranges(ds, metas, nq, (hipStream_t)stream, qr);
        if (s != FBF_OK) return s;
        const uint64_t max_batch = opts->max_batch ? opts->max_batch : 131072u;
        uint64_t fallbacks = 0;  // summed over the sub-batches
        for (uint64_t i0 = 0; i0 < nq; i0 += max_batch) {
            const uint64_t m = std::min<uint64_t>(max_batch, nq - i0);
            if (opts->mode == FBF_MODE_EXACT_BF16) {
                s = search_exact_bf16(ds, queries_dev + i0 * ds->d, nullptr, m, k, *opts, out_ids_dev + i0 * k,
                                      out_dists_dev + i0 * k, (hipStream_t)stream, qr.data() + i0);
                if (s == FBF_OK) {
                    std::lock_guard<std::mutex> lock(ds->mu);
                    fallbacks += ds->last_fallbacks;
                    ds->last_fallbacks = fallbacks;
                }
            } else {
                {
                    std::lock_guard<std::mutex> lock(ds->mu);
                    ds->last_fallbacks = 0;
                }
                s = search_impl(ds, ds->scr, true, queries_dev + i0 * ds->d, nullptr, m, k, opts, out_ids_dev + i0 * k,
                                out_dists_dev + i0 * k, (hipStream_t)stream, nullptr, nullptr, qr.data() + i0);
            }
            if (s != FBF_OK) return s;
        }
        return FBF_OK;
    }
    // Very large batches run as consecutive sub-batches (bounded list memory).
    const uint64_t max_batch = opts && opts->max_batch ? opts->max_batch : 131072u;
    if (nq > max_batch) {
        if (!metas || !queries_dev || !out_ids_dev || !out_dists_dev) return fail(FBF_ERR_INVALID_ARG, "null argument");
        uint64_t fallbacks = 0;  // summed over the sub-batches
        for (uint64_t i0 = 0; i0 < nq; i0 += max_batch) {
            const uint64_t m = std::min<uint64_t>(max_batch, nq - i0);
            const fbf_status s = fbf_search(ds, queries_dev + i0 * ds->d, metas + 4 * i0, m, k, opts,
                                            out_ids_dev + i0 * k, out_dists_dev + i0 * k, stream);
            if (s != FBF_OK) return s;
            std::lock_guard<std::mutex> lock(ds->mu);
            fallbacks += ds->last_fallbacks;
            ds->last_fallbacks = fallbacks;
        }
        return FBF_OK;
    }
    if (opts && opts->mode == FBF_MODE_EXACT_BF16)
        return search_exact_bf16(ds, queries_dev, metas, nq, k, *opts, out_ids_dev, out_dists_dev, (hipStream_t)stream);
    {
        std::lock_guard<std::mutex> lock(ds->mu);
        ds->last_fallbacks = 0;
    }
    return search_impl(ds, ds->scr, true, queries_dev, metas, nq, k, opts, out_ids_dev, out_dists_dev,
                       (hipStream_t)stream, nullptr);
}

// ---------------------------------------------------------------- captured searches
// Buffers of a plan built on the device (any-batch graphs).
struct DPlanBufs {
    DevBuf ranges, rkeys, rvals, keys, skeys, vals, order, hist, hoff, part, poff, gd, gnc, qnc, ioff_pre,
        ioff_main, items_tmp_pre, items_tmp_main, items_pre, items_main, islot_pre, islot_main, bins, bcost, boff,
        iq_pre, iq_main, n_items, xoff, status;
    void release() {
        for (DevBuf *b : {&ranges, &rkeys, &rvals, &keys, &skeys, &vals, &order, &hist, &hoff, &part, &poff, &gd,
                          &gnc, &qnc, &ioff_pre, &ioff_main, &items_tmp_pre, &items_tmp_main, &items_pre,
                          &items_main, &islot_pre, &islot_main, &bins, &bcost, &boff, &iq_pre, &iq_main, &n_items,
                          &xoff, &status})
            b->release();
    }
};

struct fbf_search_graph {
    fbf_dataset *ds = nullptr;
    Scratch scr;      // the graph's own plan, lists and thresholds
    GraphOut g;
    uint64_t nq = 0;
    bool any_batch = false;  // plans on the device from metas read at each replay
    DPlanBufs dp;
};

// Over-fetch depth of an approximate search (0 in o.kprime: the mode's default).
static fbf_status kprime_of(const fbf_search_opts &o, uint32_t k, uint32_t &K) {
    const bool fine = o.mode == FBF_MODE_BF16 || o.mode == FBF_MODE_SQ8;
    K = o.kprime ? o.kprime : std::min<uint32_t>(fine ? (3 * k + 1) / 2 : o.mode == FBF_MODE_FP8 ? 9 * k : 4 * k, 896);
    if (K < k) return fail(FBF_ERR_INVALID_ARG, "kprime must be >= k");
    if (K > 896) return fail(FBF_ERR_INVALID_ARG, "kprime must be <= 896");
    return FBF_OK;
}

// Candidate-list memory of a graph that plans on the device: room for every
// query's lists at the worst plan, at most this many bytes or three quarters
// of the device memory free at creation, whichever is more.
constexpr uint64_t kAnyBatchListBytes = 32ull << 30;

// An any-batch graph: the router kernel, the device planner (fbf_dplan.hip)
// and the seeded MFMA search captured together, so each replay reads the
// filters (metas, device memory) and the vectors of that moment.  Buffers are
// sized at creation: work items for the worst plan of nq queries over this
// dataset (each query its own group over a whole layout), list slots for the
// worst plan or the list-memory cap, whichever is smaller.  A replay whose
// plan needs more slots returns padded outputs and sets the overflow flag
// fbf_search_graph_plan_info reads.
static fbf_status capture_any_batch(fbf_dataset *ds, fbf_search_graph *G, const float *queries_dev,
                                    const float *metas_dev, uint64_t nq, uint32_t k, const fbf_search_opts &o,
                                    uint32_t *out_ids_dev, float *out_dists_dev, hipStream_t st) {
    if (!queries_dev || !metas_dev || !out_ids_dev || !out_dists_dev) return fail(FBF_ERR_INVALID_ARG, "null argument");
    if (k == 0 || k > 512) return fail(FBF_ERR_INVALID_ARG, "k must be in [1, 512]");
    const bool mfma = o.mode == FBF_MODE_BF16 || o.mode == FBF_MODE_FP8 || o.mode == FBF_MODE_SQ8 ||
                      o.mode == FBF_MODE_SQ8SYM;
    if (!mfma) return fail(FBF_ERR_INVALID_ARG, "graphs that plan on the device run the BF16, FP8, SQ8 and SQ8SYM modes");
    fbf_status s = mode_check(ds, o.mode);
    if (s != FBF_OK) return s;
    if (nq >= 0x7FFFFFFFull) return fail(FBF_ERR_INVALID_ARG, "batch too large");
    uint32_t K = 0;
    s = kprime_of(o, k, K);
    if (s != FBF_OK) return s;
    const uint32_t chunk = o.chunk_rows ? (o.chunk_rows + kRowTile - 1) / kRowTile * kRowTile : 65536u;
    const uint32_t gmax = o.group_queries ? std::min<uint32_t>(o.group_queries, 256u) : 256u;
    const uint32_t pdiv = o.prepass_div ? o.prepass_div : 16u;
    const uint32_t margin = o.prepass_margin ? o.prepass_margin : K <= 160 ? 300u : K <= 500 ? 200u : 150u;
    const uint32_t kpre = (uint32_t)std::min<uint64_t>(
        K, std::max<uint64_t>(1, ((uint64_t)K * margin + 100ull * pdiv - 1) / (100ull * pdiv)));
    const bool seeded = pdiv > 1;
    const int B = K <= 160 ? 256 : K <= 300 ? 512 : 1024;
    const int B_pre = !seeded ? B : std::min(B, kpre <= 64 ? 128 : kpre <= 160 ? 256 : kpre <= 300 ? 512 : 1024);
    // Prefix items carry at most 128 queries (two waves per tile), as in the
    // host plan: prefix lists start empty and compact often.
    const uint32_t gpre = std::min<uint32_t>(gmax, 128u);
    auto waves = [](uint32_t g) { return g <= kMfmaWaveQueries ? 1 : g <= 2 * kMfmaWaveQueries ? 2 : 4; };
    const int nw = waves(gmax), nw_pre = waves(gpre);
    // Worst case: every query its own group over a whole layout.
    const uint64_t nchunk = (ds->n + kRowTile + chunk - 1) / chunk + 1;
    const uint64_t max_items = nq * nchunk;
    HIP_TRY(hipSetDevice(ds->device));
    size_t free_b = 0, total_b = 0;
    HIP_TRY(hipMemGetInfo(&free_b, &total_b));
    const uint64_t list_cap = std::max<uint64_t>(kAnyBatchListBytes, (uint64_t)free_b / 4 * 3);
    const uint64_t max_slots =
        std::min<uint64_t>({2 * nq * nchunk, list_cap / ((uint64_t)B * sizeof(uint64_t)), 0xFFFFFFF0ull});
    if (max_items >= 0xFFFFFFF0ull) return fail(FBF_ERR_INVALID_ARG, "batch too large for a graph that plans on the device");
    const uint64_t nwin = (nq + gmax - 1) / gmax, ngslot = nwin * gmax;
    const uint64_t nblk = (nq + 255) / 256;
    const uint32_t nbins = dplan_bins(ds->n, chunk);
    Scratch &S = G->scr;
    DPlanBufs &D = G->dp;
    HIP_TRY(D.ranges.reserve(nq * 4 * sizeof(uint32_t)));
    HIP_TRY(D.rkeys.reserve(nq * 4 * sizeof(uint64_t)));
    HIP_TRY(D.rvals.reserve(nq * 2 * sizeof(uint32_t)));
    for (DevBuf *b : {&D.keys, &D.skeys}) HIP_TRY(b->reserve(nq * sizeof(uint64_t)));
    for (DevBuf *b : {&D.vals, &D.order}) HIP_TRY(b->reserve(nq * sizeof(uint32_t)));
    HIP_TRY(D.hist.reserve(256 * nblk * sizeof(uint32_t)));
    HIP_TRY(D.hoff.reserve((256 * nblk + 1) * sizeof(uint64_t)));
    const uint64_t nparts = dplan_scan_parts(nq, gmax, ds->n);
    HIP_TRY(D.part.reserve(nparts * sizeof(uint64_t)));
    HIP_TRY(D.poff.reserve((nparts + 1) * sizeof(uint64_t)));
    HIP_TRY(D.gd.reserve(ngslot * kGdStride * sizeof(uint32_t)));
    HIP_TRY(D.gnc.reserve(ngslot * kGncStride * sizeof(uint32_t)));
    HIP_TRY(D.qnc.reserve(nq * 2 * sizeof(uint32_t)));
    HIP_TRY(D.ioff_pre.reserve((ngslot + 1) * sizeof(uint64_t)));
    HIP_TRY(D.ioff_main.reserve((ngslot + 1) * sizeof(uint64_t)));
    for (DevBuf *b : {&D.items_tmp_pre, &D.items_tmp_main, &D.items_pre, &D.items_main})
        HIP_TRY(b->reserve(max_items * sizeof(WorkItem)));
    for (DevBuf *b : {&D.islot_pre, &D.islot_main}) HIP_TRY(b->reserve(max_items * sizeof(uint32_t)));
    const uint32_t nrb = dplan_row_bins(ds->n);
    HIP_TRY(D.bins.reserve(2ull * nrb * sizeof(uint32_t)));
    HIP_TRY(D.boff.reserve(2ull * (nrb + 1) * sizeof(uint64_t)));
    HIP_TRY(D.bcost.reserve(2ull * nbins * sizeof(uint64_t)));
    HIP_TRY(D.iq_pre.reserve(nq * sizeof(ItemQuery)));
    HIP_TRY(D.iq_main.reserve(nq * sizeof(ItemQuery)));
    HIP_TRY(D.n_items.reserve(2 * sizeof(uint64_t)));
    HIP_TRY(D.xoff.reserve(18 * sizeof(uint32_t)));
    HIP_TRY(D.status.reserve(4 * sizeof(uint32_t)));
    HIP_TRY(S.d_slot_off.reserve((nq + 1) * sizeof(uint64_t)));
    HIP_TRY(S.d_pre_end.reserve(nq * sizeof(uint64_t)));
    HIP_TRY(S.d_lists.reserve(max_slots * (size_t)B * sizeof(uint64_t)));
    HIP_TRY(S.d_counts.reserve(max_slots * sizeof(uint32_t)));
    HIP_TRY(S.d_queue.reserve(8 * 64));
    HIP_TRY(S.d_gtau.reserve(nq * sizeof(uint64_t)));
    HIP_TRY(S.d_tau0.reserve(nq * sizeof(uint64_t)));
    HIP_TRY(S.d_flags.reserve((nq + 1) * sizeof(uint32_t)));
    HIP_TRY(S.d_qf.reserve(nq * kQFragU4 * sizeof(uint4)));
    if (o.mode == FBF_MODE_SQ8SYM) HIP_TRY(S.d_q.reserve(nq * 112));

    DPlan dp{};
    dp.ranges = D.ranges.as<uint32_t>();
    dp.nq = nq;
    dp.n_rows = ds->n;
    dp.gmax = gmax;
    dp.gpre = gpre;
    dp.chunk = chunk;
    dp.pdiv = seeded ? pdiv : 1u;
    dp.min_prefix = (uint32_t)std::max<uint64_t>(kMinPrefixRows, 2ull * kpre);
    dp.max_items = max_items;
    dp.max_slots = max_slots;
    dp.keys = D.keys.as<uint64_t>();
    dp.skeys = D.skeys.as<uint64_t>();
    dp.vals = D.vals.as<uint32_t>();
    dp.order = D.order.as<uint32_t>();
    dp.hist = D.hist.as<uint32_t>();
    dp.hoff = D.hoff.as<uint64_t>();
    dp.part = D.part.as<uint64_t>();
    dp.poff = D.poff.as<uint64_t>();
    dp.gd = D.gd.as<uint32_t>();
    dp.gnc = D.gnc.as<uint32_t>();
    dp.qnc = D.qnc.as<uint32_t>();
    dp.ioff_pre = D.ioff_pre.as<uint64_t>();
    dp.ioff_main = D.ioff_main.as<uint64_t>();
    dp.slot_off = S.d_slot_off.as<uint64_t>();
    dp.pre_end = S.d_pre_end.as<uint64_t>();
    dp.items_tmp_pre = D.items_tmp_pre.as<WorkItem>();
    dp.items_tmp_main = D.items_tmp_main.as<WorkItem>();
    dp.items_pre = D.items_pre.as<WorkItem>();
    dp.items_main = D.items_main.as<WorkItem>();
    dp.islot_pre = D.islot_pre.as<uint32_t>();
    dp.islot_main = D.islot_main.as<uint32_t>();
    dp.bins = D.bins.as<uint32_t>();
    dp.bcost = D.bcost.as<unsigned long long>();
    dp.boff = D.boff.as<uint64_t>();
    dp.iq_pre = D.iq_pre.as<ItemQuery>();
    dp.iq_main = D.iq_main.as<ItemQuery>();
    dp.n_items = D.n_items.as<uint64_t>();
    dp.xoff_pre = D.xoff.as<uint32_t>();
    dp.xoff_main = D.xoff.as<uint32_t>() + 9;
    dp.status = D.status.as<uint32_t>();

    const int variant = o.mode == FBF_MODE_BF16 ? 0 : o.mode == FBF_MODE_FP8 ? 1 : o.mode == FBF_MODE_SQ8 ? 2 : 3;
    warm_scan_mfma(variant, B, nw);
    if (seeded) {
        warm_scan_mfma(variant, B_pre, nw_pre);
        warm_scan_mfma(variant, B, nw_pre);
    }
    HIP_TRY(hipMemsetAsync(D.status.p, 0, 4 * sizeof(uint32_t), st));  // before the capture: no memset node
    HIP_TRY(hipStreamSynchronize(st));
    HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    fbf_status sw = [&]() -> fbf_status {
        HIP_TRY(launch_route(metas_dev, nq, ds->d_lab_b, ds->d_ts_b, ds->d_ts_t, ds->n, (float)ds->n, 0.045f, 0.08f,
                             D.ranges.as<uint32_t>(), nullptr, D.rkeys.as<uint64_t>(), D.rvals.as<uint32_t>(), nullptr,
                             0, st));
        HIP_TRY(launch_fill_u32(dp.status, 0u, 4, st));
        HIP_TRY(launch_dplan(dp, st));
        uint64_t *gtau = S.d_gtau.as<uint64_t>();
        uint32_t *flags = S.d_flags.as<uint32_t>();
        uint32_t *queue = S.d_queue.as<uint32_t>();
        HIP_TRY(launch_fill_u32(reinterpret_cast<uint32_t *>(gtau), 0xFFFFFFFFu, 2 * nq, st));
        HIP_TRY(launch_fill_u32(flags, 0u, nq + 1, st));
        if (variant == 3) HIP_TRY(launch_sym_encode(queries_dev, nullptr, nq, nq, ds->alpha, S.d_q.as<int8_t>(), st));
        HIP_TRY(launch_qfrag(variant, queries_dev, ds->sq8_w, variant == 3 ? S.d_q.as<int8_t>() : nullptr, nq,
                             S.d_qf.as<uint4>(), st));
        ScanMfmaArgs a{};
        a.qfrag = S.d_qf.as<uint4>();
        for (int L = 0; L < 2; ++L) {
            const LayoutDev &Ld = ds->lay[L];
            a.lay[L].ids = Ld.ids;
            a.lay[L].tiles = variant == 0 ? Ld.bf16 : variant == 1 ? Ld.fp8 : variant == 2 ? Ld.sq8 : Ld.sym_mfma;
            a.lay[L].xnorm = variant == 1 ? Ld.xnorm : variant == 2 ? Ld.sq8_norm : Ld.sym_norm;
            a.lay[L].scale = Ld.fp8_scale;
        }
        a.lists = S.d_lists.as<uint64_t>();
        a.counts = S.d_counts.as<uint32_t>();
        a.gtau = gtau;
        a.list_stride = (uint32_t)B;
        a.queue = queue;
        a.n_items = max_items;  // grid sizing; the runs come from the device
        auto scan = [&](int pass, uint32_t kk, int cap_b, const uint32_t *mask) -> fbf_status {
            a.items = pass ? dp.items_main : dp.items_pre;
            a.iq = pass ? dp.iq_main : dp.iq_pre;
            a.xoff_dev = pass ? dp.xoff_main : dp.xoff_pre;
            a.K = kk;
            a.qmask = mask;
            a.nmask = mask ? mask + nq : nullptr;
            HIP_TRY(launch_fill_u32(queue, 0u, 8 * 16, st));
            HIP_TRY(launch_scan_mfma(variant, a, cap_b, pass ? nw : nw_pre, st));
            return FBF_OK;
        };
        MergeArgs m{};
        m.slot_off = S.d_slot_off.as<uint64_t>();
        m.lists = S.d_lists.as<uint64_t>();
        m.counts = S.d_counts.as<uint32_t>();
        m.list_cap = (uint32_t)B;
        m.nq = nq;
        m.k = k;
        m.kprime = K;
        m.queries = queries_dev;
        m.base_rows = ds->base_rows;
        m.out_ids = out_ids_dev;
        m.out_dists = out_dists_dev;
        m.rerank = 1;
        m.key_pos = 1;
        m.ids_lay[0] = ds->lay[0].ids;
        m.ids_lay[1] = ds->lay[1].ids;
        fbf_status r = FBF_OK;
        if (!seeded) {
            r = scan(1, K, B, nullptr);
            if (r != FBF_OK) return r;
            HIP_TRY(launch_merge(m, st));
            return FBF_OK;
        }
        r = scan(0, kpre, B_pre, nullptr);
        if (r != FBF_OK) return r;
        m.tau0 = S.d_tau0.as<uint64_t>();
        m.flags = flags;
        m.gtau = gtau;
        m.nflag = flags + nq;
        m.pre_end = S.d_pre_end.as<uint64_t>();
        m.pass = 3;  // seeds: the kpre-th key of each query's prefix lists
        m.kprime = kpre;
        HIP_TRY(launch_merge(m, st));
        m.kprime = K;
        r = scan(1, K, B, nullptr);
        if (r != FBF_OK) return r;
        m.pass = 1;  // verify each seeded query
        HIP_TRY(launch_merge(m, st));
        r = scan(0, K, B, flags);  // rescan the failures over their whole range
        if (r != FBF_OK) return r;
        r = scan(1, K, B, flags);
        if (r != FBF_OK) return r;
        m.pass = 2;
        HIP_TRY(launch_merge(m, st));
        return FBF_OK;
    }();
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(st, &g);
    if (sw != FBF_OK) {
        if (g) (void)hipGraphDestroy(g);
        return sw;
    }
    if (e != hipSuccess || g == nullptr) {
        if (g) (void)hipGraphDestroy(g);
        return fail(FBF_ERR_DEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    }
    hipGraphExec_t x = nullptr;
    const hipError_t ei = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    if (ei != hipSuccess) {
        (void)hipGraphDestroy(g);
        return fail(FBF_ERR_DEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
    }
    G->g.graph = g;
    G->g.exec = x;
    G->any_batch = true;
    return FBF_OK;
}

extern "C" fbf_status fbf_search_graph_create(fbf_dataset *ds, const float *queries_dev, const float *metas,
                                              uint64_t nq, uint32_t k, const fbf_search_opts *opts,
                                              uint32_t *out_ids_dev, float *out_dists_dev,
                                              fbf_search_graph **out) {
    if (!ds || !out) return fail(FBF_ERR_INVALID_ARG, "null argument");
    if (nq == 0) return fail(FBF_ERR_INVALID_ARG, "a captured search needs at least one query");
    *out = nullptr;
    HIP_TRY(hipSetDevice(ds->device));
    auto *G = new (std::nothrow) fbf_search_graph();
    if (!G) return fail(FBF_ERR_DEVICE, "out of host memory");
    G->ds = ds;
    G->nq = nq;
    hipStream_t st = nullptr;
    hipError_t e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
    fbf_status s = e == hipSuccess ? FBF_OK : fail(FBF_ERR_DEVICE, "hipStreamCreate");
    if (s == FBF_OK && opts && opts->metas_on_device) {
        s = capture_any_batch(ds, G, queries_dev, metas, nq, k, *opts, out_ids_dev, out_dists_dev, st);
    } else if (s == FBF_OK) {
        s = search_impl(ds, G->scr, false, queries_dev, metas, nq, k, opts, out_ids_dev, out_dists_dev, st, &G->g);
    }
    if (st) {
        (void)hipStreamSynchronize(st);
        (void)hipStreamDestroy(st);
    }
    if (s != FBF_OK) {
        fbf_search_graph_destroy(G);
        return s;
    }
    *out = G;
    return FBF_OK;
}

extern "C" fbf_status fbf_search_graph_launch(fbf_search_graph *g, void *stream) {
    if (!g || !g->g.exec) return fail(FBF_ERR_INVALID_ARG, "null graph");
    HIP_TRY(hipSetDevice(g->ds->device));
    HIP_TRY(hipGraphLaunch(g->g.exec, (hipStream_t)stream));
    return FBF_OK;
}

extern "C" fbf_status fbf_search_graph_destroy(fbf_search_graph *g) {
    if (!g) return FBF_OK;
    (void)hipSetDevice(g->ds->device);
    // The caller's launches must be complete; the device is drained so that
    // freeing the graph's buffers cannot race a replay still in flight.
    (void)hipDeviceSynchronize();
    if (g->g.exec) (void)hipGraphExecDestroy(g->g.exec);
    if (g->g.graph) (void)hipGraphDestroy(g->g.graph);
    g->scr.release();
    g->dp.release();
    delete g;
    return FBF_OK;
}

extern "C" fbf_status fbf_search_graph_plan_info(fbf_search_graph *g, uint64_t *items_prefix, uint64_t *items_main,
                                                 uint32_t *overflow) {
    if (!g) return fail(FBF_ERR_INVALID_ARG, "null graph");
    if (!g->any_batch) return fail(FBF_ERR_INVALID_ARG, "the graph was planned on the host");
    HIP_TRY(hipSetDevice(g->ds->device));
    uint32_t st[4] = {0, 0, 0, 0};
    HIP_TRY(hipMemcpy(st, g->dp.status.p, sizeof(st), hipMemcpyDeviceToHost));
    if (items_prefix) *items_prefix = st[0];
    if (items_main) *items_main = st[1];
    if (overflow) *overflow = st[2];
    return FBF_OK;
}

extern "C" fbf_status fbf_search_host(fbf_dataset *ds, const float *queries, const float *metas,
                                      uint64_t nq, uint32_t k, const fbf_search_opts *opts,
                                      uint32_t *out_ids, float *out_dists) {
    if (!ds || !queries || !metas || !out_ids || !out_dists) return fail(FBF_ERR_INVALID_ARG, "null argument");
    if (nq == 0) return FBF_OK;
    HIP_TRY(hipSetDevice(ds->device));
    float *dq = nullptr, *dd = nullptr;
    uint32_t *di = nullptr;
    HIP_TRY(hipMalloc(&dq, nq * ds->d * sizeof(float)));
    hipError_t e1 = hipMalloc(&di, nq * (size_t)k * sizeof(uint32_t));
    hipError_t e2 = hipMalloc(&dd, nq * (size_t)k * sizeof(float));
    fbf_status s = FBF_OK;
    if (e1 != hipSuccess || e2 != hipSuccess) s = fail(FBF_ERR_DEVICE, "hipMalloc outputs");
    hipStream_t st = nullptr;
    if (s == FBF_OK && hipStreamCreate(&st) != hipSuccess) s = fail(FBF_ERR_DEVICE, "hipStreamCreate");
    if (s == FBF_OK && hipMemcpyAsync(dq, queries, nq * ds->d * sizeof(float), hipMemcpyHostToDevice, st) != hipSuccess)
        s = fail(FBF_ERR_DEVICE, "upload queries");
    if (s == FBF_OK) s = fbf_search(ds, dq, metas, nq, k, opts, di, dd, st);
    if (s == FBF_OK && (hipMemcpyAsync(out_ids, di, nq * (size_t)k * 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
                        hipMemcpyAsync(out_dists, dd, nq * (size_t)k * 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
                        hipStreamSynchronize(st) != hipSuccess))
        s = fail(FBF_ERR_DEVICE, "download results");
    if (st) hipStreamDestroy(st);
    hipFree(dq);
    if (di) hipFree(di);
    if (dd) hipFree(dd);
    return s;
}

extern "C" fbf_status fbf_dataset_quant_params(const fbf_dataset *ds, float *alpha, float *mi, float *dif) {
    if (!ds) return fail(FBF_ERR_INVALID_ARG, "null dataset");
    if (alpha) {
        if (!(ds->repr & FBF_REPR_SQ8SYM)) return fail(FBF_ERR_NO_REPR, "dataset has no symmetric int8 codes");
        *alpha = ds->alpha;
    }
    if (mi || dif) {
        if (!(ds->repr & FBF_REPR_SQ8)) return fail(FBF_ERR_NO_REPR, "dataset has no uint8 codes");
        if (mi) std::memcpy(mi, ds->h_mi.data(), kDim * sizeof(float));
        if (dif) std::memcpy(dif, ds->h_dif.data(), kDim * sizeof(float));
    }
    return FBF_OK;
}

// Rows of one sorted layout as stored on the device: original ids and the
// quantiser codes (symmetric int8 rows of 112 bytes, or the per-dimension
// uint8 codes unpacked from their MFMA fragment order).
extern "C" fbf_status fbf_dataset_layout_rows(const fbf_dataset *ds, uint32_t layout, uint64_t row0, uint64_t nrows,
                                              uint32_t repr, uint32_t *ids, void *codes) {
    if (!ds) return fail(FBF_ERR_INVALID_ARG, "null dataset");
    if (layout > 1) return fail(FBF_ERR_INVALID_ARG, "layout must be 0 or 1");
    if (row0 > ds->n || nrows > ds->n - row0) return fail(FBF_ERR_INVALID_ARG, "rows outside the layout");
    if (codes && repr != FBF_REPR_SQ8SYM && repr != FBF_REPR_SQ8)
        return fail(FBF_ERR_INVALID_ARG, "codes are kept for FBF_REPR_SQ8SYM and FBF_REPR_SQ8 only");
    if (codes && !(ds->repr & repr)) return fail(FBF_ERR_NO_REPR, "dataset was created without these codes");
    if (nrows == 0) return FBF_OK;
    HIP_TRY(hipSetDevice(ds->device));
    const LayoutDev &D = ds->lay[layout];
    if (ids) HIP_TRY(hipMemcpy(ids, D.ids + row0, nrows * sizeof(uint32_t), hipMemcpyDeviceToHost));
    if (!codes) return FBF_OK;
    if (repr == FBF_REPR_SQ8SYM) {
        HIP_TRY(hipMemcpy(codes, D.sym + row0 * 112, nrows * 112, hipMemcpyDeviceToHost));
        return FBF_OK;
    }
    // uint8 codes: 32-row blocks of 7 steps x 64 lanes x 8 bytes; lane l of
    // step s holds row l % 32 of the block at dims 16 s + 8 (l / 32) + e.
    constexpr uint64_t kBlockBytes = 7 * 64 * 8;
    const uint64_t b0 = row0 / 32, b1 = (row0 + nrows + 31) / 32;
    std::vector<uint8_t> blk((b1 - b0) * kBlockBytes);
    HIP_TRY(hipMemcpy(blk.data(), (const uint8_t *)D.sq8 + b0 * kBlockBytes, blk.size(), hipMemcpyDeviceToHost));
    uint8_t *out = static_cast<uint8_t *>(codes);
    for (uint64_t r = 0; r < nrows; ++r) {
        const uint64_t p = row0 + r, b = p / 32 - b0, row = p % 32;
        for (int k = 0; k < kDim; ++k) {
            const int st = k / 16, half = (k % 16) / 8, e = k % 8;
            out[r * kDim + k] = blk[b * kBlockBytes + ((uint64_t)st * 64 + half * 32 + row) * 8 + e];
        }
    }
    return FBF_OK;
}

// ---------------------------------------------------------------- measurement
extern "C" fbf_status fbf_profile_enable(fbf_dataset *ds, int enable) {
    if (!ds) return fail(FBF_ERR_INVALID_ARG, "null dataset");
    ds->profile = enable != 0;
    return FBF_OK;
}

extern "C" fbf_status fbf_profile_collect(fbf_dataset *ds, const char *kernel, uint64_t *launches,
                                          double *total_ms) {
    if (!ds || !kernel) return fail(FBF_ERR_INVALID_ARG, "null argument");
    int want = -1;
    int want_pass = -1;
    if (!std::strcmp(kernel, "scan")) want = 0;
    else if (!std::strcmp(kernel, "scan_main")) want = 0, want_pass = 0;
    else if (!std::strcmp(kernel, "scan_prefix")) want = 0, want_pass = 1;
    else if (!std::strcmp(kernel, "scan_rescan")) want = 0, want_pass = 2;
    else if (!std::strcmp(kernel, "merge")) want = 1;
    else if (!std::strcmp(kernel, "rerank")) want = 2;
    else if (!std::strcmp(kernel, "build")) want = 3;
    else if (!std::strcmp(kernel, "plan")) want = 4;
    else if (std::strcmp(kernel, "all")) return fail(FBF_ERR_INVALID_ARG, "unknown kernel name");
    uint64_t cnt = 0;
    double ms = 0.0;
    std::vector<ProfEvent> keep;
    for (auto &ev : ds->events) {
        if ((want >= 0 && ev.kind != want) || (want_pass >= 0 && ev.pass != want_pass)) {
            keep.push_back(ev);
            continue;
        }
        if (ev.kind == 4) {
            ms += ev.host_ms;
            ++cnt;
            continue;
        }
        HIP_TRY(hipEventSynchronize(ev.b));
        float t = 0.f;
        HIP_TRY(hipEventElapsedTime(&t, ev.a, ev.b));
        ms += t;
        ++cnt;
        hipEventDestroy(ev.a);
        hipEventDestroy(ev.b);
    }
    ds->events.swap(keep);
    if (launches) *launches = cnt;
    if (total_ms) *total_ms = ms;
    return FBF_OK;
}

extern "C" fbf_status fbf_last_search_stats(const fbf_dataset *ds, uint64_t *pairs, uint64_t *pairs_computed,
                                            uint64_t *work_items) {
    if (!ds) return fail(FBF_ERR_INVALID_ARG, "null dataset");
    if (pairs) *pairs = ds->last_pairs;
    if (pairs_computed) *pairs_computed = ds->last_pairs_computed;
    if (work_items) *work_items = ds->last_items;
    return FBF_OK;
}

extern "C" fbf_status fbf_last_search_pass_pairs(const fbf_dataset *ds, uint64_t *prefix_pairs,
                                                 uint64_t *main_pairs) {
    if (!ds) return fail(FBF_ERR_INVALID_ARG, "null dataset");
    if (prefix_pairs) *prefix_pairs = ds->last_pairs_prefix;
    if (main_pairs) *main_pairs = ds->last_pairs - ds->last_pairs_prefix;
    return FBF_OK;
}

extern "C" fbf_status fbf_last_search_rescans(fbf_dataset *ds, uint64_t *queries) {
    if (!ds || !queries) return fail(FBF_ERR_INVALID_ARG, "null argument");
    std::lock_guard<std::mutex> lock(ds->mu);
    *queries = 0;
    if (!ds->last_seeded) return FBF_OK;
    HIP_TRY(hipEventSynchronize(ds->scr.ev_nflag));
    *queries = *static_cast<const uint32_t *>(ds->scr.h_nflag.p);
    return FBF_OK;
}

extern "C" fbf_status fbf_last_search_fallbacks(const fbf_dataset *ds, uint64_t *queries) {
    if (!ds || !queries) return fail(FBF_ERR_INVALID_ARG, "null argument");
    *queries = ds->last_fallbacks;
    return FBF_OK;
}

extern "C" const char *fbf_last_error(void) { return g_err.c_str(); }

extern "C" const char *fbf_version(void) {
    return "fbf 0.3 gfx950: exact-f32 bf16-mfma fp8-mfma sq8-mfma sq8sym subinterval noalloc";
}

// ---------------------------------------------------------------- reference-shaped entry points
//
// Registry of host code arrays uploaded once (the router's codes_both /
// codes_time / bucket codes, hybrid_graph.cpp:155-161,505-524).  A codes_in
// pointer inside a registered range resolves to the device copy.
struct CodeReg {
    uint64_t bytes;
    int device;
    int8_t *dev;
};
static std::mutex g_reg_mu;
// Never destroyed (a heap object): its device copies live until they are
// unregistered or the process ends, and no destructor runs at exit.
static std::map<uintptr_t, CodeReg> &g_reg = *new std::map<uintptr_t, CodeReg>();

extern "C" fbf_status fbf_register_codes(const void *host_base, uint64_t bytes) {
    if (!host_base || bytes == 0) return fail(FBF_ERR_INVALID_ARG, "null or empty code array");
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    int8_t *d = nullptr;
    HIP_TRY(hipMalloc(&d, bytes));
    hipError_t e = hipMemcpy(d, host_base, bytes, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(d);
        return fail(FBF_ERR_DEVICE, std::string("upload codes: ") + hipGetErrorString(e));
    }
    std::lock_guard<std::mutex> lk(g_reg_mu);
    auto it = g_reg.find((uintptr_t)host_base);
    if (it != g_reg.end()) {
        (void)hipFree(it->second.dev);
        g_reg.erase(it);
    }
    g_reg[(uintptr_t)host_base] = CodeReg{bytes, dev, d};
    return FBF_OK;
}

extern "C" fbf_status fbf_unregister_codes(const void *host_base) {
    std::lock_guard<std::mutex> lk(g_reg_mu);
    auto it = g_reg.find((uintptr_t)host_base);
    if (it == g_reg.end()) return fail(FBF_ERR_UNREGISTERED, "pointer was not registered");
    (void)hipFree(it->second.dev);
    g_reg.erase(it);
    return FBF_OK;
}

// The registered array holding [p, p + bytes): its device base and size.
static bool resolve_codes_array(const void *p, uint64_t bytes, const int8_t *&arr, uint64_t &arr_bytes,
                                uint64_t &offset) {
    std::lock_guard<std::mutex> lk(g_reg_mu);
    const uintptr_t a = (uintptr_t)p;
    auto it = g_reg.upper_bound(a);
    if (it == g_reg.begin()) return false;
    --it;
    const uintptr_t base = it->first;
    if (a < base || a + bytes > base + it->second.bytes) return false;
    arr = it->second.dev;
    arr_bytes = it->second.bytes;
    offset = a - base;
    return true;
}

// Device address of [p, p + bytes) when it lies inside one registered array.
static const int8_t *resolve_codes(const void *p, uint64_t bytes) {
    std::lock_guard<std::mutex> lk(g_reg_mu);
    const uintptr_t a = (uintptr_t)p;
    auto it = g_reg.upper_bound(a);
    if (it == g_reg.begin()) return nullptr;
    --it;
    const uintptr_t base = it->first;
    if (a < base || a + bytes > base + it->second.bytes) return nullptr;
    if ((a - base) % 16) return nullptr;  // 16-byte loads need 16-byte alignment
    return it->second.dev + (a - base);
}

// Engine of the dataset-free calls: a stream and scratch of its own, so the
// router's OpenMP threads do not serialise on one another.  A thread borrows
// an engine from a process-wide pool on its first call and hands it back when
// the thread ends; engines are never freed.  Nothing here calls HIP from a
// destructor: at process exit the HIP runtime may already be tearing down
// (thread and static destructors run after it has begun to), and the device
// memory of the pool goes with the process.
struct ThreadEngine {
    int device = -1;
    hipStream_t st = nullptr;
    Scratch scr;
    DevBuf tmp_codes;
    HostPinned h_out;
    uint32_t *done = nullptr;  // arrival counter of sym_one_kernel (zero between calls)
    DevBuf d_reqs, d_done;     // combined calls: requests; arrival counters (zero between calls), then
                               // shared thresholds (all ones between calls)
    HostPinned h_reqs;
};
static std::mutex g_eng_mu;
static std::vector<ThreadEngine *> *g_eng_free = new std::vector<ThreadEngine *>();  // never destroyed

struct EngineLease {
    ThreadEngine *e = nullptr;
    ~EngineLease() {
        if (!e) return;
        std::lock_guard<std::mutex> lk(g_eng_mu);
        g_eng_free->push_back(e);  // host bookkeeping only
    }
};
static thread_local EngineLease t_lease;

// The calling thread's engine on the current device (stream created on first use).
static fbf_status thread_engine(ThreadEngine *&out) {
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (t_lease.e && t_lease.e->device != dev) {
        std::lock_guard<std::mutex> lk(g_eng_mu);
        g_eng_free->push_back(t_lease.e);
        t_lease.e = nullptr;
    }
    if (!t_lease.e) {
        std::lock_guard<std::mutex> lk(g_eng_mu);
        for (size_t i = 0; i < g_eng_free->size(); ++i)
            if ((*g_eng_free)[i]->device == dev) {
                t_lease.e = (*g_eng_free)[i];
                g_eng_free->erase(g_eng_free->begin() + (long)i);
                break;
            }
    }
    if (!t_lease.e) {
        auto *e = new (std::nothrow) ThreadEngine();
        if (!e) return fail(FBF_ERR_DEVICE, "out of host memory");
        e->device = dev;
        const hipError_t he = hipStreamCreateWithFlags(&e->st, hipStreamNonBlocking);
        if (he != hipSuccess) {
            delete e;
            return fail(FBF_ERR_DEVICE, std::string("hipStreamCreate: ") + hipGetErrorString(he));
        }
        t_lease.e = e;
    }
    out = t_lease.e;
    return FBF_OK;
}

// Device buffers of one symmetric-int8 batch: the plan, the per-(query,
// chunk) candidate lists, the query codes and the results.
struct SymBufs {
    WorkItem *items = nullptr;
    ItemQuery *iq = nullptr;
    uint64_t *slot_off = nullptr;
    uint64_t *lists = nullptr;
    uint32_t *counts = nullptr;
    int8_t *qcodes = nullptr;
    uint32_t *out_ids = nullptr;
    float *out_d = nullptr;  // raw int distances (merge key_int)
};

// Plan of nq slices [starts[i], starts[i] + counts[i]) of one code array: up
// to 16 queries per work item, and a chunk size that gives a batch about 2048
// work items (4 Ki to 64 Ki rows each), so one call fills the device even for
// a single long slice.
static fbf_status sym_plan(const uint64_t *starts, const uint64_t *counts, uint64_t nq, uint64_t n_rows, Plan &P) {
    std::vector<QRange> qr(nq);
    uint64_t total = 0;
    for (uint64_t i = 0; i < nq; ++i) {
        if (starts[i] + counts[i] > n_rows || starts[i] + counts[i] >= 0xFFFFFFFFull)
            return fail(FBF_ERR_INVALID_ARG, "slice outside the code array");
        qr[i] = QRange{0, (uint32_t)starts[i], (uint32_t)(starts[i] + counts[i]), (uint32_t)i};
        total += counts[i];
    }
    const uint64_t c = (total / 2048 + kRowTile - 1) / kRowTile * kRowTile;
    const uint32_t chunk = (uint32_t)std::min<uint64_t>(std::max<uint64_t>(c, 4096), 65536);
    make_plan(qr, chunk, kMaxGroup, P);
    return finish_slots(P, nq);
}

// Queue the scan and the merge of a planned batch whose plan and query codes
// are in b (on st).
static fbf_status sym_launch(const int8_t *codes_dev, uint64_t n_rows, const Plan &P, uint64_t nq, uint32_t k,
                             int B, const SymBufs &b, hipStream_t st) {
    ScanSymArgs a{};
    a.items = b.items;
    a.n_items = P.items.size();
    a.iq = b.iq;
    a.qcodes = b.qcodes;
    a.codes = codes_dev;
    a.ids = nullptr;
    a.n_rows = n_rows;
    a.lists = b.lists;
    a.counts = b.counts;
    a.K = k;
    a.queue = nullptr;
    HIP_TRY(launch_scan_sq8sym(a, B, st));
    MergeArgs m{};
    m.slot_off = b.slot_off;
    m.lists = b.lists;
    m.counts = b.counts;
    m.list_cap = (uint32_t)B;
    m.nq = nq;
    m.k = k;
    m.kprime = k;
    m.out_ids = b.out_ids;
    m.out_dists = b.out_d;
    m.rerank = 0;
    m.key_int = 1;
    HIP_TRY(launch_merge(m, st));
    return FBF_OK;
}

// Results of the merge (absolute row ids, raw int distances) as slice-relative
// positions and int distances; pads are -1 / INT32_MAX.  ids may alias out_pos
// and dbits out_dist.
static void sym_unpack(const uint32_t *ids, const uint32_t *dbits, const uint64_t *starts, uint64_t nq, uint32_t k,
                       int32_t *out_pos, int32_t *out_dist) {
    for (uint64_t i = 0; i < nq * (uint64_t)k; ++i) {
        const uint32_t id = ids[i];
        const uint32_t db = dbits[i];
        out_pos[i] = id == kPadId ? -1 : (int32_t)(id - (uint32_t)starts[i / k]);
        out_dist[i] = id == kPadId ? INT32_MAX : (int32_t)db;
    }
}

// Symmetric int8 top-k over nq slices of one device code array, on the
// calling thread's engine.  out_pos / out_dist are host arrays [nq][k];
// positions are relative to each slice.
static fbf_status sym_topk(const int8_t *codes_dev, uint64_t n_rows, const uint64_t *starts,
                           const uint64_t *counts, const int8_t *qcodes_host, uint64_t nq, uint32_t k,
                           int32_t *out_pos, int32_t *out_dist) {
    ThreadEngine *Ep = nullptr;
    fbf_status se = thread_engine(Ep);
    if (se != FBF_OK) return se;
    ThreadEngine &E = *Ep;
    hipStream_t st = E.st;
    if (nq == 1 && counts[0] > 0) {
        // One query: one launch (the workgroups' lists merged by the last
        // one to finish) and one copy back -- no plan, no query upload.
        const uint64_t n = counts[0];
        if (starts[0] + n > n_rows || n >= 0xFFFFFFFFull) return fail(FBF_ERR_INVALID_ARG, "slice outside the code array");
        const int B = list_capacity((int)k);
        // About 1024 rows per workgroup, at most 256 workgroups.
        uint64_t chunk = std::max<uint64_t>(1024, (n + 255) / 256);
        chunk = (chunk + kRowTile - 1) / kRowTile * kRowTile;
        const uint64_t grid = (n + chunk - 1) / chunk;
        Scratch &S = E.scr;
        HIP_TRY(S.d_lists.reserve(grid * (uint64_t)B * sizeof(uint64_t)));
        HIP_TRY(S.d_counts.reserve(grid * sizeof(uint32_t)));
        HIP_TRY(S.d_out_ids.reserve(2 * (size_t)k * sizeof(uint32_t)));
        if (!E.done) {
            HIP_TRY(hipMalloc(&E.done, 256));
            HIP_TRY(hipMemsetAsync(E.done, 0, 256, st));
        }
        SymOneArgs a{};
        a.codes = codes_dev + starts[0] * 112;
        a.n = (uint32_t)n;
        a.chunk = (uint32_t)chunk;
        a.K = k;
        std::memcpy(a.q, qcodes_host, 112);
        a.lists = S.d_lists.as<uint64_t>();
        a.counts = S.d_counts.as<uint32_t>();
        a.done = E.done;
        a.out = S.d_out_ids.as<uint32_t>();
        HIP_TRY(launch_sym_one(a, B, st));
        // (Writing the results straight into pinned host memory from the
        // kernel instead of this copy measured 45% slower per call.)
        HIP_TRY(E.h_out.reserve(2 * (size_t)k * sizeof(uint32_t)));
        uint32_t *hp = (uint32_t *)E.h_out.p;
        HIP_TRY(hipMemcpyAsync(hp, a.out, 2 * (size_t)k * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        for (uint32_t i = 0; i < k; ++i) {
            const bool pad = hp[i] == kPadId;
            out_pos[i] = pad ? -1 : (int32_t)hp[i];
            out_dist[i] = pad ? INT32_MAX : (int32_t)hp[k + i];
        }
        return FBF_OK;
    }
    Plan P;
    fbf_status s = sym_plan(starts, counts, nq, n_rows, P);
    if (s != FBF_OK) return s;
    const int B = list_capacity((int)k);
    Scratch &S = E.scr;
    s = upload_plan(S, P, nq, B, st);
    if (s != FBF_OK) return s;
    HIP_TRY(S.d_q.reserve(nq * 112));
    HIP_TRY(S.d_out_ids.reserve(nq * (size_t)k * 4));
    HIP_TRY(S.d_out_d.reserve(nq * (size_t)k * 4));
    HIP_TRY(hipMemcpyAsync(S.d_q.p, qcodes_host, nq * 112, hipMemcpyHostToDevice, st));
    SymBufs b;
    b.items = S.d_items.as<WorkItem>();
    b.iq = S.d_iq.as<ItemQuery>();
    b.slot_off = S.d_slot_off.as<uint64_t>();
    b.lists = S.d_lists.as<uint64_t>();
    b.counts = S.d_counts.as<uint32_t>();
    b.qcodes = S.d_q.as<int8_t>();
    b.out_ids = S.d_out_ids.as<uint32_t>();
    b.out_d = S.d_out_d.as<float>();
    s = sym_launch(codes_dev, n_rows, P, nq, k, B, b, st);
    if (s != FBF_OK) return s;
    HIP_TRY(E.h_out.reserve(nq * (size_t)k * 8));
    uint32_t *hp = (uint32_t *)E.h_out.p;
    HIP_TRY(hipMemcpyAsync(hp, S.d_out_ids.p, nq * (size_t)k * 4, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(hp + nq * (size_t)k, S.d_out_d.p, nq * (size_t)k * 4, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    sym_unpack(hp, hp + nq * (size_t)k, starts, nq, k, out_pos, out_dist);
    return FBF_OK;
}

// Combining of concurrent per-call searches (the router's OpenMP threads call
// bruteforce() once per query, hybrid_graph.cpp:168-211).  A call on a
// registered code array queues its request; a waiting caller that finds
// fewer than kBfLeaders batches in flight takes every queued request of its
// device and runs them on its own engine as batches (one per code array and
// k: one plan, one scan launch, one merge, one copy back), then hands each
// request its results.  A lone request runs as today's single launch.  Each
// request's results are those of its own call: the batch computes every
// (slice, query) pair independently.
struct BfReq {
    const int8_t *arr = nullptr;    // device base of the registered code array
    uint64_t arr_rows = 0;          // its whole rows
    uint64_t start = 0;             // the slice: rows [start, start + n) of the array
    uint64_t n = 0;
    const int8_t *q = nullptr;      // the query's 112 code bytes (host)
    uint32_t k = 0;
    int32_t *pos = nullptr, *dist = nullptr;  // host outputs [k]
    fbf_status status = FBF_OK;
    std::string err;
    bool done = false;
};
// Batches in flight per device: 3 (C1-like router slices from 16 C++ threads:
// 1 -> 24.1 k calls/s, 2 -> 29.7 k, 3 -> 31.2 k, 4 -> 30.8 k).
constexpr int kBfLeaders = 3;
constexpr size_t kBfMaxBatch = 512;  // requests per batch
struct BfCombiner {
    std::mutex mu;
    std::condition_variable cv;
    std::vector<BfReq *> queue;
    int leaders = 0;
};
static BfCombiner *g_bf = new BfCombiner[64];  // per device, never destroyed

// Run one batch of requests of one device on the calling thread's engine:
// every request a one-slice search of its own, all in one launch
// (sym_multi_kernel), at the largest k of the batch (a request's own top-k is
// the head of that sorted list); one upload of the requests, one copy back.
static fbf_status bf_run_batch_(std::vector<BfReq *> &batch) {
    ThreadEngine *Ep = nullptr;
    fbf_status s = thread_engine(Ep);
    if (s != FBF_OK) return s;
    ThreadEngine &E = *Ep;
    const hipStream_t st = E.st;
    const uint32_t nreq = (uint32_t)batch.size();
    uint32_t K = 0;
    for (const BfReq *r : batch) K = std::max(K, r->k);
    const int B = list_capacity((int)K);
    HIP_TRY(E.h_reqs.reserve(nreq * sizeof(SymReq)));
    SymReq *hr = static_cast<SymReq *>(E.h_reqs.p);
    uint64_t slots = 0;
    uint32_t gx = 0;
    for (uint32_t i = 0; i < nreq; ++i) {
        const BfReq &r = *batch[i];
        // At least 2048 rows per workgroup and at most 64 workgroups: each
        // keeps its own top k and the last one merges them all, so fewer,
        // longer workgroups (their tiles prefetched) cost less in all.  On
        // slices like the router's (1 Ki to 800 Ki rows) from 16 threads:
        // 26.5 k calls/s, against 17.9 k at 1 Ki rows and up to 256.
        uint64_t chunk = std::max<uint64_t>(2048, (r.n + 63) / 64);
        chunk = (chunk + kRowTile - 1) / kRowTile * kRowTile;
        const uint32_t nblk = (uint32_t)((r.n + chunk - 1) / chunk);
        // The merging workgroup reads every list's count in one load, one lane each.
        if (nblk > 64) return fail(FBF_ERR_UNSUPPORTED, "combined call split into more than 64 workgroups");
        SymReq q{};
        q.codes = r.arr + r.start * 112;
        q.n = (uint32_t)r.n;
        q.chunk = (uint32_t)chunk;
        q.nblk = nblk;
        q.list0 = (uint32_t)slots;
        std::memcpy(q.q, r.q, 112);
        hr[i] = q;
        slots += nblk;
        gx = std::max(gx, nblk);
    }
    Scratch &S = E.scr;
    HIP_TRY(S.d_lists.reserve(slots * (uint64_t)B * sizeof(uint64_t)));
    HIP_TRY(S.d_counts.reserve(slots * sizeof(uint32_t)));
    HIP_TRY(S.d_out_ids.reserve((size_t)nreq * 2 * K * sizeof(uint32_t)));
    HIP_TRY(E.d_reqs.reserve(nreq * sizeof(SymReq)));
    // [done: kBfMaxBatch u32][gtau: kBfMaxBatch u64], set once.
    const size_t done_bytes = kBfMaxBatch * (sizeof(uint32_t) + sizeof(uint64_t));
    if (E.d_done.bytes < done_bytes) {
        HIP_TRY(E.d_done.reserve(done_bytes));
        HIP_TRY(hipMemsetAsync(E.d_done.p, 0, kBfMaxBatch * sizeof(uint32_t), st));
        HIP_TRY(hipMemsetAsync(E.d_done.as<uint32_t>() + kBfMaxBatch, 0xFF, kBfMaxBatch * sizeof(uint64_t), st));
    }
    if (nreq > kBfMaxBatch) return fail(FBF_ERR_INVALID_ARG, "too many combined requests");
    HIP_TRY(hipMemcpyAsync(E.d_reqs.p, hr, nreq * sizeof(SymReq), hipMemcpyHostToDevice, st));
    SymMultiArgs a{};
    a.reqs = E.d_reqs.as<SymReq>();
    a.nreq = nreq;
    a.K = K;
    a.lists = S.d_lists.as<uint64_t>();
    a.counts = S.d_counts.as<uint32_t>();
    a.done = E.d_done.as<uint32_t>();
    a.gtau = reinterpret_cast<uint64_t *>(E.d_done.as<uint32_t>() + kBfMaxBatch);
    a.out = S.d_out_ids.as<uint32_t>();
    HIP_TRY(launch_sym_multi(a, gx, B, st));
    HIP_TRY(E.h_out.reserve((size_t)nreq * 2 * K * sizeof(uint32_t)));
    uint32_t *ho = static_cast<uint32_t *>(E.h_out.p);
    HIP_TRY(hipMemcpyAsync(ho, a.out, (size_t)nreq * 2 * K * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    for (uint32_t i = 0; i < nreq; ++i) {
        BfReq &r = *batch[i];
        const uint32_t *o = ho + (size_t)i * 2 * K;
        for (uint32_t j = 0; j < r.k; ++j) {
            const bool pad = o[j] == kPadId;
            r.pos[j] = pad ? -1 : (int32_t)o[j];
            r.dist[j] = pad ? INT32_MAX : (int32_t)o[K + j];
        }
    }
    return FBF_OK;
}

static void bf_run_batch(std::vector<BfReq *> &batch) {
    const fbf_status s = bf_run_batch_(batch);
    const std::string msg = s == FBF_OK ? std::string() : g_err;
    for (BfReq *r : batch) {
        r->status = s;
        r->err = msg;
    }
}

// Queue a request and wait for its results (running batches meanwhile when
// fewer than kBfLeaders are in flight).
static fbf_status bf_combined(BfReq &r, int dev) {
    BfCombiner &C = g_bf[dev];
    std::unique_lock<std::mutex> lk(C.mu);
    C.queue.push_back(&r);
    while (!r.done) {
        if (C.leaders < kBfLeaders && !C.queue.empty()) {
            ++C.leaders;
            std::vector<BfReq *> batch;
            const size_t m = std::min(C.queue.size(), kBfMaxBatch);
            batch.assign(C.queue.begin(), C.queue.begin() + (long)m);
            C.queue.erase(C.queue.begin(), C.queue.begin() + (long)m);
            lk.unlock();
            bf_run_batch(batch);
            lk.lock();
            for (BfReq *b : batch) b->done = true;
            --C.leaders;
            C.cv.notify_all();
        } else {
            C.cv.wait(lk);
        }
    }
    if (r.status != FBF_OK) return fail(r.status, r.err);
    return FBF_OK;
}

static fbf_status bruteforce_common(size_t k, int *closest_points, float *distances, size_t npoints,
                                    const char *codes_in, float alpha, const int8_t *queries_in) {
    if (k == 0) return FBF_OK;  // the reference writes nothing either
    if (!closest_points || !queries_in || (npoints && !codes_in))
        return fail(FBF_ERR_INVALID_ARG, "null argument");
    // A list keeps k keys plus one tile (64) of appends: k <= 896 fits the
    // 1024-key lists (list_capacity).
    if (k > 896) return fail(FBF_ERR_INVALID_ARG, "k must be <= 896");
    std::vector<int32_t> pos(k, -1), dist(k, INT32_MAX);
    if (npoints > 0) {
        const uint64_t bytes = (uint64_t)npoints * 112;
        const int8_t *dev = resolve_codes(codes_in, bytes);
        const int8_t *arr = nullptr;
        uint64_t arr_bytes = 0, offset = 0;
        if (dev && resolve_codes_array(codes_in, bytes, arr, arr_bytes, offset) && offset % 112 == 0) {
            // A slice of whole rows of a registered array: combined with the
            // other threads' calls.
            int d = 0;
            HIP_TRY(hipGetDevice(&d));
            if (d < 0 || d >= 64) return fail(FBF_ERR_UNSUPPORTED, "device index beyond the combiner table");
            BfReq r;
            r.arr = arr;
            r.arr_rows = arr_bytes / 112;
            r.start = offset / 112;
            r.n = npoints;
            r.q = queries_in;
            r.k = (uint32_t)k;
            r.pos = pos.data();
            r.dist = dist.data();
            fbf_status s = bf_combined(r, d);
            if (s != FBF_OK) return s;
        } else if (dev) {
            const uint64_t st0 = 0, cnt = npoints;
            fbf_status s = sym_topk(dev, npoints, &st0, &cnt, queries_in, 1, (uint32_t)k, pos.data(), dist.data());
            if (s != FBF_OK) return s;
        } else {
            // Not registered: stage this slice (correct, but a copy per call).
            ThreadEngine *Ep = nullptr;
            fbf_status se = thread_engine(Ep);
            if (se != FBF_OK) return se;
            ThreadEngine &E = *Ep;
            HIP_TRY(E.tmp_codes.reserve(bytes));
            HIP_TRY(hipMemcpyAsync(E.tmp_codes.p, codes_in, bytes, hipMemcpyHostToDevice, E.st));
            const uint64_t st0 = 0, cnt = npoints;
            fbf_status s = sym_topk(E.tmp_codes.as<int8_t>(), npoints, &st0, &cnt, queries_in, 1, (uint32_t)k,
                                    pos.data(), dist.data());
            if (s != FBF_OK) return s;
        }
    }
    for (size_t j = 0; j < k; ++j) {
        closest_points[j] = pos[j];
        if (distances) distances[j] = pos[j] < 0 ? HUGE_VALF : (float)dist[j] * alpha * alpha;
    }
    return FBF_OK;
}

extern "C" fbf_status fbf_bruteforce(size_t dim, size_t k, int *closest_points, size_t npoints,
                                     const char *codes_in, float alpha, size_t nqueries,
                                     const int8_t *queries_in) {
    (void)dim; (void)nqueries;
    return bruteforce_common(k, closest_points, nullptr, npoints, codes_in, alpha, queries_in);
}

extern "C" fbf_status fbf_bruteforce_subgraph(size_t dim, size_t k, int *closest_points, float *distances,
                                              size_t npoints, const char *codes_in, float alpha,
                                              size_t nqueries, const int8_t *queries_in) {
    (void)dim; (void)nqueries;
    if (!distances) return fail(FBF_ERR_INVALID_ARG, "null distances");
    return bruteforce_common(k, closest_points, distances, npoints, codes_in, alpha, queries_in);
}

extern "C" fbf_status fbf_bruteforce_batch(const char *codes_base, const uint64_t *starts,
                                           const uint64_t *counts, const int8_t *qcodes, uint64_t nq,
                                           uint32_t k, int32_t *out_pos, int32_t *out_dist) {
    if (!codes_base || !starts || !counts || !qcodes || !out_pos || !out_dist)
        return fail(FBF_ERR_INVALID_ARG, "null argument");
    if (k == 0 || k > 896) return fail(FBF_ERR_INVALID_ARG, "k must be in [1, 896]");
    if (nq == 0) return FBF_OK;
    uint64_t rows = 0;
    for (uint64_t i = 0; i < nq; ++i) rows = std::max(rows, starts[i] + counts[i]);
    const int8_t *dev = resolve_codes(codes_base, rows * 112);
    if (!dev) return fail(FBF_ERR_UNREGISTERED, "codes_base is not inside a registered code array");
    return sym_topk(dev, rows, starts, counts, qcodes, nq, k, out_pos, out_dist);
}

// ---------------------------------------------------------------- allocation-free batch
//
// fbf_bruteforce_batch_noalloc: fbf_bruteforce_batch for callers that may not
// allocate device memory.  Its plan, candidate lists, query codes and results
// live in one fixed device array (below); the caller passes the inputs and
// outputs, and the code array must already be registered.
constexpr uint64_t kNoAllocBytes = 64ull << 20;  // FBF_NOALLOC_SCRATCH_BYTES
static_assert(kNoAllocBytes == FBF_NOALLOC_SCRATCH_BYTES, "header and library disagree on the scratch size");
// The scratch: one copy per GPU the library runs on, held until the process
// ends.  Host code never passes this name to a kernel or a copy (on the host
// it is a stand-in, not a device address); it uses the per-device address
// that hipGetSymbolAddress returns.
static __device__ __align__(256) uint8_t scratch[kNoAllocBytes];
constexpr int kMaxDevices = 64;
static std::mutex g_noalloc_mu;  // one call at a time: every call on a GPU shares its scratch
static uint8_t *g_noalloc_ptr[kMaxDevices] = {};

// Device address of the scratch on the current device (looked up once per device).
static fbf_status noalloc_scratch(int dev, uint8_t *&out) {
    if (dev < 0 || dev >= kMaxDevices) return fail(FBF_ERR_UNSUPPORTED, "device index beyond the scratch table");
    if (!g_noalloc_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || p == nullptr)
            return fail(FBF_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_noalloc_ptr[dev] = static_cast<uint8_t *>(p);
    }
    out = g_noalloc_ptr[dev];
    return FBF_OK;
}

extern "C" fbf_status fbf_bruteforce_batch_noalloc(const char *codes_base, const uint64_t *starts,
                                                   const uint64_t *counts, const int8_t *qcodes, uint64_t nq,
                                                   uint32_t k, int32_t *out_pos, int32_t *out_dist, void *stream) {
    if (!codes_base || !starts || !counts || !qcodes || !out_pos || !out_dist)
        return fail(FBF_ERR_INVALID_ARG, "null argument");
    if (k == 0 || k > 896) return fail(FBF_ERR_INVALID_ARG, "k must be in [1, 896]");
    if (nq == 0) return FBF_OK;
    uint64_t rows = 0;
    for (uint64_t i = 0; i < nq; ++i) rows = std::max(rows, starts[i] + counts[i]);
    const int8_t *dev_codes = resolve_codes(codes_base, rows * 112);
    if (!dev_codes) return fail(FBF_ERR_UNREGISTERED, "codes_base is not inside a registered code array");
    Plan P;
    fbf_status s = sym_plan(starts, counts, nq, rows, P);
    if (s != FBF_OK) return s;
    const int B = list_capacity((int)k);
    // Carve the scratch: 256-byte aligned pieces.
    const uint64_t n_slots = P.slot_off[nq];
    uint64_t off = 0;
    auto take = [&](uint64_t bytes) {
        const uint64_t o = off;
        off += (bytes + 255) / 256 * 256;
        return o;
    };
    const uint64_t o_items = take(P.items.size() * sizeof(WorkItem));
    const uint64_t o_iq = take(P.iq.size() * sizeof(ItemQuery));
    const uint64_t o_slot = take((nq + 1) * sizeof(uint64_t));
    const uint64_t o_lists = take(n_slots * (uint64_t)B * sizeof(uint64_t));
    const uint64_t o_counts = take(n_slots * sizeof(uint32_t));
    const uint64_t o_q = take(nq * 112);
    const uint64_t o_ids = take(nq * (uint64_t)k * 4);
    const uint64_t o_d = take(nq * (uint64_t)k * 4);
    if (off > kNoAllocBytes)
        return fail(FBF_ERR_INVALID_ARG, "batch needs " + std::to_string(off) + " bytes of scratch; the allocation-free "
                                         "entry point holds " + std::to_string(kNoAllocBytes) +
                                         " (FBF_NOALLOC_SCRATCH_BYTES): split the batch");
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    uint8_t *base = nullptr;
    s = noalloc_scratch(dev, base);
    if (s != FBF_OK) return s;
    hipStream_t st = (hipStream_t)stream;
    SymBufs b;
    b.items = reinterpret_cast<WorkItem *>(base + o_items);
    b.iq = reinterpret_cast<ItemQuery *>(base + o_iq);
    b.slot_off = reinterpret_cast<uint64_t *>(base + o_slot);
    b.lists = reinterpret_cast<uint64_t *>(base + o_lists);
    b.counts = reinterpret_cast<uint32_t *>(base + o_counts);
    b.qcodes = reinterpret_cast<int8_t *>(base + o_q);
    b.out_ids = reinterpret_cast<uint32_t *>(base + o_ids);
    b.out_d = reinterpret_cast<float *>(base + o_d);
    // Nothing in the scratch needs clearing: each work item writes the count
    // of every list slot it owns, and the merge reads only those.
    if (!P.items.empty())
        HIP_TRY(hipMemcpyAsync(b.items, P.items.data(), P.items.size() * sizeof(WorkItem), hipMemcpyHostToDevice, st));
    if (!P.iq.empty())
        HIP_TRY(hipMemcpyAsync(b.iq, P.iq.data(), P.iq.size() * sizeof(ItemQuery), hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(b.slot_off, P.slot_off.data(), (nq + 1) * sizeof(uint64_t), hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(b.qcodes, qcodes, nq * 112, hipMemcpyHostToDevice, st));
    s = sym_launch(dev_codes, rows, P, nq, k, B, b, st);
    if (s != FBF_OK) return s;
    // Results straight into the caller's arrays, then made slice-relative there.
    HIP_TRY(hipMemcpyAsync(out_pos, b.out_ids, nq * (uint64_t)k * 4, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(out_dist, b.out_d, nq * (uint64_t)k * 4, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    sym_unpack(reinterpret_cast<const uint32_t *>(out_pos), reinterpret_cast<const uint32_t *>(out_dist), starts, nq, k,
               out_pos, out_dist);
    return FBF_OK;
}

// ---------------------------------------------------------------- type-2 sub-interval path
//
// fbf_search_subinterval: the SMALL-bucket branch of the router's type-2
// search (hybrid_graph.cpp:317-420) and its global refine (:463-494) for a
// batch.  Per query, SplitInterval over the time buckets (host, O(10 log n));
// every SMALL (query, bucket) slice becomes one slice of the buckets' code
// array, scanned by the symmetric int8 kernel with the query encoded by that
// bucket's scale and cut to its bucket_k best; the slices' results (and any
// caller candidates) form each query's candidate pool of pool_k, and the
// merge reranks the pool with the canonical FP32 distance.

// SplitInterval (hybrid_graph.cpp:91-134) of [l, r]: types[t] 255 (bucket
// absent or not overlapped), 0 SMALL, 1 MEDIUM, 3 FULL; ls / ln the local
// start and count of a partial bucket.  thr: the SMALL threshold.
static void split_interval(const fbf_dataset *ds, float l, float r, float thr, uint8_t *types, uint64_t *ls,
                           uint64_t *ln) {
    const float *ts = ds->ts_t.data();
    for (int t = 0; t < 10; ++t) {
        types[t] = 255;
        ls[t] = ln[t] = 0;
        const uint64_t bn = ds->tb_count[t];
        if (bn == 0) continue;
        const float b_lo = (float)t / 10, b_hi = (float)(t + 1) / 10;
        if (r < b_lo || l > b_hi) continue;
        if (l <= b_lo && r >= b_hi) {
            types[t] = 3;
            ln[t] = bn;
            continue;
        }
        const float lb = std::max(b_lo, l), rb = std::min(b_hi, r);
        const float *bt = ts + ds->tb_start[t];
        const uint64_t s = std::lower_bound(bt, bt + bn, lb) - bt;
        const uint64_t e = std::max<uint64_t>(s, std::upper_bound(bt, bt + bn, rb) - bt);
        types[t] = (float)(e - s) / (float)bn < thr ? 0 : 1;
        ls[t] = s;
        ln[t] = e - s;
    }
}

extern "C" fbf_status fbf_search_subinterval(fbf_dataset *ds, const float *queries_dev, const float *metas,
                                             uint64_t nq, uint32_t k, const fbf_subinterval_opts *opts,
                                             const uint32_t *ext_pos_dev, const float *ext_dist_dev, uint32_t n_ext,
                                             uint32_t *out_ids_dev, float *out_dists_dev, uint8_t *bucket_types,
                                             void *stream) {
    if (!ds || !metas || (nq && (!queries_dev || !out_ids_dev || !out_dists_dev)))
        return fail(FBF_ERR_INVALID_ARG, "null argument");
    if (!(ds->repr & FBF_REPR_TBUCKETS)) return fail(FBF_ERR_NO_REPR, "dataset was created without FBF_REPR_TBUCKETS");
    if (k == 0 || k > 512) return fail(FBF_ERR_INVALID_ARG, "k must be in [1, 512]");
    if (n_ext && (!ext_pos_dev || !ext_dist_dev)) return fail(FBF_ERR_INVALID_ARG, "null external candidates");
    if (n_ext > 896) return fail(FBF_ERR_INVALID_ARG, "n_ext must be <= 896");
    fbf_subinterval_opts o{};
    if (opts) o = *opts;
    const uint32_t kb = o.bucket_k ? o.bucket_k : 140u;   // bf_refine_k, hybrid_graph.h:80
    const uint32_t kp = o.pool_k ? o.pool_k : 150u;       // global_graph_refine_k, hybrid_graph.h:84
    if (kb > 896 || kp > 896) return fail(FBF_ERR_INVALID_ARG, "bucket_k and pool_k must be <= 896");
    if (!(o.small_thr >= 0.f)) return fail(FBF_ERR_INVALID_ARG, "small_thr must be >= 0");
    if (nq == 0) return FBF_OK;
    // Split every query on the host.
    std::vector<SliceInfo> sl;
    std::vector<uint64_t> starts, counts;
    std::vector<uint64_t> pool_off(nq + 1, 0);
    std::vector<uint32_t> ext_slot(nq, 0);
    std::vector<uint8_t> types(nq * 10);
    uint64_t slot = 0;
    for (uint64_t q = 0; q < nq; ++q) {
        const float *m = metas + 4 * q;
        if (!(m[0] == 2.0f)) return fail(FBF_ERR_INVALID_ARG, "query " + std::to_string(q) + ": not a type-2 filter");
        const float l = m[2], r = m[3];
        // The query's selectivity over the timestamp layout picks the SMALL
        // threshold (hybrid_graph.cpp:324-333): 0.5 below 0.2, else 0.2.
        const float *ts = ds->ts_t.data();
        const uint64_t gs = std::lower_bound(ts, ts + ds->n, l) - ts;
        const uint64_t ge = std::max<uint64_t>(gs, std::upper_bound(ts, ts + ds->n, r) - ts);
        const float sel = (float)(ge - gs) / (float)ds->n;
        const float thr = o.small_thr > 0.f ? o.small_thr : ((double)sel < 0.2 ? 0.5f : 0.2f);
        uint64_t ls[10], ln[10];
        split_interval(ds, l, r, thr, &types[q * 10], ls, ln);
        pool_off[q] = slot;
        for (int t = 0; t < 10; ++t) {
            if (types[q * 10 + t] != 0 || ln[t] == 0) continue;
            SliceInfo S{};
            S.q = (uint32_t)q;
            S.alpha = ds->tb_alpha[t];
            S.row_to_pos = (int64_t)ds->tb_start[t] - (int64_t)ds->tb_off[t];
            S.pool_slot = (uint32_t)slot++;
            sl.push_back(S);
            starts.push_back(ds->tb_off[t] + ls[t]);
            counts.push_back(ln[t]);
        }
        if (n_ext) ext_slot[q] = (uint32_t)slot++;
    }
    pool_off[nq] = slot;
    if (bucket_types) std::memcpy(bucket_types, types.data(), types.size());
    const uint64_t nsl = sl.size();
    const uint32_t stride = std::max(kb, n_ext);

    std::lock_guard<std::mutex> lock(ds->mu);
    HIP_TRY(hipSetDevice(ds->device));
    hipStream_t st = (hipStream_t)stream;
    Scratch &S = ds->scr;
    struct DoneRecord {
        hipEvent_t ev = nullptr;
        hipStream_t st = nullptr;
        ~DoneRecord() {
            if (ev) (void)hipEventRecord(ev, st);
        }
    } done_rec;
    if (!S.ev_done) HIP_TRY(hipEventCreateWithFlags(&S.ev_done, hipEventDisableTiming));
    else if (S.used) HIP_TRY(hipStreamWaitEvent(st, S.ev_done, 0));
    S.used = true;
    done_rec.ev = S.ev_done;
    done_rec.st = st;

    // The slices' scan: plan, query codes, scan and per-slice merge to kb.
    Plan P;
    fbf_status s = sym_plan(starts.data(), counts.data(), nsl, ds->tb_rows, P);
    if (s != FBF_OK) return s;
    const int B = list_capacity((int)kb);
    s = upload_plan(S, P, nsl, B, st);
    if (s != FBF_OK) return s;
    HIP_TRY(S.d_q.reserve(std::max<uint64_t>(1, nsl) * 112));
    HIP_TRY(S.d_out_ids.reserve(std::max<uint64_t>(1, nsl) * kb * 4));
    HIP_TRY(S.d_out_d.reserve(std::max<uint64_t>(1, nsl) * kb * 4));
    HIP_TRY(S.d_ti_sl.reserve(std::max<uint64_t>(1, nsl) * sizeof(SliceInfo)));
    HIP_TRY(S.d_ti_lists.reserve(std::max<uint64_t>(1, slot) * stride * sizeof(uint64_t)));
    HIP_TRY(S.d_ti_counts.reserve(std::max<uint64_t>(1, slot) * sizeof(uint32_t)));
    HIP_TRY(S.d_ti_off.reserve((nq + 1) * sizeof(uint64_t)));
    HIP_TRY(S.d_ti_ext.reserve(nq * sizeof(uint32_t)));
    // Slices, pool offsets and ext slots through one pinned staging buffer,
    // refilled only after the copies of the previous call are done.
    const size_t b_sl = nsl * sizeof(SliceInfo), b_off = (nq + 1) * sizeof(uint64_t), b_ext = nq * sizeof(uint32_t);
    if (S.ev_ti) HIP_TRY(hipEventSynchronize(S.ev_ti));
    else HIP_TRY(hipEventCreateWithFlags(&S.ev_ti, hipEventDisableTiming));
    HIP_TRY(S.h_ti.reserve(b_sl + b_off + b_ext + 64));
    char *h = (char *)S.h_ti.p;
    if (b_sl) std::memcpy(h, sl.data(), b_sl);
    std::memcpy(h + b_sl, pool_off.data(), b_off);
    std::memcpy(h + b_sl + b_off, ext_slot.data(), b_ext);
    if (b_sl) HIP_TRY(hipMemcpyAsync(S.d_ti_sl.p, h, b_sl, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(S.d_ti_off.p, h + b_sl, b_off, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(S.d_ti_ext.p, h + b_sl + b_off, b_ext, hipMemcpyHostToDevice, st));
    HIP_TRY(hipEventRecord(S.ev_ti, st));
    if (nsl) {
        HIP_TRY(launch_sym_encode_slices(queries_dev, S.d_ti_sl.as<SliceInfo>(), nsl, S.d_q.as<int8_t>(), st));
        SymBufs b;
        b.items = S.d_items.as<WorkItem>();
        b.iq = S.d_iq.as<ItemQuery>();
        b.slot_off = S.d_slot_off.as<uint64_t>();
        b.lists = S.d_lists.as<uint64_t>();
        b.counts = S.d_counts.as<uint32_t>();
        b.qcodes = S.d_q.as<int8_t>();
        b.out_ids = S.d_out_ids.as<uint32_t>();
        b.out_d = S.d_out_d.as<float>();
        s = sym_launch(ds->tb_codes, ds->tb_rows, P, nsl, kb, B, b, st);
        if (s != FBF_OK) return s;
        HIP_TRY(launch_pool_keys(S.d_out_ids.as<uint32_t>(), S.d_out_d.as<uint32_t>(), S.d_ti_sl.as<SliceInfo>(), nsl,
                                 kb, stride, S.d_ti_lists.as<uint64_t>(), S.d_ti_counts.as<uint32_t>(), st));
    }
    if (n_ext)
        HIP_TRY(launch_ext_keys(ext_pos_dev, ext_dist_dev, S.d_ti_ext.as<uint32_t>(), nq, n_ext, ds->n, stride,
                                S.d_ti_lists.as<uint64_t>(), S.d_ti_counts.as<uint32_t>(), st));
    // The pool (pool_k smallest keys over the query's slots) and the global
    // refine (canonical FP32, k smallest (distance, original id)).
    MergeArgs m{};
    m.slot_off = S.d_ti_off.as<uint64_t>();
    m.lists = S.d_ti_lists.as<uint64_t>();
    m.counts = S.d_ti_counts.as<uint32_t>();
    m.list_cap = stride;
    m.nq = nq;
    m.k = k;
    m.kprime = kp;
    m.queries = queries_dev;
    m.base_rows = ds->base_rows;
    m.out_ids = out_ids_dev;
    m.out_dists = out_dists_dev;
    m.rerank = 1;
    m.key_int = 0;
    m.key_pos = 1;
    m.ids_lay[0] = ds->lay[0].ids;
    m.ids_lay[1] = ds->lay[1].ids;
    HIP_TRY(launch_merge(m, st));
    return FBF_OK;
}
