// ref_shim.cpp -- C-ABI around the reference's own header-only kernels.
//
// TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile (target "ref") against
// the reference sources where they lie, into oracle/_ref/libfbf_ref.so.  It
// contains no reference source: it includes the reference headers at compile
// time and exports thin wrappers so tests and the CPU baseline can call the
// reference's kernels directly.  Never linked into the product library.
#include <cstdint>
#include <cstring>
#include <vector>

// hybrid_graph.cpp brings hybrid_graph.h (HybridGraph::SplitInterval, the
// timestamp searchers), bruteforce.h (bruteforce, bruteforce_subgraph,
// maxPQIFCS) and utils.h (Bruteforce, GetKNNRecall, RFType; distance.h's
// normal_l2 arrives through it) -- compiled from the reference's own files.
#include "hybrid_graph.cpp"
#include "glass/quant/sq8_quant.hpp"        // SQ8Quantizer
#include "glass/quant/sq8_sym_quant.hpp"    // SQ8SymmetricQuantizer

extern "C" {

int ref_l2_sq8sym(const int8_t *x, const int8_t *y) {
    return glass::L2SqrSQ8_sym(x, y, 112);
}

float ref_l2_ref(const float *x, const float *y, int d) { return glass::L2SqrRef(x, y, d); }

float ref_normal_l2(const float *x, const float *y, int d) { return normal_l2(x, y, (unsigned)d); }

// x and mi/dif must hold 112 floats (zero padded) as the reference reads 112 lanes.
float ref_l2_sq8ext(const float *x, const uint8_t *y, const float *mi, const float *dif) {
    return glass::L2SqrSQ8_ext(x, y, 112, mi, dif);
}

// bruteforce.h:244-265 as the router calls it (one query).
void ref_bruteforce(uint64_t k, int32_t *closest_points, uint64_t npoints, const int8_t *codes,
                    float alpha, const int8_t *qcode) {
    bruteforce(100, k, closest_points, npoints, (char *)codes, alpha, 1, (int8_t *)qcode);
}

// bruteforce.h:267-289.
void ref_bruteforce_subgraph(uint64_t k, int32_t *closest_points, float *distances,
                             uint64_t npoints, const int8_t *codes, float alpha,
                             const int8_t *qcode) {
    bruteforce_subgraph(100, k, closest_points, distances, npoints, (char *)codes, alpha, 1,
                        (int8_t *)qcode);
}

// SQ8SymmetricQuantizer::train + encode (sq8_sym_quant.hpp:32-74); codes n x 112.
float ref_sq8sym_train_encode(const float *data, int n, int d, int8_t *codes_out) {
    glass::SQ8SymmetricQuantizer<glass::Metric::L2> q(d);
    q.train(data, n);
    std::memcpy(codes_out, q.codes, (size_t)n * q.code_size);
    return q.alpha;
}

void ref_sq8sym_encode(const float *x, int d, float alpha, int8_t *code_out) {
    glass::SQ8SymmetricQuantizer<glass::Metric::L2> q(d);
    q.alpha = alpha;
    std::memset(code_out, 0, q.code_size);
    q.encode(x, (char *)code_out);
}

// SQ8Quantizer::train + encode (sq8_quant.hpp:30-63); codes n x 112, mi/dif 112.
void ref_sq8_train_encode(const float *data, int n, int d, float *mi, float *dif,
                          uint8_t *codes_out) {
    glass::SQ8Quantizer<glass::Metric::L2> q(d);
    q.train(data, n);
    std::memcpy(codes_out, q.codes, (size_t)n * q.code_size);
    std::memcpy(mi, q.mi.data(), sizeof(float) * q.d_align);
    std::memcpy(dif, q.dif.data(), sizeof(float) * q.d_align);
}

// utils.h:80-158 ground truth.  nodes: n x (2 + d) [label, ts, vec]; queries:
// nq x (4 + d) [type, v, l, r, vec]; gt: nq x 100.  Every query must match at
// least 100 rows (the reference throws otherwise).
void ref_ground_truth(const float *nodes, uint32_t n, const float *queries, uint32_t nq,
                      int d, uint32_t *gt_out) {
    std::vector<std::vector<float>> nv(n, std::vector<float>(d + 2));
    for (uint32_t i = 0; i < n; ++i) std::memcpy(nv[i].data(), nodes + (size_t)i * (d + 2), 4 * (d + 2));
    std::vector<std::vector<float>> qv(nq, std::vector<float>(d + 4));
    for (uint32_t i = 0; i < nq; ++i) std::memcpy(qv[i].data(), queries + (size_t)i * (d + 4), 4 * (d + 4));
    std::vector<std::vector<uint32_t>> gt;
    Bruteforce(nv, qv, gt);
    for (uint32_t i = 0; i < nq; ++i) std::memcpy(gt_out + (size_t)i * 100, gt[i].data(), 400);
}

// The router's brute-force branch (hybrid_graph.cpp:203-228) driven with the
// reference kernels over presorted contiguous ranges, one OpenMP task per
// query as at hybrid_graph.cpp:168.  ranges: nq x [layout, s, e]; codes_b /
// codes_t are the 112-byte codes in each layout order; id_b / id_t map
// positions to original ids; base is row-major in original id order.
// Output ids follow the reference's heap-array order.
void ref_router_bf(const int8_t *codes_b, const int8_t *codes_t, const uint32_t *id_b,
                   const uint32_t *id_t, const float *base, int d, float alpha,
                   const float *queries, const int8_t *qcodes, const int64_t *ranges,
                   uint32_t nq, uint32_t kq, uint32_t k, uint32_t *out_ids) {
#pragma omp parallel for schedule(dynamic)
    for (int64_t i = 0; i < (int64_t)nq; ++i) {
        int64_t lay = ranges[i * 3], s = ranges[i * 3 + 1], e = ranges[i * 3 + 2];
        const int8_t *codes = (lay == 0 ? codes_b : codes_t) + s * 112;
        const uint32_t *idm = (lay == 0 ? id_b : id_t) + s;
        size_t K = std::min<size_t>(kq, (size_t)(e - s));
        std::vector<int> tmp(K);
        if (K > 0)
            bruteforce(d, K, tmp.data(), e - s, (char *)codes, alpha, 1,
                       (int8_t *)(qcodes + i * 112));
        for (size_t j = 0; j < K; ++j) tmp[j] = (int)idm[tmp[j]];
        maxPQIFCS<float> pd(k);
        const float *q = queries + (size_t)i * d;
        for (size_t j = 0; j < K; ++j)
            pd.maybe_pop_emplace(tmp[j], glass::L2SqrRef(q, base + (size_t)tmp[j] * d, d));
        for (size_t l = 0; l < k; ++l)
            out_ids[i * k + k - l - 1] = (l < (size_t)pd.size()) ? (uint32_t)pd.data_[l + 1].id : 0xFFFFFFFFu;
    }
}

// The type-2 sub-interval route of BatchSearch for the SMALL buckets, driven
// with the reference's own pieces: the 0.1-wide time buckets of
// ReadSortedBaseTimestamp (utils.h:470-478), one timestamp searcher per
// bucket whose SetData trains that bucket's SQ8SymmetricQuantizer
// (baseline.cpp:130-141, searcher.hpp:94-102), HybridGraph::SplitInterval
// with the router's threshold rule (hybrid_graph.cpp:324-333), the SMALL
// branch -- bucket-encoded query, bruteforce_subgraph, candidate_pool
// maybe_pop_emplace (hybrid_graph.cpp:389-419) -- in bucket order t = 0..9,
// and the global refine with L2SqrRef into maxPQIFCS<float> (:463-494).
// FULL and MEDIUM buckets (graph searches there) add nothing.
// vecs_t / ts_t / ids_t: the rows, timestamps and original ids in timestamp
// order; nb: the selectivity denominator.  out_ids: nq x k original ids in
// heap-array order (0xFFFFFFFF where the refine heap holds fewer); out_types:
// nq x 10 RFType per bucket (255: no overlap or no bucket); alphas: 10.
void ref_subinterval_small(const float *vecs_t, const float *ts_t, const uint32_t *ids_t, uint32_t n, int d,
                           const float *queries, const float *metas, uint32_t nq, uint32_t k, uint32_t nb,
                           uint32_t *out_ids, int32_t *out_types, float *alphas) {
    HybridGraph h;
    std::vector<float> tsv(ts_t, ts_t + n);
    for (int i = 0; i <= 9; i += interval) {
        auto s = std::lower_bound(tsv.begin(), tsv.end(), 0.1 * i);
        auto e = std::upper_bound(tsv.begin(), tsv.end(), 0.1 * (i + interval));
        int num = (int)(e - s);
        alphas[i] = 0.0f;
        if (num == 0) continue;
        h.timestamp_map[i] = {(uint32_t)(s - tsv.begin()), (uint32_t)num};
    }
    for (auto &[t, p] : h.timestamp_map) {
        glass::Graph<int> graph;
        h.timestamp_searcher[t] = glass::create_searcher(std::move(graph), "L2", 3);
        h.timestamp_searcher[t]->SetData(vecs_t + (size_t)p.first * d, nullptr, tsv.data() + p.first, p.second, d);
        alphas[t] = ((glass::Searcher<glass::SQ8SymmetricQuantizer<glass::Metric::L2>> *)h.timestamp_searcher[t].get())
                        ->quant.alpha;
    }
    std::vector<std::unordered_map<int, std::pair<RFType, std::pair<int, int>>>> pool_iv(nq);
    for (uint32_t i = 0; i < nq; ++i) {
        const float l = metas[i * 4 + 2], r = metas[i * 4 + 3];
        auto s = std::lower_bound(tsv.begin(), tsv.end(), l);
        auto e = std::upper_bound(tsv.begin(), tsv.end(), r);
        const float sel = (float)(e - s) / (float)nb;
        if (sel < 0.2) h.SplitInterval(std::make_pair(l, r), pool_iv[i], 0.5);
        else h.SplitInterval(std::make_pair(l, r), pool_iv[i]);
        for (int t = 0; t < 10; ++t) {
            auto it = pool_iv[i].find(t);
            out_types[i * 10 + t] = it == pool_iv[i].end() ? 255 : (int)it->second.first;
        }
    }
    std::vector<maxPQIFCS<int>> candidate_pool(nq);
    for (int t = 0; t <= 9; ++t) {
        auto itm = h.timestamp_map.find(t);
        if (itm == h.timestamp_map.end()) continue;
        auto &quant = ((glass::Searcher<glass::SQ8SymmetricQuantizer<glass::Metric::L2>> *)h.timestamp_searcher[t].get())
                          ->quant;
        const uint32_t start_id = itm->second.first;
#pragma omp parallel for schedule(dynamic)
        for (int64_t i = 0; i < (int64_t)nq; ++i) {
            auto it = pool_iv[i].find(t);
            if (it == pool_iv[i].end() || it->second.first != RFType::SMALL) continue;
            if (candidate_pool[i].size() == 0) candidate_pool[i].resize(h.global_graph_refine_k);
            const int s = it->second.second.first;
            const uint32_t n_points = it->second.second.second;
            int8_t *encode_q = (int8_t *)glass::alloc64B(quant.code_size);
            quant.encode(queries + (size_t)i * d, (char *)encode_q);
            char *codes = quant.codes + (size_t)s * quant.code_size;
            auto K = std::min(h.bf_refine_k, n_points);
            std::vector<int> tmp(K);
            std::vector<float> dists(K);
            bruteforce_subgraph(d, K, tmp.data(), dists.data(), n_points, codes, quant.alpha, 1, encode_q);
            for (size_t j = 0; j < K; ++j) candidate_pool[i].maybe_pop_emplace(tmp[j] + start_id + s, dists[j]);
            free(encode_q);
        }
    }
#pragma omp parallel for schedule(dynamic)
    for (int64_t i = 0; i < (int64_t)nq; ++i) {
        auto &cd = candidate_pool[i];
        maxPQIFCS<float> point_dist(k);
        for (int j = 1; j <= cd.size(); ++j) {
            const int id = cd.data_[j].id;
            point_dist.maybe_pop_emplace(id, glass::L2SqrRef(queries + (size_t)i * d, vecs_t + (size_t)id * d, d));
        }
        for (uint32_t l = 0; l < k; ++l)
            out_ids[i * k + k - l - 1] = l < (uint32_t)point_dist.size() ? ids_t[point_dist.data_[l + 1].id] : 0xFFFFFFFFu;
    }
}

}  // extern "C"
