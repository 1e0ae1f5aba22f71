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

#include "bruteforce.h"                     // bruteforce, bruteforce_subgraph, maxPQIFCS
// distance.h (normal_l2) arrives through utils.h, which has no include guard for it
#include "glass/quant/sq8_quant.hpp"        // SQ8Quantizer
#include "glass/quant/sq8_sym_quant.hpp"    // SQ8SymmetricQuantizer
#include "utils.h"                          // Bruteforce, GetKNNRecall

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

}  // extern "C"
