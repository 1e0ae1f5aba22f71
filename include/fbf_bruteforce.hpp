// fbf_bruteforce.hpp -- header-only C++ shims with the reference's exact
// brute-force signatures, forwarding to the C-ABI of libfbf.so.
//
//   bruteforce          replaces bruteforce.h:244-265
//   bruteforce_subgraph replaces bruteforce.h:267-289
//
// The reference has no error channel, so a nonzero status aborts with the
// library's message.  Results: closest_points[j] is a position relative to
// codes_in; the k positions come in ascending (int distance, position) order
// (the reference emits the same set in heap-array order; the set differs only
// among codes tied at the k-th distance, where the smaller position wins).
// k > npoints pads with -1 (the reference reads out of bounds there).
//
// Register the code arrays once (fbf_register_codes) so calls are zero-copy;
// an unregistered codes_in is uploaded on every call.
//
// With FBF_REPLACE_REFERENCE_BRUTEFORCE defined, the shims are also declared
// in the global namespace under the reference names, so hybrid_graph.cpp's
// unqualified calls (hybrid_graph.cpp:211, :410) bind to them once the two
// definitions are removed from bruteforce.h.
#pragma once

#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstdlib>

#include "fbf.h"

namespace fbf_ref {

inline void check_or_abort(fbf_status s, const char *what) {
    if (s != FBF_OK) {
        std::fprintf(stderr, "%s failed (status %d): %s\n", what, (int)s, fbf_last_error());
        std::abort();
    }
}

inline void bruteforce(const size_t dim, const size_t k, int *const closest_points, size_t npoints,
                       char *codes_in, float alpha, size_t nqueries, int8_t *queries_in) {
    check_or_abort(fbf_bruteforce(dim, k, closest_points, npoints, codes_in, alpha, nqueries, queries_in),
                   "fbf_bruteforce");
}

inline void bruteforce_subgraph(const size_t dim, const size_t k, int *const closest_points,
                                float *distances, size_t npoints, char *codes_in, float alpha,
                                size_t nqueries, int8_t *queries_in) {
    check_or_abort(fbf_bruteforce_subgraph(dim, k, closest_points, distances, npoints, codes_in, alpha,
                                           nqueries, queries_in),
                   "fbf_bruteforce_subgraph");
}

}  // namespace fbf_ref

#ifdef FBF_REPLACE_REFERENCE_BRUTEFORCE
inline void bruteforce(const size_t dim, const size_t k, int *const closest_points, size_t npoints,
                       char *codes_in, float alpha, size_t nqueries, int8_t *queries_in) {
    fbf_ref::bruteforce(dim, k, closest_points, npoints, codes_in, alpha, nqueries, queries_in);
}
inline void bruteforce_subgraph(const size_t dim, const size_t k, int *const closest_points,
                                float *distances, size_t npoints, char *codes_in, float alpha,
                                size_t nqueries, int8_t *queries_in) {
    fbf_ref::bruteforce_subgraph(dim, k, closest_points, distances, npoints, codes_in, alpha, nqueries,
                                 queries_in);
}
#endif
