// Device PMA store: per-vertex adjacency ranges with trailing gaps, updated by
// batches. Reference: host::PMAGraph (groute/graphs/csr_graph.cuh:315-979).
//
// Layout (SoA in HBM):
//   off[nv + 1]  u64 first slot of each vertex; off[nv] = capacity
//   deg[nv]      u32 live entries; they occupy [off[v], off[v] + deg[v])
//   col[cap]     u32 targets; kGap in [off[v] + deg[v], off[v + 1])
// The reference stores the same (index, degree) pair (vertex_sync_element,
// csr_graph.cuh:301-304) with 0xFFFFFFFF gaps.
//
// A batch is applied in two steps:
//   prepare  split + sort the pairs, list the touched source vertices, count
//            each touched vertex's new degree (inserts appended, then every
//            deleted target removed — csr_graph.cuh:658-811, framework.cuh:1001-1023)
//   commit   vertices whose new list fits their range are merged in place;
//            otherwise the smallest window of PMA leaves (up the implicit tree,
//            csr_graph.cuh:813-863) whose density stays under its threshold is
//            re-spread in proportion to degree + 1 (calculate_positions_V1,
//            csr_graph.cuh:543-560); if even the root is too dense the
//            capacity doubles (resize, csr_graph.cuh:935-958).
#ifndef SG_PMA_H
#define SG_PMA_H

#include "sg_kernels.h"

namespace sg {

// Per-touched-vertex ranges into the sorted batch.
struct BatchView {
    const uint32_t *ins_key;   // sorted sources (kGap = dropped pair)
    const uint32_t *ins_val;   // targets, file order within a source
    uint32_t nins;
    const uint64_t *del_key;   // sorted (u << 32 | v); ~0 = dropped pair
    uint32_t ndel;
};

__device__ __forceinline__ void equal_range_u32(const uint32_t *a, uint32_t n, uint32_t key, uint32_t &b, uint32_t &e) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) { uint32_t m = (lo + hi) >> 1; if (a[m] < key) lo = m + 1; else hi = m; }
    b = lo;
    hi = n;
    while (lo < hi) { uint32_t m = (lo + hi) >> 1; if (a[m] <= key) lo = m + 1; else hi = m; }
    e = lo;
}
__device__ __forceinline__ void src_range_u64(const uint64_t *a, uint32_t n, uint32_t u, uint32_t &b, uint32_t &e) {
    const uint64_t k0 = (uint64_t)u << 32, k1 = k0 | 0xFFFFFFFFull;
    uint32_t lo = 0, hi = n;
    while (lo < hi) { uint32_t m = (lo + hi) >> 1; if (a[m] < k0) lo = m + 1; else hi = m; }
    b = lo;
    hi = n;
    while (lo < hi) { uint32_t m = (lo + hi) >> 1; if (a[m] <= k1) lo = m + 1; else hi = m; }
    e = lo;
}
__device__ __forceinline__ bool in_sorted_u64(const uint64_t *a, uint32_t b, uint32_t e, uint64_t key) {
    uint32_t lo = b, hi = e;
    while (lo < hi) { uint32_t m = (lo + hi) >> 1; if (a[m] < key) lo = m + 1; else hi = m; }
    return lo < e && a[lo] == key;
}

// Writes the merged list of u (survivors of the old list, then surviving
// inserts) to dst[0, ...) with one wave; returns the count (wave-uniform).
// When in_place, dst == old list start and survivors only move left.
__device__ __forceinline__ uint32_t wave_merge_list(const uint32_t *old, uint32_t deg, const BatchView &bv,
                                                    uint32_t u, uint32_t ib, uint32_t ie, uint32_t db, uint32_t de,
                                                    uint32_t *dst, bool skip_old) {
    const uint32_t lane = lane_id();
    const uint64_t k0 = (uint64_t)u << 32;
    uint32_t w = 0;
    if (!skip_old) {
        for (uint32_t k = 0; k < deg; k += kWave) {
            const uint32_t i = k + lane;
            bool keep = false;
            uint32_t x = 0;
            if (i < deg) {
                x = old[i];
                keep = (db == de) || !in_sorted_u64(bv.del_key, db, de, k0 | x);
            }
            const uint64_t m = __ballot(keep);
            __builtin_amdgcn_wave_barrier();
            if (keep) dst[w + __popcll(m & lanemask_lt())] = x;
            w += (uint32_t)__popcll(m);
        }
    } else {
        w = deg;
    }
    for (uint32_t k = ib; k < ie; k += kWave) {
        const uint32_t i = k + lane;
        bool keep = false;
        uint32_t y = 0;
        if (i < ie) {
            y = bv.ins_val[i];
            keep = (db == de) || !in_sorted_u64(bv.del_key, db, de, k0 | y);
        }
        const uint64_t m = __ballot(keep);
        if (keep) dst[w + __popcll(m & lanemask_lt())] = y;
        w += (uint32_t)__popcll(m);
    }
    return w;
}

}  // namespace sg

#endif
