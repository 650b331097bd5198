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
// A batch (inserts first, then deletes; every copy of a deleted (u, v) goes —
// csr_graph.cuh:658-811, framework.cuh:1001-1023) is applied in chunks so that
// no wave walks a whole hub list:
//   touched   sources with an insert or a delete (run heads of the sorted batch)
//   info      per touched vertex: batch ranges and its chunk descriptors — the
//             old list in kMergeChunk pieces when it has deletes, and the
//             inserts in kMergeChunk pieces
//   count     one wave per chunk counts its survivors (delete set staged in LDS)
//   scan      exclusive prefix of the counts = each chunk's output position
//   newdeg    per vertex; a vertex that outgrows its range marks its PMA leaf
//   write     one wave per chunk writes survivors, in order, to a scratch array
//   place     one wave per chunk copies them back into the vertex's range and
//             turns the freed tail into gaps
// Vertices that outgrow their range are moved by the window re-spread of the
// engine (rebalance_wrapper / rebalance_weighted / resize, csr_graph.cuh:813-958).
#ifndef SG_PMA_H
#define SG_PMA_H

#include "sg_kernels.h"

namespace sg {

constexpr uint32_t kMergeChunk = 2048;  // entries per merge chunk
constexpr uint32_t kDelStage = 1024;    // delete targets of a vertex staged in LDS
constexpr uint32_t kLaneIns = 128;      // insert-only vertices with at most this many inserts: one lane each

// Sorted batch.
struct BatchView {
    const uint32_t *ins_key;   // sorted sources (kGap = dropped pair)
    const uint32_t *ins_val;   // targets, file order within a source
    uint32_t nins;
    const uint64_t *del_key;   // sorted (u << 32 | v); ~0 = dropped pair
    uint32_t ndel;
};

// Per touched vertex. Its merged list is the sequence (old list ++ inserts)
// with every deleted target removed; when it has deletes that sequence is
// filtered in kMergeChunk pieces ("chunked"), otherwise the inserts are just
// appended (by chunks for long insert runs, else by one lane).
struct TInfo {
    uint32_t u, d0, ni, ib, db, de, cbase, nch;
};
__device__ __forceinline__ bool has_del(const TInfo &t) { return t.de > t.db; }

struct MChunk {
    uint32_t t;       // touched index
    uint32_t lo, hi;  // positions in the vertex's sequence (old list ++ inserts, or inserts only)
    uint32_t pad;
};

__device__ __forceinline__ bool in_sorted_u64(const uint64_t *a, uint32_t b, uint32_t e, uint64_t key) {
    uint32_t lo = b, hi = e;
    while (lo < hi) { uint32_t m = (lo + hi) >> 1; if (a[m] < key) lo = m + 1; else hi = m; }
    return lo < e && a[lo] == key;
}
__device__ __forceinline__ bool in_sorted_u32(const uint32_t *a, uint32_t n, uint32_t key) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) { uint32_t m = (lo + hi) >> 1; if (a[m] < key) lo = m + 1; else hi = m; }
    return lo < n && a[lo] == key;
}

__device__ __forceinline__ bool bit_test(const uint32_t *bits, uint32_t v) { return (bits[v >> 5] >> (v & 31)) & 1u; }

// Per-vertex batch ranges [ins_b, ins_e), [del_b, del_e) of the sorted batch,
// written at run heads and tails (zero for vertices without updates), and the
// touched list: each source claimed once through the `mark` bitmap.
struct BatchRanges {
    uint32_t *ins_b, *ins_e, *del_b, *del_e;
};

__global__ void __launch_bounds__(kBlock) k_touched(BatchView bv, BatchRanges br, uint32_t *mark, uint32_t *touched,
                                                    uint32_t *tcount, uint32_t cap, unsigned long long *overflow) {
    const uint64_t total = (uint64_t)bv.nins + bv.ndel;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t base = (uint64_t)blockIdx.x * blockDim.x; base < total; base += stride) {
        const uint64_t i = base + threadIdx.x;
        bool p = false;
        uint32_t u = 0;
        if (i < total) {
            if (i < bv.nins) {
                u = bv.ins_key[i];
                if (u != kGap) {
                    p = i == 0 || bv.ins_key[i - 1] != u;
                    if (p) br.ins_b[u] = (uint32_t)i;
                    if (i + 1 == bv.nins || bv.ins_key[i + 1] != u) br.ins_e[u] = (uint32_t)i + 1;
                }
            } else {
                const uint32_t j = (uint32_t)(i - bv.nins);
                const uint64_t k = bv.del_key[j];
                u = (uint32_t)(k >> 32);
                if (k != ~0ull) {
                    p = j == 0 || (bv.del_key[j - 1] >> 32) != u;
                    if (p) br.del_b[u] = j;
                    if (j + 1 == bv.ndel || (bv.del_key[j + 1] >> 32) != u) br.del_e[u] = j + 1;
                }
            }
            if (p) {
                const uint32_t bit = 1u << (u & 31);
                p = !(atomicOr(&mark[u >> 5], bit) & bit);
            }
        }
        wave_append(p, u, touched, tcount, cap, overflow);
    }
}

__device__ __forceinline__ uint32_t seq_len(const TInfo &t) { return (has_del(t) ? t.d0 : 0) + t.ni; }
__device__ __forceinline__ bool chunked(const TInfo &t) { return has_del(t) || t.ni > kLaneIns; }

// Batch ranges and chunk descriptors of every touched vertex.
__global__ void __launch_bounds__(kBlock) k_touch_info(GraphView g, BatchRanges br, const uint32_t *touched,
                                                       const uint32_t *tcount, TInfo *tinfo, uint32_t *tidx,
                                                       MChunk *chunks, uint32_t *nchunks, uint32_t chunk_cap,
                                                       unsigned long long *overflow) {
    const uint32_t n = *tcount;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t base = (uint64_t)blockIdx.x * blockDim.x; base < n; base += stride) {
        const uint64_t t = base + threadIdx.x;
        TInfo ti{};
        uint32_t nch = 0;
        if (t < n) {
            ti.u = touched[t];
            ti.ib = br.ins_b[ti.u];
            ti.ni = br.ins_e[ti.u] - ti.ib;
            ti.db = br.del_b[ti.u];
            ti.de = br.del_e[ti.u];
            ti.d0 = g.deg[ti.u];
            nch = chunked(ti) ? (seq_len(ti) + kMergeChunk - 1) / kMergeChunk : 0;
            ti.nch = nch;
        }
        const uint32_t incl = wave_incl_scan(nch);
        const uint32_t tot = __shfl(incl, kWave - 1, kWave);
        uint32_t cb = 0;
        if (lane_id() == 0 && tot) cb = atomicAdd(nchunks, tot);
        cb = __shfl(cb, 0, kWave) + incl - nch;
        if (t < n) {
            ti.cbase = cb;
            if (cb + nch > chunk_cap) {  // never by construction; flagged, and no chunk is read
                atomicOr(overflow, 4ull);
                ti.cbase = 0;
                ti.nch = 0;
            }
            tinfo[t] = ti;
            tidx[ti.u] = (uint32_t)t;
            const uint32_t len = seq_len(ti);
            for (uint32_t c = 0; c < ti.nch; ++c) {
                MChunk m;
                m.t = (uint32_t)t;
                m.lo = c * kMergeChunk;
                m.hi = min(len, m.lo + kMergeChunk);
                m.pad = 0;
                chunks[cb + c] = m;
            }
        }
    }
}

// Zero the batch ranges of the touched vertices (after the batch is applied).
__global__ void k_clear_ranges(const TInfo *tinfo, const uint32_t *tcount, BatchRanges br) {
    const uint32_t n = *tcount;
    for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t u = tinfo[t].u;
        br.ins_b[u] = br.ins_e[u] = br.del_b[u] = br.del_e[u] = 0;
    }
}

constexpr uint32_t kDelMaskWords = 32;  // 1024-bit filter of a vertex's delete targets

__device__ __forceinline__ uint32_t del_hash(uint32_t x) { return (x * 2654435761u) >> 22; }  // 10 bits

// Stage the delete targets of ti in this wave's LDS slice when they fit, with
// a 1024-bit filter in front: most list entries are not deleted and leave
// after one LDS probe instead of a binary search.
__device__ __forceinline__ uint32_t stage_deletes(const BatchView &bv, const TInfo &ti, uint32_t *s_del,
                                                  uint32_t *s_mask) {
    const uint32_t nd = ti.de - ti.db;
    if (nd == 0 || nd > kDelStage) return 0;
    if (lane_id() < kDelMaskWords) s_mask[lane_id()] = 0;
    __builtin_amdgcn_wave_barrier();
    for (uint32_t i = lane_id(); i < nd; i += kWave) {
        const uint32_t x = (uint32_t)bv.del_key[ti.db + i];
        s_del[i] = x;
        const uint32_t h = del_hash(x);
        atomicOr(&s_mask[h >> 5], 1u << (h & 31));
    }
    __builtin_amdgcn_wave_barrier();
    return nd;
}

__device__ __forceinline__ bool survives(uint32_t x, const BatchView &bv, const TInfo &ti, const uint32_t *s_del,
                                         const uint32_t *s_mask, uint32_t nstaged) {
    if (nstaged) {
        const uint32_t h = del_hash(x);
        if (!((s_mask[h >> 5] >> (h & 31)) & 1u)) return true;
        return !in_sorted_u32(s_del, nstaged, x);
    }
    return !in_sorted_u64(bv.del_key, ti.db, ti.de, ((uint64_t)ti.u << 32) | x);
}

// One wave per chunk of a vertex with deletes: count survivors (pass 0) or
// write them in order to out[cpos[c] ...] (pass 1). Chunks of insert-only
// vertices keep everything: pass 0 counts them without reading, pass 1 skips
// them (they are placed straight from the sorted batch).
template <int kPass>
__global__ void __launch_bounds__(kBlock) k_chunk_filter(GraphView g, BatchView bv, const TInfo *tinfo,
                                                         const MChunk *chunks, const uint32_t *nchunks, uint32_t chunk_cap,
                                                         unsigned long long *ccount, const unsigned long long *cpos,
                                                         uint32_t *out) {
    __shared__ uint32_t s_del[kWavesPerBlock][kDelStage];
    __shared__ uint32_t s_mask[kWavesPerBlock][kDelMaskWords];
    const uint32_t n = min(*nchunks, chunk_cap);
    const uint32_t w = threadIdx.x / kWave, lane = lane_id();
    const uint64_t gw = (uint64_t)blockIdx.x * kWavesPerBlock + w;
    const uint64_t nw = (uint64_t)gridDim.x * kWavesPerBlock;
    for (uint64_t c = gw; c < n; c += nw) {
        const MChunk m = chunks[c];
        const TInfo ti = tinfo[m.t];
        if (!has_del(ti)) {
            if (!kPass && lane == 0) ccount[c] = m.hi - m.lo;
            continue;
        }
        const uint32_t *old = g.col + g.off[ti.u];
        const uint32_t *ins = bv.ins_val + ti.ib;
        const uint32_t ns = stage_deletes(bv, ti, s_del[w], s_mask[w]);
        uint32_t cnt = 0;
        uint32_t *dst = kPass ? out + cpos[c] : nullptr;
        for (uint32_t k0 = m.lo; k0 < m.hi; k0 += kIlp * kWave) {
            uint32_t x[kIlp];
            bool keep[kIlp];
#pragma unroll
            for (int j = 0; j < kIlp; ++j) {  // kIlp independent loads in flight
                const uint32_t i = k0 + j * kWave + lane;
                x[j] = i < m.hi ? (i < ti.d0 ? old[i] : ins[i - ti.d0]) : 0;
            }
#pragma unroll
            for (int j = 0; j < kIlp; ++j) {
                const uint32_t i = k0 + j * kWave + lane;
                keep[j] = i < m.hi && survives(x[j], bv, ti, s_del[w], s_mask[w], ns);
            }
#pragma unroll
            for (int j = 0; j < kIlp; ++j) {  // survivors stay in sequence order
                const uint64_t b = __ballot(keep[j]);
                if (kPass && keep[j]) dst[cnt + __popcll(b & lanemask_lt())] = x[j];
                cnt += (uint32_t)__popcll(b);
            }
        }
        if (!kPass && lane == 0) ccount[c] = cnt;
        __builtin_amdgcn_wave_barrier();
    }
}

// covered(u): some ancestor-or-self node of u's leaf is a maximal window.
__device__ __forceinline__ bool covered(const uint8_t *maximal, uint64_t nleaves, uint32_t L, uint32_t u) {
    for (uint64_t node = (uint64_t)(u / L) + nleaves; node >= 1; node >>= 1)
        if (maximal[node]) return true;
    return false;
}

// One wave per chunk: copy the chunk's survivors into the vertex's own range
// (from `merged` when it has deletes, from the sorted inserts otherwise) and
// gap-fill the freed tail. Vertices in a window are skipped; the window
// re-spread moves them.
__global__ void __launch_bounds__(kBlock) k_chunk_place(GraphView g, BatchView bv, uint32_t *col, const TInfo *tinfo,
                                                        const MChunk *chunks, const uint32_t *nchunks, uint32_t chunk_cap,
                                                        const unsigned long long *cpos, const uint32_t *merged,
                                                        const uint32_t *newdeg, const uint8_t *maximal,
                                                        uint64_t nleaves, uint32_t L, int skip_covered) {
    const uint32_t n = min(*nchunks, chunk_cap);
    const uint32_t w = threadIdx.x / kWave, lane = lane_id();
    const uint64_t gw = (uint64_t)blockIdx.x * kWavesPerBlock + w;
    const uint64_t nw = (uint64_t)gridDim.x * kWavesPerBlock;
    for (uint64_t c = gw; c < n; c += nw) {
        const MChunk m = chunks[c];
        const TInfo ti = tinfo[m.t];
        if (skip_covered && covered(maximal, nleaves, L, ti.u)) continue;
        uint32_t *lst = col + g.off[ti.u];
        const bool hd = has_del(ti);
        const uint64_t p0 = (hd ? 0 : ti.d0) + (cpos[c] - cpos[ti.cbase]);
        const uint64_t len = cpos[c + 1] - cpos[c];
        const uint32_t *srcp = hd ? merged + cpos[c] : bv.ins_val + ti.ib + m.lo;
        for (uint64_t k0 = 0; k0 < len; k0 += kIlp * kWave) {
            uint32_t x[kIlp];
#pragma unroll
            for (int j = 0; j < kIlp; ++j) {
                const uint64_t k = k0 + j * kWave + lane;
                x[j] = k < len ? srcp[k] : 0;
            }
#pragma unroll
            for (int j = 0; j < kIlp; ++j) {
                const uint64_t k = k0 + j * kWave + lane;
                if (k < len) lst[p0 + k] = x[j];
            }
        }
        if (hd) {
            const uint32_t nd = newdeg[ti.u];
            for (uint32_t k = max(m.lo, nd) + lane; k < min(m.hi, ti.d0); k += kWave) lst[k] = kGap;
        }
    }
}

// One lane per insert-only vertex with a short insert run: append in place.
__global__ void k_lane_place(GraphView g, BatchView bv, uint32_t *col, const TInfo *tinfo, const uint32_t *tcount,
                             const uint8_t *maximal, uint64_t nleaves, uint32_t L, int skip_covered) {
    const uint32_t n = *tcount;
    for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (uint64_t)gridDim.x * blockDim.x) {
        const TInfo ti = tinfo[t];
        if (chunked(ti) || ti.ni == 0) continue;
        if (skip_covered && covered(maximal, nleaves, L, ti.u)) continue;
        uint32_t *dst = col + g.off[ti.u] + ti.d0;
        const uint32_t *src = bv.ins_val + ti.ib;
        for (uint32_t k = 0; k < ti.ni; ++k) dst[k] = src[k];
    }
}

// ---- edge-centric sweep ----------------------------------------------------------
// A block takes a tile of kSweepTile live edges (positions in the concatenation
// of all adjacency lists, dpre = exclusive prefix of the degrees), finds the
// tile's source range once, stages that slice of dpre in LDS and maps every
// edge to its source with an LDS binary search. Every block does the same
// amount of work whatever the degree distribution.
constexpr uint32_t kSweepTile = 1024;
constexpr uint32_t kSweepStage = 2048;

__device__ __forceinline__ uint32_t upper_bound_u64(const unsigned long long *a, uint32_t n, unsigned long long key) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) { uint32_t m = (lo + hi) >> 1; if (a[m] <= key) lo = m + 1; else hi = m; }
    return lo;
}

// tile_owner[t] = source of live edge t * kSweepTile, for t in [0, ntiles];
// tile_owner[ntiles] = nv - 1. One thread per vertex writes the boundaries
// that fall in its range.
__global__ void k_tile_owners(uint32_t nv, const unsigned long long *dpre, uint32_t *tile_owner) {
    const unsigned long long E = dpre[nv];
    const unsigned long long ntiles = (E + kSweepTile - 1) / kSweepTile;
    for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < nv; v += (uint64_t)gridDim.x * blockDim.x) {
        const unsigned long long a = dpre[v], b = dpre[v + 1];
        for (unsigned long long t = (a + kSweepTile - 1) / kSweepTile; t * kSweepTile < b; ++t) tile_owner[t] = (uint32_t)v;
        if (v == 0) tile_owner[ntiles] = nv - 1;
    }
}

template <class Op, class EFn>
__global__ void __launch_bounds__(kBlock) k_sweep_edges(GraphView g, const unsigned long long *dpre,
                                                        const uint32_t *tile_owner, Op op,
                                                        const typename Op::buffer_t *buffer, EFn efn, PassIO io) {
    __shared__ long long s_pre[kSweepStage];
    __shared__ uint32_t s_out[kWavesPerBlock][kOutStage];
    __shared__ unsigned long long s_red[kWavesPerBlock];
    const unsigned long long E = dpre[g.nv];
    const uint32_t w = threadIdx.x / kWave;
    OutStage st{s_out[w], 0};
    unsigned long long n_edges = 0;
    for (unsigned long long t0 = (unsigned long long)blockIdx.x * kSweepTile; t0 < E;
         t0 += (unsigned long long)gridDim.x * kSweepTile) {
        const unsigned long long t1 = min(t0 + kSweepTile, E);
        const unsigned long long tile = t0 / kSweepTile;
        // sources of the tile: from the owner of its first edge to the owner
        // of the next tile's first edge (a superset; extra vertices start past t1)
        const uint32_t v0 = tile_owner[tile], span = tile_owner[tile + 1] - v0 + 1;
        const bool staged = span <= kSweepStage;
        if (staged)
            for (uint32_t i = threadIdx.x; i < span; i += blockDim.x) s_pre[i] = (long long)dpre[v0 + i] - (long long)t0;
        __syncthreads();
        uint32_t v[kIlp], src[kIlp];
        uint64_t peek[kIlp];
#pragma unroll
        for (int j = 0; j < kIlp; ++j) {
            const unsigned long long e = t0 + (unsigned long long)j * kBlock + threadIdx.x;
            v[j] = kGap;
            src[j] = 0;
            if (e < t1) {
                uint32_t o;
                if (staged) {
                    const long long rel = (long long)(e - t0);
                    uint32_t lo = 0, hi = span;  // largest o with s_pre[o] <= rel
                    while (hi - lo > 1) { const uint32_t m = (lo + hi) >> 1; if (s_pre[m] <= rel) lo = m; else hi = m; }
                    o = lo;
                    src[j] = v0 + o;
                    v[j] = g.col[g.off[v0 + o] + (uint64_t)(rel - s_pre[o])];
                } else {
                    o = upper_bound_u64(dpre + v0, span, e) - 1;
                    src[j] = v0 + o;
                    v[j] = g.col[g.off[v0 + o] + (e - dpre[v0 + o])];
                }
            }
        }
#pragma unroll
        for (int j = 0; j < kIlp; ++j) peek[j] = v[j] != kGap ? efn.load(v[j]) : 0;
        bool enq[kIlp];
#pragma unroll
        for (int j = 0; j < kIlp; ++j) {
            enq[j] = false;
            if (v[j] != kGap && ((peek[j] >> (v[j] & 31)) & 1u))
                enq[j] = efn.apply(src[j], v[j], op.PushValue(ld_peek(&buffer[src[j]])), peek[j]);
        }
#pragma unroll
        for (int j = 0; j < kIlp; ++j) stage_push(enq[j], v[j], st, io);
        n_edges += t1 - t0;
        __syncthreads();
    }
    stage_flush(st, io);
    if (threadIdx.x == 0 && n_edges) atomicAdd(io.stats + 0, n_edges);
    (void)s_red;
}

// Invalidated vertices that are active on their own (CC: own label) join the
// frontier: one thread per bitmap word.
template <class Op>
__global__ void k_inval_activate(uint32_t nv, Op op, const uint32_t *inval, const typename Op::value_t *value,
                                 const typename Op::buffer_t *buffer, uint32_t *stamp, uint32_t stamp_val, uint32_t *list,
                                 uint32_t *cnt, uint32_t cap, unsigned long long *overflow) {
    const uint32_t nw = (nv + 31) / 32;
    for (uint64_t wd = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; wd < nw; wd += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t bits = inval[wd];
        while (bits) {
            const uint32_t b = __builtin_ctz(bits);
            bits &= bits - 1;
            const uint32_t v = (uint32_t)wd * 32 + b;
            if (v < nv && op.IsActiveNode(v, buffer[v], value[v]) && atomicExch(&stamp[v], stamp_val) != stamp_val) {
                const uint32_t pos = atomicAdd(cnt, 1u);
                if (pos < cap) list[pos] = v; else atomicOr(overflow, 1ull);
            }
        }
    }
}

}  // namespace sg

#endif
