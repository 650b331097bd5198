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
constexpr int kMergePerLane = kMergeChunk / kWave;
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
    uint64_t base;    // off[u] when the batch was prepared
    uint32_t u;
    uint32_t lo, hi;  // positions in the vertex's sequence (old list ++ inserts, or inserts only)
    uint32_t d0, ib, db, de;  // as TInfo
    uint32_t first;   // the vertex's first chunk (TInfo::cbase)
    uint64_t ibase;   // where the filter writes this chunk's survivors (vertices with deletes)
};
// ibase of a chunk filtered in place: the vertex has deletes, its whole
// sequence is this one chunk and fits the vertex's slot range, so one wave
// holds all of it in registers and compacts it where it lies (no staging).
constexpr uint64_t kInPlace = ~0ull;

__device__ __forceinline__ TInfo chunk_info(const MChunk &m) { return TInfo{m.u, m.d0, 0, m.ib, m.db, m.de, m.first, 0}; }
__device__ __forceinline__ MChunk load_chunk(const MChunk *chunks, uint64_t c, uint64_t n) {
    if (c < n) return chunks[c];
    MChunk z{};
    return z;
}

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
    __shared__ BlockAppend sh;
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
        block_append(p, u, touched, tcount, cap, overflow, sh);
    }
}

__device__ __forceinline__ uint32_t seq_len(const TInfo &t) { return (has_del(t) ? t.d0 : 0) + t.ni; }
__device__ __forceinline__ bool chunked(const TInfo &t) { return has_del(t) || t.ni > kLaneIns; }

// Batch ranges and chunk descriptors of every touched vertex.
__global__ void __launch_bounds__(kBlock) k_touch_info(GraphView g, BatchRanges br, const uint32_t *touched,
                                                       const uint32_t *tcount, TInfo *tinfo, uint32_t *tidx,
                                                       MChunk *chunks, uint32_t *nchunks, uint32_t chunk_cap,
                                                       unsigned long long *seqcur, int allow_inplace,
                                                       unsigned long long *overflow) {
    __shared__ uint32_t s_c[kWavesPerBlock], s_cb;
    __shared__ unsigned long long s_s[kWavesPerBlock], s_sb;
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
        const bool inplace = allow_inplace && t < n && has_del(ti) && nch == 1 &&
                             g.off[ti.u + 1] - g.off[ti.u] >= seq_len(ti);
        // chunk indexes and staging ranges: wave scans, then one atomic per
        // counter and block
        const uint32_t incl = wave_incl_scan(nch);
        const unsigned long long sl = (t < n && has_del(ti) && !inplace) ? seq_len(ti) : 0ull;
        const unsigned long long sincl = wave_incl_scan(sl);
        const uint32_t w = threadIdx.x / kWave;
        if (lane_id() == kWave - 1) {
            s_c[w] = incl;
            s_s[w] = sincl;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t ct = 0;
            unsigned long long st = 0;
            for (int i = 0; i < kWavesPerBlock; ++i) {
                const uint32_t c = s_c[i];
                const unsigned long long q = s_s[i];
                s_c[i] = ct;
                s_s[i] = st;
                ct += c;
                st += q;
            }
            s_cb = ct ? atomicAdd(nchunks, ct) : 0u;
            s_sb = st ? atomicAdd(seqcur, st) : 0ull;
        }
        __syncthreads();
        const uint32_t cb = s_cb + s_c[w] + incl - nch;
        const unsigned long long sb = s_sb + s_s[w] + sincl - sl;
        __syncthreads();
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
                MChunk m{};
                m.base = g.off[ti.u];
                m.u = ti.u;
                m.lo = c * kMergeChunk;
                m.hi = min(len, m.lo + kMergeChunk);
                m.d0 = ti.d0;
                m.ib = ti.ib;
                m.db = ti.db;
                m.de = ti.de;
                m.first = ti.cbase;
                m.ibase = inplace ? kInPlace : sb + m.lo;
                chunks[ti.cbase + c] = m;
            }
        }
    }
}

// Zero the batch ranges of the touched vertices (after the batch is applied).
// Also clears the touched-vertex bitmap: every set bit belongs to a touched
// vertex, so each touched vertex may zero its whole word.
__global__ void k_clear_ranges(const TInfo *tinfo, const uint32_t *tcount, BatchRanges br, uint32_t *mark) {
    const uint32_t n = *tcount;
    for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t u = tinfo[t].u;
        br.ins_b[u] = br.ins_e[u] = br.del_b[u] = br.del_e[u] = 0;
        mark[u >> 5] = 0;
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

// One wave per chunk of a vertex with deletes: write the survivors, in
// order, to out[ibase ...] and count them; an in-place chunk (kInPlace)
// compacts its survivors into the vertex's own slots instead (every load is
// in registers before the first store, a survivor only moves left, and slots
// already holding their survivor are not rewritten) and gap-fills the freed
// tail. Chunks of insert-only vertices keep everything: counted without reads
// (they are placed straight from the sorted batch).
__global__ void __launch_bounds__(kBlock) k_chunk_filter(GraphView g, BatchView bv, const MChunk *chunks,
                                                         const uint32_t *nchunks, uint32_t chunk_cap,
                                                         unsigned long long *ccount, uint32_t *out, uint32_t *col) {
    __shared__ uint32_t s_del[kWavesPerBlock][kDelStage];
    __shared__ uint32_t s_mask[kWavesPerBlock][kDelMaskWords];
    const uint32_t n = min(*nchunks, chunk_cap);
    const uint32_t w = threadIdx.x / kWave, lane = lane_id();
    const uint64_t gw = (uint64_t)blockIdx.x * kWavesPerBlock + w;
    const uint64_t nw = (uint64_t)gridDim.x * kWavesPerBlock;
    MChunk mn = load_chunk(chunks, gw, n);  // descriptors are prefetched one iteration ahead
    for (uint64_t c = gw; c < n; c += nw) {
        const MChunk m = mn;
        mn = load_chunk(chunks, c + nw, n);
        const TInfo ti = chunk_info(m);
        if (!has_del(ti)) {
            if (lane == 0) ccount[c] = m.hi - m.lo;
            continue;
        }
        const uint32_t *old = g.col + m.base;
        const uint32_t *ins = bv.ins_val + ti.ib;
        const uint32_t ns = stage_deletes(bv, ti, s_del[w], s_mask[w]);
        // the whole chunk is loaded before the first probe: one memory latency
        // per chunk; only the nj wave rows the chunk covers are loaded and probed
        const uint32_t nj = (m.hi - m.lo + kWave - 1) / kWave;
        uint32_t x[kMergePerLane];
#pragma unroll
        for (int j = 0; j < kMergePerLane; ++j) {
            const uint32_t i = m.lo + j * kWave + lane;
            x[j] = 0;
            if (j < (int)nj && i < m.hi) x[j] = i < ti.d0 ? old[i] : ins[i - ti.d0];
        }
        uint32_t cnt = 0;
        const bool inpl = m.ibase == kInPlace;
        uint32_t *dst = inpl ? col + m.base : out + m.ibase;
#pragma unroll
        for (int j = 0; j < kMergePerLane; ++j) {  // survivors stay in sequence order
            if (j >= (int)nj) break;
            const uint32_t i = m.lo + j * kWave + lane;
            const bool keep = i < m.hi && survives(x[j], bv, ti, s_del[w], s_mask[w], ns);
            const uint64_t b = __ballot(keep);
            const uint32_t pos = cnt + (uint32_t)__popcll(b & lanemask_lt());
            if (keep && (!inpl || pos != i)) dst[pos] = x[j];
            cnt += (uint32_t)__popcll(b);
        }
        if (inpl)
            for (uint32_t k = cnt + lane; k < ti.d0; k += kWave) dst[k] = kGap;
        if (lane == 0) ccount[c] = cnt;
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
    MChunk mn = load_chunk(chunks, gw, n);
    for (uint64_t c = gw; c < n; c += nw) {
        const MChunk m = mn;
        mn = load_chunk(chunks, c + nw, n);
        const TInfo ti = chunk_info(m);
        if (m.ibase == kInPlace) continue;  // compacted by the filter
        if (skip_covered && covered(maximal, nleaves, L, ti.u)) continue;
        uint32_t *lst = col + m.base;
        const bool hd = has_del(ti);
        const uint64_t p0 = (hd ? 0 : ti.d0) + (cpos[c] - cpos[ti.cbase]);
        const uint64_t len = cpos[c + 1] - cpos[c];
        const uint32_t *srcp = hd ? merged + m.ibase : bv.ins_val + ti.ib + m.lo;
        // len <= kMergeChunk: one load per lane and covered wave row, all in flight
        const uint32_t nj = (uint32_t)((len + kWave - 1) / kWave);
        const uint32_t *sp = srcp + lane;
        uint32_t *dp = lst + p0 + lane;
        uint32_t x[kMergePerLane];
#pragma unroll
        for (int j = 0; j < kMergePerLane; ++j) {
            x[j] = 0;
            if (j < (int)nj && (uint32_t)j * kWave + lane < len) x[j] = sp[j * kWave];
        }
#pragma unroll
        for (int j = 0; j < kMergePerLane; ++j) {
            if (j >= (int)nj) break;
            if ((uint32_t)j * kWave + lane < len) dp[j * kWave] = x[j];
        }
        if (hd) {
            const uint32_t nd = newdeg[ti.u];
            for (uint32_t k = max(m.lo, nd) + lane; k < min(m.hi, ti.d0); k += kWave) lst[k] = kGap;
        }
    }
}

// One lane per touched vertex: an insert-only vertex with a short insert run
// appends it in place; every touched vertex takes its new degree (vertices
// in a re-spread window were written by the window pass).
__global__ void k_lane_place(GraphView g, BatchView bv, uint32_t *col, const TInfo *tinfo, const uint32_t *tcount,
                             const uint32_t *newdeg, uint32_t *deg, const uint8_t *maximal, uint64_t nleaves, uint32_t L,
                             int skip_covered) {
    const uint32_t n = *tcount;
    for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (uint64_t)gridDim.x * blockDim.x) {
        const TInfo ti = tinfo[t];
        if (skip_covered && covered(maximal, nleaves, L, ti.u)) continue;
        if (!chunked(ti) && ti.ni) {
            uint32_t *dst = col + g.off[ti.u] + ti.d0;
            const uint32_t *src = bv.ins_val + ti.ib;
            for (uint32_t k = 0; k < ti.ni; ++k) dst[k] = src[k];
        }
        deg[ti.u] = newdeg[ti.u];
    }
}

// ---- edge-centric sweep ----------------------------------------------------------
// Live edges are walked along a merge path over vertices and edges: path
// position v + dpre[v] (dpre = exclusive prefix of the degrees) is the first
// edge of vertex v, the position after its last edge is v's end mark. A tile
// is kSweepTile consecutive path positions, so it holds at most kSweepTile
// edges and kSweepTile + 1 sources whatever the degree distribution.
constexpr uint32_t kSweepTile = 512;   // path positions per tile (one wave)
constexpr uint32_t kScanBlock = 256;
constexpr uint32_t kScanIters = 16;
constexpr uint32_t kScanSpan = kScanBlock * kScanIters;  // vertices per block of the degree scan

struct SweepTiles {
    uint32_t *v;             // [cap] first source of tile t (tile cap - 1 is the sentinel)
    unsigned long long *e;   // [cap] first live-edge index of tile t
    uint64_t cap;            // entries of v and e
};

// Block-wide exclusive scan of one value per thread; *total gets the sum.
// s_w holds kScanBlock / kWave words; callers alternate two of them.
__device__ __forceinline__ unsigned long long block_excl_scan(unsigned long long x, unsigned long long *s_w,
                                                              unsigned long long &total) {
    const unsigned long long inc = wave_incl_scan(x);
    const uint32_t w = threadIdx.x / kWave;
    if ((threadIdx.x & (kWave - 1)) == kWave - 1) s_w[w] = inc;
    __syncthreads();
    unsigned long long before = 0;
    total = 0;
#pragma unroll
    for (uint32_t i = 0; i < kScanBlock / kWave; ++i) {
        if (i < w) before += s_w[i];
        total += s_w[i];
    }
    return before + inc - x;
}

// Degree scan, phase 1: the degree sum of each kScanSpan-vertex block.
__global__ void __launch_bounds__(kScanBlock) k_deg_block_sums(uint32_t nv, const uint32_t *deg, unsigned long long *bsum) {
    __shared__ unsigned long long s_w[kScanBlock / kWave];
    const uint64_t base = (uint64_t)blockIdx.x * kScanSpan;
    unsigned long long x = 0;
#pragma unroll 4
    for (uint32_t i = 0; i < kScanIters; ++i) {
        const uint64_t v = base + (uint64_t)i * kScanBlock + threadIdx.x;
        if (v < nv) x += deg[v];
    }
    unsigned long long total;
    block_excl_scan(x, s_w, total);
    if (threadIdx.x == 0) bsum[blockIdx.x] = total;
}

// Phase 2 (one block): exclusive scan of the nb block sums in place.
__global__ void __launch_bounds__(1024) k_deg_scan_sums(uint32_t nb, unsigned long long *bsum) {
    __shared__ unsigned long long s_w[1024 / kWave];
    __shared__ unsigned long long s_carry;
    if (threadIdx.x == 0) s_carry = 0;
    __syncthreads();
    for (uint32_t b0 = 0; b0 < nb; b0 += 1024) {
        const uint32_t i = b0 + threadIdx.x;
        const unsigned long long x = i < nb ? bsum[i] : 0;
        const unsigned long long inc = wave_incl_scan(x);
        const uint32_t w = threadIdx.x / kWave;
        if ((threadIdx.x & (kWave - 1)) == kWave - 1) s_w[w] = inc;
        __syncthreads();
        unsigned long long before = s_carry, total = 0;
        for (uint32_t j = 0; j < 1024 / kWave; ++j) {
            if (j < w) before += s_w[j];
            total += s_w[j];
        }
        if (i < nb) bsum[i] = before + inc - x;
        __syncthreads();
        if (threadIdx.x == 0) s_carry += total;
        __syncthreads();
    }
}

// Phase 3: dpre[v] for v in [0, nv].
__global__ void __launch_bounds__(kScanBlock) k_deg_scan_write(uint32_t nv, const uint32_t *deg,
                                                               const unsigned long long *bsum, unsigned long long *dpre) {
    __shared__ unsigned long long s_w[2][kScanBlock / kWave];
    const uint64_t base = (uint64_t)blockIdx.x * kScanSpan;
    unsigned long long carry = bsum[blockIdx.x];
    for (uint32_t i = 0; i < kScanIters; ++i) {
        const uint64_t v = base + (uint64_t)i * kScanBlock + threadIdx.x;
        const uint32_t d = v < nv ? deg[v] : 0u;
        unsigned long long total;
        const unsigned long long a = carry + block_excl_scan(d, s_w[i & 1], total);
        carry += total;
        if (v < nv) {
            dpre[v] = a;
            if (v + 1 == nv) dpre[nv] = a + d;
        }
    }
}

// Phase 4, one thread per tile: tile t starts at path position d = t *
// kSweepTile, inside the vertex v with v + dpre[v] <= d < v + 1 + dpre[v + 1]
// (the largest v with v + dpre[v] <= d), at edge min(d - v, dpre[v + 1]).
// Tile ntiles = ceil((nv + E) / kSweepTile) is the sentinel {nv - 1, E}.
__global__ void k_sweep_tiles(uint32_t nv, const unsigned long long *dpre, SweepTiles tiles, unsigned long long *overflow) {
    const unsigned long long E = dpre[nv];
    const unsigned long long nt = ((unsigned long long)nv + E + kSweepTile - 1) / kSweepTile;
    if (nt >= tiles.cap) {
        if (blockIdx.x == 0 && threadIdx.x == 0) atomicOr(overflow, 4ull);
        return;
    }
    for (unsigned long long t = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; t <= nt;
         t += (unsigned long long)gridDim.x * blockDim.x) {
        if (t == nt) {
            tiles.v[t] = nv - 1;
            tiles.e[t] = E;
            continue;
        }
        const unsigned long long d = t * kSweepTile;
        uint32_t lo = 0, hi = nv - 1;  // largest v in [0, nv - 1] with v + dpre[v] <= d
        while (lo < hi) {
            const uint32_t m = lo + (hi - lo + 1) / 2;
            if ((unsigned long long)m + dpre[m] <= d) lo = m; else hi = m - 1;
        }
        tiles.v[t] = lo;
        tiles.e[t] = min_u64(d - lo, dpre[lo + 1]);
    }
}

// One wave per tile (grid-stride over waves, no block-wide barrier, so the
// waves of a block overlap their memory phases). Each source of the tile gets
// its slot base relative to the tile's first source (off[v] - off[v0] -
// (dpre[v] - e0), 32 bits) in the wave's LDS, the first edge of every source
// with edges is marked with its 16-bit index, a wave max-scan fills the owners
// forward, and the tile's edges are loaded coalesced, kSweepPerLane per lane,
// all in flight. A tile whose slot range does not fit 32 bits takes the slot
// of each edge from off and dpre in global memory instead. The next tile's
// bounds are loaded while the current tile is walked.
constexpr int kSweepPerLane = kSweepTile / kWave;
constexpr uint32_t kSweepOutStage = 128;
constexpr int kSweepWaves = 4;                    // waves per block of the sweep kernels
constexpr int kSweepBlock = kSweepWaves * kWave;

struct SweepWave {
    uint16_t own[kSweepTile];
    int32_t base[kSweepTile + 1];
    uint32_t out[kSweepOutStage];
};
// base[] also holds a tile's packed words (ids of at most 31 bits, 16-byte groups)
static_assert(kSweepTile + 1 >= (kSweepTile * 31 + 31) / 32 + 8, "base[] holds a tile's packed words");

// Bounds of tile t, loaded by lanes 0-3 with vector loads so that the load
// stays in flight while the wave works on the current tile: lane 0 v[t],
// lane 1 v[t + 1], lane 2 e[t], lane 3 e[t + 1].
__device__ __forceinline__ unsigned long long tile_head_load(const SweepTiles &tiles, unsigned long long t) {
    const uint32_t lane = lane_id();
    unsigned long long x = 0;
    if (lane < 2) x = tiles.v[t + lane];
    else if (lane < 4) x = tiles.e[t + lane - 2];
    return x;
}

struct TileHead {
    uint32_t v0, span, ne;  // first source, sources (v0 + span <= nv), edges
    unsigned long long e0;  // first live-edge index
};

__device__ __forceinline__ unsigned long long read_lane_u64(unsigned long long x, int lane) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)x, lane);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(x >> 32), lane);
    return ((unsigned long long)hi << 32) | lo;
}

// The bounds as wave-uniform (scalar) values.
__device__ __forceinline__ TileHead tile_head(unsigned long long raw) {
    TileHead h;
    h.v0 = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)raw, 0);
    const uint32_t v1 = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)raw, 1);
    h.e0 = read_lane_u64(raw, 2);
    const unsigned long long e1 = read_lane_u64(raw, 3);
    h.ne = (uint32_t)min_u64(e1 - h.e0, kSweepTile);
    h.span = min<uint32_t>(v1 - h.v0 + 1, kSweepTile + 1);
    return h;
}

// Per-source words of the tile's first kWave sources (lane i: source v0 + i,
// off and dpre; position span is the vertex after the last source: its off
// is the width check and its dpre the last source's end). Loaded one tile
// ahead.
struct TilePre {
    uint64_t ov;
    unsigned long long a;
};

__device__ __forceinline__ TilePre tile_pre(const GraphView &g, const unsigned long long *dpre, const TileHead &h,
                                            uint32_t i) {
    TilePre p{0, 0};
    if (i <= h.span) {
        p.ov = g.off[h.v0 + i];
        p.a = dpre[h.v0 + i];
    }
    return p;
}

// colp / slot_lo: where slot s of the tile's sources is read (colp[s - slot_lo]:
// the edge array itself, or one streamed segment of it); only sources in
// [src_lo, src_hi) are walked. pre = tile_pre(h, lane). vv[j] = target at tile
// position j * 64 + lane (kGap when none); the owner index of a position stays
// in sh.own.
template <bool kStagePacked = false>
__device__ __forceinline__ void sweep_tile(const GraphView &g, const unsigned long long *dpre, const TileHead &h,
                                           const TilePre &pre, const uint32_t *colp, uint64_t slot_lo, uint32_t src_lo,
                                           uint32_t src_hi, SweepWave &sh, uint32_t vv[kSweepPerLane],
                                           uint32_t &n_edges) {
    const uint32_t lane = lane_id();
    const uint32_t v0 = h.v0, span = h.span, ne = h.ne;
    const unsigned long long e0 = h.e0;
#pragma unroll
    for (int j = 0; j < kSweepPerLane; ++j) sh.own[j * kWave + lane] = 0;
    __builtin_amdgcn_wave_barrier();
    const uint64_t off0 = read_lane_u64(pre.ov, 0);
    bool far = false;
    for (uint32_t i0 = 0; i0 <= span; i0 += kWave) {
        const uint32_t i = i0 + lane;
        const TilePre p = i0 == 0 ? pre : tile_pre(g, dpre, h, i);
        far |= i <= span && p.ov - off0 >= 0x40000000ull;
        // end of source i = start of source i + 1 (the next lane, or one more load on the last lane)
        unsigned long long b = __shfl_down(p.a, 1, kWave);
        if (lane == kWave - 1 && i < span) b = dpre[v0 + i + 1];
        if (i < span) {
            sh.base[i] = (int32_t)((long long)(p.ov - off0) - (long long)(p.a - e0));
            if (b > p.a && p.a >= e0 && p.a < e0 + ne) sh.own[p.a - e0] = (uint16_t)i;
        }
    }
    const bool wide = __ballot(far) != 0;
    __builtin_amdgcn_wave_barrier();
    // forward fill of the owners: lane k holds positions [k * P, k * P + P)
    uint32_t own[kSweepPerLane];
    uint32_t m = 0;
#pragma unroll
    for (int j = 0; j < kSweepPerLane; ++j) {
        own[j] = sh.own[lane * kSweepPerLane + j];
        m = max(m, own[j]);
    }
    uint32_t x = m;
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        const uint32_t y = __shfl_up(x, d, kWave);
        if (lane >= (uint32_t)d) x = max(x, y);
    }
    uint32_t prev = __shfl_up(x, 1, kWave);
    if (lane == 0) prev = 0;
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int j = 0; j < kSweepPerLane; ++j) {
        prev = max(prev, own[j]);
        sh.own[lane * kSweepPerLane + j] = (uint16_t)prev;
    }
    __builtin_amdgcn_wave_barrier();
    // kStagePacked: a tile whose edges all lie in one packed copy (the HBM
    // mirror or the host tail) loads the words it spans with at most two
    // 16-byte loads per lane into LDS (base[] is not needed for such a tile)
    // and unpacks its ids there: each word crosses the bus once, where reading
    // id by id loads two overlapping words per id. The copies keep 4 words of
    // slack past their last id for the aligned loads. Kernels that seldom
    // read a copy leave it off (it costs registers).
    const uint32_t b = g.mbits;
    const bool in_mirror = g.mlive && e0 + ne <= g.mlive;
    const bool packed_tile = kStagePacked && b < 32 && ne && (in_mirror || (g.tail && e0 >= g.mlive));
    uint32_t *const lw = reinterpret_cast<uint32_t *>(sh.base);
    uint32_t bit0 = 0;  // bit of the tile's first id in lw
    if (packed_tile) {
        const uint32_t *cp = in_mirror ? g.mirror : g.tail;
        const unsigned long long i0 = in_mirror ? e0 : e0 - g.mlive;
        const unsigned long long wa = ((i0 * b) >> 5) & ~3ull;
        // words [wa, end]: end is the word after the one holding the last bit (<= 501 words)
        const uint32_t nw = (uint32_t)(((i0 + ne) * b + 31) / 32 + 1 - wa);
        bit0 = (uint32_t)(i0 * b - wa * 32);
        const uint32_t q0 = lane * 4, q1 = q0 + kWave * 4;
        uint4 x0 = make_uint4(0u, 0u, 0u, 0u), x1 = x0;
        if (q0 < nw) x0 = *reinterpret_cast<const uint4 *>(cp + wa + q0);
        if (q1 < nw) x1 = *reinterpret_cast<const uint4 *>(cp + wa + q1);
        if (q0 < nw) {
            lw[q0] = x0.x;
            lw[q0 + 1] = x0.y;
            lw[q0 + 2] = x0.z;
            lw[q0 + 3] = x0.w;
        }
        if (q1 < nw) {
            lw[q1] = x1.x;
            lw[q1 + 1] = x1.y;
            lw[q1 + 2] = x1.z;
            lw[q1 + 3] = x1.w;
        }
        __builtin_amdgcn_wave_barrier();
    }
    if (packed_tile) {
#pragma unroll
        for (int j = 0; j < kSweepPerLane; ++j) {
            const uint32_t p = j * kWave + lane;
            vv[j] = kGap;
            if (p < ne) {
                const uint32_t src = v0 + sh.own[p];
                if (src >= src_lo && src < src_hi) {
                    ++n_edges;
                    const uint32_t bit = bit0 + p * b;
                    const uint64_t x = ((uint64_t)lw[(bit >> 5) + 1] << 32) | lw[bit >> 5];
                    vv[j] = (uint32_t)(x >> (bit & 31u)) & ((1u << b) - 1u);
                }
            }
        }
        return;
    }
#pragma unroll
    for (int j = 0; j < kSweepPerLane; ++j) {
        const uint32_t p = j * kWave + lane;
        vv[j] = kGap;
        if (p < ne) {
            const uint32_t o = sh.own[p];
            const uint32_t src = v0 + o;
            if (src >= src_lo && src < src_hi) {
                ++n_edges;
                const uint64_t slot = !wide ? off0 + (uint64_t)((long long)sh.base[o] + (long long)p)
                                            : g.off[src] + (e0 + p - dpre[src]);
                const unsigned long long k = e0 + p;
                vv[j] = k < g.mlive ? packed_at(g.mirror, k, b)
                        : g.tail    ? packed_at(g.tail, k - g.mlive, b)
                                    : colp[slot - slot_lo];
            }
        }
    }
}

// Tile pipeline of one wave over tiles t, t + nw, ... < tend: the bounds are
// loaded two tiles ahead and the first sources' words one tile ahead, so they
// are in flight while the current tile's edges are loaded and tested.
struct TilePipe {
    unsigned long long t, nw, tend;
    unsigned long long raw_next;  // tile_head_load(t + nw)
    TileHead h;                   // tile t
    TilePre pre;                  // tile_pre of tile t
    __device__ __forceinline__ void start(const GraphView &g, const unsigned long long *dpre, const SweepTiles &tiles) {
        raw_next = t + nw < tend ? tile_head_load(tiles, t + nw) : 0;
        if (t < tend) {
            h = tile_head(tile_head_load(tiles, t));
            pre = tile_pre(g, dpre, h, lane_id());
        }
    }
    // after the current tile's edges are loaded: start the next tile's loads
    __device__ __forceinline__ void advance(const GraphView &g, const unsigned long long *dpre, const SweepTiles &tiles) {
        const unsigned long long raw2 = t + 2 * nw < tend ? tile_head_load(tiles, t + 2 * nw) : 0;
        if (t + nw < tend) {
            h = tile_head(raw_next);
            pre = tile_pre(g, dpre, h, lane_id());
        }
        raw_next = raw2;
        t += nw;
    }
};

// Block-wide sum of a per-lane counter, one atomic per block of kWaves waves.
template <int kWaves>
__device__ __forceinline__ void block_count(unsigned long long x, unsigned long long *s_red, unsigned long long *dst) {
    x = wave_sum_u64(x);
    if (lane_id() == 0) s_red[threadIdx.x / kWave] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
        for (int i = 0; i < kWaves; ++i) t += s_red[i];
        if (t) atomicAdd(dst, t);
    }
}

// The block's copy of a coarse vertex filter (coarse_words(nv, shift) words).
__device__ __forceinline__ void load_filter(uint32_t *s_filter, const uint32_t *coarse, uint32_t nwords) {
    for (uint32_t i = threadIdx.x; i < nwords; i += blockDim.x) s_filter[i] = coarse[i];
    __syncthreads();
}

// Tiles in [t_begin, t_end) are walked. The edge functor's coarse filter is
// tested in LDS; only targets it passes read the functor's word (efn.load).
template <class Op, class EFn>
__global__ void __launch_bounds__(kSweepBlock) k_sweep_edges(GraphView g, const unsigned long long *dpre, SweepTiles tiles,
                                                             const uint32_t *colp, uint64_t slot_lo, uint32_t src_lo,
                                                             uint32_t src_hi, uint64_t t_begin, uint64_t t_end, Op op,
                                                             const typename Op::buffer_t *buffer, EFn efn, PassIO io) {
    __shared__ SweepWave s_wave[kSweepWaves];
    __shared__ uint32_t s_filter[kFilterWords];
    __shared__ unsigned long long s_red[kSweepWaves];
    const unsigned long long E = dpre[g.nv];
    const unsigned long long ntiles = (g.nv + E + kSweepTile - 1) / kSweepTile;
    const uint32_t w = threadIdx.x / kWave, lane = lane_id();
    SweepWave &sh = s_wave[w];
    OutStage st{sh.out, 0, kSweepOutStage};
    uint32_t n_edges = 0;
    const uint32_t cs = efn.cshift;
    load_filter(s_filter, efn.coarse, coarse_words(g.nv, cs));
    TilePipe tp;
    tp.tend = min_u64(min_u64(t_end, ntiles), tiles.cap - 1);
    tp.nw = (unsigned long long)gridDim.x * kSweepWaves;
    tp.t = t_begin + (unsigned long long)blockIdx.x * kSweepWaves + w;
    tp.start(g, dpre, tiles);
    while (tp.t < tp.tend) {
        uint32_t vv[kSweepPerLane];
        const uint32_t v0 = tp.h.v0;
        sweep_tile(g, dpre, tp.h, tp.pre, colp, slot_lo, src_lo, src_hi, sh, vv, n_edges);
        tp.advance(g, dpre, tiles);
        bool pass[kSweepPerLane];
#pragma unroll
        for (int j = 0; j < kSweepPerLane; ++j) pass[j] = vv[j] != kGap && coarse_test(s_filter, cs, vv[j]);
        decltype(efn.load(0u)) peek[kSweepPerLane];
#pragma unroll
        for (int j = 0; j < kSweepPerLane; ++j) peek[j] = pass[j] ? efn.load(vv[j]) : 0;
        // compact the positions that pass the functor's word test in place:
        // the k-th hit (position p >= k) moves its owner to own[k] and its
        // target to the (now unused) base[k]; then one loop applies them
        uint32_t nh = 0;  // wave-uniform
        uint32_t *const hv = reinterpret_cast<uint32_t *>(sh.base);
#pragma unroll
        for (int j = 0; j < kSweepPerLane; ++j) {
            const bool h = pass[j] && efn.test(vv[j], peek[j]);
            const uint64_t m = __ballot(h);
            if (m) {
                const uint32_t o = sh.own[j * kWave + lane];
                __builtin_amdgcn_wave_barrier();
                if (h) {
                    const uint32_t k = nh + (uint32_t)__popcll(m & lanemask_lt());
                    sh.own[k] = (uint16_t)o;
                    hv[k] = vv[j];
                }
                nh += (uint32_t)__popcll(m);
            }
        }
        __builtin_amdgcn_wave_barrier();
        for (uint32_t i0 = 0; i0 < nh; i0 += kWave) {
            const uint32_t i = i0 + lane;
            bool enq = false;
            uint32_t v = 0;
            if (i < nh) {
                v = hv[i];
                const uint32_t src = v0 + sh.own[i];
                enq = efn.apply(src, v, op.PushValue(ld_peek(&buffer[src])));
            }
            stage_push(enq, v, st, io);
        }
        __builtin_amdgcn_wave_barrier();
    }
    stage_flush(st, io);
    block_count<kSweepWaves>((unsigned long long)n_edges, s_red, io.stats + 0);
}

// Invalidated vertices that are active on their own (CC: own label) join the
// frontier: one thread per bitmap word. The sweep has read the bitmap and its
// coarse filter for the last time: both are cleared here for the next batch.
template <class Op>
__global__ void k_inval_activate(uint32_t nv, Op op, uint32_t *inval, uint32_t *coarse, uint32_t cshift,
                                 const typename Op::value_t *value, const typename Op::buffer_t *buffer, uint32_t *stamp,
                                 uint32_t stamp_val, uint32_t *list, uint32_t *cnt, uint32_t cap,
                                 unsigned long long *overflow) {
    const uint32_t nw = (nv + 31) / 32;
    for (uint64_t wd = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; wd < nw; wd += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t bits = inval[wd];
        if (bits) inval[wd] = 0;
        if ((wd & ((1ull << cshift) - 1)) == 0) coarse[wd >> cshift] = 0;  // coarse word c covers words c << cshift ..
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
