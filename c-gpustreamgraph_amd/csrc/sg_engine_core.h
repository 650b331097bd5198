// Engine core: the device PMA store, the frontier round machinery and the
// traversal program template, shared by the library and by engines built
// from a caller's own operator (include/sgraph_operator.h).
//
// Everything here is a template or defined inside its class, so each shared
// object that includes it compiles its own copy; the one thread-local error
// string of sg_create lives in the library.
#ifndef SG_ENGINE_CORE_H
#define SG_ENGINE_CORE_H

#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <tuple>
#include <type_traits>
#include <unordered_map>
#include <utility>
#include <vector>

#include "sgraph.h"
#include "sg_kernels.h"
#include "sg_operators.h"
#include "sg_pma.h"
#include "sg_multi.h"
#include "sg_pull.h"

namespace sg {

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
struct Error {
    sg_status status;
    std::string msg;
};

[[noreturn]] static void fail(sg_status s, const std::string &m) { throw Error{s, m}; }

#define SG_HIP(call)                                                                         \
    do {                                                                                     \
        hipError_t e_ = (call);                                                              \
        if (e_ != hipSuccess)                                                                \
            ::sg::fail(SG_EDEVICE, std::string(#call) + ": " + hipGetErrorString(e_) +       \
                                       " (" __FILE__ ":" + std::to_string(__LINE__) + ")");  \
    } while (0)
#define SG_LAUNCHED() SG_HIP(hipGetLastError())

// why the last sg_create on this thread failed (defined once, in the library)
extern thread_local std::string g_create_error;

// Scratch of sg_gather_at_noalloc: the vertex ids of one call, then their
// result words. Each device holds its own copy until the process ends; host
// code reaches it only through the address hipGetSymbolAddress returns on that
// device (the name is a host-side stand-in), resolved once per device when an
// engine is created there.
constexpr uint64_t kNoAllocGatherMax = 65536;  // vertices per call (2 x 256 KiB)
static __device__ __align__(256) uint32_t scratch[2 * kNoAllocGatherMax];
static std::mutex g_scratch_mu;                 // guards the address table
static std::vector<uint32_t *> g_scratch_addr;  // per device ordinal
static std::mutex g_scratch_call_mu[64];        // one call at a time per device (shared contents)

// ---------------------------------------------------------------------------
// PMA / batch kernels
// ---------------------------------------------------------------------------
// HBM edge-cache slots of -cache levels 0..4 (graph_datum.cuh:151-163).
constexpr uint64_t kCacheSlotsByLevel[5] = {0ull, 268435455ull, 536870921ull, 805306365ull, 1ull << 30};

struct Ctl {
    unsigned long long overflow;  // some touched vertex outgrows its range
    unsigned long long ne_delta;  // sum of (new - old) degrees (two's complement)
    unsigned long long dropped;   // pairs with an id >= nv
    unsigned long long resize;    // the root window is too dense
    unsigned long long nwin;      // maximal windows
    unsigned long long inval;     // invalidated vertices
    unsigned long long pad[2];
    unsigned long long seqcur;    // staging cursor of the filtered sequences
    uint32_t tcount;              // touched vertices
    uint32_t nchunks;             // merge chunks
    uint32_t pad32[2];
};

__device__ __forceinline__ uint64_t umin64(uint64_t a, uint64_t b) { return a < b ? a : b; }

__global__ void k_fill_u32(uint32_t *p, uint64_t n, uint32_t v) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) p[i] = v;
}
__global__ void k_fill_u8(uint8_t *p, uint64_t n, uint8_t v) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) p[i] = v;
}
__global__ void k_fill_u64(unsigned long long *p, uint64_t n, unsigned long long v) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) p[i] = v;
}

// Initial layout: gaps spread in proportion to deg + 1 (calculate_positions_V1,
// csr_graph.cuh:543-560): off[v] = row[v] + floor(gaps * (row[v] + v) / (ne + nv)).
__global__ void k_layout_offsets(uint32_t nv, const uint64_t *row, uint64_t cap, uint64_t *off, uint32_t *deg) {
    const uint64_t ne = row[nv];
    const double scale = (double)(cap - ne) / (double)(ne + nv);
    for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v <= nv; v += (uint64_t)gridDim.x * blockDim.x) {
        if (v == nv) { off[nv] = cap; continue; }
        off[v] = row[v] + (uint64_t)floor(scale * (double)(row[v] + v));
        deg[v] = (uint32_t)(row[v + 1] - row[v]);
    }
}
// One wave per vertex copies its list to its PMA range.
__global__ void __launch_bounds__(kBlock) k_layout_copy(uint32_t nv, const uint64_t *row, const uint32_t *src,
                                                        const uint64_t *off, uint32_t *col) {
    const uint64_t gw = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
    const uint64_t nw = ((uint64_t)gridDim.x * blockDim.x) / kWave;
    for (uint64_t v = gw; v < nv; v += nw) {
        const uint64_t b = row[v], d = row[v + 1] - b, o = off[v];
        for (uint64_t k = lane_id(); k < d; k += kWave) col[o + k] = src[b + k];
    }
}
// Per-leaf live-edge counts from the CSR prefix.
__global__ void k_leaf_init(uint32_t nv, uint32_t L, uint64_t nleaves, const uint64_t *row, unsigned long long *leaf_cnt) {
    for (uint64_t l = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; l <= nleaves; l += (uint64_t)gridDim.x * blockDim.x) {
        if (l == nleaves) { leaf_cnt[l] = 0; continue; }
        const uint64_t a = umin64(l * L, nv), b = umin64((l + 1) * L, nv);
        leaf_cnt[l] = row[b] - row[a];
    }
}

// SEGMENT split of Engine::LoadGraph (framework.cuh:310-362): from the first
// vertex a of a segment, take vertices while the accumulated slot span
// (gaps included) is below ceil(capacity / nseg); the next segment starts at
// the first b with off[b] - off[a] >= ceil(capacity / nseg). One thread,
// one binary search per cut. Cuts past the last vertex are nv.
__global__ void k_segment_cuts(uint32_t nv, const uint64_t *off, uint32_t nseg, uint32_t *seg_v,
                               unsigned long long *seg_slot) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    const uint64_t cap = off[nv];
    const uint64_t per = (cap + nseg - 1) / nseg;
    uint32_t a = 0;
    for (uint32_t s = 0; s < nseg; ++s) {
        seg_v[s] = a;
        seg_slot[s] = off[a];
        if (a < nv) {
            uint32_t lo = a + 1, hi = nv;  // first b in [a + 1, nv] with off[b] - off[a] >= per (nv if none)
            while (lo < hi) {
                const uint32_t m = lo + (hi - lo) / 2;
                if (off[m] - off[a] >= per) hi = m; else lo = m + 1;
            }
            a = lo;
        }
    }
    seg_v[nseg] = nv;
    seg_slot[nseg] = cap;
}
// Path position v + dpre[v] of every cut (merge-path tiles of the sweep).
__global__ void k_segment_paths(uint32_t nseg, const uint32_t *seg_v, const unsigned long long *dpre,
                                unsigned long long *seg_path) {
    for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s <= nseg; s += gridDim.x * blockDim.x)
        seg_path[s] = (unsigned long long)seg_v[s] + dpre[seg_v[s]];
}

// HBM edge cache (compute_hot_vertices / LoadCache, framework.cuh:149-279):
// rank vertices by the activations of the last 4 snapshots, cache the longest
// prefix of the ranking whose lists fit, copy each list from the old cache
// when it is there and unchanged, else from the edge array.
__global__ void k_cache_keys(uint32_t nv, const uint32_t *deg, const uint8_t *h0, const uint8_t *h1, const uint8_t *h2,
                             const uint8_t *h3, unsigned long long *keys) {
    for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < nv; v += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t score = (uint32_t)h0[v] + h1[v] + h2[v] + h3[v];
        keys[v] = (score && deg[v]) ? ((unsigned long long)(1023u - score) << 32) | v : ~0ull;
    }
}
struct CacheDegAt {
    const unsigned long long *keys;
    const uint32_t *deg;
    uint32_t n;
    __host__ __device__ unsigned long long operator()(uint32_t i) const {
        if (i >= n) return 0ull;
        const unsigned long long k = keys[i];
        return k == ~0ull ? 0ull : deg[(uint32_t)k];
    }
};
__global__ void k_cache_assign(uint32_t nv, const unsigned long long *keys, const unsigned long long *pos,
                               const uint32_t *deg, uint64_t slots, uint64_t *coff_new, unsigned long long *count) {
    unsigned long long c = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nv; i += (uint64_t)gridDim.x * blockDim.x) {
        const unsigned long long k = keys[i];
        if (k == ~0ull) continue;
        const uint32_t v = (uint32_t)k;
        if (pos[i] + deg[v] <= slots) {
            coff_new[v] = pos[i];
            ++c;
        }
    }
    if (c) atomicAdd(count, c);
}
// One wave per ranked vertex: fill its slot range of the new cache.
__global__ void __launch_bounds__(256) k_cache_fill(uint32_t nv, const unsigned long long *keys, const uint64_t *coff_new,
                                                    const uint64_t *coff_old, const uint32_t *ccol_old, const uint32_t *col,
                                                    const uint64_t *off, const uint32_t *deg, uint32_t *ccol_new,
                                                    unsigned long long *fetched) {
    const uint64_t gw = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
    const uint64_t nw = ((uint64_t)gridDim.x * blockDim.x) / kWave;
    unsigned long long f = 0;
    for (uint64_t i = gw; i < nv; i += nw) {
        const unsigned long long k = keys[i];
        if (k == ~0ull) break;  // keys are sorted: the invalid ones come last
        const uint32_t v = (uint32_t)k;
        const uint64_t dst = coff_new[v];
        if (dst == ~0ull) continue;
        const uint32_t d = deg[v];
        const uint64_t old = coff_old[v];
        const uint32_t *src = old != ~0ull ? ccol_old + old : col + off[v];
        if (old == ~0ull) f += d;
        // four loads in flight per lane: over the host link a load takes
        // microseconds, so one at a time would leave the link idle
        for (uint32_t e0 = 0; e0 < d; e0 += 4 * kWave) {
            uint32_t x[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const uint32_t e = e0 + j * kWave + lane_id();
                x[j] = e < d ? src[e] : 0u;
            }
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const uint32_t e = e0 + j * kWave + lane_id();
                if (e < d) ccol_new[dst + e] = x[j];
            }
        }
    }
    if (lane_id() == 0 && f) atomicAdd(fetched, f);
}
// Lists that the batch changed leave the cache.
__global__ void k_cache_drop(const TInfo *tinfo, const uint32_t *tcount, uint64_t *coff) {
    const uint32_t n = *tcount;
    for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (uint64_t)gridDim.x * blockDim.x)
        coff[tinfo[t].u] = ~0ull;
}

// Split (u, v) pairs into sort keys; drop ids >= nv. swap = 1 reads every
// pair as (v, u): the batch of the reverse (in-edge) store.
__global__ void k_split(uint32_t nv, const uint32_t *ins_uv, uint32_t nins, const uint32_t *del_uv, uint32_t ndel,
                        uint32_t *ins_k, uint32_t *ins_v, unsigned long long *del_k, Ctl *ctl, int swap) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const int a = swap ? 1 : 0, b = swap ? 0 : 1;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < (uint64_t)nins + ndel; i += stride) {
        if (i < nins) {
            const uint32_t u = ins_uv[2 * i + a], v = ins_uv[2 * i + b];
            const bool ok = u < nv && v < nv;
            ins_k[i] = ok ? u : kGap;
            ins_v[i] = v;
            if (!ok) atomicAdd(&ctl->dropped, 1ull);
        } else {
            const uint64_t j = i - nins;
            const uint32_t u = del_uv[2 * j + a], v = del_uv[2 * j + b];
            const bool ok = u < nv && v < nv;
            del_k[j] = ok ? (((unsigned long long)u << 32) | v) : ~0ull;
            if (!ok) atomicAdd(&ctl->dropped, 1ull);
        }
    }
}

__device__ __forceinline__ uint64_t off_at(const uint64_t *off, uint32_t nv, uint64_t cap, uint64_t v) {
    return v >= nv ? cap : off[v];
}

// New degree of every touched vertex from the chunk prefix; vertices that
// outgrow their range mark their leaf dirty. Leaf and edge totals follow.
__global__ void __launch_bounds__(kBlock) k_newdeg(GraphView g, const TInfo *tinfo, const uint32_t *tcount,
                                                   const unsigned long long *cpos, uint32_t *newdeg, uint8_t *leaf_dirty,
                                                   uint32_t L, unsigned long long *leaf_cnt, Ctl *ctl) {
    __shared__ unsigned long long s_sum[kWavesPerBlock];
    __shared__ bool s_over[kWavesPerBlock];
    const uint32_t n = *tcount;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t base = (uint64_t)blockIdx.x * blockDim.x; base < n; base += stride) {
        const uint64_t t = base + threadIdx.x;
        long long delta = 0;
        bool over = false;
        if (t < n) {
            const TInfo ti = tinfo[t];
            const uint32_t nd = has_del(ti) ? (uint32_t)(cpos[ti.cbase + ti.nch] - cpos[ti.cbase]) : ti.d0 + ti.ni;
            newdeg[ti.u] = nd;
            if (nd > g.off[ti.u + 1] - g.off[ti.u]) {
                leaf_dirty[ti.u / L] = 1;
                over = true;
            }
            delta = (long long)nd - (long long)ti.d0;
            if (delta) atomicAdd(&leaf_cnt[ti.u / L], (unsigned long long)delta);
        }
        // one atomic per counter and block
        const unsigned long long wsum = wave_sum_u64((unsigned long long)delta);
        const bool wover = __ballot(over) != 0;
        const uint32_t w = threadIdx.x / kWave;
        if (lane_id() == 0) {
            s_sum[w] = wsum;
            s_over[w] = wover;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            unsigned long long a = 0;
            bool o = false;
            for (int i = 0; i < kWavesPerBlock; ++i) {
                a += s_sum[i];
                o = o || s_over[i];
            }
            if (a) atomicAdd(&ctl->ne_delta, a);
            if (o) atomicOr(&ctl->overflow, 1ull);
        }
        __syncthreads();
    }
}

// deg[u] = new degree for touched vertices placed in their own range.
// Upper density threshold at height h of a tree of height H
// (up_0 = 1.0 at the leaves, up_h = 0.75 at the root, csr_graph.cuh:345-346, 847-851).
__device__ __host__ __forceinline__ double up_threshold(uint32_t h, uint32_t H) {
    return H == 0 ? 0.75 : 1.0 - 0.25 * (double)h / (double)H;
}

// Each dirty leaf climbs the implicit tree (heap numbering: root 1, leaves
// [nleaves, 2 nleaves)) to the first window whose density is under its
// threshold (rebalance_wrapper, csr_graph.cuh:813-863).
__global__ void k_choose_windows(uint32_t nv, uint32_t L, uint64_t nleaves, uint32_t H, const uint64_t *off, uint64_t cap,
                                 const unsigned long long *leaf_pre, uint8_t *leaf_dirty, uint8_t *chosen, Ctl *ctl) {
    for (uint64_t l = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; l < nleaves; l += (uint64_t)gridDim.x * blockDim.x) {
        if (!leaf_dirty[l]) continue;
        leaf_dirty[l] = 0;
        bool done = false;
        for (uint32_t h = 0; h <= H && !done; ++h) {
            const uint64_t lo = (l >> h) << h, hi = umin64(lo + (1ull << h), nleaves);
            const uint64_t vlo = umin64(lo * L, nv), vhi = umin64(hi * L, nv);
            if (vlo >= vhi) continue;
            const uint64_t total = off_at(off, nv, cap, vhi) - off_at(off, nv, cap, vlo);
            const uint64_t actual = leaf_pre[hi] - leaf_pre[lo];
            if (total > 0 && (double)actual < up_threshold(h, H) * (double)total) {
                chosen[(nleaves >> h) + (l >> h)] = 1;
                done = true;
            }
        }
        if (!done) atomicOr(&ctl->resize, 1ull);
    }
}

struct Window {
    uint32_t vlo, vhi;
    uint64_t slo, shi;   // slot range
    uint64_t pbase;      // P (exclusive prefix of new degrees) at vlo
    uint64_t actual;     // live edges after the batch
    uint64_t gaps;
};

// Keep only chosen nodes with no chosen ancestor; list them as windows.
__global__ void k_prune(uint32_t nv, uint32_t L, uint64_t nleaves, uint8_t *chosen, uint8_t *maximal, Window *win,
                        Ctl *ctl, uint32_t max_win) {
    for (uint64_t node = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x + 1; node < 2 * nleaves;
         node += (uint64_t)gridDim.x * blockDim.x) {
        if (!chosen[node]) continue;
        bool top = true;
        for (uint64_t a = node >> 1; a >= 1; a >>= 1)
            if (chosen[a]) { top = false; break; }
        if (!top) continue;
        // node -> leaf range
        uint32_t depth = 63 - __clzll((long long)node);
        uint32_t leaf_depth = 63 - __clzll((long long)nleaves);
        const uint32_t h = leaf_depth - depth;
        const uint64_t lo = (node - (1ull << depth)) << h, hi = umin64(lo + (1ull << h), nleaves);
        const uint64_t vlo = umin64(lo * L, nv), vhi = umin64(hi * L, nv);
        if (vlo >= vhi) continue;
        maximal[node] = 1;
        const unsigned long long w = atomicAdd(&ctl->nwin, 1ull);
        if (w < max_win) { win[w].vlo = (uint32_t)vlo; win[w].vhi = (uint32_t)vhi; }
    }
}

__global__ void k_win_sizes(const Window *win, uint32_t nwin, unsigned long long *wsz) {
    for (uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; w <= nwin; w += (uint64_t)gridDim.x * blockDim.x)
        wsz[w] = (w == nwin) ? 0ull : (unsigned long long)(win[w].vhi - win[w].vlo);
}

__device__ __forceinline__ uint32_t find_window(const unsigned long long *wv_off, uint32_t nwin, uint64_t i) {
    uint32_t lo = 0, hi = nwin;  // largest w with wv_off[w] <= i
    while (hi - lo > 1) { uint32_t m = (lo + hi) >> 1; if (wv_off[m] <= i) lo = m; else hi = m; }
    return lo;
}

__global__ void k_win_nd(const Window *win, uint32_t nwin, const unsigned long long *wv_off, uint64_t Wv,
                         const uint32_t *mark, const uint32_t *newdeg, const uint32_t *deg, unsigned long long *nd) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i <= Wv; i += (uint64_t)gridDim.x * blockDim.x) {
        if (i == Wv) { nd[i] = 0; continue; }
        const uint32_t w = find_window(wv_off, nwin, i);
        const uint32_t v = win[w].vlo + (uint32_t)(i - wv_off[w]);
        nd[i] = bit_test(mark, v) ? newdeg[v] : deg[v];
    }
}

__global__ void k_win_params(Window *win, uint32_t nwin, const unsigned long long *wv_off, const unsigned long long *P,
                             const uint64_t *off, uint32_t nv, uint64_t cap, uint64_t new_cap) {
    for (uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; w < nwin; w += (uint64_t)gridDim.x * blockDim.x) {
        Window x = win[w];
        x.pbase = P[wv_off[w]];
        x.actual = P[wv_off[w + 1]] - x.pbase;
        if (new_cap) { x.slo = 0; x.shi = new_cap; }
        else { x.slo = off_at(off, nv, cap, x.vlo); x.shi = off_at(off, nv, cap, x.vhi); }
        x.gaps = (x.shi - x.slo) - x.actual;
        win[w] = x;
    }
}

// Merged lists of window vertices, packed into out at P[i]. A touched
// vertex's merged list is in `merged` (old survivors then inserts when it had
// deletes; only its inserts otherwise, after its unchanged old list).
__global__ void __launch_bounds__(kBlock) k_win_gather(GraphView g, BatchView bv, const Window *win, uint32_t nwin,
                                                       const unsigned long long *wv_off, uint64_t Wv,
                                                       const unsigned long long *P, const uint32_t *mark,
                                                       const uint32_t *tidx, const TInfo *tinfo,
                                                       const unsigned long long *cpos, const MChunk *chunks,
                                                       const uint32_t *merged, uint32_t *out) {
    const uint64_t gw = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
    const uint64_t nw = ((uint64_t)gridDim.x * blockDim.x) / kWave;
    const uint32_t lane = lane_id();
    for (uint64_t i = gw; i < Wv; i += nw) {
        const uint32_t w = find_window(wv_off, nwin, i);
        const uint32_t v = win[w].vlo + (uint32_t)(i - wv_off[w]);
        const uint32_t *old = g.col + g.off[v];
        uint32_t *dst = out + P[i];
        uint32_t keep_old = g.deg[v];
        const uint32_t *tail = nullptr;
        uint64_t ntail = 0;
        if (bit_test(mark, v)) {
            const TInfo ti = tinfo[tidx[v]];
            if (ti.de > ti.db && chunks[ti.cbase].ibase == kInPlace) {  // compacted where it lies
                keep_old = (uint32_t)(cpos[ti.cbase + 1] - cpos[ti.cbase]);
            } else if (ti.de > ti.db) {  // survivors of old list + inserts, staged by the filter chunk by chunk
                for (uint32_t c = ti.cbase; c < ti.cbase + ti.nch; ++c) {
                    const uint32_t *src = merged + chunks[c].ibase;
                    const uint64_t at = cpos[c] - cpos[ti.cbase], cnt = cpos[c + 1] - cpos[c];
                    for (uint64_t k = lane; k < cnt; k += kWave) dst[at + k] = src[k];
                }
                continue;
            } else {  // old list unchanged, inserts straight from the sorted batch
                tail = bv.ins_val + ti.ib;
                ntail = ti.ni;
            }
        }
        for (uint32_t k = lane; k < keep_old; k += kWave) dst[k] = old[k];
        for (uint64_t k = lane; k < ntail; k += kWave) dst[keep_old + k] = tail[k];
    }
}

__device__ __forceinline__ uint64_t win_pos(const Window &x, uint64_t pl, uint64_t il, uint64_t n) {
    if (il >= n) return x.shi;
    const double denom = (double)(x.actual + n);
    return x.slo + pl + (uint64_t)floor((double)x.gaps * (double)(pl + il) / denom);
}

// Write the window back with gaps spread in proportion to new degree + 1.
__global__ void __launch_bounds__(kBlock) k_win_scatter(const Window *win, uint32_t nwin, const unsigned long long *wv_off,
                                                        uint64_t Wv, const unsigned long long *P, const uint32_t *tmp,
                                                        uint32_t *dst_col, uint64_t *off, uint32_t *deg) {
    const uint64_t gw = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
    const uint64_t nw = ((uint64_t)gridDim.x * blockDim.x) / kWave;
    const uint32_t lane = lane_id();
    for (uint64_t i = gw; i < Wv; i += nw) {
        const uint32_t w = find_window(wv_off, nwin, i);
        const Window x = win[w];
        const uint64_t il = i - wv_off[w], n = x.vhi - x.vlo;
        const uint64_t pl = P[i] - x.pbase;
        const uint64_t nd = P[i + 1] - P[i];
        const uint64_t a = win_pos(x, pl, il, n);
        const uint64_t b = win_pos(x, pl + nd, il + 1, n);
        const uint32_t *src = tmp + P[i];
        for (uint64_t k = lane; k < nd; k += kWave) dst_col[a + k] = src[k];
        for (uint64_t k = a + nd + lane; k < b; k += kWave) dst_col[k] = kGap;
        if (lane == 0) {
            off[x.vlo + il] = a;
            deg[x.vlo + il] = (uint32_t)nd;
        }
    }
}

__global__ void k_set_u64(uint64_t *p, uint64_t v) { *p = v; }

// ---------------------------------------------------------------------------
// operator-generic kernels
// ---------------------------------------------------------------------------
template <class Op>
__global__ void k_init_state(uint32_t nv, Op op, typename Op::value_t *value, typename Op::buffer_t *buffer) {
    for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < nv; v += (uint64_t)gridDim.x * blockDim.x) {
        value[v] = op.GetInitValue((uint32_t)v);
        buffer[v] = op.GetInitBuffer((uint32_t)v);
    }
}

// RebuildWorklist (driver.cuh:654-674): every active vertex joins the frontier.
template <class Op, bool kBothSigns = false>
__global__ void __launch_bounds__(kBlock) k_seed_active(uint32_t nv, Op op, const typename Op::value_t *value,
                                                        const typename Op::buffer_t *buffer, uint32_t *stamp,
                                                        uint32_t stamp_val, uint32_t *list, uint32_t *cnt,
                                                        unsigned long long *overflow) {
    __shared__ BlockAppend sh;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t base = (uint64_t)blockIdx.x * blockDim.x; base < nv; base += stride) {
        const uint64_t v = base + threadIdx.x;
        bool p = false;
        if (v < nv) {
            p = op.IsActiveNode((uint32_t)v, buffer[v], value[v]);
            if constexpr (kBothSigns) p = p || op.IsActiveNode_INC((uint32_t)v, buffer[v], value[v]);
            if (p) stamp[v] = stamp_val;
        }
        block_append(p, (uint32_t)v, list, cnt, nv, overflow, sh);
    }
}

// reset_del_edges (driver.cuh:399-430): a deleted tree edge (u, v) resets v.
template <class Op>
__global__ void __launch_bounds__(kBlock) k_inval_seed(Op op, const unsigned long long *del_k, uint32_t ndel,
                                                       typename Op::value_t *value, typename Op::buffer_t *buffer,
                                                       uint32_t *inval, uint32_t *coarse, uint32_t cshift, uint32_t *list,
                                                       uint32_t *cnt, uint32_t cap, unsigned long long *stats) {
    __shared__ BlockAppend sh;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t base = (uint64_t)blockIdx.x * blockDim.x; base < ndel; base += stride) {
        const uint64_t i = base + threadIdx.x;
        bool p = false;
        uint32_t v = 0;
        if (i < ndel && del_k[i] != ~0ull) {
            const uint32_t u = (uint32_t)(del_k[i] >> 32);
            v = (uint32_t)del_k[i];
            const uint64_t b = ld_relaxed(&buffer[v]);
            if (Op::Parent(b) == u) {
                const uint32_t bit = 1u << (v & 31);
                if (!(atomicOr(&inval[v >> 5], bit) & bit)) {
                    coarse_mark(coarse, cshift, v);
                    typename Op::buffer_t bb = b;
                    op.AccumulateBuffer_del(u, v, &bb, &value[v]);
                    buffer[v] = bb;
                    p = true;
                }
            }
        }
        block_append(p, v, list, cnt, cap, stats + 3, sh);
    }
}

// Relax each inserted edge that survived the batch's deletes.
template <class Op>
__global__ void __launch_bounds__(kBlock) k_insert_relax(Op op, BatchView bv, BatchRanges br, typename Op::buffer_t *buffer,
                                                         uint32_t *stamp, uint32_t stamp_val, uint32_t wmode,
                                                         uint64_t wseed, uint32_t *list, uint32_t *cnt, uint32_t cap,
                                                         unsigned long long *stats) {
    __shared__ BlockAppend sh;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t base = (uint64_t)blockIdx.x * blockDim.x; base < bv.nins; base += stride) {
        const uint64_t i = base + threadIdx.x;
        bool p = false;
        uint32_t v = 0;
        if (i < bv.nins && bv.ins_key[i] != kGap) {
            const uint32_t u = bv.ins_key[i];
            v = bv.ins_val[i];
            const unsigned long long key = ((unsigned long long)u << 32) | v;
            const uint32_t db = br.del_b[u], de = br.del_e[u];
            if (db == de || !in_sorted_u64(bv.del_key, db, de, key)) {
                const uint32_t w = Op::kWeighted ? sg_weight(wmode, wseed, u, v) : 0u;
                if (op.AccumulateBuffer(u, v, w, &buffer[v], op.PushValue(ld_relaxed(&buffer[u]))))
                    p = atomicExch(&stamp[v], stamp_val) != stamp_val;
            }
        }
        block_append(p, v, list, cnt, cap, stats + 3, sh);
    }
}

// FP64 residual reduction of PageRank: out[0] = sum |r|, out[1] = max |r| (bits).
__global__ void __launch_bounds__(kBlock) k_pr_residual(uint32_t nv, const float *res, double *sum, unsigned long long *mx) {
    __shared__ double s_sum[kWavesPerBlock];
    __shared__ double s_max[kWavesPerBlock];
    double a = 0.0, m = 0.0;
    for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < nv; v += (uint64_t)gridDim.x * blockDim.x) {
        const double x = fabs((double)res[v]);
        a += x;
        m = fmax(m, x);
    }
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
        a += __shfl_xor(a, d, kWave);
        m = fmax(m, __shfl_xor(m, d, kWave));
    }
    const uint32_t w = threadIdx.x / kWave;
    if (lane_id() == 0) { s_sum[w] = a; s_max[w] = m; }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int i = 1; i < kWavesPerBlock; ++i) { a += s_sum[i]; m = fmax(m, s_max[i]); }
        unsafeAtomicAdd(sum, a);
        atomicMax(mx, (unsigned long long)__double_as_longlong(m));
    }
}

template <class Op>
__global__ void k_extract(uint32_t nv, const typename Op::buffer_t *buffer, uint32_t *out, int hi) {
    for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < nv; v += (uint64_t)gridDim.x * blockDim.x)
        out[v] = hi ? (uint32_t)(buffer[v] >> 32) : (uint32_t)buffer[v];
}

// Pointwise gather: out[i] = 32-bit word base[verts[i] * stride + word].
__global__ void k_gather_words(const uint32_t *base, uint64_t stride, uint32_t word, uint32_t nv, const uint32_t *verts,
                               uint64_t n, uint32_t *out) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t v = verts[i];
        out[i] = v < nv ? base[(uint64_t)v * stride + word] : 0xFFFFFFFFu;
    }
}
__global__ void k_iota(uint32_t *p, uint32_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = (uint32_t)i;
}

// ---------------------------------------------------------------------------
// engine
// ---------------------------------------------------------------------------
static uint32_t grid_for(uint64_t n, uint32_t per_block, uint32_t cap = 65535) {
    uint64_t g = (n + per_block - 1) / per_block;
    return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(g, cap));
}

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
};

// fn(lo, hi) over [0, n) split among up to 16 host threads (large loads:
// validation and layout of edge arrays with billions of entries).
template <class Fn>
static void parallel_for(uint64_t n, Fn fn) {
    const uint64_t nt = std::max<uint64_t>(1, std::min<uint64_t>({16, std::thread::hardware_concurrency(), n / (1u << 20) + 1}));
    if (nt == 1) { fn((uint64_t)0, n); return; }
    std::vector<std::thread> th;
    for (uint64_t t = 0; t < nt; ++t) th.emplace_back(fn, n * t / nt, n * (t + 1) / nt);
    for (auto &x : th) x.join();
}

class Engine {
public:
    explicit Engine(const sg_config &cfg) : cfg_(cfg) {}
    virtual ~Engine() { release(); }

    sg_config cfg_;
    std::string last_error_ = "no error";
    hipStream_t stream_ = nullptr;
    bool own_stream_ = true;
    bool swap_pairs_ = false;  // batches are read as (v, u): the in-edge store of another engine
    sg_stats st_{};

    // --- memory -------------------------------------------------------------
    uint64_t dev_bytes_ = 0;
    std::vector<std::pair<void *, size_t>> allocs_;
    uint32_t *scratch_ = nullptr;  // this device's address of the file-scope scratch

    // An engine and the stores it owns share one device budget: each side
    // counts the other's bytes.
    const Engine *budget_peer_ = nullptr;
    uint64_t budget_used() const {
        return dev_bytes_ + aux_device_bytes() + (budget_peer_ ? budget_peer_->dev_bytes_ : 0);
    }
    template <class T>
    T *dalloc(size_t count) {
        const size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
        const double budget = cfg_.device_budget_gb * 1073741824.0;
        if (budget > 0 && (double)(budget_used() + bytes) > budget)
            fail(SG_EBUDGET, "device budget exceeded: need " + std::to_string(budget_used() + bytes) + " bytes");
        void *p = nullptr;
        hipError_t e = hipMalloc(&p, bytes);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            fail(SG_EBUDGET, std::string("hipMalloc of ") + std::to_string(bytes) + " bytes failed: " + hipGetErrorString(e));
        }
        dev_bytes_ += bytes;
        allocs_.push_back({p, bytes});
        return (T *)p;
    }
    void dfree(void *p) {
        if (!p) return;
        for (size_t i = 0; i < allocs_.size(); ++i) {
            if (allocs_[i].first == p) {
                dev_bytes_ -= allocs_[i].second;
                allocs_[i] = allocs_.back();
                allocs_.pop_back();
                (void)hipFree(p);
                return;
            }
        }
    }
    template <class T>
    void ensure(T *&p, size_t &cap, size_t need) {
        if (need <= cap && p) return;
        size_t nc = std::max<size_t>(need, cap + cap / 2);
        dfree(p);
        p = nullptr;
        p = dalloc<T>(nc);
        cap = nc;
    }
    // Pinned host memory, mapped into the device address space (the edge array
    // of the out-of-memory path and its load staging).
    uint64_t host_bytes_ = 0;
    std::vector<std::pair<void *, size_t>> hallocs_;
    template <class T>
    T *halloc(size_t count) {
        const size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
        void *p = nullptr;
        hipError_t e = hipHostMalloc(&p, bytes, hipHostMallocMapped);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            fail(SG_ENOMEM, std::string("hipHostMalloc of ") + std::to_string(bytes) + " bytes failed: " + hipGetErrorString(e));
        }
        void *dp = nullptr;
        SG_HIP(hipHostGetDevicePointer(&dp, p, 0));
        if (dp != p) {
            (void)hipHostFree(p);
            fail(SG_EDEVICE, "pinned host memory is not addressable by the device at its host address");
        }
        host_bytes_ += bytes;
        hallocs_.push_back({p, bytes});
        return (T *)p;
    }
    void hfree(void *p) {
        if (!p) return;
        for (size_t i = 0; i < hallocs_.size(); ++i) {
            if (hallocs_[i].first == p) {
                host_bytes_ -= hallocs_[i].second;
                hallocs_[i] = hallocs_.back();
                hallocs_.pop_back();
                (void)hipHostFree(p);
                return;
            }
        }
    }

    void release() {
        if (stream_) (void)hipStreamSynchronize(stream_);
        if (cstream_) (void)hipStreamSynchronize(cstream_);
        for (auto &a : allocs_) (void)hipFree(a.first);
        allocs_.clear();
        dev_bytes_ = 0;
        for (auto &a : hallocs_) (void)hipHostFree(a.first);
        hallocs_.clear();
        host_bytes_ = 0;
        for (auto &ev : seg_copied_) if (ev) (void)hipEventDestroy(ev);
        for (auto &ev : seg_used_) if (ev) (void)hipEventDestroy(ev);
        if (sync_ev_) (void)hipEventDestroy(sync_ev_);
        sync_ev_ = nullptr;
        if (cstream_) (void)hipStreamDestroy(cstream_);
        cstream_ = nullptr;
        release_graphs();
        for (auto &ev : events_) { (void)hipEventDestroy(ev.first); (void)hipEventDestroy(ev.second); }
        events_.clear();
        if (pinned_) (void)hipHostFree(pinned_);
        pinned_ = nullptr;
        if (stream_ && own_stream_) (void)hipStreamDestroy(stream_);
        stream_ = nullptr;
    }

    // --- graph --------------------------------------------------------------
    uint32_t nv_ = 0;
    uint64_t ne_ = 0, cap_ = 0;
    uint64_t *off_ = nullptr;
    uint32_t *deg_ = nullptr;
    uint32_t *col_ = nullptr;
    uint32_t L_ = 1;          // vertices per PMA leaf
    uint64_t nleaves_ = 1;
    uint32_t H_ = 0;          // tree height
    unsigned long long *leaf_cnt_ = nullptr, *leaf_pre_ = nullptr;
    uint8_t *leaf_dirty_ = nullptr, *chosen_ = nullptr, *maximal_ = nullptr;
    bool loaded_ = false, inited_ = false;

    // --- frontier -----------------------------------------------------------
    uint32_t *q_[2] = {nullptr, nullptr};
    // Control block (device, one allocation, read back by one copy):
    //   u32 [0..2] frontier counters, [3..5] spill counters, [7] always zero,
    //   u32 [8, 8 + 2 * kRoundLog) round log, then u64 stats[16]:
    //   [0] edges (light), [1] vertices, [2] edges (spilled chunks), [3] overflow
    //   flags, [4] pull rounds, [5] edges read from the HBM cache, [6..7] cache
    //   rebuild counts, [8..15] the same as [0..7] for timed rounds.
    //   Then u64 ecnt[4]: out-edges of the frontiers of rounds R % 3 (direction choice).
    static constexpr size_t kCtlBlockBytes = (8 + 2 * kRoundLog) * 4 + 16 * 8 + 4 * 8;
    uint32_t *cnt_ = nullptr;    // [3] rotating round counters
    uint32_t *hcnt_ = nullptr;   // [3] rotating spill counters
    uint32_t *rlog_ = nullptr;
    uint64_t timed_edges_ = 0, timed_verts_ = 0, timed_cedges_ = 0, timed_pulls_ = 0;  // host totals of the timed-round stats
    uint64_t last_appended_ = 0;  // vertices the last run of rounds (and its seeds) put on frontiers
    HeavyDesc *heavy_[3] = {nullptr, nullptr, nullptr};  // rotating spill queues
    size_t heavy_cap_ = 0;
    unsigned long long *stats_ = nullptr;  // [0..3] pass stats, [4] inval count
    unsigned long long *ecnt_ = nullptr;   // [3] rotating frontier out-edge counters
    uint32_t *stamp_ = nullptr;
    uint32_t *inval_ = nullptr;   // bitmap of invalidated vertices
    uint32_t *coarse_ = nullptr;  // coarse filter of inval_: one bit per 2^cshift_ vertices
    uint32_t cshift_ = 6;
    // phases whose round counts seed the next snapshot's first batch (a second
    // program on the same engine has its own pair)
    static constexpr int kPhaseInval = 0, kPhaseRelax = 1, kPhaseInval2 = 2, kPhaseRelax2 = 3;
    uint32_t last_rounds_[4] = {0, 0, 0, 0};  // rounds the phase took at the last snapshot
    std::unordered_map<const void *, int> occ_;  // resident blocks per CU of a kernel
    int num_cus_ = 0;
    uint32_t R_ = 1;             // global round counter (stamps)

    // --- batch --------------------------------------------------------------
    uint32_t *raw_ins_ = nullptr, *raw_del_ = nullptr;
    size_t raw_ins_cap_ = 0, raw_del_cap_ = 0;
    uint32_t *ins_k_ = nullptr, *ins_v_ = nullptr, *ins_k2_ = nullptr, *ins_v2_ = nullptr;
    size_t ins_cap_ = 0, ins_cap2_ = 0, ins_cap3_ = 0, ins_cap4_ = 0;
    unsigned long long *del_k_ = nullptr, *del_k2_ = nullptr;
    size_t del_cap_ = 0, del_cap2_ = 0;
    uint8_t *cub_tmp_ = nullptr;
    size_t cub_cap_ = 0;
    uint32_t *mark_ = nullptr;    // bitmap of touched sources
    BatchRanges br_{};             // per-vertex [begin, end) of its inserts / deletes in the sorted batch
    unsigned long long *dpre_ = nullptr;  // exclusive prefix of deg (edge-centric sweep)
    unsigned long long *bsum_ = nullptr;  // per-block degree sums of the scan
    size_t bsum_cap_ = 0;
    uint32_t *tile_v_ = nullptr;          // merge-path sweep tiles: first source ...
    unsigned long long *tile_e_ = nullptr;  // ... and first live edge of each tile
    size_t tile_v_cap_ = 0, tile_e_cap_ = 0;
    uint32_t *touched_ = nullptr, *newdeg_ = nullptr, *tidx_ = nullptr;
    TInfo *tinfo_ = nullptr;
    size_t tinfo_cap_ = 0;
    MChunk *chunks_ = nullptr;
    size_t chunks_cap_ = 0;
    unsigned long long *ccount_ = nullptr, *cpos_ = nullptr;
    size_t ccount_cap_ = 0, cpos_cap_ = 0;
    uint64_t chunk_bound_ = 0;
    uint32_t *tmp2_ = nullptr;
    size_t tmp2_cap_ = 0;
    Ctl *ctl_ = nullptr;
    uint32_t *tmp_ = nullptr;
    size_t tmp_cap_ = 0;
    Window *win_ = nullptr;
    size_t win_cap_ = 0;
    unsigned long long *wsz_ = nullptr, *wvoff_ = nullptr, *ndarr_ = nullptr, *P_arr_ = nullptr;
    size_t wsz_cap_ = 0, wvoff_cap_ = 0, ndarr_cap_ = 0, Parr_cap_ = 0;
    BatchView bv_{};
    uint32_t nins_ = 0, ndel_ = 0;

    // --- out-of-memory path ------------------------------------------------------
    bool host_edges_ = false;                 // col_ is pinned host memory
    uint32_t nseg_ = 0;                       // SEGMENT count
    uint32_t *seg_v_ = nullptr;               // [nseg + 1] first vertex of each segment
    unsigned long long *seg_aux_ = nullptr;   // [2 * (nseg + 1)]: slot and path position of each cut
    std::vector<uint32_t> h_seg_v_;
    std::vector<unsigned long long> h_seg_slot_, h_seg_path_;
    bool seg_dirty_ = true;
    hipStream_t cstream_ = nullptr;           // copy stream of the streamed segments
    static constexpr int kMaxStage = 8;       // staging buffers (n_stream segment copies in flight)
    hipEvent_t seg_copied_[kMaxStage] = {}, seg_used_[kMaxStage] = {};
    hipEvent_t sync_ev_ = nullptr;            // polled by sync()
    uint32_t *stage_[kMaxStage] = {};
    size_t stage_cap_[kMaxStage] = {};
    int nstage() const { return (int)std::max<uint32_t>(2, std::min<uint32_t>(cfg_.n_stream, kMaxStage)); }
    // HBM edge cache (host edges only)
    uint64_t cache_slots_ = 0;
    uint32_t *ccol_[2] = {nullptr, nullptr};
    uint64_t *coff_ = nullptr, *coff2_ = nullptr;
    uint8_t *hotw_[4] = {nullptr, nullptr, nullptr, nullptr};
    uint32_t hot_cur_ = 0, ccur_ = 0;
    unsigned long long *ckeys_ = nullptr, *ckeys2_ = nullptr, *cpos2_ = nullptr;

    // --- pinned host scratch --------------------------------------------------
    void *pinned_ = nullptr;
    Ctl *h_ctl_ = nullptr;
    uint32_t *h_blk_ = nullptr;   // host copy of the control block
    unsigned long long *h_u64_ = nullptr;  // [16]
    const unsigned long long *h_stats() const { return (const unsigned long long *)(h_blk_ + 8 + 2 * kRoundLog); }
    const unsigned long long *h_ecnt() const { return h_stats() + 16; }

    // --- kernel timing --------------------------------------------------------
    std::vector<std::pair<hipEvent_t, hipEvent_t>> events_;
    std::vector<uint32_t> ev_cls_;  // kernel class of each recorded pair
    std::vector<uint32_t> ev_n_;    // launches each pair covers
    size_t ev_used_ = 0;
    // per kernel class (sg_kernel_class): launches, summed event time, algorithmic bytes
    uint64_t kt_launches_[SG_KCLASS_COUNT] = {}, kt_bytes_[SG_KCLASS_COUNT] = {};
    double kt_ms_[SG_KCLASS_COUNT] = {};

    static constexpr uint32_t kGridList = 2048;   // light-kernel blocks for device-sized lists
    // the chunk filter: one wave per chunk of a vertex with deletes, a latency
    // chain each; 8192 blocks (about one chunk per wave at 100k deletes) beat
    // 2048 by 20 us per batch at scale 22
    static constexpr uint32_t kGridFilter = 8192;
    static constexpr uint32_t kGridHeavy = 1024;

    // shared: run on another engine's stream (a store owned by that engine,
    // whose work must be ordered with the owner's kernels); null = own stream.
    void init_device(hipStream_t shared = nullptr) {
        int ndev = 0;
        SG_HIP(hipGetDeviceCount(&ndev));
        if (cfg_.device < 0 || cfg_.device >= ndev) fail(SG_EDEVICE, "no HIP device " + std::to_string(cfg_.device));
        SG_HIP(hipSetDevice(cfg_.device));
        hipDeviceProp_t prop;
        SG_HIP(hipGetDeviceProperties(&prop, cfg_.device));
        if (prop.warpSize != 64) fail(SG_EDEVICE, "wave64 device required");
        num_cus_ = prop.multiProcessorCount;
        if (shared) {
            stream_ = shared;
            own_stream_ = false;
        } else {
            SG_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
        }
        {  // this device's copy of the no-allocation scratch (device current here)
            std::lock_guard<std::mutex> lk(g_scratch_mu);
            if ((size_t)cfg_.device >= g_scratch_addr.size()) g_scratch_addr.resize((size_t)cfg_.device + 1, nullptr);
            if (!g_scratch_addr[cfg_.device]) {
                void *p = nullptr;
                SG_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
                g_scratch_addr[cfg_.device] = (uint32_t *)p;
            }
            scratch_ = g_scratch_addr[cfg_.device];
        }
        SG_HIP(hipEventCreateWithFlags(&sync_ev_, hipEventDisableTiming));
        SG_HIP(hipHostMalloc(&pinned_, 65536, hipHostMallocDefault));
        h_ctl_ = (Ctl *)pinned_;
        h_blk_ = (uint32_t *)((char *)pinned_ + 4096);
        h_u64_ = (unsigned long long *)((char *)pinned_ + 16384);
    }

    // Wait for the stream by polling an event: the host resumes as soon as the
    // work ends instead of after a blocking wake-up (the engine syncs a few
    // times per snapshot, and the GPU idles until the next launch).
    void sync() {
        SG_HIP(hipEventRecord(sync_ev_, stream_));
        for (;;) {
            const hipError_t e = hipEventQuery(sync_ev_);
            if (e == hipSuccess) return;
            if (e != hipErrorNotReady) SG_HIP(e);
            __builtin_ia32_pause();
        }
    }

    template <class T>
    void fill(T *p, uint64_t n, T v) {
        if (n == 0) return;
        const uint32_t g = grid_for(n, 256, 8192);
        if constexpr (sizeof(T) == 4) hipLaunchKernelGGL(k_fill_u32, dim3(g), dim3(256), 0, stream_, (uint32_t *)p, n, (uint32_t)v);
        else if constexpr (sizeof(T) == 1) hipLaunchKernelGGL(k_fill_u8, dim3(g), dim3(256), 0, stream_, (uint8_t *)p, n, (uint8_t)v);
        else hipLaunchKernelGGL(k_fill_u64, dim3(g), dim3(256), 0, stream_, (unsigned long long *)p, n, (unsigned long long)v);
        SG_LAUNCHED();
    }

    // ---- load ------------------------------------------------------------------
    void load_csr(uint32_t nv, uint64_t ne, const uint64_t *row, const uint32_t *col) {
        if (loaded_) fail(SG_EINVAL, "graph already loaded");
        if (nv == 0 || nv == 0xFFFFFFFFu) fail(SG_EINVAL, "nv must be in [1, 2^32-2]");
        if (!row || (ne && !col)) fail(SG_EINVAL, "null CSR array");
        if (row[0] != 0 || row[nv] != ne) fail(SG_EINVAL, "row[0] must be 0 and row[nv] must equal ne");
        for (uint32_t v = 0; v < nv; ++v)
            if (row[v + 1] < row[v]) fail(SG_EINVAL, "row offsets must be non-decreasing");
        {
            std::atomic<uint64_t> bad{~0ull};
            parallel_for(ne, [&](uint64_t lo, uint64_t hi) {
                for (uint64_t e = lo; e < hi; ++e)
                    if (col[e] >= nv) {
                        uint64_t cur = bad.load();
                        while (e < cur && !bad.compare_exchange_weak(cur, e)) {}
                        return;
                    }
            });
            if (bad.load() != ~0ull) fail(SG_EINVAL, "col entry >= nv at " + std::to_string(bad.load()));
        }
        nv_ = nv;
        ne_ = ne;
        const double dens = (cfg_.pma_density > 0 && cfg_.pma_density <= 1) ? cfg_.pma_density : 0.25;
        cap_ = std::max<uint64_t>({(uint64_t)std::ceil((double)ne / dens), ne + nv, 64});
        // leaf geometry (compute_capacity, csr_graph.cuh:364-382)
        if (nv < 2) { L_ = 1; nleaves_ = 1; }
        else {
            uint64_t s = 64 - __builtin_clzll((uint64_t)nv - 1);
            uint64_t c = (nv + s - 1) / s;
            c = 1ull << (63 - __builtin_clzll(c));
            L_ = (uint32_t)((nv + c - 1) / c);
            nleaves_ = c;
        }
        H_ = (uint32_t)(63 - __builtin_clzll(nleaves_));

        off_ = dalloc<uint64_t>((size_t)nv + 1);
        deg_ = dalloc<uint32_t>(nv);
        host_edges_ = (cfg_.flags & SG_FLAG_HOST_EDGES) ||
                      (cfg_.edge_budget_gb > 0 && (double)cap_ * 4.0 > cfg_.edge_budget_gb * 1073741824.0);
        col_ = host_edges_ ? halloc<uint32_t>(cap_) : dalloc<uint32_t>(cap_);
        nseg_ = std::max<uint32_t>(1, std::min<uint32_t>(cfg_.segments ? cfg_.segments : 1, 65536));
        seg_v_ = dalloc<uint32_t>(nseg_ + 1);
        seg_aux_ = dalloc<unsigned long long>(2 * ((size_t)nseg_ + 1));
        if (host_edges_) {
            // no larger than twice the edge array: the rest could never be filled
            cache_slots_ = std::min<uint64_t>(kCacheSlotsByLevel[std::min<uint32_t>(cfg_.cache_level, 4)], 2 * cap_);
            SG_HIP(hipStreamCreateWithFlags(&cstream_, hipStreamNonBlocking));
            for (int b = 0; b < kMaxStage; ++b) {
                SG_HIP(hipEventCreateWithFlags(&seg_copied_[b], hipEventDisableTiming));
                SG_HIP(hipEventCreateWithFlags(&seg_used_[b], hipEventDisableTiming));
            }
        }
        if (cache_slots_) {
            ccol_[0] = dalloc<uint32_t>(cache_slots_);
            ccol_[1] = dalloc<uint32_t>(cache_slots_);
            coff_ = dalloc<uint64_t>(nv);
            coff2_ = dalloc<uint64_t>(nv);
            for (auto &h : hotw_) h = dalloc<uint8_t>(nv);
            ckeys_ = dalloc<unsigned long long>(nv);
            ckeys2_ = dalloc<unsigned long long>(nv);
            cpos2_ = dalloc<unsigned long long>((size_t)nv + 1);
        }
        leaf_cnt_ = dalloc<unsigned long long>(nleaves_ + 1);
        leaf_pre_ = dalloc<unsigned long long>(nleaves_ + 1);
        leaf_dirty_ = dalloc<uint8_t>(nleaves_);
        chosen_ = dalloc<uint8_t>(2 * nleaves_);
        maximal_ = dalloc<uint8_t>(2 * nleaves_);
        q_[0] = dalloc<uint32_t>(nv);
        q_[1] = dalloc<uint32_t>(nv);
        // control block: counters, round log and stats, read back by one copy
        cnt_ = (uint32_t *)dalloc<unsigned long long>(kCtlBlockBytes / 8);
        hcnt_ = cnt_ + 3;
        rlog_ = cnt_ + 8;
        stats_ = (unsigned long long *)(cnt_ + 8 + 2 * kRoundLog);
        ecnt_ = stats_ + 16;
        stamp_ = dalloc<uint32_t>(nv);
        inval_ = dalloc<uint32_t>(((size_t)nv + 31) / 32);
        cshift_ = coarse_shift_for(nv);
        coarse_ = dalloc<uint32_t>(coarse_words(nv, cshift_));
        mark_ = dalloc<uint32_t>(((size_t)nv + 31) / 32);
        tidx_ = dalloc<uint32_t>(nv);
        br_.ins_b = dalloc<uint32_t>(nv);
        br_.ins_e = dalloc<uint32_t>(nv);
        br_.del_b = dalloc<uint32_t>(nv);
        br_.del_e = dalloc<uint32_t>(nv);
        dpre_ = dalloc<unsigned long long>((size_t)nv + 1);
        touched_ = dalloc<uint32_t>(nv);
        newdeg_ = dalloc<uint32_t>(nv);
        ctl_ = dalloc<Ctl>(1);
        ensure_heavy(ne);

        uint64_t *d_row = dalloc<uint64_t>((size_t)nv + 1);
        SG_HIP(hipMemcpyAsync(d_row, row, ((size_t)nv + 1) * sizeof(uint64_t), hipMemcpyHostToDevice, stream_));
        hipLaunchKernelGGL(k_layout_offsets, dim3(grid_for((uint64_t)nv + 1, 256, 16384)), dim3(256), 0, stream_, nv, d_row,
                           cap_, off_, deg_);
        SG_LAUNCHED();
        if (host_edges_) {
            // the edge array is pinned host memory: lay the lists out on the host
            // (no staging copy of the CSR, no pass over the host link)
            std::vector<uint64_t> hoff((size_t)nv + 1);
            SG_HIP(hipMemcpyAsync(hoff.data(), off_, hoff.size() * 8, hipMemcpyDeviceToHost, stream_));
            sync();
            uint32_t *hc = col_;
            parallel_for(nv, [&](uint64_t a, uint64_t b) {
                for (uint64_t v = a; v < b; ++v) {
                    const uint64_t o = hoff[v], d = row[v + 1] - row[v], span = hoff[v + 1] - o;
                    if (d) std::memcpy(hc + o, col + row[v], d * 4);
                    if (span > d) std::memset(hc + o + d, 0xFF, (span - d) * 4);
                }
            });
        } else {
            ensure(tmp_, tmp_cap_, std::max<uint64_t>(ne, 1));
            if (ne) SG_HIP(hipMemcpyAsync(tmp_, col, ne * sizeof(uint32_t), hipMemcpyHostToDevice, stream_));
            fill<uint32_t>(col_, cap_, kGap);
            hipLaunchKernelGGL(k_layout_copy, dim3(grid_for(nv, kWavesPerBlock, 16384)), dim3(kBlock), 0, stream_, nv, d_row,
                               (const uint32_t *)tmp_, off_, col_);
            SG_LAUNCHED();
        }
        hipLaunchKernelGGL(k_leaf_init, dim3(grid_for(nleaves_ + 1, 256, 16384)), dim3(256), 0, stream_, nv, L_, nleaves_,
                           d_row, leaf_cnt_);
        SG_LAUNCHED();
        fill<uint8_t>(leaf_dirty_, nleaves_, 0);
        fill<uint8_t>(chosen_, 2 * nleaves_, 0);
        fill<uint8_t>(maximal_, 2 * nleaves_, 0);
        fill<uint32_t>(mark_, ((uint64_t)nv + 31) / 32, 0u);
        fill<uint32_t>(inval_, ((uint64_t)nv + 31) / 32, 0u);
        fill<uint32_t>(coarse_, coarse_words(nv, cshift_), 0u);
        fill<uint32_t>(br_.ins_b, nv, 0u);
        fill<uint32_t>(br_.ins_e, nv, 0u);
        fill<uint32_t>(br_.del_b, nv, 0u);
        fill<uint32_t>(br_.del_e, nv, 0u);
        fill<uint32_t>(stamp_, nv, 0);
        fill<uint32_t>(cnt_, kCtlBlockBytes / 4, 0u);
        fill<uint8_t>((uint8_t *)ctl_, sizeof(Ctl), 0);
        if (cache_slots_) {
            fill<uint64_t>(coff_, nv, ~0ull);
            for (auto &h : hotw_) fill<uint8_t>(h, nv, 0);
        }
        sync();
        dfree(d_row);
        seg_dirty_ = true;
        loaded_ = true;
        st_.nv = nv_;
        st_.ne = ne_;
        st_.capacity = cap_;
        st_.host_edges = host_edges_ ? 1 : 0;
        st_.cache_slots = cache_slots_;
        after_load(row, col);
    }
    // Stores an engine keeps beside its PMA (PageRank's in-edge store) are
    // built here from the loaded CSR.
    virtual void after_load(const uint64_t *, const uint32_t *) {}
    // Device and pinned host memory of those stores.
    virtual uint64_t aux_device_bytes() const { return 0; }
    virtual uint64_t aux_host_bytes() const { return 0; }
    virtual uint64_t in_capacity() const { return 0; }
    virtual void in_mirror(uint64_t &slots, uint64_t &live) { slots = live = 0; }

    void ensure_heavy(uint64_t edges) {
        // every list past a wave's budget becomes >= 1 descriptor: at most one per
        // vertex plus one per kChunk edges, per pass
        const size_t need = (size_t)(edges / kChunk + nv_ + 4096);
        if (need > heavy_cap_) {
            for (int i = 0; i < 3; ++i) {
                dfree(heavy_[i]);
                heavy_[i] = nullptr;
            }
            for (int i = 0; i < 3; ++i) {
                heavy_[i] = dalloc<HeavyDesc>(need);
                // zeroed: a slot never written holds an empty (start 0, len 1) piece
                fill<uint32_t>((uint32_t *)heavy_[i], need * (sizeof(HeavyDesc) / 4), 0u);
            }
            heavy_cap_ = need;
        }
    }

    // HBM copy of the leading live edges of a host-resident edge array (see
    // GraphStore): the tile sweeps read those edges from HBM.
    uint32_t *mirror_ = nullptr;
    uint64_t mirror_live_ = 0;
    uint32_t *tail_ = nullptr;  // gap-free host copy of the live edges past the mirror (GraphStore)
    uint32_t mbits_ = 32;       // bits per id in the mirror and the tail
    GraphView gview() const {
        GraphView g{nv_, off_, deg_, col_, nullptr, nullptr, cap_, nullptr, mirror_, std::min(mirror_live_, ne_), tail_, mbits_};
        if (cache_slots_) {
            g.ccol = ccol_[ccur_];
            g.coff = coff_;
            g.hot = hotw_[hot_cur_];
        }
        return g;
    }
    uint32_t budget() const { return cfg_.wave_budget ? cfg_.wave_budget : kWaveBudget; }
    // round kernels: 1024 blocks (4096 waves); small frontiers spread over
    // them, and fewer blocks than kGridList shorten each round's launch and drain
    static constexpr uint32_t kRoundBlocks = 1024;
    uint32_t round_blocks() const { return cfg_.round_blocks ? std::min<uint32_t>(cfg_.round_blocks, 65535) : kRoundBlocks; }

    // dpre_[v] = sum of deg[0..v) for v in [0, nv] (64-bit) and the merge-path
    // tiles of the edge-centric sweep, in three kernels.
    SweepTiles sweep_tiles() {
        const uint32_t nb = (uint32_t)(((uint64_t)nv_ + kScanSpan - 1) / kScanSpan);
        ensure(bsum_, bsum_cap_, (size_t)nb + 1);
        ensure(tile_v_, tile_v_cap_, ((uint64_t)nv_ + ne_) / kSweepTile + 2);
        ensure(tile_e_, tile_e_cap_, tile_v_cap_);
        SweepTiles t{tile_v_, tile_e_, std::min<uint64_t>(tile_v_cap_, tile_e_cap_)};
        hipLaunchKernelGGL(k_deg_block_sums, dim3(nb), dim3(kScanBlock), 0, stream_, nv_, deg_, bsum_);
        SG_LAUNCHED();
        hipLaunchKernelGGL(k_deg_scan_sums, dim3(1), dim3(1024), 0, stream_, nb, bsum_);
        SG_LAUNCHED();
        hipLaunchKernelGGL(k_deg_scan_write, dim3(nb), dim3(kScanBlock), 0, stream_, nv_, deg_, bsum_, dpre_);
        SG_LAUNCHED();
        const uint64_t ntiles = ((uint64_t)nv_ + ne_ + kSweepTile - 1) / kSweepTile;
        hipLaunchKernelGGL(k_sweep_tiles, dim3(grid_for(ntiles + 1, 256, 16384)), dim3(256), 0, stream_, nv_,
                           (const unsigned long long *)dpre_, t, stats_ + 3);
        SG_LAUNCHED();
        return t;
    }
    // Blocks of a sweep kernel for ntiles tiles: one wave per tile, at most as
    // many blocks as fit on the device at once (the kernels stride over tiles,
    // so every block stays resident and loads its LDS filter once).
    template <class K>
    uint32_t sweep_grid(K kernel, uint64_t ntiles) {
        const void *key = reinterpret_cast<const void *>(kernel);
        auto it = occ_.find(key);
        int per_cu = 0;
        if (it == occ_.end()) {
            SG_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kSweepBlock, 0));
            per_cu = std::max(per_cu, 1);
            occ_[key] = per_cu;
        } else {
            per_cu = it->second;
        }
        const uint64_t resident = (uint64_t)per_cu * (uint64_t)std::max(num_cus_, 1);
        return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((ntiles + kSweepWaves - 1) / kSweepWaves, resident));
    }
    // An engine that walks the old out-lists between prepare and commit
    // (PageRank's cancellation) keeps them intact: no chunk is compacted in
    // place during prepare.
    virtual bool reads_old_lists() const { return false; }
    uint32_t heavy_deg() const { return std::max<uint32_t>(kChunk, cfg_.heavy_degree ? cfg_.heavy_degree : kHeavyDeg); }

    // ---- passes and rounds ---------------------------------------------------------
    // Round R reads frontier q[R & 1] (count cnt[R % 3]) and the descriptors
    // spilled by round R - 1 (hq[(R + 2) % 3]); it appends to q[(R + 1) & 1]
    // and spills into hq[R % 3]; it clears cnt[(R + 2) % 3] and hc[(R + 1) % 3],
    // whose readers have finished. Seed passes before round R append to
    // round R's frontier and spill into hq[(R + 2) % 3].
    void frontier_reset() {  // word 7 stays zero: an empty list count
        fill<uint32_t>(cnt_, 8, 0);
        fill<unsigned long long>(ecnt_, 4, 0ull);
    }

    PassIO base_io() const {
        PassIO io{};
        io.in_cap = nv_;
        io.out_cap = nv_;
        io.heavy_cap = (uint32_t)std::min<size_t>(heavy_cap_, 0xFFFFFFFFu);
        io.budget = budget();
        io.heavy_deg = heavy_deg();
        io.stats = stats_;
        return io;
    }
    PassIO round_io(uint32_t R) const {
        PassIO io = base_io();
        io.in_list = q_[R & 1];
        io.in_cnt = cnt_ + (R % 3);
        io.out_list = q_[(R + 1) & 1];
        io.out_cnt = cnt_ + ((R + 1) % 3);
        io.zero_cnt = cnt_ + ((R + 2) % 3);
        io.heavy = heavy_[R % 3];
        io.heavy_cnt = hcnt_ + (R % 3);
        io.hin = heavy_[(R + 2) % 3];
        io.hin_cnt = hcnt_ + ((R + 2) % 3);
        io.heavy_zero = hcnt_ + ((R + 1) % 3);
        return io;
    }
    // A seed pass over a list (in_list, in_cnt) or the range [0, range_n).
    PassIO seed_io(uint32_t R, const uint32_t *in_list, const uint32_t *in_cnt, uint32_t range_n) const {
        PassIO io = base_io();
        io.in_list = in_list;
        io.in_cnt = in_cnt;
        io.range_n = range_n;
        io.out_list = q_[R & 1];
        io.out_cnt = cnt_ + (R % 3);
        io.heavy = heavy_[(R + 2) % 3];
        io.heavy_cnt = hcnt_ + ((R + 2) % 3);
        return io;
    }

    // HIP events around one launch (or a graph of nlaunch) on the engine's stream.
    void begin_timed(bool timed, uint32_t cls = SG_KCLASS_RELAX, uint32_t nlaunch = 1) {
        if (!timed) return;
        if (ev_used_ == events_.size()) {
            hipEvent_t a, b;
            SG_HIP(hipEventCreate(&a));
            SG_HIP(hipEventCreate(&b));
            events_.push_back({a, b});
            ev_cls_.push_back(0);
            ev_n_.push_back(1);
        }
        ev_cls_[ev_used_] = cls;
        ev_n_[ev_used_] = nlaunch;
        SG_HIP(hipEventRecord(events_[ev_used_].first, stream_));
    }
    void end_timed(bool timed) {
        if (timed) SG_HIP(hipEventRecord(events_[ev_used_++].second, stream_));
    }
    // Kernel class cls is timed: SG_FLAG_KERNEL_TIMING, narrowed to the sweep by SG_FLAG_TIME_SWEEP.
    bool timing(uint32_t cls) const {
        return (cfg_.flags & SG_FLAG_KERNEL_TIMING) && (!(cfg_.flags & SG_FLAG_TIME_SWEEP) || cls == SG_KCLASS_SWEEP);
    }

    // ---- round batches as HIP graphs (SG_FLAG_GRAPH_ROUNDS) ----------------------
    // A batch of k round launches becomes a chain of k kernel nodes. The graph
    // of a (kernel, k) pair is built and instantiated once; later batches only
    // rewrite the nodes' arguments (round-dependent queues, counters, stamps)
    // and launch it: one submission per batch instead of k.
    struct GraphLaunch {
        const void *func;
        dim3 grid, block;
        std::shared_ptr<void> args;  // tuple of the argument values
        std::vector<void *> ptrs;    // kernelParams: pointers into args
    };
    struct RoundGraph {
        hipGraph_t g = nullptr;
        hipGraphExec_t exec = nullptr;
        std::vector<const void *> funcs;  // kernel of each node
        std::vector<hipGraphNode_t> nodes;
        std::vector<std::shared_ptr<void>> keep;  // argument values of the last launch, kept past it
    };
    std::vector<GraphLaunch> *batch_ = nullptr;  // set while drive_rounds collects a batch
    std::map<std::pair<const void *, size_t>, RoundGraph> graphs_;

    template <class... Args, size_t... I>
    static std::vector<void *> tuple_ptrs(std::tuple<Args...> &t, std::index_sequence<I...>) {
        return {static_cast<void *>(&std::get<I>(t))...};
    }
    // Launch now, or add to the batch being collected.
    template <class... Args>
    void round_launch(void (*kernel)(Args...), dim3 grid, dim3 block, Args... args) {
        if (!batch_) {
            hipLaunchKernelGGL(kernel, grid, block, 0, stream_, args...);
            SG_LAUNCHED();
            return;
        }
        auto t = std::make_shared<std::tuple<Args...>>(args...);
        GraphLaunch l;
        l.func = reinterpret_cast<const void *>(kernel);
        l.grid = grid;
        l.block = block;
        l.ptrs = tuple_ptrs(*t, std::index_sequence_for<Args...>{});
        l.args = t;
        batch_->push_back(std::move(l));
    }
    void run_graph_batch(std::vector<GraphLaunch> &b) {
        if (b.empty()) return;
        auto key = std::make_pair(b[0].func, b.size());
        RoundGraph &rg = graphs_[key];
        bool same = rg.exec != nullptr;
        for (size_t i = 0; same && i < b.size(); ++i) same = rg.funcs[i] == b[i].func;
        if (rg.exec && !same) {  // another kernel sequence of the same length: rebuild
            SG_HIP(hipStreamSynchronize(stream_));
            (void)hipGraphExecDestroy(rg.exec);
            (void)hipGraphDestroy(rg.g);
            rg = RoundGraph{};
        }
        if (!rg.exec) {
            rg.funcs.clear();
            for (auto &l : b) rg.funcs.push_back(l.func);
        }
        auto params = [&](GraphLaunch &l) {
            hipKernelNodeParams kp{};
            kp.func = const_cast<void *>(l.func);
            kp.gridDim = l.grid;
            kp.blockDim = l.block;
            kp.sharedMemBytes = 0;
            kp.kernelParams = l.ptrs.data();
            kp.extra = nullptr;
            return kp;
        };
        if (!rg.exec) {
            SG_HIP(hipGraphCreate(&rg.g, 0));
            rg.nodes.resize(b.size());
            for (size_t i = 0; i < b.size(); ++i) {
                const hipKernelNodeParams kp = params(b[i]);
                SG_HIP(hipGraphAddKernelNode(&rg.nodes[i], rg.g, i ? &rg.nodes[i - 1] : nullptr, i ? 1 : 0, &kp));
            }
            SG_HIP(hipGraphInstantiate(&rg.exec, rg.g, nullptr, nullptr, 0));
        } else {
            for (size_t i = 0; i < b.size(); ++i) {
                const hipKernelNodeParams kp = params(b[i]);
                SG_HIP(hipGraphExecKernelNodeSetParams(rg.exec, rg.nodes[i], &kp));
            }
        }
        SG_HIP(hipGraphLaunch(rg.exec, stream_));
        rg.keep.clear();  // the previous launch of this graph ended at the wait after its batch
        for (auto &l : b) rg.keep.push_back(l.args);
    }
    void release_graphs() {
        for (auto &kv : graphs_) {
            if (kv.second.exec) (void)hipGraphExecDestroy(kv.second.exec);
            if (kv.second.g) (void)hipGraphDestroy(kv.second.g);
        }
        graphs_.clear();
    }

    template <class VFn, class EFn>
    void launch_pass(const PassIO &io, VFn vfn, EFn efn, bool timed) {
        launch_pass_view(gview(), io, vfn, efn, timed);
    }
    // The same over another store's lists (the in-edge store of a pull).
    template <class VFn, class EFn>
    void launch_pass_view(const GraphView &g, const PassIO &io, VFn vfn, EFn efn, bool timed) {
        const uint32_t g1 = io.in_list ? round_blocks() : grid_for(io.range_n, kWave * kWavesPerBlock, round_blocks());
        begin_timed(timed && !batch_);
        round_launch(k_round<VFn, EFn>, dim3(g1), dim3(kBlock), g, io, vfn, efn);
        end_timed(timed && !batch_);
    }

    // Walk, with efn, the descriptors that a seed pass before round R spilled
    // (hq[(R + 2) % 3]), then empty that queue. For a pass whose chunks must not
    // outlive the slot layout they index (PR cancel, before the PMA commit), or
    // whose chunks index another store (g: a pull over the in-edge store).
    template <class EFn>
    void drain_seed_spill(uint32_t R, EFn efn) { drain_seed_spill_view(gview(), R, efn); }
    template <class EFn>
    void drain_seed_spill_view(const GraphView &g, uint32_t R, EFn efn) {
        PassIO io = base_io();
        io.in_list = q_[R & 1];
        io.in_cnt = cnt_ + 7;  // zero: no light items
        io.out_list = q_[R & 1];
        io.out_cnt = cnt_ + (R % 3);
        io.hin = heavy_[(R + 2) % 3];
        io.hin_cnt = hcnt_ + ((R + 2) % 3);
        launch_pass_view(g, io, VAll{}, efn, false);
        fill<uint32_t>(hcnt_ + ((R + 2) % 3), 1, 0u);
    }

    // Harvest kernel timings of completed events (call after a sync).
    void harvest_events() {
        for (size_t i = 0; i < ev_used_; ++i) {
            float ms = 0.f;
            SG_HIP(hipEventElapsedTime(&ms, events_[i].first, events_[i].second));
            const uint32_t c = ev_cls_[i];
            kt_ms_[c] += ms;
            kt_launches_[c] += ev_n_[i];
            if (c == SG_KCLASS_RELAX) {
                st_.relax_ms += ms;
                st_.relax_launches += ev_n_[i];
            }
        }
        ev_used_ = 0;
    }

    // Copy the control block to the host and wait.
    void read_block() {
        SG_HIP(hipMemcpyAsync(h_blk_, cnt_, kCtlBlockBytes, hipMemcpyDeviceToHost, stream_));
        sync();
    }

    // Run rounds from the frontier of round R_ until a round activates nothing
    // and spills nothing. Rounds are enqueued in batches (rounds_per_sync, then
    // doubling up to 64); each round logs what the previous one produced, so a
    // batch costs one copy of the control block and one wait. Rounds past the
    // last busy one find empty inputs and return at once. make_fns(stamp_for_next)
    // returns {vfn, efn}. Timed runs count into stats [8..11]. Returns the rounds run.
    template <class MakeFns>
    uint32_t run_rounds(MakeFns make_fns, bool timed, int phase = -1) {
        const bool do_time = timed && timing(SG_KCLASS_RELAX);
        return drive_rounds([&](const PassIO &io) {
            auto fns = make_fns(io.round + 1);
            launch_pass(io, fns.first, fns.second, do_time);
        }, timed, phase);
    }

    // launch(io) enqueues round io.round (timed with begin_timed / end_timed
    // when the run is timed and SG_FLAG_KERNEL_TIMING is set). phase (kPhase*,
    // or -1) names a per-snapshot phase: with rounds_per_sync = 0 its first
    // batch is the rounds it took last time plus one, so a phase whose round
    // count is steady costs one wait and at most one empty round; a phase that
    // runs longer continues with batches of 4, 8, ...
    // limit > 0: stop after about `limit` rounds (whole batches) even if work
    // remains (stopped_ tells which); the frontier and spills stay for the next call.
    bool stopped_ = false;
    template <class Launch>
    uint32_t drive_rounds(Launch launch, bool timed, int phase = -1, uint32_t limit = 0) {
        stopped_ = false;
        uint32_t k0 = cfg_.rounds_per_sync ? cfg_.rounds_per_sync : 4;
        const bool adaptive = !cfg_.rounds_per_sync && phase >= 0 && last_rounds_[phase];
        if (adaptive) k0 = last_rounds_[phase] + 1;
        uint32_t k = std::max<uint32_t>(1, std::min<uint32_t>(k0, kRoundLog));
        const uint32_t maxr = cfg_.max_rounds ? cfg_.max_rounds : 1000000u;
        const bool do_time = timed && timing(SG_KCLASS_RELAX);
        if (timed) fill<unsigned long long>(stats_ + 8, 8, 0ull);
        uint32_t rounds = 0;
        bool first = true;
        last_appended_ = 0;
        const bool graphs = (cfg_.flags & SG_FLAG_GRAPH_ROUNDS) != 0;
        std::vector<GraphLaunch> batch;
        for (;;) {
            const uint32_t R0 = R_;
            if (limit) k = std::max<uint32_t>(1, std::min<uint32_t>(k, limit - std::min(limit - 1, rounds)));
            if (graphs) {
                batch.clear();
                batch_ = &batch;
            }
            for (uint32_t i = 0; i < k; ++i) {
                PassIO io = round_io(R_);
                io.log = rlog_;
                io.round = R_;
                if (timed) io.stats = stats_ + 8;
                launch(io);
                ++R_;
            }
            if (graphs) {
                batch_ = nullptr;
                begin_timed(do_time, SG_KCLASS_RELAX, k);
                run_graph_batch(batch);
                end_timed(do_time);
            }
            read_block();
            if (do_time) harvest_events();
            if (first) {  // what the seed passes put on the first frontier (logged by the first round)
                last_appended_ += h_blk_[8 + 2 * ((R0 + kRoundLog - 1) % kRoundLog)];
                first = false;
            }
            for (uint32_t i = 0; i < k; ++i) {
                ++rounds;
                const uint32_t R = R0 + i;
                uint32_t front, spill;
                if (i + 1 < k) {
                    front = h_blk_[8 + 2 * (R % kRoundLog)];
                    spill = h_blk_[8 + 2 * (R % kRoundLog) + 1];
                } else {
                    front = h_blk_[(R + 1) % 3];
                    spill = h_blk_[3 + R % 3];
                }
                last_appended_ += front;
                if (front == 0 && spill == 0) {
                    finish_rounds(timed);
                    if (phase >= 0) last_rounds_[phase] = rounds;
                    return rounds;
                }
            }
            if (limit && rounds >= limit) {
                finish_rounds(timed);
                stopped_ = true;
                return rounds;
            }
            if (rounds >= maxr) {
                finish_rounds(timed);
                fail(SG_EDEVICE, "no convergence within " + std::to_string(maxr) + " rounds");
            }
            // past the expected count the remaining rounds are few: restart small
            k = adaptive && rounds == k ? 4u : std::min<uint32_t>(2 * k, 64);
        }
    }

    // Uses the control block read by the last batch of rounds.
    void finish_rounds(bool timed) {
        const unsigned long long *s = h_stats();
        if (s[3] | s[11]) fail(SG_EDEVICE, "internal queue overflow (flags " + std::to_string(s[3] | s[11]) + ")");
        if (timed) {
            timed_edges_ += s[8] + s[10];
            timed_verts_ += s[9];
            if (timing(SG_KCLASS_RELAX)) {
                st_.relax_bytes += 12ull * (s[8] + s[10]) + 32ull * s[9];
                kt_bytes_[SG_KCLASS_RELAX] += 12ull * (s[8] + s[10]) + 32ull * s[9];
            }
        }
        if (timed) timed_cedges_ += s[13];
        if (timed) timed_pulls_ += s[12];
        st_.pull_rounds = s[4] + timed_pulls_;
        st_.edges_relaxed = s[0] + s[2] + timed_edges_;
        st_.vertices_processed = s[1] + timed_verts_;
        st_.edges_from_cache = s[5] + timed_cedges_;
    }

    // ---- batch prepare / commit -------------------------------------------------------
    void stage_host_batch(const uint32_t *ins, uint64_t nins, const uint32_t *del, uint64_t ndel,
                          const uint32_t *&d_ins, const uint32_t *&d_del) {
        for (uint64_t i = 0; i < 2 * nins; ++i)
            if (ins[i] >= nv_) fail(SG_EINVAL, "insert pair " + std::to_string(i / 2) + " has an id >= nv");
        for (uint64_t i = 0; i < 2 * ndel; ++i)
            if (del[i] >= nv_) fail(SG_EINVAL, "delete pair " + std::to_string(i / 2) + " has an id >= nv");
        ensure(raw_ins_, raw_ins_cap_, 2 * nins + 2);
        ensure(raw_del_, raw_del_cap_, 2 * ndel + 2);
        if (nins) SG_HIP(hipMemcpyAsync(raw_ins_, ins, 2 * nins * sizeof(uint32_t), hipMemcpyHostToDevice, stream_));
        if (ndel) SG_HIP(hipMemcpyAsync(raw_del_, del, 2 * ndel * sizeof(uint32_t), hipMemcpyHostToDevice, stream_));
        d_ins = raw_ins_;
        d_del = raw_del_;
    }

    // Sort the batch, list touched sources, cut their lists into chunks, count
    // survivors and new degrees. Leaves the graph unchanged.
    bool prepare(const uint32_t *d_ins, uint64_t nins64, const uint32_t *d_del, uint64_t ndel64) {
        if (nins64 > 0x7FFFFFFFull || ndel64 > 0x7FFFFFFFull) fail(SG_EINVAL, "batch too large (max 2^31-1 pairs)");
        const uint32_t nins = (uint32_t)nins64, ndel = (uint32_t)ndel64;
        nins_ = nins;
        ndel_ = ndel;
        ensure(ins_k_, ins_cap_, nins + 1);
        ensure(ins_v_, ins_cap2_, nins + 1);
        ensure(ins_k2_, ins_cap3_, nins + 1);
        ensure(ins_v2_, ins_cap4_, nins + 1);
        ensure(del_k_, del_cap_, ndel + 1);
        ensure(del_k2_, del_cap2_, ndel + 1);
        ensure_heavy(ne_ + nins);
        // chunks: sum over touched vertices of ceil(deg/CH) (with deletes) + ceil(ins/CH)
        const uint64_t chunk_bound = (ne_ + nins) / kMergeChunk + (uint64_t)nins + ndel + 64;
        if (chunk_bound > 0xFFFFFFF0ull) fail(SG_EINVAL, "batch too large for the merge chunk list");
        chunk_bound_ = chunk_bound;
        ensure(tinfo_, tinfo_cap_, (size_t)nins + ndel + 1);
        // filtered sequences are staged at their own ranges: at most every live edge plus the inserts
        // (with the edges in host memory the staging is sized to the batch instead, below)
        if (!host_edges_) ensure(tmp_, tmp_cap_, std::max<uint64_t>(ne_ + nins, 1));
        ensure(chunks_, chunks_cap_, chunk_bound);
        ensure(ccount_, ccount_cap_, chunk_bound + 1);
        ensure(cpos_, cpos_cap_, chunk_bound + 1);
        fill<uint8_t>((uint8_t *)ctl_, sizeof(Ctl), 0);
        fill<unsigned long long>(ccount_, chunk_bound + 1, 0ull);
        if (nins + ndel) {
            hipLaunchKernelGGL(k_split, dim3(grid_for((uint64_t)nins + ndel, 256, 8192)), dim3(256), 0, stream_, nv_, d_ins, nins,
                               d_del, ndel, ins_k_, ins_v_, del_k_, ctl_, swap_pairs_ ? 1 : 0);
            SG_LAUNCHED();
        }
        // sort only the bits ids can hold: kb = bit length of nv, so a masked
        // kGap / ~0 (dropped pair) still compares above every valid key
        const int kb = nv_ ? 32 - __builtin_clz(nv_) : 1;
        size_t b1 = 0, b2 = 0, b3 = 0;
        SG_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, b1, ins_k_, ins_k2_, ins_v_, ins_v2_, (int)nins, 0, kb, stream_));
        SG_HIP(hipcub::DeviceRadixSort::SortKeys(nullptr, b2, del_k_, del_k2_, (int)ndel, 0, 32 + kb, stream_));
        SG_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, b3, ccount_, cpos_, (int)(chunk_bound + 1), stream_));
        ensure(cub_tmp_, cub_cap_, std::max({b1, b2, b3}) + 256);
        size_t b = cub_cap_;
        if (nins) SG_HIP(hipcub::DeviceRadixSort::SortPairs(cub_tmp_, b, ins_k_, ins_k2_, ins_v_, ins_v2_, (int)nins, 0, kb, stream_));
        b = cub_cap_;
        if (ndel) SG_HIP(hipcub::DeviceRadixSort::SortKeys(cub_tmp_, b, del_k_, del_k2_, (int)ndel, 0, 32 + kb, stream_));
        bv_.ins_key = ins_k2_;
        bv_.ins_val = ins_v2_;
        bv_.nins = nins;
        bv_.del_key = (const uint64_t *)del_k2_;
        bv_.ndel = ndel;
        if (nins + ndel) {
            hipLaunchKernelGGL(k_touched, dim3(grid_for((uint64_t)nins + ndel, kBlock, 8192)), dim3(kBlock), 0, stream_, bv_,
                               br_, mark_, touched_, &ctl_->tcount, nv_, &ctl_->pad[0]);
            SG_LAUNCHED();
            const uint32_t gt = grid_for((uint64_t)nins + ndel, kBlock, 8192);
            hipLaunchKernelGGL(k_touch_info, dim3(gt), dim3(kBlock), 0, stream_, gview(), br_, touched_, &ctl_->tcount, tinfo_,
                               tidx_, chunks_, &ctl_->nchunks, (uint32_t)chunk_bound, &ctl_->seqcur,
                               reads_old_lists() ? 0 : 1, &ctl_->pad[1]);
            SG_LAUNCHED();
            if (host_edges_) {  // HBM holds only what this batch stages, not a copy of the edge array
                SG_HIP(hipMemcpyAsync(h_ctl_, ctl_, sizeof(Ctl), hipMemcpyDeviceToHost, stream_));
                sync();
                ensure(tmp_, tmp_cap_, std::max<uint64_t>(h_ctl_->seqcur, 1));
            }
            begin_timed(timing(SG_KCLASS_MERGE), SG_KCLASS_MERGE);
            hipLaunchKernelGGL(k_chunk_filter, dim3(kGridFilter), dim3(kBlock), 0, stream_, gview(), bv_,
                               (const MChunk *)chunks_, &ctl_->nchunks, (uint32_t)chunk_bound, ccount_, tmp_, col_);
            SG_LAUNCHED();
            end_timed(timing(SG_KCLASS_MERGE));
        }
        b = cub_cap_;
        SG_HIP(hipcub::DeviceScan::ExclusiveSum(cub_tmp_, b, ccount_, cpos_, (int)(chunk_bound + 1), stream_));
        if (nins + ndel) {
            hipLaunchKernelGGL(k_newdeg, dim3(grid_for((uint64_t)nins + ndel, kBlock, 8192)), dim3(kBlock), 0, stream_, gview(),
                               tinfo_, &ctl_->tcount, cpos_, newdeg_, leaf_dirty_, L_, leaf_cnt_, ctl_);
            SG_LAUNCHED();
        }
        SG_HIP(hipMemcpyAsync(h_ctl_, ctl_, sizeof(Ctl), hipMemcpyDeviceToHost, stream_));
        sync();
        if (h_ctl_->pad[0] || h_ctl_->pad[1]) fail(SG_EDEVICE, "internal touched/chunk list overflow");
        st_.dropped_updates += h_ctl_->dropped;
        return h_ctl_->overflow != 0;
    }

    void commit(bool overflow) {
        const Ctl c = *h_ctl_;
        const uint64_t ne_new = ne_ + (uint64_t)(long long)c.ne_delta;
        if (nins_ + ndel_) {
            if (overflow) rebalance(ne_new);  // moves the touched vertices of its windows
            begin_timed(timing(SG_KCLASS_MERGE), SG_KCLASS_MERGE);
            hipLaunchKernelGGL(k_chunk_place, dim3(kGridList), dim3(kBlock), 0, stream_, gview(), bv_, col_, tinfo_, chunks_,
                               &ctl_->nchunks, (uint32_t)chunk_bound_, cpos_, tmp_, newdeg_, maximal_, nleaves_, L_,
                               overflow ? 1 : 0);
            SG_LAUNCHED();
            end_timed(timing(SG_KCLASS_MERGE));
            hipLaunchKernelGGL(k_lane_place, dim3(grid_for(c.tcount, 256, 4096)), dim3(256), 0, stream_, gview(), bv_, col_,
                               tinfo_, &ctl_->tcount, newdeg_, deg_, maximal_, nleaves_, L_, overflow ? 1 : 0);
            SG_LAUNCHED();
            if (cache_slots_) {
                hipLaunchKernelGGL(k_cache_drop, dim3(1024), dim3(256), 0, stream_, tinfo_, &ctl_->tcount, coff_);
                SG_LAUNCHED();
            }
            if (overflow) fill<uint8_t>(maximal_, 2 * nleaves_, 0);
        }
        ne_ = ne_new;
        st_.ne = ne_;
        st_.capacity = cap_;
    }

    // The batch's per-vertex ranges and marks are kept until the snapshot's
    // recompute has used them; then they are cleared for the next batch.
    void finish_batch() {
        if (ev_used_) {  // timed launches the last round batch did not collect
            sync();
            harvest_events();
        }
        cache_rebuild();
        if (nins_ + ndel_ == 0) return;
        hipLaunchKernelGGL(k_clear_ranges, dim3(1024), dim3(256), 0, stream_, tinfo_, &ctl_->tcount, br_, mark_);
        SG_LAUNCHED();
    }

    void rebalance(uint64_t ne_new) {
        // prefix of per-leaf counts
        size_t b = 0;
        SG_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, b, leaf_cnt_, leaf_pre_, (int)(nleaves_ + 1), stream_));
        ensure(cub_tmp_, cub_cap_, b + 256);
        b = cub_cap_;
        SG_HIP(hipcub::DeviceScan::ExclusiveSum(cub_tmp_, b, leaf_cnt_, leaf_pre_, (int)(nleaves_ + 1), stream_));
        hipLaunchKernelGGL(k_choose_windows, dim3(grid_for(nleaves_, 256, 8192)), dim3(256), 0, stream_, nv_, L_, nleaves_, H_,
                           off_, cap_, leaf_pre_, leaf_dirty_, chosen_, ctl_);
        SG_LAUNCHED();
        SG_HIP(hipMemcpyAsync(h_ctl_, ctl_, sizeof(Ctl), hipMemcpyDeviceToHost, stream_));
        SG_HIP(hipMemcpyAsync(h_u64_, leaf_pre_ + nleaves_, sizeof(unsigned long long), hipMemcpyDeviceToHost, stream_));
        sync();
        const bool resize = h_ctl_->resize != 0;
        const uint64_t total = h_u64_[0];
        uint64_t new_cap = 0;
        uint32_t nwin = 0;
        ensure(win_, win_cap_, std::max<size_t>(nleaves_, 1));
        if (resize) {
            new_cap = cap_;
            while ((double)total >= up_threshold(H_, H_) * (double)new_cap) new_cap *= 2;
            fill<uint8_t>(chosen_, 2 * nleaves_, 0);
            Window root{};
            root.vlo = 0;
            root.vhi = nv_;
            SG_HIP(hipMemcpyAsync(win_, &root, sizeof(Window), hipMemcpyHostToDevice, stream_));
            fill<uint8_t>(maximal_ + 1, 1, 1);
            nwin = 1;
            st_.resizes += 1;
        } else {
            hipLaunchKernelGGL(k_prune, dim3(grid_for(2 * nleaves_, 256, 8192)), dim3(256), 0, stream_, nv_, L_, nleaves_,
                               chosen_, maximal_, win_, ctl_, (uint32_t)win_cap_);
            SG_LAUNCHED();
            fill<uint8_t>(chosen_, 2 * nleaves_, 0);
            SG_HIP(hipMemcpyAsync(h_ctl_, ctl_, sizeof(Ctl), hipMemcpyDeviceToHost, stream_));
            sync();
            if (h_ctl_->nwin > win_cap_) fail(SG_EDEVICE, "internal window-list overflow");
            nwin = (uint32_t)h_ctl_->nwin;
        }
        st_.rebalances += nwin;
        // window vertex offsets
        ensure(wsz_, wsz_cap_, (size_t)nwin + 1);
        ensure(wvoff_, wvoff_cap_, (size_t)nwin + 1);
        hipLaunchKernelGGL(k_win_sizes, dim3(grid_for((uint64_t)nwin + 1, 256, 8192)), dim3(256), 0, stream_, win_, nwin, wsz_);
        SG_LAUNCHED();
        b = 0;
        SG_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, b, wsz_, wvoff_, (int)(nwin + 1), stream_));
        ensure(cub_tmp_, cub_cap_, b + 256);
        b = cub_cap_;
        SG_HIP(hipcub::DeviceScan::ExclusiveSum(cub_tmp_, b, wsz_, wvoff_, (int)(nwin + 1), stream_));
        SG_HIP(hipMemcpyAsync(h_u64_, wvoff_ + nwin, sizeof(unsigned long long), hipMemcpyDeviceToHost, stream_));
        sync();
        const uint64_t Wv = h_u64_[0];
        ensure(ndarr_, ndarr_cap_, Wv + 1);
        ensure(P_arr_, Parr_cap_, Wv + 1);
        hipLaunchKernelGGL(k_win_nd, dim3(grid_for(Wv + 1, 256, 16384)), dim3(256), 0, stream_, win_, nwin, wvoff_, Wv, mark_,
                           newdeg_, deg_, ndarr_);
        SG_LAUNCHED();
        b = 0;
        SG_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, b, ndarr_, P_arr_, (int)(Wv + 1), stream_));
        ensure(cub_tmp_, cub_cap_, b + 256);
        b = cub_cap_;
        SG_HIP(hipcub::DeviceScan::ExclusiveSum(cub_tmp_, b, ndarr_, P_arr_, (int)(Wv + 1), stream_));
        hipLaunchKernelGGL(k_win_params, dim3(grid_for(nwin, 256, 8192)), dim3(256), 0, stream_, win_, nwin, wvoff_, P_arr_,
                           off_, nv_, cap_, new_cap);
        SG_LAUNCHED();
        ensure(tmp2_, tmp2_cap_, std::max<uint64_t>(ne_new, 1));
        const uint32_t gw = grid_for(Wv, kWavesPerBlock, 16384);
        hipLaunchKernelGGL(k_win_gather, dim3(gw), dim3(kBlock), 0, stream_, gview(), bv_, win_, nwin, wvoff_, Wv, P_arr_, mark_,
                           tidx_, tinfo_, cpos_, (const MChunk *)chunks_, tmp_, tmp2_);
        SG_LAUNCHED();
        uint32_t *dst = col_;
        if (resize) {
            dst = host_edges_ ? halloc<uint32_t>(new_cap) : dalloc<uint32_t>(new_cap);
        }
        hipLaunchKernelGGL(k_win_scatter, dim3(gw), dim3(kBlock), 0, stream_, win_, nwin, wvoff_, Wv, P_arr_, tmp2_, dst, off_,
                           deg_);
        SG_LAUNCHED();
        if (resize) {
            hipLaunchKernelGGL(k_set_u64, dim3(1), dim3(1), 0, stream_, off_ + nv_, new_cap);
            SG_LAUNCHED();
            sync();
            if (host_edges_) hfree(col_);
            else dfree(col_);
            col_ = dst;
            cap_ = new_cap;
        }
        seg_dirty_ = true;
    }

    // ---- out-of-memory path: segments, streamed sweep, edge cache -----------------------
    void ensure_segments() {
        if (!seg_dirty_) return;
        hipLaunchKernelGGL(k_segment_cuts, dim3(1), dim3(1), 0, stream_, nv_, off_, nseg_, seg_v_, seg_aux_);
        SG_LAUNCHED();
        h_seg_v_.resize((size_t)nseg_ + 1);
        h_seg_slot_.resize((size_t)nseg_ + 1);
        SG_HIP(hipMemcpyAsync(h_seg_v_.data(), seg_v_, ((size_t)nseg_ + 1) * 4, hipMemcpyDeviceToHost, stream_));
        SG_HIP(hipMemcpyAsync(h_seg_slot_.data(), seg_aux_, ((size_t)nseg_ + 1) * 8, hipMemcpyDeviceToHost, stream_));
        sync();
        seg_dirty_ = false;
    }

    // Edge-centric sweep with the edge array in pinned host memory: segment by
    // segment, the slot range is copied to one of n_stream HBM staging buffers
    // on the copy stream while the sweep kernel walks earlier segments
    // (framework.cuh:310-384 SEGMENT, n_stream streams). launch(colp, slot_lo,
    // v_lo, v_hi, t_lo, t_hi) enqueues the sweep of one segment on stream_.
    // streamed_bytes / stream_ms count these copies and the wall time of the
    // whole streamed sweep, nothing else.
    template <class Launch>
    void streamed_sweep(Launch launch) {
        ensure_segments();
        hipLaunchKernelGGL(k_segment_paths, dim3(grid_for((uint64_t)nseg_ + 1, 256, 64)), dim3(256), 0, stream_, nseg_,
                           seg_v_, dpre_, seg_aux_ + nseg_ + 1);
        SG_LAUNCHED();
        h_seg_path_.resize((size_t)nseg_ + 1);
        SG_HIP(hipMemcpyAsync(h_seg_path_.data(), seg_aux_ + nseg_ + 1, ((size_t)nseg_ + 1) * 8, hipMemcpyDeviceToHost,
                              stream_));
        sync();
        size_t span = 0;
        for (uint32_t k = 0; k < nseg_; ++k) span = std::max<size_t>(span, h_seg_slot_[k + 1] - h_seg_slot_[k]);
        const int nb = nstage();
        for (int b = 0; b < nb; ++b) ensure(stage_[b], stage_cap_[b], std::max<size_t>(span, 1));
        auto t0 = std::chrono::steady_clock::now();
        uint64_t bytes = 0;
        uint32_t used = 0;  // segments launched (buffer of the j-th: j % nb)
        for (uint32_t k = 0; k < nseg_; ++k) {
            const uint32_t a = h_seg_v_[k], b2 = h_seg_v_[k + 1];
            if (a >= b2) continue;
            const int b = (int)(used % (uint32_t)nb);
            const uint64_t slo = h_seg_slot_[k], shi = h_seg_slot_[k + 1];
            if (used >= (uint32_t)nb) SG_HIP(hipStreamWaitEvent(cstream_, seg_used_[b], 0));  // buffer b released
            if (shi > slo)
                SG_HIP(hipMemcpyAsync(stage_[b], col_ + slo, (shi - slo) * 4, hipMemcpyHostToDevice, cstream_));
            SG_HIP(hipEventRecord(seg_copied_[b], cstream_));
            SG_HIP(hipStreamWaitEvent(stream_, seg_copied_[b], 0));
            const uint64_t tlo = h_seg_path_[k] / kSweepTile;
            const uint64_t thi = (h_seg_path_[k + 1] + kSweepTile - 1) / kSweepTile;
            launch((const uint32_t *)stage_[b], slo, a, b2, tlo, thi);
            SG_HIP(hipEventRecord(seg_used_[b], stream_));
            bytes += (shi - slo) * 4;
            ++used;
        }
        sync();
        st_.streamed_bytes += bytes;
        st_.stream_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }

    // Sweep over all live edges: resident, or streamed by segment.
    template <class Op, class EFn>
    void sweep_all(const SweepTiles &tl, Op op, const typename Op::buffer_t *buffer, EFn ef, const PassIO &io) {
        const uint64_t ntiles = ((uint64_t)nv_ + ne_ + kSweepTile - 1) / kSweepTile;
        if (!host_edges_) {
            const uint32_t grid = sweep_grid(k_sweep_edges<Op, EFn>, ntiles);
            begin_timed(timing(SG_KCLASS_SWEEP), SG_KCLASS_SWEEP);
            hipLaunchKernelGGL((k_sweep_edges<Op, EFn>), dim3(grid), dim3(kSweepBlock), 0, stream_, gview(), dpre_, tl,
                               (const uint32_t *)col_, (uint64_t)0, 0u, nv_, (uint64_t)0, ntiles, op, buffer, ef, io);
            SG_LAUNCHED();
            end_timed(timing(SG_KCLASS_SWEEP));
            // algorithmic bytes: every live target id, and off + dpre of every source
            if (timing(SG_KCLASS_SWEEP)) kt_bytes_[SG_KCLASS_SWEEP] += 4ull * ne_ + 16ull * nv_;
            return;
        }
        streamed_sweep([&](const uint32_t *colp, uint64_t slo, uint32_t va, uint32_t vb, uint64_t tlo, uint64_t thi) {
            const uint32_t grid = sweep_grid(k_sweep_edges<Op, EFn>, thi - tlo);
            hipLaunchKernelGGL((k_sweep_edges<Op, EFn>), dim3(grid), dim3(kSweepBlock), 0, stream_, gview(), dpre_, tl, colp, slo, va,
                               vb, tlo, std::min<uint64_t>(thi, ntiles), op, buffer, ef, io);
            SG_LAUNCHED();
        });
    }

    // After a solve or a snapshot: rebuild the HBM edge cache from the last
    // four snapshots' activations and open a new hotness window.
    void cache_rebuild() {
        if (!cache_slots_) return;
        sync();  // the refill's wall time starts with the stream idle
        const auto t0 = std::chrono::steady_clock::now();
        hipLaunchKernelGGL(k_cache_keys, dim3(grid_for(nv_, 256, 16384)), dim3(256), 0, stream_, nv_, deg_, hotw_[0], hotw_[1],
                           hotw_[2], hotw_[3], ckeys_);
        SG_LAUNCHED();
        size_t b = 0;
        SG_HIP(hipcub::DeviceRadixSort::SortKeys(nullptr, b, ckeys_, ckeys2_, (int)nv_, 0, 64, stream_));
        hipcub::CountingInputIterator<uint32_t> cnt(0);
        hipcub::TransformInputIterator<unsigned long long, CacheDegAt, hipcub::CountingInputIterator<uint32_t>> in(
            cnt, CacheDegAt{ckeys2_, deg_, nv_});
        size_t b2 = 0;
        SG_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, b2, in, cpos2_, (int)nv_ + 1, stream_));
        ensure(cub_tmp_, cub_cap_, std::max(b, b2) + 256);
        b = cub_cap_;
        SG_HIP(hipcub::DeviceRadixSort::SortKeys(cub_tmp_, b, ckeys_, ckeys2_, (int)nv_, 0, 64, stream_));
        b2 = cub_cap_;
        SG_HIP(hipcub::DeviceScan::ExclusiveSum(cub_tmp_, b2, in, cpos2_, (int)nv_ + 1, stream_));
        fill<uint64_t>(coff2_, nv_, ~0ull);
        fill<unsigned long long>(stats_ + 6, 2, 0ull);
        hipLaunchKernelGGL(k_cache_assign, dim3(grid_for(nv_, 256, 16384)), dim3(256), 0, stream_, nv_,
                           (const unsigned long long *)ckeys2_, (const unsigned long long *)cpos2_, deg_, cache_slots_, coff2_,
                           stats_ + 6);
        SG_LAUNCHED();
        hipLaunchKernelGGL(k_cache_fill, dim3(grid_for(nv_, 4, 16384)), dim3(256), 0, stream_, nv_,
                           (const unsigned long long *)ckeys2_, coff2_, coff_, ccol_[ccur_], col_, off_, deg_,
                           ccol_[ccur_ ^ 1], stats_ + 7);
        SG_LAUNCHED();
        std::swap(coff_, coff2_);
        ccur_ ^= 1;
        hot_cur_ = (hot_cur_ + 1) % 4;
        fill<uint8_t>(hotw_[hot_cur_], nv_, 0);
        unsigned long long h[2];
        SG_HIP(hipMemcpyAsync(h_u64_, stats_ + 6, 16, hipMemcpyDeviceToHost, stream_));
        sync();
        std::memcpy(h, h_u64_, 16);
        st_.cached_vertices = h[0];
        st_.cache_fill_bytes += 4 * h[1];  // lists fetched into the cache over the host link
        st_.cache_fill_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }

    // ---- interface ------------------------------------------------------------------------
    virtual void init(const uint32_t *sources, uint32_t nsrc) = 0;
    virtual void apply(const uint32_t *d_ins, uint64_t nins, const uint32_t *d_del, uint64_t ndel) = 0;
    virtual void gather(uint32_t q, void *out, size_t bytes) = 0;
    virtual void init_queries(const uint32_t *, const uint32_t *, uint32_t) {
        fail(SG_EINVAL, "sg_init_queries needs an engine created with SG_ALGO_MULTI");
    }
    virtual void gather_query(uint32_t, uint32_t, void *, size_t) {
        fail(SG_EINVAL, "sg_gather_query needs an engine created with SG_ALGO_MULTI");
    }
    virtual void open_slots(uint32_t) { fail(SG_EINVAL, "query slots need an engine created with SG_ALGO_MULTI"); }
    virtual void add_queries(const uint32_t *, const uint32_t *, uint32_t, uint32_t *) {
        fail(SG_EINVAL, "query slots need an engine created with SG_ALGO_MULTI");
    }
    virtual void step_queries(uint32_t, uint32_t *, uint32_t, uint32_t *, uint32_t *) {
        fail(SG_EINVAL, "query slots need an engine created with SG_ALGO_MULTI");
    }
    virtual void release_queries(const uint32_t *, uint32_t) {
        fail(SG_EINVAL, "query slots need an engine created with SG_ALGO_MULTI");
    }
    // Where the 32-bit result word of (query, qi) for vertex v lives:
    // base[v * stride + word].
    virtual void word_source(uint32_t q, uint32_t qi, const uint32_t *&base, uint64_t &stride, uint32_t &word) {
        (void)qi;
        if (q == SG_Q_DEGREE) { base = deg_; stride = 1; word = 0; return; }
        fail(SG_EINVAL, "query not available pointwise");
    }
    virtual void topk(uint32_t, uint32_t *, float *) { fail(SG_EINVAL, "sg_pr_topk needs a PageRank engine"); }
    // Bring deferred snapshots to their fixed point (SG_FLAG_DEFER_CONVERGE).
    virtual void converge() {}

    // True while snapshots wait for a deferred convergence.
    virtual bool pending() const { return false; }

    // gather_at staged through the file-scope scratch instead of buffers the
    // engine allocates: no device allocation, n <= kNoAllocGatherMax. The
    // caller holds g_scratch_call_mu[device] for the whole call.
    void gather_at_noalloc(uint32_t q, uint32_t qi, const uint32_t *verts, uint64_t n, uint32_t *out) {
        if (!inited_ && q != SG_Q_DEGREE) fail(SG_EINVAL, "no vertex state before sg_init");
        if (n > kNoAllocGatherMax) fail(SG_EINVAL, "at most " + std::to_string(kNoAllocGatherMax) + " vertices per call");
        if (pending()) fail(SG_EINVAL, "deferred snapshots: call sg_converge first");
        if (!scratch_) fail(SG_EDEVICE, "no scratch address on this device");
        if (n == 0) return;
        const uint32_t *base = nullptr;
        uint64_t stride = 1;
        uint32_t word = 0;
        word_source(q, qi, base, stride, word);
        uint32_t *ids = scratch_, *res = scratch_ + kNoAllocGatherMax;
        SG_HIP(hipMemcpyAsync(ids, verts, n * 4, hipMemcpyHostToDevice, stream_));
        hipLaunchKernelGGL(k_gather_words, dim3(grid_for(n, 256, 256)), dim3(256), 0, stream_, base, stride, word, nv_,
                           (const uint32_t *)ids, n, res);
        SG_LAUNCHED();
        SG_HIP(hipMemcpyAsync(out, res, n * 4, hipMemcpyDeviceToHost, stream_));
        sync();
    }

    uint32_t *gat_v_ = nullptr, *gat_o_ = nullptr;
    size_t gat_v_cap_ = 0, gat_o_cap_ = 0;
    void gather_at(uint32_t q, uint32_t qi, const uint32_t *verts, uint64_t n, uint32_t *out) {
        if (!inited_ && q != SG_Q_DEGREE) fail(SG_EINVAL, "no vertex state before sg_init");
        if (n == 0) return;
        const uint32_t *base = nullptr;
        uint64_t stride = 1;
        uint32_t word = 0;
        word_source(q, qi, base, stride, word);
        ensure(gat_v_, gat_v_cap_, n);
        ensure(gat_o_, gat_o_cap_, n);
        SG_HIP(hipMemcpyAsync(gat_v_, verts, n * 4, hipMemcpyHostToDevice, stream_));
        hipLaunchKernelGGL(k_gather_words, dim3(grid_for(n, 256, 4096)), dim3(256), 0, stream_, base, stride, word, nv_, gat_v_,
                           n, gat_o_);
        SG_LAUNCHED();
        SG_HIP(hipMemcpyAsync(out, gat_o_, n * 4, hipMemcpyDeviceToHost, stream_));
        sync();
    }

    void gather_graph(uint32_t q, void *out, size_t bytes) {
        if (q == SG_Q_DEGREE) {
            if (bytes < (size_t)nv_ * 4) fail(SG_EINVAL, "output too small");
            SG_HIP(hipMemcpyAsync(out, deg_, (size_t)nv_ * 4, hipMemcpyDeviceToHost, stream_));
        } else if (q == SG_Q_OFFSETS) {
            if (bytes < ((size_t)nv_ + 1) * 8) fail(SG_EINVAL, "output too small");
            SG_HIP(hipMemcpyAsync(out, off_, ((size_t)nv_ + 1) * 8, hipMemcpyDeviceToHost, stream_));
        } else if (q == SG_Q_SLOTS) {
            if (bytes < (size_t)cap_ * 4) fail(SG_EINVAL, "output too small");
            SG_HIP(hipMemcpyAsync(out, col_, (size_t)cap_ * 4, hipMemcpyDefault, stream_));
        } else if (q == SG_Q_SEGMENTS) {
            if (bytes < ((size_t)nseg_ + 1) * 4) fail(SG_EINVAL, "output too small");
            ensure_segments();
            std::memcpy(out, h_seg_v_.data(), ((size_t)nseg_ + 1) * 4);
        } else {
            fail(SG_EINVAL, "unknown query " + std::to_string(q));
        }
        sync();
    }

    // ---- checkpoint (sg_save_state / sg_restore_state) -------------------------
    // Device arrays of the vertex state (allocated when the engine has none
    // yet) and the traversal source.
    virtual void state_arrays(std::vector<std::pair<void *, size_t>> &, uint32_t &) {
        fail(SG_EINVAL, "checkpoints are for single-program engines (PR, SSSP, BFS, CC)");
    }
    // The arrays hold a restored state: take the source, make the engine ready.
    virtual void restored(uint32_t) {}

    struct StateHeader {
        char magic[8];
        uint32_t algo, nv;
        uint64_t ne;
        uint32_t source, narr, weight_mode, pad;
        uint64_t weight_seed;
        double pr_alpha, pr_epsilon;
    };
    static constexpr char kStateMagic[8] = {'S', 'G', 'S', 'T', 'A', 'T', 'E', '1'};

    // File: StateHeader, narr x (u64 bytes, bytes of data), u64 row[nv + 1],
    // u32 col[ne] (the snapshot's adjacency lists in PMA order).
    void save_state(const char *path) {
        if (!inited_) fail(SG_EINVAL, "no vertex state before sg_init");
        converge();  // nothing deferred is left out of the saved state
        std::vector<std::pair<void *, size_t>> arrs;
        uint32_t src = 0;
        state_arrays(arrs, src);
        std::vector<uint64_t> off((size_t)nv_ + 1);
        std::vector<uint32_t> deg(nv_), slots(cap_);
        gather_graph(SG_Q_OFFSETS, off.data(), off.size() * 8);
        gather_graph(SG_Q_DEGREE, deg.data(), deg.size() * 4);
        gather_graph(SG_Q_SLOTS, slots.data(), slots.size() * 4);
        std::vector<uint64_t> row((size_t)nv_ + 1, 0);
        for (uint32_t v = 0; v < nv_; ++v) row[v + 1] = row[v] + deg[v];
        if (row[nv_] != ne_) fail(SG_EDEVICE, "degree total differs from the edge count");
        std::vector<uint32_t> col(std::max<uint64_t>(ne_, 1));
        for (uint32_t v = 0; v < nv_; ++v)
            if (deg[v]) std::memcpy(&col[row[v]], &slots[off[v]], (size_t)deg[v] * 4);
        std::vector<std::vector<uint8_t>> host(arrs.size());
        for (size_t i = 0; i < arrs.size(); ++i) {
            host[i].resize(arrs[i].second);
            SG_HIP(hipMemcpyAsync(host[i].data(), arrs[i].first, arrs[i].second, hipMemcpyDeviceToHost, stream_));
        }
        sync();
        StateHeader h{};
        std::memcpy(h.magic, kStateMagic, 8);
        h.algo = cfg_.algo;
        h.nv = nv_;
        h.ne = ne_;
        h.source = src;
        h.narr = (uint32_t)arrs.size();
        h.weight_mode = cfg_.weight_mode;
        h.weight_seed = cfg_.weight_seed;
        h.pr_alpha = cfg_.pr_alpha;
        h.pr_epsilon = cfg_.pr_epsilon;
        FILE *f = std::fopen(path, "wb");
        if (!f) fail(SG_EIO, std::string("cannot open ") + path);
        bool ok = std::fwrite(&h, sizeof h, 1, f) == 1;
        for (auto &x : host) {
            const uint64_t n = x.size();
            ok = ok && std::fwrite(&n, 8, 1, f) == 1 && (n == 0 || std::fwrite(x.data(), 1, n, f) == n);
        }
        ok = ok && std::fwrite(row.data(), 8, row.size(), f) == row.size() &&
             (ne_ == 0 || std::fwrite(col.data(), 4, ne_, f) == ne_);
        ok = (std::fclose(f) == 0) && ok;
        if (!ok) fail(SG_EIO, std::string("cannot write ") + path);
    }

    void restore_state(const char *path) {
        if (loaded_) fail(SG_EINVAL, "sg_restore_state needs a fresh engine (no sg_load_csr before it)");
        FILE *f = std::fopen(path, "rb");
        if (!f) fail(SG_EIO, std::string("cannot open ") + path);
        std::unique_ptr<FILE, int (*)(FILE *)> fc(f, std::fclose);
        StateHeader h{};
        if (std::fread(&h, sizeof h, 1, f) != 1 || std::memcmp(h.magic, kStateMagic, 8) != 0)
            fail(SG_EIO, "not a state checkpoint");
        if (h.algo != cfg_.algo) fail(SG_EINVAL, "checkpoint of another program");
        if ((h.algo == SG_ALGO_PR && (h.pr_alpha != cfg_.pr_alpha || h.pr_epsilon != cfg_.pr_epsilon)) ||
            (h.algo == SG_ALGO_SSSP && (h.weight_mode != cfg_.weight_mode || h.weight_seed != cfg_.weight_seed)))
            fail(SG_EINVAL, "engine configuration differs from the checkpoint's");
        if (h.nv == 0 || h.narr > 8) fail(SG_EIO, "bad checkpoint header");
        std::vector<std::vector<uint8_t>> host(h.narr);
        for (auto &x : host) {
            uint64_t n = 0;
            if (std::fread(&n, 8, 1, f) != 1 || n > (uint64_t)h.nv * 16) fail(SG_EIO, "bad checkpoint state array");
            x.resize(n);
            if (n && std::fread(x.data(), 1, n, f) != n) fail(SG_EIO, "truncated checkpoint");
        }
        std::vector<uint64_t> row((size_t)h.nv + 1);
        std::vector<uint32_t> col(std::max<uint64_t>(h.ne, 1));
        if (std::fread(row.data(), 8, row.size(), f) != row.size() ||
            (h.ne && std::fread(col.data(), 4, h.ne, f) != h.ne) || std::fgetc(f) != EOF)
            fail(SG_EIO, "truncated or oversized checkpoint");
        load_csr(h.nv, h.ne, row.data(), col.data());  // validates the offsets and ids
        std::vector<std::pair<void *, size_t>> arrs;
        uint32_t unused = 0;
        state_arrays(arrs, unused);
        if (arrs.size() != host.size()) fail(SG_EIO, "checkpoint state does not fit this program");
        for (size_t i = 0; i < arrs.size(); ++i) {
            if (arrs[i].second != host[i].size()) fail(SG_EIO, "checkpoint state does not fit this program");
            SG_HIP(hipMemcpyAsync(arrs[i].first, host[i].data(), host[i].size(), hipMemcpyHostToDevice, stream_));
        }
        sync();
        if (h.algo != SG_ALGO_CC && h.algo != SG_ALGO_PR && h.source >= h.nv) fail(SG_EIO, "bad checkpoint source");
        restored(h.source);
    }
};

// A batch's (u, v) pairs as a relabelled in-edge store records them: u
// replaced by its id there (ids >= nv stay, to be dropped as invalid).
__global__ void k_relabel_pairs(const uint32_t *in, uint64_t n, const uint32_t *relabel, uint32_t nv, uint32_t *out) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const uint32_t u = in[2 * i];
        out[2 * i] = u < nv ? relabel[u] : u;
        out[2 * i + 1] = in[2 * i + 1];
    }
}

// ---------------------------------------------------------------------------
// In-edge store: a PMA of the reversed graph owned by another engine, on that
// engine's stream, updated by the owner's batches read as (v, u). No vertex
// program (the reference's pull side, pull_functor.h:22-369, keeps no store
// of its own; its in-edges would come from a CSC the reference never builds).
// ---------------------------------------------------------------------------
class GraphStore final : public Engine {
public:
    using Engine::Engine;
    SweepTiles tiles_{};  // merge-path tiles over the in-lists (refreshed after every batch)
    // With relabelled sources (make_in_store): relabel_[u] is the id the
    // in-lists record for source u, unlabel_ the inverse; null: ids as given.
    uint32_t *relabel_ = nullptr, *unlabel_ = nullptr;
    uint32_t *rl_ins_ = nullptr, *rl_del_ = nullptr;
    size_t rl_ins_cap_ = 0, rl_del_cap_ = 0;
    void init(const uint32_t *, uint32_t) override { fail(SG_EINVAL, "a graph store has no vertex program"); }
    void apply(const uint32_t *d_ins, uint64_t nins, const uint32_t *d_del, uint64_t ndel) override {
        if (relabel_) {  // the batch's sources as the in-lists record them
            if (nins) {
                ensure(rl_ins_, rl_ins_cap_, 2 * nins);
                hipLaunchKernelGGL(k_relabel_pairs, dim3(grid_for(nins, 256, 4096)), dim3(256), 0, stream_, d_ins, nins,
                                   (const uint32_t *)relabel_, nv_, rl_ins_);
                SG_LAUNCHED();
                d_ins = rl_ins_;
            }
            if (ndel) {
                ensure(rl_del_, rl_del_cap_, 2 * ndel);
                hipLaunchKernelGGL(k_relabel_pairs, dim3(grid_for(ndel, 256, 4096)), dim3(256), 0, stream_, d_del, ndel,
                                   (const uint32_t *)relabel_, nv_, rl_del_);
                SG_LAUNCHED();
                d_del = rl_del_;
            }
        }
        const bool overflow = prepare(d_ins, nins, d_del, ndel);
        commit(overflow);
        finish_batch();
        st_.snapshots += 1;
        refresh_tiles();
        refresh_mirror();
    }
    // With the store in pinned host memory, the device edge budget holds an
    // HBM copy of its leading live edges, gaps left out (every pull round
    // reads the whole store; those edges then come from HBM instead of over
    // the host link). Rebuilt after every batch from the live slots (the
    // live-edge order shifts with every insert and delete).
    uint64_t mirror_extra_bytes_ = 0;  // HBM beyond the edge budget the owner gives the mirror
    // With want_tail_ (an owner whose pull rounds read every in-edge), the
    // live edges past the mirror also get a gap-free copy in pinned host
    // memory: a pull then reads them as contiguous runs over the host link
    // instead of short lists scattered between gaps. Rebuilt with the mirror.
    bool want_tail_ = false;
    uint64_t tail_cap_ = 0;
    // Ids are packed at ceil(log2 nv) bits (27 at 2^27 vertices): the same
    // HBM holds 32 / 27 as many edges, and the host link moves that much less
    // per tail edge. Packed words are ORed together at run ends, so the copies
    // are zeroed first. The tail is packed in windows of an HBM stage (at most
    // kTailStageWords, taken out of the mirror's budget) and each window
    // copied out before the next is packed, so no more of the store than the
    // budget allows is held in HBM.
    static constexpr uint64_t kTailStageWords = 16ull << 20;  // 64 MiB
    static constexpr uint64_t kTailStageMinWords = 1024;
    static constexpr uint64_t kCopySlack = 4;
    uint32_t *tail_stage_ = nullptr;
    uint64_t tail_stage_cap_ = 0;
    static uint64_t packed_words(uint64_t n, uint32_t b) { return (n * b + 31) / 32 + 1; }
    static uint64_t ids_in_words(uint64_t words, uint32_t b) { return words > 1 ? (words - 1) * 32 / b : 0; }
    void refresh_mirror() {
        mirror_live_ = 0;
        if (!host_edges_) return;
        const uint32_t b = nv_ > 1 ? 32 - (uint32_t)__builtin_clz(nv_ - 1) : 1;
        mbits_ = b > 31 ? 32 : b;
        const double bytes = std::max(0.0, cfg_.edge_budget_gb) * 1073741824.0 + (double)mirror_extra_bytes_;
        uint64_t cap_words = (uint64_t)(bytes / 4.0);
        // the tile sweeps read whole 16-byte groups: 4 words of slack past the last id
        auto fits = [&](uint64_t words) { return ids_in_words(words > kCopySlack ? words - kCopySlack : 0, mbits_); };
        // a tail needs a stage: its words come out of the mirror's share
        uint64_t stage_words = 0;
        if (want_tail_ && ne_ > fits(cap_words)) {
            stage_words = std::max(kTailStageMinWords, std::min(kTailStageWords, cap_words / 2));
            cap_words = cap_words > stage_words ? cap_words - stage_words : 0;
        }
        const uint64_t want = std::min<uint64_t>(ne_, fits(cap_words));
        const uint64_t rest = want_tail_ ? ne_ - want : 0;
        if (want == 0 && rest == 0) return;
        const uint64_t mwords = want ? packed_words(want, mbits_) : 0;
        if (mwords && mwords + kCopySlack > mirror_cap_) {
            dfree(mirror_);
            mirror_ = nullptr;
            mirror_cap_ = std::min<uint64_t>(cap_words, mwords + std::max<uint64_t>(kCopySlack, mwords / 8));  // room for growth
            mirror_ = dalloc<uint32_t>(mirror_cap_);
        }
        const uint64_t twords = rest ? packed_words(rest, mbits_) : 0;
        if (twords && twords + kCopySlack > tail_cap_) {
            hfree(tail_);
            tail_ = nullptr;
            tail_cap_ = twords + twords / 8 + kCopySlack;
            tail_ = halloc<uint32_t>(tail_cap_);
        }
        if (rest) {
            stage_words = std::min(stage_words, twords);
            if (stage_words > tail_stage_cap_) {
                dfree(tail_stage_);
                tail_stage_ = nullptr;
                tail_stage_ = dalloc<uint32_t>(stage_words);
                tail_stage_cap_ = stage_words;
            }
        }
        GraphView g = gview();  // without copies: the fill reads the host store
        g.mirror = nullptr;
        g.mlive = 0;
        g.tail = nullptr;
        const uint32_t grid = sweep_grid(k_mirror_fill, ntiles());
        if (want) {
            if (mbits_ < 32) fill<uint32_t>(mirror_, mwords, 0u);
            hipLaunchKernelGGL(k_mirror_fill, dim3(grid), dim3(kSweepBlock), 0, stream_, g,
                               (const unsigned long long *)dpre_, tiles_, mirror_, (uint64_t)want, (uint32_t *)nullptr,
                               (uint64_t)0, (uint64_t)0, mbits_);
            SG_LAUNCHED();
        }
        // tail windows of win ids (a multiple of 32: each starts on a word)
        const uint64_t win = rest ? std::max<uint64_t>(32, ids_in_words(stage_words, mbits_) & ~31ull) : 0;
        for (uint64_t lo = 0; lo < rest; lo += win) {
            const uint64_t hi = std::min(rest, lo + win);
            const uint64_t used = ((hi - lo) * mbits_ + 31) / 32;  // words holding the window's bits
            if (used > tail_stage_cap_) fail(SG_EDEVICE, "tail window exceeds its stage");
            if (mbits_ < 32) fill<uint32_t>(tail_stage_, used, 0u);
            hipLaunchKernelGGL(k_mirror_fill, dim3(grid), dim3(kSweepBlock), 0, stream_, g,
                               (const unsigned long long *)dpre_, tiles_, (uint32_t *)nullptr, (uint64_t)want,
                               tail_stage_, lo, hi, mbits_);
            SG_LAUNCHED();
            SG_HIP(hipMemcpyAsync(tail_ + lo * mbits_ / 32, tail_stage_, used * 4, hipMemcpyDeviceToHost, stream_));
        }
        if (!tsplit_) tsplit_ = dalloc<unsigned long long>(1);
        hipLaunchKernelGGL(k_tile_split, dim3(1), dim3(1), 0, stream_, tiles_, (const unsigned long long *)dpre_, nv_,
                           (unsigned long long)want, tsplit_);
        SG_LAUNCHED();
        mirror_live_ = want;
    }
    uint64_t mirror_cap_ = 0;
    unsigned long long *tsplit_ = nullptr;  // first tile past the mirror (k_tile_split)
    // The tile split for pull passes: null without a mirror, or with every
    // live edge in it.
    const unsigned long long *tile_split() const {
        return host_edges_ && mirror_live_ && mirror_live_ < ne_ ? tsplit_ : nullptr;
    }

    // Mirror capacity (ids) and the live edges it holds.
    void mirror_counts(uint64_t &slots, uint64_t &live) {
        slots = mirror_ && mirror_cap_ > 1 ? (mirror_cap_ - 1) * 32 / mbits_ : 0;
        live = std::min(mirror_live_, ne_);
    }
    void refresh_tiles() { tiles_ = sweep_tiles(); }
    uint64_t ntiles() const { return ((uint64_t)nv_ + ne_ + kSweepTile - 1) / kSweepTile; }
    void gather(uint32_t q, void *out, size_t bytes) override { gather_in(q, out, bytes); }
    // The store's arrays with sources as vertex ids (relabelled slots mapped back).
    void gather_in(uint32_t q, void *out, size_t bytes) {
        gather_graph(q, out, bytes);
        if (q != SG_Q_SLOTS || !unlabel_) return;
        std::vector<uint32_t> un(nv_);
        SG_HIP(hipMemcpyAsync(un.data(), unlabel_, (size_t)nv_ * 4, hipMemcpyDeviceToHost, stream_));
        sync();
        uint32_t *slots = (uint32_t *)out;
        for (uint64_t i = 0; i < cap_; ++i)
            if (slots[i] != kGap) slots[i] = un[slots[i]];
    }
};

// In-lists of a CSR (counting sort by target; sources ascending in each list).
static void transpose_csr(uint32_t nv, uint64_t ne, const uint64_t *row, const uint32_t *col, std::vector<uint64_t> &trow,
                          std::vector<uint32_t> &tcol) {
    trow.assign((size_t)nv + 1, 0);
    for (uint64_t e = 0; e < ne; ++e) ++trow[(size_t)col[e] + 1];
    for (uint32_t v = 0; v < nv; ++v) trow[(size_t)v + 1] += trow[v];
    std::vector<uint64_t> pos(trow.begin(), trow.end() - 1);
    tcol.resize(std::max<uint64_t>(ne, 1));
    for (uint32_t u = 0; u < nv; ++u)
        for (uint64_t e = row[u]; e < row[(size_t)u + 1]; ++e) tcol[pos[col[e]]++] = u;
}

// keys[e] = target << 32 | source of every edge e (one wave per source).
// relabel (null: identity): the id the in-lists record for source v.
__global__ void __launch_bounds__(kBlock) k_edge_keys_t(uint32_t nv, const uint64_t *row, const uint32_t *col,
                                                        const uint32_t *relabel, unsigned long long *keys) {
    const uint64_t gw = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
    const uint64_t nw = ((uint64_t)gridDim.x * blockDim.x) / kWave;
    for (uint64_t v = gw; v < nv; v += nw) {
        const uint64_t b = row[v], d = row[v + 1] - b;
        const uint32_t id = relabel ? relabel[v] : (uint32_t)v;
        for (uint64_t k = lane_id(); k < d; k += kWave) keys[b + k] = ((unsigned long long)col[b + k] << 32) | id;
    }
}
// Sorted keys (hi << 32 | lo) as a CSR: row[v] = first key with hi >= v, col[i] = lo.
__global__ void k_keys_to_csr(uint32_t nv, uint64_t ne, const unsigned long long *keys, uint64_t *row, uint32_t *col) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < ne; i += stride) col[i] = (uint32_t)keys[i];
    for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v <= nv; v += stride) {
        const unsigned long long key = v << 32;
        uint64_t lo = 0, hi = ne;
        while (lo < hi) {
            const uint64_t m = (lo + hi) >> 1;
            if (keys[m] < key) lo = m + 1;
            else hi = m;
        }
        row[v] = lo;
    }
}

// In-lists of a large CSR on the device: a radix sort of (target, source)
// keys, then the CSR of the sorted keys (sources ascending in each list, as
// transpose_csr gives). Transient device memory of about 20 bytes per edge,
// outside the engine's budget (a load-time scratch, freed before returning).
static void transpose_csr_device(hipStream_t st, uint32_t nv, uint64_t ne, const uint64_t *row, const uint32_t *col,
                                 const uint32_t *rank, std::vector<uint64_t> &trow, std::vector<uint32_t> &tcol) {
    void *d_row = nullptr, *d_col = nullptr, *d_k0 = nullptr, *d_k1 = nullptr, *d_tmp = nullptr, *d_rank = nullptr;
    auto release = [&]() {
        for (void *p : {d_row, d_col, d_k0, d_k1, d_tmp, d_rank})
            if (p) (void)hipFree(p);
    };
    try {
        SG_HIP(hipMalloc(&d_row, ((size_t)nv + 1) * 8));
        SG_HIP(hipMalloc(&d_col, (size_t)ne * 4));
        SG_HIP(hipMalloc(&d_k0, (size_t)ne * 8));
        SG_HIP(hipMalloc(&d_k1, (size_t)ne * 8));
        SG_HIP(hipMemcpyAsync(d_row, row, ((size_t)nv + 1) * 8, hipMemcpyHostToDevice, st));
        SG_HIP(hipMemcpyAsync(d_col, col, (size_t)ne * 4, hipMemcpyHostToDevice, st));
        if (rank) {
            SG_HIP(hipMalloc(&d_rank, (size_t)nv * 4));
            SG_HIP(hipMemcpyAsync(d_rank, rank, (size_t)nv * 4, hipMemcpyHostToDevice, st));
        }
        hipLaunchKernelGGL(k_edge_keys_t, dim3(grid_for(nv, kWavesPerBlock, 16384)), dim3(kBlock), 0, st, nv,
                           (const uint64_t *)d_row, (const uint32_t *)d_col, (const uint32_t *)d_rank,
                           (unsigned long long *)d_k0);
        SG_LAUNCHED();
        const int kb = 32 - __builtin_clz(std::max<uint32_t>(nv, 1));
        size_t tb = 0;
        SG_HIP(hipcub::DeviceRadixSort::SortKeys(nullptr, tb, (const unsigned long long *)d_k0, (unsigned long long *)d_k1,
                                                 ne, 0, 32 + kb, st));
        SG_HIP(hipMalloc(&d_tmp, tb + 256));
        size_t bb = tb + 256;
        SG_HIP(hipcub::DeviceRadixSort::SortKeys(d_tmp, bb, (const unsigned long long *)d_k0, (unsigned long long *)d_k1, ne,
                                                 0, 32 + kb, st));
        hipLaunchKernelGGL(k_keys_to_csr, dim3(4096), dim3(256), 0, st, nv, ne, (const unsigned long long *)d_k1,
                           (uint64_t *)d_row, (uint32_t *)d_col);
        SG_LAUNCHED();
        trow.resize((size_t)nv + 1);
        tcol.resize(std::max<uint64_t>(ne, 1));
        SG_HIP(hipMemcpyAsync(trow.data(), d_row, ((size_t)nv + 1) * 8, hipMemcpyDeviceToHost, st));
        SG_HIP(hipMemcpyAsync(tcol.data(), d_col, (size_t)ne * 4, hipMemcpyDeviceToHost, st));
        SG_HIP(hipStreamSynchronize(st));
    } catch (...) {
        release();
        throw;
    }
    release();
}

// The in-edge store of an engine, built from the CSR it loads (its own device
// budget counted with the owner's, its work on the owner's stream). In-lists
// are built on the host for small graphs, on the device from 2^24 edges.
// mirror_extra_bytes: HBM beyond the edge budget for the store's mirror when
// its edges live in host memory.
// tail: the owner's pull rounds read every in-edge (a gap-free host copy of
// the edges past the mirror pays for its rebuild).
// relabel: the in-lists record each source by its rank in descending
// out-degree instead of its id (hubs first), so the per-in-edge gathers of a
// pull pass (frontier bit and push of the source) fall on a few hot words
// that stay in the caches; the store translates batches and queries.
static std::unique_ptr<GraphStore> make_in_store(Engine &owner, const uint64_t *row, const uint32_t *col,
                                                 uint64_t mirror_extra_bytes = 0, bool tail = false, bool relabel = false) {
    const uint32_t nv = owner.nv_;
    std::vector<uint32_t> rank, unrank;
    if (relabel) {  // counting sort by out-degree, descending; ties by id
        uint64_t maxd = 0;
        for (uint32_t u = 0; u < nv; ++u) maxd = std::max<uint64_t>(maxd, row[u + 1] - row[u]);
        std::vector<uint64_t> start(maxd + 2, 0);
        for (uint32_t u = 0; u < nv; ++u) ++start[maxd - (row[u + 1] - row[u]) + 1];
        for (uint64_t d = 1; d < start.size(); ++d) start[d] += start[d - 1];
        rank.resize(nv);
        unrank.resize(nv);
        for (uint32_t u = 0; u < nv; ++u) {
            const uint32_t r = (uint32_t)start[maxd - (row[u + 1] - row[u])]++;
            rank[u] = r;
            unrank[r] = u;
        }
    }
    std::vector<uint64_t> trow;
    std::vector<uint32_t> tcol;
    if (owner.ne_ >= (1ull << 24)) {
        transpose_csr_device(owner.stream_, nv, owner.ne_, row, col, relabel ? rank.data() : nullptr, trow, tcol);
    } else {
        transpose_csr(nv, owner.ne_, row, col, trow, tcol);
        if (relabel) {
            for (uint64_t i = 0; i < owner.ne_; ++i) tcol[i] = rank[tcol[i]];
        }
    }
    sg_config rc = owner.cfg_;
    rc.flags &= SG_FLAG_HOST_EDGES;  // no timing, graphs or deferral of its own
    rc.cache_level = 0;
    std::unique_ptr<GraphStore> in(new GraphStore(rc));
    in->swap_pairs_ = true;
    in->budget_peer_ = &owner;
    in->init_device(owner.stream_);
    in->load_csr(owner.nv_, owner.ne_, trow.data(), tcol.data());
    if (relabel) {
        in->relabel_ = in->dalloc<uint32_t>(nv);
        in->unlabel_ = in->dalloc<uint32_t>(nv);
        SG_HIP(hipMemcpyAsync(in->relabel_, rank.data(), (size_t)nv * 4, hipMemcpyHostToDevice, owner.stream_));
        SG_HIP(hipMemcpyAsync(in->unlabel_, unrank.data(), (size_t)nv * 4, hipMemcpyHostToDevice, owner.stream_));
        in->sync();  // the host vectors die with this call
    }
    in->mirror_extra_bytes_ = mirror_extra_bytes;
    in->want_tail_ = tail;
    in->refresh_tiles();
    in->refresh_mirror();
    return in;
}

// ---------------------------------------------------------------------------
// traversal operators (SSSP, BFS, CC)
// ---------------------------------------------------------------------------
// One traversal program over an engine's graph: its vertex state, the static
// solve and the per-snapshot recompute. A TraversalEngine runs one; the
// PageRank + CC engine runs a CC program beside PageRank on the same stores.
// Operators with the pull-shape method (MergeBuffer) can take pull rounds.
template <class Op, class = void>
struct has_merge : std::false_type {};
template <class Op>
struct has_merge<Op, std::void_t<decltype(std::declval<const Op &>().MergeBuffer(
                         0u, (typename Op::buffer_t *)nullptr, typename Op::buffer_t{}))>> : std::true_type {};

// Frontier out-edges above which a traversal round pulls, as a share of the
// edges (E / divisor). A traversal push peeks the destination before any
// atomic, so per edge it costs about what a pull costs (one 8-byte gather);
// the pull pays when the frontier holds half the edges (the first rounds of
// a CC solve, the middle levels of a BFS from a hub).
constexpr uint32_t kTravPullDivisor = 2;

template <class Op>
struct TravProgram {
    Op op{};
    uint32_t *value = nullptr;
    unsigned long long *buffer = nullptr;
    int phase_inval = Engine::kPhaseInval, phase_relax = Engine::kPhaseRelax;
    // pull rounds (direction-optimising): the in-edge store they walk, the
    // frontier bitmaps and pushes of dense rounds; null: push rounds only
    GraphStore *pull_in = nullptr;
    uint32_t *fb[2] = {nullptr, nullptr};
    uint32_t *pv = nullptr;
    uint32_t pull_div = kTravPullDivisor;

    // Pull rounds over in (cfg.pull_divisor: 0 = default share, UINT32_MAX =
    // never, UINT32_MAX - 1... = a round pulls above E / divisor).
    void enable_pull(Engine &e, GraphStore *in, uint32_t divisor) {
        if constexpr (has_merge<Op>::value) {
            if (!in || in->relabel_ || divisor == 0xFFFFFFFFu || pv) return;
            pull_in = in;
            pull_div = divisor ? divisor : kTravPullDivisor;
            const uint64_t words = ((uint64_t)e.nv_ + 31) / 32;
            for (auto &b : fb) {
                b = e.dalloc<uint32_t>(words);
                e.template fill<uint32_t>(b, words, 0u);
            }
            pv = e.dalloc<uint32_t>(e.nv_);
        }
    }

    void alloc(Engine &e) {
        if (!value) {
            value = e.dalloc<uint32_t>(e.nv_);
            buffer = e.dalloc<unsigned long long>(e.nv_);
        }
    }
    std::pair<VCombine<Op>, ERelax<Op>> relax_fns(const Engine &e, uint32_t stamp_val) const {
        return {VCombine<Op>{op, value, buffer}, ERelax<Op>{op, buffer, e.stamp_, stamp_val, e.cfg_.weight_mode, e.cfg_.weight_seed}};
    }

    // InitGraph + Start (framework.cuh:389-439, 549-584): returns the rounds.
    uint32_t solve(Engine &e) {
        alloc(e);
        hipLaunchKernelGGL(k_init_state<Op>, dim3(grid_for(e.nv_, 256, 16384)), dim3(256), 0, e.stream_, e.nv_, op, value, buffer);
        SG_LAUNCHED();
        e.frontier_reset();
        hipLaunchKernelGGL(k_seed_active<Op>, dim3(grid_for(e.nv_, kBlock, 16384)), dim3(kBlock), 0, e.stream_, e.nv_, op, value,
                           buffer, e.stamp_, e.R_, e.q_[e.R_ & 1], e.cnt_ + (e.R_ % 3), e.stats_ + 3);
        SG_LAUNCHED();
        return relax_rounds(e, false, -1);
    }

    // Relax rounds from the frontier of round R_ to convergence; with an
    // in-edge store each round is a push or a pull, chosen on the device from
    // the frontier's out-edges.
    uint32_t relax_rounds(Engine &e, bool timed, int phase) {
        if constexpr (has_merge<Op>::value) {
            if (pv) return relax_rounds_pull(e, timed, phase);
        }
        return e.run_rounds([&](uint32_t sv) { return relax_fns(e, sv); }, timed, phase);
    }

    uint32_t relax_rounds_pull(Engine &e, bool timed, int phase) {
        // out-edges of the seeded frontier: the first round's direction
        hipLaunchKernelGGL(k_list_degsum, dim3(256), dim3(kBlock), 0, e.stream_, (const uint32_t *)e.q_[e.R_ & 1],
                           (const uint32_t *)(e.cnt_ + (e.R_ % 3)), e.nv_, (const uint32_t *)e.deg_, e.ecnt_ + (e.R_ % 3));
        SG_LAUNCHED();
        const bool do_time = timed && e.timing(SG_KCLASS_RELAX);
        const bool time_pull = timed && e.timing(SG_KCLASS_PULL);
        const unsigned long long thr = std::max<unsigned long long>(1ull, e.ne_ / pull_div);
        const uint64_t p0 = e.st_.pull_rounds;
        const uint32_t r = e.drive_rounds([&](const PassIO &io0) {
            PassIO io = io0;
            const uint32_t R = io.round;
            io.ecnt_in = e.ecnt_ + (R % 3);
            io.ecnt_out = e.ecnt_ + ((R + 1) % 3);
            io.ecnt_zero = e.ecnt_ + ((R + 2) % 3);
            io.fdeg = e.deg_;
            io.dense_thr = thr;
            // round R's frontier bitmap starts empty (it last served round R - 2)
            const uint64_t words = ((uint64_t)e.nv_ + 31) / 32;
            e.round_launch(k_fill_u32, dim3(grid_for(words, 256, 2048)), dim3(256), fb[R & 1], words, 0u);
            e.launch_pass(io, VCombinePull<Op>{op, value, buffer, PullTake{fb[R & 1], pv, nullptr}},
                          ERelax<Op>{op, buffer, e.stamp_, R + 1, e.cfg_.weight_mode, e.cfg_.weight_seed}, do_time);
            // the pull kernel returns at once unless the round is dense (device flag)
            const uint64_t nt = pull_in->ntiles();
            e.begin_timed(time_pull && !e.batch_, SG_KCLASS_PULL);
            e.round_launch(k_trav_pull<Op>, dim3(e.sweep_grid(k_trav_pull<Op>, nt)), dim3(kSweepBlock), pull_in->gview(),
                           (const unsigned long long *)pull_in->dpre_, pull_in->tiles_, (const uint32_t *)pull_in->col_,
                           (uint64_t)0,
                           PullTrav<Op>{op, fb[R & 1], pv, buffer, e.stamp_, R, e.cfg_.weight_mode, e.cfg_.weight_seed},
                           io, 0);
            e.end_timed(time_pull && !e.batch_);
        }, timed, phase);
        // 4 B per in-edge slot read, its bitmap word (4 B), 8 B buffer + 4 B
        // stamp per destination
        if (time_pull) e.kt_bytes_[SG_KCLASS_PULL] += (e.st_.pull_rounds - p0) * (8ull * e.ne_ + 12ull * e.nv_);
        return r;
    }

    // After the batch is committed: (1) reset the subtrees hanging from deleted
    // tree edges (to a fixed point); (2) seed the recompute with the best value
    // each reset vertex can take from a live in-edge (in: from its in-list in
    // the in-edge store; else from a sweep over every live edge) and with the
    // inserted edges; (3) converge. Returns the convergence rounds.
    uint32_t recompute(Engine &e, GraphStore *in) {
        // the bitmaps are zero here (cleared by the last batch's activation pass)
        e.frontier_reset();
        if (e.ndel_) {
            hipLaunchKernelGGL(k_inval_seed<Op>, dim3(grid_for(e.ndel_, kBlock, 8192)), dim3(kBlock), 0, e.stream_, op,
                               (const unsigned long long *)e.bv_.del_key, e.ndel_, value, buffer, e.inval_, e.coarse_,
                               e.cshift_, e.q_[e.R_ & 1], e.cnt_ + (e.R_ % 3), e.nv_, e.stats_);
            SG_LAUNCHED();
            e.run_rounds([&](uint32_t) {
                return std::make_pair(VAll{}, EInvalidate<Op>{op, value, buffer, e.inval_, e.coarse_, e.cshift_});
            }, false, phase_inval);
        }
        // the invalidation rounds' last read of the control block holds the count
        // every vertex the invalidation seeds and rounds appended was reset once
        const uint64_t ninval = e.ndel_ ? e.last_appended_ : 0;
        e.st_.invalidated = ninval;
        e.frontier_reset();
        const uint32_t R = e.R_;
        if (ninval) {
            if (in) {
                // the reset vertices as a list in the next round's frontier buffer
                // (free until round R runs), their in-lists walked by a round kernel
                uint32_t *list = e.q_[(R + 1) & 1], *lcnt = e.cnt_ + 6;
                hipLaunchKernelGGL(k_bits_to_list, dim3(grid_for((e.nv_ + 31) / 32, kBlock, 4096)), dim3(kBlock), 0, e.stream_,
                                   e.nv_, (const uint32_t *)e.inval_, list, lcnt, e.nv_, e.stats_ + 3);
                SG_LAUNCHED();
                PassIO io = e.seed_io(R, list, lcnt, 0);
                const EPullSeed<Op> ef{op, buffer, e.stamp_, R, e.cfg_.weight_mode, e.cfg_.weight_seed,
                                       (const uint32_t *)in->unlabel_};
                const GraphView ig = in->gview();
                e.launch_pass_view(ig, io, VAll{}, ef, false);
                e.drain_seed_spill_view(ig, R, ef);  // long in-lists cut into chunks
            } else {
                // edges into invalidated vertices: edge-centric sweep over all live edges
                const SweepTiles tl = e.sweep_tiles();
                PassIO io = e.seed_io(R, nullptr, nullptr, e.nv_);
                ERelaxInval<Op> ef{ERelax<Op>{op, buffer, e.stamp_, R, e.cfg_.weight_mode, e.cfg_.weight_seed}, e.inval_,
                                   e.coarse_, e.cshift_};
                e.sweep_all(tl, op, (const typename Op::buffer_t *)buffer, ef, io);
            }
            // invalidated vertices active on their own (CC: own label); clears the bitmaps
            hipLaunchKernelGGL(k_inval_activate<Op>, dim3(grid_for((e.nv_ + 31) / 32, 256, 4096)), dim3(256), 0, e.stream_, e.nv_,
                               op, e.inval_, e.coarse_, e.cshift_, value, buffer, e.stamp_, R, e.q_[R & 1], e.cnt_ + (R % 3),
                               e.nv_, e.stats_ + 3);
            SG_LAUNCHED();
        }
        if (e.nins_) {
            hipLaunchKernelGGL(k_insert_relax<Op>, dim3(grid_for(e.nins_, kBlock, 8192)), dim3(kBlock), 0, e.stream_, op, e.bv_,
                               e.br_, buffer, e.stamp_, R, e.cfg_.weight_mode, e.cfg_.weight_seed, e.q_[R & 1],
                               e.cnt_ + (R % 3), e.nv_, e.stats_);
            SG_LAUNCHED();
        }
        return relax_rounds(e, true, phase_relax);
    }

    void word_source(const Engine &, uint32_t q, const uint32_t *&base, uint64_t &stride, uint32_t &word) const {
        if (q == SG_Q_VALUE) { base = value; stride = 1; word = 0; return; }
        if (q == SG_Q_PARENT) { base = (const uint32_t *)buffer; stride = 2; word = 0; return; }  // low word
        if (q == SG_Q_BUFFER) { base = (const uint32_t *)buffer; stride = 2; word = 1; return; }  // high word
        fail(SG_EINVAL, "query not available pointwise");
    }

    void gather(Engine &e, uint32_t q, void *out, size_t bytes) const {
        if (!value) fail(SG_EINVAL, "no vertex state before sg_init");
        if (bytes < (size_t)e.nv_ * 4) fail(SG_EINVAL, "output too small");
        if (q == SG_Q_VALUE) {
            SG_HIP(hipMemcpyAsync(out, value, (size_t)e.nv_ * 4, hipMemcpyDeviceToHost, e.stream_));
        } else {
            e.ensure(e.tmp_, e.tmp_cap_, e.nv_);
            hipLaunchKernelGGL(k_extract<Op>, dim3(grid_for(e.nv_, 256, 16384)), dim3(256), 0, e.stream_, e.nv_, buffer, e.tmp_,
                               q == SG_Q_BUFFER ? 1 : 0);
            SG_LAUNCHED();
            SG_HIP(hipMemcpyAsync(out, e.tmp_, (size_t)e.nv_ * 4, hipMemcpyDeviceToHost, e.stream_));
        }
        e.sync();
    }
};

// Operators with a source vertex (a member m_source_node, as the reference's
// SSSP and BFS apps) take it from sg_init; the others (CC) take none.
template <class Op, class = void>
struct has_source : std::false_type {};
template <class Op>
struct has_source<Op, std::void_t<decltype(std::declval<Op &>().m_source_node)>> : std::true_type {};

template <class Op>
class TraversalEngine final : public Engine {
public:
    using Engine::Engine;
    TravProgram<Op> prog_;
    std::unique_ptr<GraphStore> in_;  // SG_FLAG_IN_EDGES: the in-edge store the recompute seeds from

    void after_load(const uint64_t *row, const uint32_t *col) override {
        if (cfg_.flags & SG_FLAG_IN_EDGES) {
            in_ = make_in_store(*this, row, col);
            prog_.enable_pull(*this, in_.get(), cfg_.pull_divisor);
        }
    }
    uint64_t aux_device_bytes() const override { return in_ ? in_->dev_bytes_ : 0; }
    uint64_t aux_host_bytes() const override { return in_ ? in_->host_bytes_ : 0; }
    uint64_t in_capacity() const override { return in_ ? in_->cap_ : 0; }
    void in_mirror(uint64_t &slots, uint64_t &live) override {
        slots = live = 0;
        if (in_) in_->mirror_counts(slots, live);
    }

    void state_arrays(std::vector<std::pair<void *, size_t>> &a, uint32_t &source) override {
        prog_.alloc(*this);
        a = {{prog_.value, (size_t)nv_ * 4}, {prog_.buffer, (size_t)nv_ * 8}};
        source = 0;
        if constexpr (has_source<Op>::value) source = prog_.op.m_source_node;
    }
    void restored(uint32_t source) override {
        if constexpr (has_source<Op>::value) prog_.op.m_source_node = source;
        cache_rebuild();
        inited_ = true;
    }

    void init(const uint32_t *sources, uint32_t nsrc) override {
        if (!loaded_) fail(SG_EINVAL, "sg_init before sg_load_csr");
        if constexpr (has_source<Op>::value) {
            if (nsrc != 1 || !sources) fail(SG_EINVAL, "this program needs exactly one source");
            if (sources[0] >= nv_) fail(SG_EINVAL, "source id >= nv");
            prog_.op.m_source_node = sources[0];
        }
        auto t0 = std::chrono::steady_clock::now();
        st_.rounds = prog_.solve(*this);
        cache_rebuild();
        inited_ = true;
        st_.last_phase_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }

    void apply(const uint32_t *d_ins, uint64_t nins, const uint32_t *d_del, uint64_t ndel) override {
        if (!inited_) fail(SG_EINVAL, "sg_apply before sg_init");
        auto t0 = std::chrono::steady_clock::now();
        const bool overflow = prepare(d_ins, nins, d_del, ndel);
        commit(overflow);
        if (in_) in_->apply(d_ins, nins, d_del, ndel);
        st_.rounds = prog_.recompute(*this, in_.get());
        finish_batch();
        st_.snapshots += 1;
        st_.last_phase_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }

    void word_source(uint32_t q, uint32_t qi, const uint32_t *&base, uint64_t &stride, uint32_t &word) override {
        if (q == SG_Q_VALUE || q == SG_Q_PARENT || q == SG_Q_BUFFER) return prog_.word_source(*this, q, base, stride, word);
        Engine::word_source(q, qi, base, stride, word);
    }

    void gather(uint32_t q, void *out, size_t bytes) override {
        if (q == SG_Q_VALUE || q == SG_Q_BUFFER || q == SG_Q_PARENT) return prog_.gather(*this, q, out, bytes);
        if (q == SG_Q_IN_DEGREE || q == SG_Q_IN_OFFSETS || q == SG_Q_IN_SLOTS) {
            if (!in_) fail(SG_EINVAL, "this engine keeps no in-edge store (SG_FLAG_IN_EDGES)");
            sync();
            return in_->gather_in(q - SG_Q_IN_DEGREE + SG_Q_DEGREE, out, bytes);
        }
        gather_graph(q, out, bytes);
    }
};

}  // namespace sg

// The C ABI handle (include/sgraph.h): one engine of any program.
struct sg_engine {
    std::unique_ptr<sg::Engine> impl;
};

#endif
