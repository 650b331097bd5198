// Device building blocks of the engine: wave64 primitives, the frontier
// expansion kernels (light + heavy tiers) and the per-vertex / per-edge
// functors that specialise them for each phase.
//
// Reference counterparts:
//   frontier append          groute::dev::Queue::append (groute/device/queue.cuh:84-88):
//                            there one atomicAdd per item; here one per wave (ballot).
//   neighbour scheduling     CTAWorkSchedulerNew::schedule (groute/device/cta_scheduler_hybrid.cuh:59-245):
//                            there block-owner election for huge lists + warp loops;
//                            here a wave stages 64 lists in LDS and walks their union,
//                            and lists longer than kHeavyDeg become fixed-size chunks
//                            taken by any wave (second kernel).
//   relax                    RelaxCTADB (framework/variants/sync_push_dd.cuh:180-268) +
//                            PushFunctorDB (framework/variants/push_functor.h:252-321)
#ifndef SG_KERNELS_H
#define SG_KERNELS_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sg_operators.h"
#include "../host/sg_rng.h"

namespace sg {

constexpr uint32_t kGap = 0xFFFFFFFFu;       // empty PMA slot (csr_graph.cuh:416, memset -1)
constexpr int kWave = 64;
constexpr int kBlock = 256;                  // threads per block of the expansion kernels
constexpr int kWavesPerBlock = kBlock / kWave;
constexpr uint32_t kHeavyDeg = 2048;         // lists longer than this go to the heavy tier
constexpr uint32_t kChunk = 1024;            // edges per heavy descriptor

// A kChunk-edge piece of an adjacency list: (start << kLenBits | len - 1),
// source, push value bits.
constexpr uint32_t kLenBits = 11;
struct HeavyDesc {
    uint64_t start_len;
    uint32_t u;
    uint32_t push;
};

// Graph as seen by kernels: PMA CSR with per-vertex gaps.
struct GraphView {
    uint32_t nv;
    const uint64_t *off;  // [nv + 1], off[nv] = capacity
    const uint32_t *deg;  // [nv]
    const uint32_t *col;  // [capacity], kGap in unused slots (HBM, or pinned host memory)
    // HBM edge cache (null when unused): copies of the lists of hot vertices.
    // A list start s >= cap addresses ccol[s - cap].
    const uint32_t *ccol;
    const uint64_t *coff;  // [nv] start of v's copy in ccol, ~0 = not cached
    uint64_t cap;
    uint8_t *hot;          // [nv] activations this snapshot (null when unused)
    // HBM copy of the live edges [0, mlive) of a col in pinned host memory, in
    // live-edge order (edge k of the merge path, vertex by vertex: no gaps),
    // ids packed at mbits bits each (packed_at), read by the tile sweeps; list
    // walks and updates use col. mlive = 0: none.
    const uint32_t *mirror;
    uint64_t mlive;
    // Gap-free packed copy of the live edges [mlive, E) in pinned host memory
    // (id k - mlive = edge k), read by the tile sweeps instead of col's gappy
    // slots; null: none.
    const uint32_t *tail;
    uint32_t mbits;  // bits per packed id (32: plain words)
};

// Id i of an array of b-bit ids packed back to back from bit 0 (b in [1, 32];
// the array holds one spare word past its last id).
__device__ __forceinline__ uint32_t packed_at(const uint32_t *p, unsigned long long i, uint32_t b) {
    const unsigned long long bit = i * b;
    const unsigned long long w = bit >> 5;
    const uint64_t x = ((uint64_t)p[w + 1] << 32) | p[w];
    const uint32_t v = (uint32_t)(x >> (bit & 31u));
    return b == 32 ? v : v & ((1u << b) - 1u);
}

// Where a relax walk reads u's list: its cached copy when there is one.
__device__ __forceinline__ uint64_t list_start(const GraphView &g, uint32_t u) {
    if (g.coff) {
        const uint64_t c = g.coff[u];
        if (c != ~0ull) return g.cap + c;
    }
    return g.off[u];
}
__device__ __forceinline__ uint32_t edge_at(const GraphView &g, uint64_t s) {
    return s < g.cap ? g.col[s] : g.ccol[s - g.cap];
}
// Activation count of the cache's hotness window (non-atomic and saturating,
// like the reference's hotness[0]++ in RelaxCTADB, sync_push_dd.cuh:196-230).
__device__ __forceinline__ void count_hot(const GraphView &g, uint32_t u) {
    if (g.hot) {
        const uint8_t h = g.hot[u];
        if (h < 255) g.hot[u] = h + 1;
    }
}

// Where a pass reads its work items and appends the vertices it activates.
struct PassIO {
    const uint32_t *in_list;  // explicit list, or null for the range [range_lo, range_lo + range_n)
    const uint32_t *in_cnt;   // device count of in_list
    uint32_t in_cap;          // capacity of in_list
    uint32_t range_lo, range_n;
    uint32_t *out_list;       // next frontier
    uint32_t *out_cnt;
    uint32_t out_cap;
    uint32_t *zero_cnt;       // frontier counter this pass clears for a later round (or null)
    HeavyDesc *heavy;         // descriptors this pass spills (processed by the next pass)
    uint32_t *heavy_cnt;
    const HeavyDesc *hin;     // descriptors spilled by the previous pass (or null)
    const uint32_t *hin_cnt;
    uint32_t *heavy_zero;     // descriptor counter this pass clears for a later round (or null)
    uint32_t heavy_cap;
    uint32_t budget;          // light-tier edges per 64-item batch of a wave
    uint32_t heavy_deg;       // lists longer than this always go to the heavy tier
    unsigned long long *stats;  // [0] edges (light), [1] vertices, [2] edges (heavy), [3] overflow
    uint32_t *log;            // round log (or null): round R records the counts its inputs had
    uint32_t round;           // R, for the log
    // Direction choice (PageRank with an in-edge store): ecnt_in holds the
    // out-edges of this round's frontier; above dense_thr the round is dense
    // and the vertex functor only takes the frontier's pushes (a pull pass
    // over the in-edges delivers them). Appended vertices add their out-degree
    // (fdeg) to ecnt_out. All null / 0 when the engine has no such choice.
    const unsigned long long *ecnt_in;
    unsigned long long *ecnt_out;
    unsigned long long *ecnt_zero;  // edge counter this pass clears for a later round (or null)
    const uint32_t *fdeg;
    unsigned long long dense_thr;
};
constexpr uint32_t kRoundLog = 256;  // entries of the round log ring

// ---- wave64 primitives -----------------------------------------------------
__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }
__device__ __forceinline__ uint64_t lanemask_lt() {
    const uint32_t l = lane_id();
    return l ? (~0ull >> (64 - l)) : 0ull;
}
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t x) {
    const uint32_t lane = lane_id();
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        const uint32_t y = __shfl_up(x, d, kWave);
        if (lane >= (uint32_t)d) x += y;
    }
    return x;
}
__device__ __forceinline__ unsigned long long wave_incl_scan(unsigned long long x) {
    const uint32_t lane = lane_id();
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        const unsigned long long y = __shfl_up(x, d, kWave);
        if (lane >= (uint32_t)d) x += y;
    }
    return x;
}
// Wave-aggregated append: one atomicAdd per wave. Every active lane of the
// wave must call it (pred may differ per lane).
__device__ __forceinline__ void wave_append(bool pred, uint32_t item, uint32_t *list, uint32_t *cnt,
                                            uint32_t cap, unsigned long long *overflow) {
    const uint64_t m = __ballot(pred);
    if (m == 0) return;
    const uint32_t leader = (uint32_t)__ffsll((unsigned long long)m) - 1;
    uint32_t base = 0;
    if (lane_id() == leader) base = atomicAdd(cnt, (uint32_t)__popcll(m));
    base = __shfl(base, (int)leader, kWave);
    if (pred) {
        const uint32_t pos = base + (uint32_t)__popcll(m & lanemask_lt());
        if (pos < cap) list[pos] = item;
        else atomicOr(overflow, 1ull);
    }
}
// Block-wide append: one global atomic per block and call instead of one per
// wave (per-wave atomics on one counter serialise across the chip). Every
// thread of a kBlock-thread block must call it; returns the block's count.
struct BlockAppend {
    uint32_t wsum[kWavesPerBlock];
    uint32_t base, total;
};
__device__ __forceinline__ uint32_t block_append(bool pred, uint32_t item, uint32_t *list, uint32_t *cnt, uint32_t cap,
                                                 unsigned long long *overflow, BlockAppend &sh) {
    const uint64_t m = __ballot(pred);
    const uint32_t w = threadIdx.x / kWave;
    if (lane_id() == 0) sh.wsum[w] = (uint32_t)__popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (int i = 0; i < kWavesPerBlock; ++i) {
            const uint32_t c = sh.wsum[i];
            sh.wsum[i] = t;  // exclusive offset of each wave
            t += c;
        }
        sh.total = t;
        sh.base = t ? atomicAdd(cnt, t) : 0u;
    }
    __syncthreads();
    if (pred) {
        const uint32_t pos = sh.base + sh.wsum[w] + (uint32_t)__popcll(m & lanemask_lt());
        if (pos < cap) list[pos] = item;
        else atomicOr(overflow, 1ull);
    }
    const uint32_t total = sh.total;
    __syncthreads();  // the LDS words are reused by the next call
    return total;
}

__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long x) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) x += __shfl_xor(x, d, kWave);
    return x;
}

// 64-bit minimum on integers (the generic min template goes through double).
__host__ __device__ __forceinline__ unsigned long long min_u64(unsigned long long a, unsigned long long b) {
    return a < b ? a : b;
}

// Coarse filter over a vertex bitmap: one bit per 2^shift vertices with
// shift = max(6, ceil(log2 nv) - 16), so the filter has at most 65536 bits
// (kFilterWords words) and a block can keep a copy in LDS.
constexpr uint32_t kFilterWords = 2048;
__host__ __device__ inline uint32_t coarse_shift_for(uint32_t nv) {
    uint32_t s = 6;
    while ((((uint64_t)nv + (1ull << s) - 1) >> s) > 32ull * kFilterWords) ++s;
    return s;
}
__host__ __device__ inline uint32_t coarse_words(uint32_t nv, uint32_t shift) {
    return (uint32_t)(((((uint64_t)nv + (1ull << shift) - 1) >> shift) + 31) / 32);
}
__device__ __forceinline__ void coarse_mark(uint32_t *coarse, uint32_t shift, uint32_t v) {
    atomicOr(&coarse[v >> (shift + 5)], 1u << ((v >> shift) & 31));
}
__device__ __forceinline__ bool coarse_test(const uint32_t *coarse, uint32_t shift, uint32_t v) {
    return (coarse[v >> (shift + 5)] >> ((v >> shift) & 31)) & 1u;
}

template <class T> __device__ __forceinline__ uint32_t to_bits(T x);
template <> __device__ __forceinline__ uint32_t to_bits<uint32_t>(uint32_t x) { return x; }
template <> __device__ __forceinline__ uint32_t to_bits<float>(float x) { return __float_as_uint(x); }
template <class T> __device__ __forceinline__ T from_bits(uint32_t x);
template <> __device__ __forceinline__ uint32_t from_bits<uint32_t>(uint32_t x) { return x; }
template <> __device__ __forceinline__ float from_bits<float>(uint32_t x) { return __uint_as_float(x); }

__device__ __forceinline__ uint64_t ld_relaxed(const unsigned long long *p) {
    return __hip_atomic_load(const_cast<unsigned long long *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Plain (cacheable) load for pre-tests of values that only decrease during a
// phase: a stale copy is never below the current value, so a pre-test that
// passes on the current value also passes on the copy.
__device__ __forceinline__ uint64_t ld_peek(const unsigned long long *p) { return *p; }

// ---- per-wave output staging -------------------------------------------------
// Activated vertices are collected in a per-wave LDS buffer and appended to
// the global frontier kOutStage at a time: one atomicAdd per flush instead of
// one per wave-iteration, and coalesced stores.
constexpr uint32_t kOutStage = 512;
constexpr uint32_t kWaveBudget = 1024;  // default light-tier edges per 64-item batch of a wave
constexpr int kIlp = 8;                 // edges per lane per iteration

struct OutStage {
    uint32_t *buf;              // this wave's LDS slice [cap]
    uint32_t n;                 // wave-uniform fill
    uint32_t cap = kOutStage;   // slice size (>= kWave)
    unsigned long long fedges = 0;  // per lane: out-degrees of the items it appended (io.fdeg)
};

__device__ __forceinline__ void stage_flush(OutStage &st, const PassIO &io) {
    if (st.n == 0) return;
    uint32_t base = 0;
    if (lane_id() == 0) base = atomicAdd(io.out_cnt, st.n);
    base = __shfl(base, 0, kWave);
    for (uint32_t i = lane_id(); i < st.n; i += kWave) {
        const uint32_t pos = base + i;
        if (pos < io.out_cap) io.out_list[pos] = st.buf[i];
        else atomicOr(io.stats + 3, 1ull);
    }
    st.n = 0;
}

__device__ __forceinline__ void stage_push(bool pred, uint32_t item, OutStage &st, const PassIO &io) {
    const uint64_t m = __ballot(pred);
    if (m == 0) return;
    if (io.fdeg && pred) st.fedges += io.fdeg[item];
    const uint32_t c = (uint32_t)__popcll(m);
    if (st.n + c > st.cap) stage_flush(st, io);
    if (pred) st.buf[st.n + (uint32_t)__popcll(m & lanemask_lt())] = item;
    st.n += c;
}

// One atomic per wave: the out-edges of what the wave appended (io.ecnt_out).
__device__ __forceinline__ void stage_count_flush(OutStage &st, const PassIO &io) {
    if (!io.ecnt_out) return;
    const unsigned long long t = wave_sum_u64(st.fedges);
    if (lane_id() == 0 && t) atomicAdd(io.ecnt_out, t);
    st.fedges = 0;
}

// Per-wave counters, reduced per block at the end of a kernel.
struct WaveStats {
    unsigned long long edges = 0, verts = 0, hedges = 0, cedges = 0;
};

// LDS of the light tier (per block).
struct LightShared {
    uint64_t start[kWavesPerBlock][kWave];
    uint32_t excl[kWavesPerBlock][kWave];
    uint32_t src[kWavesPerBlock][kWave];
    uint32_t push[kWavesPerBlock][kWave];
    uint32_t out[kWavesPerBlock][kOutStage];
    unsigned long long red[kWavesPerBlock][4];
};

// Owner of edge k among the 64 staged lists: the largest j with excl[j] <= k
// (six unrolled LDS probes, so the kIlp searches of a lane interleave).
__device__ __forceinline__ uint32_t staged_owner(const uint32_t *excl, uint32_t k) {
    uint32_t lo = 0;
#pragma unroll
    for (uint32_t step = kWave / 2; step > 0; step >>= 1)
        if (excl[lo + step] <= k) lo += step;
    return lo;
}

// Edge loop of one wave over `total` edges whose owners are staged in LDS,
// kIlp edges per lane per step: all target ids are loaded, then all functor
// peeks, then the updates.
template <class EFn>
__device__ __forceinline__ void walk_staged(const GraphView &g, const PassIO &io, const EFn &efn, uint32_t total,
                                            const uint32_t *s_excl, const uint64_t *s_start, const uint32_t *s_src,
                                            const uint32_t *s_push, OutStage &st) {
    const uint32_t lane = lane_id();
    for (uint32_t k0 = 0; k0 < total; k0 += kIlp * kWave) {
        uint32_t v[kIlp], src[kIlp], push[kIlp];
        uint64_t peek[kIlp];
        bool act[kIlp];
#pragma unroll
        for (int j = 0; j < kIlp; ++j) {
            const uint32_t k = k0 + j * kWave + lane;
            act[j] = k < total;
            const uint32_t o = staged_owner(s_excl, act[j] ? k : 0);
            src[j] = s_src[o];
            push[j] = s_push[o];
            v[j] = act[j] ? edge_at(g, s_start[o] + (k - s_excl[o])) : kGap;
        }
#pragma unroll
        for (int j = 0; j < kIlp; ++j) peek[j] = (v[j] != kGap) ? efn.load(v[j]) : 0;
        bool enq[kIlp];
#pragma unroll
        for (int j = 0; j < kIlp; ++j) enq[j] = v[j] != kGap && efn.apply(src[j], v[j], push[j], peek[j]);
#pragma unroll
        for (int j = 0; j < kIlp; ++j) stage_push(enq[j], EFn::kAppendSource ? src[j] : v[j], st, io);
    }
}

// Power of two in [1, kWave] at least ceil(n / nwaves).
__device__ __forceinline__ uint32_t light_width(uint32_t n, uint64_t nwaves) {
    const uint64_t q = (n + nwaves - 1) / nwaves;
    uint32_t per = 1;
    while (per < (uint32_t)kWave && per < q) per <<= 1;
    return per;
}

// Light tier body: waves [gwave, +nwaves) take up to 64 work items at a time
// (light_width), run the vertex functor, and walk the scheduled lists that
// fit the wave budget; longer lists, and lists past the budget, are cut into
// kChunk-edge descriptors for the heavy tier, so no wave carries more than
// budget + heavy_deg edges per batch.
template <class VFn, class EFn>
__device__ __forceinline__ void light_body(const GraphView &g, const PassIO &io, const VFn &vfn, const EFn &efn,
                                           uint32_t n, uint64_t gwave, uint64_t nwaves, LightShared &sh, OutStage &st,
                                           WaveStats &ws) {
    const uint32_t w = threadIdx.x / kWave, lane = lane_id();
    // frontier vertices per wave batch: 64 when the frontier fills every wave,
    // fewer (a power of two, down to 1) when it is small, so that its lists
    // spread over all waves instead of a few waves walking 64 lists each
    const uint32_t per = light_width(n, nwaves);
    // dense round: the functor records each frontier vertex's push for the
    // pull pass and no list is walked here
    bool dense = false;
    if constexpr (VFn::kDense) dense = io.dense_thr && io.ecnt_in && *io.ecnt_in > io.dense_thr;
    for (uint64_t base = gwave * per; base < n; base += nwaves * per) {
        const uint64_t idx = base + lane;
        uint32_t u = 0, d = 0, push = 0;
        uint64_t start = 0;
        bool self = false, sched = false;
        if (lane < per && idx < n) {
            u = io.in_list ? io.in_list[idx] : io.range_lo + (uint32_t)idx;
            const uint32_t du = g.deg[u];
            sched = vfn(u, du, push, self);
            if constexpr (VFn::kDense) {
                if (dense) {
                    if (sched && du) {
                        vfn.take(u, push);
                        count_hot(g, u);  // a pusher: its list is hot for the HBM edge cache
                    }
                    sched = false;
                }
            }
            if (sched && du) {
                d = du;
                start = list_start(g, u);
                count_hot(g, u);
            }
        }
        stage_push(self, u, st, io);
        ws.verts += sched ? 1 : 0;
        const uint32_t incl0 = wave_incl_scan(d);
        const bool heavy = d > io.heavy_deg || (d > 0 && incl0 > io.budget);
        const uint32_t nch = heavy ? (d + kChunk - 1) / kChunk : 0;
        const uint32_t nch_incl = wave_incl_scan(nch);
        const uint32_t nch_tot = __shfl(nch_incl, kWave - 1, kWave);
        if (nch_tot) {
            uint32_t hb = 0;
            if (lane == 0) hb = atomicAdd(io.heavy_cnt, nch_tot);
            hb = __shfl(hb, 0, kWave) + nch_incl - nch;
            if (heavy) {
                if (hb + nch <= io.heavy_cap) {
                    for (uint32_t c = 0; c < nch; ++c) {
                        const uint64_t len = (c + 1 < nch) ? kChunk : d - c * kChunk;
                        HeavyDesc hd;
                        hd.start_len = ((start + (uint64_t)c * kChunk) << kLenBits) | (len - 1);
                        hd.u = u;
                        hd.push = push;
                        io.heavy[hb + c] = hd;
                    }
                } else {
                    atomicOr(io.stats + 3, 2ull);
                }
            }
        }
        if (heavy) d = 0;
        const uint32_t incl = wave_incl_scan(d);
        const uint32_t total = __shfl(incl, kWave - 1, kWave);
        sh.start[w][lane] = start;
        sh.excl[w][lane] = incl - d;
        sh.src[w][lane] = u;
        sh.push[w][lane] = push;
        __builtin_amdgcn_wave_barrier();
        walk_staged(g, io, efn, total, sh.excl[w], sh.start[w], sh.src[w], sh.push[w], st);
        ws.edges += total;
        if (g.coff) ws.cedges += wave_sum_u64(start >= g.cap ? d : 0u);
        __builtin_amdgcn_wave_barrier();
    }
}

// Heavy tier body: one wave per kChunk-edge descriptor.
template <class EFn>
__device__ __forceinline__ void heavy_body(const GraphView &g, const PassIO &io, const EFn &efn, uint32_t nh,
                                           uint64_t gwave, uint64_t nwaves, OutStage &st, WaveStats &ws) {
    const uint32_t lane = lane_id();
    for (uint64_t i = gwave; i < nh; i += nwaves) {
        const HeavyDesc hd = io.hin[i];
        const uint64_t hstart = hd.start_len >> kLenBits;
        const uint32_t hlen = (uint32_t)(hd.start_len & ((1u << kLenBits) - 1)) + 1;
        for (uint32_t k0 = 0; k0 < hlen; k0 += kIlp * kWave) {
            uint32_t v[kIlp];
            uint64_t peek[kIlp];
            bool enq[kIlp];
#pragma unroll
            for (int j = 0; j < kIlp; ++j) {
                const uint32_t k = k0 + j * kWave + lane;
                v[j] = k < hlen ? edge_at(g, hstart + k) : kGap;
            }
#pragma unroll
            for (int j = 0; j < kIlp; ++j) peek[j] = v[j] != kGap ? efn.load(v[j]) : 0;
#pragma unroll
            for (int j = 0; j < kIlp; ++j) enq[j] = v[j] != kGap && efn.apply(hd.u, v[j], hd.push, peek[j]);
#pragma unroll
            for (int j = 0; j < kIlp; ++j) stage_push(enq[j], EFn::kAppendSource ? hd.u : v[j], st, io);
        }
        ws.hedges += hlen;
        if (hstart >= g.cap) ws.cedges += hlen;
    }
}

// One atomic per block for the pass counters.
__device__ __forceinline__ void block_stats(const WaveStats &ws, LightShared &sh, unsigned long long *stats) {
    const uint32_t w = threadIdx.x / kWave;
    const unsigned long long v = wave_sum_u64(ws.verts);
    if (lane_id() == 0) {
        sh.red[w][0] = ws.edges;
        sh.red[w][1] = v;
        sh.red[w][2] = ws.hedges;
        sh.red[w][3] = ws.cedges;  // lane 0 holds the wave's cached-edge total
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        unsigned long long t = 0;
        for (int i = 0; i < kWavesPerBlock; ++i) t += sh.red[i][threadIdx.x];
        // [0] edges, [1] vertices, [2] spilled-chunk edges, [5] edges read from the cache
        const int slot = threadIdx.x == 3 ? 5 : (int)threadIdx.x;
        if (t) atomicAdd(stats + slot, t);
    }
}

// ---- one pass ---------------------------------------------------------------------
// Light items of this pass, then the descriptors the previous pass spilled;
// the edge updates of every operator are order-insensitive (min-relaxations,
// residual pushes), so the spilled lists of round R-1 are walked in round R
// and whatever they activate joins round R+1.
template <class VFn, class EFn>
__global__ void __launch_bounds__(kBlock) k_round(GraphView g, PassIO io, VFn vfn, EFn efn) {
    __shared__ LightShared sh;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        if (io.zero_cnt) *io.zero_cnt = 0;
        if (io.heavy_zero) *io.heavy_zero = 0;
        if (io.ecnt_zero) *io.ecnt_zero = 0;
        if (io.log) {  // the inputs of round R are what round R - 1 produced
            const uint32_t i = 2 * ((io.round + kRoundLog - 1) % kRoundLog);
            io.log[i] = *io.in_cnt;
            io.log[i + 1] = io.hin_cnt ? *io.hin_cnt : 0u;
        }
    }
    // a list never holds more than its capacity (appends past it are dropped and flagged)
    const uint32_t n = io.in_list ? min(*io.in_cnt, io.in_cap) : io.range_n;
    const uint32_t nh = io.hin ? min(*io.hin_cnt, io.heavy_cap) : 0;
    const uint32_t w = threadIdx.x / kWave;
    const uint64_t gwave = (uint64_t)blockIdx.x * kWavesPerBlock + w, nwaves = (uint64_t)gridDim.x * kWavesPerBlock;
    OutStage st{sh.out[w], 0};
    WaveStats ws;
    light_body(g, io, vfn, efn, n, gwave, nwaves, sh, st, ws);
    heavy_body(g, io, efn, nh, gwave, nwaves, st, ws);
    stage_flush(st, io);
    stage_count_flush(st, io);
    block_stats(ws, sh, io.stats);
}

// ---- vertex functors ---------------------------------------------------------
// Signature: bool (uint32_t u, uint32_t deg, uint32_t &push_bits, bool &enqueue_self)
// returns whether u's out-edges are expanded with push_bits.

// Data-driven relax round: commit the buffer, push (RelaxCTADB, sync_push_dd.cuh:196-230).
template <class Op>
struct VCombine {
    static constexpr bool kDense = false;
    Op op;
    typename Op::value_t *value;
    typename Op::buffer_t *buffer;
    __device__ bool operator()(uint32_t u, uint32_t deg, uint32_t &push, bool &self) const {
        self = false;
        auto p = op.CombineValueBuffer(u, &value[u], &buffer[u], deg);
        push = to_bits<typename Op::push_t>(p.value);
        return p.scheduled;
    }
};

// Re-push of committed values scaled by type: PR cancel (-1) / compensate (+1)
// (RelaxCTADBAmend, sync_push_dd.cuh:375-470).
template <class Op>
struct VAmend {
    static constexpr bool kDense = false;
    Op op;
    int type;
    typename Op::value_t *value;
    typename Op::buffer_t *buffer;
    __device__ bool operator()(uint32_t u, uint32_t deg, uint32_t &push, bool &self) const {
        self = false;
        auto p = op.CombineValueBufferAmend(u, type, &value[u], &buffer[u], deg);
        push = to_bits<typename Op::push_t>(p.value);
        return p.scheduled;
    }
};

// Traversal boundary sweep after invalidation: every vertex with a finite best
// value pushes it, edges only act on invalidated targets; invalidated vertices
// that are active on their own (CC: own label) join the frontier
// (the reference pushes to every target: RelaxCTADB_all_vertices,
// sync_push_dd.cuh:1470-1557).
template <class Op>
struct VSweep {
    static constexpr bool kDense = false;
    Op op;
    typename Op::value_t *value;
    typename Op::buffer_t *buffer;
    const uint32_t *inval;   // bitmap
    uint32_t *stamp;
    uint32_t stamp_val;
    __device__ bool operator()(uint32_t u, uint32_t, uint32_t &push, bool &self) const {
        const uint64_t b = ld_relaxed(&buffer[u]);
        self = false;
        if ((inval[u >> 5] >> (u & 31)) & 1u) {
            if (op.IsActiveNode(u, b, value[u])) self = atomicExch(&stamp[u], stamp_val) != stamp_val;
        }
        push = op.PushValue(b);
        return push != api::kInf;
    }
};

// Invalidation propagation: expand every invalidated vertex.
struct VAll {
    static constexpr bool kDense = false;
    __device__ bool operator()(uint32_t u, uint32_t, uint32_t &push, bool &self) const {
        self = false;
        push = u;
        return true;
    }
};

// ---- edge functors -------------------------------------------------------------
// load(v) peeks the destination state (issued for kIlp edges at once);
// apply(u, v, push, peek) updates it and returns whether v joins the output
// (u, the walked list's owner, when kAppendSource: a pull over in-lists).

template <class Op>
struct ERelax {
    static constexpr bool kAppendSource = false;
    Op op;
    typename Op::buffer_t *buffer;
    uint32_t *stamp;
    uint32_t stamp_val;
    uint32_t wmode;
    uint64_t wseed;
    __device__ ERelax with_round(uint32_t stamp_next) const {
        ERelax e = *this;
        e.stamp_val = stamp_next;
        return e;
    }
    __device__ uint64_t load(uint32_t v) const {
        if constexpr (Op::kTraversal) return ld_peek(&buffer[v]);
        else return 0;
    }
    __device__ bool apply(uint32_t u, uint32_t v, uint32_t push, uint64_t peek) const {
        const uint32_t w = Op::kWeighted ? sg_weight(wmode, wseed, u, v) : 0u;
        const auto p = from_bits<typename Op::push_t>(push);
        if constexpr (Op::kTraversal) {
            if (!op.MayImprove(u, w, p, peek)) return false;
        }
        if (!op.AccumulateBuffer(u, v, w, &buffer[v], p)) return false;
        return atomicExch(&stamp[v], stamp_val) != stamp_val;
    }
};

// Relax restricted to invalidated targets (boundary sweep). The probe reads a
// coarse bitmap (one bit per 64 vertices, small enough to stay in the vector
// L1) before the fine one, so most edges never touch L2 for the test.
template <class Op>
struct ERelaxInval {
    static constexpr bool kAppendSource = false;
    ERelax<Op> r;
    const uint32_t *inval;
    const uint32_t *coarse;  // coarse filter of inval (the sweep tests it first, in LDS)
    uint32_t cshift;
    __device__ ERelaxInval with_round(uint32_t stamp_next) const {
        ERelaxInval e = *this;
        e.r.stamp_val = stamp_next;
        return e;
    }
    __device__ uint32_t load(uint32_t v) const { return inval[v >> 5]; }
    __device__ bool test(uint32_t v, uint32_t word) const { return (word >> (v & 31)) & 1u; }
    // v is invalidated: relax it from u
    __device__ bool apply(uint32_t u, uint32_t v, uint32_t push) const { return r.apply(u, v, push, r.load(v)); }
};

// Reset tree children of u (AccumulateBuffer_del via PushFunctorDB_Del,
// push_functor.h:960-1034), claiming each child once through the bitmap.
template <class Op>
struct EInvalidate {
    static constexpr bool kAppendSource = false;
    Op op;
    typename Op::value_t *value;
    typename Op::buffer_t *buffer;
    uint32_t *inval;
    uint32_t *coarse;
    uint32_t cshift;
    __device__ EInvalidate with_round(uint32_t) const { return *this; }
    __device__ uint64_t load(uint32_t v) const { return ld_relaxed(&buffer[v]); }
    __device__ bool apply(uint32_t u, uint32_t v, uint32_t, uint64_t b) const {
        if (Op::Parent(b) != u) return false;
        const uint32_t bit = 1u << (v & 31);
        if (atomicOr(&inval[v >> 5], bit) & bit) return false;
        coarse_mark(coarse, cshift, v);
        typename Op::buffer_t bb = b;
        op.AccumulateBuffer_del(u, v, &bb, &value[v]);
        buffer[v] = bb;
        return true;  // counted by the frontier it joins
    }
};


}  // namespace sg

#endif
