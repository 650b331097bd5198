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

struct HeavyDesc {
    uint64_t start;
    uint32_t len;
    uint32_t u;
    uint32_t push;  // push value bits
    uint32_t pad;
};

// Graph as seen by kernels: PMA CSR with per-vertex gaps.
struct GraphView {
    uint32_t nv;
    const uint64_t *off;  // [nv + 1], off[nv] = capacity
    const uint32_t *deg;  // [nv]
    const uint32_t *col;  // [capacity], kGap in unused slots
};

// Where a pass reads its work items and appends the vertices it activates.
struct PassIO {
    const uint32_t *in_list;  // explicit list, or null for the range [range_lo, range_lo + range_n)
    const uint32_t *in_cnt;   // device count of in_list
    uint32_t in_cap;          // capacity of in_list
    uint32_t range_lo, range_n;
    uint32_t *out_list;       // next frontier
    uint32_t *out_cnt;
    uint32_t out_cap;
    uint32_t *zero_cnt;       // a counter this pass clears for the pass after next (or null)
    HeavyDesc *heavy;
    uint32_t *heavy_cnt;      // appended by the light kernel, read by the heavy kernel
    uint32_t *heavy_zero;     // cleared by the heavy kernel (the next pass's heavy_cnt)
    uint32_t heavy_cap;
    unsigned long long *stats;  // [0] edges (light), [1] vertices, [2] edges (heavy), [3] overflow
};

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
__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long x) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) x += __shfl_xor(x, d, kWave);
    return x;
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

// ---- light tier ------------------------------------------------------------
// Each wave takes 64 work items, runs the vertex functor on them, stages the
// scheduled adjacency lists (start, prefix, source, push) in LDS and walks
// their concatenation 64 edges at a time. Lists longer than kHeavyDeg are cut
// into kChunk-edge descriptors for k_expand_heavy.
template <class VFn, class EFn>
__global__ void __launch_bounds__(kBlock) k_expand(GraphView g, PassIO io, VFn vfn, EFn efn) {
    __shared__ uint64_t s_start[kWavesPerBlock][kWave];
    __shared__ uint32_t s_excl[kWavesPerBlock][kWave];
    __shared__ uint32_t s_src[kWavesPerBlock][kWave];
    __shared__ uint32_t s_push[kWavesPerBlock][kWave];

    if (blockIdx.x == 0 && threadIdx.x == 0 && io.zero_cnt) *io.zero_cnt = 0;
    // a list never holds more than its capacity (appends past it are dropped and flagged)
    const uint32_t n = io.in_list ? min(*io.in_cnt, io.in_cap) : io.range_n;
    const uint32_t w = threadIdx.x / kWave, lane = lane_id();
    const uint64_t gwave = (uint64_t)blockIdx.x * kWavesPerBlock + w;
    const uint64_t nwaves = (uint64_t)gridDim.x * kWavesPerBlock;
    unsigned long long n_edges = 0, n_verts = 0;

    for (uint64_t base = gwave * kWave; base < n; base += nwaves * kWave) {
        const uint64_t idx = base + lane;
        uint32_t u = 0, d = 0, push = 0;
        uint64_t start = 0;
        bool self = false, sched = false;
        if (idx < n) {
            u = io.in_list ? io.in_list[idx] : io.range_lo + (uint32_t)idx;
            const uint32_t du = g.deg[u];
            sched = vfn(u, du, push, self);
            if (sched && du) { d = du; start = g.off[u]; }
        }
        wave_append(self, u, io.out_list, io.out_cnt, io.out_cap, io.stats + 3);
        n_verts += sched ? 1 : 0;
        if (d > kHeavyDeg) {
            const uint32_t nch = (d + kChunk - 1) / kChunk;
            const uint32_t hb = atomicAdd(io.heavy_cnt, nch);
            if (hb + nch <= io.heavy_cap) {
                for (uint32_t c = 0; c < nch; ++c) {
                    HeavyDesc hd;
                    hd.start = start + (uint64_t)c * kChunk;
                    hd.len = (c + 1 < nch) ? kChunk : d - c * kChunk;
                    hd.u = u;
                    hd.push = push;
                    hd.pad = 0;
                    io.heavy[hb + c] = hd;
                }
            } else {
                atomicOr(io.stats + 3, 2ull);
            }
            d = 0;
        }
        const uint32_t incl = wave_incl_scan(d);
        const uint32_t total = __shfl(incl, kWave - 1, kWave);
        s_start[w][lane] = start;
        s_excl[w][lane] = incl - d;
        s_src[w][lane] = u;
        s_push[w][lane] = push;
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        for (uint32_t k0 = 0; k0 < total; k0 += kWave) {
            const uint32_t k = k0 + lane;
            bool enq = false;
            uint32_t v = 0;
            if (k < total) {
                uint32_t lo = 0, hi = kWave;
                while (hi - lo > 1) {
                    const uint32_t mid = (lo + hi) >> 1;
                    if (s_excl[w][mid] <= k) lo = mid; else hi = mid;
                }
                v = g.col[s_start[w][lo] + (k - s_excl[w][lo])];
                enq = efn(s_src[w][lo], v, s_push[w][lo]);
            }
            wave_append(enq, v, io.out_list, io.out_cnt, io.out_cap, io.stats + 3);
        }
        n_edges += total;
        __builtin_amdgcn_wave_barrier();
    }
    n_verts = wave_sum_u64(n_verts);
    if (lane == 0) {
        if (n_edges) atomicAdd(io.stats + 0, n_edges);
        if (n_verts) atomicAdd(io.stats + 1, n_verts);
    }
}

// ---- heavy tier: one wave per kChunk-edge descriptor -------------------------
template <class EFn>
__global__ void __launch_bounds__(kBlock) k_expand_heavy(GraphView g, PassIO io, EFn efn) {
    if (blockIdx.x == 0 && threadIdx.x == 0 && io.heavy_zero) *io.heavy_zero = 0;
    const uint32_t nh = min(*io.heavy_cnt, io.heavy_cap);
    const uint32_t w = threadIdx.x / kWave, lane = lane_id();
    const uint64_t gwave = (uint64_t)blockIdx.x * kWavesPerBlock + w;
    const uint64_t nwaves = (uint64_t)gridDim.x * kWavesPerBlock;
    unsigned long long n_edges = 0;
    for (uint64_t i = gwave; i < nh; i += nwaves) {
        const HeavyDesc hd = io.heavy[i];
        for (uint32_t k0 = 0; k0 < hd.len; k0 += kWave) {
            const uint32_t k = k0 + lane;
            bool enq = false;
            uint32_t v = 0;
            if (k < hd.len) {
                v = g.col[hd.start + k];
                enq = efn(hd.u, v, hd.push);
            }
            wave_append(enq, v, io.out_list, io.out_cnt, io.out_cap, io.stats + 3);
        }
        n_edges += hd.len;
    }
    if (lane == 0 && n_edges) atomicAdd(io.stats + 2, n_edges);
}

// ---- vertex functors ---------------------------------------------------------
// Signature: bool (uint32_t u, uint32_t deg, uint32_t &push_bits, bool &enqueue_self)
// returns whether u's out-edges are expanded with push_bits.

// Data-driven relax round: commit the buffer, push (RelaxCTADB, sync_push_dd.cuh:196-230).
template <class Op>
struct VCombine {
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
    __device__ bool operator()(uint32_t u, uint32_t, uint32_t &push, bool &self) const {
        self = false;
        push = u;
        return true;
    }
};

// ---- edge functors -------------------------------------------------------------
// Signature: bool (uint32_t u, uint32_t v, uint32_t push_bits) -> append v.

template <class Op>
struct ERelax {
    Op op;
    typename Op::buffer_t *buffer;
    uint32_t *stamp;
    uint32_t stamp_val;
    uint32_t wmode;
    uint64_t wseed;
    __device__ bool operator()(uint32_t u, uint32_t v, uint32_t push) const {
        if (v == kGap) return false;
        const uint32_t w = Op::kWeighted ? sg_weight(wmode, wseed, u, v) : 0u;
        if (!op.AccumulateBuffer(u, v, w, &buffer[v], from_bits<typename Op::push_t>(push))) return false;
        return atomicExch(&stamp[v], stamp_val) != stamp_val;
    }
};

// Relax restricted to invalidated targets (boundary sweep).
template <class Op>
struct ERelaxInval {
    ERelax<Op> r;
    const uint32_t *inval;
    __device__ bool operator()(uint32_t u, uint32_t v, uint32_t push) const {
        if (v == kGap || !((inval[v >> 5] >> (v & 31)) & 1u)) return false;
        return r(u, v, push);
    }
};

// Reset tree children of u (AccumulateBuffer_del via PushFunctorDB_Del,
// push_functor.h:960-1034), claiming each child once through the bitmap.
template <class Op>
struct EInvalidate {
    Op op;
    typename Op::value_t *value;
    typename Op::buffer_t *buffer;
    uint32_t *inval;
    unsigned long long *count;
    __device__ bool operator()(uint32_t u, uint32_t v, uint32_t) const {
        if (v == kGap) return false;
        const uint64_t b = ld_relaxed(&buffer[v]);
        if (Op::Parent(b) != u) return false;
        const uint32_t bit = 1u << (v & 31);
        if (atomicOr(&inval[v >> 5], bit) & bit) return false;
        typename Op::buffer_t bb = b;
        op.AccumulateBuffer_del(u, v, &bb, &value[v]);
        buffer[v] = bb;
        atomicAdd(count, 1ull);
        return true;
    }
};

}  // namespace sg

#endif
