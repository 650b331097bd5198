// Pull rounds of PageRank over the in-edge store (a second PMA of the same
// graph with every edge reversed, kept up to date by the same batches).
//
// A round R takes the residual of every frontier vertex u (value[u] += x,
// push p[u] = alpha * x / outdeg(u)) and delivers p[u] to every out-neighbour.
// The push form walks u's out-list with one f32 atomic add per edge; on R-MAT
// graphs the frontier is a few hub vertices that own most of the edges for
// tens of rounds, so those rounds are atomic-bound. The pull form walks all
// in-lists instead: r[v] += sum of p[u] over in-neighbours u in the frontier.
// Every destination is owned by one wave (or, for a list cut by a tile
// boundary, finished by a fix-up pass), so no global atomic is issued per
// edge. The engine picks the form per round from the frontier's out-edge count
// (direction-optimising, as Beamer's BFS), on the device, so rounds stay
// queued ahead of the host.
//
// Reference: the pull functor shape of include/framework/variants/pull_functor.h:22-369
// and sync_pull_dd.cuh:1-357 (dead code there: its schedule calls do not match
// cta_scheduler_hybrid.cuh:63), and PageRank's operator (samples/hybrid_pr/hybrid_pr.cu:66-136).
#ifndef SG_PULL_H
#define SG_PULL_H

#include "sg_pma.h"

namespace sg {

// The pushes a dense round takes: bit u of the round's frontier bitmap and
// the push bits pv[u]. A pull pass tests the bitmap for every in-edge - one
// bit per vertex, small enough to stay in the caches (16 MiB at 2^27
// vertices) - and reads pv only for the in-neighbours in the frontier, so the
// per-edge random read no longer touches a word per vertex in HBM. Two
// bitmaps alternate between rounds; round R's is cleared before round R (it
// last served round R - 2).
struct PullTake {
    uint32_t *bits;           // [ceil(nv / 32)] this round's frontier
    uint32_t *pv;             // [nv] push bits of the frontier's vertices
    const uint32_t *relabel;  // the id the in-lists record for a vertex (null: its own)
    __device__ void take(uint32_t u, uint32_t push) const {
        const uint32_t r = relabel ? relabel[u] : u;
        pv[r] = push;
        atomicOr(&bits[r >> 5], 1u << (r & 31));
    }
};

// Vertex functor of PageRank rounds (CombineValueBuffer, hybrid_pr.cu:66-85).
// In a dense round take() records the push instead of walking the list.
struct VPageRank {
    static constexpr bool kDense = true;
    api::PageRank op;
    float *value;
    float *buffer;
    PullTake t;
    __device__ bool operator()(uint32_t u, uint32_t deg, uint32_t &push, bool &self) const {
        self = false;
        auto p = op.CombineValueBuffer(u, &value[u], &buffer[u], deg);
        push = __float_as_uint(p.value);
        return p.scheduled;
    }
    __device__ void take(uint32_t u, uint32_t push) const { t.take(u, push); }
};

// Cancel / compensate of a PageRank snapshot as a take for a pull pass: each
// touched source records the push its CombineValueBufferAmend makes (type -1
// over the old out-degree before the commit, +1 over the new one after it).
// The pull over the in-edge store as it stands then delivers it along exactly
// the lists the push form walks (old in-lists before the store's update, new
// ones after), one read of the store instead of the hub out-lists twice.
__global__ void k_pr_amend_take(const uint32_t *list, const uint32_t *cnt, uint32_t cap, api::PageRank op, int type,
                                float *value, float *buffer, const uint32_t *deg, PullTake t) {
    const uint32_t n = min(*cnt, cap);
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const uint32_t u = list[i];
        const auto p = op.CombineValueBufferAmend(u, type, &value[u], &buffer[u], deg[u]);
        if (p.scheduled) t.take(u, __float_as_uint(p.value));
    }
}

// State a pull pass reads and updates.
struct PullPR {
    const uint32_t *bits;          // this round's frontier bitmap
    const float *pv;               // pushes of the frontier's vertices
    float *res;                    // residual (the operator's buffer)
    float *accg;                   // partial sums of lists cut by tile boundaries (zero between rounds)
    uint32_t *stamp;               // round stamps (frontier de-duplication)
    uint32_t round;                // R
    float eps;                     // activity threshold (IsActiveNode / IsActiveNode_INC)
    // Store in host memory with an HBM mirror: the first tile whose edges
    // reach past the mirror (null: no split). Tiles before it read HBM only,
    // the rest mostly the host link; host_waves64 / 64 of the waves take the
    // latter, the others the former, so that both run at once instead of
    // every wave stalling on the host link in turn.
    const unsigned long long *tsplit;
    uint32_t host_waves64;
};

__device__ __forceinline__ bool pull_dense(const PassIO &io, int force) {
    return force || (io.dense_thr && io.ecnt_in && *io.ecnt_in > io.dense_thr);
}

// Residual r of v is final: v joins the next frontier when |r| > eps
// (both signs: incremental snapshots leave negative residuals).
__device__ __forceinline__ bool pull_finish(const PullPR &p, uint32_t v, float total) {
    float r = p.res[v];
    if (total != 0.0f) {
        r += total;
        p.res[v] = r;
    }
    return fabsf(r) > p.eps && atomicExch(&p.stamp[v], p.round + 1) != p.round + 1;
}

// Walk the in-edge tiles [t_begin, t_end) of the store rg (edges read from
// colp[slot - slot_lo]: the store itself or one streamed segment); only
// destinations in [dst_lo, dst_hi) are handled. One wave per tile: the
// tile's in-neighbour ids are loaded coalesced (sweep_tile), their tagged
// pushes gathered, summed per destination in LDS, and the destinations
// strictly inside the tile finished at once; the tile's first and last
// destination (whose lists may continue in the neighbouring tiles) add their
// partial sum to accg for the fix-up pass.
__global__ void __launch_bounds__(kSweepBlock) k_pr_pull(GraphView rg, const unsigned long long *dpre, SweepTiles tiles,
                                                         const uint32_t *colp, uint64_t slot_lo, uint32_t dst_lo,
                                                         uint32_t dst_hi, uint64_t t_begin, uint64_t t_end, PullPR p,
                                                         PassIO io, int force) {
    if (!pull_dense(io, force)) return;
    __shared__ SweepWave s_wave[kSweepWaves];
    __shared__ unsigned long long s_red[kSweepWaves];
    const unsigned long long E = dpre[rg.nv];
    const unsigned long long ntiles = (rg.nv + E + kSweepTile - 1) / kSweepTile;
    const uint32_t w = threadIdx.x / kWave, lane = lane_id();
    SweepWave &sh = s_wave[w];
    float *const acc = reinterpret_cast<float *>(sh.base);  // base[] is dead once a tile's edges are loaded
    OutStage st{sh.out, 0, kSweepOutStage};
    uint32_t n_edges = 0;
    unsigned long long n_done = 0;
    unsigned long long t0 = t_begin, t1 = min_u64(min_u64(t_end, ntiles), tiles.cap - 1);
    unsigned long long gw = (unsigned long long)blockIdx.x * kSweepWaves + w, nwv = (unsigned long long)gridDim.x * kSweepWaves;
    if (p.tsplit) {
        const unsigned long long ts = *p.tsplit;
        const unsigned long long nh = max(1ull, nwv * p.host_waves64 / 64);
        if (ts > t0 && ts < t1 && nh < nwv) {
            if (gw < nh) {
                t0 = ts;
                nwv = nh;
            } else {
                t1 = ts;
                gw -= nh;
                nwv -= nh;
            }
        }
    }
    TilePipe tp;
    tp.tend = t1;
    tp.nw = nwv;
    tp.t = t0 + gw;
    tp.start(rg, dpre, tiles);
    while (tp.t < tp.tend) {
        uint32_t vv[kSweepPerLane];
        const TileHead h = tp.h;
        sweep_tile<true>(rg, dpre, h, tp.pre, colp, slot_lo, dst_lo, dst_hi, sh, vv, n_edges);
        tp.advance(rg, dpre, tiles);
        uint32_t bw[kSweepPerLane];
#pragma unroll
        for (int j = 0; j < kSweepPerLane; ++j) bw[j] = vv[j] != kGap ? p.bits[vv[j] >> 5] : 0u;
        float c[kSweepPerLane];
#pragma unroll
        for (int j = 0; j < kSweepPerLane; ++j) c[j] = (bw[j] >> (vv[j] & 31)) & 1u ? p.pv[vv[j]] : 0.0f;
        __builtin_amdgcn_wave_barrier();
        for (uint32_t i = lane; i < h.span; i += kWave) acc[i] = 0.0f;
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int j = 0; j < kSweepPerLane; ++j)
            if (c[j] != 0.0f) atomicAdd(&acc[sh.own[j * kWave + lane]], c[j]);
        __builtin_amdgcn_wave_barrier();
        for (uint32_t i0 = 0; i0 < h.span; i0 += kWave) {
            const uint32_t i = i0 + lane;
            const uint32_t v = h.v0 + i;
            bool act = false;
            if (i < h.span && v >= dst_lo && v < dst_hi) {
                const float total = acc[i];
                if (i == 0 || i + 1 == h.span) {
                    if (total != 0.0f) atomicAdd(&p.accg[v], total);
                } else {
                    act = pull_finish(p, v, total);
                    ++n_done;
                }
            }
            stage_push(act, v, st, io);
        }
        __builtin_amdgcn_wave_barrier();
    }
    stage_flush(st, io);
    stage_count_flush(st, io);
    block_count<kSweepWaves>((unsigned long long)n_edges, s_red, io.stats + 0);
    __syncthreads();
    block_count<kSweepWaves>(n_done, s_red, io.stats + 1);
}

// Fix-up of a pull pass: every tile's first destination (the vertices whose
// lists can cross a tile boundary, each handled at its first tile) adds its
// partial sums and is finished. Tile ntiles is the sentinel (vertex nv - 1).
__global__ void __launch_bounds__(kBlock) k_pr_pull_fix(const unsigned long long *dpre, uint32_t nv, SweepTiles tiles,
                                                        PullPR p, PassIO io, int force) {
    if (!pull_dense(io, force)) return;
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(io.stats + 4, 1ull);  // a pull round
    const unsigned long long E = dpre[nv];
    const unsigned long long nt = min_u64(((unsigned long long)nv + E + kSweepTile - 1) / kSweepTile, tiles.cap - 1);
    unsigned long long fed = 0;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t base = (uint64_t)blockIdx.x * blockDim.x; base <= nt; base += stride) {
        const uint64_t t = base + threadIdx.x;
        bool act = false;
        uint32_t v = 0;
        if (t <= nt) {
            v = tiles.v[t];
            if (t == 0 || tiles.v[t - 1] != v) {
                const float total = p.accg[v];
                if (total != 0.0f) p.accg[v] = 0.0f;
                act = pull_finish(p, v, total);
                if (act && io.fdeg) fed += io.fdeg[v];
            }
        }
        wave_append(act, v, io.out_list, io.out_cnt, io.out_cap, io.stats + 3);
    }
    if (io.ecnt_out) {
        fed = wave_sum_u64(fed);
        if (lane_id() == 0 && fed) atomicAdd(io.ecnt_out, fed);
    }
}

// ---- pull rounds of traversal programs -------------------------------------------
// The same direction choice for min-type programs (SSSP, BFS, CC and a
// caller's operator with a MergeBuffer method). In a dense round R the round
// kernel commits every frontier vertex and records its push tagged with R
// (VCombinePull::take) instead of walking its out-list; k_trav_pull then walks
// every in-list: per destination, the offers of the in-neighbours tagged R are
// accumulated in LDS with the operator's AccumulateBuffer (from
// GetIdentityElement) and folded into the destination's buffer with
// MergeBuffer. The fold is atomic, so a destination whose in-list is split by
// a tile boundary needs no fix-up pass: each tile folds its part. Every offer
// a push round makes along u -> v is made here along the same edge, and both
// folds keep the minimum, so both directions reach the same buffers and
// schedule the same vertices.
template <class Op>
struct VCombinePull {
    static constexpr bool kDense = true;
    Op op;
    typename Op::value_t *value;
    typename Op::buffer_t *buffer;
    PullTake t;
    __device__ bool operator()(uint32_t u, uint32_t deg, uint32_t &push, bool &self) const {
        self = false;
        auto p = op.CombineValueBuffer(u, &value[u], &buffer[u], deg);
        push = to_bits<typename Op::push_t>(p.value);
        return p.scheduled;
    }
    __device__ void take(uint32_t u, uint32_t push) const { t.take(u, push); }
};

template <class Op>
struct PullTrav {
    Op op;
    const uint32_t *bits;          // this round's frontier bitmap
    const uint32_t *pv;            // pushes of the frontier's vertices
    typename Op::buffer_t *buffer;
    uint32_t *stamp;               // round stamps (frontier de-duplication)
    uint32_t round;                // R
    uint32_t wmode;
    uint64_t wseed;
};

// One wave per tile of the in-edge store (as k_pr_pull); colp/slot_lo as there.
template <class Op>
__global__ void __launch_bounds__(kSweepBlock) k_trav_pull(GraphView rg, const unsigned long long *dpre, SweepTiles tiles,
                                                           const uint32_t *colp, uint64_t slot_lo, PullTrav<Op> p,
                                                           PassIO io, int force) {
    using buf_t = typename Op::buffer_t;
    if (!pull_dense(io, force)) return;
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(io.stats + 4, 1ull);  // a pull round
    __shared__ SweepWave s_wave[kSweepWaves];
    __shared__ buf_t s_acc[kSweepWaves][kSweepTile + 1];
    __shared__ unsigned long long s_red[kSweepWaves];
    const unsigned long long E = dpre[rg.nv];
    const unsigned long long ntiles = (rg.nv + E + kSweepTile - 1) / kSweepTile;
    const uint32_t w = threadIdx.x / kWave, lane = lane_id();
    SweepWave &sh = s_wave[w];
    buf_t *const acc = s_acc[w];
    const buf_t ident = p.op.GetIdentityElement();
    OutStage st{sh.out, 0, kSweepOutStage};
    uint32_t n_edges = 0;
    unsigned long long n_done = 0;
    TilePipe tp;
    tp.tend = min_u64(ntiles, tiles.cap - 1);
    tp.nw = (unsigned long long)gridDim.x * kSweepWaves;
    tp.t = (unsigned long long)blockIdx.x * kSweepWaves + w;
    tp.start(rg, dpre, tiles);
    while (tp.t < tp.tend) {
        uint32_t vv[kSweepPerLane];
        const TileHead h = tp.h;
        sweep_tile(rg, dpre, h, tp.pre, colp, slot_lo, 0u, rg.nv, sh, vv, n_edges);
        tp.advance(rg, dpre, tiles);
        uint32_t bw[kSweepPerLane];
#pragma unroll
        for (int j = 0; j < kSweepPerLane; ++j) bw[j] = vv[j] != kGap ? p.bits[vv[j] >> 5] : 0u;
        uint32_t pw[kSweepPerLane];
#pragma unroll
        for (int j = 0; j < kSweepPerLane; ++j) {
            bw[j] = (bw[j] >> (vv[j] & 31)) & 1u;
            pw[j] = bw[j] ? p.pv[vv[j]] : 0u;
        }
        for (uint32_t i = lane; i < h.span; i += kWave) acc[i] = ident;
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int j = 0; j < kSweepPerLane; ++j) {
            if (bw[j]) {
                const uint32_t own = sh.own[j * kWave + lane];
                const uint32_t u = vv[j], v = h.v0 + own;
                const uint32_t wt = Op::kWeighted ? sg_weight(p.wmode, p.wseed, u, v) : 0u;
                p.op.AccumulateBuffer(u, v, wt, &acc[own], from_bits<typename Op::push_t>(pw[j]));
            }
        }
        __builtin_amdgcn_wave_barrier();
        for (uint32_t i0 = 0; i0 < h.span; i0 += kWave) {
            const uint32_t i = i0 + lane;
            const uint32_t v = h.v0 + i;
            bool act = false;
            if (i < h.span) {
                const buf_t a = acc[i];
                if (a != ident) {
                    act = p.op.MergeBuffer(v, &p.buffer[v], a) && atomicExch(&p.stamp[v], p.round + 1) != p.round + 1;
                    ++n_done;
                }
            }
            stage_push(act, v, st, io);
        }
        __builtin_amdgcn_wave_barrier();
    }
    stage_flush(st, io);
    stage_count_flush(st, io);
    block_count<kSweepWaves>((unsigned long long)n_edges, s_red, io.stats + 0);
    __syncthreads();
    block_count<kSweepWaves>(n_done, s_red, io.stats + 1);
}

// ---- read copies of a host-resident in-edge store ----------------------------------
// Lanes [l0, l0 + n) of the wave hold (in v) the ids of indices i0 .. i0 + n - 1
// of an array of b-bit ids packed back to back (b <= 31 here, so the words
// they cover fit one per lane; the array is zeroed before the fill). Each
// lane assembles one word from the up to 32 / b + 2 ids that overlap it;
// whole words are stored, the two end words - shared with the neighbouring
// runs - are ORed in. Every lane of the wave calls it (the ids move by
// shuffles).
__device__ __forceinline__ void pack_run(uint32_t *arr, unsigned long long i0, uint32_t l0, uint32_t n, uint32_t v,
                                         uint32_t b) {
    if (n == 0) return;
    const uint32_t lane = lane_id();
    const unsigned long long bit0 = i0 * b, bit1 = (i0 + n) * b;
    const unsigned long long w0 = bit0 >> 5, w1 = (bit1 - 1) >> 5;
    const unsigned long long w = w0 + lane;
    const unsigned long long wb = w << 5;
    unsigned long long q = wb / b;  // first id overlapping word w
    if (q < i0) q = i0;
    uint32_t word = 0;
    const uint32_t nt = 32 / b + 2;
    for (uint32_t t = 0; t < nt; ++t, ++q) {
        const long long rel = (long long)(q - i0);
        const uint32_t src = l0 + (uint32_t)(rel < (long long)n ? rel : (long long)n - 1);
        const uint32_t x = __shfl(v, (int)src, kWave);
        if (w <= w1 && q < i0 + n && q * b < wb + 32) {
            const long long sh = (long long)(q * b) - (long long)wb;
            word |= sh >= 0 ? (x << (uint32_t)sh) : (x >> (uint32_t)(-sh));
        }
    }
    if (w <= w1) {
        if (w == w0 || w == w1) atomicOr(&arr[w], word);
        else arr[w] = word;
    }
}

// Store the ids of lanes [l0, l0 + n) at indices i0 .. of a copy: packed at
// b bits, or plain words (b = 32).
__device__ __forceinline__ void copy_run(uint32_t *arr, unsigned long long i0, uint32_t l0, uint32_t n, uint32_t v,
                                         uint32_t b) {
    if (b == 32) {
        const uint32_t lane = lane_id();
        if (lane >= l0 && lane < l0 + n) arr[i0 + (lane - l0)] = v;
    } else {
        pack_run(arr, i0, l0, n, v, b);
    }
}

// First tile whose edges reach past live-edge index x (every tile before it
// ends at or below x); nt when none.
__device__ __forceinline__ unsigned long long tile_reaching(const SweepTiles &tiles, unsigned long long nt,
                                                            unsigned long long x) {
    unsigned long long lo = 0, hi = nt;
    while (lo < hi) {
        const unsigned long long mid = (lo + hi) / 2;
        if (tiles.e[mid + 1] > x) hi = mid;
        else lo = mid + 1;
    }
    return lo;
}

// Copies of the store's live edges, read from the store in host memory
// through the tile sweep (live slots only, coalesced); g must carry no copy
// itself, and packed arrays must be zero.
//  - mirror (tail null): mirror id k = live edge k (merge-path order) for
//    k < mlive; the tiles are in edge order, so a wave stops at its first
//    tile past mlive.
//  - tail window (mirror null): stage id i = live edge mlive + wlo + i for
//    mlive + wlo <= k < mlive + whi; waves start at the first tile reaching
//    the window. wlo is a multiple of 32, so the window's packed bits start
//    on a word boundary and windows share no word.
__global__ void __launch_bounds__(kSweepBlock) k_mirror_fill(GraphView g, const unsigned long long *dpre, SweepTiles tiles,
                                                             uint32_t *mirror, uint64_t mlive, uint32_t *tail,
                                                             uint64_t wlo, uint64_t whi, uint32_t b) {
    __shared__ SweepWave s_wave[kSweepWaves];
    const unsigned long long E = dpre[g.nv];
    const unsigned long long ntiles = min_u64((g.nv + E + kSweepTile - 1) / kSweepTile, tiles.cap - 1);
    const uint32_t w = threadIdx.x / kWave;
    SweepWave &sh = s_wave[w];
    uint32_t n_edges = 0;
    const unsigned long long k_lo = tail ? mlive + wlo : 0, k_hi = tail ? mlive + whi : mlive;
    TilePipe tp;
    tp.tend = ntiles;
    tp.nw = (unsigned long long)gridDim.x * kSweepWaves;
    tp.t = (tail ? tile_reaching(tiles, ntiles, k_lo) : 0ull) + (unsigned long long)blockIdx.x * kSweepWaves + w;
    tp.start(g, dpre, tiles);
    while (tp.t < tp.tend) {
        const TileHead h = tp.h;
        if (h.e0 >= k_hi) break;
        uint32_t vv[kSweepPerLane];
        sweep_tile(g, dpre, h, tp.pre, g.col, 0, 0u, g.nv, sh, vv, n_edges);
        tp.advance(g, dpre, tiles);
#pragma unroll
        for (int j = 0; j < kSweepPerLane; ++j) {
            // positions [j * 64, j * 64 + nrun) of the tile: edges k0 .. k0 + nrun - 1, of which
            // lanes [l0, l1) fall in [k_lo, k_hi)
            const uint32_t pj = j * kWave;
            const uint32_t nrun = h.ne > pj ? min(h.ne - pj, (uint32_t)kWave) : 0u;
            const unsigned long long k0 = h.e0 + pj;
            const uint32_t l0 = k0 < k_lo ? (uint32_t)min_u64(nrun, k_lo - k0) : 0u;
            const uint32_t l1 = k0 < k_hi ? (uint32_t)min_u64(nrun, k_hi - k0) : 0u;
            if (l1 > l0) copy_run(tail ? tail : mirror, k0 + l0 - k_lo, l0, l1 - l0, vv[j], b);
        }
        __builtin_amdgcn_wave_barrier();
    }
}

// First tile whose edges reach edge index mlive (every tile before it lies
// wholly in the mirror); ntiles when none. One thread.
__global__ void k_tile_split(SweepTiles tiles, const unsigned long long *dpre, uint32_t nv, unsigned long long mlive,
                             unsigned long long *out) {
    const unsigned long long E = dpre[nv];
    const unsigned long long nt = min_u64((nv + E + kSweepTile - 1) / kSweepTile, tiles.cap - 1);
    *out = tile_reaching(tiles, nt, mlive);
}

// ---- traversal recompute from the in-lists of invalidated vertices -------------
// After deletions, every invalidated vertex v must see the best value any
// in-neighbour u offers: PushValue(best(u)) + w(u, v) (for CC the label of u).
// The forward form sweeps every live edge and relaxes those whose target is
// invalidated (the reference re-pushes from every vertex:
// RelaxCTADB_all_vertices, sync_push_dd.cuh:1470-1557); with an in-edge store
// only the in-lists of the invalidated vertices are read. A round kernel walks
// those lists (in-store view, the invalidated vertices as items) with this
// functor; whatever improves v appends v itself.
template <class Op>
struct EPullSeed {
    static constexpr bool kAppendSource = true;
    Op op;
    typename Op::buffer_t *buffer;
    uint32_t *stamp;
    uint32_t stamp_val;
    uint32_t wmode;
    uint64_t wseed;
    const uint32_t *unlabel;  // vertex of each in-list id (null: ids are vertices)
    __device__ uint32_t vertex(uint32_t x) const { return unlabel ? unlabel[x] : x; }
    __device__ uint64_t load(uint32_t x) const { return ld_peek(&buffer[vertex(x)]); }
    // v: the invalidated vertex whose in-list is walked; x: one in-list entry
    __device__ bool apply(uint32_t v, uint32_t x, uint32_t, uint64_t bu) const {
        const uint32_t u = vertex(x);
        const uint32_t w = Op::kWeighted ? sg_weight(wmode, wseed, u, v) : 0u;
        const auto p = op.PushValue(bu);
        if (!op.MayImprove(u, w, p, ld_peek(&buffer[v]))) return false;
        if (!op.AccumulateBuffer(u, v, w, &buffer[v], p)) return false;
        return atomicExch(&stamp[v], stamp_val) != stamp_val;
    }
};

// The vertices of a bitmap as a list (one thread per word; one atomic per wave).
__global__ void __launch_bounds__(kBlock) k_bits_to_list(uint32_t nv, const uint32_t *bits, uint32_t *list, uint32_t *cnt,
                                                         uint32_t cap, unsigned long long *overflow) {
    const uint32_t nw = (nv + 31) / 32;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t base = (uint64_t)blockIdx.x * blockDim.x; base < nw; base += stride) {
        const uint64_t wd = base + threadIdx.x;
        uint32_t b = wd < nw ? bits[wd] : 0u;
        if (wd + 1 == nw && (nv & 31)) b &= (1u << (nv & 31)) - 1;
        const uint32_t c = __popc(b);
        const uint32_t incl = wave_incl_scan(c);
        const uint32_t total = __shfl(incl, kWave - 1, kWave);
        if (total == 0) continue;
        uint32_t at = 0;
        if (lane_id() == 0) at = atomicAdd(cnt, total);
        at = __shfl(at, 0, kWave) + incl - c;
        while (b) {
            const uint32_t v = (uint32_t)wd * 32 + __builtin_ctz(b);
            b &= b - 1;
            if (at < cap) list[at] = v;
            else atomicOr(overflow, 1ull);
            ++at;
        }
    }
}

// Out-edges of a frontier list (the first round's direction choice).
__global__ void __launch_bounds__(kBlock) k_list_degsum(const uint32_t *list, const uint32_t *cnt, uint32_t cap,
                                                        const uint32_t *deg, unsigned long long *out) {
    const uint32_t n = min(*cnt, cap);
    unsigned long long s = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        s += deg[list[i]];
    s = wave_sum_u64(s);
    if (lane_id() == 0 && s) atomicAdd(out, s);
}

}  // namespace sg

#endif
