// Batched BFS/SSSP queries over one shared PMA graph: many sources answered by
// one launch per round, one PMA update per batch and one boundary sweep for
// all queries. No reference counterpart (the reference runs one query per
// process); each query follows the single-source semantics of the SSSP/BFS
// operators (hybrid_sssp.cu:28-194, hybrid_bfs.cu:26-167): a BFS query is an
// SSSP with unit weights, and the pair (distance, parent) is kept in one
// 64-bit word updated by a lexicographic atomicMin, so a query gives the same
// distances and parents alone and in a batch.
//
// Layout: queries are cut into G groups of QS slots (QS = 64, or the next
// power of two >= the query count when there are fewer than 64); a vertex
// holds G * QS words of state and one 64-bit mask per group for each of
// act / inv / ninv. A wave step covers 64 / QS items (edges, pairs) times QS
// queries: lane = item * QS + slot.
#ifndef SG_MULTI_H
#define SG_MULTI_H

#include "sg_operators.h"
#include "sg_pma.h"

namespace sg {

using api::hi32;
using api::kInf;
using api::kNone;
using api::lo32;
using api::pack;

constexpr uint32_t kMEdgeChunk = 512;  // edges a wave walks for one (vertex, group) item; the rest is spilled
constexpr int kMUnroll = 8;            // wave steps whose loads are in flight together

enum : uint8_t { kQWeighted = 0, kQUnit = 1, kQNone = 2 };

struct MultiState {
    unsigned long long *best;  // [nv][KP]: distance << 32 | parent; ~0 = unreached
    unsigned long long *act;   // [nv][G]: slots whose distance dropped (to push)
    unsigned long long *inv;   // [nv][G]: slots reset by this batch's deletions
    unsigned long long *ninv;  // [nv][G]: slots reset and not yet propagated to tree children
    uint32_t *invbit;          // [nv / 32]: v has a reset slot in some group
    uint32_t *coarse;          // coarse filter of invbit: one bit per 2^cshift vertices
    uint32_t cshift;
    const uint8_t *kind;       // [KP]: kQWeighted, kQUnit or kQNone (padding)
    uint32_t G, KP, QS, qs_log;
    uint32_t wmode;
    uint64_t wseed;
};

// Spilled piece of an adjacency list: walk [start, start + len) of u for the
// slots in mask of group grp.
struct MDesc {
    uint64_t start;
    uint64_t mask;
    uint32_t u, len, grp, pad;
};

struct QLane {
    uint32_t sub;    // item of the wave step
    uint32_t slot;   // query slot in the group
    uint32_t per;    // items per wave step (64 / QS)
    uint64_t smask;  // the QS bits of this lane's item in a ballot
};
__device__ __forceinline__ QLane qlane(const MultiState &ms) {
    QLane q;
    const uint32_t l = lane_id();
    q.sub = l >> ms.qs_log;
    q.slot = l & (ms.QS - 1);
    q.per = kWave >> ms.qs_log;
    const uint64_t ones = ms.QS == 64 ? ~0ull : ((1ull << ms.QS) - 1);
    q.smask = ones << (q.sub * ms.QS);
    return q;
}

__device__ __forceinline__ uint32_t qweight(const MultiState &ms, uint8_t kind, uint32_t u, uint32_t v) {
    return kind == kQUnit ? 1u : sg_weight(ms.wmode, ms.wseed, u, v);
}

// Relax slot (grp, slot) of v from u at distance du; true when v's distance
// dropped.
__device__ __forceinline__ bool mrelax(const MultiState &ms, uint32_t u, uint32_t v, uint32_t grp, uint32_t slot,
                                       uint32_t du, uint8_t kind) {
    const uint32_t cand = du + qweight(ms, kind, u, v);
    unsigned long long *p = &ms.best[(uint64_t)v * ms.KP + grp * ms.QS + slot];
    const unsigned long long nw = pack(cand, u);
    if (nw >= ld_peek(p)) return false;
    const unsigned long long old = atomicMin(p, nw);
    return cand < hi32(old);
}

// After a wave step: for every item whose ballot bits are set, OR them into
// the item's mask word and put its vertex on the frontier once per round.
__device__ __forceinline__ void mark_items(const MultiState &ms, const PassIO &io, OutStage &st, const QLane &q,
                                           unsigned long long *masks, uint32_t grp, uint32_t v, bool hit, uint32_t *stamp,
                                           uint32_t stamp_val) {
    const uint64_t b = __ballot(hit) & q.smask;
    bool app = false;
    if (q.slot == 0 && b) {
        atomicOr(&masks[(uint64_t)v * ms.G + grp], b >> (q.sub * ms.QS));
        app = atomicExch(&stamp[v], stamp_val) != stamp_val;
    }
    stage_push(app, v, st, io);
}

// Reset the slots of v whose parent is the walking source (child = true in
// the slot's lane): claim them in inv, queue them in ninv, flag v in the
// vertex bitmaps, reset the state words and put v on the frontier.
__device__ __forceinline__ void reset_children(const MultiState &ms, const PassIO &io, OutStage &st, const QLane &q,
                                               uint32_t grp, uint32_t v, bool child, uint32_t *stamp, uint32_t stamp_val) {
    const uint64_t b = __ballot(child) & q.smask;
    uint64_t fresh = 0;
    if (q.slot == 0 && b) {
        const uint64_t m = b >> (q.sub * ms.QS);
        fresh = m & ~atomicOr(&ms.inv[(uint64_t)v * ms.G + grp], m);
    }
    fresh = __shfl(fresh, (int)(q.sub * ms.QS), kWave);
    bool app = false;
    if (q.slot == 0 && fresh) {
        atomicOr(&ms.ninv[(uint64_t)v * ms.G + grp], fresh);
        atomicOr(&ms.invbit[v >> 5], 1u << (v & 31));
        coarse_mark(ms.coarse, ms.cshift, v);
        app = atomicExch(&stamp[v], stamp_val) != stamp_val;
    }
    if ((fresh >> q.slot) & 1u) ms.best[(uint64_t)v * ms.KP + grp * ms.QS + q.slot] = ~0ull;
    stage_push(app, v, st, io);
}

// Walk for a sparse slot mask (at most kSparseSlots queries active at u):
// one lane per edge, each lane trying the few active slots in turn, instead
// of QS lanes per edge of which most would idle. kMode as mwalk.
constexpr int kSparseSlots = 4;
constexpr int kSparseUnroll = 4;  // edges per lane in flight
template <int kMode>
__device__ void mwalk_sparse(const GraphView &g, const MultiState &ms, const PassIO &io, OutStage &st, uint32_t *stamp,
                             uint32_t stamp_val, uint32_t u, uint32_t grp, uint64_t mask, uint64_t start, uint32_t len,
                             unsigned long long &edges) {
    uint32_t sl[kSparseSlots], dus[kSparseSlots];
    bool unit[kSparseSlots], live[kSparseSlots];
    bool any = false;
    uint64_t m = mask;
#pragma unroll
    for (int a = 0; a < kSparseSlots; ++a) {  // the active slots and u's distances (wave-uniform)
        sl[a] = 0;
        dus[a] = kInf;
        unit[a] = false;
        live[a] = false;
        if (m) {
            const uint32_t s = (uint32_t)__builtin_ctzll(m);
            m &= m - 1;
            const uint32_t qi = grp * ms.QS + s;
            const uint8_t kind = ms.kind[qi];
            sl[a] = s;
            unit[a] = kind == kQUnit;
            if (kMode == 0) {
                dus[a] = kind == kQNone ? kInf : hi32(ld_relaxed(&ms.best[(uint64_t)u * ms.KP + qi]));
                live[a] = dus[a] != kInf;
            } else {
                live[a] = kind != kQNone;
            }
            any = any || live[a];
        }
    }
    if (!any) return;
    edges += len;
    const uint32_t lane = lane_id();
    for (uint32_t e0 = 0; e0 < len; e0 += kWave * kSparseUnroll) {
        uint32_t vv[kSparseUnroll];
#pragma unroll
        for (int j = 0; j < kSparseUnroll; ++j) {
            const uint32_t e = e0 + j * kWave + lane;
            vv[j] = e < len ? edge_at(g, start + e) : kGap;
        }
#pragma unroll
        for (int j = 0; j < kSparseUnroll; ++j) {
            const uint32_t v = vv[j];
            bool app = false;
            if (v != kGap) {
                if (kMode == 0) {
                    const uint32_t w = sg_weight(ms.wmode, ms.wseed, u, v);
                    uint64_t better = 0;
#pragma unroll
                    for (int a = 0; a < kSparseSlots; ++a) {
                        if (!live[a]) continue;
                        const uint32_t cand = dus[a] + (unit[a] ? 1u : w);
                        unsigned long long *p = &ms.best[(uint64_t)v * ms.KP + grp * ms.QS + sl[a]];
                        const unsigned long long nw = pack(cand, u);
                        if (nw < ld_peek(p) && cand < hi32(atomicMin(p, nw))) better |= 1ull << sl[a];
                    }
                    if (better) {
                        atomicOr(&ms.act[(uint64_t)v * ms.G + grp], better);
                        app = atomicExch(&stamp[v], stamp_val) != stamp_val;
                    }
                } else {
                    // slots of v whose tree parent is u: claim them once, reset them
                    uint64_t child = 0;
#pragma unroll
                    for (int a = 0; a < kSparseSlots; ++a) {
                        if (!live[a]) continue;
                        const unsigned long long b = ld_relaxed(&ms.best[(uint64_t)v * ms.KP + grp * ms.QS + sl[a]]);
                        if (lo32(b) == u && hi32(b) != kInf) child |= 1ull << sl[a];
                    }
                    uint64_t fresh = 0;
                    if (child) fresh = child & ~atomicOr(&ms.inv[(uint64_t)v * ms.G + grp], child);
                    if (fresh) {
                        atomicOr(&ms.ninv[(uint64_t)v * ms.G + grp], fresh);
                        atomicOr(&ms.invbit[v >> 5], 1u << (v & 31));
                        coarse_mark(ms.coarse, ms.cshift, v);
                        app = atomicExch(&stamp[v], stamp_val) != stamp_val;
                        for (uint64_t f = fresh; f; f &= f - 1)
                            ms.best[(uint64_t)v * ms.KP + grp * ms.QS + (uint32_t)__builtin_ctzll(f)] = ~0ull;
                    }
                }
            }
            stage_push(app, v, st, io);
        }
    }
}

// Position of the r-th (0-based) set bit of m; r < popcount(m).
__device__ __forceinline__ uint32_t nth_set_bit(uint64_t m, uint32_t r) {
    uint32_t pos = 0;
#pragma unroll
    for (uint32_t width = 32; width >= 1; width >>= 1) {
        const uint32_t c = (uint32_t)__popcll(m & ((1ull << width) - 1));
        if (r >= c) {
            r -= c;
            m >>= width;
            pos += width;
        }
    }
    return pos;
}

// Walk for a mid-density slot mask (W/2 < popcount <= W, W = 8, 16 or 32
// below the group's QS): the active slots are packed into W lanes, so a wave
// step covers 64 / W edges instead of 64 / QS with most lanes idle. Lane
// rank r takes the r-th active slot; the slot bits of an edge's hits are
// OR-reduced over its W lanes before the one atomic per item. kMode as mwalk.
template <int kMode>
__device__ __forceinline__ void mwalk_compact(const GraphView &g, const MultiState &ms, const PassIO &io, OutStage &st, uint32_t *stamp,
                              uint32_t stamp_val, uint32_t u, uint32_t grp, uint64_t mask, uint32_t wlog, uint64_t start,
                              uint32_t len, unsigned long long &edges) {
    const uint32_t lane = lane_id();
    const uint32_t W = 1u << wlog;
    const uint32_t r = lane & (W - 1);  // rank among the active slots
    const uint32_t sub = lane >> wlog;   // edge of the wave step
    const uint32_t per = kWave >> wlog;  // edges per wave step
    const bool on = r < (uint32_t)__popcll(mask);
    const uint32_t slot = on ? nth_set_bit(mask, r) : 0u;
    const uint32_t qi = grp * ms.QS + slot;
    const uint8_t kind = on ? ms.kind[qi] : (uint8_t)kQNone;
    uint32_t du = kInf;
    if (kMode == 0 && on) du = hi32(ld_relaxed(&ms.best[(uint64_t)u * ms.KP + qi]));
    const bool live = on && kind != kQNone && (kMode == 1 || du != kInf);
    if (!__any(live)) return;
    edges += len;
    for (uint32_t e0 = 0; e0 < len; e0 += per * kMUnroll) {
        uint32_t vv[kMUnroll];
#pragma unroll
        for (int j = 0; j < kMUnroll; ++j) {
            const uint32_t e = e0 + j * per + sub;
            vv[j] = e < len ? edge_at(g, start + e) : kGap;
        }
        uint64_t pk[kMUnroll];
        uint32_t cand[kMUnroll];
#pragma unroll
        for (int j = 0; j < kMUnroll; ++j) {
            const bool ok = live && vv[j] != kGap;
            if (kMode == 0) {
                cand[j] = ok ? du + qweight(ms, kind, u, vv[j]) : kInf;
                pk[j] = ok ? ld_peek(&ms.best[(uint64_t)vv[j] * ms.KP + qi]) : 0ull;
            } else {
                cand[j] = 0;
                pk[j] = ok ? ld_relaxed(&ms.best[(uint64_t)vv[j] * ms.KP + qi]) : 0ull;
            }
        }
#pragma unroll
        for (int j = 0; j < kMUnroll; ++j) {
            const uint32_t v = vv[j];
            bool hit = false;
            if (kMode == 0) {
                const unsigned long long nw = pack(cand[j], u);
                if (cand[j] != kInf && nw < pk[j]) {
                    const unsigned long long old = atomicMin(&ms.best[(uint64_t)v * ms.KP + qi], nw);
                    hit = cand[j] < hi32(old);
                }
            } else {
                hit = live && v != kGap && lo32(pk[j]) == u && hi32(pk[j]) != kInf;
            }
            // slot bits of this edge's hits, gathered in every lane of its W-lane group
            unsigned long long bits = hit ? (1ull << slot) : 0ull;
            for (uint32_t o = 1; o < W; o <<= 1) bits |= __shfl_xor(bits, (int)o, kWave);
            bool app = false;
            if (kMode == 0) {
                if (r == 0 && bits) {
                    atomicOr(&ms.act[(uint64_t)v * ms.G + grp], bits);
                    app = atomicExch(&stamp[v], stamp_val) != stamp_val;
                }
            } else {
                unsigned long long fresh = 0;
                if (r == 0 && bits) fresh = bits & ~atomicOr(&ms.inv[(uint64_t)v * ms.G + grp], bits);
                fresh = __shfl(fresh, (int)(sub << wlog), kWave);
                if (r == 0 && fresh) {
                    atomicOr(&ms.ninv[(uint64_t)v * ms.G + grp], fresh);
                    atomicOr(&ms.invbit[v >> 5], 1u << (v & 31));
                    coarse_mark(ms.coarse, ms.cshift, v);
                    app = atomicExch(&stamp[v], stamp_val) != stamp_val;
                }
                if (on && ((fresh >> slot) & 1u)) ms.best[(uint64_t)v * ms.KP + qi] = ~0ull;
            }
            stage_push(app, v, st, io);
        }
    }
}

// Walk len edges of u from slot position start for the slots in mask.
// kMode 0: relax (activations go to act); kMode 1: reset tree children whose
// parent is u (resets go to inv / ninv and the vertex bitmaps).
template <int kMode>
__device__ void mwalk(const GraphView &g, const MultiState &ms, const PassIO &io, OutStage &st, uint32_t *stamp,
                      uint32_t stamp_val, uint32_t u, uint32_t grp, uint64_t mask, uint64_t start, uint32_t len,
                      unsigned long long &edges) {
    const uint32_t np = (uint32_t)__popcll(mask);
    if (ms.QS > kSparseSlots && np <= kSparseSlots) {
        mwalk_sparse<kMode>(g, ms, io, st, stamp, stamp_val, u, grp, mask, start, len, edges);
        return;
    }
    if (ms.QS >= 16 && 2 * np <= ms.QS) {  // 8, 16 or 32 lanes per edge instead of QS
        const uint32_t wlog = np <= 8 ? 3u : np <= 16 ? 4u : 5u;
        mwalk_compact<kMode>(g, ms, io, st, stamp, stamp_val, u, grp, mask, wlog, start, len, edges);
        return;
    }
    const QLane q = qlane(ms);
    const bool on = (mask >> q.slot) & 1u;
    const uint32_t qi = grp * ms.QS + q.slot;
    const uint8_t kind = ms.kind[qi];
    uint32_t du = kInf;
    if (kMode == 0 && on) du = hi32(ld_relaxed(&ms.best[(uint64_t)u * ms.KP + qi]));
    const bool live = on && kind != kQNone && (kMode == 1 || du != kInf);
    if (!__any(live)) return;
    edges += len;
    for (uint32_t e0 = 0; e0 < len; e0 += q.per * kMUnroll) {
        uint32_t vv[kMUnroll];
#pragma unroll
        for (int j = 0; j < kMUnroll; ++j) {
            const uint32_t e = e0 + j * q.per + q.sub;
            vv[j] = e < len ? edge_at(g, start + e) : kGap;
        }
        if (kMode == 0) {
            // all peeks in flight before the first atomic
            uint64_t pk[kMUnroll];
            uint32_t cand[kMUnroll];
#pragma unroll
            for (int j = 0; j < kMUnroll; ++j) {
                const bool ok = live && vv[j] != kGap;
                cand[j] = ok ? du + qweight(ms, kind, u, vv[j]) : kInf;
                pk[j] = ok ? ld_peek(&ms.best[(uint64_t)vv[j] * ms.KP + qi]) : 0ull;
            }
#pragma unroll
            for (int j = 0; j < kMUnroll; ++j) {
                bool hit = false;
                const unsigned long long nw = pack(cand[j], u);
                if (cand[j] != kInf && nw < pk[j]) {
                    const unsigned long long old = atomicMin(&ms.best[(uint64_t)vv[j] * ms.KP + qi], nw);
                    hit = cand[j] < hi32(old);
                }
                mark_items(ms, io, st, q, ms.act, grp, vv[j], hit, stamp, stamp_val);
            }
        } else {
            uint64_t peek[kMUnroll];
#pragma unroll
            for (int j = 0; j < kMUnroll; ++j)
                peek[j] = (live && vv[j] != kGap) ? ld_relaxed(&ms.best[(uint64_t)vv[j] * ms.KP + qi]) : 0ull;
#pragma unroll
            for (int j = 0; j < kMUnroll; ++j) {
                const uint32_t v = vv[j];
                const bool child = live && v != kGap && lo32(peek[j]) == u && hi32(peek[j]) != kInf;
                reset_children(ms, io, st, q, grp, v, child, stamp, stamp_val);
            }
        }
    }
}

// LDS of the light path (per wave): 64 items of a wave batch.
constexpr uint32_t kMLightQS = 16;  // light path for one group of at most this many slots
struct MLight {
    uint32_t excl[kWave];
    uint32_t u[kWave];
    uint64_t start[kWave];
    uint32_t mask[kWave];
    uint32_t du[kWave * kMLightQS];
};

// Light path (one group, QS <= kMLightQS): a wave takes 64 frontier vertices
// at once, stages their lists (at most kMEdgeChunk edges each, the rest is
// spilled) and walks the concatenation, 64 / QS edges per wave step.
template <int kMode>
__device__ void mlight(const GraphView &g, const MultiState &ms, const PassIO &io, OutStage &st, MLight &sh,
                       MDesc *spill, uint32_t spill_cap, uint32_t *stamp, uint32_t n, uint64_t gw, uint64_t nw,
                       unsigned long long &edges, unsigned long long &verts) {
    unsigned long long *masks = kMode == 0 ? ms.act : ms.ninv;
    const QLane q = qlane(ms);
    const uint32_t lane = lane_id();
    const uint32_t stamp_val = io.round + 1;
    const uint8_t kind = ms.kind[q.slot];
    const uint32_t per = light_width(n, nw);  // small frontiers spread over all waves
    for (uint64_t b = gw * per; b < n; b += nw * per) {
        const uint64_t idx = b + lane;
        uint32_t u = 0, m = 0, len = 0, deg = 0;
        uint64_t start = 0;
        if (lane < per && idx < n) {
            u = io.in_list[idx];
            m = (uint32_t)atomicExch(&masks[u], 0ull);
            if (m) {
                deg = g.deg[u];
                start = list_start(g, u);
                len = min(deg, kMEdgeChunk);
                count_hot(g, u);
            }
        }
        verts += (uint32_t)__popcll(__ballot(m != 0));
        // lists past the wave's edge budget go to other waves whole, as
        // descriptors; of the others only the tail past kMEdgeChunk does
        const uint32_t pre = wave_incl_scan(len);
        if (len && pre > io.budget) len = 0;
        const uint32_t nch = deg > len ? (deg - len + kMEdgeChunk - 1) / kMEdgeChunk : 0u;
        const uint32_t incl = wave_incl_scan(nch);
        const uint32_t tot = __shfl(incl, kWave - 1, kWave);
        if (tot) {
            uint32_t base = 0;
            if (lane == 0) base = atomicAdd(io.heavy_cnt, tot);
            base = __shfl(base, 0, kWave) + incl - nch;
            for (uint32_t c = 0; c < nch; ++c) {
                const uint32_t lo = len + c * kMEdgeChunk;
                if (base + c < spill_cap) spill[base + c] = MDesc{start + lo, m, u, min(kMEdgeChunk, deg - lo), 0, 0};
                else atomicOr(io.stats + 3, 2ull);
            }
        }
        const uint32_t lincl = wave_incl_scan(len);
        const uint32_t total = __shfl(lincl, kWave - 1, kWave);
        sh.excl[lane] = lincl - len;
        sh.u[lane] = u;
        sh.start[lane] = start;
        sh.mask[lane] = m;
        if (kMode == 0) {
            for (uint32_t s2 = 0; s2 < ms.QS; ++s2)
                sh.du[lane * ms.QS + s2] = ((m >> s2) & 1u) ? hi32(ld_relaxed(&ms.best[(uint64_t)u * ms.KP + s2])) : kInf;
        }
        __builtin_amdgcn_wave_barrier();
        if (total == 0) continue;
        edges += total;
        for (uint32_t k0 = 0; k0 < total; k0 += q.per * kMUnroll) {
            uint32_t vv[kMUnroll], src[kMUnroll], dv[kMUnroll];
            bool on[kMUnroll];
#pragma unroll
            for (int j = 0; j < kMUnroll; ++j) {
                const uint32_t k = k0 + j * q.per + q.sub;
                const bool ok = k < total;
                const uint32_t o = staged_owner(sh.excl, ok ? k : 0u);
                src[j] = sh.u[o];
                on[j] = ok && ((sh.mask[o] >> q.slot) & 1u) && kind != kQNone;
                dv[j] = kMode == 0 ? sh.du[o * ms.QS + q.slot] : 0u;
                vv[j] = ok ? edge_at(g, sh.start[o] + (k - sh.excl[o])) : kGap;
            }
            if (kMode == 0) {
                uint64_t pk[kMUnroll];
                uint32_t cand[kMUnroll];
#pragma unroll
                for (int j = 0; j < kMUnroll; ++j) {
                    const bool ok = on[j] && vv[j] != kGap && dv[j] != kInf;
                    cand[j] = ok ? dv[j] + qweight(ms, kind, src[j], vv[j]) : kInf;
                    pk[j] = ok ? ld_peek(&ms.best[(uint64_t)vv[j] * ms.KP + q.slot]) : 0ull;
                }
#pragma unroll
                for (int j = 0; j < kMUnroll; ++j) {
                    bool hit = false;
                    const unsigned long long nwv = pack(cand[j], src[j]);
                    if (cand[j] != kInf && nwv < pk[j]) {
                        const unsigned long long old = atomicMin(&ms.best[(uint64_t)vv[j] * ms.KP + q.slot], nwv);
                        hit = cand[j] < hi32(old);
                    }
                    mark_items(ms, io, st, q, ms.act, 0, vv[j], hit, stamp, stamp_val);
                }
            } else {
                uint64_t pk[kMUnroll];
#pragma unroll
                for (int j = 0; j < kMUnroll; ++j)
                    pk[j] = (on[j] && vv[j] != kGap) ? ld_relaxed(&ms.best[(uint64_t)vv[j] * ms.KP + q.slot]) : 0ull;
#pragma unroll
                for (int j = 0; j < kMUnroll; ++j) {
                    const uint32_t v = vv[j];
                    const bool child = on[j] && v != kGap && lo32(pk[j]) == src[j] && hi32(pk[j]) != kInf;
                    reset_children(ms, io, st, q, 0, v, child, stamp, stamp_val);
                }
            }
        }
        __builtin_amdgcn_wave_barrier();
    }
}

// One convergence (kMode 0) or invalidation (kMode 1) round: items are
// (frontier vertex, group); the mask of the item is taken (and cleared) from
// act or ninv; lists longer than kMEdgeChunk leave their tail to the next
// round as descriptors; then the descriptors the previous round spilled.
template <int kMode>
__global__ void __launch_bounds__(kBlock) k_mround(GraphView g, MultiState ms, PassIO io, MDesc *spill, const MDesc *sin,
                                                   uint32_t spill_cap, uint32_t *stamp) {
    __shared__ uint32_t s_out[kWavesPerBlock][kOutStage];
    __shared__ unsigned long long s_red[kWavesPerBlock], s_red2[kWavesPerBlock];
    __shared__ MLight s_light[kWavesPerBlock];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        if (io.zero_cnt) *io.zero_cnt = 0;
        if (io.heavy_zero) *io.heavy_zero = 0;
        if (io.log) {
            const uint32_t i = 2 * ((io.round + kRoundLog - 1) % kRoundLog);
            io.log[i] = *io.in_cnt;
            io.log[i + 1] = io.hin_cnt ? *io.hin_cnt : 0u;
        }
    }
    unsigned long long *masks = kMode == 0 ? ms.act : ms.ninv;
    const uint32_t n = min(*io.in_cnt, io.in_cap);
    const uint32_t nh = sin ? min(*io.hin_cnt, spill_cap) : 0u;
    const uint32_t w = threadIdx.x / kWave, lane = lane_id();
    const uint64_t gw = (uint64_t)blockIdx.x * kWavesPerBlock + w, nw = (uint64_t)gridDim.x * kWavesPerBlock;
    OutStage st{s_out[w], 0};
    unsigned long long edges = 0, verts = 0;
    const bool light = ms.G == 1 && ms.QS <= kMLightQS;
    if (light) mlight<kMode>(g, ms, io, st, s_light[w], spill, spill_cap, stamp, n, gw, nw, edges, verts);
    // (frontier vertex, group) items, then the previous round's spilled
    // descriptors, in one loop with one walk call site
    const uint64_t items = light ? 0 : (uint64_t)n * ms.G;
    for (uint64_t it = gw; it < items + nh; it += nw) {
        uint32_t u, grp, len;
        unsigned long long m;
        uint64_t start;
        if (it < items) {
            u = io.in_list[it / ms.G];
            grp = (uint32_t)(it % ms.G);
            m = 0;
            if (lane == 0) m = atomicExch(&masks[(uint64_t)u * ms.G + grp], 0ull);
            m = __shfl(m, 0, kWave);
            if (m == 0) continue;
            ++verts;
            start = list_start(g, u);
            if (lane == 0) count_hot(g, u);
            const uint32_t deg = g.deg[u];
            len = min(deg, kMEdgeChunk);
            if (deg > len) {
                const uint32_t nch = (deg - len + kMEdgeChunk - 1) / kMEdgeChunk;
                uint32_t base = 0;
                if (lane == 0) base = atomicAdd(io.heavy_cnt, nch);
                base = __shfl(base, 0, kWave);
                for (uint32_t c = lane; c < nch; c += kWave) {
                    const uint32_t lo = len + c * kMEdgeChunk;
                    if (base + c < spill_cap) spill[base + c] = MDesc{start + lo, m, u, min(kMEdgeChunk, deg - lo), grp, 0};
                    else atomicOr(io.stats + 3, 2ull);
                }
            }
        } else {
            const MDesc d = sin[it - items];
            u = d.u;
            grp = d.grp;
            m = d.mask;
            start = d.start;
            len = d.len;
        }
        mwalk<kMode>(g, ms, io, st, stamp, io.round + 1, u, grp, m, start, len, edges);
    }
    stage_flush(st, io);
    // one atomic per counter and block
    if (lane == 0) {
        s_red[w] = edges;
        s_red2[w] = verts;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long a = 0, b = 0;
        for (int i = 0; i < kWavesPerBlock; ++i) {
            a += s_red[i];
            b += s_red2[i];
        }
        if (a) atomicAdd(io.stats + 0, a);
        if (b) atomicAdd(io.stats + 1, b);
    }
}

// Pairs (u, v) processed per wave step (one per item, QS slots each):
// kMode 0 relaxes all slots of inserted edges that survive the batch;
// kMode 1 resets the slots of v whose parent is u for deleted edges.
template <int kMode>
__global__ void __launch_bounds__(kBlock) k_mpairs(MultiState ms, BatchView bv, BatchRanges br, PassIO io,
                                                   uint32_t *stamp, uint32_t stamp_val) {
    __shared__ uint32_t s_out[kWavesPerBlock][kOutStage];
    const uint32_t w = threadIdx.x / kWave;
    const QLane q = qlane(ms);
    OutStage st{s_out[w], 0};
    const uint64_t n = kMode == 0 ? bv.nins : bv.ndel;
    const uint64_t gw = (uint64_t)blockIdx.x * kWavesPerBlock + w, nw = (uint64_t)gridDim.x * kWavesPerBlock;
    for (uint64_t b0 = gw * q.per; b0 < n; b0 += nw * q.per) {
        const uint64_t i = b0 + q.sub;
        bool ok = i < n;
        uint32_t u = 0, v = 0;
        if (ok) {
            if (kMode == 0) {
                u = bv.ins_key[i];
                v = bv.ins_val[i];
                ok = u != kGap;
                if (ok) {
                    const uint32_t db = br.del_b[u], de = br.del_e[u];
                    ok = db == de || !in_sorted_u64(bv.del_key, db, de, ((uint64_t)u << 32) | v);
                }
            } else {
                const uint64_t k = bv.del_key[i];
                ok = k != ~0ull;
                u = (uint32_t)(k >> 32);
                v = (uint32_t)k;
            }
        }
        for (uint32_t grp = 0; grp < ms.G; ++grp) {
            const uint32_t qi = grp * ms.QS + q.slot;
            const uint8_t kind = ms.kind[qi];
            const bool on = ok && kind != kQNone;
            if (kMode == 1) {
                const uint64_t cur = on ? ld_relaxed(&ms.best[(uint64_t)v * ms.KP + qi]) : 0ull;
                const bool child = on && lo32(cur) == u && hi32(cur) != kInf;
                reset_children(ms, io, st, q, grp, v, child, stamp, stamp_val);
            } else {
                const uint32_t du = on ? hi32(ld_relaxed(&ms.best[(uint64_t)u * ms.KP + qi])) : kInf;
                const bool hit = on && du != kInf && mrelax(ms, u, v, grp, q.slot, du, kind);
                mark_items(ms, io, st, q, ms.act, grp, v, hit, stamp, stamp_val);
            }
        }
    }
    stage_flush(st, io);
}

// Edge-centric sweep for the batch: every live edge (u, v) whose target has a
// reset slot relaxes v's reset slots from u (the sweep hits). Same wave tiles
// as k_sweep_edges: the tile's hits are compacted in place in the wave's LDS
// (owner into own[k], target into the dead base words) and relaxed one hit
// per lane, over the reset slots of the target only.
__global__ void __launch_bounds__(kSweepBlock) k_msweep(GraphView g, const unsigned long long *dpre, SweepTiles tiles,
                                                        MultiState ms, PassIO io, uint32_t *stamp, uint32_t stamp_val) {
    __shared__ SweepWave s_wave[kSweepWaves];
    __shared__ uint32_t s_filter[kFilterWords];
    const unsigned long long E = dpre[g.nv];
    const unsigned long long ntiles = (g.nv + E + kSweepTile - 1) / kSweepTile;
    const uint32_t w = threadIdx.x / kWave, lane = lane_id();
    SweepWave &sh = s_wave[w];
    OutStage st{sh.out, 0, kSweepOutStage};
    uint32_t n_edges = 0;
    const uint32_t cs = ms.cshift;
    load_filter(s_filter, ms.coarse, coarse_words(g.nv, cs));
    TilePipe tp;
    tp.tend = min_u64(ntiles, tiles.cap - 1);
    tp.nw = (unsigned long long)gridDim.x * kSweepWaves;
    tp.t = (unsigned long long)blockIdx.x * kSweepWaves + w;
    tp.start(g, dpre, tiles);
    while (tp.t < tp.tend) {
        uint32_t vv[kSweepPerLane];
        const uint32_t v0 = tp.h.v0;
        sweep_tile(g, dpre, tp.h, tp.pre, g.col, 0, 0, g.nv, sh, vv, n_edges);
        tp.advance(g, dpre, tiles);
        bool pass[kSweepPerLane];
#pragma unroll
        for (int j = 0; j < kSweepPerLane; ++j) pass[j] = vv[j] != kGap && coarse_test(s_filter, cs, vv[j]);
        uint32_t word[kSweepPerLane];
#pragma unroll
        for (int j = 0; j < kSweepPerLane; ++j) word[j] = pass[j] ? ms.invbit[vv[j] >> 5] : 0u;
        uint32_t nh = 0;  // wave-uniform
        uint32_t *const hv = reinterpret_cast<uint32_t *>(sh.base);
#pragma unroll
        for (int j = 0; j < kSweepPerLane; ++j) {
            const bool h = pass[j] && ((word[j] >> (vv[j] & 31)) & 1u);
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
        // one lane per hit: relax exactly the reset slots of its target (usually
        // a few of the group's QS), then flag the improved ones and queue v once
        for (uint32_t i0 = 0; i0 < nh; i0 += kWave) {
            const uint32_t i = i0 + lane;
            bool app = false;
            uint32_t v = 0;
            if (i < nh) {
                v = hv[i];
                const uint32_t u = v0 + sh.own[i];
                const uint32_t wt = sg_weight(ms.wmode, ms.wseed, u, v);
                bool any = false;
                for (uint32_t grp = 0; grp < ms.G; ++grp) {
                    uint64_t m = ms.inv[(uint64_t)v * ms.G + grp];
                    uint64_t better = 0;
                    while (m) {
                        const uint32_t slot = (uint32_t)__builtin_ctzll(m);
                        m &= m - 1;
                        const uint32_t qi = grp * ms.QS + slot;
                        const uint8_t kind = ms.kind[qi];
                        if (kind == kQNone) continue;
                        const uint32_t du = hi32(ld_relaxed(&ms.best[(uint64_t)u * ms.KP + qi]));
                        if (du == kInf) continue;
                        const uint32_t cand = du + (kind == kQUnit ? 1u : wt);
                        unsigned long long *p = &ms.best[(uint64_t)v * ms.KP + qi];
                        const unsigned long long nw = pack(cand, u);
                        if (nw >= ld_peek(p)) continue;
                        if (cand < hi32(atomicMin(p, nw))) better |= 1ull << slot;
                    }
                    if (better) {
                        atomicOr(&ms.act[(uint64_t)v * ms.G + grp], better);
                        any = true;
                    }
                }
                app = any && atomicExch(&stamp[v], stamp_val) != stamp_val;
            }
            stage_push(app, v, st, io);
        }
        __builtin_amdgcn_wave_barrier();
    }
    stage_flush(st, io);
}

// Clear the masks of the vertices with a reset slot (one thread per bitmap word).
// The vertex bitmap and its coarse filter are cleared on the way.
__global__ void k_mclear_inv(uint32_t nv, uint32_t G, uint32_t *invbit, uint32_t *coarse, uint32_t cshift,
                             unsigned long long *inv) {
    const uint32_t nw = (nv + 31) / 32;
    for (uint64_t wd = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; wd < nw; wd += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t bits = invbit[wd];
        if (bits) invbit[wd] = 0;
        if ((wd & ((1ull << cshift) - 1)) == 0) coarse[wd >> cshift] = 0;
        while (bits) {
            const uint32_t b = __builtin_ctz(bits);
            bits &= bits - 1;
            const uint64_t v = wd * 32 + b;
            for (uint32_t grp = 0; grp < G; ++grp) inv[v * G + grp] = 0ull;
        }
    }
}

// Initial state: every slot unreached, each query's source at distance 0 and
// active; the sources form the first frontier (each vertex once).
__global__ void k_minit_sources(MultiState ms, const uint32_t *src, uint32_t nq, uint32_t *stamp, uint32_t stamp_val,
                                uint32_t *list, uint32_t *cnt, uint32_t cap, unsigned long long *overflow) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nq; i += gridDim.x * blockDim.x) {
        const uint32_t s = src[i], grp = i / ms.QS, slot = i % ms.QS;
        ms.best[(uint64_t)s * ms.KP + i] = pack(0, kNone);
        atomicOr(&ms.act[(uint64_t)s * ms.G + grp], 1ull << slot);
        if (atomicExch(&stamp[s], stamp_val) != stamp_val) {
            const uint32_t pos = atomicAdd(cnt, 1u);
            if (pos < cap) list[pos] = s;
            else atomicOr(overflow, 1ull);
        }
    }
}

// ---- continuous batching: queries join and leave fixed slots between round batches ----
// New slots: every vertex's word unreached, no pending or reset bits.
__global__ void k_slot_reset(MultiState ms, uint32_t nv, const uint32_t *slots, uint32_t n) {
    for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < nv; v += (uint64_t)gridDim.x * blockDim.x) {
        for (uint32_t i = 0; i < n; ++i) {
            const uint32_t s = slots[i], grp = s / ms.QS;
            const unsigned long long keep = ~(1ull << (s % ms.QS));
            ms.best[v * ms.KP + s] = ~0ull;
            if (ms.act[v * ms.G + grp] & ~keep) ms.act[v * ms.G + grp] &= keep;
        }
    }
}
// Each new query's source at distance 0 and active, appended once to the
// frontier the next round reads (stamp_val = that round).
__global__ void k_slot_seed(MultiState ms, const uint32_t *slots, const uint32_t *src, uint32_t n, uint32_t *stamp,
                            uint32_t stamp_val, uint32_t *list, uint32_t *cnt, uint32_t cap, unsigned long long *overflow) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t s = src[i], q = slots[i], grp = q / ms.QS, slot = q % ms.QS;
        ms.best[(uint64_t)s * ms.KP + q] = pack(0, kNone);
        atomicOr(&ms.act[(uint64_t)s * ms.G + grp], 1ull << slot);
        if (atomicExch(&stamp[s], stamp_val) != stamp_val) {
            const uint32_t pos = atomicAdd(cnt, 1u);
            if (pos < cap) list[pos] = s;
            else atomicOr(overflow, 1ull);
        }
    }
}
// Slots with work still pending (per group): the masks of the frontier the
// next round reads and of the descriptors the last round spilled.
__global__ void __launch_bounds__(kBlock) k_slot_pending(MultiState ms, const uint32_t *list, const uint32_t *cnt,
                                                         uint32_t cap, const MDesc *spill, const uint32_t *scnt,
                                                         uint32_t scap, unsigned long long *out) {
    const uint32_t n = min(*cnt, cap), ns = min(*scnt, scap);
    for (uint32_t grp = 0; grp < ms.G; ++grp) {
        unsigned long long m = 0;
        for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
            m |= ms.act[(uint64_t)list[i] * ms.G + grp];
        for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < ns; i += (uint64_t)gridDim.x * blockDim.x)
            if (spill[i].grp == grp) m |= spill[i].mask;
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) m |= __shfl_xor(m, d, kWave);
        if (lane_id() == 0 && m) atomicOr(&out[grp], m);
    }
}

// Column q of the state: distance (what = 0) or parent (what = 1).
__global__ void k_mextract(uint32_t nv, const unsigned long long *best, uint32_t KP, uint32_t q, uint32_t *out, int what) {
    for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < nv; v += (uint64_t)gridDim.x * blockDim.x) {
        const unsigned long long b = best[v * KP + q];
        out[v] = what ? lo32(b) : hi32(b);
    }
}

}  // namespace sg

#endif
