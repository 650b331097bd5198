// Widest path from one source, as a caller's own operator: built outside the
// library with include/sgraph_operator.h into libsg_widest_path.so, and run
// through the ordinary C ABI (sg_create_widest_path returns an sg_engine).
// No counterpart in the reference's samples; it plays the part of a user's
// App struct given to Engine<..., TAppImpl> (framework.cuh:77).
//
// width(src) = UINT32_MAX (unbounded), width(v) = the largest, over paths
// src -> v, of the smallest edge weight on the path; 0 when v is unreachable.
// Edge weights are the engine's weight function (sg_config.weight_mode).
//
// The pending word holds key = ~width, so that a smaller key is better as in
// the packed (key << 32 | parent) words of the built-in programs, and the
// parent. The offer along an edge of weight w is key' = max(key(u), ~w), never
// better than u's own key. Widths tie along paths, so a word changes only on
// a strict improvement (a CAS loop, as CC's labels): a parent then always held
// its offer before its child took it, and the parent edges form a forest.
#include "sgraph_operator.h"

namespace {

using sg::api::hi32;
using sg::api::kInf;
using sg::api::kNone;
using sg::api::lo32;
using sg::api::pack;

struct WidestPath : sg::api::TraversalBase {
    uint32_t m_source_node;
    static constexpr bool kWeighted = true;

    __device__ static uint32_t offer(uint32_t key, uint32_t w) { return key > ~w ? key : ~w; }

    __device__ value_t GetInitValue(uint32_t) const { return 0u; }  // width 0: not reached
    __device__ buffer_t GetInitBuffer(uint32_t v) const {
        return v == m_source_node ? pack(0u, kNone) : pack(kInf, kNone);
    }
    // commit a wider pending width, offer its key
    __device__ sg::api::Push<push_t> CombineValueBuffer(uint32_t, value_t *p_value, buffer_t *p_buffer, uint32_t) const {
        const uint32_t k = hi32(*p_buffer);
        if (~k > *p_value) {
            *p_value = ~k;
            return {k, true};
        }
        return {k, false};
    }
    __device__ sg::api::Push<push_t> CombineValueBufferAmend(uint32_t, int, value_t *, buffer_t *p_buffer, uint32_t) const {
        const uint32_t k = hi32(*p_buffer);
        return {k, k != kInf};
    }
    // the pending width is larger than the committed one
    __device__ bool IsActiveNode(uint32_t, buffer_t buf, value_t val) const { return ~hi32(buf) > val; }
    __device__ bool IsActiveNode_INC(uint32_t v, buffer_t buf, value_t val) const { return IsActiveNode(v, buf, val); }
    __device__ bool MayImprove(uint32_t, uint32_t w, push_t pushed, uint64_t peek) const {
        return pushed != kInf && offer(pushed, w) < hi32(peek);
    }
    __device__ bool AccumulateBuffer(uint32_t src, uint32_t, uint32_t w, buffer_t *p_buffer, push_t pushed) const {
        if (pushed == kInf) return false;
        return MergeBuffer(0u, p_buffer, pack(offer(pushed, w), src));
    }
    // pull shape: the widest offer of the in-edges, with its in-neighbour as
    // the parent, on a strict improvement only
    __device__ bool MergeBuffer(uint32_t, buffer_t *p_buffer, buffer_t acc) const {
        unsigned long long cur = __hip_atomic_load(p_buffer, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        while (hi32(acc) < hi32(cur)) {
            const unsigned long long prev = atomicCAS(p_buffer, cur, acc);
            if (prev == cur) return true;
            cur = prev;
        }
        return false;
    }
    // the tree edge src -> dst is gone: dst is no longer reached through it
    __device__ bool AccumulateBuffer_del(uint32_t src, uint32_t dst, buffer_t *p_buffer, value_t *p_value) const {
        if (lo32(*p_buffer) != src) return false;
        *p_buffer = GetInitBuffer(dst);
        *p_value = GetInitValue(dst);
        return true;
    }
    __device__ push_t PushValue(buffer_t buf) const { return hi32(buf); }
};

}  // namespace

SG_OPERATOR_ENGINE(sg_create_widest_path, WidestPath)
