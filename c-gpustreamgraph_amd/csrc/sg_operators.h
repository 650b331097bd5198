// Per-algorithm operator API of the engine, and the four operators.
//
// The method names and meanings are the reference's AppBase
// (include/framework/variants/api.cuh:22-160); dispatch is static (every engine
// kernel is a template over the operator type), so there are no virtuals.
//
//   GetInitValue(v)                  initial committed value            api.cuh:41-42
//   GetInitBuffer(v)                 initial pending buffer             api.cuh:44-45
//   GetIdentityElement()             "nothing pending"                  api.cuh:47-48
//   CombineValueBuffer(v,&val,&buf,deg) -> {push, scheduled}           api.cuh:56-59
//                                    commit the buffer into the value and say what
//                                    to push along out-edges
//   CombineValueBufferAmend(v,type,&val,&buf,deg) -> {push, scheduled} api.cuh:60-63
//                                    re-push the committed value scaled by type
//                                    (PR cancel = -1 / compensate = +1)
//   AccumulateBuffer(src,dst,w,&buf_dst,push) -> dst became active     api.cuh:65-78
//   AccumulateBuffer_del(src,dst,&buf,&val) -> dst was reset           api.cuh:93-103
//   IsActiveNode(v,buf,val), IsActiveNode_INC(v,buf,val)               api.cuh:111-122
//   MergeBuffer(dst,&buf_dst,acc) -> dst improved (pull shape)         pull_functor.h:137-175
//                                    the destination side of a pull round: what
//                                    dst's in-edges offered, accumulated with
//                                    AccumulateBuffer into a word of its own that
//                                    started at GetIdentityElement(), folded into
//                                    dst's buffer (atomically: the in-list of a
//                                    vertex can be split between waves)
//
// The contract the public operator header (include/sgraph_operator.h) states
// for a caller's own operator is the one these four follow.
//
// Differences from the reference, each deliberate:
//  * The traversal operators keep parent and pending value in one 64-bit buffer
//    word (value << 32 | parent). One 64-bit atomic then updates both, where the
//    reference does atomicMin + an atomicCAS loop on a separate parent array
//    (samples/hybrid_sssp/hybrid_sssp.cu:97-118). The reference's TValue*
//    p_parent argument is therefore folded into p_buffer.
//  * AccumulateBuffer returns whether the destination's pending value strictly
//    improved (it must then be scheduled); the reference returns an unused
//    status code. The ACCUMULATE_* codes are kept for source compatibility.
//  * CombineValueBuffer receives the out-degree instead of reading it from the
//    graph object the reference stores in the operator (hybrid_pr.cu:79).
//  * CC's AccumulateBuffer_del resets a child to its own label (GetInitBuffer),
//    not to UINT32_MAX as hybrid_cc.cu:145-150 does; the latter loses labels
//    of vertices whose own id is their minimum ancestor.
#ifndef SG_OPERATORS_H
#define SG_OPERATORS_H

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sg {
namespace api {

// api.cuh:26-29
constexpr int ACCUMULATE_SUCCESS_BREAK = 0;
constexpr int ACCUMULATE_SUCCESS_CONTINUE = 1;
constexpr int ACCUMULATE_FAILURE_BREAK = 2;
constexpr int ACCUMULATE_FAILURE_CONTINUE = 3;

constexpr uint32_t kInf = 0xFFFFFFFFu;   // IDENTITY_ELEMENT of SSSP/BFS/CC
constexpr uint32_t kNone = 0xFFFFFFFFu;  // no parent

template <class T>
struct Push {
    T value;
    bool scheduled;
};

__device__ __forceinline__ uint64_t pack(uint32_t hi, uint32_t lo) {
    return ((uint64_t)hi << 32) | lo;
}
__device__ __forceinline__ uint32_t hi32(uint64_t x) { return (uint32_t)(x >> 32); }
__device__ __forceinline__ uint32_t lo32(uint64_t x) { return (uint32_t)x; }

// ---------------------------------------------------------------------------
// Traversal operators share the packed (value << 32 | parent) buffer and the
// "commit if better" combine (hybrid_sssp.cu:77-89, hybrid_cc.cu:62-77).
// ---------------------------------------------------------------------------
struct TraversalBase {
    using value_t = uint32_t;
    using buffer_t = unsigned long long;
    using push_t = uint32_t;
    static constexpr bool kTraversal = true;

    __host__ __device__ buffer_t GetIdentityElement() const { return ((uint64_t)kInf << 32) | kNone; }

    __device__ bool IsActiveNode(uint32_t, buffer_t buf, value_t val) const { return hi32(buf) < val; }
    __device__ bool IsActiveNode_INC(uint32_t v, buffer_t buf, value_t val) const {
        return IsActiveNode(v, buf, val);
    }

    // Parent of v recorded in its buffer.
    __device__ static uint32_t Parent(buffer_t buf) { return lo32(buf); }
    // Best value known for v (pending or committed).
    __device__ static uint32_t Best(buffer_t buf) { return hi32(buf); }
};

// Lexicographic atomic minimum of (value, parent): dist strictly increases
// along parent edges when every weight is >= 1, so parents form a forest.
__device__ __forceinline__ bool accumulate_min(unsigned long long *p_buffer, uint32_t src, uint32_t cand) {
    const unsigned long long nv = pack(cand, src);
    const unsigned long long cur = __hip_atomic_load(p_buffer, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (nv >= cur) return false;
    const unsigned long long old = atomicMin(p_buffer, nv);
    return cand < hi32(old);
}

struct SSSP : TraversalBase {
    uint32_t m_source_node;  // hybrid_sssp.cu:31
    static constexpr bool kWeighted = true;

    __device__ value_t GetInitValue(uint32_t) const { return kInf; }
    __device__ buffer_t GetInitBuffer(uint32_t v) const {
        return v == m_source_node ? pack(0, kNone) : pack(kInf, kNone);
    }
    // hybrid_sssp.cu:77-89: commit a smaller pending distance and push it
    __device__ Push<push_t> CombineValueBuffer(uint32_t, value_t *p_value, buffer_t *p_buffer, uint32_t) const {
        const uint32_t b = hi32(*p_buffer);
        if (*p_value > b) { *p_value = b; return {b, true}; }
        return {b, false};
    }
    // hybrid_sssp.cu:68-74: re-push the current buffer (full sweep after updates)
    __device__ Push<push_t> CombineValueBufferAmend(uint32_t, int, value_t *, buffer_t *p_buffer, uint32_t) const {
        const uint32_t b = hi32(*p_buffer);
        return {b, b != kInf};
    }
    // Cheap pre-test against a peeked buffer word: can (pushed + w, src) win?
    __device__ bool MayImprove(uint32_t src, uint32_t weight, push_t pushed, uint64_t peek) const {
        return pushed != kInf && pack(pushed + weight, src) < peek;
    }
    // hybrid_sssp.cu:97-118: atomicMin(dst, src_dist + w) and parent = src
    __device__ bool AccumulateBuffer(uint32_t src, uint32_t, uint32_t weight, buffer_t *p_buffer, push_t pushed) const {
        if (pushed == kInf) return false;
        return accumulate_min(p_buffer, src, pushed + weight);
    }
    // hybrid_sssp.cu:154-167: reset a child whose tree edge came from src
    __device__ bool AccumulateBuffer_del(uint32_t src, uint32_t dst, buffer_t *p_buffer, value_t *p_value) const {
        if (lo32(*p_buffer) != src) return false;
        *p_buffer = GetInitBuffer(dst);
        *p_value = GetInitValue(dst);
        return true;
    }
    // What v pushes along an edge right now (edges added by an update).
    __device__ push_t PushValue(buffer_t buf) const { return hi32(buf); }
    // pull shape: the best (distance, parent) the in-edges offered
    __device__ bool MergeBuffer(uint32_t, buffer_t *p_buffer, buffer_t acc) const {
        return accumulate_min(p_buffer, lo32(acc), hi32(acc));
    }
};

struct BFS : TraversalBase {
    uint32_t m_source_node;  // hybrid_bfs.cu:29
    static constexpr bool kWeighted = false;

    __device__ value_t GetInitValue(uint32_t) const { return kInf; }
    __device__ buffer_t GetInitBuffer(uint32_t v) const {
        return v == m_source_node ? pack(0, kNone) : pack(kInf, kNone);
    }
    // hybrid_bfs.cu:76-92: commit the level and push level + 1
    __device__ Push<push_t> CombineValueBuffer(uint32_t, value_t *p_value, buffer_t *p_buffer, uint32_t) const {
        const uint32_t b = hi32(*p_buffer);
        if (*p_value > b) { *p_value = b; return {b + 1, true}; }
        return {b == kInf ? kInf : b + 1, false};
    }
    __device__ Push<push_t> CombineValueBufferAmend(uint32_t, int, value_t *, buffer_t *p_buffer, uint32_t) const {
        const uint32_t b = hi32(*p_buffer);
        return {b == kInf ? kInf : b + 1, b != kInf};
    }
    __device__ bool MayImprove(uint32_t src, uint32_t, push_t pushed, uint64_t peek) const {
        return pushed != kInf && pack(pushed, src) < peek;
    }
    // hybrid_bfs.cu:102-109 (atomicMin of the pushed level), plus the parent word
    __device__ bool AccumulateBuffer(uint32_t src, uint32_t, uint32_t, buffer_t *p_buffer, push_t pushed) const {
        if (pushed == kInf) return false;
        return accumulate_min(p_buffer, src, pushed);
    }
    __device__ bool AccumulateBuffer_del(uint32_t src, uint32_t dst, buffer_t *p_buffer, value_t *p_value) const {
        if (lo32(*p_buffer) != src) return false;
        *p_buffer = GetInitBuffer(dst);
        *p_value = GetInitValue(dst);
        return true;
    }
    __device__ push_t PushValue(buffer_t buf) const { return hi32(buf) == kInf ? kInf : hi32(buf) + 1; }
    __device__ bool MergeBuffer(uint32_t, buffer_t *p_buffer, buffer_t acc) const {
        return accumulate_min(p_buffer, lo32(acc), hi32(acc));
    }
};

// Directed min-ancestor labels: label(v) = min{u : u reaches v}.
struct CC : TraversalBase {
    static constexpr bool kWeighted = false;

    __device__ value_t GetInitValue(uint32_t) const { return kInf; }
    __device__ buffer_t GetInitBuffer(uint32_t v) const { return pack(v, kNone); }  // hybrid_cc.cu:48-52
    // hybrid_cc.cu:62-77
    __device__ Push<push_t> CombineValueBuffer(uint32_t, value_t *p_value, buffer_t *p_buffer, uint32_t) const {
        const uint32_t b = hi32(*p_buffer);
        if (*p_value > b) { *p_value = b; return {b, true}; }
        return {b, false};
    }
    __device__ Push<push_t> CombineValueBufferAmend(uint32_t, int, value_t *, buffer_t *p_buffer, uint32_t) const {
        return {hi32(*p_buffer), true};
    }
    __device__ bool MayImprove(uint32_t, uint32_t, push_t pushed, uint64_t peek) const {
        return pushed < hi32(peek);
    }
    // hybrid_cc.cu:96-110: atomicMin of the label; the parent is recorded only
    // on a strict label decrease (a CAS on the packed word), so parent edges
    // always point to a vertex that held the label earlier: a forest.
    __device__ bool AccumulateBuffer(uint32_t src, uint32_t, uint32_t, buffer_t *p_buffer, push_t pushed) const {
        return MergeBuffer(0, p_buffer, pack(pushed, src));
    }
    // pull shape: the smallest label an in-edge offered, with that in-neighbour
    // as the parent, on a strict decrease only (as above)
    __device__ bool MergeBuffer(uint32_t, buffer_t *p_buffer, buffer_t acc) const {
        unsigned long long cur = __hip_atomic_load(p_buffer, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        while (hi32(acc) < hi32(cur)) {
            const unsigned long long prev = atomicCAS(p_buffer, cur, acc);
            if (prev == cur) return true;
            cur = prev;
        }
        return false;
    }
    __device__ bool AccumulateBuffer_del(uint32_t src, uint32_t dst, buffer_t *p_buffer, value_t *p_value) const {
        if (lo32(*p_buffer) != src) return false;
        *p_buffer = GetInitBuffer(dst);
        *p_value = GetInitValue(dst);
        return true;
    }
    __device__ push_t PushValue(buffer_t buf) const { return hi32(buf); }
};

// ---------------------------------------------------------------------------
// PageRank, delta-push form (samples/hybrid_pr/hybrid_pr.cu:25-152): value is
// the accumulated rank, buffer the residual still to be pushed.
// ---------------------------------------------------------------------------
struct PageRank {
    using value_t = float;
    using buffer_t = float;
    using push_t = float;
    static constexpr bool kTraversal = false;
    static constexpr bool kWeighted = false;
    double m_alpha;  // ALPHA, hybrid_pr_common.h:38
    double m_error;  // FLAGS_error, hybrid_pr/main.cu:37

    __device__ value_t GetInitValue(uint32_t) const { return 0.0f; }                      // hybrid_pr.cu:43-47
    __device__ buffer_t GetInitBuffer(uint32_t) const { return (float)(1.0 - m_alpha); }  // hybrid_pr.cu:49-54
    __host__ __device__ buffer_t GetIdentityElement() const { return 0.0f; }

    // hybrid_pr.cu:66-85: take the whole residual, add it to the rank, push
    // alpha * residual / out-degree
    __device__ Push<push_t> CombineValueBuffer(uint32_t, value_t *p_value, buffer_t *p_buffer, uint32_t deg) const {
        const float x = atomicExch(p_buffer, 0.0f);
        if (x == 0.0f) return {0.0f, false};
        *p_value += x;
        if (deg == 0) return {0.0f, false};  // dangling: mass dropped
        return {(float)(m_alpha * (double)x / (double)deg), true};
    }
    // hybrid_pr.cu:88-106: push alpha * type * value / out-degree
    __device__ Push<push_t> CombineValueBufferAmend(uint32_t, int type, value_t *p_value, buffer_t *, uint32_t deg) const {
        const float x = *p_value * (float)type;
        if (x == 0.0f || deg == 0) return {0.0f, false};
        return {(float)(m_alpha * (double)x / (double)deg), true};
    }
    // hybrid_pr.cu:109-122: residual += pushed; the destination becomes active
    // when |residual| crosses the threshold (the crossing rule of
    // hybrid_pr_host.cu:96-99, extended to negative residuals)
    __device__ bool AccumulateBuffer(uint32_t, uint32_t, uint32_t, buffer_t *p_buffer, push_t pushed) const {
        const float old = unsafeAtomicAdd(p_buffer, pushed);
        const float now = old + pushed;
        const float e = (float)m_error;
        return fabsf(now) > e && !(fabsf(old) > e);
    }
    __device__ bool AccumulateBuffer_del(uint32_t, uint32_t, buffer_t *, value_t *) const { return false; }
    __device__ bool IsActiveNode(uint32_t, buffer_t buf, value_t) const { return buf > (float)m_error; }      // :126-129
    __device__ bool IsActiveNode_INC(uint32_t, buffer_t buf, value_t) const { return buf < -(float)m_error; }  // :133-136
};

}  // namespace api
}  // namespace sg

#endif
