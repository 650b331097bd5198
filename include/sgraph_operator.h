/* sgraph_operator.h - a caller's own vertex program compiled into an engine.
 *
 * Reference interface: the operator is a template argument of the engine,
 * sepgraph::engine::Engine<TValue, TBuffer, TWeight, TAppImpl>
 * (include/framework/framework.cuh:77), and implements the AppBase methods
 * (include/framework/variants/api.cuh:22-160); a sample compiles the
 * framework's headers with its own App struct (samples/hybrid_sssp/hybrid_sssp.cu).
 *
 * Here a caller's HIP source file includes this header, defines its operator
 * struct and names a create function with SG_OPERATOR_ENGINE. The file is
 * built for gfx950 into a shared library of its own, linked to libsgraph.so:
 *
 *     make -C c-gpustreamgraph_amd operator OP=/path/widest.hip OPLIB=/path/libwidest.so
 *
 * The create function returns an ordinary sg_engine handle: every sg_* call of
 * sgraph.h (load, init with a source, apply, gather, checkpoint, stats,
 * sg_destroy) drives it, and sg_last_error(NULL) explains a failed create.
 * The engine is the one the built-in SSSP/BFS/CC programs run on: device PMA
 * updates, invalidation of the subtrees under deleted tree edges, recompute
 * seeded from the in-lists (SG_FLAG_IN_EDGES) or a sweep, host-memory edges,
 * the HBM edge cache, direction-optimising rounds.
 *
 * The operator: a struct (static dispatch, no virtuals) for a min-type
 * program, i.e. one whose pending state per vertex is a 64-bit word that
 * only improves until a deletion resets it:
 *
 *   using value_t = uint32_t;            committed value per vertex
 *   using buffer_t = unsigned long long; pending word: (key << 32 | parent)
 *   using push_t = uint32_t;             what a vertex offers along its out-edges
 *   static constexpr bool kTraversal = true;
 *   static constexpr bool kWeighted;     offers depend on the edge weight
 *                                        (sg_config.weight_mode)
 *   uint32_t m_source_node;              optional: the program takes one source
 *                                        from sg_init (hybrid_sssp.cu:31)
 *
 *   value_t  GetInitValue(v)                                    api.cuh:41-42
 *   buffer_t GetInitBuffer(v)                                   api.cuh:44-45
 *   buffer_t GetIdentityElement()      "no offer" (host and device) api.cuh:47-48
 *   Push<push_t> CombineValueBuffer(v, &val, &buf, outdeg)      api.cuh:56-59
 *            commit buf into val when it is better; {offer, true} schedules
 *            the offer along v's out-edges
 *   Push<push_t> CombineValueBufferAmend(v, type, &val, &buf, outdeg)
 *            re-offer the current buffer (recompute seeds)      api.cuh:60-63
 *   bool MayImprove(src, w, offer, peek)  cheap test against a relaxed read
 *            of the destination's word; false must mean AccumulateBuffer
 *            would not improve it
 *   bool AccumulateBuffer(src, dst, w, &buf_dst, offer)         api.cuh:65-78
 *            fold one offer into dst's word atomically; true iff the word
 *            strictly improved (dst is then scheduled)
 *   bool AccumulateBuffer_del(src, dst, &buf, &val)             api.cuh:93-103
 *            dst's parent is src and the edge is gone: reset dst (to
 *            GetInitBuffer / GetInitValue) and return true
 *   bool IsActiveNode(v, buf, val), IsActiveNode_INC(v, buf, val)
 *            the pending word is better than the committed value api.cuh:111-122
 *   push_t PushValue(buf)               the offer of a vertex whose word is
 *            buf; api::kInf = nothing to offer
 *   static uint32_t Parent(buf)         the in-neighbour whose offer set buf
 *   bool MergeBuffer(dst, &buf_dst, acc)  optional, the pull shape
 *            (pull_functor.h:137-175): fold acc - the offers of several of
 *            dst's in-edges, accumulated with AccumulateBuffer into a word that
 *            started at GetIdentityElement() - into dst's word atomically; true
 *            iff it strictly improved. Operators with it take pull rounds.
 *
 * What the engine relies on (the four built-in programs follow it):
 *   - a word changes only to a strictly better one, except through
 *     AccumulateBuffer_del; the fixed point does not depend on the order of
 *     the offers (a lexicographic minimum, or strict improvements only);
 *   - Parent edges form a forest: a vertex's parent held its offer before
 *     the vertex took it (record the parent on strict improvements only when
 *     offers can tie along a path, as CC does);
 *   - offers are monotone along edges (an offer is never better than the
 *     offering vertex's own key), so a reset subtree can only be re-reached
 *     from outside it.
 * sg::api::TraversalBase gives the packed-word helpers (value_t, buffer_t,
 * push_t, identity, IsActiveNode for "smaller key is better", Parent).
 *
 * C++17, gfx950 (hipcc); the package's csrc/ directory on the include path.
 */
#ifndef SGRAPH_OPERATOR_H
#define SGRAPH_OPERATOR_H

#include "sgraph.h"
#include "sg_engine_core.h"

namespace sg {

// The template entry point: an engine running Op, configured by cfg (algo is
// ignored and set to SG_ALGO_CUSTOM), with op as the operator's initial state
// (its parameters; m_source_node is set by sg_init). Returns SG_EINVAL for a
// bad configuration, SG_EDEVICE without a usable device.
template <class Op>
sg_status create_engine(const sg_config *cfg, const Op &op, sg_engine **out) noexcept {
    static_assert(std::is_same<typename Op::value_t, uint32_t>::value, "value_t must be uint32_t");
    static_assert(std::is_same<typename Op::buffer_t, unsigned long long>::value, "buffer_t must be a 64-bit word");
    static_assert(std::is_same<typename Op::push_t, uint32_t>::value, "push_t must be uint32_t");
    static_assert(Op::kTraversal, "the engine runs min-type (traversal) operators");
    if (!cfg || !out) {
        g_create_error = "null argument";
        return SG_EINVAL;
    }
    *out = nullptr;
    if (cfg->weight_mode > SG_WEIGHT_ONE) {
        g_create_error = "unknown weight mode";
        return SG_EINVAL;
    }
    try {
        sg_config c = *cfg;
        c.algo = SG_ALGO_CUSTOM;
        std::unique_ptr<TraversalEngine<Op>> impl(new TraversalEngine<Op>(c));
        impl->prog_.op = op;
        impl->init_device();
        *out = new sg_engine{std::move(impl)};
        return SG_OK;
    } catch (const Error &e) {
        g_create_error = e.msg;
        return e.status;
    } catch (const std::bad_alloc &) {
        g_create_error = "host allocation failed";
        return SG_ENOMEM;
    } catch (...) {
        g_create_error = "unexpected exception";
        return SG_EDEVICE;
    }
}

}  // namespace sg

/* Export `sg_status fn(const sg_config *cfg, sg_engine **out)` from the
 * operator's shared library: a default-constructed Op. */
#define SG_OPERATOR_ENGINE(fn, Op)                                                               \
    extern "C" __attribute__((visibility("default"))) sg_status fn(const sg_config *cfg,        \
                                                                   sg_engine **out) {           \
        return ::sg::create_engine<Op>(cfg, Op{}, out);                                          \
    }

#endif
