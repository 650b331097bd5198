// MI355X streaming-graph engine: device PMA store, frontier rounds, the static
// solve and the per-snapshot incremental phases of the four operators, behind
// the C ABI of include/sgraph.h.
//
// Reference path (C-GpuStreamGraph):
//   Engine                          include/framework/framework.cuh:77-2182
//   host PMA mutation               include/groute/graphs/csr_graph.cuh:658-979
//   PR cancel/compensate            framework.cuh:1110-1301
//   traversal del/add               framework.cuh:1507-1655
//   kernels                         include/framework/variants/driver.cuh, sync_push_dd.cuh
// Design differences (each kept deliberately):
//   * the graph is updated on the GPU in batches (sorted, merged in place,
//     windows re-spread) instead of one host insert at a time;
//   * deletion invalidation is iterated to a fixed point (the reference runs one
//     round, framework.cuh:1629-1648 with the counts of 1913-1925 commented out);
//   * after invalidation only edges into invalidated vertices are re-pushed, and
//     PR's cancel/compensate only touches sources whose out-lists changed; both
//     equal the reference's full sweeps in exact arithmetic;
//   * frontiers are deduplicated with per-vertex round stamps and appended with
//     one atomic per wave; the host checks convergence every rounds_per_sync
//     rounds instead of one blocking copy per segment per round.
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
#include "sg_engine_core.h"

// Waves of a pull pass, in 64ths, that take the tiles beyond the in-store's
// HBM mirror (a tuning constant; the host link saturates with few waves).
#ifndef SG_PULL_HOST_WAVES64
#define SG_PULL_HOST_WAVES64 8
#endif

static_assert(sizeof(sg_config) == 104, "sg_config layout");
static_assert(sizeof(sg_stats) == 240, "sg_stats layout");

namespace sg {

thread_local std::string g_create_error = "no error";

// ---------------------------------------------------------------------------
// PageRank
// ---------------------------------------------------------------------------
class PageRankEngine : public Engine {
public:
    // With the edges in host memory, every dense round pulls over all
    // in-edges, while push rounds touch few lists: the -cache HBM budget then
    // extends the in-store's HBM copy (read by every pull round) instead of
    // caching hot out-lists of the forward store (the reference's cache holds
    // out-lists for its push rounds, graph_datum.cuh:151-163). The HBM spent
    // on edges stays the edge budget plus the cache.
    uint64_t in_cache_bytes_ = 0;
    explicit PageRankEngine(const sg_config &c) : Engine(c) {
        if (c.pull_divisor != 0xFFFFFFFFu && c.cache_level) {
            in_cache_bytes_ = 4ull * kCacheSlotsByLevel[std::min<uint32_t>(c.cache_level, 4)];
            cfg_.cache_level = 0;
        }
    }
    bool reads_old_lists() const override { return true; }
    api::PageRank op_{};
    float *value_ = nullptr, *buffer_ = nullptr;
    double *d_sum_ = nullptr;
    // pull rounds (cfg.pull_divisor != UINT32_MAX): the in-edge store, the
    // frontier bitmaps and pushes of dense rounds, the partial sums of cut
    // in-lists
    std::unique_ptr<GraphStore> rev_;
    uint32_t *fb_[2] = {nullptr, nullptr};
    uint32_t *pv_ = nullptr;
    float *accg_ = nullptr;

    uint64_t aux_device_bytes() const override { return rev_ ? rev_->dev_bytes_ : 0; }
    uint64_t aux_host_bytes() const override { return rev_ ? rev_->host_bytes_ : 0; }
    uint64_t in_capacity() const override { return rev_ ? rev_->cap_ : 0; }
    void in_mirror(uint64_t &slots, uint64_t &live) override {
        slots = live = 0;
        if (rev_) rev_->mirror_counts(slots, live);
    }

    // Whether the engine keeps an in-edge store.
    virtual bool wants_in_store() const { return cfg_.pull_divisor != 0xFFFFFFFFu; }
    void after_load(const uint64_t *row, const uint32_t *col) override {
        if (!wants_in_store()) return;
        const bool pulling = cfg_.pull_divisor != 0xFFFFFFFFu;
        rev_ = make_in_store(*this, row, col, in_cache_bytes_, pulling, pulling);
        if (cfg_.pull_divisor == 0xFFFFFFFFu) return;  // the store serves another program only
        const uint64_t words = ((uint64_t)nv_ + 31) / 32;
        for (auto &b : fb_) {
            b = dalloc<uint32_t>(words);
            fill<uint32_t>(b, words, 0u);
        }
        pv_ = dalloc<uint32_t>(nv_);
        accg_ = dalloc<float>(nv_);
        fill<uint32_t>((uint32_t *)accg_, nv_, 0u);
        sync();
    }

    // Frontier out-edges above which a round is a pull. A resident pull reads
    // every in-edge at HBM speed without atomics (about 40x cheaper per edge
    // than an atomic push): it pays above E / 24. With the edges in host
    // memory a push reads the frontier's out-lists over the host link while a
    // pull reads the in-store mostly from its HBM copy; measured at 2^27
    // vertices, a pull pays above E / 8.
    unsigned long long dense_thr() const {
        const uint64_t div = cfg_.pull_divisor ? cfg_.pull_divisor : (rev_ && rev_->host_edges_ ? 8 : 24);
        return std::max<unsigned long long>(1ull, ne_ / div);
    }
    bool pulls() const { return pv_ != nullptr; }
    PullTake take_state(uint32_t R) const { return PullTake{fb_[R & 1], pv_, rev_ ? rev_->relabel_ : nullptr}; }
    // Round R's frontier bitmap starts empty (it last served round R - 2).
    void clear_take(uint32_t R) {
        if (!pulls()) return;
        const uint64_t words = ((uint64_t)nv_ + 31) / 32;
        round_launch(k_fill_u32, dim3(grid_for(words, 256, 2048)), dim3(256), fb_[R & 1], words, 0u);
    }
    PassIO pr_io(PassIO io) const {
        if (!pulls()) return io;
        const uint32_t R = io.round;
        io.ecnt_in = ecnt_ + (R % 3);
        io.ecnt_out = ecnt_ + ((R + 1) % 3);
        io.ecnt_zero = ecnt_ + ((R + 2) % 3);
        io.fdeg = deg_;
        io.dense_thr = dense_thr();
        return io;
    }
    // Share of a pull pass's waves that take the tiles beyond the HBM mirror
    // (64ths): the host link saturates with far fewer waves than the HBM
    // gathers need.
    static constexpr uint32_t kHostWaves64 = SG_PULL_HOST_WAVES64;
    PullPR pull_state(uint32_t R, bool split = true) const {
        return PullPR{fb_[R & 1], (const float *)pv_, buffer_, accg_, stamp_, R, (float)cfg_.pr_epsilon,
                      split && rev_ ? rev_->tile_split() : nullptr, kHostWaves64};
    }

    // Pull pass of round R over the whole in-edge store (resident store).
    // force = 0: the kernels return unless the round is dense (device flag).
    void pull_resident(const PassIO &io, int force, bool timed) {
        const uint64_t nt = rev_->ntiles();
        const uint32_t grid = sweep_grid(k_pr_pull, nt);
        begin_timed(timed && !batch_, SG_KCLASS_PULL);
        round_launch(k_pr_pull, dim3(grid), dim3(kSweepBlock), rev_->gview(), (const unsigned long long *)rev_->dpre_,
                     rev_->tiles_, (const uint32_t *)rev_->col_, (uint64_t)0, 0u, nv_, (uint64_t)0, nt, pull_state(io.round),
                     io, force);
        round_launch(k_pr_pull_fix, dim3(grid_for(nt + 1, kBlock, 4096)), dim3(kBlock),
                     (const unsigned long long *)rev_->dpre_, nv_, rev_->tiles_, pull_state(io.round), io, force);
        end_timed(timed && !batch_);
    }
    // The same with the store's edges in pinned host memory: its SEGMENT slot
    // ranges stream through HBM staging buffers while earlier ones are pulled.
    void pull_streamed(const PassIO &io, bool timed) {
        const uint64_t nt = rev_->ntiles();
        const PullPR ps = pull_state(io.round, false);  // segments are staged in HBM: no split
        GraphView rg = rev_->gview();
        rg.tail = nullptr;  // the staged segments stand in for the host copy
        const unsigned long long *rdpre = rev_->dpre_;
        const SweepTiles rt = rev_->tiles_;
        begin_timed(timed, SG_KCLASS_PULL);
        const uint64_t b0 = rev_->st_.streamed_bytes;
        const double ms0 = rev_->st_.stream_ms;
        rev_->streamed_sweep([&](const uint32_t *colp, uint64_t slo, uint32_t va, uint32_t vb, uint64_t tlo, uint64_t thi) {
            const uint32_t grid = sweep_grid(k_pr_pull, thi - tlo);
            hipLaunchKernelGGL(k_pr_pull, dim3(grid), dim3(kSweepBlock), 0, stream_, rg, rdpre, rt, colp, slo, va, vb, tlo,
                               std::min<uint64_t>(thi, nt), ps, io, 1);
            SG_LAUNCHED();
        });
        hipLaunchKernelGGL(k_pr_pull_fix, dim3(grid_for(nt + 1, kBlock, 4096)), dim3(kBlock), 0, stream_, rdpre, nv_, rt, ps,
                           io, 1);
        SG_LAUNCHED();
        end_timed(timed);
        st_.streamed_bytes += rev_->st_.streamed_bytes - b0;  // the in-store's segments crossed the host link
        st_.stream_ms += rev_->st_.stream_ms - ms0;
    }
    // The touched sources' amend pushes (type -1: cancel, +1: compensate)
    // delivered by one pull over the in-store as it stands; what crosses
    // epsilon joins round R's frontier (stamp R).
    void amend_pull(uint32_t R, int type) {
        const uint64_t words = ((uint64_t)nv_ + 31) / 32;
        const PullTake t = take_state(R - 1);  // the pull of "round R - 1" reads it and stamps R
        fill<uint32_t>(t.bits, words, 0u);
        hipLaunchKernelGGL(k_pr_amend_take, dim3(grid_for(nv_, 256, 4096)), dim3(256), 0, stream_,
                           (const uint32_t *)touched_, (const uint32_t *)&ctl_->tcount, nv_, op_, type, value_, buffer_,
                           (const uint32_t *)deg_, t);
        SG_LAUNCHED();
        PassIO io = seed_io(R, nullptr, nullptr, 0);
        io.round = R - 1;
        // counted as a pull round on the device: the next pr_rounds adds its bytes
        pull_resident(io, 1, timing(SG_KCLASS_PULL));
    }

    // One round: push (k_round) or, when the frontier's out-edges pass the
    // threshold, take + pull + fix-up (the round kernel only takes the
    // frontier's pushes then).
    void pr_round(const PassIO &io0, bool do_time) {
        const PassIO io = pr_io(io0);
        const uint32_t R = io.round;
        clear_take(R);
        launch_pass(io, VPageRank{op_, value_, buffer_, take_state(R)},
                    ERelax<api::PageRank>{op_, buffer_, stamp_, R + 1, 0, 0}, do_time);
        if (pulls()) pull_resident(io, 0, do_time && timing(SG_KCLASS_PULL));
    }

    // Rounds from the frontier of round R_ to convergence.
    uint32_t pr_rounds(bool timed, int phase = -1) {
        if (pulls()) {  // out-edges of the seeded frontier: the first round's direction
            hipLaunchKernelGGL(k_list_degsum, dim3(256), dim3(kBlock), 0, stream_, (const uint32_t *)q_[R_ & 1],
                               (const uint32_t *)(cnt_ + (R_ % 3)), nv_, (const uint32_t *)deg_, ecnt_ + (R_ % 3));
            SG_LAUNCHED();
        }
        if (pulls() && rev_->host_edges_ && (cfg_.flags & SG_FLAG_PULL_STREAM)) return pr_rounds_host(timed);
        const bool do_time = timed && timing(SG_KCLASS_RELAX);
        const uint64_t p0 = st_.pull_rounds;
        const uint32_t r = drive_rounds([&](const PassIO &io) { pr_round(io, do_time); }, timed, phase);
        // the pull kernels of push rounds returned at once: bytes only for the rounds that pulled
        if (timed && timing(SG_KCLASS_PULL))
            kt_bytes_[SG_KCLASS_PULL] += (st_.pull_rounds - p0) * (8ull * ne_ + 24ull * nv_);
        return r;
    }

    // Host-driven rounds for an in-edge store in pinned host memory: the
    // direction is chosen on the host before each round (a streamed pull moves
    // the whole store over the host link, so a push round must not start one).
    uint32_t pr_rounds_host(bool timed) {
        const bool do_time = timed && timing(SG_KCLASS_RELAX);
        const uint64_t p0 = st_.pull_rounds;
        if (timed) fill<unsigned long long>(stats_ + 8, 8, 0ull);
        const uint32_t maxr = cfg_.max_rounds ? cfg_.max_rounds : 1000000u;
        read_block();
        uint32_t rounds = 0;
        last_appended_ = h_blk_[R_ % 3];
        for (;;) {
            const uint32_t R = R_;
            PassIO io = round_io(R);
            io.log = rlog_;
            io.round = R;
            if (timed) io.stats = stats_ + 8;
            io = pr_io(io);
            const bool dense = h_ecnt()[R % 3] > io.dense_thr;
            clear_take(R);
            launch_pass(io, VPageRank{op_, value_, buffer_, take_state(R)},
                        ERelax<api::PageRank>{op_, buffer_, stamp_, R + 1, 0, 0}, do_time);
            if (dense) pull_streamed(io, do_time && timing(SG_KCLASS_PULL));
            ++R_;
            ++rounds;
            read_block();
            if (do_time) harvest_events();
            const uint32_t front = h_blk_[(R + 1) % 3], spill = h_blk_[3 + R % 3];
            last_appended_ += front;
            if (front == 0 && spill == 0) break;
            if (rounds >= maxr) {
                finish_rounds(timed);
                fail(SG_EDEVICE, "no convergence within " + std::to_string(maxr) + " rounds");
            }
        }
        finish_rounds(timed);
        if (timed && timing(SG_KCLASS_PULL))
            kt_bytes_[SG_KCLASS_PULL] += (st_.pull_rounds - p0) * (8ull * ne_ + 24ull * nv_);
        return rounds;
    }

    void residual() {
        SG_HIP(hipMemsetAsync(d_sum_, 0, 2 * sizeof(double), stream_));
        hipLaunchKernelGGL(k_pr_residual, dim3(grid_for(nv_, kBlock, 2048)), dim3(kBlock), 0, stream_, nv_, buffer_, d_sum_,
                           (unsigned long long *)(d_sum_ + 1));
        SG_LAUNCHED();
        double h[2];
        SG_HIP(hipMemcpyAsync(h_u64_, d_sum_, 2 * sizeof(double), hipMemcpyDeviceToHost, stream_));
        sync();
        std::memcpy(h, h_u64_, sizeof h);
        st_.pr_residual_l1 = h[0];
        st_.pr_residual_max = h[1];
    }

    void state_arrays(std::vector<std::pair<void *, size_t>> &a, uint32_t &source) override {
        if (!value_) {
            value_ = dalloc<float>(nv_);
            buffer_ = dalloc<float>(nv_);
            d_sum_ = dalloc<double>(2);
        }
        a = {{value_, (size_t)nv_ * 4}, {buffer_, (size_t)nv_ * 4}};
        source = 0;
    }
    void restored(uint32_t) override {
        op_.m_alpha = cfg_.pr_alpha;
        op_.m_error = cfg_.pr_epsilon;
        pending_ = false;
        residual();
        cache_rebuild();
        inited_ = true;
    }

    void init(const uint32_t *, uint32_t) override {
        if (!loaded_) fail(SG_EINVAL, "sg_init before sg_load_csr");
        op_.m_alpha = cfg_.pr_alpha;
        op_.m_error = cfg_.pr_epsilon;
        if (!value_) {
            value_ = dalloc<float>(nv_);
            buffer_ = dalloc<float>(nv_);
            d_sum_ = dalloc<double>(2);
        }
        auto t0 = std::chrono::steady_clock::now();
        hipLaunchKernelGGL(k_init_state<api::PageRank>, dim3(grid_for(nv_, 256, 16384)), dim3(256), 0, stream_, nv_, op_, value_,
                           buffer_);
        SG_LAUNCHED();
        frontier_reset();
        hipLaunchKernelGGL(k_seed_active<api::PageRank>, dim3(grid_for(nv_, kBlock, 16384)), dim3(kBlock), 0, stream_, nv_, op_,
                           value_, buffer_, stamp_, R_, q_[R_ & 1], cnt_ + (R_ % 3), stats_ + 3);
        SG_LAUNCHED();
        st_.rounds = pr_rounds(false);
        residual();
        init_others();
        cache_rebuild();
        inited_ = true;
        st_.last_phase_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    // Programs that share this engine's stores: their static solve, and their
    // recompute after each committed batch (PageRank's rounds done).
    virtual void init_others() {}
    virtual void recompute_others() {}

    // Cancelation + Compensate (framework.cuh:1110-1301) restricted to the sources
    // whose out-lists change: push -alpha*value/deg over the old list, update the
    // graph (and the in-edge store), push +alpha*value/deg' over the new list,
    // converge on |residual| > eps.
    void apply(const uint32_t *d_ins, uint64_t nins, const uint32_t *d_del, uint64_t ndel) override {
        if (!inited_) fail(SG_EINVAL, "sg_apply before sg_init");
        auto t0 = std::chrono::steady_clock::now();
        const bool overflow = prepare(d_ins, nins, d_del, ndel);
        frontier_reset();
        const uint32_t R = R_;
        const uint32_t *tc = &ctl_->tcount;
        if (pulls()) {
            // cancel over the old in-lists, compensate over the new ones: two pulls
            amend_pull(R, -1);
            commit(overflow);
            rev_->apply(d_ins, nins, d_del, ndel);
            amend_pull(R, +1);
        } else {
            {
                PassIO io = seed_io(R, touched_, tc, 0);
                launch_pass(io, VAmend<api::PageRank>{op_, -1, value_, buffer_},
                            ERelax<api::PageRank>{op_, buffer_, stamp_, R, 0, 0}, false);
                // long old lists were cut into chunks: walk them before the commit moves them
                drain_seed_spill(R, ERelax<api::PageRank>{op_, buffer_, stamp_, R, 0, 0});
            }
            commit(overflow);
            if (rev_) rev_->apply(d_ins, nins, d_del, ndel);
            {
                PassIO io = seed_io(R, touched_, tc, 0);
                launch_pass(io, VAmend<api::PageRank>{op_, +1, value_, buffer_},
                            ERelax<api::PageRank>{op_, buffer_, stamp_, R, 0, 0}, false);
            }
        }
        if (cfg_.flags & SG_FLAG_DEFER_CONVERGE) {
            // the long new lists were cut into chunks for the first round: walk
            // them now. The residual then holds alpha * (A' - A) * value on top
            // of the old one, exact for any committed value: later snapshots add
            // theirs and one convergence serves them all
            drain_seed_spill(R, ERelax<api::PageRank>{op_, buffer_, stamp_, R, 0, 0});
            pending_ = true;
            st_.rounds = 0;
        } else {
            st_.rounds = pr_rounds(true, kPhaseRelax);
        }
        recompute_others();
        finish_batch();
        st_.snapshots += 1;
        st_.last_phase_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        if (!pending_) residual();
    }

    // Deferred snapshots: every vertex whose residual is beyond +-epsilon
    // starts the rounds (the frontiers of the deferred passes were not kept;
    // cancellations leave negative residuals).
    bool pending_ = false;
    bool pending() const override { return pending_; }
    void converge() override {
        if (!pending_) return;
        auto t0 = std::chrono::steady_clock::now();
        frontier_reset();
        hipLaunchKernelGGL((k_seed_active<api::PageRank, true>), dim3(grid_for(nv_, kBlock, 16384)), dim3(kBlock), 0, stream_, nv_, op_,
                           value_, buffer_, stamp_, R_, q_[R_ & 1], cnt_ + (R_ % 3), stats_ + 3);
        SG_LAUNCHED();
        st_.rounds = pr_rounds(true, kPhaseRelax);
        pending_ = false;
        residual();
        st_.last_phase_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    void word_source(uint32_t q, uint32_t qi, const uint32_t *&base, uint64_t &stride, uint32_t &word) override {
        if (q == SG_Q_VALUE) { base = (const uint32_t *)value_; stride = 1; word = 0; return; }
        if (q == SG_Q_BUFFER) { base = (const uint32_t *)buffer_; stride = 1; word = 0; return; }
        Engine::word_source(q, qi, base, stride, word);
    }

    // The k highest ranks (ties by smaller id), by a descending radix sort of
    // (rank, id) on the device.
    float *tk_keys_ = nullptr;
    uint32_t *tk_ids_ = nullptr, *tk_ids2_ = nullptr;
    size_t tk_cap_[3] = {0, 0, 0};
    void topk(uint32_t k, uint32_t *ids, float *ranks) override {
        if (!inited_) fail(SG_EINVAL, "no ranks before sg_init");
        k = std::min<uint32_t>(k, nv_);
        if (k == 0) return;
        ensure(tk_keys_, tk_cap_[0], nv_);
        ensure(tk_ids_, tk_cap_[1], nv_);
        ensure(tk_ids2_, tk_cap_[2], nv_);
        hipLaunchKernelGGL(k_iota, dim3(grid_for(nv_, 256, 16384)), dim3(256), 0, stream_, tk_ids_, nv_);
        SG_LAUNCHED();
        size_t b = 0;
        SG_HIP(hipcub::DeviceRadixSort::SortPairsDescending(nullptr, b, value_, tk_keys_, tk_ids_, tk_ids2_, (int)nv_, 0, 32,
                                                            stream_));
        ensure(cub_tmp_, cub_cap_, b + 256);
        b = cub_cap_;
        SG_HIP(hipcub::DeviceRadixSort::SortPairsDescending(cub_tmp_, b, value_, tk_keys_, tk_ids_, tk_ids2_, (int)nv_, 0, 32,
                                                            stream_));
        SG_HIP(hipMemcpyAsync(ids, tk_ids2_, (size_t)k * 4, hipMemcpyDeviceToHost, stream_));
        if (ranks) SG_HIP(hipMemcpyAsync(ranks, tk_keys_, (size_t)k * 4, hipMemcpyDeviceToHost, stream_));
        sync();
    }

    void gather(uint32_t q, void *out, size_t bytes) override {
        if (q == SG_Q_VALUE || q == SG_Q_BUFFER) {
            if (!value_) fail(SG_EINVAL, "no vertex state before sg_init");
            if (bytes < (size_t)nv_ * 4) fail(SG_EINVAL, "output too small");
            SG_HIP(hipMemcpyAsync(out, q == SG_Q_VALUE ? value_ : buffer_, (size_t)nv_ * 4, hipMemcpyDeviceToHost, stream_));
            sync();
            return;
        }
        if (q == SG_Q_PARENT) {
            if (bytes < (size_t)nv_ * 4) fail(SG_EINVAL, "output too small");
            std::memset(out, 0xFF, (size_t)nv_ * 4);
            return;
        }
        if (q == SG_Q_IN_DEGREE || q == SG_Q_IN_OFFSETS || q == SG_Q_IN_SLOTS) {
            if (!rev_) fail(SG_EINVAL, "this engine keeps no in-edge store (pull_divisor = UINT32_MAX)");
            sync();
            rev_->gather_in(q - SG_Q_IN_DEGREE + SG_Q_DEGREE, out, bytes);
            return;
        }
        gather_graph(q, out, bytes);
    }
};

// ---------------------------------------------------------------------------
// PageRank and CC labels on one graph: one forward store and one in-edge
// store serve both programs, so each batch updates each store once; CC's
// recompute after deletions reads the in-lists of its reset vertices instead
// of streaming every live edge. The reference runs HybridPageRank and
// HybridCC as separate processes, each with its own PMA
// (samples/hybrid_pr/hybrid_pr.cu:154-201, samples/hybrid_cc/hybrid_cc.cu:172-208).
// ---------------------------------------------------------------------------
class PrCcEngine final : public PageRankEngine {
public:
    using PageRankEngine::PageRankEngine;
    TravProgram<api::CC> cc_;

    bool wants_in_store() const override { return true; }
    void init_others() override {
        cc_.phase_inval = kPhaseInval2;
        cc_.phase_relax = kPhaseRelax2;
        cc_.solve(*this);
    }
    void recompute_others() override { cc_.recompute(*this, rev_.get()); }

    void state_arrays(std::vector<std::pair<void *, size_t>> &, uint32_t &) override {
        fail(SG_EINVAL, "checkpoints are for single-program engines (PR, SSSP, BFS, CC)");
    }
    void gather_query(uint32_t qi, uint32_t what, void *out, size_t bytes) override {
        if (!inited_) fail(SG_EINVAL, "no vertex state before sg_init");
        if (qi == 0) return PageRankEngine::gather(what, out, bytes);
        if (qi == 1) {
            if (what != SG_Q_VALUE && what != SG_Q_PARENT && what != SG_Q_BUFFER) fail(SG_EINVAL, "CC gives values, parents, buffers");
            return cc_.gather(*this, what, out, bytes);
        }
        fail(SG_EINVAL, "PR + CC engines: qi 0 = PageRank, 1 = CC");
    }
    void word_source(uint32_t q, uint32_t qi, const uint32_t *&base, uint64_t &stride, uint32_t &word) override {
        if (qi == 1 && (q == SG_Q_VALUE || q == SG_Q_PARENT || q == SG_Q_BUFFER))
            return cc_.word_source(*this, q, base, stride, word);
        PageRankEngine::word_source(q, qi, base, stride, word);
    }
};

// ---------------------------------------------------------------------------
// Batched BFS/SSSP queries on one graph (no reference counterpart: the
// reference answers one source per process). The PMA update, the boundary
// sweep and every round are shared by all queries.
// ---------------------------------------------------------------------------
class MultiEngine final : public Engine {
public:
    using Engine::Engine;
    uint32_t nq_ = 0, G_ = 0, KP_ = 0, QS_ = 0, qs_log_ = 0;
    unsigned long long *best_ = nullptr, *act_ = nullptr, *minv_ = nullptr, *ninv_ = nullptr;
    uint8_t *kind_ = nullptr;
    uint32_t *src_ = nullptr;
    MDesc *mspill_[3] = {nullptr, nullptr, nullptr};
    size_t mspill_cap_ = 0;
    size_t state_nv_ = 0, state_kp_ = 0;

    MultiState mstate() const {
        MultiState m;
        m.best = best_;
        m.act = act_;
        m.inv = minv_;
        m.ninv = ninv_;
        m.invbit = inval_;
        m.coarse = coarse_;
        m.cshift = cshift_;
        m.kind = kind_;
        m.G = G_;
        m.KP = KP_;
        m.QS = QS_;
        m.qs_log = qs_log_;
        m.wmode = cfg_.weight_mode;
        m.wseed = cfg_.weight_seed;
        return m;
    }

    void ensure_mspill() {
        // per round: one descriptor per kMEdgeChunk edges of a list, and at least
        // one per frontier vertex whose list passes a wave's budget
        const size_t need = (size_t)((ne_ + nins_) / kMEdgeChunk + nv_ + 1) * G_ + 4096;
        if (need <= mspill_cap_) return;
        for (auto &p : mspill_) { dfree(p); p = nullptr; }
        for (auto &p : mspill_) p = dalloc<MDesc>(need);
        mspill_cap_ = need;
    }

    template <int kMode>
    void mround(const PassIO &io, bool timed) {
        begin_timed(timed && !batch_);
        round_launch(k_mround<kMode>, dim3(round_blocks()), dim3(kBlock), gview(), mstate(), io, mspill_[io.round % 3],
                     (const MDesc *)mspill_[(io.round + 2) % 3], (uint32_t)std::min<size_t>(mspill_cap_, 0xFFFFFFFFu),
                     stamp_);
        end_timed(timed && !batch_);
    }

    template <int kMode>
    uint32_t mrounds(bool timed, int phase = -1, uint32_t limit = 0) {
        const bool do_time = timed && timing(SG_KCLASS_RELAX);
        return drive_rounds([&](const PassIO &io) { mround<kMode>(io, do_time); }, timed, phase, limit);
    }

    // Slot layout for nq queries: groups of QS slots (QS = 64, or the next
    // power of two >= nq below 64); the state only grows.
    void layout(uint32_t nq) {
        QS_ = 1;
        while (QS_ < std::min<uint32_t>(nq, 64)) QS_ <<= 1;
        qs_log_ = 0;
        while ((1u << qs_log_) < QS_) ++qs_log_;
        G_ = (nq + QS_ - 1) / QS_;
        KP_ = G_ * QS_;
        if (state_nv_ != nv_ || state_kp_ < KP_) {
            dfree(best_); dfree(act_); dfree(minv_); dfree(ninv_); dfree(kind_); dfree(src_);
            best_ = act_ = minv_ = ninv_ = nullptr;
            kind_ = nullptr;
            src_ = nullptr;
            const uint32_t G = (KP_ + 63) / 64 > G_ ? (KP_ + 63) / 64 : G_;
            best_ = dalloc<unsigned long long>((size_t)nv_ * KP_);
            act_ = dalloc<unsigned long long>((size_t)nv_ * G);
            minv_ = dalloc<unsigned long long>((size_t)nv_ * G);
            ninv_ = dalloc<unsigned long long>((size_t)nv_ * G);
            kind_ = dalloc<uint8_t>(KP_);
            src_ = dalloc<uint32_t>(std::max<uint32_t>(KP_, 2 * 4096));
            state_nv_ = nv_;
            state_kp_ = KP_;
        }
    }

    // ---- continuous batching (sg_query_open / _add / _step / _release) ----
    // A fixed set of KP slots; queries join free slots between round batches
    // (their sources seeded into the frontier the next round reads), run
    // together, and leave when no work is pending for them.
    bool cont_ = false;
    std::vector<uint8_t> live_, done_, h_kind_;
    unsigned long long *pend_ = nullptr;
    size_t pend_cap_ = 0;

    void open_slots(uint32_t k) override {
        if (!loaded_) fail(SG_EINVAL, "sg_query_open before sg_load_csr");
        if (k == 0 || k > 4096) fail(SG_EINVAL, "slots must be in [1, 4096]");
        layout(k);
        nq_ = KP_;
        h_kind_.assign(KP_, kQNone);
        live_.assign(KP_, 0);
        done_.assign(KP_, 0);
        SG_HIP(hipMemcpyAsync(kind_, h_kind_.data(), KP_, hipMemcpyHostToDevice, stream_));
        fill<unsigned long long>(best_, (uint64_t)nv_ * KP_, ~0ull);
        fill<unsigned long long>(act_, (uint64_t)nv_ * G_, 0ull);
        fill<unsigned long long>(minv_, (uint64_t)nv_ * G_, 0ull);
        fill<unsigned long long>(ninv_, (uint64_t)nv_ * G_, 0ull);
        fill<uint32_t>(inval_, ((uint64_t)nv_ + 31) / 32, 0u);
        fill<uint32_t>(coarse_, coarse_words(nv_, cshift_), 0u);
        ensure(pend_, pend_cap_, G_);
        nins_ = 0;
        ensure_mspill();
        frontier_reset();
        sync();
        cont_ = true;
        inited_ = true;
    }
    uint32_t live_count(bool only_running) const {
        uint32_t n = 0;
        for (size_t i = 0; i < live_.size(); ++i) n += live_[i] && !(only_running && done_[i]);
        return n;
    }
    void add_queries(const uint32_t *sources, const uint32_t *kinds, uint32_t n, uint32_t *slots) override {
        if (!cont_) fail(SG_EINVAL, "sg_query_add needs sg_query_open first");
        if (n == 0) return;
        if (!sources || !kinds || !slots) fail(SG_EINVAL, "null sources, kinds or slots");
        std::vector<uint32_t> hs, hsrc;
        for (uint32_t i = 0; i < n; ++i) {
            if (sources[i] >= nv_) fail(SG_EINVAL, "query " + std::to_string(i) + ": source id >= nv");
            if (kinds[i] != SG_ALGO_SSSP && kinds[i] != SG_ALGO_BFS)
                fail(SG_EINVAL, "query " + std::to_string(i) + ": kind must be SG_ALGO_SSSP or SG_ALGO_BFS");
        }
        for (uint32_t q = 0; q < KP_ && hs.size() < n; ++q)
            if (!live_[q]) hs.push_back(q);
        if (hs.size() < n) fail(SG_EINVAL, "not enough free query slots");
        for (uint32_t i = 0; i < n; ++i) {
            h_kind_[hs[i]] = kinds[i] == SG_ALGO_BFS ? kQUnit : kQWeighted;
            hsrc.push_back(sources[i]);
        }
        uint32_t *d = src_;  // n slots then n sources (n <= KP <= 4096)
        SG_HIP(hipMemcpyAsync(kind_, h_kind_.data(), KP_, hipMemcpyHostToDevice, stream_));
        SG_HIP(hipMemcpyAsync(d, hs.data(), (size_t)n * 4, hipMemcpyHostToDevice, stream_));
        SG_HIP(hipMemcpyAsync(d + n, hsrc.data(), (size_t)n * 4, hipMemcpyHostToDevice, stream_));
        hipLaunchKernelGGL(k_slot_reset, dim3(grid_for(nv_, 256, 16384)), dim3(256), 0, stream_, mstate(), nv_,
                           (const uint32_t *)d, n);
        SG_LAUNCHED();
        hipLaunchKernelGGL(k_slot_seed, dim3(grid_for(n, 256, 64)), dim3(256), 0, stream_, mstate(), (const uint32_t *)d,
                           (const uint32_t *)(d + n), n, stamp_, R_, q_[R_ & 1], cnt_ + (R_ % 3), nv_, stats_ + 3);
        SG_LAUNCHED();
        sync();  // the host vectors above were read by the copies
        for (uint32_t i = 0; i < n; ++i) {
            live_[hs[i]] = 1;
            done_[hs[i]] = 0;
            slots[i] = hs[i];
        }
    }
    // Up to max_rounds rounds, then the slots whose work is done (converged
    // since the last call) in done[0..*ndone); *nactive = slots still running.
    void step_queries(uint32_t max_rounds, uint32_t *done, uint32_t cap, uint32_t *ndone, uint32_t *nactive) override {
        if (!cont_) fail(SG_EINVAL, "sg_query_step needs sg_query_open first");
        *ndone = 0;
        if (live_count(true)) mrounds<0>(false, -1, std::max<uint32_t>(1, max_rounds));
        fill<unsigned long long>(pend_, G_, 0ull);
        const uint32_t R = R_;
        hipLaunchKernelGGL(k_slot_pending, dim3(64), dim3(kBlock), 0, stream_, mstate(), (const uint32_t *)q_[R & 1],
                           (const uint32_t *)(cnt_ + (R % 3)), nv_, (const MDesc *)mspill_[(R + 2) % 3],
                           (const uint32_t *)(hcnt_ + ((R + 2) % 3)), (uint32_t)std::min<size_t>(mspill_cap_, 0xFFFFFFFFu),
                           pend_);
        SG_LAUNCHED();
        std::vector<unsigned long long> pend(G_);
        SG_HIP(hipMemcpyAsync(pend.data(), pend_, (size_t)G_ * 8, hipMemcpyDeviceToHost, stream_));
        sync();
        uint32_t act = 0;
        for (uint32_t q = 0; q < KP_; ++q) {
            if (!live_[q] || done_[q]) continue;
            if ((pend[q / QS_] >> (q % QS_)) & 1ull) { ++act; continue; }
            done_[q] = 1;
            if (*ndone < cap) done[(*ndone)++] = q;
        }
        *nactive = act;
    }
    void release_queries(const uint32_t *slots, uint32_t n) override {
        if (!cont_) fail(SG_EINVAL, "sg_query_release needs sg_query_open first");
        for (uint32_t i = 0; i < n; ++i) {
            if (slots[i] >= KP_ || !live_[slots[i]]) fail(SG_EINVAL, "slot " + std::to_string(slots[i]) + " holds no query");
            if (!done_[slots[i]]) fail(SG_EINVAL, "slot " + std::to_string(slots[i]) + " is still converging");
        }
        for (uint32_t i = 0; i < n; ++i) {
            live_[slots[i]] = done_[slots[i]] = 0;
            h_kind_[slots[i]] = kQNone;
        }
        SG_HIP(hipMemcpyAsync(kind_, h_kind_.data(), KP_, hipMemcpyHostToDevice, stream_));
        sync();
    }

    template <int kMode>
    void mpairs(uint64_t n, uint32_t R) {
        if (n == 0) return;
        PassIO io = seed_io(R, nullptr, nullptr, 0);
        const uint64_t per = kWave / QS_;
        const uint32_t grid = grid_for((n + per - 1) / per, kWavesPerBlock, 8192);
        hipLaunchKernelGGL((k_mpairs<kMode>), dim3(grid), dim3(kBlock), 0, stream_, mstate(), bv_, br_, io, stamp_, R);
        SG_LAUNCHED();
    }

    void init(const uint32_t *sources, uint32_t nsrc) override {
        std::vector<uint32_t> kinds(nsrc, cfg_.weight_mode == SG_WEIGHT_ONE ? (uint32_t)SG_ALGO_BFS : (uint32_t)SG_ALGO_SSSP);
        init_queries(sources, kinds.data(), nsrc);
    }

    void init_queries(const uint32_t *sources, const uint32_t *kinds, uint32_t nq) override {
        if (!loaded_) fail(SG_EINVAL, "sg_init before sg_load_csr");
        if (nq == 0) {  // drop the queries: later snapshots only update the graph
            nq_ = 0;
            cont_ = false;
            inited_ = true;
            return;
        }
        if (!sources || !kinds) fail(SG_EINVAL, "null sources or kinds");
        if (nq > 4096) fail(SG_EINVAL, "at most 4096 queries per engine");
        for (uint32_t i = 0; i < nq; ++i) {
            if (sources[i] >= nv_) fail(SG_EINVAL, "query " + std::to_string(i) + ": source id >= nv");
            if (kinds[i] != SG_ALGO_SSSP && kinds[i] != SG_ALGO_BFS)
                fail(SG_EINVAL, "query " + std::to_string(i) + ": kind must be SG_ALGO_SSSP or SG_ALGO_BFS");
        }
        nq_ = nq;
        cont_ = false;
        // state for KP slots; the allocation only grows, a smaller KP uses a prefix of it
        layout(nq);
        std::vector<uint8_t> hk(KP_, kQNone);
        for (uint32_t i = 0; i < nq; ++i) hk[i] = kinds[i] == SG_ALGO_BFS ? kQUnit : kQWeighted;
        auto t0 = std::chrono::steady_clock::now();
        SG_HIP(hipMemcpyAsync(kind_, hk.data(), KP_, hipMemcpyHostToDevice, stream_));
        SG_HIP(hipMemcpyAsync(src_, sources, (size_t)nq * 4, hipMemcpyHostToDevice, stream_));
        fill<unsigned long long>(best_, (uint64_t)nv_ * KP_, ~0ull);
        fill<unsigned long long>(act_, (uint64_t)nv_ * G_, 0ull);
        fill<unsigned long long>(minv_, (uint64_t)nv_ * G_, 0ull);
        fill<unsigned long long>(ninv_, (uint64_t)nv_ * G_, 0ull);
        fill<uint32_t>(inval_, ((uint64_t)nv_ + 31) / 32, 0u);
        fill<uint32_t>(coarse_, coarse_words(nv_, cshift_), 0u);
        nins_ = 0;
        ensure_mspill();
        frontier_reset();
        hipLaunchKernelGGL(k_minit_sources, dim3(grid_for(nq, 256, 64)), dim3(256), 0, stream_, mstate(), src_, nq, stamp_,
                           R_, q_[R_ & 1], cnt_ + (R_ % 3), nv_, stats_ + 3);
        SG_LAUNCHED();
        st_.rounds = mrounds<0>(false);
        sync();  // host vectors above were read by the async copies
        cache_rebuild();
        inited_ = true;
        st_.last_phase_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }

    void apply(const uint32_t *d_ins, uint64_t nins, const uint32_t *d_del, uint64_t ndel) override {
        if (!inited_) fail(SG_EINVAL, "sg_apply before sg_init");
        auto t0 = std::chrono::steady_clock::now();
        if (cont_ && live_count(true))
            fail(SG_EINVAL, "queries are still converging: step them to convergence before an update");
        const bool overflow = prepare(d_ins, nins, d_del, ndel);
        commit(overflow);
        if (nq_ == 0 || (cont_ && live_count(false) == 0)) {  // no queries: the PMA update is the whole snapshot
            st_.invalidated = 0;
            st_.rounds = 0;
            finish_batch();
            st_.snapshots += 1;
            st_.last_phase_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
            return;
        }
        ensure_mspill();
        // 1. reset the slots whose tree edge was deleted, then their tree children (fixed point)
        frontier_reset();
        uint64_t ninval = 0;
        if (ndel_) {
            mpairs<1>(ndel_, R_);
            mrounds<1>(false, kPhaseInval);
            ninval = last_appended_;  // vertices put on the invalidation frontiers (> 0: some slot reset)
        }
        st_.invalidated = ninval;
        // 2. seed the recompute: live edges into reset slots (one sweep for all queries) + inserted edges
        frontier_reset();
        const uint32_t R = R_;
        if (ninval) {
            const SweepTiles tl = sweep_tiles();
            const uint64_t ntiles = ((uint64_t)nv_ + ne_ + kSweepTile - 1) / kSweepTile;
            const uint32_t grid = sweep_grid(k_msweep, ntiles);
            PassIO io = seed_io(R, nullptr, nullptr, 0);
            hipLaunchKernelGGL(k_msweep, dim3(grid), dim3(kSweepBlock), 0, stream_, gview(), dpre_, tl, mstate(), io, stamp_,
                               R);
            SG_LAUNCHED();
            hipLaunchKernelGGL(k_mclear_inv, dim3(grid_for((nv_ + 31) / 32, 256, 4096)), dim3(256), 0, stream_, nv_, G_,
                               inval_, coarse_, cshift_, minv_);
            SG_LAUNCHED();
        }
        mpairs<0>(nins_, R);
        // 3. converge
        st_.rounds = mrounds<0>(true, kPhaseRelax);
        finish_batch();
        st_.snapshots += 1;
        st_.last_phase_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }

    void gather_query(uint32_t qi, uint32_t what, void *out, size_t bytes) override {
        if (!inited_) fail(SG_EINVAL, "no queries before sg_init_queries");
        if (qi >= nq_) fail(SG_EINVAL, "query index out of range");
        if (cont_ && !live_[qi]) fail(SG_EINVAL, "slot holds no query");
        if (what != SG_Q_VALUE && what != SG_Q_PARENT) fail(SG_EINVAL, "sg_gather_query returns SG_Q_VALUE or SG_Q_PARENT");
        if (bytes < (size_t)nv_ * 4) fail(SG_EINVAL, "output too small");
        ensure(tmp_, tmp_cap_, nv_);
        hipLaunchKernelGGL(k_mextract, dim3(grid_for(nv_, 256, 16384)), dim3(256), 0, stream_, nv_,
                           (const unsigned long long *)best_, KP_, qi, tmp_, what == SG_Q_PARENT ? 1 : 0);
        SG_LAUNCHED();
        SG_HIP(hipMemcpyAsync(out, tmp_, (size_t)nv_ * 4, hipMemcpyDeviceToHost, stream_));
        sync();
    }

    void word_source(uint32_t q, uint32_t qi, const uint32_t *&base, uint64_t &stride, uint32_t &word) override {
        if (q == SG_Q_VALUE || q == SG_Q_PARENT) {
            if (qi >= nq_) fail(SG_EINVAL, "query index out of range");
            if (cont_ && !live_[qi]) fail(SG_EINVAL, "slot holds no query");
            base = (const uint32_t *)best_;
            stride = 2ull * KP_;
            word = 2 * qi + (q == SG_Q_VALUE ? 1 : 0);
            return;
        }
        Engine::word_source(q, qi, base, stride, word);
    }

    void gather(uint32_t q, void *out, size_t bytes) override {
        if (q == SG_Q_VALUE || q == SG_Q_PARENT) return gather_query(0, q, out, bytes);
        if (q == SG_Q_BUFFER) fail(SG_EINVAL, "batched queries keep no separate buffer");
        gather_graph(q, out, bytes);
    }
};

}  // namespace sg

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------

template <class F>
static sg_status guarded(sg_engine *eng, F &&f) {
    if (!eng || !eng->impl) return SG_EINVAL;
    try {
        // the calling thread's current device may differ from the one the
        // engine was created on (a request loop calls from its worker thread)
        if (eng->impl->stream_) SG_HIP(hipSetDevice(eng->impl->cfg_.device));
        f(*eng->impl);
        eng->impl->last_error_ = "no error";
        return SG_OK;
    } catch (const sg::Error &e) {
        eng->impl->last_error_ = e.msg;
        return e.status;
    } catch (const std::bad_alloc &) {
        eng->impl->last_error_ = "host allocation failed";
        return SG_ENOMEM;
    } catch (...) {
        eng->impl->last_error_ = "unexpected exception";
        return SG_EDEVICE;
    }
}

extern "C" {

void sg_config_default(sg_config *c) {
    if (!c) return;
    std::memset(c, 0, sizeof *c);
    c->algo = SG_ALGO_PR;
    c->weight_mode = SG_WEIGHT_HASH128;
    c->weight_seed = 27491095ull;
    c->pr_alpha = 0.85;
    c->pr_epsilon = 1e-6;
    c->pma_density = 0.25;
    c->segments = 1;
    c->n_stream = 3;
    c->pull_divisor = 0;
    c->max_rounds = 0;
    c->rounds_per_sync = 0;  // adaptive
    c->flags = 0;
    c->device_budget_gb = 0;
    c->device = 0;
    c->cache_level = 4;     // app_skeleton.h:100 (-cache 4); used only with host edges
    c->edge_budget_gb = 0;
}

namespace sg {
// ---- seeded R-MAT on the device (sg_gen_rmat_device) ---------------------------
// Draw i is sg_rmat_draw of the splitmix stream (seed, i), as on the host;
// self-loops become ~0 and sort past every edge.
__global__ void k_rmat_keys(uint64_t ndraw, uint32_t scale, uint64_t seed, SgRmatThresh th, const uint32_t *perm,
                            unsigned long long *keys) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < ndraw; i += stride) {
        SgRng rng(seed, i);
        uint32_t s, d;
        sg_rmat_draw(rng, scale, th, s, d);
        if (perm) {
            s = perm[s];
            d = perm[d];
        }
        keys[i] = s == d ? ~0ull : (((unsigned long long)s << 32) | d);
    }
}

// keep[i]: first copy of a key that is an edge.
__global__ void k_rmat_keep(uint64_t n, const unsigned long long *keys, uint8_t *keep) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        keep[i] = keys[i] != ~0ull && (i == 0 || keys[i] != keys[i - 1]);
}

// row[v] = first edge of source v (lower bound of v << 32 in the sorted keys),
// col[i] = target of edge i.
__global__ void k_rmat_csr(uint32_t nv, uint64_t ne, const unsigned long long *keys, uint64_t *row, uint32_t *col) {
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
}  // namespace sg

sg_status sg_gen_rmat_device(int32_t device, uint32_t scale, uint32_t edge_factor, uint64_t seed, double a, double b,
                             double c, int permute, uint32_t *nv_out, uint64_t *ne_out, uint64_t **row_out,
                             uint32_t **col_out) {
    if (!nv_out || !ne_out || !row_out || !col_out) return SG_EINVAL;
    if (scale < 1 || scale > 31 || a < 0 || b < 0 || c < 0 || a + b + c > 1.0) return SG_EINVAL;
    const uint32_t nv = 1u << scale;
    const uint64_t ndraw = (uint64_t)edge_factor * nv;
    SgRmatThresh th;
    auto q = [](double x) { return (uint64_t)std::llround(std::min(1.0, std::max(0.0, x)) * 4294967296.0); };
    th.a = q(a);
    th.ab = q(a + b);
    th.abc = q(a + b + c);
    int prev = 0;
    if (hipGetDevice(&prev) != hipSuccess || hipSetDevice(device) != hipSuccess) {
        sg::g_create_error = "no usable device " + std::to_string(device);
        return SG_EDEVICE;
    }
    hipStream_t st = nullptr;
    void *d_keys = nullptr, *d_sorted = nullptr, *d_keep = nullptr, *d_perm = nullptr, *d_tmp = nullptr, *d_cnt = nullptr;
    void *d_row = nullptr, *d_col = nullptr;
    uint64_t *row = nullptr;
    uint32_t *col = nullptr;
    sg_status status = SG_OK;
    try {
        std::vector<uint32_t> perm;
        if (permute) {  // the host generator's seeded Fisher-Yates permutation
            perm.resize(nv);
            for (uint32_t i = 0; i < nv; ++i) perm[i] = i;
            SgRng rng(seed ^ 0x5045524d55544531ull, 0);
            for (uint32_t i = nv - 1; i > 0; --i) std::swap(perm[i], perm[(uint32_t)(rng.next() % ((uint64_t)i + 1))]);
        }
        SG_HIP(hipStreamCreate(&st));
        SG_HIP(hipMalloc(&d_keys, std::max<uint64_t>(ndraw, 1) * 8));
        SG_HIP(hipMalloc(&d_sorted, std::max<uint64_t>(ndraw, 1) * 8));
        SG_HIP(hipMalloc(&d_keep, std::max<uint64_t>(ndraw, 1)));
        SG_HIP(hipMalloc(&d_cnt, 8));
        if (permute) {
            SG_HIP(hipMalloc(&d_perm, (size_t)nv * 4));
            SG_HIP(hipMemcpyAsync(d_perm, perm.data(), (size_t)nv * 4, hipMemcpyHostToDevice, st));
        }
        const uint32_t grid = 4096;
        hipLaunchKernelGGL(sg::k_rmat_keys, dim3(grid), dim3(256), 0, st, ndraw, scale, seed, th, (const uint32_t *)d_perm,
                           (unsigned long long *)d_keys);
        SG_LAUNCHED();
        size_t b1 = 0, b2 = 0;
        SG_HIP(hipcub::DeviceRadixSort::SortKeys(nullptr, b1, (const unsigned long long *)d_keys,
                                                 (unsigned long long *)d_sorted, ndraw, 0, 64, st));
        SG_HIP(hipcub::DeviceSelect::Flagged(nullptr, b2, (const unsigned long long *)d_sorted, (const uint8_t *)d_keep,
                                             (unsigned long long *)d_keys, (unsigned long long *)d_cnt, (int64_t)ndraw, st));
        size_t tb = std::max(b1, b2) + 256;
        SG_HIP(hipMalloc(&d_tmp, tb));
        size_t bb = tb;
        SG_HIP(hipcub::DeviceRadixSort::SortKeys(d_tmp, bb, (const unsigned long long *)d_keys,
                                                 (unsigned long long *)d_sorted, ndraw, 0, 64, st));
        hipLaunchKernelGGL(sg::k_rmat_keep, dim3(grid), dim3(256), 0, st, ndraw, (const unsigned long long *)d_sorted,
                           (uint8_t *)d_keep);
        SG_LAUNCHED();
        bb = tb;
        SG_HIP(hipcub::DeviceSelect::Flagged(d_tmp, bb, (const unsigned long long *)d_sorted, (const uint8_t *)d_keep,
                                             (unsigned long long *)d_keys, (unsigned long long *)d_cnt, (int64_t)ndraw, st));
        unsigned long long ne = 0;
        SG_HIP(hipMemcpyAsync(&ne, d_cnt, 8, hipMemcpyDeviceToHost, st));
        SG_HIP(hipStreamSynchronize(st));
        if (ne > ndraw) throw sg::Error{SG_EDEVICE, "edge count beyond the draws"};
        SG_HIP(hipMalloc(&d_row, ((size_t)nv + 1) * 8));
        SG_HIP(hipMalloc(&d_col, std::max<uint64_t>(ne, 1) * 4));
        hipLaunchKernelGGL(sg::k_rmat_csr, dim3(grid), dim3(256), 0, st, nv, (uint64_t)ne, (const unsigned long long *)d_keys,
                           (uint64_t *)d_row, (uint32_t *)d_col);
        SG_LAUNCHED();
        row = (uint64_t *)std::malloc(((size_t)nv + 1) * 8);
        col = (uint32_t *)std::malloc(std::max<uint64_t>(ne, 1) * 4);
        if (!row || !col) throw std::bad_alloc();
        SG_HIP(hipMemcpyAsync(row, d_row, ((size_t)nv + 1) * 8, hipMemcpyDeviceToHost, st));
        SG_HIP(hipMemcpyAsync(col, d_col, (size_t)ne * 4, hipMemcpyDeviceToHost, st));
        SG_HIP(hipStreamSynchronize(st));
        *nv_out = nv;
        *ne_out = ne;
        *row_out = row;
        *col_out = col;
        row = nullptr;
        col = nullptr;
    } catch (const sg::Error &e) {
        sg::g_create_error = e.msg;
        status = e.status;
    } catch (const std::bad_alloc &) {
        sg::g_create_error = "host allocation failed";
        status = SG_ENOMEM;
    } catch (...) {
        sg::g_create_error = "unexpected exception";
        status = SG_EDEVICE;
    }
    std::free(row);
    std::free(col);
    for (void *p : {d_keys, d_sorted, d_keep, d_perm, d_tmp, d_cnt, d_row, d_col})
        if (p) (void)hipFree(p);
    if (st) (void)hipStreamDestroy(st);
    (void)hipSetDevice(prev);
    return status;
}

sg_status sg_create(const sg_config *cfg, sg_engine **out) {
    if (!cfg || !out) { sg::g_create_error = "null argument"; return SG_EINVAL; }
    *out = nullptr;
    std::unique_ptr<sg::Engine> impl;
    try {
        switch (cfg->algo) {
            case SG_ALGO_PR: impl.reset(new sg::PageRankEngine(*cfg)); break;
            case SG_ALGO_SSSP: impl.reset(new sg::TraversalEngine<sg::api::SSSP>(*cfg)); break;
            case SG_ALGO_BFS: impl.reset(new sg::TraversalEngine<sg::api::BFS>(*cfg)); break;
            case SG_ALGO_CC: impl.reset(new sg::TraversalEngine<sg::api::CC>(*cfg)); break;
            case SG_ALGO_MULTI: impl.reset(new sg::MultiEngine(*cfg)); break;
            case SG_ALGO_PR_CC: impl.reset(new sg::PrCcEngine(*cfg)); break;
            case SG_ALGO_CUSTOM:
                sg::g_create_error = "SG_ALGO_CUSTOM engines come from an operator library's create function";
                return SG_EINVAL;
            default: sg::g_create_error = "unknown algo"; return SG_EINVAL;
        }
        if (cfg->weight_mode > SG_WEIGHT_ONE) { sg::g_create_error = "unknown weight mode"; return SG_EINVAL; }
        if ((cfg->algo == SG_ALGO_PR || cfg->algo == SG_ALGO_PR_CC) && !(cfg->pr_alpha > 0 && cfg->pr_alpha < 1)) {
            sg::g_create_error = "pr_alpha must be in (0, 1)";
            return SG_EINVAL;
        }
        impl->init_device();
    } catch (const sg::Error &e) {
        sg::g_create_error = e.msg;
        return e.status;
    } catch (const std::bad_alloc &) {
        sg::g_create_error = "host allocation failed";
        return SG_ENOMEM;
    } catch (...) {
        sg::g_create_error = "unexpected exception";
        return SG_EDEVICE;
    }
    *out = new (std::nothrow) sg_engine{std::move(impl)};
    if (!*out) { sg::g_create_error = "host allocation failed"; return SG_ENOMEM; }
    return SG_OK;
}

void sg_destroy(sg_engine *eng) { delete eng; }

const char *sg_last_error(const sg_engine *eng) {
    if (!eng || !eng->impl) return sg::g_create_error.c_str();
    return eng->impl->last_error_.c_str();
}

sg_status sg_load_csr(sg_engine *eng, uint32_t nv, uint64_t ne, const uint64_t *row, const uint32_t *col) {
    return guarded(eng, [&](sg::Engine &e) { e.load_csr(nv, ne, row, col); });
}

sg_status sg_init(sg_engine *eng, const uint32_t *sources, uint32_t nsrc) {
    return guarded(eng, [&](sg::Engine &e) { e.init(sources, nsrc); });
}

sg_status sg_save_state(sg_engine *eng, const char *path) {
    return guarded(eng, [&](sg::Engine &e) {
        if (!path) sg::fail(SG_EINVAL, "null path");
        e.save_state(path);
    });
}

sg_status sg_restore_state(sg_engine *eng, const char *path) {
    return guarded(eng, [&](sg::Engine &e) {
        if (!path) sg::fail(SG_EINVAL, "null path");
        e.restore_state(path);
    });
}

sg_status sg_init_queries(sg_engine *eng, const uint32_t *sources, const uint32_t *kinds, uint32_t nq) {
    return guarded(eng, [&](sg::Engine &e) { e.init_queries(sources, kinds, nq); });
}

sg_status sg_query_open(sg_engine *eng, uint32_t slots) {
    return guarded(eng, [&](sg::Engine &e) { e.open_slots(slots); });
}

sg_status sg_query_add(sg_engine *eng, const uint32_t *sources, const uint32_t *kinds, uint32_t n, uint32_t *slots) {
    return guarded(eng, [&](sg::Engine &e) { e.add_queries(sources, kinds, n, slots); });
}

sg_status sg_query_step(sg_engine *eng, uint32_t max_rounds, uint32_t *done, uint32_t cap, uint32_t *ndone,
                        uint32_t *nactive) {
    return guarded(eng, [&](sg::Engine &e) {
        if (!ndone || !nactive || (cap && !done)) sg::fail(SG_EINVAL, "null output");
        e.step_queries(max_rounds, done, cap, ndone, nactive);
    });
}

sg_status sg_query_release(sg_engine *eng, const uint32_t *slots, uint32_t n) {
    return guarded(eng, [&](sg::Engine &e) {
        if (n && !slots) sg::fail(SG_EINVAL, "null slot list");
        e.release_queries(slots, n);
    });
}

sg_status sg_gather_query(sg_engine *eng, uint32_t qi, uint32_t what, void *out, size_t bytes) {
    return guarded(eng, [&](sg::Engine &e) {
        if (!out) sg::fail(SG_EINVAL, "null output");
        e.converge();  // deferred PageRank snapshots of a PR + CC engine
        e.gather_query(qi, what, out, bytes);
    });
}

sg_status sg_apply_snapshot(sg_engine *eng, const uint32_t *ins_uv, uint64_t nins, const uint32_t *del_uv, uint64_t ndel) {
    return guarded(eng, [&](sg::Engine &e) {
        if ((nins && !ins_uv) || (ndel && !del_uv)) sg::fail(SG_EINVAL, "null update array");
        if (!e.inited_) sg::fail(SG_EINVAL, "sg_apply_snapshot before sg_init");
        const uint32_t *di = nullptr, *dd = nullptr;
        e.stage_host_batch(ins_uv, nins, del_uv, ndel, di, dd);
        e.apply(di, nins, dd, ndel);
    });
}

sg_status sg_apply_snapshot_device(sg_engine *eng, const uint32_t *d_ins, uint64_t nins, const uint32_t *d_del,
                                   uint64_t ndel) {
    return guarded(eng, [&](sg::Engine &e) {
        if ((nins && !d_ins) || (ndel && !d_del)) sg::fail(SG_EINVAL, "null update array");
        e.apply(d_ins, nins, d_del, ndel);
    });
}

sg_status sg_gather(sg_engine *eng, uint32_t query, void *out, size_t bytes) {
    return guarded(eng, [&](sg::Engine &e) {
        if (!out) sg::fail(SG_EINVAL, "null output");
        if (!e.loaded_) sg::fail(SG_EINVAL, "no graph loaded");
        e.converge();
        e.gather(query, out, bytes);
    });
}

sg_status sg_converge(sg_engine *eng) {
    return guarded(eng, [&](sg::Engine &e) { e.converge(); });
}

sg_status sg_gather_at(sg_engine *eng, uint32_t query, uint32_t qi, const uint32_t *vertices, uint64_t n, uint32_t *out) {
    return guarded(eng, [&](sg::Engine &e) {
        if (n && (!vertices || !out)) sg::fail(SG_EINVAL, "null vertex list or output");
        if (!e.loaded_) sg::fail(SG_EINVAL, "no graph loaded");
        e.converge();
        e.gather_at(query, qi, vertices, n, out);
    });
}

sg_status sg_gather_at_noalloc(sg_engine *eng, uint32_t query, uint32_t qi, const uint32_t *vertices, uint64_t n,
                               uint32_t *out) {
    return guarded(eng, [&](sg::Engine &e) {
        if (n && (!vertices || !out)) sg::fail(SG_EINVAL, "null vertex list or output");
        if (!e.loaded_) sg::fail(SG_EINVAL, "no graph loaded");
        const int dev = e.cfg_.device;
        if (dev < 0 || dev >= 64) sg::fail(SG_EDEVICE, "device ordinal beyond the scratch table");
        std::lock_guard<std::mutex> lk(sg::g_scratch_call_mu[dev]);
        e.gather_at_noalloc(query, qi, vertices, n, out);
    });
}

sg_status sg_device_memory(int32_t device, uint64_t *free_bytes, uint64_t *total_bytes) {
    if (!free_bytes || !total_bytes) return SG_EINVAL;
    int cur = 0;
    if (hipGetDevice(&cur) != hipSuccess) return SG_EDEVICE;
    if (hipSetDevice(device) != hipSuccess) { (void)hipGetLastError(); return SG_EDEVICE; }
    size_t f = 0, t = 0;
    const hipError_t e = hipMemGetInfo(&f, &t);
    (void)hipSetDevice(cur);
    if (e != hipSuccess) { (void)hipGetLastError(); return SG_EDEVICE; }
    *free_bytes = f;
    *total_bytes = t;
    return SG_OK;
}

sg_status sg_pr_topk(sg_engine *eng, uint32_t k, uint32_t *ids, float *ranks) {
    return guarded(eng, [&](sg::Engine &e) {
        if (k && !ids) sg::fail(SG_EINVAL, "null id output");
        e.converge();
        e.topk(k, ids, ranks);
    });
}

sg_status sg_kernel_times(sg_engine *eng, uint32_t cls, uint64_t *launches, double *ms, uint64_t *bytes) {
    return guarded(eng, [&](sg::Engine &e) {
        if (cls >= SG_KCLASS_COUNT || !launches || !ms || !bytes) sg::fail(SG_EINVAL, "bad kernel class or null output");
        *launches = e.kt_launches_[cls];
        *ms = e.kt_ms_[cls];
        *bytes = e.kt_bytes_[cls];
    });
}

sg_status sg_get_stats(sg_engine *eng, sg_stats *out) {
    return guarded(eng, [&](sg::Engine &e) {
        if (!out) sg::fail(SG_EINVAL, "null output");
        e.st_.device_bytes = e.dev_bytes_ + e.aux_device_bytes();
        e.st_.host_bytes = e.host_bytes_ + e.aux_host_bytes();
        e.st_.in_capacity = e.in_capacity();
        e.in_mirror(e.st_.in_mirror_slots, e.st_.in_mirror_live);
        *out = e.st_;
    });
}

sg_status sg_reset_stats(sg_engine *eng) {
    return guarded(eng, [&](sg::Engine &e) {
        e.st_.edges_relaxed = e.st_.vertices_processed = 0;
        e.st_.rebalances = e.st_.resizes = e.st_.dropped_updates = 0;
        e.st_.relax_launches = 0;
        e.st_.relax_ms = 0;
        e.st_.relax_bytes = 0;
        e.timed_edges_ = e.timed_verts_ = e.timed_cedges_ = e.timed_pulls_ = 0;
        e.st_.pull_rounds = 0;
        e.st_.cache_fill_bytes = 0;
        e.st_.cache_fill_ms = 0;
        for (int c = 0; c < SG_KCLASS_COUNT; ++c) {
            e.kt_launches_[c] = e.kt_bytes_[c] = 0;
            e.kt_ms_[c] = 0;
        }
        e.st_.edges_from_cache = e.st_.streamed_bytes = 0;
        e.st_.stream_ms = 0;
        if (e.stats_) {
            e.fill<unsigned long long>(e.stats_, 16, 0ull);
            e.sync();
        }
    });
}

sg_status sg_device_alloc(sg_engine *eng, size_t bytes, void **dptr) {
    return guarded(eng, [&](sg::Engine &e) {
        if (!dptr) sg::fail(SG_EINVAL, "null output");
        *dptr = e.dalloc<uint8_t>(bytes);
    });
}

sg_status sg_device_free(sg_engine *eng, void *dptr) {
    return guarded(eng, [&](sg::Engine &e) {
        e.sync();
        e.dfree(dptr);
    });
}

sg_status sg_copy_to_device(sg_engine *eng, void *dst, const void *src, size_t bytes) {
    return guarded(eng, [&](sg::Engine &e) {
        if (!dst || (!src && bytes)) sg::fail(SG_EINVAL, "null pointer");
        SG_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, e.stream_));
        e.sync();
    });
}

sg_status sg_synchronize(sg_engine *eng) {
    return guarded(eng, [&](sg::Engine &e) { e.sync(); });
}

}  // extern "C"
