// Request loop over one dynamic graph (declared in sgraph_serve.h). No
// counterpart in the reference, whose drivers run a fixed snapshot list with
// one source per process (samples/hybrid_*/main.cu). The worker thread owns
// the batched-query and CC engines; a second thread owns the PageRank engine,
// so its long convergences overlap the queries on the device instead of
// delaying them. Submitters only touch the request tables under the mutex.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "sgraph.h"
#include "sgraph_serve.h"

namespace {

using Clock = std::chrono::steady_clock;

struct Request {
    uint32_t kind = 0, arg = 0, target = 0;
    std::vector<uint32_t> ins, del;
    std::vector<uint8_t> result;
    sg_result_info info{};
    sg_status status = SG_OK;
    std::string error;
    bool done = false;
    bool claimed = false;  // a thread waits for it (one waiter per ticket)
    Clock::time_point t_submit;
};

double pct(std::vector<double> v, double p) {
    if (v.empty()) return 0.0;
    std::sort(v.begin(), v.end());
    const double r = p / 100.0 * (double)(v.size() - 1);
    const size_t i = (size_t)r;
    const double f = r - (double)i;
    return i + 1 < v.size() ? v[i] * (1 - f) + v[i + 1] * f : v[i];
}

}  // namespace

// Last error of a failed sg_server_* call, per calling thread (several
// client threads submit and wait at once).
thread_local std::string t_error;

// An update batch the worker has applied, on its way to the PageRank thread.
struct PrUpdate {
    uint64_t snapshot = 0;
    std::vector<uint32_t> ins, del;
};

struct sg_server {
    sg_server_config cfg{};
    uint32_t nv = 0;
    sg_engine *multi = nullptr, *cc = nullptr, *pr = nullptr;
    uint32_t slots = 0;  // query slots opened on the batched engine (0: no HBM for them)

    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<uint64_t> q_updates, q_queries, q_lookups;
    std::unordered_map<uint64_t, std::unique_ptr<Request>> reqs;
    uint64_t next_id = 1;
    bool stop = false;
    std::thread worker;

    // After an engine fails to apply a snapshot the engines may hold different
    // graphs: every later request fails with the saved error.
    sg_status failed_status = SG_OK;
    std::string failed;

    // PageRank thread: applies the worker's snapshots in order with deferred
    // convergence and answers top-k requests on the snapshot it has reached
    // (converging once for all snapshots since the last answer).
    std::deque<PrUpdate> q_pr;
    std::deque<uint64_t> q_topk;
    std::condition_variable cv_pr;
    std::thread pr_worker;
    bool pr_stop = false;
    uint64_t pr_snapshot = 0;
    std::string pr_error;

    // queries in the batched engine's slots (worker thread only)
    std::unordered_map<uint32_t, uint64_t> running;  // slot -> request id

    uint64_t snapshot = 0, step = 0;
    std::vector<double> lat_all, lat_q, lat_u;
    sg_server_stats st{};
    Clock::time_point t_start;

    void fail_request(Request *r, sg_status code, const std::string &msg) {
        r->status = code;
        r->error = msg;
    }

    std::string engine_error(sg_engine *e) { return std::string(sg_last_error(e)); }

    Request *req(uint64_t id) {
        std::lock_guard<std::mutex> lk(mu);
        return reqs[id].get();
    }

    // Complete requests (under the lock) and wake waiters.
    void complete(const std::vector<std::pair<uint64_t, Request *>> &done) {
        if (done.empty()) return;
        const auto now = Clock::now();
        std::lock_guard<std::mutex> lk(mu);
        for (auto &d : done) {
            Request *r = d.second;
            r->info.latency_ms = std::chrono::duration<double, std::milli>(now - r->t_submit).count();
            r->info.step = step;
            r->done = true;
            lat_all.push_back(r->info.latency_ms);
            if (r->kind == SG_REQ_UPDATE) lat_u.push_back(r->info.latency_ms);
            else if (r->kind == SG_REQ_BFS || r->kind == SG_REQ_SSSP) lat_q.push_back(r->info.latency_ms);
        }
        cv_done.notify_all();
    }

    // Query slots the free HBM allows beside what the engines hold, keeping
    // cfg.hbm_reserve_gb free: a slot's state is about nv * 8 bytes plus three
    // 64-bit masks per vertex per 64 slots.
    uint32_t affordable_slots(uint32_t want) {
        uint64_t fr = 0, tot = 0;
        if (sg_device_memory(cfg.engine.device, &fr, &tot) != SG_OK) return want;
        const double reserve = cfg.hbm_reserve_gb * 1073741824.0;
        auto bytes = [&](uint32_t k) {
            uint32_t qs = 1;
            while (qs < std::min<uint32_t>(k, 64)) qs <<= 1;
            const uint64_t g = (k + qs - 1) / qs, kp = g * qs;
            return (double)nv * (double)kp * 8.0 + 3.0 * (double)nv * (double)g * 8.0;
        };
        uint32_t k = want;
        while (k > 0 && (double)fr - bytes(k) < reserve) k /= 2;
        return k;
    }

    // Answer the queries whose slots converged, then free the slots.
    void finish_slots(const uint32_t *done, uint32_t n, uint64_t ne) {
        std::vector<std::pair<uint64_t, Request *>> out;
        std::vector<uint32_t> rel;
        for (uint32_t i = 0; i < n; ++i) {
            auto it = running.find(done[i]);
            if (it == running.end()) continue;
            Request *r = req(it->second);
            sg_status g;
            if (r->target == 0xFFFFFFFFu) {
                r->result.resize((size_t)nv * 4);
                g = sg_gather_query(multi, done[i], SG_Q_VALUE, r->result.data(), r->result.size());
            } else {
                r->result.resize(4);
                g = sg_gather_at(multi, SG_Q_VALUE, done[i], &r->target, 1, (uint32_t *)r->result.data());
            }
            if (g != SG_OK) fail_request(r, g, engine_error(multi));
            r->info.snapshot = snapshot;
            out.push_back({it->second, r});
            rel.push_back(done[i]);
            running.erase(it);
        }
        if (!rel.empty() && sg_query_release(multi, rel.data(), (uint32_t)rel.size()) != SG_OK)
            set_failed(SG_EDEVICE, engine_error(multi));
        {
            std::lock_guard<std::mutex> lk(mu);
            st.queries += out.size();
            st.edges_answered += ne * out.size();
        }
        complete(out);
    }

    // Rounds for the running queries (all of them when `drain`), answering
    // each as it converges.
    void run_queries(bool drain, uint64_t ne) {
        std::vector<uint32_t> done(std::max<uint32_t>(slots, 1));
        do {
            if (running.empty()) return;
            uint32_t nd = 0, na = 0;
            if (sg_query_step(multi, drain ? 64 : kRoundsPerStep, done.data(), (uint32_t)done.size(), &nd, &na) != SG_OK) {
                set_failed(SG_EDEVICE, engine_error(multi));
                fail_running();
                return;
            }
            finish_slots(done.data(), nd, ne);
        } while (drain);
    }
    static constexpr uint32_t kRoundsPerStep = 8;  // rounds between admissions

    void set_failed(sg_status code, const std::string &msg) {
        std::lock_guard<std::mutex> lk(mu);
        if (failed_status == SG_OK) {
            failed_status = code;
            failed = msg;
        }
    }
    void fail_running() {
        std::vector<std::pair<uint64_t, Request *>> out;
        for (auto &kv : running) {
            Request *r = req(kv.second);
            fail_request(r, failed_status, failed);
            out.push_back({kv.second, r});
        }
        running.clear();
        complete(out);
    }

    uint64_t graph_edges() {
        sg_stats es{};
        return sg_get_stats(multi, &es) == SG_OK ? es.ne : 0;
    }

    // One step of the worker: update batches (at a point where no query is
    // converging), CC lookups, admission of waiting queries into free slots,
    // and a few rounds for every running query.
    void run_step(std::vector<uint64_t> &U, std::vector<uint64_t> &L) {
        std::vector<std::pair<uint64_t, Request *>> done;
        uint64_t ne = graph_edges();
        if (!U.empty()) {
            run_queries(true, ne);  // running queries finish on the snapshot they started on
            for (uint64_t id : U) {
                Request *r = req(id);
                const uint64_t ni = r->ins.size() / 2, nd = r->del.size() / 2;
                sg_status fs;
                {
                    std::lock_guard<std::mutex> lk(mu);
                    fs = failed_status;
                }
                if (fs != SG_OK) {
                    fail_request(r, fs, failed);
                    done.push_back({id, r});
                    continue;
                }
                for (sg_engine *e : {multi, cc}) {
                    if (!e) continue;
                    const sg_status s = sg_apply_snapshot(e, r->ins.data(), ni, r->del.data(), nd);
                    if (s != SG_OK) {
                        // this engine (and the ones after it) missed the batch: no
                        // later answer can name one snapshot for every program
                        set_failed(s, engine_error(e));
                        fail_request(r, s, engine_error(e));
                        break;
                    }
                }
                if (r->status == SG_OK) {
                    ++snapshot;
                    r->info.snapshot = snapshot;
                    if (pr) {
                        PrUpdate pu;
                        pu.snapshot = snapshot;
                        pu.ins = std::move(r->ins);
                        pu.del = std::move(r->del);
                        std::lock_guard<std::mutex> lk(mu);
                        q_pr.push_back(std::move(pu));
                        cv_pr.notify_one();
                    }
                    std::lock_guard<std::mutex> lk(mu);
                    st.edge_updates += ni + nd;
                }
                r->ins.clear();
                r->ins.shrink_to_fit();
                r->del.clear();
                r->del.shrink_to_fit();
                done.push_back({id, r});
            }
            {
                std::lock_guard<std::mutex> lk(mu);
                st.snapshots = snapshot;
            }
            ne = graph_edges();
        }
        // CC label lookups on the current snapshot
        std::vector<uint32_t> verts;
        std::vector<Request *> ccr;
        for (uint64_t id : L) {
            Request *r = req(id);
            r->info.snapshot = snapshot;
            verts.push_back(r->arg);
            ccr.push_back(r);
            done.push_back({id, r});
        }
        if (!ccr.empty()) {
            std::vector<uint32_t> out(verts.size());
            sg_status fs;
            {
                std::lock_guard<std::mutex> lk(mu);
                fs = failed_status;
            }
            const sg_status s = fs != SG_OK ? fs : sg_gather_at(cc, SG_Q_VALUE, 0, verts.data(), verts.size(), out.data());
            for (size_t i = 0; i < ccr.size(); ++i) {
                if (s != SG_OK) { fail_request(ccr[i], s, fs != SG_OK ? failed : engine_error(cc)); continue; }
                ccr[i]->result.resize(4);
                std::memcpy(ccr[i]->result.data(), &out[i], 4);
            }
            std::lock_guard<std::mutex> lk(mu);
            st.lookups += L.size();
        }
        complete(done);
        // admission: waiting queries into free slots, as the free HBM allows
        std::vector<uint64_t> Q;
        uint32_t deferred = 0;
        {
            std::lock_guard<std::mutex> lk(mu);
            const uint32_t free_slots = slots - (uint32_t)running.size();
            const uint32_t want = (uint32_t)std::min<size_t>(q_queries.size(), free_slots);
            uint32_t k = want;
            if (want && reserve_short()) k = 0;
            deferred = want - k;
            for (uint32_t i = 0; i < k; ++i) {
                Q.push_back(q_queries.front());
                q_queries.pop_front();
            }
            st.deferred_for_memory += deferred;
            if (failed_status != SG_OK || slots == 0 || (deferred && running.empty())) {
                // nothing can run them: a failed server, no slots at all, or HBM
                // short with no running query whose end would free any
                std::vector<uint64_t> all(q_queries.begin(), q_queries.end());
                q_queries.clear();
                for (uint64_t id : all) Q.push_back(id);
                if (failed_status == SG_OK) st.deferred_for_memory += all.size();  // refused for memory
                refuse_ = true;
            }
        }
        if (refuse_) {
            std::vector<std::pair<uint64_t, Request *>> out;
            sg_status code;
            std::string msg;
            {
                std::lock_guard<std::mutex> lk(mu);
                code = failed_status != SG_OK ? failed_status : SG_EBUDGET;
                msg = failed_status != SG_OK ? failed
                                             : "no device memory for query state under the HBM reserve (hbm_reserve_gb)";
            }
            for (uint64_t id : Q) {
                Request *r = req(id);
                r->info.snapshot = snapshot;
                fail_request(r, code, msg);
                out.push_back({id, r});
            }
            refuse_ = false;
            complete(out);
            Q.clear();
        }
        if (!Q.empty()) {
            std::vector<uint32_t> src, kinds, slot(Q.size());
            for (uint64_t id : Q) {
                Request *r = req(id);
                src.push_back(r->arg);
                kinds.push_back(r->kind == SG_REQ_BFS ? (uint32_t)SG_ALGO_BFS : (uint32_t)SG_ALGO_SSSP);
            }
            const sg_status s = sg_query_add(multi, src.data(), kinds.data(), (uint32_t)src.size(), slot.data());
            if (s != SG_OK) {
                std::vector<std::pair<uint64_t, Request *>> out;
                for (uint64_t id : Q) {
                    Request *r = req(id);
                    fail_request(r, s, engine_error(multi));
                    out.push_back({id, r});
                }
                complete(out);
            } else {
                for (size_t i = 0; i < Q.size(); ++i) running[slot[i]] = Q[i];
                std::lock_guard<std::mutex> lk(mu);
                st.max_batch = std::max<uint64_t>(st.max_batch, running.size());
            }
        }
        run_queries(false, ne);
    }
    bool refuse_ = false;
    // Free HBM below the reserve (the other engines' batch scratch can grow).
    bool reserve_short() {
        uint64_t fr = 0, tot = 0;
        if (sg_device_memory(cfg.engine.device, &fr, &tot) != SG_OK) return false;
        return (double)fr < cfg.hbm_reserve_gb * 1073741824.0;
    }

    void pr_loop() {
        for (;;) {
            std::vector<PrUpdate> U;
            std::vector<uint64_t> K;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv_pr.wait(lk, [&] { return pr_stop || !q_pr.empty() || !q_topk.empty(); });
                if (pr_stop && q_pr.empty() && q_topk.empty()) break;
                for (auto &u : q_pr) U.push_back(std::move(u));
                q_pr.clear();
                K.assign(q_topk.begin(), q_topk.end());
                q_topk.clear();
            }
            for (auto &u : U) {
                if (!pr_error.empty()) break;
                const sg_status s = sg_apply_snapshot(pr, u.ins.data(), u.ins.size() / 2, u.del.data(), u.del.size() / 2);
                if (s != SG_OK) {
                    pr_error = engine_error(pr);
                    set_failed(s, pr_error);
                } else {
                    pr_snapshot = u.snapshot;
                }
            }
            if (K.empty()) continue;
            std::vector<std::pair<uint64_t, Request *>> done;
            uint32_t kmax = 0;
            {
                std::lock_guard<std::mutex> lk(mu);
                for (uint64_t id : K) {
                    Request *r = reqs[id].get();
                    kmax = std::max(kmax, std::min<uint32_t>(r->arg, nv));
                    done.push_back({id, r});
                }
            }
            std::vector<uint32_t> ids(kmax);
            std::vector<float> rk(kmax);
            sg_status s = SG_EDEVICE;
            if (pr_error.empty()) {
                s = sg_pr_topk(pr, kmax, ids.data(), rk.data());  // converges the deferred snapshots
                if (s != SG_OK) {
                    pr_error = engine_error(pr);
                    set_failed(s, pr_error);
                }
            }
            for (auto &d : done) {
                Request *r = d.second;
                r->info.snapshot = pr_snapshot;
                if (s != SG_OK) { fail_request(r, s, pr_error); continue; }
                const uint32_t k = std::min<uint32_t>(r->arg, nv);
                r->result.resize((size_t)k * 8);
                std::memcpy(r->result.data(), ids.data(), (size_t)k * 4);
                std::memcpy(r->result.data() + (size_t)k * 4, rk.data(), (size_t)k * 4);
            }
            {
                std::lock_guard<std::mutex> lk(mu);
                st.lookups += done.size();
            }
            complete(done);
        }
    }

    void loop() {
        for (;;) {
            std::vector<uint64_t> U, L;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv_work.wait(lk, [&] {
                    return stop || !q_updates.empty() || !q_queries.empty() || !q_lookups.empty() || !running.empty();
                });
                if (stop && q_updates.empty() && q_queries.empty() && q_lookups.empty() && running.empty()) break;
                U.assign(q_updates.begin(), q_updates.end());
                q_updates.clear();
                L.assign(q_lookups.begin(), q_lookups.end());
                q_lookups.clear();
                ++step;
            }
            const auto t0 = Clock::now();
            run_step(U, L);
            std::lock_guard<std::mutex> lk(mu);
            st.busy_ms += std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
            st.steps = step;
        }
    }
};

extern "C" {

void sg_server_config_default(sg_server_config *cfg) {
    if (!cfg) return;
    std::memset(cfg, 0, sizeof *cfg);
    sg_config_default(&cfg->engine);
    cfg->query_slots = 64;
    cfg->standing = SG_SERVE_CC | SG_SERVE_PR;
    cfg->hbm_reserve_gb = 1.0;
}

static sg_status make_engine(const sg_server_config *cfg, uint32_t algo, uint32_t nv, uint64_t ne, const uint64_t *row,
                             const uint32_t *col, sg_engine **out, std::string &err) {
    sg_config c = cfg->engine;
    c.algo = algo;
    sg_status s = sg_create(&c, out);
    if (s != SG_OK) { err = sg_last_error(nullptr); return s; }
    s = sg_load_csr(*out, nv, ne, row, col);
    if (s == SG_OK && algo != SG_ALGO_MULTI) s = sg_init(*out, nullptr, 0);
    if (s != SG_OK) err = sg_last_error(*out);
    return s;
}

sg_status sg_server_create(const sg_server_config *cfg, uint32_t nv, uint64_t ne, const uint64_t *row,
                           const uint32_t *col, sg_server **out) {
    if (!cfg || !out) return SG_EINVAL;
    *out = nullptr;
    std::unique_ptr<sg_server> s(new (std::nothrow) sg_server());
    if (!s) return SG_ENOMEM;
    s->cfg = *cfg;
    s->nv = nv;
    s->slots = cfg->query_slots ? std::min<uint32_t>(cfg->query_slots, 4096) : 64;  // asked; capped by HBM below
    std::string err;
    sg_status st = SG_OK;
    if (cfg->standing & SG_SERVE_CC) st = make_engine(cfg, SG_ALGO_CC, nv, ne, row, col, &s->cc, err);
    if (st == SG_OK && (cfg->standing & SG_SERVE_PR)) {
        sg_server_config pc = *cfg;
        pc.engine.flags |= SG_FLAG_DEFER_CONVERGE;  // one convergence per top-k answer
        st = make_engine(&pc, SG_ALGO_PR, nv, ne, row, col, &s->pr, err);
    }
    if (st == SG_OK) st = make_engine(cfg, SG_ALGO_MULTI, nv, ne, row, col, &s->multi, err);
    if (st == SG_OK) {  // query slots: as many as asked and as the free HBM allows over the reserve
        s->slots = s->affordable_slots(s->slots);
        if (s->slots) {
            st = sg_query_open(s->multi, s->slots);
            if (st != SG_OK) err = sg_last_error(s->multi);
        } else {
            st = sg_init_queries(s->multi, nullptr, nullptr, 0);  // updates only change the graph
            if (st != SG_OK) err = sg_last_error(s->multi);
        }
    }
    if (st != SG_OK) {
        t_error = err;
        sg_destroy(s->multi);
        sg_destroy(s->cc);
        sg_destroy(s->pr);
        return st;
    }
    s->t_start = Clock::now();
    sg_server *raw = s.get();
    try {
        if (s->pr) s->pr_worker = std::thread([raw] { raw->pr_loop(); });
        s->worker = std::thread([raw] { raw->loop(); });
    } catch (...) {
        if (s->pr_worker.joinable()) {
            {
                std::lock_guard<std::mutex> lk(s->mu);
                s->pr_stop = true;
            }
            s->cv_pr.notify_all();
            s->pr_worker.join();
        }
        sg_destroy(s->multi);
        sg_destroy(s->cc);
        sg_destroy(s->pr);
        return SG_ENOMEM;
    }
    *out = s.release();
    return SG_OK;
}

sg_status sg_server_submit(sg_server *s, uint32_t kind, uint32_t arg, uint32_t target, const uint32_t *ins_uv,
                           uint64_t nins, const uint32_t *del_uv, uint64_t ndel, uint64_t *ticket) {
    if (!s || !ticket) return SG_EINVAL;
    std::unique_ptr<Request> r(new (std::nothrow) Request());
    if (!r) return SG_ENOMEM;
    r->kind = kind;
    r->arg = arg;
    r->target = target;
    switch (kind) {
        case SG_REQ_UPDATE:
            if ((nins && !ins_uv) || (ndel && !del_uv)) { t_error = "null update array"; return SG_EINVAL; }
            for (uint64_t i = 0; i < 2 * nins; ++i)
                if (ins_uv[i] >= s->nv) { t_error = "insert id >= nv"; return SG_EINVAL; }
            for (uint64_t i = 0; i < 2 * ndel; ++i)
                if (del_uv[i] >= s->nv) { t_error = "delete id >= nv"; return SG_EINVAL; }
            r->ins.assign(ins_uv, ins_uv + 2 * nins);
            r->del.assign(del_uv, del_uv + 2 * ndel);
            break;
        case SG_REQ_BFS:
        case SG_REQ_SSSP:
            if (arg >= s->nv || (target != 0xFFFFFFFFu && target >= s->nv)) { t_error = "vertex id >= nv"; return SG_EINVAL; }
            break;
        case SG_REQ_CC:
            if (!s->cc) { t_error = "CC labels are not maintained (SG_SERVE_CC)"; return SG_EINVAL; }
            if (arg >= s->nv) { t_error = "vertex id >= nv"; return SG_EINVAL; }
            break;
        case SG_REQ_PR_TOPK:
            if (!s->pr) { t_error = "PageRank is not maintained (SG_SERVE_PR)"; return SG_EINVAL; }
            if (arg == 0) { t_error = "k must be at least 1"; return SG_EINVAL; }
            break;
        default:
            t_error = "unknown request kind";
            return SG_EINVAL;
    }
    r->t_submit = Clock::now();
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop) { t_error = "server is stopping"; return SG_EINVAL; }
    if (s->failed_status != SG_OK) { t_error = "server failed earlier: " + s->failed; return s->failed_status; }
    const uint64_t id = s->next_id++;
    s->reqs[id] = std::move(r);
    *ticket = id;
    if (kind == SG_REQ_PR_TOPK) {
        s->q_topk.push_back(id);
        s->cv_pr.notify_one();
        return SG_OK;
    }
    if (kind == SG_REQ_UPDATE) s->q_updates.push_back(id);
    else if (kind == SG_REQ_BFS || kind == SG_REQ_SSSP) s->q_queries.push_back(id);
    else s->q_lookups.push_back(id);
    s->cv_work.notify_one();
    return SG_OK;
}

sg_status sg_server_wait(sg_server *s, uint64_t ticket, void *out, size_t bytes, sg_result_info *info) {
    if (!s) return SG_EINVAL;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) { t_error = "unknown ticket"; return SG_EINVAL; }
    Request *r = it->second.get();
    if (r->claimed) { t_error = "another thread is waiting for this ticket"; return SG_EINVAL; }
    r->claimed = true;  // the request lives until this waiter releases it
    s->cv_done.wait(lk, [&] { return r->done; });
    sg_status st = r->status;
    if (st != SG_OK) t_error = r->error;
    else if (!r->result.empty()) {
        if (!out || bytes < r->result.size()) {
            t_error = "output smaller than the result (" + std::to_string(r->result.size()) + " bytes)";
            st = SG_EINVAL;
        } else {
            std::memcpy(out, r->result.data(), r->result.size());
        }
    }
    if (info) *info = r->info;
    if (st != SG_EINVAL || r->status != SG_OK) s->reqs.erase(it);  // a too-small buffer may retry
    else r->claimed = false;
    return st;
}

sg_status sg_server_get_stats(sg_server *s, sg_server_stats *out) {
    if (!s || !out) return SG_EINVAL;
    std::lock_guard<std::mutex> lk(s->mu);
    *out = s->st;
    out->p50_ms = pct(s->lat_all, 50);
    out->p99_ms = pct(s->lat_all, 99);
    out->query_p50_ms = pct(s->lat_q, 50);
    out->query_p99_ms = pct(s->lat_q, 99);
    out->update_p50_ms = pct(s->lat_u, 50);
    out->update_p99_ms = pct(s->lat_u, 99);
    out->wall_ms = std::chrono::duration<double, std::milli>(Clock::now() - s->t_start).count();
    return SG_OK;
}

const char *sg_server_last_error(const sg_server *) { return t_error.c_str(); }

void sg_server_destroy(sg_server *s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    if (s->worker.joinable()) s->worker.join();  // may still hand snapshots to the PageRank thread
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->pr_stop = true;
    }
    s->cv_pr.notify_all();
    if (s->pr_worker.joinable()) s->pr_worker.join();
    sg_destroy(s->multi);
    sg_destroy(s->cc);
    sg_destroy(s->pr);
    delete s;
}

}  // extern "C"
