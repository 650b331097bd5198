// Command-line front end: hybrid_pr, hybrid_sssp, hybrid_bfs, hybrid_cc.
//
// Same binaries, flag names and snapshot loop as the reference:
//   flags          include/utils/app_skeleton.h:42-100, samples/hybrid_pr/main.cu:37,
//                  samples/hybrid_sssp/hybrid_sssp.cu:16-19, hybrid_pr_host.cu:36-38
//   driver loops   samples/hybrid_pr/hybrid_pr.cu:154-201, hybrid_sssp.cu:204-297,
//                  hybrid_cc.cu:172-208, hybrid_bfs.cu:169-208
//   output files   hybrid_sssp_host.cu:169-188, hybrid_bfs_host.cpp:105-127 ("%u %u" / "%u INF"),
//                  hybrid_pr_host.cu:161-229 (ALPHA/EPSILON header, ranks sorted)
// The flag syntax is gflags' (-name value, -name=value, --name, -noname for
// booleans). Reference flags that steer nothing in this engine are accepted and
// ignored. The number of snapshots is the number of lines of the stream-size
// file (the reference always runs 10, hybrid_pr.cu:175), capped by -snapshots.
// Exit status: 0 on success, 1 on a bad flag or input, 2 on an engine error.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <numeric>
#include <string>
#include <vector>

#include "sgraph.h"
#include "sgraph_io.h"

#ifndef SG_CLI_ALGO
#define SG_CLI_ALGO SG_ALGO_SSSP
#endif

namespace {

struct Flag {
    enum Kind { STR, I64, F64, BOOL } kind;
    std::string s;
    long long i = 0;
    double f = 0;
    bool b = false;
    const char *help;
};

std::map<std::string, Flag> &flags() {
    static std::map<std::string, Flag> m;
    return m;
}
void def_s(const char *n, const char *v, const char *h) { flags()[n] = Flag{Flag::STR, v, 0, 0, false, h}; }
void def_i(const char *n, long long v, const char *h) { flags()[n] = Flag{Flag::I64, "", v, 0, false, h}; }
void def_f(const char *n, double v, const char *h) { flags()[n] = Flag{Flag::F64, "", 0, v, false, h}; }
void def_b(const char *n, bool v, const char *h) { flags()[n] = Flag{Flag::BOOL, "", 0, 0, v, h}; }
const std::string &S(const char *n) { return flags()[n].s; }
long long I(const char *n) { return flags()[n].i; }
double F(const char *n) { return flags()[n].f; }
bool B(const char *n) { return flags()[n].b; }

void define_flags() {
    // input (app_skeleton.h:63-68, 97-100)
    def_s("graphfile", "", "edge-list graph file: 'u v [w]' per line");
    def_s("format", "market_big", "graph format (market_big)");
    def_s("updatefile", "", "update file: 'a u v' / 'd u v' lines");
    def_s("update_size", "", "stream-size file: 'n_add n_del' per snapshot");
    def_s("snapshot_file", "", "binary snapshot file (graph, stream sizes, updates) read instead of the three text files");
    def_s("write_snapshot_file", "", "write the inputs read as a binary snapshot file, then run");
    def_i("cache", 4, "HBM edge cache level 0..4 (graph_datum.cuh:151-163); used when the edge array is in host memory");
    def_i("SEGMENT", 1, "vertex segments streamed from host memory (framework.cuh:310-384)");
    def_i("n_stream", 3, "number of streams");
    def_i("block_size", 512, "accepted for compatibility; launch shapes are the engine's own");
    def_s("lb_push", "", "load-balancing policy for push");
    def_s("lb_pull", "", "load-balancing policy for pull");
    def_i("weight_num", 0, "1: synthesise graph-file weights");
    def_b("weight", true, "read weighted update lists");
    def_i("hybrid", 0, "0:zerocopy 1:explicit 2:hybrid (accepted)");
    // algorithms
    def_i("source_node", 0, "source of SSSP/BFS (clamped to [0, nnodes-1])");
    def_f("error", 0.000001, "PR error tolerance");
    def_i("top_ranks", 10, "PR ranks to print");
    def_b("print_ranks", false, "write ranks to the output file");
    def_b("norm", false, "normalise PR output ranks (L1)");
    // general (app_skeleton.h:53-60)
    def_s("output", "", "file to store the final result to");
    def_b("check", true, "accepted (results are checked by the test suite)");
    def_b("verbose", true, "verbose prints");
    def_b("trace", false, "accepted");
    def_s("json", "", "accepted");
    def_b("stats", false, "accepted");
    def_b("undirected", false, "accepted");
    def_b("interactive", false, "accepted");
    def_s("cmdfile", "", "accepted");
    def_i("max_iteration", 50, "accepted");
    def_b("priority", false, "accepted");
    def_i("prio_delta", 0, "accepted");
    def_i("priority_a", 1, "accepted");
    def_i("residence", 0, "accepted");
    def_f("edge_factor", 1.0, "accepted");
    def_f("alpha", 0.8, "accepted (direction-optimising heuristic of the reference)");
    def_f("beta", 0.4, "accepted");
    def_f("wl_alloc_factor", 0.2, "accepted");
    def_b("wl_sort", false, "accepted");
    def_b("wl_unique", false, "accepted");
    def_i("grid_size", 20, "accepted");
    def_b("gen_graph", false, "accepted");
    def_i("gen_nnodes", 100000, "accepted");
    def_i("gen_factor", 10, "accepted");
    def_i("gen_method", 0, "accepted");
    def_b("gen_weights", false, "accepted");
    def_i("gen_weight_range", 100, "accepted");
    def_b("sparse", false, "accepted");
    def_b("non_atomic", true, "accepted");
    def_f("cut_threshold", 0, "accepted");
    // additions of this engine
    def_s("weight_mode", "hash128", "SSSP weight: hash128 ((u+v)%128+1), seeded ([1,63]), one");
    def_i("weight_seed", 27491095, "seed of -weight_mode seeded");
    def_i("snapshots", -1, "apply at most this many snapshots (-1: all lines of -update_size)");
    def_s("dump_dir", "", "write the result after every snapshot to <dir>/snapshot_<t>.txt");
    def_f("device_graph_budget_gb", 0, "device budget of the edge array in GiB; a larger edge array stays in pinned host memory (0: none)");
    def_f("pma_density", 0.25, "initial PMA edge density");
    def_i("device", 0, "HIP device");
}

bool parse_bool(const std::string &v, bool &out) {
    if (v == "1" || v == "true" || v == "t" || v == "yes" || v == "y") { out = true; return true; }
    if (v == "0" || v == "false" || v == "f" || v == "no" || v == "n") { out = false; return true; }
    return false;
}

bool parse_flags(int argc, char **argv) {
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        if (a.size() < 2 || a[0] != '-') continue;  // positional arguments are left alone, as gflags does
        a = a.substr(a[1] == '-' ? 2 : 1);
        std::string name = a, val;
        bool has_val = false;
        const size_t eq = a.find('=');
        if (eq != std::string::npos) { name = a.substr(0, eq); val = a.substr(eq + 1); has_val = true; }
        if (name == "help") {
            for (auto &kv : flags()) std::printf("  -%s: %s\n", kv.first.c_str(), kv.second.help);
            std::exit(0);
        }
        auto it = flags().find(name);
        if (it == flags().end() && name.rfind("no", 0) == 0) {
            auto jt = flags().find(name.substr(2));
            if (jt != flags().end() && jt->second.kind == Flag::BOOL && !has_val) { jt->second.b = false; continue; }
        }
        if (it == flags().end()) { std::fprintf(stderr, "ERROR: unknown command line flag '%s'\n", name.c_str()); return false; }
        Flag &f = it->second;
        if (f.kind == Flag::BOOL && !has_val) {
            // gflags: a following bare true/false is not consumed; "-flag" alone means true
            f.b = true;
            continue;
        }
        if (!has_val) {
            if (i + 1 >= argc) { std::fprintf(stderr, "ERROR: flag '-%s' is missing its argument\n", name.c_str()); return false; }
            val = argv[++i];
        }
        char *end = nullptr;
        switch (f.kind) {
            case Flag::STR: f.s = val; break;
            case Flag::I64: f.i = std::strtoll(val.c_str(), &end, 10); if (!end || *end) goto bad; break;
            case Flag::F64: f.f = std::strtod(val.c_str(), &end); if (!end || *end) goto bad; break;
            case Flag::BOOL: if (!parse_bool(val, f.b)) goto bad; break;
        }
        continue;
    bad:
        std::fprintf(stderr, "ERROR: illegal value '%s' specified for flag '%s'\n", val.c_str(), name.c_str());
        return false;
    }
    return true;
}

double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

bool write_result(const char *path, uint32_t algo, const std::vector<uint32_t> &u32, const std::vector<float> &f32) {
    FILE *f = std::fopen(path, "w");
    if (!f) { std::fprintf(stderr, "Could not open '%s' for writing\n", path); return false; }
    if (algo == SG_ALGO_PR) {
        // hybrid_pr_host.cu:161-229: ranks sorted descending (stable on id)
        std::vector<uint32_t> ord(f32.size());
        std::iota(ord.begin(), ord.end(), 0u);
        std::stable_sort(ord.begin(), ord.end(), [&](uint32_t a, uint32_t b) { return f32[a] < f32[b]; });
        double sum = 0;
        for (float x : f32) sum += x;
        std::fprintf(f, "ALPHA %*e EPSILON %*e\n", 6, 0.85, 6, F("error"));
        std::fprintf(f, "RANKS 1--%d of %d\n", (int)f32.size(), (int)f32.size());
        for (size_t i = 1; i <= f32.size(); ++i) {
            const uint32_t v = ord[f32.size() - i];
            if (!B("print_ranks")) std::fprintf(f, "%zu %u\n", i, v);
            else if (B("norm")) std::fprintf(f, "%u %*e\n", v, 6, f32[v] / sum);
            else std::fprintf(f, "%u %*e\n", v, 6, (double)f32[v]);
        }
    } else {
        for (size_t i = 0; i < u32.size(); ++i) {
            if (u32[i] == 0xFFFFFFFFu) std::fprintf(f, "%zu INF\n", i);
            else std::fprintf(f, "%zu %u\n", i, u32[i]);
        }
    }
    return std::fclose(f) == 0;
}

int fail_engine(sg_engine *e, const char *what, sg_status st) {
    std::fprintf(stderr, "ERROR: %s failed (status %d): %s\n", what, (int)st, sg_last_error(e));
    sg_destroy(e);
    return 2;
}

}  // namespace

int main(int argc, char **argv) {
    define_flags();
    if (!parse_flags(argc, argv)) return 1;
    const uint32_t algo = SG_CLI_ALGO;
    static const char *names[] = {"page rank", "sssp", "bfs", "cc"};
    if (S("graphfile").empty() && S("snapshot_file").empty()) {
        std::fprintf(stderr, "ERROR: -graphfile (or -snapshot_file) is required\n");
        return 1;
    }
    if (S("format") != "market_big" && S("format") != "market") {
        std::fprintf(stderr, "ERROR: -format %s is not supported (market_big)\n", S("format").c_str());
        return 1;
    }
    std::printf("\nTesting single GPU %s\n--------------------\n\n", names[algo]);
    uint32_t nv = 0;
    uint64_t ne = 0, *row = nullptr;
    uint32_t *col = nullptr;
    // stream: sizes[2 t] / sizes[2 t + 1] inserts / deletes of snapshot t, taken in order from ins / del
    uint32_t nsnap = 0;
    uint64_t *sizes = nullptr, nins = 0, ndel = 0;
    uint32_t *ins = nullptr, *del = nullptr;
    bool have_updates = false;
    double t0 = now_ms();
    sg_status st;
    if (!S("snapshot_file").empty()) {
        st = sg_read_snapshot_file(S("snapshot_file").c_str(), &nv, &ne, &row, &col, &nsnap, &sizes, &ins, &del);
        if (st != SG_OK || nv == 0) {
            std::fprintf(stderr, "ERROR: cannot read snapshot file '%s'\n", S("snapshot_file").c_str());
            return 1;
        }
        for (uint32_t t = 0; t < nsnap; ++t) {
            nins += sizes[2 * t];
            ndel += sizes[2 * t + 1];
        }
        have_updates = nsnap > 0;
    } else {
        st = sg_read_edge_list(S("graphfile").c_str(), &nv, &ne, &row, &col);
        if (st != SG_OK || nv == 0) { std::fprintf(stderr, "ERROR: cannot read graph '%s'\n", S("graphfile").c_str()); return 1; }
        if (!S("updatefile").empty()) {
            if (S("update_size").empty() || sg_read_stream_sizes(S("update_size").c_str(), &nsnap, &sizes) != SG_OK) {
                std::fprintf(stderr, "ERROR: cannot read -update_size file\n");
                return 1;
            }
            if (sg_read_updates(S("updatefile").c_str(), &nins, &ins, &ndel, &del) != SG_OK) {
                std::fprintf(stderr, "ERROR: cannot read -updatefile\n");
                return 1;
            }
            have_updates = true;
        }
    }
    if (!S("write_snapshot_file").empty()) {
        // the stream the run would apply: the snapshots the updates cover
        uint32_t ns = 0;
        uint64_t ta = 0, td = 0;
        while (ns < nsnap && ta + sizes[2 * ns] <= nins && td + sizes[2 * ns + 1] <= ndel) {
            ta += sizes[2 * ns];
            td += sizes[2 * ns + 1];
            ++ns;
        }
        if (sg_write_snapshot_file(S("write_snapshot_file").c_str(), nv, ne, row, col, ns, sizes, ins, del) != SG_OK) {
            std::fprintf(stderr, "ERROR: cannot write snapshot file '%s'\n", S("write_snapshot_file").c_str());
            return 1;
        }
    }
    if (B("verbose")) std::printf("graph: %u vertices, %llu edges (read %.1f ms)\n", nv, (unsigned long long)ne, now_ms() - t0);

    sg_config cfg;
    sg_config_default(&cfg);
    cfg.algo = algo;
    const std::string wm = S("weight_mode");
    cfg.weight_mode = wm == "seeded" ? SG_WEIGHT_SEEDED : wm == "one" ? SG_WEIGHT_ONE : SG_WEIGHT_HASH128;
    if (wm != "seeded" && wm != "one" && wm != "hash128") { std::fprintf(stderr, "ERROR: bad -weight_mode\n"); return 1; }
    cfg.weight_seed = (uint64_t)I("weight_seed");
    cfg.pr_epsilon = F("error");
    cfg.segments = (uint32_t)std::max<long long>(1, I("SEGMENT"));
    cfg.n_stream = (uint32_t)std::max<long long>(1, I("n_stream"));
    cfg.edge_budget_gb = F("device_graph_budget_gb");  // an edge array above this stays in pinned host memory
    cfg.cache_level = (uint32_t)std::max<long long>(0, std::min<long long>(4, I("cache")));
    cfg.pma_density = F("pma_density");
    cfg.device = (int32_t)I("device");
    sg_engine *eng = nullptr;
    if ((st = sg_create(&cfg, &eng)) != SG_OK) {
        std::fprintf(stderr, "ERROR: sg_create failed (status %d): %s\n", (int)st, sg_last_error(nullptr));
        return 2;
    }
    if ((st = sg_load_csr(eng, nv, ne, row, col)) != SG_OK) return fail_engine(eng, "sg_load_csr", st);
    sg_free(row);
    sg_free(col);
    // hybrid_sssp.cu:222: source clamped to [0, nnodes-1]
    uint32_t src = (uint32_t)std::min<long long>(std::max<long long>(0, I("source_node")), (long long)nv - 1);
    t0 = now_ms();
    st = (algo == SG_ALGO_SSSP || algo == SG_ALGO_BFS) ? sg_init(eng, &src, 1) : sg_init(eng, nullptr, 0);
    if (st != SG_OK) return fail_engine(eng, "sg_init", st);
    if (B("verbose")) std::printf("static solve: %.3f ms\n", now_ms() - t0);

    auto gather = [&](std::vector<uint32_t> &u32, std::vector<float> &f32) -> sg_status {
        if (algo == SG_ALGO_PR) { f32.resize(nv); return sg_gather(eng, SG_Q_VALUE, f32.data(), f32.size() * 4); }
        u32.resize(nv);
        return sg_gather(eng, SG_Q_VALUE, u32.data(), u32.size() * 4);
    };
    std::vector<uint32_t> u32;
    std::vector<float> f32;

    if (have_updates) {
        if (I("snapshots") >= 0) nsnap = (uint32_t)std::min<long long>(nsnap, I("snapshots"));
        uint64_t ia = 0, id = 0;
        double total = 0;
        for (uint32_t t = 0; t < nsnap; ++t) {
            const uint64_t na = sizes[2 * t], nd = sizes[2 * t + 1];
            if (ia + na > nins || id + nd > ndel) {
                std::fprintf(stderr, "ERROR: snapshot %u needs more updates than the update file holds\n", t);
                sg_destroy(eng);
                return 1;
            }
            const double ts = now_ms();
            st = sg_apply_snapshot(eng, ins + 2 * ia, na, del + 2 * id, nd);
            if (st != SG_OK) return fail_engine(eng, "sg_apply_snapshot", st);
            const double ms = now_ms() - ts;
            total += ms;
            ia += na;
            id += nd;
            sg_stats s;
            sg_get_stats(eng, &s);
            std::printf("batch number %u: +%llu -%llu edges, %llu live, %.3f ms, %u rounds\n", t, (unsigned long long)na,
                        (unsigned long long)nd, (unsigned long long)s.ne, ms, s.rounds);
            if (!S("dump_dir").empty()) {
                if (gather(u32, f32) != SG_OK) return fail_engine(eng, "sg_gather", SG_EDEVICE);
                const std::string p = S("dump_dir") + "/snapshot_" + std::to_string(t) + ".txt";
                if (!write_result(p.c_str(), algo, u32, f32)) { sg_destroy(eng); return 1; }
            }
        }
        std::printf("%u snapshots in %.3f ms\n", nsnap, total);
    }
    sg_free(sizes);
    sg_free(ins);
    sg_free(del);
    if ((st = gather(u32, f32)) != SG_OK) return fail_engine(eng, "sg_gather", st);
    if (algo == SG_ALGO_PR) {
        for (uint32_t i = 0; i < std::min<uint32_t>(20, nv); ++i) std::printf("v %u data %f\n", i, f32[i]);  // hybrid_pr.cu:190-193
    } else if (B("verbose")) {
        uint64_t reached = 0;
        for (uint32_t x : u32) reached += x != 0xFFFFFFFFu;
        std::printf("reached %llu of %u vertices\n", (unsigned long long)reached, nv);
    }
    if (!S("output").empty() && !write_result(S("output").c_str(), algo, u32, f32)) { sg_destroy(eng); return 1; }
    sg_destroy(eng);
    std::printf("Overall: Test passed\n");
    return 0;
}
