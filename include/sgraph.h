/*
 * sgraph.h — C ABI of the MI355X streaming-graph engine (libsgraph.so).
 *
 * The engine keeps a dynamic directed graph resident in HBM as a packed-memory
 * array (PMA) CSR with per-vertex gaps, applies update batches (inserts first,
 * then deletes) with HIP kernels, and recomputes one of four vertex programs
 * incrementally per snapshot: PageRank, SSSP, BFS, or directed min-ancestor
 * component labels.
 *
 * Every call returns an sg_status. No call exits the process and no C++
 * exception crosses this boundary. Host buffers passed in are read (or
 * written) during the call only; the engine owns all device memory it
 * allocates. An engine handle is single-threaded.
 *
 * Each entry point cites the reference interface it replaces (paths relative
 * to the C-GpuStreamGraph source tree).
 */
#ifndef SGRAPH_H
#define SGRAPH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum sg_status {
    SG_OK = 0,
    SG_EINVAL = 1,  /* bad argument, id out of range, wrong call order        */
    SG_ENOMEM = 2,  /* host allocation failed                                  */
    SG_EBUDGET = 3, /* device allocation would exceed the device budget        */
    SG_EDEVICE = 4, /* HIP runtime error or no usable device                   */
    SG_EIO = 5      /* file could not be opened, read or parsed                */
} sg_status;

typedef enum sg_algo {
    SG_ALGO_PR = 0,   /* operator PageRank, samples/hybrid_pr/hybrid_pr.cu:25-152   */
    SG_ALGO_SSSP = 1, /* operator SSSP, samples/hybrid_sssp/hybrid_sssp.cu:28-194   */
    SG_ALGO_BFS = 2,  /* operator BFS, samples/hybrid_bfs/hybrid_bfs.cu:26-167      */
    SG_ALGO_CC = 3,   /* operator CC, samples/hybrid_cc/hybrid_cc.cu:26-170         */
    SG_ALGO_MULTI = 4, /* batched BFS/SSSP queries on one graph (sg_init_queries); no
                          reference counterpart: the reference answers one source per
                          process (samples/hybrid_sssp/main.cu, samples/hybrid_bfs/main.cu) */
    SG_ALGO_PR_CC = 5, /* PageRank and CC labels kept together on one graph: one forward
                          store and one in-edge store serve both programs, each batch is
                          applied to them once, and CC's recompute after deletions reads
                          the in-lists of the reset vertices instead of sweeping every
                          edge. Results: sg_gather_query / sg_gather_at with qi 0 =
                          PageRank (f32 ranks), qi 1 = CC (u32 labels, parents). No
                          reference counterpart (HybridPageRank and HybridCC run in
                          separate processes on separate graphs) */
    SG_ALGO_CUSTOM = 6 /* an engine built from a caller's own operator by a create
                          function of sgraph_operator.h (framework.cuh:77 takes the
                          operator as a template argument); sg_create refuses it */
} sg_algo;

/* Edge-weight function of SSSP. Weights are computed from (u, v) on the fly
 * and never stored, as in the reference (push_functor.h:306). */
typedef enum sg_weight_mode {
    SG_WEIGHT_HASH128 = 0, /* w = (u + v) % 128 + 1, push_functor.h:306          */
    SG_WEIGHT_SEEDED = 1,  /* w = 1 + splitmix64(seed ^ (u << 32 | v)) % 63      */
    SG_WEIGHT_ONE = 2      /* w = 1                                              */
} sg_weight_mode;

typedef struct sg_config {
    uint32_t algo;            /* sg_algo                                                  */
    uint32_t weight_mode;     /* sg_weight_mode (SSSP only)                               */
    uint64_t weight_seed;     /* seed of SG_WEIGHT_SEEDED                                 */
    double pr_alpha;          /* damping 0.85, samples/hybrid_pr/hybrid_pr_common.h:38    */
    double pr_epsilon;        /* residual threshold 1e-6, samples/hybrid_pr/main.cu:37    */
    double pma_density;       /* initial PMA edge density 0.25, csr_graph.cuh:349-352,379 */
    uint32_t segments;        /* -SEGMENT vertex segments, framework.cuh:310-384          */
    uint32_t n_stream;        /* -n_stream (app_skeleton.h:89): segment copies in flight
                                 (staging buffers) of a streamed sweep, 2..8               */
    uint32_t pull_divisor;    /* a round whose frontier's out-edges exceed E /
                                 pull_divisor is a pull round over the in-edge store
                                 (pull_functor.h:22-369, sync_pull_dd.cuh:1-357).
                                 PageRank: 0 = 24, or 8 with the store in host memory;
                                 UINT32_MAX = push rounds only, no in-edge store.
                                 SSSP / BFS / CC and custom operators with the pull
                                 method, with SG_FLAG_IN_EDGES: 0 = 2; UINT32_MAX =
                                 push rounds only (the store still seeds recomputes) */
    uint32_t max_rounds;      /* cap on convergence rounds per phase (0 = 1000000)        */
    uint32_t rounds_per_sync; /* rounds enqueued between host convergence checks; 0 =
                                 adaptive: a snapshot phase first enqueues the rounds it
                                 took last time + 1 (4 at first), then doubles         */
    uint32_t flags;           /* SG_FLAG_*                                                */
    double device_budget_gb;  /* cap on engine device memory in GiB (0 = no cap)          */
    int32_t device;           /* HIP device ordinal                                       */
    uint32_t wave_budget;     /* light-tier edges per 64-vertex batch of a wave (0 = 1024) */
    uint32_t heavy_degree;    /* lists longer than this always become chunks (0 = 2048)  */
    uint32_t cache_level;     /* -cache: HBM edge cache of {0, 268435455, 536870921,
                                 805306365, 2^30} u32 slots for levels 0..4
                                 (graph_datum.cuh:151-163); used with host edges.
                                 PageRank engines with pull rounds spend it on the
                                 in-edge store's HBM copy instead (every pull round
                                 reads all in-edges); the rest cache hot out-lists */
    double edge_budget_gb;    /* device budget of the edge array in GiB
                                 (-device_graph_budget_gb): an edge array larger than
                                 this lives in pinned host memory (0 = no cap)       */
    uint32_t round_blocks;    /* blocks of a round kernel launch (0 = 1024)             */
    uint32_t reserved[1];
} sg_config;

/* Time the kernel classes (sg_kernel_class) with HIP events recorded on the
 * engine's stream around their launches. */
#define SG_FLAG_KERNEL_TIMING 0x1u
/* Keep the edge array in pinned host memory whatever its size (the
 * out-of-memory path of -cache / -SEGMENT, framework.cuh:310-384). */
#define SG_FLAG_HOST_EDGES 0x2u
/* PageRank only: sg_apply_snapshot updates the graph and the residuals
 * (cancel over the old out-lists, compensate over the new ones) but runs no
 * convergence rounds; sg_converge, or the next sg_gather / sg_gather_at /
 * sg_pr_topk, converges once for all deferred snapshots. The fixed point is
 * the same as converging after each one. Other algorithms ignore the flag. */
#define SG_FLAG_DEFER_CONVERGE 0x4u
/* Submit each batch of convergence rounds as one HIP graph (a chain of
 * kernel nodes built once per batch length and re-parameterised per batch)
 * instead of one launch per round. Results are identical. */
#define SG_FLAG_GRAPH_ROUNDS 0x8u
/* With SG_FLAG_KERNEL_TIMING: time only SG_KCLASS_SWEEP (one event pair per
 * snapshot); the other classes report no launches. Events between round
 * launches and their harvest after each round batch cost the timed snapshot
 * several percent, which this keeps out of the measurement. */
#define SG_FLAG_TIME_SWEEP 0x10u
/* SSSP / BFS / CC: keep an in-edge store (the reversed graph as a second PMA,
 * updated by every batch) and seed the recompute after deletions from the
 * in-lists of the reset vertices instead of a sweep over every live edge
 * (the reference re-pushes from every vertex, sync_push_dd.cuh:1470-1557).
 * Results are identical; the sweep's cost moves to the second PMA update.
 * The store also makes rounds direction-optimising (pull_divisor): a round
 * whose frontier holds a large share of the edges walks every in-list and
 * folds, per destination, what its in-neighbours in the frontier offer
 * (the operator's pull method), instead of pushing along every out-edge of
 * the frontier; both directions give the same values and parents. */
#define SG_FLAG_IN_EDGES 0x20u
/* PageRank with its in-edge store in pinned host memory: a pull round
 * streams every SEGMENT's slot range (gaps included) through HBM staging
 * buffers, the direction being chosen on the host before each round. Without
 * the flag a pull round reads the live slots in place over the host link and
 * rounds stay queued on the device. */
#define SG_FLAG_PULL_STREAM 0x40u

typedef struct sg_stats {
    uint32_t nv;                  /* vertices                                          */
    uint32_t rounds;              /* convergence rounds of the last phase              */
    uint64_t ne;                  /* live edges after the last snapshot                */
    uint64_t capacity;            /* PMA slots                                         */
    uint64_t snapshots;           /* snapshots applied                                 */
    uint64_t edges_relaxed;       /* edges traversed by relax kernels (cumulative)     */
    uint64_t vertices_processed;  /* frontier vertices expanded (cumulative)           */
    uint64_t invalidated;         /* vertices reset by deletions (last snapshot)       */
    uint64_t rebalances;          /* PMA windows re-spread (cumulative)                */
    uint64_t resizes;             /* PMA capacity doublings (cumulative)               */
    uint64_t dropped_updates;     /* update pairs with an id >= nv, ignored (cumul.)   */
    uint64_t relax_launches;      /* launches of the timed relax kernel                */
    double relax_ms;              /* summed HIP-event time of those launches           */
    uint64_t relax_bytes;         /* algorithmic bytes of those launches               */
    double last_phase_ms;         /* host wall time of the last sg_init/sg_apply_* call */
    double pr_residual_l1;        /* PR: sum |residual| after convergence (FP64 on GPU) */
    double pr_residual_max;       /* PR: max |residual| after convergence (FP64 on GPU) */
    uint64_t device_bytes;        /* device memory held by the engine                  */
    uint64_t host_edges;          /* 1 when the edge array is in pinned host memory    */
    uint64_t cache_slots;         /* HBM edge cache capacity (u32 slots)               */
    uint64_t cached_vertices;     /* vertices whose list is in the cache now           */
    uint64_t edges_from_cache;    /* relaxed edges read from the cache (cumulative)     */
    uint64_t streamed_bytes;      /* host-to-device bytes of streamed segments (cumul.) */
    double stream_ms;             /* wall time of the streamed sweeps (cumulative)      */
    uint64_t host_bytes;          /* pinned host memory held by the engine             */
    uint64_t cache_fill_bytes;    /* bytes copied over the host link into the HBM edge
                                     cache (cumulative; not in streamed_bytes)          */
    double cache_fill_ms;         /* wall time of those cache refills (cumulative)     */
    uint64_t pull_rounds;         /* PageRank rounds run as pulls over the in-edge
                                     store (cumulative)                                 */
    uint64_t in_capacity;         /* slots of the in-edge store (0 = none)             */
    uint64_t in_mirror_slots;     /* entries of the HBM copy of the in-store's leading
                                     live edges (host-resident store under an edge
                                     budget; 0 = none)                                 */
    uint64_t in_mirror_live;      /* live edges it holds (the first ones in vertex order) */
} sg_stats;

/* What sg_gather copies out. */
typedef enum sg_query {
    SG_Q_VALUE = 0,   /* per-vertex result: u32 levels/distances/labels (UINT32_MAX =
                         unreached), f32 ranks; GatherValue, graph_datum.cuh:297-315   */
    SG_Q_BUFFER = 1,  /* per-vertex buffer: f32 PR residual, u32 pending dist/label;
                         GatherBuffer, graph_datum.cuh:337-355                         */
    SG_Q_PARENT = 2,  /* per-vertex u32 parent (UINT32_MAX = none); GatherParent,
                         graph_datum.cuh:317-335                                       */
    SG_Q_DEGREE = 3,  /* per-vertex u32 out-degree                                      */
    SG_Q_OFFSETS = 4, /* u64[nv + 1] PMA slot offsets (last entry = capacity)           */
    SG_Q_SLOTS = 5,   /* u32[capacity] PMA slots (UINT32_MAX = gap)                     */
    SG_Q_SEGMENTS = 6, /* u32[segments + 1] first vertex of each SEGMENT (last = nv), the
                          greedy split by PMA span of Engine::LoadGraph, framework.cuh:310-362 */
    /* The in-edge store of an engine that keeps one (PageRank unless
     * pull_divisor is UINT32_MAX): the same three arrays for the reversed
     * graph, whose lists hold the sources of each vertex's in-edges. */
    SG_Q_IN_DEGREE = 7,  /* u32[nv] in-degrees                                          */
    SG_Q_IN_OFFSETS = 8, /* u64[nv + 1] in-store slot offsets (last entry = in_capacity)  */
    SG_Q_IN_SLOTS = 9    /* u32[in_capacity] in-store slots (UINT32_MAX = gap)            */
} sg_query;

typedef struct sg_engine sg_engine;

/* Fill *cfg with the reference's defaults (app_skeleton.h:42-100,
 * samples/hybrid_pr/main.cu:37, hybrid_pr_common.h:38). */
void sg_config_default(sg_config *cfg);

/* Create an engine on cfg->device: validates the device, creates the streams.
 * Replaces Engine::Engine (include/framework/framework.cuh:124-140). */
sg_status sg_create(const sg_config *cfg, sg_engine **out);

/* Release every device and host resource of the engine. NULL is a no-op. */
void sg_destroy(sg_engine *eng);

/* Text of the last error on this engine (or of the last sg_create failure
 * when eng is NULL). Never NULL. Errors print nothing by themselves;
 * replaces GROUTE_CUDA_CHECK's print-and-exit (groute/internal/cuda_utils.h:41-65). */
const char *sg_last_error(const sg_engine *eng);

/* Load a CSR snapshot G_0 (row[nv+1] offsets, col[ne] destinations, adjacency
 * order kept) into a device PMA with gaps spread in proportion to deg+1.
 * Replaces PMAGraph::Bind + spread_weighted_V1 (groute/graphs/csr_graph.cuh:384-511),
 * Engine::LoadGraph (framework.cuh:310-384) and PMAGraphAllocator
 * (csr_graph.cuh:1768-1905). */
sg_status sg_load_csr(sg_engine *eng, uint32_t nv, uint64_t ne,
                      const uint64_t *row, const uint32_t *col);

/* Initialise vertex state and run the static solve to convergence.
 * sources: SSSP/BFS source vertex (nsrc must be 1); ignored by PR and CC.
 * Replaces Engine::InitGraph + Engine::Start (framework.cuh:389-439, 549-584). */
sg_status sg_init(sg_engine *eng, const uint32_t *sources, uint32_t nsrc);

/* The seeded R-MAT graph of sg_gen_rmat (sgraph_io.h), the same vertices and
 * edges, drawn, sorted and de-duplicated on device `device` (the host
 * generator takes ≈110 s at scale 27 EF32). Arrays as sg_gen_rmat's (malloc,
 * sg_free). Needs about 17 bytes of device memory per draw (edge_factor <<
 * scale draws). No counterpart in the reference, which reads text files. */
sg_status sg_gen_rmat_device(int32_t device, uint32_t scale, uint32_t edge_factor, uint64_t seed, double a,
                             double b, double c, int permute, uint32_t *nv, uint64_t *ne, uint64_t **row,
                             uint32_t **col);

/* Checkpoint of a single-program engine (PR, SSSP, BFS, CC): sg_save_state
 * writes the current snapshot's adjacency lists (in PMA order), the vertex
 * state arrays and the traversal source to path (deferred PageRank snapshots
 * are converged first); sg_restore_state loads such a file into a fresh
 * engine of the same program and configuration (no sg_load_csr / sg_init
 * before it) without a static solve. Later snapshots then give the results
 * the saved engine would have given. No counterpart in the reference, whose
 * state lives only in the running process (GraphDatum, graph_datum.cuh:33-357). */
sg_status sg_save_state(sg_engine *eng, const char *path);
sg_status sg_restore_state(sg_engine *eng, const char *path);

/* SG_ALGO_MULTI engines: start nq queries on the loaded graph, query i from
 * sources[i], of kind kinds[i] (SG_ALGO_SSSP: weights of cfg.weight_mode;
 * SG_ALGO_BFS: unit weights, i.e. levels), and run them to convergence
 * together. Every later sg_apply_snapshot* updates the graph once and brings
 * all queries to their fixed points with one launch per round. Each query's
 * distances and parents equal those of a single-source SSSP/BFS engine.
 * sg_init on such an engine starts nsrc queries of one kind (BFS when
 * cfg.weight_mode is SG_WEIGHT_ONE, else SSSP). At most 4096 queries.
 * nq = 0 (sources and kinds may be NULL) drops the queries: later snapshots
 * only update the graph, which is how a request loop keeps answered one-shot
 * queries from being maintained.
 * Batches the per-source runs of Engine::InitGraph + Engine::Start
 * (framework.cuh:389-439, 549-584); no reference counterpart as one call. */
sg_status sg_init_queries(sg_engine *eng, const uint32_t *sources, const uint32_t *kinds, uint32_t nq);

/* Continuous batching on an SG_ALGO_MULTI engine: queries join free slots
 * while others are converging, and leave when their own work is done.
 *   sg_query_open     fixes `slots` query slots (<= 4096) and drops every
 *                     query (the graph stays);
 *   sg_query_add      puts n BFS/SSSP queries (kinds as sg_init_queries) into
 *                     free slots, written to slots[0..n); they start with the
 *                     next round, beside the queries already running;
 *   sg_query_step     runs up to max_rounds rounds for all running queries,
 *                     then lists in done[0..*ndone) (at most cap) the slots
 *                     whose query has converged since the last step
 *                     (*nactive = slots still running); a converged slot's
 *                     result is read with sg_gather_query / sg_gather_at
 *                     (qi = slot) and stays until released;
 *   sg_query_release  frees converged slots.
 * An update (sg_apply_snapshot*) needs every running query converged first
 * (SG_EINVAL otherwise); with no slot in use it only updates the graph, and
 * converged queries still held are brought to the new snapshot. Each query's
 * result equals a single-source engine's (the lexicographic (distance, parent)
 * minimum has one fixed point whatever joins or leaves meanwhile). No
 * reference counterpart (one query per process). */
sg_status sg_query_open(sg_engine *eng, uint32_t slots);
sg_status sg_query_add(sg_engine *eng, const uint32_t *sources, const uint32_t *kinds, uint32_t n, uint32_t *slots);
sg_status sg_query_step(sg_engine *eng, uint32_t max_rounds, uint32_t *done, uint32_t cap, uint32_t *ndone,
                        uint32_t *nactive);
sg_status sg_query_release(sg_engine *eng, const uint32_t *slots, uint32_t n);

/* Per-query result of an SG_ALGO_MULTI engine: what = SG_Q_VALUE (u32
 * distance or level, UINT32_MAX = unreached) or SG_Q_PARENT; bytes >= 4 * nv.
 * For an SG_ALGO_PR_CC engine qi selects the program: 0 = PageRank
 * (SG_Q_VALUE / SG_Q_BUFFER, f32; deferred snapshots are converged first),
 * 1 = CC (SG_Q_VALUE labels, SG_Q_PARENT, SG_Q_BUFFER).
 * GatherValue / GatherParent (graph_datum.cuh:297-335) for query qi. */
sg_status sg_gather_query(sg_engine *eng, uint32_t qi, uint32_t what, void *out, size_t bytes);

/* Apply one snapshot from host memory: ins_uv/del_uv are (u, v) pairs, i.e.
 * 2*nins / 2*ndel uint32 values. Inserts are applied first, then deletes
 * (every copy of a deleted (u, v) is removed); then the vertex program is
 * brought to its fixed point on the new graph. Pairs are validated on the host
 * first: an id >= nv returns SG_EINVAL and changes nothing.
 * Replaces, for SSSP/BFS/CC, Engine::del_edge + Engine::add_edge
 * (framework.cuh:975-979, 1025-1038, 1507-1655) and, for PR,
 * Engine::Cancelation + Engine::Compensate (framework.cuh:1110-1301). */
sg_status sg_apply_snapshot(sg_engine *eng,
                            const uint32_t *ins_uv, uint64_t nins,
                            const uint32_t *del_uv, uint64_t ndel);

/* Same as sg_apply_snapshot with the pairs already in device memory (for
 * example from sg_device_alloc + sg_copy_to_device). Pairs with an id >= nv
 * are ignored and counted in sg_stats.dropped_updates. */
sg_status sg_apply_snapshot_device(sg_engine *eng,
                                   const uint32_t *d_ins_uv, uint64_t nins,
                                   const uint32_t *d_del_uv, uint64_t ndel);

/* Copy per-vertex state or the PMA arrays to host memory. bytes must be at
 * least the size of the requested array (see sg_query).
 * Replaces GatherValue/GatherParent/GatherBuffer (graph_datum.cuh:297-355). */
sg_status sg_gather(sg_engine *eng, uint32_t query, void *out, size_t bytes);

/* Pointwise gather: out[i] = the u32 result word of vertices[i] (UINT32_MAX
 * for an id >= nv). query: SG_Q_VALUE, SG_Q_PARENT, SG_Q_BUFFER or
 * SG_Q_DEGREE (PageRank values and residuals are f32 bit patterns); qi is
 * the query index of an SG_ALGO_MULTI engine (ignored otherwise). For
 * lookups that need a few vertices rather than all of them (serving).
 * GatherValue / GatherParent (graph_datum.cuh:297-335) restricted to a list. */
sg_status sg_gather_at(sg_engine *eng, uint32_t query, uint32_t qi, const uint32_t *vertices, uint64_t n,
                       uint32_t *out);

/* sg_gather_at without any device allocation: the ids and results are
 * staged through a fixed scratch array in the library's device image (512 KiB
 * per device, held until the process ends) instead of buffers the engine
 * allocates. n <= 65536 (SG_EINVAL beyond); SG_EINVAL while PageRank
 * snapshots wait for a deferred convergence (sg_converge first). Every call
 * on one device shares that device's scratch: the library serialises them
 * with a per-device lock, so concurrent calls wait for each other. No
 * counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc). */
sg_status sg_gather_at_noalloc(sg_engine *eng, uint32_t query, uint32_t qi, const uint32_t *vertices, uint64_t n,
                               uint32_t *out);

/* PageRank engines: ids[0..k) = the k vertices of highest rank (ties: smaller
 * id first), ranks (optional) their values. The reference prints the top
 * ranks on the host after a full copy (hybrid_pr_host.cu:161-229, -top_ranks);
 * here the ranking is a device sort and only k entries cross the link. */
sg_status sg_pr_topk(sg_engine *eng, uint32_t k, uint32_t *ids, float *ranks);

/* Converge the snapshots deferred by SG_FLAG_DEFER_CONVERGE (a no-op
 * otherwise). No counterpart in the reference, which converges (or stops
 * after 100 rounds) inside every snapshot. */
sg_status sg_converge(sg_engine *eng);

/* Kernel classes timed with HIP events when SG_FLAG_KERNEL_TIMING is set
 * (events recorded on the engine's stream around each launch). */
typedef enum sg_kernel_class {
    SG_KCLASS_RELAX = 0,  /* convergence rounds of a snapshot (k_round / k_mround);
                             bytes: 12 per relaxed edge + 32 per expanded vertex   */
    SG_KCLASS_SWEEP = 1,  /* boundary sweep over all live edges (resident edges);
                             bytes: 4 per live edge + 16 per vertex (off, dpre)    */
    SG_KCLASS_MERGE = 2,  /* PMA batch merge: chunk count, chunk write, place       */
    SG_KCLASS_PULL = 3,   /* PageRank pull rounds over the in-edge store (k_pr_pull +
                             fix-up); bytes: 8 per in-edge (id + pushed value) + 24 per
                             vertex (in-store off and dpre, residual read and write) */
    SG_KCLASS_COUNT = 4
} sg_kernel_class;

/* Launches, summed HIP-event milliseconds and algorithmic bytes (0 where the
 * class has no byte model) of a kernel class since the last sg_reset_stats. */
sg_status sg_kernel_times(sg_engine *eng, uint32_t cls, uint64_t *launches, double *ms, uint64_t *bytes);

/* Counters of the engine (see sg_stats). */
sg_status sg_get_stats(sg_engine *eng, sg_stats *out);

/* Zero the cumulative counters and kernel timers. */
sg_status sg_reset_stats(sg_engine *eng);

/* Device-memory helpers for callers without an allocator of their own. The
 * allocation counts against the engine's device budget. */
sg_status sg_device_alloc(sg_engine *eng, size_t bytes, void **dptr);
sg_status sg_device_free(sg_engine *eng, void *dptr);
sg_status sg_copy_to_device(sg_engine *eng, void *dst, const void *src, size_t bytes);

/* Free and total memory of HIP device `device` (admission control of the
 * request loop). */
sg_status sg_device_memory(int32_t device, uint64_t *free_bytes, uint64_t *total_bytes);

/* Wait for all work of the engine. */
sg_status sg_synchronize(sg_engine *eng);

#ifdef __cplusplus
}
#endif
#endif /* SGRAPH_H */
