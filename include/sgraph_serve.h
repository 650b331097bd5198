/*
 * sgraph_serve.h — request loop over one dynamic graph (libsgraph.so).
 *
 * No counterpart in the reference: its drivers apply a fixed list of
 * snapshots and answer one source per process (samples/hybrid_{pr,sssp,bfs,cc}/main.cu,
 * HybridSSSP/HybridBFS/HybridCC/HybridPageRank). The server keeps the graph
 * and vertex state resident on the device and takes requests as they come:
 *
 *   update batches   applied in submission order, one snapshot each (the
 *                    engine's sg_apply_snapshot semantics: inserts, then
 *                    deletes);
 *   BFS / SSSP       one-shot queries from a source on the current snapshot,
 *                    answered by a batched engine (SG_ALGO_MULTI) with
 *                    continuous batching: a query takes a free slot between
 *                    round batches (every 8 rounds), runs beside the queries
 *                    already converging, and is answered as soon as its own
 *                    work is done; the slots are opened at creation, as many
 *                    as query_slots asks and the free HBM allows over
 *                    hbm_reserve_gb (none: BFS/SSSP requests fail with
 *                    SG_EBUDGET; so do waiting ones when the free HBM falls
 *                    under the reserve with no query running);
 *   CC label lookup  label of one vertex from a standing CC computation that
 *                    every update keeps at its fixed point;
 *   PR top-k         the k highest ranks of a standing PageRank, answered on
 *                    the latest snapshot its own thread has reached (the
 *                    snapshot is reported with the answer).
 *
 * A worker thread runs steps: apply pending updates (first letting the
 * running queries converge, so each answer belongs to one snapshot), answer
 * lookups, admit waiting queries into free slots, run a few rounds, answer the
 * queries that converged. If an engine fails to apply a snapshot the server
 * enters a failed state: the engines may hold different graphs, so that
 * request and every later one fail with the saved error.
 * A second thread applies the same snapshots, in order, to the PageRank
 * engine with deferred convergence (SG_FLAG_DEFER_CONVERGE) and converges
 * once per batch of top-k requests, on its own HIP stream, so PageRank work
 * overlaps the queries on the device. Submission and waiting are
 * thread-safe; every call returns an sg_status.
 */
#ifndef SGRAPH_SERVE_H
#define SGRAPH_SERVE_H

#include "sgraph.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef enum sg_request_kind {
    SG_REQ_UPDATE = 0,  /* update batch; no result                                   */
    SG_REQ_BFS = 1,     /* levels from arg; u32[nv], or u32 at `target`               */
    SG_REQ_SSSP = 2,    /* distances from arg (cfg weights); u32[nv], or u32 at target */
    SG_REQ_CC = 3,      /* CC label of vertex arg; u32                                */
    SG_REQ_PR_TOPK = 4  /* arg = k: u32[k] ids then f32[k] ranks                       */
} sg_request_kind;

#define SG_SERVE_CC 0x1u  /* maintain CC labels (needed by SG_REQ_CC)         */
#define SG_SERVE_PR 0x2u  /* maintain PageRank (needed by SG_REQ_PR_TOPK)      */

typedef struct sg_server_config {
    sg_config engine;       /* device, weights, PMA density, budgets (algo ignored) */
    uint32_t query_slots;   /* BFS/SSSP queries per step (1..4096, 0 = 64)          */
    uint32_t standing;      /* SG_SERVE_* programs kept at their fixed point        */
    double hbm_reserve_gb;  /* admission leaves this much HBM free                  */
    uint32_t reserved[4];
} sg_server_config;

typedef struct sg_result_info {
    uint64_t snapshot;   /* snapshots applied before the answer (updates: after) */
    double latency_ms;   /* submission to completion                            */
    uint64_t step;       /* worker step that completed the request              */
} sg_result_info;

typedef struct sg_server_stats {
    uint64_t steps;              /* worker steps run                              */
    uint64_t snapshots;          /* update batches applied                        */
    uint64_t queries;            /* BFS/SSSP queries answered                     */
    uint64_t lookups;            /* CC / PR requests answered                     */
    uint64_t edge_updates;       /* inserted + deleted pairs applied              */
    uint64_t edges_answered;     /* sum over answered queries of the edges of the
                                    snapshot they were answered on                */
    uint64_t max_batch;          /* most queries admitted in one step             */
    uint64_t deferred_for_memory;/* admissions postponed by the HBM reserve       */
    double p50_ms, p99_ms;       /* request latency, all completed requests        */
    double query_p50_ms, query_p99_ms;
    double update_p50_ms, update_p99_ms;
    double busy_ms;              /* time the worker spent in steps                */
    double wall_ms;              /* since the server started                      */
    uint64_t reserved[4];
} sg_server_stats;

typedef struct sg_server sg_server;

void sg_server_config_default(sg_server_config *cfg);

/* Load G_0, bring the standing programs to their fixed points, start the
 * worker. row/col as sg_load_csr. */
sg_status sg_server_create(const sg_server_config *cfg, uint32_t nv, uint64_t ne, const uint64_t *row,
                           const uint32_t *col, sg_server **out);

/* Queue a request. UPDATE: (ins_uv, nins, del_uv, ndel) pairs, copied; ids >=
 * nv are rejected here. BFS/SSSP: arg = source, target = a vertex for a
 * single-value answer or UINT32_MAX for the whole vector. CC: arg = vertex.
 * PR_TOPK: arg = k. *ticket identifies the request for sg_server_wait. */
sg_status sg_server_submit(sg_server *s, uint32_t kind, uint32_t arg, uint32_t target, const uint32_t *ins_uv,
                           uint64_t nins, const uint32_t *del_uv, uint64_t ndel, uint64_t *ticket);

/* Block until the request completes, copy its result to out (bytes at least
 * the result's size; NULL for updates) and release it. One thread may wait
 * for a ticket (a second waiter gets SG_EINVAL); after SG_EINVAL for a too
 * small buffer the ticket can be waited for again. */
sg_status sg_server_wait(sg_server *s, uint64_t ticket, void *out, size_t bytes, sg_result_info *info);

sg_status sg_server_get_stats(sg_server *s, sg_server_stats *out);

/* Text of the calling thread's last failed sg_server_* call (for a failed
 * request, the worker's message; for a failed sg_server_create, s may be NULL). */
const char *sg_server_last_error(const sg_server *s);

/* Finish the queued requests, stop the worker, release everything. */
void sg_server_destroy(sg_server *s);

#ifdef __cplusplus
}
#endif
#endif /* SGRAPH_SERVE_H */
