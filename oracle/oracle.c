/*
 * oracle.c — CPU restatement of the reference's algorithms for the
 * streaming-graph hot path. TEST INFRASTRUCTURE ONLY: it is the checker the
 * parity tests compare the HIP engine against, and the CPU baseline that
 * bench.py times. The product (libsgraph.so, the CLI binaries) never links,
 * loads or calls it.
 *
 * All graphs are CSR: row[nv + 1] (u64 offsets), col[row[nv]] (u32 targets).
 * "Unreached" is UINT32_MAX, as IDENTITY_ELEMENT in the reference
 * (samples/hybrid_sssp/hybrid_sssp_common.h:35).
 *
 * Function                 follows                                              pinned by (tests/)
 * or_bfs_levels            samples/hybrid_bfs/hybrid_bfs_host.cpp:33-60 BFSHost  README fixture, path/star closed forms, brute force
 * or_sssp_dijkstra         textbook Dijkstra, weights of push_functor.h:306      README fixture, Floyd-Warshall on tiny graphs
 * or_sssp_fifo             samples/hybrid_sssp/hybrid_sssp_host.cu:34-63         equality with or_sssp_dijkstra
 *                          (SSSPHostNaive, weight function as a parameter)
 * or_cc_min_ancestor       label(v) = min{u : u reaches v}, the fixed point of   README fixture, transitive-closure brute force
 *                          samples/hybrid_cc/hybrid_cc.cu:62-110
 * or_cc_label_prop         FIFO min-label propagation (no reference CPU path)    equality with or_cc_min_ancestor
 * or_pr_fp64               fixed point of samples/hybrid_pr/hybrid_pr.cu:43-136  dense linear solve, cycle/star closed forms
 *                          r = (1-a) + a * A^T D^-1 r, dangling mass dropped
 * or_pr_host_f32           samples/hybrid_pr/hybrid_pr_host.cu:43-124 PageRankHost  agreement with or_pr_fp64 (rel 1e-4)
 * or_apply_updates         csr_graph.cuh:658-811 (insert appends at the end of the  README fixture graphs G1/G2, brute-force
 *                          source's list; del_edge removes matches by left shift), edge-multiset model
 *                          framework.cuh:1001-1023 (all inserts, then all deletes)
 * or_pma_geometry          csr_graph.cuh:364-382 compute_capacity                 hand-computed values
 * or_segment_partition     framework.cuh:310-362 LoadGraph greedy split          hand-computed values
 *
 * Certificates (checks of a result in O(V + E), for graphs where the from-
 * scratch oracles above are too slow; each returns the number of violations):
 * or_sp_certificate        d(src) = 0; d(v) <= d(u) + w(u, v) on every edge;     oracle outputs pass, planted
 *                          every reached v != src has a parent edge (p, v) with   mistakes (a raised or lowered
 *                          d(v) = d(p) + w(p, v): with w >= 1 d is the distance   distance, a wrong parent) fail
 * or_cc_certificate        label(v) <= v; label(v) <= label(u) on every edge      oracle outputs pass, planted
 *                          u -> v; parent edges exist, keep the label, and every  mistakes fail
 *                          parent chain ends at the vertex label(v) (so label(v)
 *                          reaches v and is below every ancestor: the min ancestor)
 * or_pr_residual           FP64 residual of r = (1-a) + a A^T D^-1 r at a given  zero (to rounding) at the
 *                          rank vector: sum and max |res|. With M = a A^T D^-1     or_pr_fp64 fixed point; the
 *                          (column sums <= a), ||r - r*||_1 <= ||res||_1/(1 - a)  bound holds on random vectors
 *
 * Programs with no reference counterpart (operators a caller compiles in):
 * or_widest_path           widest (maximum-bottleneck) path from a source:       brute force over all simple
 *                          textbook Dijkstra variant (largest width first);       paths on tiny graphs; an
 *                          width(src) = UINT32_MAX, unreachable 0                 independent fixed-point
 *                                                                                 iteration (tests/)
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define OR_INF 0xFFFFFFFFu

/* ---- weights (identical formulas to the engine's; see sgraph.h) -------- */
static uint64_t or_splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

uint32_t or_weight(uint32_t mode, uint64_t seed, uint32_t u, uint32_t v) {
    if (mode == 0) return (uint32_t)(((uint64_t)u + v) % 128u) + 1u; /* push_functor.h:306 */
    if (mode == 1) return 1u + (uint32_t)(or_splitmix64(seed ^ (((uint64_t)u << 32) | v)) % 63ull);
    if (mode == 3) return (uint32_t)(((uint64_t)u + v) % 128u);       /* hybrid_sssp_host.cu:52 as written */
    return 1u;
}

/* ---- BFS: hybrid_bfs_host.cpp:33-60 (FIFO, first visit sets the level) -- */
int or_bfs_levels(uint32_t nv, const uint64_t *row, const uint32_t *col, uint32_t src, uint32_t *out) {
    uint32_t *q;
    uint64_t head = 0, tail = 0;
    for (uint32_t i = 0; i < nv; ++i) out[i] = OR_INF;
    if (src >= nv) return 1;
    q = (uint32_t *)malloc((size_t)(nv ? nv : 1) * sizeof(uint32_t));
    if (!q) return 2;
    out[src] = 0;
    q[tail++] = src;
    while (head < tail) {
        uint32_t u = q[head++];
        for (uint64_t e = row[u]; e < row[u + 1]; ++e) {
            uint32_t v = col[e];
            if (out[v] == OR_INF) { out[v] = out[u] + 1; q[tail++] = v; }
        }
    }
    free(q);
    return 0;
}

/* ---- Dijkstra with a binary heap keyed by (dist, vertex) ---------------- */
/* heap entries are dist << 32 | vertex */

static void heap_push(uint64_t *h, uint64_t *n, uint64_t k) {
    uint64_t i = (*n)++;
    h[i] = k;
    while (i > 0) {
        uint64_t p = (i - 1) / 2;
        if (h[p] <= h[i]) break;
        uint64_t t = h[p]; h[p] = h[i]; h[i] = t; i = p;
    }
}

static uint64_t heap_pop(uint64_t *h, uint64_t *n) {
    uint64_t top = h[0], i = 0;
    h[0] = h[--(*n)];
    for (;;) {
        uint64_t l = 2 * i + 1, r = l + 1, m = i;
        if (l < *n && h[l] < h[m]) m = l;
        if (r < *n && h[r] < h[m]) m = r;
        if (m == i) break;
        uint64_t t = h[m]; h[m] = h[i]; h[i] = t; i = m;
    }
    return top;
}

int or_sssp_dijkstra(uint32_t nv, const uint64_t *row, const uint32_t *col, uint32_t src,
                     uint32_t wmode, uint64_t seed, uint32_t *out) {
    uint64_t ne = nv ? row[nv] : 0, n = 0;
    uint64_t *h;
    for (uint32_t i = 0; i < nv; ++i) out[i] = OR_INF;
    if (src >= nv) return 1;
    h = (uint64_t *)malloc((size_t)(ne + 1) * sizeof(uint64_t));
    if (!h) return 2;
    out[src] = 0;
    heap_push(h, &n, (uint64_t)src);
    while (n) {
        uint64_t k = heap_pop(h, &n);
        uint32_t d = (uint32_t)(k >> 32), u = (uint32_t)k;
        if (d != out[u]) continue; /* stale entry */
        for (uint64_t e = row[u]; e < row[u + 1]; ++e) {
            uint32_t v = col[e];
            uint64_t nd = (uint64_t)d + or_weight(wmode, seed, u, v);
            if (nd < out[v]) { out[v] = (uint32_t)nd; heap_push(h, &n, (nd << 32) | v); }
        }
    }
    free(h);
    return 0;
}

/* ---- widest path: Dijkstra on widths, widest first ---------------------
 * Heap keys (~width) << 32 | vertex pop the widest vertex first; a popped
 * width is final because widths only shrink along a path. */
int or_widest_path(uint32_t nv, const uint64_t *row, const uint32_t *col, uint32_t src,
                   uint32_t wmode, uint64_t seed, uint32_t *out) {
    uint64_t ne = nv ? row[nv] : 0, n = 0;
    uint64_t *h;
    for (uint32_t i = 0; i < nv; ++i) out[i] = 0;
    if (src >= nv) return 1;
    h = (uint64_t *)malloc((size_t)(ne + 1) * sizeof(uint64_t));
    if (!h) return 2;
    out[src] = OR_INF;
    heap_push(h, &n, (uint64_t)src); /* key ~UINT32_MAX = 0 */
    while (n) {
        uint64_t k = heap_pop(h, &n);
        uint32_t wu = ~(uint32_t)(k >> 32), u = (uint32_t)k;
        if (wu != out[u]) continue; /* stale entry */
        for (uint64_t e = row[u]; e < row[u + 1]; ++e) {
            uint32_t v = col[e];
            uint32_t w = or_weight(wmode, seed, u, v);
            uint32_t nw = w < wu ? w : wu;
            if (nw > out[v]) { out[v] = nw; heap_push(h, &n, ((uint64_t)(uint32_t)~nw << 32) | v); }
        }
    }
    free(h);
    return 0;
}

/* ---- SSSPHostNaive: hybrid_sssp_host.cu:34-63 (FIFO Bellman-Ford) -------- */
int or_sssp_fifo(uint32_t nv, const uint64_t *row, const uint32_t *col, uint32_t src,
                 uint32_t wmode, uint64_t seed, uint32_t *out) {
    /* ring buffer queue; a vertex may be queued many times, as in the reference */
    uint64_t cap = 1024, head = 0, size = 0;
    uint32_t *q;
    for (uint32_t i = 0; i < nv; ++i) out[i] = OR_INF;
    if (src >= nv) return 1;
    q = (uint32_t *)malloc((size_t)cap * sizeof(uint32_t));
    if (!q) return 2;
    out[src] = 0;
    q[0] = src; size = 1;
    while (size) {
        uint32_t u = q[head];
        head = (head + 1) & (cap - 1); --size;
        uint32_t d = out[u];
        for (uint64_t e = row[u]; e < row[u + 1]; ++e) {
            uint32_t v = col[e];
            uint64_t nd = (uint64_t)d + or_weight(wmode, seed, u, v);
            if (out[v] > nd) {
                out[v] = (uint32_t)nd;
                if (size == cap) { /* grow: unroll the ring into a doubled buffer */
                    uint32_t *nq = (uint32_t *)malloc((size_t)cap * 2 * sizeof(uint32_t));
                    if (!nq) { free(q); return 2; }
                    for (uint64_t i = 0; i < size; ++i) nq[i] = q[(head + i) & (cap - 1)];
                    free(q); q = nq; head = 0; cap *= 2;
                }
                q[(head + size) & (cap - 1)] = v; ++size;
            }
        }
    }
    free(q);
    return 0;
}

/* ---- directed min-ancestor labels ---------------------------------------
 * label(v) = min{u : u reaches v} (u = v included). Visiting roots in
 * increasing id order and stopping at already-labelled vertices is exact: a
 * labelled vertex was reached from a smaller root, and so was everything it
 * reaches. */
int or_cc_min_ancestor(uint32_t nv, const uint64_t *row, const uint32_t *col, uint32_t *out) {
    uint32_t *stack = (uint32_t *)malloc((size_t)(nv ? nv : 1) * sizeof(uint32_t));
    if (!stack) return 2;
    for (uint32_t i = 0; i < nv; ++i) out[i] = OR_INF;
    for (uint32_t r = 0; r < nv; ++r) {
        uint64_t top = 0;
        if (out[r] != OR_INF) continue;
        out[r] = r;
        stack[top++] = r;
        while (top) {
            uint32_t u = stack[--top];
            for (uint64_t e = row[u]; e < row[u + 1]; ++e) {
                uint32_t v = col[e];
                if (out[v] == OR_INF) { out[v] = r; stack[top++] = v; }
            }
        }
    }
    free(stack);
    return 0;
}

/* Textbook FIFO min-label propagation (CPU baseline for CC; the reference has
 * no CPU CC path). Every vertex starts active with its own id. */
int or_cc_label_prop(uint32_t nv, const uint64_t *row, const uint32_t *col, uint32_t *out) {
    uint32_t *q = (uint32_t *)malloc((size_t)(nv ? nv : 1) * sizeof(uint32_t));
    uint8_t *inq = (uint8_t *)calloc(nv ? nv : 1, 1);
    uint64_t head = 0, size = 0;
    if (!q || !inq) { free(q); free(inq); return 2; }
    for (uint32_t i = 0; i < nv; ++i) { out[i] = i; q[i] = i; inq[i] = 1; }
    size = nv;
    while (size) {
        uint32_t u = q[head];
        head = (head + 1 == nv) ? 0 : head + 1; --size;
        inq[u] = 0;
        for (uint64_t e = row[u]; e < row[u + 1]; ++e) {
            uint32_t v = col[e];
            if (out[u] < out[v]) {
                out[v] = out[u];
                if (!inq[v]) { inq[v] = 1; q[(head + size) % nv] = v; ++size; }
            }
        }
    }
    free(q); free(inq);
    return 0;
}

/* ---- PageRank fixed point in FP64 ----------------------------------------
 * r = (1 - alpha) * 1 + alpha * A^T D^-1 r with D = out-degree (duplicates
 * counted) and the mass of dangling vertices dropped — the fixed point of the
 * delta-push operator (hybrid_pr.cu:43-136: value 0, buffer 1-alpha, push
 * alpha*x/deg). Jacobi iteration until max |r_{k+1} - r_k| <= tol. Returns the
 * iteration count, or -1 when max_iter is reached. */
int or_pr_fp64(uint32_t nv, const uint64_t *row, const uint32_t *col, double alpha,
               double tol, int max_iter, double *out) {
    double *nx = (double *)malloc((size_t)(nv ? nv : 1) * sizeof(double));
    double *c = (double *)malloc((size_t)(nv ? nv : 1) * sizeof(double));
    int it;
    if (!nx || !c) { free(nx); free(c); return -2; }
    for (uint32_t i = 0; i < nv; ++i) out[i] = 1.0 - alpha;
    for (it = 1; it <= max_iter; ++it) {
        double diff = 0.0;
        for (uint32_t u = 0; u < nv; ++u) {
            uint64_t d = row[u + 1] - row[u];
            c[u] = d ? alpha * out[u] / (double)d : 0.0;
        }
        for (uint32_t i = 0; i < nv; ++i) nx[i] = 1.0 - alpha;
        for (uint32_t u = 0; u < nv; ++u)
            for (uint64_t e = row[u]; e < row[u + 1]; ++e) nx[col[e]] += c[u];
        for (uint32_t i = 0; i < nv; ++i) {
            double dd = fabs(nx[i] - out[i]);
            if (dd > diff) diff = dd;
            out[i] = nx[i];
        }
        if (diff <= tol) break;
    }
    free(nx); free(c);
    return it > max_iter ? -1 : it;
}

/* ---- PageRankHost: hybrid_pr_host.cu:43-124 -------------------------------
 * f32 ranks start at 1-alpha, residuals at the first push of every vertex;
 * FIFO worklist; a vertex is queued when its residual crosses eps from below. */
int or_pr_host_f32(uint32_t nv, const uint64_t *row, const uint32_t *col, double alpha,
                   double eps, float *ranks) {
    /* the reference's std::queue may hold a vertex more than once per pass (a
     * vertex reset to 0 later in the same pass can cross eps again), so the
     * worklists grow on demand */
    float *res = (float *)calloc(nv ? nv : 1, sizeof(float));
    uint64_t cap_in = nv ? nv : 1, cap_out = nv ? nv : 1, n_in = 0, n_out = 0;
    uint32_t *in_wl = (uint32_t *)malloc((size_t)cap_in * sizeof(uint32_t));
    uint32_t *out_wl = (uint32_t *)malloc((size_t)cap_out * sizeof(uint32_t));
    int iterations = 0;
    if (!res || !in_wl || !out_wl) { free(res); free(in_wl); free(out_wl); return -2; }
    for (uint32_t u = 0; u < nv; ++u) ranks[u] = (float)(1.0 - alpha);
    for (uint32_t u = 0; u < nv; ++u) {
        uint64_t d = row[u + 1] - row[u];
        if (!d) continue;
        float upd = (float)(((1.0 - alpha) * alpha) / (double)d);
        for (uint64_t e = row[u]; e < row[u + 1]; ++e) res[col[e]] += upd;
    }
    for (uint32_t u = 0; u < nv; ++u) in_wl[n_in++] = u;
    while (n_in) {
        n_out = 0;
        for (uint64_t i = 0; i < n_in; ++i) {
            uint32_t u = in_wl[i];
            float r = res[u];
            ranks[u] += r;
            res[u] = 0.0f;
            uint64_t d = row[u + 1] - row[u];
            if (!d) continue;
            float upd = (float)((double)r * alpha / (double)d);
            for (uint64_t e = row[u]; e < row[u + 1]; ++e) {
                uint32_t v = col[e];
                float prev = res[v];
                res[v] += upd;
                if (prev + upd > (float)eps && prev < (float)eps) {
                    if (n_out == cap_out) {
                        uint32_t *g = (uint32_t *)realloc(out_wl, (size_t)cap_out * 2 * sizeof(uint32_t));
                        if (!g) { free(res); free(in_wl); free(out_wl); return -2; }
                        out_wl = g; cap_out *= 2;
                    }
                    out_wl[n_out++] = v;
                }
            }
        }
        ++iterations;
        { uint32_t *t = in_wl; in_wl = out_wl; out_wl = t; }
        { uint64_t t = cap_in; cap_in = cap_out; cap_out = t; }
        n_in = n_out;
    }
    free(res); free(in_wl); free(out_wl);
    return iterations;
}

/* ---- update model: csr_graph.cuh:658-811 + framework.cuh:1001-1023 -------
 * Per source vertex the adjacency sequence is: surviving old entries in
 * order, then inserted targets in file order; every entry equal to a deleted
 * target is removed (del_edge's left shift keeps the order of the rest).
 * Inserts are applied before deletes, so an edge inserted and deleted in the
 * same snapshot is absent afterwards. Output arrays are malloc'd. */
static int cmp_u64(const void *a, const void *b) {
    uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
    return x < y ? -1 : x > y;
}

int or_apply_updates(uint32_t nv, const uint64_t *row, const uint32_t *col,
                     const uint32_t *ins_uv, uint64_t nins, const uint32_t *del_uv, uint64_t ndel,
                     uint64_t **row_out, uint32_t **col_out) {
    uint64_t *irow = (uint64_t *)calloc((size_t)nv + 1, sizeof(uint64_t));
    uint32_t *icol = (uint32_t *)malloc((size_t)(nins ? nins : 1) * sizeof(uint32_t));
    uint64_t *dkeys = (uint64_t *)malloc((size_t)(ndel ? ndel : 1) * sizeof(uint64_t));
    uint64_t *nrow = (uint64_t *)calloc((size_t)nv + 1, sizeof(uint64_t));
    uint32_t *ncol;
    uint64_t *pos, total = 0;
    if (!irow || !icol || !dkeys || !nrow) goto oom;
    for (uint64_t i = 0; i < nins; ++i) {
        if (ins_uv[2 * i] >= nv || ins_uv[2 * i + 1] >= nv) goto bad;
        irow[ins_uv[2 * i] + 1]++;
    }
    for (uint32_t u = 0; u < nv; ++u) irow[u + 1] += irow[u];
    pos = (uint64_t *)malloc((size_t)(nv ? nv : 1) * sizeof(uint64_t));
    if (!pos) goto oom;
    memcpy(pos, irow, (size_t)nv * sizeof(uint64_t));
    for (uint64_t i = 0; i < nins; ++i) icol[pos[ins_uv[2 * i]]++] = ins_uv[2 * i + 1]; /* stable */
    free(pos);
    for (uint64_t i = 0; i < ndel; ++i) {
        if (del_uv[2 * i] >= nv || del_uv[2 * i + 1] >= nv) goto bad;
        dkeys[i] = ((uint64_t)del_uv[2 * i] << 32) | del_uv[2 * i + 1];
    }
    qsort(dkeys, (size_t)ndel, sizeof(uint64_t), cmp_u64);
    ncol = (uint32_t *)malloc((size_t)((row[nv] + nins) ? (row[nv] + nins) : 1) * sizeof(uint32_t));
    if (!ncol) goto oom;
    for (uint32_t u = 0; u < nv; ++u) {
        /* deleted targets of u: dkeys[lo, hi) */
        uint64_t lo = 0, hi = ndel, k0 = (uint64_t)u << 32;
        while (lo < hi) { uint64_t m = (lo + hi) / 2; if (dkeys[m] < k0) lo = m + 1; else hi = m; }
        hi = lo;
        while (hi < ndel && (dkeys[hi] >> 32) == u) ++hi;
        nrow[u] = total;
        for (int pass = 0; pass < 2; ++pass) {
            uint64_t b = pass ? irow[u] : row[u], e = pass ? irow[u + 1] : row[u + 1];
            const uint32_t *src = pass ? icol : col;
            for (uint64_t j = b; j < e; ++j) {
                uint64_t key = k0 | src[j], l2 = lo, h2 = hi;
                while (l2 < h2) { uint64_t m = (l2 + h2) / 2; if (dkeys[m] < key) l2 = m + 1; else h2 = m; }
                if (l2 < hi && dkeys[l2] == key) continue; /* deleted */
                ncol[total++] = src[j];
            }
        }
    }
    nrow[nv] = total;
    free(irow); free(icol); free(dkeys);
    *row_out = nrow; *col_out = ncol;
    return 0;
bad:
    free(irow); free(icol); free(dkeys); free(nrow);
    return 1;
oom:
    free(irow); free(icol); free(dkeys); free(nrow);
    return 2;
}

void or_free(void *p) { free(p); }

/* ---- PMA leaf geometry: csr_graph.cuh:364-382 compute_capacity -----------
 * segment_size = ceil_log2(nv); segment_count = hyperfloor(ceil_div(nv, size));
 * segment_size = ceil_div(nv, segment_count). (groute/graphs/util.h:28-66) */
static uint64_t last_bit_set(uint64_t x) { return x ? 64u - (uint64_t)__builtin_clzll(x) : 0; }

void or_pma_geometry(uint32_t nv, uint32_t *seg_size, uint64_t *seg_count) {
    uint64_t s, c;
    if (nv < 2) { *seg_size = 1; *seg_count = 1; return; }
    s = last_bit_set((uint64_t)nv - 1);                  /* ceil_log2 */
    c = (nv + s - 1) / s;                                /* ceil_div */
    c = 1ull << (last_bit_set(c) - 1);                   /* hyperfloor */
    s = (nv + c - 1) / c;
    *seg_size = (uint32_t)s; *seg_count = c;
}

/* ---- vertex SEGMENT split: framework.cuh:310-362 LoadGraph ----------------
 * Walk vertices accumulating their slot span off[v+1]-off[v] (gaps included)
 * until it reaches ceil(capacity / nseg); record each segment's first vertex.
 * seg_snode[nseg + 1] receives the boundaries (last = nv); returns the number
 * of segments produced (may be fewer than nseg). */
uint32_t or_segment_partition(uint32_t nv, const uint64_t *off, uint32_t nseg, uint32_t *seg_snode) {
    uint64_t cap = off[nv];
    uint64_t per = (cap + nseg - 1) / (nseg ? nseg : 1);
    uint32_t node = 0, s = 0;
    if (nseg == 0) return 0;
    while (node < nv && s < nseg) {
        uint64_t acc = 0;
        seg_snode[s++] = node;
        while (acc < per && node < nv) { acc += off[node + 1] - off[node]; ++node; }
    }
    seg_snode[s] = nv;
    return s;
}

/* ---- certificates ------------------------------------------------------ */
/* Shortest-path certificate: d = dist[] and parent[] of a single-source run
 * (UINT32_MAX = unreached / no parent) with weights w(mode, seed) >= 1. */
uint64_t or_sp_certificate(uint32_t nv, const uint64_t *row, const uint32_t *col, uint32_t src, uint32_t wmode,
                           uint64_t seed, const uint32_t *dist, const uint32_t *parent) {
    uint64_t bad = 0;
    uint8_t *found;
    if (src >= nv) return 1;
    if (dist[src] != 0) ++bad;
    found = (uint8_t *)calloc((size_t)(nv ? nv : 1), 1);
    if (!found) return ~0ull;
    /* found[v] is written only by the thread that scans v's parent's list (the
     * one source u with parent[v] == u), so the parallel scan has no races */
#pragma omp parallel for schedule(dynamic, 4096) reduction(+ : bad)
    for (int64_t ui = 0; ui < (int64_t)nv; ++ui) {
        const uint32_t u = (uint32_t)ui;
        if (dist[u] == OR_INF) continue;
        for (uint64_t e = row[u]; e < row[u + 1]; ++e) {
            const uint32_t v = col[e];
            const uint64_t nd = (uint64_t)dist[u] + or_weight(wmode, seed, u, v);
            if ((uint64_t)dist[v] > nd) ++bad;                   /* an edge still relaxes */
            if (parent[v] == u && (uint64_t)dist[v] == nd) found[v] = 1;  /* a tight parent edge */
        }
    }
    for (uint32_t v = 0; v < nv; ++v)
        if (v != src && dist[v] != OR_INF && !found[v]) ++bad;  /* reached without a tight parent edge */
    free(found);
    return bad;
}

/* Min-ancestor label certificate: label[] and parent[] (UINT32_MAX = none). */
uint64_t or_cc_certificate(uint32_t nv, const uint64_t *row, const uint32_t *col, const uint32_t *label,
                           const uint32_t *parent) {
    uint64_t bad = 0;
    uint8_t *ok, *state;
    uint32_t *stack;
    for (uint32_t v = 0; v < nv; ++v)
        if (label[v] > v) ++bad;
    ok = (uint8_t *)calloc((size_t)(nv ? nv : 1), 1);     /* parent edge (parent[v], v) exists */
    state = (uint8_t *)calloc((size_t)(nv ? nv : 1), 1);  /* chain check: 0 new, 1 on stack, 2 good, 3 bad */
    stack = (uint32_t *)malloc((size_t)(nv ? nv : 1) * sizeof(uint32_t));
    if (!ok || !state || !stack) { free(ok); free(state); free(stack); return ~0ull; }
#pragma omp parallel for schedule(dynamic, 4096) reduction(+ : bad)
    for (int64_t ui = 0; ui < (int64_t)nv; ++ui) {  /* ok[v]: written by the scan of parent[v]'s list only */
        const uint32_t u = (uint32_t)ui;
        for (uint64_t e = row[u]; e < row[u + 1]; ++e) {
            const uint32_t v = col[e];
            if (label[v] > label[u]) ++bad;  /* the label did not reach v along this edge */
            if (parent[v] == u) ok[v] = 1;
        }
    }
    /* every chain v -> parent[v] -> ... stays on one label and ends at the
     * vertex label[v] (its own label, no parent); chains are walked once each */
    for (uint32_t v = 0; v < nv; ++v) {
        uint64_t n = 0;
        uint32_t x = v;
        while (state[x] == 0) {
            state[x] = 1;
            stack[n++] = x;
            if (parent[x] == OR_INF) break;
            if (parent[x] >= nv || !ok[x] || label[parent[x]] != label[x]) { state[x] = 3; break; }
            x = parent[x];
        }
        uint8_t r;
        if (state[x] == 1 && parent[x] == OR_INF) r = (label[x] == x) ? 2 : 3;  /* a root */
        else if (state[x] == 1) r = 3;                                          /* a cycle */
        else r = state[x];
        while (n) state[stack[--n]] = r;
        if (r == 3) ++bad;
    }
    free(ok);
    free(state);
    free(stack);
    return bad;
}

/* FP64 residual res_v = (1 - a) + a * sum_{u -> v} r_u / deg(u) - r_v of a
 * rank vector (float32 ranks as the engine keeps them). */
void or_pr_residual(uint32_t nv, const uint64_t *row, const uint32_t *col, double alpha, const float *rank,
                    double *l1_out, double *max_out, double *sum_rank_out) {
    double *acc = (double *)calloc((size_t)(nv ? nv : 1), sizeof(double));
    double l1 = 0, mx = 0, sr = 0;
    if (!acc) { *l1_out = *max_out = *sum_rank_out = -1; return; }
    for (uint32_t u = 0; u < nv; ++u) {
        const uint64_t d = row[u + 1] - row[u];
        if (!d) continue;
        const double c = alpha * (double)rank[u] / (double)d;
        for (uint64_t e = row[u]; e < row[u + 1]; ++e) acc[col[e]] += c;
    }
    for (uint32_t v = 0; v < nv; ++v) {
        const double r = (1.0 - alpha) + acc[v] - (double)rank[v];
        l1 += fabs(r);
        if (fabs(r) > mx) mx = fabs(r);
        sr += (double)rank[v];
    }
    free(acc);
    *l1_out = l1;
    *max_out = mx;
    *sum_rank_out = sr;
}
