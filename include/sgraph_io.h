/*
 * sgraph_io.h — host-side input formats and synthetic inputs of libsgraph.so.
 *
 * These calls need no GPU. Arrays they return are allocated with malloc and
 * released with sg_free.
 *
 * Formats (the reference's, Appendix-B of its README and loaders):
 *   graph file   "u v [w]" per line, whitespace separated, no header; nv = max id + 1;
 *                edges grouped by source in file order; duplicates and self-loops kept
 *                (src/utils/parser.cpp:296-406, ReadGraphMarket_bigdata).
 *   update file  "a u v [w]" inserts, "d u v [w]" deletes; other tags ignored; all "a"
 *                lines in file order form the insert sequence, all "d" lines the delete
 *                sequence (include/framework/Loader.h:64-100, ReadWeightList).
 *   stream sizes "n_add n_del" per line, trailing text ignored; line i sizes snapshot i
 *                (include/framework/Loader.h:27-62).
 */
#ifndef SGRAPH_IO_H
#define SGRAPH_IO_H

#include "sgraph.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Release an array returned by this header's functions. NULL is a no-op. */
void sg_free(void *p);

/* Read an edge-list graph file into CSR (row[nv+1], col[ne]).
 * Replaces ReadGraphMarket_bigdata (src/utils/parser.cpp:296-406); unlike it,
 * nv is sized by max(src, dst) + 1 so a max id seen only as a destination is kept. */
sg_status sg_read_edge_list(const char *path, uint32_t *nv, uint64_t *ne,
                            uint64_t **row, uint32_t **col);

/* Write CSR as "u v" lines in adjacency order. */
sg_status sg_write_edge_list(const char *path, uint32_t nv,
                             const uint64_t *row, const uint32_t *col);

/* Read the stream-size file: *nsnap lines -> sizes[2*i] = n_add, sizes[2*i+1] = n_del.
 * Replaces Loader::Loader(size_file, weight) (include/framework/Loader.h:27-62). */
sg_status sg_read_stream_sizes(const char *path, uint32_t *nsnap, uint64_t **sizes);

/* Read the update file into the insert and delete sequences as (u, v) pairs.
 * Replaces Loader::ReadWeightList / ReadList (include/framework/Loader.h:64-131). */
sg_status sg_read_updates(const char *path, uint64_t *nins, uint32_t **ins_uv,
                          uint64_t *ndel, uint32_t **del_uv);

/* Write an update file ("a u v" then "d u v" lines per snapshot) and its
 * stream-size file. sizes as returned by sg_read_stream_sizes. */
sg_status sg_write_updates(const char *update_path, const char *size_path,
                           uint32_t nsnap, const uint64_t *sizes,
                           const uint32_t *ins_uv, const uint32_t *del_uv);

/* Seeded R-MAT graph: 2^scale vertices, edge_factor * 2^scale draws with
 * quadrant probabilities a/b/c (d = 1-a-b-c), 32 random bits per level from a
 * splitmix64 stream keyed by (seed, draw index), so the output does not depend
 * on the thread count. Self-loops and duplicates are dropped; the CSR is sorted
 * by (src, dst). With permute != 0 vertex ids are shuffled by a seeded
 * Fisher-Yates permutation (generator.h:53-63 does the same with mt19937).
 * Constants default in the reference: a/b/c = 0.57/0.19/0.19 (generator.h:83). */
sg_status sg_gen_rmat(uint32_t scale, uint32_t edge_factor, uint64_t seed,
                      double a, double b, double c, int permute,
                      uint32_t *nv, uint64_t *ne, uint64_t **row, uint32_t **col);

/* Seeded update stream over G_0 = (row, col): nsnap snapshots of n_ins inserts
 * and n_del deletes each. Inserts are fresh R-MAT draws (same scale and a/b/c)
 * absent from G_{t-1}, self-loops excluded, no pair repeated within a batch; a
 * fraction hub_frac of them takes its source uniformly from the top hub_top
 * fraction of vertices by G_0 out-degree. Deletes are sampled uniformly from the
 * edges of G_{t-1}, without repetition. Snapshot t applies its inserts, then its
 * deletes. Output: ins_uv[2 * nsnap * n_ins], del_uv[2 * nsnap * n_del]. */
sg_status sg_gen_updates(uint32_t nv, const uint64_t *row, const uint32_t *col,
                         uint32_t scale, double a, double b, double c,
                         uint32_t nsnap, uint64_t n_ins, uint64_t n_del,
                         double hub_frac, double hub_top, uint64_t seed,
                         uint32_t **ins_uv, uint32_t **del_uv);

/* Binary snapshot file: G_0 as CSR, the stream sizes and the insert / delete
 * sequences in one file, read without parsing. No counterpart in the
 * reference, which re-parses its text files on every run (the graph twice,
 * src/utils/parser.cpp:324-401). Layout, little-endian: "SGSNAP01", u32 nv,
 * u32 nsnap, u64 ne, u64 total inserts, u64 total deletes, u64 row[nv + 1],
 * u32 col[ne], u64 sizes[2 * nsnap] (as sg_read_stream_sizes), u32 pairs of
 * the inserts, then of the deletes. Ids are checked by sg_load_csr and the
 * apply calls, as for the text formats; the reader checks the header against
 * the file length, the offsets and the size totals (SG_EIO otherwise). */
sg_status sg_write_snapshot_file(const char *path, uint32_t nv, uint64_t ne, const uint64_t *row,
                                 const uint32_t *col, uint32_t nsnap, const uint64_t *sizes,
                                 const uint32_t *ins_uv, const uint32_t *del_uv);
sg_status sg_read_snapshot_file(const char *path, uint32_t *nv, uint64_t *ne, uint64_t **row,
                                uint32_t **col, uint32_t *nsnap, uint64_t **sizes,
                                uint32_t **ins_uv, uint32_t **del_uv);

/* SSSP edge weight of the engine for (u, v) (sg_weight_mode). */
uint32_t sg_edge_weight(uint32_t mode, uint64_t seed, uint32_t u, uint32_t v);

#ifdef __cplusplus
}
#endif
#endif /* SGRAPH_IO_H */
