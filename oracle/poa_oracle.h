/*
 * TEST INFRASTRUCTURE ONLY. A plain-C restatement of the reference's
 * sequence-to-DAG POA path (minipoa-rs: src/align.rs, src/graph.rs,
 * src/lib.rs). Only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline leg may load it, and only as the checker. The product library
 * never links or calls it.
 *
 * Every function names the reference file:line it restates.
 */
#ifndef POA_ORACLE_H
#define POA_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* src/lib.rs:47-71 Params (every field the path and validate() read). */
typedef struct {
    int32_t match_score, mismatch, gap_open, gap_ext;
    uint64_t b, f;
    int32_t adaptive, seeding;
    uint64_t k, w;
    int32_t progressive;
    uint64_t threads, win_band, win_max;
    double mm_filter_ratio;
    uint64_t max_occ, tr_dist;
    uint8_t result, verbose;
    int32_t is_rna;
    /* Cargo feature `scalar-dp` (Cargo.toml:19-25): force the i64 kernel. */
    int32_t scalar_dp;
    uint32_t alphabet_len;
    const uint8_t *alphabet;
    const int32_t *matrix; /* alphabet_len * alphabet_len, row-major a*n+b */
} op_params;

/* src/lib.rs:140-145 AlignedBase; -1 encodes None. */
typedef struct {
    int64_t node;
    int64_t base;
    int64_t aligned;
} ob_base;

typedef struct og_graph og_graph;

const char *o_last_error(void);

/* src/graph.rs:30-178 */
og_graph *og_new(const op_params *p);
/* Test infrastructure: a graph copied from flat arrays (node-indexed fields,
 * adjacency in list order with weights, rank order, paths as seq id + node
 * ids); coverage = sum of in-weights. */
og_graph *og_from_arrays(const op_params *p, size_t n_nodes, const uint64_t *rank,
                         const uint64_t *parent, const uint64_t *base, const uint64_t *in_off,
                         const uint64_t *in_node, const uint64_t *in_w, const uint64_t *out_off,
                         const uint64_t *out_node, const uint64_t *out_w, const int64_t *aligned,
                         size_t n_order, const uint64_t *order, size_t n_paths,
                         const uint64_t *path_seq, const uint64_t *path_off,
                         const uint64_t *path_ids);
void og_free(og_graph *g);
size_t og_add_node(og_graph *g, size_t base);
void og_add_edge(og_graph *g, size_t from, size_t to);
int og_add_path(og_graph *g, size_t seq_id, const ob_base *aln, size_t len);
int og_toposort(og_graph *g, int grouped);
size_t og_n_nodes(const og_graph *g);
/* Node fields: rank, parent, base, coverage, n_in, n_out. */
void og_node_info(const og_graph *g, size_t id, uint64_t out6[6]);
int64_t og_node_aligned(const og_graph *g, size_t id, size_t base);
void og_node_edges(const og_graph *g, size_t id, uint64_t *in_ids, uint64_t *in_w,
                   uint64_t *out_ids, uint64_t *out_w);
size_t og_rank_len(const og_graph *g);
void og_rank(const og_graph *g, uint64_t *out);
size_t og_n_paths(const og_graph *g);
/* src/graph.rs:181-331; strings are malloc'd, release with o_free. */
size_t og_consensus(const og_graph *g, const op_params *p, uint8_t *seq, uint64_t *ids,
                    uint64_t *cov, size_t cap);
int og_msa(og_graph *g, const op_params *p, const char *const *names,
           const char *const *comments, size_t nseq, char **out);
int og_gfa(const og_graph *g, const op_params *p, const char *const *names, size_t nseq,
           char **out);
void o_free(void *ptr);

/* src/lib.rs:109-138 */
size_t o_encode(const op_params *p, uint8_t base);
int o_validate(const op_params *p);

/* src/align.rs:459-491: out = I[j+1..j+8] given md[0..8] and carry I[j]. */
void o_insertion_prefix(const int32_t md[8], int32_t carry, int32_t open, int32_t ext,
                        int32_t out[8]);
/* src/align.rs:493-529 */
int o_i32_scores_safe(const op_params *p, const og_graph *g, size_t beg_id, size_t end_id,
                      size_t query_len);

/* src/minimizer.rs:20-26 Anchor (adaptive_before 0/1). */
typedef struct {
    uint64_t target_end, query_end, span;
    int64_t adaptive_before;
} o_anchor;

/*
 * Minimizer seeding, src/minimizer.rs:32-488. Each returns the full count and
 * writes up to `cap` entries. Minimizer positions are parity unpinned (the
 * reference ranks k-mers inside the simd-minimizers crate).
 */
/* src/minimizer.rs:32-76: value = exact 2-bit k-mer, end = last base. */
size_t o_sketch(const uint8_t *seq, size_t len, size_t k, size_t w, uint64_t *value,
                uint64_t *end, size_t cap);
/* src/minimizer.rs:78-110 */
size_t o_repeat_intervals(const uint64_t *value, const uint64_t *end, const uint64_t *span,
                          size_t n, size_t len, size_t max_dist, uint64_t *out_start,
                          uint64_t *out_end, size_t cap);
/* src/minimizer.rs:128-176 */
size_t o_collect_anchors(const op_params *p, const uint8_t *q, size_t qlen, const uint8_t *t,
                         size_t tlen, o_anchor *out, size_t cap);
/* src/minimizer.rs:199-377 */
size_t o_chain(const o_anchor *anchors, size_t n, size_t k, o_anchor *out, size_t cap);
/* src/minimizer.rs:379-467 */
size_t o_partition_windows(const o_anchor *chain, size_t n, size_t tlen, size_t qlen,
                           size_t max_drift, size_t max_win, o_anchor *out, size_t cap,
                           int *tail);
/* src/minimizer.rs:489-578: order[0..n) = the progressive guide order
 * (input order unless p->progressive, n >= 2 and a DNA alphabet). */
void o_guide_order(const op_params *p, const uint8_t *const *seqs, const size_t *lens, size_t n,
                   size_t *order);
/* src/minimizer.rs:475-488 */
size_t o_anchor_windows(const op_params *p, const uint8_t *q, size_t qlen, const uint8_t *t,
                        size_t tlen, o_anchor *out, size_t cap, int *tail);

/*
 * All alignment entry points write up to `cap` AlignedBase records to `out`,
 * set *out_len (the full length, even if it exceeds cap) and return 0, or
 * return 1 with the reference's error message in o_last_error().
 */
/* src/align.rs:81-188 (empty-graph, unseeded and seeded branches). Seeded
 * windows run on p->threads threads when it is above 1, results in window
 * order (src/parallel.rs:46-67). */
int o_align(const op_params *p, const og_graph *g, const uint8_t *q, size_t qlen, ob_base *out,
            size_t cap, size_t *out_len);
/* src/align.rs:192-211 */
int o_poa(const op_params *p, const og_graph *g, size_t beg_id, size_t end_id,
          const uint8_t *q, size_t qlen, int adaptive, size_t band_b, ob_base *out, size_t cap,
          size_t *out_len);
/* src/align.rs:531-545 */
int o_poa_inner(const op_params *p, const og_graph *g, size_t beg_id, size_t end_id,
                const uint8_t *q, size_t qlen, int adaptive, size_t band_b, ob_base *out,
                size_t cap, size_t *out_len);
/* src/align.rs:547-567 + 587-911 (8-lane i32 kernel). */
int o_poa_simd_i32(const op_params *p, const og_graph *g, size_t beg_id, size_t end_id,
                   const uint8_t *q, size_t qlen, int adaptive, size_t band_b, ob_base *out,
                   size_t cap, size_t *out_len);
/* src/align.rs:913-1162 (i64 scalar kernel). */
int o_poa_scalar_i64(const op_params *p, const og_graph *g, size_t beg_id, size_t end_id,
                     const uint8_t *q, size_t qlen, int adaptive, size_t band_b, ob_base *out,
                     size_t cap, size_t *out_len);

/* Cells (sum of computed row lengths) of DP calls on this thread since reset. */
uint64_t o_cells(void);
/* M at (n-1, m-1) of this thread's last DP fill (the global optimum). */
int64_t o_last_score(void);
void o_reset_cells(void);
/* Capacity of the calling thread's i32 arena (src/align.rs:222-226). */
size_t o_arena_capacity(void);
void o_arena_clear(void);

/* src/lib.rs:177-204 align_sequences (input order, or the guide order with -p). */
og_graph *o_align_sequences(const op_params *p, const uint8_t *const *seqs,
                            const size_t *lens, size_t nseq);
/* Continue progressive alignment of seqs[existing..] into g (src/lib.rs:184-204). */
int o_extend_graph(const op_params *p, og_graph *g, const uint8_t *const *seqs,
                   const size_t *lens, size_t nseq, size_t existing);

/*
 * Job-parallel CPU driver for the throughput baseline: job i aligns reads
 * read_off[job_off[i]..job_off[i+1]] progressively and renders consensus and
 * MSA (reads named read<k>, empty comments). out_hash[i] is the FNV-1a 64 of
 * the consensus FASTA followed by the MSA text, or 0 on error. Returns cells.
 */
uint64_t o_batch_run(const op_params *p, size_t njobs, const uint64_t *job_off,
                     const uint64_t *read_off, const uint8_t *bases, int nthreads,
                     uint64_t *out_hash);
/* The same, plus out_gfa[i]: the FNV-1a 64 of job i's GFA text (-r2), or 0. */
uint64_t o_batch_run_gfa(const op_params *p, size_t njobs, const uint64_t *job_off,
                         const uint64_t *read_off, const uint8_t *bases, int nthreads,
                         uint64_t *out_hash, uint64_t *out_gfa);

#ifdef __cplusplus
}
#endif
#endif
