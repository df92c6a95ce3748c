/*
 * minipoa_hip -- MI355X (gfx950) backend for minipoa-rs's affine-gap
 * sequence-to-DAG alignment DP, as a C ABI.
 *
 * Plain pointers, sizes and status codes only: no C++ types, no exceptions
 * and no GPU types cross this boundary, so a Rust `-sys` crate can bind it
 * (minipoa itself is `#![forbid(unsafe_code)]`, src/lib.rs:2). Every
 * function returns MPOA_OK (0) or a status; the message of the calling
 * thread's last failure is mpoa_last_error(). A host allocation failure
 * inside the library returns MPOA_ERR_ALLOC (in a batch or server job: that
 * job's status).
 *
 * usize fields of the reference are uint64_t here; Option<usize> is int64_t
 * with -1 for None.
 */
#ifndef MINIPOA_HIP_H
#define MINIPOA_HIP_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. Each DP error has the reference's message (src/align.rs),
 * so `poa`'s banded -> full-DP retry fires on exactly the same cases. */
typedef enum {
    MPOA_OK = 0,
    MPOA_ERR_ANCHOR_NOT_IN_GRAPH = 1,     /* align.rs:597-598 */
    MPOA_ERR_ANCHORS_NOT_ORDERED = 2,     /* align.rs:602-603 */
    MPOA_ERR_COORDINATE_LIMITS = 3,       /* align.rs:607-608 */
    MPOA_ERR_NO_PATH = 4,                 /* align.rs:679-680 */
    MPOA_ERR_BAND_NO_INTERSECT = 5,       /* align.rs:730-731 */
    MPOA_ERR_BAND_EXCLUDES_SOURCE = 6,    /* align.rs:738-739 */
    MPOA_ERR_TRACEBACK_LEFT_BAND = 7,     /* align.rs:1190-1191 */
    MPOA_ERR_TRACEBACK_NO_ADVANCE = 8,    /* align.rs:1184-1185 */
    MPOA_ERR_TRACEBACK_NO_PREDECESSOR = 9,/* align.rs:1304-1306 */
    MPOA_ERR_ALLOC = 10,                  /* host or device allocation */
    MPOA_ERR_INVALID_PARAMS = 11,         /* Params::validate, lib.rs:149-173 */
    MPOA_ERR_GRAPH = 12,                  /* add_path / toposort / msa / gfa errors */
    MPOA_ERR_DEVICE = 13,                 /* HIP runtime error */
    MPOA_ERR_NO_DEVICE = 14,              /* no gfx950 device / library not loadable */
    MPOA_ERR_UNSUPPORTED = 15,            /* a reference feature outside this backend */
    MPOA_ERR_BUFFER_TOO_SMALL = 16,       /* *out_len holds the size needed */
    MPOA_ERR_INVALID_ARGUMENT = 17,
    MPOA_ERR_PENDING = 18,                /* a served job has not finished yet */
    MPOA_ERR_SEED_SPAN = 19,              /* align.rs:116-130 invalid minimizer span */
    MPOA_ERR_SCRATCH_LIMIT = 20           /* a DP's buffers exceed mpoa_poa_noalloc's scratch */
} mpoa_status;

/* src/lib.rs:47-71 Params. `alphabet` and `matrix` are borrowed for the
 * duration of each call; matrix is alphabet_len^2, row-major (a*n + b). */
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
    const int32_t *matrix;
} mpoa_params;

/* src/lib.rs:140-145 AlignedBase. */
typedef struct {
    int64_t node;    /* -1 = None */
    int64_t base;    /* alphabet index */
    int64_t aligned; /* -1 = None */
} mpoa_aligned_base;

/* Read-only flat view of a graph::Graph (src/graph.rs:6-27), node-indexed.
 * Adjacency lists keep the reference's insertion order: it decides every
 * tie in the band heuristics and in traceback. */
typedef struct {
    uint64_t n_nodes;
    uint64_t alphabet_len;
    const uint64_t *rank;       /* nodes[i].rank */
    const uint64_t *parent;     /* nodes[i].parent */
    const uint64_t *base;       /* nodes[i].base */
    const uint64_t *in_off;     /* n_nodes + 1 offsets into in_node/in_weight */
    const uint64_t *in_node;    /* nodes[i].incoming */
    const uint64_t *in_weight;  /* nodes[i].in_weights */
    const uint64_t *out_off;    /* n_nodes + 1 offsets into out_node/out_weight */
    const uint64_t *out_node;   /* nodes[i].outgoing */
    const uint64_t *out_weight; /* nodes[i].out_weights */
    const int64_t *aligned;     /* n_nodes * alphabet_len; nodes[i].aligned[b], -1 = None */
    uint64_t n_order;           /* graph.rank.len() */
    const uint64_t *order;      /* graph.rank: rank -> node id */
} mpoa_graph_view;

typedef struct mpoa_ctx mpoa_ctx;
typedef struct mpoa_graph mpoa_graph;

/* ---- backend selection (src/lib.rs:15-37 dp_backend) ---- */

/* "hip-gfx950 (wave64 x i32)"; printed where the reference prints its
 * `DP backend:` line (src/cli.rs:112-121). */
const char *mpoa_backend_name(void);
/* The calling thread's last failure message. */
const char *mpoa_last_error(void);

/* A context owns one device's stream, device arenas and staging buffers. It
 * plays the part of the reference's per-worker DP_ARENA (src/align.rs:222-226):
 * use one context per host thread. */
int mpoa_ctx_create(int device, mpoa_ctx **out);
void mpoa_ctx_destroy(mpoa_ctx *ctx);
/* No counterpart in the reference (whose DP matrix lives in host memory):
 * how a context's DPs lay out their traceback data in device memory.
 * VALUES keeps M, D and I of every cell (12 bytes per i32 cell; the fill
 * does the least work per cell, the fastest on the batched workloads
 * measured); RECORDS keeps a 1-8 byte traceback record per cell with the
 * `choose` winners already resolved, plus M and D of rows read beyond the
 * next (about 1.5-2 bytes per cell, at half again the fill's integer work).
 * AUTO (the default) takes VALUES, and RECORDS in mpoa_poa_noalloc, whose
 * fixed scratch then holds several times larger DPs. Results are identical
 * in every layout. The environment variable MPOA_LAYOUT (values, records)
 * sets the default of contexts created afterwards. Fails while a batch or a
 * server holds the context. */
enum {
    MPOA_LAYOUT_AUTO = 0,
    MPOA_LAYOUT_VALUES = 1,
    MPOA_LAYOUT_RECORDS = 2
};
int mpoa_ctx_set_layout(mpoa_ctx *ctx, int layout);
/* Bytes of device memory the context currently holds. */
uint64_t mpoa_ctx_device_bytes(const mpoa_ctx *ctx);

/* ---- drop-in DP entry points (src/align.rs) ---- */

/* src/align.rs:192-211 `poa`: banded attempt, then an exact full-DP retry
 * with f = 0 when the banded attempt fails and f != 0. Writes up to out_cap
 * records; *out_len is the full length (MPOA_ERR_BUFFER_TOO_SMALL if it
 * exceeds out_cap). *out_score (may be NULL) receives M at (n-1, m-1). */
int mpoa_poa(mpoa_ctx *ctx, const mpoa_params *params, const mpoa_graph_view *graph,
             uint64_t beg_id, uint64_t end_id, const uint8_t *query, uint64_t qlen,
             int adaptive, uint64_t band_b, mpoa_aligned_base *out, uint64_t out_cap,
             uint64_t *out_len, int64_t *out_score);
/* No counterpart in the reference: mpoa_poa for callers that cannot allocate
 * device memory (code around HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). Same arguments, results and errors as
 * mpoa_poa, but every device buffer of its launches (inputs, results, row
 * records, score arena) is carved from a fixed 64 MiB array that each GPU
 * the library runs on holds from its first context until the process ends,
 * and host staging uses pageable memory: the call makes no hipMalloc,
 * hipFree or hipHostMalloc. A DP whose buffers exceed 64 MiB returns
 * MPOA_ERR_SCRATCH_LIMIT (for the banded attempt or for poa's full-DP
 * retry). The array keeps the last call's contents; calls on one GPU share
 * it, so the library runs them one at a time (a per-device lock), and two
 * processes each have their own copy. */
int mpoa_poa_noalloc(mpoa_ctx *ctx, const mpoa_params *params, const mpoa_graph_view *graph,
                     uint64_t beg_id, uint64_t end_id, const uint8_t *query, uint64_t qlen,
                     int adaptive, uint64_t band_b, mpoa_aligned_base *out, uint64_t out_cap,
                     uint64_t *out_len, int64_t *out_score);
/* src/align.rs:531-545 `poa_inner`: one attempt, no retry; the i32 kernel
 * under the score guard (src/align.rs:493-529), else the i64 kernel. */
int mpoa_poa_inner(mpoa_ctx *ctx, const mpoa_params *params, const mpoa_graph_view *graph,
                   uint64_t beg_id, uint64_t end_id, const uint8_t *query, uint64_t qlen,
                   int adaptive, uint64_t band_b, mpoa_aligned_base *out, uint64_t out_cap,
                   uint64_t *out_len, int64_t *out_score);
/* Force one kernel: score_bits 32 or 64 (the i32 fast path and the i64 path
 * of src/align.rs:587-911 and 913-1162). The i32 kernel refuses inputs that
 * fail the score guard (MPOA_ERR_INVALID_ARGUMENT). */
int mpoa_poa_kernel(mpoa_ctx *ctx, const mpoa_params *params, const mpoa_graph_view *graph,
                    uint64_t beg_id, uint64_t end_id, const uint8_t *query, uint64_t qlen,
                    int adaptive, uint64_t band_b, int score_bits, mpoa_aligned_base *out,
                    uint64_t out_cap, uint64_t *out_len, int64_t *out_score);
/* src/align.rs:493-529 */
int mpoa_i32_scores_safe(const mpoa_params *params, const mpoa_graph_view *graph,
                         uint64_t beg_id, uint64_t end_id, uint64_t query_len);
/* src/align.rs:81-188 `align`: empty-graph shortcut, unseeded DP, or with
 * params.seeding the windows between chained minimizer anchors of the query
 * and the graph's consensus, all windows of the query in one launch (each
 * with poa's retry), concatenated with the anchors' own nodes. */
int mpoa_align(mpoa_ctx *ctx, const mpoa_params *params, const mpoa_graph_view *graph,
               const uint8_t *query, uint64_t qlen, mpoa_aligned_base *out, uint64_t out_cap,
               uint64_t *out_len);

/* ---- minimizer seeding (src/minimizer.rs), host side ----
 * Each writes up to cap entries and sets *out_len to the full count
 * (MPOA_ERR_BUFFER_TOO_SMALL when it exceeds cap). k-mers are ranked by the
 * upper 16 bits of a 32-bit forward ntHash, leftmost minimum per window of w
 * k-mers: the ordering of the simd-minimizers crate the reference uses. */

/* src/minimizer.rs:20-26 Anchor (adaptive_before 0 or 1). */
typedef struct {
    uint64_t target_end, query_end, span;
    int64_t adaptive_before;
} mpoa_anchor;

/* src/minimizer.rs:32-76 `sketch`: value = exact 2-bit k-mer (A 0, C 1,
 * T/U 2, G 3; first base lowest), end = inclusive position of its last base. */
int mpoa_sketch(const uint8_t *seq, uint64_t len, uint64_t k, uint64_t w, uint64_t *value,
                uint64_t *end, uint64_t cap, uint64_t *out_len);
/* src/minimizer.rs:128-176 `collect_anchors` (query vs target). */
int mpoa_collect_anchors(const mpoa_params *params, const uint8_t *query, uint64_t qlen,
                         const uint8_t *target, uint64_t tlen, mpoa_anchor *out, uint64_t cap,
                         uint64_t *out_len);
/* src/minimizer.rs:199-377 `chain` (anchors sorted by target, then query end). */
int mpoa_chain(const mpoa_anchor *anchors, uint64_t n, uint64_t k, mpoa_anchor *out, uint64_t cap,
               uint64_t *out_len);
/* src/minimizer.rs:379-467 `partition_windows`; *tail_adaptive: the last
 * window's adaptive flag. */
int mpoa_partition_windows(const mpoa_anchor *chain, uint64_t n, uint64_t tlen, uint64_t qlen,
                           uint64_t max_drift, uint64_t max_win, mpoa_anchor *out, uint64_t cap,
                           uint64_t *out_len, int *tail_adaptive);
/* src/minimizer.rs:489-578 `guide_order`: order[0..nseq) is the -p
 * progressive order (input order unless params.progressive, nseq >= 2 and a
 * DNA alphabet). mpoa_align_sequences, mpoa_extend_graph (existing == 0),
 * batch jobs and served jobs without a graph align in this order. */
int mpoa_guide_order(const mpoa_params *params, const uint8_t *const *seqs, const uint64_t *lens,
                     uint64_t nseq, uint64_t *order);
/* src/minimizer.rs:475-488 `anchor_windows`. */
int mpoa_anchor_windows(const mpoa_params *params, const uint8_t *query, uint64_t qlen,
                        const uint8_t *consensus, uint64_t clen, mpoa_anchor *out, uint64_t cap,
                        uint64_t *out_len, int *tail_adaptive);

/* ---- library-owned graph (src/graph.rs), for callers without one ---- */

mpoa_graph *mpoa_graph_new(const mpoa_params *params);                 /* graph.rs:30-40 */
void mpoa_graph_free(mpoa_graph *g);
/* graph.rs:42-61: the new node's id, or UINT64_MAX (message set) when base
 * is outside the alphabet (the reference panics on that index). */
uint64_t mpoa_graph_add_node(mpoa_graph *g, uint64_t base);
/* graph.rs:63-80: MPOA_ERR_INVALID_ARGUMENT, graph unchanged, when an id is
 * not a node of g (the reference panics on that index). */
int mpoa_graph_add_edge(mpoa_graph *g, uint64_t from, uint64_t to);
int mpoa_graph_add_path(mpoa_graph *g, uint64_t seq_id, const mpoa_aligned_base *alignment,
                        uint64_t len);                                 /* graph.rs:82-122 */
int mpoa_graph_toposort(mpoa_graph *g, int grouped);                   /* graph.rs:124-178 */
/* A view of the graph, valid until the graph is next modified or freed. */
int mpoa_graph_view_of(mpoa_graph *g, mpoa_graph_view *out);
uint64_t mpoa_graph_n_paths(const mpoa_graph *g);
/* graph.paths[k] (src/graph.rs:20-27): its sequence id and node ids (up to
 * cap; *out_len is the full length). */
int mpoa_graph_path(const mpoa_graph *g, uint64_t k, uint64_t *seq_id, uint64_t *ids, uint64_t cap,
                    uint64_t *out_len);
/* graph.rs:181-237: writes up to cap bases (decoded) / node ids / coverage
 * (any pointer may be NULL) and returns the consensus length. */
uint64_t mpoa_graph_consensus(const mpoa_graph *g, const mpoa_params *params, uint8_t *seq,
                              uint64_t *ids, uint64_t *coverage, uint64_t cap);
/* graph.rs:239-331: text is written to out (up to cap bytes, no terminator);
 * *out_len is the full size. names/comments are indexed by sequence id. */
int mpoa_graph_msa(mpoa_graph *g, const mpoa_params *params, const char *const *names,
                   const char *const *comments, uint64_t nseq, char *out, uint64_t cap,
                   uint64_t *out_len);
int mpoa_graph_gfa(const mpoa_graph *g, const mpoa_params *params, const char *const *names,
                   uint64_t nseq, char *out, uint64_t cap, uint64_t *out_len);

/* src/lib.rs:149-173 */
int mpoa_params_validate(const mpoa_params *params);
/* `io::load_matrix` for the -m option (src/lib.rs:209-211; src/io.rs is not
 * in the reference's tree, so the format follows tests/io_cli.rs:326-372 and
 * test-data/HOXD70.mtx): '#' lines and blank lines are skipped; the first
 * other line lists one single-character label per column; then one row per
 * label, in any order: the label and one integer per column. Writes the
 * alphabet in header order and the alphabet_len^2 matrix row-major in that
 * order (cap: capacity of `alphabet`; `matrix` holds cap * cap). Malformed
 * text returns MPOA_ERR_INVALID_PARAMS and writes nothing. */
int mpoa_load_matrix(const char *text, uint64_t len, uint8_t *alphabet, int32_t *matrix, uint32_t cap,
                     uint32_t *alphabet_len);
/* ---- input formats: the reference's io module (src/io.rs is not in its
 * tree; behaviour from tests/io_cli.rs and README.md) ----
 * Records (src/lib.rs:39-45 Sequence) live in an mpoa_records handle; the
 * pointers mpoa_record_get returns stay valid until mpoa_records_free. */
typedef struct mpoa_records mpoa_records;
/* io::read_sequences / io::parse_sequences (tests/io_cli.rs:34-87): FASTA or
 * FASTQ text (multi-line records, CRLF, header comments, empty records,
 * quality lines that start with '@' or '+'), gzip-compressed when the data
 * starts with the gzip magic (concatenated members). Malformed input returns
 * MPOA_ERR_INVALID_ARGUMENT with the reason; empty input gives no records. */
int mpoa_read_sequences(const uint8_t *data, uint64_t len, mpoa_records **out);
uint64_t mpoa_records_len(const mpoa_records *r);
int mpoa_record_get(const mpoa_records *r, uint64_t k, const char **name, uint64_t *name_len,
                    const char **comment, uint64_t *comment_len, const uint8_t **seq,
                    uint64_t *seq_len, const uint8_t **qual, uint64_t *qual_len);
void mpoa_records_free(mpoa_records *r);
/* io::read_gfa for -i (src/lib.rs:214-216; tests/io_cli.rs:374-419, 563-610):
 * a forward, acyclic GFA1 graph (H/S/L/P, explicit sequences, 0M overlaps)
 * as a graph plus one record per path (in file order; sequence id k is the
 * k-th path). Each segment becomes one node per base, chained; links add
 * edges; paths are replayed as reads, so edge weights count the paths that
 * use them (a link no path uses weighs 1); nodes without a predecessor or
 * successor are joined to the source or sink. Every node is its own
 * alignment column (GFA carries no column groups). Continue it with
 * mpoa_extend_graph(existing = number of records). Errors: MPOA_ERR_GRAPH. */
int mpoa_read_gfa(const mpoa_params *params, const uint8_t *data, uint64_t len, mpoa_graph **graph,
                  mpoa_records **records);

/* src/lib.rs:177-204 `align_sequences` (input order, or the guide order with
 * params.progressive) on the device. */
int mpoa_align_sequences(mpoa_ctx *ctx, const mpoa_params *params, const uint8_t *const *seqs,
                         const uint64_t *lens, uint64_t nseq, mpoa_graph **out);
/* src/lib.rs:184-204 `extend_graph` from sequence index `existing` (the -i
 * incremental mode continues a loaded graph this way). */
int mpoa_extend_graph(mpoa_ctx *ctx, const mpoa_params *params, mpoa_graph *g,
                      const uint8_t *const *seqs, const uint64_t *lens, uint64_t nseq,
                      uint64_t existing);

/* ---- batched progressive POA (additive; no reference counterpart) ----
 * Job i aligns reads read_off[job_off[i]] .. read_off[job_off[i+1]] of
 * `bases` progressively (src/lib.rs:184-204 per job); every DP of one
 * progressive step of all jobs runs in one device launch. Results keep job
 * order (the ordering contract of src/parallel.rs:46-67). */
typedef struct {
    uint64_t cells;        /* DP cells computed (sum of row band widths) */
    uint64_t dp_calls;     /* DP alignments run */
    uint64_t retries;      /* banded failures re-run as full DP */
    uint64_t i64_calls;    /* alignments routed to the i64 kernel */
    uint64_t pred_cells;   /* sum over rows and predecessors of band overlaps */
    uint64_t fill_launches;/* fill kernel launches */
    double device_ms;      /* fill + traceback kernel time (HIP events) */
    double fill_ms;        /* fill kernel time alone (HIP events, summed) */
    double host_ms;        /* host preparation + fusion time */
    double wall_ms;
    /* Parts of host_ms: planning, result consumption (path assembly),
     * fusion (add_path + toposort) and rendering, per host worker; and the
     * launch slots' input packing, per slot. */
    double plan_ms, consume_ms, fuse_ms, render_ms, pack_ms;
    /* DPs whose adaptive rows outgrew their arena and were re-run with
     * full-width rows (the same DP, so the same result). */
    uint64_t arena_retries;
    /* Device bytes of traceback data the fills wrote (completed DPs), and
     * the fill launches that used the records layout (mpoa_ctx_set_layout). */
    uint64_t arena_bytes;
    uint64_t record_launches;
    /* mpoa_batch_run_resident only: jobs that finished on the device, jobs
     * that left it and finished on the host engine, captured step-loop
     * replays, length buckets, DP placements postponed to a later step,
     * wall time of the device-resident part and its step loops' device
     * time (HIP events). Zero for mpoa_batch_run. */
    uint64_t resident_jobs, evicted_jobs, step_replays, buckets, deferred;
    double resident_ms, resident_device_ms;
} mpoa_batch_stats;

typedef struct mpoa_batch mpoa_batch;
/* Runs all jobs to completion. names: reads are named read<k> within a job. */
int mpoa_batch_run(mpoa_ctx *ctx, const mpoa_params *params, uint64_t njobs,
                   const uint64_t *job_off, const uint64_t *read_off, const uint8_t *bases,
                   mpoa_batch **out);
/* mpoa_batch_run with every job kept on the device between reads (no
 * reference counterpart; the same per-job semantics, src/lib.rs:184-204):
 * each job's graph, its next DP's descriptors and its traceback live in
 * device memory; one progressive step of a length bucket of jobs -- path
 * fusion (src/graph.rs:82-122), the LIFO re-rank (src/graph.rs:124-178)
 * and the next read's planning (src/align.rs:587-688), then fill and
 * traceback -- is one captured HIP graph, replayed until the bucket's jobs
 * finish; only final graphs come back. Jobs outside it (seeding, the -p
 * guide order, scores that need the i64 kernel, reads over 31,998 bases,
 * alphabets over 25 letters, no reads) and jobs whose DP outgrows the
 * device arena continue on the host-driven engine from the read where they
 * left. Outputs are byte-identical to mpoa_batch_run's. */
int mpoa_batch_run_resident(mpoa_ctx *ctx, const mpoa_params *params, uint64_t njobs,
                            const uint64_t *job_off, const uint64_t *read_off, const uint8_t *bases,
                            mpoa_batch **out);
void mpoa_batch_free(mpoa_batch *b);
int mpoa_batch_stats_get(const mpoa_batch *b, mpoa_batch_stats *out);
/* Per-job status (MPOA_OK or the job's first error). */
int mpoa_batch_job_status(const mpoa_batch *b, uint64_t job);
/* Per-job rendered output: the `-r0` consensus FASTA (mode 0), the `-r1`
 * MSA (mode 1) or the `-r2` GFA (mode 2), exactly as src/lib.rs:233-241
 * would print them. */
int mpoa_batch_job_output(const mpoa_batch *b, uint64_t job, int mode, char *out, uint64_t cap,
                          uint64_t *out_len);

/* Self-check of the device-resident step's code (no reference counterpart):
 * the step's graph fusion, re-rank and planning templates run on one host
 * thread beside this library's host graph and planner (src/graph.rs:82-178,
 * src/align.rs:587-688), applying read r's path paths[path_off[r] ..
 * path_off[r+1]) in turn, and compare nodes, adjacency order, ranks, paths
 * and every read's DP plan (banded and full). lds_bytes: the step's LDS
 * staging (0 = global arrays). MPOA_OK when equal; otherwise the first
 * difference is written to report (cap bytes, NUL-terminated). */
int mpoa_resident_check(const mpoa_params *params, const uint8_t *bases, const uint64_t *read_off,
                        uint64_t nreads, const mpoa_aligned_base *paths, const uint64_t *path_off,
                        uint32_t lds_bytes, char *report, uint64_t cap);

/* ---- serving (additive; no reference counterpart) ----
 * A server admits POA jobs as they arrive and runs them in the same
 * continuous engine as a batch: every job advances one progressive step per
 * DP (src/lib.rs:184-204), and the DPs of all admitted jobs that are ready
 * share launches. A job is admitted when the device footprint of the
 * admitted jobs (their next DP's score arena, or an estimate before the
 * first) leaves room for it; others wait in arrival order. A job's results
 * do not depend on what else is running: each DP is the same computation
 * alone or in a launch with others. While a server exists, its context's
 * launch slots are busy: mpoa_batch_run and a second server fail. */
typedef struct mpoa_server mpoa_server;
typedef struct {
    uint64_t admit_bytes; /* footprint bound for admission (0 = the launch slots' arena budget) */
    uint32_t slots;       /* launch slots, streams of their own (0 = 4: two take DPs of
                             6,000 rows or more first, two take only shorter DPs) */
    int32_t render;       /* nonzero: render the -r0 and -r1 text of each finished job */
    int32_t resident;     /* nonzero: serve on device-resident graphs. A job is admitted into a
                             slot of a step table of its size class (longest read up to 2,000,
                             8,000 or 31,998 bases) once the table's device heap holds its graph,
                             reads and DP descriptors; they stay on the device until it finishes,
                             and the table's captured step graph (fusion, re-rank and planning on
                             the device, then fill and traceback) advances every admitted job one
                             read per replay, new jobs joining between replays. Jobs outside the
                             resident path (seeding, -p, the i64 kernel) run host-driven. Outputs
                             are the same either way. */
} mpoa_server_opts;
typedef struct {
    double submit_ms, admit_ms, done_ms; /* on the server's clock (ms since creation) */
    uint64_t cells;                      /* DP cells of the job's alignments */
} mpoa_job_times;

int mpoa_server_create(mpoa_ctx *ctx, const mpoa_params *params, const mpoa_server_opts *opts,
                       mpoa_server **out);
/* One job: nreads reads, read k = bases[read_off[k] .. read_off[k+1]) (copied).
 * extend: NULL for a new graph, or a graph holding `existing` sequences that
 * the job extends as src/lib.rs:184-204 `extend_graph` does for the -i mode
 * (the job works on a copy; `extend` is not modified). *ticket: the job id. */
int mpoa_server_submit(mpoa_server *s, const uint8_t *bases, const uint64_t *read_off, uint64_t nreads,
                       const mpoa_graph *extend, uint64_t existing, uint64_t *ticket);
/* 1 when the job finished, 0 while it runs or waits for admission. */
int mpoa_server_done(mpoa_server *s, uint64_t ticket);
/* Blocks until the job finished (MPOA_OK) or the server failed. */
int mpoa_server_wait(mpoa_server *s, uint64_t ticket);
/* The finished job's status (MPOA_ERR_PENDING before it finished). */
int mpoa_server_job_status(mpoa_server *s, uint64_t ticket);
/* As mpoa_batch_job_output. */
int mpoa_server_job_output(mpoa_server *s, uint64_t ticket, int mode, char *out, uint64_t cap,
                           uint64_t *out_len);
int mpoa_server_job_times(mpoa_server *s, uint64_t ticket, mpoa_job_times *out);
/* A copy of the finished job's graph (to extend later, or to render). */
int mpoa_server_job_graph(mpoa_server *s, uint64_t ticket, mpoa_graph **out);
int mpoa_server_stats_get(mpoa_server *s, mpoa_batch_stats *out);
/* The server's clock (ms since creation). */
double mpoa_server_now_ms(mpoa_server *s);
/* Finishes every submitted job, then releases the context's launch slots. */
void mpoa_server_free(mpoa_server *s);

#ifdef __cplusplus
}
#endif
#endif
