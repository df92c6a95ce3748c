/*
 * TEST INFRASTRUCTURE ONLY -- the oracle. A plain-C restatement of the
 * reference's POA hot path (minipoa-rs src/align.rs:16-1347), the graph it
 * reads and writes (src/graph.rs:30-331) and the library driver
 * (src/lib.rs:109-204). It exists to check the HIP product path and to time
 * the reference's algorithm on host cores; the product never links it.
 *
 * Pinning (what fixes each part independently of this file):
 *   insertion prefix      serial recurrence I<-max(I+e, md+o) (tests)
 *   DP optimum            textbook Gotoh DP + path rescoring on the
 *                         reference's LCG seeds (tests/test_oracle.py)
 *   branched optimum      brute force over all DAG paths (tiny graphs)
 *   i32 kernel == i64     cross-implementation equality incl. errors
 *   graph/consensus       reference KATs (src/graph.rs:344-392)
 *   pipeline bytes        reference KATs (tests/io_cli.rs:421-451, regressions)
 *   MSA/GFA structure     invariants of tests/avx512.rs:27-91
 *   seeding               the reference's src/minimizer.rs tests (k-mer
 *                         values, run splitting, repeat filter, interval
 *                         merge, chain monotonicity, drift split, minimum
 *                         chain score) and its seeded-alignment invariants
 *                         (src/align.rs:1525-1559, tests/avx512.rs:145-180)
 *   ordering rules        hand-derived known answers (tests/rule_kats.py):
 *                         choose ties and weights, the sum(w)/10 filter,
 *                         open-before-extend for deletions and insertions,
 *                         head/tail support ties, streak-30 recentring,
 *                         extra growth and its 5 x band cap, LIFO toposort,
 *                         consensus ties; each fails when its rule is
 *                         mutated here (ORACLE_MUTANT, tests/test_oracle_mutants.py)
 * Parity unpinned: minimizer positions (see the seeding section); the reset
 * of `extra` on recentring and its zero-propagation to successors, and the
 * sharing of `support` between the head and tail passes, are pinned only by
 * this restatement being line-faithful.
 */
#include "poa_oracle.h"

#include <pthread.h>
#include <ctype.h>
#include <math.h>
#include <stdarg.h>
#include <stdatomic.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <immintrin.h>

/* Mutation testing: built with -DORACLE_MUTANT=k (tests/test_oracle_mutants.py
 * does so in a scratch directory), one ordering rule below is replaced by a
 * plausible wrong one, so that the tests can show their known answers fail
 * on it. The default, 0, is the faithful restatement. */
#ifndef ORACLE_MUTANT
#define ORACLE_MUTANT 0
#endif
#define MUTANT(k) (ORACLE_MUTANT == (k))

/* src/align.rs:16-23, 216-220 */
#define NEG (-(INT64_C(1) << 50))
#define WIDTH 8
#define MATCH 1u
#define DELETE 2u
#define INSERT 4u
#define ALL (MATCH | DELETE | INSERT)
#define NEG32 (-(INT32_C(1) << 29))
#define SCORE_LIMIT (UINT64_C(1) << 26)
#define SHIFT_FILL (NEG32 - 2 * (int32_t)SCORE_LIMIT)

static __thread char g_err[512];
static __thread uint64_t g_cells;
static __thread int64_t g_last_score;

static int fail(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return 1;
}
const char *o_last_error(void) { return g_err; }
uint64_t o_cells(void) { return g_cells; }
int64_t o_last_score(void) { return g_last_score; }
void o_reset_cells(void) { g_cells = 0; }
void o_free(void *ptr) { free(ptr); }

static void *xmalloc(size_t n) {
    void *p = malloc(n ? n : 1);
    if (!p) {
        fprintf(stderr, "oracle: out of memory\n");
        abort();
    }
    return p;
}
static void *xcalloc(size_t n, size_t s) {
    void *p = calloc(n ? n : 1, s ? s : 1);
    if (!p) {
        fprintf(stderr, "oracle: out of memory\n");
        abort();
    }
    return p;
}
static void *xrealloc(void *p, size_t n) {
    p = realloc(p, n ? n : 1);
    if (!p) {
        fprintf(stderr, "oracle: out of memory\n");
        abort();
    }
    return p;
}

/* Wrapping i32 arithmetic: wide::i32x8 lanes add with two's-complement wrap. */
static inline int32_t wadd(int32_t a, int32_t b) { return (int32_t)((uint32_t)a + (uint32_t)b); }
static inline int32_t max32(int32_t a, int32_t b) { return a > b ? a : b; }
static inline int64_t max64(int64_t a, int64_t b) { return a > b ? a : b; }
static inline int64_t min64(int64_t a, int64_t b) { return a < b ? a : b; }

typedef struct {
    size_t *v;
    size_t n, cap;
} vecz;
static void vz_push(vecz *a, size_t x) {
    if (a->n == a->cap) {
        a->cap = a->cap ? 2 * a->cap : 4;
        a->v = xrealloc(a->v, a->cap * sizeof(size_t));
    }
    a->v[a->n++] = x;
}

typedef struct {
    char *s;
    size_t n, cap;
} sbuf;
static void sb_putn(sbuf *b, const char *s, size_t n) {
    if (b->n + n + 1 > b->cap) {
        size_t c = b->cap ? b->cap : 256;
        while (c < b->n + n + 1) c *= 2;
        b->s = xrealloc(b->s, c);
        b->cap = c;
    }
    memcpy(b->s + b->n, s, n);
    b->n += n;
    b->s[b->n] = 0;
}
static void sb_puts(sbuf *b, const char *s) { sb_putn(b, s, strlen(s)); }
static void sb_putc(sbuf *b, char c) { sb_putn(b, &c, 1); }
static void sb_printf(sbuf *b, const char *fmt, ...) {
    char tmp[128];
    va_list ap;
    va_start(ap, fmt);
    int k = vsnprintf(tmp, sizeof tmp, fmt, ap);
    va_end(ap);
    sb_putn(b, tmp, (size_t)k);
}

typedef struct {
    ob_base *v;
    size_t n, cap;
} obvec;
static void ob_push(obvec *a, int64_t node, int64_t base, int64_t aligned) {
    if (a->n == a->cap) {
        a->cap = a->cap ? 2 * a->cap : 64;
        a->v = xrealloc(a->v, a->cap * sizeof(ob_base));
    }
    a->v[a->n].node = node;
    a->v[a->n].base = base;
    a->v[a->n].aligned = aligned;
    a->n++;
}

/* ---------------- Params (src/lib.rs:109-173) ---------------- */

static size_t p_alpha(const op_params *p) { return p->alphabet_len; }

/* src/lib.rs:110-126 */
size_t o_encode(const op_params *p, uint8_t base) {
    uint8_t b = (base >= 'a' && base <= 'z') ? (uint8_t)(base - 32) : base;
    if (p->alphabet_len <= 5 && b == 'U') b = 'T';
    for (size_t i = 0; i < p->alphabet_len; i++)
        if (p->alphabet[i] == b) return i;
    for (size_t i = 0; i < p->alphabet_len; i++)
        if (p->alphabet[i] == 'N') return i;
    return p->alphabet_len - 1;
}
/* src/lib.rs:127-134 */
static uint8_t o_decode(const op_params *p, size_t base) {
    uint8_t c = p->alphabet[base];
    if (p->is_rna && p->alphabet_len <= 5 && c == 'T') return 'U';
    return c;
}
/* src/lib.rs:135-137 */
static inline int32_t o_score(const op_params *p, size_t a, size_t b) {
    return p->matrix[a * p->alphabet_len + b];
}

/* src/lib.rs:149-173 */
int o_validate(const op_params *p) {
    size_t n = p->alphabet_len;
    if (n == 0 || n > 256 || p->matrix == NULL) return fail("invalid scoring matrix dimensions");
    if (p->threads == 0) return fail("thread count must be positive");
    if (!(p->k >= 1 && p->k <= 31) || !(p->w >= 1 && p->w < 32768))
        return fail("k must be 1..31 and minimizer window must be 1..32767");
    if (p->gap_open > 0 || p->gap_ext > 0) return fail("gap penalties must be zero or negative");
    if (p->result > 2 || p->verbose > 2) return fail("result format and verbosity must be 0..2");
    double r = p->mm_filter_ratio;
    if (!(r == r) || r - r != 0.0 || !(r >= 0.0 && r <= 1.0))
        return fail("minimizer filter ratio must be between 0 and 1");
    if (p->win_max == 0) return fail("window maximum must be positive");
    return 0;
}

/* ---------------- Graph (src/graph.rs:6-331) ---------------- */

typedef struct {
    size_t id, rank, parent, base, coverage;
    vecz incoming, in_w, outgoing, out_w;
    int64_t *aligned;
} onode;

typedef struct {
    size_t seq_id;
    vecz ids;
} opath;

struct og_graph {
    onode *nodes;
    size_t n_nodes, cap_nodes;
    vecz rank;
    opath *paths;
    size_t n_paths, cap_paths;
    size_t alphabet_size;
};

/* src/graph.rs:42-61 */
size_t og_add_node(og_graph *g, size_t base) {
    size_t id = g->n_nodes;
    if (g->n_nodes == g->cap_nodes) {
        g->cap_nodes = g->cap_nodes ? 2 * g->cap_nodes : 64;
        g->nodes = xrealloc(g->nodes, g->cap_nodes * sizeof(onode));
    }
    onode *nd = &g->nodes[id];
    memset(nd, 0, sizeof *nd);
    nd->id = id;
    nd->rank = 0;
    nd->parent = id;
    nd->base = base;
    nd->aligned = xmalloc(g->alphabet_size * sizeof(int64_t));
    for (size_t a = 0; a < g->alphabet_size; a++) nd->aligned[a] = -1;
    nd->aligned[base] = (int64_t)id;
    g->n_nodes++;
    return id;
}

/* Test infrastructure: a copy of a graph given as flat arrays (node fields,
 * weighted adjacency in list order, rank order, paths), so that a graph
 * built elsewhere (for instance loaded from GFA by the library under test)
 * can be continued by the oracle. Node coverage is the sum of in-weights,
 * which is what add_edge leaves (src/graph.rs:63-80). */
og_graph *og_from_arrays(const op_params *p, size_t n_nodes, const uint64_t *rank,
                         const uint64_t *parent, const uint64_t *base, const uint64_t *in_off,
                         const uint64_t *in_node, const uint64_t *in_w, const uint64_t *out_off,
                         const uint64_t *out_node, const uint64_t *out_w, const int64_t *aligned,
                         size_t n_order, const uint64_t *order, size_t n_paths,
                         const uint64_t *path_seq, const uint64_t *path_off,
                         const uint64_t *path_ids) {
    og_graph *g = xcalloc(1, sizeof *g);
    g->alphabet_size = p->alphabet_len;
    for (size_t i = 0; i < n_nodes; i++) {
        size_t id = og_add_node(g, base[i]);
        onode *nd = &g->nodes[id];
        nd->rank = rank[i];
        nd->parent = parent[i];
        for (size_t a = 0; a < g->alphabet_size; a++) nd->aligned[a] = aligned[i * g->alphabet_size + a];
        for (uint64_t k = in_off[i]; k < in_off[i + 1]; k++) {
            vz_push(&nd->incoming, in_node[k]);
            vz_push(&nd->in_w, in_w[k]);
            nd->coverage += in_w[k];
        }
        for (uint64_t k = out_off[i]; k < out_off[i + 1]; k++) {
            vz_push(&nd->outgoing, out_node[k]);
            vz_push(&nd->out_w, out_w[k]);
        }
    }
    for (size_t r = 0; r < n_order; r++) vz_push(&g->rank, order[r]);
    for (size_t k = 0; k < n_paths; k++) {
        if (g->n_paths == g->cap_paths) {
            g->cap_paths = g->cap_paths ? 2 * g->cap_paths : 16;
            g->paths = xrealloc(g->paths, g->cap_paths * sizeof(opath));
        }
        opath *pa = &g->paths[g->n_paths++];
        memset(pa, 0, sizeof *pa);
        pa->seq_id = path_seq[k];
        for (uint64_t t = path_off[k]; t < path_off[k + 1]; t++) vz_push(&pa->ids, path_ids[t]);
    }
    return g;
}

/* src/graph.rs:30-40 */
og_graph *og_new(const op_params *p) {
    og_graph *g = xcalloc(1, sizeof *g);
    g->alphabet_size = p->alphabet_len;
    og_add_node(g, o_encode(p, 'N'));
    og_add_node(g, o_encode(p, 'N'));
    return g;
}

void og_free(og_graph *g) {
    if (!g) return;
    for (size_t i = 0; i < g->n_nodes; i++) {
        onode *nd = &g->nodes[i];
        free(nd->incoming.v);
        free(nd->in_w.v);
        free(nd->outgoing.v);
        free(nd->out_w.v);
        free(nd->aligned);
    }
    for (size_t i = 0; i < g->n_paths; i++) free(g->paths[i].ids.v);
    free(g->nodes);
    free(g->paths);
    free(g->rank.v);
    free(g);
}

/* src/graph.rs:63-80 */
void og_add_edge(og_graph *g, size_t from, size_t to) {
    if (from == to) return;
    onode *f = &g->nodes[from], *t = &g->nodes[to];
    size_t i;
    for (i = 0; i < f->outgoing.n; i++)
        if (f->outgoing.v[i] == to) break;
    if (i < f->outgoing.n) {
        f->out_w.v[i] += 1;
    } else {
        vz_push(&f->outgoing, to);
        vz_push(&f->out_w, 1);
    }
    for (i = 0; i < t->incoming.n; i++)
        if (t->incoming.v[i] == from) break;
    if (i < t->incoming.n) {
        t->in_w.v[i] += 1;
    } else {
        vz_push(&t->incoming, from);
        vz_push(&t->in_w, 1);
    }
    t->coverage += 1;
}

/* src/graph.rs:82-122 */
int og_add_path(og_graph *g, size_t seq_id, const ob_base *aln, size_t len) {
    for (size_t k = 0; k < len; k++) {
        if (aln[k].base < 0 || (uint64_t)aln[k].base >= g->alphabet_size)
            return fail("alignment base outside alphabet");
        if ((aln[k].node >= 0 && (uint64_t)aln[k].node >= g->n_nodes) ||
            (aln[k].aligned >= 0 && (uint64_t)aln[k].aligned >= g->n_nodes))
            return fail("alignment references an unknown node");
    }
    int has_prev = 0;
    size_t previous = 0;
    vecz path = {0};
    for (size_t k = 0; k < len; k++) {
        const ob_base *it = &aln[k];
        size_t id;
        size_t base = (size_t)it->base;
        if (it->node >= 0) {
            id = (size_t)it->node;
        } else if (it->aligned >= 0) {
            size_t parent = g->nodes[(size_t)it->aligned].parent;
            if (g->nodes[parent].aligned[base] >= 0) {
                id = (size_t)g->nodes[parent].aligned[base];
            } else {
                id = og_add_node(g, base);
                g->nodes[id].parent = parent;
                g->nodes[parent].aligned[base] = (int64_t)id;
            }
        } else {
            id = og_add_node(g, base);
        }
        if (has_prev) og_add_edge(g, previous, id);
        previous = id;
        has_prev = 1;
        if (id >= 2) vz_push(&path, id);
    }
    og_add_edge(g, has_prev ? previous : 0, 1);
    if (g->n_paths == g->cap_paths) {
        g->cap_paths = g->cap_paths ? 2 * g->cap_paths : 16;
        g->paths = xrealloc(g->paths, g->cap_paths * sizeof(opath));
    }
    g->paths[g->n_paths].seq_id = seq_id;
    g->paths[g->n_paths].ids = path;
    g->n_paths++;
    return 0;
}

/* src/graph.rs:124-178 -- LIFO Kahn order; `grouped` merges column groups. */
int og_toposort(og_graph *g, int grouped) {
    size_t nn = g->n_nodes;
    size_t *degree = xcalloc(nn, sizeof(size_t));
    vecz stack = {0};
    for (size_t id = 0; id < nn; id++) {
        onode *n = &g->nodes[id];
        if (grouped && n->parent != n->id) continue;
        if (grouped) {
            size_t s = 0;
            for (size_t a = 0; a < g->alphabet_size; a++)
                if (n->aligned[a] >= 0) s += g->nodes[(size_t)n->aligned[a]].incoming.n;
            degree[id] = s;
        } else {
            degree[id] = n->incoming.n;
        }
        if (degree[id] == 0) vz_push(&stack, id);
    }
    g->rank.n = 0;
    size_t visited = 0;
    int rc = 0;
    size_t fifo_head = 0; /* MUTANT(16): first in, first out */
    while (MUTANT(16) ? fifo_head < stack.n : stack.n > 0) {
        size_t id = MUTANT(16) ? stack.v[fifo_head++] : stack.v[--stack.n];
        size_t rank = g->rank.n;
        vz_push(&g->rank, id);
        size_t members[256];
        size_t nm = 0;
        if (grouped) {
            for (size_t a = 0; a < g->alphabet_size; a++)
                if (g->nodes[id].aligned[a] >= 0) members[nm++] = (size_t)g->nodes[id].aligned[a];
        } else {
            members[nm++] = id;
        }
        for (size_t mi = 0; mi < nm; mi++) {
            onode *mem = &g->nodes[members[mi]];
            mem->rank = rank;
            visited++;
            for (size_t e = 0; e < mem->outgoing.n; e++) {
                size_t next = mem->outgoing.v[e];
                if (grouped) next = g->nodes[next].parent;
                if (degree[next] == 0) {
                    rc = fail("invalid graph column dependencies");
                    goto done;
                }
                degree[next]--;
                if (degree[next] == 0) vz_push(&stack, next);
            }
        }
    }
    if (visited != nn) rc = fail("graph contains a cycle (possibly between alignment columns)");
done:
    free(degree);
    free(stack.v);
    return rc;
}

size_t og_n_nodes(const og_graph *g) { return g->n_nodes; }
void og_node_info(const og_graph *g, size_t id, uint64_t out6[6]) {
    const onode *n = &g->nodes[id];
    out6[0] = n->rank;
    out6[1] = n->parent;
    out6[2] = n->base;
    out6[3] = n->coverage;
    out6[4] = n->incoming.n;
    out6[5] = n->outgoing.n;
}
int64_t og_node_aligned(const og_graph *g, size_t id, size_t base) {
    return g->nodes[id].aligned[base];
}
void og_node_edges(const og_graph *g, size_t id, uint64_t *in_ids, uint64_t *in_w,
                   uint64_t *out_ids, uint64_t *out_w) {
    const onode *n = &g->nodes[id];
    for (size_t k = 0; k < n->incoming.n; k++) {
        in_ids[k] = n->incoming.v[k];
        in_w[k] = n->in_w.v[k];
    }
    for (size_t k = 0; k < n->outgoing.n; k++) {
        out_ids[k] = n->outgoing.v[k];
        out_w[k] = n->out_w.v[k];
    }
}
size_t og_rank_len(const og_graph *g) { return g->rank.n; }
void og_rank(const og_graph *g, uint64_t *out) {
    for (size_t k = 0; k < g->rank.n; k++) out[k] = g->rank.v[k];
}
size_t og_n_paths(const og_graph *g) { return g->n_paths; }

/* src/graph.rs:180-237 -- heaviest incoming edge, ties by path score then
 * by lower incoming index; FIFO Kahn from the source, stop at the sink. */
size_t og_consensus(const og_graph *g, const op_params *p, uint8_t *seq, uint64_t *ids_out,
                    uint64_t *cov, size_t cap) {
    size_t nn = g->n_nodes;
    size_t *degree = xmalloc(nn * sizeof(size_t));
    int64_t *previous = xmalloc(nn * sizeof(int64_t));
    size_t *score = xcalloc(nn, sizeof(size_t));
    size_t *queue = xmalloc(nn * sizeof(size_t));
    size_t qh = 0, qt = 0;
    for (size_t i = 0; i < nn; i++) {
        degree[i] = g->nodes[i].incoming.n;
        previous[i] = -1;
    }
    queue[qt++] = 0;
    while (qh < qt) {
        size_t id = queue[qh++];
        const onode *node = &g->nodes[id];
        if (id != 0 && node->incoming.n > 0) {
            size_t bi = 0;
            for (size_t k = 1; k < node->incoming.n; k++) {
                size_t a = node->incoming.v[bi], b = node->incoming.v[k];
                /* max_by keeps the accumulator only when it compares Greater. */
                int acc_greater;
                if (node->in_w.v[bi] != node->in_w.v[k])
                    acc_greater = node->in_w.v[bi] > node->in_w.v[k];
                else if (score[a] != score[b])
                    acc_greater = MUTANT(18) ? score[a] < score[b] : score[a] > score[b];
                else
                    acc_greater = MUTANT(17) ? k < bi : k > bi;
                if (!acc_greater) bi = k;
            }
            size_t pr = node->incoming.v[bi];
            previous[id] = (int64_t)pr;
            score[id] = score[pr] + node->in_w.v[bi];
        }
        if (id == 1) break;
        for (size_t e = 0; e < node->outgoing.n; e++) {
            size_t next = node->outgoing.v[e];
            degree[next]--;
            if (degree[next] == 0) queue[qt++] = next;
        }
    }
    vecz ids = {0};
    int64_t cur = previous[1];
    while (cur >= 0) {
        if (cur == 0) break;
        vz_push(&ids, (size_t)cur);
        cur = previous[cur];
    }
    size_t len = ids.n;
    for (size_t k = 0; k < len && k < cap; k++) {
        size_t id = ids.v[len - 1 - k];
        const onode *n = &g->nodes[id];
        if (seq) seq[k] = o_decode(p, n->base);
        if (ids_out) ids_out[k] = id;
        if (cov) {
            size_t s = 0;
            for (size_t e = 0; e < n->out_w.n; e++) s += n->out_w.v[e];
            cov[k] = n->coverage > s ? n->coverage : s;
        }
    }
    free(ids.v);
    free(degree);
    free(previous);
    free(score);
    free(queue);
    return len;
}

/* src/graph.rs:255-279 */
static int msa_grouped(const og_graph *g, const op_params *p, const char *const *names,
                       const char *const *comments, size_t nseq, sbuf *out) {
    size_t width = g->rank.n >= 2 ? g->rank.n - 2 : 0;
    char *row = xmalloc(width + 1);
    for (size_t pi = 0; pi < g->n_paths; pi++) {
        const opath *pa = &g->paths[pi];
        if (pa->seq_id >= nseq) {
            free(row);
            return fail("unknown sequence ID in graph path");
        }
        memset(row, '-', width);
        row[width] = 0;
        for (size_t k = 0; k < pa->ids.n; k++) {
            size_t id = pa->ids.v[k];
            size_t rank = g->nodes[g->nodes[id].parent].rank;
            if (rank == 0 || rank > width) {
                free(row);
                return fail("graph path is outside MSA columns");
            }
            row[rank - 1] = (char)o_decode(p, g->nodes[id].base);
        }
        sb_putc(out, '>');
        sb_puts(out, names[pa->seq_id]);
        sb_putc(out, ' ');
        sb_puts(out, comments ? comments[pa->seq_id] : "");
        sb_putc(out, '\n');
        sb_putn(out, row, width);
        sb_putc(out, '\n');
    }
    free(row);
    return 0;
}

/* src/graph.rs:239-253 -- grouped toposort, render, restore ranks. */
int og_msa(og_graph *g, const op_params *p, const char *const *names,
           const char *const *comments, size_t nseq, char **out) {
    sbuf b = {0};
    sb_puts(&b, "");
    if (g->n_paths == 0) {
        *out = b.s;
        return 0;
    }
    size_t *rank = xmalloc((g->rank.n + 1) * sizeof(size_t));
    size_t nrank = g->rank.n;
    memcpy(rank, g->rank.v, nrank * sizeof(size_t));
    size_t *node_ranks = xmalloc(g->n_nodes * sizeof(size_t));
    for (size_t i = 0; i < g->n_nodes; i++) node_ranks[i] = g->nodes[i].rank;
    int rc = og_toposort(g, 1);
    if (rc == 0) rc = msa_grouped(g, p, names, comments, nseq, &b);
    g->rank.n = 0;
    for (size_t k = 0; k < nrank; k++) vz_push(&g->rank, rank[k]);
    for (size_t i = 0; i < g->n_nodes; i++) g->nodes[i].rank = node_ranks[i];
    free(rank);
    free(node_ranks);
    if (rc) {
        free(b.s);
        *out = NULL;
        return rc;
    }
    *out = b.s;
    return 0;
}

/* src/graph.rs:281-331 */
int og_gfa(const og_graph *g, const op_params *p, const char *const *names, size_t nseq,
           char **out) {
    sbuf b = {0};
    sb_puts(&b, "");
    if (g->n_nodes <= 2 && g->n_paths == 0) {
        *out = b.s;
        return 0;
    }
    sb_puts(&b, "H\tVN:Z:1.0\n");
    size_t nn = g->n_nodes;
    size_t *degree = xmalloc(nn * sizeof(size_t));
    size_t *queue = xmalloc(nn * sizeof(size_t));
    size_t qh = 0, qt = 0;
    for (size_t i = 0; i < nn; i++) degree[i] = g->nodes[i].incoming.n;
    queue[qt++] = 0;
    while (qh < qt) {
        size_t id = queue[qh++];
        if (id == 1) break;
        const onode *node = &g->nodes[id];
        if (id >= 2) {
            sb_printf(&b, "S\t%zu\t%c\n", id - 1, (char)o_decode(p, node->base));
            for (size_t e = 0; e < node->incoming.n; e++) {
                size_t prev = node->incoming.v[e];
                if (prev != 0) sb_printf(&b, "L\t%zu\t+\t%zu\t+\t0M\n", prev - 1, id - 1);
            }
        }
        for (size_t e = 0; e < node->outgoing.n; e++) {
            size_t next = node->outgoing.v[e];
            degree[next]--;
            if (degree[next] == 0) queue[qt++] = next;
        }
    }
    free(degree);
    free(queue);
    for (size_t pi = 0; pi < g->n_paths; pi++) {
        const opath *pa = &g->paths[pi];
        if (pa->seq_id >= nseq) {
            free(b.s);
            *out = NULL;
            return fail("unknown sequence ID in graph path");
        }
        sb_puts(&b, "P\t");
        sb_puts(&b, names[pa->seq_id]);
        sb_putc(&b, '\t');
        if (pa->ids.n == 0) sb_putc(&b, '*');
        for (size_t k = 0; k < pa->ids.n; k++) {
            if (k > 0) sb_putc(&b, ',');
            sb_printf(&b, "%zu+", pa->ids.v[k] - 1);
        }
        sb_puts(&b, "\t*\n");
    }
    *out = b.s;
    return 0;
}

/* ---------------- DP common setup (src/align.rs:598-688, 913-1011) ------- */

typedef struct {
    size_t beg, end, n, m, padded, unknown;
    size_t *seq;
    int adaptive;
    int64_t band;
    int64_t *head, *tail, *support;
    int64_t *peak, *streak, *peak_left, *peak_right, *offset_left, *offset_right, *extra;
} dpsetup;

static void setup_free(dpsetup *s) {
    free(s->seq);
    free(s->head);
    free(s->tail);
    free(s->support);
    free(s->peak);
    free(s->streak);
    free(s->peak_left);
    free(s->peak_right);
    free(s->offset_left);
    free(s->offset_right);
    free(s->extra);
    memset(s, 0, sizeof *s);
}

/* Anchor checks, encoded/padded query, band, and the head/tail/support
 * most-supported-path distances. src/align.rs:598-688 == 921-1011. */
static int dp_setup(const op_params *p, const og_graph *g, size_t beg_id, size_t end_id,
                    const uint8_t *q, size_t qlen, int adaptive, size_t band_b, dpsetup *s) {
    memset(s, 0, sizeof *s);
    if (beg_id >= g->n_nodes || end_id >= g->n_nodes) return fail("alignment anchor is not in graph");
    size_t beg = g->nodes[beg_id].rank, end = g->nodes[end_id].rank;
    if (end <= beg || end >= g->rank.n) return fail("alignment anchors are not topologically ordered");
    size_t n = end - beg + 1;
    size_t m = qlen + 2;
    if (n > (size_t)INT32_MAX || m > (size_t)INT32_MAX)
        return fail("alignment exceeds coordinate limits");
    s->beg = beg;
    s->end = end;
    s->n = n;
    s->m = m;
    s->unknown = o_encode(p, 'N');
    s->padded = (m / WIDTH + 1) * WIDTH;
    s->seq = xmalloc(s->padded * sizeof(size_t));
    s->seq[0] = s->unknown;
    for (size_t k = 0; k < qlen; k++) s->seq[k + 1] = o_encode(p, q[k]);
    for (size_t k = qlen + 1; k < s->padded; k++) s->seq[k] = s->unknown;
    s->adaptive = adaptive || p->adaptive;
    uint64_t mdivf = p->f ? (uint64_t)m / p->f : 0;
    uint64_t bw = (uint64_t)band_b + mdivf;
    if (bw < (uint64_t)band_b) bw = UINT64_MAX; /* saturating_add */
    if (bw > (uint64_t)INT32_MAX) bw = INT32_MAX;
    s->band = (int64_t)bw;
    s->head = xmalloc(n * sizeof(int64_t));
    s->tail = xmalloc(n * sizeof(int64_t));
    s->support = xcalloc(n, sizeof(int64_t));
    for (size_t i = 0; i < n; i++) s->head[i] = s->tail[i] = NEG;
    s->head[0] = 0;
    for (size_t i = 1; i < n; i++) {
        const onode *node = &g->nodes[g->rank.v[beg + i]];
        int have = 0;
        size_t bp = 0, bw2 = 0;
        for (size_t k = 0; k < node->incoming.n; k++) {
            size_t rank = g->nodes[node->incoming.v[k]].rank;
            if (rank < beg || rank >= beg + i) continue;
            size_t pr = rank - beg;
            if (s->head[pr] == NEG) continue;
            size_t w = node->in_w.v[k];
            const int better = MUTANT(9) ? s->support[pr] < s->support[bp]
                                         : s->support[pr] > s->support[bp];
            if (!have || w > bw2 || (w == bw2 && better)) {
                have = 1;
                bp = pr;
                bw2 = w;
            }
        }
        if (have) {
            s->head[i] = s->head[bp] + 1;
            s->support[i] = s->support[bp] + (int64_t)bw2;
        }
    }
    s->tail[n - 1] = 1;
    for (size_t ii = n - 1; ii-- > 0;) {
        size_t i = ii;
        const onode *node = &g->nodes[g->rank.v[beg + i]];
        int have = 0;
        size_t bs = 0, bw2 = 0;
        for (size_t k = 0; k < node->outgoing.n; k++) {
            size_t rank = g->nodes[node->outgoing.v[k]].rank;
            if (rank <= beg + i || rank > end) continue;
            size_t sc = rank - beg;
            if (s->tail[sc] == NEG) continue;
            size_t w = node->out_w.v[k];
            const int better = MUTANT(10) ? s->support[sc] < s->support[bs]
                                          : s->support[sc] > s->support[bs];
            if (!have || w > bw2 || (w == bw2 && better)) {
                have = 1;
                bs = sc;
                bw2 = w;
            }
        }
        if (have) {
            s->tail[i] = s->tail[bs] + 1;
            s->support[i] = s->support[bs] + (int64_t)bw2;
        }
    }
    if (s->head[n - 1] == NEG) {
        setup_free(s);
        return fail("no graph path connects alignment anchors");
    }
    s->peak = xcalloc(n, sizeof(int64_t));
    s->streak = xcalloc(n, sizeof(int64_t));
    s->peak_left = xmalloc(n * sizeof(int64_t));
    s->peak_right = xcalloc(n, sizeof(int64_t));
    s->offset_left = xcalloc(n, sizeof(int64_t));
    s->offset_right = xcalloc(n, sizeof(int64_t));
    s->extra = xcalloc(n, sizeof(int64_t));
    for (size_t i = 0; i < n; i++) s->peak_left[i] = (int64_t)m;
    return 0;
}

/* Row band, src/align.rs:711-734 == 1017-1040. */
static int row_band(const op_params *p, const dpsetup *s, size_t i, size_t *start, size_t *stop) {
    int64_t m = (int64_t)s->m, left, right;
    if (p->f == 0) {
        left = 0;
        right = m;
    } else if (s->adaptive) {
        int64_t a = min64(min64(s->peak_left[i], s->head[i] + s->offset_left[i]), m - s->tail[i]);
        int64_t b = max64(max64(s->peak_right[i], s->head[i] + s->offset_right[i]), m - s->tail[i]);
        left = max64(a - s->band - s->extra[i], 0);
        right = min64(b + s->band + s->extra[i], m);
    } else {
        left = max64(min64(s->head[i], m - s->tail[i]) - s->band, 0);
        right = min64(max64(s->head[i], m - s->tail[i]) + s->band, m);
    }
    if (left > right || right < 0) return fail("alignment band does not intersect graph path");
    *start = (size_t)left / WIDTH * WIDTH;
    *stop = ((size_t)right / WIDTH + 1) * WIDTH;
    return 0;
}

/* Adaptive band bookkeeping after row i, src/align.rs:829-897 == 1088-1146.
 * `row_m` are the row's M scores as i64; argmax is the first maximum. */
static void adaptive_update(const og_graph *g, dpsetup *s, size_t i, size_t start,
                            int64_t maximum_col) {
    const onode *node = &g->nodes[g->rank.v[s->beg + i]];
    int64_t maximum = maximum_col;
    s->peak[i] = maximum;
    int64_t max_streak = 0;
    int is_match = 0;
    (void)start;
    if (node->base == s->seq[(size_t)maximum]) {
        for (size_t k = 0; k < node->incoming.n; k++) {
            size_t rank = g->nodes[node->incoming.v[k]].rank;
            if (rank < s->beg || rank >= s->beg + i) continue;
            size_t pr = rank - s->beg;
            if (s->head[pr] == NEG) continue;
            max_streak = max64(max_streak, s->streak[pr]);
            if (s->peak[pr] + 1 == maximum) {
                s->streak[i] = max64(s->streak[i], s->streak[pr] + 1);
                is_match = 1;
            }
        }
    }
    if (!is_match && max_streak < 30)
        s->extra[i] = min64(s->extra[i] + max_streak + (MUTANT(13) ? 0 : 1),
                            (MUTANT(14) ? 4 : 5) * s->band);
    if (!MUTANT(11) && s->streak[i] >= (MUTANT(12) ? 31 : 30)) {
        s->offset_left[i] = maximum - s->head[i];
        s->offset_right[i] = maximum - s->head[i];
        if (!MUTANT(15)) s->extra[i] = 0;
        for (size_t k = 0; k < node->outgoing.n; k++) {
            size_t rank = g->nodes[node->outgoing.v[k]].rank;
            if (rank > s->beg + i && rank <= s->end) {
                size_t sc = rank - s->beg;
                s->offset_left[sc] = s->offset_left[i];
                s->offset_right[sc] = s->offset_right[i];
            }
        }
    }
    for (size_t k = 0; k < node->outgoing.n; k++) {
        size_t rank = g->nodes[node->outgoing.v[k]].rank;
        if (rank <= s->beg + i || rank > s->end) continue;
        size_t sc = rank - s->beg;
        s->offset_left[sc] = min64(s->offset_left[sc], s->offset_left[i]);
        s->offset_right[sc] = max64(s->offset_right[sc], s->offset_right[i]);
        s->peak_left[sc] = min64(s->peak_left[sc], maximum);
        s->peak_right[sc] = max64(s->peak_right[sc], maximum + 1);
        s->extra[sc] = (MUTANT(19) || s->extra[i] != 0) ? max64(s->extra[sc], s->extra[i]) : 0;
    }
}

/* ---------------- Traceback (src/align.rs:276-353, 1164-1347) ------------ */

typedef struct omat omat;
struct omat {
    int (*contains)(const omat *, size_t r, size_t c);
    int64_t (*m)(const omat *, size_t r, size_t c);
    int64_t (*d)(const omat *, size_t r, size_t c);
    int64_t (*ins)(const omat *, size_t r, size_t c);
    int64_t negative;
    /* i64 rows (src/align.rs:25-77) */
    size_t *r_start, *r_len;
    int64_t **r_m, **r_d, **r_i;
    /* i32 arena rows (src/align.rs:228-274) */
    const int32_t *data;
    size_t *f_start, *f_len, *f_m, *f_d, *f_i;
};

static int s_contains(const omat *x, size_t r, size_t c) {
    return c >= x->r_start[r] && c - x->r_start[r] < x->r_len[r];
}
static int64_t s_m(const omat *x, size_t r, size_t c) {
    return s_contains(x, r, c) ? x->r_m[r][c - x->r_start[r]] : NEG;
}
static int64_t s_d(const omat *x, size_t r, size_t c) {
    return s_contains(x, r, c) ? x->r_d[r][c - x->r_start[r]] : NEG;
}
static int64_t s_i(const omat *x, size_t r, size_t c) {
    return s_contains(x, r, c) ? x->r_i[r][c - x->r_start[r]] : NEG;
}
static int f_contains(const omat *x, size_t r, size_t c) {
    return c >= x->f_start[r] && c - x->f_start[r] < x->f_len[r];
}
static int64_t f_m(const omat *x, size_t r, size_t c) {
    return f_contains(x, r, c) ? (int64_t)x->data[x->f_m[r] + c - x->f_start[r]] : (int64_t)NEG32;
}
static int64_t f_d(const omat *x, size_t r, size_t c) {
    return f_contains(x, r, c) ? (int64_t)x->data[x->f_d[r] + c - x->f_start[r]] : (int64_t)NEG32;
}
static int64_t f_i(const omat *x, size_t r, size_t c) {
    return f_contains(x, r, c) ? (int64_t)x->data[x->f_i[r] + c - x->f_start[r]] : (int64_t)NEG32;
}

/* src/align.rs:1315-1347 */
static void push_diagonal(obvec *res, const og_graph *g, size_t i, size_t n, size_t j, size_t m,
                          size_t query_base, size_t node_id, size_t node_base) {
    if (i == n - 1 && j != m - 1) {
        ob_push(res, -1, (int64_t)query_base, -1);
    } else if (node_base == query_base) {
        ob_push(res, (int64_t)node_id, (int64_t)g->nodes[node_id].base, -1);
    } else {
        size_t parent = g->nodes[node_id].parent;
        ob_push(res, g->nodes[parent].aligned[query_base], (int64_t)query_base, (int64_t)parent);
    }
}

/* src/align.rs:1201-1225 `choose`: max in_weight (strict >, first wins) among
 * in-range predecessors whose state equals target - penalty. */
static int64_t tb_choose(const og_graph *g, const dpsetup *s, const omat *x, const onode *node,
                         size_t i, size_t j, unsigned state, int64_t penalty, size_t pj) {
    int64_t best = -1;
    size_t weight = 0;
    for (size_t k = 0; k < node->incoming.n; k++) {
        size_t rank = g->nodes[node->incoming.v[k]].rank;
        if (rank < s->beg || rank >= s->beg + i) continue;
        size_t pr = rank - s->beg;
        if (s->head[pr] == NEG || !x->contains(x, pr, pj)) continue;
        /* MUTANT(20): the diagonal only from predecessors whose band also
         * holds column j (the fill's overlap), not one ending at j - 1 */
        if (MUTANT(20) && state == MATCH && !x->contains(x, pr, j)) continue;
        int64_t source = state == DELETE ? x->d(x, pr, pj) : x->m(x, pr, pj);
        int64_t target = state == MATCH ? x->m(x, i, j) : x->d(x, i, j);
        const int heavier = MUTANT(1) ? node->in_w.v[k] >= weight
                            : MUTANT(2) ? 0
                                        : node->in_w.v[k] > weight;
        if (target == source + penalty && (best < 0 || heavier)) {
            best = (int64_t)k;
            weight = node->in_w.v[k];
        }
    }
    return best;
}

/* src/align.rs:1164-1313 */
static int traceback(const op_params *p, const og_graph *g, const dpsetup *s, const omat *x,
                     int64_t gap_open, int64_t gap_ext, obvec *result) {
    size_t n = s->n, m = s->m, beg = s->beg;
    size_t i = n - 1, j = m - 1;
    unsigned op = ALL;
    size_t steps = 0;
    result->n = 0;
    while (i > 0 || j > 0) {
        steps++;
        if (steps > n + m) return fail("alignment traceback failed to advance");
        if (!x->contains(x, i, j) || x->m(x, i, j) <= x->negative / 2)
            return fail("alignment traceback left active band");
        const onode *node = &g->nodes[g->rank.v[beg + i]];
        size_t base = i == n - 1 ? o_encode(p, 'N') : node->base;
        int64_t score = o_score(p, base, s->seq[j]);
        int64_t pref = -1;
        if ((op & MATCH) && j > 0 && base == s->seq[j])
            pref = tb_choose(g, s, x, node, i, j, MATCH, score, j - 1);
        if (pref >= 0) {
            size_t tot = 0;
            for (size_t k = 0; k < node->in_w.n; k++) tot += node->in_w.v[k];
            const int keep = MUTANT(3) ? 1
                             : MUTANT(4) ? node->in_w.v[pref] > tot / 10
                                         : node->in_w.v[pref] >= tot / 10;
            if (!keep) pref = -1;
        }
        if (pref >= 0) {
            push_diagonal(result, g, i, n, j, m, s->seq[j], node->id, base);
            i = g->nodes[node->incoming.v[pref]].rank - beg;
            j -= 1;
            op = ALL;
            continue;
        }
        if ((op & DELETE) && (op == DELETE || x->m(x, i, j) == x->d(x, i, j))) {
            int64_t k;
            if (MUTANT(5)) {
                k = tb_choose(g, s, x, node, i, j, DELETE, gap_ext, j);
                if (k >= 0) {
                    i = g->nodes[node->incoming.v[k]].rank - beg;
                    op = DELETE;
                    continue;
                }
            }
            k = tb_choose(g, s, x, node, i, j, INSERT, gap_open, j);
            if (k >= 0) {
                i = g->nodes[node->incoming.v[k]].rank - beg;
                op = (p->gap_open == 0 && !MUTANT(6)) ? ALL : (MATCH | INSERT);
                continue;
            }
            k = tb_choose(g, s, x, node, i, j, DELETE, gap_ext, j);
            if (k >= 0) {
                i = g->nodes[node->incoming.v[k]].rank - beg;
                op = DELETE;
                continue;
            }
        }
        if ((op & INSERT) && j > 0 && x->contains(x, i, j - 1) &&
            (op == INSERT || x->m(x, i, j) == x->ins(x, i, j))) {
            if (MUTANT(8) && x->ins(x, i, j) == x->ins(x, i, j - 1) + gap_ext) {
                ob_push(result, -1, (int64_t)s->seq[j], -1);
                j -= 1;
                op = INSERT;
                continue;
            }
            if (x->ins(x, i, j) == x->m(x, i, j - 1) + gap_open) {
                ob_push(result, -1, (int64_t)s->seq[j], -1);
                j -= 1;
                op = (p->gap_open == 0 && !MUTANT(7)) ? ALL : (MATCH | DELETE);
                continue;
            }
            if (x->ins(x, i, j) == x->ins(x, i, j - 1) + gap_ext) {
                ob_push(result, -1, (int64_t)s->seq[j], -1);
                j -= 1;
                op = INSERT;
                continue;
            }
        }
        if ((op & MATCH) && j > 0) {
            int64_t k = tb_choose(g, s, x, node, i, j, MATCH, score, j - 1);
            if (k >= 0) {
                push_diagonal(result, g, i, n, j, m, s->seq[j], node->id, base);
                i = g->nodes[node->incoming.v[k]].rank - beg;
                j -= 1;
                op = ALL;
                continue;
            }
        }
        return fail("alignment traceback has no predecessor at graph row %zu, query column %zu", i, j);
    }
    size_t start = g->rank.v[beg];
    ob_push(result, (int64_t)start, (int64_t)g->nodes[start].base, -1);
    /* reverse, then pop the last element */
    for (size_t a = 0, b = result->n - 1; a < b; a++, b--) {
        ob_base t = result->v[a];
        result->v[a] = result->v[b];
        result->v[b] = t;
    }
    result->n -= 1;
    return 0;
}

static int emit(const obvec *res, ob_base *out, size_t cap, size_t *out_len) {
    *out_len = res->n;
    for (size_t k = 0; k < res->n && k < cap; k++) out[k] = res->v[k];
    return 0;
}

/* ---------------- i64 scalar kernel (src/align.rs:913-1162) ------------- */

int o_poa_scalar_i64(const op_params *p, const og_graph *g, size_t beg_id, size_t end_id,
                     const uint8_t *q, size_t qlen, int adaptive, size_t band_b, ob_base *out,
                     size_t cap, size_t *out_len) {
    dpsetup s;
    if (dp_setup(p, g, beg_id, end_id, q, qlen, adaptive, band_b, &s)) return 1;
    size_t n = s.n, beg = s.beg;
    int64_t gap_ext = p->gap_ext;
    int64_t gap_open = (int64_t)p->gap_open + gap_ext;
    omat x;
    memset(&x, 0, sizeof x);
    x.contains = s_contains;
    x.m = s_m;
    x.d = s_d;
    x.ins = s_i;
    x.negative = NEG;
    x.r_start = xcalloc(n, sizeof(size_t));
    x.r_len = xcalloc(n, sizeof(size_t));
    x.r_m = xcalloc(n, sizeof(int64_t *));
    x.r_d = xcalloc(n, sizeof(int64_t *));
    x.r_i = xcalloc(n, sizeof(int64_t *));
    int rc = 0;
    obvec res = {0};
    for (size_t i = 0; i < n; i++) {
        if (s.head[i] == NEG || s.tail[i] == NEG) continue;
        size_t start, stop;
        if (row_band(p, &s, i, &start, &stop)) {
            rc = 1;
            goto out;
        }
        size_t len = stop - start;
        int64_t *rm = xmalloc(len * sizeof(int64_t)), *rd = xmalloc(len * sizeof(int64_t)),
                *ri = xmalloc(len * sizeof(int64_t));
        for (size_t k = 0; k < len; k++) rm[k] = rd[k] = ri[k] = NEG;
        x.r_start[i] = start;
        x.r_len[i] = len;
        x.r_m[i] = rm;
        x.r_d[i] = rd;
        x.r_i[i] = ri;
        g_cells += len;
        if (i == 0) {
            if (!s_contains(&x, 0, 0)) {
                rc = fail("alignment band excludes source");
                goto out;
            }
            rm[0] = o_score(p, s.unknown, s.unknown);
            for (size_t j = 1; j < len; j++) {
                ri[j] = max64(ri[j - 1] + gap_ext, rm[j - 1] + gap_open);
                rm[j] = max64(rm[j], ri[j]);
            }
            continue;
        }
        const onode *node = &g->nodes[g->rank.v[beg + i]];
        size_t base = i == n - 1 ? s.unknown : node->base;
        for (size_t j = start, local = 0; j < stop; j++, local++) {
            size_t block = j / WIDTH * WIDTH;
            int64_t diagonal = NEG, deletion = NEG;
            for (size_t k = 0; k < node->incoming.n; k++) {
                size_t rank = g->nodes[node->incoming.v[k]].rank;
                if (rank < beg || rank >= beg + i) continue;
                size_t pr = rank - beg;
                if (s.head[pr] == NEG) continue;
                if (!s_contains(&x, pr, block)) continue;
                if (j > 0) diagonal = max64(diagonal, s_m(&x, pr, j - 1));
                deletion = max64(max64(deletion, s_d(&x, pr, j) + gap_ext), s_m(&x, pr, j) + gap_open);
            }
            rd[local] = deletion;
            int64_t md = max64(diagonal + (int64_t)o_score(p, base, s.seq[j]), deletion);
            rm[local] = max64(md, ri[local]);
            if (local + 1 < len) ri[local + 1] = max64(ri[local] + gap_ext, md + gap_open);
        }
        if (p->f != 0 && s.adaptive) {
            size_t argmax = 0;
            for (size_t j = 1; j < len; j++)
                if (rm[j] > rm[argmax]) argmax = j;
            adaptive_update(g, &s, i, start, (int64_t)(start + argmax));
        }
    }
    g_last_score = s_m(&x, n - 1, s.m - 1);
    rc = traceback(p, g, &s, &x, gap_open, gap_ext, &res);
    if (!rc) emit(&res, out, cap, out_len);
out:
    for (size_t i = 0; i < n; i++) {
        free(x.r_m[i]);
        free(x.r_d[i]);
        free(x.r_i[i]);
    }
    free(x.r_start);
    free(x.r_len);
    free(x.r_m);
    free(x.r_d);
    free(x.r_i);
    free(res.v);
    setup_free(&s);
    return rc;
}

/* ---------------- i32 8-lane kernel (src/align.rs:222-911) -------------- */

/* src/align.rs:222-226 worker-local arena, retained across calls. */
static __thread int32_t *t_arena;
static __thread size_t t_arena_len, t_arena_cap;
size_t o_arena_capacity(void) { return t_arena_cap; }
void o_arena_clear(void) {
    free(t_arena);
    t_arena = NULL;
    t_arena_len = t_arena_cap = 0;
}

/* i32x8 lanes as AVX2 registers: lane-wise wrapping add and max, constant
 * lane shifts (the `wide::i32x8` operations of src/align.rs:356-375). */
typedef __m256i v8;
static inline v8 ld8(const int32_t *p) { return _mm256_loadu_si256((const __m256i *)p); }
static inline void st8(int32_t *p, v8 v) { _mm256_storeu_si256((__m256i *)p, v); }
static inline v8 splat8(int32_t x) { return _mm256_set1_epi32(x); }
static inline v8 add8(v8 a, v8 b) { return _mm256_add_epi32(a, b); }
static inline v8 max8(v8 a, v8 b) { return _mm256_max_epi32(a, b); }
/* [fill x k, a0 .. a(7-k)] */
static inline v8 shift1(v8 a) {
    v8 t = _mm256_permutevar8x32_epi32(a, _mm256_setr_epi32(0, 0, 1, 2, 3, 4, 5, 6));
    return _mm256_blend_epi32(t, splat8(SHIFT_FILL), 0x01);
}
static inline v8 shift2(v8 a) {
    v8 t = _mm256_permutevar8x32_epi32(a, _mm256_setr_epi32(0, 0, 0, 1, 2, 3, 4, 5));
    return _mm256_blend_epi32(t, splat8(SHIFT_FILL), 0x03);
}
static inline v8 shift4(v8 a) {
    v8 t = _mm256_permutevar8x32_epi32(a, _mm256_setr_epi32(0, 0, 0, 0, 0, 1, 2, 3));
    return _mm256_blend_epi32(t, splat8(SHIFT_FILL), 0x0F);
}

/* src/align.rs:459-491: three shift-max stages padded with SHIFT_FILL, then
 * the carry extended across the eight lanes. `ext_steps` = ext * (1..8). */
static inline v8 insertion_prefix8(v8 md, int32_t carry, v8 open, v8 e1, v8 e2, v8 e4,
                                   v8 ext_steps) {
    v8 v = add8(md, open);
    v = max8(v, add8(shift1(v), e1));
    v = max8(v, add8(shift2(v), e2));
    v = max8(v, add8(shift4(v), e4));
    return max8(v, add8(splat8(carry), ext_steps));
}

static inline v8 ext_steps8(int32_t ext) {
    int32_t t[8];
    for (int l = 0; l < 8; l++) t[l] = (int32_t)((uint32_t)ext * (uint32_t)(l + 1));
    return ld8(t);
}

void o_insertion_prefix(const int32_t md[8], int32_t carry, int32_t open, int32_t ext,
                        int32_t out[8]) {
    v8 r = insertion_prefix8(ld8(md), carry, splat8(open), splat8(ext),
                             splat8((int32_t)((uint32_t)ext * 2u)),
                             splat8((int32_t)((uint32_t)ext * 4u)), ext_steps8(ext));
    st8(out, r);
}

/* src/align.rs:493-529 */
int o_i32_scores_safe(const op_params *p, const og_graph *g, size_t beg_id, size_t end_id,
                      size_t query_len) {
    if (p->scalar_dp || beg_id >= g->n_nodes || end_id >= g->n_nodes) return 0;
    size_t re = g->nodes[end_id].rank, rb = g->nodes[beg_id].rank;
    if (re < rb) return 0;
    uint64_t rows = re - rb;
    uint64_t steps = rows + query_len;
    if (steps < rows) return 0;
    if (steps + WIDTH * 2 + 4 < steps) return 0;
    steps += WIDTH * 2 + 4;
    int64_t opening = (int64_t)p->gap_open + (int64_t)p->gap_ext;
    uint64_t largest = 0;
    size_t nm = (size_t)p->alphabet_len * p->alphabet_len;
    for (size_t k = 0; k < nm; k++) {
        int64_t v = p->matrix[k];
        uint64_t a = (uint64_t)(v < 0 ? -v : v);
        if (a > largest) largest = a;
    }
    uint64_t a = (uint64_t)(opening < 0 ? -opening : opening);
    if (a > largest) largest = a;
    int64_t ge = p->gap_ext;
    a = (uint64_t)(ge < 0 ? -ge : ge);
    if (a > largest) largest = a;
    if (largest != 0 && steps > UINT64_MAX / largest) return 0;
    return largest * steps <= SCORE_LIMIT;
}

static void arena_reserve(size_t need) {
    if (need > t_arena_len) {
        if (need > t_arena_cap) {
            size_t c = t_arena_cap ? t_arena_cap : 1024;
            while (c < need) c *= 2;
            t_arena = xrealloc(t_arena, c * sizeof(int32_t));
            t_arena_cap = c;
        }
        for (size_t k = t_arena_len; k < need; k++) t_arena[k] = NEG32;
        t_arena_len = need;
    }
}

/* src/align.rs:587-911 */
int o_poa_simd_i32(const op_params *p, const og_graph *g, size_t beg_id, size_t end_id,
                   const uint8_t *q, size_t qlen, int adaptive, size_t band_b, ob_base *out,
                   size_t cap, size_t *out_len) {
    dpsetup s;
    if (dp_setup(p, g, beg_id, end_id, q, qlen, adaptive, band_b, &s)) return 1;
    size_t n = s.n, beg = s.beg, padded = s.padded;
    int32_t gap_ext = p->gap_ext;
    int32_t gap_open = p->gap_open + gap_ext;
    const v8 ext = splat8(gap_ext), open = splat8(gap_open);
    const v8 ext2 = splat8((int32_t)((uint32_t)gap_ext * 2u)), ext4 = splat8((int32_t)((uint32_t)gap_ext * 4u));
    const v8 steps = ext_steps8(gap_ext);
    size_t alpha = p_alpha(p);
    int32_t *profiles = xmalloc(alpha * padded * sizeof(int32_t));
    for (size_t a = 0; a < alpha; a++)
        for (size_t c = 0; c < padded; c++) profiles[a * padded + c] = o_score(p, a, s.seq[c]);
    omat x;
    memset(&x, 0, sizeof x);
    x.contains = f_contains;
    x.m = f_m;
    x.d = f_d;
    x.ins = f_i;
    x.negative = NEG32;
    x.f_start = xcalloc(n, sizeof(size_t));
    x.f_len = xcalloc(n, sizeof(size_t));
    x.f_m = xcalloc(n, sizeof(size_t));
    x.f_d = xcalloc(n, sizeof(size_t));
    x.f_i = xcalloc(n, sizeof(size_t));
    size_t cursor = 0;
    int rc = 0;
    obvec res = {0};
    for (size_t i = 0; i < n; i++) {
        if (s.head[i] == NEG || s.tail[i] == NEG) continue;
        size_t start, stop;
        if (row_band(p, &s, i, &start, &stop)) {
            rc = 1;
            goto out;
        }
        /* FastMatrix::allocate, src/align.rs:235-273: [8 pad][M][8 pad][D][I] */
        size_t len = stop - start;
        size_t size = len * 3 + WIDTH * 2;
        size_t rstop = cursor + size;
        arena_reserve(rstop);
        int32_t *data = t_arena;
        size_t mo = cursor + WIDTH, dof = cursor + WIDTH * 2 + len, io = cursor + WIDTH * 2 + len * 2;
        for (size_t k = cursor; k < dof + len; k++) data[k] = NEG32;
        cursor = rstop;
        x.f_start[i] = start;
        x.f_len[i] = len;
        x.f_m[i] = mo;
        x.f_d[i] = dof;
        x.f_i[i] = io;
        g_cells += len;
        if (i == 0) {
            if (start != 0) {
                rc = fail("alignment band excludes source");
                goto out;
            }
            data[mo] = o_score(p, s.unknown, s.unknown);
            data[io] = NEG32;
            for (size_t j = 1; j < len; j++) {
                int32_t ins = max32(wadd(data[io + j - 1], gap_ext), wadd(data[mo + j - 1], gap_open));
                data[io + j] = ins;
                data[mo + j] = max32(data[mo + j], ins);
            }
            continue;
        }
        const onode *node = &g->nodes[g->rank.v[beg + i]];
        for (size_t k = 0; k < node->incoming.n; k++) {
            size_t rank = g->nodes[node->incoming.v[k]].rank;
            if (rank < beg || rank >= beg + i) continue;
            size_t pr = rank - beg;
            if (s.head[pr] == NEG) continue;
            size_t ps = x.f_start[pr], pl = x.f_len[pr], pm = x.f_m[pr], pd = x.f_d[pr];
            size_t ov_start = start > ps ? start : ps;
            size_t ov_stop = stop < ps + pl ? stop : ps + pl;
            for (size_t j = ov_start; j < ov_stop; j += WIDTH) {
                size_t local = j - start, prior = j - ps;
                v8 diagonal = ld8(data + pm + prior - 1);
                v8 prev_m = ld8(data + pm + prior);
                v8 prev_d = ld8(data + pd + prior);
                v8 mm = max8(ld8(data + mo + local), diagonal);
                v8 dd = max8(max8(ld8(data + dof + local), add8(prev_d, ext)), add8(prev_m, open));
                st8(data + mo + local, mm);
                st8(data + dof + local, dd);
            }
        }
        size_t base = i == n - 1 ? s.unknown : node->base;
        const int32_t *profile = profiles + base * padded;
        int32_t carry = NEG32;
        for (size_t local = 0; local < len; local += WIDTH) {
            v8 diagonal = add8(ld8(data + mo + local), ld8(profile + start + local));
            v8 md = max8(diagonal, ld8(data + dof + local));
            v8 next = insertion_prefix8(md, carry, open, ext, ext2, ext4, steps);
            /* current = [carry, next0 .. next6] */
            v8 current = _mm256_permutevar8x32_epi32(next, _mm256_setr_epi32(0, 0, 1, 2, 3, 4, 5, 6));
            current = _mm256_blend_epi32(current, splat8(carry), 0x01);
            st8(data + io + local, current);
            st8(data + mo + local, max8(md, current));
            carry = _mm256_extract_epi32(next, 7);
        }
        if (p->f != 0 && s.adaptive) {
            /* fold over whole 8-lane chunks from splat(NEG32), then the first
             * position holding the maximum (rows are multiples of eight). */
            v8 bv = splat8(NEG32);
            for (size_t c = 0; c + WIDTH <= len; c += WIDTH) bv = max8(bv, ld8(data + mo + c));
            int32_t lanes[8];
            st8(lanes, bv);
            int32_t best = lanes[0];
            for (int l = 1; l < 8; l++) best = max32(best, lanes[l]);
            size_t argmax = 0;
            while (argmax + 1 < len && data[mo + argmax] != best) argmax++;
            adaptive_update(g, &s, i, start, (int64_t)(start + argmax));
        }
    }
    x.data = t_arena;
    g_last_score = f_m(&x, n - 1, s.m - 1);
    rc = traceback(p, g, &s, &x, gap_open, gap_ext, &res);
    if (!rc) emit(&res, out, cap, out_len);
out:
    free(profiles);
    free(x.f_start);
    free(x.f_len);
    free(x.f_m);
    free(x.f_d);
    free(x.f_i);
    free(res.v);
    setup_free(&s);
    return rc;
}

/* ---------------- Dispatch and driver (src/align.rs:81-211, lib.rs) ----- */

/* src/align.rs:531-545 */
int o_poa_inner(const op_params *p, const og_graph *g, size_t beg_id, size_t end_id,
                const uint8_t *q, size_t qlen, int adaptive, size_t band_b, ob_base *out,
                size_t cap, size_t *out_len) {
    if (o_i32_scores_safe(p, g, beg_id, end_id, qlen))
        return o_poa_simd_i32(p, g, beg_id, end_id, q, qlen, adaptive, band_b, out, cap, out_len);
    return o_poa_scalar_i64(p, g, beg_id, end_id, q, qlen, adaptive, band_b, out, cap, out_len);
}

/* src/align.rs:192-211 */
int o_poa(const op_params *p, const og_graph *g, size_t beg_id, size_t end_id,
          const uint8_t *q, size_t qlen, int adaptive, size_t band_b, ob_base *out, size_t cap,
          size_t *out_len) {
    int rc = o_poa_inner(p, g, beg_id, end_id, q, qlen, adaptive, band_b, out, cap, out_len);
    if (rc && p->f != 0) {
        op_params full = *p;
        full.f = 0;
        rc = o_poa_inner(&full, g, beg_id, end_id, q, qlen, 0, band_b, out, cap, out_len);
    }
    return rc;
}

/* ---------------- Minimizer seeding (src/minimizer.rs) ------------------
 *
 * Parity unpinned for minimizer POSITIONS: the reference takes them from the
 * simd-minimizers 3.0.0 crate (not vendored), which ranks k-mers by a
 * forward ntHash truncated inside its SIMD scan and keeps the leftmost of
 * equal keys (src/minimizer.rs:1-6, 39-76). This restatement ranks each
 * k-mer by the upper 16 bits of a 32-bit forward ntHash whose per-base seeds
 * are the published 64-bit ntHash seeds truncated to their low 32 bits, and
 * keeps the leftmost minimum of each window of w k-mers. The k-mer VALUES,
 * the run splitting, the window clamp and everything downstream (repeat
 * filter, anchors, chaining, windows) follow the reference exactly.
 */

typedef struct {
    uint64_t value;
    size_t end, span;
} o_mm;

typedef struct {
    o_mm *v;
    size_t n, cap;
} mmvec;

static void mm_push(mmvec *a, o_mm x) {
    if (a->n == a->cap) {
        a->cap = a->cap ? 2 * a->cap : 64;
        a->v = xrealloc(a->v, a->cap * sizeof(o_mm));
    }
    a->v[a->n++] = x;
}

typedef struct {
    o_anchor *v;
    size_t n, cap;
} anvec;

static void an_push(anvec *a, o_anchor x) {
    if (a->n == a->cap) {
        a->cap = a->cap ? 2 * a->cap : 64;
        a->v = xrealloc(a->v, a->cap * sizeof(o_anchor));
    }
    a->v[a->n++] = x;
}

/* src/minimizer.rs:28-30 */
static int is_dna(uint8_t c) {
    c = (uint8_t)toupper(c);
    return c == 'A' || c == 'C' || c == 'G' || c == 'T' || c == 'U';
}

/* 2-bit code of packed-seq: (base >> 1) & 3 (A 0, C 1, T/U 2, G 3). */
static inline uint32_t pack2(uint8_t c) { return (uint32_t)(c >> 1) & 3u; }

static inline uint32_t rotl32(uint32_t x, unsigned r) {
    r &= 31u;
    return r ? (x << r) | (x >> (32u - r)) : x;
}

static const uint32_t kNtSeed[4] = {0x95c60474u, 0x62a02b4cu, 0x4be24456u, 0x82572324u};

/* Minimizers of one run of DNA (len >= k) at offset `base_pos`. */
static void sketch_run(const uint8_t *s, size_t len, size_t k, size_t w, size_t base_pos,
                       mmvec *out) {
    size_t nk = len - k + 1;
    uint32_t *key = xmalloc(nk * sizeof(uint32_t));
    uint32_t h = 0;
    for (size_t i = 0; i < k; i++) h ^= rotl32(kNtSeed[pack2(s[i])], (unsigned)(k - 1 - i));
    key[0] = h >> 16;
    for (size_t t = 1; t < nk; t++) {
        h = rotl32(h, 1) ^ rotl32(kNtSeed[pack2(s[t - 1])], (unsigned)k) ^ kNtSeed[pack2(s[t + k - 1])];
        key[t] = h >> 16;
    }
    /* Sliding leftmost minimum over windows [st, st + w), monotone deque. */
    size_t *dq = xmalloc(nk * sizeof(size_t));
    size_t head = 0, tail = 0;
    size_t last = (size_t)-1;
    for (size_t t = 0; t < nk; t++) {
        while (tail > head && key[dq[tail - 1]] > key[t]) tail--;
        dq[tail++] = t;
        if (t + 1 < w) continue;
        size_t st = t + 1 - w;
        while (dq[head] < st) head++;
        size_t pos = dq[head];
        if (pos == last) continue;
        last = pos;
        uint64_t v = 0;
        for (size_t i = 0; i < k; i++) v |= (uint64_t)pack2(s[pos + i]) << (2 * i);
        o_mm m = {v, base_pos + pos + k - 1, k};
        mm_push(out, m);
    }
    free(dq);
    free(key);
}

/* src/minimizer.rs:32-76 */
static void sketch_into(const uint8_t *seq, size_t len, size_t k, size_t w, mmvec *out) {
    out->n = 0;
    if (k < 1 || k > 31 || w < 1 || w >= 32768 || len >= (size_t)UINT32_MAX) return;
    size_t start = 0;
    while (start < len) {
        while (start < len && !is_dna(seq[start])) start++;
        size_t end = start;
        while (end < len && is_dna(seq[end])) end++;
        if (end - start >= k) {
            size_t run = end - start;
            size_t win = w < run - k + 1 ? w : run - k + 1;
            sketch_run(seq + start, run, k, win, start, out);
        }
        start = end + 1;
    }
}

size_t o_sketch(const uint8_t *seq, size_t len, size_t k, size_t w, uint64_t *value,
                uint64_t *end, size_t cap) {
    mmvec mm = {0};
    sketch_into(seq, len, k, w, &mm);
    for (size_t i = 0; i < mm.n && i < cap; i++) {
        value[i] = mm.v[i].value;
        end[i] = mm.v[i].end;
    }
    size_t n = mm.n;
    free(mm.v);
    return n;
}

typedef struct {
    size_t start, end;
} o_iv;

static int iv_cmp(const void *a, const void *b) {
    const o_iv *x = a, *y = b;
    if (x->start != y->start) return x->start < y->start ? -1 : 1;
    if (x->end != y->end) return x->end < y->end ? -1 : 1;
    return 0;
}

/* Sort order of minimizer indices by (value, index). */
/* Per-thread: the job-parallel batch driver sorts on several threads. */
static __thread const o_mm *g_sort_mm;
static int mmidx_cmp(const void *a, const void *b) {
    size_t i = *(const size_t *)a, j = *(const size_t *)b;
    uint64_t vi = g_sort_mm[i].value, vj = g_sort_mm[j].value;
    if (vi != vj) return vi < vj ? -1 : 1;
    return i < j ? -1 : (i > j ? 1 : 0);
}

static size_t *mm_order(const o_mm *mm, size_t n) {
    size_t *ord = xmalloc((n ? n : 1) * sizeof(size_t));
    for (size_t i = 0; i < n; i++) ord[i] = i;
    g_sort_mm = mm;
    qsort(ord, n, sizeof(size_t), mmidx_cmp);
    return ord;
}

static inline size_t sat_sub(size_t a, size_t b) { return a > b ? a - b : 0; }

/* src/minimizer.rs:78-110; returns the merged intervals (malloc'd). */
static o_iv *repeat_intervals(const o_mm *mm, size_t n, size_t len, size_t max_dist,
                              size_t *out_n) {
    *out_n = 0;
    if (max_dist == 0) return NULL;
    /* prev[i]: the latest earlier minimizer with the same value (the
     * HashMap `last` of the reference, read before its insert). */
    size_t *ord = mm_order(mm, n);
    size_t *prev = xmalloc((n ? n : 1) * sizeof(size_t));
    for (size_t r = 0; r < n; r++) {
        size_t i = ord[r];
        prev[i] = (r > 0 && mm[ord[r - 1]].value == mm[i].value) ? ord[r - 1] : (size_t)-1;
    }
    o_iv *iv = xmalloc((n ? n : 1) * sizeof(o_iv));
    size_t ni = 0;
    for (size_t i = 0; i < n; i++) {
        if (prev[i] == (size_t)-1) continue;
        size_t previous = mm[prev[i]].end;
        size_t period = sat_sub(mm[i].end, previous);
        if (period > 0 && period <= max_dist) {
            size_t start = sat_sub(sat_sub(previous + 1, mm[i].span), period);
            size_t end = mm[i].end + 1 + period;
            if (end > len) end = len;
            iv[ni].start = start;
            iv[ni].end = end;
            ni++;
        }
    }
    qsort(iv, ni, sizeof(o_iv), iv_cmp);
    size_t nm = 0;
    for (size_t k = 0; k < ni; k++) {
        if (nm > 0 && iv[k].start <= iv[nm - 1].end + 50) {
            if (iv[k].end > iv[nm - 1].end) iv[nm - 1].end = iv[k].end;
            continue;
        }
        iv[nm++] = iv[k];
    }
    free(prev);
    free(ord);
    *out_n = nm;
    return iv;
}

size_t o_repeat_intervals(const uint64_t *value, const uint64_t *end, const uint64_t *span,
                          size_t n, size_t len, size_t max_dist, uint64_t *out_start,
                          uint64_t *out_end, size_t cap) {
    o_mm *mm = xmalloc((n ? n : 1) * sizeof(o_mm));
    for (size_t i = 0; i < n; i++) {
        mm[i].value = value[i];
        mm[i].end = end[i];
        mm[i].span = span[i];
    }
    size_t ni = 0;
    o_iv *iv = repeat_intervals(mm, n, len, max_dist, &ni);
    for (size_t k = 0; k < ni && k < cap; k++) {
        out_start[k] = iv[k].start;
        out_end[k] = iv[k].end;
    }
    free(iv);
    free(mm);
    return ni;
}

/* src/minimizer.rs:112-126: prefix counts of repeat-covered positions. */
static size_t *repeat_coverage(const o_mm *mm, size_t n, size_t len, size_t max_dist) {
    size_t *prefix = xcalloc(len + 1, sizeof(size_t));
    size_t ni = 0;
    o_iv *iv = repeat_intervals(mm, n, len, max_dist, &ni);
    for (size_t k = 0; k < ni; k++)
        for (size_t x = iv[k].start + 1; x <= iv[k].end; x++) prefix[x] = 1;
    for (size_t i = 0; i < len; i++) prefix[i + 1] += prefix[i];
    free(iv);
    return prefix;
}

static int overlaps_repeat(const o_mm *m, const size_t *prefix) {
    return prefix[m->end + 1] > prefix[m->end + 1 - m->span];
}

static int anchor_cmp(const void *a, const void *b) {
    const o_anchor *x = a, *y = b;
    if (x->target_end != y->target_end) return x->target_end < y->target_end ? -1 : 1;
    if (x->query_end != y->query_end) return x->query_end < y->query_end ? -1 : 1;
    return 0;
}

/* src/minimizer.rs:128-176 */
static void collect_anchors(const op_params *p, const uint8_t *q, size_t qlen,
                            const uint8_t *t, size_t tlen, anvec *out) {
    out->n = 0;
    if (p->alphabet_len > 5) return;
    mmvec qm = {0}, tm = {0};
    sketch_into(q, qlen, p->k, p->w, &qm);
    sketch_into(t, tlen, p->k, p->w, &tm);
    size_t *qcov = repeat_coverage(qm.v, qm.n, qlen, p->tr_dist);
    size_t *tcov = repeat_coverage(tm.v, tm.n, tlen, p->tr_dist);
    size_t *qo = mm_order(qm.v, qm.n), *to = mm_order(tm.v, tm.n);
    size_t a = 0, b = 0;
    while (a < tm.n && b < qm.n) {
        uint64_t tv = tm.v[to[a]].value, qv = qm.v[qo[b]].value;
        if (tv < qv) {
            a++;
            continue;
        }
        if (qv < tv) {
            b++;
            continue;
        }
        size_t a1 = a, b1 = b;
        while (a1 < tm.n && tm.v[to[a1]].value == tv) a1++;
        while (b1 < qm.n && qm.v[qo[b1]].value == tv) b1++;
        int keep = !(p->max_occ > 0 && (a1 - a > p->max_occ || b1 - b > p->max_occ));
        for (size_t x = a; keep && x < a1; x++) {
            const o_mm *tt = &tm.v[to[x]];
            if (overlaps_repeat(tt, tcov)) continue;
            for (size_t y = b; y < b1; y++) {
                const o_mm *qq = &qm.v[qo[y]];
                if (overlaps_repeat(qq, qcov)) continue;
                o_anchor an = {tt->end, qq->end, p->k, 0};
                an_push(out, an);
            }
        }
        a = a1;
        b = b1;
    }
    qsort(out->v, out->n, sizeof(o_anchor), anchor_cmp);
    size_t w = 0;
    for (size_t i = 0; i < out->n; i++)
        if (w == 0 || anchor_cmp(&out->v[w - 1], &out->v[i]) != 0) out->v[w++] = out->v[i];
    out->n = w;
    free(qo);
    free(to);
    free(qcov);
    free(tcov);
    free(qm.v);
    free(tm.v);
}

/* src/minimizer.rs:178-197; returns 0 for None. */
static int chain_score(const o_anchor *a, const o_anchor *b, size_t k, int32_t *score) {
    if (a->query_end < b->query_end || a->target_end < b->target_end) return 0;
    size_t dq = a->query_end - b->query_end, dr = a->target_end - b->target_end;
    if (dq == 0 || dq > 1000) return 0;
    size_t delta = dr > dq ? dr - dq : dq - dr;
    size_t distance = dr < dq ? dr : dq;
    int32_t s = (int32_t)(a->span < distance ? a->span : distance);
    if (delta > 0 || distance > a->span) {
        if (dr == 0) {
            s += 1;
        } else {
            double linear = 0.01 * (double)k * (double)delta;
            double logarithmic = log2((double)(delta + 1));
            s -= (int32_t)(linear < logarithmic ? linear : logarithmic);
        }
    }
    *score = s;
    return 1;
}

typedef struct {
    size_t *v;
    size_t n;
} o_local;

static __thread const o_anchor *g_sort_an;
static int local_cmp(const void *a, const void *b) {
    const o_local *x = a, *y = b;
    return anchor_cmp(&g_sort_an[x->v[x->n - 1]], &g_sort_an[y->v[y->n - 1]]);
}

typedef struct {
    int32_t score;
    size_t i;
} o_term;
static int term_cmp(const void *a, const void *b) {
    const o_term *x = a, *y = b;
    if (x->score != y->score) return x->score < y->score ? -1 : 1;
    return x->i < y->i ? -1 : (x->i > y->i ? 1 : 0);
}

/* src/minimizer.rs:199-377 */
static void chain(const o_anchor *an, size_t n, size_t k, anvec *out) {
    out->n = 0;
    if (n == 0) return;
    const size_t NONE = (size_t)-1;
    int32_t *scores = xcalloc(n, sizeof(int32_t));
    size_t *parent = xmalloc(n * sizeof(size_t));
    size_t *touched = xmalloc(n * sizeof(size_t));
    for (size_t i = 0; i < n; i++) parent[i] = touched[i] = NONE;
    size_t start = 0;
    for (size_t i = 0; i < n; i++) {
        while (start < i && sat_sub(an[i].target_end, an[start].target_end) > 1000) start++;
        if (start < sat_sub(i, 5000)) start = sat_sub(i, 5000);
        int32_t best = (int32_t)an[i].span;
        size_t skip = 0, iterations = 0;
        for (size_t j = i; j-- > start;) {
            int32_t step;
            if (!chain_score(&an[i], &an[j], k, &step)) continue;
            int32_t cand = step + scores[j];
            if (cand > best) {
                best = cand;
                parent[i] = j;
                skip = sat_sub(skip, 1);
            } else {
                if (cand != best) {
                    iterations++;
                    if (iterations > 50) break;
                }
                if (touched[j] == i) {
                    skip++;
                    if (skip > 25) break;
                }
            }
            if (parent[j] != NONE) touched[parent[j]] = i;
        }
        scores[i] = best;
    }
    uint8_t *has_child = xcalloc(n, 1);
    for (size_t i = 0; i < n; i++)
        if (parent[i] != NONE) has_child[parent[i]] = 1;
    o_term *terms = xmalloc(n * sizeof(o_term));
    size_t nt = 0;
    for (size_t i = 0; i < n; i++)
        if (!has_child[i] && scores[i] >= 100) {
            terms[nt].score = scores[i];
            terms[nt].i = i;
            nt++;
        }
    qsort(terms, nt, sizeof(o_term), term_cmp);
    uint8_t *used = xcalloc(n, 1);
    o_local *locals = xmalloc((nt ? nt : 1) * sizeof(o_local));
    size_t nl = 0;
    size_t *idx = xmalloc(n * sizeof(size_t));
    for (size_t r = nt; r-- > 0;) {
        size_t end = terms[r].i;
        if (used[end]) continue;
        size_t cur = end, ni = 0, best_cut = NONE;
        int32_t best_gain = 0;
        while (cur != NONE) {
            size_t i = cur;
            if (used[i]) break;
            idx[ni++] = i;
            cur = parent[i];
            int32_t gain = scores[end] - (cur == NONE ? 0 : scores[cur]);
            if (gain > best_gain) {
                best_gain = gain;
                best_cut = cur;
            } else if (best_gain - gain > 1000) {
                cur = best_cut;
                break;
            }
        }
        if (cur != NONE)
            while (ni > 0 && idx[ni - 1] <= cur) ni--;
        for (size_t x = 0; x < ni; x++) used[idx[x]] = 1;
        int32_t local_score = scores[end] - (cur == NONE ? 0 : scores[cur]);
        if (ni >= 3 && local_score >= 100) {
            o_local L;
            L.n = ni;
            L.v = xmalloc(ni * sizeof(size_t));
            for (size_t x = 0; x < ni; x++) L.v[x] = idx[ni - 1 - x];
            locals[nl++] = L;
        }
    }
    if (nl > 0) {
        g_sort_an = an;
        qsort(locals, nl, sizeof(o_local), local_cmp);
        int32_t *totals = xcalloc(nl, sizeof(int32_t));
        size_t *previous = xmalloc(nl * sizeof(size_t));
        size_t *trim = xcalloc(nl, sizeof(size_t));
        for (size_t i = 0; i < nl; i++) {
            previous[i] = NONE;
            const o_local *L = &locals[i];
            int32_t *prefix = xmalloc(L->n * sizeof(int32_t));
            prefix[0] = (int32_t)an[L->v[0]].span;
            for (size_t x = 1; x < L->n; x++) {
                int32_t s = 0;
                if (!chain_score(&an[L->v[x]], &an[L->v[x - 1]], k, &s)) s = 0;
                prefix[x] = prefix[x - 1] + s;
            }
            totals[i] = prefix[L->n - 1];
            const o_anchor *first = &an[L->v[0]];
            for (size_t j = i; j-- > 0;) {
                if (first->target_end < an[locals[j].v[0]].target_end) continue;
                const o_anchor *lastj = &an[locals[j].v[locals[j].n - 1]];
                size_t off = NONE;
                for (size_t x = 0; x < L->n; x++)
                    if (an[L->v[x]].target_end > lastj->target_end &&
                        an[L->v[x]].query_end > lastj->query_end) {
                        off = x;
                        break;
                    }
                if (off == NONE) continue;
                int32_t remaining = prefix[L->n - 1] - prefix[off] + (int32_t)an[L->v[off]].span;
                int32_t cand = totals[j] + remaining;
                if (cand > totals[i]) {
                    totals[i] = cand;
                    previous[i] = j;
                    trim[i] = off;
                }
            }
            free(prefix);
        }
        size_t best = 0;
        for (size_t i = 1; i < nl; i++)
            if (totals[i] > totals[best]) best = i;
        size_t *ids = xmalloc(nl * sizeof(size_t));
        size_t nc = 0;
        for (size_t i = best; i != NONE; i = previous[i]) ids[nc++] = i;
        for (size_t c = nc; c-- > 0;) {
            const o_local *L = &locals[ids[c]];
            for (size_t x = trim[ids[c]]; x < L->n; x++) {
                const o_anchor *a = &an[L->v[x]];
                if (out->n == 0 || (a->target_end > out->v[out->n - 1].target_end &&
                                    a->query_end > out->v[out->n - 1].query_end))
                    an_push(out, *a);
            }
        }
        free(ids);
        free(totals);
        free(previous);
        free(trim);
    }
    for (size_t i = 0; i < nl; i++) free(locals[i].v);
    free(locals);
    free(idx);
    free(used);
    free(terms);
    free(has_child);
    free(touched);
    free(parent);
    free(scores);
}

/* src/minimizer.rs:379-467 */
static void partition_windows(const o_anchor *ch, size_t n, size_t tlen, size_t qlen,
                              size_t max_drift, size_t max_win, anvec *out, int *tail) {
    out->n = 0;
    *tail = 0;
#define TE(i) ((i) < 0 ? (int64_t)-1 : (int64_t)ch[(i)].target_end)
#define QE(i) ((i) < 0 ? (int64_t)-1 : (int64_t)ch[(i)].query_end)
#define TS(i) ((size_t)(i) == n ? (int64_t)tlen : (int64_t)ch[(i)].target_end - (int64_t)ch[(i)].span + 1)
#define QS(i) ((size_t)(i) == n ? (int64_t)qlen : (int64_t)ch[(i)].query_end - (int64_t)ch[(i)].span + 1)
#define DIAG(i) ((i) < 0 ? (int64_t)0 : ((size_t)(i) == n ? (int64_t)tlen - (int64_t)qlen : TE(i) - QE(i)))
    int64_t s = -1;
    while (s < (int64_t)n) {
        int64_t boundary = -1;
        for (int64_t i = s + 1; i <= (int64_t)n; i++) {
            int64_t gt = TS(i) - TE(s) - 1;
            int64_t gq = QS(i) - QE(s) - 1;
            int64_t dd = DIAG(i) - DIAG(s);
            uint64_t drift = (uint64_t)(dd < 0 ? -dd : dd);
            int64_t g1 = TS(i) - TE(i - 1) - 1, g2 = QS(i) - QE(i - 1) - 1;
            int64_t gap = g1 > g2 ? g1 : g2;
            int64_t gm = gt > gq ? gt : gq;
            if (drift > (uint64_t)max_drift || gap > (int64_t)max_drift || gm > (int64_t)max_win)
                break;
            if (gt >= 0 && gq >= 0) boundary = i;
        }
        int adaptive;
        int64_t e;
        if (boundary >= 0) {
            adaptive = 0;
            e = boundary;
        } else {
            e = s + 1;
            while (e < (int64_t)n && (TS(e) - TE(s) - 1 < 0 || QS(e) - QE(s) - 1 < 0)) e++;
            int64_t dd = DIAG(e) - DIAG(s);
            uint64_t drift = (uint64_t)(dd < 0 ? -dd : dd);
            int64_t g1 = TS(e) - TE(s) - 1, g2 = QS(e) - QE(s) - 1;
            int64_t gap = g1 > g2 ? g1 : g2;
            adaptive = drift > (uint64_t)max_drift || gap > (int64_t)max_drift;
        }
        if (e < (int64_t)n) {
            o_anchor a = ch[e];
            a.adaptive_before = adaptive;
            an_push(out, a);
        } else {
            *tail = adaptive;
        }
        s = e;
    }
#undef TE
#undef QE
#undef TS
#undef QS
#undef DIAG
}

/* src/minimizer.rs:475-488 */
static void anchor_windows(const op_params *p, const uint8_t *q, size_t qlen,
                           const uint8_t *cons, size_t clen, anvec *out, int *tail) {
    anvec m = {0}, c = {0};
    collect_anchors(p, q, qlen, cons, clen, &m);
    chain(m.v, m.n, p->k, &c);
    size_t mw = qlen / 25 > 500 ? qlen / 25 : 500;
    if (p->win_max < mw) mw = p->win_max;
    partition_windows(c.v, c.n, clen, qlen, p->win_band, mw, out, tail);
    free(m.v);
    free(c.v);
}

static size_t copy_anchors(const anvec *a, o_anchor *out, size_t cap) {
    for (size_t i = 0; i < a->n && i < cap; i++) out[i] = a->v[i];
    return a->n;
}

size_t o_collect_anchors(const op_params *p, const uint8_t *q, size_t qlen, const uint8_t *t,
                         size_t tlen, o_anchor *out, size_t cap) {
    anvec a = {0};
    collect_anchors(p, q, qlen, t, tlen, &a);
    size_t n = copy_anchors(&a, out, cap);
    free(a.v);
    return n;
}

size_t o_chain(const o_anchor *anchors, size_t n, size_t k, o_anchor *out, size_t cap) {
    anvec a = {0};
    chain(anchors, n, k, &a);
    size_t r = copy_anchors(&a, out, cap);
    free(a.v);
    return r;
}

size_t o_partition_windows(const o_anchor *chain_in, size_t n, size_t tlen, size_t qlen,
                           size_t max_drift, size_t max_win, o_anchor *out, size_t cap,
                           int *tail) {
    anvec a = {0};
    partition_windows(chain_in, n, tlen, qlen, max_drift, max_win, &a, tail);
    size_t r = copy_anchors(&a, out, cap);
    free(a.v);
    return r;
}

size_t o_anchor_windows(const op_params *p, const uint8_t *q, size_t qlen, const uint8_t *t,
                        size_t tlen, o_anchor *out, size_t cap, int *tail) {
    anvec a = {0};
    anchor_windows(p, q, qlen, t, tlen, &a, tail);
    size_t r = copy_anchors(&a, out, cap);
    free(a.v);
    return r;
}

/* src/minimizer.rs:489-578 guide_order: minimizer multiset Jaccard
 * similarity weighted 0.51/0.49 by length, the most similar ordered pair
 * first, then the read with the largest summed similarity to those placed
 * (ties to the lower index). Inherits the sketch's unpinned positions. */
typedef struct {
    uint64_t value;
    size_t rid;
} o_vr;

static int vr_cmp(const void *a, const void *b) {
    const o_vr *x = a, *y = b;
    if (x->value != y->value) return x->value < y->value ? -1 : 1;
    return x->rid < y->rid ? -1 : (x->rid > y->rid ? 1 : 0);
}

void o_guide_order(const op_params *p, const uint8_t *const *seqs, const size_t *lens, size_t n,
                   size_t *order) {
    for (size_t i = 0; i < n; i++) order[i] = i;
    if (!p->progressive || n < 2 || p->alphabet_len > 5) return;
    /* (value, read) pairs sorted: each value's group lists its reads in
     * increasing id, a read repeated once per occurrence. */
    o_vr *vr = NULL;
    size_t nvr = 0, cvr = 0;
    mmvec mm = {0};
    for (size_t r = 0; r < n; r++) {
        sketch_into(seqs[r], lens[r], p->k, p->w, &mm);
        for (size_t i = 0; i < mm.n; i++) {
            if (nvr == cvr) {
                cvr = cvr ? 2 * cvr : 1024;
                vr = xrealloc(vr, cvr * sizeof(o_vr));
            }
            vr[nvr].value = mm.v[i].value;
            vr[nvr].rid = r;
            nvr++;
        }
    }
    free(mm.v);
    if (nvr == 0) {
        free(vr);
        return;
    }
    qsort(vr, nvr, sizeof(o_vr), vr_cmp);
    size_t *total = xcalloc(n, sizeof(size_t));
    size_t *shared = xcalloc(n * n, sizeof(size_t));
    size_t *rid = xmalloc(n * sizeof(size_t)), *cnt = xmalloc(n * sizeof(size_t));
    for (size_t a = 0; a < nvr;) {
        size_t b = a;
        while (b < nvr && vr[b].value == vr[a].value) b++;
        /* distinct reads of the group with their counts, increasing id */
        size_t nd = 0;
        for (size_t x = a; x < b; x++) {
            if (nd > 0 && rid[nd - 1] == vr[x].rid) {
                cnt[nd - 1]++;
            } else {
                rid[nd] = vr[x].rid;
                cnt[nd] = 1;
                nd++;
            }
        }
        if (!((double)nd > (double)n * p->mm_filter_ratio)) {
            for (size_t x = 0; x < nd; x++) {
                total[rid[x]] += cnt[x];
                for (size_t y = 0; y < x; y++) {
                    size_t m = cnt[x] < cnt[y] ? cnt[x] : cnt[y];
                    shared[rid[x] * n + rid[y]] += m;
                    shared[rid[y] * n + rid[x]] += m;
                }
            }
        }
        a = b;
    }
    size_t maxl = 1;
    for (size_t i = 0; i < n; i++)
        if (lens[i] > maxl) maxl = lens[i];
    double max_len = (double)maxl;
    double *sim = xcalloc(n * n, sizeof(double));
    double best = -1.0;
    size_t bi = 0, bj = 1;
    for (size_t i = 0; i < n; i++)
        for (size_t j = 0; j < n; j++) {
            if (i == j) continue;
            size_t common = shared[i * n + j];
            size_t uni = total[i] + total[j] - common;
            double jac = uni == 0 ? 0.0
                                  : (double)common / (double)uni *
                                        (0.51 * (double)lens[i] / max_len + 0.49 * (double)lens[j] / max_len);
            sim[i * n + j] = jac;
            if (jac > best) {
                best = jac;
                bi = i;
                bj = j;
            }
        }
    uint8_t *used = xcalloc(n, 1);
    double *sums = xmalloc(n * sizeof(double));
    order[0] = bi;
    order[1] = bj;
    used[bi] = used[bj] = 1;
    for (size_t v = 0; v < n; v++) sums[v] = sim[bi * n + v] + sim[bj * n + v];
    for (size_t k = 2; k < n; k++) {
        size_t next = (size_t)-1;
        for (size_t v = 0; v < n; v++)
            if (!used[v] && (next == (size_t)-1 || sums[v] > sums[next])) next = v;
        order[k] = next;
        used[next] = 1;
        for (size_t v = 0; v < n; v++)
            if (!used[v]) sums[v] += sim[next * n + v];
    }
    free(sums);
    free(used);
    free(sim);
    free(rid);
    free(cnt);
    free(shared);
    free(total);
    free(vr);
}

/* One seeded window (src/align.rs:104-111) and its outcome. */
typedef struct {
    size_t beg, end, qbeg, qend;
    int adaptive;
    const uint64_t *anchor;
    size_t nanchor;
    int rc;
    char err[512];
    obvec path;
    uint64_t cells;
} o_window;

typedef struct {
    const op_params *p;
    const og_graph *g;
    const uint8_t *q;
    o_window *w;
    size_t n;
    atomic_size_t next;
} o_winpool;

/* src/align.rs:162-181 `process`: poa over the window, then its anchors. */
static void process_window(const op_params *p, const og_graph *g, const uint8_t *q, o_window *w) {
    size_t qn = w->qend - w->qbeg;
    size_t cap = qn + 8;
    ob_base *buf = xmalloc(cap * sizeof(ob_base));
    size_t len = 0;
    uint64_t c0 = g_cells;
    w->rc = o_poa(p, g, w->beg, w->end, q + w->qbeg, qn, w->adaptive,
                  w->adaptive ? p->b : p->win_band, buf, cap, &len);
    w->cells = g_cells - c0;
    if (w->rc == 0 && len > cap) w->rc = fail("oracle path buffer overflow");
    if (w->rc) {
        snprintf(w->err, sizeof w->err, "%s", g_err);
    } else {
        for (size_t k = 0; k < len; k++) ob_push(&w->path, buf[k].node, buf[k].base, buf[k].aligned);
        for (size_t k = 0; k < w->nanchor; k++)
            ob_push(&w->path, (int64_t)w->anchor[k], (int64_t)g->nodes[w->anchor[k]].base, -1);
    }
    free(buf);
}

static void *window_worker(void *arg) {
    o_winpool *wp = arg;
    for (;;) {
        size_t i = atomic_fetch_add(&wp->next, 1);
        if (i >= wp->n) break;
        process_window(wp->p, wp->g, wp->q, &wp->w[i]);
    }
    o_arena_clear();
    return NULL;
}

/* src/align.rs:99-188: windows between chained anchors, each aligned by
 * poa(beg, end) and followed by the anchor's own graph nodes. With
 * p->threads > 1 the windows run on that many threads, results in window
 * order and the first error in window order (src/parallel.rs:46-67). */
static int seeded_align(const op_params *p, const og_graph *g, const uint8_t *q, size_t qlen,
                        ob_base *out, size_t cap, size_t *out_len) {
    size_t nn = g->n_nodes;
    uint8_t *cons = xmalloc(nn + 1);
    uint64_t *ids = xmalloc((nn + 1) * sizeof(uint64_t));
    size_t clen = og_consensus(g, p, cons, ids, NULL, nn);
    anvec an = {0};
    int tail = 0;
    anchor_windows(p, q, qlen, cons, clen, &an, &tail);
    int rc = 0;
    o_window *w = NULL;
    size_t nw = 0;
    if (an.n == 0) {
        rc = o_poa(p, g, 0, 1, q, qlen, 1, p->b, out, cap, out_len);
        goto done;
    }
    /* Validate every anchor first: any invalid one means the whole query
     * takes the full-graph fallback. */
    size_t qbeg = 0, prev_te = (size_t)-1;
    for (size_t a = 0; a < an.n; a++) {
        const o_anchor *x = &an.v[a];
        if (x->target_end + 1 < x->span) {
            rc = fail("invalid minimizer target span");
            goto done;
        }
        if (x->query_end + 1 < x->span) {
            rc = fail("invalid minimizer query span");
            goto done;
        }
        size_t ts = x->target_end + 1 - x->span, qs = x->query_end + 1 - x->span;
        int valid = x->span > 0 && x->target_end < clen && x->query_end < qlen && qs >= qbeg &&
                    (prev_te == (size_t)-1 || ts > prev_te);
        for (size_t o = 0; valid && o < x->span; o++)
            if (o_encode(p, cons[ts + o]) != o_encode(p, q[qs + o])) valid = 0;
        if (!valid) {
            rc = o_poa(p, g, 0, 1, q, qlen, 1, p->b, out, cap, out_len);
            goto done;
        }
        qbeg = x->query_end + 1;
        prev_te = x->target_end;
    }
    nw = an.n + 1;
    w = xcalloc(nw, sizeof(o_window));
    {
        size_t beg = 0;
        qbeg = 0;
        for (size_t a = 0; a < nw; a++) {
            o_window *x = &w[a];
            x->beg = beg;
            x->qbeg = qbeg;
            if (a < an.n) {
                const o_anchor *y = &an.v[a];
                size_t ts = y->target_end + 1 - y->span;
                x->end = ids[ts];
                x->qend = y->query_end + 1 - y->span;
                x->adaptive = (int)y->adaptive_before;
                x->anchor = ids + ts;
                x->nanchor = y->target_end - ts;
                beg = ids[y->target_end];
                qbeg = y->query_end + 1;
            } else {
                x->end = 1;
                x->qend = qlen;
                x->adaptive = tail;
            }
        }
    }
    {
        size_t workers = p->threads > 1 ? (p->threads < nw ? p->threads : nw) : 1;
        if (workers <= 1) {
            /* Sequential map stops at the first error. */
            for (size_t a = 0; a < nw; a++) {
                process_window(p, g, q, &w[a]);
                if (w[a].rc) break;
            }
        } else {
            o_winpool wp;
            wp.p = p;
            wp.g = g;
            wp.q = q;
            wp.w = w;
            wp.n = nw;
            atomic_init(&wp.next, 0);
            pthread_t *th = xmalloc(workers * sizeof(pthread_t));
            for (size_t t = 0; t < workers; t++) pthread_create(&th[t], NULL, window_worker, &wp);
            for (size_t t = 0; t < workers; t++) pthread_join(th[t], NULL);
            free(th);
            for (size_t a = 0; a < nw; a++) g_cells += w[a].cells;
        }
        obvec res = {0};
        for (size_t a = 0; a < nw && rc == 0; a++) {
            if (w[a].rc) {
                rc = fail("%s", w[a].err);
                break;
            }
            for (size_t k = 0; k < w[a].path.n; k++)
                ob_push(&res, w[a].path.v[k].node, w[a].path.v[k].base, w[a].path.v[k].aligned);
        }
        if (rc == 0) rc = emit(&res, out, cap, out_len);
        free(res.v);
    }
done:
    for (size_t a = 0; a < nw; a++) free(w[a].path.v);
    free(w);
    free(an.v);
    free(cons);
    free(ids);
    return rc;
}

/* src/align.rs:81-188 (empty graph, unseeded and seeded branches). */
int o_align(const op_params *p, const og_graph *g, const uint8_t *q, size_t qlen, ob_base *out,
            size_t cap, size_t *out_len) {
    if (g->n_nodes <= 2) {
        *out_len = qlen + 1;
        if (cap > 0) {
            out[0].node = 0;
            out[0].base = (int64_t)o_encode(p, 'N');
            out[0].aligned = -1;
        }
        for (size_t k = 0; k < qlen && k + 1 < cap; k++) {
            out[k + 1].node = -1;
            out[k + 1].base = (int64_t)o_encode(p, q[k]);
            out[k + 1].aligned = -1;
        }
        return 0;
    }
    if (!p->seeding) return o_poa(p, g, 0, 1, q, qlen, p->adaptive, p->b, out, cap, out_len);
    return seeded_align(p, g, q, qlen, out, cap, out_len);
}

/* src/lib.rs:184-204 (input order). */
int o_extend_graph(const op_params *p, og_graph *g, const uint8_t *const *seqs,
                   const size_t *lens, size_t nseq, size_t existing) {
    size_t *order = xmalloc((nseq ? nseq : 1) * sizeof(size_t));
    if (p->progressive && existing == 0) {
        o_guide_order(p, seqs, lens, nseq, order);
    } else {
        for (size_t i = 0; i < nseq; i++) order[i] = i;
    }
    size_t cap = 0;
    ob_base *buf = NULL;
    for (size_t index = existing; index < nseq; index++) {
        size_t id = order[index];
        size_t need = lens[id] + 8;
        if (need > cap) {
            cap = need * 2;
            buf = xrealloc(buf, cap * sizeof(ob_base));
        }
        size_t len = 0;
        if (o_align(p, g, seqs[id], lens[id], buf, cap, &len)) {
            free(buf);
            free(order);
            return 1;
        }
        if (len > cap) {
            free(buf);
            free(order);
            return fail("oracle path buffer overflow");
        }
        if (og_add_path(g, id, buf, len) || og_toposort(g, 0)) {
            free(buf);
            free(order);
            return 1;
        }
    }
    free(buf);
    free(order);
    return 0;
}

/* src/lib.rs:177-182 */
og_graph *o_align_sequences(const op_params *p, const uint8_t *const *seqs,
                            const size_t *lens, size_t nseq) {
    if (o_validate(p)) return NULL;
    og_graph *g = og_new(p);
    if (o_extend_graph(p, g, seqs, lens, nseq, 0)) {
        og_free(g);
        return NULL;
    }
    return g;
}

/* ---------------- Job-parallel CPU baseline driver ---------------------- */

static uint64_t fnv1a(uint64_t h, const void *data, size_t n) {
    const uint8_t *b = data;
    for (size_t k = 0; k < n; k++) {
        h ^= b[k];
        h *= UINT64_C(1099511628211);
    }
    return h;
}

typedef struct {
    const op_params *p;
    size_t njobs;
    const uint64_t *job_off, *read_off;
    const uint8_t *bases;
    uint64_t *out_hash;
    uint64_t *out_gfa; /* NULL: no GFA */
    atomic_size_t next;
    atomic_uint_fast64_t cells;
} batch_ctx;

static uint64_t run_job(const op_params *p, const batch_ctx *c, size_t job, uint64_t *gfa_hash) {
    size_t r0 = c->job_off[job], r1 = c->job_off[job + 1], nr = r1 - r0;
    const uint8_t **seqs = xmalloc((nr + 1) * sizeof(uint8_t *));
    size_t *lens = xmalloc((nr + 1) * sizeof(size_t));
    char **names = xmalloc((nr + 1) * sizeof(char *));
    for (size_t k = 0; k < nr; k++) {
        seqs[k] = c->bases + c->read_off[r0 + k];
        lens[k] = c->read_off[r0 + k + 1] - c->read_off[r0 + k];
        names[k] = xmalloc(32);
        snprintf(names[k], 32, "read%zu", k);
    }
    uint64_t h = 0;
    og_graph *g = o_align_sequences(p, seqs, lens, nr);
    if (g) {
        size_t total = 0;
        for (size_t k = 0; k < nr; k++) total += lens[k];
        uint8_t *cons = xmalloc(g->n_nodes + 1);
        size_t cl = og_consensus(g, p, cons, NULL, NULL, g->n_nodes);
        char *msa = NULL;
        if (og_msa(g, p, (const char *const *)names, NULL, nr, &msa) == 0) {
            h = UINT64_C(14695981039346656037);
            const char *hdr = ">consensus sequence\n";
            h = fnv1a(h, hdr, strlen(hdr));
            h = fnv1a(h, cons, cl);
            h = fnv1a(h, "\n", 1);
            h = fnv1a(h, msa, strlen(msa));
            if (h == 0) h = 1;
        }
        char *gfa = NULL;
        if (gfa_hash && og_gfa(g, p, (const char *const *)names, nr, &gfa) == 0) {
            uint64_t x = fnv1a(UINT64_C(14695981039346656037), gfa, strlen(gfa));
            *gfa_hash = x ? x : 1;
        }
        free(gfa);
        free(msa);
        free(cons);
        og_free(g);
        (void)total;
    }
    for (size_t k = 0; k < nr; k++) free(names[k]);
    free(names);
    free(seqs);
    free(lens);
    return h;
}

static void *batch_worker(void *arg) {
    batch_ctx *c = arg;
    g_cells = 0;
    for (;;) {
        size_t job = atomic_fetch_add(&c->next, 1);
        if (job >= c->njobs) break;
        c->out_hash[job] = run_job(c->p, c, job, c->out_gfa ? c->out_gfa + job : NULL);
    }
    atomic_fetch_add(&c->cells, g_cells);
    o_arena_clear();
    return NULL;
}

uint64_t o_batch_run(const op_params *p, size_t njobs, const uint64_t *job_off,
                     const uint64_t *read_off, const uint8_t *bases, int nthreads,
                     uint64_t *out_hash) {
    return o_batch_run_gfa(p, njobs, job_off, read_off, bases, nthreads, out_hash, NULL);
}

uint64_t o_batch_run_gfa(const op_params *p, size_t njobs, const uint64_t *job_off,
                         const uint64_t *read_off, const uint8_t *bases, int nthreads,
                         uint64_t *out_hash, uint64_t *out_gfa) {
    batch_ctx c;
    c.p = p;
    c.njobs = njobs;
    c.job_off = job_off;
    c.read_off = read_off;
    c.bases = bases;
    c.out_hash = out_hash;
    c.out_gfa = out_gfa;
    if (out_gfa) memset(out_gfa, 0, njobs * sizeof(uint64_t));
    atomic_init(&c.next, 0);
    atomic_init(&c.cells, 0);
    if (nthreads < 1) nthreads = 1;
    pthread_t *th = xmalloc((size_t)nthreads * sizeof(pthread_t));
    for (int t = 0; t < nthreads; t++) pthread_create(&th[t], NULL, batch_worker, &c);
    for (int t = 0; t < nthreads; t++) pthread_join(th[t], NULL);
    free(th);
    return (uint64_t)atomic_load(&c.cells);
}
