// Device-resident progressive POA jobs: each job's graph, its next DP's
// descriptors and its traceback results stay in device memory between
// reads, and one step of a job -- consume the last DP's traceback, fuse the
// path into the graph (src/graph.rs:82-122), re-rank it with the LIFO Kahn
// walk (src/graph.rs:124-178) and plan the next read's DP
// (src/align.rs:587-688) -- runs on the device, one wave per job, between
// the fill and traceback launches of the same captured step loop.
//
// The step is written once, as templates over a Team: a wave of 64 lanes on
// the device (resident.hip), or one host thread in the CPU check that holds
// it against the host Graph and planner (resident.cpp). Team-parallel loops
// visit items in chunks of team.size lanes and combine them with exclusive
// scans, so with one lane they are the reference's own serial loops; the
// parts the reference's ordering rules make serial (the Kahn walk, the
// head and tail passes) run as uniform code on every lane.
#pragma once
#include <cstdint>

#include "dp_types.hpp"
#include "minipoa_hip.h"

#if defined(__HIPCC__)
#define MPOA_HD __host__ __device__
#else
#define MPOA_HD
#endif

namespace mpoa {

constexpr uint32_t kNil = 0xFFFFFFFFu;

// One edge, linked into its tail's out-list and its head's in-list (both in
// insertion order, as the reference's adjacency vectors). The weight is the
// same in both lists (src/graph.rs:63-80 increments both).
struct REdge {
    uint32_t from, to, w, next_out, next_in;
};

// An in-range successor of a planned row (the tail pass), in out-list order.
struct RSucc {
    uint32_t row, w;
};

// Scoring parameters of a resident batch (src/lib.rs:47-71, 110-137).
struct RParams {
    uint8_t codes[256];   // encode() of every byte
    const int32_t *matrix;  // alpha x alpha
    uint64_t b, f;
    uint64_t largest;     // max |score|, |gap_open + gap_ext|, |gap_ext| (i32 guard)
    int32_t gap_open, gap_ext, score_nn;
    uint32_t alpha, unknown, adaptive, scalar_dp;
};

// Job states and the phases of a running job's current read.
enum RState : uint32_t { kRsRunning = 0, kRsDone = 1, kRsFailed = 2, kRsEvicted = 3 };
enum RPhase : uint32_t { kRpIdle = 0, kRpPlanned = 1, kRpLaunched = 2 };
// Why a job left the device (host engine continues it from read `cur`).
enum REvict : uint32_t {
    kEvNone = 0,
    kEvGuard = 1,        // the i32 score guard fails (the i64 kernel runs on the host path)
    kEvArena = 2,        // the DP's arena does not fit the batch's arena
    kEvUnreachable = 3,  // a row has a predecessor the head pass cannot reach
    kEvCapacity = 4,     // node, edge or coordinate capacity
    kEvStalled = 5,      // the step loop's replay bound was reached (host side)
};

// One resident job. Pointers are device addresses (host addresses in the
// CPU check). Capacities: nc nodes, ec edges (sum over reads of len + 1,
// the most edges its paths can add), maxm query columns.
struct RJob {
    const uint8_t *bases;      // the job's reads, concatenated
    const uint64_t *read_off;  // nreads + 1, relative to bases
    uint32_t nreads, first;    // first: sequence id of read 0
    uint32_t nc, ec, maxm, pad0_;
    // graph
    uint8_t *base;
    uint32_t *parent, *rank, *order;
    int32_t *aligned;  // nc x alpha
    uint32_t *out_head, *out_tail, *in_head, *in_tail, *in_cnt, *cover, *mark;
    REdge *edges;
    uint32_t *paths, *path_off;  // node ids (>= 2) of each read's path; nreads + 1 offsets
    // planning scratch in global memory (used when a graph outgrows LDS)
    uint32_t *csr_off, *csr_adj, *deg, *stack, *succ_beg;
    RSucc *succ;
    int32_t *head, *tail;
    uint32_t *support;
    uint32_t *win;  // the serial passes' staging windows when LDS is not used
    // the job's regions of the launch pools (DevJob offsets)
    uint64_t rows_off, aux_off, preds_off, seq_off, meta_off, adapt_off, tb_off;
    // state
    uint32_t nn, ne, cur, phase, state, retry, full_arena, epoch;
    int32_t err;       // MPOA status of a failed job
    int32_t dev_status;  // DevStatus of a failed DP (kOk when the failure was not a DP's)
    uint32_t err_i, err_j;
    uint32_t evict, beg;  // eviction reason; rank of the planned DP's first row
    uint32_t npath, pad1_;  // path ids written (path_off[cur])
    uint64_t cells, dp_calls, retries, arena_retries, deferred, pred_cells, arena_bytes;
    uint64_t prev_used, prev_cells, prev_n;  // the last completed DP (arena hint)
    uint64_t est, emax;  // arena bytes of the planned DP: estimate and full-width bound
    DevJob dj;           // the planned DP (arena fields set when it is placed)
};

// Kernel arguments of one resident batch's step.
struct RPools {
    RJob *jobs;
    DevJob *dj;
    DevResult *res;
    RowIn *rows;
    RowAux *aux;
    PredE *preds;
    uint8_t *seq;
    uint32_t *tb;
    unsigned long long *arena_ctr;  // bytes handed out this step (reset before each step)
    uint64_t alloc_limit;           // the step's placements stay below this; fill chunks above
    uint64_t job_cap;               // a DP needing more arena than this leaves the device
    uint32_t *done;                 // jobs that left the running state
    const RParams *prm;             // in device memory
    uint32_t njobs, lds_bytes;
};

// Arguments of the persistent resident kernel: a queue of job indices
// taken in order (the host sorts the jobs longest first), and one arena
// region of slot_bytes per workgroup.
struct RPersist {
    uint32_t *next;       // queue position: jobs taken so far
    uint64_t arena_off;   // byte offset of workgroup 0's region in the arena
    uint64_t slot_bytes;  // bytes per region (the largest DP arena a job may use)
    uint32_t step_lds;    // LDS bytes of the step (the rest of the kernel's LDS aliases it)
    uint32_t prio;        // bit 0: the step's wave, bit 1: the traceback's, at raised issue priority
    // Streaming (sbuf non-null): a workgroup packs each job it finishes into
    // sbuf at an offset taken from *scur (while it fits in scap bytes) and
    // then publishes sent[4k..4k+3] = {1, offset low, offset high, bytes},
    // so the host can copy and render finished jobs while the launch runs.
    uint8_t *sbuf;
    unsigned long long *scur;
    uint64_t scap;
    uint32_t *sent;
};

// ---------------------------------------------------------------------------
// Team primitives used below (device: resident.hip; host: resident.cpp):
//   lane, size               this lane and the team's width
//   sync()                   barrier with memory ordering across the team
//   excl_scan(v, total)      exclusive sum over lanes, total over lanes
//   reduce_add64(v), reduce_max(v), any(p)
//   xchg(p, v), cas(p, cmp, v), add(p, v), umax(p, v)   atomics
// ---------------------------------------------------------------------------

template <class T, class F>
MPOA_HD inline void team_for(T &t, uint32_t n, F &&f) {
    for (uint32_t k = t.lane; k < n; k += t.size) f(k);
}

// Items 0..n-1 in chunks of t.size: cnt(k) items each, placed at the
// running exclusive sum: put(k, offset). Returns the total. Item order is
// preserved, so one lane gives the serial order.
template <class T, class C, class P>
MPOA_HD inline uint32_t team_place(T &t, uint32_t n, C &&cnt, P &&put) {
    uint32_t carry = 0;
    for (uint32_t b0 = 0; b0 < n; b0 += t.size) {
        const uint32_t k = b0 + t.lane;
        const uint32_t v = k < n ? cnt(k) : 0u;
        uint32_t tot;
        const uint32_t ex = t.excl_scan(v, tot);
        if (k < n) put(k, carry + ex);
        carry += tot;
    }
    return carry;
}

// ---------------------------------------------------------------------------
// Graph (src/graph.rs:30-122)
// ---------------------------------------------------------------------------

MPOA_HD inline void rj_node_init(const RParams &P, RJob &J, uint32_t id, uint32_t b, uint32_t parent) {
    J.base[id] = uint8_t(b);
    J.parent[id] = parent;
    J.rank[id] = 0;
    for (uint32_t a = 0; a < P.alpha; a++) J.aligned[uint64_t(id) * P.alpha + a] = -1;
    J.aligned[uint64_t(id) * P.alpha + b] = int32_t(id);
    J.out_head[id] = J.out_tail[id] = J.in_head[id] = J.in_tail[id] = kNil;
    J.in_cnt[id] = 0;
    J.cover[id] = 0;
    J.mark[id] = 0;
}

// A new edge record at slot e, appended to both lists.
MPOA_HD inline void rj_edge_append(RJob &J, uint32_t e, uint32_t from, uint32_t to) {
    J.edges[e] = REdge{from, to, 1u, kNil, kNil};
    if (J.out_tail[from] == kNil) J.out_head[from] = e;
    else J.edges[J.out_tail[from]].next_out = e;
    J.out_tail[from] = e;
    if (J.in_tail[to] == kNil) J.in_head[to] = e;
    else J.edges[J.in_tail[to]].next_in = e;
    J.in_tail[to] = e;
    J.in_cnt[to] += 1;
}

// List walks stop after ne + 1 records: a list is never longer than the
// edge count, so the bound only matters if a list were ever corrupted.
MPOA_HD inline uint32_t rj_find_out(const RJob &J, uint32_t from, uint32_t to) {
    for (uint32_t e = J.out_head[from], g_ = 0; e != kNil && g_ <= J.ne; e = J.edges[e].next_out, g_++)
        if (J.edges[e].to == to) return e;
    return kNil;
}

// src/graph.rs:63-80 (serial form): self edges ignored, repeats add weight.
MPOA_HD inline void rj_add_edge_serial(RJob &J, uint32_t from, uint32_t to) {
    if (from == to) return;
    const uint32_t e = rj_find_out(J, from, to);
    if (e != kNil) J.edges[e].w += 1;
    else rj_edge_append(J, J.ne++, from, to);
    J.cover[to] += 1;
}

// Path elements: (node, base, aligned) as AlignedBase, -1 = None.
struct RElem {
    int64_t node, base, aligned;
};

// src/graph.rs:82-122, one element at a time: the reference's own loop
// (used when two elements of a path resolve to the same node).
template <class E>
MPOA_HD inline void rj_add_path_serial(const RParams &P, RJob &J, E &&elem, uint32_t L) {
    bool has_prev = false;
    uint32_t prev = 0, np = 0;
    const uint32_t po = J.path_off[J.cur];
    for (uint32_t k = 0; k < L; k++) {
        const RElem a = elem(k);
        const uint32_t b = uint32_t(a.base);
        uint32_t id;
        if (a.node >= 0) {
            id = uint32_t(a.node);
        } else if (a.aligned >= 0) {
            const uint32_t par = J.parent[uint32_t(a.aligned)];
            const int32_t col = J.aligned[uint64_t(par) * P.alpha + b];
            if (col >= 0) {
                id = uint32_t(col);
            } else {
                id = J.nn++;
                rj_node_init(P, J, id, b, par);
                J.aligned[uint64_t(par) * P.alpha + b] = int32_t(id);
            }
        } else {
            id = J.nn++;
            rj_node_init(P, J, id, b, id);
        }
        if (has_prev) rj_add_edge_serial(J, prev, id);
        prev = id;
        has_prev = true;
        if (id >= 2) J.paths[po + np++] = id;
    }
    rj_add_edge_serial(J, has_prev ? prev : 0u, 1u);
    J.path_off[J.cur + 1] = po + np;
    J.npath = po + np;
}

// src/graph.rs:82-122 across the team. Elements resolve in parallel; when
// their nodes are distinct (always, unless a path passes two members of
// one aligned column), every new node, weight increment and list append is
// independent: new ids follow path order (a scan), each list gains at most
// one edge, and no two lanes touch one record field. Otherwise the serial
// form runs. ids and kinds are scratch arrays of at least L entries.
enum : uint32_t { kElExist = 0, kElNewAligned = 1, kElNew = 2, kElClaimed = 4 };

template <class T, class E>
MPOA_HD inline void rj_add_path(T &t, const RParams &P, RJob &J, E &&elem, uint32_t L, uint32_t *ids,
                                uint32_t *kinds) {
    const uint32_t alpha = P.alpha;
    const uint32_t epoch = J.epoch + 1;
    t.sync();
    if (t.lane == 0) J.epoch = epoch;
    // Resolve: an existing node, a new member of column `parent` or a new node.
    team_for(t, L, [&](uint32_t k) {
        const RElem a = elem(k);
        const uint32_t b = uint32_t(a.base);
        if (a.node >= 0) {
            ids[k] = uint32_t(a.node);
            kinds[k] = kElExist;
        } else if (a.aligned >= 0) {
            const uint32_t par = J.parent[uint32_t(a.aligned)];
            const int32_t col = J.aligned[uint64_t(par) * alpha + b];
            if (col >= 0) {
                ids[k] = uint32_t(col);
                kinds[k] = kElExist;
            } else {
                ids[k] = par;
                kinds[k] = kElNewAligned | b << 8;
            }
        } else {
            ids[k] = 0;
            kinds[k] = kElNew | b << 8;
        }
    });
    t.sync();
    // Distinctness: existing ids stamp mark[] with this path's epoch; new
    // column members claim their (column, base) slot of aligned[].
    bool clash = false;
    team_for(t, L, [&](uint32_t k) {
        const uint32_t kd = kinds[k];
        if ((kd & 3u) == kElExist) {
            // The sink also ends the path's last edge: an element naming it
            // would give it two new in-edges.
            if (t.xchg(&J.mark[ids[k]], epoch) == epoch || ids[k] == 1u) clash = true;
        } else if ((kd & 3u) == kElNewAligned) {
            int32_t *slot = &J.aligned[uint64_t(ids[k]) * alpha + (kd >> 8)];
            if (t.cas(slot, -1, -2) == -1) kinds[k] = kd | kElClaimed;
            else clash = true;
        }
    });
    t.sync();
    if (t.any(clash)) {
        team_for(t, L, [&](uint32_t k) {
            if (kinds[k] & kElClaimed) J.aligned[uint64_t(ids[k]) * alpha + (kinds[k] >> 8)] = -1;
        });
        t.sync();
        if (t.lane == 0) rj_add_path_serial(P, J, elem, L);
        t.sync();
        return;
    }
    // New nodes, numbered in path order.
    const uint32_t nn0 = J.nn;
    const uint32_t nnew = team_place(
        t, L, [&](uint32_t k) { return (kinds[k] & 3u) != kElExist ? 1u : 0u; },
        [&](uint32_t k, uint32_t at) {
            const uint32_t kd = kinds[k];
            if ((kd & 3u) == kElExist) return;
            const uint32_t id = nn0 + at, b = kd >> 8;
            if ((kd & 3u) == kElNewAligned) {
                const uint32_t par = ids[k];
                rj_node_init(P, J, id, b, par);
                J.aligned[uint64_t(par) * alpha + b] = int32_t(id);
            } else {
                rj_node_init(P, J, id, b, id);
            }
            ids[k] = id;
        });
    t.sync();
    // Edges id[k-1] -> id[k], then the last -> sink: increments in place,
    // new records in path order.
    const uint32_t ne0 = J.ne;
    const uint32_t nedge = team_place(
        t, L,
        [&](uint32_t k) {
            const uint32_t from = ids[k], to = k + 1 < L ? ids[k + 1] : 1u;
            kinds[k] = 0;
            if (from == to) return 0u;
            const uint32_t e = rj_find_out(J, from, to);
            J.cover[to] += 1;
            if (e != kNil) {
                J.edges[e].w += 1;
                return 0u;
            }
            kinds[k] = 1;
            return 1u;
        },
        [&](uint32_t k, uint32_t at) {
            if (kinds[k]) rj_edge_append(J, ne0 + at, ids[k], k + 1 < L ? ids[k + 1] : 1u);
        });
    t.sync();
    // The read's path: its nodes other than source and sink, in order.
    const uint32_t po = J.path_off[J.cur];
    const uint32_t npath = team_place(
        t, L, [&](uint32_t k) { return ids[k] >= 2u ? 1u : 0u; },
        [&](uint32_t k, uint32_t at) {
            if (ids[k] >= 2u) J.paths[po + at] = ids[k];
        });
    t.sync();
    if (t.lane == 0) {
        J.nn = nn0 + nnew;
        J.ne = ne0 + nedge;
        J.path_off[J.cur + 1] = po + npath;
        J.npath = po + npath;
    }
    t.sync();
}

// ---------------------------------------------------------------------------
// toposort(false) (src/graph.rs:124-178): Kahn's walk with a LIFO stack
// seeded in id order. The out-lists are flattened into CSR arrays first
// (team-parallel), then the walk runs as uniform code. I: index type of the
// staged arrays (uint16_t in LDS for graphs below 65,536 nodes).
// Returns false on a cycle.
// ---------------------------------------------------------------------------
template <class T, class I>
MPOA_HD inline bool rj_toposort(T &t, RJob &J, uint32_t *off, I *adj, I *deg, I *stack) {
    const uint32_t nn = J.nn;
    team_place(
        t, nn,
        [&](uint32_t id) {
            uint32_t c = 0;
            for (uint32_t e = J.out_head[id], g_ = 0; e != kNil && g_ <= J.ne; e = J.edges[e].next_out, g_++) c++;
            return c;
        },
        [&](uint32_t id, uint32_t at) { off[id] = at; });
    t.sync();
    team_for(t, nn, [&](uint32_t id) {
        uint32_t p = off[id];
        for (uint32_t e = J.out_head[id], g_ = 0; e != kNil && g_ <= J.ne; e = J.edges[e].next_out, g_++) adj[p++] = I(J.edges[e].to);
        if (id + 1 == nn) off[nn] = p;
        deg[id] = I(J.in_cnt[id]);
    });
    t.sync();
    uint32_t sp = team_place(
        t, nn, [&](uint32_t id) { return deg[id] == 0 ? 1u : 0u; },
        [&](uint32_t id, uint32_t at) {
            if (deg[id] == 0) stack[at] = I(id);
        });
    t.sync();
    uint32_t r = 0;
    while (sp > 0) {
        const uint32_t id = stack[--sp];
        if (t.lane == 0) {
            J.rank[id] = r;
            J.order[r] = id;
        }
        r++;
        const uint32_t e1 = off[id + 1];
        for (uint32_t k = off[id]; k < e1; k++) {
            const uint32_t v = adj[k];
            const uint32_t d = uint32_t(deg[v]) - 1u;
            deg[v] = I(d);
            if (d == 0) stack[sp++] = I(v);
        }
    }
    t.sync();
    return r == nn;
}

// ---------------------------------------------------------------------------
// Planning one `poa_inner` call over ranks [rank(0), rank(1)] (the unseeded
// whole-graph DP of src/align.rs:96-98) into the job's launch-pool regions:
// src/align.rs:587-688 as plan_job (plan.hpp) does it on the host.
// Returns MPOA_OK, a plan error (MPOA_ERR_*), or -1 when the job must leave
// the device (J.evict says why).
// ---------------------------------------------------------------------------

struct RPlanArrays {
    RowIn *rows;
    RowAux *aux;
    PredE *preds;
    uint8_t *seq;
};

MPOA_HD inline bool rj_scores_safe(const RParams &P, uint64_t rows, uint64_t qlen) {
    // src/align.rs:493-529 with rows = rank(end) - rank(beg).
    if (P.scalar_dp) return false;
    uint64_t steps = rows + qlen;
    if (steps < rows || steps + 20 < steps) return false;
    steps += 20;
    if (P.largest != 0 && steps > ~uint64_t(0) / P.largest) return false;
    return P.largest * steps <= (uint64_t(1) << 26);
}

// head, tail, support: per-row arrays (LDS or global) of at least n entries.
// The serial head and tail passes read each row's list through windows
// staged in LDS by the whole team (kRowWin rows' list offsets, kEntWin list
// entries), and keep the previous row's distance and support in registers
// (its predecessor or successor is usually the adjacent row).
constexpr uint32_t kRowWin = 512, kEntWin = 1024;
struct RWin {
    uint32_t *rb;  // kRowWin + 1 list offsets
    uint32_t *er;  // kEntWin entry rows
    uint32_t *ew;  // kEntWin entry weights
};
constexpr uint32_t kWinBytes = (kRowWin + 1 + 3) / 4 * 16 + 8 * kEntWin;

// src/align.rs:623-650: the heaviest in-range predecessor, ties to the
// larger support, then the first in `incoming` (the choose-order list keeps
// equal weights in incoming order). False when a listed predecessor has no
// head distance.
template <class T>
MPOA_HD inline bool rj_head_pass(T &t, uint32_t n, const RowIn *rows, const PredE *preds, int32_t *head,
                                 uint32_t *support, const RWin &W) {
    const uint32_t ntot = rows[n].pred_beg;
    bool unreachable = false;
    uint32_t r0 = 0, r1 = 0, e0 = 0, e1 = 0;
    int32_t hprev = 0;
    uint32_t sprev = 0;
    head[0] = 0;
    for (uint32_t i = 1; i < n; i++) {
        if (i + 1 > r1) {
            r0 = i;
            r1 = i + kRowWin < n ? i + kRowWin : n;
            t.sync();
            team_for(t, r1 - r0 + 1, [&](uint32_t k) { W.rb[k] = rows[r0 + k].pred_beg; });
            t.sync();
        }
        const uint32_t p0 = W.rb[i - r0], p1 = W.rb[i + 1 - r0];
        const bool direct = p1 - p0 > kEntWin;
        if (!direct && (p0 < e0 || p1 > e1)) {
            e0 = p0;
            e1 = p0 + kEntWin < ntot ? p0 + kEntWin : ntot;
            t.sync();
            team_for(t, e1 - e0, [&](uint32_t k) {
                const PredE pe = preds[e0 + k];
                W.er[k] = pe.row;
                W.ew[k] = pe.weight;
            });
            t.sync();
        }
        bool have = false;
        uint32_t bwt = 0, sb = 0;
        int32_t hb = 0;
        for (uint32_t a = p0; a < p1; a++) {
            const uint32_t row = direct ? preds[a].row : W.er[a - e0];
            const uint32_t w = direct ? preds[a].weight : W.ew[a - e0];
            const int32_t h = row + 1 == i ? hprev : head[row];
            if (h == kHeadNone) {
                unreachable = true;
                continue;
            }
            const uint32_t sp = row + 1 == i ? sprev : support[row];
            if (!have || w > bwt || (w == bwt && sp > sb)) {
                have = true;
                hb = h;
                sb = sp;
                bwt = w;
            }
        }
        if (have) {
            hprev = hb + 1;
            sprev = sb + bwt;
            head[i] = hprev;
            support[i] = sprev;
        } else {
            hprev = kHeadNone;
            sprev = 0;
        }
    }
    t.sync();
    return !unreachable;
}

// src/align.rs:651-678: tail[n-1] = 1, then rows in reverse, the heaviest
// in-range successor (out-list order) with a tail distance, ties to the
// larger support.
template <class T>
MPOA_HD inline void rj_tail_pass(T &t, uint32_t n, const uint32_t *sbeg, const RSucc *succ, int32_t *tail,
                                 uint32_t *support, const RWin &W) {
    const uint32_t ntot = sbeg[n];
    uint32_t r0 = n, r1 = n, e0 = 0, e1 = 0;
    tail[n - 1] = 1;
    int32_t tprev = 1;
    uint32_t sprev = support[n - 1];
    for (uint32_t ii = n - 1; ii-- > 0;) {
        if (ii < r0) {
            r1 = ii + 1;
            r0 = r1 >= kRowWin ? r1 - kRowWin : 0;
            t.sync();
            team_for(t, r1 - r0 + 1, [&](uint32_t k) { W.rb[k] = sbeg[r0 + k]; });
            t.sync();
        }
        const uint32_t s0 = W.rb[ii - r0], s1 = W.rb[ii + 1 - r0];
        const bool direct = s1 - s0 > kEntWin;
        if (!direct && (s0 < e0 || s1 > e1)) {
            e1 = s1;
            e0 = s1 >= kEntWin ? s1 - kEntWin : 0;
            (void)ntot;
            t.sync();
            team_for(t, e1 - e0, [&](uint32_t k) {
                const RSucc sc = succ[e0 + k];
                W.er[k] = sc.row;
                W.ew[k] = sc.w;
            });
            t.sync();
        }
        bool have = false;
        uint32_t bwt = 0, sb = 0;
        int32_t tb = 0;
        for (uint32_t a = s0; a < s1; a++) {
            const uint32_t row = direct ? succ[a].row : W.er[a - e0];
            const uint32_t w = direct ? succ[a].w : W.ew[a - e0];
            const int32_t tl = row == ii + 1 ? tprev : tail[row];
            if (tl == kHeadNone) continue;
            const uint32_t sp = row == ii + 1 ? sprev : support[row];
            if (!have || w > bwt || (w == bwt && sp > sb)) {
                have = true;
                tb = tl;
                sb = sp;
                bwt = w;
            }
        }
        if (have) {
            tprev = tb + 1;
            sprev = sb + bwt;
            tail[ii] = tprev;
            support[ii] = sprev;
        } else {
            tprev = kHeadNone;
            sprev = support[ii];
        }
    }
    t.sync();
}

template <class T>
MPOA_HD inline int rj_plan(T &t, const RParams &P, RJob &J, const RPlanArrays &A, const uint8_t *q,
                           uint64_t qlen, bool full, int32_t *head, int32_t *tail, uint32_t *support,
                           const RWin &W) {
    const uint32_t beg = J.rank[0], end = J.rank[1];
    if (end <= beg || end >= J.nn) return MPOA_ERR_ANCHORS_NOT_ORDERED;
    const uint64_t n64 = uint64_t(end) - beg + 1, m64 = qlen + 2;
    if (n64 > uint64_t(INT32_MAX) || m64 > uint64_t(INT32_MAX) || n64 + (m64 / 8 + 1) * 8 > kMaxSpan)
        return MPOA_ERR_COORDINATE_LIMITS;
    if (m64 > J.maxm) {
        J.evict = kEvCapacity;
        return -1;
    }
    if (!rj_scores_safe(P, n64 - 1, qlen)) {
        J.evict = kEvGuard;
        return -1;
    }
    const uint32_t n = uint32_t(n64), m = uint32_t(m64);
    const uint32_t padded = (m / 8 + 1) * 8;
    const bool adapt = !full && P.adaptive;
    // b + m / f, saturating (src/align.rs:619-620), then within int32.
    const uint64_t mdivf = (full || P.f == 0) ? 0 : m / P.f;
    uint64_t band = P.b + mdivf;
    if (band < P.b || band > uint64_t(INT32_MAX)) band = uint64_t(INT32_MAX);
    const uint32_t flags = (full ? kFull : 0u) | (adapt ? kAdaptive : 0u) | (P.gap_open == 0 ? kGapOpenZero : 0u);
    // Encoded query padded with N.
    team_for(t, padded, [&](uint32_t c) {
        A.seq[c] = (c >= 1 && c <= qlen) ? P.codes[q[c - 1]] : uint8_t(P.unknown);
    });
    // Per row: its in-range predecessors (rank order of the row before it)
    // and in-range successors, counted in list order; the filter
    // "head != NEG" of src/align.rs:632-640 is checked by the head pass.
    const uint32_t *order = J.order + beg;
    const uint32_t np_tot = team_place(
        t, n,
        [&](uint32_t i) {
            const uint32_t id = order[i];
            uint32_t c = 0;
            uint64_t tot = 0;
            for (uint32_t e = J.in_head[id], g_ = 0; e != kNil && g_ <= J.ne; e = J.edges[e].next_in, g_++) {
                const REdge ed = J.edges[e];
                tot += ed.w;
                const uint32_t rk = J.rank[ed.from];
                if (rk >= beg && rk < beg + i) c++;
            }
            A.aux[i] = RowAux{uint32_t(tot / 10 > 0xFFFFFFFFull ? 0xFFFFFFFFull : tot / 10), J.base[id]};
            A.rows[i] = RowIn{0, kHeadNone, kHeadNone, 0};
            return c;
        },
        [&](uint32_t i, uint32_t at) { A.rows[i].pred_beg = at; });
    const uint32_t ns_tot = team_place(
        t, n,
        [&](uint32_t i) {
            uint32_t c = 0;
            for (uint32_t e = J.out_head[order[i]], g_ = 0; e != kNil && g_ <= J.ne; e = J.edges[e].next_out, g_++) {
                const uint32_t rk = J.rank[J.edges[e].to];
                if (rk > beg + i && rk <= end) c++;
            }
            return c;
        },
        [&](uint32_t i, uint32_t at) { J.succ_beg[i] = at; });
    if (np_tot > J.ec || ns_tot > J.ec) {
        J.evict = kEvCapacity;
        return -1;
    }
    if (t.lane == 0) {
        A.rows[n] = RowIn{np_tot, kHeadNone, kHeadNone, 0};
        J.succ_beg[n] = ns_tot;
    }
    t.sync();
    // The lists: predecessors in `choose` order (heaviest edge first, then
    // the earliest in `incoming`: a stable insertion sort of the incoming-
    // order entries), rperm by counting; successors in out-list order.
    uint32_t max_np = 0;
    team_for(t, n, [&](uint32_t i) {
        const uint32_t id = order[i];
        PredE *pe = A.preds + A.rows[i].pred_beg;
        uint32_t c = 0, ord = 0;
        for (uint32_t e = J.in_head[id], g_ = 0; e != kNil && g_ <= J.ne; e = J.edges[e].next_in, g_++) {
            const REdge ed = J.edges[e];
            const uint32_t k = ord++;
            const uint32_t rk = J.rank[ed.from];
            if (rk < beg || rk >= beg + i) continue;
            const PredE v{rk - beg, ed.w, k, 0};
            uint32_t b = c++;
            for (; b > 0 && pe[b - 1].weight < v.weight; b--) pe[b] = pe[b - 1];
            pe[b] = v;
        }
        for (uint32_t a = 0; a < c; a++) {
            uint32_t r = 0;
            for (uint32_t b = 0; b < c; b++) r += pe[b].row < pe[a].row;
            pe[r].rperm = a;
        }
        if (c > max_np) max_np = c;
        RSucc *sv = J.succ + J.succ_beg[i];
        uint32_t s = 0;
        for (uint32_t e = J.out_head[id], g_ = 0; e != kNil && g_ <= J.ne; e = J.edges[e].next_out, g_++) {
            const REdge ed = J.edges[e];
            const uint32_t rk = J.rank[ed.to];
            if (rk > beg + i && rk <= end) sv[s++] = RSucc{rk - beg, ed.w};
        }
    });
    max_np = t.reduce_max(max_np);
    t.sync();
    // last_use: the last row that reads each row.
    team_for(t, n, [&](uint32_t i) {
        const uint32_t p0 = A.rows[i].pred_beg, p1 = A.rows[i + 1].pred_beg;
        for (uint32_t a = p0; a < p1; a++) t.umax(&A.rows[A.preds[a].row].last_use, i);
    });
    // Head distances (src/align.rs:623-650), then tail distances
    // (src/align.rs:651-678) sharing `support`, uniform on every lane.
    team_for(t, n, [&](uint32_t i) {
        head[i] = kHeadNone;
        tail[i] = kHeadNone;
        support[i] = 0;
    });
    t.sync();
    if (!rj_head_pass(t, n, A.rows, A.preds, head, support, W)) {
        // The reference drops such predecessors from every list; resident
        // graphs (built from reads) never have one.
        J.evict = kEvUnreachable;
        return -1;
    }
    rj_tail_pass(t, n, J.succ_beg, J.succ, tail, support, W);
    if (head[n - 1] == kHeadNone) return MPOA_ERR_NO_PATH;
    // Rows, and the arena bytes of the values layout (M, D, I of every
    // cell; adaptive rows sized at their static width plus two bands).
    const int64_t mi = int64_t(m), bnd = int64_t(band);
    uint64_t cells = 0, est = 0, nrows = 0;
    team_for(t, n, [&](uint32_t i) {
        const int32_t h = head[i], tl = tail[i];
        A.rows[i].head = h;
        A.rows[i].tail = tl;
        if (h == kHeadNone || tl == kHeadNone) return;
        nrows++;
        uint64_t w = padded;
        int64_t left = 0, right = mi;
        if (!full) {
            left = int64_t(h) < mi - tl ? int64_t(h) : mi - tl;
            left = left - bnd > 0 ? left - bnd : 0;
            right = int64_t(h) > mi - tl ? int64_t(h) : mi - tl;
            right = right + bnd < mi ? right + bnd : mi;
        }
        if (!(left > right || right < 0)) w = (uint64_t(right) / 8 + 1) * 8 - uint64_t(left) / 8 * 8;
        cells += w;
        uint64_t we = w;
        if (adapt) {
            const uint64_t x = w + 2 * uint64_t(band) + 16;
            we = x < padded ? x : padded;
        }
        est += 3 * we * 4;
    });
    cells = t.reduce_add64(cells);
    est = t.reduce_add64(est);
    nrows = t.reduce_add64(nrows);
    t.sync();
    uint64_t emax = 3 * nrows * uint64_t(padded) * 4;
    // Start adaptive DPs from what the last DP used, scaled to this graph
    // (the batch engine's hint), with eight full-width rows of slack.
    if (adapt && J.prev_n > 0) {
        uint64_t hint = J.prev_used * n / J.prev_n;
        hint += hint / 8 + uint64_t(padded) * 8 * 12;
        if (hint > est) est = hint;
        if (est > emax) est = emax;
    }
    if (t.lane == 0) {
        J.beg = beg;
        J.est = est < 256 ? 256 : est;
        J.emax = emax < 256 ? 256 : emax;
        DevJob &d = J.dj;
        d = DevJob{};
        d.rows = J.rows_off;
        d.aux = J.aux_off;
        d.preds = J.preds_off;
        d.seq = J.seq_off;
        d.matrix = 0;
        d.rowmeta = J.meta_off;
        d.adapt = J.adapt_off;
        d.tb = J.tb_off;
        d.n = n;
        d.m = m;
        d.padded = padded;
        d.alpha = P.alpha;
        d.band = int64_t(band);
        d.e = P.gap_ext;
        d.o = P.gap_open + P.gap_ext;
        d.unknown = P.unknown;
        d.score_nn = P.score_nn;
        d.flags = flags;
        d.band32 = int32_t(int64_t(band) < int64_t(kBandCap) ? int64_t(band) : int64_t(kBandCap));
        d.cap5 = int32_t(5 * int64_t(band) < int64_t(kBandCap) ? 5 * int64_t(band) : int64_t(kBandCap));
    }
    (void)max_np;
    t.sync();
    return MPOA_OK;
}

// Path elements of a completed DP (src/align.rs:1293-1347), as assemble()
// in plan.hpp: tb[j] is the row of the diagonal step that consumed column j
// or kTbInsert; element 0 is the source, element j the query column j.
struct RTbElems {
    const RJob *J;
    const uint8_t *seq;
    const uint32_t *tb;
    const uint32_t *row_node;
    uint32_t n, m, alpha;
    MPOA_HD RElem operator()(uint32_t j) const {
        if (j == 0) {
            const uint32_t s = row_node[0];
            return RElem{int64_t(s), int64_t(J->base[s]), -1};
        }
        const uint32_t qb = seq[j];
        const uint32_t i = tb[j];
        if (i == kTbInsert || (i == n - 1 && j != m - 1)) return RElem{-1, int64_t(qb), -1};
        const uint32_t id = row_node[i];
        const uint32_t nb = J->base[id];
        if (nb == qb) return RElem{int64_t(id), int64_t(nb), -1};
        const uint32_t par = J->parent[id];
        return RElem{int64_t(J->aligned[uint64_t(par) * alpha + qb]), int64_t(qb), int64_t(par)};
    }
};

// The first read of an empty graph (src/align.rs:82-95): the source, then
// every query base as an insertion.
struct RFirstElems {
    const RParams *P;
    const uint8_t *q;
    MPOA_HD RElem operator()(uint32_t j) const {
        if (j == 0) return RElem{0, int64_t(P->unknown), -1};
        return RElem{-1, int64_t(P->codes[q[j - 1]]), -1};
    }
};

// The parts of a finished (or evicted) job's graph the host rebuilds it from,
// packed per job: node bases, parents, ranks, coverage, list heads, edges,
// and the reads' paths.
struct RPackLayout {
    uint64_t base, parent, rank, cover, out_head, in_head, edges, path_off, paths, bytes;
};
MPOA_HD inline RPackLayout rj_pack_layout(uint32_t nn, uint32_t ne, uint32_t nreads, uint32_t npath) {
    auto up = [](uint64_t x) { return (x + 15) / 16 * 16; };
    RPackLayout L;
    L.base = 0;
    L.parent = up(L.base + nn);
    L.rank = up(L.parent + 4ull * nn);
    L.cover = up(L.rank + 4ull * nn);
    L.out_head = up(L.cover + 4ull * nn);
    L.in_head = up(L.out_head + 4ull * nn);
    L.edges = up(L.in_head + 4ull * nn);
    L.path_off = up(L.edges + sizeof(REdge) * uint64_t(ne));
    L.paths = up(L.path_off + 4ull * (nreads + 1));
    L.bytes = up(L.paths + 4ull * npath);
    return L;
}

// ---------------------------------------------------------------------------
// One step of job k (src/lib.rs:195-202 for one read, with the retry of
// src/align.rs:202-207): consume the DP launched in the previous step (or
// retry it), fuse its path and re-rank, plan the next read and place its
// DP in the step's arena. lds: per-team staging of lds_bytes for the Kahn
// walk and the head/tail passes when the graph fits.
// ---------------------------------------------------------------------------

// The job's graph re-ranked, its staging in LDS when it fits.
template <class T>
MPOA_HD inline bool rj_rerank(T &t, RJob &J, char *lds, uint32_t lds_bytes) {
    const uint64_t nn = J.nn, ne = J.ne;
    const uint64_t need = 4 * (nn + 1) + 2 * (ne + 1) + 2 * nn + 2 * nn + 64;
    if (nn < 65536 && need <= lds_bytes) {
        uint32_t *off = reinterpret_cast<uint32_t *>(lds);
        uint16_t *adj = reinterpret_cast<uint16_t *>(off + nn + 1);
        uint16_t *deg = adj + ne + 1;
        uint16_t *stk = deg + nn;
        return rj_toposort<T, uint16_t>(t, J, off, adj, deg, stk);
    }
    return rj_toposort<T, uint32_t>(t, J, J.csr_off, J.csr_adj, J.deg, J.stack);
}

template <class T>
MPOA_HD inline int rj_plan_read(T &t, const RPools &R, const RParams &P, RJob &J, bool full, char *lds,
                                uint32_t lds_bytes) {
    const uint8_t *q = J.bases + J.read_off[J.cur];
    const uint64_t qlen = J.read_off[J.cur + 1] - J.read_off[J.cur];
    const RPlanArrays A{R.rows + J.rows_off, R.aux + J.aux_off, R.preds + J.preds_off, R.seq + J.seq_off};
    // LDS: the passes' windows, then head, tail and support when they fit;
    // with lds_bytes below the windows (the CPU check's global form) every
    // array is the job's own.
    RWin W{J.win, J.win + kRowWin + 1, J.win + kRowWin + 1 + kEntWin};
    uint64_t used = 0;
    if (lds_bytes >= kWinBytes) {
        W.rb = reinterpret_cast<uint32_t *>(lds);
        W.er = reinterpret_cast<uint32_t *>(lds + (kRowWin + 1 + 3) / 4 * 16);
        W.ew = W.er + kEntWin;
        used = kWinBytes;
    }
    if (used > 0 && used + uint64_t(J.nn) * 12 + 64 <= lds_bytes) {
        int32_t *hd = reinterpret_cast<int32_t *>(lds + used);
        int32_t *tl = hd + J.nn;
        uint32_t *sp = reinterpret_cast<uint32_t *>(tl + J.nn);
        return rj_plan(t, P, J, A, q, qlen, full, hd, tl, sp, W);
    }
    return rj_plan(t, P, J, A, q, qlen, full, J.head, J.tail, J.support, W);
}

// slot_cap != 0: the DP's arena is the caller's own region [slot_off,
// slot_off + slot_cap) (the persistent kernel, one region per workgroup),
// which the fill extends with chunks of the launch's shared pool when it
// outgrows it; a DP whose estimate exceeds R.job_cap, or that finds the pool
// spent, leaves the device. Otherwise the DP is placed in the step's shared
// arena.
template <class T>
MPOA_HD inline void rj_job_step(T &t, const RPools &R, uint32_t k, char *lds, uint32_t lds_bytes,
                                uint64_t slot_off = 0, uint64_t slot_cap = 0) {
    RJob &J = R.jobs[k];
    const RParams &P = *R.prm;
    DevJob &out = R.dj[k];
    if (J.state != kRsRunning) {
        if (t.lane == 0) out.n = 0;
        return;
    }
    auto leave = [&](uint32_t state) {
        if (t.lane == 0) {
            J.state = state;
            t.add(R.done, 1u);
            out.n = 0;
        }
        t.sync();
    };
    auto fail = [&](int32_t err, int32_t dev, uint32_t ei, uint32_t ej) {
        if (t.lane == 0) {
            J.err = err;
            J.dev_status = dev;
            J.err_i = ei;
            J.err_j = ej;
        }
        leave(kRsFailed);
    };
    if (J.nn == 0) {
        // src/graph.rs:30-40: source (id 0) and sink (id 1), both base N.
        if (t.lane == 0) {
            rj_node_init(P, J, 0, P.unknown, 0);
            rj_node_init(P, J, 1, P.unknown, 1);
            J.nn = 2;
            J.path_off[0] = 0;
        }
        t.sync();
    }
    uint32_t phase = J.phase;
    // Plans the current read: banded (unless f == 0), then poa's full-DP
    // retry on any error. False when the job left the device.
    auto plan_with_retry = [&](bool retry_now) -> bool {
        int rc = -2;
        if (!retry_now) {
            if (t.lane == 0) {
                J.retry = 0;
                J.full_arena = 0;
            }
            t.sync();
            rc = rj_plan_read(t, R, P, J, P.f == 0, lds, lds_bytes);
            if (rc == -1) {
                leave(kRsEvicted);
                return false;
            }
            if (rc == MPOA_OK) return true;
            if (P.f == 0) {
                fail(rc, kOk, 0, 0);
                return false;
            }
        }
        if (t.lane == 0) {
            J.retry = 1;
            J.retries += 1;
            J.full_arena = 0;
        }
        t.sync();
        rc = rj_plan_read(t, R, P, J, true, lds, lds_bytes);
        if (rc == -1) {
            leave(kRsEvicted);
            return false;
        }
        if (rc != MPOA_OK) {
            fail(rc, kOk, 0, 0);
            return false;
        }
        return true;
    };
    if (phase == kRpLaunched) {
        const DevResult res = R.res[k];
        if (res.status == kArenaOverflow) {
            if (slot_cap) {
                // Its region and the shared pool are spent: a larger
                // estimate would not give it more.
                if (t.lane == 0) J.evict = kEvArena;
                leave(kRsEvicted);
                return;
            }
            // The same DP again, its arena sized for full-width rows.
            if (t.lane == 0) {
                J.arena_retries += 1;
                J.full_arena = 1;
            }
            t.sync();
            phase = kRpPlanned;
        } else if (res.status != kOk) {
            if (P.f != 0 && !J.retry) {
                if (!plan_with_retry(true)) return;
                phase = kRpPlanned;
            } else {
                fail(0, res.status, res.err_i, res.err_j);
                return;
            }
        } else {
            if (t.lane == 0) {
                J.cells += res.cells;
                J.pred_cells += res.pred_cells;
                J.arena_bytes += res.arena_used;
                J.dp_calls += 1;
                if (!J.retry) {
                    J.prev_used = res.arena_used;
                    J.prev_cells = res.cells;
                    J.prev_n = J.dj.n;
                }
            }
            const RTbElems el{&J, R.seq + J.seq_off, R.tb + J.tb_off, J.order + J.beg, J.dj.n, J.dj.m, P.alpha};
            rj_add_path(t, P, J, el, J.dj.m - 1, J.stack, J.deg);
            if (!rj_rerank(t, J, lds, lds_bytes)) {
                fail(MPOA_ERR_GRAPH, kOk, 0, 0);
                return;
            }
            if (t.lane == 0) J.cur += 1;
            t.sync();
            phase = kRpIdle;
        }
    }
    if (phase == kRpIdle) {
        // An empty graph takes the read as insertions after the source
        // (src/align.rs:82-95), with no DP.
        while (J.cur < J.nreads && J.nn <= 2) {
            const uint8_t *q = J.bases + J.read_off[J.cur];
            const uint32_t qlen = uint32_t(J.read_off[J.cur + 1] - J.read_off[J.cur]);
            rj_add_path(t, P, J, RFirstElems{&P, q}, qlen + 1, J.stack, J.deg);
            if (!rj_rerank(t, J, lds, lds_bytes)) {
                fail(MPOA_ERR_GRAPH, kOk, 0, 0);
                return;
            }
            if (t.lane == 0) J.cur += 1;
            t.sync();
        }
        if (J.cur >= J.nreads) {
            leave(kRsDone);
            return;
        }
        if (!plan_with_retry(false)) return;
    }
    // Place the DP: its arena bytes from the step's arena, or wait for the
    // next step when this one's placements are spent.
    const uint64_t want = ((J.full_arena ? J.emax : J.est) + 255) / 256 * 256;
    if (want > R.job_cap) {
        if (t.lane == 0) J.evict = kEvArena;
        leave(kRsEvicted);
        return;
    }
    if (slot_cap) {
        if (t.lane == 0) {
            J.dj.arena = slot_off;
            J.dj.arena_cap = slot_cap;
            out = J.dj;
            J.phase = kRpLaunched;
        }
        t.sync();
        return;
    }
    if (t.lane == 0) {
        unsigned long long old = *R.arena_ctr;
        bool placed = false;
        while (old + want <= R.alloc_limit) {
            const unsigned long long got = t.cas64(R.arena_ctr, old, old + want);
            if (got == old) {
                placed = true;
                break;
            }
            old = got;
        }
        if (placed) {
            J.dj.arena = old;
            J.dj.arena_cap = want;
            out = J.dj;
            J.phase = kRpLaunched;
        } else {
            J.deferred += 1;
            out.n = 0;
            J.phase = kRpPlanned;
        }
    }
    t.sync();
}

}  // namespace mpoa
