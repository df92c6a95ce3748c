// Host planning of one DP job: anchor checks, query encoding, band size,
// the most-supported-path head/tail distances and the rank-ordered row pools
// the device kernels read; plus assembly of the device traceback record into
// AlignedBase records. Semantics: minipoa-rs src/align.rs (cited inline).
#pragma once
#include <algorithm>
#include <cstdint>
#include <string>
#include <vector>

#include "dp_types.hpp"
#include "graph.hpp"
#include "minipoa_hip.h"

namespace mpoa {

// Graph accessors: the library-owned Graph and the caller's flat view expose
// the same read-only surface to the planner.
struct GraphAcc {
    const Graph &g;
    uint64_t size() const { return g.size(); }
    uint64_t rank(uint64_t id) const { return g.rank_of(id); }
    uint64_t order_size() const { return g.order().size(); }
    uint64_t order(uint64_t r) const { return g.order()[r]; }
    uint32_t base(uint64_t id) const { return g.node(id).base; }
    uint64_t parent(uint64_t id) const { return g.node(id).parent; }
    int64_t aligned(uint64_t id, uint64_t b) const { return g.aligned(id, b); }
    template <class F> void for_in(uint64_t id, F &&f) const {
        for (const Edge &ed : g.node(id).in) f(uint64_t(ed.node), uint64_t(ed.w));
    }
    template <class F> void for_out(uint64_t id, F &&f) const {
        for (const Edge &ed : g.node(id).out) f(uint64_t(ed.node), uint64_t(ed.w));
    }
};

struct ViewAcc {
    const mpoa_graph_view &v;
    uint64_t size() const { return v.n_nodes; }
    uint64_t rank(uint64_t id) const { return v.rank[id]; }
    uint64_t order_size() const { return v.n_order; }
    uint64_t order(uint64_t r) const { return v.order[r]; }
    uint32_t base(uint64_t id) const { return uint32_t(v.base[id]); }
    uint64_t parent(uint64_t id) const { return v.parent[id]; }
    int64_t aligned(uint64_t id, uint64_t b) const { return v.aligned[id * v.alphabet_len + b]; }
    template <class F> void for_in(uint64_t id, F &&f) const {
        for (uint64_t k = v.in_off[id]; k < v.in_off[id + 1]; k++) f(v.in_node[k], v.in_weight[k]);
    }
    template <class F> void for_out(uint64_t id, F &&f) const {
        for (uint64_t k = v.out_off[id]; k < v.out_off[id + 1]; k++) f(v.out_node[k], v.out_weight[k]);
    }
};

struct JobPlan {
    int status = MPOA_OK;  // host-side error before any device work
    std::string err;
    uint32_t n = 0, m = 0, padded = 0, beg = 0;
    bool use_i64 = false;
    uint32_t flags = 0, unknown = 0;
    int64_t band = 0;
    int32_t o = 0, e = 0, score_nn = 0;
    uint64_t static_cells = 0;  // cells of the rows under the static/full band
    // Arena bytes of the DP in each layout: values ([M][D][I] of every
    // row) and records (records of every row, M and D of rows read beyond
    // the next); est as planned, max with every row at full width and its
    // scores kept. use_layout() picks one into arena_est / arena_max.
    uint64_t est_val = 0, max_val = 0, est_rec = 0, max_rec = 0;
    bool records = false;       // the layout the arena is sized for
    uint64_t arena_est = 0;     // arena bytes requested
    uint64_t arena_max = 0;     // arena bytes of every row at full width
    uint32_t max_width = 0;     // widest static row (chooses columns per lane)
    uint32_t max_np = 0;        // most predecessors of one row (record width)
    uint32_t width_hint = 0;    // mean adaptive row width seen on this job's last DP
    bool full_arena = false;    // size the arena for full-width rows (after an overflow)
    std::vector<RowIn> rows;    // n + 1
    std::vector<RowAux> aux;    // n
    std::vector<PredE> preds;
    std::vector<uint8_t> seq;   // encoded query, padded with N
    std::vector<uint32_t> row_node;

    // Back to the default state, keeping the pools' capacity.
    void reset() {
        status = MPOA_OK;
        err.clear();
        n = m = padded = beg = 0;
        use_i64 = false;
        flags = unknown = 0;
        band = 0;
        o = e = score_nn = 0;
        static_cells = arena_est = arena_max = 0;
        est_val = max_val = est_rec = max_rec = 0;
        records = false;
        max_width = width_hint = max_np = 0;
        full_arena = false;
        rows.clear();
        aux.clear();
        preds.clear();
        seq.clear();
        row_node.clear();
    }
};

// Most-supported-path scratch of one row (src/align.rs:623-678).
struct PathDist {
    int64_t head, tail, support;
};

// Sizes the DP's arena for a layout: records (rec) or values. Records
// name at most kMaxRecPreds predecessors of a row; past that the DP keeps
// values. Returns the layout taken.
inline bool use_layout(JobPlan &J, bool rec) {
    rec = rec && J.max_np <= kMaxRecPreds;
    J.records = rec;
    J.arena_est = rec ? J.est_rec : J.est_val;
    J.arena_max = rec ? J.max_rec : J.max_val;
    return rec;
}

constexpr uint64_t kWidth = 8;
constexpr uint64_t kScoreLimit = uint64_t(1) << 26;  // src/align.rs:217

// src/align.rs:493-529: i32 states are lossless when every |score| times the
// number of DP steps stays within 2^26.
template <class G>
bool i32_scores_safe(const Params &p, const G &g, uint64_t beg_id, uint64_t end_id,
                     uint64_t qlen) {
    if (p.scalar_dp || beg_id >= g.size() || end_id >= g.size()) return false;
    uint64_t re = g.rank(end_id), rb = g.rank(beg_id);
    if (re < rb) return false;
    uint64_t rows = re - rb, steps = rows + qlen;
    if (steps < rows || steps + kWidth * 2 + 4 < steps) return false;
    steps += kWidth * 2 + 4;
    int64_t opening = int64_t(p.gap_open) + int64_t(p.gap_ext);
    uint64_t largest = 0;
    for (int32_t v : p.matrix) largest = std::max<uint64_t>(largest, uint64_t(std::llabs(int64_t(v))));
    largest = std::max<uint64_t>(largest, uint64_t(std::llabs(opening)));
    largest = std::max<uint64_t>(largest, uint64_t(std::llabs(int64_t(p.gap_ext))));
    if (largest != 0 && steps > UINT64_MAX / largest) return false;
    return largest * steps <= kScoreLimit;
}

// Static/full row band (src/align.rs:711-734 without the adaptive branch).
inline bool static_band(const JobPlan &J, uint32_t i, uint64_t &start, uint64_t &stop) {
    int64_t m = J.m, left, right;
    if (J.flags & kFull) {
        left = 0;
        right = m;
    } else {
        int64_t h = J.rows[i].head, t = J.rows[i].tail;
        left = std::max<int64_t>(std::min(h, m - t) - J.band, 0);
        right = std::min<int64_t>(std::max(h, m - t) + J.band, m);
    }
    if (left > right || right < 0) return false;
    start = uint64_t(left) / kWidth * kWidth;
    stop = (uint64_t(right) / kWidth + 1) * kWidth;
    return true;
}

// Plans one `poa_inner` call (src/align.rs:587-688 == 913-1011 up to the row
// loop). force_bits: 0 = guard decides, 32 / 64 = that kernel.
template <class G>
void plan_job(const Params &p, const G &g, uint64_t beg_id, uint64_t end_id, const uint8_t *q,
              uint64_t qlen, bool adaptive, uint64_t band_b, int force_bits, JobPlan &J) {
    J.reset();
    if (beg_id >= g.size() || end_id >= g.size()) {
        J.status = MPOA_ERR_ANCHOR_NOT_IN_GRAPH;
        J.err = "alignment anchor is not in graph";
        return;
    }
    uint64_t beg = g.rank(beg_id), end = g.rank(end_id);
    if (end <= beg || end >= g.order_size()) {
        J.status = MPOA_ERR_ANCHORS_NOT_ORDERED;
        J.err = "alignment anchors are not topologically ordered";
        return;
    }
    uint64_t n = end - beg + 1, m = qlen + 2;
    // The device keeps rows, columns and band offsets in 32 bits (RowRec).
    if (n > uint64_t(INT32_MAX) || m > uint64_t(INT32_MAX) || n + (m / kWidth + 1) * kWidth > kMaxSpan) {
        J.status = MPOA_ERR_COORDINATE_LIMITS;
        J.err = "alignment exceeds coordinate limits";
        return;
    }
    J.use_i64 = force_bits == 64 || (force_bits == 0 && !i32_scores_safe(p, g, beg_id, end_id, qlen));
    J.n = uint32_t(n);
    J.m = uint32_t(m);
    J.beg = uint32_t(beg);
    J.unknown = p.encode('N');
    J.padded = uint32_t((m / kWidth + 1) * kWidth);
    J.seq.assign(J.padded, uint8_t(J.unknown));
    for (uint64_t k = 0; k < qlen; k++) J.seq[k + 1] = uint8_t(p.encode(q[k]));
    bool adapt = adaptive || p.adaptive;
    uint64_t mdivf = p.f ? m / p.f : 0;
    uint64_t bw = band_b + mdivf;
    if (bw < band_b) bw = UINT64_MAX;  // saturating_add
    J.band = int64_t(std::min<uint64_t>(bw, uint64_t(INT32_MAX)));
    J.e = p.gap_ext;
    J.o = p.gap_open + p.gap_ext;
    J.score_nn = p.score(J.unknown, J.unknown);
    J.flags = (p.f == 0 ? kFull : 0) | ((p.f != 0 && adapt) ? kAdaptive : 0) |
              (p.gap_open == 0 ? kGapOpenZero : 0);

    // One pass over `incoming` per row in rank order builds the head
    // distances (src/align.rs:623-650: heaviest edge, ties to larger
    // support) and the row's predecessor list. Both keep only in-range
    // predecessors with head != NEG, and a predecessor's head is final
    // before its successors are visited.
    thread_local std::vector<PathDist> dist;
    dist.assign(n, PathDist{INT64_MIN, INT64_MIN, 0});
    J.row_node.resize(n);
    for (uint64_t i = 0; i < n; i++) J.row_node[i] = uint32_t(g.order(beg + i));
    J.rows.assign(n + 1, RowIn{0, 0, 0, 0});
    J.aux.resize(n);
    uint32_t max_np = 0;
    dist[0].head = 0;
    for (uint64_t i = 0; i < n; i++) {
        const uint64_t id = J.row_node[i];
        const uint32_t pbeg = uint32_t(J.preds.size());
        bool have = false;
        uint64_t bp = 0, bwt = 0, tot = 0;
        uint32_t ord = 0;
        g.for_in(id, [&](uint64_t pred, uint64_t w) {
            tot += w;
            uint32_t k = ord++;
            uint64_t rk = g.rank(pred);
            if (rk < beg || rk >= beg + i) return;
            uint64_t pr = rk - beg;
            if (dist[pr].head == INT64_MIN) return;
            J.preds.push_back(PredE{uint32_t(pr), uint32_t(std::min<uint64_t>(w, UINT32_MAX)), k, 0});
            if (!have || w > bwt || (w == bwt && dist[pr].support > dist[bp].support)) {
                have = true;
                bp = pr;
                bwt = w;
            }
        });
        if (i > 0 && have) {
            dist[i].head = dist[bp].head + 1;
            dist[i].support = dist[bp].support + int64_t(bwt);
        }
        // The list in `choose` order (src/align.rs:1201-1225): heaviest edge
        // first, then the earliest in `incoming` (entries arrive in `incoming`
        // order, so a stable sort on weight gives it). rperm walks the same
        // entries in rank order, which the adaptive band fold needs.
        PredE *pe = J.preds.data();
        const uint32_t pend = uint32_t(J.preds.size()), np = pend - pbeg;
        if (np <= 16) {
            // Short lists (nearly all): insertion sort, which is stable, and
            // rank positions by counting.
            for (uint32_t a = pbeg + 1; a < pend; a++) {
                const PredE v = pe[a];
                uint32_t b = a;
                for (; b > pbeg && pe[b - 1].weight < v.weight; b--) pe[b] = pe[b - 1];
                pe[b] = v;
            }
            for (uint32_t a = pbeg; a < pend; a++) {
                uint32_t t = 0;
                for (uint32_t b = pbeg; b < pend; b++) t += pe[b].row < pe[a].row;
                pe[pbeg + t].rperm = a - pbeg;
            }
        } else {
            std::stable_sort(pe + pbeg, pe + pend, [](const PredE &x, const PredE &y) { return x.weight > y.weight; });
            thread_local std::vector<uint32_t> by_row;
            by_row.resize(np);
            for (uint32_t a = 0; a < np; a++) by_row[a] = a;
            std::sort(by_row.begin(), by_row.end(),
                      [&](uint32_t x, uint32_t y) { return pe[pbeg + x].row < pe[pbeg + y].row; });
            for (uint32_t a = 0; a < np; a++) pe[pbeg + a].rperm = by_row[a];
        }
        for (uint32_t a = pbeg; a < pend; a++) J.rows[pe[a].row].last_use = uint32_t(i);
        max_np = std::max(max_np, np);
        J.rows[i].pred_beg = pbeg;
        J.aux[i].thresh = uint32_t(std::min<uint64_t>(tot / 10, UINT32_MAX));
        J.aux[i].base = g.base(id);
    }
    // Tail distances (src/align.rs:651-678), sharing `support` with the
    // head pass as the reference does.
    dist[n - 1].tail = 1;
    for (uint64_t ii = n - 1; ii-- > 0;) {
        bool have = false;
        uint64_t bs = 0, bwt = 0;
        g.for_out(J.row_node[ii], [&](uint64_t succ, uint64_t w) {
            uint64_t r = g.rank(succ);
            if (r <= beg + ii || r > end) return;
            uint64_t sc = r - beg;
            if (dist[sc].tail == INT64_MIN) return;
            if (!have || w > bwt || (w == bwt && dist[sc].support > dist[bs].support)) {
                have = true;
                bs = sc;
                bwt = w;
            }
        });
        if (have) {
            dist[ii].tail = dist[bs].tail + 1;
            dist[ii].support = dist[bs].support + int64_t(bwt);
        }
    }
    if (dist[n - 1].head == INT64_MIN) {
        J.status = MPOA_ERR_NO_PATH;
        J.err = "no graph path connects alignment anchors";
        return;
    }
    for (uint64_t i = 0; i < n; i++) {
        J.rows[i].head = dist[i].head == INT64_MIN ? kHeadNone : int32_t(dist[i].head);
        J.rows[i].tail = dist[i].tail == INT64_MIN ? kHeadNone : int32_t(dist[i].tail);
    }
    J.rows[n].pred_beg = uint32_t(J.preds.size());
    J.rows[n].head = J.rows[n].tail = kHeadNone;
    J.rows[n].last_use = 0;
    J.max_np = max_np;

    // Arena bytes. Values: M, D and I of every cell. Records: one record
    // per cell (rec_bytes of the row's predecessor count), rounded to 16
    // bytes, and M and D of a row read beyond the fill's LDS history, or
    // read later and wider than the smallest LDS ring's half (no ring slot),
    // or -- an estimate of the fill's spills -- read more than four rows
    // later. Static/full bands are exact in values; adaptive rows start from
    // the static width plus two band widths; rows beyond the estimate take
    // pool chunks, and an overflow relaunches with the full-width bound.
    const uint64_t tsz = J.use_i64 ? 8 : 4;
    auto rec_row = [&](uint32_t i, uint64_t w, bool all_scores) {
        const uint32_t np = J.rows[i + 1].pred_beg - J.rows[i].pred_beg;
        uint64_t b = (w * rec_bytes(np) + 15) / 16 * 16;
        const uint32_t lu = J.rows[i].last_use;
        if (all_scores || lu >= i + kHistRows || lu > i + 4 || (lu > i && 2 * w * tsz > 8192)) b += 2 * w * tsz;
        return b;
    };
    uint64_t cells = 0, est_r = 0, est_v = 0;
    uint32_t maxw = 0;
    uint64_t nrows = 0;
    for (uint32_t i = 0; i < n; i++) {
        if (J.rows[i].head == kHeadNone || J.rows[i].tail == kHeadNone) continue;
        nrows++;
        uint64_t s0, t0;
        uint64_t w;
        if (static_band(J, i, s0, t0))
            w = t0 - s0;
        else
            w = J.padded;
        cells += w;
        maxw = std::max<uint32_t>(maxw, uint32_t(w));
        const uint64_t we = (J.flags & kAdaptive) ? std::min<uint64_t>(J.padded, w + 2 * uint64_t(J.band) + 16) : w;
        est_r += rec_row(i, we, false);
        est_v += 3 * we * tsz;
    }
    J.static_cells = cells;
    J.max_width = maxw;
    J.est_rec = est_r;
    J.est_val = est_v;
    J.max_val = 3 * nrows * J.padded * tsz;
    uint64_t full_r = 0;
    for (uint32_t i = 0; i < n; i++)
        if (J.rows[i].head != kHeadNone && J.rows[i].tail != kHeadNone) full_r += rec_row(i, J.padded, true);
    J.max_rec = full_r;
    use_layout(J, false);
}

// Full-width arena bound of the chosen layout, in bytes.
inline uint64_t arena_full(const JobPlan &J) { return J.arena_max; }

// Device traceback record -> AlignedBase list (src/align.rs:1293-1347): the
// record holds, per column j = 1 .. m-1, the row of the diagonal step that
// consumed j or kTbInsert. Reversal then dropping the first traceback step
// (column m-1) leaves [source, col 1, .., col m-2].
template <class G>
void assemble(const G &g, const JobPlan &J, const uint32_t *tb, std::vector<AlignedBase> &out) {
    out.clear();
    out.reserve(J.m);
    uint64_t start = J.row_node[0];
    out.push_back(AlignedBase{int64_t(start), int64_t(g.base(start)), -1});
    for (uint32_t j = 1; j + 1 < J.m; j++) {
        uint32_t qb = J.seq[j];
        uint32_t i = tb[j];
        if (i == kTbInsert || (i == J.n - 1 && j != J.m - 1)) {
            out.push_back(AlignedBase{-1, int64_t(qb), -1});
            continue;
        }
        uint64_t id = J.row_node[i];
        uint32_t nb = g.base(id);
        if (nb == qb) {
            out.push_back(AlignedBase{int64_t(id), int64_t(nb), -1});
        } else {
            uint64_t parent = g.parent(id);
            out.push_back(AlignedBase{g.aligned(parent, qb), int64_t(qb), int64_t(parent)});
        }
    }
}

}  // namespace mpoa
