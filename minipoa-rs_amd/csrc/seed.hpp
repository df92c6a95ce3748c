// Minimizer seeding and seeded windows of the query (minipoa-rs
// src/minimizer.rs:32-488, src/align.rs:99-188), on the host: they decide
// which anchored sub-DPs a seeded alignment runs, and every window of one
// read is an independent DP the device runs side by side with the others.
#pragma once
#include <algorithm>
#include <cstdint>
#include <string>
#include <vector>

#include "graph.hpp"
#include "minipoa_hip.h"
#include "plan.hpp"

namespace mpoa {

// src/minimizer.rs:11-18: exact forward 2-bit k-mer value (first base in
// the low bits), inclusive end position, span k.
struct Minimizer {
    uint64_t value;
    uint64_t end;
    uint32_t span;
};

// src/minimizer.rs:20-26
struct Anchor {
    uint64_t target_end, query_end, span;
    bool adaptive_before;
};

// src/minimizer.rs:32-76. Runs of A/C/G/T/U (any case) are sketched apart;
// each k-mer is ranked by the upper 16 bits of its 32-bit forward ntHash and
// every window of w k-mers keeps its leftmost minimum (the simd-minimizers
// crate's ordering; the exact key the crate truncates to is not pinned).
void sketch(const uint8_t *seq, uint64_t len, uint64_t k, uint64_t w, std::vector<Minimizer> &out);
// src/minimizer.rs:78-110
void repeat_intervals(const std::vector<Minimizer> &mm, uint64_t len, uint64_t max_dist,
                      std::vector<std::pair<uint64_t, uint64_t>> &out);
// src/minimizer.rs:128-176
void collect_anchors(const Params &p, const uint8_t *q, uint64_t qlen, const uint8_t *t,
                     uint64_t tlen, std::vector<Anchor> &out);
// src/minimizer.rs:199-377
void chain(const std::vector<Anchor> &anchors, uint64_t k, std::vector<Anchor> &out);
// src/minimizer.rs:379-467; returns the tail window's adaptive flag.
bool partition_windows(const std::vector<Anchor> &chain, uint64_t tlen, uint64_t qlen,
                       uint64_t max_drift, uint64_t max_win, std::vector<Anchor> &out);
// src/minimizer.rs:475-488; returns the tail window's adaptive flag.
bool anchor_windows(const Params &p, const uint8_t *q, uint64_t qlen, const uint8_t *cons,
                    uint64_t clen, std::vector<Anchor> &out);

// src/minimizer.rs:489-578: the progressive guide order of n reads (input
// order unless p.progressive, n >= 2 and at most five symbols).
void guide_order(const Params &p, const uint8_t *const *seqs, const uint64_t *lens, uint64_t n,
                 std::vector<uint32_t> &order);

// One anchored sub-DP of a seeded alignment: poa(beg, end) over
// query[qbeg, qend), followed in the result by the anchor's own graph nodes.
struct SeedWindow {
    uint64_t beg = 0, end = 1, qbeg = 0, qend = 0;
    bool adaptive = true;
    uint64_t band = 0;
    std::vector<uint32_t> anchor;
};

// src/graph.rs:180-237 over any graph accessor: node ids of the heaviest
// path, source and sink excluded.
template <class G>
void heaviest_path(const G &g, std::vector<uint32_t> &ids) {
    const uint64_t nn = g.size();
    std::vector<uint32_t> indeg(nn, 0), queue;
    std::vector<int64_t> prev(nn, -1);
    std::vector<uint64_t> score(nn, 0);
    for (uint64_t v = 0; v < nn; v++) g.for_in(v, [&](uint64_t, uint64_t) { indeg[v]++; });
    queue.reserve(nn);
    queue.push_back(0);
    for (size_t h = 0; h < queue.size(); h++) {
        const uint32_t v = queue[h];
        if (v != 0) {
            // Heaviest incoming edge; ties by the predecessor's path score,
            // then the earlier edge.
            bool any = false;
            uint64_t bw = 0, bu = 0;
            g.for_in(v, [&](uint64_t u, uint64_t w) {
                if (!any || w > bw || (w == bw && score[u] > score[bu])) {
                    any = true;
                    bw = w;
                    bu = u;
                }
            });
            if (any) {
                prev[v] = int64_t(bu);
                score[v] = score[bu] + bw;
            }
        }
        if (v == 1) break;
        g.for_out(v, [&](uint64_t u, uint64_t) {
            if (--indeg[u] == 0) queue.push_back(uint32_t(u));
        });
    }
    ids.clear();
    for (int64_t c = nn > 1 ? prev[1] : -1; c > 0; c = prev[size_t(c)]) ids.push_back(uint32_t(c));
    std::reverse(ids.begin(), ids.end());
}

// The consensus path's ids: the library graph's own (flat-array) walk, or
// the generic walk over a caller's view.
template <class G>
void consensus_ids(const G &g, std::vector<uint32_t> &ids) { heaviest_path(g, ids); }
inline void consensus_ids(const GraphAcc &g, std::vector<uint32_t> &ids) { g.g.consensus(ids); }

// src/align.rs:99-162: the windows of a seeded alignment of q against the
// graph's consensus. An empty chain or an unusable anchor gives one
// whole-graph window with the adaptive band. Returns MPOA_OK, or
// MPOA_ERR_SEED_SPAN with the reference's message in err.
template <class G>
int seed_windows(const Params &p, const G &g, const uint8_t *q, uint64_t qlen,
                 std::vector<SeedWindow> &out, std::string &err) {
    out.clear();
    std::vector<uint32_t> ids;
    consensus_ids(g, ids);
    std::string cons(ids.size(), 'N');
    for (size_t k = 0; k < ids.size(); k++) cons[k] = char(p.decode(g.base(ids[k])));
    std::vector<Anchor> an;
    const bool tail = anchor_windows(p, q, qlen, reinterpret_cast<const uint8_t *>(cons.data()),
                                     cons.size(), an);
    auto whole = [&] {
        out.clear();
        SeedWindow w;
        w.qend = qlen;
        w.band = p.b;
        out.push_back(std::move(w));
        return int(MPOA_OK);
    };
    if (an.empty()) return whole();
    uint64_t beg = 0, qbeg = 0;
    bool have_prev = false;
    uint64_t prev_te = 0;
    for (const Anchor &a : an) {
        if (a.target_end + 1 < a.span) {
            err = "invalid minimizer target span";
            return MPOA_ERR_SEED_SPAN;
        }
        if (a.query_end + 1 < a.span) {
            err = "invalid minimizer query span";
            return MPOA_ERR_SEED_SPAN;
        }
        const uint64_t ts = a.target_end + 1 - a.span, qs = a.query_end + 1 - a.span;
        bool ok = a.span > 0 && a.target_end < ids.size() && a.query_end < qlen && qs >= qbeg &&
                  (!have_prev || ts > prev_te);
        for (uint64_t o = 0; ok && o < a.span; o++)
            ok = p.encode(uint8_t(cons[ts + o])) == p.encode(q[qs + o]);
        // Hash collisions or overlapping anchors cannot constrain the DP.
        if (!ok) return whole();
        SeedWindow w;
        w.beg = beg;
        w.end = ids[ts];
        w.qbeg = qbeg;
        w.qend = qs;
        w.adaptive = a.adaptive_before;
        w.band = w.adaptive ? p.b : p.win_band;
        w.anchor.assign(ids.begin() + ts, ids.begin() + a.target_end);
        out.push_back(std::move(w));
        beg = ids[a.target_end];
        qbeg = a.query_end + 1;
        have_prev = true;
        prev_te = a.target_end;
    }
    SeedWindow last;
    last.beg = beg;
    last.end = 1;
    last.qbeg = qbeg;
    last.qend = qlen;
    last.adaptive = tail;
    last.band = tail ? p.b : p.win_band;
    out.push_back(std::move(last));
    return MPOA_OK;
}

// src/align.rs:162-187: a window's path followed by its anchor nodes.
template <class G>
void append_window(const G &g, const SeedWindow &w, const std::vector<AlignedBase> &path,
                   std::vector<AlignedBase> &out) {
    out.insert(out.end(), path.begin(), path.end());
    for (uint32_t id : w.anchor) out.push_back(AlignedBase{int64_t(id), int64_t(g.base(id)), -1});
}

}  // namespace mpoa
