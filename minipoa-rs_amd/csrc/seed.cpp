// Minimizer seeding (seed.hpp). Semantics: minipoa-rs src/minimizer.rs.
#include "seed.hpp"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <unordered_map>

namespace mpoa {

namespace {

bool dna_base(uint8_t c) {
    switch (c | 0x20) {  // lower case
    case 'a': case 'c': case 'g': case 't': case 'u': return true;
    default: return false;
    }
}

// packed-seq's 2-bit code: (byte >> 1) & 3 gives A 0, C 1, T/U 2, G 3 in
// either case.
inline uint32_t code2(uint8_t c) { return (c >> 1) & 3u; }

inline uint32_t rot(uint32_t x, uint32_t r) {
    r &= 31u;
    return r ? (x << r) | (x >> (32u - r)) : x;
}

// ntHash per-base seeds (the published 64-bit seeds, low 32 bits), by code.
constexpr uint32_t kSeed[4] = {0x95c60474u, 0x62a02b4cu, 0x4be24456u, 0x82572324u};

// One run of DNA, n >= k bases at offset `at`: keys packed with their k-mer
// index (key << 32 | index) so that the minimum of a window is its leftmost
// smallest key; windows use block prefix/suffix minima (van Herk /
// Gil-Werman), three comparisons per k-mer whatever w is.
void sketch_run(const uint8_t *s, uint64_t n, uint32_t k, uint64_t w, uint64_t at,
                std::vector<Minimizer> &out) {
    const uint64_t nk = n - k + 1;
    std::vector<uint64_t> key(nk);
    uint32_t h = 0;
    for (uint32_t i = 0; i < k; i++) h ^= rot(kSeed[code2(s[i])], k - 1 - i);
    key[0] = uint64_t(h >> 16) << 32;
    for (uint64_t t = 1; t < nk; t++) {
        h = rot(h, 1) ^ rot(kSeed[code2(s[t - 1])], k) ^ kSeed[code2(s[t + k - 1])];
        key[t] = (uint64_t(h >> 16) << 32) | t;
    }
    std::vector<uint64_t> pre(nk), suf(nk);
    for (uint64_t b = 0; b < nk; b += w) {
        const uint64_t e = std::min(nk, b + w);
        pre[b] = key[b];
        for (uint64_t t = b + 1; t < e; t++) pre[t] = std::min(pre[t - 1], key[t]);
        suf[e - 1] = key[e - 1];
        for (uint64_t t = e - 1; t-- > b;) suf[t] = std::min(suf[t + 1], key[t]);
    }
    uint64_t last = ~0ull;
    for (uint64_t st = 0; st + w <= nk; st++) {
        const uint64_t en = st + w - 1;
        // [st, en] spans at most two blocks: suffix of st's, prefix of en's.
        const uint64_t m = st % w == 0 ? suf[st] : std::min(suf[st], pre[en]);
        const uint64_t pos = m & 0xffffffffull;
        if (pos == last) continue;
        last = pos;
        uint64_t v = 0;
        for (uint32_t i = 0; i < k; i++) v |= uint64_t(code2(s[pos + i])) << (2 * i);
        out.push_back(Minimizer{v, at + pos + k - 1, k});
    }
}

// Prefix counts of repeat-covered positions (src/minimizer.rs:112-126).
std::vector<uint64_t> coverage(const std::vector<Minimizer> &mm, uint64_t len, uint64_t max_dist) {
    std::vector<std::pair<uint64_t, uint64_t>> iv;
    repeat_intervals(mm, len, max_dist, iv);
    std::vector<uint64_t> c(len + 1, 0);
    for (const auto &x : iv)
        for (uint64_t i = x.first + 1; i <= x.second; i++) c[i] = 1;
    for (uint64_t i = 0; i < len; i++) c[i + 1] += c[i];
    return c;
}

bool in_repeat(const Minimizer &m, const std::vector<uint64_t> &c) {
    return c[m.end + 1] > c[m.end + 1 - m.span];
}

// src/minimizer.rs:178-197; false for None.
bool link_score(const Anchor &a, const Anchor &b, uint64_t k, int32_t &out) {
    if (a.query_end < b.query_end || a.target_end < b.target_end) return false;
    const uint64_t dq = a.query_end - b.query_end, dr = a.target_end - b.target_end;
    if (dq == 0 || dq > 1000) return false;
    const uint64_t delta = dr > dq ? dr - dq : dq - dr;
    const uint64_t dist = std::min(dr, dq);
    int32_t s = int32_t(std::min<uint64_t>(a.span, dist));
    if (delta > 0 || dist > a.span) {
        if (dr == 0) {
            s += 1;
        } else {
            const double lin = 0.01 * double(k) * double(delta);
            const double lg = std::log2(double(delta + 1));
            s -= int32_t(std::min(lin, lg));
        }
    }
    out = s;
    return true;
}

bool by_pos(const Anchor &x, const Anchor &y) {
    return x.target_end != y.target_end ? x.target_end < y.target_end : x.query_end < y.query_end;
}

}  // namespace

void sketch(const uint8_t *seq, uint64_t len, uint64_t k, uint64_t w, std::vector<Minimizer> &out) {
    out.clear();
    if (k < 1 || k > 31 || w < 1 || w >= 32768 || len >= 0xffffffffull) return;
    uint64_t i = 0;
    while (i < len) {
        while (i < len && !dna_base(seq[i])) i++;
        uint64_t j = i;
        while (j < len && dna_base(seq[j])) j++;
        if (j - i >= k) sketch_run(seq + i, j - i, uint32_t(k), std::min(w, j - i - k + 1), i, out);
        i = j + 1;
    }
}

void repeat_intervals(const std::vector<Minimizer> &mm, uint64_t len, uint64_t max_dist,
                      std::vector<std::pair<uint64_t, uint64_t>> &out) {
    out.clear();
    if (max_dist == 0) return;
    std::unordered_map<uint64_t, uint64_t> last;
    last.reserve(mm.size() * 2);
    std::vector<std::pair<uint64_t, uint64_t>> iv;
    for (const Minimizer &m : mm) {
        auto it = last.find(m.value);
        if (it != last.end()) {
            const uint64_t prev = it->second;
            const uint64_t period = m.end > prev ? m.end - prev : 0;
            if (period > 0 && period <= max_dist) {
                uint64_t st = prev + 1;
                st = st > m.span ? st - m.span : 0;
                st = st > period ? st - period : 0;
                iv.emplace_back(st, std::min(m.end + 1 + period, len));
            }
            it->second = m.end;
        } else {
            last.emplace(m.value, m.end);
        }
    }
    std::sort(iv.begin(), iv.end());
    for (const auto &x : iv) {
        if (!out.empty() && x.first <= out.back().second + 50) {
            out.back().second = std::max(out.back().second, x.second);
        } else {
            out.push_back(x);
        }
    }
}

void collect_anchors(const Params &p, const uint8_t *q, uint64_t qlen, const uint8_t *t,
                     uint64_t tlen, std::vector<Anchor> &out) {
    out.clear();
    if (p.alpha() > 5) return;
    std::vector<Minimizer> qm, tm;
    sketch(q, qlen, p.k, p.w, qm);
    sketch(t, tlen, p.k, p.w, tm);
    const std::vector<uint64_t> qc = coverage(qm, qlen, p.tr_dist), tc = coverage(tm, tlen, p.tr_dist);
    // (value, end) sorted; equal values form one group on each side.
    auto keyed = [](const std::vector<Minimizer> &v) {
        std::vector<std::pair<uint64_t, uint32_t>> r(v.size());
        for (uint32_t i = 0; i < v.size(); i++) r[i] = {v[i].value, i};
        std::sort(r.begin(), r.end());
        return r;
    };
    const auto qk = keyed(qm), tk = keyed(tm);
    size_t a = 0, b = 0;
    while (a < tk.size() && b < qk.size()) {
        const uint64_t v = tk[a].first;
        if (v < qk[b].first) {
            a++;
            continue;
        }
        if (qk[b].first < v) {
            b++;
            continue;
        }
        size_t ae = a, be = b;
        while (ae < tk.size() && tk[ae].first == v) ae++;
        while (be < qk.size() && qk[be].first == v) be++;
        if (!(p.max_occ > 0 && (ae - a > p.max_occ || be - b > p.max_occ))) {
            for (size_t x = a; x < ae; x++) {
                const Minimizer &tmz = tm[tk[x].second];
                if (in_repeat(tmz, tc)) continue;
                for (size_t y = b; y < be; y++) {
                    const Minimizer &qmz = qm[qk[y].second];
                    if (in_repeat(qmz, qc)) continue;
                    out.push_back(Anchor{tmz.end, qmz.end, p.k, false});
                }
            }
        }
        a = ae;
        b = be;
    }
    std::sort(out.begin(), out.end(), by_pos);
    out.erase(std::unique(out.begin(), out.end(),
                          [](const Anchor &x, const Anchor &y) {
                              return x.target_end == y.target_end && x.query_end == y.query_end;
                          }),
              out.end());
}

void chain(const std::vector<Anchor> &A, uint64_t k, std::vector<Anchor> &out) {
    out.clear();
    const size_t n = A.size();
    if (n == 0) return;
    constexpr size_t kNone = ~size_t(0);
    std::vector<int32_t> score(n, 0);
    std::vector<size_t> parent(n, kNone), touched(n, kNone);
    // Sparse DP: predecessors within 1000 target bases and 5000 anchors,
    // stopping after 50 non-improving or 25 already-reached candidates.
    size_t lo = 0;
    for (size_t i = 0; i < n; i++) {
        while (lo < i && A[i].target_end > A[lo].target_end + 1000) lo++;
        if (i > 5000) lo = std::max(lo, i - 5000);
        int32_t best = int32_t(A[i].span);
        size_t skip = 0, misses = 0;
        for (size_t j = i; j-- > lo;) {
            int32_t st;
            if (!link_score(A[i], A[j], k, st)) continue;
            const int32_t c = st + score[j];
            if (c > best) {
                best = c;
                parent[i] = j;
                if (skip) skip--;
            } else {
                if (c != best && ++misses > 50) break;
                if (touched[j] == i && ++skip > 25) break;
            }
            if (parent[j] != kNone) touched[parent[j]] = i;
        }
        score[i] = best;
    }
    // Local chains from the best-scoring ends, cut at a 1000-point drop.
    std::vector<uint8_t> child(n, 0), used(n, 0);
    for (size_t i = 0; i < n; i++)
        if (parent[i] != kNone) child[parent[i]] = 1;
    std::vector<size_t> ends;
    for (size_t i = 0; i < n; i++)
        if (!child[i] && score[i] >= 100) ends.push_back(i);
    std::sort(ends.begin(), ends.end(), [&](size_t x, size_t y) {
        return score[x] != score[y] ? score[x] > score[y] : x > y;
    });
    std::vector<std::vector<size_t>> locals;
    std::vector<size_t> run;
    for (size_t e : ends) {
        if (used[e]) continue;
        run.clear();
        size_t cur = e, cut = kNone;
        int32_t best_gain = 0;
        while (cur != kNone && !used[cur]) {
            run.push_back(cur);
            cur = parent[cur];
            const int32_t gain = score[e] - (cur == kNone ? 0 : score[cur]);
            if (gain > best_gain) {
                best_gain = gain;
                cut = cur;
            } else if (best_gain - gain > 1000) {
                cur = cut;
                break;
            }
        }
        if (cur != kNone)
            while (!run.empty() && run.back() <= cur) run.pop_back();
        for (size_t i : run) used[i] = 1;
        const int32_t local = score[e] - (cur == kNone ? 0 : score[cur]);
        if (run.size() >= 3 && local >= 100) locals.emplace_back(run.rbegin(), run.rend());
    }
    if (locals.empty()) return;
    std::sort(locals.begin(), locals.end(),
              [&](const std::vector<size_t> &x, const std::vector<size_t> &y) { return by_pos(A[x.back()], A[y.back()]); });
    // Join compatible local chains; a later chain is trimmed to its first
    // seed beyond both ends of the earlier one.
    const size_t L = locals.size();
    std::vector<int32_t> total(L);
    std::vector<size_t> from(L, kNone), trim(L, 0);
    std::vector<int32_t> pre;
    for (size_t i = 0; i < L; i++) {
        const std::vector<size_t> &c = locals[i];
        pre.assign(c.size(), int32_t(A[c[0]].span));
        for (size_t x = 1; x < c.size(); x++) {
            int32_t st = 0;
            if (!link_score(A[c[x]], A[c[x - 1]], k, st)) st = 0;
            pre[x] = pre[x - 1] + st;
        }
        total[i] = pre.back();
        for (size_t j = i; j-- > 0;) {
            if (A[c[0]].target_end < A[locals[j][0]].target_end) continue;
            const Anchor &tail = A[locals[j].back()];
            size_t off = 0;
            while (off < c.size() && !(A[c[off]].target_end > tail.target_end && A[c[off]].query_end > tail.query_end))
                off++;
            if (off == c.size()) continue;
            const int32_t cand = total[j] + pre.back() - pre[off] + int32_t(A[c[off]].span);
            if (cand > total[i]) {
                total[i] = cand;
                from[i] = j;
                trim[i] = off;
            }
        }
    }
    size_t best = 0;
    for (size_t i = 1; i < L; i++)
        if (total[i] > total[best]) best = i;
    std::vector<size_t> picked;
    for (size_t i = best; i != kNone; i = from[i]) picked.push_back(i);
    for (size_t r = picked.size(); r-- > 0;) {
        const std::vector<size_t> &c = locals[picked[r]];
        for (size_t x = trim[picked[r]]; x < c.size(); x++) {
            const Anchor &a = A[c[x]];
            // Strictly increasing on both sequences.
            if (out.empty() || (a.target_end > out.back().target_end && a.query_end > out.back().query_end))
                out.push_back(a);
        }
    }
}

bool partition_windows(const std::vector<Anchor> &ch, uint64_t tlen, uint64_t qlen,
                       uint64_t max_drift, uint64_t max_win, std::vector<Anchor> &out) {
    out.clear();
    const int64_t n = int64_t(ch.size());
    // Drift compares unsigned; gaps compare against the bounds as signed
    // 64-bit values (the reference's `as i64`).
    const int64_t gap_max = int64_t(max_drift), win_max = int64_t(max_win);
    // Boundary coordinates; index -1 is the start, n the end of both sequences.
    auto t_end = [&](int64_t i) { return i < 0 ? -1 : int64_t(ch[size_t(i)].target_end); };
    auto q_end = [&](int64_t i) { return i < 0 ? -1 : int64_t(ch[size_t(i)].query_end); };
    auto t_beg = [&](int64_t i) { return i == n ? int64_t(tlen) : int64_t(ch[size_t(i)].target_end) - int64_t(ch[size_t(i)].span) + 1; };
    auto q_beg = [&](int64_t i) { return i == n ? int64_t(qlen) : int64_t(ch[size_t(i)].query_end) - int64_t(ch[size_t(i)].span) + 1; };
    auto diag = [&](int64_t i) { return i < 0 ? 0 : (i == n ? int64_t(tlen) - int64_t(qlen) : t_end(i) - q_end(i)); };
    auto absd = [](int64_t x) { return x < 0 ? uint64_t(0) - uint64_t(x) : uint64_t(x); };
    bool tail = false;
    int64_t s = -1;
    while (s < n) {
        int64_t last_ok = -1;
        for (int64_t i = s + 1; i <= n; i++) {
            const int64_t gt = t_beg(i) - t_end(s) - 1, gq = q_beg(i) - q_end(s) - 1;
            const uint64_t drift = absd(diag(i) - diag(s));
            const int64_t gap = std::max(t_beg(i) - t_end(i - 1) - 1, q_beg(i) - q_end(i - 1) - 1);
            if (drift > max_drift || gap > gap_max || std::max(gt, gq) > win_max) break;
            if (gt >= 0 && gq >= 0) last_ok = i;
        }
        int64_t e;
        bool adaptive = false;
        if (last_ok >= 0) {
            e = last_ok;
        } else {
            e = s + 1;
            while (e < n && (t_beg(e) - t_end(s) - 1 < 0 || q_beg(e) - q_end(s) - 1 < 0)) e++;
            const uint64_t drift = absd(diag(e) - diag(s));
            const int64_t gap = std::max(t_beg(e) - t_end(s) - 1, q_beg(e) - q_end(s) - 1);
            adaptive = drift > max_drift || gap > gap_max;
        }
        if (e < n) {
            Anchor a = ch[size_t(e)];
            a.adaptive_before = adaptive;
            out.push_back(a);
        } else {
            tail = adaptive;
        }
        s = e;
    }
    return tail;
}

void guide_order(const Params &p, const uint8_t *const *seqs, const uint64_t *lens, uint64_t n,
                 std::vector<uint32_t> &order) {
    order.resize(n);
    for (uint64_t i = 0; i < n; i++) order[i] = uint32_t(i);
    if (!p.progressive || n < 2 || p.alpha() > 5) return;
    // Every (k-mer value, read) occurrence, sorted: a value's group lists
    // its reads in increasing id, once per occurrence.
    std::vector<std::pair<uint64_t, uint32_t>> occ;
    std::vector<Minimizer> mm;
    for (uint64_t r = 0; r < n; r++) {
        sketch(seqs[r], lens[r], p.k, p.w, mm);
        for (const Minimizer &m : mm) occ.emplace_back(m.value, uint32_t(r));
    }
    if (occ.empty()) return;
    std::sort(occ.begin(), occ.end());
    std::vector<uint64_t> total(n, 0), shared(n * n, 0);
    std::vector<std::pair<uint32_t, uint64_t>> grp;  // (read, count)
    for (size_t a = 0; a < occ.size();) {
        size_t b = a;
        grp.clear();
        for (; b < occ.size() && occ[b].first == occ[a].first; b++) {
            if (!grp.empty() && grp.back().first == occ[b].second) grp.back().second++;
            else grp.emplace_back(occ[b].second, 1);
        }
        a = b;
        // Values shared by more than mm_filter_ratio of the reads are too
        // frequent to tell reads apart.
        if (double(grp.size()) > double(n) * p.mm_filter_ratio) continue;
        for (size_t x = 0; x < grp.size(); x++) {
            total[grp[x].first] += grp[x].second;
            for (size_t y = 0; y < x; y++) {
                const uint64_t c = std::min(grp[x].second, grp[y].second);
                shared[size_t(grp[x].first) * n + grp[y].first] += c;
                shared[size_t(grp[y].first) * n + grp[x].first] += c;
            }
        }
    }
    uint64_t longest = 1;
    for (uint64_t i = 0; i < n; i++) longest = std::max(longest, lens[i]);
    const double max_len = double(longest);
    std::vector<double> sim(n * n, 0.0);
    double best = -1.0;
    uint64_t bi = 0, bj = 1;
    for (uint64_t i = 0; i < n; i++) {
        for (uint64_t j = 0; j < n; j++) {
            if (i == j) continue;
            const uint64_t common = shared[i * n + j];
            const uint64_t uni = total[i] + total[j] - common;
            // Multiset Jaccard, weighted toward the longer read of the pair
            // (evaluated in the reference's order of operations).
            const double jac = uni == 0 ? 0.0
                                        : double(common) / double(uni) *
                                              (0.51 * double(lens[i]) / max_len + 0.49 * double(lens[j]) / max_len);
            sim[i * n + j] = jac;
            if (jac > best) {
                best = jac;
                bi = i;
                bj = j;
            }
        }
    }
    std::vector<uint8_t> placed(n, 0);
    std::vector<double> sum(n);
    order[0] = uint32_t(bi);
    order[1] = uint32_t(bj);
    placed[bi] = placed[bj] = 1;
    for (uint64_t v = 0; v < n; v++) sum[v] = sim[bi * n + v] + sim[bj * n + v];
    for (uint64_t k = 2; k < n; k++) {
        uint64_t next = n;
        for (uint64_t v = 0; v < n; v++)
            if (!placed[v] && (next == n || sum[v] > sum[next])) next = v;
        order[k] = uint32_t(next);
        placed[next] = 1;
        for (uint64_t v = 0; v < n; v++)
            if (!placed[v]) sum[v] += sim[next * n + v];
    }
}

bool anchor_windows(const Params &p, const uint8_t *q, uint64_t qlen, const uint8_t *cons,
                    uint64_t clen, std::vector<Anchor> &out) {
    std::vector<Anchor> m, c;
    collect_anchors(p, q, qlen, cons, clen, m);
    chain(m, p.k, c);
    const uint64_t max_win = std::min<uint64_t>(p.win_max, std::max<uint64_t>(500, qlen / 25));
    return partition_windows(c, clen, qlen, p.win_band, max_win, out);
}

}  // namespace mpoa
