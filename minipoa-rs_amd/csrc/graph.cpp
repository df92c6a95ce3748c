// Host POA graph: fusion, LIFO-Kahn ranking and rendering with the exact
// ordering rules of minipoa-rs src/graph.rs (cited per function).
#include "graph.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <deque>

namespace mpoa {

Edges &Edges::operator=(const Edges &o) {
    if (this == &o) return *this;
    n_ = 0;
    while (cap_ < o.n_) grow();
    const Edge *src = o.data();
    Edge *dst = data();
    for (uint32_t k = 0; k < o.n_; k++) dst[k] = src[k];
    n_ = o.n_;
    return *this;
}

Edges &Edges::operator=(Edges &&o) noexcept {
    if (this == &o) return *this;
    if (cap_ > kInline) free(u_.heap);
    n_ = o.n_;
    cap_ = o.cap_;
    u_ = o.u_;
    o.n_ = 0;
    o.cap_ = kInline;
    return *this;
}

void Edges::grow() {
    const uint32_t ncap = cap_ * 2;
    Edge *h = static_cast<Edge *>(malloc(sizeof(Edge) * ncap));
    if (!h) abort();
    const Edge *d = data();
    for (uint32_t k = 0; k < n_; k++) h[k] = d[k];
    if (cap_ > kInline) free(u_.heap);
    u_.heap = h;
    cap_ = ncap;
}

Params Params::from_c(const mpoa_params &c) {
    Params p;
    p.match_score = c.match_score;
    p.mismatch = c.mismatch;
    p.gap_open = c.gap_open;
    p.gap_ext = c.gap_ext;
    p.b = c.b;
    p.f = c.f;
    p.adaptive = c.adaptive != 0;
    p.seeding = c.seeding != 0;
    p.progressive = c.progressive != 0;
    p.k = c.k;
    p.w = c.w;
    p.threads = c.threads;
    p.win_band = c.win_band;
    p.win_max = c.win_max;
    p.mm_filter_ratio = c.mm_filter_ratio;
    p.max_occ = c.max_occ;
    p.tr_dist = c.tr_dist;
    p.result = c.result;
    p.verbose = c.verbose;
    p.is_rna = c.is_rna != 0;
    p.scalar_dp = c.scalar_dp != 0;
    if (c.alphabet && c.alphabet_len) p.alphabet.assign(c.alphabet, c.alphabet + c.alphabet_len);
    p.matrix_valid = c.matrix != nullptr || c.alphabet_len == 0;
    if (c.matrix && c.alphabet_len)
        p.matrix.assign(c.matrix, c.matrix + size_t(c.alphabet_len) * c.alphabet_len);
    if (!p.alphabet.empty() && p.alphabet.size() <= 256) {
        for (int b = 0; b < 256; b++) p.codes[b] = uint8_t(p.encode_slow(uint8_t(b)));
        p.coded = true;
    }
    return p;
}

uint32_t Params::encode_slow(uint8_t c) const {
    if (c >= 'a' && c <= 'z') c = uint8_t(c - 'a' + 'A');
    if (alphabet.size() <= 5 && c == 'U') c = 'T';
    for (size_t i = 0; i < alphabet.size(); i++)
        if (alphabet[i] == c) return uint32_t(i);
    for (size_t i = 0; i < alphabet.size(); i++)
        if (alphabet[i] == 'N') return uint32_t(i);
    return uint32_t(alphabet.size() - 1);
}

uint8_t Params::decode(uint32_t b) const {
    uint8_t c = alphabet[b];
    return (is_rna && alphabet.size() <= 5 && c == 'T') ? uint8_t('U') : c;
}

std::string Params::validate() const {
    size_t n = alphabet.size();
    if (n == 0 || n > 256 || !matrix_valid || matrix.size() != n * n)
        return "invalid scoring matrix dimensions";
    if (threads == 0) return "thread count must be positive";
    if (!(k >= 1 && k <= 31) || !(w >= 1 && w < 32768))
        return "k must be 1..31 and minimizer window must be 1..32767";
    if (gap_open > 0 || gap_ext > 0) return "gap penalties must be zero or negative";
    if (result > 2 || verbose > 2) return "result format and verbosity must be 0..2";
    if (!std::isfinite(mm_filter_ratio) || mm_filter_ratio < 0.0 || mm_filter_ratio > 1.0)
        return "minimizer filter ratio must be between 0 and 1";
    if (win_max == 0) return "window maximum must be positive";
    return {};
}

// graph.rs:30-40: source (id 0) and sink (id 1), both base N.
Graph::Graph(const Params &p) : alpha_(p.alpha()) {
    uint32_t n = p.alpha() ? p.encode('N') : 0;
    add_node(n);
    add_node(n);
}

uint32_t Graph::add_node(uint32_t base) {
    uint32_t id = uint32_t(nodes_.size());
    Node nd;
    nd.parent = id;
    nd.base = base;
    nodes_.push_back(std::move(nd));
    rank_flat_.push_back(0);
    aligned_.resize(aligned_.size() + alpha_, -1);
    if (alpha_) aligned_[size_t(id) * alpha_ + base] = id;
    return id;
}

// graph.rs:63-80: self edges ignored; repeated edges add weight.
void Graph::add_edge(uint32_t from, uint32_t to) {
    if (from == to) return;
    Node &f = nodes_[from];
    int k = f.out.find(to);
    if (k >= 0)
        f.out[size_t(k)].w += 1;
    else
        f.out.push(to, 1);
    Node &t = nodes_[to];
    k = t.in.find(from);
    if (k >= 0)
        t.in[size_t(k)].w += 1;
    else
        t.in.push(from, 1);
    t.coverage += 1;
}

void Graph::link(uint32_t from, uint32_t to, uint32_t w) {
    if (from == to) return;
    Node &f = nodes_[from];
    int k = f.out.find(to);
    if (k >= 0)
        f.out[size_t(k)].w += w;
    else
        f.out.push(to, w);
    Node &t = nodes_[to];
    k = t.in.find(from);
    if (k >= 0)
        t.in[size_t(k)].w += w;
    else
        t.in.push(from, w);
    t.coverage += w;
}

void Graph::weigh_unused_edges() {
    for (Node &n : nodes_) {
        for (uint32_t k = 0; k < n.out.size(); k++)
            if (n.out[k].w == 0) n.out[k].w = 1;
        for (uint32_t k = 0; k < n.in.size(); k++)
            if (n.in[k].w == 0) {
                n.in[k].w = 1;
                n.coverage += 1;
            }
    }
}

// graph.rs:82-122: reuse the column node of `aligned`'s parent for the base
// or create it; the path runs from the source and ends with an edge to the sink.
std::string Graph::add_path(uint64_t seq_id, const AlignedBase *a, size_t len) {
    for (size_t k = 0; k < len; k++) {
        if (a[k].base < 0 || uint64_t(a[k].base) >= alpha_) return "alignment base outside alphabet";
        if ((a[k].node >= 0 && uint64_t(a[k].node) >= nodes_.size()) ||
            (a[k].aligned >= 0 && uint64_t(a[k].aligned) >= nodes_.size()))
            return "alignment references an unknown node";
    }
    bool has_prev = false;
    uint32_t prev = 0;
    std::vector<uint32_t> path;
    path.reserve(len);
    for (size_t k = 0; k < len; k++) {
        uint32_t base = uint32_t(a[k].base);
        uint32_t id;
        if (a[k].node >= 0) {
            id = uint32_t(a[k].node);
        } else if (a[k].aligned >= 0) {
            uint32_t parent = nodes_[size_t(a[k].aligned)].parent;
            int64_t col = aligned_[size_t(parent) * alpha_ + base];
            if (col >= 0) {
                id = uint32_t(col);
            } else {
                id = add_node(base);
                nodes_[id].parent = parent;
                aligned_[size_t(parent) * alpha_ + base] = id;
            }
        } else {
            id = add_node(base);
        }
        if (has_prev) add_edge(prev, id);
        prev = id;
        has_prev = true;
        if (id >= 2) path.push_back(id);
    }
    add_edge(has_prev ? prev : 0, 1);
    paths_.emplace_back(seq_id, std::move(path));
    return {};
}

// graph.rs:124-178: Kahn's algorithm with a LIFO stack seeded in id order.
// `grouped` ranks whole column groups (parents) for MSA rendering.
std::string Graph::toposort(bool grouped) {
    const size_t nn = nodes_.size();
    if (!grouped) {
        // The same LIFO Kahn walk as below, on a flat copy of the out-edges
        // and in-degrees made by one sequential pass: the walk then touches
        // 4-byte entries instead of whole node records in pop order, and
        // ranks go back to the nodes in id order.
        ts_off_.resize(nn + 1);
        ts_deg_.resize(nn);
        ts_adj_.clear();
        ts_stack_.clear();
        for (size_t id = 0; id < nn; id++) {
            const Node &n = nodes_[id];
            ts_off_[id] = uint32_t(ts_adj_.size());
            for (const Edge &ed : n.out) ts_adj_.push_back(ed.node);
            ts_deg_[id] = uint32_t(n.in.size());
            if (ts_deg_[id] == 0) ts_stack_.push_back(uint32_t(id));
        }
        ts_off_[nn] = uint32_t(ts_adj_.size());
        rank_.clear();
        rank_.reserve(nn);
        while (!ts_stack_.empty()) {
            const uint32_t id = ts_stack_.back();
            ts_stack_.pop_back();
            rank_flat_[id] = uint32_t(rank_.size());
            rank_.push_back(id);
            for (uint32_t k = ts_off_[id]; k < ts_off_[id + 1]; k++)
                if (--ts_deg_[ts_adj_[k]] == 0) ts_stack_.push_back(ts_adj_[k]);
        }
        for (size_t id = 0; id < nn; id++) nodes_[id].rank = rank_flat_[id];
        if (rank_.size() != nn) return "graph contains a cycle (possibly between alignment columns)";
        return {};
    }
    std::vector<uint32_t> degree(nn, 0);
    std::vector<uint32_t> stack;
    stack.reserve(64);
    for (size_t id = 0; id < nn; id++) {
        const Node &n = nodes_[id];
        if (grouped && n.parent != id) continue;
        if (grouped) {
            uint32_t s = 0;
            for (size_t b = 0; b < alpha_; b++) {
                int64_t mbr = aligned_[id * alpha_ + b];
                if (mbr >= 0) s += uint32_t(nodes_[size_t(mbr)].in.size());
            }
            degree[id] = s;
        } else {
            degree[id] = uint32_t(n.in.size());
        }
        if (degree[id] == 0) stack.push_back(uint32_t(id));
    }
    rank_.clear();
    rank_.reserve(nn);
    size_t visited = 0;
    uint32_t single[1];
    std::vector<uint32_t> members;
    while (!stack.empty()) {
        uint32_t id = stack.back();
        stack.pop_back();
        uint32_t rank = uint32_t(rank_.size());
        rank_.push_back(id);
        const uint32_t *mb;
        size_t nm;
        if (grouped) {
            members.clear();
            for (size_t b = 0; b < alpha_; b++) {
                int64_t mbr = aligned_[size_t(id) * alpha_ + b];
                if (mbr >= 0) members.push_back(uint32_t(mbr));
            }
            mb = members.data();
            nm = members.size();
        } else {
            single[0] = id;
            mb = single;
            nm = 1;
        }
        for (size_t k = 0; k < nm; k++) {
            Node &mem = nodes_[mb[k]];
            mem.rank = rank;
            visited++;
            rank_flat_[mb[k]] = rank;
            for (const Edge &ed : mem.out) {
                uint32_t next = ed.node;
                if (grouped) next = nodes_[next].parent;
                if (degree[next] == 0) return "invalid graph column dependencies";
                if (--degree[next] == 0) stack.push_back(next);
            }
        }
    }
    if (visited != nn) return "graph contains a cycle (possibly between alignment columns)";
    return {};
}

// graph.rs:180-237: FIFO Kahn from the source; each node keeps the incoming
// edge of largest weight, then largest accumulated score, then lowest index.
void Graph::consensus(std::vector<uint32_t> &ids) const {
    const size_t nn = nodes_.size();
    std::vector<uint32_t> degree(nn);
    std::vector<int64_t> previous(nn, -1);
    std::vector<uint64_t> score(nn, 0);
    std::vector<uint32_t> queue;
    queue.reserve(nn);
    for (size_t i = 0; i < nn; i++) degree[i] = uint32_t(nodes_[i].in.size());
    queue.push_back(0);
    for (size_t qh = 0; qh < queue.size(); qh++) {
        uint32_t id = queue[qh];
        const Node &node = nodes_[id];
        if (id != 0 && !node.in.empty()) {
            size_t best = 0;
            for (size_t k = 1; k < node.in.size(); k++) {
                uint32_t wb = node.in[best].w, wk = node.in[k].w;
                uint64_t sb = score[node.in[best].node], sk = score[node.in[k].node];
                bool keep = wb != wk ? wb > wk : (sb != sk ? sb > sk : true);
                if (!keep) best = k;
            }
            previous[id] = node.in[best].node;
            score[id] = score[node.in[best].node] + node.in[best].w;
        }
        if (id == 1) break;
        for (const Edge &ed : node.out)
            if (--degree[ed.node] == 0) queue.push_back(ed.node);
    }
    ids.clear();
    for (int64_t cur = previous[1]; cur > 0; cur = previous[size_t(cur)]) ids.push_back(uint32_t(cur));
    std::reverse(ids.begin(), ids.end());
}

std::string Graph::consensus_bytes(const Params &p) const {
    std::vector<uint32_t> ids;
    consensus(ids);
    std::string s(ids.size(), 'N');
    for (size_t k = 0; k < ids.size(); k++) s[k] = char(p.decode(nodes_[ids[k]].base));
    return s;
}

// graph.rs:239-279: rank column groups, render one row per path, then restore
// the per-node ranks and rank order whatever the outcome.
std::string Graph::msa(const Params &p, const std::vector<std::string> &names,
                       const std::vector<std::string> &comments, std::string &out) {
    out.clear();
    if (paths_.empty()) return {};
    std::vector<uint32_t> saved_order = rank_;
    std::vector<uint32_t> saved_rank(nodes_.size());
    for (size_t i = 0; i < nodes_.size(); i++) saved_rank[i] = nodes_[i].rank;
    std::string err = toposort(true);
    if (err.empty()) {
        size_t width = rank_.size() >= 2 ? rank_.size() - 2 : 0;
        std::string row;
        for (const auto &pa : paths_) {
            if (pa.first >= names.size()) {
                err = "unknown sequence ID in graph path";
                break;
            }
            row.assign(width, '-');
            bool bad = false;
            for (uint32_t id : pa.second) {
                uint32_t r = nodes_[nodes_[id].parent].rank;
                if (r == 0 || r > width) {
                    bad = true;
                    break;
                }
                row[r - 1] = char(p.decode(nodes_[id].base));
            }
            if (bad) {
                err = "graph path is outside MSA columns";
                break;
            }
            out += '>';
            out += names[pa.first];
            out += ' ';
            if (pa.first < comments.size()) out += comments[pa.first];
            out += '\n';
            out += row;
            out += '\n';
        }
    }
    rank_ = std::move(saved_order);
    for (size_t i = 0; i < nodes_.size(); i++) nodes_[i].rank = rank_flat_[i] = saved_rank[i];
    if (!err.empty()) out.clear();
    return err;
}

// graph.rs:281-331: segments and links in FIFO-Kahn order, then one P line
// per path; GFA ids are node ids minus one.
std::string Graph::gfa(const Params &p, const std::vector<std::string> &names,
                       std::string &out) const {
    out.clear();
    if (nodes_.size() <= 2 && paths_.empty()) return {};
    out = "H\tVN:Z:1.0\n";
    const size_t nn = nodes_.size();
    std::vector<uint32_t> degree(nn);
    std::vector<uint32_t> queue;
    queue.reserve(nn);
    for (size_t i = 0; i < nn; i++) degree[i] = uint32_t(nodes_[i].in.size());
    queue.push_back(0);
    char buf[96];
    for (size_t qh = 0; qh < queue.size(); qh++) {
        uint32_t id = queue[qh];
        if (id == 1) break;
        const Node &node = nodes_[id];
        if (id >= 2) {
            int k = snprintf(buf, sizeof buf, "S\t%u\t%c\n", id - 1, char(p.decode(node.base)));
            out.append(buf, size_t(k));
            for (const Edge &ed : node.in) {
                const uint32_t prev = ed.node;
                if (prev == 0) continue;
                k = snprintf(buf, sizeof buf, "L\t%u\t+\t%u\t+\t0M\n", prev - 1, id - 1);
                out.append(buf, size_t(k));
            }
        }
        for (const Edge &ed : node.out)
            if (--degree[ed.node] == 0) queue.push_back(ed.node);
    }
    for (const auto &pa : paths_) {
        if (pa.first >= names.size()) {
            out.clear();
            return "unknown sequence ID in graph path";
        }
        out += "P\t";
        out += names[pa.first];
        out += '\t';
        if (pa.second.empty()) out += '*';
        for (size_t k = 0; k < pa.second.size(); k++) {
            if (k) out += ',';
            int n = snprintf(buf, sizeof buf, "%u+", pa.second[k] - 1);
            out.append(buf, size_t(n));
        }
        out += "\t*\n";
    }
    return {};
}

mpoa_graph_view Graph::view() {
    const size_t nn = nodes_.size();
    v_rank_.resize(nn);
    v_parent_.resize(nn);
    v_base_.resize(nn);
    v_in_off_.assign(nn + 1, 0);
    v_out_off_.assign(nn + 1, 0);
    v_in_.clear();
    v_in_w_.clear();
    v_out_.clear();
    v_out_w_.clear();
    for (size_t i = 0; i < nn; i++) {
        const Node &n = nodes_[i];
        v_rank_[i] = n.rank;
        v_parent_[i] = n.parent;
        v_base_[i] = n.base;
        for (const Edge &ed : n.in) {
            v_in_.push_back(ed.node);
            v_in_w_.push_back(ed.w);
        }
        for (const Edge &ed : n.out) {
            v_out_.push_back(ed.node);
            v_out_w_.push_back(ed.w);
        }
        v_in_off_[i + 1] = v_in_.size();
        v_out_off_[i + 1] = v_out_.size();
    }
    v_order_.assign(rank_.begin(), rank_.end());
    mpoa_graph_view v{};
    v.n_nodes = nn;
    v.alphabet_len = alpha_;
    v.rank = v_rank_.data();
    v.parent = v_parent_.data();
    v.base = v_base_.data();
    v.in_off = v_in_off_.data();
    v.in_node = v_in_.data();
    v.in_weight = v_in_w_.data();
    v.out_off = v_out_off_.data();
    v.out_node = v_out_.data();
    v.out_weight = v_out_w_.data();
    v.aligned = aligned_.data();
    v.n_order = v_order_.size();
    v.order = v_order_.data();
    return v;
}

void Graph::assign_packed(uint32_t nn, const uint8_t *base, const uint32_t *parent, const uint32_t *rank,
                          const uint32_t *cover, const uint32_t *out_head, const uint32_t *in_head,
                          const uint32_t *edges5, const uint32_t *path_off, const uint32_t *paths,
                          uint32_t nread, uint32_t first,
                          const std::vector<std::pair<uint64_t, std::vector<uint32_t>>> *prior) {
    constexpr uint32_t nil = 0xFFFFFFFFu;
    nodes_.clear();
    nodes_.resize(nn);
    aligned_.assign(size_t(nn) * alpha_, -1);
    rank_.assign(nn, 0);
    rank_flat_.assign(nn, 0);
    for (uint32_t id = 0; id < nn; id++) {
        Node &nd = nodes_[id];
        nd.base = base[id];
        nd.parent = parent[id];
        nd.rank = rank[id];
        nd.coverage = cover[id];
        for (uint32_t e = out_head[id]; e != nil; e = edges5[size_t(e) * 5 + 3])
            nd.out.push(edges5[size_t(e) * 5 + 1], edges5[size_t(e) * 5 + 2]);
        for (uint32_t e = in_head[id]; e != nil; e = edges5[size_t(e) * 5 + 4])
            nd.in.push(edges5[size_t(e) * 5 + 0], edges5[size_t(e) * 5 + 2]);
        aligned_[size_t(id) * alpha_ + nd.base] = id;
        if (nd.parent != id) aligned_[size_t(nd.parent) * alpha_ + nd.base] = id;
        rank_flat_[id] = rank[id];
        if (rank[id] < nn) rank_[rank[id]] = id;
    }
    paths_.clear();
    if (prior) paths_ = *prior;
    for (uint32_t k = 0; k < nread; k++)
        paths_.emplace_back(uint64_t(first) + k,
                            std::vector<uint32_t>(paths + path_off[k], paths + path_off[k + 1]));
}

}  // namespace mpoa
