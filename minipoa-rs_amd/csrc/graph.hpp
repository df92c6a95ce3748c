// Host-side POA graph and parameters: the data the device DP reads and the
// fusion/rendering that consume its paths. Semantics follow minipoa-rs
// src/graph.rs and src/lib.rs; adjacency order and ranks decide DP row order
// and every tie, so they are kept exactly.
#pragma once
#include <cstdint>
#include <cstdlib>
#include <string>
#include <vector>

#include "minipoa_hip.h"

namespace mpoa {

// src/lib.rs:47-71 (the fields the DP path reads), owned copy.
struct Params {
    int32_t match_score = 2, mismatch = -4, gap_open = -4, gap_ext = -2;
    uint64_t b = 100, f = 40;
    bool adaptive = false, seeding = false, progressive = false;
    uint64_t k = 19, w = 10, threads = 1, win_band = 500, win_max = 10000;
    double mm_filter_ratio = 0.25;
    uint64_t max_occ = 0, tr_dist = 1000;
    uint8_t result = 0, verbose = 0;
    bool is_rna = false, scalar_dp = false;
    std::vector<uint8_t> alphabet;
    std::vector<int32_t> matrix;
    bool matrix_valid = true;
    uint8_t codes[256] = {};  // encode() of every byte, filled by from_c
    bool coded = false;

    static Params from_c(const mpoa_params &p);
    size_t alpha() const { return alphabet.size(); }
    // src/lib.rs:110-126
    uint32_t encode(uint8_t c) const { return coded ? codes[c] : encode_slow(c); }
    uint32_t encode_slow(uint8_t c) const;
    // src/lib.rs:127-134
    uint8_t decode(uint32_t b) const;
    // src/lib.rs:135-137
    int32_t score(uint32_t a, uint32_t b) const { return matrix[size_t(a) * alpha() + b]; }
    // src/lib.rs:149-173; returns empty string when valid.
    std::string validate() const;
};

// src/lib.rs:140-145 (node/aligned -1 = None).
using AlignedBase = mpoa_aligned_base;

// One adjacency entry: neighbour id and edge weight.
struct Edge {
    uint32_t node, w;
};

// Adjacency list in insertion order with three entries inline: most POA
// nodes have one to three neighbours, so fusion allocates nothing per node.
class Edges {
public:
    Edges() : n_(0), cap_(kInline) {}
    Edges(const Edges &o) : n_(0), cap_(kInline) { *this = o; }
    Edges(Edges &&o) noexcept : n_(o.n_), cap_(o.cap_), u_(o.u_) {
        o.n_ = 0;
        o.cap_ = kInline;
    }
    Edges &operator=(const Edges &o);
    Edges &operator=(Edges &&o) noexcept;
    ~Edges() {
        if (cap_ > kInline) free(u_.heap);
    }
    uint32_t size() const { return n_; }
    bool empty() const { return n_ == 0; }
    Edge *data() { return cap_ > kInline ? u_.heap : u_.inl; }
    const Edge *data() const { return cap_ > kInline ? u_.heap : u_.inl; }
    Edge &operator[](size_t k) { return data()[k]; }
    const Edge &operator[](size_t k) const { return data()[k]; }
    const Edge *begin() const { return data(); }
    const Edge *end() const { return data() + n_; }
    int find(uint32_t node) const {
        const Edge *d = data();
        for (uint32_t k = 0; k < n_; k++)
            if (d[k].node == node) return int(k);
        return -1;
    }
    void push(uint32_t node, uint32_t w) {
        if (n_ == cap_) grow();
        data()[n_++] = Edge{node, w};
    }

private:
    static constexpr uint32_t kInline = 3;
    void grow();
    uint32_t n_, cap_;
    union U {
        Edge inl[kInline];
        Edge *heap;
    } u_;
};

// src/graph.rs:6-27. Node ids are dense indices; adjacency keeps insertion
// order. `aligned` lives in one flat table of n_nodes x alphabet entries.
class Graph {
public:
    explicit Graph(const Params &p);

    struct Node {
        uint32_t rank = 0, parent = 0, base = 0;
        uint64_t coverage = 0;
        Edges in, out;  // (neighbour, weight) in insertion order
    };

    size_t size() const { return nodes_.size(); }
    const Node &node(size_t id) const { return nodes_[id]; }
    // nodes[id].rank from a flat array (cache-friendly for planning).
    uint32_t rank_of(size_t id) const { return rank_flat_[id]; }
    int64_t aligned(size_t id, size_t b) const { return aligned_[id * alpha_ + b]; }
    const std::vector<uint32_t> &order() const { return rank_; }
    const std::vector<std::pair<uint64_t, std::vector<uint32_t>>> &paths() const { return paths_; }
    size_t alpha() const { return alpha_; }

    uint32_t add_node(uint32_t base);                                  // graph.rs:42-61
    void add_edge(uint32_t from, uint32_t to);                         // graph.rs:63-80
    // add_edge with weight w (0 allowed: the edge exists before any read
    // traverses it); used to load graphs from GFA links.
    void link(uint32_t from, uint32_t to, uint32_t w);
    // Every edge of weight 0 gets weight 1 (links that no read traversed).
    void weigh_unused_edges();
    std::string add_path(uint64_t seq_id, const AlignedBase *a, size_t len);  // graph.rs:82-122
    std::string toposort(bool grouped);                                // graph.rs:124-178
    // graph.rs:180-237: ids of the heaviest path (source/sink excluded).
    void consensus(std::vector<uint32_t> &ids) const;
    std::string consensus_bytes(const Params &p) const;
    // graph.rs:239-279 (restores ranks afterwards, success or not).
    std::string msa(const Params &p, const std::vector<std::string> &names,
                    const std::vector<std::string> &comments, std::string &out);
    // graph.rs:281-331
    std::string gfa(const Params &p, const std::vector<std::string> &names, std::string &out) const;

    // Flat node-indexed view (mpoa_graph_view) rebuilt on demand.
    mpoa_graph_view view();

    // Replaces the graph with a device-resident job's packed arrays (prior:
    // paths of sequences the graph held before, kept first): per node
    // its base, parent, rank, coverage and the first records of its out- and
    // in-list; edge records of five words (from, to, weight, next out, next
    // in; 0xFFFFFFFF ends a list) linking each list in insertion order; and
    // nread paths (path k: ids path_off[k] .. path_off[k + 1]) with sequence
    // ids first + k. `aligned` follows from parents and bases: every node
    // holds itself under its base, and a node whose parent is another node
    // is that column's member for its base (the only writes of
    // src/graph.rs:42-61, 95-104).
    void assign_packed(uint32_t nn, const uint8_t *base, const uint32_t *parent, const uint32_t *rank,
                       const uint32_t *cover, const uint32_t *out_head, const uint32_t *in_head,
                       const uint32_t *edges5, const uint32_t *path_off, const uint32_t *paths,
                       uint32_t nread, uint32_t first,
                       const std::vector<std::pair<uint64_t, std::vector<uint32_t>>> *prior = nullptr);

private:
    size_t alpha_;
    std::vector<Node> nodes_;
    std::vector<int64_t> aligned_;
    std::vector<uint32_t> rank_;
    std::vector<uint32_t> rank_flat_;
    std::vector<std::pair<uint64_t, std::vector<uint32_t>>> paths_;
    // toposort(false) scratch: out-edges as flat arrays, in-degrees, stack
    std::vector<uint32_t> ts_off_, ts_adj_, ts_deg_, ts_stack_;
    // view buffers
    std::vector<uint64_t> v_rank_, v_parent_, v_base_, v_in_off_, v_in_, v_in_w_, v_out_off_,
        v_out_, v_out_w_, v_order_;
};

}  // namespace mpoa
