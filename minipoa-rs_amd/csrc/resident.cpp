// Host side of the device-resident batch (resident.hpp): jobs are grouped
// into length buckets; each bucket's jobs live in device memory (reads,
// graphs, DP descriptors) for their whole run, and one progressive step of
// all of them -- device fusion and planning, fill, traceback -- is captured
// once as a HIP graph and replayed until every job has finished. Only the
// final graphs come back, packed, for rendering. Jobs the device path does
// not take (seeding, the guide order, the i64 kernel, oversized DPs) run on
// the continuous host-driven engine, from the read where they left.
//
// Also the CPU check of the resident step's graph and planning code: the
// same templates on one host thread, held against Graph and plan_job.
#include "resident.hpp"
#include "resident_host.hpp"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <thread>

#include "engine.hpp"
#include "kernels.hpp"
#include "plan.hpp"
#include "runtime.hpp"

namespace mpoa {

void launch_rj_step(const RPools &R, hipStream_t s);
void launch_rj_reset(unsigned long long *ctr, hipStream_t s);
void launch_rj_pack(const RJob *jobs, const uint32_t *sel, uint32_t njobs, const uint64_t *off, uint8_t *out,
                    hipStream_t s);


RParams make_rparams(const Params &p, const int32_t *matrix) {
    RParams r{};
    for (int c = 0; c < 256; c++) r.codes[c] = uint8_t(p.encode(uint8_t(c)));
    r.matrix = matrix;
    r.b = p.b;
    r.f = p.f;
    uint64_t largest = 0;
    for (int32_t v : p.matrix) largest = std::max<uint64_t>(largest, uint64_t(std::llabs(int64_t(v))));
    largest = std::max<uint64_t>(largest, uint64_t(std::llabs(int64_t(p.gap_open) + p.gap_ext)));
    largest = std::max<uint64_t>(largest, uint64_t(std::llabs(int64_t(p.gap_ext))));
    r.largest = largest;
    r.gap_open = p.gap_open;
    r.gap_ext = p.gap_ext;
    r.alpha = uint32_t(p.alpha());
    r.unknown = p.encode('N');
    r.score_nn = p.score(r.unknown, r.unknown);
    r.adaptive = p.adaptive ? 1u : 0u;
    r.scalar_dp = p.scalar_dp ? 1u : 0u;
    return r;
}

RCaps job_caps(const uint64_t *roff, uint32_t nreads) {
    RCaps c;
    c.nreads = nreads;
    uint64_t maxlen = 0;
    for (uint32_t r = 0; r < nreads; r++) {
        const uint64_t len = roff[r + 1] - roff[r];
        c.total += len;
        maxlen = std::max(maxlen, len);
    }
    c.nc = uint32_t(2 + c.total);
    c.ec = uint32_t(c.total + nreads + 1);
    c.maxm = uint32_t(maxlen + 2);
    c.padm = (c.maxm / 8 + 1) * 8;
    return c;
}

RGraphLayout graph_layout(const RCaps &c, uint32_t alpha) {
    RGraphLayout L{};
    uint64_t at = 0;
    auto take = [&](uint64_t bytes) {
        const uint64_t o = at;
        at = up16(at + bytes);
        return o;
    };
    const uint64_t nc = c.nc, ec = c.ec;
    L.base = take(nc);
    L.parent = take(4 * nc);
    L.rank = take(4 * nc);
    L.order = take(4 * nc);
    L.aligned = take(4 * nc * alpha);
    L.out_head = take(4 * nc);
    L.out_tail = take(4 * nc);
    L.in_head = take(4 * nc);
    L.in_tail = take(4 * nc);
    L.in_cnt = take(4 * nc);
    L.cover = take(4 * nc);
    L.mark = take(4 * nc);
    L.edges = take(sizeof(REdge) * ec);
    L.paths = take(4 * std::max<uint64_t>(c.total, 1));
    L.path_off = take(4 * (uint64_t(c.nreads) + 1));
    L.csr_off = take(4 * (nc + 1));
    L.csr_adj = take(4 * ec);
    L.deg = take(4 * nc);
    L.stack = take(4 * nc);
    L.succ_beg = take(4 * (nc + 1));
    L.succ = take(sizeof(RSucc) * ec);
    L.head = take(4 * nc);
    L.tail = take(4 * nc);
    L.support = take(4 * nc);
    L.win = take(4 * (kRowWin + 1 + 2 * kEntWin));
    L.bytes = at;
    return L;
}

// Points J's arrays into a block at `base` (device or host memory).
void bind_graph(RJob &J, char *blk, const RGraphLayout &L) {
    J.base = reinterpret_cast<uint8_t *>(blk + L.base);
    J.parent = reinterpret_cast<uint32_t *>(blk + L.parent);
    J.rank = reinterpret_cast<uint32_t *>(blk + L.rank);
    J.order = reinterpret_cast<uint32_t *>(blk + L.order);
    J.aligned = reinterpret_cast<int32_t *>(blk + L.aligned);
    J.out_head = reinterpret_cast<uint32_t *>(blk + L.out_head);
    J.out_tail = reinterpret_cast<uint32_t *>(blk + L.out_tail);
    J.in_head = reinterpret_cast<uint32_t *>(blk + L.in_head);
    J.in_tail = reinterpret_cast<uint32_t *>(blk + L.in_tail);
    J.in_cnt = reinterpret_cast<uint32_t *>(blk + L.in_cnt);
    J.cover = reinterpret_cast<uint32_t *>(blk + L.cover);
    J.mark = reinterpret_cast<uint32_t *>(blk + L.mark);
    J.edges = reinterpret_cast<REdge *>(blk + L.edges);
    J.paths = reinterpret_cast<uint32_t *>(blk + L.paths);
    J.path_off = reinterpret_cast<uint32_t *>(blk + L.path_off);
    J.csr_off = reinterpret_cast<uint32_t *>(blk + L.csr_off);
    J.csr_adj = reinterpret_cast<uint32_t *>(blk + L.csr_adj);
    J.deg = reinterpret_cast<uint32_t *>(blk + L.deg);
    J.stack = reinterpret_cast<uint32_t *>(blk + L.stack);
    J.succ_beg = reinterpret_cast<uint32_t *>(blk + L.succ_beg);
    J.succ = reinterpret_cast<RSucc *>(blk + L.succ);
    J.head = reinterpret_cast<int32_t *>(blk + L.head);
    J.tail = reinterpret_cast<int32_t *>(blk + L.tail);
    J.support = reinterpret_cast<uint32_t *>(blk + L.support);
    J.win = reinterpret_cast<uint32_t *>(blk + L.win);
}

// Message of a failed resident job (the reference's text for its error).
int failure_of(const RJob &J, std::string &msg) {
    switch (J.err) {
        case 0: {
            DevResult r{};
            r.status = J.dev_status;
            r.err_i = J.err_i;
            r.err_j = J.err_j;
            return device_status(r, msg);
        }
        case MPOA_ERR_ANCHORS_NOT_ORDERED: msg = "alignment anchors are not topologically ordered"; break;
        case MPOA_ERR_COORDINATE_LIMITS: msg = "alignment exceeds coordinate limits"; break;
        case MPOA_ERR_NO_PATH: msg = "no graph path connects alignment anchors"; break;
        case MPOA_ERR_GRAPH: msg = "graph contains a cycle (possibly between alignment columns)"; break;
        default: msg = "device-resident job failed"; break;
    }
    return J.err;
}

namespace {

// The team of resident.hpp on one host thread: every team loop is the
// serial loop, every scan the running sum.
struct SoloTeam {
    static constexpr uint32_t lane = 0, size = 1;
    void sync() {}
    uint32_t excl_scan(uint32_t v, uint32_t &tot) {
        tot = v;
        return 0;
    }
    uint64_t reduce_add64(uint64_t v) { return v; }
    uint32_t reduce_max(uint32_t v) { return v; }
    bool any(bool p) { return p; }
    uint32_t xchg(uint32_t *p, uint32_t v) {
        const uint32_t o = *p;
        *p = v;
        return o;
    }
    int32_t cas(int32_t *p, int32_t c, int32_t v) {
        const int32_t o = *p;
        if (o == c) *p = v;
        return o;
    }
    unsigned long long cas64(unsigned long long *p, unsigned long long c, unsigned long long v) {
        const unsigned long long o = *p;
        if (o == c) *p = v;
        return o;
    }
    void add(uint32_t *p, uint32_t v) { *p += v; }
    void umax(uint32_t *p, uint32_t v) {
        if (v > *p) *p = v;
    }
};

// ---------------------------------------------------------------- CPU check

std::string graph_diff(const Graph &g, const RJob &J, uint32_t alpha) {
    char buf[256];
    if (g.size() != J.nn) {
        snprintf(buf, sizeof buf, "node count %zu vs %u", g.size(), J.nn);
        return buf;
    }
    for (uint32_t id = 0; id < J.nn; id++) {
        const Graph::Node &nd = g.node(id);
        if (nd.base != J.base[id] || nd.parent != J.parent[id] || g.rank_of(id) != J.rank[id] ||
            nd.coverage != J.cover[id]) {
            snprintf(buf, sizeof buf, "node %u: base/parent/rank/coverage %u/%u/%u/%llu vs %u/%u/%u/%u", id,
                     nd.base, nd.parent, g.rank_of(id), (unsigned long long)nd.coverage, J.base[id],
                     J.parent[id], J.rank[id], J.cover[id]);
            return buf;
        }
        for (uint32_t a = 0; a < alpha; a++)
            if (g.aligned(id, a) != J.aligned[uint64_t(id) * alpha + a]) {
                snprintf(buf, sizeof buf, "node %u aligned[%u]", id, a);
                return buf;
            }
        uint32_t k = 0;
        for (uint32_t e = J.out_head[id]; e != kNil; e = J.edges[e].next_out, k++)
            if (k >= nd.out.size() || nd.out[k].node != J.edges[e].to || nd.out[k].w != J.edges[e].w) {
                snprintf(buf, sizeof buf, "node %u out-list entry %u", id, k);
                return buf;
            }
        if (k != nd.out.size()) {
            snprintf(buf, sizeof buf, "node %u out-list length", id);
            return buf;
        }
        k = 0;
        for (uint32_t e = J.in_head[id]; e != kNil; e = J.edges[e].next_in, k++)
            if (k >= nd.in.size() || nd.in[k].node != J.edges[e].from || nd.in[k].w != J.edges[e].w) {
                snprintf(buf, sizeof buf, "node %u in-list entry %u", id, k);
                return buf;
            }
        if (k != nd.in.size() || k != J.in_cnt[id]) {
            snprintf(buf, sizeof buf, "node %u in-list length", id);
            return buf;
        }
    }
    for (uint32_t r = 0; r < J.nn; r++)
        if (g.order()[r] != J.order[r]) {
            snprintf(buf, sizeof buf, "order[%u] %u vs %u", r, g.order()[r], J.order[r]);
            return buf;
        }
    const auto &ps = g.paths();
    if (ps.size() != J.cur) return "path count";
    for (uint32_t r = 0; r < J.cur; r++) {
        const auto &ids = ps[r].second;
        const uint32_t b = J.path_off[r], e = J.path_off[r + 1];
        if (ids.size() != e - b || !std::equal(ids.begin(), ids.end(), J.paths + b)) {
            snprintf(buf, sizeof buf, "path %u", r);
            return buf;
        }
    }
    return {};
}

std::string plan_diff(const JobPlan &H, const RJob &J, const RowIn *rows, const RowAux *aux, const PredE *preds,
                      const uint8_t *seq) {
    char buf[256];
    const DevJob &d = J.dj;
    if (H.n != d.n || H.m != d.m || H.padded != d.padded || H.band != d.band || H.flags != d.flags ||
        H.o != d.o || H.e != d.e || H.unknown != d.unknown || H.score_nn != d.score_nn || H.beg != J.beg) {
        snprintf(buf, sizeof buf, "plan header n/m/band/flags %u/%u/%lld/%u vs %u/%u/%lld/%u", H.n, H.m,
                 (long long)H.band, H.flags, d.n, d.m, (long long)d.band, d.flags);
        return buf;
    }
    for (uint32_t i = 0; i <= H.n; i++) {
        const RowIn &a = H.rows[i], &b = rows[i];
        if (a.pred_beg != b.pred_beg || a.head != b.head || a.tail != b.tail || a.last_use != b.last_use) {
            snprintf(buf, sizeof buf, "row %u: pred_beg/head/tail/last_use %u/%d/%d/%u vs %u/%d/%d/%u", i,
                     a.pred_beg, a.head, a.tail, a.last_use, b.pred_beg, b.head, b.tail, b.last_use);
            return buf;
        }
        if (i < H.n && (H.aux[i].thresh != aux[i].thresh || H.aux[i].base != aux[i].base)) {
            snprintf(buf, sizeof buf, "row %u aux", i);
            return buf;
        }
    }
    for (size_t k = 0; k < H.preds.size(); k++) {
        const PredE &a = H.preds[k], &b = preds[k];
        if (a.row != b.row || a.weight != b.weight || a.ord != b.ord || a.rperm != b.rperm) {
            snprintf(buf, sizeof buf, "pred %zu: %u/%u/%u/%u vs %u/%u/%u/%u", k, a.row, a.weight, a.ord,
                     a.rperm, b.row, b.weight, b.ord, b.rperm);
            return buf;
        }
    }
    if (memcmp(H.seq.data(), seq, H.padded) != 0) return "encoded query";
    if (H.est_val != J.est && !(H.est_val < 256 && J.est == 256)) {
        snprintf(buf, sizeof buf, "arena estimate %llu vs %llu", (unsigned long long)H.est_val,
                 (unsigned long long)J.est);
        return buf;
    }
    if (H.max_val != J.emax && !(H.max_val < 256 && J.emax == 256)) return "full-width arena bound";
    return {};
}

}  // namespace

int resident_check(const Params &p, const uint8_t *bases, const uint64_t *read_off, uint32_t nreads,
                   const AlignedBase *paths, const uint64_t *path_off, uint32_t lds_bytes, std::string &report) {
    const uint32_t alpha = uint32_t(p.alpha());
    RCaps c = job_caps(read_off, nreads);
    // Paths are the caller's: size the graph for every element of them.
    uint64_t elems = 0;
    for (uint32_t r = 0; r < nreads; r++) elems += path_off[r + 1] - path_off[r];
    c.nc = uint32_t(std::max<uint64_t>(c.nc, 2 + elems + c.total));
    c.ec = uint32_t(std::max<uint64_t>(c.ec, elems + c.total + nreads + 1));
    c.total = std::max<uint64_t>(c.total, elems);
    const RGraphLayout GL = graph_layout(c, alpha);
    std::vector<char> blk(GL.bytes + 16);
    std::vector<RowIn> rows(size_t(c.nc) + 1);
    std::vector<RowAux> aux(size_t(c.nc) + 1);
    std::vector<PredE> preds(std::max<size_t>(c.ec, 1));
    std::vector<uint8_t> seq(c.padm + 16);
    std::vector<char> lds(std::max<uint32_t>(lds_bytes, 16));
    const RParams P = make_rparams(p, p.matrix.data());
    RJob J{};
    J.bases = bases;
    J.read_off = read_off;
    J.nreads = nreads;
    J.nc = c.nc;
    J.ec = c.ec;
    J.maxm = c.maxm;
    bind_graph(J, blk.data(), GL);
    DevJob dj{};
    DevResult res{};
    unsigned long long ctr = 0;
    uint32_t done = 0;
    RPools R{};
    R.jobs = &J;
    R.dj = &dj;
    R.res = &res;
    R.rows = rows.data();
    R.aux = aux.data();
    R.preds = preds.data();
    R.seq = seq.data();
    R.arena_ctr = &ctr;
    R.done = &done;
    R.prm = &P;
    R.njobs = 1;
    R.lds_bytes = lds_bytes;
    SoloTeam t;
    Params full = p;
    full.f = 0;
    Graph g(p);
    rj_node_init(P, J, 0, P.unknown, 0);
    rj_node_init(P, J, 1, P.unknown, 1);
    J.nn = 2;
    J.path_off[0] = 0;
    char buf[160];
    for (uint32_t r = 0; r < nreads; r++) {
        J.cur = r;
        const uint8_t *q = bases + read_off[r];
        const uint64_t qlen = read_off[r + 1] - read_off[r];
        const AlignedBase *pa = paths + path_off[r];
        const uint32_t L = uint32_t(path_off[r + 1] - path_off[r]);
        if (g.size() > 2) {
            // Both planners on the graph before this read: banded, then the
            // full-DP retry.
            for (int pass = 0; pass < 2; pass++) {
                const bool f0 = pass == 1 || p.f == 0;
                JobPlan H;
                if (f0) plan_job(full, GraphAcc{g}, 0, 1, q, qlen, false, p.b, 0, H);
                else plan_job(p, GraphAcc{g}, 0, 1, q, qlen, p.adaptive, p.b, 0, H);
                const int rc = rj_plan_read(t, R, P, J, f0, lds.data(), lds_bytes);
                if (rc == -1) {
                    if (!(H.status == MPOA_OK && H.use_i64)) {
                        snprintf(buf, sizeof buf, "read %u: device plan left the device (reason %u), host status %d",
                                 r, J.evict, H.status);
                        report = buf;
                        return 1;
                    }
                } else if (rc != H.status || (rc == MPOA_OK && H.use_i64)) {
                    snprintf(buf, sizeof buf, "read %u: plan status %d vs host %d (i64 %d)", r, rc, H.status,
                             int(H.use_i64));
                    report = buf;
                    return 1;
                } else if (rc == MPOA_OK) {
                    std::string d = plan_diff(H, J, rows.data(), aux.data(), preds.data(), seq.data());
                    if (!d.empty()) {
                        report = "read " + std::to_string(r) + (f0 ? " (full DP): " : ": ") + d;
                        return 1;
                    }
                }
                if (p.f == 0) break;
            }
        }
        std::string e = g.add_path(r, pa, L);
        if (!e.empty()) {
            report = "host add_path: " + e;
            return 2;
        }
        const std::string te = g.toposort(false);
        auto elem = [pa](uint32_t k) { return RElem{pa[k].node, pa[k].base, pa[k].aligned}; };
        rj_add_path(t, P, J, elem, L, J.stack, J.deg);
        const bool ok = rj_rerank(t, J, lds.data(), lds_bytes);
        if (ok != te.empty()) {
            report = "read " + std::to_string(r) + ": cycle verdicts differ";
            return 1;
        }
        if (!ok) return 0;  // both found the cycle
        J.cur = r + 1;
        std::string d = graph_diff(g, J, alpha);
        if (!d.empty()) {
            report = "after read " + std::to_string(r) + ": " + d;
            return 1;
        }
    }
    report.clear();
    return 0;
}

// ---------------------------------------------------------------- batch

// Per launch slot: the device blocks and host staging of its buckets,
// sized before the first bucket starts and kept across buckets and runs
// (hipMalloc and hipFree while other streams run stall them).
struct RSlot {
    Ctx *ctx = nullptr;
    DevBuf dg, dp, da, dk;
    PinBuf hin, hjobs, hpack, hoff, hsent, hsjobs;
    hipEvent_t ev[2] = {nullptr, nullptr};
    hipStream_t s2 = nullptr;  // copies of finished jobs while the persistent launch runs
    ~RSlot() {
        for (auto &e : ev)
            if (e) (void)hipEventDestroy(e);
        if (s2) (void)hipStreamDestroy(s2);
        hsent.release();
        hsjobs.release();
        dg.release();
        dp.release();
        da.release();
        dk.release();
        hin.release();
        hjobs.release();
        hpack.release();
        hoff.release();
    }
};

struct ResidentCache {
    std::vector<std::unique_ptr<RSlot>> slots;
    uint64_t arena_bytes = 0;
};

ResidentCache *resident_cache_new() { return new ResidentCache(); }
void resident_cache_free(ResidentCache *c) { delete c; }
void resident_cache_trim(ResidentCache *c) {
    if (!c) return;
    for (auto &s : c->slots) {
        s->dg.release();
        s->dp.release();
        s->da.release();
        s->dk.release();
    }
}
uint64_t resident_cache_bytes(const ResidentCache *c) {
    uint64_t t = 0;
    if (c)
        for (const auto &s : c->slots) t += s->dg.cap + s->dp.cap + s->da.cap + s->dk.cap;
    return t;
}

namespace {

// The persistent kernel's launch: workgroups, bytes of each one's arena
// region, and the step's LDS.
struct PersistCfg {
    uint32_t grid = 0;
    uint64_t slot_bytes = 0;
    uint64_t pool_bytes = 0;  // shared by DPs that outgrow their region
};

struct RShared {
    const Params *p = nullptr;
    const BatchInput *in = nullptr;
    TaskQueue *tasks = nullptr;
    BatchOutput *out = nullptr;
    uint64_t arena_bytes = 0;
    uint32_t nslots = 1;
    // Persistent mode: the stream of a launch slot that the launch leaves
    // idle (its own hardware queue), for copies while the launch runs.
    hipStream_t copy_stream = nullptr;
    std::mutex mu;
    std::vector<std::vector<uint64_t>> buckets;
    std::atomic<size_t> next{0};
    // evicted jobs: (job, cur, graph)
    std::vector<std::tuple<uint64_t, uint32_t, std::unique_ptr<Graph>>> evicted;
    ResidentStats st;
    int rc = MPOA_OK;
    std::string err;
    // MPOA_RES_TRACE=<file>: one CSV line per bucket (host clock in ms from
    // the start of the run at each phase's end), written at the end.
    std::chrono::steady_clock::time_point t0;
    struct Trace {
        size_t slot;
        uint32_t njobs, rmax;
        uint64_t lmax, replays;
        double t[6];
        float dev_ms;
    };
    std::vector<Trace> trace;
    double ms() const {
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
};

#define RJ_TRY(expr)                                                                        \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess) {                                                             \
            set_error(std::string("HIP error: ") + hipGetErrorString(e_) + " at " #expr);   \
            return MPOA_ERR_DEVICE;                                                         \
        }                                                                                   \
    } while (0)

// Launch shape of a bucket from its longest template: the adaptive band's
// expected width (two bands plus a quarter of the query, as measured on
// DNA jobs), or the padded query with full rows.
uint32_t expected_width(const Params &p, uint64_t L) {
    const uint64_t m = L + 2, padded = (m / 8 + 1) * 8;
    if (p.f == 0) return uint32_t(padded);
    const uint64_t band = p.b + m / p.f;
    return uint32_t(std::min<uint64_t>(padded, 2 * band + m / 4 + 16));
}

// Layout of one bucket: block A (parameters, counters, job states, the
// launch's job and result arrays, reads, graphs) and block B (the launch
// pools of the DP descriptors), in bytes.
struct BucketLayout {
    std::vector<RCaps> caps;
    std::vector<RGraphLayout> gl;
    uint32_t maxpad = 8, rmax = 0;
    uint64_t lmax = 0, nmax = 0;
    uint64_t a_prm, a_mat, a_ctr, a_jobs, a_dj, a_res, a_poff, a_sent, a_bases, a_roff, a_graph, a_end;
    uint64_t b_rows, b_aux, b_preds, b_seq, b_meta, b_adapt, b_tb, b_end;
};

BucketLayout layout_bucket(const Params &p, const BatchInput &in, const std::vector<uint64_t> &jobs) {
    BucketLayout B;
    const uint32_t N = uint32_t(jobs.size());
    const uint32_t alpha = uint32_t(p.alpha());
    B.caps.resize(N);
    B.gl.resize(N);
    uint64_t nbases = 0, noffs = 0, gbytes = 0, rows = 0, preds = 0, seqb = 0, meta = 0, tbn = 0;
    for (uint32_t k = 0; k < N; k++) {
        const uint64_t j = jobs[k];
        const uint64_t *ro = in.read_off + in.job_off[j];
        const uint32_t nr = uint32_t(in.job_off[j + 1] - in.job_off[j]);
        B.caps[k] = job_caps(ro, nr);
        B.gl[k] = graph_layout(B.caps[k], alpha);
        nbases += B.caps[k].total;
        noffs += nr + 1;
        gbytes += B.gl[k].bytes;
        rows += uint64_t(B.caps[k].nc) + 1;
        preds += B.caps[k].ec;
        seqb += B.caps[k].padm;
        meta += B.caps[k].nc;
        tbn += B.caps[k].maxm;
        B.maxpad = std::max(B.maxpad, B.caps[k].padm);
        B.rmax = std::max(B.rmax, nr);
        B.nmax = std::max<uint64_t>(B.nmax, B.caps[k].nc);
        if (nr) B.lmax = std::max<uint64_t>(B.lmax, ro[1] - ro[0]);
    }
    B.a_prm = 0;
    B.a_mat = up256(B.a_prm + sizeof(RParams));
    B.a_ctr = up256(B.a_mat + 4ull * alpha * alpha);
    B.a_jobs = up256(B.a_ctr + 256);
    B.a_dj = up256(B.a_jobs + sizeof(RJob) * N);
    B.a_res = up256(B.a_dj + sizeof(DevJob) * N);
    B.a_poff = up256(B.a_res + sizeof(DevResult) * N);
    B.a_sent = up256(B.a_poff + 8ull * N);
    B.a_bases = up256(B.a_sent + 16ull * N);
    B.a_roff = up256(B.a_bases + nbases);
    B.a_graph = up256(B.a_roff + 8 * noffs);
    B.a_end = up256(B.a_graph + gbytes);
    B.b_rows = 0;
    B.b_aux = up256(B.b_rows + sizeof(RowIn) * rows);
    B.b_preds = up256(B.b_aux + sizeof(RowAux) * rows);
    B.b_seq = up256(B.b_preds + sizeof(PredE) * std::max<uint64_t>(preds, 1));
    B.b_meta = up256(B.b_seq + seqb + 16);
    B.b_adapt = up256(B.b_meta + sizeof(RowMeta) * meta);
    B.b_tb = up256(B.b_adapt + sizeof(RowRec) * meta);
    B.b_end = up256(B.b_tb + 4 * tbn);
    return B;
}

// Jobs rendered while the persistent launch still runs: each finished
// job's packed graph and record come back as soon as its workgroup
// publishes it (RPersist::sent), and a host worker rebuilds and renders it.
struct StreamState {
    std::vector<uint8_t> taken;                  // job handed to a worker
    std::vector<std::unique_ptr<Graph>> graphs;  // jobs that leave the device
    std::vector<uint8_t> leaves;
    std::vector<std::string> names;
    std::atomic<uint32_t> pending{0};
    uint64_t end = 0;  // end of the streamed bytes in the pack buffers
    uint32_t *d_sent = nullptr;
    hipStream_t s2 = nullptr;
};

int finish_bucket(RShared &S, RSlot &slot, const std::vector<uint64_t> &jobs, const BucketLayout &BL,
                  const RPools &R, char *d_poff, RShared::Trace &tr, uint64_t replays, StreamState *ss = nullptr);
int stream_poll(RShared &S, RSlot &slot, const std::vector<uint64_t> &jobs, const RPools &R, StreamState &ss);

int run_bucket(RShared &S, RSlot &slot, size_t si, const std::vector<uint64_t> &jobs,
               const PersistCfg *pc = nullptr) {
    RShared::Trace tr{};
    tr.slot = si;
    tr.t[0] = S.ms();
    const Params &p = *S.p;
    const BatchInput &in = *S.in;
    Ctx &ctx = *slot.ctx;
    const uint32_t N = uint32_t(jobs.size());
    const uint32_t alpha = uint32_t(p.alpha());
    RJ_TRY(hipSetDevice(ctx.device));
    hipStream_t s = ctx.stream;
    const BucketLayout BL = layout_bucket(p, in, jobs);
    const std::vector<RCaps> &caps = BL.caps;
    const std::vector<RGraphLayout> &gl = BL.gl;
    const uint32_t maxpad = BL.maxpad, rmax = BL.rmax;
    const uint64_t lmax = BL.lmax;
    const uint64_t a_prm = BL.a_prm, a_mat = BL.a_mat, a_ctr = BL.a_ctr, a_jobs = BL.a_jobs, a_dj = BL.a_dj,
                   a_res = BL.a_res, a_poff = BL.a_poff, a_bases = BL.a_bases, a_roff = BL.a_roff,
                   a_graph = BL.a_graph, a_end = BL.a_end;
    const uint64_t b_rows = BL.b_rows, b_aux = BL.b_aux, b_preds = BL.b_preds, b_seq = BL.b_seq,
                   b_meta = BL.b_meta, b_adapt = BL.b_adapt, b_tb = BL.b_tb, b_end = BL.b_end;
    RJ_TRY(slot.dg.ensure(a_end));
    RJ_TRY(slot.dp.ensure(b_end));
    RJ_TRY(slot.da.ensure_exact(S.arena_bytes));
    RJ_TRY(slot.hin.ensure(a_graph));
    char *dA = static_cast<char *>(slot.dg.p);
    char *dB = static_cast<char *>(slot.dp.p);
    char *h = static_cast<char *>(slot.hin.p);
    memset(h, 0, a_graph);
    const RParams prm = make_rparams(p, reinterpret_cast<const int32_t *>(dA + a_mat));
    memcpy(h + a_prm, &prm, sizeof prm);
    memcpy(h + a_mat, p.matrix.data(), 4ull * alpha * alpha);
    if (pc) {
        const unsigned long long pool0 = pc->slot_bytes * pc->grid;
        memcpy(h + a_ctr + 128, &pool0, sizeof pool0);
    }
    uint64_t bo = 0, oo = 0, go = 0, ro_ = 0, po = 0, so = 0, mo = 0, to = 0;
    for (uint32_t k = 0; k < N; k++) {
        const uint64_t j = jobs[k];
        const uint64_t *ro = in.read_off + in.job_off[j];
        const RCaps &c = caps[k];
        RJob J{};
        J.bases = reinterpret_cast<const uint8_t *>(dA + a_bases + bo);
        J.read_off = reinterpret_cast<const uint64_t *>(dA + a_roff + 8 * oo);
        J.nreads = c.nreads;
        J.first = 0;
        J.nc = c.nc;
        J.ec = c.ec;
        J.maxm = c.maxm;
        bind_graph(J, dA + a_graph + go, gl[k]);
        J.rows_off = ro_;
        J.aux_off = ro_;
        J.preds_off = po;
        J.seq_off = so;
        J.meta_off = mo;
        J.adapt_off = mo;
        J.tb_off = to;
        memcpy(h + a_jobs + sizeof(RJob) * k, &J, sizeof J);
        uint64_t *hoffs = reinterpret_cast<uint64_t *>(h + a_roff) + oo;
        for (uint32_t r = 0; r <= c.nreads; r++) hoffs[r] = ro[r] - ro[0];
        if (c.total) memcpy(h + a_bases + bo, in.bases + ro[0], c.total);
        bo += c.total;
        oo += c.nreads + 1;
        go += gl[k].bytes;
        ro_ += uint64_t(c.nc) + 1;
        po += c.ec;
        so += c.padm;
        mo += c.nc;
        to += c.maxm;
    }
    RJ_TRY(hipMemcpyAsync(dA, h, a_graph, hipMemcpyHostToDevice, s));
    tr.t[1] = S.ms();

    RPools R{};
    R.jobs = reinterpret_cast<RJob *>(dA + a_jobs);
    R.dj = reinterpret_cast<DevJob *>(dA + a_dj);
    R.res = reinterpret_cast<DevResult *>(dA + a_res);
    R.rows = reinterpret_cast<RowIn *>(dB + b_rows);
    R.aux = reinterpret_cast<RowAux *>(dB + b_aux);
    R.preds = reinterpret_cast<PredE *>(dB + b_preds);
    R.seq = reinterpret_cast<uint8_t *>(dB + b_seq);
    R.tb = reinterpret_cast<uint32_t *>(dB + b_tb);
    R.arena_ctr = reinterpret_cast<unsigned long long *>(dA + a_ctr);
    R.done = reinterpret_cast<uint32_t *>(dA + a_ctr + 64);
    R.prm = reinterpret_cast<const RParams *>(dA + a_prm);
    R.njobs = N;
    R.alloc_limit = S.arena_bytes / 4 * 3;
    R.job_cap = R.alloc_limit / 2;
    // The step's LDS: the serial passes' windows and per-row arrays for a
    // graph of about three times the longest template (what the batched
    // workload's graphs reach), at most 64 KB; larger graphs use the jobs'
    // global arrays. Smaller buckets leave LDS to the fills beside them.
    R.lds_bytes = step_lds_bytes(lmax);

    DevPools P{};
    P.jobs = R.dj;
    P.res = R.res;
    P.rows = R.rows;
    P.aux = R.aux;
    P.preds = R.preds;
    P.seq = R.seq;
    P.matrix = reinterpret_cast<const int32_t *>(dA + a_mat);
    P.arena = slot.da.p;
    P.meta = reinterpret_cast<RowMeta *>(dB + b_meta);
    P.adapt = reinterpret_cast<RowRec *>(dB + b_adapt);
    P.tb = R.tb;
    P.pool_next = R.arena_ctr;
    P.pool_end = S.arena_bytes;
    P.chunk = up256(std::max<uint64_t>(262144, 3ull * maxpad * 12));
    P.seq_lds = maxpad;
    P.mat_lds = uint32_t(alpha * alpha * 4);
    const uint32_t wexp = expected_width(p, lmax);
    FillShape shape = ctx.force_shape.cols ? ctx.force_shape : fill_shape_for_width(wexp);
    if (const char *e = getenv("MPOA_FILL_SHAPE")) {
        FillShape f{0, 0};
        if (sscanf(e, "%d,%d", &f.cols, &f.waves) == 2 && fill_shape_ok(f)) shape = f;
    }
    P.ring_bytes = ring_bytes_for_width(wexp, false);
    // A bucket is one length class: the tuning table's entry for its width
    // (minipoa_hip/tune.py) decides its shape and ring when one is named.
    if (!ctx.force_shape.cols && !getenv("MPOA_FILL_SHAPE")) (void)tuned_shape(wexp, shape, P.ring_bytes);
    {
        const uint32_t static_lds = 10240, dyn_other = ((P.seq_lds + 15u) & ~15u) + P.mat_lds;
        const uint32_t room = 65536u - static_lds - dyn_other;
        if (P.ring_bytes > room) P.ring_bytes = room & ~63u;
    }
    P.prio = 0;

    if (pc) {
        // One launch of the persistent kernel runs every job to its end.
        // Its DPs use their workgroup's arena region; the launch-wide chunk
        // pool is closed (pool_end 0), so a DP that outgrows its region
        // returns kArenaOverflow and is re-planned full-width, or leaves the
        // device when that exceeds the region too.
        RPersist Q{};
        Q.next = reinterpret_cast<uint32_t *>(dA + a_ctr + 192);  // zero from the staging copy
        Q.arena_off = 0;
        Q.slot_bytes = pc->slot_bytes;
        Q.prio = getenv("MPOA_PERSIST_PRIO") ? uint32_t(atoi(getenv("MPOA_PERSIST_PRIO"))) & 3u : 3u;
        // The shared pool follows the regions: its counter starts at the
        // pool's first byte (written with the staging copy).
        P.pool_next = reinterpret_cast<unsigned long long *>(dA + a_ctr + 128);
        P.pool_end = pc->slot_bytes * pc->grid + pc->pool_bytes;
        P.ring_bytes = ring_bytes_for_width(wexp, false);
        {
            const uint32_t dyn_other = ((P.seq_lds + 15u) & ~15u) + P.mat_lds;
            const uint32_t room = 65536u - 10240u - dyn_other;
            if (P.ring_bytes > room) P.ring_bytes = room & ~63u;
        }
        // The step gets what four workgroups per CU leave (a gfx950 CU holds
        // 160 KB: 40 KB each, less the kernel's 16 static bytes), or what the
        // fill and the traceback need if that is more. Measured on the
        // 4,096-job workload: 40 KB instead of the ~34.5 KB the other phases
        // need, launch 998 -> 990 ms (larger graphs re-rank and plan in LDS).
        Q.step_lds = std::max(rj_persist_lds(P, 0), 40944u) & ~15u;
        if (const char *e = getenv("MPOA_PERSIST_STEP_LDS")) Q.step_lds = std::max(Q.step_lds, uint32_t(atoi(e)) & ~15u);
        R.lds_bytes = Q.step_lds;
        R.job_cap = pc->slot_bytes + pc->pool_bytes / 4;
        if (!slot.ev[0]) RJ_TRY(hipEventCreate(&slot.ev[0]));
        if (!slot.ev[1]) RJ_TRY(hipEventCreate(&slot.ev[1]));
        // Streaming (MPOA_PERSIST_STREAM=0 turns it off): workgroups pack
        // the jobs they finish into the pack buffer and publish them; this
        // thread copies them back on a second stream while the launch runs
        // and host workers render them, so little is left after it ends.
        const bool stream = [] {
            const char *e = getenv("MPOA_PERSIST_STREAM");
            return !e || atoi(e) != 0;
        }();
        StreamState ss;
        if (stream) {
            // A stream that shares the launch's hardware queue would run the
            // copies only after the launch: another slot's stream has a
            // queue of its own.
            if (!S.copy_stream && !slot.s2) RJ_TRY(hipStreamCreateWithFlags(&slot.s2, hipStreamNonBlocking));
            ss.s2 = S.copy_stream ? S.copy_stream : slot.s2;
            RJ_TRY(slot.hsent.ensure(16ull * N));
            RJ_TRY(slot.hsjobs.ensure(sizeof(RJob) * N));
            ss.taken.assign(N, 0);
            ss.graphs.resize(N);
            ss.leaves.assign(N, 0);
            ss.names.resize(rmax);
            for (uint32_t k = 0; k < rmax; k++) ss.names[k] = "read" + std::to_string(k);
            ss.d_sent = reinterpret_cast<uint32_t *>(dA + BL.a_sent);  // zero from the staging copy
            Q.sbuf = static_cast<uint8_t *>(slot.dk.p);
            Q.scap = std::min<uint64_t>(slot.dk.cap, slot.hpack.cap);
            // Test knob: a smaller streaming share, so that later jobs are
            // packed after the launch instead.
            if (const char *e = getenv("MPOA_PERSIST_SCAP")) Q.scap = std::min<uint64_t>(Q.scap, strtoull(e, nullptr, 10));
            Q.scur = reinterpret_cast<unsigned long long *>(dA + a_ctr + 224);  // zero likewise
            Q.sent = ss.d_sent;
        }
        tr.t[2] = S.ms();
        RJ_TRY(hipEventRecord(slot.ev[0], s));
        // Device-to-host copies issued while a launch holds every workgroup
        // slot start only as it drains (measured here). Streaming leaves
        // eight slots free (measured on the 4,096-job workload: 4 free slots
        // still hold the copies back, 8 let them run beside the launch for
        // +1.2% device time, 16 cost +2.3%); MPOA_PERSIST_FREE overrides.
        uint32_t free_slots = stream ? 8u : 0u;
        if (const char *e = getenv("MPOA_PERSIST_FREE")) free_slots = uint32_t(atoi(e));
        const uint32_t grid = std::min<uint32_t>(N, pc->grid > free_slots ? pc->grid - free_slots : 1u);
        if (launch_rj_persist(R, P, Q, grid, FillShape{4, 4}, s) != 0) {
            set_error("no persistent kernel for the launch shape");
            return MPOA_ERR_DEVICE;
        }
        RJ_TRY(hipGetLastError());
        RJ_TRY(hipEventRecord(slot.ev[1], s));
        if (stream) {
            // Until the launch's end event completes (never on a flag alone).
            int rc = MPOA_OK;
            for (;;) {
                const hipError_t q = hipEventQuery(slot.ev[1]);
                if (q == hipSuccess) break;
                if (q != hipErrorNotReady) {
                    set_error(std::string("HIP error: ") + hipGetErrorString(q) + " waiting for the persistent launch");
                    rc = MPOA_ERR_DEVICE;
                    break;
                }
                std::this_thread::sleep_for(std::chrono::milliseconds(2));
                rc = stream_poll(S, slot, jobs, R, ss);
                if (rc != MPOA_OK) break;
            }
            if (rc != MPOA_OK) {
                (void)hipEventSynchronize(slot.ev[1]);
                while (ss.pending.load() != 0) std::this_thread::sleep_for(std::chrono::microseconds(200));
                return rc;
            }
        }
        // Every streamed render has ended before ss goes away, whatever the
        // outcome.
        const hipError_t se = hipEventSynchronize(slot.ev[1]);
        int frc = MPOA_OK;
        if (se != hipSuccess) {
            set_error(std::string("HIP error: ") + hipGetErrorString(se) + " waiting for the persistent launch");
            frc = MPOA_ERR_DEVICE;
        } else {
            frc = finish_bucket(S, slot, jobs, BL, R, dA + a_poff, tr, 1, stream ? &ss : nullptr);
        }
        while (ss.pending.load() != 0) std::this_thread::sleep_for(std::chrono::microseconds(200));
        return frc;
    }

    // The step, captured once and replayed.
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    RJ_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    launch_rj_reset(R.arena_ctr, s);
    launch_rj_step(R, s);
    const int lf = launch_fill(P, N, false, false, false, shape, s);
    launch_traceback(P, N, false, false, s);
    const hipError_t ce = hipStreamEndCapture(s, &graph);
    if (lf != 0 || ce != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        set_error(lf != 0 ? "no fill kernel for the bucket's launch shape"
                          : std::string("HIP error: ") + hipGetErrorString(ce) + " in step capture");
        return MPOA_ERR_DEVICE;
    }
    hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    if (ie != hipSuccess) {
        (void)hipGraphDestroy(graph);
        set_error(std::string("HIP error: ") + hipGetErrorString(ie) + " instantiating the step graph");
        return MPOA_ERR_DEVICE;
    }
    if (!slot.ev[0]) RJ_TRY(hipEventCreate(&slot.ev[0]));
    if (!slot.ev[1]) RJ_TRY(hipEventCreate(&slot.ev[1]));
    RJ_TRY(slot.hoff.ensure(4096));
    uint32_t *hdone = static_cast<uint32_t *>(slot.hoff.p);
    tr.t[2] = S.ms();
    RJ_TRY(hipEventRecord(slot.ev[0], s));
    // Every read is one replay (the first read needs none of its own, the
    // last replay fuses the last read); retries and deferred DPs add more.
    uint64_t replays = 0;
    const uint64_t bound = 3ull * rmax + 64;
    uint32_t batch = rmax + 1;
    int rc = MPOA_OK;
    for (;;) {
        for (uint32_t r = 0; r < batch; r++) {
            const hipError_t le = hipGraphLaunch(exec, s);
            if (le != hipSuccess) {
                set_error(std::string("HIP error: ") + hipGetErrorString(le) + " launching the step graph");
                rc = MPOA_ERR_DEVICE;
                break;
            }
        }
        if (rc != MPOA_OK) break;
        replays += batch;
        const hipError_t me = hipMemcpyAsync(hdone, R.done, 4, hipMemcpyDeviceToHost, s);
        const hipError_t se = me == hipSuccess ? hipStreamSynchronize(s) : me;
        if (se != hipSuccess) {
            set_error(std::string("HIP error: ") + hipGetErrorString(se) + " in the step loop");
            rc = MPOA_ERR_DEVICE;
            break;
        }
        if (*hdone >= N || replays >= bound) break;
        batch = 4;
    }
    if (rc == MPOA_OK) {
        const hipError_t e1 = hipEventRecord(slot.ev[1], s);
        if (e1 != hipSuccess) rc = MPOA_ERR_DEVICE;
    }
    (void)hipGraphExecDestroy(exec);
    (void)hipGraphDestroy(graph);
    if (rc != MPOA_OK) return rc;
    RJ_TRY(hipEventSynchronize(slot.ev[1]));
    return finish_bucket(S, slot, jobs, BL, R, dA + a_poff, tr, replays);
}

// Rebuilds job k's graph from its packed bytes and renders its outputs (or
// records its failure); a job that leaves the device keeps its graph in
// graphs[k] for the host engine.
void render_job_body(RShared &S, const std::vector<uint64_t> &jobs, uint32_t k, const RJob &J, const char *b,
                     const std::vector<std::string> &names, std::vector<std::unique_ptr<Graph>> &graphs,
                     std::vector<uint8_t> &leaves) {
    const Params &p = *S.p;
    const BatchInput &in = *S.in;
    const uint64_t j = jobs[k];
    const RPackLayout L = rj_pack_layout(J.nn, J.ne, J.nreads, J.npath);
    auto g = std::make_unique<Graph>(p);
    const uint32_t nfused = J.cur;
    g->assign_packed(J.nn, reinterpret_cast<const uint8_t *>(b + L.base),
                     reinterpret_cast<const uint32_t *>(b + L.parent), reinterpret_cast<const uint32_t *>(b + L.rank),
                     reinterpret_cast<const uint32_t *>(b + L.cover),
                     reinterpret_cast<const uint32_t *>(b + L.out_head),
                     reinterpret_cast<const uint32_t *>(b + L.in_head), reinterpret_cast<const uint32_t *>(b + L.edges),
                     reinterpret_cast<const uint32_t *>(b + L.path_off), reinterpret_cast<const uint32_t *>(b + L.paths),
                     nfused, 0);
    if (J.state == kRsDone) {
        if (in.render) {
            std::string out0 = ">consensus sequence\n" + g->consensus_bytes(p) + "\n";
            std::vector<std::string> nm(names.begin(), names.begin() + J.nreads), cm(J.nreads);
            std::string out1;
            std::string e = g->msa(p, nm, cm, out1);
            (*S.out->out0)[j] = std::move(out0);
            (*S.out->out1)[j] = std::move(out1);
            if (!e.empty()) {
                (*S.out->status)[j] = MPOA_ERR_GRAPH;
                (*S.out->err)[j] = e;
            }
        }
        (*S.out->graphs)[j] = std::move(g);
    } else if (J.state == kRsFailed) {
        std::string msg;
        (*S.out->status)[j] = failure_of(J, msg);
        (*S.out->err)[j] = msg;
        (*S.out->graphs)[j] = std::move(g);
    } else {
        leaves[k] = 1;
        graphs[k] = std::move(g);
    }
}

// render_job_body that never throws (it runs on host workers, streamed
// jobs while the launch is still running): a job whose rendering fails ends
// with an error instead.
void render_job(RShared &S, const std::vector<uint64_t> &jobs, uint32_t k, const RJob &J, const char *b,
                const std::vector<std::string> &names, std::vector<std::unique_ptr<Graph>> &graphs,
                std::vector<uint8_t> &leaves) {
    const uint64_t j = jobs[k];
    try {
        render_job_body(S, jobs, k, J, b, names, graphs, leaves);
    } catch (const std::bad_alloc &) {
        leaves[k] = 0;
        (*S.out->status)[j] = MPOA_ERR_ALLOC;
        (*S.out->err)[j] = "host memory allocation failed";
    } catch (const std::exception &e) {
        leaves[k] = 0;
        (*S.out->status)[j] = MPOA_ERR_INVALID_ARGUMENT;
        (*S.out->err)[j] = std::string("unexpected error: ") + e.what();
    }
}

// One look at the published entries: the records and packed bytes of jobs
// not yet taken come back on the slot's second stream (contiguous ranges in
// one copy each), then each goes to a host worker.
int stream_poll(RShared &S, RSlot &slot, const std::vector<uint64_t> &jobs, const RPools &R, StreamState &ss) {
    const uint32_t N = uint32_t(jobs.size());
    hipStream_t s2 = ss.s2;
    uint32_t *hs = static_cast<uint32_t *>(slot.hsent.p);
    RJ_TRY(hipMemcpyAsync(hs, ss.d_sent, 16ull * N, hipMemcpyDeviceToHost, s2));
    RJ_TRY(hipStreamSynchronize(s2));
    struct Item {
        uint32_t k;
        uint64_t off, bytes;
    };
    std::vector<Item> got;
    for (uint32_t k = 0; k < N; k++) {
        // An entry whose flag arrived without its fields (a copy that read
        // the entry between the kernel's two releases) waits for the next
        // look; every packed graph has a nonzero size.
        if (ss.taken[k] || hs[4ull * k] != 1u || hs[4ull * k + 3] == 0u) continue;
        const uint64_t off = uint64_t(hs[4ull * k + 1]) | uint64_t(hs[4ull * k + 2]) << 32;
        got.push_back(Item{k, off, hs[4ull * k + 3]});
    }
    if (got.empty()) return MPOA_OK;
    std::sort(got.begin(), got.end(), [](const Item &x, const Item &y) { return x.off < y.off; });
    RJob *hj = static_cast<RJob *>(slot.hsjobs.p);
    char *hp = static_cast<char *>(slot.hpack.p);
    const char *dk = static_cast<const char *>(slot.dk.p);
    for (size_t a = 0; a < got.size();) {
        size_t b = a + 1;
        uint64_t end = got[a].off + got[a].bytes;
        while (b < got.size() && got[b].off == end) end += got[b++].bytes;
        RJ_TRY(hipMemcpyAsync(hp + got[a].off, dk + got[a].off, end - got[a].off, hipMemcpyDeviceToHost, s2));
        a = b;
    }
    for (const Item &it : got)
        RJ_TRY(hipMemcpyAsync(hj + it.k, R.jobs + it.k, sizeof(RJob), hipMemcpyDeviceToHost, s2));
    RJ_TRY(hipStreamSynchronize(s2));
    for (const Item &it : got) {
        ss.taken[it.k] = 1;
        ss.end = std::max(ss.end, it.off + it.bytes);
        ss.pending.fetch_add(1);
        const RJob J = hj[it.k];
        const char *b = hp + it.off;
        const uint32_t k = it.k;
        S.tasks->push([&S, &jobs, &ss, k, J, b] {
            render_job(S, jobs, k, J, b, ss.names, ss.graphs, ss.leaves);
            ss.pending.fetch_sub(1);
        });
    }
    return MPOA_OK;
}

// After a bucket's device run: final job states, the graphs of jobs not yet
// streamed packed into one copy and rebuilt on the host, outputs of
// finished jobs, evicted jobs queued for the host engine, statistics.
int finish_bucket(RShared &S, RSlot &slot, const std::vector<uint64_t> &jobs, const BucketLayout &BL,
                  const RPools &R, char *d_poff, RShared::Trace &tr, uint64_t replays, StreamState *ss) {
    hipStream_t s = slot.ctx->stream;
    const uint32_t N = uint32_t(jobs.size());
    const uint32_t rmax = BL.rmax;
    const uint64_t lmax = BL.lmax;
    float dev_ms = 0;
    RJ_TRY(hipEventElapsedTime(&dev_ms, slot.ev[0], slot.ev[1]));
    tr.t[3] = S.ms();

    StreamState local;
    if (!ss) {
        ss = &local;
        ss->taken.assign(N, 0);
        ss->graphs.resize(N);
        ss->leaves.assign(N, 0);
        ss->names.resize(rmax);
        for (uint32_t k = 0; k < rmax; k++) ss->names[k] = "read" + std::to_string(k);
    } else {
        // What the workgroups published last.
        const int rc = stream_poll(S, slot, jobs, R, *ss);
        if (rc != MPOA_OK) {
            while (ss->pending.load() != 0) std::this_thread::sleep_for(std::chrono::microseconds(200));
            return rc;
        }
    }
    auto drain = [&] {
        while (ss->pending.load() != 0) std::this_thread::sleep_for(std::chrono::microseconds(200));
    };

    // Final states, then the remaining jobs' graphs packed into one copy,
    // after the streamed bytes (or from the start once their renders ended,
    // if they would not fit).
    RJ_TRY(slot.hjobs.ensure(sizeof(RJob) * N));
    RJob *hj = static_cast<RJob *>(slot.hjobs.p);
    const hipError_t je = hipMemcpyAsync(hj, R.jobs, sizeof(RJob) * N, hipMemcpyDeviceToHost, s);
    const hipError_t js = je == hipSuccess ? hipStreamSynchronize(s) : je;
    if (js != hipSuccess) {
        drain();
        RJ_TRY(js);
    }
    std::vector<uint32_t> rest;
    for (uint32_t k = 0; k < N; k++)
        if (!ss->taken[k]) rest.push_back(k);
    const uint32_t nr = uint32_t(rest.size());
    uint64_t base = (ss->end + 255) / 256 * 256, pbytes = 0;
    for (uint32_t k : rest) pbytes += rj_pack_layout(hj[k].nn, hj[k].ne, hj[k].nreads, hj[k].npath).bytes;
    if (base + pbytes > std::min<uint64_t>(slot.dk.cap, slot.hpack.cap)) {
        drain();
        base = 0;
    }
    hipError_t pe = slot.hoff.ensure(8ull * N + 4096);
    uint64_t *hoff = static_cast<uint64_t *>(slot.hoff.p);
    if (pe == hipSuccess && nr) {
        uint64_t at = base;
        for (uint32_t a = 0; a < nr; a++) {
            hoff[a] = at;
            const uint32_t k = rest[a];
            at += rj_pack_layout(hj[k].nn, hj[k].ne, hj[k].nreads, hj[k].npath).bytes;
        }
        if (base == 0) {
            drain();
            pe = slot.dk.ensure(std::max<uint64_t>(pbytes, 256));
            if (pe == hipSuccess) pe = slot.hpack.ensure(std::max<uint64_t>(pbytes, 256));
        }
        // The selection goes where the published entries were (read above).
        uint32_t *d_sel = reinterpret_cast<uint32_t *>(d_poff + (BL.a_sent - BL.a_poff));
        if (pe == hipSuccess) pe = hipMemcpyAsync(d_poff, hoff, 8ull * nr, hipMemcpyHostToDevice, s);
        if (pe == hipSuccess) pe = hipMemcpyAsync(d_sel, rest.data(), 4ull * nr, hipMemcpyHostToDevice, s);
        if (pe == hipSuccess) {
            launch_rj_pack(R.jobs, d_sel, nr, reinterpret_cast<const uint64_t *>(d_poff), static_cast<uint8_t *>(slot.dk.p), s);
            pe = hipGetLastError();
        }
        if (pe == hipSuccess)
            pe = hipMemcpyAsync(static_cast<char *>(slot.hpack.p) + base, static_cast<char *>(slot.dk.p) + base, pbytes,
                                hipMemcpyDeviceToHost, s);
        if (pe == hipSuccess) pe = hipStreamSynchronize(s);
    }
    if (pe != hipSuccess) {
        drain();
        RJ_TRY(pe);
    }

    tr.t[4] = S.ms();
    // Graphs, outputs and the jobs that continue on the host engine.
    const char *pk = static_cast<const char *>(slot.hpack.p);
    par_for(*S.tasks, nr, [&](size_t a) {
        const uint32_t k = rest[a];
        render_job(S, jobs, k, hj[k], pk + hoff[a], ss->names, ss->graphs, ss->leaves);
    });
    drain();
    tr.t[5] = S.ms();
    tr.njobs = N;
    tr.rmax = rmax;
    tr.lmax = lmax;
    tr.replays = replays;
    tr.dev_ms = dev_ms;
    std::lock_guard<std::mutex> lk(S.mu);
    S.trace.push_back(tr);
    ResidentStats &st = S.st;
    st.buckets += 1;
    st.replays += replays;
    st.device_ms += dev_ms;
    for (uint32_t k = 0; k < N; k++) {
        const RJob &J = hj[k];
        st.cells += J.cells;
        st.pred_cells += J.pred_cells;
        st.arena_bytes += J.arena_bytes;
        st.dp_calls += J.dp_calls;
        st.retries += J.retries;
        st.arena_retries += J.arena_retries;
        st.deferred += J.deferred;
        if (ss->leaves[k]) {
            st.evicted += 1;
            S.evicted.emplace_back(jobs[k], J.cur, std::move(ss->graphs[k]));
        } else {
            st.resident_done += 1;
        }
    }
    return MPOA_OK;
}

}  // namespace

bool resident_eligible(const Params &p, const uint64_t *roff, uint32_t nreads) {
    if (p.seeding || p.progressive || p.scalar_dp || p.alpha() == 0 || p.alpha() * p.alpha() > kMatLdsEntries ||
        nreads == 0)
        return false;
    uint64_t total = 0;
    for (uint32_t r = 0; r < nreads; r++) {
        const uint64_t len = roff[r + 1] - roff[r];
        if (len + 2 > 32000) return false;  // query codes staged in LDS
        total += len;
    }
    return total + nreads + 2 < (uint64_t(1) << 30);
}

int run_resident(std::vector<Ctx *> &slots, TaskQueue &tasks, ResidentCache &cache, const Params &p,
                 const BatchInput &in, BatchOutput &out, EngineStats &est, ResidentStats &rst) {
    const uint64_t N = in.njobs;
    RShared S;
    S.p = &p;
    S.in = &in;
    S.tasks = &tasks;
    S.out = &out;
    S.nslots = uint32_t(slots.size());
    std::vector<uint64_t> host_jobs, res_jobs;
    std::vector<uint64_t> L(N, 0);
    for (uint64_t j = 0; j < N; j++) {
        const uint32_t nr = uint32_t(in.job_off[j + 1] - in.job_off[j]);
        const uint64_t *ro = in.read_off + in.job_off[j];
        L[j] = nr ? ro[1] - ro[0] : 0;
        (resident_eligible(p, ro, nr) ? res_jobs : host_jobs).push_back(j);
    }
    // Length buckets, longest templates first (they bound the makespan).
    std::stable_sort(res_jobs.begin(), res_jobs.end(), [&](uint64_t a, uint64_t b) { return L[a] > L[b]; });
    // About two buckets per slot: fewer, larger buckets leave more of each
    // step's lockstep idle time unfilled at the end; more, smaller ones pay
    // each step's fixed latency (the serial planning passes, three launches)
    // for fewer jobs. Measured on the 4,096-job workload (wall time per
    // run): 7 slots x 256 jobs 1.63 s, 6 x 344 1.62 s, 7 x 512 2.08 s,
    // 7 x 128 2.47 s, 7 x 64 4.52 s.
    uint64_t bsize = (res_jobs.size() + 2 * slots.size() - 1) / std::max<size_t>(1, 2 * slots.size());
    bsize = std::min<uint64_t>(std::max<uint64_t>(bsize, 32), 384);
    if (const char *e = getenv("MPOA_RES_BUCKET")) bsize = std::max<uint64_t>(1, strtoull(e, nullptr, 10));
    // Persistent mode (the default; MPOA_RES_PERSIST=0 selects the captured
    // step loops): every job in one launch of the persistent kernel (one
    // bucket, the first slot), the longest first.
    PersistCfg pcfg;
    const bool persist = [] {
        const char *e = getenv("MPOA_RES_PERSIST");
        return !e || atoi(e) != 0;
    }();
    if (persist) {
        bsize = std::max<uint64_t>(res_jobs.size(), 1);
        if (!getenv("MPOA_PERSIST_ORDER") || atoi(getenv("MPOA_PERSIST_ORDER")) != 0) {
            // Longest expected chain first: the reads' DPs grow with the
            // template and, as the graph takes in reads, with their count.
            std::vector<double> key(N, 0.0);
            for (uint64_t j : res_jobs) {
                const double nr = double(in.job_off[j + 1] - in.job_off[j]);
                key[j] = double(L[j]) * nr * (1.0 + 0.05 * nr);
            }
            std::stable_sort(res_jobs.begin(), res_jobs.end(), [&](uint64_t a, uint64_t b) { return key[a] > key[b]; });
        }
    }
    for (uint64_t b0 = 0; b0 < res_jobs.size(); b0 += bsize)
        S.buckets.emplace_back(res_jobs.begin() + b0, res_jobs.begin() + std::min<uint64_t>(res_jobs.size(), b0 + bsize));
    // Arena per slot: half the device memory over the slots, at most 6 GiB
    // (MPOA_RES_ARENA_MB overrides; the squeeze knob of the tests makes DPs
    // outgrow it and leave the device). Slot buffers stay with the context
    // and are sized for the largest bucket before any bucket starts.
    uint64_t held = 0;
    for (const auto &x : cache.slots) held += x->da.cap;
    const uint64_t free_b = device_free_bytes(slots.empty() ? 0 : slots[0]->device) + held;
    uint64_t arena = std::min<uint64_t>(free_b / 2 / std::max<size_t>(1, slots.size()), uint64_t(6) << 30);
    if (const char *e = getenv("MPOA_RES_ARENA_MB")) arena = strtoull(e, nullptr, 10) << 20;
    if (getenv("MPOA_ARENA_SQUEEZE")) arena = uint64_t(1) << 20;
    S.arena_bytes = std::max<uint64_t>(arena, uint64_t(1) << 20) / 256 * 256;
    if (persist && !res_jobs.empty()) {
        // Four 4x4 workgroups per CU; each region holds the largest DP the
        // longest template is expected to reach (a graph of three times the
        // template, rows of the expected width, 12 bytes per cell, half again
        // as margin), within half the free memory, and a shared pool of an
        // eighth of that takes what larger DPs need beyond their region.
        // MPOA_RES_ARENA_MB sets the region size (the pool stays an eighth).
        int cus = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, slots[0]->device) != hipSuccess ||
            cus <= 0) {
            (void)hipGetLastError();
            cus = 256;
        }
        uint64_t per_cu = 4;
        if (const char *e = getenv("MPOA_PERSIST_WGS")) per_cu = std::max<uint64_t>(1, strtoull(e, nullptr, 10));
        pcfg.grid = uint32_t(std::min<uint64_t>(res_jobs.size(), per_cu * uint64_t(cus)));
        uint64_t lmax = 0;  // the longest template (the queue is ordered by chain length)
        for (uint64_t j : res_jobs) lmax = std::max(lmax, L[j]);
        uint64_t want = up256(3 * (lmax + 2) * uint64_t(expected_width(p, lmax)) * 12 * 3 / 2);
        want = std::max<uint64_t>(want, uint64_t(16) << 20);
        want = std::min<uint64_t>(want, free_b / 2 / pcfg.grid / 9 * 8);
        if (const char *e = getenv("MPOA_RES_ARENA_MB")) want = strtoull(e, nullptr, 10) << 20;
        if (getenv("MPOA_ARENA_SQUEEZE")) want = uint64_t(1) << 20;
        pcfg.slot_bytes = std::max<uint64_t>(want, uint64_t(1) << 20) / 256 * 256;
        pcfg.pool_bytes = pcfg.slot_bytes * pcfg.grid / 8 / 256 * 256;
        if (const char *e = getenv("MPOA_RES_POOL_MB")) pcfg.pool_bytes = strtoull(e, nullptr, 10) << 20;
        S.arena_bytes = pcfg.slot_bytes * pcfg.grid + pcfg.pool_bytes;
    }
    while (cache.slots.size() < slots.size()) cache.slots.push_back(std::make_unique<RSlot>());
    uint64_t need_a = 0, need_b = 0, need_h = 0, need_pack = 0, need_n = 0;
    for (const auto &bk : S.buckets) {
        const BucketLayout BL = layout_bucket(p, in, bk);
        need_a = std::max(need_a, BL.a_end);
        need_b = std::max(need_b, BL.b_end);
        need_h = std::max(need_h, BL.a_graph);
        need_n = std::max<uint64_t>(need_n, bk.size());
        uint64_t pk = 0;
        for (const RCaps &c : BL.caps) pk += rj_pack_layout(c.nc, c.ec, c.nreads, uint32_t(c.total)).bytes;
        need_pack = std::max(need_pack, pk);
    }
    std::vector<RSlot *> rs;
    for (size_t k = 0; k < slots.size(); k++) {
        RSlot &x = *cache.slots[k];
        x.ctx = slots[k];
        if (persist && k > 0) {
            // Only the first slot runs the persistent launch.
            x.da.release();
        } else if (!S.buckets.empty()) {
            hipError_t e = hipSetDevice(x.ctx->device);
            if (e == hipSuccess) e = x.dg.ensure(need_a);
            if (e == hipSuccess) e = x.dp.ensure(need_b);
            if (e == hipSuccess) e = x.da.ensure_exact(S.arena_bytes);
            // An arena the device cannot give (free memory is only an
            // estimate, other allocations may come between): smaller regions
            // (or a smaller bucket arena), down to 16 MiB; DPs that outgrow
            // them leave the device.
            while (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) {
                (void)hipGetLastError();
                if (persist) {
                    if (pcfg.slot_bytes / 2 < (uint64_t(1) << 24)) break;
                    pcfg.slot_bytes = pcfg.slot_bytes / 2 / 256 * 256;
                    pcfg.pool_bytes = pcfg.pool_bytes / 2 / 256 * 256;
                    S.arena_bytes = pcfg.slot_bytes * pcfg.grid + pcfg.pool_bytes;
                } else {
                    if (S.arena_bytes / 2 < (uint64_t(1) << 24)) break;
                    S.arena_bytes = S.arena_bytes / 2 / 256 * 256;
                }
                e = x.da.ensure_exact(S.arena_bytes);
            }
            if (e == hipSuccess) e = x.dk.ensure(std::max<uint64_t>(need_pack, 256));
            if (e == hipSuccess) e = x.hin.ensure(need_h);
            if (e == hipSuccess) e = x.hjobs.ensure(sizeof(RJob) * need_n);
            if (e == hipSuccess) e = x.hpack.ensure(std::max<uint64_t>(need_pack, 256));
            if (e == hipSuccess) e = x.hoff.ensure(8 * need_n + 4096);
            if (e != hipSuccess) {
                (void)hipGetLastError();
                set_error(std::string("HIP error: ") + hipGetErrorString(e) + " sizing the resident buffers");
                return MPOA_ERR_ALLOC;
            }
        }
        rs.push_back(&x);
    }
    if (persist && rs.size() > 1) S.copy_stream = rs[1]->ctx->stream;
    const auto t0 = std::chrono::steady_clock::now();
    S.t0 = t0;
    std::vector<std::thread> th;
    for (size_t si = 0; si < (persist ? std::min<size_t>(1, rs.size()) : rs.size()) && !S.buckets.empty(); si++)
        th.emplace_back([&, si] {
            try {
                for (size_t b; (b = S.next++) < S.buckets.size();) {
                    const int rc = run_bucket(S, *rs[si], si, S.buckets[b], persist ? &pcfg : nullptr);
                    if (rc != MPOA_OK) {
                        std::lock_guard<std::mutex> lk(S.mu);
                        if (S.rc == MPOA_OK) {
                            S.rc = rc;
                            S.err = last_error();
                        }
                        S.next = S.buckets.size();
                        return;
                    }
                }
            } catch (...) {
                std::lock_guard<std::mutex> lk(S.mu);
                if (S.rc == MPOA_OK) {
                    S.rc = MPOA_ERR_ALLOC;
                    S.err = "host allocation failed in the resident batch";
                }
                S.next = S.buckets.size();
            }
        });
    for (auto &t : th) t.join();
    if (const char *tp = getenv("MPOA_RES_TRACE")) {
        if (FILE *f = fopen(tp, "w")) {
            fprintf(f, "slot,jobs,rmax,lmax,replays,staged,captured,looped,packed,rendered,start,dev_ms\n");
            for (const auto &x : S.trace)
                fprintf(f, "%zu,%u,%u,%llu,%llu,%.2f,%.2f,%.2f,%.2f,%.2f,%.2f,%.2f\n", x.slot, x.njobs, x.rmax,
                        (unsigned long long)x.lmax, (unsigned long long)x.replays, x.t[1], x.t[2], x.t[3], x.t[4],
                        x.t[5], x.t[0], x.dev_ms);
            fclose(f);
        }
    }
    if (S.rc != MPOA_OK) {
        set_error(S.err);
        return S.rc;
    }
    rst = S.st;
    rst.resident_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    // Everything else continues on the host-driven engine: jobs the device
    // path does not take from the start, and evicted jobs from the read
    // where they left (extend_graph semantics: read k keeps sequence id k).
    if (host_jobs.empty() && S.evicted.empty()) {
        est.run.cells = rst.cells;
        est.run.pred_cells = rst.pred_cells;
        est.run.arena_bytes = rst.arena_bytes;
        est.run.dp_calls = rst.dp_calls;
        est.run.retries = rst.retries;
        est.run.arena_retries = rst.arena_retries;
        est.run.fill_launches = rst.replays;
        est.run.device_ms = rst.device_ms;
        return MPOA_OK;
    }
    EngineOpts o;
    o.render = in.render;
    o.admit_bytes = ~0ull;
    std::vector<std::pair<uint64_t, uint64_t>> ids;
    int rc;
    {
        Engine E(slots, tasks, p, o);
        for (uint64_t j : host_jobs) {
            JobIn ji;
            ji.bases = in.bases;
            ji.read_off = in.read_off + in.job_off[j];
            ji.nreads = uint32_t(in.job_off[j + 1] - in.job_off[j]);
            ji.guide = p.progressive;
            ids.emplace_back(j, E.submit(std::move(ji)));
        }
        for (auto &ev : S.evicted) {
            const uint64_t j = std::get<0>(ev);
            const uint32_t cur = std::get<1>(ev);
            JobIn ji;
            ji.bases = in.bases;
            ji.read_off = in.read_off + in.job_off[j] + cur;
            ji.nreads = uint32_t(in.job_off[j + 1] - in.job_off[j]) - cur;
            ji.first = cur;
            ji.graph = std::move(std::get<2>(ev));
            ids.emplace_back(j, E.submit(std::move(ji)));
        }
        rc = E.wait_all();
        if (rc == MPOA_OK) {
            for (auto &pr : ids) {
                JobOut &jo = E.out(pr.second);
                const uint64_t j = pr.first;
                (*out.status)[j] = jo.status;
                (*out.err)[j] = std::move(jo.err);
                (*out.out0)[j] = std::move(jo.out0);
                (*out.out1)[j] = std::move(jo.out1);
                (*out.graphs)[j] = std::move(jo.graph);
            }
        }
        est = E.stats();
    }
    est.run.cells += rst.cells;
    est.run.pred_cells += rst.pred_cells;
    est.run.arena_bytes += rst.arena_bytes;
    est.run.dp_calls += rst.dp_calls;
    est.run.retries += rst.retries;
    est.run.arena_retries += rst.arena_retries;
    est.run.fill_launches += rst.replays;
    rst.host_jobs = host_jobs.size();
    return rc;
}

}  // namespace mpoa
