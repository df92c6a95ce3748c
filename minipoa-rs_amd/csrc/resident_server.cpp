// Serving on device-resident graphs. Jobs arrive at any time; each is
// admitted into a slot of a step table of its size class when the table's
// device heap has room for its region (graph, reads, DP descriptors), and
// its graph and descriptors then stay in that region until it finishes. The
// table's captured step graph (resident.hpp: fusion, re-rank and planning
// on the device, then fill and traceback) advances every occupied slot by
// one read per replay. Between replays the table's thread reads back the
// slots' states, collects finished jobs (one packed download; the host
// rebuilds and renders them) and admits waiting jobs into the freed slots
// and heap, so a request that arrives while others run joins the next
// replay. Jobs outside the resident path, and jobs that leave it, run on the
// host-driven engine (src/lib.rs:184-226 semantics either way).
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <unordered_map>

#include "resident_host.hpp"
#include "runtime.hpp"

namespace mpoa {

namespace {

#define RS_TRY(expr)                                                                        \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess) {                                                             \
            set_error(std::string("HIP error: ") + hipGetErrorString(e_) + " at " #expr);   \
            return MPOA_ERR_DEVICE;                                                         \
        }                                                                                   \
    } while (0)

// First-fit ranges of a device heap (host bookkeeping), with coalescing.
class RangeAlloc {
public:
    void reset(uint64_t size) {
        free_.clear();
        used_.clear();
        if (size) free_[0] = size;
    }
    bool take(uint64_t bytes, uint64_t align, uint64_t &off) {
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            const uint64_t b = it->first, e = b + it->second;
            const uint64_t a = (b + align - 1) / align * align;
            if (a + bytes > e) continue;
            free_.erase(it);
            if (a > b) free_[b] = a - b;
            if (a + bytes < e) free_[a + bytes] = e - (a + bytes);
            used_[a] = bytes;
            off = a;
            return true;
        }
        return false;
    }
    void give(uint64_t off) {
        auto u = used_.find(off);
        if (u == used_.end()) return;
        uint64_t b = off, len = u->second;
        used_.erase(u);
        auto nx = free_.lower_bound(b);
        if (nx != free_.end() && nx->first == b + len) {
            len += nx->second;
            nx = free_.erase(nx);
        }
        if (nx != free_.begin()) {
            auto pv = std::prev(nx);
            if (pv->first + pv->second == b) {
                pv->second += len;
                return;
            }
        }
        free_[b] = len;
    }

private:
    std::map<uint64_t, uint64_t> free_, used_;
};

// Offsets of one job's parts inside its heap region. The DP descriptor
// arrays sit at multiples of their element sizes from the heap's base (the
// DevJob indices count elements from there); RowRec (20 bytes) is placed
// once the region's base is known.
struct RRegion {
    RCaps caps;
    RGraphLayout gl;
    uint64_t graph, bases, roff, rows, aux, preds, seq, meta, adapt, tb, bytes;
};

RRegion region_of(const RCaps &c, uint32_t alpha) {
    RRegion R;
    R.caps = c;
    R.gl = graph_layout(c, alpha);
    uint64_t at = 0;
    auto take = [&](uint64_t bytes) {
        const uint64_t o = at;
        at = (at + bytes + 63) / 64 * 64;
        return o;
    };
    R.graph = take(R.gl.bytes);
    R.bases = take(std::max<uint64_t>(c.total, 1));
    R.roff = take(8ull * (c.nreads + 1));
    R.rows = take(sizeof(RowIn) * (uint64_t(c.nc) + 1));
    R.aux = take(sizeof(RowAux) * (uint64_t(c.nc) + 1));
    R.preds = take(sizeof(PredE) * std::max<uint64_t>(c.ec, 1));
    R.seq = take(c.padm + 16);
    R.meta = take(sizeof(RowMeta) * c.nc);
    R.adapt = take(sizeof(RowRec) * c.nc + 64);
    R.tb = take(4ull * c.maxm);
    R.bytes = at;
    return R;
}

// Size classes: tables per class, slots per table and the longest read a
// class takes. Jobs of one table share each step's duration (its slowest
// DP), so classes keep short jobs from waiting on long jobs' DPs.
struct RClass {
    uint32_t tables, K, maxlen;
    double heap_share, arena_share;  // of free device memory, over the class's tables
};
const RClass kClasses[] = {
    {2, 64, 2000, 0.04, 0.02},
    {2, 32, 8000, 0.06, 0.04},
    {2, 16, 31998, 0.10, 0.06},
};
constexpr int kNClasses = 3;

}  // namespace

struct ResidentServer::Impl {
    struct Job {
        JobIn in;
        JobOut out;
        int cls = -1;
        int64_t engine_id = -1;  // run (from the start or after leaving the device) on the host engine
        bool engine_copied = false;
        RRegion reg;
        std::vector<std::pair<uint64_t, std::vector<uint32_t>>> prior;  // paths of an extended graph
    };
    struct Table {
        int device = 0;
        hipStream_t stream = nullptr;
        int cls = 0;
        uint32_t K = 0, maxlen = 0;
        DevBuf blk, heap, arena, pack;
        PinBuf hstage, hjobs, hpack, hsel;
        RangeAlloc alloc;
        uint64_t heap_bytes = 0, arena_bytes = 0;
        uint64_t b_mat = 0, b_ctr = 0, b_jobs = 0, b_poff = 0, b_sel = 0;
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        std::vector<int64_t> slot_job;      // ticket per slot, -1 = free
        std::vector<Job *> slot_ptr;        // its job
        std::vector<uint64_t> slot_region;  // heap offset per slot
        uint32_t active = 0;
        RPools R{};
        std::thread th;
        uint64_t replays = 0;
        ~Table() {
            if (exec) (void)hipGraphExecDestroy(exec);
            if (graph) (void)hipGraphDestroy(graph);
            if (stream) (void)hipStreamDestroy(stream);
            blk.release();
            heap.release();
            arena.release();
            pack.release();
            hstage.release();
            hjobs.release();
            hpack.release();
            hsel.release();
        }
    };

    TaskQueue &tasks;
    Params p;
    EngineOpts o;
    std::chrono::steady_clock::time_point t0;
    std::unique_ptr<Engine> host;
    std::vector<std::unique_ptr<Table>> tables;
    std::mutex mu;
    std::condition_variable cv;
    std::deque<std::unique_ptr<Job>> jobs;
    std::vector<std::deque<uint64_t>> queues;  // waiting tickets per class
    uint64_t pending = 0;                      // resident jobs not yet finished or handed to the host
    bool stop = false;
    int rc = 0;
    std::string err;
    ResidentStats st;

    Impl(std::vector<Ctx *> &slots, TaskQueue &tq, const Params &pp, const EngineOpts &oo)
        : tasks(tq), p(pp), o(oo), t0(std::chrono::steady_clock::now()) {}

    double now_ms() const {
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }

    void fail_all(int code, const std::string &e) {
        std::lock_guard<std::mutex> lk(mu);
        if (rc == 0) {
            rc = code;
            err = e;
        }
        cv.notify_all();
    }

    int build_table(Table &T, uint32_t alpha);
    int class_of(const JobIn &in) const;
    int admit(Table &T, uint32_t k, uint64_t t, Job &S, uint64_t region);
    int collect(Table &T, const std::vector<uint32_t> &fin);
    void table_loop(Table &T);
};

int ResidentServer::Impl::build_table(Table &T, uint32_t alpha) {
    RS_TRY(hipSetDevice(T.device));
    RS_TRY(hipStreamCreateWithFlags(&T.stream, hipStreamNonBlocking));
    const uint32_t K = T.K;
    // [params][matrix][counters][RJob K][DevJob K][DevResult K][pack offsets K][pack selection K]
    T.b_mat = up256(sizeof(RParams));
    T.b_ctr = up256(T.b_mat + 4ull * alpha * alpha);
    T.b_jobs = up256(T.b_ctr + 256);
    const uint64_t b_dj = up256(T.b_jobs + sizeof(RJob) * K);
    const uint64_t b_res = up256(b_dj + sizeof(DevJob) * K);
    T.b_poff = up256(b_res + sizeof(DevResult) * K);
    T.b_sel = up256(T.b_poff + 8ull * K);
    const uint64_t b_end = up256(T.b_sel + 4ull * K);
    RS_TRY(T.blk.ensure(b_end));
    RS_TRY(T.heap.ensure_exact(T.heap_bytes));
    RS_TRY(T.arena.ensure_exact(T.arena_bytes));
    RS_TRY(T.hstage.ensure(b_end));
    RS_TRY(T.hjobs.ensure(sizeof(RJob) * K + 4096));
    RS_TRY(T.hsel.ensure(12ull * K + 4096));
    T.alloc.reset(T.heap_bytes);
    T.slot_job.assign(K, -1);
    T.slot_ptr.assign(K, nullptr);
    T.slot_region.assign(K, 0);
    char *d = static_cast<char *>(T.blk.p);
    char *h = static_cast<char *>(T.hstage.p);
    memset(h, 0, b_end);
    const RParams prm = make_rparams(p, reinterpret_cast<const int32_t *>(d + T.b_mat));
    memcpy(h, &prm, sizeof prm);
    memcpy(h + T.b_mat, p.matrix.data(), 4ull * alpha * alpha);
    for (uint32_t k = 0; k < K; k++) {
        RJob J{};
        J.state = kRsDone;  // a vacant slot: every step skips it
        memcpy(h + T.b_jobs + sizeof(RJob) * k, &J, sizeof J);
    }
    RS_TRY(hipMemcpyAsync(d, h, b_end, hipMemcpyHostToDevice, T.stream));
    RS_TRY(hipStreamSynchronize(T.stream));
    char *heap = static_cast<char *>(T.heap.p);
    RPools &R = T.R;
    R.jobs = reinterpret_cast<RJob *>(d + T.b_jobs);
    R.dj = reinterpret_cast<DevJob *>(d + b_dj);
    R.res = reinterpret_cast<DevResult *>(d + b_res);
    R.rows = reinterpret_cast<RowIn *>(heap);
    R.aux = reinterpret_cast<RowAux *>(heap);
    R.preds = reinterpret_cast<PredE *>(heap);
    R.seq = reinterpret_cast<uint8_t *>(heap);
    R.tb = reinterpret_cast<uint32_t *>(heap);
    R.arena_ctr = reinterpret_cast<unsigned long long *>(d + T.b_ctr);
    R.done = reinterpret_cast<uint32_t *>(d + T.b_ctr + 64);
    R.prm = reinterpret_cast<const RParams *>(d);
    R.njobs = K;
    R.alloc_limit = T.arena_bytes / 4 * 3;
    R.job_cap = R.alloc_limit / 2;
    R.lds_bytes = step_lds_bytes(T.maxlen);
    const uint32_t maxpad = ((T.maxlen + 2) / 8 + 1) * 8;
    DevPools P{};
    P.jobs = R.dj;
    P.res = R.res;
    P.rows = R.rows;
    P.aux = R.aux;
    P.preds = R.preds;
    P.seq = R.seq;
    P.matrix = reinterpret_cast<const int32_t *>(d + T.b_mat);
    P.arena = T.arena.p;
    P.meta = reinterpret_cast<RowMeta *>(heap);
    P.adapt = reinterpret_cast<RowRec *>(heap);
    P.tb = R.tb;
    P.pool_next = R.arena_ctr;
    P.pool_end = T.arena_bytes;
    P.chunk = up256(std::max<uint64_t>(262144, 3ull * maxpad * 12));
    P.seq_lds = maxpad;
    P.mat_lds = alpha * alpha * 4;
    uint32_t wexp = maxpad;
    if (p.f != 0) {
        const uint64_t m = T.maxlen + 2, band = p.b + m / p.f;
        wexp = uint32_t(std::min<uint64_t>(maxpad, 2 * band + m / 4 + 16));
    }
    FillShape shape = fill_shape_for_width(wexp);
    P.ring_bytes = ring_bytes_for_width(wexp, false);
    if (const char *e = getenv("MPOA_FILL_SHAPE")) {
        FillShape f{0, 0};
        if (sscanf(e, "%d,%d", &f.cols, &f.waves) == 2 && fill_shape_ok(f)) shape = f;
    } else {
        (void)tuned_shape(wexp, shape, P.ring_bytes);
    }
    {
        const uint32_t static_lds = 10240, dyn_other = ((P.seq_lds + 15u) & ~15u) + P.mat_lds;
        const uint32_t room = 65536u - static_lds - dyn_other;
        if (P.ring_bytes > room) P.ring_bytes = room & ~63u;
    }
    RS_TRY(hipStreamBeginCapture(T.stream, hipStreamCaptureModeThreadLocal));
    launch_rj_reset(R.arena_ctr, T.stream);
    launch_rj_step(R, T.stream);
    const int lf = launch_fill(P, K, false, false, false, shape, T.stream);
    launch_traceback(P, K, false, false, T.stream);
    const hipError_t ce = hipStreamEndCapture(T.stream, &T.graph);
    if (lf != 0 || ce != hipSuccess) {
        set_error(lf != 0 ? "no fill kernel for a step table's launch shape"
                          : std::string("HIP error: ") + hipGetErrorString(ce) + " capturing a step table");
        return MPOA_ERR_DEVICE;
    }
    RS_TRY(hipGraphInstantiate(&T.exec, T.graph, nullptr, nullptr, 0));
    return MPOA_OK;
}

int ResidentServer::Impl::class_of(const JobIn &in) const {
    if (in.guide || p.seeding) return -1;
    if (!resident_eligible(p, in.read_off, in.nreads)) return -1;
    uint64_t maxlen = 0;
    for (uint32_t r = 0; r < in.nreads; r++) maxlen = std::max<uint64_t>(maxlen, in.read_off[r + 1] - in.read_off[r]);
    if (in.graph) {
        // The device rebuilds `aligned` from parents when it hands a graph
        // back: take only graphs whose columns follow that rule (every graph
        // built by add_path or loaded from GFA).
        const Graph &g = *in.graph;
        if (g.size() + 2 >= (uint64_t(1) << 30) || g.order().size() != g.size()) return -1;
        for (size_t id = 0; id < g.size(); id++) {
            const Graph::Node &nd = g.node(id);
            if (nd.parent >= g.size() || (nd.parent != id && g.node(nd.parent).parent != nd.parent)) return -1;
            for (size_t a = 0; a < g.alpha(); a++) {
                const int64_t v = g.aligned(id, a);
                if (v < 0) continue;
                const bool own = size_t(v) == id && a == nd.base;
                const bool member = size_t(v) < g.size() && g.node(size_t(v)).parent == id &&
                                    size_t(v) != id && g.node(size_t(v)).base == a;
                if (!own && !member) return -1;
            }
        }
    }
    for (int c = 0; c < kNClasses; c++)
        if (maxlen <= kClasses[c].maxlen) return c;
    return -1;
}

// Stages ticket t into slot k of T: its reads, an extended graph's arrays
// (one edge record per edge, linked in each list's insertion order) and its
// state. The region at `region` of T's heap is already reserved.
int ResidentServer::Impl::admit(Table &T, uint32_t k, uint64_t t, Job &S, uint64_t region) {
    const JobIn &in = S.in;
    const RRegion &RR = S.reg;
    const RCaps &c = RR.caps;
    char *heap = static_cast<char *>(T.heap.p);
    const uint64_t base = region;
    RJob J{};
    J.bases = reinterpret_cast<const uint8_t *>(heap + base + RR.bases);
    J.read_off = reinterpret_cast<const uint64_t *>(heap + base + RR.roff);
    J.nreads = c.nreads;
    J.first = in.first;
    J.nc = c.nc;
    J.ec = c.ec;
    J.maxm = c.maxm;
    bind_graph(J, heap + base + RR.graph, RR.gl);
    J.rows_off = (base + RR.rows) / sizeof(RowIn);
    J.aux_off = (base + RR.aux) / sizeof(RowAux);
    J.preds_off = (base + RR.preds) / sizeof(PredE);
    J.seq_off = base + RR.seq;
    J.meta_off = (base + RR.meta) / sizeof(RowMeta);
    J.adapt_off = (base + RR.adapt + sizeof(RowRec) - 1) / sizeof(RowRec);
    J.tb_off = (base + RR.tb) / 4;
    const uint64_t gimg = in.graph ? up16(RR.gl.bytes) : 0;
    const uint64_t rb = up16(std::max<uint64_t>(c.total, 1));
    const uint64_t img = gimg + rb + 8ull * (c.nreads + 1);
    RS_TRY(T.hstage.ensure(std::max<uint64_t>(img, 256)));
    char *h = static_cast<char *>(T.hstage.p);
    const uint64_t b0 = c.nreads ? in.read_off[0] : 0;
    if (c.total) memcpy(h + gimg, in.bases + b0, c.total);
    uint64_t *ro = reinterpret_cast<uint64_t *>(h + gimg + rb);
    for (uint32_t r = 0; r <= c.nreads; r++) ro[r] = in.read_off[r] - b0;
    if (in.graph) {
        const Graph &g = *in.graph;
        RJob Hj{};
        bind_graph(Hj, h, RR.gl);
        const uint32_t nn = uint32_t(g.size()), alpha = uint32_t(g.alpha());
        std::unordered_map<uint64_t, uint32_t> rec;
        uint32_t ne = 0;
        for (uint32_t id = 0; id < nn; id++) {
            const Graph::Node &nd = g.node(id);
            Hj.base[id] = uint8_t(nd.base);
            Hj.parent[id] = nd.parent;
            Hj.rank[id] = g.rank_of(id);
            Hj.cover[id] = uint32_t(nd.coverage);
            Hj.mark[id] = 0;
            for (uint32_t a = 0; a < alpha; a++) Hj.aligned[uint64_t(id) * alpha + a] = int32_t(g.aligned(id, a));
            Hj.out_head[id] = Hj.out_tail[id] = Hj.in_head[id] = Hj.in_tail[id] = kNil;
            Hj.in_cnt[id] = nd.in.size();
            Hj.order[id] = g.order()[id];
        }
        for (uint32_t id = 0; id < nn; id++)
            for (const Edge &ed : g.node(id).out) {
                const uint32_t e = ne++;
                Hj.edges[e] = REdge{id, ed.node, ed.w, kNil, kNil};
                if (Hj.out_tail[id] == kNil) Hj.out_head[id] = e;
                else Hj.edges[Hj.out_tail[id]].next_out = e;
                Hj.out_tail[id] = e;
                rec[uint64_t(id) << 32 | ed.node] = e;
            }
        for (uint32_t id = 0; id < nn; id++)
            for (const Edge &ed : g.node(id).in) {
                auto it = rec.find(uint64_t(ed.node) << 32 | id);
                if (it == rec.end()) {
                    set_error("an extended graph's in- and out-lists disagree");
                    return MPOA_ERR_GRAPH;
                }
                const uint32_t e = it->second;
                if (Hj.in_tail[id] == kNil) Hj.in_head[id] = e;
                else Hj.edges[Hj.in_tail[id]].next_in = e;
                Hj.in_tail[id] = e;
            }
        Hj.path_off[0] = 0;
        J.nn = nn;
        J.ne = ne;
        S.prior = g.paths();
        RS_TRY(hipMemcpyAsync(heap + base + RR.graph, h, RR.gl.bytes, hipMemcpyHostToDevice, T.stream));
    }
    RS_TRY(hipMemcpyAsync(heap + base + RR.bases, h + gimg, std::max<uint64_t>(c.total, 1), hipMemcpyHostToDevice,
                          T.stream));
    RS_TRY(hipMemcpyAsync(heap + base + RR.roff, ro, 8ull * (c.nreads + 1), hipMemcpyHostToDevice, T.stream));
    RJob *hj = static_cast<RJob *>(T.hjobs.p);
    hj[k] = J;
    RS_TRY(hipMemcpyAsync(static_cast<char *>(T.blk.p) + T.b_jobs + sizeof(RJob) * k, &hj[k], sizeof(RJob),
                          hipMemcpyHostToDevice, T.stream));
    // The staging buffers are reused by the next admission.
    RS_TRY(hipStreamSynchronize(T.stream));
    T.slot_job[k] = int64_t(t);
    T.slot_ptr[k] = &S;
    T.slot_region[k] = region;
    T.active++;
    return MPOA_OK;
}

// Finished (or departing) slots fin: one packed download, then per job the
// host graph, its outputs, or its continuation on the host engine.
int ResidentServer::Impl::collect(Table &T, const std::vector<uint32_t> &fin) {
    const uint32_t n = uint32_t(fin.size());
    const RJob *hj = static_cast<const RJob *>(T.hjobs.p);
    uint64_t *hoff = static_cast<uint64_t *>(T.hsel.p);
    uint32_t *hsel = reinterpret_cast<uint32_t *>(hoff + T.K);
    uint64_t bytes = 0;
    for (uint32_t q = 0; q < n; q++) {
        const RJob &J = hj[fin[q]];
        hoff[q] = bytes;
        hsel[q] = fin[q];
        bytes += rj_pack_layout(J.nn, J.ne, J.nreads, J.npath).bytes;
    }
    RS_TRY(T.pack.ensure(std::max<uint64_t>(bytes, 256)));
    RS_TRY(T.hpack.ensure(std::max<uint64_t>(bytes, 256)));
    char *d = static_cast<char *>(T.blk.p);
    RS_TRY(hipMemcpyAsync(d + T.b_poff, hoff, 8ull * n, hipMemcpyHostToDevice, T.stream));
    RS_TRY(hipMemcpyAsync(d + T.b_sel, hsel, 4ull * n, hipMemcpyHostToDevice, T.stream));
    launch_rj_pack(T.R.jobs, reinterpret_cast<const uint32_t *>(d + T.b_sel), n,
                   reinterpret_cast<const uint64_t *>(d + T.b_poff), static_cast<uint8_t *>(T.pack.p), T.stream);
    RS_TRY(hipGetLastError());
    RS_TRY(hipMemcpyAsync(T.hpack.p, T.pack.p, bytes, hipMemcpyDeviceToHost, T.stream));
    RS_TRY(hipStreamSynchronize(T.stream));
    const char *pk = static_cast<const char *>(T.hpack.p);
    for (uint32_t q = 0; q < n; q++) {
        const uint32_t k = fin[q];
        const RJob &J = hj[k];
        Job &S = *T.slot_ptr[k];
        const RPackLayout L = rj_pack_layout(J.nn, J.ne, J.nreads, J.npath);
        const char *b = pk + hoff[q];
        auto g = std::make_unique<Graph>(p);
        g->assign_packed(J.nn, reinterpret_cast<const uint8_t *>(b + L.base),
                         reinterpret_cast<const uint32_t *>(b + L.parent), reinterpret_cast<const uint32_t *>(b + L.rank),
                         reinterpret_cast<const uint32_t *>(b + L.cover),
                         reinterpret_cast<const uint32_t *>(b + L.out_head),
                         reinterpret_cast<const uint32_t *>(b + L.in_head),
                         reinterpret_cast<const uint32_t *>(b + L.edges),
                         reinterpret_cast<const uint32_t *>(b + L.path_off),
                         reinterpret_cast<const uint32_t *>(b + L.paths), J.cur, J.first,
                         S.prior.empty() ? nullptr : &S.prior);
        T.alloc.give(T.slot_region[k]);
        T.slot_job[k] = -1;
        T.slot_ptr[k] = nullptr;
        T.active--;
        {
            std::lock_guard<std::mutex> lk(mu);
            st.cells += J.cells;
            st.pred_cells += J.pred_cells;
            st.arena_bytes += J.arena_bytes;
            st.dp_calls += J.dp_calls;
            st.retries += J.retries;
            st.arena_retries += J.arena_retries;
            st.deferred += J.deferred;
            if (J.state == kRsEvicted || J.state == kRsRunning) st.evicted += 1;
            else st.resident_done += 1;
        }
        if (J.state == kRsEvicted || J.state == kRsRunning) {
            // Continue on the host engine from read `cur` (extend_graph:
            // sequence ids stay first + k).
            JobIn ji;
            ji.nreads = J.nreads - J.cur;
            ji.first = J.first + J.cur;
            ji.own_off.assign(S.in.read_off + J.cur, S.in.read_off + J.nreads + 1);
            ji.own_bases.assign(S.in.bases, S.in.bases + (J.nreads ? S.in.read_off[J.nreads] : 0));
            ji.graph = std::move(g);
            const uint64_t id = host->submit(std::move(ji));
            std::lock_guard<std::mutex> lk(mu);
            S.engine_id = int64_t(id);
            S.out.cells = J.cells;
            pending--;
            cv.notify_all();
            continue;
        }
        S.out.cells = J.cells;
        const double t_done = now_ms();
        std::string msg;
        const int fst = J.state == kRsFailed ? failure_of(J, msg) : MPOA_OK;
        Job *sp = &S;
        Graph *gp = g.release();
        tasks.push([this, sp, gp, fst, msg, t_done] {
            std::unique_ptr<Graph> gg(gp);
            JobOut &out = sp->out;
            if (fst != MPOA_OK) {
                out.status = fst;
                out.err = msg;
            } else if (o.render) {
                out.out0 = ">consensus sequence\n" + gg->consensus_bytes(p) + "\n";
                const uint32_t ns = out.nseq;
                std::vector<std::string> names(ns), comments(ns);
                for (uint32_t q = 0; q < ns; q++) names[q] = "read" + std::to_string(q);
                std::string e = gg->msa(p, names, comments, out.out1);
                if (!e.empty()) {
                    out.status = MPOA_ERR_GRAPH;
                    out.err = e;
                }
            }
            out.graph = std::move(gg);
            std::lock_guard<std::mutex> lk(mu);
            out.t_done = std::max(t_done, now_ms());
            out.done = true;
            pending--;
            cv.notify_all();
        });
    }
    return MPOA_OK;
}

void ResidentServer::Impl::table_loop(Table &T) {
    (void)hipSetDevice(T.device);
    const uint32_t K = T.K;
    RJob *hj = static_cast<RJob *>(T.hjobs.p);
    auto fail = [&](int code) { fail_all(code, last_error()); };
    for (;;) {
        // Admission: the class's waiting jobs in arrival order, while a slot
        // is free and the heap has room (first come, first served).
        for (;;) {
            uint64_t t = 0;
            uint64_t region = 0;
            uint32_t slot = K;
            Job *sj = nullptr;
            {
                std::unique_lock<std::mutex> lk(mu);
                if (stop || rc) return;
                auto &q = queues[size_t(T.cls)];
                if (q.empty() || T.active == K) {
                    if (T.active > 0) break;
                    cv.wait(lk, [&] { return stop || rc || !queues[size_t(T.cls)].empty(); });
                    if (stop || rc) return;
                    continue;
                }
                t = q.front();
                Job &S = *jobs[t];
                if (!T.alloc.take(S.reg.bytes, 256, region)) {
                    if (T.active > 0) break;
                    // Alone, and still too large for this table's heap: the host engine.
                    q.pop_front();
                    lk.unlock();
                    JobIn ji = std::move(S.in);
                    const uint64_t id = host->submit(std::move(ji));
                    lk.lock();
                    S.engine_id = int64_t(id);
                    pending--;
                    st.host_jobs += 1;
                    cv.notify_all();
                    continue;
                }
                q.pop_front();
                sj = &S;
                S.out.t_admit = now_ms();
                for (uint32_t k = 0; k < K; k++)
                    if (T.slot_job[k] < 0) {
                        slot = k;
                        break;
                    }
            }
            const int arc = admit(T, slot, t, *sj, region);
            if (arc != MPOA_OK) {
                fail(arc);
                return;
            }
        }
        // One step of every occupied slot, then the slots' states.
        hipError_t e = hipGraphLaunch(T.exec, T.stream);
        if (e == hipSuccess) e = hipMemcpyAsync(hj, T.R.jobs, sizeof(RJob) * K, hipMemcpyDeviceToHost, T.stream);
        if (e == hipSuccess) e = hipStreamSynchronize(T.stream);
        if (e != hipSuccess) {
            set_error(std::string("HIP error: ") + hipGetErrorString(e) + " in a step table");
            fail(MPOA_ERR_DEVICE);
            return;
        }
        T.replays++;
        std::vector<uint32_t> fin;
        for (uint32_t k = 0; k < K; k++)
            if (T.slot_job[k] >= 0 && hj[k].state != kRsRunning) fin.push_back(k);
        if (!fin.empty()) {
            const int crc = collect(T, fin);
            if (crc != MPOA_OK) {
                fail(crc);
                return;
            }
        }
    }
}

ResidentServer::ResidentServer(std::vector<Ctx *> slots, TaskQueue &tasks, const Params &p, const EngineOpts &o)
    : d_(std::make_unique<Impl>(slots, tasks, p, o)) {
    Impl &I = *d_;
    I.host = std::make_unique<Engine>(slots, tasks, p, o);
    I.queues.resize(kNClasses);
    const int device = slots.empty() ? 0 : slots[0]->device;
    const uint64_t free_b = device_free_bytes(device);
    const uint32_t alpha = uint32_t(p.alpha());
    int rc = MPOA_OK;
    for (int c = 0; c < kNClasses && rc == MPOA_OK; c++)
        for (uint32_t k = 0; k < kClasses[c].tables && rc == MPOA_OK; k++) {
            auto T = std::make_unique<Impl::Table>();
            T->device = device;
            T->cls = c;
            T->K = kClasses[c].K;
            T->maxlen = kClasses[c].maxlen;
            T->heap_bytes = std::min<uint64_t>(uint64_t(double(free_b) * kClasses[c].heap_share / kClasses[c].tables),
                                               uint64_t(24) << 30) / 256 * 256;
            T->arena_bytes = std::min<uint64_t>(uint64_t(double(free_b) * kClasses[c].arena_share / kClasses[c].tables),
                                                uint64_t(12) << 30) / 256 * 256;
            if (const char *e = getenv("MPOA_RES_ARENA_MB")) T->arena_bytes = strtoull(e, nullptr, 10) << 20;
            if (const char *e = getenv("MPOA_RES_HEAP_MB")) T->heap_bytes = strtoull(e, nullptr, 10) << 20;
            T->arena_bytes = std::max<uint64_t>(T->arena_bytes, uint64_t(1) << 20);
            T->heap_bytes = std::max<uint64_t>(T->heap_bytes, uint64_t(1) << 20);
            rc = I.build_table(*T, alpha);
            I.tables.push_back(std::move(T));
        }
    if (rc != MPOA_OK) {
        I.rc = rc;
        I.err = last_error();
        I.tables.clear();
        return;
    }
    for (auto &T : I.tables) {
        Impl::Table *tp = T.get();
        T->th = std::thread([&I, tp] { I.table_loop(*tp); });
    }
}

ResidentServer::~ResidentServer() {
    Impl &I = *d_;
    {
        std::unique_lock<std::mutex> lk(I.mu);
        I.cv.wait(lk, [&] { return I.pending == 0 || I.rc != 0; });
        I.stop = true;
    }
    I.cv.notify_all();
    for (auto &T : I.tables)
        if (T->th.joinable()) T->th.join();
    I.tables.clear();
    I.host.reset();  // finishes the host engine's jobs
}

int ResidentServer::error() {
    std::lock_guard<std::mutex> lk(d_->mu);
    return d_->rc;
}

uint64_t ResidentServer::submit(JobIn in) {
    Impl &I = *d_;
    auto J = std::make_unique<Impl::Job>();
    J->in = std::move(in);
    if (!J->in.own_bases.empty() || !J->in.own_off.empty()) {
        J->in.bases = J->in.own_bases.data();
        J->in.read_off = J->in.own_off.data();
    }
    J->out.t_submit = I.now_ms();
    J->out.nseq = J->in.first + J->in.nreads;
    J->cls = I.class_of(J->in);
    if (J->cls >= 0) {
        RCaps c = job_caps(J->in.read_off, J->in.nreads);
        if (J->in.graph) {
            // room for the extended graph's nodes and edges too
            uint64_t ne = 0;
            for (size_t id = 0; id < J->in.graph->size(); id++) ne += J->in.graph->node(id).out.size();
            c.nc = uint32_t(c.nc + J->in.graph->size());
            c.ec = uint32_t(c.ec + ne);
        }
        J->reg = region_of(c, uint32_t(I.p.alpha()));
    }
    std::unique_lock<std::mutex> lk(I.mu);
    const uint64_t t = I.jobs.size();
    Impl::Job *raw = J.get();
    I.jobs.push_back(std::move(J));
    if (raw->cls < 0 || I.rc != 0) {
        I.st.host_jobs += 1;
        lk.unlock();
        JobIn h = std::move(raw->in);
        const uint64_t id = I.host->submit(std::move(h));
        lk.lock();
        raw->engine_id = int64_t(id);
        I.cv.notify_all();
        return t;
    }
    I.pending++;
    I.queues[size_t(raw->cls)].push_back(t);
    lk.unlock();
    I.cv.notify_all();
    return t;
}

bool ResidentServer::done(uint64_t t) {
    Impl &I = *d_;
    int64_t eid;
    {
        std::lock_guard<std::mutex> lk(I.mu);
        if (t >= I.jobs.size()) return false;
        Impl::Job &S = *I.jobs[t];
        if (S.out.done) return true;
        eid = S.engine_id;
    }
    return eid >= 0 && I.host->done(uint64_t(eid));
}

bool ResidentServer::wait(uint64_t t) {
    Impl &I = *d_;
    int64_t eid;
    {
        std::unique_lock<std::mutex> lk(I.mu);
        if (t >= I.jobs.size()) return false;
        Impl::Job &S = *I.jobs[t];
        I.cv.wait(lk, [&] { return S.out.done || S.engine_id >= 0 || I.rc != 0; });
        if (S.out.done) return true;
        if (S.engine_id < 0) return false;
        eid = S.engine_id;
    }
    return I.host->wait(uint64_t(eid));
}

JobOut &ResidentServer::out(uint64_t t) {
    Impl &I = *d_;
    std::lock_guard<std::mutex> lk(I.mu);
    Impl::Job &S = *I.jobs[t];
    if (S.engine_id >= 0 && !S.engine_copied && I.host->done(uint64_t(S.engine_id))) {
        // The host engine's result, on this server's clock and with the
        // cells of the part that ran on the device.
        JobOut &h = I.host->out(uint64_t(S.engine_id));
        const double t_submit = S.out.t_submit, t_admit = S.out.t_admit;
        const uint64_t cells = S.out.cells;
        const uint32_t nseq = S.out.nseq;
        S.out = std::move(h);
        S.out.t_submit = t_submit;
        S.out.t_admit = t_admit > 0 ? t_admit : S.out.t_admit;
        S.out.cells += cells;
        S.out.nseq = nseq;
        S.engine_copied = true;
    }
    return S.out;
}

uint64_t ResidentServer::submitted() {
    std::lock_guard<std::mutex> lk(d_->mu);
    return d_->jobs.size();
}

double ResidentServer::now_ms() const { return d_->now_ms(); }

EngineStats ResidentServer::stats(ResidentStats *rs) {
    Impl &I = *d_;
    EngineStats es = I.host->stats();
    std::lock_guard<std::mutex> lk(I.mu);
    es.run.cells += I.st.cells;
    es.run.pred_cells += I.st.pred_cells;
    es.run.arena_bytes += I.st.arena_bytes;
    es.run.dp_calls += I.st.dp_calls;
    es.run.retries += I.st.retries;
    es.run.arena_retries += I.st.arena_retries;
    uint64_t rep = 0;
    for (const auto &T : I.tables) rep += T->replays;
    es.run.fill_launches += rep;
    if (rs) {
        *rs = I.st;
        rs->replays = rep;
        rs->buckets = I.tables.size();
    }
    return es;
}

}  // namespace mpoa
