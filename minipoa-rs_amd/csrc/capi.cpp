// C ABI of the gfx950 POA backend (include/minipoa_hip.h): the drop-in DP
// entry points, the library-owned graph and the batched progressive engine.
#include <algorithm>
#include <cctype>
#include <cerrno>
#include <climits>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <new>
#include <memory>
#include <mutex>
#include <thread>
#include <string>
#include <vector>

#include "engine.hpp"
#include "graph.hpp"
#include "io.hpp"
#include "minipoa_hip.h"
#include "plan.hpp"
#include "resident_host.hpp"
#include "runtime.hpp"
#include "seed.hpp"

using namespace mpoa;

struct mpoa_ctx {
    Ctx *c = nullptr;
    std::mutex mu;
    std::vector<std::unique_ptr<Ctx>> lanes;  // batched-engine launch slots
    std::unique_ptr<TaskQueue> tasks;         // batched-engine host workers
    bool engine_busy = false;                 // a batch or a server holds the slots (under mu)
    ResidentCache *resident = nullptr;        // resident batches' per-slot buffers
};
struct mpoa_graph {
    Graph g;
    explicit mpoa_graph(const Params &p) : g(p) {}
};
struct mpoa_records {
    std::vector<SeqRecord> r;
};
struct mpoa_server {
    mpoa_ctx *ctx = nullptr;
    Params params;
    std::unique_ptr<Engine> engine;          // host-driven serving
    std::unique_ptr<ResidentServer> res;     // serving on device-resident graphs
    uint64_t submitted() { return res ? res->submitted() : engine->submitted(); }
    bool done(uint64_t t) { return res ? res->done(t) : engine->done(t); }
    bool wait(uint64_t t) { return res ? res->wait(t) : engine->wait(t); }
    int error() { return res ? res->error() : engine->error(); }
    JobOut &out(uint64_t t) { return res ? res->out(t) : engine->out(t); }
    double now_ms() { return res ? res->now_ms() : engine->now_ms(); }
    uint64_t submit(JobIn in) { return res ? res->submit(std::move(in)) : engine->submit(std::move(in)); }
};
struct mpoa_batch {
    std::vector<int> status;
    std::vector<std::string> err, out0, out1;
    std::vector<std::unique_ptr<Graph>> graphs;
    std::vector<uint32_t> nreads;
    Params params;
    mpoa_batch_stats stats{};
};

namespace {

int fail(int status, const std::string &msg) {
    set_error(msg);
    return status;
}

// No C++ exception crosses the C ABI: an entry point that throws (host
// allocation, in practice) returns a status with the message set.
int exception_status() {
    try {
        throw;
    } catch (const std::bad_alloc &) {
        return fail(MPOA_ERR_ALLOC, "host memory allocation failed");
    } catch (const std::exception &e) {
        return fail(MPOA_ERR_INVALID_ARGUMENT, std::string("unexpected error: ") + e.what());
    } catch (...) {
        return fail(MPOA_ERR_INVALID_ARGUMENT, "unexpected error");
    }
}

// Reference-visible DP errors (src/align.rs returns Err for each of them), on
// which `poa` retries with the exact full matrix.
bool is_dp_error(int s) { return s >= MPOA_ERR_ANCHOR_NOT_IN_GRAPH && s <= MPOA_ERR_ALLOC; }

template <class G>
int poa_once(Ctx &ctx, const Params &p, const G &g, uint64_t beg, uint64_t end, const uint8_t *q,
             uint64_t qlen, bool adaptive, uint64_t band_b, int force_bits,
             std::vector<AlignedBase> &out, int64_t *score, const FixedScratch *fx = nullptr) {
    JobPlan J;
    plan_job(p, g, beg, end, q, qlen, adaptive, band_b, force_bits, J);
    if (J.status != MPOA_OK) return fail(J.status, J.err);
    use_layout(J, fx != nullptr || ctx.layout == kLayoutRecords);
    std::vector<JobPlan *> plans{&J};
    std::vector<DevResult> res;
    std::vector<uint32_t> tb;
    std::vector<uint64_t> off;
    int rc = run_plans(ctx, p, plans, res, tb, off, nullptr, fx);
    if (rc != MPOA_OK) return rc;
    std::string msg;
    int s = device_status(res[0], msg);
    if (s != MPOA_OK) return fail(s, msg);
    assemble(g, J, tb.data() + off[0], out);
    if (score) *score = res[0].score;
    return MPOA_OK;
}

// src/align.rs:192-211
template <class G>
int poa_retry(Ctx &ctx, const Params &p, const G &g, uint64_t beg, uint64_t end,
              const uint8_t *q, uint64_t qlen, bool adaptive, uint64_t band_b,
              std::vector<AlignedBase> &out, int64_t *score, const FixedScratch *fx = nullptr) {
    int rc = poa_once(ctx, p, g, beg, end, q, qlen, adaptive, band_b, 0, out, score, fx);
    if (rc != MPOA_OK && p.f != 0 && is_dp_error(rc)) {
        Params full = p;
        full.f = 0;
        rc = poa_once(ctx, full, g, beg, end, q, qlen, false, band_b, 0, out, score, fx);
    }
    return rc;
}

// src/align.rs:162-187: every window of one query in one launch; windows
// whose DP fails retry with the full matrix (poa's retry) in a second
// launch; the first failing window in query order decides the error.
template <class G>
int run_windows(Ctx &ctx, const Params &p, const G &g, const uint8_t *q,
                const std::vector<SeedWindow> &win, std::vector<AlignedBase> &out) {
    const size_t W = win.size();
    Params full = p;
    full.f = 0;
    std::vector<JobPlan> plan(W);
    std::vector<int> st(W, MPOA_OK);
    std::vector<std::string> msg(W);
    std::vector<std::vector<AlignedBase>> path(W);
    std::vector<size_t> todo(W);
    for (size_t k = 0; k < W; k++) todo[k] = k;
    for (int attempt = 0; attempt < 2 && !todo.empty(); attempt++) {
        const Params &pp = attempt ? full : p;
        std::vector<JobPlan *> launch;
        std::vector<size_t> who;
        for (size_t k : todo) {
            const SeedWindow &w = win[k];
            plan_job(pp, g, w.beg, w.end, q + w.qbeg, w.qend - w.qbeg, attempt ? false : w.adaptive, w.band, 0,
                     plan[k]);
            if (plan[k].status != MPOA_OK) {
                st[k] = plan[k].status;
                msg[k] = plan[k].err;
                continue;
            }
            use_layout(plan[k], ctx.layout == kLayoutRecords);
            launch.push_back(&plan[k]);
            who.push_back(k);
        }
        if (!launch.empty()) {
            std::vector<DevResult> res;
            std::vector<uint32_t> tb;
            std::vector<uint64_t> off;
            const int rc = run_plans(ctx, pp, launch, res, tb, off, nullptr);
            if (rc != MPOA_OK) return rc;
            for (size_t i = 0; i < who.size(); i++) {
                const size_t k = who[i];
                st[k] = device_status(res[i], msg[k]);
                if (st[k] == MPOA_OK) assemble(g, plan[k], tb.data() + off[i], path[k]);
            }
        }
        std::vector<size_t> again;
        if (attempt == 0 && p.f != 0)
            for (size_t k : todo)
                if (st[k] != MPOA_OK && is_dp_error(st[k])) again.push_back(k);
        todo.swap(again);
    }
    for (size_t k = 0; k < W; k++)
        if (st[k] != MPOA_OK) return fail(st[k], msg[k]);
    out.clear();
    for (size_t k = 0; k < W; k++) append_window(g, win[k], path[k], out);
    return MPOA_OK;
}

// src/align.rs:81-188
template <class G>
int align_one(Ctx &ctx, const Params &p, const G &g, const uint8_t *q, uint64_t qlen,
              std::vector<AlignedBase> &out) {
    if (g.size() <= 2) {
        out.clear();
        out.push_back(AlignedBase{0, int64_t(p.encode('N')), -1});
        for (uint64_t k = 0; k < qlen; k++) out.push_back(AlignedBase{-1, int64_t(p.encode(q[k])), -1});
        return MPOA_OK;
    }
    if (!p.seeding) return poa_retry(ctx, p, g, 0, 1, q, qlen, p.adaptive, p.b, out, nullptr);
    std::vector<SeedWindow> win;
    std::string err;
    const int rc = seed_windows(p, g, q, qlen, win, err);
    if (rc != MPOA_OK) return fail(rc, err);
    return run_windows(ctx, p, g, q, win, out);
}

int emit(const std::vector<AlignedBase> &v, mpoa_aligned_base *out, uint64_t cap, uint64_t *len) {
    if (len) *len = v.size();
    if (v.size() > cap) return fail(MPOA_ERR_BUFFER_TOO_SMALL, "output buffer is too small");
    if (!v.empty()) memcpy(out, v.data(), v.size() * sizeof(AlignedBase));
    return MPOA_OK;
}

int emit_text(const std::string &s, char *out, uint64_t cap, uint64_t *len) {
    if (len) *len = s.size();
    if (s.size() > cap) return fail(MPOA_ERR_BUFFER_TOO_SMALL, "output buffer is too small");
    if (!s.empty()) memcpy(out, s.data(), s.size());
    return MPOA_OK;
}

bool check_params(const mpoa_params *params, Params &p, int &rc) {
    if (!params) {
        rc = fail(MPOA_ERR_INVALID_ARGUMENT, "params is NULL");
        return false;
    }
    p = Params::from_c(*params);
    std::string v = p.validate();
    if (!v.empty()) {
        rc = fail(MPOA_ERR_INVALID_PARAMS, v);
        return false;
    }
    return true;
}

bool view_ok(const mpoa_graph_view *g, const Params &p, int &rc) {
    if (!g || !g->rank || !g->base || !g->in_off || !g->out_off || !g->order || !g->parent ||
        !g->aligned) {
        rc = fail(MPOA_ERR_INVALID_ARGUMENT, "incomplete graph view");
        return false;
    }
    if (g->alphabet_len != p.alpha()) {
        rc = fail(MPOA_ERR_INVALID_ARGUMENT, "graph alphabet differs from params");
        return false;
    }
    return true;
}

std::vector<std::string> read_names(size_t n) {
    std::vector<std::string> v(n);
    for (size_t k = 0; k < n; k++) v[k] = "read" + std::to_string(k);
    return v;
}

double ms_since(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

// Launch slots (a stream with its own buffers each) and the host workers of
// the batched engine are created on first use and kept with the context.
int ensure_engine(mpoa_ctx *ctx, int slots) {
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (ctx->engine_busy) return fail(MPOA_ERR_INVALID_ARGUMENT, "the context's launch slots are in use by a batch or a server");
    while (int(ctx->lanes.size()) < slots) {
        Ctx *c = nullptr;
        int rc = ctx_create(ctx->c->device, &c, 2);  // 2 threads pack each launch
        if (rc != MPOA_OK) return rc;
        c->layout = ctx->c->layout;
        ctx->lanes.emplace_back(c);
    }
    if (!ctx->tasks) ctx->tasks = std::make_unique<TaskQueue>(std::max(1u, default_threads() - 1));
    ctx->engine_busy = true;
    return MPOA_OK;
}

void release_engine(mpoa_ctx *ctx) {
    std::lock_guard<std::mutex> lk(ctx->mu);
    ctx->engine_busy = false;
}

// Hardware queues the HIP runtime opens per process: GPU_MAX_HW_QUEUES (4
// when unset). Streams beyond that share queues, and a launch slot whose
// stream shares a queue waits behind another stream's work.
int hw_queues() {
    const char *s = getenv("GPU_MAX_HW_QUEUES");
    const int q = s ? atoi(s) : 0;
    return q > 0 ? q : 4;
}

// Launch slots of the batched engine. With four queues: three, so that with
// the context's own stream every stream has a queue. With eight or more:
// five (measured on the 4,096-job workload with eight queues, wall time per
// run: 3 slots 1.43-1.50 s, 4 slots 1.34-1.39, 5 slots 1.29, 6 slots 1.52).
int engine_slots(int want) {
    int n = want > 0 ? want : (hw_queues() >= 8 ? 5 : 3);
    if (const char *s = getenv("MPOA_SLOTS")) n = atoi(s);
    return std::max(1, std::min(8, n));
}

void fill_stats(const EngineStats &es, mpoa_batch_stats &S) {
    S.cells = es.run.cells;
    S.dp_calls = es.run.dp_calls;
    S.retries = es.run.retries;
    S.i64_calls = es.run.i64_calls;
    S.pred_cells = es.run.pred_cells;
    S.fill_launches = es.run.fill_launches;
    S.device_ms = es.run.device_ms;
    S.fill_ms = es.run.fill_ms;
    S.host_ms = es.host_ms;
    S.plan_ms = es.plan_ms;
    S.consume_ms = es.consume_ms;
    S.fuse_ms = es.fuse_ms;
    S.render_ms = es.render_ms;
    S.pack_ms = es.pack_ms;
    S.arena_retries = es.run.arena_retries;
    S.arena_bytes = es.run.arena_bytes;
    S.record_launches = es.run.record_launches;
}

}  // namespace

extern "C" {

const char *mpoa_backend_name(void) { return "hip-gfx950 (wave64 x i32)"; }
const char *mpoa_last_error(void) { return last_error(); }

int mpoa_ctx_create(int device, mpoa_ctx **out) try {
    if (!out) return fail(MPOA_ERR_INVALID_ARGUMENT, "out is NULL");
    *out = nullptr;
    Ctx *c = nullptr;
    int rc = ctx_create(device, &c);
    if (rc != MPOA_OK) return rc;
    // MPOA_LAYOUT (values / records) sets the DP layout of every context the
    // process creates, for side-by-side runs; mpoa_ctx_set_layout overrides it.
    if (const char *lay = getenv("MPOA_LAYOUT")) {
        if (!strcmp(lay, "values")) c->layout = kLayoutValues;
        if (!strcmp(lay, "records")) c->layout = kLayoutRecords;
    }
    auto *h = new mpoa_ctx();
    h->c = c;
    *out = h;
    return MPOA_OK;
} catch (...) { return exception_status(); }

int mpoa_ctx_set_layout(mpoa_ctx *ctx, int layout) try {
    if (!ctx) return fail(MPOA_ERR_INVALID_ARGUMENT, "ctx is NULL");
    if (layout != MPOA_LAYOUT_AUTO && layout != MPOA_LAYOUT_VALUES && layout != MPOA_LAYOUT_RECORDS)
        return fail(MPOA_ERR_INVALID_ARGUMENT, "unknown DP layout");
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (ctx->engine_busy) return fail(MPOA_ERR_INVALID_ARGUMENT, "the context's launch slots are in use by a batch or a server");
    ctx->c->layout = layout;
    for (auto &c : ctx->lanes) c->layout = layout;
    return MPOA_OK;
} catch (...) { return exception_status(); }

void mpoa_ctx_destroy(mpoa_ctx *ctx) {
    if (!ctx) return;
    resident_cache_free(ctx->resident);
    ctx->lanes.clear();
    delete ctx->c;
    delete ctx;
}

uint64_t mpoa_ctx_device_bytes(const mpoa_ctx *ctx) try {
    if (!ctx) return 0;
    uint64_t total = 0;
    auto add = [&](const Ctx &c) {
        total += c.d_in.cap + c.d_res.cap + c.d_meta.cap + c.d_adapt.cap + c.d_tb.cap + c.d_arena.cap;
    };
    add(*ctx->c);
    for (const auto &l : ctx->lanes) add(*l);
    return total + resident_cache_bytes(ctx->resident);
} catch (...) {
    (void)exception_status();
    return 0;
}

int mpoa_load_matrix(const char *text, uint64_t len, uint8_t *alphabet, int32_t *matrix, uint32_t cap,
                     uint32_t *alphabet_len) try {
    if (!text || !alphabet_len || (cap && (!alphabet || !matrix)))
        return fail(MPOA_ERR_INVALID_ARGUMENT, "NULL argument");
    *alphabet_len = 0;
    // Lines: '#' comments and blank lines skipped; the first other line is
    // the header (one single-character label per column); then one row per
    // label, in any order: the label and one integer per column.
    std::vector<std::vector<std::string>> lines;
    {
        size_t at = 0;
        while (at < len) {
            size_t end = at;
            while (end < len && text[end] != '\n') end++;
            std::string line(text + at, end - at);
            at = end + 1;
            if (!line.empty() && line.back() == '\r') line.pop_back();
            size_t k = 0;
            while (k < line.size() && isspace((unsigned char)line[k])) k++;
            if (k == line.size() || line[k] == '#') continue;
            std::vector<std::string> tok;
            size_t i = 0;
            while (i < line.size()) {
                while (i < line.size() && isspace((unsigned char)line[i])) i++;
                size_t j = i;
                while (j < line.size() && !isspace((unsigned char)line[j])) j++;
                if (j > i) tok.emplace_back(line.substr(i, j - i));
                i = j;
            }
            lines.push_back(std::move(tok));
        }
    }
    if (lines.empty()) return fail(MPOA_ERR_INVALID_PARAMS, "scoring matrix has no header");
    const std::vector<std::string> &head = lines[0];
    const size_t n = head.size();
    int pos[256];
    for (int &x : pos) x = -1;
    for (size_t c = 0; c < n; c++) {
        if (head[c].size() != 1) return fail(MPOA_ERR_INVALID_PARAMS, "scoring matrix labels must be single characters");
        const unsigned char ch = (unsigned char)head[c][0];
        if (pos[ch] >= 0) return fail(MPOA_ERR_INVALID_PARAMS, "duplicate scoring matrix label " + head[c]);
        pos[ch] = int(c);
    }
    if (lines.size() - 1 != n) return fail(MPOA_ERR_INVALID_PARAMS, "scoring matrix needs one row per label");
    std::vector<int32_t> m(n * n, 0);
    std::vector<uint8_t> seen(n, 0);
    for (size_t r = 1; r < lines.size(); r++) {
        const std::vector<std::string> &row = lines[r];
        if (row[0].size() != 1 || pos[(unsigned char)row[0][0]] < 0)
            return fail(MPOA_ERR_INVALID_PARAMS, "scoring matrix row label " + row[0] + " is not in the header");
        const size_t a = size_t(pos[(unsigned char)row[0][0]]);
        if (seen[a]) return fail(MPOA_ERR_INVALID_PARAMS, "duplicate scoring matrix row " + row[0]);
        seen[a] = 1;
        if (row.size() != n + 1) return fail(MPOA_ERR_INVALID_PARAMS, "scoring matrix row " + row[0] + " has the wrong length");
        for (size_t c = 0; c < n; c++) {
            const std::string &v = row[c + 1];
            char *endp = nullptr;
            errno = 0;
            const long x = strtol(v.c_str(), &endp, 10);
            if (v.empty() || *endp != '\0' || errno != 0 || x < INT32_MIN || x > INT32_MAX)
                return fail(MPOA_ERR_INVALID_PARAMS, "scoring matrix entry " + v + " is not an integer");
            m[a * n + c] = int32_t(x);
        }
    }
    *alphabet_len = uint32_t(n);
    if (n > cap) return fail(MPOA_ERR_BUFFER_TOO_SMALL, "alphabet buffer is too small");
    for (size_t c = 0; c < n; c++) alphabet[c] = (uint8_t)head[c][0];
    memcpy(matrix, m.data(), m.size() * sizeof(int32_t));
    return MPOA_OK;
} catch (...) { return exception_status(); }

// ---- input formats (the reference's io module) ----

int mpoa_read_sequences(const uint8_t *data, uint64_t len, mpoa_records **out) try {
    if (!out || (!data && len)) return fail(MPOA_ERR_INVALID_ARGUMENT, "NULL argument");
    *out = nullptr;
    std::vector<uint8_t> text;
    std::string e = gunzip_if_gzip(data, len, text);
    auto r = std::make_unique<mpoa_records>();
    if (e.empty()) e = parse_sequences(text.data(), text.size(), r->r);
    if (!e.empty()) return fail(MPOA_ERR_INVALID_ARGUMENT, e);
    *out = r.release();
    return MPOA_OK;
} catch (...) { return exception_status(); }

uint64_t mpoa_records_len(const mpoa_records *r) { return r ? r->r.size() : 0; }

int mpoa_record_get(const mpoa_records *r, uint64_t k, const char **name, uint64_t *name_len,
                    const char **comment, uint64_t *comment_len, const uint8_t **seq,
                    uint64_t *seq_len, const uint8_t **qual, uint64_t *qual_len) try {
    if (!r || k >= r->r.size()) return fail(MPOA_ERR_INVALID_ARGUMENT, "record out of range");
    const SeqRecord &x = r->r[k];
    if (name) *name = x.name.c_str();
    if (name_len) *name_len = x.name.size();
    if (comment) *comment = x.comment.c_str();
    if (comment_len) *comment_len = x.comment.size();
    if (seq) *seq = x.seq.data();
    if (seq_len) *seq_len = x.seq.size();
    if (qual) *qual = x.qual.data();
    if (qual_len) *qual_len = x.qual.size();
    return MPOA_OK;
} catch (...) { return exception_status(); }

void mpoa_records_free(mpoa_records *r) { delete r; }

int mpoa_read_gfa(const mpoa_params *params, const uint8_t *data, uint64_t len, mpoa_graph **graph,
                  mpoa_records **records) try {
    if (!graph || !records || (!data && len)) return fail(MPOA_ERR_INVALID_ARGUMENT, "NULL argument");
    *graph = nullptr;
    *records = nullptr;
    Params p;
    int rc = MPOA_OK;
    if (!check_params(params, p, rc)) return rc;
    std::vector<uint8_t> text;
    std::string e = gunzip_if_gzip(data, len, text);
    if (!e.empty()) return fail(MPOA_ERR_INVALID_ARGUMENT, e);
    auto g = std::make_unique<mpoa_graph>(p);
    auto r = std::make_unique<mpoa_records>();
    e = read_gfa(p, text.data(), text.size(), g->g, r->r);
    if (!e.empty()) return fail(MPOA_ERR_GRAPH, e);
    *graph = g.release();
    *records = r.release();
    return MPOA_OK;
} catch (...) { return exception_status(); }

int mpoa_params_validate(const mpoa_params *params) try {
    Params p;
    int rc = MPOA_OK;
    check_params(params, p, rc);
    return rc;
} catch (...) { return exception_status(); }

int mpoa_i32_scores_safe(const mpoa_params *params, const mpoa_graph_view *graph,
                         uint64_t beg_id, uint64_t end_id, uint64_t query_len) try {
    if (!params || !graph) return 0;
    Params p = Params::from_c(*params);
    return i32_scores_safe(p, ViewAcc{*graph}, beg_id, end_id, query_len) ? 1 : 0;
} catch (...) { return exception_status(); }

static int poa_entry(mpoa_ctx *ctx, const mpoa_params *params, const mpoa_graph_view *graph,
                     uint64_t beg_id, uint64_t end_id, const uint8_t *query, uint64_t qlen,
                     int adaptive, uint64_t band_b, int mode, mpoa_aligned_base *out,
                     uint64_t out_cap, uint64_t *out_len, int64_t *out_score) {
    if (!ctx || (!query && qlen)) return fail(MPOA_ERR_INVALID_ARGUMENT, "NULL argument");
    Params p;
    int rc = MPOA_OK;
    if (!check_params(params, p, rc) || !view_ok(graph, p, rc)) return rc;
    ViewAcc g{*graph};
    std::vector<AlignedBase> v;
    if (mode == 0) {
        rc = poa_retry(*ctx->c, p, g, beg_id, end_id, query, qlen, adaptive != 0, band_b, v, out_score);
    } else if (mode == 1) {
        rc = poa_once(*ctx->c, p, g, beg_id, end_id, query, qlen, adaptive != 0, band_b, 0, v, out_score);
    } else {
        if (mode == 32 && !i32_scores_safe(p, g, beg_id, end_id, qlen))
            return fail(MPOA_ERR_INVALID_ARGUMENT, "scores exceed the i32 guard");
        rc = poa_once(*ctx->c, p, g, beg_id, end_id, query, qlen, adaptive != 0, band_b, mode, v,
                      out_score);
    }
    if (rc != MPOA_OK) return rc;
    return emit(v, out, out_cap, out_len);
}

int mpoa_poa(mpoa_ctx *ctx, const mpoa_params *params, const mpoa_graph_view *graph,
             uint64_t beg_id, uint64_t end_id, const uint8_t *query, uint64_t qlen, int adaptive,
             uint64_t band_b, mpoa_aligned_base *out, uint64_t out_cap, uint64_t *out_len,
             int64_t *out_score) try {
    return poa_entry(ctx, params, graph, beg_id, end_id, query, qlen, adaptive, band_b, 0, out,
                     out_cap, out_len, out_score);
} catch (...) { return exception_status(); }

// One lock per device: calls of the no-allocation entry point on one GPU
// share its scratch array, so they run one at a time within the process.
static std::mutex g_noalloc_mu[kScratchDevices];

int mpoa_poa_noalloc(mpoa_ctx *ctx, const mpoa_params *params, const mpoa_graph_view *graph,
                     uint64_t beg_id, uint64_t end_id, const uint8_t *query, uint64_t qlen,
                     int adaptive, uint64_t band_b, mpoa_aligned_base *out, uint64_t out_cap,
                     uint64_t *out_len, int64_t *out_score) try {
    if (!ctx || (!query && qlen)) return fail(MPOA_ERR_INVALID_ARGUMENT, "NULL argument");
    Params p;
    int rc = MPOA_OK;
    if (!check_params(params, p, rc) || !view_ok(graph, p, rc)) return rc;
    const int dev = ctx->c->device;
    FixedScratch fx;
    fx.dev = static_cast<char *>(scratch_device_ptr(dev));
    fx.bytes = kScratchBytes;
    if (!fx.dev) return fail(MPOA_ERR_DEVICE, "the device scratch of this GPU is not available");
    std::vector<AlignedBase> v;
    {
        std::lock_guard<std::mutex> lk(g_noalloc_mu[dev]);
        rc = poa_retry(*ctx->c, p, ViewAcc{*graph}, beg_id, end_id, query, qlen, adaptive != 0,
                       band_b, v, out_score, &fx);
    }
    if (rc != MPOA_OK) return rc;
    return emit(v, out, out_cap, out_len);
} catch (...) { return exception_status(); }

int mpoa_poa_inner(mpoa_ctx *ctx, const mpoa_params *params, const mpoa_graph_view *graph,
                   uint64_t beg_id, uint64_t end_id, const uint8_t *query, uint64_t qlen,
                   int adaptive, uint64_t band_b, mpoa_aligned_base *out, uint64_t out_cap,
                   uint64_t *out_len, int64_t *out_score) try {
    return poa_entry(ctx, params, graph, beg_id, end_id, query, qlen, adaptive, band_b, 1, out,
                     out_cap, out_len, out_score);
} catch (...) { return exception_status(); }

int mpoa_poa_kernel(mpoa_ctx *ctx, const mpoa_params *params, const mpoa_graph_view *graph,
                    uint64_t beg_id, uint64_t end_id, const uint8_t *query, uint64_t qlen,
                    int adaptive, uint64_t band_b, int score_bits, mpoa_aligned_base *out,
                    uint64_t out_cap, uint64_t *out_len, int64_t *out_score) try {
    if (score_bits != 32 && score_bits != 64)
        return fail(MPOA_ERR_INVALID_ARGUMENT, "score_bits must be 32 or 64");
    return poa_entry(ctx, params, graph, beg_id, end_id, query, qlen, adaptive, band_b, score_bits,
                     out, out_cap, out_len, out_score);
} catch (...) { return exception_status(); }

int mpoa_align(mpoa_ctx *ctx, const mpoa_params *params, const mpoa_graph_view *graph,
               const uint8_t *query, uint64_t qlen, mpoa_aligned_base *out, uint64_t out_cap,
               uint64_t *out_len) try {
    if (!ctx || (!query && qlen)) return fail(MPOA_ERR_INVALID_ARGUMENT, "NULL argument");
    Params p;
    int rc = MPOA_OK;
    if (!check_params(params, p, rc) || !view_ok(graph, p, rc)) return rc;
    std::vector<AlignedBase> v;
    rc = align_one(*ctx->c, p, ViewAcc{*graph}, query, qlen, v);
    if (rc != MPOA_OK) return rc;
    return emit(v, out, out_cap, out_len);
} catch (...) { return exception_status(); }

// ---- library-owned graph ----

mpoa_graph *mpoa_graph_new(const mpoa_params *params) try {
    if (!params) return nullptr;
    Params p = Params::from_c(*params);
    if (p.alpha() == 0) {
        set_error("invalid scoring matrix dimensions");
        return nullptr;
    }
    return new mpoa_graph(p);
} catch (...) {
    (void)exception_status();
    return nullptr;
}

void mpoa_graph_free(mpoa_graph *g) { delete g; }

// The reference panics on an out-of-range base or node id (Rust bounds
// checks, graph.rs:42-80); here they are refused before the graph changes.
uint64_t mpoa_graph_add_node(mpoa_graph *g, uint64_t base) try {
    if (!g) {
        set_error("graph is NULL");
        return UINT64_MAX;
    }
    if (base >= g->g.alpha()) {
        set_error("node base outside the alphabet");
        return UINT64_MAX;
    }
    if (g->g.size() >= uint64_t(UINT32_MAX)) {
        set_error("graph holds the maximum number of nodes");
        return UINT64_MAX;
    }
    return g->g.add_node(uint32_t(base));
} catch (...) {
    (void)exception_status();
    return UINT64_MAX;
}

int mpoa_graph_add_edge(mpoa_graph *g, uint64_t from, uint64_t to) try {
    if (!g) return fail(MPOA_ERR_INVALID_ARGUMENT, "graph is NULL");
    if (from >= g->g.size() || to >= g->g.size())
        return fail(MPOA_ERR_INVALID_ARGUMENT, "edge references an unknown node");
    g->g.add_edge(uint32_t(from), uint32_t(to));
    return MPOA_OK;
} catch (...) { return exception_status(); }

int mpoa_graph_add_path(mpoa_graph *g, uint64_t seq_id, const mpoa_aligned_base *a, uint64_t len) try {
    std::string e = g->g.add_path(seq_id, a, len);
    return e.empty() ? MPOA_OK : fail(MPOA_ERR_GRAPH, e);
} catch (...) { return exception_status(); }

int mpoa_graph_toposort(mpoa_graph *g, int grouped) try {
    std::string e = g->g.toposort(grouped != 0);
    return e.empty() ? MPOA_OK : fail(MPOA_ERR_GRAPH, e);
} catch (...) { return exception_status(); }

int mpoa_graph_view_of(mpoa_graph *g, mpoa_graph_view *out) try {
    if (!g || !out) return fail(MPOA_ERR_INVALID_ARGUMENT, "NULL argument");
    *out = g->g.view();
    return MPOA_OK;
} catch (...) { return exception_status(); }

int mpoa_graph_path(const mpoa_graph *g, uint64_t k, uint64_t *seq_id, uint64_t *ids, uint64_t cap,
                    uint64_t *out_len) try {
    if (!g || k >= g->g.paths().size()) return fail(MPOA_ERR_INVALID_ARGUMENT, "path out of range");
    const auto &pa = g->g.paths()[k];
    if (seq_id) *seq_id = pa.first;
    if (out_len) *out_len = pa.second.size();
    if (pa.second.size() > cap) return fail(MPOA_ERR_BUFFER_TOO_SMALL, "output buffer is too small");
    for (size_t t = 0; t < pa.second.size(); t++) ids[t] = pa.second[t];
    return MPOA_OK;
} catch (...) { return exception_status(); }

uint64_t mpoa_graph_n_paths(const mpoa_graph *g) try { return g->g.paths().size(); } catch (...) {
    (void)exception_status();
    return 0;
}

uint64_t mpoa_graph_consensus(const mpoa_graph *g, const mpoa_params *params, uint8_t *seq,
                              uint64_t *ids, uint64_t *coverage, uint64_t cap) try {
    Params p = Params::from_c(*params);
    std::vector<uint32_t> v;
    g->g.consensus(v);
    for (size_t k = 0; k < v.size() && k < cap; k++) {
        const Graph::Node &n = g->g.node(v[k]);
        if (seq) seq[k] = p.decode(n.base);
        if (ids) ids[k] = v[k];
        if (coverage) {
            uint64_t s = 0;
            for (const Edge &ed : n.out) s += ed.w;
            coverage[k] = std::max<uint64_t>(n.coverage, s);
        }
    }
    return v.size();
} catch (...) {
    (void)exception_status();
    return 0;
}

int mpoa_graph_msa(mpoa_graph *g, const mpoa_params *params, const char *const *names,
                   const char *const *comments, uint64_t nseq, char *out, uint64_t cap,
                   uint64_t *out_len) try {
    Params p = Params::from_c(*params);
    std::vector<std::string> nm(nseq), cm(nseq);
    for (uint64_t k = 0; k < nseq; k++) {
        nm[k] = names && names[k] ? names[k] : "";
        cm[k] = comments && comments[k] ? comments[k] : "";
    }
    std::string text;
    std::string e = g->g.msa(p, nm, cm, text);
    if (!e.empty()) return fail(MPOA_ERR_GRAPH, e);
    return emit_text(text, out, cap, out_len);
} catch (...) { return exception_status(); }

int mpoa_graph_gfa(const mpoa_graph *g, const mpoa_params *params, const char *const *names,
                   uint64_t nseq, char *out, uint64_t cap, uint64_t *out_len) try {
    Params p = Params::from_c(*params);
    std::vector<std::string> nm(nseq);
    for (uint64_t k = 0; k < nseq; k++) nm[k] = names && names[k] ? names[k] : "";
    std::string text;
    std::string e = g->g.gfa(p, nm, text);
    if (!e.empty()) return fail(MPOA_ERR_GRAPH, e);
    return emit_text(text, out, cap, out_len);
} catch (...) { return exception_status(); }

int mpoa_extend_graph(mpoa_ctx *ctx, const mpoa_params *params, mpoa_graph *g,
                      const uint8_t *const *seqs, const uint64_t *lens, uint64_t nseq,
                      uint64_t existing) try {
    if (!ctx || !g) return fail(MPOA_ERR_INVALID_ARGUMENT, "NULL argument");
    Params p;
    int rc = MPOA_OK;
    if (!check_params(params, p, rc)) return rc;
    // src/lib.rs:189-193: the guide order with -p on a new graph, else input
    // order; the read keeps its input index as sequence id.
    std::vector<uint32_t> order;
    if (p.progressive && existing == 0) {
        if (nseq > UINT32_MAX) return fail(MPOA_ERR_INVALID_ARGUMENT, "too many sequences");
        guide_order(p, seqs, lens, nseq, order);
    }
    std::vector<AlignedBase> path;
    for (uint64_t index = existing; index < nseq; index++) {
        const uint64_t id = order.empty() ? index : order[index];
        rc = align_one(*ctx->c, p, GraphAcc{g->g}, seqs[id], lens[id], path);
        if (rc != MPOA_OK) return rc;
        std::string e = g->g.add_path(id, path.data(), path.size());
        if (e.empty()) e = g->g.toposort(false);
        if (!e.empty()) return fail(MPOA_ERR_GRAPH, e);
    }
    return MPOA_OK;
} catch (...) { return exception_status(); }

int mpoa_align_sequences(mpoa_ctx *ctx, const mpoa_params *params, const uint8_t *const *seqs,
                         const uint64_t *lens, uint64_t nseq, mpoa_graph **out) try {
    if (!out) return fail(MPOA_ERR_INVALID_ARGUMENT, "out is NULL");
    *out = nullptr;
    Params p;
    int rc = MPOA_OK;
    if (!check_params(params, p, rc)) return rc;
    auto *g = new mpoa_graph(p);
    rc = mpoa_extend_graph(ctx, params, g, seqs, lens, nseq, 0);
    if (rc != MPOA_OK) {
        delete g;
        return rc;
    }
    *out = g;
    return MPOA_OK;
} catch (...) { return exception_status(); }

// ---- batched progressive POA ----

int mpoa_batch_run(mpoa_ctx *ctx, const mpoa_params *params, uint64_t njobs,
                   const uint64_t *job_off, const uint64_t *read_off, const uint8_t *bases,
                   mpoa_batch **out) try {
    if (!ctx || !out || (njobs && (!job_off || !read_off || !bases)))
        return fail(MPOA_ERR_INVALID_ARGUMENT, "NULL argument");
    *out = nullptr;
    Params p;
    int rc = MPOA_OK;
    if (!check_params(params, p, rc)) return rc;
    auto t_all = std::chrono::steady_clock::now();
    auto B = std::make_unique<mpoa_batch>();
    B->params = p;
    B->status.assign(njobs, MPOA_OK);
    B->err.assign(njobs, "");
    B->out0.assign(njobs, "");
    B->out1.assign(njobs, "");
    B->graphs.resize(njobs);
    B->nreads.resize(njobs);
    for (uint64_t j = 0; j < njobs; j++) {
        B->graphs[j] = std::make_unique<Graph>(p);
        B->nreads[j] = uint32_t(job_off[j + 1] - job_off[j]);
    }
    // Jobs advance independently; launch slots take whichever jobs are
    // ready (engine.hpp).
    const int nslots = engine_slots(0);
    resident_cache_trim(ctx->resident);
    rc = ensure_engine(ctx, nslots);
    if (rc != MPOA_OK) return rc;
    std::vector<Ctx *> slots;
    for (int k = 0; k < nslots; k++) slots.push_back(ctx->lanes[size_t(k)].get());
    BatchInput in{njobs, job_off, read_off, bases, true};
    BatchOutput eo{&B->status, &B->err, &B->out0, &B->out1, &B->graphs};
    EngineStats es;
    rc = run_engine(slots, *ctx->tasks, p, in, eo, es);
    release_engine(ctx);
    if (rc != MPOA_OK) return rc;
    mpoa_batch_stats &S = B->stats;
    fill_stats(es, S);
    S.wall_ms = ms_since(t_all);
    *out = B.release();
    return MPOA_OK;
} catch (...) { return exception_status(); }

int mpoa_batch_run_resident(mpoa_ctx *ctx, const mpoa_params *params, uint64_t njobs,
                            const uint64_t *job_off, const uint64_t *read_off, const uint8_t *bases,
                            mpoa_batch **out) try {
    if (!ctx || !out || (njobs && (!job_off || !read_off || !bases)))
        return fail(MPOA_ERR_INVALID_ARGUMENT, "NULL argument");
    *out = nullptr;
    Params p;
    int rc = MPOA_OK;
    if (!check_params(params, p, rc)) return rc;
    auto t_all = std::chrono::steady_clock::now();
    auto B = std::make_unique<mpoa_batch>();
    B->params = p;
    B->status.assign(njobs, MPOA_OK);
    B->err.assign(njobs, "");
    B->out0.assign(njobs, "");
    B->out1.assign(njobs, "");
    B->graphs.resize(njobs);
    B->nreads.resize(njobs);
    for (uint64_t j = 0; j < njobs; j++) B->nreads[j] = uint32_t(job_off[j + 1] - job_off[j]);
    // A resident bucket's step loop keeps one stream busy for its whole
    // run, and its lockstep steps leave CUs idle that other buckets fill:
    // one slot per hardware queue but one (measured on the 4,096-job
    // workload with eight queues, wall time per run: 5 slots 1.95 s, 7
    // slots 1.63 s, 8 slots 1.62 s).
    const int nslots = getenv("MPOA_SLOTS") ? engine_slots(0) : std::max(engine_slots(0), std::min(8, hw_queues() - 1));
    rc = ensure_engine(ctx, nslots);
    if (rc != MPOA_OK) return rc;
    std::vector<Ctx *> slots;
    for (int k = 0; k < nslots; k++) slots.push_back(ctx->lanes[size_t(k)].get());
    BatchInput in{njobs, job_off, read_off, bases, true};
    BatchOutput eo{&B->status, &B->err, &B->out0, &B->out1, &B->graphs};
    EngineStats es;
    ResidentStats rs;
    if (!ctx->resident) ctx->resident = resident_cache_new();
    rc = run_resident(slots, *ctx->tasks, *ctx->resident, p, in, eo, es, rs);
    release_engine(ctx);
    if (rc != MPOA_OK) return rc;
    for (uint64_t j = 0; j < njobs; j++)
        if (!B->graphs[j]) B->graphs[j] = std::make_unique<Graph>(p);
    mpoa_batch_stats &S = B->stats;
    fill_stats(es, S);
    S.resident_jobs = rs.resident_done;
    S.evicted_jobs = rs.evicted;
    S.step_replays = rs.replays;
    S.buckets = rs.buckets;
    S.deferred = rs.deferred;
    S.resident_ms = rs.resident_ms;
    S.resident_device_ms = rs.device_ms;
    S.wall_ms = ms_since(t_all);
    *out = B.release();
    return MPOA_OK;
} catch (...) { return exception_status(); }

int mpoa_resident_check(const mpoa_params *params, const uint8_t *bases, const uint64_t *read_off,
                        uint64_t nreads, const mpoa_aligned_base *paths, const uint64_t *path_off,
                        uint32_t lds_bytes, char *report, uint64_t cap) try {
    if (!params || !read_off || !path_off || (nreads && (!bases || !paths)) || nreads > UINT32_MAX)
        return fail(MPOA_ERR_INVALID_ARGUMENT, "NULL argument");
    Params p;
    int rc = MPOA_OK;
    if (!check_params(params, p, rc)) return rc;
    std::string rep;
    const int r = resident_check(p, bases, read_off, uint32_t(nreads), paths, path_off, lds_bytes, rep);
    if (report && cap) {
        const size_t n = std::min<size_t>(rep.size(), size_t(cap - 1));
        memcpy(report, rep.data(), n);
        report[n] = 0;
    }
    if (r == 0) return MPOA_OK;
    return fail(r == 2 ? MPOA_ERR_GRAPH : MPOA_ERR_DEVICE, rep);
} catch (...) { return exception_status(); }

void mpoa_batch_free(mpoa_batch *b) { delete b; }

int mpoa_batch_stats_get(const mpoa_batch *b, mpoa_batch_stats *out) try {
    if (!b || !out) return fail(MPOA_ERR_INVALID_ARGUMENT, "NULL argument");
    *out = b->stats;
    return MPOA_OK;
} catch (...) { return exception_status(); }

int mpoa_batch_job_status(const mpoa_batch *b, uint64_t job) try {
    if (!b || job >= b->status.size()) return fail(MPOA_ERR_INVALID_ARGUMENT, "job out of range");
    if (b->status[job] != MPOA_OK) set_error(b->err[job]);
    return b->status[job];
} catch (...) { return exception_status(); }

int mpoa_batch_job_output(const mpoa_batch *b, uint64_t job, int mode, char *out, uint64_t cap,
                          uint64_t *out_len) try {
    if (!b || job >= b->status.size()) return fail(MPOA_ERR_INVALID_ARGUMENT, "job out of range");
    if (b->status[job] != MPOA_OK) return fail(b->status[job], b->err[job]);
    if (mode == 0) return emit_text(b->out0[job], out, cap, out_len);
    if (mode == 1) return emit_text(b->out1[job], out, cap, out_len);
    if (mode == 2) {
        std::string text;
        std::string e = b->graphs[job]->gfa(b->params, read_names(b->nreads[job]), text);
        if (!e.empty()) return fail(MPOA_ERR_GRAPH, e);
        return emit_text(text, out, cap, out_len);
    }
    return fail(MPOA_ERR_INVALID_ARGUMENT, "mode must be 0, 1 or 2");
} catch (...) { return exception_status(); }

// ---- serving: jobs admitted as they arrive ----

int mpoa_server_create(mpoa_ctx *ctx, const mpoa_params *params, const mpoa_server_opts *opts,
                       mpoa_server **out) try {
    if (!ctx || !out) return fail(MPOA_ERR_INVALID_ARGUMENT, "NULL argument");
    *out = nullptr;
    Params p;
    int rc = MPOA_OK;
    if (!check_params(params, p, rc)) return rc;
    // Four slots by default: two for long DPs, two for short ones.
    const int nslots = engine_slots(opts && opts->slots ? int(opts->slots) : 4);
    resident_cache_trim(ctx->resident);
    rc = ensure_engine(ctx, nslots);
    if (rc != MPOA_OK) return rc;
    std::vector<Ctx *> slots;
    for (int k = 0; k < nslots; k++) slots.push_back(ctx->lanes[size_t(k)].get());
    EngineOpts o;
    o.render = opts ? opts->render != 0 : true;
    o.crit_lane = false;
    o.admit_bytes = opts ? opts->admit_bytes : 0;
    // Served jobs range from hundreds to tens of thousands of rows per DP:
    // DPs of 6,000 rows or more (about 25 ms of fill) share half the slots.
    o.long_rows = 6000;
    o.long_slots = uint32_t(nslots / 2);
    auto S = std::make_unique<mpoa_server>();
    S->ctx = ctx;
    S->params = p;
    if (opts && opts->resident) {
        S->res = std::make_unique<ResidentServer>(slots, *ctx->tasks, p, o);
        if (S->res->error() != MPOA_OK) {
            const int e = S->res->error();
            const std::string msg = last_error();
            S->res.reset();
            release_engine(ctx);
            return fail(e, msg);
        }
    } else {
        S->engine = std::make_unique<Engine>(slots, *ctx->tasks, p, o);
    }
    *out = S.release();
    return MPOA_OK;
} catch (...) { return exception_status(); }

int mpoa_server_submit(mpoa_server *s, const uint8_t *bases, const uint64_t *read_off, uint64_t nreads,
                       const mpoa_graph *extend, uint64_t existing, uint64_t *ticket) try {
    if (!s || !ticket || (nreads && (!bases || !read_off)))
        return fail(MPOA_ERR_INVALID_ARGUMENT, "NULL argument");
    if (nreads > UINT32_MAX || existing > UINT32_MAX) return fail(MPOA_ERR_INVALID_ARGUMENT, "too many reads");
    if (extend && extend->g.alpha() != s->params.alpha())
        return fail(MPOA_ERR_INVALID_ARGUMENT, "graph alphabet differs from params");
    if (!extend && existing != 0) return fail(MPOA_ERR_INVALID_ARGUMENT, "existing reads need a graph to extend");
    JobIn ji;
    ji.nreads = uint32_t(nreads);
    ji.first = uint32_t(existing);
    const uint64_t b0 = nreads ? read_off[0] : 0;
    ji.own_off.resize(nreads + 1);
    for (uint64_t k = 0; k <= nreads; k++) ji.own_off[k] = nreads ? read_off[k] - b0 : 0;
    if (nreads) ji.own_bases.assign(bases + b0, bases + read_off[nreads]);
    if (extend) ji.graph = std::make_unique<Graph>(extend->g);
    const uint64_t L = nreads ? ji.own_off[1] : 0;
    ji.work = L * nreads * nreads;
    ji.guide = s->params.progressive && existing == 0;
    *ticket = s->submit(std::move(ji));
    return MPOA_OK;
} catch (...) { return exception_status(); }

int mpoa_server_done(mpoa_server *s, uint64_t ticket) try {
    if (!s || ticket >= s->submitted()) return fail(MPOA_ERR_INVALID_ARGUMENT, "unknown ticket");
    return s->done(ticket) ? 1 : 0;
} catch (...) { return exception_status(); }

int mpoa_server_wait(mpoa_server *s, uint64_t ticket) try {
    if (!s || ticket >= s->submitted()) return fail(MPOA_ERR_INVALID_ARGUMENT, "unknown ticket");
    if (!s->wait(ticket)) return fail(s->error(), "the engine stopped: " + std::string(last_error()));
    return MPOA_OK;
} catch (...) { return exception_status(); }

int mpoa_server_job_status(mpoa_server *s, uint64_t ticket) try {
    if (!s || ticket >= s->submitted()) return fail(MPOA_ERR_INVALID_ARGUMENT, "unknown ticket");
    if (!s->done(ticket)) return fail(MPOA_ERR_PENDING, "the job has not finished");
    const JobOut &o = s->out(ticket);
    if (o.status != MPOA_OK) set_error(o.err);
    return o.status;
} catch (...) { return exception_status(); }

int mpoa_server_job_output(mpoa_server *s, uint64_t ticket, int mode, char *out, uint64_t cap,
                           uint64_t *out_len) try {
    const int st = mpoa_server_job_status(s, ticket);
    if (st != MPOA_OK) return st;
    JobOut &o = s->out(ticket);
    if (mode == 0) return emit_text(o.out0, out, cap, out_len);
    if (mode == 1) return emit_text(o.out1, out, cap, out_len);
    if (mode == 2) {
        std::string text;
        std::string e = o.graph->gfa(s->params, read_names(o.nseq), text);
        if (!e.empty()) return fail(MPOA_ERR_GRAPH, e);
        return emit_text(text, out, cap, out_len);
    }
    return fail(MPOA_ERR_INVALID_ARGUMENT, "mode must be 0, 1 or 2");
} catch (...) { return exception_status(); }

int mpoa_server_job_times(mpoa_server *s, uint64_t ticket, mpoa_job_times *out) try {
    if (!out) return fail(MPOA_ERR_INVALID_ARGUMENT, "NULL argument");
    if (!s || ticket >= s->submitted()) return fail(MPOA_ERR_INVALID_ARGUMENT, "unknown ticket");
    if (!s->done(ticket)) return fail(MPOA_ERR_PENDING, "the job has not finished");
    const JobOut &o = s->out(ticket);
    out->submit_ms = o.t_submit;
    out->admit_ms = o.t_admit;
    out->done_ms = o.t_done;
    out->cells = o.cells;
    return MPOA_OK;
} catch (...) { return exception_status(); }

int mpoa_server_job_graph(mpoa_server *s, uint64_t ticket, mpoa_graph **out) try {
    if (!out) return fail(MPOA_ERR_INVALID_ARGUMENT, "NULL argument");
    *out = nullptr;
    const int st = mpoa_server_job_status(s, ticket);
    if (st != MPOA_OK) return st;
    auto *g = new mpoa_graph(s->params);
    g->g = *s->out(ticket).graph;
    *out = g;
    return MPOA_OK;
} catch (...) { return exception_status(); }

int mpoa_server_stats_get(mpoa_server *s, mpoa_batch_stats *out) try {
    if (!s || !out) return fail(MPOA_ERR_INVALID_ARGUMENT, "NULL argument");
    *out = mpoa_batch_stats{};
    if (s->res) {
        ResidentStats rs;
        fill_stats(s->res->stats(&rs), *out);
        out->resident_jobs = rs.resident_done;
        out->evicted_jobs = rs.evicted;
        out->step_replays = rs.replays;
        out->buckets = rs.buckets;
        out->deferred = rs.deferred;
    } else {
        fill_stats(s->engine->stats(), *out);
    }
    out->wall_ms = s->now_ms();
    return MPOA_OK;
} catch (...) { return exception_status(); }

double mpoa_server_now_ms(mpoa_server *s) { return s ? s->now_ms() : 0.0; }

void mpoa_server_free(mpoa_server *s) {
    if (!s) return;
    mpoa_ctx *ctx = s->ctx;
    s->res.reset();     // finishes every submitted job first
    s->engine.reset();
    delete s;
    release_engine(ctx);
}

}  // extern "C"

// ---- minimizer seeding (host) ----

namespace {

template <class T, class F>
int emit_list(const std::vector<T> &v, uint64_t cap, uint64_t *len, F &&put) {
    if (len) *len = v.size();
    if (v.size() > cap) return fail(MPOA_ERR_BUFFER_TOO_SMALL, "output buffer is too small");
    for (size_t i = 0; i < v.size(); i++) put(i, v[i]);
    return MPOA_OK;
}

void anchors_in(const mpoa_anchor *a, uint64_t n, std::vector<Anchor> &v) {
    v.resize(n);
    for (uint64_t i = 0; i < n; i++) v[i] = Anchor{a[i].target_end, a[i].query_end, a[i].span, a[i].adaptive_before != 0};
}

int anchors_out(const std::vector<Anchor> &v, mpoa_anchor *out, uint64_t cap, uint64_t *len) {
    return emit_list(v, cap, len, [&](size_t i, const Anchor &a) {
        out[i] = mpoa_anchor{a.target_end, a.query_end, a.span, a.adaptive_before ? 1 : 0};
    });
}

}  // namespace

extern "C" {

int mpoa_sketch(const uint8_t *seq, uint64_t len, uint64_t k, uint64_t w, uint64_t *value,
                uint64_t *end, uint64_t cap, uint64_t *out_len) try {
    if ((!seq && len) || (cap && (!value || !end))) return fail(MPOA_ERR_INVALID_ARGUMENT, "NULL argument");
    std::vector<Minimizer> mm;
    sketch(seq, len, k, w, mm);
    return emit_list(mm, cap, out_len, [&](size_t i, const Minimizer &m) {
        value[i] = m.value;
        end[i] = m.end;
    });
} catch (...) { return exception_status(); }

int mpoa_collect_anchors(const mpoa_params *params, const uint8_t *query, uint64_t qlen,
                         const uint8_t *target, uint64_t tlen, mpoa_anchor *out, uint64_t cap,
                         uint64_t *out_len) try {
    if ((!query && qlen) || (!target && tlen) || (cap && !out)) return fail(MPOA_ERR_INVALID_ARGUMENT, "NULL argument");
    Params p;
    int rc = MPOA_OK;
    if (!check_params(params, p, rc)) return rc;
    std::vector<Anchor> a;
    collect_anchors(p, query, qlen, target, tlen, a);
    return anchors_out(a, out, cap, out_len);
} catch (...) { return exception_status(); }

int mpoa_chain(const mpoa_anchor *anchors, uint64_t n, uint64_t k, mpoa_anchor *out, uint64_t cap,
               uint64_t *out_len) try {
    if ((!anchors && n) || (cap && !out)) return fail(MPOA_ERR_INVALID_ARGUMENT, "NULL argument");
    std::vector<Anchor> in, c;
    anchors_in(anchors, n, in);
    chain(in, k, c);
    return anchors_out(c, out, cap, out_len);
} catch (...) { return exception_status(); }

int mpoa_partition_windows(const mpoa_anchor *chain_in, uint64_t n, uint64_t tlen, uint64_t qlen,
                           uint64_t max_drift, uint64_t max_win, mpoa_anchor *out, uint64_t cap,
                           uint64_t *out_len, int *tail_adaptive) try {
    if ((!chain_in && n) || (cap && !out)) return fail(MPOA_ERR_INVALID_ARGUMENT, "NULL argument");
    std::vector<Anchor> in, w;
    anchors_in(chain_in, n, in);
    const bool tail = partition_windows(in, tlen, qlen, max_drift, max_win, w);
    if (tail_adaptive) *tail_adaptive = tail ? 1 : 0;
    return anchors_out(w, out, cap, out_len);
} catch (...) { return exception_status(); }

int mpoa_guide_order(const mpoa_params *params, const uint8_t *const *seqs, const uint64_t *lens,
                     uint64_t nseq, uint64_t *order) try {
    if (nseq && (!seqs || !lens || !order)) return fail(MPOA_ERR_INVALID_ARGUMENT, "NULL argument");
    if (nseq > UINT32_MAX) return fail(MPOA_ERR_INVALID_ARGUMENT, "too many sequences");
    Params p;
    int rc = MPOA_OK;
    if (!check_params(params, p, rc)) return rc;
    std::vector<uint32_t> o;
    guide_order(p, seqs, lens, nseq, o);
    for (uint64_t i = 0; i < nseq; i++) order[i] = o[i];
    return MPOA_OK;
} catch (...) { return exception_status(); }

int mpoa_anchor_windows(const mpoa_params *params, const uint8_t *query, uint64_t qlen,
                        const uint8_t *consensus, uint64_t clen, mpoa_anchor *out, uint64_t cap,
                        uint64_t *out_len, int *tail_adaptive) try {
    if ((!query && qlen) || (!consensus && clen) || (cap && !out))
        return fail(MPOA_ERR_INVALID_ARGUMENT, "NULL argument");
    Params p;
    int rc = MPOA_OK;
    if (!check_params(params, p, rc)) return rc;
    std::vector<Anchor> w;
    const bool tail = anchor_windows(p, query, qlen, consensus, clen, w);
    if (tail_adaptive) *tail_adaptive = tail ? 1 : 0;
    return anchors_out(w, out, cap, out_len);
} catch (...) { return exception_status(); }

}  // extern "C"
