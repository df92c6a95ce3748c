// Device context, staging and launch orchestration for the DP kernels.
#include "runtime.hpp"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>

namespace mpoa {

static thread_local std::string t_error;
void set_error(const std::string &msg) { t_error = msg; }
const char *last_error() { return t_error.c_str(); }

// ---------------- thread pool ----------------

ThreadPool::ThreadPool(unsigned n) {
    for (unsigned k = 1; k < n; k++) workers_.emplace_back([this] { loop(); });
}

ThreadPool::~ThreadPool() {
    {
        std::lock_guard<std::mutex> lk(mu_);
        stop_ = true;
    }
    cv_.notify_all();
    for (auto &t : workers_) t.join();
}

void ThreadPool::loop() {
    size_t seen = 0;
    for (;;) {
        const std::function<void(size_t)> *fn;
        size_t n;
        {
            std::unique_lock<std::mutex> lk(mu_);
            cv_.wait(lk, [&] { return stop_ || generation_ != seen; });
            if (stop_) return;
            seen = generation_;
            fn = fn_;
            n = n_;
        }
        for (size_t i; (i = next_.fetch_add(1)) < n;) (*fn)(i);
        {
            std::lock_guard<std::mutex> lk(mu_);
            if (--active_ == 0) done_cv_.notify_all();
        }
    }
}

void ThreadPool::parallel_for(size_t n, const std::function<void(size_t)> &f) {
    if (n == 0) return;
    if (workers_.empty() || n == 1) {
        for (size_t i = 0; i < n; i++) f(i);
        return;
    }
    {
        std::lock_guard<std::mutex> lk(mu_);
        fn_ = &f;
        n_ = n;
        next_.store(0);
        active_ = unsigned(workers_.size());
        generation_++;
    }
    cv_.notify_all();
    for (size_t i; (i = next_.fetch_add(1)) < n;) f(i);
    std::unique_lock<std::mutex> lk(mu_);
    done_cv_.wait(lk, [&] { return active_ == 0; });
}

unsigned default_threads() {
    if (const char *s = getenv("MPOA_THREADS")) {
        int v = atoi(s);
        if (v > 0) return unsigned(v);
    }
    unsigned hw = std::thread::hardware_concurrency();
    if (hw == 0) hw = 1;
    return std::min(hw, 16u);
}

// ---------------- buffers ----------------

hipError_t DevBuf::ensure(size_t bytes) {
    if (bytes <= cap) return hipSuccess;
    // Grow with 50% headroom: hipFree synchronises the whole device, stalling
    // the other streams, so reallocations should be rare. If the headroom does
    // not fit, take the exact size.
    const size_t old = cap;
    release();
    size_t want = std::max(bytes, old) / 2 * 3;
    want = (want + 4095) & ~size_t(4095);
    hipError_t e = hipMalloc(&p, want);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        p = nullptr;
        cap = 0;
        e = hipMalloc(&p, bytes);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            p = nullptr;
            return e;
        }
        cap = bytes;
        return hipSuccess;
    }
    cap = want;
    return hipSuccess;
}

hipError_t DevBuf::ensure_exact(size_t bytes) {
    if (bytes <= cap) return hipSuccess;
    release();
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        p = nullptr;
        return e;
    }
    cap = bytes;
    return hipSuccess;
}

void DevBuf::release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
}

hipError_t PinBuf::ensure(size_t bytes) {
    if (bytes <= cap) return hipSuccess;
    const size_t old = cap;
    release();
    size_t want = std::max(bytes, old) / 2 * 3;
    want = (want + 4095) & ~size_t(4095);
    hipError_t e = hipHostMalloc(&p, want, hipHostMallocDefault);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        p = nullptr;
        want = bytes;
        e = hipHostMalloc(&p, want, hipHostMallocDefault);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            p = nullptr;
            return e;
        }
    }
    cap = want;
    return hipSuccess;
}

void PinBuf::release() {
    if (p) (void)hipHostFree(p);
    p = nullptr;
    cap = 0;
}

Ctx::~Ctx() {
    (void)hipSetDevice(device);
    if (stream) (void)hipStreamSynchronize(stream);
    d_in.release();
    d_res.release();
    d_meta.release();
    d_adapt.release();
    d_tb.release();
    d_arena.release();
    h_in.release();
    h_res.release();
    h_tb.release();
    for (auto &e : ev)
        if (e) (void)hipEventDestroy(e);
    if (stream) (void)hipStreamDestroy(stream);
    delete pool;
}

#define HIP_TRY(expr)                                                                 \
    do {                                                                              \
        hipError_t e_ = (expr);                                                       \
        if (e_ != hipSuccess) {                                                       \
            set_error(std::string("HIP error: ") + hipGetErrorString(e_) + " at " #expr); \
            return MPOA_ERR_DEVICE;                                                   \
        }                                                                             \
    } while (0)

int ctx_create(int device, Ctx **out, unsigned threads) {
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
        (void)hipGetLastError();
        set_error("no HIP device is available");
        return MPOA_ERR_NO_DEVICE;
    }
    if (device < 0 || device >= count) {
        set_error("device index out of range");
        return MPOA_ERR_NO_DEVICE;
    }
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, device));
    if (strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
        set_error(std::string("device is ") + prop.gcnArchName + ", this library is built for gfx950");
        return MPOA_ERR_NO_DEVICE;
    }
    Ctx *c = new Ctx();
    c->device = device;
    hipError_t e = hipSetDevice(device);
    // The no-allocation entry point's scratch address on this device (looked
    // up once per device, at the first context).
    if (e == hipSuccess) e = scratch_init(device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    for (auto &ev : c->ev)
        if (e == hipSuccess) e = hipEventCreate(&ev);
    if (e != hipSuccess) {
        set_error(std::string("HIP error: ") + hipGetErrorString(e));
        delete c;
        return MPOA_ERR_DEVICE;
    }
    c->pool = new ThreadPool(threads ? threads : default_threads());
    *out = c;
    return MPOA_OK;
}

int device_status(const DevResult &r, std::string &msg) {
    switch (r.status) {
        case kOk: msg.clear(); return MPOA_OK;
        case kBandNoIntersect: msg = "alignment band does not intersect graph path"; return MPOA_ERR_BAND_NO_INTERSECT;
        case kBandExcludesSource: msg = "alignment band excludes source"; return MPOA_ERR_BAND_EXCLUDES_SOURCE;
        case kTbLeftBand: msg = "alignment traceback left active band"; return MPOA_ERR_TRACEBACK_LEFT_BAND;
        case kTbNoAdvance: msg = "alignment traceback failed to advance"; return MPOA_ERR_TRACEBACK_NO_ADVANCE;
        case kTbNoPredecessor: {
            char buf[128];
            snprintf(buf, sizeof buf,
                     "alignment traceback has no predecessor at graph row %u, query column %u",
                     r.err_i, r.err_j);
            msg = buf;
            return MPOA_ERR_TRACEBACK_NO_PREDECESSOR;
        }
        case kArenaOverflow: msg = "alignment arena overflow"; return MPOA_ERR_ALLOC;
        case kTooManyPreds:
            msg = "a graph node has more predecessors than a device traceback record can name";
            return MPOA_ERR_UNSUPPORTED;
        case kLostScores: msg = "device DP lost a predecessor row's scores"; return MPOA_ERR_DEVICE;
        default: msg = "device DP returned an unknown status"; return MPOA_ERR_DEVICE;
    }
}

static inline size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

namespace {

// Launch-shape table measured on the device by minipoa_hip/tune.py and named
// by MPOA_TUNING: one line per row-width bucket, "max_width cols waves
// ring_kb", in increasing max_width. The first bucket that covers a launch's
// expected row width decides its i32 fill shape and LDS ring; wider rows, i64
// launches and the global-table kernel keep the built-in choice. The file is
// read again when the variable names another file.
struct TuneEntry {
    uint32_t max_width;
    FillShape shape;
    uint32_t ring_bytes;
};
std::mutex g_tune_mu;
std::string g_tune_path;
std::vector<TuneEntry> g_tune;

bool tuned_for(uint32_t width, TuneEntry &out) {
    const char *e = getenv("MPOA_TUNING");
    std::lock_guard<std::mutex> lk(g_tune_mu);
    const std::string path = e ? e : "";
    if (path != g_tune_path) {
        g_tune_path = path;
        g_tune.clear();
        if (FILE *f = path.empty() ? nullptr : fopen(path.c_str(), "r")) {
            unsigned w, c, v, r;
            char line[256];
            while (fgets(line, sizeof line, f)) {
                if (sscanf(line, "%u %u %u %u", &w, &c, &v, &r) != 4) continue;
                const FillShape sh{int(c), int(v)};
                if (!fill_shape_ok(sh) || r < 1 || r > 48) continue;
                g_tune.push_back(TuneEntry{w, sh, r * 1024u});
            }
            fclose(f);
        }
    }
    for (const TuneEntry &t : g_tune)
        if (width <= t.max_width) {
            out = t;
            return true;
        }
    return false;
}

}  // namespace

bool tuned_shape(uint32_t width, FillShape &shape, uint32_t &ring_bytes) {
    TuneEntry t{};
    if (!tuned_for(width, t)) return false;
    shape = t.shape;
    ring_bytes = t.ring_bytes;
    return true;
}

int launch_begin(Ctx &ctx, const Params &p, const std::vector<JobPlan *> &plans, bool i64,
                 bool full_arena, LaunchInfo &li, const FixedScratch *fx) {
    const size_t N = plans.size();
    const size_t tsz = i64 ? 8 : 4;
    li.n = N;
    li.i64 = i64;
    li.tb_at.assign(N, 0);
    struct Off {
        uint64_t rows, preds, seq, arena, cap, meta, adapt, tb;
    };
    // MPOA_ARENA_SQUEEZE (tests): private arena regions of 256 bytes and a
    // one-chunk pool, so that DPs overflow and take the full-arena relaunch.
    // Arena offsets and sizes are in bytes.
    const bool squeeze = getenv("MPOA_ARENA_SQUEEZE") != nullptr;
    // One layout per launch: records only if every DP of it was sized for
    // records (a row past kMaxRecPreds predecessors keeps a DP in values).
    bool rec = N > 0;
    for (size_t k = 0; k < N; k++) rec = rec && plans[k]->records;
    for (size_t k = 0; k < N; k++)
        if (!rec && plans[k]->records) use_layout(*plans[k], false);
    li.records = rec;
    std::vector<Off> off(N);
    uint64_t rows = 0, aux = 0, preds = 0, seq = 0, arena = 0, meta = 0, adapt = 0,
             tbn = 0;
    uint32_t maxw = 0, maxpad = 0;
    for (size_t k = 0; k < N; k++) {
        const JobPlan &J = *plans[k];
        Off &o = off[k];
        o.rows = rows;
        o.preds = preds;
        o.seq = seq;
        o.cap = align_up((full_arena || J.full_arena) ? arena_full(J)
                         : squeeze                  ? 256
                                                    : std::max<uint64_t>(J.arena_est, 256),
                         256);
        o.arena = arena;
        o.meta = meta;
        o.adapt = adapt;
        o.tb = tbn;
        rows += J.n + 1;
        aux += J.n + 1;
        preds += J.preds.size();
        seq += J.padded;
        arena += o.cap;
        meta += J.n;
        if (J.flags & kAdaptive) adapt += J.n;
        tbn += J.m;
        maxw = std::max(maxw, J.max_width);
        maxpad = std::max(maxpad, J.padded);
    }
    // Arena chunk pool after the private regions (rows that outgrow their
    // estimate take chunks of at least three full-width rows: M, D and I, or
    // 8-byte records and M and D).
    const uint64_t row_max = rec ? 8 + 2 * tsz : 3 * tsz;
    const uint64_t chunk = align_up(std::max<uint64_t>(262144, 3ull * maxpad * row_max), 256);
    const uint64_t pool_beg = arena;
    uint64_t pool_len = squeeze ? chunk : std::max<uint64_t>(arena / 2, 4 * chunk);
    const size_t alpha = p.alpha();
    size_t b_ctr = 0;  // the pool's next-chunk counter
    size_t b_jobs = 256;
    size_t b_rows = align_up(b_jobs + N * sizeof(DevJob), 256);
    size_t b_aux = align_up(b_rows + rows * sizeof(RowIn), 256);
    size_t b_preds = align_up(b_aux + aux * sizeof(RowAux), 256);
    size_t b_seq = align_up(b_preds + preds * sizeof(PredE), 256);
    size_t b_mat = align_up(b_seq + seq, 256);
    size_t b_end = align_up(b_mat + alpha * alpha * sizeof(int32_t), 256);

    // Fixed scratch: [inputs][results][row meta][adaptive records][traceback
    // records][arena], the pool taking whatever the scratch has left.
    const uint64_t s_res = align_up(b_end, 256);
    const uint64_t s_meta = align_up(s_res + N * sizeof(DevResult), 256);
    const uint64_t s_adapt = align_up(s_meta + std::max<uint64_t>(meta, 1) * sizeof(RowMeta), 256);
    const uint64_t s_tb = align_up(s_adapt + std::max<uint64_t>(adapt, 1) * sizeof(RowRec), 256);
    const uint64_t s_arena = align_up(s_tb + std::max<uint64_t>(tbn, 1) * sizeof(uint32_t), 256);
    if (fx) {
        const uint64_t need = s_arena + arena;
        if (need > fx->bytes) {
            char buf[200];
            snprintf(buf, sizeof buf,
                     "the DP needs %llu bytes of device buffers; the no-allocation scratch holds %llu",
                     (unsigned long long)need, (unsigned long long)fx->bytes);
            set_error(buf);
            return MPOA_ERR_SCRATCH_LIMIT;
        }
        if (!squeeze) pool_len = (fx->bytes - s_arena - arena) / 256 * 256;
    }
    arena += pool_len;

    HIP_TRY(hipSetDevice(ctx.device));
    li.grown = 0;
    li.arena_bytes = 0;
    int nbuf = 0;
    auto grow = [&li, &nbuf](auto &buf, size_t bytes) {
        const size_t cap = buf.cap;
        const hipError_t e = buf.ensure(bytes);
        if (buf.cap != cap) li.grown |= 1u << nbuf;
        nbuf++;
        return e;
    };
    // Host staging: the context's pinned buffers, or pageable memory of the
    // calling thread for a fixed-scratch launch.
    static thread_local std::vector<char> t_in, t_res, t_tb;
    char *h;
    if (fx) {
        if (t_in.size() < b_end) t_in.resize(b_end);
        h = t_in.data();
    } else {
        HIP_TRY(grow(ctx.h_in, b_end));
        h = static_cast<char *>(ctx.h_in.p);
    }
    ctx.pool->parallel_for(N, [&](size_t k) {
        const JobPlan &J = *plans[k];
        const Off &o = off[k];
        DevJob d{};
        d.rows = o.rows;
        d.aux = o.rows;
        d.preds = o.preds;
        d.seq = o.seq;
        d.matrix = 0;
        d.arena = o.arena;
        d.arena_cap = o.cap;
        d.rowmeta = o.meta;
        d.adapt = o.adapt;
        d.tb = o.tb;
        d.n = J.n;
        d.m = J.m;
        d.padded = J.padded;
        d.alpha = uint32_t(alpha);
        d.band = J.band;
        d.o = J.o;
        d.e = J.e;
        d.unknown = J.unknown;
        d.score_nn = J.score_nn;
        d.flags = J.flags;
        d.band32 = int32_t(std::min<int64_t>(J.band, kBandCap));
        d.cap5 = int32_t(std::min<int64_t>(5 * J.band, kBandCap));
        memcpy(h + b_jobs + k * sizeof(DevJob), &d, sizeof d);
        memcpy(h + b_rows + o.rows * sizeof(RowIn), J.rows.data(), J.rows.size() * sizeof(RowIn));
        memcpy(h + b_aux + o.rows * sizeof(RowAux), J.aux.data(), J.aux.size() * sizeof(RowAux));
        if (!J.preds.empty())
            memcpy(h + b_preds + o.preds * sizeof(PredE), J.preds.data(), J.preds.size() * sizeof(PredE));
        memcpy(h + b_seq + o.seq, J.seq.data(), J.seq.size());
    });
    memcpy(h + b_mat, p.matrix.data(), alpha * alpha * sizeof(int32_t));
    const unsigned long long ctr0 = pool_beg;
    memcpy(h + b_ctr, &ctr0, sizeof ctr0);

    char *d;
    void *d_res, *d_meta, *d_adapt, *d_tb, *d_arena;
    if (fx) {
        if (t_res.size() < N * sizeof(DevResult)) t_res.resize(N * sizeof(DevResult));
        if (t_tb.size() < std::max<uint64_t>(tbn, 1) * sizeof(uint32_t))
            t_tb.resize(std::max<uint64_t>(tbn, 1) * sizeof(uint32_t));
        li.hres = t_res.data();
        li.htb = t_tb.data();
        d = fx->dev;
        d_res = fx->dev + s_res;
        d_meta = fx->dev + s_meta;
        d_adapt = fx->dev + s_adapt;
        d_tb = fx->dev + s_tb;
        d_arena = fx->dev + s_arena;
    } else {
        HIP_TRY(grow(ctx.d_in, b_end));
        HIP_TRY(grow(ctx.d_res, N * sizeof(DevResult)));
        HIP_TRY(grow(ctx.d_meta, std::max<uint64_t>(meta, 1) * sizeof(RowMeta)));
        HIP_TRY(grow(ctx.d_adapt, std::max<uint64_t>(adapt, 1) * sizeof(RowRec)));
        HIP_TRY(grow(ctx.d_tb, std::max<uint64_t>(tbn, 1) * sizeof(uint32_t)));
        HIP_TRY(grow(ctx.d_arena, std::max<uint64_t>(arena, 256)));
        HIP_TRY(grow(ctx.h_res, N * sizeof(DevResult)));
        HIP_TRY(grow(ctx.h_tb, std::max<uint64_t>(tbn, 1) * sizeof(uint32_t)));
        li.hres = ctx.h_res.p;
        li.htb = ctx.h_tb.p;
        d = static_cast<char *>(ctx.d_in.p);
        d_res = ctx.d_res.p;
        d_meta = ctx.d_meta.p;
        d_adapt = ctx.d_adapt.p;
        d_tb = ctx.d_tb.p;
        d_arena = ctx.d_arena.p;
    }
    li.arena_bytes = std::max<uint64_t>(arena, 256);
    li.rows = meta;

    DevPools P;
    P.jobs = reinterpret_cast<const DevJob *>(d + b_jobs);
    P.res = static_cast<DevResult *>(d_res);
    P.rows = reinterpret_cast<const RowIn *>(d + b_rows);
    P.aux = reinterpret_cast<const RowAux *>(d + b_aux);
    P.preds = reinterpret_cast<const PredE *>(d + b_preds);
    P.seq = reinterpret_cast<const uint8_t *>(d + b_seq);
    P.matrix = reinterpret_cast<const int32_t *>(d + b_mat);
    P.arena = d_arena;
    P.meta = static_cast<RowMeta *>(d_meta);
    P.adapt = static_cast<RowRec *>(d_adapt);
    P.tb = static_cast<uint32_t *>(d_tb);
    P.pool_next = reinterpret_cast<unsigned long long *>(d + b_ctr);
    P.pool_end = pool_beg + pool_len;
    P.chunk = chunk;
    // Query codes up to 32 KB stage in LDS (the ring shrinks to fit below);
    // longer queries read the global-table kernel.
    P.seq_lds = maxpad <= 32768 ? maxpad : 0;
    P.mat_lds = alpha * alpha <= kMatLdsEntries ? uint32_t(alpha * alpha * 4) : 0;
    if (P.seq_lds == 0 || P.mat_lds == 0) P.seq_lds = P.mat_lds = 0;  // both tables in global memory

    // Lane width and ring size follow the expected row width: the widest
    // static row, or the mean adaptive width seen in earlier steps.
    uint32_t wexp = maxw;
    {
        uint64_t num = 0, den = 0;
        for (size_t k = 0; k < N; k++) {
            const JobPlan &J = *plans[k];
            if (J.width_hint) {
                num += J.width_hint;
                den += 1;
            }
        }
        if (den) wexp = std::max<uint32_t>(wexp, uint32_t(num / den));
    }
    FillShape shape = ctx.force_shape.cols ? ctx.force_shape : fill_shape_for_width(wexp);
    P.ring_bytes = ring_bytes_for_width(wexp, i64);
    // A launch of few DPs is bound by their row chains, not by how many
    // workgroups fit a CU. It gets the largest ring (clamped to the LDS left
    // below), so fewer predecessor rows come from HBM: on one seeded 100-read
    // job (~7 window DPs per launch) 48 KB instead of 16 KB cut the fill
    // time by 16%. Narrow rows also take two waves of four columns per lane
    // instead of one wave of eight (a shorter row chain): one 2 kb job of
    // 10 reads, one DP per launch, ran in 72 ms instead of 84.
    constexpr size_t kFewDps = 64;
    if (N <= kFewDps) {
        P.ring_bytes = std::max<uint32_t>(P.ring_bytes, 49152u);
        if (!ctx.force_shape.cols && !i64 && wexp <= 384u) shape = FillShape{4, 2};
    }
    // The tuning table is measured on full launches (calibration batches of
    // hundreds of jobs); launches of few DPs keep the rule above.
    TuneEntry tuned{};
    if (N > kFewDps && !ctx.force_shape.cols && !i64 && P.seq_lds != 0 && tuned_for(wexp, tuned)) {
        shape = tuned.shape;
        P.ring_bytes = tuned.ring_bytes;
    }
    if (const char *e = getenv("MPOA_FILL_SHAPE")) {
        FillShape f{0, 0};
        if (sscanf(e, "%d,%d", &f.cols, &f.waves) == 2 && fill_shape_ok(f)) shape = f;
    }
    // Rows with more than kMaxNarrowPreds predecessors take the wide-record
    // kernel, which like the global-table kernel has one shape.
    bool wide = false;
    for (size_t k = 0; k < N; k++) wide = wide || (rec && plans[k]->max_np > kMaxNarrowPreds);
    if (P.seq_lds == 0 || wide) shape = FillShape{8, 1};
    if (const char *e = getenv("MPOA_RING_KB")) P.ring_bytes = uint32_t(atoi(e)) * 1024u;
    {
        // Stay within 64 KB of LDS per workgroup with the fill's static arrays.
        const uint32_t static_lds = 10240, dyn_other = ((P.seq_lds + 15u) & ~15u) + P.mat_lds;
        const uint32_t room = 65536u - static_lds - dyn_other;
        if (P.ring_bytes > room) P.ring_bytes = room & ~63u;
    }
    HIP_TRY(hipEventRecord(ctx.ev[4], ctx.stream));
    HIP_TRY(hipMemcpyAsync(d, h, b_end, hipMemcpyHostToDevice, ctx.stream));
    HIP_TRY(hipEventRecord(ctx.ev[0], ctx.stream));
    P.prio = ctx.prio ? 1u : 0u;
    li.ring_bytes = P.ring_bytes;
    li.shape = shape;
    li.width = wexp;
    if (launch_fill(P, uint32_t(N), i64, rec, wide, shape, ctx.stream) != 0) {
        set_error("no fill kernel for the requested launch shape");
        return MPOA_ERR_DEVICE;
    }
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(ctx.ev[1], ctx.stream));
    launch_traceback(P, uint32_t(N), i64, rec, ctx.stream);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(ctx.ev[2], ctx.stream));
    HIP_TRY(hipMemcpyAsync(const_cast<void *>(li.hres), P.res, N * sizeof(DevResult),
                           hipMemcpyDeviceToHost, ctx.stream));
    HIP_TRY(hipMemcpyAsync(const_cast<void *>(li.htb), P.tb, tbn * sizeof(uint32_t),
                           hipMemcpyDeviceToHost, ctx.stream));
    HIP_TRY(hipEventRecord(ctx.ev[3], ctx.stream));
    for (size_t k = 0; k < N; k++) li.tb_at[k] = off[k].tb;
    return MPOA_OK;
}

bool launch_done(Ctx &ctx) {
    const hipError_t e = hipEventQuery(ctx.ev[3]);
    if (e == hipErrorNotReady) return false;
    if (e != hipSuccess) (void)hipGetLastError();  // launch_finish reports it
    return true;
}

uint64_t job_arena_bytes(const JobPlan &J, bool) {
    const uint64_t cap = align_up(J.full_arena ? arena_full(J) : std::max<uint64_t>(J.arena_est, 256), 256);
    return cap + cap / 2;
}

uint64_t device_free_bytes(int device) {
    size_t free = 0, total = 0;
    if (hipSetDevice(device) != hipSuccess || hipMemGetInfo(&free, &total) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return free;
}

int launch_finish(Ctx &ctx, LaunchInfo &li, RunStats *st) {
    HIP_TRY(hipSetDevice(ctx.device));
    HIP_TRY(hipEventSynchronize(ctx.ev[3]));
    float fill_ms = 0, all_ms = 0;
    HIP_TRY(hipEventElapsedTime(&fill_ms, ctx.ev[0], ctx.ev[1]));
    HIP_TRY(hipEventElapsedTime(&all_ms, ctx.ev[0], ctx.ev[2]));
    HIP_TRY(hipEventElapsedTime(&li.h2d_ms, ctx.ev[4], ctx.ev[0]));
    HIP_TRY(hipEventElapsedTime(&li.d2h_ms, ctx.ev[2], ctx.ev[3]));
    li.fill_ms = fill_ms;
    li.all_ms = all_ms;
    const DevResult *hr = launch_results(li);
    if (st) {
        // Overflowed attempts are re-run with the full arena; count the cells
        // of the run that completes.
        st->fill_ms += fill_ms;
        st->device_ms += all_ms;
        st->fill_launches += 1;
        if (li.records) st->record_launches += 1;
        for (size_t k = 0; k < li.n; k++) {
            if (hr[k].status == kArenaOverflow) {
                st->arena_retries += 1;
                continue;
            }
            st->cells += hr[k].cells;
            st->pred_cells += hr[k].pred_cells;
            st->arena_bytes += hr[k].arena_used;
            st->dp_calls += 1;
            if (li.i64) st->i64_calls += 1;
        }
    }
    return MPOA_OK;
}

const DevResult *launch_results(const LaunchInfo &li) { return static_cast<const DevResult *>(li.hres); }
const uint32_t *launch_tb(const Ctx &, const LaunchInfo &li, size_t k) {
    return static_cast<const uint32_t *>(li.htb) + li.tb_at[k];
}

// One synchronous launch over plans[idx[*]] that share the score width.
static int run_group(Ctx &ctx, const Params &p, std::vector<JobPlan *> &plans,
                     const std::vector<size_t> &idx, bool i64, bool full_arena,
                     std::vector<DevResult> &res, std::vector<uint32_t> &tb,
                     const std::vector<uint64_t> &tb_off, RunStats *st, const FixedScratch *fx) {
    std::vector<JobPlan *> group(idx.size());
    for (size_t k = 0; k < idx.size(); k++) group[k] = plans[idx[k]];
    LaunchInfo li;
    int rc = launch_begin(ctx, p, group, i64, full_arena, li, fx);
    if (rc != MPOA_OK) return rc;
    rc = launch_finish(ctx, li, st);
    if (rc != MPOA_OK) return rc;
    const DevResult *hr = launch_results(li);
    for (size_t k = 0; k < idx.size(); k++) res[idx[k]] = hr[k];
    ctx.pool->parallel_for(idx.size(), [&](size_t k) {
        if (hr[k].status == kOk)
            memcpy(tb.data() + tb_off[idx[k]], launch_tb(ctx, li, k),
                   size_t(group[k]->m) * sizeof(uint32_t));
    });
    return MPOA_OK;
}

int run_plans(Ctx &ctx, const Params &p, std::vector<JobPlan *> &plans,
              std::vector<DevResult> &res, std::vector<uint32_t> &tb,
              std::vector<uint64_t> &tb_off, RunStats *st, const FixedScratch *fx) {
    const size_t N = plans.size();
    res.assign(N, DevResult{});
    tb_off.assign(N + 1, 0);
    for (size_t k = 0; k < N; k++)
        tb_off[k + 1] = tb_off[k] + (plans[k]->status == MPOA_OK ? plans[k]->m : 0);
    tb.resize(tb_off[N]);
    for (int pass = 0; pass < 2; pass++) {
        for (int wide = 0; wide < 2; wide++) {
            std::vector<size_t> idx;
            for (size_t k = 0; k < N; k++) {
                if (plans[k]->status != MPOA_OK || plans[k]->use_i64 != (wide == 1)) continue;
                if (pass == 1 && res[k].status != kArenaOverflow) continue;
                idx.push_back(k);
            }
            if (idx.empty()) continue;
            // Largest jobs first: workgroups dispatch in order, and a step
            // lasts as long as its longest row chain.
            std::stable_sort(idx.begin(), idx.end(), [&](size_t x, size_t y) {
                return plans[x]->static_cells > plans[y]->static_cells;
            });
            int rc = run_group(ctx, p, plans, idx, wide == 1, pass == 1, res, tb, tb_off, st, fx);
            if (rc != MPOA_OK) return rc;
        }
    }
    return MPOA_OK;
}

}  // namespace mpoa
