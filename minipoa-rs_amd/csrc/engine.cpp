// Continuous batched POA engine (engine.hpp).
#include "engine.hpp"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <exception>
#include <new>

#include "minipoa_hip.h"
#include "plan.hpp"
#include "seed.hpp"

namespace mpoa {

TaskQueue::TaskQueue(unsigned n) {
    for (unsigned k = 0; k < std::max(1u, n); k++) workers_.emplace_back([this] { loop(); });
}

TaskQueue::~TaskQueue() {
    {
        std::lock_guard<std::mutex> lk(mu_);
        stop_ = true;
    }
    cv_.notify_all();
    for (auto &t : workers_) t.join();
}

void TaskQueue::push(std::function<void()> f, bool urgent) {
    {
        std::lock_guard<std::mutex> lk(mu_);
        (urgent ? urgent_ : normal_).q.push_back(std::move(f));
    }
    cv_.notify_one();
}

void TaskQueue::loop() {
    for (;;) {
        std::function<void()> f;
        {
            std::unique_lock<std::mutex> lk(mu_);
            cv_.wait(lk, [&] { return stop_ || !urgent_.empty() || !normal_.empty(); });
            Fifo &src = !urgent_.empty() ? urgent_ : normal_;
            if (src.empty()) return;  // stopping and drained
            f = std::move(src.q[src.head++]);
            if (src.empty()) {
                src.q.clear();
                src.head = 0;
            }
        }
        f();
    }
}

void par_for(TaskQueue &tq, size_t n, const std::function<void(size_t)> &f) {
    if (n == 0) return;
    struct State {
        std::atomic<size_t> next{0}, done{0};
        size_t n = 0;
        const std::function<void(size_t)> *f = nullptr;
        std::mutex mu;
        std::condition_variable cv;
    };
    auto st = std::make_shared<State>();
    st->n = n;
    st->f = &f;
    auto run = [st] {
        for (;;) {
            const size_t i = st->next++;
            if (i >= st->n) return;
            (*st->f)(i);  // f outlives every claimed index: the caller waits for them
            if (++st->done == st->n) {
                std::lock_guard<std::mutex> lk(st->mu);
                st->cv.notify_all();
            }
        }
    };
    const size_t helpers = std::min<size_t>(tq.size() > 1 ? tq.size() - 1 : 0, n - 1);
    for (size_t k = 0; k < helpers; k++) tq.push(run, true);
    run();
    std::unique_lock<std::mutex> lk(st->mu);
    st->cv.wait(lk, [&] { return st->done.load() == st->n; });
}

// One DP of a job's step: the whole-graph DP, or one seeded window.
struct Engine::Dp {
    Job *job = nullptr;
    uint32_t idx = 0;        // window index within the step
    JobPlan plan;
    DevResult res{};
    std::vector<uint32_t> tb;
    bool retry = false;      // the planned DP is poa's full-matrix retry
    bool fresh = false;      // res (or a planning error) awaits consumption
    int status = MPOA_OK;    // final outcome once consumed
    std::string err;
    uint64_t cost = 0;       // expected cells of the planned DP
    uint64_t rem = ~0ull;    // the job's cells left when last queued
    std::vector<AlignedBase> path;
};

// One job's progressive state.
struct Engine::Job {
    uint64_t id = 0;
    JobIn in;
    JobOut out;
    uint32_t step = 0;       // reads of `in` aligned so far
    std::vector<std::unique_ptr<Dp>> dp;  // the step's DPs are dp[0 .. ndp)
    uint32_t ndp = 0;
    std::vector<SeedWindow> win;          // the step's seeded windows
    bool seeded_ahead = false;            // win (or seed_rc) already holds the next step's windows
    int seed_rc = MPOA_OK;
    std::string seed_err;
    uint32_t pending = 0;    // DPs of the step on the device (under the engine lock)
    uint64_t prev_used = 0, prev_cells = 0, prev_n = 0;
    uint64_t cost = 0;       // expected cells of the step
    uint64_t rem = ~0ull;    // cells left when last queued (0 once finished; under the engine lock)
    uint64_t footprint = 0;  // bytes counted in active_bytes_ (under the engine lock)
    bool admitted = false;
    std::vector<AlignedBase> path;
    std::vector<uint32_t> order;  // -p: read aligned at each step (empty: input order)

    // The read of step s: its index in `in` and its bases.
    uint32_t read_index(uint32_t s) const { return order.empty() ? s : order[s]; }
    const uint8_t *read(uint32_t s, uint64_t &len) const {
        const uint32_t r = read_index(s);
        len = in.read_off[r + 1] - in.read_off[r];
        return in.bases + in.read_off[r];
    }
};

// One line per launch for MPOA_TRACE=<file>: scheduling diagnostics.
struct Engine::TraceRec {
    int slot;
    double t_begin, t_launched, t_done;  // host clock: before packing, after enqueueing, results on host
    float fill_ms, all_ms, h2d_ms, d2h_ms;  // device clock
    uint32_t grown;
    uint64_t arena_bytes, arena_cap;
    size_t njobs;
    uint64_t max_cost, sum_cost;
    int64_t watched_step;  // step of the job with the most work if it is in this launch, else -1
    uint32_t ring_bytes, width;
    FillShape shape;
    uint64_t rows;
};

namespace {

uint64_t env_u64(const char *name, uint64_t dflt, uint64_t lo, uint64_t hi) {
    const char *s = getenv(name);
    if (!s) return dflt;
    const long long v = atoll(s);
    if (v < (long long)lo) return lo;
    if (v > (long long)hi) return hi;
    return uint64_t(v);
}

}  // namespace

Engine::Engine(std::vector<Ctx *> slots, TaskQueue &tasks, const Params &p, const EngineOpts &o)
    : slots_(std::move(slots)), tasks_(tasks), p_(p), full_(p), o_(o), t0_(std::chrono::steady_clock::now()) {
    full_.f = 0;
    const size_t nslots = slots_.size();
    o_.crit_lane = o_.crit_lane && nslots >= 2;
    // The priority slot's waves issue ahead of the other slots' waves on
    // shared SIMDs: its jobs are the makespan's critical chain.
    const bool crit_prio = o_.crit_lane && env_u64("MPOA_CRIT_PRIO", 1, 0, 1) != 0;
    for (size_t si = 0; si < nslots; si++) slots_[si]->prio = crit_prio && si == 0;
    crit_wait_ms_ = double(env_u64("MPOA_CRIT_WAIT_US", 50000, 0, 10000000)) / 1000.0;
    // Once the largest job is ready, the priority slot waits this long for
    // the other critical jobs' host steps before it launches.
    crit_gather_ms_ = double(env_u64("MPOA_CRIT_GATHER_US", 3000, 0, 10000000)) / 1000.0;
    max_launch_ = env_u64("MPOA_MAX_LAUNCH", 1024, 1, 1u << 20);
    // Each slot's score arena is bounded so that every slot's buffers (which
    // grow with 50% headroom) fit the device together; a launch takes jobs
    // until its arena reaches the bound (at least one job).
    arena_budget_ = env_u64("MPOA_SLOT_ARENA_MB", 0, 0, 1ull << 30) << 20;
    if (!arena_budget_ && nslots) {
        const uint64_t free = device_free_bytes(slots_[0]->device);
        arena_budget_ = std::max<uint64_t>(free / 10 * 6 / nslots / 3 * 2, 256ull << 20);
    }
    admit_budget_ = o_.admit_bytes ? o_.admit_bytes : arena_budget_ * std::max<size_t>(nslots, 1);
    const size_t nord = o_.crit_lane ? nslots - 1 : nslots;
    if (const char *e = getenv("MPOA_LONG_ROWS")) o_.long_rows = uint32_t(atoll(e));
    // At least one slot for each queue; with a single ordinary slot there
    // is one queue.
    if (nord < 2) o_.long_rows = 0;
    if (o_.long_rows) o_.long_slots = uint32_t(std::min<size_t>(std::max<uint32_t>(o_.long_slots, 1), nord - 1));
    else o_.long_slots = 0;
    idle_[1] = o_.long_slots;
    idle_[0] = nord - o_.long_slots;
    sst_.resize(nslots);
    strace_.resize(nslots);
    trace_path_ = getenv("MPOA_TRACE");
    for (size_t si = 0; si < nslots; si++) threads_.emplace_back([this, si] { slot_main(si); });
}

Engine::~Engine() {
    wait_all();
    {
        std::lock_guard<std::mutex> lk(mu_);
        stop_ = true;
    }
    cv_.notify_all();
    for (auto &t : threads_) t.join();
    // Let every queued task finish before the state they reference goes away.
    {
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [&] { return outstanding_ == 0; });
    }
    if (trace_path_) {
        std::vector<TraceRec> trace;
        for (const auto &v : strace_) trace.insert(trace.end(), v.begin(), v.end());
        std::sort(trace.begin(), trace.end(), [](const TraceRec &x, const TraceRec &y) { return x.t_begin < y.t_begin; });
        if (FILE *f = fopen(trace_path_, "w")) {
            fprintf(f, "slot,t_begin_ms,t_launched_ms,t_done_ms,fill_ms,all_ms,h2d_ms,d2h_ms,grown,arena_bytes,arena_cap,njobs,max_cost,sum_cost,"
                       "watched_step,ring_bytes,width,cols,waves,rows\n");
            for (const TraceRec &t : trace)
                fprintf(f, "%d,%.3f,%.3f,%.3f,%.3f,%.3f,%.3f,%.3f,%u,%llu,%llu,%zu,%llu,%llu,%lld,%u,%u,%d,%d,%llu\n", t.slot, t.t_begin,
                        t.t_launched, t.t_done, t.fill_ms, t.all_ms, t.h2d_ms, t.d2h_ms, t.grown, (unsigned long long)t.arena_bytes,
                        (unsigned long long)t.arena_cap, t.njobs, (unsigned long long)t.max_cost, (unsigned long long)t.sum_cost,
                        (long long)t.watched_step, t.ring_bytes, t.width, t.shape.cols, t.shape.waves, (unsigned long long)t.rows);
            fclose(f);
        }
    }
}

double Engine::now_ms() const {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0_).count();
}

uint64_t Engine::submitted() {
    std::lock_guard<std::mutex> lk(mu_);
    return jobs_.size();
}

int Engine::error() {
    std::lock_guard<std::mutex> lk(mu_);
    return rc_;
}

bool Engine::done(uint64_t id) {
    std::lock_guard<std::mutex> lk(mu_);
    return id < jobs_.size() && jobs_[id]->out.done;
}

bool Engine::wait(uint64_t id) {
    std::unique_lock<std::mutex> lk(mu_);
    if (id >= jobs_.size()) return false;
    Job *J = jobs_[id].get();
    cv_.wait(lk, [&] { return J->out.done || rc_ != 0; });
    return J->out.done;
}

int Engine::wait_all() {
    std::unique_lock<std::mutex> lk(mu_);
    cv_.wait(lk, [&] { return finished_ == jobs_.size() || rc_ != 0; });
    return rc_;
}

JobOut &Engine::out(uint64_t id) {
    std::lock_guard<std::mutex> lk(mu_);
    return jobs_[id]->out;
}

EngineStats Engine::stats() {
    EngineStats st;
    std::lock_guard<std::mutex> lk(mu_);
    for (const RunStats &x : sst_) {
        st.run.cells += x.cells;
        st.run.dp_calls += x.dp_calls;
        st.run.i64_calls += x.i64_calls;
        st.run.pred_cells += x.pred_cells;
        st.run.fill_launches += x.fill_launches;
        st.run.device_ms += x.device_ms;
        st.run.fill_ms += x.fill_ms;
        st.run.arena_retries += x.arena_retries;
        st.run.arena_bytes += x.arena_bytes;
        st.run.record_launches += x.record_launches;
    }
    st.run.retries = retries_.load();
    const double nw = double(std::max(1u, tasks_.size()));
    st.host_ms = double(task_ns_.load()) / 1e6 / nw;
    st.plan_ms = double(plan_ns_.load()) / 1e6 / nw;
    st.consume_ms = double(consume_ns_.load()) / 1e6 / nw;
    st.fuse_ms = double(fuse_ns_.load()) / 1e6 / nw;
    st.render_ms = double(render_ns_.load()) / 1e6 / nw;
    st.pack_ms = double(pack_ns_.load()) / 1e6 / double(std::max<size_t>(1, slots_.size()));
    return st;
}

// Device bytes a job is expected to need before its first DP is planned:
// score rows (M, D, I) for a graph of about twice the first read's length,
// each row as wide as the band allows (the whole query when f == 0).
uint64_t Engine::footprint_estimate(const Job &J) const {
    if (J.in.nreads == 0) return 0;
    const uint64_t L = J.in.read_off[1] - J.in.read_off[0];
    uint64_t width = L + 16;
    if (p_.f != 0) width = std::min<uint64_t>(width, 2 * (p_.b + L / p_.f) + 16);
    uint64_t rows = 2 * L + 2;
    if (J.out.graph) rows = std::max<uint64_t>(rows, J.out.graph->size() * 3 / 2);
    return 12 * rows * width;
}

void Engine::admit_locked(Job &J) {
    J.admitted = true;
    J.out.t_admit = now_ms();
    J.footprint = footprint_estimate(J);
    active_bytes_ += J.footprint;
}

uint64_t Engine::submit(JobIn in) {
    auto up = std::make_unique<Job>();
    Job &J = *up;
    J.in = std::move(in);
    if (!J.in.own_bases.empty() || !J.in.own_off.empty()) {
        J.in.bases = J.in.own_bases.data();
        J.in.read_off = J.in.own_off.data();
    }
    J.out.graph = J.in.graph ? std::move(J.in.graph) : std::make_unique<Graph>(p_);
    J.out.nseq = J.in.first + J.in.nreads;
    bool start = false;
    {
        std::lock_guard<std::mutex> lk(mu_);
        J.id = jobs_.size();
        J.out.t_submit = now_ms();
        if (!o_.crit_lane) J.in.crit = false;
        if (J.in.crit) {
            crit_jobs_.push_back(&J);
            std::stable_sort(crit_jobs_.begin(), crit_jobs_.end(),
                             [](const Job *x, const Job *y) { return x->in.work > y->in.work; });
            active_crit_++;
            if (!top_job_ || J.in.work > top_job_->in.work) top_job_ = &J;
        }
        if (!watched_ || J.in.work > watched_->in.work) watched_ = &J;
        jobs_.push_back(std::move(up));
        // Admission by device footprint: first come, first served; a job
        // that does not fit waits for finished jobs to release theirs.
        const uint64_t est = footprint_estimate(J);
        if (waiting_.empty() && (active_bytes_ == 0 || active_bytes_ + est <= admit_budget_)) {
            admit_locked(J);
            start = true;
        } else {
            waiting_.push_back(&J);
        }
    }
    if (start) task(J, false);
    return J.id;
}

void Engine::finish(Job &J) {
    std::vector<Job *> admitted;
    {
        std::lock_guard<std::mutex> lk(mu_);
        finished_++;
        J.rem = 0;
        J.out.t_done = now_ms();
        J.out.done = true;
        active_bytes_ -= std::min(active_bytes_, J.footprint);
        J.footprint = 0;
        if (J.in.crit) {
            active_crit_--;
            if (&J == top_job_) {
                top_job_ = nullptr;  // crit_jobs_ is in decreasing order of work
                for (const Job *K : crit_jobs_)
                    if (!K->out.done) {
                        top_job_ = K;
                        break;
                    }
            }
        }
        while (!waiting_.empty()) {
            Job *W = waiting_.front();
            if (active_bytes_ != 0 && active_bytes_ + footprint_estimate(*W) > admit_budget_) break;
            waiting_.pop_front();
            admit_locked(*W);
            admitted.push_back(W);
        }
        cv_.notify_all();
    }
    for (Job *W : admitted) task(*W, false);
}

void Engine::requeue(Job &J, const std::vector<Dp *> &dps) {
    std::lock_guard<std::mutex> lk(mu_);
    J.pending = uint32_t(dps.size());
    J.rem = J.cost * uint64_t(J.in.nreads - J.step);
    uint64_t bytes = 0;
    for (const Dp *D : dps) bytes += job_arena_bytes(D->plan, D->plan.use_i64);
    const uint64_t fp = std::max(bytes, footprint_estimate(J));
    active_bytes_ = active_bytes_ - std::min(active_bytes_, J.footprint) + fp;
    J.footprint = fp;
    const bool crit = J.in.crit && (!crit_busy_ || &J == top_job_);
    if (crit && ready_crit_.empty()) crit_since_ = now_ms();
    if (crit && &J == top_job_) top_since_ = now_ms();
    for (Dp *D : dps) {
        D->rem = J.rem;
        if (crit) {
            ready_crit_.push_back(D);
        } else {
            // A critical job that missed the priority slot's launch takes an
            // ordinary slot for this step rather than holding the priority
            // slot when the largest job comes back.
            (o_.long_rows && D->plan.n >= o_.long_rows ? ready_long_ : ready_).push_back(D);
        }
    }
    cv_.notify_all();
}

void Engine::fail_job(Job &J, int status, const std::string &e) {
    J.out.status = status;
    J.out.err = e;
    finish(J);
}

// graph.rs:82-178 per read; false when the job ended with an error. With
// seeding, the next read's windows (src/align.rs:99-162: consensus, anchors,
// chain) are computed while the graph is re-sorted: seeding reads edges and
// bases, the sort writes only ranks.
bool Engine::fuse(Job &J) {
    Graph &g = *J.out.graph;
    std::string e = g.add_path(J.in.first + J.read_index(J.step), J.path.data(), J.path.size());
    J.seeded_ahead = false;
    const uint32_t next = J.step + 1;
    if (e.empty() && p_.seeding && next < J.in.nreads && g.size() > 2) {
        uint64_t qlen;
        const uint8_t *q = J.read(next, qlen);
        par_for(tasks_, 2, [&](size_t i) {
            if (i == 0) {
                e = g.toposort(false);
            } else {
                J.seed_err.clear();
                J.seed_rc = seed_windows(p_, GraphAcc{g}, q, qlen, J.win, J.seed_err);
            }
        });
        J.seeded_ahead = true;
    } else if (e.empty()) {
        e = g.toposort(false);
    }
    if (!e.empty()) {
        fail_job(J, MPOA_ERR_GRAPH, e);
        return false;
    }
    J.step++;
    return true;
}

void Engine::render(Job &J) {
    Graph &g = *J.out.graph;
    J.out.out0 = ">consensus sequence\n" + g.consensus_bytes(p_) + "\n";
    const uint32_t ns = J.out.nseq;
    std::vector<std::string> names(ns), comments(ns);
    for (uint32_t k = 0; k < ns; k++) names[k] = "read" + std::to_string(k);
    std::string e = g.msa(p_, names, comments, J.out.out1);
    if (!e.empty()) {
        J.out.status = MPOA_ERR_GRAPH;
        J.out.err = e;
    }
}

// Consumes the step's fresh DP results. A DP that overflowed its arena, or
// failed with f != 0 on its first attempt (poa's full-matrix retry,
// src/align.rs:202-207), goes into `again`; every other DP gets its final
// status and path. False when the job ended with an error.
bool Engine::consume(Job &J, std::vector<Dp *> &again) {
    Graph &g = *J.out.graph;
    uint64_t qlen;
    const uint8_t *q = J.read(J.step, qlen);
    again.clear();
    for (uint32_t k = 0; k < J.ndp; k++) {
        Dp &D = *J.dp[k];
        if (!D.fresh) continue;
        D.fresh = false;
        if (D.plan.status == MPOA_OK && D.res.status == kArenaOverflow) {
            D.plan.full_arena = true;  // same DP, arena for full rows
            again.push_back(&D);
            continue;
        }
        const bool failed = D.plan.status != MPOA_OK || D.res.status != kOk;
        if (failed && p_.f != 0 && !D.retry) {
            D.retry = true;
            retries_++;
            if (J.win.empty()) {
                plan_job(full_, GraphAcc{g}, 0, 1, q, qlen, false, p_.b, 0, D.plan);
                use_layout(D.plan, slots_[0]->layout == kLayoutRecords);
            } else {
                const SeedWindow &w = J.win[k];
                plan_job(full_, GraphAcc{g}, w.beg, w.end, q + w.qbeg, w.qend - w.qbeg, false, w.band, 0, D.plan);
                use_layout(D.plan, slots_[0]->layout == kLayoutRecords);
            }
            if (D.plan.status == MPOA_OK) {
                D.cost = D.plan.static_cells;
                again.push_back(&D);
            } else {
                D.status = D.plan.status;
                D.err = D.plan.err;
            }
            continue;
        }
        if (D.plan.status != MPOA_OK) {
            D.status = D.plan.status;
            D.err = D.plan.err;
            continue;
        }
        D.status = device_status(D.res, D.err);
        if (D.status != MPOA_OK) continue;
        if (!D.retry && J.win.empty()) {
            J.prev_used = D.res.arena_used;
            J.prev_cells = D.res.cells;
            J.prev_n = D.plan.n;
        }
        J.out.cells += D.res.cells;
        assemble(GraphAcc{g}, D.plan, D.tb.data(), D.path);
    }
    if (!again.empty()) return true;
    // Every DP of the step is final: the first failing window (in query
    // order) fails the job, as map_ordered returns the first error.
    for (uint32_t k = 0; k < J.ndp; k++) {
        const Dp &D = *J.dp[k];
        if (D.status != MPOA_OK) {
            fail_job(J, D.status, D.err);
            return false;
        }
    }
    if (J.win.empty()) {
        J.path.swap(J.dp[0]->path);
    } else {
        J.path.clear();
        for (uint32_t k = 0; k < J.ndp; k++) append_window(GraphAcc{g}, J.win[k], J.dp[k]->path, J.path);
    }
    return true;
}

// Plans the DPs of the job's next step (src/align.rs:96-188). DPs whose
// planning failed are marked fresh for consume(); the others go to `launch`.
// False when the job ended with an error.
bool Engine::plan_step(Job &J, std::vector<Dp *> &launch) {
    Graph &g = *J.out.graph;
    uint64_t qlen;
    const uint8_t *q = J.read(J.step, qlen);
    if (p_.seeding) {
        if (!J.seeded_ahead) {
            J.seed_err.clear();
            J.seed_rc = seed_windows(p_, GraphAcc{g}, q, qlen, J.win, J.seed_err);
        }
        J.seeded_ahead = false;
        if (J.seed_rc != MPOA_OK) {
            fail_job(J, J.seed_rc, J.seed_err);
            return false;
        }
    } else {
        J.win.clear();
    }
    J.ndp = uint32_t(std::max<size_t>(1, J.win.size()));
    while (J.dp.size() < J.ndp) {
        J.dp.push_back(std::make_unique<Dp>());
        J.dp.back()->job = &J;
        J.dp.back()->idx = uint32_t(J.dp.size() - 1);
    }
    launch.clear();
    J.cost = 0;
    auto plan_one = [&](size_t k) {
        Dp &D = *J.dp[k];
        D.retry = false;
        D.fresh = false;
        D.status = MPOA_OK;
        D.err.clear();
        if (J.win.empty()) {
            plan_job(p_, GraphAcc{g}, 0, 1, q, qlen, p_.adaptive, p_.b, 0, D.plan);
            use_layout(D.plan, slots_[0]->layout == kLayoutRecords);
        } else {
            const SeedWindow &w = J.win[k];
            plan_job(p_, GraphAcc{g}, w.beg, w.end, q + w.qbeg, w.qend - w.qbeg, w.adaptive, w.band, 0, D.plan);
            use_layout(D.plan, slots_[0]->layout == kLayoutRecords);
        }
    };
    // Windows plan independently (each writes its own plan only).
    if (J.ndp > 1) par_for(tasks_, J.ndp, plan_one);
    else plan_one(0);
    for (uint32_t k = 0; k < J.ndp; k++) {
        Dp &D = *J.dp[k];
        JobPlan &P = D.plan;
        if (P.status != MPOA_OK) {
            D.fresh = true;
            continue;
        }
        D.cost = P.static_cells;
        // Adaptive bands: start from what the previous step's DP used,
        // scaled to the grown graph.
        if (J.win.empty() && (P.flags & kAdaptive) && J.prev_n > 0) {
            uint64_t hint = J.prev_used * P.n / J.prev_n;
            // Slack: eight full-width rows (records with scores, or M/D/I).
            const uint64_t tsz = P.use_i64 ? 8 : 4;
            hint += hint / 8 + uint64_t(P.padded) * 8 * (P.records ? 2 + 2 * tsz : 3 * tsz);
            P.arena_est = std::min(std::max(P.arena_est, hint), arena_full(P));
            P.width_hint = uint32_t(J.prev_cells / J.prev_n);
            D.cost = uint64_t(P.width_hint) * P.n;
        }
        J.cost += D.cost;
        launch.push_back(&D);
    }
    return true;
}

// Consumes the job's DP results (have), fuses the step's path and plans the
// next step. src/lib.rs:195-202 (one read per step), src/align.rs:81-188.
void Engine::advance(Job &J, bool have) {
    Graph &g = *J.out.graph;
    std::vector<Dp *> again;
    if (J.in.guide && J.step == 0 && !have && J.order.empty() && J.in.nreads > 0) {
        // src/lib.rs:189-193: the guide order of all the job's reads.
        std::vector<const uint8_t *> seqs(J.in.nreads);
        std::vector<uint64_t> lens(J.in.nreads);
        for (uint32_t r = 0; r < J.in.nreads; r++) {
            seqs[r] = J.in.bases + J.in.read_off[r];
            lens[r] = J.in.read_off[r + 1] - J.in.read_off[r];
        }
        guide_order(p_, seqs.data(), lens.data(), J.in.nreads, J.order);
    }
    using clk = std::chrono::steady_clock;
    auto ns_since = [](clk::time_point t) {
        return std::chrono::duration_cast<std::chrono::nanoseconds>(clk::now() - t).count();
    };
    for (;;) {
        if (have) {
            have = false;
            auto t = clk::now();
            const bool ok = consume(J, again);
            consume_ns_ += ns_since(t);
            if (!ok) return;
            if (!again.empty()) {
                requeue(J, again);
                return;
            }
            t = clk::now();
            const bool fused = fuse(J);
            fuse_ns_ += ns_since(t);
            if (!fused) return;
        }
        if (J.step >= J.in.nreads) {
            if (o_.render) {
                const auto t = clk::now();
                render(J);
                render_ns_ += ns_since(t);
            }
            finish(J);
            return;
        }
        uint64_t qlen;
        const uint8_t *q = J.read(J.step, qlen);
        if (g.size() <= 2) {
            J.path.clear();
            J.path.push_back(AlignedBase{0, int64_t(p_.encode('N')), -1});
            for (uint64_t t = 0; t < qlen; t++) J.path.push_back(AlignedBase{-1, int64_t(p_.encode(q[t])), -1});
            if (!fuse(J)) return;
            continue;
        }
        const auto tp = clk::now();
        const bool planned = plan_step(J, again);
        plan_ns_ += ns_since(tp);
        if (!planned) return;
        if (again.empty()) {  // every DP failed at planning
            have = true;
            continue;
        }
        requeue(J, again);
        return;
    }
}

void Engine::task(Job &J, bool have) {
    {
        std::lock_guard<std::mutex> lk(mu_);
        outstanding_++;
    }
    tasks_.push([this, jp = &J, have] {
        const auto t0 = std::chrono::steady_clock::now();
        try {
            advance(*jp, have);
        } catch (const std::bad_alloc &) {
            // The job's step did not complete: it ends with an error.
            if (!jp->out.done) fail_job(*jp, MPOA_ERR_ALLOC, "host memory allocation failed");
        } catch (const std::exception &e) {
            if (!jp->out.done) fail_job(*jp, MPOA_ERR_INVALID_ARGUMENT, std::string("unexpected error: ") + e.what());
        }
        task_ns_ += std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now() - t0).count();
        std::lock_guard<std::mutex> lk(mu_);
        outstanding_--;
        cv_.notify_all();
    }, J.in.crit);
}

// One thread per slot: take a batch, launch it, wait for it on the slot's own
// stream and hand its jobs back to the host workers. A slot's packing and
// waiting never hold up another slot.
void Engine::slot_main(size_t si) {
    try {
        slot_loop(si);
    } catch (...) {
        // A slot that cannot go on (host allocation) stops the engine.
        std::lock_guard<std::mutex> lk(mu_);
        rc_ = MPOA_ERR_ALLOC;
        stop_ = true;
        cv_.notify_all();
    }
}

void Engine::slot_loop(size_t si) {
    Ctx &c = *slots_[si];
    const bool crit_slot = o_.crit_lane && si == 0;
    // Ordinary slots: the first long_slots take long DPs first, then short
    // ones; the others take short DPs only.
    const size_t ord = o_.crit_lane ? si - 1 : si;
    const int role = !crit_slot && ord < o_.long_slots ? 1 : 0;
    // Serving (long and short lanes): launches group DPs of similar rows,
    // and within each lane every second slot starts from the smallest DPs,
    // so short chains are not held behind the largest jobs' launches.
    const bool by_size = o_.long_rows != 0 && !crit_slot;
    const size_t lane_first = role == 1 ? 0 : o_.long_slots;
    const bool from_small = by_size && (ord - lane_first) % 2 == 1;
    std::vector<uint8_t> taken;
    std::vector<Dp *> batch;
    std::vector<Job *> back;
    std::vector<JobPlan *> plans;
    LaunchInfo li;
    for (;;) {
        batch.clear();
        {
            std::unique_lock<std::mutex> lk(mu_);
            for (;;) {
                if (stop_ || rc_ != 0) return;
                if (crit_slot && !ready_crit_.empty()) {
                    // Launch once the critical job with the most work is
                    // ready and the others have had a short while to join; a
                    // job that misses the launch takes an ordinary slot for
                    // that step (requeue).
                    const bool top_ready = std::any_of(ready_crit_.begin(), ready_crit_.end(),
                                                       [&](const Dp *D) { return D->job == top_job_; });
                    const double t = now_ms();
                    if ((top_ready && (ready_crit_.size() >= active_crit_ || t - top_since_ >= crit_gather_ms_)) ||
                        t - crit_since_ >= crit_wait_ms_)
                        break;
                    cv_.wait_for(lk, std::chrono::microseconds(200));
                } else if (crit_slot || (ready_.empty() && (role == 0 || ready_long_.empty()))) {
                    cv_.wait(lk);
                } else {
                    break;
                }
            }
            std::vector<Dp *> &q = crit_slot ? ready_crit_ : (role == 1 && !ready_long_.empty() ? ready_long_ : ready_);
            const size_t idle = std::max<size_t>(1, &q == &ready_long_ ? idle_[1] : idle_[0] + (role == 1 ? 1 : 0));
            if (!crit_slot) {
                if (by_size) {
                    // Serving: by rows, so that a launch holds DPs of similar
                    // length (a DP waits for the longest one in its launch).
                    std::sort(q.begin(), q.end(), [](const Dp *x, const Dp *y) {
                        return x->plan.n != y->plan.n ? x->plan.n > y->plan.n : x->rem > y->rem;
                    });
                } else {
                    // Longest remaining chain first (remaining reads times
                    // the current step's cells); neighbours in that order
                    // have similar step sizes, and a launch lasts as long as
                    // its longest job.
                    std::sort(q.begin(), q.end(), [](const Dp *x, const Dp *y) { return x->rem > y->rem; });
                }
            }
            const uint64_t want = crit_slot ? q.size() : std::min<uint64_t>(max_launch_, (q.size() + idle - 1) / idle);
            const size_t nq = q.size();
            const Dp *first = from_small ? q[nq - 1] : q[0];
            const bool i64 = first->plan.use_i64;  // one score width per launch
            const uint64_t n0 = first->plan.n;
            size_t w = 0;
            uint64_t arena = 0;
            taken.assign(nq, 0);
            for (size_t t = 0; t < nq; t++) {
                const size_t k = from_small ? nq - 1 - t : t;
                const Dp *D = q[k];
                const uint64_t a = job_arena_bytes(D->plan, i64);
                // by_size: within a factor of two of the launch's first DP.
                const bool similar = !by_size || (from_small ? D->plan.n <= 2 * n0 : 2 * uint64_t(D->plan.n) >= n0);
                if (batch.size() < want && D->plan.use_i64 == i64 && similar &&
                    (batch.empty() || arena + a <= arena_budget_)) {
                    batch.push_back(q[k]);
                    taken[k] = 1;
                    arena += a;
                }
            }
            for (size_t k = 0; k < nq; k++)
                if (!taken[k]) q[w++] = q[k];
            q.resize(w);
            if (crit_slot) {
                crit_busy_ = true;
                // Jobs left behind (another score width) stay queued.
                if (w) crit_since_ = now_ms();
            } else {
                idle_[role]--;
            }
        }
        plans.resize(batch.size());
        for (size_t k = 0; k < batch.size(); k++) plans[k] = &batch[k]->plan;
        const double t0 = now_ms();
        RunStats rs;
        int r = launch_begin(c, p_, plans, batch[0]->plan.use_i64, false, li);
        const double t1 = now_ms();
        pack_ns_ += int64_t((t1 - t0) * 1e6);
        if (r == MPOA_OK) r = launch_finish(c, li, &rs);
        if (r != MPOA_OK) {
            std::lock_guard<std::mutex> lk(mu_);
            rc_ = r;
            stop_ = true;
            cv_.notify_all();
            return;
        }
        const DevResult *hr = launch_results(li);
        for (size_t k = 0; k < batch.size(); k++) {
            Dp &D = *batch[k];
            D.res = hr[k];
            D.fresh = true;
            if (hr[k].status == kOk) {
                const uint32_t *t = launch_tb(c, li, k);
                D.tb.assign(t, t + D.plan.m);
            }
        }
        {
            std::lock_guard<std::mutex> lk(mu_);
            RunStats &x = sst_[si];
            x.cells += rs.cells;
            x.dp_calls += rs.dp_calls;
            x.i64_calls += rs.i64_calls;
            x.pred_cells += rs.pred_cells;
            x.fill_launches += rs.fill_launches;
            x.device_ms += rs.device_ms;
            x.fill_ms += rs.fill_ms;
            x.arena_retries += rs.arena_retries;
            x.arena_bytes += rs.arena_bytes;
            x.record_launches += rs.record_launches;
            if (trace_path_) {
                TraceRec t{int(si), t0, t1, now_ms(), li.fill_ms, li.all_ms, li.h2d_ms, li.d2h_ms, li.grown, li.arena_bytes,
                           c.d_arena.cap, batch.size(), 0, 0, -1, li.ring_bytes, li.width, li.shape, li.rows};
                for (const Dp *D : batch) {
                    t.max_cost = std::max(t.max_cost, D->cost);
                    t.sum_cost += D->cost;
                    if (D->job == watched_) t.watched_step = D->job->step;
                }
                strace_[si].push_back(t);
            }
            // A job goes back to the host once every DP of its step is back.
            back.clear();
            for (Dp *D : batch)
                if (--D->job->pending == 0) back.push_back(D->job);
            if (crit_slot) {
                // The top job's host step first, then the rest by cells left.
                std::stable_sort(back.begin(), back.end(), [&](const Job *x, const Job *y) {
                    if ((x == top_job_) != (y == top_job_)) return x == top_job_;
                    return x->rem > y->rem;
                });
                crit_busy_ = false;
            } else {
                idle_[role]++;
            }
        }
        for (Job *J : back) task(*J, true);
    }
}

int run_engine(std::vector<Ctx *> &slots, TaskQueue &tasks, const Params &p, const BatchInput &in,
               BatchOutput &out, EngineStats &st) {
    const uint64_t N = in.njobs;
    // Largest jobs first: they bound the makespan through their chain of
    // steps. With two or more slots they get slot 0 to themselves.
    std::vector<uint64_t> order(N), work(N);
    for (uint64_t j = 0; j < N; j++) {
        const uint64_t R = in.job_off[j + 1] - in.job_off[j];
        const uint64_t r0 = in.job_off[j];
        const uint64_t L = R ? in.read_off[r0 + 1] - in.read_off[r0] : 0;
        work[j] = L * R * R;
        order[j] = j;
    }
    std::stable_sort(order.begin(), order.end(), [&](uint64_t a, uint64_t b) { return work[a] > work[b]; });
    std::vector<uint8_t> crit(N, 0);
    if (slots.size() >= 2 && N > 1) {
        const double frac = double(env_u64("MPOA_CRIT_PCT", 50, 0, 100)) / 100.0;
        const uint64_t cap = env_u64("MPOA_CRIT_MAX", 256, 0, N);
        for (uint64_t k = 0; k < N && k < cap; k++) {
            if (double(work[order[k]]) < frac * double(work[order[0]])) break;
            crit[order[k]] = 1;
        }
    }
    EngineOpts o;
    o.render = in.render;
    o.crit_lane = true;
    o.admit_bytes = ~0ull;  // a batch admits every job at once
    int rc;
    std::vector<uint64_t> id(N);
    {
        Engine E(slots, tasks, p, o);
        for (uint64_t j : order) {
            JobIn ji;
            ji.bases = in.bases;
            ji.read_off = in.read_off + in.job_off[j];
            ji.nreads = uint32_t(in.job_off[j + 1] - in.job_off[j]);
            ji.crit = crit[j] != 0;
            ji.guide = p.progressive;
            ji.work = work[j];
            id[j] = E.submit(std::move(ji));
        }
        rc = E.wait_all();
        if (rc == MPOA_OK) {
            for (uint64_t j = 0; j < N; j++) {
                JobOut &jo = E.out(id[j]);
                (*out.status)[j] = jo.status;
                (*out.err)[j] = std::move(jo.err);
                (*out.out0)[j] = std::move(jo.out0);
                (*out.out1)[j] = std::move(jo.out1);
                (*out.graphs)[j] = std::move(jo.graph);
            }
        }
        st = E.stats();
    }
    return rc;
}

}  // namespace mpoa
