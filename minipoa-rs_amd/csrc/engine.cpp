// Continuous batched POA engine (engine.hpp).
#include "engine.hpp"

#include <algorithm>
#include <array>
#include <thread>
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>

#include "plan.hpp"

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

namespace {

// One job's progressive state.
struct EJob {
    uint64_t j = 0;
    uint32_t step = 0, nreads = 0;
    bool retry = false;      // the planned DP is poa's full-matrix retry
    JobPlan plan;
    uint64_t prev_used = 0, prev_n = 0;
    uint64_t cost = 0;       // expected cells of the planned DP
    bool crit = false;       // one of the largest jobs: runs in the priority slot
    uint64_t rem = ~0ull;    // cells left when last queued (0 once finished; under the engine lock)
    std::vector<AlignedBase> path;
    DevResult res{};
    std::vector<uint32_t> tb;
};

// One line per launch for MPOA_TRACE=<file>: scheduling diagnostics.
struct TraceRec {
    int slot;
    double t_begin, t_launched, t_done;  // host clock: before packing, after enqueueing, results on host
    float fill_ms, all_ms, h2d_ms, d2h_ms;  // device clock
    uint32_t grown;
    uint64_t arena_bytes, arena_cap;
    size_t njobs;
    uint64_t max_cost, sum_cost;
    int64_t watched_step;  // step of the largest job if it is in this launch, else -1
    uint32_t ring_bytes, width;
    FillShape shape;
};

uint64_t env_u64(const char *name, uint64_t dflt, uint64_t lo, uint64_t hi) {
    const char *s = getenv(name);
    if (!s) return dflt;
    const long long v = atoll(s);
    if (v < (long long)lo) return lo;
    if (v > (long long)hi) return hi;
    return uint64_t(v);
}

}  // namespace

int run_engine(std::vector<Ctx *> &slots, TaskQueue &tasks, const Params &p, const EngineInput &in,
               EngineOutput &out, EngineStats &st) {
    const uint64_t N = in.njobs;
    Params full = p;
    full.f = 0;
    std::vector<EJob> jobs(N);
    std::mutex mu;
    std::condition_variable cv;
    std::vector<EJob *> ready, ready_crit, crit_jobs;
    uint64_t finished = 0, active_crit = 0;
    bool stop = false;  // every job finished, or a launch failed
    const auto t_start = std::chrono::steady_clock::now();
    auto now_ms = [&] {
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
    };
    double crit_since = 0;  // when the first of the waiting critical jobs became ready
    uint64_t outstanding = 0;  // queued or running host tasks
    std::atomic<uint64_t> retries{0};
    std::atomic<int64_t> task_ns{0};

    auto finish = [&](EJob &J) {
        std::lock_guard<std::mutex> lk(mu);
        finished++;
        J.rem = 0;
        if (J.crit) active_crit--;
        if (finished == N) stop = true;
        cv.notify_all();
    };
    auto requeue = [&](EJob &J) {
        std::lock_guard<std::mutex> lk(mu);
        J.rem = J.cost * uint64_t(J.nreads - J.step);
        if (J.crit) {
            if (ready_crit.empty()) crit_since = now_ms();
            ready_crit.push_back(&J);
        } else {
            ready.push_back(&J);
        }
        cv.notify_all();
    };
    auto fail_job = [&](EJob &J, int status, const std::string &e) {
        (*out.status)[J.j] = status;
        (*out.err)[J.j] = e;
        finish(J);
    };
    auto query = [&](const EJob &J, const uint8_t *&q, uint64_t &qlen) {
        const uint64_t r = in.job_off[J.j] + J.step;
        q = in.bases + in.read_off[r];
        qlen = in.read_off[r + 1] - in.read_off[r];
    };
    // graph.rs:82-178 per read; false when the job ended with an error.
    auto fuse = [&](EJob &J) {
        Graph &g = *(*out.graphs)[J.j];
        std::string e = g.add_path(J.step, J.path.data(), J.path.size());
        if (e.empty()) e = g.toposort(false);
        if (!e.empty()) {
            fail_job(J, MPOA_ERR_GRAPH, e);
            return false;
        }
        J.step++;
        return true;
    };
    auto render = [&](EJob &J) {
        Graph &g = *(*out.graphs)[J.j];
        (*out.out0)[J.j] = ">consensus sequence\n" + g.consensus_bytes(p) + "\n";
        std::vector<std::string> names(J.nreads), comments(J.nreads);
        for (uint32_t k = 0; k < J.nreads; k++) names[k] = "read" + std::to_string(k);
        std::string e = g.msa(p, names, comments, (*out.out1)[J.j]);
        if (!e.empty()) {
            (*out.status)[J.j] = MPOA_ERR_GRAPH;
            (*out.err)[J.j] = e;
        }
    };
    // Consumes the job's DP result (have), fuses it and plans the next DP.
    // src/lib.rs:195-202 (one read per step), src/align.rs:81-98 (empty
    // graph), src/align.rs:202-207 (full-matrix retry on any error).
    auto advance = [&](EJob &J, bool have) {
        Graph &g = *(*out.graphs)[J.j];
        for (;;) {
            if (have) {
                have = false;
                if (J.plan.status == MPOA_OK && J.res.status == kArenaOverflow) {
                    J.plan.full_arena = true;  // same DP, arena for full rows
                    requeue(J);
                    return;
                }
                const bool failed = J.plan.status != MPOA_OK || J.res.status != kOk;
                if (failed && p.f != 0 && !J.retry) {
                    J.retry = true;
                    retries++;
                    const uint8_t *q;
                    uint64_t qlen;
                    query(J, q, qlen);
                    plan_job(full, GraphAcc{g}, 0, 1, q, qlen, false, p.b, 0, J.plan);
                    if (J.plan.status != MPOA_OK) {
                        have = true;
                        continue;
                    }
                    J.cost = J.plan.static_cells;
                    requeue(J);
                    return;
                }
                if (J.plan.status != MPOA_OK) {
                    fail_job(J, J.plan.status, J.plan.err);
                    return;
                }
                std::string msg;
                const int sst = device_status(J.res, msg);
                if (sst != MPOA_OK) {
                    fail_job(J, sst, msg);
                    return;
                }
                if (!J.retry) {
                    J.prev_used = J.res.arena_used;
                    J.prev_n = J.plan.n;
                }
                J.retry = false;
                assemble(GraphAcc{g}, J.plan, J.tb.data(), J.path);
                if (!fuse(J)) return;
            }
            if (J.step >= J.nreads) {
                if (in.render) render(J);
                finish(J);
                return;
            }
            const uint8_t *q;
            uint64_t qlen;
            query(J, q, qlen);
            if (g.size() <= 2) {
                J.path.clear();
                J.path.push_back(AlignedBase{0, int64_t(p.encode('N')), -1});
                for (uint64_t t = 0; t < qlen; t++)
                    J.path.push_back(AlignedBase{-1, int64_t(p.encode(q[t])), -1});
                if (!fuse(J)) return;
                continue;
            }
            JobPlan &P = J.plan;
            plan_job(p, GraphAcc{g}, 0, 1, q, qlen, p.adaptive, p.b, 0, P);
            if (P.status != MPOA_OK) {
                have = true;
                continue;
            }
            J.cost = P.static_cells;
            // Adaptive bands: start from what the previous step's DP used,
            // scaled to the grown graph.
            if ((P.flags & kAdaptive) && J.prev_n > 0) {
                uint64_t hint = J.prev_used * P.n / J.prev_n;
                hint += hint / 8 + 3ull * P.padded * 8;
                P.arena_est = std::min(std::max(P.arena_est, hint), arena_full(P));
                P.width_hint = uint32_t(J.prev_used / (3 * J.prev_n));
                J.cost = uint64_t(P.width_hint) * P.n;
            }
            requeue(J);
            return;
        }
    };
    // MPOA_TRACE also records the host side of the largest job's steps:
    // result on the host, task start, task end (host clock, ms).
    const EJob *watched = nullptr;
    std::vector<std::array<double, 3>> wtrace;
    auto submit = [&](EJob &J, bool have) {
        {
            std::lock_guard<std::mutex> lk(mu);
            outstanding++;
        }
        tasks.push([&, jp = &J, have, ts = now_ms()] {
            const auto t0 = std::chrono::steady_clock::now();
            const double tb = jp == watched ? now_ms() : 0;
            advance(*jp, have);
            if (jp == watched) wtrace.push_back({ts, tb, now_ms()});
            task_ns += std::chrono::duration_cast<std::chrono::nanoseconds>(
                           std::chrono::steady_clock::now() - t0)
                           .count();
            std::lock_guard<std::mutex> lk(mu);
            outstanding--;
            cv.notify_all();
        }, J.crit);
    };

    // Largest jobs first: they bound the makespan.
    std::vector<uint64_t> order(N), work(N);
    for (uint64_t j = 0; j < N; j++) {
        EJob &J = jobs[j];
        J.j = j;
        J.nreads = uint32_t(in.job_off[j + 1] - in.job_off[j]);
        const uint64_t r0 = in.job_off[j];
        const uint64_t L = J.nreads ? in.read_off[r0 + 1] - in.read_off[r0] : 0;
        work[j] = L * J.nreads * J.nreads;
        order[j] = j;
    }
    std::stable_sort(order.begin(), order.end(), [&](uint64_t a, uint64_t b) { return work[a] > work[b]; });
    // The largest jobs bound the makespan through their chain of steps. With
    // two or more slots they get slot 0 to themselves and advance in
    // lockstep: relaunched as soon as all of them are ready (or after a short
    // wait), instead of queueing behind other launches.
    const bool crit_lane = slots.size() >= 2 && N > 1;
    if (crit_lane) {
        const double frac = double(env_u64("MPOA_CRIT_PCT", 50, 0, 100)) / 100.0;
        const uint64_t cap = env_u64("MPOA_CRIT_MAX", 256, 0, N);
        for (uint64_t k = 0; k < N && k < cap; k++) {
            if (double(work[order[k]]) < frac * double(work[order[0]])) break;
            jobs[order[k]].crit = true;
            crit_jobs.push_back(&jobs[order[k]]);
            active_crit++;
        }
    }
    const double crit_wait_ms = double(env_u64("MPOA_CRIT_WAIT_US", 50000, 0, 10000000)) / 1000.0;
    const char *trace_path = getenv("MPOA_TRACE");
    if (trace_path && N) watched = &jobs[order[0]];
    for (uint64_t j : order) submit(jobs[j], false);

    const uint64_t max_launch = env_u64("MPOA_MAX_LAUNCH", 1024, 1, 1u << 20);
    const size_t nslots = slots.size();
    std::vector<RunStats> sst(nslots);
    std::vector<std::vector<TraceRec>> strace(nslots);
    // Each slot's score arena is bounded so that every slot's buffers (which
    // grow with 50% headroom) fit the device together; a launch takes jobs
    // until its arena reaches the bound (at least one job).
    uint64_t arena_budget = env_u64("MPOA_SLOT_ARENA_MB", 0, 0, 1ull << 30) << 20;
    if (!arena_budget && nslots) {
        const uint64_t free = device_free_bytes(slots[0]->device);
        arena_budget = std::max<uint64_t>(free / 10 * 6 / nslots / 3 * 2, 256ull << 20);
    }
    int rc = MPOA_OK;  // under mu
    size_t rest_idle = crit_lane ? nslots - 1 : nslots;  // under mu
    {
        std::lock_guard<std::mutex> lk(mu);
        if (finished == N) stop = true;
    }
    // One thread per slot: take a batch, launch it, wait for it on the slot's
    // own stream and hand its jobs back to the host workers. A slot's
    // packing and waiting never hold up another slot.
    auto slot_main = [&](size_t si) {
        Ctx &c = *slots[si];
        const bool crit_slot = crit_lane && si == 0;
        std::vector<EJob *> batch;
        std::vector<JobPlan *> plans;
        LaunchInfo li;
        for (;;) {
            batch.clear();
            {
                std::unique_lock<std::mutex> lk(mu);
                for (;;) {
                    if (stop) return;
                    if (crit_slot && !ready_crit.empty()) {
                        // Launch once the critical job with the most cells
                        // left is ready, taking whichever others are ready
                        // too: a job that misses a launch has fewer cells
                        // left than the one it would hold up.
                        uint64_t top = 0, top_ready = 0;
                        for (const EJob *J : crit_jobs) top = std::max(top, J->rem);
                        for (const EJob *J : ready_crit) top_ready = std::max(top_ready, J->rem);
                        if (top_ready >= top || ready_crit.size() >= active_crit ||
                            now_ms() - crit_since >= crit_wait_ms)
                            break;
                        cv.wait_for(lk, std::chrono::microseconds(200));
                    } else if (crit_slot || ready.empty()) {
                        cv.wait(lk);
                    } else {
                        break;
                    }
                }
                std::vector<EJob *> &q = crit_slot ? ready_crit : ready;
                if (!crit_slot) {
                    // Longest remaining chain first (remaining reads times the
                    // current step's cells); neighbours in that order have
                    // similar step sizes, and a launch lasts as long as its
                    // longest job.
                    auto rem = [](const EJob *J) { return J->cost * uint64_t(J->nreads - J->step); };
                    std::sort(q.begin(), q.end(), [&](const EJob *x, const EJob *y) { return rem(x) > rem(y); });
                }
                const uint64_t want =
                    crit_slot ? q.size() : std::min<uint64_t>(max_launch, (q.size() + rest_idle - 1) / rest_idle);
                const bool i64 = q[0]->plan.use_i64;  // one score width per launch
                size_t w = 0;
                uint64_t arena = 0;
                for (size_t k = 0; k < q.size(); k++) {
                    const uint64_t a = job_arena_bytes(q[k]->plan, i64);
                    if (batch.size() < want && q[k]->plan.use_i64 == i64 &&
                        (batch.empty() || arena + a <= arena_budget)) {
                        batch.push_back(q[k]);
                        arena += a;
                    } else {
                        q[w++] = q[k];
                    }
                }
                q.resize(w);
                if (crit_slot && w) crit_since = now_ms();
                if (!crit_slot) rest_idle--;
            }
            plans.resize(batch.size());
            for (size_t k = 0; k < batch.size(); k++) plans[k] = &batch[k]->plan;
            const double t0 = now_ms();
            int r = launch_begin(c, p, plans, batch[0]->plan.use_i64, false, li);
            const double t1 = now_ms();
            if (r == MPOA_OK) r = launch_finish(c, li, &sst[si]);
            if (r != MPOA_OK) {
                std::lock_guard<std::mutex> lk(mu);
                rc = r;
                stop = true;
                cv.notify_all();
                return;
            }
            if (trace_path) {
                TraceRec t{int(si), t0, t1, now_ms(), li.fill_ms, li.all_ms, li.h2d_ms, li.d2h_ms, li.grown, li.arena_bytes, c.d_arena.cap, batch.size(), 0, 0, -1,
                           li.ring_bytes, li.width, li.shape};
                for (const EJob *J : batch) {
                    t.max_cost = std::max(t.max_cost, J->cost);
                    t.sum_cost += J->cost;
                    if (J == watched) t.watched_step = J->step;
                }
                strace[si].push_back(t);
            }
            const DevResult *hr = launch_results(c);
            for (size_t k = 0; k < batch.size(); k++) {
                EJob &J = *batch[k];
                J.res = hr[k];
                if (hr[k].status == kOk) {
                    const uint32_t *t = launch_tb(c, li, k);
                    J.tb.assign(t, t + J.plan.m);
                }
                submit(J, true);
            }
            if (!crit_slot) {
                std::lock_guard<std::mutex> lk(mu);
                rest_idle++;
            }
        }
    };
    std::vector<std::thread> threads;
    for (size_t si = 0; si < nslots; si++) threads.emplace_back(slot_main, si);
    for (auto &t : threads) t.join();
    // Let every queued task finish before the state they reference goes away.
    {
        std::unique_lock<std::mutex> lk(mu);
        cv.wait(lk, [&] { return outstanding == 0; });
    }
    std::vector<TraceRec> trace;
    for (size_t si = 0; si < nslots; si++) {
        const RunStats &x = sst[si];
        st.run.cells += x.cells;
        st.run.dp_calls += x.dp_calls;
        st.run.i64_calls += x.i64_calls;
        st.run.pred_cells += x.pred_cells;
        st.run.fill_launches += x.fill_launches;
        st.run.device_ms += x.device_ms;
        st.run.fill_ms += x.fill_ms;
        trace.insert(trace.end(), strace[si].begin(), strace[si].end());
    }
    std::sort(trace.begin(), trace.end(), [](const TraceRec &x, const TraceRec &y) { return x.t_begin < y.t_begin; });
    if (trace_path) {
        if (FILE *f = fopen(trace_path, "w")) {
            fprintf(f, "slot,t_begin_ms,t_launched_ms,t_done_ms,fill_ms,all_ms,h2d_ms,d2h_ms,grown,arena_bytes,arena_cap,njobs,max_cost,sum_cost,"
                       "watched_step,ring_bytes,width,cols,waves\n");
            for (const TraceRec &t : trace)
                fprintf(f, "%d,%.3f,%.3f,%.3f,%.3f,%.3f,%.3f,%.3f,%u,%llu,%llu,%zu,%llu,%llu,%lld,%u,%u,%d,%d\n", t.slot, t.t_begin,
                        t.t_launched, t.t_done, t.fill_ms, t.all_ms, t.h2d_ms, t.d2h_ms, t.grown, (unsigned long long)t.arena_bytes,
                        (unsigned long long)t.arena_cap,
                        t.njobs, (unsigned long long)t.max_cost, (unsigned long long)t.sum_cost,
                        (long long)t.watched_step, t.ring_bytes, t.width, t.shape.cols, t.shape.waves);
            fclose(f);
        }
        const std::string wpath = std::string(trace_path) + ".watch";
        if (FILE *f = fopen(wpath.c_str(), "w")) {
            fprintf(f, "t_result_ms,t_task_begin_ms,t_task_end_ms\n");
            for (const auto &w : wtrace) fprintf(f, "%.3f,%.3f,%.3f\n", w[0], w[1], w[2]);
            fclose(f);
        }
    }
    st.run.retries += retries.load();
    st.host_ms += double(task_ns.load()) / 1e6 / double(std::max(1u, tasks.size()));
    return rc;
}

}  // namespace mpoa
