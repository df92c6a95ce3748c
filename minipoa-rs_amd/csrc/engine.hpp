// Continuous batched POA engine. Every job advances through its progressive
// steps on its own (src/lib.rs:184-204 per job): when a job's DPs come back
// from the device, a host worker fuses the path into its graph, plans the
// job's next step and hands its DPs back to the dispatcher. A step is one
// whole-graph DP, or with minimizer seeding one DP per anchored window of
// the read (src/align.rs:99-188), all launched side by side. The dispatcher keeps
// several launch slots (streams with their own buffers) busy with whatever
// jobs are ready, grouping jobs of similar size, so no job waits for a
// longer one to finish its step.
//
// Jobs may arrive at any time (serving): a job is admitted when the device
// footprint of the admitted jobs leaves room for it, and joins the launches
// already in flight from its first DP on. A batch is the same engine with
// every job submitted up front.
#pragma once
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <functional>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "graph.hpp"
#include "runtime.hpp"

namespace mpoa {

// Host tasks run by a fixed set of workers: two FIFOs, the urgent one first.
class TaskQueue {
public:
    explicit TaskQueue(unsigned n);
    ~TaskQueue();
    unsigned size() const { return unsigned(workers_.size()); }
    void push(std::function<void()> f, bool urgent = false);

private:
    struct Fifo {
        std::vector<std::function<void()>> q;
        size_t head = 0;
        bool empty() const { return head == q.size(); }
    };
    void loop();
    std::vector<std::thread> workers_;
    std::mutex mu_;
    std::condition_variable cv_;
    Fifo urgent_, normal_;
    bool stop_ = false;
};

// Runs f(0) .. f(n-1) on the calling thread and on idle workers of `tq`, and
// returns once every index ran. Helpers that start after the indices are
// taken return at once, so the caller never waits on a queued task.
void par_for(TaskQueue &tq, size_t n, const std::function<void(size_t)> &f);

// One job: reads read_off[k] .. read_off[k+1] of `bases`, k < nreads, aligned
// progressively into `graph` (a new graph when null). `first` is the number
// of sequences already in the graph (extend_graph's `existing`): read k gets
// sequence id first + k, also when `guide` changes the order in which the
// reads are aligned. The reads are borrowed unless `own` holds them.
struct JobIn {
    const uint8_t *bases = nullptr;
    const uint64_t *read_off = nullptr;
    uint32_t nreads = 0;
    uint32_t first = 0;
    std::unique_ptr<Graph> graph;
    std::vector<uint8_t> own_bases;
    std::vector<uint64_t> own_off;
    bool guide = false;  // align in the -p guide order (src/minimizer.rs:489-578); first must be 0
    bool crit = false;   // runs in the priority slot (batch mode)
    uint64_t work = 0;   // L x R^2: orders the priority slot's jobs
};

struct JobOut {
    int status = 0;
    std::string err, out0, out1;  // -r0 consensus FASTA, -r1 MSA
    std::unique_ptr<Graph> graph;
    double t_submit = 0, t_admit = 0, t_done = 0;  // engine clock, ms
    uint64_t cells = 0;   // DP cells of the job's alignments
    uint32_t nseq = 0;    // sequences in the final graph
    bool done = false;
};

struct EngineOpts {
    bool render = true;          // render -r0 and -r1 for each finished job
    bool crit_lane = false;      // slot 0 runs only jobs submitted with crit
    uint64_t admit_bytes = 0;    // device footprint bound for admission (0: slots' budget)
    // DPs of at least long_rows rows queue apart from the others, and only
    // the first long_slots ordinary slots take them: a launch lasts as long
    // as its longest DP, so short DPs keep slots of their own (0: one queue).
    uint32_t long_rows = 0, long_slots = 0;
};

struct EngineStats {
    RunStats run;
    double host_ms = 0;  // host task time divided by the number of workers
    // Parts of the host task time (also per worker): planning the next DP
    // (head/tail distances, predecessor lists), consuming results (path
    // assembly), fusion (add_path + toposort) and rendering the outputs;
    // and the launch slots' packing of input blobs (per slot).
    double plan_ms = 0, consume_ms = 0, fuse_ms = 0, render_ms = 0, pack_ms = 0;
};

class Engine {
public:
    Engine(std::vector<Ctx *> slots, TaskQueue &tasks, const Params &p, const EngineOpts &o);
    ~Engine();  // finishes every submitted job, then stops the slots
    Engine(const Engine &) = delete;
    Engine &operator=(const Engine &) = delete;

    // Thread-safe. Returns the job's id (ids count from 0).
    uint64_t submit(JobIn in);
    // Blocks until the job finished; false if the engine failed first.
    bool wait(uint64_t id);
    bool done(uint64_t id);
    // Blocks until every submitted job finished (or the engine failed).
    int wait_all();
    // The job's results; valid once done(id).
    JobOut &out(uint64_t id);
    uint64_t submitted();
    double now_ms() const;
    EngineStats stats();
    int error();

    struct Job;
    struct Dp;

private:
    void finish(Job &J);
    void requeue(Job &J, const std::vector<Dp *> &dps);
    void fail_job(Job &J, int status, const std::string &e);
    bool fuse(Job &J);
    void render(Job &J);
    bool consume(Job &J, std::vector<Dp *> &again);
    bool plan_step(Job &J, std::vector<Dp *> &launch);
    void advance(Job &J, bool have);
    void task(Job &J, bool have);
    void admit_locked(Job &J);
    void slot_main(size_t si);
    void slot_loop(size_t si);
    uint64_t footprint_estimate(const Job &J) const;

    std::vector<Ctx *> slots_;
    TaskQueue &tasks_;
    Params p_, full_;
    EngineOpts o_;
    std::chrono::steady_clock::time_point t0_;

    std::mutex mu_;
    std::condition_variable cv_;
    std::deque<std::unique_ptr<Job>> jobs_;
    // DPs ready to launch (one per job step, or one per seeded window).
    std::vector<Dp *> ready_, ready_long_, ready_crit_;
    std::vector<Job *> crit_jobs_;
    std::deque<Job *> waiting_;  // submitted, not yet admitted (FIFO)
    uint64_t finished_ = 0, active_crit_ = 0, outstanding_ = 0;
    uint64_t active_bytes_ = 0, admit_budget_ = 0;
    double crit_since_ = 0, top_since_ = 0;
    bool crit_busy_ = false, stop_ = false;
    const Job *top_job_ = nullptr;
    const Job *watched_ = nullptr;  // the job with the most work (trace column)
    int rc_ = 0;
    size_t idle_[2] = {0, 0};  // idle ordinary slots: short-only, long-first
    uint64_t arena_budget_ = 0, max_launch_ = 0;
    double crit_wait_ms_ = 0, crit_gather_ms_ = 0;
    std::vector<RunStats> sst_;
    std::atomic<uint64_t> retries_{0};
    std::atomic<int64_t> task_ns_{0};
    std::atomic<int64_t> plan_ns_{0}, consume_ns_{0}, fuse_ns_{0}, render_ns_{0}, pack_ns_{0};
    std::vector<std::thread> threads_;

    // MPOA_TRACE=<file>: one CSV line per launch, written at shutdown.
    struct TraceRec;
    std::vector<std::vector<TraceRec>> strace_;
    const char *trace_path_ = nullptr;
};

// Runs the batch job_off/read_off/bases to completion (every job submitted
// at once, the largest ones in the priority slot).
struct BatchInput {
    uint64_t njobs = 0;
    const uint64_t *job_off = nullptr;
    const uint64_t *read_off = nullptr;
    const uint8_t *bases = nullptr;
    bool render = true;
};
struct BatchOutput {
    std::vector<int> *status;
    std::vector<std::string> *err, *out0, *out1;
    std::vector<std::unique_ptr<Graph>> *graphs;
};
int run_engine(std::vector<Ctx *> &slots, TaskQueue &tasks, const Params &p, const BatchInput &in,
               BatchOutput &out, EngineStats &st);

}  // namespace mpoa
