// Continuous batched POA engine. Every job advances through its progressive
// steps on its own (src/lib.rs:184-204 per job): when a job's DP comes back
// from the device, a host worker fuses the path into its graph, plans the
// job's next DP and hands it back to the dispatcher. The dispatcher keeps
// several launch slots (streams with their own buffers) busy with whatever
// jobs are ready, grouping jobs of similar size, so no job waits for a
// longer one to finish its step.
#pragma once
#include <condition_variable>
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

// Inputs: job i holds reads job_off[i] .. job_off[i+1] of read_off / bases.
struct EngineInput {
    uint64_t njobs = 0;
    const uint64_t *job_off = nullptr;
    const uint64_t *read_off = nullptr;
    const uint8_t *bases = nullptr;
    // Render the -r0 consensus FASTA and the -r1 MSA of each finished job.
    bool render = true;
};

// Outputs, one entry per job (sized by the caller).
struct EngineOutput {
    std::vector<int> *status;
    std::vector<std::string> *err, *out0, *out1;
    std::vector<std::unique_ptr<Graph>> *graphs;
};

struct EngineStats {
    RunStats run;
    double host_ms = 0;  // host task time divided by the number of workers
};

// slots: launch contexts (one stream each); tasks: host workers.
int run_engine(std::vector<Ctx *> &slots, TaskQueue &tasks, const Params &p, const EngineInput &in,
               EngineOutput &out, EngineStats &st);

}  // namespace mpoa
