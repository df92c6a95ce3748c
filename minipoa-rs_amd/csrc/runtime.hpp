// Device context and DP orchestration: plans go up as one blob per launch,
// the fill and traceback kernels run on the context's stream, and results
// plus per-column traceback records come back for host assembly.
#pragma once
#include <hip/hip_runtime_api.h>

#include <atomic>
#include <condition_variable>
#include <cstdint>
#include <functional>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "dp_types.hpp"
#include "graph.hpp"
#include "kernels.hpp"
#include "plan.hpp"

namespace mpoa {

void set_error(const std::string &msg);
const char *last_error();

// Fixed pool of host workers for per-job planning, assembly and fusion.
class ThreadPool {
public:
    explicit ThreadPool(unsigned n);
    ~ThreadPool();
    unsigned size() const { return unsigned(workers_.size()) + 1; }
    // Runs f(i) for i in [0, n) across the pool and the calling thread.
    void parallel_for(size_t n, const std::function<void(size_t)> &f);

private:
    void loop();
    std::vector<std::thread> workers_;
    std::mutex mu_;
    std::condition_variable cv_, done_cv_;
    const std::function<void(size_t)> *fn_ = nullptr;
    size_t n_ = 0;
    std::atomic<size_t> next_{0};
    size_t generation_ = 0;
    unsigned active_ = 0;
    bool stop_ = false;
};

struct DevBuf {
    void *p = nullptr;
    size_t cap = 0;
    hipError_t ensure(size_t bytes);
    // At least `bytes`, reallocated to exactly that size when smaller (for
    // buffers sized once, such as the resident arenas).
    hipError_t ensure_exact(size_t bytes);
    void release();
};

struct PinBuf {
    void *p = nullptr;
    size_t cap = 0;
    hipError_t ensure(size_t bytes);
    void release();
};

struct RunStats {
    uint64_t cells = 0, dp_calls = 0, retries = 0, i64_calls = 0, pred_cells = 0;
    uint64_t fill_launches = 0;
    uint64_t arena_retries = 0;  // DPs re-run with full-width arena rows after an overflow
    uint64_t arena_bytes = 0;    // traceback data the fills wrote (arena bytes of completed DPs)
    uint64_t record_launches = 0;  // fill launches in the records layout
    double device_ms = 0, fill_ms = 0, host_ms = 0;
};

constexpr int kLayoutAuto = 0, kLayoutValues = 1, kLayoutRecords = 2;

struct Ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    // fill start, fill end, traceback end, results on the host, inputs copy start
    hipEvent_t ev[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    DevBuf d_in, d_res, d_meta, d_adapt, d_tb, d_arena;
    PinBuf h_in, h_res, h_tb;
    ThreadPool *pool = nullptr;
    FillShape force_shape{0, 0};  // cols 0: choose from the expected row width
    // Launches of this context run their waves at raised issue priority
    // (s_setprio): the engine's priority slot, whose jobs bound the makespan.
    bool prio = false;
    // DP layout in HBM (kLayout*, the C ABI's MPOA_LAYOUT_*): auto takes
    // values, except for the no-allocation entry point, whose fixed scratch
    // holds several times more cells as records.
    int layout = 0;
    ~Ctx();
};

// Device memory that a launch carves all of its device buffers from instead
// of the context's own growable buffers (the no-allocation entry point's
// scratch). Host staging then uses pageable memory of the calling thread, so
// such a launch calls no HIP allocator at all.
struct FixedScratch {
    char *dev = nullptr;
    uint64_t bytes = 0;
};

// threads: host workers of the context's pool (0 = default_threads()).
int ctx_create(int device, Ctx **out, unsigned threads = 0);
unsigned default_threads();

// Runs every plan with status MPOA_OK on the device (i32 and i64 groups in
// separate launches; arena overflows are re-run with the full-width bound).
// Fills res[k] and tb (flattened, tb_off[k] .. tb_off[k] + m).
// With fx, every launch takes its device buffers from fx (a DP whose
// buffers do not fit returns MPOA_ERR_SCRATCH_LIMIT).
int run_plans(Ctx &ctx, const Params &p, std::vector<JobPlan *> &plans,
              std::vector<DevResult> &res, std::vector<uint32_t> &tb,
              std::vector<uint64_t> &tb_off, RunStats *stats,
              const FixedScratch *fx = nullptr);

// An asynchronous launch: one fill and one traceback over plans that share
// the score width, with staging in and results out on ctx's stream. The plans
// must stay alive until launch_finish; results stay in ctx's pinned buffers
// until its next launch.
struct LaunchInfo {
    size_t n = 0;
    bool i64 = false;
    std::vector<uint64_t> tb_at;  // per job: offset of its traceback record
    uint32_t ring_bytes = 0;      // the fill's LDS ring
    FillShape shape{0, 0};        // the fill's launch shape
    uint32_t width = 0;           // expected row width the shape was chosen for
    uint64_t rows = 0;            // DP rows over the launch's jobs
    uint32_t grown = 0;           // bit mask of the buffers this launch reallocated
    uint64_t arena_bytes = 0;     // score arena the launch asked for
    bool records = false;         // the launch's DP layout (else values)
    float fill_ms = 0, all_ms = 0;  // device time of the fill, and of the fill and traceback
    float h2d_ms = 0, d2h_ms = 0;   // device time of the input and result copies
    const void *hres = nullptr;     // host copies of the results and traceback records
    const void *htb = nullptr;
};
int launch_begin(Ctx &ctx, const Params &p, const std::vector<JobPlan *> &plans, bool i64,
                 bool full_arena, LaunchInfo &li, const FixedScratch *fx = nullptr);
// The launch shape and LDS ring the tuning table named by MPOA_TUNING
// (minipoa_hip/tune.py) gives an expected row width; false without a table
// or an entry.
bool tuned_shape(uint32_t width, FillShape &shape, uint32_t &ring_bytes);
// Score-arena bytes job J adds to a launch, its share of the overflow pool
// included.
uint64_t job_arena_bytes(const JobPlan &J, bool i64);
// Free device memory (0 if it cannot be queried).
uint64_t device_free_bytes(int device);
// True once the launch's results are on the host (or it failed).
bool launch_done(Ctx &ctx);
// Waits for the launch and adds its timings and cells to *st.
int launch_finish(Ctx &ctx, LaunchInfo &li, RunStats *st);
const DevResult *launch_results(const LaunchInfo &li);
const uint32_t *launch_tb(const Ctx &ctx, const LaunchInfo &li, size_t k);

// Maps a failed DevResult to (status, message) as the reference reports it.
int device_status(const DevResult &r, std::string &msg);

}  // namespace mpoa
