// Host entry points of the device-resident batch (resident.cpp).
#pragma once
#include <memory>
#include <string>
#include <vector>

#include "engine.hpp"
#include "resident.hpp"

namespace mpoa {

struct ResidentStats {
    uint64_t buckets = 0, replays = 0, cells = 0, dp_calls = 0, retries = 0, arena_retries = 0;
    uint64_t pred_cells = 0, arena_bytes = 0;  // band overlaps and traceback data of completed DPs
    uint64_t deferred = 0;       // DP placements postponed to a later step (arena spent)
    uint64_t evicted = 0;        // jobs that left the device and finished on the host engine
    uint64_t resident_done = 0;  // jobs that finished (or failed) on the device
    uint64_t host_jobs = 0;      // jobs the device path does not take (seeding, -p, i64, ...)
    double device_ms = 0;        // step loops, summed over buckets (HIP events)
    double resident_ms = 0;      // wall time of the device-resident part
};

// Helpers shared by the resident batch and server (resident.cpp).
inline uint64_t up16(uint64_t x) { return (x + 15) / 16 * 16; }
inline uint64_t up256(uint64_t x) { return (x + 255) / 256 * 256; }

// LDS of the step kernel for jobs whose longest template is lmax: its
// windows and the per-row arrays of a graph of about 3 x lmax nodes (12 B
// per row), between 16 KB and 64 KB. A gfx950 workgroup may hold 160 KB,
// but a step holding more than 64 KB keeps other buckets' fills off its CU:
// with up to 160 KB the 4,096-job batch ran 1.92-2.09 s per run instead of
// 1.53-1.63 s, and the serving stream's latencies did not improve.
inline uint32_t step_lds_bytes(uint64_t lmax) {
    const uint64_t want = kWinBytes + 1024 + 36 * lmax;
    return uint32_t(want < 16384 ? 16384 : want > 65536 ? 65536 : want);
}

// Scoring parameters of the step kernel; matrix: the matrix's address there.
RParams make_rparams(const Params &p, const int32_t *matrix);

// Capacities of one job: reads, nodes (2 + bases), edges (bases + reads +
// 1), query columns (longest read + 2), padded query bytes.
struct RCaps {
    uint32_t nreads = 0, nc = 0, ec = 0, maxm = 0, padm = 0;
    uint64_t total = 0;  // bases over all reads
};
RCaps job_caps(const uint64_t *roff, uint32_t nreads);

// Byte offsets of one job's graph arrays inside a block (each 16-aligned).
struct RGraphLayout {
    uint64_t base, parent, rank, order, aligned, out_head, out_tail, in_head, in_tail, in_cnt, cover, mark,
        edges, paths, path_off, csr_off, csr_adj, deg, stack, succ_beg, succ, head, tail, support, win, bytes;
};
RGraphLayout graph_layout(const RCaps &c, uint32_t alpha);
void bind_graph(RJob &J, char *blk, const RGraphLayout &L);
// Status and the reference's message of a failed resident job.
int failure_of(const RJob &J, std::string &msg);

void launch_rj_step(const RPools &R, hipStream_t s);
// The persistent resident kernel (kernels.hip): `grid` workgroups of the
// fill shape f (4x4 only), LDS from rj_persist_lds. Returns -1 for a shape
// without a persistent kernel.
uint32_t rj_persist_lds(const DevPools &P, uint32_t step_lds);
int launch_rj_persist(const RPools &R, const DevPools &P, const RPersist &Q, uint32_t grid, FillShape f,
                      hipStream_t s);
void launch_rj_reset(unsigned long long *ctr, hipStream_t s);
void launch_rj_pack(const RJob *jobs, const uint32_t *sel, uint32_t njobs, const uint64_t *off, uint8_t *out,
                    hipStream_t s);

// Device blocks and host staging of resident batches, one set per launch
// slot, kept with a context across runs.
struct ResidentCache;
ResidentCache *resident_cache_new();
void resident_cache_free(ResidentCache *c);
uint64_t resident_cache_bytes(const ResidentCache *c);
// Releases the device blocks (the persistent launch's arena is large), so
// that a host-driven run on the same context budgets its launches from the
// whole free memory; the next resident run allocates them again.
void resident_cache_trim(ResidentCache *c);

// Whether a job runs device-resident from its first read.
bool resident_eligible(const Params &p, const uint64_t *roff, uint32_t nreads);

// The batch of `in` with every eligible job device-resident (length
// buckets, captured step loops on the slots' streams); the rest, and jobs
// that leave the device, on the host-driven engine over the same slots.
int run_resident(std::vector<Ctx *> &slots, TaskQueue &tasks, ResidentCache &cache, const Params &p,
                 const BatchInput &in, BatchOutput &out, EngineStats &est, ResidentStats &rst);

// Serving on device-resident graphs (resident_server.cpp): the Engine's
// submit / done / wait / out surface over size-class step tables whose slots
// hold admitted jobs' graphs and DP descriptors in device memory between
// reads; jobs outside the resident path, and jobs that leave it, run on a
// host-driven Engine over `slots`.
class ResidentServer {
public:
    ResidentServer(std::vector<Ctx *> slots, TaskQueue &tasks, const Params &p, const EngineOpts &o);
    ~ResidentServer();  // finishes every submitted job
    ResidentServer(const ResidentServer &) = delete;
    ResidentServer &operator=(const ResidentServer &) = delete;
    uint64_t submit(JobIn in);
    bool done(uint64_t ticket);
    bool wait(uint64_t ticket);
    JobOut &out(uint64_t ticket);  // valid once done(ticket)
    uint64_t submitted();
    double now_ms() const;
    EngineStats stats(ResidentStats *rs = nullptr);
    int error();

private:
    struct Impl;
    std::unique_ptr<Impl> d_;
};

// Applies the reads' paths (path r: paths[path_off[r] .. path_off[r+1]])
// to a host Graph and to the resident step's code on one host thread, and
// compares the two after every read (nodes, lists, ranks, paths) and their
// plans of every read's DP (banded and full). lds_bytes: the staging the
// step would have (0: every array in global memory). Returns 0 when equal,
// 1 with *report naming the first difference, 2 when the host rejects a path.
int resident_check(const Params &p, const uint8_t *bases, const uint64_t *read_off, uint32_t nreads,
                   const AlignedBase *paths, const uint64_t *path_off, uint32_t lds_bytes, std::string &report);

}  // namespace mpoa
