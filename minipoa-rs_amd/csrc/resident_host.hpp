// Host entry points of the device-resident batch (resident.cpp).
#pragma once
#include <string>
#include <vector>

#include "engine.hpp"

namespace mpoa {

struct ResidentStats {
    uint64_t buckets = 0, replays = 0, cells = 0, dp_calls = 0, retries = 0, arena_retries = 0;
    uint64_t deferred = 0;       // DP placements postponed to a later step (arena spent)
    uint64_t evicted = 0;        // jobs that left the device and finished on the host engine
    uint64_t resident_done = 0;  // jobs that finished (or failed) on the device
    uint64_t host_jobs = 0;      // jobs the device path does not take (seeding, -p, i64, ...)
    double device_ms = 0;        // step loops, summed over buckets (HIP events)
    double resident_ms = 0;      // wall time of the device-resident part
};

// Device blocks and host staging of resident batches, one set per launch
// slot, kept with a context across runs.
struct ResidentCache;
ResidentCache *resident_cache_new();
void resident_cache_free(ResidentCache *c);
uint64_t resident_cache_bytes(const ResidentCache *c);

// Whether a job runs device-resident from its first read.
bool resident_eligible(const Params &p, const uint64_t *roff, uint32_t nreads);

// The batch of `in` with every eligible job device-resident (length
// buckets, captured step loops on the slots' streams); the rest, and jobs
// that leave the device, on the host-driven engine over the same slots.
int run_resident(std::vector<Ctx *> &slots, TaskQueue &tasks, ResidentCache &cache, const Params &p,
                 const BatchInput &in, BatchOutput &out, EngineStats &est, ResidentStats &rst);

// Applies the reads' paths (path r: paths[path_off[r] .. path_off[r+1]])
// to a host Graph and to the resident step's code on one host thread, and
// compares the two after every read (nodes, lists, ranks, paths) and their
// plans of every read's DP (banded and full). lds_bytes: the staging the
// step would have (0: every array in global memory). Returns 0 when equal,
// 1 with *report naming the first difference, 2 when the host rejects a path.
int resident_check(const Params &p, const uint8_t *bases, const uint64_t *read_off, uint32_t nreads,
                   const AlignedBase *paths, const uint64_t *path_off, uint32_t lds_bytes, std::string &report);

}  // namespace mpoa
