// Device side of the resident jobs (resident.hpp): the step kernel, one
// wave per job, and the per-step reset of the arena counter. A graph node
// clears the counter rather than a memset node (memset nodes in captured
// graphs are avoided on this platform).
#include <hip/hip_runtime.h>

#include "kernels.hpp"
#include "resident.hpp"
#include "resident_dev.hpp"

namespace mpoa {

namespace {

__global__ __launch_bounds__(64) void rj_step_kernel(RPools R) {
    extern __shared__ __align__(16) char rj_lds[];
    WaveTeam t{threadIdx.x};
    if (blockIdx.x >= R.njobs) return;
    rj_job_step(t, R, blockIdx.x, rj_lds, R.lds_bytes);
}

__global__ void rj_reset_kernel(unsigned long long *ctr) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *ctr = 0;
}

// Job k's graph into out + off[k] (layout rj_pack_layout), one wave per job.
// sel: the jobs to pack (k-th packed job jobs[sel[k]]), or null for all.
__global__ __launch_bounds__(64) void rj_pack_kernel(const RJob *jobs, const uint32_t *sel, uint32_t njobs,
                                                     const uint64_t *off, uint8_t *out) {
    const uint32_t k = blockIdx.x;
    if (k >= njobs) return;
    rj_pack_job(jobs[sel ? sel[k] : k], out + off[k], threadIdx.x, 64);
}

}  // namespace

void launch_rj_pack(const RJob *jobs, const uint32_t *sel, uint32_t njobs, const uint64_t *off, uint8_t *out,
                    hipStream_t s) {
    if (njobs == 0) return;
    hipLaunchKernelGGL(rj_pack_kernel, dim3(njobs), dim3(64), 0, s, jobs, sel, njobs, off, out);
}

void launch_rj_step(const RPools &R, hipStream_t s) {
    if (R.njobs == 0) return;
    hipLaunchKernelGGL(rj_step_kernel, dim3(R.njobs), dim3(64), R.lds_bytes, s, R);
}

void launch_rj_reset(unsigned long long *ctr, hipStream_t s) {
    hipLaunchKernelGGL(rj_reset_kernel, dim3(1), dim3(64), 0, s, ctr);
}

}  // namespace mpoa
