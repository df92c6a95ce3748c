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
    const RJob &J = jobs[sel ? sel[k] : k];
    const uint32_t nn = J.nn, ne = J.ne, nr = J.nreads, np = J.npath;
    const RPackLayout L = rj_pack_layout(nn, ne, nr, np);
    uint8_t *o = out + off[k];
    for (uint32_t i = threadIdx.x; i < nn; i += 64) {
        o[L.base + i] = J.base[i];
        reinterpret_cast<uint32_t *>(o + L.parent)[i] = J.parent[i];
        reinterpret_cast<uint32_t *>(o + L.rank)[i] = J.rank[i];
        reinterpret_cast<uint32_t *>(o + L.cover)[i] = J.cover[i];
        reinterpret_cast<uint32_t *>(o + L.out_head)[i] = J.out_head[i];
        reinterpret_cast<uint32_t *>(o + L.in_head)[i] = J.in_head[i];
    }
    for (uint32_t i = threadIdx.x; i < ne; i += 64) reinterpret_cast<REdge *>(o + L.edges)[i] = J.edges[i];
    for (uint32_t i = threadIdx.x; i <= nr; i += 64)
        reinterpret_cast<uint32_t *>(o + L.path_off)[i] = i <= J.cur ? J.path_off[i] : 0u;
    for (uint32_t i = threadIdx.x; i < np; i += 64) reinterpret_cast<uint32_t *>(o + L.paths)[i] = J.paths[i];
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
