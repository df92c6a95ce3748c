// Achievable int32 VALU rate of the device: the denominator of the DP's
// int-VALU roofline. Every lane runs kChains independent chains of
// x <- max(x + a, c) (one v_add_u32 and one v_max_i32 per step); eight
// chains per lane and eight waves per SIMD keep every issue slot busy, and
// nothing but those two instructions sits in the loop body. Built as its own
// library (libmpoa_probe.so) for the benchmark; it is not part of the DP
// backend's C ABI.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

namespace {

constexpr int kChains = 8;
constexpr int kThreads = 256;

__global__ __launch_bounds__(kThreads) void int_valu_chains(int32_t *out, int iters, int32_t a,
                                                            int32_t c) {
    const int32_t t = int32_t(blockIdx.x * kThreads + threadIdx.x);
    int32_t x[kChains];
#pragma unroll
    for (int k = 0; k < kChains; k++) x[k] = t * (k + 3);
#pragma unroll 4
    for (int it = 0; it < iters; it++) {
#pragma unroll
        for (int k = 0; k < kChains; k++) {
            const int32_t s = int32_t(uint32_t(x[k]) + uint32_t(a));
            x[k] = s > c ? s : c;
        }
    }
    int32_t r = 0;
#pragma unroll
    for (int k = 0; k < kChains; k++) r ^= x[k];
    out[t] = r;
}

}  // namespace

extern "C" {

// Runs the probe `reps` times on `device` with `blocks` workgroups of 256
// lanes and `iters` steps per chain; *ops_per_s receives the best rate in
// int32 lane-operations per second (2 per step: the add and the max), *ms
// that run's kernel time. Returns 0, or the HIP error code.
int mpoa_probe_int_valu(int device, int blocks, int iters, int reps, double *ops_per_s,
                        double *ms) {
    if (blocks <= 0 || iters <= 0 || reps <= 0 || !ops_per_s || !ms) return int(hipErrorInvalidValue);
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) return int(e);
    int32_t *out = nullptr;
    hipEvent_t b = nullptr, f = nullptr;
    e = hipMalloc(&out, size_t(blocks) * kThreads * sizeof(int32_t));
    if (e == hipSuccess) e = hipEventCreate(&b);
    if (e == hipSuccess) e = hipEventCreate(&f);
    float best = 0.f;
    for (int r = 0; r < reps + 1 && e == hipSuccess; r++) {
        e = hipEventRecord(b, nullptr);
        if (e != hipSuccess) break;
        hipLaunchKernelGGL(int_valu_chains, dim3(blocks), dim3(kThreads), 0, nullptr, out, iters, -1,
                           int32_t(-(1 << 30)));
        e = hipGetLastError();
        if (e == hipSuccess) e = hipEventRecord(f, nullptr);
        if (e == hipSuccess) e = hipEventSynchronize(f);
        float t = 0.f;
        if (e == hipSuccess) e = hipEventElapsedTime(&t, b, f);
        if (e == hipSuccess && r > 0 && (best == 0.f || t < best)) best = t;  // run 0 warms up
    }
    if (b) (void)hipEventDestroy(b);
    if (f) (void)hipEventDestroy(f);
    if (out) (void)hipFree(out);
    if (e != hipSuccess) return int(e);
    const double ops = double(blocks) * kThreads * double(iters) * kChains * 2.0;
    *ms = best;
    *ops_per_s = best > 0.f ? ops / (double(best) * 1e-3) : 0.0;
    return 0;
}

}  // extern "C"
