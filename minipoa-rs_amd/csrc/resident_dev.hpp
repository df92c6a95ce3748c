// Device side of the resident step shared by the kernels that run it (the
// step kernel of the captured step loops, resident.hip, and the persistent
// resident kernel, kernels.hip): the team of resident.hpp as one wave, and
// the serial head and tail passes on that wave.
#pragma once
#include <hip/hip_runtime.h>

#include "resident.hpp"

namespace mpoa {

namespace {

// The team of resident.hpp on the device: one wave of 64 lanes.
struct WaveTeam {
    uint32_t lane;
    static constexpr uint32_t size = 64;
    // The team is one wave, which may be one of several in its workgroup
    // (the persistent kernel runs the step on wave 0 alone): ordering of the
    // wave's own LDS and global accesses, no workgroup barrier.
    __device__ void sync() {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    }
    __device__ uint32_t excl_scan(uint32_t v, uint32_t &tot) {
        uint32_t x = v;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const uint32_t y = __shfl_up(x, d, 64);
            if (lane >= uint32_t(d)) x += y;
        }
        tot = __shfl(x, 63, 64);
        return x - v;
    }
    __device__ uint64_t reduce_add64(uint64_t v) {
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
        return v;
    }
    __device__ uint32_t reduce_max(uint32_t v) {
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) {
            const uint32_t y = __shfl_xor(v, d, 64);
            v = y > v ? y : v;
        }
        return v;
    }
    __device__ bool any(bool p) { return __any(p); }
    __device__ uint32_t xchg(uint32_t *p, uint32_t v) { return atomicExch(p, v); }
    __device__ int32_t cas(int32_t *p, int32_t c, int32_t v) { return atomicCAS(p, c, v); }
    __device__ unsigned long long cas64(unsigned long long *p, unsigned long long c, unsigned long long v) {
        return atomicCAS(p, c, v);
    }
    __device__ void add(uint32_t *p, uint32_t v) { atomicAdd(p, v); }
    __device__ void umax(uint32_t *p, uint32_t v) { atomicMax(p, v); }
};

__device__ __forceinline__ uint32_t rl(uint32_t v, uint32_t k) {
    return uint32_t(__builtin_amdgcn_readlane(int32_t(v), int32_t(k)));
}
__device__ __forceinline__ int32_t rl(int32_t v, uint32_t k) { return __builtin_amdgcn_readlane(v, int32_t(k)); }

// Job J's graph in the packed layout (rj_pack_layout) at o, by threads t0,
// t0 + nt, ... of the caller.
__device__ void rj_pack_job(const RJob &J, uint8_t *o, uint32_t t0, uint32_t nt) {
    const uint32_t nn = J.nn, ne = J.ne, nr = J.nreads, np = J.npath;
    const RPackLayout L = rj_pack_layout(nn, ne, nr, np);
    for (uint32_t i = t0; i < nn; i += nt) {
        o[L.base + i] = J.base[i];
        reinterpret_cast<uint32_t *>(o + L.parent)[i] = J.parent[i];
        reinterpret_cast<uint32_t *>(o + L.rank)[i] = J.rank[i];
        reinterpret_cast<uint32_t *>(o + L.cover)[i] = J.cover[i];
        reinterpret_cast<uint32_t *>(o + L.out_head)[i] = J.out_head[i];
        reinterpret_cast<uint32_t *>(o + L.in_head)[i] = J.in_head[i];
    }
    for (uint32_t i = t0; i < ne; i += nt) reinterpret_cast<REdge *>(o + L.edges)[i] = J.edges[i];
    for (uint32_t i = t0; i <= nr; i += nt)
        reinterpret_cast<uint32_t *>(o + L.path_off)[i] = i <= J.cur ? J.path_off[i] : 0u;
    for (uint32_t i = t0; i < np; i += nt) reinterpret_cast<uint32_t *>(o + L.paths)[i] = J.paths[i];
}

// The serial head and tail passes on one wave, chosen over the generic
// forms of resident.hpp by overload resolution. Rows go in blocks of 64:
// lane k first loads row B + k's list offsets and up to kRegEntries of its
// entries into registers; then the rows are walked in order on uniform
// values, each entry and each recent row's distance and support taken from
// a lane register by readlane (the current and the previous block are held
// that way), so a row costs scalar work instead of LDS round trips. Rows
// further back come from the arrays, written once per block.
constexpr uint32_t kRegEntries = 4;

__device__ bool rj_head_pass(WaveTeam &t, uint32_t n, const RowIn *rows, const PredE *preds, int32_t *head,
                             uint32_t *support, const RWin &) {
    const uint32_t lane = t.lane;
    bool unreachable = false;
    int32_t hp = kHeadNone, hc = kHeadNone;
    uint32_t sp = 0, sc = 0;
    for (uint32_t B = 0; B < n; B += 64) {
        const uint32_t i = B + lane;
        uint32_t p0 = 0, np = 0, er[kRegEntries], ew[kRegEntries];
        if (i < n) {
            p0 = rows[i].pred_beg;
            np = rows[i + 1].pred_beg - p0;
        }
#pragma unroll
        for (uint32_t a = 0; a < kRegEntries; a++) {
            er[a] = 0;
            ew[a] = 0;
            if (a < np) {
                const PredE pe = preds[p0 + a];
                er[a] = pe.row;
                ew[a] = pe.weight;
            }
        }
        hc = kHeadNone;
        sc = 0;
        const uint32_t cnt = n - B < 64 ? n - B : 64;
        for (uint32_t k = 0; k < cnt; k++) {
            int32_t hnew = 0;
            uint32_t snew = 0;
            if (B + k > 0) {
                const uint32_t npk = rl(np, k);
                bool have = false;
                uint32_t bwt = 0, sb = 0;
                int32_t hb = 0;
                auto take = [&](uint32_t pr, uint32_t w) {
                    int32_t h;
                    uint32_t s;
                    if (pr >= B) {
                        h = rl(hc, pr - B);
                        s = rl(sc, pr - B);
                    } else if (pr + 64 >= B) {
                        h = rl(hp, pr + 64 - B);
                        s = rl(sp, pr + 64 - B);
                    } else {
                        h = head[pr];
                        s = support[pr];
                    }
                    if (h == kHeadNone) {
                        unreachable = true;
                        return;
                    }
                    if (!have || w > bwt || (w == bwt && s > sb)) {
                        have = true;
                        hb = h;
                        sb = s;
                        bwt = w;
                    }
                };
#pragma unroll
                for (uint32_t a = 0; a < kRegEntries; a++)
                    if (a < npk) take(rl(er[a], k), rl(ew[a], k));
                if (npk > kRegEntries) {
                    const uint32_t pk = rl(p0, k);
                    for (uint32_t a = kRegEntries; a < npk; a++) {
                        const PredE pe = preds[pk + a];
                        take(pe.row, pe.weight);
                    }
                }
                if (have) {
                    hnew = hb + 1;
                    snew = sb + bwt;
                } else {
                    hnew = kHeadNone;
                    snew = 0;
                }
            }
            if (lane == k) {
                hc = hnew;
                sc = snew;
            }
        }
        if (i < n) {
            head[i] = hc;
            support[i] = sc;
        }
        hp = hc;
        sp = sc;
        t.sync();
    }
    return !unreachable;
}

__device__ void rj_tail_pass(WaveTeam &t, uint32_t n, const uint32_t *sbeg, const RSucc *succ, int32_t *tail,
                             uint32_t *support, const RWin &) {
    const uint32_t lane = t.lane;
    if (lane == 0) tail[n - 1] = 1;
    t.sync();
    int32_t tn = kHeadNone, tc = kHeadNone;
    uint32_t sn = 0, sc = 0;
    // Blocks [B, B + 64) from the top; rows in decreasing order.
    const uint32_t nb = (n + 63) / 64;
    for (uint32_t bi = nb; bi-- > 0;) {
        const uint32_t B = bi * 64;
        const uint32_t i = B + lane;
        uint32_t s0 = 0, ns = 0, er[kRegEntries], ew[kRegEntries];
        tc = kHeadNone;
        sc = 0;
        if (i < n) {
            s0 = sbeg[i];
            ns = sbeg[i + 1] - s0;
            tc = tail[i];
            sc = support[i];
        }
#pragma unroll
        for (uint32_t a = 0; a < kRegEntries; a++) {
            er[a] = 0;
            ew[a] = 0;
            if (a < ns) {
                const RSucc v = succ[s0 + a];
                er[a] = v.row;
                ew[a] = v.w;
            }
        }
        const uint32_t top = n - 1 < B + 64 ? n - 1 : B + 64;  // first row to compute is top - 1
        for (uint32_t ii = top; ii-- > B;) {
            const uint32_t k = ii - B;
            const uint32_t nsk = rl(ns, k);
            bool have = false;
            uint32_t bwt = 0, sb = 0;
            int32_t tb = 0;
            auto take = [&](uint32_t row, uint32_t w) {
                int32_t tl;
                uint32_t s;
                if (row < B + 64) {
                    tl = rl(tc, row - B);
                    s = rl(sc, row - B);
                } else if (row < B + 128) {
                    tl = rl(tn, row - B - 64);
                    s = rl(sn, row - B - 64);
                } else {
                    tl = tail[row];
                    s = support[row];
                }
                if (tl == kHeadNone) return;
                if (!have || w > bwt || (w == bwt && s > sb)) {
                    have = true;
                    tb = tl;
                    sb = s;
                    bwt = w;
                }
            };
#pragma unroll
            for (uint32_t a = 0; a < kRegEntries; a++)
                if (a < nsk) take(rl(er[a], k), rl(ew[a], k));
            if (nsk > kRegEntries) {
                const uint32_t sk = rl(s0, k);
                for (uint32_t a = kRegEntries; a < nsk; a++) {
                    const RSucc v = succ[sk + a];
                    take(v.row, v.w);
                }
            }
            if (have && lane == k) {
                tc = tb + 1;
                sc = sb + bwt;
            }
        }
        if (i < n) {
            tail[i] = tc;
            support[i] = sc;
        }
        tn = tc;
        sn = sc;
        t.sync();
    }
}

}  // namespace

}  // namespace mpoa
