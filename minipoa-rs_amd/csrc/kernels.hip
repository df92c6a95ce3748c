// gfx950 kernels for the affine-gap sequence-to-DAG DP (minipoa-rs
// src/align.rs:587-1347), integer max-plus only: no MFMA.
//
// Fill: one wave64 per job walks the rows in rank order. A row's band is
// computed on the device (static, full or adaptive), then the wave covers it
// in segments of 64*C columns, lane l owning C contiguous columns. Per lane:
// predecessor max-plus (M <- max M_p[j-1], D <- max(D_p+e, M_p+o)), profile
// add, and a serial in-lane insertion prefix; one Kogge-Stone max-plus scan
// across the 64 lanes resolves the insertion carries exactly (integer
// max-plus is associative, so any bracketing equals the serial recurrence
// I[j+1] = max(I[j]+e, md[j]+o) that the reference pins in
// src/align.rs:1568-1601). Adaptive bands take a wave first-argmax of M and
// lane 0 updates the per-row band state (src/align.rs:829-897).
//
// Traceback: one lane per job replays the weighted deterministic traceback
// (src/align.rs:1164-1313) over the stored M/D/I and writes, per query
// column, the row of its diagonal step or an insertion mark.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "dp_types.hpp"
#include "kernels.hpp"

namespace mpoa {

template <typename T> struct Lim;
template <> struct Lim<int32_t> {
    static constexpr int32_t neg = kNeg32;
    // "No contribution" for the scan; stays below every guarded state and
    // above INT32_MIN after the bounded additions of one row.
    static constexpr int32_t ninf = -(1 << 30) - (1 << 29) - (1 << 28);
};
template <> struct Lim<int64_t> {
    static constexpr int64_t neg = kNeg64;
    static constexpr int64_t ninf = -(int64_t(1) << 61);
};

template <typename T> __device__ __forceinline__ T tmax(T a, T b) { return a > b ? a : b; }
template <typename T> __device__ __forceinline__ T tmin(T a, T b) { return a < b ? a : b; }

constexpr uint32_t kMatch = 1, kDelete = 2, kInsert = 4, kAll = 7;

template <typename T, int C>
__global__ __launch_bounds__(64) void poa_fill(DevPools P) {
    const DevJob J = P.jobs[blockIdx.x];
    const int lane = threadIdx.x;
    const uint32_t n = J.n, m = J.m;
    const RowIn *rin = P.rows + J.rows;
    const RowAux *aux = P.aux + J.rows;
    const PredE *pr = P.preds + J.preds;
    const uint32_t *sc = P.succs + J.succs;
    const uint8_t *seq = P.seq + J.seq;
    const int32_t *mat = P.matrix + J.matrix;
    T *arena = reinterpret_cast<T *>(P.arena) + J.arena;
    RowMeta *meta = P.meta + J.rowmeta;
    const bool adaptive = (J.flags & kAdaptive) != 0;
    const bool full = (J.flags & kFull) != 0;
    int64_t *peak = P.adapt + J.adapt;
    int64_t *streak = peak + n, *pkl = peak + 2 * size_t(n), *pkr = peak + 3 * size_t(n);
    int64_t *ofl = peak + 4 * size_t(n), *ofr = peak + 5 * size_t(n), *extra = peak + 6 * size_t(n);
    const T e = T(J.e), o = T(J.o);
    const T a = T(C) * e;
    constexpr T NEG = Lim<T>::neg, NINF = Lim<T>::ninf;
    const int64_t mm = int64_t(m);

    if (adaptive) {
        for (uint32_t r = lane; r < n; r += 64) {
            peak[r] = 0;
            streak[r] = 0;
            pkl[r] = mm;
            pkr[r] = 0;
            ofl[r] = 0;
            ofr[r] = 0;
            extra[r] = 0;
        }
        __syncthreads();
    }

    uint64_t cursor = 0, cells = 0, pcells = 0;
    int32_t status = kOk;
    for (uint32_t i = 0; i < n; i++) {
        const RowIn ri = rin[i];
        if (ri.head == kHeadNone || ri.tail == kHeadNone) {
            if (lane == 0) meta[i] = RowMeta{0, 0, 0};
            continue;
        }
        // Row band (src/align.rs:711-734).
        int64_t left, right;
        const int64_t h = ri.head, t = ri.tail;
        if (full) {
            left = 0;
            right = mm;
        } else if (adaptive) {
            int64_t lo = tmin(tmin(pkl[i], h + ofl[i]), mm - t);
            int64_t hi = tmax(tmax(pkr[i], h + ofr[i]), mm - t);
            left = tmax<int64_t>(lo - J.band - extra[i], 0);
            right = tmin<int64_t>(hi + J.band + extra[i], mm);
        } else {
            left = tmax<int64_t>(tmin(h, mm - t) - J.band, 0);
            right = tmin<int64_t>(tmax(h, mm - t) + J.band, mm);
        }
        if (left > right || right < 0) {
            status = kBandNoIntersect;
            break;
        }
        const uint32_t start = uint32_t(left) / 8 * 8;
        const uint32_t stop = (uint32_t(right) / 8 + 1) * 8;
        const uint32_t len = stop - start;
        if (cursor + 3ull * len > J.arena_cap) {
            status = kArenaOverflow;
            break;
        }
        const uint64_t off = cursor;
        cursor += 3ull * len;
        cells += len;
        if (lane == 0) meta[i] = RowMeta{start, len, off};
        T *Mr = arena + off;
        T *Dr = Mr + len;
        T *Ir = Dr + len;

        if (i == 0) {
            // Source row (src/align.rs:737-750): M[0] = score(N,N), then
            // insertions only; D stays NEG.
            if (start != 0) {
                status = kBandExcludesSource;
                break;
            }
            if (lane == 0) {
                T mprev = T(J.score_nn), iprev = NEG;
                Mr[0] = mprev;
                Dr[0] = NEG;
                Ir[0] = NEG;
                for (uint32_t j = 1; j < len; j++) {
                    T ins = tmax(iprev + e, mprev + o);
                    T mj = tmax(NEG, ins);
                    Mr[j] = mj;
                    Dr[j] = NEG;
                    Ir[j] = ins;
                    iprev = ins;
                    mprev = mj;
                }
            }
            __syncthreads();
            continue;
        }

        const uint32_t sbase = (i == n - 1) ? J.unknown : aux[i].base;
        const int32_t *mrow = mat + size_t(sbase) * J.alpha;
        const uint32_t pb = ri.pred_beg, pe = rin[i + 1].pred_beg;
        T carry = NEG;
        T best = NINF;
        uint32_t bestc = 0xFFFFFFFFu;
        for (uint32_t seg = start; seg < stop; seg += 64u * C) {
            const uint32_t c0 = seg + uint32_t(lane) * C;
            T Mv[C], Dv[C];
#pragma unroll
            for (int k = 0; k < C; k++) {
                Mv[k] = NEG;
                Dv[k] = NEG;
            }
            // Predecessor transitions on the band overlap (src/align.rs:751-797);
            // the diagonal into the predecessor's first column reads NEG.
            for (uint32_t q = pb; q < pe; q++) {
                const PredE pq = pr[q];
                const RowMeta pm = meta[pq.row];
                const uint32_t ovs = start > pm.start ? start : pm.start;
                const uint32_t ove = stop < pm.start + pm.len ? stop : pm.start + pm.len;
                if (ovs >= ove) continue;
                if (seg == start) pcells += ove - ovs;
                const T *PM = arena + pm.off;
                const T *PD = PM + pm.len;
#pragma unroll
                for (int k = 0; k < C; k++) {
                    const uint32_t col = c0 + k;
                    if (col >= ovs && col < ove) {
                        const uint32_t rel = col - pm.start;
                        const T dg = rel > 0 ? PM[rel - 1] : NEG;
                        const T pmv = PM[rel];
                        const T pdv = PD[rel];
                        Mv[k] = tmax(Mv[k], dg);
                        Dv[k] = tmax(Dv[k], tmax(pdv + e, pmv + o));
                    }
                }
            }
            // Profile + in-lane insertion prefix without carry.
            T md[C], Lk[C];
            T L = NINF;
#pragma unroll
            for (int k = 0; k < C; k++) {
                const uint32_t col = c0 + k;
                T v = NINF;
                if (col < stop) v = tmax(Mv[k] + T(mrow[seq[col]]), Dv[k]);
                md[k] = v;
                Lk[k] = L;
                L = tmax(L + e, v + o);
            }
            // Wave inclusive scan S_l = max_{j<=l} (L_j + (l-j)*C*e).
            T S = L;
#pragma unroll
            for (int d = 1; d < 64; d <<= 1) {
                const T tv = __shfl_up(S, d, 64);
                if (lane >= d) S = tmax(S, tv + T(d) * a);
            }
            const T Sp = __shfl_up(S, 1, 64);
            T X = carry + T(lane) * a;
            if (lane > 0) X = tmax(X, Sp);
#pragma unroll
            for (int k = 0; k < C; k++) {
                const uint32_t col = c0 + k;
                if (col < stop) {
                    const T Ik = tmax(X + T(k) * e, Lk[k]);
                    const T Mk = tmax(md[k], Ik);
                    const uint32_t rel = col - start;
                    Mr[rel] = Mk;
                    Dr[rel] = Dv[k];
                    Ir[rel] = Ik;
                    if (Mk > best) {
                        best = Mk;
                        bestc = col;
                    }
                }
            }
            const T cout = tmax(X + a, L);
            carry = __shfl(cout, 63, 64);
        }
        __syncthreads();

        if (adaptive) {
            // First argmax of M over the whole row (src/align.rs:830-844).
#pragma unroll
            for (int d = 32; d >= 1; d >>= 1) {
                const T ob = __shfl_xor(best, d, 64);
                const uint32_t oc = __shfl_xor(bestc, d, 64);
                if (ob > best || (ob == best && oc < bestc)) {
                    best = ob;
                    bestc = oc;
                }
            }
            if (lane == 0) {
                const int64_t maximum = int64_t(bestc);
                peak[i] = maximum;
                int64_t max_streak = 0;
                bool is_match = false;
                if (aux[i].base == seq[bestc]) {
                    for (uint32_t q = pb; q < pe; q++) {
                        const uint32_t p = pr[q].row;
                        max_streak = tmax(max_streak, streak[p]);
                        if (peak[p] + 1 == maximum) {
                            streak[i] = tmax(streak[i], streak[p] + 1);
                            is_match = true;
                        }
                    }
                }
                if (!is_match && max_streak < 30)
                    extra[i] = tmin(extra[i] + max_streak + 1, 5 * J.band);
                const uint32_t sb = ri.succ_beg, se = rin[i + 1].succ_beg;
                if (streak[i] >= 30) {
                    ofl[i] = maximum - h;
                    ofr[i] = maximum - h;
                    extra[i] = 0;
                    for (uint32_t q = sb; q < se; q++) {
                        ofl[sc[q]] = ofl[i];
                        ofr[sc[q]] = ofr[i];
                    }
                }
                for (uint32_t q = sb; q < se; q++) {
                    const uint32_t s = sc[q];
                    ofl[s] = tmin(ofl[s], ofl[i]);
                    ofr[s] = tmax(ofr[s], ofr[i]);
                    pkl[s] = tmin(pkl[s], maximum);
                    pkr[s] = tmax(pkr[s], maximum + 1);
                    extra[s] = extra[i] != 0 ? tmax(extra[s], extra[i]) : 0;
                }
            }
            __syncthreads();
        }
    }
    if (lane == 0) {
        DevResult R{};
        R.status = status;
        R.cells = cells;
        R.arena_used = cursor;
        R.pred_cells = pcells;
        R.score = int64_t(NEG);
        if (status == kOk) {
            const RowMeta rm = meta[n - 1];
            if (m - 1 >= rm.start && m - 1 - rm.start < rm.len)
                R.score = int64_t(arena[rm.off + (m - 1 - rm.start)]);
        }
        P.res[blockIdx.x] = R;
    }
}

// Weighted deterministic traceback (src/align.rs:1164-1313), one lane per job.
template <typename T>
__global__ __launch_bounds__(64) void poa_traceback(DevPools P, uint32_t njobs) {
    const uint32_t jid = blockIdx.x * 64u + threadIdx.x;
    if (jid >= njobs) return;
    DevResult *R = P.res + jid;
    if (R->status != kOk) return;
    const DevJob J = P.jobs[jid];
    const uint32_t n = J.n, m = J.m;
    const RowIn *rin = P.rows + J.rows;
    const RowAux *aux = P.aux + J.rows;
    const PredE *pr = P.preds + J.preds;
    const uint8_t *seq = P.seq + J.seq;
    const int32_t *mat = P.matrix + J.matrix;
    const T *arena = reinterpret_cast<const T *>(P.arena) + J.arena;
    const RowMeta *meta = P.meta + J.rowmeta;
    uint32_t *tb = P.tb + J.tb;
    const int64_t o = J.o, e = J.e;
    const int64_t half = int64_t(Lim<T>::neg) / 2;
    const bool gap0 = (J.flags & kGapOpenZero) != 0;

    auto inside = [](const RowMeta &r, uint32_t c) { return c >= r.start && c - r.start < r.len; };
    uint32_t i = n - 1, j = m - 1, op = kAll;
    uint64_t steps = 0;
    int32_t status = kOk;
    while (i > 0 || j > 0) {
        if (++steps > uint64_t(n) + m) {
            status = kTbNoAdvance;
            break;
        }
        const RowMeta rm = meta[i];
        if (!inside(rm, j)) {
            status = kTbLeftBand;
            break;
        }
        const T *Mr = arena + rm.off;
        const T *Dr = Mr + rm.len;
        const T *Ir = Dr + rm.len;
        const int64_t Mij = int64_t(Mr[j - rm.start]);
        if (Mij <= half) {
            status = kTbLeftBand;
            break;
        }
        const int64_t Dij = int64_t(Dr[j - rm.start]);
        const uint32_t base = (i == n - 1) ? J.unknown : aux[i].base;
        const int64_t score = mat[size_t(base) * J.alpha + seq[j]];
        const uint32_t pb = rin[i].pred_beg, pe = rin[i + 1].pred_beg;
        // `choose` (src/align.rs:1201-1225): heaviest in-range predecessor
        // (strict >, first wins) whose state plus penalty equals the target.
        auto choose = [&](uint32_t state, int64_t penalty, uint32_t pj) -> int64_t {
            const int64_t target = state == kMatch ? Mij : Dij;
            int64_t bestq = -1;
            uint32_t w = 0;
            for (uint32_t q = pb; q < pe; q++) {
                const PredE pq = pr[q];
                const RowMeta pm = meta[pq.row];
                if (!inside(pm, pj)) continue;
                const T *PM = arena + pm.off;
                const int64_t src = state == kDelete ? int64_t(PM[pm.len + pj - pm.start])
                                                     : int64_t(PM[pj - pm.start]);
                if (target == src + penalty && (bestq < 0 || pq.weight > w)) {
                    bestq = q;
                    w = pq.weight;
                }
            }
            return bestq;
        };
        int64_t pref = -1;
        if ((op & kMatch) && j > 0 && base == seq[j]) pref = choose(kMatch, score, j - 1);
        if (pref >= 0 && pr[pref].weight < aux[i].thresh) pref = -1;
        if (pref >= 0) {
            tb[j] = i;
            i = pr[pref].row;
            j -= 1;
            op = kAll;
            continue;
        }
        if ((op & kDelete) && (op == kDelete || Mij == Dij)) {
            int64_t k = choose(kInsert, o, j);
            if (k >= 0) {
                i = pr[k].row;
                op = gap0 ? kAll : (kMatch | kInsert);
                continue;
            }
            k = choose(kDelete, e, j);
            if (k >= 0) {
                i = pr[k].row;
                op = kDelete;
                continue;
            }
        }
        if ((op & kInsert) && j > 0 && inside(rm, j - 1)) {
            const int64_t Iij = int64_t(Ir[j - rm.start]);
            if (op == kInsert || Mij == Iij) {
                if (Iij == int64_t(Mr[j - 1 - rm.start]) + o) {
                    tb[j] = kTbInsert;
                    j -= 1;
                    op = gap0 ? kAll : (kMatch | kDelete);
                    continue;
                }
                if (Iij == int64_t(Ir[j - 1 - rm.start]) + e) {
                    tb[j] = kTbInsert;
                    j -= 1;
                    op = kInsert;
                    continue;
                }
            }
        }
        if ((op & kMatch) && j > 0) {
            const int64_t k = choose(kMatch, score, j - 1);
            if (k >= 0) {
                tb[j] = i;
                i = pr[k].row;
                j -= 1;
                op = kAll;
                continue;
            }
        }
        status = kTbNoPredecessor;
        break;
    }
    if (status != kOk) {
        R->status = status;
        R->err_i = i;
        R->err_j = j;
    }
}

template <typename T, int C> static void launch_fill_t(const DevPools &P, uint32_t njobs, hipStream_t s) {
    hipLaunchKernelGGL((poa_fill<T, C>), dim3(njobs), dim3(64), 0, s, P);
}

template <typename T> static void launch_fill_c(const DevPools &P, uint32_t njobs, int c, hipStream_t s) {
    switch (c) {
        case 1: launch_fill_t<T, 1>(P, njobs, s); break;
        case 2: launch_fill_t<T, 2>(P, njobs, s); break;
        case 4: launch_fill_t<T, 4>(P, njobs, s); break;
        case 6: launch_fill_t<T, 6>(P, njobs, s); break;
        case 8: launch_fill_t<T, 8>(P, njobs, s); break;
        case 12: launch_fill_t<T, 12>(P, njobs, s); break;
        default: launch_fill_t<T, 16>(P, njobs, s); break;
    }
}

int lanes_columns_for_width(uint32_t width) {
    static const int cs[] = {1, 2, 4, 6, 8, 12, 16};
    for (int c : cs)
        if (uint32_t(64 * c) >= width) return c;
    return 16;
}

void launch_fill(const DevPools &P, uint32_t njobs, bool i64, int cols, hipStream_t s) {
    if (njobs == 0) return;
    if (i64)
        launch_fill_c<int64_t>(P, njobs, cols, s);
    else
        launch_fill_c<int32_t>(P, njobs, cols, s);
}

void launch_traceback(const DevPools &P, uint32_t njobs, bool i64, hipStream_t s) {
    if (njobs == 0) return;
    dim3 grid((njobs + 63) / 64);
    if (i64)
        hipLaunchKernelGGL((poa_traceback<int64_t>), grid, dim3(64), 0, s, P, njobs);
    else
        hipLaunchKernelGGL((poa_traceback<int32_t>), grid, dim3(64), 0, s, P, njobs);
}

}  // namespace mpoa
