// gfx950 kernels for the affine-gap sequence-to-DAG DP (minipoa-rs
// src/align.rs:587-1347), integer max-plus only: no MFMA.
//
// Fill: one wave64 per job walks the rows in rank order; a row is a serial
// dependency of its predecessors, so the design goal is a short per-row
// critical path with no global-memory round trip on it:
//  * recent rows' M and D stay in a per-wave LDS byte ring; predecessors
//    (85% within 4 ranks, all but a few within 16) read from there, older or
//    evicted rows from the global arena;
//  * row descriptors and predecessor lists arrive in 64-row blocks, loaded
//    coalesced into LDS;
//  * the adaptive band state is pulled at each row from its predecessors'
//    final per-row records, folded in rank order (the order in which the
//    reference pushes them to successors, src/align.rs:866-896), so no row
//    writes another row's state;
//  * the row is covered in segments of 64*C columns, lane l owning C
//    contiguous columns; the insertion recurrence I[j+1] = max(I[j]+e,
//    md[j]+o) runs serially inside a lane and across lanes as a DPP max-plus
//    Kogge-Stone scan (integer max-plus is associative, so this equals the
//    serial recurrence exactly, as src/align.rs:1568-1601 requires);
//  * the first-argmax of M for adaptive bands is two DPP reductions.
// What leaves the chip per cell is a traceback record of 1-8 bytes (the
// cell's flags and the `choose` winners among its predecessors, resolved
// while they are folded in); M and D go to HBM only for rows read by a row
// other than the next.
//
// Traceback: one wave per job replays the weighted deterministic traceback
// (src/align.rs:1164-1313) over those records and writes, per query column,
// the row of its diagonal step or an insertion mark.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <type_traits>

#include "dp_types.hpp"
#include "kernels.hpp"
#include "resident_dev.hpp"

namespace mpoa {

template <typename T> struct Lim;
template <> struct Lim<int32_t> {
    static constexpr int32_t neg = kNeg32;
    // "No contribution" for scans: below every guarded state, and above
    // INT32_MIN after the bounded additions of one row.
    static constexpr int32_t ninf = -(1 << 30) - (1 << 29) - (1 << 28);
};
template <> struct Lim<int64_t> {
    static constexpr int64_t neg = kNeg64;
    static constexpr int64_t ninf = -(int64_t(1) << 61);
};

// Global data is read through address_space(1) pointers, so the compiler
// never merges an LDS load and a global load into one generic (flat) load:
// gfx950's backend rejects a select between an LDS and a global generic
// pointer. ldg() copies a record of 4-byte words.
#define AS_GLOBAL __attribute__((address_space(1)))
template <typename T> using gptr = T AS_GLOBAL *;
template <typename T> __device__ __forceinline__ gptr<T> glb(T *p) { return (gptr<T>)p; }
template <typename T> __device__ __forceinline__ T ldg(const T *p) {
    static_assert(sizeof(T) % 4 == 0, "records are whole 4-byte words");
    const gptr<const uint32_t> w = glb(reinterpret_cast<const uint32_t *>(p));
    uint32_t t[sizeof(T) / 4];
#pragma unroll
    for (unsigned k = 0; k < sizeof(T) / 4; k++) t[k] = w[k];
    T v;
    __builtin_memcpy(&v, t, sizeof(T));
    return v;
}

template <typename T> __device__ __forceinline__ T tmax(T a, T b) { return a > b ? a : b; }
template <typename T> __device__ __forceinline__ T tmin(T a, T b) { return a < b ? a : b; }

constexpr uint32_t kMatch = 1, kDelete = 2, kInsert = 4, kAll = 7;

// s_waitcnt immediate (GFX9 layout: vmcnt [3:0]+[15:14], expcnt [6:4],
// lgkmcnt [11:8]) that waits for vmcnt == 0 only.
constexpr int kWaitVmcnt0 = 0x0F70;

// DPP controls (GFX9 encoding).
constexpr int kRowShr1 = 0x111, kRowShr2 = 0x112, kRowShr4 = 0x114, kRowShr8 = 0x118;
constexpr int kBcast15 = 0x142, kBcast31 = 0x143, kWaveShr1 = 0x138;

// DPP move with `old` in lanes whose source is out of range or masked off.
template <int CTRL, int ROWMASK>
__device__ __forceinline__ int32_t dpp(int32_t old, int32_t v) {
    return __builtin_amdgcn_update_dpp(old, v, CTRL, ROWMASK, 0xF, false);
}
template <int CTRL, int ROWMASK>
__device__ __forceinline__ uint32_t dpp(uint32_t old, uint32_t v) {
    return uint32_t(__builtin_amdgcn_update_dpp(int32_t(old), int32_t(v), CTRL, ROWMASK, 0xF, false));
}
template <int CTRL, int ROWMASK>
__device__ __forceinline__ int64_t dpp(int64_t old, int64_t v) {
    const int32_t lo = __builtin_amdgcn_update_dpp(int32_t(uint64_t(old)), int32_t(uint64_t(v)),
                                                   CTRL, ROWMASK, 0xF, false);
    const int32_t hi = __builtin_amdgcn_update_dpp(int32_t(uint64_t(old) >> 32),
                                                   int32_t(uint64_t(v) >> 32), CTRL, ROWMASK,
                                                   0xF, false);
    return int64_t((uint64_t(uint32_t(hi)) << 32) | uint32_t(lo));
}

__device__ __forceinline__ int32_t lane63(int32_t v) { return __builtin_amdgcn_readlane(v, 63); }
__device__ __forceinline__ uint32_t lane63(uint32_t v) {
    return uint32_t(__builtin_amdgcn_readlane(int32_t(v), 63));
}
__device__ __forceinline__ int64_t lane63(int64_t v) {
    const int32_t lo = __builtin_amdgcn_readlane(int32_t(uint64_t(v)), 63);
    const int32_t hi = __builtin_amdgcn_readlane(int32_t(uint64_t(v) >> 32), 63);
    return int64_t((uint64_t(uint32_t(hi)) << 32) | uint32_t(lo));
}

// Inclusive max-plus scan S_l = max_{j<=l} (v_j + (l-j)*a) over the wave.
// d15/d31 = ((l&15)+1)*a and ((l&31)+1)*a: the decay from the broadcast lane.
template <typename T>
__device__ __forceinline__ T maxplus_scan(T v, T a, T d15, T d31) {
    constexpr T NI = Lim<T>::ninf;
    v = tmax(v, dpp<kRowShr1, 0xF>(NI, v) + a);
    v = tmax(v, dpp<kRowShr2, 0xF>(NI, v) + 2 * a);
    v = tmax(v, dpp<kRowShr4, 0xF>(NI, v) + 4 * a);
    v = tmax(v, dpp<kRowShr8, 0xF>(NI, v) + 8 * a);
    v = tmax(v, dpp<kBcast15, 0xA>(NI, v) + d15);
    v = tmax(v, dpp<kBcast31, 0xC>(NI, v) + d31);
    return v;
}

// Wave max / min reductions; the result is uniform.
template <typename T> __device__ __forceinline__ T wave_max(T v, T lo) {
    v = tmax(v, dpp<kRowShr1, 0xF>(lo, v));
    v = tmax(v, dpp<kRowShr2, 0xF>(lo, v));
    v = tmax(v, dpp<kRowShr4, 0xF>(lo, v));
    v = tmax(v, dpp<kRowShr8, 0xF>(lo, v));
    v = tmax(v, dpp<kBcast15, 0xA>(lo, v));
    v = tmax(v, dpp<kBcast31, 0xC>(lo, v));
    return lane63(v);
}
__device__ __forceinline__ uint32_t wave_min(uint32_t v) {
    constexpr uint32_t hi = 0xFFFFFFFFu;
    v = tmin(v, dpp<kRowShr1, 0xF>(hi, v));
    v = tmin(v, dpp<kRowShr2, 0xF>(hi, v));
    v = tmin(v, dpp<kRowShr4, 0xF>(hi, v));
    v = tmin(v, dpp<kRowShr8, 0xF>(hi, v));
    v = tmin(v, dpp<kBcast15, 0xA>(hi, v));
    v = tmin(v, dpp<kBcast31, 0xC>(hi, v));
    return lane63(v);
}

// Per-wave LDS budget (static part; the ring, query codes and score matrix
// live in dynamic LDS sized per launch).
constexpr uint32_t kBlk = 64;           // rows per descriptor block
constexpr uint32_t kPredCache = 256;    // predecessor entries cached per block
constexpr uint32_t kHist = kHistRows;   // rows of metadata/records kept in LDS
constexpr uint32_t kNoSlot = 0xFFFFFFFFu;

uint32_t fill_dynamic_lds(uint32_t ring_bytes, uint32_t seq_lds, uint32_t mat_lds) {
    return ring_bytes + ((seq_lds + 15u) & ~15u) + mat_lds;
}

struct LMeta {
    uint32_t start, len;
    uint32_t slot;   // byte offset in the ring, or kNoSlot
    uint32_t wpos;   // ring write position (monotonic bytes) of the slot
    uint64_t voff;   // arena byte offset of the row's kept M and D, or kNoValues
    int64_t mlast;   // M of the row's last column
    uint32_t last_use;  // last row that reads this row (0: none)
};

// Stores a lane's C contiguous values; whole runs use the widest aligned
// vector store (row starts are 8-element aligned, lane runs C elements).
template <typename T, int C>
__device__ __forceinline__ void store_run(T *dst, const T (&v)[C], uint32_t valid) {
    if (valid >= uint32_t(C)) {
        if constexpr (sizeof(T) == 4 && C % 4 == 0) {
#pragma unroll
            for (int k = 0; k < C; k += 4)
                *reinterpret_cast<int4 *>(dst + k) = make_int4(v[k], v[k + 1], v[k + 2], v[k + 3]);
        } else if constexpr (sizeof(T) == 4 && C % 2 == 0) {
#pragma unroll
            for (int k = 0; k < C; k += 2)
                *reinterpret_cast<int2 *>(dst + k) = make_int2(v[k], v[k + 1]);
        } else if constexpr (sizeof(T) == 8 && C % 2 == 0) {
#pragma unroll
            for (int k = 0; k < C; k += 2)
                *reinterpret_cast<longlong2 *>(dst + k) = make_longlong2(v[k], v[k + 1]);
        } else {
#pragma unroll
            for (int k = 0; k < C; k++) dst[k] = v[k];
        }
    } else {
#pragma unroll
        for (int k = 0; k < C; k++)
            if (uint32_t(k) < valid) dst[k] = v[k];
    }
}

// Loads G (4 or 8) consecutive values from a 16-byte aligned address (LDS or
// global, resolved at the call site) with 16-byte vector loads.
template <typename T, int G> __device__ __forceinline__ void loadg(const T *src, T (&v)[G]) {
    static_assert(G == 4 || G == 8, "pieces are half or whole 8-column blocks");
    if constexpr (sizeof(T) == 4) {
#pragma unroll
        for (int k = 0; k < G; k += 4) {
            const int4 x = reinterpret_cast<const int4 *>(src)[k / 4];
            v[k] = x.x;
            v[k + 1] = x.y;
            v[k + 2] = x.z;
            v[k + 3] = x.w;
        }
    } else {
#pragma unroll
        for (int k = 0; k < G; k += 2) {
            const longlong2 x = reinterpret_cast<const longlong2 *>(src)[k / 2];
            v[k] = x.x;
            v[k + 1] = x.y;
        }
    }
}

// The same from global memory (clang vector types: HIP's vector classes
// cannot be copied out of an address_space(1) object).
typedef int32_t v4i32 __attribute__((ext_vector_type(4)));
typedef int64_t v2i64 __attribute__((ext_vector_type(2)));
template <typename T, int G> __device__ __forceinline__ void loadg(gptr<const T> src, T (&v)[G]) {
    if constexpr (sizeof(T) == 4) {
#pragma unroll
        for (int k = 0; k < G; k += 4) {
            const v4i32 x = reinterpret_cast<gptr<const v4i32>>(src)[k / 4];
            v[k] = x.x;
            v[k + 1] = x.y;
            v[k + 2] = x.z;
            v[k + 3] = x.w;
        }
    } else {
#pragma unroll
        for (int k = 0; k < G; k += 2) {
            const v2i64 x = reinterpret_cast<gptr<const v2i64>>(src)[k / 2];
            v[k] = x.x;
            v[k + 1] = x.y;
        }
    }
}

// Workgroup barrier over LDS only: this wave's LDS writes are complete and
// the compiler keeps LDS accesses on their side of the barrier. Global
// stores stay in flight (no vmcnt wait); the rare global reads of data
// another wave wrote use sync_global() first. With one wave the fences
// remain: lanes exchange data through LDS, and without them the compiler
// may move a lane's LDS read above another lane's write.
template <int W> __device__ __forceinline__ void sync_lds() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    if constexpr (W > 1) __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}
// Full workgroup barrier: every wave's global stores are complete and
// visible to the workgroup afterwards.
template <int W> __device__ __forceinline__ void sync_global() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    if constexpr (W > 1) __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// Running maximum of one kind of candidate with its `choose` winner, kept
// as a field of the cell's packed record p. The predecessors are folded in
// list order, which is `choose` order (src/align.rs:1201-1225: the heaviest
// edge, then the first in `incoming`), so among tied candidates the first
// one stays. `mask` is the field, `ins` the candidate's list position
// shifted into it.
template <typename T, typename R>
__device__ __forceinline__ void track(T &v, R &p, T c, R mask, R ins) {
    p = c > v ? ((p & ~mask) | ins) : p;
    v = tmax(v, c);
}

// 1 if a < b, from the sign of a - b: exact while |a - b| < 2^(bits-1),
// which holds for every pair of DP values compared here (they lie within
// the score guard's range around NEG). Plain VALU arithmetic: no compare
// mask is held in scalar registers.
template <typename T> __device__ __forceinline__ uint32_t lt01(T a, T b) {
    using U = typename std::make_unsigned<T>::type;
    return uint32_t((U(a) - U(b)) >> (8 * sizeof(T) - 1));
}
// All ones if a < b, else zero.
template <typename R, typename T> __device__ __forceinline__ R ltmask(T a, T b) {
    return R(0) - R(lt01(a, b));
}

// Stores a lane's C traceback records of rb (1, 2, 4 or 8) bytes each. Row
// records start 16-byte aligned and a lane's run at a multiple of C records,
// so every group of four records is one aligned vector store; `valid` is a
// multiple of four (the 8-column band grid). 8-byte records come only from
// the wide-record kernel (R = uint64_t).
template <int C, typename R>
__device__ __forceinline__ void store_recs(uint8_t *dst, const R (&r)[C], uint32_t rb, uint32_t valid) {
#pragma unroll
    for (int g = 0; g < C; g += 4) {
        if (uint32_t(g) >= valid) break;
        uint8_t *d = dst + size_t(g) * rb;
        if (rb == 1) {
            *reinterpret_cast<uint32_t *>(d) = uint32_t(r[g]) | uint32_t(r[g + 1]) << 8 |
                                               uint32_t(r[g + 2]) << 16 | uint32_t(r[g + 3]) << 24;
        } else if (rb == 2) {
            *reinterpret_cast<uint2 *>(d) = make_uint2(uint32_t(r[g]) | uint32_t(r[g + 1]) << 16,
                                                       uint32_t(r[g + 2]) | uint32_t(r[g + 3]) << 16);
        } else if (sizeof(R) == 4 || rb == 4) {
            *reinterpret_cast<uint4 *>(d) = make_uint4(uint32_t(r[g]), uint32_t(r[g + 1]),
                                                       uint32_t(r[g + 2]), uint32_t(r[g + 3]));
        } else {
            *reinterpret_cast<ulonglong2 *>(d) = make_ulonglong2(r[g], r[g + 1]);
            *reinterpret_cast<ulonglong2 *>(d + 16) = make_ulonglong2(r[g + 2], r[g + 3]);
        }
    }
}

// Fill: a workgroup of W waves per job. The row is covered in stripes of
// 64*C*W columns; in each stripe wave w owns the 64*C columns from
// w*64*C and lane l of it the C columns from l*C. Every wave computes the
// row's band, arena and ring placement redundantly from the same LDS
// state; the waves meet twice per row: to chain the insertion carry
// across waves (per stripe) and to reduce the row's first argmax.
// LT: the score matrix and the query codes are staged in LDS (every launch
// whose alphabet and queries fit); otherwise they are read from global memory.
//
// What a row leaves in HBM depends on the launch's layout (REC):
//  * values (REC false): M, D and I of every cell, [M len][D len][I len]
//    (12 bytes per i32 cell); poa_traceback_values re-evaluates `choose`
//    from them. The fill does the least work per cell.
//  * records (REC true): one traceback record per cell (bit layout in
//    dp_types.hpp) -- for the cell's own M, D and I, which term the next
//    column's I takes, and the `choose` winners of the diagonal, the
//    deletion open and the deletion extension, tracked while the
//    predecessors are folded in -- and, only for a row that a later row
//    other than the next one reads (its last use, from the host) or that
//    has no ring slot, its M and D scores: 1-2 bytes per cell, at about
//    half again the fill's VALU work. The next row always finds its
//    immediate predecessor in the ring: a row whose placement would
//    overwrite the previous row's slot gets no slot (and keeps its scores in
//    HBM if any row reads it).
//
// R is the register type of a packed record: uint32_t when every row has at
// most kMaxNarrowPreds predecessors (records of 1, 2 or 4 bytes), uint64_t
// for the wide-record kernel a launch takes when some row has more.
//
// amdgpu_waves_per_eu(4) for the i32 shapes that fit 128 VGPRs: four waves
// share each SIMD. Rows are latency chains, so throughput comes from the
// number of resident jobs. Values: 8x1, 4x2 and 4x4; records: 4x2 and 4x4
// (8x1 would spill). Other kernels would spill in the row loop under the
// limit (on gfx9 a scratch reload waits for the wave's outstanding stores)
// and keep the default.
template <typename T, int C, int W, bool LT, bool REC> constexpr int fill_waves_per_eu() {
    return sizeof(T) == 4 && LT && (C == 4 || (!REC && C == 8 && W == 1)) ? 4 : 1;
}
// The fill's fixed LDS state, at the start of the job's LDS region; the
// ring, the query codes and the score matrix follow it.
template <typename T, int W> struct FillLds {
    LMeta hmeta[kHist];
    RowRec hrec[kHist];
    RowIn drow[kBlk + 1];
    RowAux daux[kBlk];
    PredE dpred[kPredCache];
    T xcar[2][W];          // per-wave scan totals, by stripe parity
    T xbest[W];            // per-wave row maximum of M
    uint32_t xbestc[W];    // its first column
    uint32_t xbestq[W];    // the query code there
    uint64_t xpool;        // arena chunk taken by thread 0
    int64_t xscore;        // M at (n-1, m-1)
    // The DP's statistics, added by thread 0 (kept out of the registers
    // that the row loop carries).
    unsigned long long xcells, xpcells, xused;
    alignas(16) T negrow[16];  // NEG: reads outside a predecessor's band
};
template <typename T, int W> constexpr uint32_t fill_fixed_lds() {
    return uint32_t((sizeof(FillLds<T, W>) + 15u) & ~size_t(15));
}

// One job's DP (the job at index b of the launch pools) by the W waves of
// the calling workgroup, with its LDS region at `lds` (16-byte aligned,
// fill_fixed_lds + ring + query codes + matrix bytes).
template <typename T, int C, int W, bool LT, typename R, bool REC>
__device__ __forceinline__ void fill_job(const DevPools &P, uint32_t b, char *lds) {
    static_assert(C == 4 || C % 8 == 0, "lane runs tile the 8-column band grid");
    constexpr int G = C < 8 ? C : 8;  // columns per piece (inside one 8-block)
    FillLds<T, W> &FL = *reinterpret_cast<FillLds<T, W> *>(lds);
    auto &hmeta = FL.hmeta;
    auto &hrec = FL.hrec;
    auto &drow = FL.drow;
    auto &daux = FL.daux;
    auto &dpred = FL.dpred;
    auto &xcar = FL.xcar;
    auto &xbest = FL.xbest;
    auto &xbestc = FL.xbestc;
    auto &xbestq = FL.xbestq;
    auto &xpool = FL.xpool;
    auto &xscore = FL.xscore;
    auto &xcells = FL.xcells;
    auto &xpcells = FL.xpcells;
    auto &xused = FL.xused;
    auto &negrow = FL.negrow;
    char *dyn_lds = lds + fill_fixed_lds<T, W>();
    const uint32_t kRingBytes = P.ring_bytes;  // recent rows' M and D
    T *ring = reinterpret_cast<T *>(dyn_lds);  // element-indexed
    uint8_t *sseq = reinterpret_cast<uint8_t *>(dyn_lds + kRingBytes);
    int32_t *smat = reinterpret_cast<int32_t *>(dyn_lds + kRingBytes + ((P.seq_lds + 15u) & ~15u));

    if (P.prio) __builtin_amdgcn_s_setprio(3);
    const DevJob J = ldg(P.jobs + b);
    if (J.n == 0) {
        // No DP for this job in this launch (a resident step loop replays
        // one launch shape over every job of its batch): uniform exit.
        if (threadIdx.x == 0) P.res[b] = DevResult{kSkipped, 0, 0, 0, 0, 0, 0, 0};
        return;
    }
    const uint32_t tid = threadIdx.x;
    const int lane = int(tid & 63u);
    const uint32_t wave = uint32_t(__builtin_amdgcn_readfirstlane(int(tid >> 6)));
    const bool leader = tid == 0;
    const uint32_t n = J.n, m = J.m;
    const RowIn *rin = P.rows + J.rows;
    const RowAux *aux = P.aux + J.aux;
    const PredE *pr = P.preds + J.preds;
    const uint8_t *seq = P.seq + J.seq;
    const int32_t *gmat = P.matrix + J.matrix;
    uint8_t *arena = static_cast<uint8_t *>(P.arena);  // RowMeta offsets are absolute bytes
    RowMeta *meta = P.meta + J.rowmeta;
    RowRec *grec = P.adapt + J.adapt;
    const bool adaptive = (J.flags & kAdaptive) != 0;
    const bool full = (J.flags & kFull) != 0;
    const T e = T(J.e), o = T(J.o);
    const T a = T(C) * e;      // decay over one lane run
    const T a64 = T(64) * a;   // decay over one wave's columns
    const T d15 = T((lane & 15) + 1) * a, d31 = T((lane & 31) + 1) * a;
    constexpr T NEG = Lim<T>::neg, NI = Lim<T>::ninf;
    constexpr T HALF = NEG / 2;
    const uint32_t alpha = J.alpha;
    constexpr bool mat_lds = LT, seq_lds = LT;
    if (tid < 16) negrow[tid] = NEG;
    if (leader) {
        xscore = int64_t(NEG);
        xcells = 0;
        xpcells = 0;
        xused = 0;
    }
    if (mat_lds)
        for (uint32_t k = tid; k < alpha * alpha; k += 64u * W) smat[k] = glb(gmat)[k];
    if (seq_lds)
        for (uint32_t k = tid; k < J.padded; k += 64u * W) sseq[k] = glb(seq)[k];

    uint64_t cur = J.arena, rend = J.arena + J.arena_cap;  // current arena region (bytes)
    uint32_t wpos = 0;      // ring bytes handed out so far
    uint32_t ring_at = 0;   // wpos modulo the ring size
    uint32_t prev_slot = kNoSlot;  // ring slot of row i - 1 (kNoSlot: none)
    uint32_t pb0 = 0, pcnt = 0;
    int32_t status = kOk;
    // Record of the previous row, kept in registers: the leader writes it
    // to LDS after the row's last barrier, so the next row reads it here.
    RowRec last_rec{0, 0, 0, 0, 0};

    for (uint32_t i = 0; i < n; i++) {
        if (i % kBlk == 0) {
            // Descriptor blocks. The global loads also retire this wave's
            // earlier stores, so after the barrier every row before i is in
            // global memory for all waves (the far-row fallbacks rely on it).
            sync_global<W>();
            const uint32_t cnt = tmin(kBlk, n - i);
            for (uint32_t k = tid; k <= cnt; k += 64u * W) drow[k] = ldg(rin + i + k);
            for (uint32_t k = tid; k < cnt; k += 64u * W) daux[k] = ldg(aux + i + k);
            pb0 = ldg(rin + i).pred_beg;
            pcnt = tmin(ldg(rin + i + cnt).pred_beg - pb0, kPredCache);
            for (uint32_t k = tid; k < pcnt; k += 64u * W) dpred[k] = ldg(pr + pb0 + k);
            sync_global<W>();
        }
        const RowIn ri = drow[i % kBlk];
        const uint32_t pb = ri.pred_beg, pe = drow[i % kBlk + 1].pred_beg;
        if (ri.head == kHeadNone || ri.tail == kHeadNone) {
            // Skipped row (src/align.rs:708-710): no band, never pushes state.
            if (leader) {
                hmeta[i % kHist] = LMeta{0, 0, kNoSlot, 0, kNoValues, int64_t(NEG)};
                hrec[i % kHist] = RowRec{0, 0, 0, 0, 0};
                meta[i] = RowMeta{0, 0, 0, kNoValues, int64_t(NEG)};
                if (adaptive) grec[i] = RowRec{0, 0, 0, 0, 0};
            }
            last_rec = RowRec{0, 0, 0, 0, 0};
            prev_slot = kNoSlot;
            sync_lds<W>();
            continue;
        }
        const uint32_t np = pe - pb;
        if (REC && np > (sizeof(R) == 4 ? kMaxNarrowPreds : kMaxRecPreds)) {
            status = kTooManyPreds;
            break;
        }
        // Record layout of the row: flags in bits 0-4, then the diagonal,
        // deletion-open and deletion-extend winner fields of fb bits each.
        const uint32_t rb = rec_bytes(np), fb = rec_field_bits(np);
        const uint32_t ffn = (1u << fb) - 1u;  // "no winner"
        const uint32_t so = 5 + fb, se = 5 + 2 * fb;
        const R Mm = R(ffn) << 5, Mo = R(ffn) << so, Me = R(ffn) << se;

        // Gather the predecessors in parallel, lane q <-> predecessor q (in
        // groups of 64): row, geometry, ring slot, adaptive record, list
        // position and rank permutation. Rows older than the LDS history come
        // from global memory.
        uint32_t g_row = 0, g_rperm = 0;
        LMeta g_meta{0, 0, kNoSlot, 0, kNoValues, 0};
        RowRec g_rec{0, 0, 0, 0, 0};
        auto gather = [&](uint32_t g0) {
            const uint32_t q = g0 + uint32_t(lane);
            g_meta = LMeta{0, 0, kNoSlot, 0, kNoValues, 0};
            g_rec = RowRec{0, 0, 0, 0, 0};
            g_row = 0;
            g_rperm = 0;
            // The global fallbacks below wait for their loads where they are
            // issued: a load still counted at a control-flow join makes the
            // compiler put vmcnt waits on the hot path, and on gfx9 vmcnt
            // also counts the wave's outstanding stores.
            if (q < pe) {
                const uint32_t k = q - pb0;
                PredE pq;
                if (k < pcnt) {
                    pq = dpred[k];
                } else {
                    pq = ldg(pr + q);
                    __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
                }
                g_row = pq.row;
                g_rperm = pq.rperm;
            }
            const bool far = q < pe && i - g_row >= kHist;
            if (q < pe) {
                if (!far) {
                    g_meta = hmeta[g_row % kHist];
                    if (adaptive) g_rec = g_row + 1 == i ? last_rec : hrec[g_row % kHist];
                } else {
                    const RowMeta gm = ldg(meta + g_row);
                    g_meta = LMeta{gm.start, gm.len, kNoSlot, 0, gm.voff, gm.mlast};
                    if (adaptive) g_rec = ldg(grec + g_row);
                    __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
                }
            }
        };
        auto rl32 = [](uint32_t v, uint32_t q) { return uint32_t(__builtin_amdgcn_readlane(int32_t(v), q)); };
        auto rli = [](int32_t v, uint32_t q) { return __builtin_amdgcn_readlane(v, q); };
        auto rl64 = [](int64_t v, uint32_t q) {
            const uint32_t lo = uint32_t(__builtin_amdgcn_readlane(int32_t(uint64_t(v)), q));
            const uint32_t hi = uint32_t(__builtin_amdgcn_readlane(int32_t(uint64_t(v) >> 32), q));
            return int64_t((uint64_t(hi) << 32) | lo);
        };
        if (np <= 64) gather(pb);

        // Adaptive state of row i: the reference's pushes from its computed
        // predecessors (rows >= 1), replayed in rank order.
        int32_t ofl = 0, ofr = 0, pkl = int32_t(m), pkr = 0, extra = 0;
        auto push = [&](const RowRec &r) {
            if (r.streak >= 30) {
                ofl = r.ofl;
                ofr = r.ofr;
            }
            ofl = tmin(ofl, r.ofl);
            ofr = tmax(ofr, r.ofr);
            pkl = tmin(pkl, r.peak);
            pkr = tmax(pkr, r.peak + 1);
            extra = r.extra != 0 ? tmax(extra, r.extra) : 0;
        };
        if (adaptive && np <= 64) {
            for (uint32_t t = 0; t < np; t++) {
                const uint32_t q = rl32(g_rperm, t);
                if (rl32(g_row, q) == 0 || rl32(g_meta.len, q) == 0) continue;
                push(RowRec{rli(g_rec.peak, q), rli(g_rec.streak, q), rli(g_rec.ofl, q),
                            rli(g_rec.ofr, q), rli(g_rec.extra, q)});
            }
        } else if (adaptive) {
            // More than 64 predecessors: one at a time, uniform loads.
            for (uint32_t t = 0; t < np; t++) {
                const uint32_t q = ldg(pr + pb + t).rperm;
                const uint32_t p = ldg(pr + pb + q).row;
                uint32_t plen;
                RowRec r;
                if (i - p < kHist) {
                    plen = hmeta[p % kHist].len;
                    r = p + 1 == i ? last_rec : hrec[p % kHist];
                } else {
                    plen = ldg(meta + p).len;
                    r = ldg(grec + p);
                    __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
                }
                if (p != 0 && plen != 0) push(r);
            }
        }

        // Row band (src/align.rs:711-734), in 32 bits: n + padded <= 2^29,
        // and band + extra saturates at 2^30, past which the row is full.
        const int32_t mi = int32_t(m), hh = ri.head, tt = ri.tail;
        int32_t left, right;
        if (full) {
            left = 0;
            right = mi;
        } else if (adaptive) {
            const int32_t lo = tmin(tmin(pkl, hh + ofl), mi - tt);
            const int32_t hi = tmax(tmax(pkr, hh + ofr), mi - tt);
            const int32_t be = extra > kBandCap - J.band32 ? kBandCap : J.band32 + extra;
            left = tmax(lo - be, 0);
            right = tmin(hi + be, mi);
        } else {
            left = tmax(tmin(hh, mi - tt) - J.band32, 0);
            right = tmin(tmax(hh, mi - tt) + J.band32, mi);
        }
        if (left > right || right < 0) {
            status = kBandNoIntersect;
            break;
        }
        const uint32_t start = uint32_t(left) / 8 * 8;
        const uint32_t stop = (uint32_t(right) / 8 + 1) * 8;
        const uint32_t len = stop - start;

        // Ring slot for this row's M and D: never over the previous row's
        // slot, which this row reads.
        const uint32_t bytes = 2u * len * uint32_t(sizeof(T));
        uint32_t slot = kNoSlot, swpos = 0;
        if (bytes <= kRingBytes / 2) {
            const bool wrap = ring_at + bytes > kRingBytes;
            if (!wrap || prev_slot == kNoSlot || bytes <= prev_slot) {
                if (wrap) {
                    wpos += kRingBytes - ring_at;
                    ring_at = 0;
                }
                slot = ring_at;
                swpos = wpos;
                wpos += bytes;
                ring_at += bytes;
            }
        }
        // Records: M and D go to HBM at once for a row that some later row
        // reads and that has no ring slot, or whose last reader is beyond the
        // LDS history (it reads the row's geometry from global memory); a
        // row read only within the history is copied out of the ring when a
        // later row's slot would overwrite it before its last reader (the
        // spill below). Values: [M][D][I] of every row.
        const bool keep = !REC || (ri.last_use > i && slot == kNoSlot) || ri.last_use >= i + kHist;
        const uint64_t rec_sz = REC ? (uint64_t(len) * rb + 15u) / 16u * 16u : 0u;
        const uint64_t need = rec_sz + (keep ? uint64_t(bytes) : 0u) + (REC ? 0u : uint64_t(bytes) / 2);
        // `want` arena bytes at cur, taking a chunk from the launch's pool
        // when the current region is exhausted (uniform across the
        // workgroup). The leader's atomic result reaches its wave by readlane
        // and other waves through LDS behind a barrier.
        auto take = [&](uint64_t want) -> bool {
            if (cur + want <= rend) return true;
            uint64_t got = 0;
            if (leader) got = atomicAdd(P.pool_next, (unsigned long long)P.chunk);
            if constexpr (W > 1) {
                if (leader) xpool = got;
                sync_lds<W>();
                got = xpool;
                sync_lds<W>();
            } else {
                got = uint64_t(rl64(int64_t(got), 0));
            }
            if (want > P.chunk || got + P.chunk > P.pool_end) return false;
            cur = got;
            rend = got + P.chunk;
            return true;
        };
        if (!take(need)) {
            status = kArenaOverflow;
            break;
        }
        const uint64_t off = cur;
        const uint64_t voff = keep ? off + rec_sz : kNoValues;
        cur += need;
        if (leader) {
            atomicAdd(&xused, (unsigned long long)need);
            atomicAdd(&xcells, (unsigned long long)len);
        }
        uint8_t *Rr = arena + off;
        T *Mk = keep ? reinterpret_cast<T *>(arena + voff) : nullptr;
        T *Dk = keep ? Mk + len : nullptr;
        T *Ik = REC ? nullptr : Mk + 2 * len;
        const uint32_t sm = (slot == kNoSlot ? 0u : slot) / uint32_t(sizeof(T));
        const uint32_t sd = sm + len;
        // The lane owning the row's last column writes its RowMeta and the
        // mlast of this LDS copy once the row is computed; later rows read
        // the copy only after this row's first barrier (mlast after its last).
        if (leader) hmeta[i % kHist] = LMeta{start, len, slot, swpos, voff, int64_t(NEG), ri.last_use};

        if constexpr (REC) {
            // Spill: rows of the history whose ring bytes this row's slot
            // overwrites (the readers' own test: wpos - slot's wpos > ring)
            // and that a row from this one on still reads get M and D
            // copied from the ring to the arena, by every thread. Lane q
            // checks row i - 1 - q; every wave finds the same rows.
            bool spilled = false;
            if (i > 0) {
                const uint32_t q = uint32_t(lane);
                bool sp = false;
                if (q < tmin(i, kHist - 1)) {  // row i - kHist shares row i's entry
                    const LMeta h = hmeta[(i - 1 - q) % kHist];
                    sp = h.len != 0 && h.slot != kNoSlot && h.voff == kNoValues && h.last_use >= i &&
                         wpos - h.wpos > kRingBytes;
                }
                uint64_t mask = __builtin_amdgcn_ballot_w64(sp);
                spilled = mask != 0;
                // Every wave reads the history before the leader updates it
                // below, so all waves spill the same rows (and meet the same
                // barriers in take()).
                if (spilled) sync_lds<W>();
                while (mask) {
                    const uint32_t qq = uint32_t(__builtin_ctzll(mask));
                    mask &= mask - 1;
                    const uint32_t r = i - 1 - qq;
                    const LMeta h = hmeta[r % kHist];
                    const uint64_t sb = 2ull * h.len * sizeof(T);
                    if (!take(sb)) {
                        status = kArenaOverflow;
                        break;
                    }
                    const uint64_t dst = cur;
                    cur += sb;
                    if (leader) atomicAdd(&xused, (unsigned long long)sb);
                    const int4 *src4 = reinterpret_cast<const int4 *>(reinterpret_cast<const char *>(ring) + h.slot);
                    int4 *dst4 = reinterpret_cast<int4 *>(arena + dst);
                    for (uint32_t k = tid; k < uint32_t(sb / 16); k += 64u * W) dst4[k] = src4[k];
                    // Read by the next rows only behind the barrier below.
                    if (leader) hmeta[r % kHist].voff = dst;
                }
                if (status != kOk) break;
            }
            if (spilled) {
                // The copies' ring reads before this row's ring writes, and the
                // new offsets before any wave reads them.
                sync_lds<W>();
                const uint32_t q = pb + uint32_t(lane);
                if (np <= 64 && q < pe && g_meta.len != 0 && i - g_row < kHist)
                    g_meta.voff = hmeta[g_row % kHist].voff;
            }
        }

        if (i == 0) {
            // Source row (src/align.rs:737-750): M[0] = score(N,N), then
            // insertions only; D stays NEG; no predecessors, so every winner
            // field is none.
            if (start != 0) {
                status = kBandExcludesSource;
                break;
            }
            if (leader) {
                const uint32_t wnone = ffn << 5 | ffn << so | ffn << se;
                T mprev = T(J.score_nn), iprev = NEG;
                for (uint32_t j = 0; j < len; j++) {
                    T mj = mprev, ins = iprev;
                    if (j > 0) {
                        ins = tmax(iprev + e, mprev + o);
                        mj = tmax(NEG, ins);
                    }
                    const uint32_t f = uint32_t(mj > HALF) | uint32_t(mj == NEG) << 1 |
                                       uint32_t(mj == ins) << 2 | uint32_t(mj + o >= ins + e) << 3 |
                                       uint32_t(ins + e >= mj + o) << 4;
                    if constexpr (REC) Rr[j] = uint8_t(f | wnone);  // no predecessors: 1-byte records
                    if (keep) {
                        Mk[j] = mj;
                        Dk[j] = NEG;
                    }
                    if constexpr (!REC) Ik[j] = ins;
                    if (slot != kNoSlot) {
                        ring[sm + j] = mj;
                        ring[sd + j] = NEG;
                    }
                    if (n == 1 && j == m - 1) xscore = int64_t(mj);
                    iprev = ins;
                    mprev = mj;
                }
                meta[0] = RowMeta{start, len, off, voff, int64_t(mprev)};
                hmeta[0].mlast = int64_t(mprev);
                hrec[0] = RowRec{0, 0, 0, 0, 0};
                if (adaptive) grec[0] = RowRec{0, 0, 0, 0, 0};
            }
            last_rec = RowRec{0, 0, 0, 0, 0};
            prev_slot = slot;
            sync_global<W>();  // the source row's arena values too
            continue;
        }

        // Predecessors outside the ring are read back from global memory,
        // where any wave may have written them: retire and publish every
        // wave's stores first. Uniform across the workgroup.
        bool anymiss = np > 64;
        if (!anymiss) {
            const bool miss = pb + uint32_t(lane) < pe && g_meta.len != 0 &&
                              (g_meta.slot == kNoSlot || wpos - g_meta.wpos > kRingBytes);
            anymiss = __builtin_amdgcn_ballot_w64(miss) != 0;
        }
        if (anymiss) sync_global<W>();

        const uint32_t rbase = daux[i % kBlk].base;
        const uint32_t srow = ((i == n - 1) ? J.unknown : rbase) * alpha;
        T carry = NEG;  // insertion carry into the stripe's first column
        T best = NI;
        uint32_t bestc = 0xFFFFFFFFu;
        uint32_t bestq = 0;  // LT: the query code of column bestc
        uint32_t parity = 0;
        bool lost = false;  // a predecessor's scores were in neither ring nor HBM
        // The stripes of the row, compiled separately for rows with one
        // predecessor (no tie rule: its fields hold position 0 until a gate
        // clears them) and with several.
        auto stripes = [&](auto one_c) {
            constexpr bool one = decltype(one_c)::value;
            for (uint32_t seg = start; seg < stop; seg += 64u * C * W, parity ^= 1u) {
                const uint32_t c0 = seg + wave * 64u * C + uint32_t(lane) * C;
                // A wave whose columns all lie past the row's end only joins the
                // carry exchange: its issue slots go to other workgroups.
                const bool live = seg + wave * 64u * C < stop;
                // Per column: the diagonal maximum and the two deletion maxima,
                // the packed record p (winner fields, then flags), D,
                // md = max(diagonal + profile score, D) and the in-lane insertion
                // prefix.
                T Mdg[C], Do[C], De[C], Dv[C], md[C], Lk[C];
                R pk[C];
                // LT: the lane's query codes, four per word (the argmax
                // column's code is then at hand for the adaptive update).
                uint32_t qw[C / 4];
#pragma unroll
                for (int k = 0; k < C / 4; k++) qw[k] = 0;
                T S = NI, Sp = NI;
                const uint32_t valid = c0 < stop ? tmin(stop - c0, uint32_t(C)) : 0u;
                const uint32_t rel = c0 - start;
                if (live) {
#pragma unroll
                    for (int k = 0; k < C; k++) {
                        Mdg[k] = NEG;
                        Do[k] = NEG;
                        De[k] = NEG;
                        // One predecessor: its fields hold position 0 until a gate
                        // clears them.
                        pk[k] = one ? R(0) : (Mm | Mo | Me);
                    }
                    // Predecessor transitions on the band overlap (src/align.rs:751-797).
                    // Bands and lane runs sit on the same 8-column grid, so each
                    // piece of G columns is wholly inside or outside an overlap; the
                    // diagonal into the predecessor's first column reads NEG. Pieces
                    // outside an overlap fold the NEG row: such a candidate is far
                    // below any cell a path can reach, so the winner it may become
                    // is never read. A row with one predecessor needs no tie rule.
                    auto fold = [&](int sg, const T (&pm)[G], const T (&pd)[G], T dg0, uint32_t idx) {
                        const R im = R(idx) << 5, io = R(idx) << so, ie = R(idx) << se;
#pragma unroll
                        for (int k = 0; k < G; k++) {
                            const T dg = k == 0 ? dg0 : pm[k - 1];
                            if constexpr (one) {
                                Mdg[sg + k] = dg;
                                Do[sg + k] = pm[k] + o;
                                De[sg + k] = pd[k] + e;
                            } else if constexpr (!REC) {
                                Mdg[sg + k] = tmax(Mdg[sg + k], dg);
                                Do[sg + k] = tmax(Do[sg + k], T(pm[k] + o));
                                De[sg + k] = tmax(De[sg + k], T(pd[k] + e));
                            } else {
                                track(Mdg[sg + k], pk[sg + k], dg, Mm, im);
                                track(Do[sg + k], pk[sg + k], T(pm[k] + o), Mo, io);
                                track(De[sg + k], pk[sg + k], T(pd[k] + e), Me, ie);
                            }
                        }
                    };
                    if (!anymiss) {
                        // Every predecessor row is in the LDS ring: this loop has no
                        // global loads, so nothing in it waits on vmcnt (which on
                        // gfx9 also counts the wave's outstanding stores).
                        for (uint32_t q = 0; q < np; q++) {
                            const uint32_t ps = rl32(g_meta.start, q), pl = rl32(g_meta.len, q);
                            const uint32_t ovs = tmax(start, ps), ove = tmin(stop, ps + pl);
                            if (ovs >= ove) continue;
                            if (leader && seg == start) atomicAdd(&xpcells, (unsigned long long)(ove - ovs));
                            const uint32_t pmo = rl32(g_meta.slot, q) / uint32_t(sizeof(T));
#pragma unroll
                            for (int sg = 0; sg < C; sg += G) {
                                const uint32_t cs = c0 + sg;
                                const bool in = cs >= ovs && cs < ove;
                                const uint32_t prel = cs - ps;
                                const T *mp = in ? ring + pmo + prel : negrow;
                                const T *dp = in ? ring + pmo + pl + prel : negrow;
                                const T *gp0 = in && prel > 0 ? ring + pmo + prel - 1 : negrow;
                                T pm[G], pd[G];
                                loadg<T, G>(mp, pm);
                                loadg<T, G>(dp, pd);
                                fold(sg, pm, pd, *gp0, q);
                            }
                        }
                    } else {
                    for (uint32_t g0 = pb; g0 < pe; g0 += 64) {
                        if (np > 64) gather(g0);
                        const uint32_t cnt = tmin(64u, pe - g0);
                        for (uint32_t q = 0; q < cnt; q++) {
                            const uint32_t ps = rl32(g_meta.start, q), pl = rl32(g_meta.len, q);
                            const uint32_t ovs = tmax(start, ps), ove = tmin(stop, ps + pl);
                            if (ovs >= ove) continue;
                            if (leader && seg == start) atomicAdd(&xpcells, (unsigned long long)(ove - ovs));
                            const uint32_t pslot = rl32(g_meta.slot, q), pw = rl32(g_meta.wpos, q);
                            const uint32_t idx = g0 - pb + q;
                            const bool in_ring = pslot != kNoSlot && wpos - pw <= kRingBytes;
                            if (in_ring) {
                                const uint32_t pmo = pslot / uint32_t(sizeof(T));
#pragma unroll
                                for (int sg = 0; sg < C; sg += G) {
                                    const uint32_t cs = c0 + sg;
                                    const bool in = cs >= ovs && cs < ove;
                                    const uint32_t prel = cs - ps;
                                    const T *mp = in ? ring + pmo + prel : negrow;
                                    const T *dp = in ? ring + pmo + pl + prel : negrow;
                                    const T *gp0 = in && prel > 0 ? ring + pmo + prel - 1 : negrow;
                                    T pm[G], pd[G];
                                    loadg<T, G>(mp, pm);
                                    loadg<T, G>(dp, pd);
                                    fold(sg, pm, pd, *gp0, idx);
                                }
                                continue;
                            }
                            const uint64_t pvoff = uint64_t(rl64(int64_t(g_meta.voff), q));
                            if (pvoff == kNoValues) {
                                // Every row read by a row other than the next one
                                // keeps its scores, and the next row always finds
                                // its predecessor in the ring; refuse rather than
                                // read outside the arena if that ever fails.
                                lost = true;
                                continue;
                            }
                            const gptr<const T> PM = glb(reinterpret_cast<const T *>(arena + pvoff));
#pragma unroll
                            for (int sg = 0; sg < C; sg += G) {
                                const uint32_t cs = c0 + sg;
                                if (cs < ovs || cs >= ove) continue;
                                const uint32_t prel = cs - ps;
                                T pm[G], pd[G], dg0;
                                loadg<T, G>(PM + prel, pm);
                                loadg<T, G>(PM + pl + prel, pd);
                                dg0 = prel > 0 ? PM[prel - 1] : NEG;
                                // Retire these loads inside the branch. Otherwise the
                                // registers reach the join with loads in flight and
                                // the ring path gets vmcnt waits, which also wait for
                                // every store of the wave.
                                __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
                                fold(sg, pm, pd, dg0, idx);
                            }
                        }
                    }
                    }
                    // D = max(open, extend), final here: a deletion winner stands
                    // only where its kind gives D. D goes out now.
#pragma unroll
                    for (int k = 0; k < C; k++) {
                        Dv[k] = tmax(Do[k], De[k]);
                        if constexpr (REC) pk[k] |= (ltmask<R>(Do[k], De[k]) & Mo) | (ltmask<R>(De[k], Do[k]) & Me);
                    }
                    if (valid > 0) {
                        if (keep) store_run<T, C>(Dk + rel, Dv, valid);
                        if (slot != kNoSlot) store_run<T, C>(ring + sd + rel, Dv, valid);
                    }
                    T L = NI;
#pragma unroll
                    for (int sg = 0; sg < C; sg += G) {
                        const uint32_t cs = c0 + sg;
                        const bool piece_live = cs < stop;
                        uint32_t codes[G];
                        T pf[G];
                        if constexpr (LT) {
                            // Dead pieces read column 0's codes and are discarded.
                            const uint32_t cr = piece_live ? cs : 0u;
                            if constexpr (G == 8) {
                                const uint2 w2 = *reinterpret_cast<const uint2 *>(sseq + cr);
#pragma unroll
                                for (int k = 0; k < 8; k++)
                                    codes[k] = ((k < 4 ? w2.x : w2.y) >> (8 * (k & 3))) & 0xFFu;
                                qw[sg / 4] = w2.x;
                                qw[sg / 4 + 1] = w2.y;
                            } else {
                                const uint32_t w1 = *reinterpret_cast<const uint32_t *>(sseq + cr);
#pragma unroll
                                for (int k = 0; k < 4; k++) codes[k] = (w1 >> (8 * k)) & 0xFFu;
                                qw[sg / 4] = w1;
                            }
#pragma unroll
                            for (int k = 0; k < G; k++) pf[k] = T(smat[srow + codes[k]]);
                        } else {
                            if (piece_live) {
#pragma unroll
                                for (int k = 0; k < G; k++) codes[k] = glb(seq)[cs + k];
                            }
#pragma unroll
                            for (int k = 0; k < G; k++) pf[k] = piece_live ? T(glb(gmat)[srow + codes[k]]) : T(0);
                        }
#pragma unroll
                        for (int k = 0; k < G; k++) {
                            const T dgp = Mdg[sg + k] + pf[k];
                            const T v = piece_live ? tmax(dgp, Dv[sg + k]) : NI;
                            // The diagonal winner stands only where the diagonal
                            // gives M: here against D, after the scan against I.
                            // M == D so far where D reaches md (md >= D).
                            if constexpr (REC)
                                pk[sg + k] |= (ltmask<R>(dgp, Dv[sg + k]) & Mm) | R(1u - lt01(Dv[sg + k], v)) << 1;
                            md[sg + k] = v;
                            Lk[sg + k] = L;
                            L = tmax(L + e, v + o);
                        }
                    }
                    // Carry into each lane: X_l = max(c_w + l*a, S_{l-1}), where c_w
                    // enters wave w: c_0 = carry, c_{w+1} = max(c_w + 64a, U_w) with
                    // U_w the wave's inclusive scan total (integer max-plus is
                    // associative, so this is the serial recurrence exactly).
                    S = maxplus_scan<T>(L, a, d15, d31);
                    Sp = dpp<kWaveShr1, 0xF>(NI, S);
                }
                T cw = carry;
                if constexpr (W > 1) {
                    if (lane == 63) xcar[parity][wave] = S;
                    sync_lds<W>();
#pragma unroll
                    for (uint32_t w = 0; w + 1 < W; w++) {
                        const T u = xcar[parity][w];
                        if (w < wave) cw = tmax(cw + a64, u);
                        carry = tmax(carry + a64, u);
                    }
                    carry = tmax(carry + a64, xcar[parity][W - 1]);
                } else {
                    carry = tmax(carry + a64, lane63(S));
                }
                if (!live) continue;
                const T X = tmax(cw + T(lane) * a, Sp);
                T Mv[C], Ivs[C];
#pragma unroll
                for (int k = 0; k < C; k++) {
                    const T Iv = tmax(X + T(k) * e, Lk[k]);
                    Mv[k] = tmax(md[k], Iv);
                    if (uint32_t(k) < valid && Mv[k] > best) {
                        best = Mv[k];
                        bestc = c0 + k;
                        if constexpr (LT) bestq = (qw[k / 4] >> (8 * (k & 3))) & 0xFFu;
                    }
                    if constexpr (!REC) {
                        Ivs[k] = Iv;
                        continue;
                    }
                    // Where I exceeds md, M comes from I: no diagonal winner, and
                    // M is above D. Then the flags: reachable, M == I, and which
                    // term I of the next column takes (I' = max(I + e, M + o)).
                    const R over = ltmask<R>(md[k], Iv);
                    const uint32_t f = lt01(HALF, Mv[k]) | (1u - lt01(Iv, md[k])) << 2 |
                                       (1u - lt01(T(Mv[k] + o), T(Iv + e))) << 3 |
                                       (1u - lt01(T(Iv + e), T(Mv[k] + o))) << 4;
                    pk[k] = ((pk[k] | (over & Mm)) & ~(over & R(2))) | R(f);
                }
                // The diagonal `choose` also sees a predecessor whose band ends
                // at the column before (its last column is j - 1, so the fill's
                // overlap leaves j out): it names that predecessor where its last
                // M plus the profile score equals M, if it comes first in list
                // order.
                for (uint32_t g0 = pb; REC && g0 < pe; g0 += 64) {
                    if (np > 64) gather(g0);
                    const uint32_t cnt = tmin(64u, pe - g0);
                    for (uint32_t q = 0; q < cnt; q++) {
                        const uint32_t pl = rl32(g_meta.len, q), bc = rl32(g_meta.start, q) + pl;
                        if (pl == 0 || bc < start || bc >= stop) continue;
                        const T ml = T(rl64(g_meta.mlast, q));
                        const uint32_t idx = g0 - pb + q;
                        if (bc >= c0 && bc < c0 + uint32_t(C)) {
                            // Only the lane holding column bc: its profile score
                            // again, rather than keeping C of them live.
                            T pfb;
                            if constexpr (LT) pfb = T(smat[srow + sseq[bc]]);
                            else pfb = T(glb(gmat)[srow + glb(seq)[bc]]);
#pragma unroll
                            for (int k = 0; k < C; k++) {
                                if (c0 + k == bc && Mv[k] == ml + pfb && idx < (uint32_t(pk[k] >> 5) & ffn))
                                    pk[k] = (pk[k] & ~Mm) | (R(idx) << 5);
                            }
                        }
                    }
                }
                if (valid > 0) {
                    if constexpr (REC) {
                        store_recs<C, R>(Rr + size_t(rel) * rb, pk, rb, valid);
                    } else {
                        store_run<T, C>(Ik + rel, Ivs, valid);
                    }
                    if (keep) store_run<T, C>(Mk + rel, Mv, valid);
                    if (slot != kNoSlot) store_run<T, C>(ring + sm + rel, Mv, valid);
                    // The row's geometry with its last M (read by successors'
                    // diagonal choice), and the DP's final score.
#pragma unroll
                    for (int k = 0; k < C; k++) {
                        if (c0 + k == stop - 1) {
                            hmeta[i % kHist].mlast = int64_t(Mv[k]);
                            meta[i] = RowMeta{start, len, off, voff, int64_t(Mv[k])};
                        }
                        if (i == n - 1 && c0 + k == m - 1) xscore = int64_t(Mv[k]);
                    }
                }
            }
        };
        if (np == 1)
            stripes(std::true_type{});
        else
            stripes(std::false_type{});
        if (lost) {
            status = kLostScores;
            break;
        }
        prev_slot = slot;

        // One wave: order this row's LDS writes (ring, geometry) before the
        // next row's reads. Several waves meet at the barrier below instead.
        if constexpr (W == 1) sync_lds<W>();
        RowRec rec{0, 0, ofl, ofr, extra};
        if (adaptive) {
            // First argmax of M over the row (src/align.rs:830-844), then the
            // streak / extra / offset update of src/align.rs:845-865.
            T gmax = wave_max<T>(best, NI);
            uint32_t maxc, maxq;
            if (len <= 64u * C * W) {
                // One stripe: lanes own increasing columns, so the lowest
                // lane holding the maximum has its first column.
                const uint32_t hl = uint32_t(__builtin_ctzll(__builtin_amdgcn_ballot_w64(best == gmax)));
                maxc = rl32(bestc, hl);
                maxq = rl32(bestq, hl);
            } else {
                maxc = wave_min(best == gmax ? bestc : 0xFFFFFFFFu);
                // Columns are unique to lanes: one lane holds maxc.
                const uint64_t at = __builtin_amdgcn_ballot_w64(best == gmax && bestc == maxc);
                maxq = rl32(bestq, at ? uint32_t(__builtin_ctzll(at)) : 0u);
            }
            if constexpr (W > 1) {
                if (lane == 0) {
                    xbest[wave] = gmax;
                    xbestc[wave] = maxc;
                    xbestq[wave] = maxq;
                }
                sync_lds<W>();  // also publishes the ring writes
                gmax = xbest[0];
                maxc = xbestc[0];
                maxq = xbestq[0];
#pragma unroll
                for (uint32_t w = 1; w < W; w++) {
                    const T b = xbest[w];
                    const uint32_t c = xbestc[w];
                    if (b > gmax || (b == gmax && c < maxc)) {
                        gmax = b;
                        maxc = c;
                        maxq = xbestq[w];
                    }
                }
            }
            const int32_t maximum = int32_t(maxc);
            int32_t streak = 0, max_streak = 0;
            bool is_match = false;
            uint32_t qbase;
            if constexpr (LT)
                qbase = maxq;
            else
                qbase = glb(seq)[maxc];
            if (rbase == qbase) {
                for (uint32_t g0 = pb; g0 < pe; g0 += 64) {
                    if (np > 64) gather(g0);
                    const uint32_t cnt = tmin(64u, pe - g0);
                    for (uint32_t q = 0; q < cnt; q++) {
                        const int32_t r_streak = rli(g_rec.streak, q), r_peak = rli(g_rec.peak, q);
                        max_streak = tmax(max_streak, r_streak);
                        if (r_peak + 1 == maximum) {
                            streak = tmax(streak, r_streak + 1);
                            is_match = true;
                        }
                    }
                }
            }
            if (!is_match && max_streak < 30) extra = tmin(extra + max_streak + 1, J.cap5);
            if (streak >= 30) {
                ofl = maximum - hh;
                ofr = maximum - hh;
                extra = 0;
            }
            rec = RowRec{maximum, streak, ofl, ofr, extra};
        } else if (W > 1) {
            sync_lds<W>();  // publishes the ring writes
        }
        if (leader) {
            // Read by rows after the next one, past its first barrier.
            hrec[i % kHist] = rec;
            if (adaptive) grec[i] = rec;
        }
        last_rec = rec;
    }
    sync_global<W>();  // every wave's stores and the score in LDS
    if (leader) {
        DevResult res{};
        res.status = status;
        res.cells = xcells;
        res.pred_cells = xpcells;
        res.arena_used = xused;
        res.score = status == kOk ? xscore : int64_t(NEG);
        P.res[b] = res;
    }
}

template <typename T, int C, int W, bool LT, typename R, bool REC>
__global__ __launch_bounds__(64 * W) __attribute__((amdgpu_waves_per_eu(fill_waves_per_eu<T, C, W, LT, REC>()))) void poa_fill(DevPools P) {
    extern __shared__ __align__(16) char fill_lds[];
    fill_job<T, C, W, LT, R, REC>(P, blockIdx.x, fill_lds);
}

// Weighted deterministic traceback (src/align.rs:1164-1313), one wave per job,
// over the fill's per-cell records: a step reads the record of its cell (and,
// for an insertion, of the cell to its left), which already names the
// `choose` winners, so no predecessor scores are read. The walk is serial and
// its cost is the latency of those reads, so the wave stages a window in LDS
// -- kTbRows rows ending at the current row (geometry, predecessor rows and
// weights) and their records over kTbCols columns ending at the current
// column -- loaded with all lanes and all loads in flight, and reloads when
// the path leaves it. Records of rows with more than 511 predecessors (8
// bytes) and predecessor lists longer than kTbPred are read from global
// memory. The query codes and the score matrix are staged in LDS once (LT),
// and the per-column results collect in LDS and go out as one coalesced
// store per 64 columns: on gfx9 a global load's vmcnt wait also waits for the
// wave's stores, so a step that stored to and loaded from global memory
// would pay a full memory round trip.
constexpr uint32_t kTbRows = 64, kTbCols = 64, kTbPred = 8;

struct WRow {
    uint32_t start, len;
    uint64_t off;   // byte offset of the row's records
    uint32_t pb, np, thresh, base;
};
struct WPred {
    uint32_t row, weight;
};

// One traceback record from global memory (rb = 1, 2, 4 or 8 bytes).
__device__ __forceinline__ uint64_t load_rec(const uint8_t *p, uint32_t rb) {
    if (rb == 8) return *glb(reinterpret_cast<const uint64_t *>(p));
    const uint64_t a = reinterpret_cast<uint64_t>(p);
    const uint32_t w = *glb(reinterpret_cast<const uint32_t *>(a & ~uint64_t(3)));
    const uint32_t v = w >> ((a & 3u) * 8u);
    return rb == 4 ? v : (v & ((1u << (8u * rb)) - 1u));
}

template <bool LT>
__global__ __launch_bounds__(64) void poa_traceback(DevPools P, uint32_t njobs) {
    __shared__ WRow wrow[kTbRows];
    __shared__ WPred wpred[kTbRows * kTbPred];
    __shared__ uint32_t wrec[kTbRows * kTbCols];
    __shared__ uint32_t tbw[64];  // results of the columns of the current 64-column chunk
    extern __shared__ __align__(16) char tb_lds[];  // LT: query codes, then the matrix
    const uint32_t jid = blockIdx.x;
    if (jid >= njobs) return;
    if (P.prio) __builtin_amdgcn_s_setprio(3);
    DevResult *R = P.res + jid;
    if (R->status != kOk) return;
    const DevJob J = P.jobs[jid];
    const int lane = threadIdx.x;
    const uint32_t n = J.n, m = J.m;
    const RowIn *rin = P.rows + J.rows;
    const RowAux *aux = P.aux + J.aux;
    const PredE *pr = P.preds + J.preds;
    const uint8_t *seq = P.seq + J.seq;
    const uint8_t *arena = static_cast<const uint8_t *>(P.arena);
    const RowMeta *meta = P.meta + J.rowmeta;
    uint32_t *tb = P.tb + J.tb;
    const bool gap0 = (J.flags & kGapOpenZero) != 0;
    uint8_t *sseq = reinterpret_cast<uint8_t *>(tb_lds);
    if constexpr (LT) {
        for (uint32_t k = uint32_t(lane); k < m; k += 64u) sseq[k] = glb(seq)[k];
    }
    auto qcode = [&](uint32_t c) -> uint32_t {
        if constexpr (LT) return sseq[c];
        else return glb(seq)[c];
    };
    // Columns are consumed in decreasing order, each exactly once, so a
    // chunk is complete when the walk moves below it.
    uint32_t chunk = (m - 1) / 64u;
    auto flush = [&]() {
        sync_lds<1>();
        const uint32_t c = chunk * 64u + uint32_t(lane);
        if (c >= 1 && c < m) tb[c] = tbw[lane];
        sync_lds<1>();
    };
    auto put = [&](uint32_t col, uint32_t v) {
        if (col / 64u != chunk) {
            flush();
            chunk = col / 64u;
        }
        if (lane == 0) tbw[col & 63u] = v;
    };

    uint32_t i0 = 0, j0 = 0, cw0 = 0;  // window: rows (i0-kTbRows, i0], columns [cw0, j0]
    bool have = false;
    auto load_window = [&](uint32_t ai, uint32_t aj) {
        sync_lds<1>();
        i0 = ai;
        j0 = aj;
        cw0 = aj + 1 > kTbCols ? aj + 1 - kTbCols : 0;
        have = true;
        // Geometry and predecessor lists: lane r holds row i0 - r. Lanes past
        // row 0 load row 0 again and store nothing, so no load sits behind a
        // branch (a load in flight at a join costs a wait on every path).
        {
            const uint32_t r = uint32_t(lane);
            const uint32_t ri = r <= i0 ? i0 - r : 0u;
            const RowMeta rm = ldg(meta + ri);
            const RowIn a = ldg(rin + ri), b = ldg(rin + ri + 1);
            const RowAux x = ldg(aux + ri);
            const uint32_t np = b.pred_beg - a.pred_beg;
            const uint32_t nq = tmin(np, kTbPred);
            PredE pe[kTbPred];
#pragma unroll
            for (uint32_t q = 0; q < kTbPred; q++)
                pe[q] = ldg(nq ? pr + a.pred_beg + (q < nq ? q : 0u) : P.preds);
            if (r <= i0) {
                wrow[r] = WRow{rm.start, rm.len, rm.off, a.pred_beg, np, x.thresh, x.base};
#pragma unroll
                for (uint32_t q = 0; q < kTbPred; q++)
                    if (q < nq) wpred[r * kTbPred + q] = WPred{pe[q].row, pe[q].weight};
            }
        }
        sync_lds<1>();
        // Records: lane k holds column cw0 + k; 16 rows per batch, every load
        // of a batch in flight before the first is used. A record of 1, 2 or
        // 4 bytes lies inside one aligned word (records are rb-aligned from a
        // 16-byte aligned row start), so every row loads that word.
        const uint32_t c = cw0 + uint32_t(lane);
        const bool col_ok = uint32_t(lane) < kTbCols && c <= j0;
        const gptr<const uint32_t> A = glb(reinterpret_cast<const uint32_t *>(arena));
        constexpr uint32_t kB = 16;
        for (uint32_t r0 = 0; r0 < kTbRows && r0 <= i0; r0 += kB) {
            uint32_t vw[kB], sh[kB];
            bool ok[kB];
#pragma unroll
            for (uint32_t u = 0; u < kB; u++) {
                const uint32_t r = r0 + u;
                const WRow g = wrow[r <= i0 ? r : 0u];
                const uint32_t rb = rec_bytes(g.np);
                ok[u] = col_ok && r <= i0 && rb <= 4u && c >= g.start && c - g.start < g.len;
                const uint64_t a = ok[u] ? g.off + uint64_t(c - g.start) * rb : 0ull;
                vw[u] = A[a >> 2];
                sh[u] = (uint32_t(a) & 3u) * 8u | (rb << 8);
            }
#pragma unroll
            for (uint32_t u = 0; u < kB; u++) {
                if (ok[u]) {
                    const uint32_t rb = sh[u] >> 8;
                    const uint32_t v = vw[u] >> (sh[u] & 31u);
                    wrec[(r0 + u) * kTbCols + uint32_t(lane)] = rb == 4 ? v : (v & ((1u << (8u * rb)) - 1u));
                }
            }
        }
        sync_lds<1>();
    };
    auto inside = [](const WRow &r, uint32_t c) { return c >= r.start && c - r.start < r.len; };

    uint32_t i = n - 1, j = m - 1, op = kAll;
    uint32_t prev_i = n - 1;
    uint64_t steps = 0;
    int32_t status = kOk;
    while (i > 0 || j > 0) {
        if (++steps > uint64_t(n) + m || i > prev_i) {
            status = kTbNoAdvance;
            break;
        }
        prev_i = i;
        // The step reads row i at columns j and j - 1.
        if (!have || i0 - i >= kTbRows || (j > 0 ? j - 1 : 0u) < cw0) load_window(i, j);
        const uint32_t wr = i0 - i;
        const WRow rg = wrow[wr];
        if (!inside(rg, j)) {
            status = kTbLeftBand;
            break;
        }
        const uint32_t rb = rec_bytes(rg.np), fb = rec_field_bits(rg.np);
        const uint64_t fnone = (uint64_t(1) << fb) - 1u;
        auto rec_at = [&](uint32_t c) -> uint64_t {
            if (rb <= 4u) return wrec[wr * kTbCols + (c - cw0)];
            return load_rec(arena + rg.off + uint64_t(c - rg.start) * rb, rb);
        };
        const uint64_t rec = rec_at(j);
        if (!(rec & kRecFlagR)) {
            status = kTbLeftBand;
            break;
        }
        const uint32_t win_m = uint32_t((rec >> 5) & fnone);
        const uint32_t win_o = uint32_t((rec >> (5 + fb)) & fnone);
        const uint32_t win_e = uint32_t((rec >> (5 + 2 * fb)) & fnone);
        // A winner names a position in the row's list; the walk only moves
        // to earlier rows.
        if ((win_m != fnone && win_m >= rg.np) || (win_o != fnone && win_o >= rg.np) ||
            (win_e != fnone && win_e >= rg.np)) {
            status = kTbNoPredecessor;
            break;
        }
        // Predecessor `idx` of row i (its position in the row's list).
        auto pred = [&](uint32_t idx) -> WPred {
            if (rg.np <= kTbPred) return wpred[wr * kTbPred + idx];
            const PredE pq = ldg(pr + rg.pb + idx);
            return WPred{pq.row, pq.weight};
        };
        const uint32_t base = (i == n - 1) ? J.unknown : rg.base;
        // Preferred match: the diagonal winner, if the query base equals the
        // node's and its edge carries at least a tenth of the node's support.
        if ((op & kMatch) && j > 0 && win_m != fnone && base == qcode(j)) {
            const WPred p = pred(win_m);
            if (p.weight >= rg.thresh) {
                put(j, i);
                i = p.row;
                j -= 1;
                op = kAll;
                continue;
            }
        }
        if ((op & kDelete) && (op == kDelete || (rec & kRecFlagMD))) {
            if (win_o != fnone) {
                i = pred(win_o).row;
                op = gap0 ? kAll : (kMatch | kInsert);
                continue;
            }
            if (win_e != fnone) {
                i = pred(win_e).row;
                op = kDelete;
                continue;
            }
        }
        if ((op & kInsert) && j > 0 && inside(rg, j - 1) && (op == kInsert || (rec & kRecFlagMI))) {
            // The record of column j - 1 says which term I at column j took.
            const uint64_t left = rec_at(j - 1);
            if (left & kRecFlagNO) {
                put(j, kTbInsert);
                j -= 1;
                op = gap0 ? kAll : (kMatch | kDelete);
                continue;
            }
            if (left & kRecFlagNE) {
                put(j, kTbInsert);
                j -= 1;
                op = kInsert;
                continue;
            }
        }
        if ((op & kMatch) && j > 0 && win_m != fnone) {
            put(j, i);
            i = pred(win_m).row;
            j -= 1;
            op = kAll;
            continue;
        }
        status = kTbNoPredecessor;
        break;
    }
    if (status == kOk) {
        flush();
    } else if (lane == 0) {
        R->status = status;
        R->err_i = i;
        R->err_j = j;
    }
}

// Weighted deterministic traceback over the values layout (src/align.rs:
// 1164-1313), one wave per job: `choose` is re-evaluated from the stored M
// and D of the predecessors, and the insertion rules from I. The walk is
// serial, so its cost is the latency of each step's reads; the wave stages a
// window of the matrix in LDS -- kTbRows rows ending at the current row
// (geometry, predecessor lists and M/D/I over kVCols columns ending at the
// current column), loaded with all lanes and all loads in flight at once --
// and walks inside it, reloading when the path nears the window's top-left
// edge. Predecessors older than the window, or rows with more than kTbPred
// predecessors, read global memory directly. Query codes, score matrix and
// the per-column results use LDS as in poa_traceback.
constexpr uint32_t kTbReach = 16;

template <typename T> constexpr uint32_t tbv_cols() { return sizeof(T) == 4 ? 32 : 16; }
// The values traceback's fixed LDS state, at the start of the job's LDS
// region; LT: the query codes and the score matrix follow it.
template <typename T> struct TbvLds {
    WRow wrow[kTbRows];
    PredE wpred[kTbRows * kTbPred];
    T wm[kTbRows * tbv_cols<T>()], wd[kTbRows * tbv_cols<T>()], wi[kTbRows * tbv_cols<T>()];
    uint32_t tbw[64];  // records of the columns of the current 64-column chunk
};
template <typename T> constexpr uint32_t tbv_fixed_lds() {
    return uint32_t((sizeof(TbvLds<T>) + 15u) & ~size_t(15));
}

// The traceback of the job at index jid by the calling wave (lanes
// threadIdx.x & 63), with its LDS region at `lds`.
template <typename T, bool LT>
__device__ __forceinline__ void tbv_job(const DevPools &P, uint32_t jid, char *lds) {
    constexpr uint32_t kVCols = tbv_cols<T>();
    TbvLds<T> &TL = *reinterpret_cast<TbvLds<T> *>(lds);
    auto &wrow = TL.wrow;
    auto &wpred = TL.wpred;
    auto &wm = TL.wm;
    auto &wd = TL.wd;
    auto &wi = TL.wi;
    auto &tbw = TL.tbw;
    char *tb_lds = lds + tbv_fixed_lds<T>();  // LT: query codes, then the matrix
    if (P.prio) __builtin_amdgcn_s_setprio(3);
    DevResult *R = P.res + jid;
    if (R->status != kOk) return;
    const DevJob J = P.jobs[jid];
    const int lane = int(threadIdx.x & 63u);
    const uint32_t n = J.n, m = J.m;
    const RowIn *rin = P.rows + J.rows;
    const RowAux *aux = P.aux + J.aux;
    const PredE *pr = P.preds + J.preds;
    const uint8_t *seq = P.seq + J.seq;
    const int32_t *mat = P.matrix + J.matrix;
    const T *arena = reinterpret_cast<const T *>(P.arena);
    const RowMeta *meta = P.meta + J.rowmeta;
    uint32_t *tb = P.tb + J.tb;
    const int64_t o = J.o, e = J.e;
    const int64_t half = int64_t(Lim<T>::neg) / 2;
    const bool gap0 = (J.flags & kGapOpenZero) != 0;
    uint8_t *sseq = reinterpret_cast<uint8_t *>(tb_lds);
    int32_t *smat = reinterpret_cast<int32_t *>(tb_lds + ((P.seq_lds + 15u) & ~15u));
    if constexpr (LT) {
        for (uint32_t k = uint32_t(lane); k < J.alpha * J.alpha; k += 64u) smat[k] = glb(mat)[k];
        for (uint32_t k = uint32_t(lane); k < m; k += 64u) sseq[k] = glb(seq)[k];
    }
    auto qcode = [&](uint32_t c) -> uint32_t {
        if constexpr (LT) return sseq[c];
        else return glb(seq)[c];
    };
    auto score_of = [&](uint32_t a, uint32_t b) -> int64_t {
        if constexpr (LT) return smat[a * J.alpha + b];
        else return glb(mat)[size_t(a) * J.alpha + b];
    };
    // Columns are consumed in decreasing order, each exactly once, so a
    // chunk is complete when the walk moves below it.
    uint32_t chunk = (m - 1) / 64u;
    auto flush = [&]() {
        sync_lds<1>();
        const uint32_t c = chunk * 64u + uint32_t(lane);
        if (c >= 1 && c < m) tb[c] = tbw[lane];
        sync_lds<1>();
    };
    auto put = [&](uint32_t col, uint32_t v) {
        if (col / 64u != chunk) {
            flush();
            chunk = col / 64u;
        }
        if (lane == 0) tbw[col & 63u] = v;
    };

    uint32_t i0 = 0, j0 = 0, cw0 = 0;  // window: rows (i0-kTbRows, i0], columns [cw0, j0]
    bool have = false;
    auto load_window = [&](uint32_t ai, uint32_t aj) {
        sync_lds<1>();
        i0 = ai;
        j0 = aj;
        cw0 = aj + 1 > kVCols ? aj + 1 - kVCols : 0;
        have = true;
        // Geometry and predecessor lists: lane r holds row i0 - r. Lanes past
        // row 0 load row 0 again and store nothing, so no load sits behind a
        // branch (a load in flight at a join costs a wait on every path).
        {
            const uint32_t r = uint32_t(lane);
            const uint32_t ri = r <= i0 ? i0 - r : 0u;
            const RowMeta rm = ldg(meta + ri);
            const RowIn a = ldg(rin + ri), b = ldg(rin + ri + 1);
            const RowAux x = ldg(aux + ri);
            const uint32_t np = tmin(b.pred_beg - a.pred_beg, kTbPred);
            PredE pe[kTbPred];
#pragma unroll
            for (uint32_t q = 0; q < kTbPred; q++) pe[q] = ldg(np ? pr + a.pred_beg + (q < np ? q : 0u) : P.preds);
            if (r <= i0) {
                wrow[r] = WRow{rm.start, rm.len, rm.voff / sizeof(T), a.pred_beg, b.pred_beg - a.pred_beg, x.thresh, x.base};
#pragma unroll
                for (uint32_t q = 0; q < kTbPred; q++)
                    if (q < np) wpred[r * kTbPred + q] = pe[q];
            }
        }
        sync_lds<1>();
        // M, D and I: lane k holds column cw0 + (k & 31) of the even rows
        // (k < 32) or the odd rows (k >= 32); 32 rows per batch, every load
        // of a batch in flight before the first is used.
        const uint32_t hl = uint32_t(lane) >> 5, cl = uint32_t(lane) & 31u;
        const uint32_t c = cw0 + cl;
        const bool col_ok = cl < kVCols && c <= j0;
        const gptr<const T> A = glb(static_cast<const T *>(arena));
        constexpr uint32_t kB = 16;
        for (uint32_t r0 = 0; r0 < kTbRows && r0 <= i0; r0 += 2 * kB) {
            T vm[kB], vd[kB], vi[kB];
            bool ok[kB];
#pragma unroll
            for (uint32_t u = 0; u < kB; u++) {
                const uint32_t r = r0 + 2 * u + hl;
                const bool rin = r < kTbRows && r <= i0;
                const WRow g = wrow[rin ? r : 0u];
                ok[u] = col_ok && rin && c >= g.start && c - g.start < g.len;
                const uint64_t k = ok[u] ? g.off + (c - g.start) : 0ull;
                const uint64_t l = ok[u] ? uint64_t(g.len) : 0ull;
                vm[u] = A[k];
                vd[u] = A[k + l];
                vi[u] = A[k + 2ull * l];
            }
#pragma unroll
            for (uint32_t u = 0; u < kB; u++) {
                if (ok[u]) {
                    const uint32_t w = (r0 + 2 * u + hl) * kVCols + cl;
                    wm[w] = vm[u];
                    wd[w] = vd[u];
                    wi[w] = vi[u];
                }
            }
        }
        sync_lds<1>();
    };
    auto row_in_win = [&](uint32_t ri) { return have && ri <= i0 && i0 - ri < kTbRows; };
    // Geometry of a row: window copy, else global.
    auto geo = [&](uint32_t ri) -> WRow {
        if (row_in_win(ri)) return wrow[i0 - ri];
        const RowMeta rm = ldg(meta + ri);
        const RowIn a = ldg(rin + ri), b = ldg(rin + ri + 1);
        return WRow{rm.start, rm.len, rm.voff / sizeof(T), a.pred_beg, b.pred_beg - a.pred_beg, 0, 0};
    };
    auto inside = [](const WRow &r, uint32_t c) { return c >= r.start && c - r.start < r.len; };
    // State s (0 = M, 1 = D, 2 = I) of row ri at an in-band column c.
    auto val = [&](uint32_t ri, const WRow &g, uint32_t c, int s) -> int64_t {
        if (row_in_win(ri) && c >= cw0 && c <= j0) {
            const uint32_t k = (i0 - ri) * kVCols + (c - cw0);
            if (s == 0) return int64_t(wm[k]);
            if (s == 1) return int64_t(wd[k]);
            return int64_t(wi[k]);
        }
        return int64_t(glb(arena)[g.off + uint64_t(s) * g.len + (c - g.start)]);
    };
    auto pred = [&](uint32_t ri, const WRow &g, uint32_t q) -> PredE {
        if (row_in_win(ri) && g.np <= kTbPred) return wpred[(i0 - ri) * kTbPred + q];
        return ldg(pr + g.pb + q);
    };

    uint32_t i = n - 1, j = m - 1, op = kAll;
    uint64_t steps = 0;
    int32_t status = kOk;
    while (i > 0 || j > 0) {
        if (++steps > uint64_t(n) + m) {
            status = kTbNoAdvance;
            break;
        }
        if (!have || i0 - i + kTbReach > kTbRows || (j > 0 && j - 1 < cw0)) load_window(i, j);
        const WRow rg = geo(i);
        if (!inside(rg, j)) {
            status = kTbLeftBand;
            break;
        }
        const int64_t Mij = val(i, rg, j, 0);
        if (Mij <= half) {
            status = kTbLeftBand;
            break;
        }
        const int64_t Dij = val(i, rg, j, 1);
        const uint32_t base = (i == n - 1) ? J.unknown : rg.base;
        const uint32_t qj = qcode(j);
        const int64_t score = score_of(base, qj);
        // `choose` (src/align.rs:1201-1225): among in-range predecessors whose
        // state plus penalty equals the target, the largest in_weight; ties go
        // to the earliest in `incoming` order. Returns the predecessor row.
        auto choose = [&](uint32_t state, int64_t penalty, uint32_t pj, uint32_t &w_out) -> int64_t {
            const int64_t target = state == kMatch ? Mij : Dij;
            int64_t best = -1;
            uint32_t w = 0, bord = 0;
            for (uint32_t q = 0; q < rg.np; q++) {
                const PredE pq = pred(i, rg, q);
                const WRow pg = geo(pq.row);
                if (!inside(pg, pj)) continue;
                const int64_t src = val(pq.row, pg, pj, state == kDelete ? 1 : 0);
                if (target == src + penalty &&
                    (best < 0 || pq.weight > w || (pq.weight == w && pq.ord < bord))) {
                    best = pq.row;
                    w = pq.weight;
                    bord = pq.ord;
                }
            }
            w_out = w;
            return best;
        };
        uint32_t w = 0;
        int64_t pref = -1;
        if ((op & kMatch) && j > 0 && base == qj) pref = choose(kMatch, score, j - 1, w);
        if (pref >= 0 && w < rg.thresh) pref = -1;
        if (pref >= 0) {
            put(j, i);
            i = uint32_t(pref);
            j -= 1;
            op = kAll;
            continue;
        }
        if ((op & kDelete) && (op == kDelete || Mij == Dij)) {
            int64_t k = choose(kInsert, o, j, w);
            if (k >= 0) {
                i = uint32_t(k);
                op = gap0 ? kAll : (kMatch | kInsert);
                continue;
            }
            k = choose(kDelete, e, j, w);
            if (k >= 0) {
                i = uint32_t(k);
                op = kDelete;
                continue;
            }
        }
        if ((op & kInsert) && j > 0 && inside(rg, j - 1)) {
            const int64_t Iij = val(i, rg, j, 2);
            if (op == kInsert || Mij == Iij) {
                if (Iij == val(i, rg, j - 1, 0) + o) {
                    put(j, kTbInsert);
                    j -= 1;
                    op = gap0 ? kAll : (kMatch | kDelete);
                    continue;
                }
                if (Iij == val(i, rg, j - 1, 2) + e) {
                    put(j, kTbInsert);
                    j -= 1;
                    op = kInsert;
                    continue;
                }
            }
        }
        if ((op & kMatch) && j > 0) {
            const int64_t k = choose(kMatch, score, j - 1, w);
            if (k >= 0) {
                put(j, i);
                i = uint32_t(k);
                j -= 1;
                op = kAll;
                continue;
            }
        }
        status = kTbNoPredecessor;
        break;
    }
    if (status == kOk) {
        flush();
    } else if (lane == 0) {
        R->status = status;
        R->err_i = i;
        R->err_j = j;
    }
}

template <typename T, bool LT>
__global__ __launch_bounds__(64) void poa_traceback_values(DevPools P, uint32_t njobs) {
    extern __shared__ __align__(16) char tbv_lds[];
    if (blockIdx.x >= njobs) return;
    tbv_job<T, LT>(P, blockIdx.x, tbv_lds);
}

template <typename T, int C, int W, bool LT, bool REC, typename R = uint32_t>
static void launch_fill_t(const DevPools &P, uint32_t njobs, hipStream_t s) {
    const uint32_t lds = fill_fixed_lds<T, W>() + fill_dynamic_lds(P.ring_bytes, P.seq_lds, P.mat_lds);
    hipLaunchKernelGGL((poa_fill<T, C, W, LT, R, REC>), dim3(njobs), dim3(64 * W), lds, s, P);
}

template <typename T, bool REC>
static int launch_fill_l(const DevPools &P, uint32_t njobs, bool wide, FillShape f, hipStream_t s) {
    const bool lt = P.seq_lds != 0 && P.mat_lds != 0;
    if (!lt || wide) {
        // Tables in global memory, or rows with more than kMaxNarrowPreds
        // predecessors in the records layout (8-byte records): one shape.
        if (f.cols != 8 || f.waves != 1) return -1;
        if constexpr (REC) {
            if (wide && lt) launch_fill_t<T, 8, 1, true, true, uint64_t>(P, njobs, s);
            else if (wide) launch_fill_t<T, 8, 1, false, true, uint64_t>(P, njobs, s);
            else launch_fill_t<T, 8, 1, false, true>(P, njobs, s);
        } else {
            if (wide) return -1;
            launch_fill_t<T, 8, 1, false, false>(P, njobs, s);
        }
        return 0;
    }
    switch (f.cols * 8 + f.waves) {
        case 8 * 8 + 1: launch_fill_t<T, 8, 1, true, REC>(P, njobs, s); return 0;
        case 16 * 8 + 1: launch_fill_t<T, 16, 1, true, REC>(P, njobs, s); return 0;
        case 4 * 8 + 2: launch_fill_t<T, 4, 2, true, REC>(P, njobs, s); return 0;
        case 8 * 8 + 2: launch_fill_t<T, 8, 2, true, REC>(P, njobs, s); return 0;
        case 4 * 8 + 4: launch_fill_t<T, 4, 4, true, REC>(P, njobs, s); return 0;
        case 8 * 8 + 4: launch_fill_t<T, 8, 4, true, REC>(P, njobs, s); return 0;
        default: return -1;
    }
}

template <typename T>
static int launch_fill_s(const DevPools &P, uint32_t njobs, bool rec, bool wide, FillShape f, hipStream_t s) {
    return rec ? launch_fill_l<T, true>(P, njobs, wide, f, s) : launch_fill_l<T, false>(P, njobs, wide, f, s);
}

bool fill_shape_ok(FillShape f) {
    const int k = f.cols * 8 + f.waves;
    return k == 8 * 8 + 1 || k == 16 * 8 + 1 || k == 4 * 8 + 2 || k == 8 * 8 + 2 ||
           k == 4 * 8 + 4 || k == 8 * 8 + 4;
}

// Tuned on the 4,096-job DNA workload (mean adaptive width ~544): a 16 KB
// ring with more resident workgroups beats deeper rings.
uint32_t ring_bytes_for_width(uint32_t width, bool i64) {
    uint32_t row = 2u * ((width + 7u) & ~7u) * (i64 ? 8u : 4u);
    uint32_t want = 3u * row;
    want = want < 16384u ? 16384u : (want > 49152u ? 49152u : want);
    return (want + 63u) & ~63u;
}

// More waves shorten a row's serial path (each wave pays the row's fixed
// work in parallel and owns fewer columns); fewer registers per wave let
// more jobs stay resident. Measured on the 4,096-job DNA workload (mean
// adaptive width ~544, rows up to ~1,500), wall time per run with four
// waves per SIMD: 4x4 1451 ms, 8x4 1565, 8x2 1568, 4x2 1649. Narrow rows
// keep one wave; rows wider than one 4x4 stripe (1,024 columns) take 8x4.
FillShape fill_shape_for_width(uint32_t width) {
    if (width <= 384u) return FillShape{8, 1};
    if (width <= 1024u) return FillShape{4, 4};
    return FillShape{8, 4};
}

int launch_fill(const DevPools &P, uint32_t njobs, bool i64, bool rec, bool wide, FillShape f, hipStream_t s) {
    if (njobs == 0) return 0;
    return i64 ? launch_fill_s<int64_t>(P, njobs, rec, wide, f, s)
               : launch_fill_s<int32_t>(P, njobs, rec, wide, f, s);
}

// Static LDS of the records traceback (window rows, predecessor lists,
// records, column results) and of the values traceback (window rows,
// predecessor lists, M/D/I window, column results); the query codes (and
// for values the score matrix) come on top when LT.
constexpr uint32_t kTbStaticLds =
    uint32_t(sizeof(WRow) * kTbRows + sizeof(WPred) * kTbRows * kTbPred +
             sizeof(uint32_t) * kTbRows * kTbCols + 64 * sizeof(uint32_t));

void launch_traceback(const DevPools &P, uint32_t njobs, bool i64, bool rec, hipStream_t s) {
    if (njobs == 0) return;
    dim3 grid(njobs);
    if (rec) {
        const uint32_t dyn = (P.seq_lds + 15u) & ~15u;
        const bool lt = P.seq_lds != 0 && kTbStaticLds + dyn <= 65536u;
        if (lt) hipLaunchKernelGGL((poa_traceback<true>), grid, dim3(64), dyn, s, P, njobs);
        else hipLaunchKernelGGL((poa_traceback<false>), grid, dim3(64), 0, s, P, njobs);
        return;
    }
    const uint32_t dyn = fill_dynamic_lds(0, P.seq_lds, P.mat_lds);
    const uint32_t stat = i64 ? tbv_fixed_lds<int64_t>() : tbv_fixed_lds<int32_t>();
    const bool lt = P.seq_lds != 0 && P.mat_lds != 0 && stat + dyn <= 65536u;
    const uint32_t lds = stat + (lt ? dyn : 0);
    // Both variants hold their windows in dynamic LDS (stat bytes).
    if (i64) {
        if (lt) hipLaunchKernelGGL((poa_traceback_values<int64_t, true>), grid, dim3(64), lds, s, P, njobs);
        else hipLaunchKernelGGL((poa_traceback_values<int64_t, false>), grid, dim3(64), lds, s, P, njobs);
    } else {
        if (lt) hipLaunchKernelGGL((poa_traceback_values<int32_t, true>), grid, dim3(64), lds, s, P, njobs);
        else hipLaunchKernelGGL((poa_traceback_values<int32_t, false>), grid, dim3(64), lds, s, P, njobs);
    }
}

// Persistent resident kernel: each workgroup takes jobs from a device queue
// (longest first) and runs each to its end -- the resident step (consume the
// last traceback, fuse, re-rank, plan; wave 0), the fill (all waves), the
// traceback (wave 0), once per read -- with every DP in the workgroup's own
// arena region. There is no launch per step and no lockstep between jobs: a
// workgroup whose job ends starts the next one at once, so the chip stays
// full until the queue runs dry. The three phases alias one LDS region and
// meet at workgroup barriers; decisions that end a loop are read after a
// barrier and made uniform, so every wave meets every barrier. Workgroups
// never wait for one another (the queue is one atomic counter), so the
// kernel completes whatever number of them is resident at once.
template <int C, int W>
__global__ __launch_bounds__(64 * W) __attribute__((amdgpu_waves_per_eu(4))) void rj_persist_kernel(
    RPools R, DevPools Pin, RPersist Q) {
    // The launch's pools as the fill and traceback take them; the pointers
    // both structures hold (the host sets them equal) come from R, so the
    // kernel keeps one copy of each.
    DevPools P = Pin;
    P.jobs = R.dj;
    P.res = R.res;
    P.rows = R.rows;
    P.aux = R.aux;
    P.preds = R.preds;
    P.seq = R.seq;
    P.tb = R.tb;
    P.prio = 0;  // the persistent launch sets priorities per phase (Q.prio)
    extern __shared__ __align__(16) char pl_lds[];
    __shared__ uint32_t xjob;
    __shared__ unsigned long long xsoff;
    const uint32_t tid = threadIdx.x;
    const uint32_t wave = uint32_t(__builtin_amdgcn_readfirstlane(int(tid >> 6)));
    const uint64_t slot_off = Q.arena_off + uint64_t(blockIdx.x) * Q.slot_bytes;
    for (;;) {
        if (tid == 0) xjob = atomicAdd(Q.next, 1u);
        sync_lds<W>();
        const uint32_t k = uint32_t(__builtin_amdgcn_readfirstlane(int(xjob)));
        sync_lds<W>();  // xjob read by every wave before the next job's write
        if (k >= R.njobs) break;
        // A read takes one step, plus at most two for poa's full-DP retry and
        // a full-width arena; past this bound the job leaves for the host
        // engine, which continues it from its next unfused read.
        const uint32_t bound = 4u * R.jobs[k].nreads + 8u;
        for (uint32_t it = 0;; it++) {
            if (wave == 0) {
                WaveTeam t{tid};
                // The step is a serial chain on one wave while other
                // workgroups fill on the same SIMDs: at raised priority it
                // takes their issue slots first.
                if (Q.prio & 1u) __builtin_amdgcn_s_setprio(2);
                if (it < bound) {
                    rj_job_step(t, R, k, pl_lds, Q.step_lds, slot_off, Q.slot_bytes);
                } else if (tid == 0) {
                    RJob &J = R.jobs[k];
                    if (J.state == kRsRunning) {
                        J.state = kRsEvicted;
                        J.evict = kEvStalled;
                        atomicAdd(R.done, 1u);
                    }
                    R.dj[k].n = 0;
                }
                if (Q.prio & 1u) __builtin_amdgcn_s_setprio(0);
            }
            sync_global<W>();  // the planned DP (or n == 0: the job left) for every wave
            const uint32_t n = uint32_t(__builtin_amdgcn_readfirstlane(int(glb(&R.dj[k].n)[0])));
            if (n == 0) break;
            fill_job<int32_t, C, W, true, uint32_t, false>(P, k, pl_lds);
            sync_global<W>();
            if (wave == 0) {
                if (Q.prio & 2u) __builtin_amdgcn_s_setprio(2);
                tbv_job<int32_t, true>(P, k, pl_lds);
                if (Q.prio & 2u) __builtin_amdgcn_s_setprio(0);
            }
            sync_global<W>();
        }
        if (Q.sbuf) {
            // The job has left the loop with its final state and graph
            // (every wave's stores published by the last barrier): pack it
            // and publish its entry. The releases are system-wide so that a
            // copy engine reading device memory sees the bytes before the
            // entry.
            const RJob &J = R.jobs[k];
            const uint64_t bytes = rj_pack_layout(J.nn, J.ne, J.nreads, J.npath).bytes;
            if (tid == 0) xsoff = atomicAdd(Q.scur, (unsigned long long)bytes);
            sync_lds<W>();
            const uint64_t off = xsoff;
            if (off + bytes <= Q.scap) {
                rj_pack_job(J, Q.sbuf + off, tid, 64u * W);
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "");
                sync_lds<W>();
                if (tid == 0) {
                    uint32_t *e = Q.sent + 4ull * k;
                    e[1] = uint32_t(off);
                    e[2] = uint32_t(off >> 32);
                    e[3] = uint32_t(bytes);
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "");
                    __hip_atomic_store(e, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "");
                }
            }
            sync_lds<W>();  // xsoff read by every wave before the next job's write
        }
    }
}

uint32_t rj_persist_lds(const DevPools &P, uint32_t step_lds) {
    const uint32_t fill = fill_fixed_lds<int32_t, 4>() + fill_dynamic_lds(P.ring_bytes, P.seq_lds, P.mat_lds);
    const uint32_t tb = tbv_fixed_lds<int32_t>() + fill_dynamic_lds(0, P.seq_lds, P.mat_lds);
    uint32_t m = fill > tb ? fill : tb;
    return m > step_lds ? m : step_lds;
}

int launch_rj_persist(const RPools &R, const DevPools &P, const RPersist &Q, uint32_t grid, FillShape f,
                      hipStream_t s) {
    if (f.cols != 4 || f.waves != 4) return -1;
    if (grid == 0) return 0;
    hipLaunchKernelGGL((rj_persist_kernel<4, 4>), dim3(grid), dim3(256), rj_persist_lds(P, Q.step_lds), s, R, P,
                       Q);
    return 0;
}

// Device scratch of the no-allocation entry point: one fixed array per GPU,
// loaded with this code object and held until the process ends. Its name in
// host code is a host-side stand-in, not a device address; the address is
// looked up once per device (hipGetSymbolAddress, device current) and only
// that address is handed to kernels and copies.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
std::mutex g_scratch_mu;
void *g_scratch_ptr[kScratchDevices];
bool g_scratch_done[kScratchDevices];
hipError_t g_scratch_err[kScratchDevices];
}  // namespace

hipError_t scratch_init(int device) {
    if (device < 0 || device >= int(kScratchDevices)) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_done[device]) return g_scratch_err[device];
    int prev = 0;
    hipError_t e = hipGetDevice(&prev);
    if (e == hipSuccess) e = hipSetDevice(device);
    void *ptr = nullptr;
    if (e == hipSuccess) {
        e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        (void)hipSetDevice(prev);
    }
    if (e != hipSuccess) {
        (void)hipGetLastError();
        ptr = nullptr;
    }
    g_scratch_ptr[device] = ptr;
    g_scratch_err[device] = e;
    g_scratch_done[device] = true;
    return e;
}

void *scratch_device_ptr(int device) {
    if (device < 0 || device >= int(kScratchDevices)) return nullptr;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return g_scratch_done[device] ? g_scratch_ptr[device] : nullptr;
}

}  // namespace mpoa
