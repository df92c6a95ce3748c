// Device-side job layout shared by the host planner and the HIP kernels.
//
// One DP job = one `poa_inner` call (src/align.rs:531-545): a graph rank range
// [beg, end] (n rows, in rank order) against one query (m = qlen + 2 columns).
// Everything the kernels read is flattened into pools in rank order: per-row
// records and each row's in-range predecessors.
#pragma once
#include <cstdint>

namespace mpoa {

constexpr int32_t kNeg32 = -(1 << 29);               // src/align.rs:216
constexpr int64_t kNeg64 = -(int64_t(1) << 50);      // src/align.rs:16
constexpr int32_t kHeadNone = INT32_MIN;             // head/tail == NEG
constexpr uint32_t kTbInsert = 0xFFFFFFFFu;          // traceback column was an insertion

enum JobFlags : uint32_t {
    kFull = 1u << 0,        // params.f == 0: every row spans [0, m]
    kAdaptive = 1u << 1,    // f != 0 and (adaptive || params.adaptive)
    kGapOpenZero = 1u << 2, // params.gap_open == 0 (traceback state rule)
};

// Per row (n + 1 entries per job; entry n carries the end offset).
struct RowIn {
    uint32_t pred_beg;  // into the job's predecessor list
    int32_t head;       // most-supported distance from beg (kHeadNone = NEG)
    int32_t tail;       // most-supported distance to end (kHeadNone = NEG)
    uint32_t last_use;  // last row that has this row as a predecessor (0: none)
};

struct RowAux {
    uint32_t thresh;    // sum(in_weights) / 10: preferred-match filter
    uint32_t base;      // node base (alphabet index)
};

// In-range predecessors of a row (head != NEG), in the traceback's `choose`
// order (heaviest edge first, then the earliest in `incoming`,
// src/align.rs:1201-1225): among tied candidates the winner is the one of
// lowest position, and records name winners by position. The adaptive band
// state is folded from the same entries in rank order, which rperm gives:
// entry rperm of the list is the t-th predecessor by rank, for the t-th
// entry.
struct PredE {
    uint32_t row;       // predecessor row (rank - beg)
    uint32_t weight;    // in_weight of that edge
    uint32_t ord;       // position in `incoming`
    uint32_t rperm;     // list position of the t-th predecessor in rank order
};

// Traceback record of one cell (the fill writes it, the traceback walks it):
//   bit 0  R    M > NEG/2 (the cell can be on a path)
//   bit 1  MD   M == D
//   bit 2  MI   M == I
//   bit 3  NO   I of the next column takes the open term:  I[j+1] == M[j] + o
//   bit 4  NE   I of the next column takes the extend term: I[j+1] == I[j] + e
//   then three fields of b bits: the list position of the `choose` winner
//   for the diagonal (M[p][j-1] + score == M), the deletion open
//   (M[p][j] + o == D) and the deletion extension (D[p][j] + e == D), or all
//   ones for none.
// b and the record's width follow the row's predecessor count np.
constexpr uint32_t kRecFlagR = 1, kRecFlagMD = 2, kRecFlagMI = 4, kRecFlagNO = 8, kRecFlagNE = 16;
constexpr uint32_t kMaxRecPreds = 65534;  // predecessors of one row a record can name
constexpr uint32_t kMaxNarrowPreds = 511;  // at most this many: records of 1, 2 or 4 bytes
// (constexpr: usable in host and device code)
constexpr uint32_t rec_field_bits(uint32_t np) {
    return np <= 1 ? 1u : np <= 7 ? 3u : np <= 511 ? 9u : 16u;
}
constexpr uint32_t rec_bytes(uint32_t np) {
    return np <= 1 ? 1u : np <= 7 ? 2u : np <= 511 ? 4u : 8u;
}

// Final adaptive band state of a computed row (src/align.rs:829-897): its
// M argmax column, match streak, band offsets and extra width. Successors
// fold these in rank order, which reproduces the reference's pushes.
// All fields are exact as int32 because the device path takes jobs with
// n + padded <= 2^29 (kMaxSpan): columns, ranks and offsets stay below 2^30,
// and `extra` saturates at 2^30 (kBandCap), past which every row is full
// width, so saturation changes no band.
struct RowRec {
    int32_t peak, streak, ofl, ofr, extra;
};

constexpr uint64_t kMaxSpan = uint64_t(1) << 29;  // n + padded of a device DP
// Rows of geometry (and adaptive records) the fill keeps in LDS; a row read
// from further back is found through global memory.
constexpr uint32_t kHistRows = 64;
constexpr int32_t kBandCap = int32_t(1) << 30;

// Row geometry written by the fill kernel, read by traceback.
constexpr uint64_t kNoValues = ~uint64_t(0);
struct RowMeta {
    uint32_t start;     // first column (multiple of 8)
    uint32_t len;       // width (multiple of 8); 0 for skipped rows
    uint64_t off;       // byte offset of the row's len traceback records in the arena
    uint64_t voff;      // byte offset of [M len][D len] scores, or kNoValues
    int64_t mlast;      // M of the row's last column (start + len - 1)
};

struct DevJob {
    uint64_t rows;      // RowIn index of row 0
    uint64_t aux;       // RowAux index of row 0
    uint64_t preds;     // PredE index of the job's first entry
    uint64_t seq;       // byte offset of the encoded, padded query
    uint64_t matrix;    // int32 index of the alpha x alpha matrix
    uint64_t arena;     // byte offset of the job's own arena region
    uint64_t arena_cap; // bytes in that region (more come from the pool)
    uint64_t rowmeta;   // RowMeta index of row 0
    uint64_t adapt;     // RowRec index of row 0 (adaptive jobs)
    uint64_t tb;        // uint32 index of the m-entry traceback record
    uint32_t n, m, padded, alpha;
    int64_t band;
    int32_t o, e;       // gap_open + gap_ext, gap_ext (src/align.rs:621-622)
    uint32_t unknown;   // encode('N')
    int32_t score_nn;   // score(N, N)
    uint32_t flags;
    int32_t band32;     // min(band, kBandCap)
    int32_t cap5;       // min(5 * band, kBandCap): the bound of `extra`
    uint32_t pad_;
};

enum DevStatus : int32_t {
    kOk = 0,
    kBandNoIntersect = 5,
    kBandExcludesSource = 6,
    kTbLeftBand = 7,
    kTbNoAdvance = 8,
    kTbNoPredecessor = 9,
    kArenaOverflow = 100,
    kTooManyPreds = 101,  // a row with more than kMaxRecPreds predecessors
    kLostScores = 102,    // a predecessor's scores were neither in LDS nor kept in HBM
    kSkipped = 103,       // no DP in this launch for the job (n == 0: resident step loops)
};

struct DevResult {
    int32_t status;
    uint32_t err_i, err_j;
    uint32_t pad_;
    int64_t score;      // M at (n-1, m-1)
    uint64_t cells;     // sum of computed row widths
    uint64_t arena_used;  // arena bytes the DP took
    uint64_t pred_cells;  // sum over rows and predecessors of the band overlap
};

}  // namespace mpoa
