// Launch interface of the gfx950 DP kernels (kernels.hip).
#pragma once
#include <hip/hip_runtime_api.h>

#include <cstdint>

#include "dp_types.hpp"

namespace mpoa {

// Device addresses of one launch's pools (all indices inside DevJob are
// relative to these bases).
struct DevPools {
    const DevJob *jobs;
    DevResult *res;
    const RowIn *rows;
    const RowAux *aux;
    const PredE *preds;
    const uint8_t *seq;
    const int32_t *matrix;
    void *arena;         // byte-addressed: traceback records and kept scores
    RowMeta *meta;
    RowRec *adapt;
    uint32_t *tb;
    // Arena chunk pool shared by the launch: a job whose own region fills up
    // takes further chunks from *pool_next (element offsets) up to pool_end.
    unsigned long long *pool_next;
    uint64_t pool_end;
    uint64_t chunk;      // bytes per chunk (>= 3 widest padded rows with scores)
    uint32_t seq_lds;    // bytes of LDS for the query codes (0 = global)
    uint32_t mat_lds;    // bytes of LDS for the score matrix (0 = global)
    uint32_t ring_bytes; // LDS ring of recent rows (multiple of 64 bytes)
    uint32_t prio;       // nonzero: the launch's waves issue at raised priority
};

// Dynamic LDS bytes a fill launch needs besides its static arrays.
uint32_t fill_dynamic_lds(uint32_t ring_bytes, uint32_t seq_lds, uint32_t mat_lds);
// Largest alphabet whose matrix is staged in LDS.
constexpr uint32_t kMatLdsEntries = 640;

// Fill launch shape: `waves` waves per job, `cols` columns per lane; one
// stripe of a row spans 64 * cols * waves columns.
struct FillShape {
    int cols, waves;
};
// Shapes with a compiled kernel: (8,1) (16,1) (4,2) (8,2) (4,4) (8,4).
bool fill_shape_ok(FillShape f);
// Shape for an expected row width.
FillShape fill_shape_for_width(uint32_t width);
// Default ring bytes for rows of a given width (about three rows, 16-48 KB).
uint32_t ring_bytes_for_width(uint32_t width, bool i64);
// wide: some row of the launch has more than kMaxNarrowPreds predecessors
// (8-byte records; shape (8,1) only). Returns -1 for a shape without a
// compiled kernel.
// rec: the records layout, else values.
int launch_fill(const DevPools &P, uint32_t njobs, bool i64, bool rec, bool wide, FillShape f, hipStream_t s);
void launch_traceback(const DevPools &P, uint32_t njobs, bool i64, bool rec, hipStream_t s);

// Fixed device scratch of the no-allocation entry point (64 MiB per GPU,
// a static device array of kernels.hip). scratch_init looks up its address
// on `device` (once per device; later calls return the first result) and
// scratch_device_ptr returns it (nullptr before a successful scratch_init).
constexpr uint64_t kScratchBytes = uint64_t(64) << 20;
constexpr uint32_t kScratchDevices = 64;
hipError_t scratch_init(int device);
void *scratch_device_ptr(int device);

}  // namespace mpoa
