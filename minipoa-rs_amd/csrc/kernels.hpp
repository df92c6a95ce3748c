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
    void *arena;
    RowMeta *meta;
    RowRec *adapt;
    uint32_t *tb;
};

// Columns per lane for a row width (64 lanes x C columns per segment).
int lanes_columns_for_width(uint32_t width);
void launch_fill(const DevPools &P, uint32_t njobs, bool i64, int cols, hipStream_t s);
void launch_traceback(const DevPools &P, uint32_t njobs, bool i64, hipStream_t s);

}  // namespace mpoa
