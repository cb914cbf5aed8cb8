// mmf_batch.hpp -- device structures of the batched (many independent grids) mode.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mmf_internal.hpp"

namespace mmf {

// One independent case of a batch: its grid, its own time-control block, where
// its SoA state starts in the batch's shared state buffers, and its tiling.
struct BatchCase {
  GridDev g;       // g.fs is this case's field stride
  Ctrl *ctrl;      // device pointer into the batch's control array
  int64_t offset;  // first double of this case in U / R / W
  int32_t ntx;     // strips along x
  int32_t chunk;   // rows per item along y
};

hipError_t launch_stage_march2d(const GridDev &g, int chunk, int stage, int mode, const double *U, const double *R,
                                const double *W, double *oR, double *oW, double *oU, Ctrl *ctrl, hipStream_t s);
hipError_t launch_stage_march2d_o2(const GridDev &g, int chunk, int stage, int mode, const double *U, const double *R,
                                   const double *W, double *oR, double *oW, double *oU, Ctrl *ctrl, hipStream_t s);
int march2d_tiles(int64_t nx);
// Item queues of the persistent batch kernel: one per XCD (a block takes
// items from the queue of blockIdx.x % BATCH_QUEUES, the XCD it runs on, and
// then from the others); queue q holds items [beg[q], beg[q + 1]).  Each
// queue's counter sits on its own 128-byte line, per stage.
constexpr int BATCH_QUEUES = 8;
constexpr int BATCH_COUNTER_LINE = 32;  // unsigned ints per line
constexpr int BATCH_COUNTERS = 3 * BATCH_QUEUES * BATCH_COUNTER_LINE;
struct BatchQueues {
  int beg[BATCH_QUEUES + 1];
};
hipError_t launch_batch_stage(int stage, const BatchCase *cases, const int *items, const BatchQueues &queues,
                              unsigned int *counters, double *U, double *R, double *W, int nblocks, hipStream_t s);
// Resident blocks per CU of the stage's batch kernel (0 if unknown).
int batch_stage_occupancy(int stage);
hipError_t launch_batch_advance(const BatchCase *cases, int ncases, unsigned int *counters, hipStream_t s);

}  // namespace mmf
