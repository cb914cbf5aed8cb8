// mmf_ctx.hpp -- the device-resident state of one grid behind mmf_ctx.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <utility>
#include <vector>

#include "mmf.h"
#include "mmf_internal.hpp"
#include "mmf_learn.hpp"

using mmf::Ctrl;
using mmf::GridDev;

struct mmf_output;  // VTK writer state (mmf_output.cpp)

struct mmf_ctx {
  mmf_grid_desc desc;
  GridDev g;
  int device = 0;
  int num_cus = 0;
  hipStream_t stream = nullptr;
  double *dU = nullptr, *dR = nullptr, *dW = nullptr;
  uint8_t *dflags = nullptr;
  Ctrl *dctrl = nullptr;
  Ctrl *hctrl = nullptr;  // pinned mirror for read-back
  double *dpartials = nullptr;
  int64_t nerr_blocks = 0;
  double *derr = nullptr;
  std::vector<int64_t> raw;  // cell_of_raw, empty when identity
  int64_t *draw = nullptr;   // the same on the device (null when identity)
  int kernel = mmf::KERNEL_AUTO;
  mmf::MarchShape shape{8, 32};
  std::string err;
  // kernel timing: 0 off, 1 an event pair around every stage launch, 2 one
  // pair around each mmf_advance call (steps still replayed from graphs)
  int timing = 0;
  struct TimedSpan {
    hipEvent_t start, end;
    int64_t launches;  // stage launches between the two events
  };
  std::vector<TimedSpan> events;
  double timed_ms = 0.0;
  int64_t timed_launches = 0;
  // learned face reconstruction (2D, order 2): active when learn_on
  bool learn_on = false;
  mmf::LearnRun learn{};
  bool learn_opts_set = false;
  // captured step graphs (1 step, GRAPH_STEPS steps) and the launch
  // configuration they were captured with
  hipGraph_t step_graph[2] = {nullptr, nullptr};
  hipGraphExec_t step_exec[2] = {nullptr, nullptr};
  std::vector<char> graph_key;
  bool use_graphs = true;
  // Ctrl::lam_now holds the early maximum of the current state (the 3D
  // marching kernel keeps it up to date from step to step; anything else
  // that changes the state, the time control or the kernel clears this)
  bool lam_valid = false;
  // cell flags (cell order) kept on the host for the 3D marching kernel's
  // tile table (empty when every cell is fluid and solved), and the table
  std::vector<uint8_t> hflags;
  int32_t *dtiles = nullptr;
  int64_t dtiles_n = 0;
  bool tiles_valid = false;
  // per-face border conditions (GridDev::bslot and its slot tables)
  uint8_t *dbslot = nullptr;
  int32_t *dbslot_bc = nullptr;
  double *dbslot_info = nullptr;
  // one stage at a time (mmf_rk_*): the stage maxima of the current step
  // and the stages updated so far (0-2)
  double rk_eig[3] = {0.0, 0.0, 0.0};
  int rk_done = 0;
  // ghost-list staging (mmf_pack_cells / mmf_unpack_cells)
  void *dpack = nullptr;
  size_t dpack_bytes = 0;
  double *dFS = nullptr;
  unsigned long long *dlcount = nullptr;
  mmf_output *out = nullptr;
};

// Enqueues the RHS of the current state into dR (stage 1, RHS only).
int mmf_ctx_enqueue_rhs(mmf_ctx *ctx);
// Waits for and releases the VTK writer of a context (mmf_output.cpp).
void mmf_output_release(mmf_ctx *ctx);

// releases the captured step graphs of a context (defined with the C-ABI)
extern "C" void mmf_drop_step_graphs(mmf_ctx *ctx);
