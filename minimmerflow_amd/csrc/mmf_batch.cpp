// mmf_batch.cpp -- C-ABI of the batched mode: many independent 2D grids of
// uneven sizes advanced together by one persistent kernel per RK stage, the
// whole step captured once in a HIP graph and replayed.
//
// Layout: the cases' SoA states are concatenated in three device buffers
// (U, R, W); case i starts at offset_i with its own field stride.  Each case
// has its own time-control block, so its dt follows its own maximum
// eigenvalue (main.cpp:398-402) and a case that reaches its end time stops
// while the others continue.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "mmf.h"
#include "mmf_batch.hpp"
#include "mmf_internal.hpp"

using mmf::BatchCase;
using mmf::Ctrl;
using mmf::GridDev;

struct mmf_batch {
  int ncases = 0;
  std::vector<mmf_grid_desc> desc;
  std::vector<BatchCase> hcases;
  std::vector<int64_t> ncell;
  int64_t total_doubles = 0;
  int64_t total_cells = 0;
  int num_cus = 0;
  hipStream_t stream = nullptr;
  double *dU = nullptr, *dR = nullptr, *dW = nullptr;
  Ctrl *dctrl = nullptr;
  Ctrl *hctrl = nullptr;
  BatchCase *dcases = nullptr;
  int *ditems = nullptr;
  int nitems = 0;
  mmf::BatchQueues queues{};
  unsigned int *dcounters = nullptr;
  int chunk = 16;
  int tail = 0;  // rows of the small items at the end of the list (0: none)
  int blocks = 0;
  int occ[2] = {0, 0};  // resident blocks per CU of stages 1 and 2
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  double *dpartials = nullptr;
  double *derr = nullptr;
  std::string err;
  bool timing = false;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> events;
  double timed_ms = 0.0;
  int64_t timed_steps = 0;
};

namespace {

int bfail(mmf_batch *b, int code, const std::string &m) {
  if (b) b->err = m;
  return code;
}

#define BCHK(b, expr)                                                                 \
  do {                                                                                \
    hipError_t e_ = (expr);                                                           \
    if (e_ != hipSuccess) return bfail(b, MMF_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

void drop_graph(mmf_batch *b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
}

// (re)build the work list and the case table for the current shape.  An item
// is (case, strip, first row, end row).  With a tail, the last 'chunk' rows of
// every strip become items of 'tail' rows, queued after all full items, so
// each stage drains in small pieces instead of ending on the slowest full
// item.
// The work items (case, strip, first row, end row) in BATCH_QUEUES queues,
// one per XCD.  In case order, row band by row band, strip by strip, the
// items are cut into contiguous runs of equal cost (cells plus the two halo
// rows an item reads), one run per queue, so the strips of a row band and
// the bands above and below it go to one XCD, whose L2 then serves the lines
// they share.  Each queue holds its full items (chunk rows) first and then,
// with a tail, each strip's last chunk rows as items of tail rows.
int build_items(mmf_batch *b) {
  struct Item { int c, t, m0, m1; bool small; };
  std::vector<Item> all;
  int64_t total = 0;
  for (int i = 0; i < b->ncases; ++i) {
    BatchCase &c = b->hcases[(size_t)i];
    c.ntx = mmf::march2d_tiles(c.g.n[0]);
    c.chunk = b->chunk;
    const int ny = (int)c.g.n[1];
    const int split = (b->tail > 0) ? std::max(0, ny - b->chunk) : ny;  // rows of full items
    for (int m0 = 0; m0 < split; m0 += b->chunk)
      for (int t = 0; t < c.ntx; ++t) all.push_back({i, t, m0, std::min(m0 + b->chunk, split), false});
    for (int m0 = split; m0 < ny; m0 += b->tail)
      for (int t = 0; t < c.ntx; ++t) all.push_back({i, t, m0, std::min(m0 + b->tail, ny), true});
  }
  for (const Item &it : all) total += it.m1 - it.m0 + 2;
  std::vector<int> q[mmf::BATCH_QUEUES][2];  // per queue: full, then small items
  int64_t acc = 0;
  for (const Item &it : all) {
    const int k = (int)std::min<int64_t>(mmf::BATCH_QUEUES - 1, acc * mmf::BATCH_QUEUES / std::max<int64_t>(total, 1));
    q[k][it.small ? 1 : 0].insert(q[k][it.small ? 1 : 0].end(), {it.c, it.t, it.m0, it.m1});
    acc += it.m1 - it.m0 + 2;
  }
  std::vector<int> items;
  for (int k = 0; k < mmf::BATCH_QUEUES; ++k) {
    b->queues.beg[k] = (int)(items.size() / 4);
    for (int h = 0; h < 2; ++h) items.insert(items.end(), q[k][h].begin(), q[k][h].end());
  }
  b->queues.beg[mmf::BATCH_QUEUES] = (int)(items.size() / 4);
  b->nitems = (int)(items.size() / 4);
  if (b->ditems) (void)hipFree(b->ditems);
  b->ditems = nullptr;
  BCHK(b, hipMalloc(&b->ditems, sizeof(int) * std::max<size_t>(items.size(), 4)));
  if (!items.empty())
    BCHK(b, hipMemcpy(b->ditems, items.data(), sizeof(int) * items.size(), hipMemcpyHostToDevice));
  BCHK(b, hipMemcpy(b->dcases, b->hcases.data(), sizeof(BatchCase) * (size_t)b->ncases, hipMemcpyHostToDevice));
  drop_graph(b);
  return MMF_OK;
}

// Persistent blocks of stage s: 'blocks' for stages 2 and 3; stage 1, which
// holds three waves per SIMD where they hold two, proportionally more.
int stage_blocks(const mmf_batch *b, int s) {
  if (s != 1 || b->occ[0] <= 0 || b->occ[1] <= 0) return b->blocks;
  return (int)((int64_t)b->blocks * b->occ[0] / b->occ[1]);
}

int enqueue_step(mmf_batch *b) {
  for (int s = 1; s <= 3; ++s) {
    hipError_t e = mmf::launch_batch_stage(s, b->dcases, b->ditems, b->queues, b->dcounters, b->dU, b->dR, b->dW,
                                           stage_blocks(b, s), b->stream);
    if (e != hipSuccess) return bfail(b, MMF_ERR_HIP, std::string("batch stage: ") + hipGetErrorString(e));
  }
  hipError_t e = mmf::launch_batch_advance(b->dcases, b->ncases, b->dcounters, b->stream);
  if (e != hipSuccess) return bfail(b, MMF_ERR_HIP, std::string("batch advance: ") + hipGetErrorString(e));
  return MMF_OK;
}

// One RK3 step (three persistent stage launches + the per-case t update),
// captured once; the graph holds kernel nodes only.
int ensure_graph(mmf_batch *b) {
  if (b->exec) return MMF_OK;
  BCHK(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  int rc = enqueue_step(b);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(b->stream, &g);
  if (rc != MMF_OK) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return bfail(b, MMF_ERR_HIP, std::string("end capture: ") + hipGetErrorString(e));
  b->graph = g;
  BCHK(b, hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
  return MMF_OK;
}

int read_ctrls(mmf_batch *b) {
  BCHK(b, hipMemcpyAsync(b->hctrl, b->dctrl, sizeof(Ctrl) * (size_t)b->ncases, hipMemcpyDeviceToHost, b->stream));
  BCHK(b, hipStreamSynchronize(b->stream));
  return MMF_OK;
}

// MMF_ERR_STATE once a case's step met a NaN or infinite stage maximum (the
// control blocks read last; that case's time no longer advances).
int batch_nonfinite_status(mmf_batch *b) {
  for (int i = 0; i < b->ncases; ++i)
    if (b->hctrl[i].nonfinite)
      return bfail(b, MMF_ERR_STATE, "case " + std::to_string(i) +
                                         " met a non-finite state (NaN or infinity); its time stays at " +
                                         std::to_string(b->hctrl[i].t));
  return MMF_OK;
}

int collect(mmf_batch *b) {
  for (auto &p : b->events) {
    BCHK(b, hipEventSynchronize(p.second));
    float ms = 0.f;
    BCHK(b, hipEventElapsedTime(&ms, p.first, p.second));
    b->timed_ms += ms;
    (void)hipEventDestroy(p.first);
    (void)hipEventDestroy(p.second);
  }
  b->events.clear();
  return MMF_OK;
}

}  // namespace

extern "C" {

int mmf_batch_create(const mmf_grid_desc *cases, int32_t ncases, mmf_batch **out) {
  if (!cases || ncases < 1 || !out) return MMF_ERR_INVALID;
  *out = nullptr;
  for (int i = 0; i < ncases; ++i) {
    const mmf_grid_desc &d = cases[i];
    if (d.dim != 2 || !(d.h > 0.0)) return MMF_ERR_INVALID;
    if (d.order != 1) return MMF_ERR_UNSUPPORTED_ORDER;
    for (int a = 0; a < 2; ++a) {
      if (d.n[a] < 1 || d.n[a] >= (int64_t)1 << 24) return MMF_ERR_INVALID;
      int lo = d.side_bc[2 * a], hi = d.side_bc[2 * a + 1];
      if (lo < MMF_BC_FREE_FLOW || lo > MMF_BC_PERIODIC || hi < MMF_BC_FREE_FLOW || hi > MMF_BC_PERIODIC)
        return MMF_ERR_INVALID;
      if ((lo == MMF_BC_PERIODIC) != (hi == MMF_BC_PERIODIC)) return MMF_ERR_INVALID;
      if (lo == MMF_BC_PERIODIC && d.n[a] < 2) return MMF_ERR_INVALID;
    }
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return MMF_ERR_NO_DEVICE;
  mmf_batch *b = new (std::nothrow) mmf_batch();
  if (!b) return MMF_ERR_OUT_OF_MEMORY;
  auto cleanup = [&](int code) {
    mmf_batch_destroy(b);
    return code;
  };
  b->ncases = ncases;
  b->desc.assign(cases, cases + ncases);
  b->hcases.resize((size_t)ncases);
  b->ncell.resize((size_t)ncases);
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return cleanup(MMF_ERR_NO_DEVICE);
  if (hipDeviceGetAttribute(&b->num_cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) b->num_cus = 256;
  b->blocks = 8 * b->num_cus;
  b->occ[0] = mmf::batch_stage_occupancy(1);
  b->occ[1] = mmf::batch_stage_occupancy(2);
  int64_t off = 0;
  for (int i = 0; i < ncases; ++i) {
    const mmf_grid_desc &d = cases[i];
    BatchCase &c = b->hcases[(size_t)i];
    std::memset(&c, 0, sizeof(c));
    GridDev &g = c.g;
    g.dim = 2;
    g.order = 1;
    g.n[0] = d.n[0];
    g.n[1] = d.n[1];
    g.n[2] = 1;
    g.N = g.n[0] * g.n[1];
    g.fs = (g.N + 63) / 64 * 64;
    g.h = d.h;
    g.V = d.h * d.h;
    g.A = d.h;
    for (int s = 0; s < 6; ++s) {
      g.bc[s] = (s < 4) ? d.side_bc[s] : MMF_BC_FREE_FLOW;
      for (int k = 0; k < 5; ++k) g.info[s][k] = d.side_info[s][k];
    }
    g.periodic[0] = d.side_bc[0] == MMF_BC_PERIODIC;
    g.periodic[1] = d.side_bc[2] == MMF_BC_PERIODIC;
    g.flags = nullptr;
    int e2 = 0;
    g.pow2V = std::frexp(g.V, &e2) == 0.5 ? 1 : 0;
    g.invV = 1.0 / g.V;
    g.pow2A = std::frexp(g.A, &e2) == 0.5 ? 1 : 0;
    g.halfA = 0.5 * g.A;
    c.offset = off;
    off += 5 * g.fs;
    b->ncell[(size_t)i] = g.N;
    b->total_cells += g.N;
  }
  b->total_doubles = off;
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) return cleanup(MMF_ERR_HIP);
  const size_t bytes = sizeof(double) * (size_t)off;
  if (hipMalloc(&b->dU, bytes) != hipSuccess || hipMalloc(&b->dR, bytes) != hipSuccess ||
      hipMalloc(&b->dW, bytes) != hipSuccess)
    return cleanup(MMF_ERR_OUT_OF_MEMORY);
  if (hipMalloc(&b->dctrl, sizeof(Ctrl) * (size_t)ncases) != hipSuccess) return cleanup(MMF_ERR_OUT_OF_MEMORY);
  if (hipHostMalloc(&b->hctrl, sizeof(Ctrl) * (size_t)ncases, hipHostMallocDefault) != hipSuccess)
    return cleanup(MMF_ERR_OUT_OF_MEMORY);
  if (hipMalloc(&b->dcases, sizeof(BatchCase) * (size_t)ncases) != hipSuccess) return cleanup(MMF_ERR_OUT_OF_MEMORY);
  if (hipMalloc(&b->dcounters, sizeof(unsigned int) * mmf::BATCH_COUNTERS) != hipSuccess)
    return cleanup(MMF_ERR_OUT_OF_MEMORY);
  if (hipMemset(b->dU, 0, bytes) != hipSuccess || hipMemset(b->dR, 0, bytes) != hipSuccess ||
      hipMemset(b->dW, 0, bytes) != hipSuccess || hipMemset(b->dctrl, 0, sizeof(Ctrl) * (size_t)ncases) != hipSuccess ||
      hipMemset(b->dcounters, 0, sizeof(unsigned int) * mmf::BATCH_COUNTERS) != hipSuccess)
    return cleanup(MMF_ERR_HIP);
  std::memset(b->hctrl, 0, sizeof(Ctrl) * (size_t)ncases);
  for (int i = 0; i < ncases; ++i) b->hcases[(size_t)i].ctrl = b->dctrl + i;
  if (build_items(b) != MMF_OK) return cleanup(MMF_ERR_HIP);
  *out = b;
  return MMF_OK;
}

int mmf_batch_destroy(mmf_batch *b) {
  if (!b) return MMF_OK;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  for (auto &p : b->events) {
    (void)hipEventDestroy(p.first);
    (void)hipEventDestroy(p.second);
  }
  drop_graph(b);
  for (void *p : {(void *)b->dU, (void *)b->dR, (void *)b->dW, (void *)b->dctrl, (void *)b->dcases,
                  (void *)b->ditems, (void *)b->dcounters, (void *)b->dpartials, (void *)b->derr})
    if (p) (void)hipFree(p);
  if (b->hctrl) (void)hipHostFree(b->hctrl);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return MMF_OK;
}

const char *mmf_batch_last_error(const mmf_batch *b) { return b ? b->err.c_str() : "null batch"; }

int mmf_batch_sizes(const mmf_batch *b, int64_t *total_cells, int32_t *nitems) {
  if (!b) return MMF_ERR_INVALID;
  if (total_cells) *total_cells = b->total_cells;
  if (nitems) *nitems = b->nitems;
  return MMF_OK;
}

int mmf_batch_set_state(mmf_batch *b, int32_t ci, const double *U) {
  if (!b || !U || ci < 0 || ci >= b->ncases) return MMF_ERR_INVALID;
  const BatchCase &c = b->hcases[(size_t)ci];
  const int64_t N = c.g.N, fs = c.g.fs;
  std::vector<double> soa((size_t)(5 * fs), 0.0);
  for (int64_t i = 0; i < N; ++i)
    for (int k = 0; k < 5; ++k) soa[(size_t)(k * fs + i)] = U[5 * i + k];
  BCHK(b, hipStreamSynchronize(b->stream));
  BCHK(b, hipMemcpy(b->dU + c.offset, soa.data(), sizeof(double) * soa.size(), hipMemcpyHostToDevice));
  BCHK(b, hipMemcpy(b->dW + c.offset, b->dU + c.offset, sizeof(double) * soa.size(), hipMemcpyDeviceToDevice));
  return MMF_OK;
}

int mmf_batch_get_state(mmf_batch *b, int32_t ci, double *U) {
  if (!b || !U || ci < 0 || ci >= b->ncases) return MMF_ERR_INVALID;
  const BatchCase &c = b->hcases[(size_t)ci];
  const int64_t N = c.g.N, fs = c.g.fs;
  std::vector<double> soa((size_t)(5 * fs));
  BCHK(b, hipStreamSynchronize(b->stream));
  BCHK(b, hipMemcpy(soa.data(), b->dU + c.offset, sizeof(double) * soa.size(), hipMemcpyDeviceToHost));
  for (int64_t i = 0; i < N; ++i)
    for (int k = 0; k < 5; ++k) U[5 * i + k] = soa[(size_t)(k * fs + i)];
  return MMF_OK;
}

int mmf_batch_set_initial_states(mmf_batch *b, const mmf_problem *problems) {
  if (!b || !problems) return MMF_ERR_INVALID;
  for (int i = 0; i < b->ncases; ++i) {
    const mmf_grid_desc &d = b->desc[(size_t)i];
    const int64_t N = b->ncell[(size_t)i];
    std::vector<double> pts((size_t)(3 * N)), U((size_t)(5 * N));
    for (int64_t c = 0; c < N; ++c) {
      pts[(size_t)(3 * c)] = d.origin[0] + ((c % d.n[0]) + 0.5) * d.h;
      pts[(size_t)(3 * c + 1)] = d.origin[1] + ((c / d.n[0]) + 0.5) * d.h;
      pts[(size_t)(3 * c + 2)] = d.origin[2];
    }
    int rc = mmf_problem_exact(&problems[i], pts.data(), N, 0.0, U.data());
    if (rc != MMF_OK) return bfail(b, rc, "problem of case " + std::to_string(i));
    rc = mmf_batch_set_state(b, i, U.data());
    if (rc != MMF_OK) return rc;
  }
  return MMF_OK;
}

int mmf_batch_set_time(mmf_batch *b, const double *t0, const double *tmax, const double *cfl) {
  if (!b || !t0 || !tmax || !cfl) return MMF_ERR_INVALID;
  for (int i = 0; i < b->ncases; ++i) {
    if (!(cfl[i] > 0.0)) return bfail(b, MMF_ERR_INVALID, "cfl must be positive");
    Ctrl &c = b->hctrl[i];
    std::memset(&c, 0, sizeof(c));
    c.t = t0[i];
    c.tmax = tmax[i];
    c.cfl = cfl[i];
    c.hmin = b->hcases[(size_t)i].g.h;
  }
  BCHK(b, hipMemcpyAsync(b->dctrl, b->hctrl, sizeof(Ctrl) * (size_t)b->ncases, hipMemcpyHostToDevice, b->stream));
  BCHK(b, hipStreamSynchronize(b->stream));
  return MMF_OK;
}

int mmf_batch_set_shape(mmf_batch *b, int32_t chunk, int32_t blocks) {
  if (!b) return MMF_ERR_INVALID;
  if (chunk > 0 && chunk != b->chunk) {
    b->chunk = chunk;
    int rc = build_items(b);
    if (rc != MMF_OK) return rc;
  }
  if (blocks > 0 && blocks != b->blocks) {
    b->blocks = blocks;
    drop_graph(b);
  }
  return MMF_OK;
}

int mmf_batch_set_tail(mmf_batch *b, int32_t rows) {
  if (!b || rows < 0) return MMF_ERR_INVALID;
  if (rows != b->tail) {
    b->tail = rows;
    return build_items(b);
  }
  return MMF_OK;
}

int mmf_batch_get_tail(const mmf_batch *b, int32_t *rows) {
  if (!b || !rows) return MMF_ERR_INVALID;
  *rows = b->tail;
  return MMF_OK;
}

int mmf_batch_get_shape(const mmf_batch *b, int32_t *chunk, int32_t *blocks) {
  if (!b) return MMF_ERR_INVALID;
  if (chunk) *chunk = b->chunk;
  if (blocks) *blocks = b->blocks;
  return MMF_OK;
}

int mmf_batch_advance(mmf_batch *b, int64_t nsteps) {
  if (!b || nsteps < 0) return MMF_ERR_INVALID;
  int rc = ensure_graph(b);
  if (rc != MMF_OK) return rc;
  for (int64_t s = 0; s < nsteps; ++s) {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (b->timing) {
      BCHK(b, hipEventCreate(&e0));
      BCHK(b, hipEventCreate(&e1));
      BCHK(b, hipEventRecord(e0, b->stream));
    }
    BCHK(b, hipGraphLaunch(b->exec, b->stream));
    if (b->timing) {
      BCHK(b, hipEventRecord(e1, b->stream));
      b->events.emplace_back(e0, e1);
      b->timed_steps += 1;
    }
  }
  return MMF_OK;
}

int mmf_batch_run(mmf_batch *b, int64_t max_steps, int64_t *steps_done) {
  if (!b) return MMF_ERR_INVALID;
  int64_t done = 0;
  for (;;) {
    int rc = read_ctrls(b);
    if (rc != MMF_OK) return rc;
    double rem = 0.0;
    bool any = false;
    for (int i = 0; i < b->ncases; ++i) {
      const Ctrl &c = b->hctrl[i];
      if (c.nonfinite) {
        if (steps_done) *steps_done = done;
        return batch_nonfinite_status(b);
      }
      if (c.t < c.tmax) {
        any = true;
        rem = std::max(rem, c.dt > 0.0 ? std::ceil((c.tmax - c.t) / c.dt) : 1.0);
      }
    }
    if (!any || (max_steps >= 0 && done >= max_steps)) break;
    int64_t batch = (int64_t)std::min(std::max(rem, 1.0), 64.0);
    if (max_steps >= 0) batch = std::min(batch, max_steps - done);
    rc = mmf_batch_advance(b, batch);
    if (rc != MMF_OK) return rc;
    done += batch;
  }
  if (steps_done) *steps_done = done;
  return MMF_OK;
}

int mmf_batch_get_times(mmf_batch *b, double *t, int64_t *steps) {
  if (!b) return MMF_ERR_INVALID;
  int rc = read_ctrls(b);
  if (rc != MMF_OK) return rc;
  for (int i = 0; i < b->ncases; ++i) {
    if (t) t[i] = b->hctrl[i].t;
    if (steps) steps[i] = b->hctrl[i].step;
  }
  return batch_nonfinite_status(b);
}

int mmf_batch_error_vs_exact(mmf_batch *b, const mmf_problem *problems, double *l1, double *l2) {
  if (!b || !problems) return MMF_ERR_INVALID;
  int rc = read_ctrls(b);
  if (rc != MMF_OK) return rc;
  for (int i = 0; i < b->ncases; ++i) {
    const mmf_grid_desc &d = b->desc[(size_t)i];
    const int64_t N = b->ncell[(size_t)i];
    std::vector<double> U((size_t)(5 * N)), pts((size_t)(3 * N)), ex((size_t)(5 * N));
    rc = mmf_batch_get_state(b, i, U.data());
    if (rc != MMF_OK) return rc;
    for (int64_t c = 0; c < N; ++c) {
      pts[(size_t)(3 * c)] = d.origin[0] + ((c % d.n[0]) + 0.5) * d.h;
      pts[(size_t)(3 * c + 1)] = d.origin[1] + ((c / d.n[0]) + 0.5) * d.h;
      pts[(size_t)(3 * c + 2)] = d.origin[2];
    }
    rc = mmf_problem_exact(&problems[i], pts.data(), N, b->hctrl[i].t, ex.data());
    if (rc != MMF_OK) return rc;
    const double V = d.h * d.h;
    double e1 = 0.0, e2 = 0.0;
    for (int64_t c = 0; c < N; ++c) {  // main.cpp:553-561, in cell order
      double dd = U[(size_t)(5 * c)] - ex[(size_t)(5 * c)];
      e1 += std::fabs(dd) * V;
      e2 += dd * dd * V;
    }
    if (l1) l1[i] = e1;
    if (l2) l2[i] = std::sqrt(e2);
  }
  return MMF_OK;
}

int mmf_batch_enable_timing(mmf_batch *b, int32_t on) {
  if (!b) return MMF_ERR_INVALID;
  int rc = collect(b);
  if (rc != MMF_OK) return rc;
  b->timing = on != 0;
  b->timed_ms = 0.0;
  b->timed_steps = 0;
  return MMF_OK;
}

int mmf_batch_read_timing(mmf_batch *b, double *ms, int64_t *steps) {
  if (!b) return MMF_ERR_INVALID;
  int rc = collect(b);
  if (rc != MMF_OK) return rc;
  if (ms) *ms = b->timed_ms;
  if (steps) *steps = b->timed_steps;
  return MMF_OK;
}

// Launch-shape tuner: times each candidate (rows per item, persistent blocks)
// over a few graph-replayed steps on a snapshot of the state, restores the
// state and keeps the fastest.  Results do not depend on the shape.
int mmf_batch_autotune(mmf_batch *b, int32_t steps_per_trial, int32_t *best_chunk, int32_t *best_blocks,
                       double *best_ms_per_step) {
  if (!b || steps_per_trial < 1) return MMF_ERR_INVALID;
  const size_t bytes = sizeof(double) * (size_t)b->total_doubles;
  double *snapU = nullptr;
  BCHK(b, hipStreamSynchronize(b->stream));
  BCHK(b, hipMalloc(&snapU, bytes));
  std::vector<Ctrl> snapC((size_t)b->ncases);
  hipError_t e = hipMemcpy(snapU, b->dU, bytes, hipMemcpyDeviceToDevice);
  if (e == hipSuccess) e = hipMemcpy(snapC.data(), b->dctrl, sizeof(Ctrl) * snapC.size(), hipMemcpyDeviceToHost);
  if (e != hipSuccess) {
    (void)hipFree(snapU);
    return bfail(b, MMF_ERR_HIP, "autotune snapshot");
  }
  const int chunks[] = {4, 8, 16, 32, 64};
  const int per_cu[] = {4, 8, 12, 16};
  const int keep_chunk = b->chunk, keep_blocks = b->blocks, keep_tail = b->tail;
  const bool keep_timing = b->timing;
  int bc = keep_chunk, bb = keep_blocks, bt = keep_tail;
  double best = 1e300;
  int rc = MMF_OK;
  for (int trial = 0; trial < 2 * 5 * 4 && rc == MMF_OK; ++trial) {
    const int ch = chunks[trial / 8], pc = per_cu[trial % 4];
    const int tl = ((trial / 4) % 2) ? std::max(4, ch / 4) : 0;  // no tail, or items of a quarter
    if (tl >= ch) continue;
    {
      rc = mmf_batch_set_tail(b, tl);
      if (rc == MMF_OK) rc = mmf_batch_set_shape(b, ch, pc * b->num_cus);
      if (rc == MMF_OK) rc = mmf_batch_enable_timing(b, 1);
      if (rc == MMF_OK) rc = mmf_batch_advance(b, 1);  // warm-up (and graph build)
      if (rc == MMF_OK) rc = mmf_batch_enable_timing(b, 1);
      if (rc == MMF_OK) rc = mmf_batch_advance(b, steps_per_trial);
      double ms = 0.0;
      int64_t n = 0;
      if (rc == MMF_OK) rc = mmf_batch_read_timing(b, &ms, &n);
      if (rc == MMF_OK && n > 0 && ms / n < best) {
        best = ms / n;
        bc = ch;
        bb = pc * b->num_cus;
        bt = tl;
      }
      // restore the state for the next trial
      if (hipMemcpyAsync(b->dU, snapU, bytes, hipMemcpyDeviceToDevice, b->stream) != hipSuccess ||
          hipMemcpyAsync(b->dctrl, snapC.data(), sizeof(Ctrl) * snapC.size(), hipMemcpyHostToDevice, b->stream) !=
              hipSuccess ||
          hipStreamSynchronize(b->stream) != hipSuccess)
        rc = bfail(b, MMF_ERR_HIP, "autotune restore");
    }
  }
  (void)hipFree(snapU);
  if (rc != MMF_OK) return rc;
  rc = mmf_batch_set_tail(b, bt);
  if (rc == MMF_OK) rc = mmf_batch_set_shape(b, bc, bb);
  if (rc == MMF_OK) rc = mmf_batch_enable_timing(b, keep_timing ? 1 : 0);
  if (rc != MMF_OK) return rc;
  if (best_chunk) *best_chunk = bc;
  if (best_blocks) *best_blocks = bb;
  if (best_ms_per_step) *best_ms_per_step = best;
  return MMF_OK;
}

}  // extern "C"
