// mmf_api.cpp -- C-ABI of minimmerflow_amd (include/mmf.h).
//
// Owns the device-resident state of one structured grid: SoA conservative
// state U, residual R, stage state W, optional cell flags and the time-control
// block.  No exception leaves this file: every entry point returns a status.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <map>
#include <new>
#include <string>
#include <utility>
#include <vector>

#include "mmf.h"
#include "mmf_ctx.hpp"
#include "mmf_internal.hpp"
#include "mmf_learn.hpp"


namespace {

// Every fluid cell is solved (no ghost cells): the marching kernel's early dt applies.
int32_t ghost_free_of(const std::vector<uint8_t> &flags) {
  for (uint8_t v : flags)
    if ((v & MMF_CELL_FLUID) && !(v & MMF_CELL_SOLVED)) return 0;
  return 1;
}

int fail(mmf_ctx *ctx, int code, const std::string &msg) {
  if (ctx) ctx->err = msg;
  return code;
}

int hip_fail(mmf_ctx *ctx, hipError_t e, const char *what) {
  std::string m = std::string(what) + ": " + hipGetErrorString(e);
  return fail(ctx, MMF_ERR_HIP, m);
}

#define HIPCHK(ctx, expr)                               \
  do {                                                  \
    hipError_t e_ = (expr);                             \
    if (e_ != hipSuccess) return hip_fail(ctx, e_, #expr); \
  } while (0)

bool valid_bc(int b) { return b >= MMF_BC_FREE_FLOW && b <= MMF_BC_PERIODIC; }

// Kernel actually used: MARCH where it applies; AUTO takes it only from the
// size where it beats the per-cell gather kernel (measured on MI355X; below
// it a grid gives the marching kernels too few blocks or rows to march):
// 2D order 1 from 2^18 cells (64^2 56 vs 28 us per step, 512^2 equal, 1024^2
// 137 vs 188), 3D order 1 from 2^20 (32^3 273 vs 37 us, 96^3 319 vs 193,
// 128^3 271 vs 379), 2D order 2 from 2^20 (below it the tile kernel, then the
// gather kernel).
bool use_march(const mmf_ctx *ctx) {
  if (ctx->kernel == mmf::KERNEL_GATHER || !mmf::march_shape_supported(ctx->g, ctx->shape)) return false;
  if (ctx->kernel == mmf::KERNEL_AUTO) {
    const int shift = (ctx->g.dim == 2 && ctx->g.order == 1) ? 18 : 20;
    if (ctx->g.N < ((int64_t)1 << shift)) return false;
  }
  return true;
}

bool tile_applies(const mmf_ctx *ctx) {
  return ctx->g.dim == 2 && ctx->g.order == 2 && !ctx->g.flags && !ctx->g.bslot;
}

// AUTO takes the tile kernel for 2D order-2 grids below this size (measured on
// MI355X, vortex: 128^2 25.7 vs 50.0 us per step for the gather kernel, 256^2
// 48.5 vs 58.0; at 512^2 the gather and strip kernels are faster, 125-128 vs
// 146, the tile kernel's fixed per-block latency no longer hidden).
constexpr int64_t TILE_AUTO_MAX_N = (int64_t)1 << 18;

// The stage kernel of the next launches.  With the learned reconstruction on:
// the tile kernel, or (GATHER) the face-state + flux kernel pair.
int kernel_used(const mmf_ctx *ctx) {
  if (ctx->learn_on) return ctx->kernel == mmf::KERNEL_GATHER ? mmf::KERNEL_GATHER : mmf::KERNEL_TILE;
  if (ctx->kernel == mmf::KERNEL_TILE) return tile_applies(ctx) ? mmf::KERNEL_TILE : mmf::KERNEL_GATHER;
  if (ctx->kernel == mmf::KERNEL_AUTO && tile_applies(ctx) && ctx->g.N < TILE_AUTO_MAX_N) return mmf::KERNEL_TILE;
  return use_march(ctx) ? mmf::KERNEL_MARCH : mmf::KERNEL_GATHER;
}

// One stage launch on states U (stage input of stage 1 and the RK base), R and
// W: stage 1 writes R, stage 2 W, stage 3 U.
int launch_stage_on(mmf_ctx *ctx, int stage, int mode, double *U, double *R, double *W) {
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (ctx->timing == 1) {
    HIPCHK(ctx, hipEventCreate(&e0));
    HIPCHK(ctx, hipEventCreate(&e1));
    HIPCHK(ctx, hipEventRecord(e0, ctx->stream));
  }
  hipError_t e;
  const int kind = kernel_used(ctx);
  if (ctx->learn_on && kind == mmf::KERNEL_GATHER) {
    // face states (learned or plain) through HBM, then flux + RK stage update
    e = mmf::launch_recon_learned(ctx->g, ctx->learn, stage, mode, U, R, W, ctx->dctrl, ctx->stream);
    if (e == hipSuccess)
      e = mmf::launch_flux_fs(ctx->g, ctx->learn, stage, mode, U, R, W, R, W, U, ctx->dctrl, ctx->stream);
  } else if (kind == mmf::KERNEL_TILE) {
    // one launch per stage, face states in LDS; plain face values without a network
    mmf::LearnRun plain{};
    plain.blend = 1.0;
    e = mmf::launch_stage_learned(ctx->g, ctx->learn_on ? ctx->learn : plain, stage, mode, U, R, W, R, W, U, ctx->dctrl,
                                  ctx->stream);
  } else if (kind == mmf::KERNEL_MARCH) {
    e = mmf::launch_stage_march(ctx->g, ctx->shape, stage, mode, U, R, W, R, W, U, ctx->dctrl, ctx->stream);
  } else {
    e = mmf::launch_stage(ctx->g, stage, mode, U, R, W, R, W, U, ctx->dctrl, ctx->stream);
  }
  if (e != hipSuccess) return hip_fail(ctx, e, "stage kernel launch");
  if (ctx->timing == 1) {
    HIPCHK(ctx, hipEventRecord(e1, ctx->stream));
    ctx->events.push_back({e0, e1, 1});
  }
  return MMF_OK;
}

int launch_stage_timed(mmf_ctx *ctx, int stage, int mode) {
  return launch_stage_on(ctx, stage, mode, ctx->dU, ctx->dR, ctx->dW);
}

int collect_timing(mmf_ctx *ctx) {
  for (auto &p : ctx->events) {
    HIPCHK(ctx, hipEventSynchronize(p.end));
    float ms = 0.f;
    HIPCHK(ctx, hipEventElapsedTime(&ms, p.start, p.end));
    ctx->timed_ms += ms;
    ctx->timed_launches += p.launches;
    (void)hipEventDestroy(p.start);
    (void)hipEventDestroy(p.end);
  }
  ctx->events.clear();
  return MMF_OK;
}

int read_ctrl(mmf_ctx *ctx) {
  HIPCHK(ctx, hipMemcpyAsync(ctx->hctrl, ctx->dctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return MMF_OK;
}

// MMF_ERR_STATE once a step met a NaN or infinite stage maximum (the control
// block read last; such steps leave the time unchanged).
int nonfinite_status(mmf_ctx *ctx) {
  const Ctrl &c = *ctx->hctrl;
  if (!c.nonfinite) return MMF_OK;
  char buf[160];
  std::snprintf(buf, sizeof(buf), "%llu step(s) met a non-finite state (NaN or infinity); t stays at %.17g",
                (unsigned long long)c.nonfinite, c.t);
  return fail(ctx, MMF_ERR_STATE, buf);
}

}  // namespace

namespace {
int ensure_tiles(mmf_ctx *ctx);
}  // namespace

namespace mmf {

// The closed-form grid of a descriptor (validated as mmf_create documents);
// no flags, no tile table, field stride = cells rounded up to 64.
int grid_from_desc(const mmf_grid_desc *d, GridDev &g) {
  if (!d) return MMF_ERR_INVALID;
  if (d->dim != 2 && d->dim != 3) return MMF_ERR_INVALID;
  if (d->order != 1 && d->order != 2) return MMF_ERR_UNSUPPORTED_ORDER;
  if (!(d->h > 0.0)) return MMF_ERR_INVALID;
  int64_t ncells = 1;
  for (int a = 0; a < d->dim; ++a) {
    // per-axis counts fit the kernels' 32-bit coordinates; the cell count
    // bounds the state allocation (5 doubles per cell, three buffers)
    if (d->n[a] < 1 || d->n[a] > ((int64_t)1 << 30)) return MMF_ERR_INVALID;
    if (ncells > ((int64_t)1 << 36) / d->n[a]) return MMF_ERR_INVALID;
    ncells *= d->n[a];
    int lo = d->side_bc[2 * a], hi = d->side_bc[2 * a + 1];
    if (!valid_bc(lo) || !valid_bc(hi)) return MMF_ERR_INVALID;
    if ((lo == MMF_BC_PERIODIC) != (hi == MMF_BC_PERIODIC)) return MMF_ERR_INVALID;
    if (lo == MMF_BC_PERIODIC && d->n[a] < 2) return MMF_ERR_INVALID;
  }
  std::memset(&g, 0, sizeof(g));
  g.dim = d->dim;
  g.order = d->order;
  for (int a = 0; a < 3; ++a) g.n[a] = (a < d->dim) ? d->n[a] : 1;
  g.N = g.n[0] * g.n[1] * g.n[2];
  g.fs = (g.N + 63) / 64 * 64;
  g.h = d->h;
  g.V = (d->dim == 3) ? d->h * d->h * d->h : d->h * d->h;
  g.A = (d->dim == 3) ? d->h * d->h : d->h;
  for (int s = 0; s < 6; ++s) {
    g.bc[s] = (s / 2 < d->dim) ? d->side_bc[s] : MMF_BC_FREE_FLOW;
    for (int k = 0; k < 5; ++k) g.info[s][k] = d->side_info[s][k];
  }
  for (int a = 0; a < 3; ++a) g.periodic[a] = (a < d->dim) && d->side_bc[2 * a] == MMF_BC_PERIODIC;
  g.flags = nullptr;
  g.ghost_free = 1;
  int e2 = 0;
  g.pow2V = std::frexp(g.V, &e2) == 0.5 ? 1 : 0;
  g.invV = 1.0 / g.V;
  g.pow2A = std::frexp(g.A, &e2) == 0.5 ? 1 : 0;
  g.halfA = 0.5 * g.A;
  return MMF_OK;
}

}  // namespace mmf

int mmf_ctx_enqueue_rhs(mmf_ctx *ctx) {
  int rc = ensure_tiles(ctx);
  if (rc != MMF_OK) return rc;
  hipError_t e = mmf::launch_zero_slot(ctx->dctrl, 3, ctx->stream);
  if (e != hipSuccess) return hip_fail(ctx, e, "zero slot");
  return launch_stage_timed(ctx, 1, mmf::MODE_RHS_ONLY);
}

extern "C" {

const char *mmf_status_string(int s) {
  switch (s) {
    case MMF_OK: return "ok";
    case MMF_ERR_INVALID: return "invalid argument";
    case MMF_ERR_UNSUPPORTED_ORDER: return "unsupported reconstruction order";
    case MMF_ERR_HIP: return "HIP runtime error";
    case MMF_ERR_NO_DEVICE: return "no HIP device";
    case MMF_ERR_OUT_OF_MEMORY: return "out of device memory";
    case MMF_ERR_STATE: return "invalid state";
    default: return "unknown status";
  }
}

const char *mmf_last_error(const mmf_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int mmf_create(const mmf_grid_desc *d, mmf_ctx **out) {
  if (!d || !out) return MMF_ERR_INVALID;
  *out = nullptr;
  GridDev gd;
  const int rcg = mmf::grid_from_desc(d, gd);
  if (rcg != MMF_OK) return rcg;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return MMF_ERR_NO_DEVICE;

  mmf_ctx *ctx = new (std::nothrow) mmf_ctx();
  if (!ctx) return MMF_ERR_OUT_OF_MEMORY;
  ctx->desc = *d;
  GridDev &g = ctx->g;
  g = gd;
  auto cleanup = [&](int code) {
    mmf_destroy(ctx);
    return code;
  };
  if (hipGetDevice(&ctx->device) != hipSuccess) return cleanup(MMF_ERR_NO_DEVICE);
  if (hipDeviceGetAttribute(&ctx->num_cus, hipDeviceAttributeMultiprocessorCount, ctx->device) != hipSuccess)
    ctx->num_cus = 0;
  ctx->shape = mmf::default_march_shape(g, ctx->num_cus);
  mmf::balance_march_shape(g, ctx->num_cus, ctx->shape);
  if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) return cleanup(MMF_ERR_HIP);
  const size_t state_bytes = sizeof(double) * 5 * (size_t)g.fs;
  if (hipMalloc(&ctx->dU, state_bytes) != hipSuccess) return cleanup(MMF_ERR_OUT_OF_MEMORY);
  if (hipMalloc(&ctx->dR, state_bytes) != hipSuccess) return cleanup(MMF_ERR_OUT_OF_MEMORY);
  if (hipMalloc(&ctx->dW, state_bytes) != hipSuccess) return cleanup(MMF_ERR_OUT_OF_MEMORY);
  if (hipMalloc(&ctx->dctrl, sizeof(Ctrl)) != hipSuccess) return cleanup(MMF_ERR_OUT_OF_MEMORY);
  if (hipHostMalloc(&ctx->hctrl, sizeof(Ctrl), hipHostMallocDefault) != hipSuccess) return cleanup(MMF_ERR_OUT_OF_MEMORY);
  ctx->nerr_blocks = mmf::error_norm_blocks(g.N);
  if (hipMalloc(&ctx->dpartials, sizeof(double) * 2 * (size_t)ctx->nerr_blocks) != hipSuccess)
    return cleanup(MMF_ERR_OUT_OF_MEMORY);
  if (hipMalloc(&ctx->derr, sizeof(double) * 2) != hipSuccess) return cleanup(MMF_ERR_OUT_OF_MEMORY);
  if (hipMemsetAsync(ctx->dU, 0, state_bytes, ctx->stream) != hipSuccess ||
      hipMemsetAsync(ctx->dR, 0, state_bytes, ctx->stream) != hipSuccess ||
      hipMemsetAsync(ctx->dW, 0, state_bytes, ctx->stream) != hipSuccess ||
      hipMemsetAsync(ctx->dctrl, 0, sizeof(Ctrl), ctx->stream) != hipSuccess ||
      hipStreamSynchronize(ctx->stream) != hipSuccess)
    return cleanup(MMF_ERR_HIP);
  std::memset(ctx->hctrl, 0, sizeof(Ctrl));
  *out = ctx;
  return MMF_OK;
}

int mmf_destroy(mmf_ctx *ctx) {
  if (!ctx) return MMF_OK;
  mmf_output_release(ctx);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  for (auto &p : ctx->events) {
    (void)hipEventDestroy(p.start);
    (void)hipEventDestroy(p.end);
  }
  // the executable step graphs hold kernel nodes whose arguments point at the
  // buffers below: release them first, while every pointer is still valid
  mmf_drop_step_graphs(ctx);
  if (ctx->dU) (void)hipFree(ctx->dU);
  if (ctx->dR) (void)hipFree(ctx->dR);
  if (ctx->dW) (void)hipFree(ctx->dW);
  if (ctx->dflags) (void)hipFree(ctx->dflags);
  if (ctx->dctrl) (void)hipFree(ctx->dctrl);
  if (ctx->hctrl) (void)hipHostFree(ctx->hctrl);
  if (ctx->dpartials) (void)hipFree(ctx->dpartials);
  if (ctx->derr) (void)hipFree(ctx->derr);
  if (ctx->dFS) (void)hipFree(ctx->dFS);
  if (ctx->dlcount) (void)hipFree(ctx->dlcount);
  if (ctx->draw) (void)hipFree(ctx->draw);
  if (ctx->dtiles) (void)hipFree(ctx->dtiles);
  if (ctx->dbslot) (void)hipFree(ctx->dbslot);
  if (ctx->dbslot_bc) (void)hipFree(ctx->dbslot_bc);
  if (ctx->dbslot_info) (void)hipFree(ctx->dbslot_info);
  if (ctx->dpack) (void)hipFree(ctx->dpack);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return MMF_OK;
}

int mmf_set_kernel(mmf_ctx *ctx, int32_t kind, int32_t tile_rows, int32_t chunk) {
  if (!ctx) return MMF_ERR_INVALID;
  ctx->lam_valid = false;
  ctx->tiles_valid = false;
  if (kind < mmf::KERNEL_AUTO || kind > mmf::KERNEL_TILE) return fail(ctx, MMF_ERR_INVALID, "unknown kernel kind");
  if (kind == mmf::KERNEL_TILE && !tile_applies(ctx))
    return fail(ctx, MMF_ERR_INVALID, "tile kernel: order 2 in 2D without bodies");
  mmf::MarchShape sh = mmf::default_march_shape(ctx->g, ctx->num_cus);
  if (tile_rows > 0) sh.ty = tile_rows;
  if (chunk > 0) sh.chunk = chunk;
  if (tile_rows > 0 && sh.tx == 64 && tile_rows == 16 && ctx->g.dim == 3) sh.tx = 32;  // 64 x 16 does not fit; 32 x 16
  if (chunk <= 0) mmf::balance_march_shape(ctx->g, ctx->num_cus, sh);
  if (kind == mmf::KERNEL_MARCH && !mmf::march_shape_supported(ctx->g, sh))
    return fail(ctx, MMF_ERR_INVALID, "marching kernel: order 1, or order 2 in 2D without bodies; fewer than 2^29 cells; tile rows 4 or 8 in 3D, 1 in 2D");
  ctx->kernel = kind;
  ctx->shape = sh;
  return MMF_OK;
}

int mmf_set_march_shape(mmf_ctx *ctx, int32_t tile_x, int32_t tile_y, int32_t chunk) {
  if (!ctx) return MMF_ERR_INVALID;
  mmf::MarchShape sh = ctx->shape;
  const mmf::MarchShape def = mmf::default_march_shape(ctx->g, ctx->num_cus);
  sh.tx = tile_x > 0 ? tile_x : def.tx;
  sh.ty = tile_y > 0 ? tile_y : def.ty;
  sh.chunk = chunk > 0 ? chunk : def.chunk;
  if (chunk > 0) sh.seg_blocks = 0;
  else mmf::balance_march_shape(ctx->g, ctx->num_cus, sh);
  if (!mmf::march_shape_supported(ctx->g, sh))
    return fail(ctx, MMF_ERR_INVALID, "marching shape: 3D tiles 64 x 4, 64 x 8, 32 x 8 or 32 x 16; 2D 64 x 1");
  ctx->lam_valid = false;
  ctx->tiles_valid = false;
  ctx->shape = sh;
  return MMF_OK;
}

int mmf_set_march_blocks(mmf_ctx *ctx, int32_t blocks) {
  if (!ctx || blocks < 0) return MMF_ERR_INVALID;
  mmf::MarchShape sh = ctx->shape;
  if (blocks == 0) {
    sh = mmf::default_march_shape(ctx->g, ctx->num_cus);
    sh.tx = ctx->shape.tx;
    sh.ty = ctx->shape.ty;
    mmf::balance_march_shape(ctx->g, ctx->num_cus, sh);
  } else {
    if (ctx->g.dim != 3 || ctx->g.order != 1 || sh.tx != 64 || sh.ty != 8)
      return fail(ctx, MMF_ERR_INVALID, "march blocks: the 3D order-1 marching kernel, 64 x 8 tiles");
    const int64_t ntiles = ((ctx->g.n[0] + sh.tx - 1) / sh.tx) * ((ctx->g.n[1] + sh.ty - 1) / sh.ty);
    if (blocks < ntiles || blocks > ntiles * ctx->g.n[2])
      return fail(ctx, MMF_ERR_INVALID, "march blocks: between the tile count and the tile planes");
    sh.seg_blocks = blocks;
  }
  ctx->lam_valid = false;
  ctx->tiles_valid = false;
  ctx->shape = sh;
  return MMF_OK;
}

int mmf_get_kernel(const mmf_ctx *ctx, int32_t *kind, int32_t *tile_rows, int32_t *chunk) {
  if (!ctx) return MMF_ERR_INVALID;
  if (kind) *kind = kernel_used(ctx);
  if (tile_rows) *tile_rows = ctx->shape.ty;
  // 0: the planes of all tiles spread evenly over the blocks (3D order 1)
  if (chunk) *chunk = (ctx->shape.seg_blocks > 0 && ctx->g.dim == 3 && ctx->g.order == 1) ? 0 : ctx->shape.chunk;
  return MMF_OK;
}

int mmf_get_sizes(const mmf_ctx *ctx, int64_t *ncells, int64_t *nifaces) {
  if (!ctx) return MMF_ERR_INVALID;
  const GridDev &g = ctx->g;
  if (ncells) *ncells = g.N;
  if (nifaces) {
    int64_t nf = 0;
    for (int a = 0; a < g.dim; ++a) nf += (g.periodic[a] ? g.n[a] : g.n[a] + 1) * (g.N / g.n[a]);
    *nifaces = nf;
  }
  return MMF_OK;
}

int mmf_set_cell_flags(mmf_ctx *ctx, const uint8_t *fluid, const uint8_t *solved) {
  if (!ctx) return MMF_ERR_INVALID;
  ctx->lam_valid = false;
  if (ctx->learn_on) return fail(ctx, MMF_ERR_STATE, "the learned reconstruction runs on grids without bodies");
  const int64_t N = ctx->g.N;
  std::vector<uint8_t> f((size_t)N);
  bool all = true;
  for (int64_t r = 0; r < N; ++r) {
    uint8_t fl = fluid ? (fluid[r] ? 1 : 0) : 1;
    uint8_t so = solved ? (solved[r] ? 1 : 0) : fl;
    if (so && !fl) return fail(ctx, MMF_ERR_INVALID, "a solved cell must be fluid (main.cpp:230-236)");
    uint8_t v = (uint8_t)((fl ? MMF_CELL_FLUID : 0) | (so ? MMF_CELL_SOLVED : 0));
    int64_t c = ctx->raw.empty() ? r : ctx->raw[(size_t)r];
    f[(size_t)c] = v;
    all = all && (v == (MMF_CELL_FLUID | MMF_CELL_SOLVED));
  }
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  ctx->tiles_valid = false;
  if (all) {
    if (ctx->dflags) (void)hipFree(ctx->dflags);
    ctx->dflags = nullptr;
    ctx->g.flags = nullptr;
    ctx->g.ghost_free = 1;
    ctx->hflags.clear();
    return MMF_OK;
  }
  if (!ctx->dflags) HIPCHK(ctx, hipMalloc(&ctx->dflags, (size_t)N));
  HIPCHK(ctx, hipMemcpy(ctx->dflags, f.data(), (size_t)N, hipMemcpyHostToDevice));
  ctx->g.flags = ctx->dflags;
  ctx->g.ghost_free = ghost_free_of(f);
  ctx->hflags.swap(f);
  return MMF_OK;
}

int mmf_border_face_count(const mmf_grid_desc *desc, int64_t *count) {
  GridDev g;
  const int rc = mmf::grid_from_desc(desc, g);
  if (rc != MMF_OK) return rc;
  if (!count) return MMF_ERR_INVALID;
  int64_t c = 0;
  for (int s = 0; s < 2 * g.dim; ++s) c += mmf::border_faces_of_side(g.n, s);
  *count = c;
  return MMF_OK;
}

// main.cpp:251-260 (interfaceBCs per border interface).  Each distinct
// (code, Dirichlet data) becomes a slot; a side whose faces share one slot
// keeps the per-side form, which every stage kernel reads.
int mmf_set_border_bcs(mmf_ctx *ctx, const int32_t *bc, const double *info) {
  if (!ctx || !bc) return MMF_ERR_INVALID;
  if (ctx->learn_on) return fail(ctx, MMF_ERR_STATE, "border conditions: the learned reconstruction is on");
  GridDev &g = ctx->g;
  int64_t off[6] = {0, 0, 0, 0, 0, 0}, total = 0;
  for (int s = 0; s < 2 * g.dim; ++s) {
    off[s] = total;
    total += mmf::border_faces_of_side(g.n, s);
  }
  std::map<std::vector<uint64_t>, int> slot_of;
  std::vector<int32_t> sbc;
  std::vector<double> sinfo;
  std::vector<uint8_t> slot((size_t)total, 0);
  bool uniform[6] = {true, true, true, true, true, true};
  for (int s = 0; s < 2 * g.dim; ++s) {
    if (g.periodic[s / 2]) continue;  // interior faces: entries ignored
    const int64_t nfs = mmf::border_faces_of_side(g.n, s);
    for (int64_t e = off[s]; e < off[s] + nfs; ++e) {
      const int32_t b = bc[e];
      if (b < MMF_BC_FREE_FLOW || b > MMF_BC_DIRICHLET) return fail(ctx, MMF_ERR_INVALID, "border conditions: code");
      if (b == MMF_BC_DIRICHLET && !info) return fail(ctx, MMF_ERR_INVALID, "border conditions: Dirichlet data");
      std::vector<uint64_t> key(1 + 5, 0);
      key[0] = (uint64_t)b;
      if (b == MMF_BC_DIRICHLET) std::memcpy(&key[1], info + 5 * e, 5 * sizeof(double));
      auto it = slot_of.find(key);
      int k;
      if (it == slot_of.end()) {
        k = (int)sbc.size();
        if (k > 255) return fail(ctx, MMF_ERR_INVALID, "border conditions: more than 256 distinct conditions");
        slot_of.emplace(key, k);
        sbc.push_back(b);
        for (int q = 0; q < 5; ++q) sinfo.push_back(b == MMF_BC_DIRICHLET ? info[5 * e + q] : 0.0);
      } else {
        k = it->second;
      }
      slot[(size_t)e] = (uint8_t)k;
      if (slot[(size_t)e] != slot[(size_t)off[s]]) uniform[s] = false;
    }
  }
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));  // no launch in flight reads the old table
  ctx->lam_valid = false;
  ctx->tiles_valid = false;
  bool all_uniform = true;
  for (int s = 0; s < 2 * g.dim; ++s) {
    if (g.periodic[s / 2]) continue;
    const int k = slot[(size_t)off[s]];
    g.bc[s] = sbc[(size_t)k];  // a non-uniform side: its first face's (not read)
    for (int q = 0; q < 5; ++q) g.info[s][q] = sinfo[(size_t)(5 * k + q)];
    if (uniform[s]) {
      ctx->desc.side_bc[s] = g.bc[s];
      for (int q = 0; q < 5; ++q) ctx->desc.side_info[s][q] = g.info[s][q];
    }
    all_uniform = all_uniform && uniform[s];
  }
  if (ctx->dbslot) (void)hipFree(ctx->dbslot);
  if (ctx->dbslot_bc) (void)hipFree(ctx->dbslot_bc);
  if (ctx->dbslot_info) (void)hipFree(ctx->dbslot_info);
  ctx->dbslot = nullptr;
  ctx->dbslot_bc = nullptr;
  ctx->dbslot_info = nullptr;
  g.bslot = nullptr;
  g.bslot_bc = nullptr;
  g.bslot_info = nullptr;
  for (int s = 0; s < 6; ++s) g.border_off[s] = 0;
  if (all_uniform) return MMF_OK;
  hipError_t e = hipMalloc(&ctx->dbslot, (size_t)total);
  if (e == hipSuccess) e = hipMalloc(&ctx->dbslot_bc, sizeof(int32_t) * sbc.size());
  if (e == hipSuccess) e = hipMalloc(&ctx->dbslot_info, sizeof(double) * sinfo.size());
  if (e == hipSuccess) e = hipMemcpy(ctx->dbslot, slot.data(), (size_t)total, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(ctx->dbslot_bc, sbc.data(), sizeof(int32_t) * sbc.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(ctx->dbslot_info, sinfo.data(), sizeof(double) * sinfo.size(), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    if (ctx->dbslot) (void)hipFree(ctx->dbslot);
    if (ctx->dbslot_bc) (void)hipFree(ctx->dbslot_bc);
    if (ctx->dbslot_info) (void)hipFree(ctx->dbslot_info);
    ctx->dbslot = nullptr;
    ctx->dbslot_bc = nullptr;
    ctx->dbslot_info = nullptr;
    return hip_fail(ctx, e, "border conditions");
  }
  g.bslot = ctx->dbslot;
  g.bslot_bc = ctx->dbslot_bc;
  g.bslot_info = ctx->dbslot_info;
  for (int s = 0; s < 6; ++s) g.border_off[s] = off[s];
  return MMF_OK;
}

int mmf_set_bodies(mmf_ctx *ctx, const double *boxes, int32_t nb, const double *spheres, int32_t ns) {
  if (!ctx || nb < 0 || ns < 0 || (nb && !boxes) || (ns && !spheres)) return fail(ctx, MMF_ERR_INVALID, "bodies");
  ctx->lam_valid = false;
  if (ctx->learn_on) return fail(ctx, MMF_ERR_STATE, "the learned reconstruction runs on grids without bodies");
  const int64_t N = ctx->g.N;
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  uint8_t *dfl = ctx->dflags;
  double *dbox = nullptr, *dsph = nullptr;
  // every temporary is released on every path below
  hipError_t e = hipSuccess;
  if (!dfl) e = hipMalloc(&dfl, (size_t)N);
  if (e == hipSuccess && nb) e = hipMalloc(&dbox, sizeof(double) * 6 * nb);
  if (e == hipSuccess && nb) e = hipMemcpy(dbox, boxes, sizeof(double) * 6 * nb, hipMemcpyHostToDevice);
  if (e == hipSuccess && ns) e = hipMalloc(&dsph, sizeof(double) * 4 * ns);
  if (e == hipSuccess && ns) e = hipMemcpy(dsph, spheres, sizeof(double) * 4 * ns, hipMemcpyHostToDevice);
  GridDev g = ctx->g;
  if (e == hipSuccess) e = mmf::launch_mark_bodies(g, ctx->desc.origin, dbox, nb, dsph, ns, dfl, ctx->stream);
  std::vector<uint8_t> h((size_t)N);
  if (e == hipSuccess) e = hipMemcpyAsync(h.data(), dfl, (size_t)N, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (dbox) (void)hipFree(dbox);
  if (dsph) (void)hipFree(dsph);
  if (e != hipSuccess) {
    if (dfl && dfl != ctx->dflags) (void)hipFree(dfl);
    return hip_fail(ctx, e, "body marking");
  }
  bool all = std::all_of(h.begin(), h.end(), [](uint8_t v) { return v == (MMF_CELL_FLUID | MMF_CELL_SOLVED); });
  ctx->tiles_valid = false;
  if (all) {
    (void)hipFree(dfl);
    ctx->dflags = nullptr;
    ctx->g.flags = nullptr;
    ctx->g.ghost_free = 1;
    ctx->hflags.clear();
  } else {
    ctx->dflags = dfl;
    ctx->g.flags = dfl;
    ctx->g.ghost_free = ghost_free_of(h);
    ctx->hflags.swap(h);
  }
  return MMF_OK;
}

int mmf_get_cell_flags(mmf_ctx *ctx, uint8_t *flags) {
  if (!ctx || !flags) return MMF_ERR_INVALID;
  const int64_t N = ctx->g.N;
  std::vector<uint8_t> h((size_t)N, (uint8_t)(MMF_CELL_FLUID | MMF_CELL_SOLVED));
  if (ctx->dflags) {
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    HIPCHK(ctx, hipMemcpy(h.data(), ctx->dflags, (size_t)N, hipMemcpyDeviceToHost));
  }
  for (int64_t r = 0; r < N; ++r) flags[r] = h[(size_t)(ctx->raw.empty() ? r : ctx->raw[(size_t)r])];
  return MMF_OK;
}

int mmf_get_centroids(const mmf_ctx *ctx, double *xyz) {
  if (!ctx || !xyz) return MMF_ERR_INVALID;
  const GridDev &g = ctx->g;
  const double *o = ctx->desc.origin;
  for (int64_t r = 0; r < g.N; ++r) {
    int64_t c = ctx->raw.empty() ? r : ctx->raw[(size_t)r];
    int64_t q[3] = {c % g.n[0], (c / g.n[0]) % g.n[1], c / (g.n[0] * g.n[1])};
    for (int d = 0; d < 3; ++d) xyz[3 * r + d] = (d < g.dim) ? o[d] + (q[d] + 0.5) * g.h : o[d];
  }
  return MMF_OK;
}

int mmf_set_raw_order(mmf_ctx *ctx, const int64_t *cell_of_raw, int64_t nraw) {
  if (!ctx) return MMF_ERR_INVALID;
  if (!cell_of_raw) {
    ctx->raw.clear();
    if (ctx->draw) {
      HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
      HIPCHK(ctx, hipFree(ctx->draw));
      ctx->draw = nullptr;
    }
    return MMF_OK;
  }
  if (nraw != ctx->g.N) return fail(ctx, MMF_ERR_INVALID, "raw order size must equal the cell count");
  std::vector<uint8_t> seen((size_t)nraw, 0);
  for (int64_t r = 0; r < nraw; ++r) {
    int64_t c = cell_of_raw[r];
    if (c < 0 || c >= nraw || seen[(size_t)c]) return fail(ctx, MMF_ERR_INVALID, "raw order is not a permutation");
    seen[(size_t)c] = 1;
  }
  ctx->raw.assign(cell_of_raw, cell_of_raw + nraw);
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  if (!ctx->draw) HIPCHK(ctx, hipMalloc(&ctx->draw, sizeof(int64_t) * (size_t)nraw));
  HIPCHK(ctx, hipMemcpy(ctx->draw, cell_of_raw, sizeof(int64_t) * (size_t)nraw, hipMemcpyHostToDevice));
  return MMF_OK;
}

// The host AoS buffer crosses PCIe as it is; the transpose to or from the SoA
// planes runs on the GPU, through the residual buffer R (dead between steps;
// it holds the 5N doubles of an AoS copy since fs >= N).
int mmf_set_state(mmf_ctx *ctx, const double *U) {
  if (!ctx || !U) return MMF_ERR_INVALID;
  ctx->lam_valid = false;
  const int64_t N = ctx->g.N, fs = ctx->g.fs;
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  HIPCHK(ctx, hipMemcpy(ctx->dR, U, sizeof(double) * 5 * (size_t)N, hipMemcpyHostToDevice));
  HIPCHK(ctx, mmf::launch_aos_to_soa(ctx->dR, ctx->draw, N, fs, ctx->dU, ctx->stream));
  // the stage buffer starts as a copy so cells never solved hold finite values
  HIPCHK(ctx, hipMemcpyAsync(ctx->dW, ctx->dU, sizeof(double) * 5 * (size_t)fs, hipMemcpyDeviceToDevice, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return MMF_OK;
}

int mmf_get_state(mmf_ctx *ctx, double *U) {
  if (!ctx || !U) return MMF_ERR_INVALID;
  const int64_t N = ctx->g.N, fs = ctx->g.fs;
  HIPCHK(ctx, mmf::launch_soa_to_aos(ctx->dU, ctx->draw, N, fs, ctx->dR, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  HIPCHK(ctx, hipMemcpy(U, ctx->dR, sizeof(double) * 5 * (size_t)N, hipMemcpyDeviceToHost));
  return MMF_OK;
}

int mmf_compute_rhs(mmf_ctx *ctx, double *rhs, double *max_eig) {
  if (!ctx) return MMF_ERR_INVALID;
  ctx->lam_valid = false;
  int rc = mmf_ctx_enqueue_rhs(ctx);
  if (rc != MMF_OK) return rc;
  if (rhs) {
    // AoS residual through W (dead between steps), which is then restored to
    // the copy of U it holds after mmf_set_state
    const int64_t N = ctx->g.N, fs = ctx->g.fs;
    HIPCHK(ctx, mmf::launch_soa_to_aos(ctx->dR, ctx->draw, N, fs, ctx->dW, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    HIPCHK(ctx, hipMemcpy(rhs, ctx->dW, sizeof(double) * 5 * (size_t)N, hipMemcpyDeviceToHost));
    HIPCHK(ctx, hipMemcpyAsync(ctx->dW, ctx->dU, sizeof(double) * 5 * (size_t)fs, hipMemcpyDeviceToDevice, ctx->stream));
  }
  int r2 = read_ctrl(ctx);
  if (r2 != MMF_OK) return r2;
  if (max_eig) {
    *max_eig = mmf::ctrl_stage_max(*ctx->hctrl, 3);
  }
  return MMF_OK;
}

// euler::computeRHS with the caller's solved flags and order (euler.hpp:41-43
// takes both on every call): flags and order change only where they differ
// from the context's.
int mmf_compute_rhs_ex(mmf_ctx *ctx, const uint8_t *solved, int32_t order, const double *U, double *rhs,
                       double *max_eig) {
  if (!ctx) return MMF_ERR_INVALID;
  if (order != 1 && order != 2) return fail(ctx, MMF_ERR_UNSUPPORTED_ORDER, "reconstruction order (reconstruction.cpp:75-76)");
  if (order != ctx->g.order) {
    if (ctx->learn_on) return fail(ctx, MMF_ERR_STATE, "order: the learned reconstruction is on");
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    ctx->g.order = order;
    ctx->desc.order = order;
    ctx->lam_valid = false;
    ctx->tiles_valid = false;
  }
  if (solved) {
    const int64_t N = ctx->g.N;
    std::vector<uint8_t> fluid((size_t)N), want((size_t)N);
    bool same = true;
    for (int64_t r = 0; r < N; ++r) {
      const int64_t c = ctx->raw.empty() ? r : ctx->raw[(size_t)r];
      const uint8_t cur = ctx->hflags.empty() ? (uint8_t)(MMF_CELL_FLUID | MMF_CELL_SOLVED) : ctx->hflags[(size_t)c];
      fluid[(size_t)r] = (cur & MMF_CELL_FLUID) ? 1 : 0;
      want[(size_t)r] = solved[r] ? 1 : 0;
      same = same && (((cur & MMF_CELL_SOLVED) ? 1 : 0) == want[(size_t)r]);
    }
    if (!same) {
      const int rc = mmf_set_cell_flags(ctx, fluid.data(), want.data());
      if (rc != MMF_OK) return rc;
    }
  }
  if (U) {
    const int rc = mmf_set_state(ctx, U);
    if (rc != MMF_OK) return rc;
  }
  return mmf_compute_rhs(ctx, rhs, max_eig);
}

// ---- one RK3 stage at a time (main.cpp:377-506) ----------------------------

namespace {
double *state_ptr(mmf_ctx *ctx, int32_t state) {
  return state == MMF_STATE_U ? ctx->dU : state == MMF_STATE_W ? ctx->dW : state == MMF_STATE_RHS ? ctx->dR : nullptr;
}
}  // namespace

// reconstruction::computePolynomials + euler::computeRHS of U or W into R
// (main.cpp:388-389, 432-433, 468-469); the stage maximum of the faces.
int mmf_rk_rhs(mmf_ctx *ctx, int32_t state, double *max_eig) {
  if (!ctx || (state != MMF_STATE_U && state != MMF_STATE_W)) return MMF_ERR_INVALID;
  ctx->lam_valid = false;
  int rc = ensure_tiles(ctx);
  if (rc != MMF_OK) return rc;
  hipError_t e = mmf::launch_zero_slot(ctx->dctrl, 3, ctx->stream);
  if (e != hipSuccess) return hip_fail(ctx, e, "zero slot");
  rc = launch_stage_on(ctx, 1, mmf::MODE_RHS_ONLY, state_ptr(ctx, state), ctx->dR, ctx->dW);
  if (rc != MMF_OK) return rc;
  rc = read_ctrl(ctx);
  if (rc != MMF_OK) return rc;
  const double m = mmf::ctrl_stage_max(*ctx->hctrl, 3);
  ctx->rk_eig[ctx->rk_done % 3] = m;
  if (max_eig) *max_eig = m;
  return MMF_OK;
}

// main.cpp:398-402 for the step's (possibly reduced over ranks) maximum.
int mmf_rk_dt(mmf_ctx *ctx, double max_eig, double *dt) {
  if (!ctx || !dt) return MMF_ERR_INVALID;
  if (!(max_eig > 0.0) || !std::isfinite(max_eig))
    return fail(ctx, MMF_ERR_STATE, "the step's maximum eigenvalue is not finite and positive");
  int rc = read_ctrl(ctx);
  if (rc != MMF_OK) return rc;
  const Ctrl &c = *ctx->hctrl;
  double d = 0.9 * c.cfl * c.hmin / max_eig;
  if (c.t + d > c.tmax) d = c.tmax - c.t;
  *dt = d;
  return MMF_OK;
}

// The update loop of stage 1 (main.cpp:409-422), 2 (445-458) or 3 (481-494);
// after stage 3, t += dt and the step is counted (main.cpp:505-506).
int mmf_rk_update(mmf_ctx *ctx, int32_t stage, double dt) {
  if (!ctx || stage < 1 || stage > 3 || !std::isfinite(dt)) return MMF_ERR_INVALID;
  ctx->lam_valid = false;
  hipError_t e = mmf::launch_rk_update(ctx->g, stage, ctx->dU, ctx->dR, ctx->dW, dt, ctx->stream);
  if (e == hipSuccess && stage == 3) e = mmf::launch_rk_close(ctx->dctrl, dt, ctx->rk_eig, ctx->stream);
  if (e != hipSuccess) return hip_fail(ctx, e, "stage update");
  ctx->rk_done = stage % 3;
  return MMF_OK;
}

// Ghost lists: rows (host order, as mmf_set_state) of a state to or from a
// host AoS buffer [n][5] (communications.tpp:164-199).
static int cells_transfer(mmf_ctx *ctx, int32_t state, const int64_t *cells, int64_t n, double *buf, bool pack) {
  double *X = ctx ? state_ptr(ctx, state) : nullptr;
  if (!X || n < 0 || (n && (!cells || !buf))) return ctx ? fail(ctx, MMF_ERR_INVALID, "cell list") : MMF_ERR_INVALID;
  if (n == 0) return MMF_OK;
  std::vector<int64_t> c((size_t)n);
  for (int64_t j = 0; j < n; ++j) {
    if (cells[j] < 0 || cells[j] >= ctx->g.N) return fail(ctx, MMF_ERR_INVALID, "cell list: index out of range");
    c[(size_t)j] = ctx->raw.empty() ? cells[j] : ctx->raw[(size_t)cells[j]];
  }
  const size_t need = (size_t)n * (5 * sizeof(double) + sizeof(int64_t));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  if (ctx->dpack_bytes < need) {
    if (ctx->dpack) (void)hipFree(ctx->dpack);
    ctx->dpack = nullptr;
    ctx->dpack_bytes = 0;
    HIPCHK(ctx, hipMalloc(&ctx->dpack, need));
    ctx->dpack_bytes = need;
  }
  double *dbuf = (double *)ctx->dpack;
  int64_t *dcells = (int64_t *)(dbuf + 5 * n);
  HIPCHK(ctx, hipMemcpyAsync(dcells, c.data(), sizeof(int64_t) * (size_t)n, hipMemcpyHostToDevice, ctx->stream));
  if (pack) {
    HIPCHK(ctx, mmf::launch_pack_cells(X, ctx->g.fs, dcells, n, dbuf, ctx->stream));
    HIPCHK(ctx, hipMemcpyAsync(buf, dbuf, sizeof(double) * 5 * (size_t)n, hipMemcpyDeviceToHost, ctx->stream));
  } else {
    HIPCHK(ctx, hipMemcpyAsync(dbuf, buf, sizeof(double) * 5 * (size_t)n, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(ctx, mmf::launch_unpack_cells(X, ctx->g.fs, dcells, n, dbuf, ctx->stream));
    ctx->lam_valid = false;
  }
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return MMF_OK;
}

int mmf_pack_cells(mmf_ctx *ctx, int32_t state, const int64_t *cells, int64_t n, double *buf) {
  return cells_transfer(ctx, state, cells, n, buf, true);
}

int mmf_unpack_cells(mmf_ctx *ctx, int32_t state, const int64_t *cells, int64_t n, const double *buf) {
  return cells_transfer(ctx, state, cells, n, const_cast<double *>(buf), false);
}

int mmf_get_device_state(mmf_ctx *ctx, int32_t state, double **ptr, int64_t *field_stride) {
  if (!ctx || !ptr || !state_ptr(ctx, state)) return MMF_ERR_INVALID;
  *ptr = state_ptr(ctx, state);
  if (field_stride) *field_stride = ctx->g.fs;
  return MMF_OK;
}

int mmf_set_time(mmf_ctx *ctx, double t, double tmax, double cfl) {
  if (!ctx || !(cfl > 0.0)) return fail(ctx, MMF_ERR_INVALID, "cfl must be positive");
  ctx->lam_valid = false;
  int rc = read_ctrl(ctx);
  if (rc != MMF_OK) return rc;
  Ctrl c;
  std::memset(&c, 0, sizeof(c));
  c.t = t;
  c.tmax = tmax;
  c.cfl = cfl;
  c.hmin = ctx->g.h;
  c.neg_p = ctx->hctrl->neg_p;
  c.neg_rho = ctx->hctrl->neg_rho;
  *ctx->hctrl = c;
  HIPCHK(ctx, hipMemcpyAsync(ctx->dctrl, ctx->hctrl, sizeof(Ctrl), hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return MMF_OK;
}

}  // extern "C"

namespace {

constexpr int GRAPH_STEPS = 16;

// Before steps of the 3D marching kernel with early dt: the maximum of the
// current state, unless the previous steps left it (outside any capture).
// The 3D marching kernel's tile table on grids with flags (the masked tiles
// first; the others run the unmasked code).
int ensure_tiles(mmf_ctx *ctx) {
  const bool march3 = ctx->g.dim == 3 && ctx->g.order == 1 && kernel_used(ctx) == mmf::KERNEL_MARCH;
  const bool segs = march3 && ctx->shape.seg_blocks > 0;
  const bool want = segs || (march3 && ctx->g.flags != nullptr && !ctx->hflags.empty());
  if (!want) {
    ctx->g.march_tiles = nullptr;
    return MMF_OK;
  }
  if (ctx->tiles_valid && ctx->g.march_tiles) return MMF_OK;
  // the segment table (3 * MARCH_SEGS entries per block) or the tile table
  // (one); a share that needs more segments than a block has: the chunks
  const uint8_t *fl = ctx->hflags.empty() ? nullptr : ctx->hflags.data();
  std::vector<int32_t> t;
  if (segs) {
    t.resize((size_t)(ctx->shape.seg_blocks * 3 * mmf::MARCH_SEGS));
    if (!mmf::build_march_segments(ctx->g, ctx->shape, fl, t.data())) {
      ctx->shape.seg_blocks = 0;
      if (ctx->g.flags == nullptr || ctx->hflags.empty()) {
        ctx->g.march_tiles = nullptr;
        return MMF_OK;
      }
      t.clear();
    }
  }
  if (t.empty()) {
    t.resize((size_t)mmf::march_blocks(ctx->g, ctx->shape));
    mmf::build_march_tiles(ctx->g, ctx->shape, fl, t.data());
  }
  const int64_t nb = (int64_t)t.size();
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  if (ctx->dtiles_n != nb) {
    if (ctx->dtiles) (void)hipFree(ctx->dtiles);
    ctx->dtiles = nullptr;
    ctx->dtiles_n = 0;
    ctx->g.march_tiles = nullptr;
    HIPCHK(ctx, hipMalloc(&ctx->dtiles, sizeof(int32_t) * (size_t)nb));
    ctx->dtiles_n = nb;
  }
  HIPCHK(ctx, hipMemcpy(ctx->dtiles, t.data(), sizeof(int32_t) * (size_t)nb, hipMemcpyHostToDevice));
  ctx->g.march_tiles = ctx->dtiles;
  ctx->tiles_valid = true;
  return MMF_OK;
}

int ensure_lam(mmf_ctx *ctx) {
  const bool early = kernel_used(ctx) == mmf::KERNEL_MARCH && mmf::march_early(ctx->g);
  if (!early) {
    ctx->lam_valid = false;
    return MMF_OK;
  }
  if (ctx->lam_valid) return MMF_OK;
  HIPCHK(ctx, hipMemsetAsync(&ctx->dctrl->lam_now, 0, sizeof(unsigned long long), ctx->stream));
  hipError_t e = mmf::launch_lam_init(ctx->g, ctx->dU, ctx->dctrl, ctx->stream);
  if (e != hipSuccess) return hip_fail(ctx, e, "early dt");
  ctx->lam_valid = true;
  return MMF_OK;
}

int enqueue_steps(mmf_ctx *ctx, int64_t nsteps) {
  for (int64_t s = 0; s < nsteps; ++s) {
    for (int stage = 1; stage <= 3; ++stage) {
      int rc = launch_stage_timed(ctx, stage, mmf::MODE_STEP);
      if (rc != MMF_OK) return rc;
    }
    // the tile kernel's and the 3D marching kernel's stage-3 launches advance the time themselves
    const int used = kernel_used(ctx);
    if (used == mmf::KERNEL_TILE || (used == mmf::KERNEL_MARCH && ctx->g.dim == 3)) continue;
    hipError_t e = mmf::launch_advance(ctx->dctrl, ctx->stream);
    if (e != hipSuccess) return hip_fail(ctx, e, "advance kernel");
  }
  return MMF_OK;
}

// everything the captured launches depend on (kernel arguments by value)
std::vector<char> launch_key(const mmf_ctx *ctx) {
  std::vector<char> k(sizeof(GridDev) + sizeof(int) + sizeof(mmf::MarchShape) + sizeof(bool) + sizeof(mmf::LearnRun));
  char *p = k.data();
  std::memcpy(p, &ctx->g, sizeof(GridDev));
  p += sizeof(GridDev);
  std::memcpy(p, &ctx->kernel, sizeof(int));
  p += sizeof(int);
  std::memcpy(p, &ctx->shape, sizeof(mmf::MarchShape));
  p += sizeof(mmf::MarchShape);
  std::memcpy(p, &ctx->learn_on, sizeof(bool));
  p += sizeof(bool);
  std::memcpy(p, &ctx->learn, sizeof(mmf::LearnRun));
  return k;
}

int capture_steps(mmf_ctx *ctx, int slot, int64_t nsteps) {
  HIPCHK(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
  int rc = enqueue_steps(ctx, nsteps);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(ctx->stream, &g);
  if (rc != MMF_OK) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return hip_fail(ctx, e, "end of step capture");
  ctx->step_graph[slot] = g;
  HIPCHK(ctx, hipGraphInstantiate(&ctx->step_exec[slot], g, nullptr, nullptr, 0));
  return MMF_OK;
}

}  // namespace

extern "C" {

void mmf_drop_step_graphs(mmf_ctx *ctx) {
  for (int i = 0; i < 2; ++i) {
    if (ctx->step_exec[i]) (void)hipGraphExecDestroy(ctx->step_exec[i]);
    if (ctx->step_graph[i]) (void)hipGraphDestroy(ctx->step_graph[i]);
    ctx->step_exec[i] = nullptr;
    ctx->step_graph[i] = nullptr;
  }
  ctx->graph_key.clear();
}

namespace {

// Steps replayed from captured HIP graphs (GRAPH_STEPS steps per replay, then
// single steps), re-captured whenever the launch configuration changed.
int replay_steps(mmf_ctx *ctx, int64_t nsteps) {
  std::vector<char> key = launch_key(ctx);
  if (key != ctx->graph_key) {
    mmf_drop_step_graphs(ctx);
    ctx->graph_key = key;
  }
  while (nsteps > 0) {
    const int slot = (nsteps >= GRAPH_STEPS) ? 1 : 0;
    if (!ctx->step_exec[slot]) {
      int rc = capture_steps(ctx, slot, slot ? GRAPH_STEPS : 1);
      if (rc != MMF_OK) {
        mmf_drop_step_graphs(ctx);
        return rc;
      }
    }
    HIPCHK(ctx, hipGraphLaunch(ctx->step_exec[slot], ctx->stream));
    nsteps -= slot ? GRAPH_STEPS : 1;
  }
  return MMF_OK;
}

}  // namespace

// Steps are replayed from graphs, or launched one by one with per-launch
// kernel timing on; region timing brackets the whole call with one event pair.
int mmf_advance(mmf_ctx *ctx, int64_t nsteps) {
  if (!ctx || nsteps < 0) return MMF_ERR_INVALID;
  int rct = ensure_tiles(ctx);
  if (rct != MMF_OK) return rct;
  int rc0 = ensure_lam(ctx);
  if (rc0 != MMF_OK) return rc0;
  hipEvent_t r0 = nullptr, r1 = nullptr;
  if (ctx->timing == 2 && nsteps > 0) {
    HIPCHK(ctx, hipEventCreate(&r0));
    if (hipEventCreate(&r1) != hipSuccess) {
      (void)hipEventDestroy(r0);
      return fail(ctx, MMF_ERR_HIP, "timing event");
    }
    ctx->events.push_back({r0, r1, 3 * nsteps});
    HIPCHK(ctx, hipEventRecord(r0, ctx->stream));
  }
  const int rc = (ctx->timing == 1 || !ctx->use_graphs || nsteps == 0) ? enqueue_steps(ctx, nsteps)
                                                                        : replay_steps(ctx, nsteps);
  if (r1) HIPCHK(ctx, hipEventRecord(r1, ctx->stream));
  return rc;
}

int mmf_set_step_graphs(mmf_ctx *ctx, int32_t on) {
  if (!ctx) return MMF_ERR_INVALID;
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  ctx->use_graphs = on != 0;
  if (!ctx->use_graphs) mmf_drop_step_graphs(ctx);
  return MMF_OK;
}

int mmf_run(mmf_ctx *ctx, int64_t max_steps, int64_t *steps_done, double *t_final) {
  if (!ctx) return MMF_ERR_INVALID;
  int rc = read_ctrl(ctx);
  if (rc != MMF_OK) return rc;
  const int64_t step0 = ctx->hctrl->step;
  for (;;) {
    const Ctrl &c = *ctx->hctrl;
    int64_t done = c.step - step0;
    if (c.nonfinite) break;
    if (!(c.t < c.tmax)) break;
    if (max_steps >= 0 && done >= max_steps) break;
    int64_t batch = 1;
    if (c.dt > 0.0) {
      double rem = std::ceil((c.tmax - c.t) / c.dt);
      batch = (int64_t)std::min(std::max(rem, 1.0), 256.0);
    }
    if (max_steps >= 0) batch = std::min(batch, max_steps - done);
    rc = mmf_advance(ctx, batch);
    if (rc != MMF_OK) return rc;
    rc = read_ctrl(ctx);
    if (rc != MMF_OK) return rc;
  }
  if (steps_done) *steps_done = ctx->hctrl->step - step0;
  if (t_final) *t_final = ctx->hctrl->t;
  return nonfinite_status(ctx);
}

int mmf_get_time(mmf_ctx *ctx, double *t, double *dt, int64_t *step, double *stage_eig) {
  if (!ctx) return MMF_ERR_INVALID;
  int rc = read_ctrl(ctx);
  if (rc != MMF_OK) return rc;
  const Ctrl &c = *ctx->hctrl;
  if (t) *t = c.t;
  if (dt) *dt = c.dt;
  if (step) *step = c.step;
  if (stage_eig)
    for (int s = 0; s < 3; ++s) stage_eig[s] = c.last_eig[s];
  return nonfinite_status(ctx);
}

int mmf_get_nonfinite_steps(mmf_ctx *ctx, int64_t *steps) {
  if (!ctx || !steps) return MMF_ERR_INVALID;
  int rc = read_ctrl(ctx);
  if (rc != MMF_OK) return rc;
  *steps = (int64_t)ctx->hctrl->nonfinite;
  return MMF_OK;
}

int mmf_synchronize(mmf_ctx *ctx) {
  if (!ctx) return MMF_ERR_INVALID;
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return MMF_OK;
}

int mmf_error_norms(mmf_ctx *ctx, const double *rho_exact, double *l1, double *l2) {
  if (!ctx || !rho_exact) return MMF_ERR_INVALID;
  const int64_t N = ctx->g.N;
  std::vector<double> ex((size_t)N);
  for (int64_t r = 0; r < N; ++r) ex[(size_t)(ctx->raw.empty() ? r : ctx->raw[(size_t)r])] = rho_exact[r];
  double *dex = nullptr;
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  HIPCHK(ctx, hipMalloc(&dex, sizeof(double) * (size_t)N));
  hipError_t e = hipMemcpy(dex, ex.data(), sizeof(double) * (size_t)N, hipMemcpyHostToDevice);
  double out[2] = {0.0, 0.0};
  if (e == hipSuccess)
    e = mmf::launch_error_norms(ctx->g, ctx->dU, dex, ctx->dpartials, ctx->nerr_blocks, ctx->derr, ctx->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(out, ctx->derr, sizeof(out), hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  (void)hipFree(dex);
  if (e != hipSuccess) return hip_fail(ctx, e, "error norms");
  if (l1) *l1 = out[0];
  if (l2) *l2 = std::sqrt(out[1]);
  return MMF_OK;
}

int mmf_get_positivity_counters(mmf_ctx *ctx, int64_t *neg_p, int64_t *neg_rho) {
  if (!ctx) return MMF_ERR_INVALID;
  int rc = read_ctrl(ctx);
  if (rc != MMF_OK) return rc;
  if (neg_p) *neg_p = (int64_t)ctx->hctrl->neg_p;
  if (neg_rho) *neg_rho = (int64_t)ctx->hctrl->neg_rho;
  return MMF_OK;
}

int mmf_reset_positivity_counters(mmf_ctx *ctx) {
  if (!ctx) return MMF_ERR_INVALID;
  ctx->lam_valid = false;
  int rc = read_ctrl(ctx);
  if (rc != MMF_OK) return rc;
  ctx->hctrl->neg_p = 0;
  ctx->hctrl->neg_rho = 0;
  ctx->hctrl->nonfinite = 0;
  HIPCHK(ctx, hipMemcpyAsync(ctx->dctrl, ctx->hctrl, sizeof(Ctrl), hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return MMF_OK;
}

int mmf_enable_timing(mmf_ctx *ctx, int32_t on) {
  if (!ctx || on < 0 || on > 2) return MMF_ERR_INVALID;
  int rc = collect_timing(ctx);
  if (rc != MMF_OK) return rc;
  ctx->timing = on;
  ctx->timed_ms = 0.0;
  ctx->timed_launches = 0;
  return MMF_OK;
}

int mmf_read_timing(mmf_ctx *ctx, double *ms, int64_t *launches) {
  if (!ctx) return MMF_ERR_INVALID;
  int rc = collect_timing(ctx);
  if (rc != MMF_OK) return rc;
  if (ms) *ms = ctx->timed_ms;
  if (launches) *launches = ctx->timed_launches;
  return MMF_OK;
}

// ---- batched point functions ------------------------------------------------

namespace {

struct DevBuf {
  void *p = nullptr;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
};

int batch_run(int64_t count, std::initializer_list<std::pair<const double *, size_t>> ins,
              std::initializer_list<std::pair<double *, size_t>> outs,
              const std::function<hipError_t(double **, double **, int64_t)> &launch) {
  if (count < 0) return MMF_ERR_INVALID;
  if (count == 0) return MMF_OK;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return MMF_ERR_NO_DEVICE;
  std::vector<DevBuf> bi(ins.size()), bo(outs.size());
  std::vector<double *> pi, po;
  size_t i = 0;
  for (auto &in : ins) {
    if (!in.first) {
      pi.push_back(nullptr);
      ++i;
      continue;
    }
    size_t bytes = sizeof(double) * in.second * (size_t)count;
    if (hipMalloc(&bi[i].p, bytes) != hipSuccess) return MMF_ERR_OUT_OF_MEMORY;
    if (hipMemcpy(bi[i].p, in.first, bytes, hipMemcpyHostToDevice) != hipSuccess) return MMF_ERR_HIP;
    pi.push_back((double *)bi[i].p);
    ++i;
  }
  i = 0;
  for (auto &o : outs) {
    size_t bytes = sizeof(double) * o.second * (size_t)count;
    if (hipMalloc(&bo[i].p, bytes) != hipSuccess) return MMF_ERR_OUT_OF_MEMORY;
    po.push_back((double *)bo[i].p);
    ++i;
  }
  if (launch(pi.data(), po.data(), count) != hipSuccess) return MMF_ERR_HIP;
  if (hipDeviceSynchronize() != hipSuccess) return MMF_ERR_HIP;
  i = 0;
  for (auto &o : outs) {
    size_t bytes = sizeof(double) * o.second * (size_t)count;
    if (hipMemcpy(o.first, bo[i].p, bytes, hipMemcpyDeviceToHost) != hipSuccess) return MMF_ERR_HIP;
    ++i;
  }
  return MMF_OK;
}

}  // namespace

int mmf_batch_cons2prim(const double *cons, int64_t n, double *prim) {
  if (!cons || !prim) return MMF_ERR_INVALID;
  return batch_run(n, {{cons, 5}}, {{prim, 5}}, [](double **i, double **o, int64_t c) {
    return mmf::launch_batch_cons2prim(i[0], c, o[0], nullptr);
  });
}

int mmf_batch_prim2cons(const double *prim, int64_t n, double *cons) {
  if (!cons || !prim) return MMF_ERR_INVALID;
  return batch_run(n, {{prim, 5}}, {{cons, 5}}, [](double **i, double **o, int64_t c) {
    return mmf::launch_batch_prim2cons(i[0], c, o[0], nullptr);
  });
}

int mmf_batch_eval_fluxes(const double *cons, const double *prim, const double *nrm, int64_t n, double *flux) {
  if (!cons || !prim || !nrm || !flux) return MMF_ERR_INVALID;
  return batch_run(n, {{cons, 5}, {prim, 5}, {nrm, 3}}, {{flux, 5}}, [](double **i, double **o, int64_t c) {
    return mmf::launch_batch_eval_fluxes(i[0], i[1], i[2], c, o[0], nullptr);
  });
}

int mmf_batch_eval_splitting(const double *L, const double *R, const double *nrm, int64_t n, double *F,
                             double *lam) {
  if (!L || !R || !nrm || !F || !lam) return MMF_ERR_INVALID;
  return batch_run(n, {{L, 5}, {R, 5}, {nrm, 3}}, {{F, 5}, {lam, 1}}, [](double **i, double **o, int64_t c) {
    return mmf::launch_batch_eval_splitting(i[0], i[1], i[2], c, o[0], o[1], nullptr);
  });
}

int mmf_batch_eval_bc_values(int32_t bc, const double *nrm, const double *info, const double *cons, int64_t n,
                             double *out) {
  if (!nrm || !cons || !out || !valid_bc(bc) || bc == MMF_BC_PERIODIC) return MMF_ERR_INVALID;
  if (bc == MMF_BC_DIRICHLET && !info) return MMF_ERR_INVALID;
  return batch_run(n, {{nrm, 3}, {info, 5}, {cons, 5}}, {{out, 5}}, [bc](double **i, double **o, int64_t c) {
    return mmf::launch_batch_bc_values(bc, i[0], i[1], i[2], c, o[0], nullptr);
  });
}

int mmf_calibration_copy(int64_t nbytes, int32_t reps, double *ms) {
  if (nbytes < 32 || reps < 1 || !ms) return MMF_ERR_INVALID;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return MMF_ERR_NO_DEVICE;
  int64_t n4 = nbytes / 32;
  DevBuf a, b;
  if (hipMalloc(&a.p, (size_t)n4 * 32) != hipSuccess) return MMF_ERR_OUT_OF_MEMORY;
  if (hipMalloc(&b.p, (size_t)n4 * 32) != hipSuccess) return MMF_ERR_OUT_OF_MEMORY;
  if (hipMemset(a.p, 0, (size_t)n4 * 32) != hipSuccess) return MMF_ERR_HIP;
  struct Events {  // destroyed on every return path
    hipEvent_t e[2] = {nullptr, nullptr};
    ~Events() {
      for (hipEvent_t x : e)
        if (x) (void)hipEventDestroy(x);
    }
  } ev;
  if (hipEventCreate(&ev.e[0]) != hipSuccess || hipEventCreate(&ev.e[1]) != hipSuccess) return MMF_ERR_HIP;
  // warm-up
  if (mmf::launch_copy((const double4 *)a.p, (double4 *)b.p, n4, nullptr) != hipSuccess) return MMF_ERR_HIP;
  if (hipEventRecord(ev.e[0], nullptr) != hipSuccess) return MMF_ERR_HIP;
  for (int r = 0; r < reps; ++r)
    if (mmf::launch_copy((const double4 *)a.p, (double4 *)b.p, n4, nullptr) != hipSuccess) return MMF_ERR_HIP;
  if (hipEventRecord(ev.e[1], nullptr) != hipSuccess || hipEventSynchronize(ev.e[1]) != hipSuccess) return MMF_ERR_HIP;
  float t = 0.f;
  if (hipEventElapsedTime(&t, ev.e[0], ev.e[1]) != hipSuccess) return MMF_ERR_HIP;
  *ms = (double)t / reps;
  return MMF_OK;
}

}  // extern "C"
