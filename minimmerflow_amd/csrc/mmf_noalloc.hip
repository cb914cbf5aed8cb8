// mmf_noalloc.hip -- euler::computeRHS on caller-owned device buffers,
// allocating nothing (mmf_compute_rhs_noalloc, include/mmf.h).
//
// The context path (mmf_create + mmf_compute_rhs) keeps its stage maxima and
// positivity counters in a control block the library allocates.  This entry
// point is for callers that cannot allocate (code under HIP graph capture,
// serving loops, harnesses that forbid hipMalloc): its control block is the
// file-scope device array 'scratch' below, one copy per GPU, held until the
// process ends.  The caller passes the state, the residual and the outputs;
// the grid comes by value.  Every call on one GPU shares that GPU's scratch,
// which keeps the last call's contents: each call clears what it reads, and
// calls that overlap on one GPU (different streams, no ordering) race.
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <mutex>

#include "mmf.h"
#include "mmf_internal.hpp"
#include "mmf_stage.hpp"

namespace {

using mmf::Ctrl;
using mmf::GridDev;

// Maximum size of the scratch: one control block (stage-maximum lines,
// positivity counters), 12,672 bytes at most.
constexpr size_t SCRATCH_BYTES = 12672;
static_assert(sizeof(Ctrl) <= SCRATCH_BYTES, "the control block must fit the scratch array");

static __device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

// Per-device address of 'scratch' (the array's name in host code is only a
// stand-in) and CU count, set once per device on its first call.
constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
Ctrl *g_scratch[MAX_DEVICES] = {};
int g_cus[MAX_DEVICES] = {};

thread_local char g_err[256] = "";

int nfail(int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

int device_scratch(int dev, Ctrl **ptr, int *cus) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_scratch[dev]) {
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return nfail(MMF_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
    int n = 0;
    if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) n = 0;
    g_cus[dev] = n;
    g_scratch[dev] = reinterpret_cast<Ctrl *>(p);
  }
  *ptr = g_scratch[dev];
  *cus = g_cus[dev];
  return MMF_OK;
}

// Clears what the RHS launch accumulates into: the compute_rhs stage-maximum
// lines and the positivity counters (a kernel, so the call can be captured).
__global__ void k_noalloc_begin(Ctrl *c) {
  if (threadIdx.x < mmf::EIG_SPLIT) c->eig[3][threadIdx.x][0] = 0ull;
  if (threadIdx.x == 0) {
    c->neg_p = 0ull;
    c->neg_rho = 0ull;
  }
}

// The stage maximum (largest over the lines) and the counters to the
// caller's device outputs.
__global__ void k_noalloc_end(const Ctrl *c, double *max_eig, int64_t *counts) {
  unsigned long long v = (threadIdx.x < mmf::EIG_SPLIT) ? c->eig[3][threadIdx.x][0] : 0ull;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned long long o = __shfl_xor(v, off, 64);
    v = (o > v) ? o : v;
  }
  if (threadIdx.x == 0) {
    if (max_eig) *max_eig = __longlong_as_double((long long)v);
    if (counts) {
      counts[0] = (int64_t)c->neg_p;
      counts[1] = (int64_t)c->neg_rho;
    }
  }
}

}  // namespace

extern "C" {

const char *mmf_noalloc_last_error(void) { return g_err; }

int mmf_compute_rhs_noalloc(const mmf_grid_desc *desc, const uint8_t *flags, const double *U, int64_t field_stride,
                            double *rhs, double *max_eig, int64_t *counts, void *stream) {
  g_err[0] = '\0';
  if (!desc || !U || !rhs) return nfail(MMF_ERR_INVALID, "desc, U and rhs are required");
  GridDev g;
  const int rc = mmf::grid_from_desc(desc, g);
  if (rc != MMF_OK)
    return nfail(rc, "grid descriptor rejected: dim 2 or 3, order 1 or 2, h > 0, 1..2^30 cells per axis, at most "
                     "2^36 cells, periodic on both sides of an axis of >= 2 cells");
  if (field_stride < g.N || field_stride % 64 != 0)
    return nfail(MMF_ERR_INVALID, "field_stride %lld: at least the cell count %lld and a multiple of 64",
                 (long long)field_stride, (long long)g.N);
  g.fs = field_stride;
  g.flags = flags;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return nfail(MMF_ERR_NO_DEVICE, "hipGetDevice: %s", hipGetErrorString(e));
  if (dev < 0 || dev >= MAX_DEVICES) return nfail(MMF_ERR_NO_DEVICE, "device %d beyond %d", dev, MAX_DEVICES);
  Ctrl *ctrl = nullptr;
  int cus = 0;
  const int rs = device_scratch(dev, &ctrl, &cus);
  if (rs != MMF_OK) return rs;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);

  hipLaunchKernelGGL(k_noalloc_begin, dim3(1), dim3(64), 0, s, ctrl);
  e = hipGetLastError();
  if (e != hipSuccess) return nfail(MMF_ERR_HIP, "clearing the scratch: %s", hipGetErrorString(e));
  // the marching kernels where the context's AUTO takes them on a grid without
  // cell flags (3D order 1 from 2^20 cells, 2D order 1 from 2^18, 2D order 2
  // from 2^20); else the per-cell gather kernel.  Every variant gives the same bits.
  const mmf::MarchShape shape = mmf::default_march_shape(g, cus);
  const int shift = (g.dim == 2 && g.order == 1) ? 18 : 20;
  const bool march = !flags && g.N >= ((int64_t)1 << shift) && mmf::march_shape_supported(g, shape);
  // inputs U (R and W are not read by an RHS launch); every output slot is the residual
  e = march ? mmf::launch_stage_march(g, shape, 1, mmf::MODE_RHS_ONLY, U, U, U, rhs, rhs, rhs, ctrl, s)
            : mmf::launch_stage(g, 1, mmf::MODE_RHS_ONLY, U, U, U, rhs, rhs, rhs, ctrl, s);
  if (e != hipSuccess) return nfail(MMF_ERR_HIP, "stage kernel launch: %s", hipGetErrorString(e));
  if (max_eig || counts) {
    hipLaunchKernelGGL(k_noalloc_end, dim3(1), dim3(64), 0, s, ctrl, max_eig, counts);
    e = hipGetLastError();
    if (e != hipSuccess) return nfail(MMF_ERR_HIP, "reading the scratch: %s", hipGetErrorString(e));
  }
  return MMF_OK;
}

}  // extern "C"
