// mmf_stage.hpp -- device helpers shared by the stage kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "mmf_device.hpp"
#include "mmf_internal.hpp"

namespace mmf {

// x / V of the RK update (main.cpp:421,457,493).  When V is a power of two the
// product with 1/V is the same double, and costs one multiply instead of a
// division sequence.
__device__ __forceinline__ double div_vol(const GridDev &g, double x) { return g.pow2V ? x * g.invV : x / g.V; }

// Linear index of this block in its grid.
__device__ __forceinline__ unsigned int block_linear() {
  return blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
}

// Stage maximum s formed by earlier launches (plain loads; largest over the lines).
__device__ __forceinline__ double stage_max(const Ctrl *c, int s) {
  unsigned long long m = 0;
#pragma unroll
  for (int k = 0; k < EIG_SPLIT; ++k) m = (c->eig[s][k][0] > m) ? c->eig[s][k][0] : m;
  return __longlong_as_double((long long)m);
}

// This block's contribution (bits of a non-negative double) to stage maximum s.
__device__ __forceinline__ void stage_max_put(Ctrl *c, int s, unsigned long long bits) {
  if (bits) atomicMax(&c->eig[s][block_linear() % EIG_SPLIT][0], bits);
}

// dt rule of main.cpp:398-402, evaluated identically by every thread.
__device__ __forceinline__ double step_dt(const Ctrl *c) {
  double me = stage_max(c, 0);
  double dt = 0.9 * c->cfl * c->hmin / me;
  if (c->t + dt > c->tmax) dt = c->tmax - c->t;
  return dt;
}

// dt of the current step from the early maximum (the same rule).
__device__ __forceinline__ double step_dt_early(const Ctrl *c) {
  const double me = __longlong_as_double((long long)c->lam_now);
  double dt = 0.9 * c->cfl * c->hmin / me;
  if (c->t + dt > c->tmax) dt = c->tmax - c->t;
  return dt;
}

// This block's contribution (bits of a non-negative double) to the next
// step's early maximum.
__device__ __forceinline__ void lam_put(Ctrl *c, unsigned long long bits) {
  if (bits) atomicMax(&c->lam[block_linear() % EIG_SPLIT][0], bits);
}

// Batched cases keep their stage maxima on line 0 of their control block
// alone (a case's blocks are few, so one line does not serialise them).
__device__ __forceinline__ double step_dt_line0(const Ctrl *c) {
  const double me = __longlong_as_double((long long)c->eig[0][0][0]);
  double dt = 0.9 * c->cfl * c->hmin / me;
  if (c->t + dt > c->tmax) dt = c->tmax - c->t;
  return dt;
}

// The end of a step (main.cpp:505-506), by one whole wave: t += dt, the
// stage maxima kept as the last step's and reset.  Lane 16 s + k reads line
// k of stage s with an atomic, so a launch whose other blocks formed them may
// call it.
// EARLY: dt comes from lam_now, and lanes 48-63 also gather the next step's
// early maximum from the lam lines into lam_now and reset them.
template <bool EARLY = false>
__device__ __forceinline__ void advance_ctrl_wave(Ctrl *c, int lane) {
  unsigned long long v = 0;
  if (lane < 3 * EIG_SPLIT) v = atomicAdd(&c->eig[lane / EIG_SPLIT][lane % EIG_SPLIT][0], 0ull);
  else if (EARLY) v = atomicAdd(&c->lam[lane - 3 * EIG_SPLIT][0], 0ull);
#pragma unroll
  for (int off = EIG_SPLIT / 2; off > 0; off >>= 1) {
    const unsigned long long o = __shfl_xor(v, off, 64);
    v = (o > v) ? o : v;
  }
  const unsigned long long e0 = __shfl(v, 0, 64), e1 = __shfl(v, EIG_SPLIT, 64), e2 = __shfl(v, 2 * EIG_SPLIT, 64);
  const unsigned long long e3 = __shfl(v, 3 * EIG_SPLIT, 64);
  if (lane == 0 && c->t < c->tmax) {
    const unsigned long long me = EARLY ? c->lam_now : e0;
    c->last_eig[0] = __longlong_as_double((long long)e0);
    c->last_eig[1] = __longlong_as_double((long long)e1);
    c->last_eig[2] = __longlong_as_double((long long)e2);
    if (EARLY) c->lam_now = e3;
    if (e0 >= NONFINITE_BITS || e1 >= NONFINITE_BITS || e2 >= NONFINITE_BITS || me >= NONFINITE_BITS) {
      c->nonfinite += 1;  // a NaN or infinite state: the time does not advance
    } else {
      double dt = 0.9 * c->cfl * c->hmin / __longlong_as_double((long long)me);
      if (c->t + dt > c->tmax) dt = c->tmax - c->t;
      c->dt = dt;
      c->t += dt;
      c->step += 1;
    }
  }
  if (lane < 3 * EIG_SPLIT) atomicExch(&c->eig[lane / EIG_SPLIT][lane % EIG_SPLIT][0], 0ull);
  else if (EARLY) atomicExch(&c->lam[lane - 3 * EIG_SPLIT][0], 0ull);
}

// Called by every thread of every block at the very end of a launch that
// closes a step, after block_max_to: the last block to finish advances the
// time for the whole grid.  Thread 0 counts its block on its line only once
// its stage-maximum atomic has completed (s_waitcnt 0; device-scope atomics
// are performed at the point of coherence of all XCDs); the block completing
// a line counts the line, and the block completing the last line reads every
// block's maximum with its own atomics.  Only atomics carry the hand-over,
// so no L2 write-back fence is needed.
template <bool EARLY = false>
__device__ __forceinline__ void close_step(Ctrl *ctrl, int tid) {
  __shared__ int last;
  if (tid == 0) {
    const unsigned int nb = gridDim.x * gridDim.y * gridDim.z, k = block_linear() % EIG_SPLIT;
    const unsigned long long nk = (nb - k + EIG_SPLIT - 1) / EIG_SPLIT;  // blocks on line k
    const unsigned long long nl = (nb < (unsigned)EIG_SPLIT) ? nb : EIG_SPLIT;  // lines in use
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0);
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    int l = 0;
    if (atomicAdd(&ctrl->done[k][0], 1ull) == nk - 1) l = atomicAdd(&ctrl->done[EIG_SPLIT][0], 1ull) == nl - 1;
    last = l;
  }
  __syncthreads();
  if (last && tid < 64) {
    advance_ctrl_wave<EARLY>(ctrl, tid);
    if (tid <= EIG_SPLIT) atomicExch(&ctrl->done[tid][0], 0ull);
  }
}

// Stage input state of one cell: U (stage 1), W1 = U + dt R / V rebuilt on the
// fly (stage 2), W (stage 3).
template <int STAGE>
struct StageIn {
  const double *X;
  const double *U;
  const double *R;
  int64_t fs;
  double dt;
  __device__ __forceinline__ void load(const GridDev &g, int64_t idx, double *c) const {
    if (STAGE == 2) {
#pragma unroll
      for (int k = 0; k < NF; ++k) c[k] = U[k * fs + idx] + div_vol(g, dt * R[k * fs + idx]);  // main.cpp:421
    } else {
#pragma unroll
      for (int k = 0; k < NF; ++k) c[k] = X[k * fs + idx];
    }
  }
};

// Block max of non-negative doubles (compared as u64 bits) into stage
// maximum s, one atomic per block (on the block's line).
template <int NT>
__device__ __forceinline__ void block_max_to(Ctrl *ctrl, int s, double v, int tid) {
  __shared__ unsigned long long red[(NT + 63) / 64];
  unsigned long long b = dbits(v);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    unsigned long long o = __shfl_xor(b, off, 64);
    b = (o > b) ? o : b;
  }
  const int lane = tid & 63, wid = tid >> 6;
  if (lane == 0) red[wid] = b;
  __syncthreads();
  if (tid == 0) {
    unsigned long long m = red[0];
#pragma unroll
    for (int w = 1; w < (NT + 63) / 64; ++w) m = (red[w] > m) ? red[w] : m;
    stage_max_put(ctrl, s, m);
  }
}

// Block max of non-negative doubles (compared as u64 bits) into the next
// step's early maximum, one atomic per block (on the block's line).
template <int NT>
__device__ __forceinline__ void block_lam_to(Ctrl *ctrl, double v, int tid) {
  __shared__ unsigned long long red[(NT + 63) / 64];
  unsigned long long b = dbits(v);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    unsigned long long o = __shfl_xor(b, off, 64);
    b = (o > b) ? o : b;
  }
  const int lane = tid & 63, wid = tid >> 6;
  if (lane == 0) red[wid] = b;
  __syncthreads();
  if (tid == 0) {
    unsigned long long m = red[0];
#pragma unroll
    for (int w = 1; w < (NT + 63) / 64; ++w) m = (red[w] > m) ? red[w] : m;
    lam_put(ctrl, m);
  }
}

// Neighbour of cell q along axis a (dir +-1); false outside a non-periodic border.
__device__ __forceinline__ bool neighbour(const GridDev &g, const int64_t *q, int a, int dir, int64_t &nb) {
  int64_t v = q[a] + dir;
  if (v < 0 || v >= g.n[a]) {
    if (!g.periodic[a]) return false;
    v = (v + g.n[a]) % g.n[a];
  }
  int64_t r[3] = {q[0], q[1], q[2]};
  r[a] = v;
  nb = r[0] + g.n[0] * (r[1] + g.n[1] * r[2]);
  return true;
}

template <bool MASK>
__device__ __forceinline__ uint8_t cell_flags(const GridDev &g, int64_t idx) {
  return MASK ? g.flags[idx] : (uint8_t)(CELL_FLUID | CELL_SOLVED);
}

__device__ __forceinline__ double minmod(double a, double b) {
  if (a > 0.0 && b > 0.0) return (a < b) ? a : b;
  if (a < 0.0 && b < 0.0) return (a > b) ? a : b;
  return 0.0;
}

// Border virtual state (euler.cpp:261-376) of condition bc (Dirichlet data
// info) from the fluid-side face value.
__device__ __forceinline__ void border_state_bc(int bc, const double *info, int a, const double *fluid, double *out) {
  if (bc == BC_DIRICHLET) {
    prim2cons(info, out);
  } else if (bc == BC_REFLECTING || bc == BC_WALL) {
    reflect_axis(fluid, a, out);
  } else {
#pragma unroll
    for (int k = 0; k < NF; ++k) out[k] = fluid[k];
  }
}

// The same with the side's condition (kernels that run on per-side
// conditions only).
__device__ __forceinline__ void border_state(const GridDev &g, int side, int a, const double *fluid, double *out) {
  border_state_bc(g.bc[side], g.info[side], a, fluid, out);
}

// The same for the border face of cell q on side 'side' (axis a): its own
// condition when the grid has per-face conditions.
__device__ __forceinline__ void border_state(const GridDev &g, int side, int a, const int64_t *q,
                                             const double *fluid, double *out) {
  if (g.bslot) {
    // the face's index on the side from the two other axes, the lower fastest
    const int b0 = (a == 0) ? 1 : 0, b1 = (a == 2) ? 1 : 2;
    const int s = g.bslot[g.border_off[side] + q[b0] + g.n[b0] * q[b1]];
    border_state_bc(g.bslot_bc[s], g.bslot_info + NF * s, a, fluid, out);
  } else {
    border_state_bc(g.bc[side], g.info[side], a, fluid, out);
  }
}

// Count negative p / rho of one face's two sides (euler.cpp:94-101).
__device__ __forceinline__ void count_neg(const Tup &L, const Tup &R, Ctrl *ctrl) {
  int np = (L.p < 0.) + (R.p < 0.);
  int nr = (L.c[0] < 0.) + (R.c[0] < 0.);
  if (np) atomicAdd(&ctrl->neg_p, (unsigned long long)np);
  if (nr) atomicAdd(&ctrl->neg_rho, (unsigned long long)nr);
}

// Per-thread counts of non-positive p / rho met on face sides (the same
// totals as count_neg), added to the control block once per wave at the end
// of a launch instead of with atomics inside the loop.
struct NegCount {
  unsigned int p = 0, r = 0;
  __device__ __forceinline__ void add(const Tup &L, const Tup &R) {
    p += (unsigned)(L.p < 0.) + (unsigned)(R.p < 0.);
    r += (unsigned)(L.c[0] < 0.) + (unsigned)(R.c[0] < 0.);
  }
  // add(L, R) where c holds, without a branch
  __device__ __forceinline__ void add_if(bool c, const Tup &L, const Tup &R) {
    p += c ? (unsigned)(L.p < 0.) + (unsigned)(R.p < 0.) : 0u;
    r += c ? (unsigned)(L.c[0] < 0.) + (unsigned)(R.c[0] < 0.) : 0u;
  }
  __device__ __forceinline__ void flush(Ctrl *ctrl, int tid) const {
    unsigned int sp = p, sr = r;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      sp += __shfl_xor(sp, off, 64);
      sr += __shfl_xor(sr, off, 64);
    }
    if ((tid & 63) == 0) {
      if (sp) atomicAdd(&ctrl->neg_p, (unsigned long long)sp);
      if (sr) atomicAdd(&ctrl->neg_rho, (unsigned long long)sr);
    }
  }
};

// RK stage update of one field (main.cpp:421,457,493).
template <int STAGE>
__device__ __forceinline__ double rk_update(const GridDev &g, double u, double x, double rhs, double dt) {
  if (STAGE == 1) return rhs;
  if (STAGE == 2) return 0.75 * u + 0.25 * (x + div_vol(g, dt * rhs));
  return (1. / 3) * u + (2. / 3) * (x + div_vol(g, dt * rhs));
}

}  // namespace mmf
