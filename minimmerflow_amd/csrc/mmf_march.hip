// mmf_march.hip -- marching stage kernel for order-1 reconstruction (gfx950).
//
// A block owns a tile of 64 (x, one wave per row) by TY (y) cells and marches
// along the last axis (z in 3D, y in 2D) over a chunk of planes.  Per plane:
//   1. every thread loads its cell (coalesced SoA rows of 512 B per field),
//      computes the cell's flux tuple once (cons2prim, sound speed, total
//      enthalpy: every division and the sqrt of the reference's per-face
//      evaluation, done once per cell) and stores it in LDS; a few threads do
//      the same for the tile's one-cell halo ring;
//   2. the x and y faces of each cell read the neighbour tuples from LDS;
//      the face between this plane and the previous one uses the previous
//      tuple kept in registers, so each z face is evaluated once;
//   3. the previous cell's residual is complete: its RK stage update is
//      applied and written.
// Residual contributions are accumulated per cell in the reference's face
// order (x lower, x upper, y lower, y upper, z lower, z upper), so the result
// is bit-identical to the gather kernel and to the oracle; no float atomics.
// Blocks are numbered so that each of the 8 XCDs takes a contiguous range of
// tiles (neighbouring tiles share that XCD's L2 for the halo reads).
#include <hip/hip_runtime.h>

#include "mmf_device.hpp"
#include "mmf_internal.hpp"
#include "mmf_stage.hpp"

namespace mmf {

namespace {

constexpr int TX = 64;
constexpr int NTUP = 11;  // c[5], vel[3], p, a, H

template <int PL>
__device__ __forceinline__ void tup_store(double *lds, int slot, const Tup &t) {
#pragma unroll
  for (int k = 0; k < NF; ++k) lds[k * PL + slot] = t.c[k];
  lds[5 * PL + slot] = t.vel[0];
  lds[6 * PL + slot] = t.vel[1];
  lds[7 * PL + slot] = t.vel[2];
  lds[8 * PL + slot] = t.p;
  lds[9 * PL + slot] = t.a;
  lds[10 * PL + slot] = t.H;
}

template <int PL>
__device__ __forceinline__ void tup_load(const double *lds, int slot, Tup &t) {
#pragma unroll
  for (int k = 0; k < NF; ++k) t.c[k] = lds[k * PL + slot];
  t.vel[0] = lds[5 * PL + slot];
  t.vel[1] = lds[6 * PL + slot];
  t.vel[2] = lds[7 * PL + slot];
  t.p = lds[8 * PL + slot];
  t.a = lds[9 * PL + slot];
  t.H = lds[10 * PL + slot];
}

// Virtual state across a border side or a wall, as a tuple.
__device__ __forceinline__ void virtual_tup(const GridDev &g, bool wall, int side, int a, const Tup &fluid, Tup &out) {
  if (!wall && g.bc[side] == BC_FREE_FLOW) {  // euler.cpp:298-308: a copy
    out = fluid;
    return;
  }
  double v[NF];
  if (wall) reflect_axis(fluid.c, a, v);  // euler.cpp:347-356
  else border_state(g, side, a, fluid.c, v);
  make_tup(v, out);
}

// Lower and upper faces of this cell along an in-plane axis a; the
// neighbour tuples come from LDS slots sm / sp.  Adds the two contributions
// to acc in reference order.
template <int PL, bool MASK>
__device__ __forceinline__ void inplane_faces(const GridDev &g, const double *lds, const uint8_t *lfl, int a,
                                              bool has_m, bool has_p, int slot_m, int slot_p, bool wrap_first,
                                              const Tup &Tc, uint8_t fc, double *acc, double &lam_max,
                                              Ctrl *ctrl, bool count) {
  double Fm[NF], Fp[NF];
  const uint8_t fm = has_m ? (MASK ? lfl[slot_m] : (uint8_t)(CELL_FLUID | CELL_SOLVED)) : 0;
  const uint8_t fp = has_p ? (MASK ? lfl[slot_p] : (uint8_t)(CELL_FLUID | CELL_SOLVED)) : 0;
  {
    const bool wall = MASK && has_m && ((fm ^ fc) & CELL_FLUID);
    Tup L;
    if (has_m && !wall) tup_load<PL>(lds, slot_m, L);
    else virtual_tup(g, wall, 2 * a, a, Tc, L);
    double lm = face_flux(L, Tc, a, Fm);
    lam_max = (lm < lam_max) ? lam_max : lm;
    const bool other_counts = has_m && !wall && (fm & CELL_SOLVED);
    if (count && !other_counts) count_neg(L, Tc, ctrl);
  }
  {
    const bool wall = MASK && has_p && ((fp ^ fc) & CELL_FLUID);
    Tup R;
    if (has_p && !wall) tup_load<PL>(lds, slot_p, R);
    else virtual_tup(g, wall, 2 * a + 1, a, Tc, R);
    double lp = face_flux(Tc, R, a, Fp);
    lam_max = (lp < lam_max) ? lam_max : lp;
    if (count) count_neg(Tc, R, ctrl);
  }
  const double A = g.A;
  if (wrap_first) {
#pragma unroll
    for (int k = 0; k < NF; ++k) { acc[k] -= A * Fp[k]; acc[k] += A * Fm[k]; }
  } else {
#pragma unroll
    for (int k = 0; k < NF; ++k) { acc[k] += A * Fm[k]; acc[k] -= A * Fp[k]; }
  }
}

// Face between the previous (L) and the current (R) cell along the march
// axis M.  Returns false when the face is inactive (neither side solved).
template <bool MASK>
__device__ __forceinline__ bool march_face(const GridDev &g, int M, bool have_l, bool have_r, const Tup &Tl,
                                           uint8_t fl, const Tup &Tr, uint8_t fr, double *F, double &lam_max,
                                           Ctrl *ctrl, bool count) {
  Tup L, R;
  if (!have_l) {  // lower border of the march axis: virtual from R
    if (!(fr & CELL_SOLVED)) return false;
    virtual_tup(g, false, 2 * M, M, Tr, L);
    R = Tr;
  } else if (!have_r) {  // upper border
    if (!(fl & CELL_SOLVED)) return false;
    L = Tl;
    virtual_tup(g, false, 2 * M + 1, M, Tl, R);
  } else {
    if (!((fl | fr) & CELL_SOLVED)) return false;
    const bool wall = MASK && ((fl ^ fr) & CELL_FLUID);
    if (wall && (fr & CELL_FLUID)) {
      virtual_tup(g, true, 2 * M, M, Tr, L);
      R = Tr;
    } else if (wall) {
      L = Tl;
      virtual_tup(g, true, 2 * M + 1, M, Tl, R);
    } else {
      L = Tl;
      R = Tr;
    }
  }
  double lam = face_flux(L, R, M, F);
  lam_max = (lam < lam_max) ? lam_max : lam;
  if (count) count_neg(L, R, ctrl);
  return true;
}

}  // namespace

template <int DIM, int TY, int STAGE, bool MASK>
__global__ __launch_bounds__(TX *TY) void k_march(GridDev g, StageIn<STAGE> in, const double *U, double *out,
                                                  Ctrl *ctrl, int mode, int chunk, int ntx, int nty) {
  constexpr int NT = TX * TY;
  constexpr int M = DIM - 1;  // march axis
  constexpr int RW = TX + 2;
  constexpr int PL = (TY + 2) * RW;
  __shared__ double lds[NTUP * PL];
  __shared__ uint8_t lfl[MASK ? PL : 1];

  const int slot_eig = (mode == MODE_RHS_ONLY) ? 3 : STAGE - 1;
  if (mode == MODE_STEP && !(ctrl->t < ctrl->tmax)) return;  // main.cpp:377 loop guard
  if (STAGE >= 2) in.dt = step_dt(ctrl);
  const double dt = in.dt;

  // XCD-aware tile numbering
  const int nblk = gridDim.x;
  const int b = blockIdx.x;
  const int tile = (nblk % 8 == 0) ? (b % 8) * (nblk / 8) + b / 8 : b;
  const int bx = tile % ntx;
  const int by = (tile / ntx) % nty;
  const int bz = tile / (ntx * nty);

  const int lane = threadIdx.x, row = threadIdx.y;
  const int tid = row * TX + lane;
  const int64_t nx = g.n[0];
  const int64_t ny = (DIM == 3) ? g.n[1] : 1;
  const int64_t nm = g.n[M];
  const int64_t fs = g.fs;
  const int64_t x0 = (int64_t)bx * TX;
  const int64_t y0 = (int64_t)by * TY;
  const int64_t m0 = (int64_t)bz * chunk;
  const int64_t m1 = (m0 + chunk < nm) ? m0 + chunk : nm;
  const int w = (int)((nx - x0 < TX) ? nx - x0 : TX);
  const int hgt = (DIM == 3) ? (int)((ny - y0 < TY) ? ny - y0 : TY) : 1;
  const bool valid = lane < w && row < hgt;
  const int64_t gx = x0 + lane;
  const int64_t gy = (DIM == 3) ? y0 + row : 0;
  const int64_t sm = (DIM == 3) ? nx * ny : nx;  // stride of the march axis
  const int64_t col = gx + nx * gy;
  const int myslot = (row + 1) * RW + lane + 1;

  // halo ring cell handled by this thread (one per plane), if any
  int64_t hcol = -1;
  int hslot = -1;
  {
    const int H = (DIM == 3) ? 2 * w + 2 * hgt : 2;
    if (tid < H) {
      int64_t hx, hy = 0;
      int sr = 1, sc;
      if (DIM == 3) {
        if (tid < w) { hx = x0 + tid; hy = y0 - 1; sr = 0; sc = tid + 1; }
        else if (tid < 2 * w) { hx = x0 + tid - w; hy = y0 + hgt; sr = hgt + 1; sc = tid - w + 1; }
        else if (tid < 2 * w + hgt) { hx = x0 - 1; hy = y0 + tid - 2 * w; sr = tid - 2 * w + 1; sc = 0; }
        else { hx = x0 + w; hy = y0 + tid - 2 * w - hgt; sr = tid - 2 * w - hgt + 1; sc = w + 1; }
      } else {
        if (tid == 0) { hx = x0 - 1; sc = 0; }
        else { hx = x0 + w; sc = w + 1; }
      }
      bool ok = true;
      if (hx < 0 || hx >= nx) {
        if (g.periodic[0]) hx = (hx + nx) % nx;
        else ok = false;
      }
      if (DIM == 3 && (hy < 0 || hy >= ny)) {
        if (g.periodic[1]) hy = (hy + ny) % ny;
        else ok = false;
      }
      if (ok) {
        hcol = hx + nx * hy;
        hslot = sr * RW + sc;
      }
    }
  }

  const bool has_xm = valid && (gx > 0 || g.periodic[0]);
  const bool has_xp = valid && (gx + 1 < nx || g.periodic[0]);
  const bool wrap_x = g.periodic[0] && gx == nx - 1;
  const bool has_ym = DIM == 3 && valid && (gy > 0 || g.periodic[1]);
  const bool has_yp = DIM == 3 && valid && (gy + 1 < ny || g.periodic[1]);
  const bool wrap_y = DIM == 3 && g.periodic[1] && gy == ny - 1;
  const bool count = true;

  double lam_max = 0.0;
  Tup Tp;                                   // previous plane's tuple (this column)
  uint8_t fprev = CELL_FLUID | CELL_SOLVED;
  double pxy[NF], zlo[NF], Up[NF];
#pragma unroll
  for (int k = 0; k < NF; ++k) pxy[k] = zlo[k] = Up[k] = 0.0;

  // prologue: the plane below the chunk
  int64_t mprev = m0 - 1;
  bool have_prev = true;
  if (mprev < 0) {
    if (g.periodic[M]) mprev += nm;
    else have_prev = false;
  }
  if (valid && have_prev) {
    double X[NF];
    in.load(g, col + mprev * sm, X);
    make_tup(X, Tp);
    fprev = cell_flags<MASK>(g, col + mprev * sm);
  }

  for (int64_t m = m0; m < m1; ++m) {
    const int64_t idx = col + m * sm;
    Tup Tc;
    uint8_t fc = CELL_FLUID | CELL_SOLVED;
    double Uc[NF];
    if (valid) {
      double X[NF];
      if (STAGE == 2) {
#pragma unroll
        for (int k = 0; k < NF; ++k) {
          Uc[k] = U[k * fs + idx];
          X[k] = Uc[k] + div_vol(g, dt * in.R[k * fs + idx]);  // main.cpp:421
        }
      } else {
        in.load(g, idx, X);
      }
      make_tup(X, Tc);
      fc = cell_flags<MASK>(g, idx);
      tup_store<PL>(lds, myslot, Tc);
      if (MASK) lfl[myslot] = fc;
    }
    if (hslot >= 0) {
      double X[NF];
      const int64_t hidx = hcol + m * sm;
      in.load(g, hidx, X);
      Tup Th;
      make_tup(X, Th);
      tup_store<PL>(lds, hslot, Th);
      if (MASK) lfl[hslot] = g.flags[hidx];
    }
    __syncthreads();

    if (valid) {
      // face between plane m-1 and plane m (evaluated once, shared)
      double Fz[NF];
      const bool hl = (m > m0) || have_prev;
      const bool zact = march_face<MASK>(g, M, hl, true, Tp, fprev, Tc, fc, Fz, lam_max, ctrl, count);

      // the previous cell is complete
      if (m > m0) {
        const int64_t pidx = idx - sm;
        if (fprev & CELL_SOLVED) {
          const double A = g.A;
#pragma unroll
          for (int k = 0; k < NF; ++k) {
            double rhs = (pxy[k] + zlo[k]) - A * Fz[k];
            double u = (STAGE == 1) ? 0.0 : (STAGE == 2) ? Up[k] : U[k * fs + pidx];
            out[k * fs + pidx] = rk_update<STAGE>(g, u, Tp.c[k], rhs, dt);
          }
        } else if (STAGE == 1) {
#pragma unroll
          for (int k = 0; k < NF; ++k) out[k * fs + pidx] = 0.0;
        } else if (STAGE == 2) {
#pragma unroll
          for (int k = 0; k < NF; ++k) out[k * fs + pidx] = Up[k];
        }
      }

      // in-plane faces of this cell
      if (fc & CELL_SOLVED) {
        double acc[NF] = {0., 0., 0., 0., 0.};  // euler.cpp:142-148
        inplane_faces<PL, MASK>(g, lds, lfl, 0, has_xm, has_xp, myslot - 1, myslot + 1, wrap_x, Tc, fc, acc,
                                lam_max, ctrl, count);
        if (DIM == 3)
          inplane_faces<PL, MASK>(g, lds, lfl, 1, has_ym, has_yp, myslot - RW, myslot + RW, wrap_y, Tc, fc, acc,
                                  lam_max, ctrl, count);
        const double A = g.A;
#pragma unroll
        for (int k = 0; k < NF; ++k) {
          pxy[k] = acc[k];
          zlo[k] = zact ? A * Fz[k] : 0.0;
        }
      }
      Tp = Tc;
      fprev = fc;
      if (STAGE == 2) {
#pragma unroll
        for (int k = 0; k < NF; ++k) Up[k] = Uc[k];
      }
    }
    __syncthreads();
  }

  // epilogue: the upper face of the chunk's last cell
  if (valid) {
    const int64_t last = m1 - 1;
    const int64_t pidx = col + last * sm;
    int64_t mnext = m1;
    bool have_next = true;
    if (mnext >= nm) {
      if (g.periodic[M]) mnext -= nm;
      else have_next = false;
    }
    Tup Tn;
    uint8_t fn = CELL_FLUID | CELL_SOLVED;
    if (have_next) {
      double X[NF];
      in.load(g, col + mnext * sm, X);
      make_tup(X, Tn);
      fn = cell_flags<MASK>(g, col + mnext * sm);
    }
    double Fz[NF];
    // counted here only at a non-periodic border: otherwise the next chunk's block counts it
    march_face<MASK>(g, M, true, have_next, Tp, fprev, Tn, fn, Fz, lam_max, ctrl, !have_next);
    if (fprev & CELL_SOLVED) {
      const double A = g.A;
      const bool wrap_z = g.periodic[M] && last == nm - 1;
#pragma unroll
      for (int k = 0; k < NF; ++k) {
        double rhs = wrap_z ? (pxy[k] - A * Fz[k]) + zlo[k] : (pxy[k] + zlo[k]) - A * Fz[k];
        double u = (STAGE == 1) ? 0.0 : (STAGE == 2) ? Up[k] : U[k * fs + pidx];
        out[k * fs + pidx] = rk_update<STAGE>(g, u, Tp.c[k], rhs, dt);
      }
    } else if (STAGE == 1) {
#pragma unroll
      for (int k = 0; k < NF; ++k) out[k * fs + pidx] = 0.0;
    } else if (STAGE == 2) {
#pragma unroll
      for (int k = 0; k < NF; ++k) out[k * fs + pidx] = Up[k];
    }
  }
  block_max_to<NT>(&ctrl->eig[slot_eig], lam_max, tid);
}

template <int DIM, int TY, int STAGE, bool MASK>
static hipError_t launch_march_t(const GridDev &g, const MarchShape &sh, int mode, const double *U, const double *R,
                                 const double *W, double *oR, double *oW, double *oU, Ctrl *ctrl, hipStream_t s) {
  StageIn<STAGE> in;
  in.fs = g.fs;
  in.dt = 0.0;
  in.U = U;
  in.R = R;
  in.X = (STAGE == 3) ? W : U;
  double *out = (STAGE == 1) ? oR : (STAGE == 2) ? oW : oU;
  const int M = DIM - 1;
  const int ntx = (int)((g.n[0] + TX - 1) / TX);
  const int nty = (DIM == 3) ? (int)((g.n[1] + TY - 1) / TY) : 1;
  const int nch = (int)((g.n[M] + sh.chunk - 1) / sh.chunk);
  dim3 grid((unsigned)(ntx * nty * nch));
  dim3 block(TX, TY);
  hipLaunchKernelGGL((k_march<DIM, TY, STAGE, MASK>), grid, block, 0, s, g, in, U, out, ctrl, mode, sh.chunk, ntx, nty);
  return hipGetLastError();
}

template <int DIM, int TY, bool MASK>
static hipError_t launch_march_s(const GridDev &g, const MarchShape &sh, int stage, int mode, const double *U,
                                 const double *R, const double *W, double *oR, double *oW, double *oU, Ctrl *ctrl,
                                 hipStream_t s) {
  switch (stage) {
    case 1: return launch_march_t<DIM, TY, 1, MASK>(g, sh, mode, U, R, W, oR, oW, oU, ctrl, s);
    case 2: return launch_march_t<DIM, TY, 2, MASK>(g, sh, mode, U, R, W, oR, oW, oU, ctrl, s);
    default: return launch_march_t<DIM, TY, 3, MASK>(g, sh, mode, U, R, W, oR, oW, oU, ctrl, s);
  }
}

bool march_shape_supported(const GridDev &g, const MarchShape &sh) {
  if (g.order != 1 || sh.chunk < 1) return false;
  if (g.dim == 2) return sh.ty == 1;
  return sh.ty == 4 || sh.ty == 8 || sh.ty == 16;
}

MarchShape default_march_shape(const GridDev &g) {
  MarchShape sh;
  if (g.dim == 2) {
    sh.ty = 1;
    sh.chunk = 16;
  } else {
    sh.ty = 8;
    sh.chunk = 32;
  }
  return sh;
}

hipError_t launch_stage_march(const GridDev &g, const MarchShape &sh, int stage, int mode, const double *U,
                              const double *R, const double *W, double *oR, double *oW, double *oU, Ctrl *ctrl,
                              hipStream_t s) {
  const bool mask = g.flags != nullptr;
  if (g.dim == 2) {
    return mask ? launch_march_s<2, 1, true>(g, sh, stage, mode, U, R, W, oR, oW, oU, ctrl, s)
                : launch_march_s<2, 1, false>(g, sh, stage, mode, U, R, W, oR, oW, oU, ctrl, s);
  }
  switch (sh.ty) {
    case 4:
      return mask ? launch_march_s<3, 4, true>(g, sh, stage, mode, U, R, W, oR, oW, oU, ctrl, s)
                  : launch_march_s<3, 4, false>(g, sh, stage, mode, U, R, W, oR, oW, oU, ctrl, s);
    case 16:
      return mask ? launch_march_s<3, 16, true>(g, sh, stage, mode, U, R, W, oR, oW, oU, ctrl, s)
                  : launch_march_s<3, 16, false>(g, sh, stage, mode, U, R, W, oR, oW, oU, ctrl, s);
    default:
      return mask ? launch_march_s<3, 8, true>(g, sh, stage, mode, U, R, W, oR, oW, oU, ctrl, s)
                  : launch_march_s<3, 8, false>(g, sh, stage, mode, U, R, W, oR, oW, oU, ctrl, s);
  }
}

}  // namespace mmf
