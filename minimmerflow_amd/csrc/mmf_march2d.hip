// mmf_march2d.hip -- 2D order-1 stage kernels: single grid and persistent batch.
//
// One wave owns an item: a strip of 62 cells along x plus one halo cell on
// each side (lanes 0 and 63; x0 = 62 * tile), marching along y over a chunk
// of rows.  Every lane builds exactly one flux tuple per row (interior or
// halo), so no lane does a second pass.  Tuple rows are double-buffered in
// LDS; one pass over row m does all of:
//   - the tuples of row m + 1 (loaded during row m - 1) into the other buffer;
//   - the two x faces of the cell in row m from the LDS neighbours;
//   - the y face between rows m and m + 1 from the two tuples in registers
//     (each y face once), which completes the cell: its RK stage update.
// Each cell's residual is summed in the reference's face order (x lower, x
// upper, y lower, y upper), bit-identical to the gather kernel and the oracle.
//
// The batch kernel is persistent: a fixed grid of one-wave blocks pulls
// (case, strip, row range) items from an atomic counter until the list is
// exhausted; each case has its own grid, time-control block and dt.
#include <hip/hip_runtime.h>

#include "mmf_batch.hpp"
#include "mmf_device.hpp"
#include "mmf_internal.hpp"
#include "mmf_stage.hpp"

namespace mmf {

namespace {

constexpr int NTUP = 11;
constexpr int W2 = 64;        // lanes
constexpr int INNER = W2 - 2;  // interior cells per strip

__device__ __forceinline__ void tstore(double *L, int slot, const Tup &t) {
#pragma unroll
  for (int k = 0; k < NF; ++k) L[k * W2 + slot] = t.c[k];
  L[5 * W2 + slot] = t.vel[0];
  L[6 * W2 + slot] = t.vel[1];
  L[7 * W2 + slot] = t.vel[2];
  L[8 * W2 + slot] = t.p;
  L[9 * W2 + slot] = t.a;
  L[10 * W2 + slot] = t.H;
}

__device__ __forceinline__ void tload(const double *L, int slot, Tup &t) {
#pragma unroll
  for (int k = 0; k < NF; ++k) t.c[k] = L[k * W2 + slot];
  t.vel[0] = L[5 * W2 + slot];
  t.vel[1] = L[6 * W2 + slot];
  t.vel[2] = L[7 * W2 + slot];
  t.p = L[8 * W2 + slot];
  t.a = L[9 * W2 + slot];
  t.H = L[10 * W2 + slot];
}

__device__ __forceinline__ double ld(const double *base, int64_t fs, int k, uint32_t i) { return base[k * fs + i]; }

template <int STAGE, bool P2>
struct In2 {
  const double *X, *U, *R;
  int64_t fs;
  __device__ __forceinline__ void state(const GridDev &g, double dt, uint32_t i, double *c) const {
#pragma unroll
    for (int k = 0; k < NF; ++k) c[k] = (STAGE == 2) ? ld(U, fs, k, i) + div_vol(g, dt * ld(R, fs, k, i)) : ld(X, fs, k, i);
  }
};

template <int STAGE>
struct Pf2 {
  double a[NF], b[NF];
  template <bool P2>
  __device__ __forceinline__ void issue(const In2<STAGE, P2> &in, uint32_t i) {
#pragma unroll
    for (int k = 0; k < NF; ++k) {
      if (STAGE == 2) {
        a[k] = ld(in.U, in.fs, k, i);
        b[k] = ld(in.R, in.fs, k, i);
      } else {
        a[k] = ld(in.X, in.fs, k, i);
      }
    }
  }
  __device__ __forceinline__ void state(const GridDev &g, double dt, double *X) const {
#pragma unroll
    for (int k = 0; k < NF; ++k) X[k] = (STAGE == 2) ? a[k] + div_vol(g, dt * b[k]) : a[k];
  }
};

__device__ __forceinline__ double vface(const GridDev &g, bool wall, int side, int a, const Tup &Fl, bool fluid_lower,
                                        double *AF, NegCount &neg, bool count) {
  Tup V;
  if (!wall && g.bc[side] == BC_FREE_FLOW) {
    V = Fl;
  } else {
    double v[NF];
    if (wall) reflect_axis(Fl.c, a, v);
    else border_state(g, side, a, Fl.c, v);
    make_tup(v, V);
  }
  double lam;
  if (fluid_lower) {
    lam = face_flux_area(Fl, V, a, g.pow2A, g.A, g.halfA, AF);
    if (count) neg.add(Fl, V);
  } else {
    lam = face_flux_area(V, Fl, a, g.pow2A, g.A, g.halfA, AF);
    if (count) neg.add(V, Fl);
  }
  return lam;
}

// One item: strip 'tile' of rows [m0, m1) of grid g.  U/out are the state
// bases of this grid.  Returns the item's maximum eigenvalue.
template <int STAGE, bool MASK, bool P2>
__device__ double march2d_item(const GridDev &g0, const In2<STAGE, P2> &in, const double *Ub, double *outb, double dt,
                               Ctrl *ctrl, int tile, int m0, int m1, double *ldsb, uint8_t *lflb) {
  GridDev g = g0;
  g.pow2A = P2;
  g.pow2V = P2;
  // P2: faces hand over unscaled sums, the cell's total is scaled by 0.5 A
  // once (the same double: a power-of-two factor commutes with the roundings)
  if (P2) g.halfA = 1.0;
  const double rhs_scale = g0.halfA;
  const int lane = threadIdx.x;
  const int nx = (int)g.n[0];
  const int nm = (int)g.n[1];
  const int64_t fs = in.fs;
  const int x0 = tile * INNER;
  const int w = (nx - x0 < INNER) ? nx - x0 : INNER;
  const int gx = x0 - 1 + lane;
  const bool valid = lane >= 1 && lane <= w;
  bool halo = (lane == 0) || (lane == w + 1);
  int cx = gx;
  if (halo && (cx < 0 || cx >= nx)) {
    if (g.periodic[0]) cx = (cx + nx) % nx;
    else halo = false;
  }
  const bool has_tuple = valid || halo;
  const bool left_halo = x0 > 0 || g.periodic[0];  // lane 0 holds a tuple
  const uint32_t col = (uint32_t)cx;
  const uint32_t sm = (uint32_t)nx;
  const bool has_xm = valid && (gx > 0 || g.periodic[0]);
  const bool has_xp = valid && (gx + 1 < nx || g.periodic[0]);
  // the row above the chunk's last one: the next chunk's first, or row 0
  // across a periodic top (then the chunk's last cell is the periodic top cell)
  const bool top_wrap_here = g.periodic[1] && m1 == nm;
  const bool have_beyond = (m1 < nm) || g.periodic[1];
  const uint32_t beyond = (m1 < nm) ? (uint32_t)m1 : 0u;
  const double A = g.A, halfA = g.halfA;

  double lam_max = 0.0;
  NegCount neg;  // non-positive states met, added to the item's control block at its end
  Tup Tc;  // tuple of this lane's cell in row m
  uint8_t fc = CELL_FLUID | CELL_SOLVED;
  bool bad_c = false;
  double AFlo[NF], Ucur[NF];  // A F of the face below row m; U of the cell in row m
#pragma unroll
  for (int k = 0; k < NF; ++k) AFlo[k] = Ucur[k] = 0.0;

  // ---- prologue: row m0 into LDS buffer 0, and the face below it ----
  {
    const uint32_t idx0 = col + (uint32_t)m0 * sm;
    if (has_tuple) {
      Pf2<STAGE> p0;
      double X[NF];
      p0.issue(in, idx0);
      p0.state(g, dt, X);
      if (STAGE == 2) {
#pragma unroll
        for (int k = 0; k < NF; ++k) Ucur[k] = p0.a[k];
      }
      make_tup(X, Tc);
      bad_c = tup_negative(Tc);
      fc = cell_flags<MASK>(g, idx0);
      tstore(ldsb, lane, Tc);
      if (MASK) lflb[lane] = fc;
    }
    if (valid) {
      int mprev = m0 - 1;
      bool have_prev = true;
      if (mprev < 0) {
        if (g.periodic[1]) mprev += nm;
        else have_prev = false;
      }
      if (!have_prev) {  // lower border
        if (fc & CELL_SOLVED) {
          const double l = vface(g, false, 2, 1, Tc, false, AFlo, neg, true);
          lam_max = lam_acc(lam_max, l);
        }
      } else {
        double Xp[NF];
        Tup Tp;
        const uint32_t pidx = col + (uint32_t)mprev * sm;
        in.state(g, dt, pidx, Xp);
        make_tup(Xp, Tp);
        const uint8_t fp = cell_flags<MASK>(g, pidx);
        // counted here: the item below does not count its chunk's top face
        if ((fp | fc) & CELL_SOLVED) {
          double l;
          if (MASK && ((fp ^ fc) & CELL_FLUID)) {
            if (fc & CELL_FLUID) l = vface(g, true, 2, 1, Tc, false, AFlo, neg, true);
            else l = vface(g, true, 3, 1, Tp, true, AFlo, neg, true);
          } else {
            l = face_flux_area(Tp, Tc, 1, P2, A, halfA, AFlo);
            if (tup_negative(Tp) | bad_c) neg.add(Tp, Tc);
          }
          lam_max = lam_acc(lam_max, l);
        }
      }
    }
  }

  // loads of row m0 + 1 (halo lanes only inside the chunk), the cell flags with them
  Pf2<STAGE> pf;
  uint8_t ff = CELL_FLUID | CELL_SOLVED;
  if ((valid && (m0 + 1 < m1 || have_beyond)) || (halo && m0 + 1 < m1)) {
    const uint32_t i1 = col + ((m0 + 1 < m1) ? (uint32_t)(m0 + 1) : beyond) * sm;
    pf.issue(in, i1);
    ff = cell_flags<MASK>(g, i1);
  }
  __syncthreads();

  int buf = 0;
  for (int m = m0; m < m1; ++m) {
    const uint32_t idx = col + (uint32_t)m * sm;
    const double *L = ldsb + buf * (NTUP * W2);
    const uint8_t *LF = lflb + (MASK ? buf * W2 : 0);
    double *Ln = ldsb + (buf ^ 1) * (NTUP * W2);
    uint8_t *LFn = lflb + (MASK ? (buf ^ 1) * W2 : 0);
    const bool in_chunk = m + 1 < m1;                // row m + 1 is this chunk's
    const bool has_next = in_chunk || have_beyond;  // a cell above this one
    if (STAGE == 3 && valid) {
#pragma unroll
      for (int k = 0; k < NF; ++k) Ucur[k] = ld(Ub, fs, k, idx);
    }

    // ---- row m + 1: tuples into the other LDS buffer (its last readers, row
    // m - 1's x faces, finished before the barrier that ended row m - 1) ----
    Tup Tn;
    uint8_t fn = CELL_FLUID | CELL_SOLVED;
    bool bad_n = false;
    double Unext[NF];
    const bool build = (valid && has_next) || (halo && in_chunk);
    if (build) {
      double X[NF];
      pf.state(g, dt, X);
      if (STAGE == 2) {
#pragma unroll
        for (int k = 0; k < NF; ++k) Unext[k] = pf.a[k];
      }
      make_tup(X, Tn);
      bad_n = tup_negative(Tn);
      fn = ff;
      if (in_chunk) {
        tstore(Ln, lane, Tn);
        if (MASK) LFn[lane] = fn;
      }
    }
    // the loads of row m + 2
    if ((valid && (m + 2 < m1 || (m + 2 == m1 && have_beyond))) || (halo && m + 2 < m1)) {
      const uint32_t i2 = col + ((m + 2 < m1) ? (uint32_t)(m + 2) : beyond) * sm;
      pf.issue(in, i2);
      ff = cell_flags<MASK>(g, i2);
    }

    // ---- without a mask: each x face once.  Lane k (1 <= k <= w + 1)
    // evaluates the face between lanes k - 1 and k, the halo lanes and the
    // lane beyond a border included; a cell takes its upper face from the
    // next lane.  The face a strip shares with the strip on its left is
    // evaluated by both, counted by the left one (as the reference counts
    // each face once). ----
    double AFx[NF] = {0., 0., 0., 0., 0.};
    if (!MASK) {
      if (lane >= 1 && lane <= w + 1) {
        const bool left_t = lane >= 2 || left_halo;  // lane k - 1 holds a tuple
        const bool own_t = lane <= w || halo;        // lane k holds one
        double l;
        if (left_t && own_t) {
          Tup Nb;
          tload(L, lane - 1, Nb);
          l = face_flux_area(Nb, Tc, 0, P2, A, halfA, AFx);
          if ((lane >= 2 || !left_halo) && (tup_negative(Nb) | bad_c)) neg.add(Nb, Tc);
        } else if (own_t) {  // the left border
          l = vface(g, false, 0, 0, Tc, false, AFx, neg, true);
        } else {  // the right border: the fluid cell is the lane below
          Tup Nb;
          tload(L, lane - 1, Nb);
          l = vface(g, false, 1, 0, Nb, true, AFx, neg, true);
        }
        lam_max = lam_acc(lam_max, l);
      }
    }
    double AFup[NF];
#pragma unroll
    for (int k = 0; k < NF; ++k) AFup[k] = MASK ? 0.0 : __shfl_down(AFx[k], 1, W2);

    if (valid) {
      // ---- x faces of this cell from the LDS neighbours, reference order ----
      double acc[NF] = {0., 0., 0., 0., 0.};
      if (!MASK) {
        // first two contributions from zero: their order does not change the bits
#pragma unroll
        for (int k = 0; k < NF; ++k) {
          acc[k] += AFx[k];
          acc[k] -= AFup[k];
        }
      } else if (fc & CELL_SOLVED) {
        double AFm[NF], AFp[NF];
        {
          const uint8_t fm = MASK && has_xm ? LF[lane - 1] : (uint8_t)(CELL_FLUID | CELL_SOLVED);
          const bool wall = MASK && has_xm && ((fm ^ fc) & CELL_FLUID);
          double l;
          if (has_xm && !wall) {
            Tup Nb;
            tload(L, lane - 1, Nb);
            l = face_flux_area(Nb, Tc, 0, P2, A, halfA, AFm);
            if (!(fm & CELL_SOLVED) && (bad_c | tup_negative(Nb))) neg.add(Nb, Tc);
          } else {
            l = vface(g, wall, 0, 0, Tc, false, AFm, neg, true);
          }
          lam_max = lam_acc(lam_max, l);
        }
        {
          const uint8_t fp = MASK && has_xp ? LF[lane + 1] : (uint8_t)(CELL_FLUID | CELL_SOLVED);
          const bool wall = MASK && has_xp && ((fp ^ fc) & CELL_FLUID);
          double l;
          if (has_xp && !wall) {
            Tup Nb;
            tload(L, lane + 1, Nb);
            l = face_flux_area(Tc, Nb, 0, P2, A, halfA, AFp);
            if (bad_c | tup_negative(Nb)) neg.add(Tc, Nb);
          } else {
            l = vface(g, wall, 1, 0, Tc, true, AFp, neg, true);
          }
          lam_max = lam_acc(lam_max, l);
        }
        // first two contributions from zero: their order does not change the bits
#pragma unroll
        for (int k = 0; k < NF; ++k) {
          acc[k] += AFm[k];
          acc[k] -= AFp[k];
        }
      }

      // ---- the y face above this cell (each y face once) ----
      double AFhi[NF] = {0., 0., 0., 0., 0.};
      if (has_next) {
        if ((fc | fn) & CELL_SOLVED) {
          // the chunk's top face and the periodic wrap face are counted by
          // the item that holds the row above
          double l;
          if (MASK && ((fc ^ fn) & CELL_FLUID)) {
            if (fn & CELL_FLUID) l = vface(g, true, 2, 1, Tn, false, AFhi, neg, in_chunk);
            else l = vface(g, true, 3, 1, Tc, true, AFhi, neg, in_chunk);
          } else {
            l = face_flux_area(Tc, Tn, 1, P2, A, halfA, AFhi);
            if (in_chunk && (bad_c | bad_n)) neg.add(Tc, Tn);
          }
          lam_max = lam_acc(lam_max, l);
        }
      } else if (fc & CELL_SOLVED) {  // upper border
        const double l = vface(g, false, 3, 1, Tc, true, AFhi, neg, true);
        lam_max = lam_acc(lam_max, l);
      }

      // ---- this cell is complete: its RK stage update ----
      if (fc & CELL_SOLVED) {
        // the periodic top cell: its upper face (the wrap face) comes first
        const bool wrap_top = top_wrap_here && m == nm - 1;
#pragma unroll
        for (int k = 0; k < NF; ++k) {
          double rhs = wrap_top ? (acc[k] - AFhi[k]) + AFlo[k] : (acc[k] + AFlo[k]) - AFhi[k];
          if (P2) rhs = rhs_scale * rhs;
          outb[k * fs + idx] = rk_update<STAGE>(g, Ucur[k], Tc.c[k], rhs, dt);
        }
      } else if (STAGE != 3) {
#pragma unroll
        for (int k = 0; k < NF; ++k) outb[k * fs + idx] = (STAGE == 1) ? 0.0 : Ucur[k];
      }
#pragma unroll
      for (int k = 0; k < NF; ++k) AFlo[k] = AFhi[k];
    }
    if (build) {
      Tc = Tn;
      fc = fn;
      bad_c = bad_n;
      if (STAGE == 2) {
#pragma unroll
        for (int k = 0; k < NF; ++k) Ucur[k] = Unext[k];
      }
    }
    __syncthreads();
    buf ^= 1;
  }
  neg.flush(ctrl, lane);
  return lam_max;
}

// An unmasked item whose x borders are copies (free flow), periodic or absent,
// with a branch-free row loop: every lane loads (lanes without a cell on a
// clamped column), builds the next row's tuple and evaluates the x face below
// it, and results of lanes without a cell are discarded, so the next row's
// division chain, this row's faces and its update share one basic block and
// the scheduler interleaves them.  A free-flow border face is the Rusanov
// flux between the fluid tuple and itself (euler.cpp:298-308).  The same
// faces, sums and counts as march2d_item<STAGE, false, P2>.
template <int STAGE, bool P2>
__device__ double march2d_item_flat(const GridDev &g0, const In2<STAGE, P2> &in, const double *Ub, double *outb,
                                    double dt, Ctrl *ctrl, int tile, int m0, int m1, double *ldsb) {
  GridDev g = g0;
  g.pow2A = P2;
  g.pow2V = P2;
  if (P2) g.halfA = 1.0;  // unscaled face sums, as in march2d_item
  const double rhs_scale = g0.halfA;
  const int lane = threadIdx.x;
  const int nx = (int)g.n[0];
  const int nm = (int)g.n[1];
  const int64_t fs = in.fs;
  const int x0 = tile * INNER;
  const int w = (nx - x0 < INNER) ? nx - x0 : INNER;
  const int gx = x0 - 1 + lane;
  const bool valid = lane >= 1 && lane <= w;
  bool halo = (lane == 0) || (lane == w + 1);
  int cx = gx;
  if (halo && (cx < 0 || cx >= nx)) {
    if (g.periodic[0]) cx = (cx + nx) % nx;
    else halo = false;
  }
  if (!valid && !halo) cx = (gx < 0) ? 0 : ((gx >= nx) ? nx - 1 : gx);  // a clamped column, results unused
  const bool left_halo = x0 > 0 || g.periodic[0];
  const uint32_t col = (uint32_t)cx;
  const uint32_t sm = (uint32_t)nx;
  const bool top_wrap_here = g.periodic[1] && m1 == nm;
  const bool have_beyond = (m1 < nm) || g.periodic[1];
  const uint32_t beyond = (m1 < nm) ? (uint32_t)m1 : 0u;
  const double A = g.A, halfA = g.halfA;
  const bool face_lane = lane >= 1 && lane <= w + 1;  // evaluates the x face below it
  const bool left_t = lane >= 2 || left_halo;         // lane - 1 holds a tuple
  const bool own_t = lane <= w || halo;               // this lane holds one
  // each face counted once: a strip's first face belongs to the strip on its left
  const bool count_x = face_lane && (lane >= 2 || !left_halo);

  double lam_max = 0.0;
  NegCount neg;
  Tup Tc;
  bool bad_c;
  double AFlo[NF], Ucur[NF];
#pragma unroll
  for (int k = 0; k < NF; ++k) AFlo[k] = Ucur[k] = 0.0;

  // ---- prologue: row m0 into LDS buffer 0, and the face below it ----
  {
    const uint32_t idx0 = col + (uint32_t)m0 * sm;
    Pf2<STAGE> p0;
    double X[NF];
    p0.issue(in, idx0);
    p0.state(g, dt, X);
    if (STAGE == 2) {
#pragma unroll
      for (int k = 0; k < NF; ++k) Ucur[k] = p0.a[k];
    }
    make_tup(X, Tc);
    bad_c = tup_negative(Tc);
    tstore(ldsb, lane, Tc);
    if (valid) {
      int mprev = m0 - 1;
      bool have_prev = true;
      if (mprev < 0) {
        if (g.periodic[1]) mprev += nm;
        else have_prev = false;
      }
      double l;
      if (!have_prev) {  // lower border
        l = vface(g, false, 2, 1, Tc, false, AFlo, neg, true);
      } else {
        double Xp[NF];
        Tup Tp;
        in.state(g, dt, col + (uint32_t)mprev * sm, Xp);
        make_tup(Xp, Tp);
        // counted here: the item below does not count its chunk's top face
        l = face_flux_area(Tp, Tc, 1, P2, A, halfA, AFlo);
        if (tup_negative(Tp) | bad_c) neg.add(Tp, Tc);
      }
      lam_max = lam_acc(lam_max, l);
    }
  }

  // loads of row m0 + 1 (every lane)
  Pf2<STAGE> pf;
#pragma unroll
  for (int k = 0; k < NF; ++k) pf.a[k] = pf.b[k] = 0.0;
  if (m0 + 1 < m1 || have_beyond) pf.issue(in, col + ((m0 + 1 < m1) ? (uint32_t)(m0 + 1) : beyond) * sm);
  __syncthreads();

  int buf = 0;
  for (int m = m0; m < m1; ++m) {
    const uint32_t idx = col + (uint32_t)m * sm;
    const double *L = ldsb + buf * (NTUP * W2);
    double *Ln = ldsb + (buf ^ 1) * (NTUP * W2);
    const bool in_chunk = m + 1 < m1;                // row m + 1 is this chunk's
    const bool has_next = in_chunk || have_beyond;  // a cell above this one
    if (STAGE == 3) {
#pragma unroll
      for (int k = 0; k < NF; ++k) Ucur[k] = ld(Ub, fs, k, idx);
    }
    // ---- row m + 1: tuples into the other LDS buffer ----
    Tup Tn;
    bool bad_n = false;
    double Unext[NF];
    if (has_next) {
      double X[NF];
      pf.state(g, dt, X);
      if (STAGE == 2) {
#pragma unroll
        for (int k = 0; k < NF; ++k) Unext[k] = pf.a[k];
      }
      make_tup(X, Tn);
      bad_n = tup_negative(Tn);
      if (in_chunk) tstore(Ln, lane, Tn);
    }
    // the loads of row m + 2
    if (m + 2 < m1 || (m + 2 == m1 && have_beyond))
      pf.issue(in, col + ((m + 2 < m1) ? (uint32_t)(m + 2) : beyond) * sm);

    // ---- the x face below this lane; a border's copy state is the fluid tuple ----
    double AFx[NF];
    {
      Tup Nb;
      tload(L, lane > 0 ? lane - 1 : 0, Nb);
      Tup Lt, Rt;
#pragma unroll
      for (int k = 0; k < NF; ++k) {
        Lt.c[k] = left_t ? Nb.c[k] : Tc.c[k];
        Rt.c[k] = own_t ? Tc.c[k] : Nb.c[k];
      }
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        Lt.vel[k] = left_t ? Nb.vel[k] : Tc.vel[k];
        Rt.vel[k] = own_t ? Tc.vel[k] : Nb.vel[k];
      }
      Lt.p = left_t ? Nb.p : Tc.p;
      Rt.p = own_t ? Tc.p : Nb.p;
      Lt.a = left_t ? Nb.a : Tc.a;
      Rt.a = own_t ? Tc.a : Nb.a;
      Lt.H = left_t ? Nb.H : Tc.H;
      Rt.H = own_t ? Tc.H : Nb.H;
      const double l = face_flux_area(Lt, Rt, 0, P2, A, halfA, AFx);
      lam_max = lam_acc(lam_max, face_lane ? l : 0.0);
      if (count_x) neg.add(Lt, Rt);
    }
    double AFup[NF];
#pragma unroll
    for (int k = 0; k < NF; ++k) AFup[k] = __shfl_down(AFx[k], 1, W2);

    // ---- the y face above this cell (each y face once) ----
    double AFhi[NF];
    if (has_next) {
      const double l = face_flux_area(Tc, Tn, 1, P2, A, halfA, AFhi);
      lam_max = lam_acc(lam_max, valid ? l : 0.0);
      // the chunk's top face and the periodic wrap face are counted by the item above
      if (valid && in_chunk && (bad_c | bad_n)) neg.add(Tc, Tn);
    } else {  // upper border
      const double l = vface(g, false, 3, 1, Tc, true, AFhi, neg, valid);
      lam_max = lam_acc(lam_max, valid ? l : 0.0);
    }

    // ---- this cell is complete: its RK stage update ----
    const bool wrap_top = top_wrap_here && m == nm - 1;
    double nu[NF];
#pragma unroll
    for (int k = 0; k < NF; ++k) {
      double acc = 0.0;  // first two contributions from zero: their order does not change the bits
      acc += AFx[k];
      acc -= AFup[k];
      double rhs = wrap_top ? (acc - AFhi[k]) + AFlo[k] : (acc + AFlo[k]) - AFhi[k];
      if (P2) rhs = rhs_scale * rhs;
      nu[k] = rk_update<STAGE>(g, Ucur[k], Tc.c[k], rhs, dt);
    }
    if (valid) {
#pragma unroll
      for (int k = 0; k < NF; ++k) outb[k * fs + idx] = nu[k];
    }
#pragma unroll
    for (int k = 0; k < NF; ++k) AFlo[k] = AFhi[k];
    if (has_next) {
      Tc = Tn;
      bad_c = bad_n;
      if (STAGE == 2) {
#pragma unroll
        for (int k = 0; k < NF; ++k) Ucur[k] = Unext[k];
      }
    }
    __syncthreads();
    buf ^= 1;
  }
  neg.flush(ctrl, lane);
  return lam_max;
}

// The flat item applies: the strip's x borders are copies, periodic or absent.
__device__ __forceinline__ bool flat_item(const GridDev &g, int tile) {
  const int nx = (int)g.n[0];
  const int x0 = tile * INNER;
  const bool lo = x0 > 0 || g.periodic[0] || g.bc[0] == BC_FREE_FLOW;
  const bool hi = x0 + INNER < nx || g.periodic[0] || g.bc[1] == BC_FREE_FLOW;
  return lo && hi;
}

// wave-wide max of non-negative doubles as u64 bits; lane 0 holds the result
__device__ __forceinline__ unsigned long long wave_max_bits(double v) {
  unsigned long long b = dbits(v);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    unsigned long long o = __shfl_xor(b, off, 64);
    b = (o > b) ? o : b;
  }
  return b;
}

}  // namespace

// ---- single grid -------------------------------------------------------------
template <int STAGE, bool MASK, bool P2>
__global__ __launch_bounds__(W2) void k_march2d(GridDev g, const double *X, const double *Uin, const double *R,
                                                double *out, Ctrl *ctrl, int mode, int chunk, int ntx) {
  __shared__ double lds[2 * NTUP * W2];
  __shared__ uint8_t lfl[MASK ? 2 * W2 : 1];
  if (mode == MODE_STEP && !(ctrl->t < ctrl->tmax)) return;
  const double dt = (STAGE >= 2) ? step_dt(ctrl) : 0.0;
  const int tile = blockIdx.x % ntx;
  const int bz = blockIdx.x / ntx;
  const int nm = (int)g.n[1];
  const int m0 = bz * chunk;
  const int m1 = (m0 + chunk < nm) ? m0 + chunk : nm;
  In2<STAGE, P2> in{X, Uin, R, g.fs};
  const double lam = (!MASK && flat_item(g, tile))
                         ? march2d_item_flat<STAGE, P2>(g, in, Uin, out, dt, ctrl, tile, m0, m1, lds)
                         : march2d_item<STAGE, MASK, P2>(g, in, Uin, out, dt, ctrl, tile, m0, m1, lds, lfl);
  unsigned long long b = wave_max_bits(lam);
  if (threadIdx.x == 0) stage_max_put(ctrl, (mode == MODE_RHS_ONLY) ? 3 : STAGE - 1, b);
}

template <int STAGE, bool MASK, bool P2>
static hipError_t launch2d_t(const GridDev &g, int chunk, int mode, const double *U, const double *R, const double *W,
                             double *oR, double *oW, double *oU, Ctrl *ctrl, hipStream_t s) {
  const double *X = (STAGE == 3) ? W : U;
  double *out = (STAGE == 1) ? oR : (STAGE == 2) ? oW : oU;
  const int ntx = (int)((g.n[0] + INNER - 1) / INNER);
  const int nch = (int)((g.n[1] + chunk - 1) / chunk);
  hipLaunchKernelGGL((k_march2d<STAGE, MASK, P2>), dim3((unsigned)(ntx * nch)), dim3(W2), 0, s, g, X, U, R, out,
                     ctrl, mode, chunk, ntx);
  return hipGetLastError();
}

template <bool MASK, bool P2>
static hipError_t launch2d_s(const GridDev &g, int chunk, int stage, int mode, const double *U, const double *R,
                             const double *W, double *oR, double *oW, double *oU, Ctrl *ctrl, hipStream_t s) {
  switch (stage) {
    case 1: return launch2d_t<1, MASK, P2>(g, chunk, mode, U, R, W, oR, oW, oU, ctrl, s);
    case 2: return launch2d_t<2, MASK, P2>(g, chunk, mode, U, R, W, oR, oW, oU, ctrl, s);
    default: return launch2d_t<3, MASK, P2>(g, chunk, mode, U, R, W, oR, oW, oU, ctrl, s);
  }
}

hipError_t launch_stage_march2d(const GridDev &g, int chunk, int stage, int mode, const double *U, const double *R,
                                const double *W, double *oR, double *oW, double *oU, Ctrl *ctrl, hipStream_t s) {
  const bool mask = g.flags != nullptr;
  const bool p2 = g.pow2A && g.pow2V;
  if (mask) return p2 ? launch2d_s<true, true>(g, chunk, stage, mode, U, R, W, oR, oW, oU, ctrl, s)
                      : launch2d_s<true, false>(g, chunk, stage, mode, U, R, W, oR, oW, oU, ctrl, s);
  return p2 ? launch2d_s<false, true>(g, chunk, stage, mode, U, R, W, oR, oW, oU, ctrl, s)
            : launch2d_s<false, false>(g, chunk, stage, mode, U, R, W, oR, oW, oU, ctrl, s);
}

int march2d_tiles(int64_t nx) { return (int)((nx + INNER - 1) / INNER); }

// ---- order 2 (limited linear) --------------------------------------------------
// One wave per item: a strip of 60 cells along x plus two halo lanes on each
// side (lanes 0, 1 and w+2, w+3; x0 = 60 * tile), marching along y over a
// chunk of rows with a window of three rows of the stage state in registers.
// For each row c: the y slope of the column (rows c-1, c, c+1) gives the
// cell's lower and upper y face values; the y face between rows c-1 and c is
// evaluated once, from the saved upper value of row c-1 and the lower value
// of row c; the row is staged in LDS for the x slopes, and each x face is
// evaluated once, by the lane on its lower side, from its upper x value and
// the next lane's lower x value (through LDS).  A row is completed one step
// later, when its upper y face is known.  Slopes, face values with their
// positivity fallback (reconstruction.cpp:67-97), border states and the
// residual order are those of the gather kernel: results are bit-identical.
constexpr int W2O = 64;
constexpr int INNER2 = W2O - 4;

namespace {

__device__ __forceinline__ void recon_side(const double *mean, const double *sl, bool upper, double *v) {
#pragma unroll
  for (int k = 0; k < NF; ++k) v[k] = upper ? mean[k] + 0.5 * sl[k] : mean[k] - 0.5 * sl[k];
  double pr[NF];
  cons2prim(v, pr);
  if (!(v[FID_RHO] > 0.0) || !(pr[FID_P] > 0.0)) {
#pragma unroll
    for (int k = 0; k < NF; ++k) v[k] = mean[k];
  }
}

// Area-weighted flux of a border face from the fluid side's face value.
__device__ __forceinline__ double border_face2(const GridDev &g, int side, int a, const Tup &Fl, bool fluid_lower,
                                               double *AF, Ctrl *ctrl) {
  double v[NF];
  border_state(g, side, a, Fl.c, v);
  Tup V;
  make_tup(v, V);
  double lam;
  if (fluid_lower) {
    lam = face_flux_area(Fl, V, a, g.pow2A, g.A, g.halfA, AF);
    count_neg(Fl, V, ctrl);
  } else {
    lam = face_flux_area(V, Fl, a, g.pow2A, g.A, g.halfA, AF);
    count_neg(V, Fl, ctrl);
  }
  return lam;
}

}  // namespace

template <int STAGE, bool P2>
__global__ __launch_bounds__(W2O) void k_march2d_o2(GridDev g, const double *X, const double *Uin, const double *R,
                                                    double *out, Ctrl *ctrl, int mode, int chunk, int ntx) {
  __shared__ double srow[NF][W2O];   // the row's states, for the x slopes
  __shared__ double sfl[NTUP][W2O];  // each lane's lower x face value (tuple)
  __shared__ double saf[NF][W2O];    // each lane's upper x face flux
  if (mode == MODE_STEP && !(ctrl->t < ctrl->tmax)) return;
  const double dt = (STAGE >= 2) ? step_dt(ctrl) : 0.0;
  GridDev gg = g;
  gg.pow2A = P2;
  gg.pow2V = P2;
  const double A = g.A, halfA = g.halfA;
  const int lane = threadIdx.x;
  const int tile = blockIdx.x % ntx, bz = blockIdx.x / ntx;
  const int nx = (int)g.n[0], ny = (int)g.n[1];
  const int m0 = bz * chunk;
  const int m1 = (m0 + chunk < ny) ? m0 + chunk : ny;
  const int x0 = tile * INNER2;
  const int w = (nx - x0 < INNER2) ? nx - x0 : INNER2;
  const int64_t fs = g.fs;
  auto col_of = [&](int l, bool &e) -> int {
    int c = x0 - 2 + l;
    e = l >= 0 && l <= w + 3;
    if (c < 0 || c >= nx) {
      if (g.periodic[0]) c = ((c % nx) + nx) % nx;
      else e = false;
    }
    return c;
  };
  bool ex, exl, exr;
  const int cx = col_of(lane, ex);
  (void)col_of(lane - 1, exl);
  (void)col_of(lane + 1, exr);
  const bool interior = lane >= 2 && lane <= w + 1;
  const bool owner = lane >= 1 && lane <= w + 1;  // x face between this column and the next
  const bool wrap_x = g.periodic[0] && (x0 - 2 + lane) == nx - 1;
  const uint32_t col = (uint32_t)cx;
  auto row_of = [&](int c, bool &e) -> int {
    e = true;
    if (c < 0 || c >= ny) {
      if (g.periodic[1]) return ((c % ny) + ny) % ny;
      e = false;
      return 0;
    }
    return c;
  };
  In2<STAGE, P2> in{X, Uin, R, fs};
  const int last = m1 + 1;  // last row loaded

  // window: rows c-1 (a), c (b), c+1 (cc) of this column
  double Xa[NF], Xb[NF], Xc[NF];
  bool ea = false, eb = false, ec = false;
  {
    bool e;
    int r = row_of(m0 - 2, e);
    ea = e && ex;
    if (ea) in.state(gg, dt, col + (uint32_t)r * (uint32_t)nx, Xa);
    r = row_of(m0 - 1, e);
    eb = e && ex;
    if (eb) in.state(gg, dt, col + (uint32_t)r * (uint32_t)nx, Xb);
    r = row_of(m0, e);
    ec = e && ex;
    if (ec) in.state(gg, dt, col + (uint32_t)r * (uint32_t)nx, Xc);
  }
  Pf2<STAGE> pf;
  bool epf = false;
  {
    bool e;
    const int r = row_of(m0 + 1, e);
    epf = e && ex && (m0 + 1 <= last);
    if (epf) pf.issue(in, col + (uint32_t)r * (uint32_t)nx);
  }

  double lam_max = 0.0;
  Tup TyU;                // upper y face value of row c-1
  double px[NF], pym[NF];  // pending row: x part of its residual, its lower y face
#pragma unroll
  for (int k = 0; k < NF; ++k) px[k] = pym[k] = 0.0;

  for (int c = m0 - 1; c <= m1; ++c) {
    // U of row c-1, completed in this step (stages 2, 3)
    double Uc1[NF];
    if (STAGE >= 2 && interior && c - 1 >= m0) {
#pragma unroll
      for (int k = 0; k < NF; ++k) Uc1[k] = ld(Uin, fs, k, col + (uint32_t)(c - 1) * (uint32_t)nx);
    }
    // y face values of row c
    double FyL[NF], FyU[NF];
    if (eb) {
      double sy[NF];
#pragma unroll
      for (int k = 0; k < NF; ++k) {
        const double dm = ea ? Xb[k] - Xa[k] : 0.0;
        const double dp = ec ? Xc[k] - Xb[k] : 0.0;
        sy[k] = minmod(dp, dm);
      }
      recon_side(Xb, sy, false, FyL);
      recon_side(Xb, sy, true, FyU);
    }
    // y face between rows c-1 and c, once per face
    double AFy[NF];
    if (interior && c >= m0) {
      double l;
      if (ea && eb) {
        Tup TL;
        make_tup(FyL, TL);
        l = face_flux_area(TyU, TL, 1, P2, A, halfA, AFy);
        if (tup_negative(TyU) || tup_negative(TL)) count_neg(TyU, TL, ctrl);
      } else if (eb) {  // lower border of the y axis
        Tup TL;
        make_tup(FyL, TL);
        l = border_face2(gg, 2, 1, TL, false, AFy, ctrl);
      } else {  // upper border of the y axis (c == m1 == ny)
        l = border_face2(gg, 3, 1, TyU, true, AFy, ctrl);
      }
      lam_max = lam_acc(lam_max, l);
      // row c-1 is complete
      if (c - 1 >= m0) {
        const int r = c - 1;
        const bool wrap_y = g.periodic[1] && r == ny - 1;
        const uint32_t idx = col + (uint32_t)r * (uint32_t)nx;
#pragma unroll
        for (int k = 0; k < NF; ++k) {
          const double rhs = wrap_y ? (px[k] - AFy[k]) + pym[k] : (px[k] + pym[k]) - AFy[k];
          out[k * fs + idx] = rk_update<STAGE>(gg, (STAGE == 1) ? 0.0 : Uc1[k], Xa[k], rhs, dt);
        }
      }
    }
    // x faces of row c
    if (c >= m0 && c < m1) {
#pragma unroll
      for (int k = 0; k < NF; ++k) srow[k][lane] = Xb[k];
      __syncthreads();
      double FxL[NF], FxU[NF];
      {
        double sxl[NF];
#pragma unroll
        for (int k = 0; k < NF; ++k) {
          const double xl = (lane > 0) ? srow[k][lane - 1] : 0.0;
          const double xr = (lane < W2O - 1) ? srow[k][lane + 1] : 0.0;
          const double dm = exl ? Xb[k] - xl : 0.0;
          const double dp = exr ? xr - Xb[k] : 0.0;
          sxl[k] = minmod(dp, dm);
        }
        recon_side(Xb, sxl, false, FxL);
        recon_side(Xb, sxl, true, FxU);
      }
      Tup TxL;
      make_tup(FxL, TxL);
      sfl[0][lane] = TxL.c[0];
      sfl[1][lane] = TxL.c[1];
      sfl[2][lane] = TxL.c[2];
      sfl[3][lane] = TxL.c[3];
      sfl[4][lane] = TxL.c[4];
      sfl[5][lane] = TxL.vel[0];
      sfl[6][lane] = TxL.vel[1];
      sfl[7][lane] = TxL.vel[2];
      sfl[8][lane] = TxL.p;
      sfl[9][lane] = TxL.a;
      sfl[10][lane] = TxL.H;
      __syncthreads();
      double AFx[NF];
#pragma unroll
      for (int k = 0; k < NF; ++k) AFx[k] = 0.0;
      if (owner) {
        Tup TR;
        if (lane < W2O - 1) {
#pragma unroll
          for (int k = 0; k < NF; ++k) TR.c[k] = sfl[k][lane + 1];
          TR.vel[0] = sfl[5][lane + 1];
          TR.vel[1] = sfl[6][lane + 1];
          TR.vel[2] = sfl[7][lane + 1];
          TR.p = sfl[8][lane + 1];
          TR.a = sfl[9][lane + 1];
          TR.H = sfl[10][lane + 1];
        }
        double l;
        if (ex && exr) {
          Tup TU;
          make_tup(FxU, TU);
          l = face_flux_area(TU, TR, 0, P2, A, halfA, AFx);
          if (tup_negative(TU) || tup_negative(TR)) count_neg(TU, TR, ctrl);
        } else if (exr) {  // this column is outside a non-periodic x border
          l = border_face2(gg, 0, 0, TR, false, AFx, ctrl);
        } else {  // the next column is outside
          Tup TU;
          make_tup(FxU, TU);
          l = border_face2(gg, 1, 0, TU, true, AFx, ctrl);
        }
        lam_max = lam_acc(lam_max, l);
      }
#pragma unroll
      for (int k = 0; k < NF; ++k) saf[k][lane] = AFx[k];
      __syncthreads();
      if (interior) {
#pragma unroll
        for (int k = 0; k < NF; ++k) {
          const double am = saf[k][lane - 1];
          px[k] = wrap_x ? (0.0 - AFx[k]) + am : (0.0 + am) - AFx[k];
          pym[k] = AFy[k];
        }
      }
    }
    // keep row c's upper y face value, slide the window
    if (eb) make_tup(FyU, TyU);
#pragma unroll
    for (int k = 0; k < NF; ++k) {
      Xa[k] = Xb[k];
      Xb[k] = Xc[k];
    }
    ea = eb;
    eb = ec;
    ec = epf;
    if (epf) pf.state(gg, dt, Xc);
    {
      bool e;
      const int r = row_of(c + 3, e);
      epf = e && ex && (c + 3 <= last);
      if (epf) pf.issue(in, col + (uint32_t)r * (uint32_t)nx);
    }
    __syncthreads();
  }
  const unsigned long long b = wave_max_bits(lam_max);
  if (threadIdx.x == 0) stage_max_put(ctrl, (mode == MODE_RHS_ONLY) ? 3 : STAGE - 1, b);
}

template <int STAGE, bool P2>
static hipError_t launch2d_o2_t(const GridDev &g, int chunk, int mode, const double *U, const double *R,
                                const double *W, double *oR, double *oW, double *oU, Ctrl *ctrl, hipStream_t s) {
  const double *X = (STAGE == 3) ? W : U;
  double *out = (STAGE == 1) ? oR : (STAGE == 2) ? oW : oU;
  const int ntx = (int)((g.n[0] + INNER2 - 1) / INNER2);
  const int nch = (int)((g.n[1] + chunk - 1) / chunk);
  hipLaunchKernelGGL((k_march2d_o2<STAGE, P2>), dim3((unsigned)(ntx * nch)), dim3(W2O), 0, s, g, X, U, R, out, ctrl,
                     mode, chunk, ntx);
  return hipGetLastError();
}

hipError_t launch_stage_march2d_o2(const GridDev &g, int chunk, int stage, int mode, const double *U, const double *R,
                                   const double *W, double *oR, double *oW, double *oU, Ctrl *ctrl, hipStream_t s) {
  const bool p2 = g.pow2A && g.pow2V;
  switch (stage) {
    case 1: return p2 ? launch2d_o2_t<1, true>(g, chunk, mode, U, R, W, oR, oW, oU, ctrl, s)
                      : launch2d_o2_t<1, false>(g, chunk, mode, U, R, W, oR, oW, oU, ctrl, s);
    case 2: return p2 ? launch2d_o2_t<2, true>(g, chunk, mode, U, R, W, oR, oW, oU, ctrl, s)
                      : launch2d_o2_t<2, false>(g, chunk, mode, U, R, W, oR, oW, oU, ctrl, s);
    default: return p2 ? launch2d_o2_t<3, true>(g, chunk, mode, U, R, W, oR, oW, oU, ctrl, s)
                       : launch2d_o2_t<3, false>(g, chunk, mode, U, R, W, oR, oW, oU, ctrl, s);
  }
}

// ---- persistent batch ---------------------------------------------------------
// Stage 1 is held to three waves per SIMD (168 registers, 20 B/lane of
// scratch): 2% faster per step on the 512-case set; stages 2 and 3 need more
// registers and measured slower at three.
template <int STAGE>
__global__ __launch_bounds__(W2, (STAGE == 1) ? 3 : 2) void k_batch2d(const BatchCase *cases, const int *items,
                                                                      BatchQueues qs, unsigned int *counters,
                                                                      double *U, double *R, double *W) {
  __shared__ double lds[2 * NTUP * W2];
  __shared__ uint8_t lfl[1];
  __shared__ int item_s;
  // this block's XCD's queue first (its neighbouring strips share lines in
  // that XCD's L2), then the others in turn once it is drained
  const int home = (int)(blockIdx.x % BATCH_QUEUES);
  int k = 0;
  for (;;) {
    const int q = (home + k) % BATCH_QUEUES;
    if (threadIdx.x == 0) {
      const int n = qs.beg[q + 1] - qs.beg[q];
      const int i = (int)atomicAdd(counters + q * BATCH_COUNTER_LINE, 1u);
      item_s = (i < n) ? qs.beg[q] + i : -1;
    }
    __syncthreads();
    const int it = item_s;
    __syncthreads();
    if (it < 0) {  // queue q is drained: the next one, until all are
      if (++k == BATCH_QUEUES) break;  // every wave of the block reaches this exit
      continue;
    }
    // item: case, strip, rows [m0, m1)
    const int ci = items[4 * it], tile = items[4 * it + 1], m0 = items[4 * it + 2], m1 = items[4 * it + 3];
    const BatchCase &bc = cases[ci];
    Ctrl *ctrl = bc.ctrl;
    if (!(ctrl->t < ctrl->tmax)) continue;  // this case has reached its end time
    const GridDev &g = bc.g;
    const double dt = (STAGE >= 2) ? step_dt_line0(ctrl) : 0.0;
    double *Ub = U + bc.offset, *Rb = R + bc.offset, *Wb = W + bc.offset;
    const double *X = (STAGE == 3) ? Wb : Ub;
    double *out = (STAGE == 1) ? Rb : (STAGE == 2) ? Wb : Ub;
    // batch cases carry no body mask (all cells fluid and solved)
    double lam;
    const bool flat = flat_item(g, tile);
    if (g.pow2A && g.pow2V)
      lam = flat ? march2d_item_flat<STAGE, true>(g, In2<STAGE, true>{X, Ub, Rb, g.fs}, Ub, out, dt, ctrl, tile, m0, m1, lds)
                 : march2d_item<STAGE, false, true>(g, In2<STAGE, true>{X, Ub, Rb, g.fs}, Ub, out, dt, ctrl, tile, m0, m1,
                                                    lds, lfl);
    else
      lam = flat ? march2d_item_flat<STAGE, false>(g, In2<STAGE, false>{X, Ub, Rb, g.fs}, Ub, out, dt, ctrl, tile, m0, m1,
                                                   lds)
                 : march2d_item<STAGE, false, false>(g, In2<STAGE, false>{X, Ub, Rb, g.fs}, Ub, out, dt, ctrl, tile, m0,
                                                     m1, lds, lfl);
    unsigned long long b = wave_max_bits(lam);
    if (threadIdx.x == 0 && b) atomicMax(&ctrl->eig[STAGE - 1][0][0], b);
  }
}

// t += dt per case (main.cpp:505-506), resets the stage maxima and the item counters.
__global__ void k_batch_advance(const BatchCase *cases, int ncases, unsigned int *counters) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < 3 * BATCH_QUEUES) counters[i * BATCH_COUNTER_LINE] = 0u;
  if (i >= ncases) return;
  Ctrl *c = cases[i].ctrl;
  if (c->t < c->tmax) {
    const unsigned long long e0 = c->eig[0][0][0], e1 = c->eig[1][0][0], e2 = c->eig[2][0][0];
    if (e0 >= NONFINITE_BITS || e1 >= NONFINITE_BITS || e2 >= NONFINITE_BITS) {
      c->nonfinite += 1;  // a NaN or infinite state: the time does not advance
    } else {
      double dt = step_dt_line0(c);
      c->dt = dt;
      c->t += dt;
      c->step += 1;
    }
#pragma unroll
    for (int s = 0; s < 3; ++s) c->last_eig[s] = __longlong_as_double((long long)c->eig[s][0][0]);
  }
#pragma unroll
  for (int s = 0; s < 3; ++s) c->eig[s][0][0] = 0ull;
}

hipError_t launch_batch_stage(int stage, const BatchCase *cases, const int *items, const BatchQueues &qs,
                              unsigned int *counters, double *U, double *R, double *W, int nblocks, hipStream_t s) {
  unsigned int *c = counters + (stage - 1) * BATCH_QUEUES * BATCH_COUNTER_LINE;
  switch (stage) {
    case 1: hipLaunchKernelGGL(k_batch2d<1>, dim3(nblocks), dim3(W2), 0, s, cases, items, qs, c, U, R, W); break;
    case 2: hipLaunchKernelGGL(k_batch2d<2>, dim3(nblocks), dim3(W2), 0, s, cases, items, qs, c, U, R, W); break;
    default: hipLaunchKernelGGL(k_batch2d<3>, dim3(nblocks), dim3(W2), 0, s, cases, items, qs, c, U, R, W); break;
  }
  return hipGetLastError();
}

int batch_stage_occupancy(int stage) {
  int n = 0;
  const void *k = (stage == 1) ? (const void *)k_batch2d<1> : (stage == 2) ? (const void *)k_batch2d<2>
                                                                          : (const void *)k_batch2d<3>;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k, W2, 0) != hipSuccess) return 0;
  return n;
}

hipError_t launch_batch_advance(const BatchCase *cases, int ncases, unsigned int *counters, hipStream_t s) {
  const int nb = (ncases + 255) / 256;
  hipLaunchKernelGGL(k_batch_advance, dim3(nb > 0 ? nb : 1), dim3(256), 0, s, cases, ncases, counters);
  return hipGetLastError();
}

}  // namespace mmf
