// mmf_march.hip -- marching stage kernel for order-1 reconstruction (gfx950).
//
// A block owns a tile of 64 (x, one wave per row) by TY (y) cells of a 3D grid
// and marches along z over a chunk of planes.  Each thread turns its cell's
// state into the cell's flux tuple -- cons2prim, sound speed, total enthalpy:
// every division and the sqrt of the reference's per-face evaluation, done
// once per cell -- and the first threads do the same for the tile's one-cell
// halo ring.  Tuple planes are double-buffered in LDS, and one pass over plane
// m, between two barriers, does all of:
//   - the tuples of plane m + 1 (its state loaded during plane m - 1, so the
//     HBM latency hides behind a whole plane of work) into the other buffer,
//     whose last readers finished before the barrier that ended plane m - 1;
//   - the x and y faces of the cell in plane m, from the neighbour tuples in
//     LDS, evaluated while plane m + 1's division chains are in flight;
//   - the z face between planes m and m + 1 from the two tuples in registers
//     (each z face once), which completes the cell: its RK stage update is
//     written in the same pass.
// Residual contributions are accumulated per cell in the reference's face
// order (x lower, x upper, y lower, y upper, z lower, z upper), so the result
// is bit-identical to the gather kernel and to the oracle; no float atomics.
// Blocks are numbered so that each of the 8 XCDs takes a contiguous range of
// tiles (neighbouring tiles share that XCD's L2 for the halo reads).
// Cell offsets are 32-bit on uniform per-field base pointers (grids below
// 2^29 cells; larger grids use the gather kernel).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "mmf_device.hpp"
#include "mmf_internal.hpp"
#include "mmf_stage.hpp"
#include "mmf_batch.hpp"

namespace mmf {

namespace {

constexpr int TX2D = 64;  // the 2D strip kernels' wave width (default_march_shape)
constexpr int NTUP = 11;  // c[5], vel[3], p, a, H
constexpr int64_t MARCH_MAX_CELLS = (int64_t)1 << 29;

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

// One SoA state: a uniform base and field stride; the per-lane part of an
// address is a 32-bit byte offset (scalar-base + vector-offset addressing).
struct Fields {
  const char *base;
  uint64_t fsb;  // field stride in bytes
  __device__ __forceinline__ void set(const double *b, int64_t fs) {
    base = (const char *)b;
    fsb = (uint64_t)fs * sizeof(double);
  }
  __device__ __forceinline__ double at(int k, uint32_t i) const {
    return *(const double *)(base + k * fsb + (uint32_t)(i * (uint32_t)sizeof(double)));
  }
};

struct OutFields {
  char *base;
  uint64_t fsb;
  __device__ __forceinline__ void set(double *b, int64_t fs) {
    base = (char *)b;
    fsb = (uint64_t)fs * sizeof(double);
  }
  __device__ __forceinline__ void put(int k, uint32_t i, double v) const {
    *(double *)(base + k * fsb + (uint32_t)(i * (uint32_t)sizeof(double))) = v;
  }
};

// Stage input of one cell from the stage's source fields.  RB: stage 2
// rebuilding W1 = U + dt R / V from U and R; otherwise the state X (U in
// stage 1, W1 in stage 2 with early dt, W in stage 3).
template <bool RB>
struct Source {
  Fields X;
  Fields Uf, Rf;
  __device__ __forceinline__ void load(const GridDev &g, double dt, uint32_t i, double *c) const {
#pragma unroll
    for (int k = 0; k < NF; ++k)
      c[k] = RB ? Uf.at(k, i) + div_vol(g, dt * Rf.at(k, i)) : X.at(k, i);  // main.cpp:421
  }
};

template <bool RB>
struct Prefetch {
  double a[NF];  // X, or U (RB)
  double b[NF];  // R (RB)
  __device__ __forceinline__ void issue(const Source<RB> &src, uint32_t i) {
#pragma unroll
    for (int k = 0; k < NF; ++k) {
      if (RB) {
        a[k] = src.Uf.at(k, i);
        b[k] = src.Rf.at(k, i);
      } else {
        a[k] = src.X.at(k, i);
      }
    }
  }
  __device__ __forceinline__ void state(const GridDev &g, double dt, double *X) const {
#pragma unroll
    for (int k = 0; k < NF; ++k) X[k] = RB ? a[k] + div_vol(g, dt * b[k]) : a[k];
  }
};

// The tuple of each Dirichlet border's virtual state (euler.cpp:261-296: the
// side's prescribed primitives, whatever the fluid state), built once per
// block into bct[side * NTUP ...] by the first threads.
__device__ __forceinline__ void dirichlet_tuples(const GridDev &g, double *bct, int tid) {
  if (tid < 6 && g.bc[tid] == BC_DIRICHLET) {
    double v[NF];
    Tup T;
    prim2cons(g.info[tid], v);
    make_tup(v, T);
    double *o = bct + tid * NTUP;
#pragma unroll
    for (int k = 0; k < NF; ++k) o[k] = T.c[k];
    o[5] = T.vel[0];
    o[6] = T.vel[1];
    o[7] = T.vel[2];
    o[8] = T.p;
    o[9] = T.a;
    o[10] = T.H;
  }
}

// Area-weighted flux AF of a face whose other side is a virtual state (border
// side 'side' or a wall), built from the fluid tuple Fl.  fluid_lower: the
// fluid cell is the lower (L) side of the face.  REFL false: the grid has no
// body and no reflecting or wall border, so only the copy and Dirichlet forms
// occur (the reflection, a state conversion per face, is compiled out).
template <bool REFL>
__device__ __forceinline__ double virtual_face(const GridDev &g, const double *bct, bool wall, int side, int a,
                                               const Tup &Fl, bool fluid_lower, double *AF, NegCount &neg, bool count) {
  Tup V;
  if (!wall && g.bc[side] == BC_FREE_FLOW) {  // euler.cpp:298-308: a copy
    V = Fl;
  } else if (!wall && g.bc[side] == BC_DIRICHLET) {
    const double *o = bct + side * NTUP;
#pragma unroll
    for (int k = 0; k < NF; ++k) V.c[k] = o[k];
    V.vel[0] = o[5];
    V.vel[1] = o[6];
    V.vel[2] = o[7];
    V.p = o[8];
    V.a = o[9];
    V.H = o[10];
  } else if (REFL) {
    double v[NF];
    if (wall) reflect_axis(Fl.c, a, v);  // euler.cpp:347-356
    else border_state(g, side, a, Fl.c, v);
    make_tup(v, V);
  } else {
    V = Fl;  // not reached: no other virtual state in a REFL-free instance
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

// The largest of a cell's per-axis eigenvalues |u_a| + sqrt(gamma T)
// (euler.cpp:57-66 for n = +-e_a) with four divisions instead of the tuple's
// seven, the same double: a correctly rounded division is odd in its
// divisor and non-decreasing in its dividend, so max_a |fl(m_a / rho)| =
// fl(max_a |m_a| / |rho|), and fl(x + a) is non-decreasing in x; T is
// cons2prim's (utils.cpp:48-63) to the bit.  A NaN momentum wins the u64
// maximum and stays NaN.
template <int DIM>
__device__ __forceinline__ double cell_max_eig(const double *c) {
  double m = fabs(c[FID_RHO_U]);
#pragma unroll
  for (int a = 1; a < DIM; ++a) m = lam_acc(m, fabs(c[FID_RHO_U + a]));
  const double umax = m / fabs(c[FID_RHO]);
  const double K = (c[FID_RHO_U] * c[FID_RHO_U] + c[FID_RHO_V] * c[FID_RHO_V] + c[FID_RHO_W] * c[FID_RHO_W]) /
                   (c[FID_RHO] * c[FID_RHO]);
  const double T = (2.0 * c[FID_RHO_E] / c[FID_RHO] - K) / (2.0 / (GAMMA - 1.0));
  return umax + sqrt(GAMMA * T);
}

}  // namespace

// One block's march over tile 'tile' (x, y tile and chunk).  MASK: the cell
// flags are read (bodies).  EARLY is the launch's data flow (early dt), so a
// block of a masked launch that runs the unmasked code keeps it.  lds, lfl
// and bct are the kernel's LDS arrays.
template <int DIM, int TX, int TY, int STAGE, bool MASK, bool P2, bool REFL, bool EARLY>
__device__ __forceinline__ void march_block(const GridDev &g, StageIn<STAGE> in, const double *U, double *out,
                                            Ctrl *ctrl, int mode, int bx, int by, int m0, int m1, double *lds,
                                            uint8_t *lfl, double *bct, double &lam_max, double &lam_nx,
                                            NegCount &neg) {
  constexpr int M = DIM - 1;  // march axis
  constexpr int RW = TX + 2;
  constexpr int PL = (TY + 2) * RW;
  // field stride of a tuple plane in LDS: the tile and its halo ring, then one
  // slot per border side holding that side's Dirichlet tuple (both buffers)
  constexpr int PS = PL + 8;
  // Early dt (no body, no reflected border): the step's dt is known before
  // stage 1 (Ctrl::lam_now), stage 1 writes W1 = U + dt R / V itself, stage 2
  // reads W1 like stage 3 reads W, and stage 3 gathers the next step's
  // maximum from the state it writes.  Otherwise stage 1 writes R and stage 2
  // rebuilds W1 from U and R (RB).
  constexpr bool RB = STAGE == 2 && !EARLY;
  const bool w1 = EARLY && STAGE == 1 && mode == MODE_STEP;
  const double dt = EARLY ? ((mode == MODE_STEP) ? step_dt_early(ctrl) : 0.0) : ((STAGE >= 2) ? step_dt(ctrl) : 0.0);
  dirichlet_tuples(g, bct, threadIdx.x + TX * threadIdx.y);
  __syncthreads();
  if constexpr (DIM == 3 && !MASK && !REFL) {
    // the Dirichlet tuples as the border slots of both buffers (the flat
    // plane pass below loads a border face's virtual state like a neighbour)
    const int t = threadIdx.x + TX * threadIdx.y;
    if (t < 6 * NTUP) {
      const int side = t / NTUP, k = t % NTUP;
      lds[k * PS + PL + side] = bct[side * NTUP + k];
      lds[NTUP * PS + k * PS + PL + side] = bct[side * NTUP + k];
    }
  }

  Source<RB> src;
  src.X.set(in.X, g.fs);
  src.Uf.set(in.U, g.fs);
  src.Rf.set(in.R, g.fs);
  Fields Uo;  // U for the RK update
  Uo.set(U, g.fs);
  OutFields outf;
  outf.set(out, g.fs);

  const int lane = threadIdx.x, row = threadIdx.y;
  const int tid = row * TX + lane;
  const int nx = (int)g.n[0];
  const int ny = (DIM == 3) ? (int)g.n[1] : 1;
  const int nm = (int)g.n[M];
  const int x0 = bx * TX;
  const int y0 = by * TY;
  const int w = (nx - x0 < TX) ? nx - x0 : TX;
  const int hgt = (DIM == 3) ? ((ny - y0 < TY) ? ny - y0 : TY) : 1;
  const bool valid = lane < w && row < hgt;
  const int gx = x0 + lane;
  const int gy = (DIM == 3) ? y0 + row : 0;
  const uint32_t sm = (uint32_t)nx * (uint32_t)ny;  // stride of the march axis
  const uint32_t col = (uint32_t)gx + (uint32_t)nx * (uint32_t)gy;
  const int myslot = (row + 1) * RW + lane + 1;

  // halo ring cell handled by this thread (one per plane), if any
  uint32_t hcol = 0;
  int hslot = -1;
  {
    const int H = (DIM == 3) ? 2 * w + 2 * hgt : 2;
    if (tid < H) {
      int hx, hy = 0;
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
        hcol = (uint32_t)hx + (uint32_t)nx * (uint32_t)hy;
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
  // the plane above the chunk's last one: the next chunk's first, or plane 0
  // across a periodic top (then the chunk's last cell is the periodic top cell)
  const bool top_wrap_here = g.periodic[M] && m1 == nm;
  const bool have_beyond = (m1 < nm) || g.periodic[M];
  const uint32_t beyond = (m1 < nm) ? (uint32_t)m1 : 0u;
  // P2: A and V are powers of two (compile-time, so no select between the forms)
  const bool pow2A = P2;
  GridDev gp = g;
  gp.pow2A = P2;
  gp.pow2V = P2;
  // P2: the faces hand over their unscaled sums X (halfA = 1, the multiply
  // folds away) and the cell's total is scaled by 0.5 A once at the update.
  // A power-of-two factor commutes with every rounding of the sum, so this is
  // the double of sum(0.5 A X) unless a partial sum leaves the normal range.
  if (P2) gp.halfA = 1.0;
  const double A = g.A, halfA = P2 ? 1.0 : g.halfA;
  const double rhs_scale = g.halfA;

  // Early dt with a body, a reflecting or wall border or a Dirichlet border
  // on a grid with flags: the next step's stage-1 faces of this solved cell
  // also take, along each axis with a wall face (a non-fluid neighbour) or
  // a reflecting / wall border, the reflected state's |u_a| + c (the virtual
  // state of virtual_face, euler.cpp:318-356), and at a Dirichlet border on a
  // grid with flags the border state's (a grid without flags adds those once
  // per block: every border cell is solved there).
  auto virtual_max_eig = [&](const GridDev &gg, const double *c, const uint8_t *LFp, int mm, uint8_t fbelow,
                             uint8_t fabove, bool above) {
    double l = 0.0;
#pragma unroll
    for (int a = 0; a < DIM; ++a) {
      bool refl = false;
#pragma unroll
      for (int d = 0; d < 2; ++d) {
        const int side = 2 * a + d;
        bool has;
        uint8_t fnb = CELL_FLUID | CELL_SOLVED;
        if (a == M) {
          has = d ? above : (mm > 0 || gg.periodic[M]);
          fnb = d ? fabove : fbelow;
        } else {
          has = (a == 0) ? (d ? has_xp : has_xm) : (d ? has_yp : has_ym);
          if (MASK && has) fnb = LFp[myslot + (d ? 1 : -1) * ((a == 0) ? 1 : RW)];
        }
        if (has) {
          refl = refl || (MASK && !(fnb & CELL_FLUID));
        } else {
          refl = refl || (REFL && (gg.bc[side] == BC_REFLECTING || gg.bc[side] == BC_WALL));
          if (gg.flags != nullptr && gg.bc[side] == BC_DIRICHLET)
            l = lam_acc(l, fabs(bct[side * NTUP + 5 + a]) + bct[side * NTUP + 9]);
        }
      }
      if (refl) {
        double v[NF];
        Tup V;
        reflect_axis(c, a, v);
        make_tup(v, V);
        l = lam_acc(l, fabs(V.vel[a]) + V.a);
      }
    }
    return l;
  };

  Tup Tc;  // tuple of this column's cell in plane m
  uint8_t fc = CELL_FLUID | CELL_SOLVED;
  uint8_t fb = CELL_FLUID | CELL_SOLVED;  // the cell below plane m (MASK)
  bool bad_c = false;   // Tc has a negative p or rho
  double AFlo[NF];      // A F of the face below plane m
  double Ucur[NF];      // U of the cell in plane m (stage 2: carried from its load)
#pragma unroll
  for (int k = 0; k < NF; ++k) AFlo[k] = Ucur[k] = 0.0;

  // ---- prologue: plane m0 into LDS buffer 0, and the face below it ----
  {
    const uint32_t idx0 = col + (uint32_t)m0 * sm;
    if (valid) {
      Prefetch<RB> p0;
      double X[NF];
      p0.issue(src, idx0);
      p0.state(gp, dt, X);
      if (RB) {
#pragma unroll
        for (int k = 0; k < NF; ++k) Ucur[k] = p0.a[k];
      }
      make_tup(X, Tc);
      bad_c = tup_negative(Tc);
      fc = cell_flags<MASK>(g, idx0);
      tup_store<PS>(lds, myslot, Tc);
      if (MASK) lfl[myslot] = fc;
    }
    if (hslot >= 0) {
      double Xh[NF];
      Tup Th;
      src.load(gp, dt, hcol + (uint32_t)m0 * sm, Xh);
      make_tup(Xh, Th);
      tup_store<PS>(lds, hslot, Th);
      if (MASK) lfl[hslot] = g.flags[hcol + (uint32_t)m0 * sm];
    }
    if (valid) {
      int mprev = m0 - 1;
      bool have_prev = true;
      if (mprev < 0) {
        if (g.periodic[M]) mprev += nm;
        else have_prev = false;
      }
      if (!have_prev) {  // lower border of the march axis
        if (fc & CELL_SOLVED) {
          const double l = virtual_face<REFL>(gp, bct, false, 2 * M, M, Tc, false, AFlo, neg, true);
          lam_max = lam_acc(lam_max, l);
        }
      } else {
        double Xp[NF];
        Tup Tp;
        const uint32_t pidx = col + (uint32_t)mprev * sm;
        src.load(gp, dt, pidx, Xp);
        make_tup(Xp, Tp);
        const uint8_t fp = cell_flags<MASK>(g, pidx);
        fb = fp;
        // counted here: the block below does not count its chunk's top face
        if ((fp | fc) & CELL_SOLVED) {
          double l;
          if (MASK && ((fp ^ fc) & CELL_FLUID)) {
            if (fc & CELL_FLUID) l = virtual_face<REFL>(gp, bct, true, 2 * M, M, Tc, false, AFlo, neg, true);
            else l = virtual_face<REFL>(gp, bct, true, 2 * M + 1, M, Tp, true, AFlo, neg, true);
          } else {
            l = face_flux_area(Tp, Tc, M, pow2A, A, halfA, AFlo);
            if (tup_negative(Tp) | bad_c) neg.add(Tp, Tc);
          }
          lam_max = lam_acc(lam_max, l);
        }
      }
    }
  }

  // loads of plane m0 + 1 (or of the plane beyond a one-plane chunk), the
  // cell flags with them
  Prefetch<RB> pf, ph;
  uint8_t ff = CELL_FLUID | CELL_SOLVED, fh = CELL_FLUID | CELL_SOLVED;
  if (valid && (m0 + 1 < m1 || have_beyond)) {
    const uint32_t i1 = col + ((m0 + 1 < m1) ? (uint32_t)(m0 + 1) : beyond) * sm;
    pf.issue(src, i1);
    ff = cell_flags<MASK>(g, i1);
  }
  if (hslot >= 0 && m0 + 1 < m1) {
    ph.issue(src, hcol + (uint32_t)(m0 + 1) * sm);
    if (MASK) fh = g.flags[hcol + (uint32_t)(m0 + 1) * sm];
  }
  __syncthreads();

  int buf = 0;
  int m = m0;
  if constexpr (DIM == 3 && !MASK && !REFL && !RB) {
    // Flat pass: a full tile whose x and y borders are free flow or Dirichlet
    // (no periodic wrap), every plane but the chunk's last.  Every lane is a
    // cell, every face of the plane is evaluated from a tuple in LDS (a
    // border face's virtual state sits in a slot of its own: the cell's own
    // tuple for free flow, the side's Dirichlet tuple), the plane above is in
    // the chunk, and the negative-state counts are added without branches, so
    // the pass (next plane's tuple, the in-plane faces, the face above, the
    // update) is one straight block whose division chains and face fluxes the
    // compiler interleaves.  The operations and their order are those of the
    // general pass below, so the results are the same doubles and counts.
    bool flat = w == TX && hgt == TY && !g.periodic[0] && !g.periodic[1];
#pragma unroll
    for (int sd = 0; sd < 4; ++sd) flat = flat && (g.bc[sd] == BC_FREE_FLOW || g.bc[sd] == BC_DIRICHLET);
    if (flat) {
      const int s_nb[4] = {has_xm ? myslot - 1 : (g.bc[0] == BC_DIRICHLET ? PL + 0 : myslot),
                           has_xp ? myslot + 1 : (g.bc[1] == BC_DIRICHLET ? PL + 1 : myslot),
                           has_ym ? myslot - RW : (g.bc[2] == BC_DIRICHLET ? PL + 2 : myslot),
                           has_yp ? myslot + RW : (g.bc[3] == BC_DIRICHLET ? PL + 3 : myslot)};
      const bool border_m[2] = {!has_xm, !has_ym};
      for (; m + 1 < m1; ++m) {
        const uint32_t idx = col + (uint32_t)m * sm;
        const double *L = lds + buf * (NTUP * PS);
        double *Ln = lds + (buf ^ 1) * (NTUP * PS);
        if (STAGE >= 2) {
#pragma unroll
          for (int k = 0; k < NF; ++k) Ucur[k] = Uo.at(k, idx);
        }
        // plane m + 1's tuple, then the loads of plane m + 2 (plane m + 1 again
        // where no plane m + 2 is needed: an in-bounds load left unused)
        Tup Tn;
        {
          double X[NF];
          pf.state(gp, dt, X);
          make_tup(X, Tn);
          tup_store<PS>(Ln, myslot, Tn);
          const uint32_t p2 = (m + 2 < m1) ? (uint32_t)(m + 2) : (have_beyond ? beyond : (uint32_t)(m + 1));
          pf.issue(src, col + p2 * sm);
        }
        // in-plane faces, reference order
        double acc[NF] = {0., 0., 0., 0., 0.};
#pragma unroll
        for (int a = 0; a < DIM - 1; ++a) {
          double AFm[NF], AFp[NF];
          Tup Nm, Np;
          tup_load<PS>(L, s_nb[2 * a], Nm);
          tup_load<PS>(L, s_nb[2 * a + 1], Np);
          const double lm = face_flux_area(Nm, Tc, a, pow2A, A, halfA, AFm);
          neg.add_if(border_m[a], Nm, Tc);
          const double lp = face_flux_area(Tc, Np, a, pow2A, A, halfA, AFp);
          neg.add(Tc, Np);
          lam_max = lam_acc(lam_max, lm);
          lam_max = lam_acc(lam_max, lp);
#pragma unroll
          for (int k = 0; k < NF; ++k) { acc[k] += AFm[k]; acc[k] -= AFp[k]; }
        }
        // the face above, which completes the cell
        double AFhi[NF];
        {
          const double l = face_flux_area(Tc, Tn, M, pow2A, A, halfA, AFhi);
          neg.add(Tc, Tn);
          lam_max = lam_acc(lam_max, l);
        }
        double nu[NF];
#pragma unroll
        for (int k = 0; k < NF; ++k) {
          double rhs = (acc[k] + AFlo[k]) - AFhi[k];
          if (P2) rhs = rhs_scale * rhs;
          nu[k] = w1 ? Tc.c[k] + div_vol(gp, dt * rhs) : rk_update<STAGE>(gp, Ucur[k], Tc.c[k], rhs, dt);
          outf.put(k, idx, nu[k]);
        }
        if (EARLY && STAGE == 3) lam_nx = lam_acc(lam_nx, cell_max_eig<DIM>(nu));
#pragma unroll
        for (int k = 0; k < NF; ++k) AFlo[k] = AFhi[k];
        Tc = Tn;
        // halo ring of plane m + 1 (waves 0-2) and its loads of plane m + 2,
        // after the update: those waves start the plane's critical chain (next
        // tuple, face above, update) right after the barrier, and the halo
        // tuple is read only after the next one
        if (hslot >= 0) {
          double Xh[NF];
          Tup Th;
          ph.state(gp, dt, Xh);
          make_tup(Xh, Th);
          tup_store<PS>(Ln, hslot, Th);
          if (m + 2 < m1) ph.issue(src, hcol + (uint32_t)(m + 2) * sm);
        }
        __syncthreads();
        buf ^= 1;
      }
      bad_c = tup_negative(Tc);
    }
  }
  for (; m < m1; ++m) {
    const uint32_t idx = col + (uint32_t)m * sm;
    const double *L = lds + buf * (NTUP * PS);
    const uint8_t *LF = lfl + (MASK ? buf * PL : 0);
    double *Ln = lds + (buf ^ 1) * (NTUP * PS);
    uint8_t *LFn = lfl + (MASK ? (buf ^ 1) * PL : 0);
    const bool in_chunk = m + 1 < m1;                // plane m + 1 is this chunk's
    const bool has_next = in_chunk || have_beyond;  // a cell above this one
    if (STAGE >= 2 && !RB && valid) {
#pragma unroll
      for (int k = 0; k < NF; ++k) Ucur[k] = Uo.at(k, idx);
    }

    // ---- plane m + 1: tuples into the other LDS buffer, whose last readers
    // (plane m - 1's faces) finished before the barrier that ended plane m - 1;
    // built while this plane's faces are evaluated ----
    Tup Tn;
    uint8_t fn = CELL_FLUID | CELL_SOLVED;
    bool bad_n = false;
    double Unext[NF];
    if (valid && has_next) {
      double X[NF];
      pf.state(gp, dt, X);
      if (RB) {
#pragma unroll
        for (int k = 0; k < NF; ++k) Unext[k] = pf.a[k];
      }
      make_tup(X, Tn);
      bad_n = tup_negative(Tn);
      fn = ff;
      if (in_chunk) {
        tup_store<PS>(Ln, myslot, Tn);
        if (MASK) LFn[myslot] = fn;
      }
    }
    if (hslot >= 0 && in_chunk) {
      double Xh[NF];
      Tup Th;
      ph.state(gp, dt, Xh);
      make_tup(Xh, Th);
      tup_store<PS>(Ln, hslot, Th);
      if (MASK) LFn[hslot] = fh;
    }
    // the loads of plane m + 2 (a whole plane of work to arrive)
    if (valid && (m + 2 < m1 || (m + 2 == m1 && have_beyond))) {
      const uint32_t i2 = col + ((m + 2 < m1) ? (uint32_t)(m + 2) : beyond) * sm;
      pf.issue(src, i2);
      ff = cell_flags<MASK>(g, i2);
    }
    if (hslot >= 0 && m + 2 < m1) {
      ph.issue(src, hcol + (uint32_t)(m + 2) * sm);
      if (MASK) fh = g.flags[hcol + (uint32_t)(m + 2) * sm];
    }

    if (valid) {
      // ---- in-plane faces of this cell, reference order ----
      double acc[NF] = {0., 0., 0., 0., 0.};  // euler.cpp:142-148
      if (fc & CELL_SOLVED) {
#pragma unroll
        for (int a = 0; a < DIM - 1; ++a) {
          const bool hm = (a == 0) ? has_xm : has_ym;
          const bool hp = (a == 0) ? has_xp : has_yp;
          const int sm_ = (a == 0) ? myslot - 1 : myslot - RW;
          const int sp_ = (a == 0) ? myslot + 1 : myslot + RW;
          const bool wrap = (a == 0) ? wrap_x : wrap_y;
          double AFm[NF], AFp[NF];
          {
            const uint8_t fm = MASK && hm ? LF[sm_] : (uint8_t)(CELL_FLUID | CELL_SOLVED);
            const bool wall = MASK && hm && ((fm ^ fc) & CELL_FLUID);
            double l;
            if (hm && !wall) {
              Tup Nb;
              tup_load<PS>(L, sm_, Nb);
              l = face_flux_area(Nb, Tc, a, pow2A, A, halfA, AFm);
              // counted here only when the neighbour does not evaluate this face
              if (!(fm & CELL_SOLVED) && (bad_c | tup_negative(Nb))) neg.add(Nb, Tc);
            } else {
              l = virtual_face<REFL>(gp, bct, wall, 2 * a, a, Tc, false, AFm, neg, true);
            }
            lam_max = lam_acc(lam_max, l);
          }
          {
            const uint8_t fp = MASK && hp ? LF[sp_] : (uint8_t)(CELL_FLUID | CELL_SOLVED);
            const bool wall = MASK && hp && ((fp ^ fc) & CELL_FLUID);
            double l;
            if (hp && !wall) {
              Tup Nb;
              tup_load<PS>(L, sp_, Nb);
              l = face_flux_area(Tc, Nb, a, pow2A, A, halfA, AFp);
              if (bad_c | tup_negative(Nb)) neg.add(Tc, Nb);
            } else {
              l = virtual_face<REFL>(gp, bct, wall, 2 * a + 1, a, Tc, true, AFp, neg, true);
            }
            lam_max = lam_acc(lam_max, l);
          }
          if (wrap) {
#pragma unroll
            for (int k = 0; k < NF; ++k) { acc[k] -= AFp[k]; acc[k] += AFm[k]; }
          } else {
#pragma unroll
            for (int k = 0; k < NF; ++k) { acc[k] += AFm[k]; acc[k] -= AFp[k]; }
          }
        }
      }

      // ---- the face above this cell (each march-axis face once) ----
      double AFhi[NF] = {0., 0., 0., 0., 0.};
      if (has_next) {
        if ((fc | fn) & CELL_SOLVED) {
          // the chunk's top face and the periodic wrap face are counted by
          // the block that holds the plane above
          double l;
          if (MASK && ((fc ^ fn) & CELL_FLUID)) {
            if (fn & CELL_FLUID) l = virtual_face<REFL>(gp, bct, true, 2 * M, M, Tn, false, AFhi, neg, in_chunk);
            else l = virtual_face<REFL>(gp, bct, true, 2 * M + 1, M, Tc, true, AFhi, neg, in_chunk);
          } else {
            l = face_flux_area(Tc, Tn, M, pow2A, A, halfA, AFhi);
            if (in_chunk && (bad_c | bad_n)) neg.add(Tc, Tn);
          }
          lam_max = lam_acc(lam_max, l);
        }
      } else if (fc & CELL_SOLVED) {  // upper border of the march axis
        const double l = virtual_face<REFL>(gp, bct, false, 2 * M + 1, M, Tc, true, AFhi, neg, true);
        lam_max = lam_acc(lam_max, l);
      }

      // ---- this cell is complete: its RK stage update ----
      if (fc & CELL_SOLVED) {
        // the periodic top cell: its upper face (the wrap face) precedes its
        // lower face in the reference's face list
        const bool wrap_top = top_wrap_here && m == nm - 1;
        double nu[NF];
#pragma unroll
        for (int k = 0; k < NF; ++k) {
          double rhs = wrap_top ? (acc[k] - AFhi[k]) + AFlo[k] : (acc[k] + AFlo[k]) - AFhi[k];
          if (P2) rhs = rhs_scale * rhs;
          nu[k] = w1 ? Tc.c[k] + div_vol(gp, dt * rhs) : rk_update<STAGE>(gp, Ucur[k], Tc.c[k], rhs, dt);
          outf.put(k, idx, nu[k]);
        }
        if (EARLY && STAGE == 3) {
          lam_nx = lam_acc(lam_nx, cell_max_eig<DIM>(nu));
          if (MASK || REFL) lam_nx = lam_acc(lam_nx, virtual_max_eig(gp, nu, LF, m, fb, fn, has_next));
        }
      } else if (STAGE != 3) {
#pragma unroll
        for (int k = 0; k < NF; ++k)
          outf.put(k, idx, (STAGE == 1) ? (w1 ? Tc.c[k] + div_vol(gp, dt * 0.0) : 0.0) : Ucur[k]);
      }

      // the next cell's lower face, tuple and U
#pragma unroll
      for (int k = 0; k < NF; ++k) AFlo[k] = AFhi[k];
      Tc = Tn;
      fb = fc;
      fc = fn;
      bad_c = bad_n;
      if (RB) {
#pragma unroll
        for (int k = 0; k < NF; ++k) Ucur[k] = Unext[k];
      }
    }
    __syncthreads();
    buf ^= 1;
  }
}

// TX x TY tiles: 64 x 4 / 64 x 8 (one wave per row) and 32 x 8 (four waves,
// 60 KB of LDS: two independent blocks per CU, each SIMD holding a wave of
// each, so one block's plane barrier and division chains overlap the other's
// work) or 32 x 16 (eight waves, a halo ring of 19% of the tile's cells
// instead of 28% for 64 x 8).
template <int DIM, int TX, int TY, int STAGE, bool MASK, bool P2, bool REFL, bool SEGS, bool EARLY>
__global__ __launch_bounds__(TX *TY, 2) void k_march(GridDev g, StageIn<STAGE> in, const double *U, double *out,
                                                     Ctrl *ctrl, int mode, int chunk, int ntx, int nty) {
  constexpr int PL = (TY + 2) * (TX + 2);
  constexpr int PS = PL + 8;  // march_block's field stride
  static_assert(2 * NTUP * PS * 8 <= 120 * 1024, "two tuple planes must fit in LDS");
  __shared__ double lds[2 * NTUP * PS];
  __shared__ uint8_t lfl[MASK ? 2 * PL : 1];
  __shared__ double bct[6 * NTUP];
  constexpr int NT = TX * TY;
  if (mode == MODE_STEP && !(ctrl->t < ctrl->tmax)) return;  // main.cpp:377 loop guard
  const int tid = threadIdx.y * TX + threadIdx.x;
  const int nm = (int)g.n[DIM - 1];
  double lam_max = 0.0;
  double lam_nx = 0.0;  // early dt, stage 3: the next step's maximum over this thread's cells
  NegCount neg;         // non-positive states met, summed into the control block at the end
  bool any = false;     // this block marched a segment (and built bct)
  if constexpr (SEGS) {
    // the segment table: up to MARCH_SEGS (tile | masked, first plane, end
    // plane) per block, unused ones -1 at the end.  One call site per
    // variant: the march is inlined once masked and once not.
#pragma unroll 1
    for (int k = 0; k < MARCH_SEGS; ++k) {
      const int32_t *t = g.march_tiles + 3 * (MARCH_SEGS * (int64_t)blockIdx.x + k);
      if (t[0] < 0) break;
      any = true;
      const int tile = t[0] & ~MARCH_TILE_MASKED;
      const int bx = tile % ntx, by = tile / ntx;
      if (MASK && (t[0] & MARCH_TILE_MASKED))
        march_block<DIM, TX, TY, STAGE, MASK, P2, REFL, EARLY>(g, in, U, out, ctrl, mode, bx, by, t[1], t[2], lds,
                                                               lfl, bct, lam_max, lam_nx, neg);
      else
        march_block<DIM, TX, TY, STAGE, false, P2, REFL, EARLY>(g, in, U, out, ctrl, mode, bx, by, t[1], t[2], lds,
                                                                lfl, bct, lam_max, lam_nx, neg);
    }
  } else {
    // one (tile, chunk) per block, from the tile table (masked tiles first; a
    // tile whose cells, halo ring and neighbour planes are all fluid and
    // solved runs the unmasked code) or the XCD-aware numbering
    any = true;
    const int t = (MASK && g.march_tiles != nullptr) ? g.march_tiles[blockIdx.x]
                                                     : march_tile_of_block((int)blockIdx.x, (int)gridDim.x);
    const int tt = t & ~MARCH_TILE_MASKED;
    const int a = tt / (ntx * nty) * chunk;
    const int b = (a + chunk < nm) ? a + chunk : nm;
    const int bx = tt % ntx, by = (tt / ntx) % nty;
    if (MASK && (g.march_tiles == nullptr || (t & MARCH_TILE_MASKED)))
      march_block<DIM, TX, TY, STAGE, MASK, P2, REFL, EARLY>(g, in, U, out, ctrl, mode, bx, by, a, b, lds, lfl, bct,
                                                             lam_max, lam_nx, neg);
    else
      march_block<DIM, TX, TY, STAGE, false, P2, REFL, EARLY>(g, in, U, out, ctrl, mode, bx, by, a, b, lds, lfl, bct,
                                                              lam_max, lam_nx, neg);
  }
  neg.flush(ctrl, tid);
  block_max_to<NT>(ctrl, (mode == MODE_RHS_ONLY) ? 3 : STAGE - 1, lam_max, tid);
  if (EARLY && STAGE == 3 && mode == MODE_STEP) {
    // a Dirichlet border's virtual state enters the next maximum too (on a
    // grid with flags the border cells add it, where they are solved)
    if (tid == 0 && any && g.flags == nullptr) {
#pragma unroll
      for (int side = 0; side < 2 * DIM; ++side)
        if (g.bc[side] == BC_DIRICHLET && !g.periodic[side / 2]) {
          const double l = fabs(bct[side * NTUP + 5 + side / 2]) + bct[side * NTUP + 9];
          lam_nx = lam_acc(lam_nx, l);
        }
    }
    block_lam_to<NT>(ctrl, lam_nx, tid);
  }
  // stage 3 of a step: the last block advances the time (no separate launch)
  if (STAGE == 3 && mode == MODE_STEP) close_step<EARLY>(ctrl, tid);
}

template <int DIM, int TX, int TY, int STAGE, bool MASK, bool P2, bool REFL, bool EARLY>
static hipError_t launch_march_t(const GridDev &g, const MarchShape &sh, int mode, const double *U, const double *R,
                                 const double *W, double *oR, double *oW, double *oU, Ctrl *ctrl, hipStream_t s) {
  StageIn<STAGE> in;
  in.fs = g.fs;
  in.dt = 0.0;
  in.U = U;
  in.R = R;
  // with early dt stage 2 reads the W1 that stage 1 wrote into R's buffer
  in.X = (STAGE == 3) ? W : (STAGE == 2 && EARLY) ? R : U;
  double *out = (STAGE == 1) ? oR : (STAGE == 2) ? oW : oU;
  const int M = DIM - 1;
  const int ntx = (int)((g.n[0] + TX - 1) / TX);
  const int nty = (DIM == 3) ? (int)((g.n[1] + TY - 1) / TY) : 1;
  dim3 block(TX, TY);
  // the segment table when one was built for this shape (64 x 8 tiles), else
  // one (tile, chunk) per block
  if constexpr (DIM == 3 && TX == 64 && TY == 8) {
    if (sh.seg_blocks > 0 && g.march_tiles != nullptr) {
      hipLaunchKernelGGL((k_march<DIM, TX, TY, STAGE, MASK, P2, REFL, true, EARLY>), dim3((unsigned)sh.seg_blocks),
                         block, 0, s, g, in, U, out, ctrl, mode, 0, ntx, nty);
      return hipGetLastError();
    }
  }
  const int nch = (int)((g.n[M] + sh.chunk - 1) / sh.chunk);
  hipLaunchKernelGGL((k_march<DIM, TX, TY, STAGE, MASK, P2, REFL, false, EARLY>), dim3((unsigned)(ntx * nty * nch)),
                     block, 0, s, g, in, U, out, ctrl, mode, sh.chunk, ntx, nty);
  return hipGetLastError();
}

template <int DIM, int TX, int TY, bool MASK, bool P2, bool REFL, bool EARLY>
static hipError_t launch_march_p(const GridDev &g, const MarchShape &sh, int stage, int mode, const double *U,
                                 const double *R, const double *W, double *oR, double *oW, double *oU, Ctrl *ctrl,
                                 hipStream_t s) {
  switch (stage) {
    case 1: return launch_march_t<DIM, TX, TY, 1, MASK, P2, REFL, EARLY>(g, sh, mode, U, R, W, oR, oW, oU, ctrl, s);
    case 2: return launch_march_t<DIM, TX, TY, 2, MASK, P2, REFL, EARLY>(g, sh, mode, U, R, W, oR, oW, oU, ctrl, s);
    default: return launch_march_t<DIM, TX, TY, 3, MASK, P2, REFL, EARLY>(g, sh, mode, U, R, W, oR, oW, oU, ctrl, s);
  }
}

// Instances: without flags, early dt always (REFL: a reflecting or wall
// border); with flags (MASK, which implies REFL), early dt where every fluid
// cell is solved, else stage 1 writes R and stage 2 rebuilds W1.
template <int DIM, int TX, int TY, bool MASK>
static hipError_t launch_march_s(const GridDev &g, const MarchShape &sh, int stage, int mode, const double *U,
                                 const double *R, const double *W, double *oR, double *oW, double *oU, Ctrl *ctrl,
                                 hipStream_t s) {
  const bool p2 = g.pow2A && g.pow2V;
  if constexpr (MASK) {
    if (march_early(g))
      return p2 ? launch_march_p<DIM, TX, TY, true, true, true, true>(g, sh, stage, mode, U, R, W, oR, oW, oU, ctrl, s)
                : launch_march_p<DIM, TX, TY, true, false, true, true>(g, sh, stage, mode, U, R, W, oR, oW, oU, ctrl, s);
    return p2 ? launch_march_p<DIM, TX, TY, true, true, true, false>(g, sh, stage, mode, U, R, W, oR, oW, oU, ctrl, s)
              : launch_march_p<DIM, TX, TY, true, false, true, false>(g, sh, stage, mode, U, R, W, oR, oW, oU, ctrl, s);
  } else {
    // a reflecting / wall border needs the reflected virtual state
    bool refl = false;
    for (int side = 0; side < 2 * DIM; ++side)
      refl = refl || ((g.bc[side] == BC_REFLECTING || g.bc[side] == BC_WALL) && !g.periodic[side / 2]);
    if (refl)
      return p2 ? launch_march_p<DIM, TX, TY, false, true, true, true>(g, sh, stage, mode, U, R, W, oR, oW, oU, ctrl, s)
                : launch_march_p<DIM, TX, TY, false, false, true, true>(g, sh, stage, mode, U, R, W, oR, oW, oU, ctrl, s);
    return p2 ? launch_march_p<DIM, TX, TY, false, true, false, true>(g, sh, stage, mode, U, R, W, oR, oW, oU, ctrl, s)
              : launch_march_p<DIM, TX, TY, false, false, false, true>(g, sh, stage, mode, U, R, W, oR, oW, oU, ctrl, s);
  }
}

int64_t march_blocks(const GridDev &g, const MarchShape &sh) {
  if (sh.seg_blocks > 0) return sh.seg_blocks;
  const int TX = sh.tx;
  const int64_t ntx = (g.n[0] + TX - 1) / TX, nty = (g.n[1] + sh.ty - 1) / sh.ty;
  return ntx * nty * ((g.n[2] + sh.chunk - 1) / sh.chunk);
}

void build_march_tiles(const GridDev &g, const MarchShape &sh, const uint8_t *flags, int32_t *tiles) {
  const int64_t nx = g.n[0], ny = g.n[1], nz = g.n[2];
  const int TX = sh.tx;
  const int64_t ntx = (nx + TX - 1) / TX, nty = (ny + sh.ty - 1) / sh.ty;
  const int nblk = (int)march_blocks(g, sh);
  const uint8_t all = CELL_FLUID | CELL_SOLVED;
  // the cells a block reads: its tile and halo ring over its chunk and the
  // planes below and above it, wrapped across periodic sides
  auto wrap = [](int64_t i, int64_t n, bool per, int64_t &out) {
    if (i >= 0 && i < n) { out = i; return true; }
    if (!per) return false;
    out = (i + n) % n;
    return true;
  };
  std::vector<uint8_t> masked((size_t)nblk, 0);
  for (int t = 0; t < nblk; ++t) {
    const int64_t bx = t % ntx, by = (t / ntx) % nty, bz = t / (ntx * nty);
    const int64_t x0 = bx * TX, y0 = by * sh.ty, m0 = bz * sh.chunk;
    const int64_t x1 = std::min(x0 + TX, nx), y1 = std::min(y0 + sh.ty, ny), m1 = std::min(m0 + sh.chunk, nz);
    bool m = false;
    for (int64_t z = m0 - 1; z <= m1 && !m; ++z) {
      int64_t zz;
      if (!wrap(z, nz, g.periodic[2], zz)) continue;
      for (int64_t y = y0 - 1; y <= y1 && !m; ++y) {
        int64_t yy;
        if (!wrap(y, ny, g.periodic[1], yy)) continue;
        const uint8_t *row = flags + (yy + ny * zz) * nx;
        for (int64_t x = x0 - 1; x <= x1; ++x) {
          int64_t xx;
          if (!wrap(x, nx, g.periodic[0], xx)) continue;
          if (row[xx] != all) { m = true; break; }
        }
      }
    }
    masked[(size_t)t] = m;
  }
  // masked tiles on the first blocks of their XCD (the blocks b with the same
  // b % 8): they start in the first round and take longer, so their CUs are
  // not the ones given a second block; every tile stays on the XCD of the
  // XCD-aware numbering, next to its neighbours' halos in that XCD's L2
  for (int x = 0; x < 8 && x < nblk; ++x) {
    int k = x;
    for (int pass = 0; pass < 2; ++pass)
      for (int b = x; b < nblk; b += 8) {
        const int t = march_tile_of_block(b, nblk);
        if ((masked[(size_t)t] != 0) == (pass == 0)) {
          tiles[k] = pass == 0 ? (t | MARCH_TILE_MASKED) : t;
          k += 8;
        }
      }
  }
}

// Per plane m of tile (bx, by): the tile's cells and halo ring in planes m -
// 1, m and m + 1 (wrapped across periodic sides) are not all fluid and solved,
// i.e. marching plane m needs the masked code.
static void plane_masks(const GridDev &g, int TX, int TY, int64_t bx, int64_t by, const uint8_t *flags,
                        std::vector<uint8_t> &pm) {
  const int64_t nx = g.n[0], ny = g.n[1], nz = g.n[2];
  pm.assign((size_t)nz, 0);
  if (!flags) return;
  const uint8_t all = CELL_FLUID | CELL_SOLVED;
  auto wrap = [](int64_t i, int64_t n, bool per, int64_t &out) {
    if (i >= 0 && i < n) { out = i; return true; }
    if (!per) return false;
    out = (i + n) % n;
    return true;
  };
  const int64_t x0 = bx * TX, y0 = by * TY;
  const int64_t x1 = std::min(x0 + TX, nx), y1 = std::min(y0 + TY, ny);
  std::vector<uint8_t> ring((size_t)nz, 0);  // plane z's tile and ring
  for (int64_t z = 0; z < nz; ++z) {
    bool m = false;
    for (int64_t y = y0 - 1; y <= y1 && !m; ++y) {
      int64_t yy;
      if (!wrap(y, ny, g.periodic[1], yy)) continue;
      const uint8_t *row = flags + (yy + ny * z) * nx;
      for (int64_t x = x0 - 1; x <= x1; ++x) {
        int64_t xx;
        if (!wrap(x, nx, g.periodic[0], xx)) continue;
        if (row[xx] != all) { m = true; break; }
      }
    }
    ring[(size_t)z] = m;
  }
  for (int64_t m = 0; m < nz; ++m)
    for (int64_t z = m - 1; z <= m + 1; ++z) {
      int64_t zz;
      if (wrap(z, nz, g.periodic[2], zz) && ring[(size_t)zz]) pm[(size_t)m] = 1;
    }
}

// Cost of a plane marched by the unmasked and by the masked code (flags,
// body faces), in quarters of the former: on a 192^3 grid with a sphere the
// stage time is lowest with masked planes at 2.5 to 3.5 times an unmasked one.
constexpr int64_t PLANE_COST = 4, PLANE_COST_MASKED = 10;

bool build_march_segments(const GridDev &g, const MarchShape &sh, const uint8_t *flags, int32_t *table) {
  const int64_t nm = g.n[2];
  const int64_t ntx = (g.n[0] + sh.tx - 1) / sh.tx, nty = (g.n[1] + sh.ty - 1) / sh.ty;
  const int64_t W = ntx * nty * nm;  // tile planes, tile by tile
  const int B = sh.seg_blocks;
  // masked planes and the cost prefix over the tile planes
  std::vector<uint8_t> masked((size_t)W), pm;
  std::vector<int64_t> pre((size_t)W + 1, 0);
  for (int64_t t = 0; t < ntx * nty; ++t) {
    plane_masks(g, sh.tx, sh.ty, t % ntx, t / ntx, flags, pm);
    for (int64_t m = 0; m < nm; ++m) masked[(size_t)(t * nm + m)] = pm[(size_t)m];
  }
  for (int64_t i = 0; i < W; ++i) pre[(size_t)i + 1] = pre[(size_t)i] + (masked[(size_t)i] ? PLANE_COST_MASKED : PLANE_COST);
  const int64_t total = pre[(size_t)W];
  // first tile plane of block share L: the first whose cost prefix reaches L / B of the total
  auto first = [&](int64_t L) {
    const int64_t goal = (L * total + B - 1) / B;  // ceil: shares of equal cost, boundaries exact at L = B
    return (int64_t)(std::lower_bound(pre.begin(), pre.end(), goal) - pre.begin());
  };
  struct Seg { int64_t tile, m0, m1; bool masked; };
  std::vector<Seg> segs;
  for (int b = 0; b < B; ++b) {
    // XCD-aware: the blocks of one XCD (b % 8) take a contiguous range of tiles
    const int64_t L = march_tile_of_block(b, B);
    const int64_t start = std::min(first(L), W), end = std::min(first(L + 1), W);
    // runs of one tile and one code (masked or not)
    segs.clear();
    for (int64_t i = start; i < end; ++i) {
      const int64_t tile = i / nm, m = i % nm;
      const bool mk = masked[(size_t)i];
      if (!segs.empty() && segs.back().tile == tile && segs.back().masked == mk) segs.back().m1 = m + 1;
      else segs.push_back({tile, m, m + 1, mk});
    }
    // too many: merge the shortest neighbour pair of one tile (masked if either is)
    while ((int)segs.size() > MARCH_SEGS) {
      int best = -1;
      int64_t len = 0;
      for (size_t k = 0; k + 1 < segs.size(); ++k)
        if (segs[k].tile == segs[k + 1].tile && (best < 0 || segs[k + 1].m1 - segs[k].m0 < len)) {
          best = (int)k;
          len = segs[k + 1].m1 - segs[k].m0;
        }
      if (best < 0) return false;  // more tiles than segments
      segs[(size_t)best].m1 = segs[(size_t)best + 1].m1;
      segs[(size_t)best].masked = segs[(size_t)best].masked || segs[(size_t)best + 1].masked;
      segs.erase(segs.begin() + best + 1);
    }
    int32_t *t = table + 3 * MARCH_SEGS * (int64_t)b;
    for (int k = 0; k < 3 * MARCH_SEGS; ++k) t[k] = -1;
    for (size_t k = 0; k < segs.size(); ++k) {
      t[3 * k] = (int32_t)segs[k].tile | (segs[k].masked ? MARCH_TILE_MASKED : 0);
      t[3 * k + 1] = (int32_t)segs[k].m0;
      t[3 * k + 2] = (int32_t)segs[k].m1;
    }
  }
  return true;
}

void balance_march_shape(const GridDev &g, int num_cus, MarchShape &sh) {
  sh.seg_blocks = 0;
  if (g.dim != 3 || g.order != 1 || num_cus <= 0 || sh.tx != 64 || sh.ty != 8) return;
  const int64_t nb = march_blocks(g, sh);
  if (nb % num_cus == 0) return;  // the chunks make whole rounds
  const int64_t ntiles = ((g.n[0] + sh.tx - 1) / sh.tx) * ((g.n[1] + sh.ty - 1) / sh.ty);
  // a block's share within about two tiles, and at least 8 planes
  if (ntiles > num_cus || ntiles * g.n[2] < 8 * (int64_t)num_cus) return;
  sh.seg_blocks = num_cus;
}

bool march_shape_supported(const GridDev &g, const MarchShape &sh) {
  if (sh.chunk < 1 || g.N >= MARCH_MAX_CELLS || g.bslot) return false;  // per-face borders: the gather kernel
  // order 2: the 2D strip kernel (grids without bodies) and the 3D
  // LDS-staged kernel (64 x 8 tiles)
  if (g.order == 2 && g.dim == 3) return sh.tx == 64 && sh.ty == 8;
  if (g.order != 1 && !(g.order == 2 && g.dim == 2 && g.flags == nullptr)) return false;
  if (g.dim == 2) return sh.ty == 1;
  return (sh.tx == 64 && (sh.ty == 4 || sh.ty == 8)) || (sh.tx == 32 && (sh.ty == 8 || sh.ty == 16));
}

bool march_early(const GridDev &g) {
  return g.dim == 3 && g.order == 1 && (g.flags == nullptr || g.ghost_free);
}

// Longest chunk (fewest boundary planes re-read) that still gives every CU a
// block: one 64x8-tile block per CU is resident at a time (LDS and VGPRs).
MarchShape default_march_shape(const GridDev &g, int num_cus) {
  MarchShape sh;
  sh.tx = TX2D;
  sh.ty = (g.dim == 2) ? 1 : 8;
  const int M = g.dim - 1;
  // 2D strips are one wave, 62 interior cells wide; several fit on a CU
  const int64_t ntx = (g.dim == 2) ? (g.n[0] + 61) / 62 : (g.n[0] + sh.tx - 1) / sh.tx;
  const int64_t nty = (g.dim == 3) ? (g.n[1] + sh.ty - 1) / sh.ty : 1;
  const int64_t want = ((num_cus > 0) ? num_cus : 256) * ((g.dim == 2) ? 4 : 1);
  sh.chunk = 128;
  while (sh.chunk > 8 && ntx * nty * ((g.n[M] + sh.chunk - 1) / sh.chunk) < want) sh.chunk /= 2;
  return sh;
}

hipError_t launch_stage_march(const GridDev &g, const MarchShape &sh, int stage, int mode, const double *U,
                              const double *R, const double *W, double *oR, double *oW, double *oU, Ctrl *ctrl,
                              hipStream_t s) {
  const bool mask = g.flags != nullptr;
  if (g.order == 2 && g.dim == 3) return launch_stage_march3d_o2(g, sh.chunk, stage, mode, U, R, W, oR, oW, oU, ctrl, s);
  if (g.dim == 2) {
    if (g.order == 2) return launch_stage_march2d_o2(g, sh.chunk, stage, mode, U, R, W, oR, oW, oU, ctrl, s);
    return launch_stage_march2d(g, sh.chunk, stage, mode, U, R, W, oR, oW, oU, ctrl, s);
  }
  if (sh.tx == 32) {
    if (sh.ty == 16)
      return mask ? launch_march_s<3, 32, 16, true>(g, sh, stage, mode, U, R, W, oR, oW, oU, ctrl, s)
                  : launch_march_s<3, 32, 16, false>(g, sh, stage, mode, U, R, W, oR, oW, oU, ctrl, s);
    return mask ? launch_march_s<3, 32, 8, true>(g, sh, stage, mode, U, R, W, oR, oW, oU, ctrl, s)
                : launch_march_s<3, 32, 8, false>(g, sh, stage, mode, U, R, W, oR, oW, oU, ctrl, s);
  }
  switch (sh.ty) {
    case 4:
      return mask ? launch_march_s<3, 64, 4, true>(g, sh, stage, mode, U, R, W, oR, oW, oU, ctrl, s)
                  : launch_march_s<3, 64, 4, false>(g, sh, stage, mode, U, R, W, oR, oW, oU, ctrl, s);
    default:
      return mask ? launch_march_s<3, 64, 8, true>(g, sh, stage, mode, U, R, W, oR, oW, oU, ctrl, s)
                  : launch_march_s<3, 64, 8, false>(g, sh, stage, mode, U, R, W, oR, oW, oU, ctrl, s);
  }
}

}  // namespace mmf
