// mmf_march_o2.hip -- 3D order-2 marching stage kernel: the limited linear
// reconstruction (reconstruction.cpp:67-97 extension) from LDS-staged planes.
//
// A block of 64 x 8 threads (one wave per row) owns a 64 x 8 tile of a 3D
// grid and marches along z over a chunk of planes.  A face value needs the
// slope of its cell, and the slope the next cell, so every face of a cell
// reaches two cells out: each plane is staged in LDS once with a two-cell
// ring (68 x 12 cells).  Four plane buffers rotate: pass m reads planes m,
// m + 1 and m + 2 while plane m + 3, whose loads were issued a whole pass
// earlier, goes into the buffer plane m - 1 left.  One pass per plane, one
// barrier:
//   - the x and y faces of the cell in plane m, both sides reconstructed
//     from the LDS ring exactly as the gather kernel does (the slopes of the
//     cell and of its neighbours, the per-face-side positivity fallback to
//     the mean, border and wall states);
//   - the z face between planes m and m + 1 from the column's cells in
//     planes m - 1 .. m + 2 (each z face once; the face below is the
//     previous pass's);
//   - the cell's RK stage update.
// Residuals are summed in the reference's face order (x, y, z; lower, upper),
// so the result is bit-identical to the gather kernel and to the oracle.
#include <hip/hip_runtime.h>

#include "mmf_device.hpp"
#include "mmf_internal.hpp"
#include "mmf_stage.hpp"

namespace mmf {

namespace {

constexpr int OX = 64, OY = 8;           // tile
constexpr int RX = OX + 4, RY = OY + 4;  // ring plane
constexpr int RP = RX * RY;              // 816 cells
constexpr int NBUF = 4;                  // plane buffers
constexpr int NT = OX * OY;              // 512 threads
constexpr int PER = 2;                   // ring cells per thread (816 = 512 + 304)
constexpr uint8_t PRESENT = 0x80;        // flag bit: the ring cell exists (in the grid or wrapped)

// reconstruction.cpp:67-97 extension: mean +/- slope / 2, the mean where the
// value has a non-positive density or pressure; the tuple from the primitives
// of that check
__device__ __forceinline__ void recon_tup(const double *mean, const double *s, bool upper, double *v, Tup &T) {
#pragma unroll
  for (int k = 0; k < NF; ++k) v[k] = upper ? mean[k] + 0.5 * s[k] : mean[k] - 0.5 * s[k];
  double pr[NF];
  cons2prim(v, pr);
  if (!(v[FID_RHO] > 0.0) || !(pr[FID_P] > 0.0)) {
#pragma unroll
    for (int k = 0; k < NF; ++k) v[k] = mean[k];
    cons2prim(v, pr);
  }
  make_tup_prim(v, pr, T);
}

// The same with the mean read again (by again(v)) for the fallback, so that
// it need not stay in registers across the positivity check.
template <class Again>
__device__ __forceinline__ void recon_tup_r(const double *mean, const double *s, bool upper, double *v, Tup &T,
                                            Again again) {
#pragma unroll
  for (int k = 0; k < NF; ++k) v[k] = upper ? mean[k] + 0.5 * s[k] : mean[k] - 0.5 * s[k];
  double pr[NF];
  cons2prim(v, pr);
  if (!(v[FID_RHO] > 0.0) || !(pr[FID_P] > 0.0)) {
    again(v);
    cons2prim(v, pr);
  }
  make_tup_prim(v, pr, T);
}

// the limited slope of a cell from its lower and upper differences
// (minmod(upper, lower), a missing or solid neighbour's difference 0)
__device__ __forceinline__ void slope(const double *xm, const double *x0, const double *xp, bool okm, bool okp,
                                      double *s) {
#pragma unroll
  for (int k = 0; k < NF; ++k) {
    const double dm = okm ? x0[k] - xm[k] : 0.0;
    const double dp = okp ? xp[k] - x0[k] : 0.0;
    s[k] = minmod(dp, dm);
  }
}

// a compiler-only memory fence: LDS values are read again after it instead of
// being kept in registers across it
__device__ __forceinline__ void ld_fence() { asm volatile("" ::: "memory"); }

constexpr int NTUP = 11;  // c[5], vel[3], p, a, H

// The tuple of each Dirichlet border's virtual state (euler.cpp:366-376), built
// once per block into LDS: the loop then reads it instead of carrying six
// uniform tuples in registers.
__device__ __forceinline__ void dirichlet_tuples_o2(const GridDev &g, double *bct, int tid) {
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

// The tuple of the virtual state beyond border side 'side' (axis a) of the
// fluid face value v (tuple Tv): euler.cpp:261-376 with n = +-e_a.
__device__ __forceinline__ void border_tup(const GridDev &g, const double *bct, int side, int a, const double *v,
                                           const Tup &Tv, Tup &V) {
  const int bc = g.bc[side];
  if (bc == BC_DIRICHLET) {
    const double *o = bct + side * NTUP;
#pragma unroll
    for (int k = 0; k < NF; ++k) V.c[k] = o[k];
    V.vel[0] = o[5];
    V.vel[1] = o[6];
    V.vel[2] = o[7];
    V.p = o[8];
    V.a = o[9];
    V.H = o[10];
  } else if (bc == BC_REFLECTING || bc == BC_WALL) {
    double w[NF];
    reflect_axis(v, a, w);
    make_tup(w, V);
  } else {
    V = Tv;  // a copy: the same state, the same tuple
  }
}

template <int STAGE>
struct RingPf {
  double a[PER][NF];  // X, or U (stage 2)
  double b[PER][NF];  // R (stage 2)
};

}  // namespace

template <int STAGE, bool MASK, bool P2>
__global__ __launch_bounds__(NT, 1) void k_march_o2(GridDev g, StageIn<STAGE> in, const double *U, double *out,
                                                     Ctrl *ctrl, int mode, int chunk, int ntx, int nty) {
  __shared__ double S[NBUF][NF][RP];
  __shared__ uint8_t FL[NBUF][RP];
  __shared__ double bct[6 * NTUP];
  const int slot_eig = (mode == MODE_RHS_ONLY) ? 3 : STAGE - 1;
  if (mode == MODE_STEP && !(ctrl->t < ctrl->tmax)) return;  // main.cpp:377 loop guard
  const double dt = (STAGE >= 2) ? step_dt(ctrl) : 0.0;
  in.dt = dt;  // the stage-2 input rebuilt with this step's dt (StageIn::load)
  GridDev gp = g;
  gp.pow2A = P2;
  gp.pow2V = P2;

  const int tile = march_tile_of_block((int)blockIdx.x, (int)gridDim.x);
  const int bx = tile % ntx, by = (tile / ntx) % nty, bz = tile / (ntx * nty);
  const int lane = threadIdx.x, row = threadIdx.y;
  const int tid = row * OX + lane;
  const int nx = (int)g.n[0], ny = (int)g.n[1], nz = (int)g.n[2];
  const int x0 = bx * OX, y0 = by * OY;
  const int m0 = bz * chunk;
  const int m1 = (m0 + chunk < nz) ? m0 + chunk : nz;
  const int gx = x0 + lane, gy = y0 + row;
  const bool valid = gx < nx && gy < ny;
  const int64_t fs = g.fs;
  const int64_t sm = (int64_t)nx * ny;
  const int64_t colidx = (int64_t)gx + (int64_t)nx * gy;
  const int cpos = (row + 2) * RX + lane + 2;  // this cell in a ring plane
  const bool pz = g.periodic[2] != 0;

  // ring cells of this thread: its own cell, then one of the 304 cells of
  // the ring proper (rows 0-1 and 10-11, columns 0-1 and 66-67 of rows
  // 2-9); ring index (RP: none) and grid column (-1: absent).  Each thread
  // writes its own cell's slot, so a pass may read its own cell of plane
  // m - 1 while plane m + 3 goes into that buffer.
  int rr[PER];
  int64_t rcol[PER];
  rr[0] = cpos;
  if (tid < 2 * RX) rr[1] = tid;
  else if (tid < 4 * RX) rr[1] = (RY - 2) * RX + (tid - 2 * RX);
  else if (tid < 4 * RX + 4 * OY) {
    const int u = tid - 4 * RX, c = u % 4;
    rr[1] = (2 + u / 4) * RX + ((c < 2) ? c : OX + c);
  } else rr[1] = RP;
#pragma unroll
  for (int i = 0; i < PER; ++i) {
    const int r = rr[i];
    rcol[i] = -1;
    if (r < RP) {
      int x = x0 - 2 + r % RX, y = y0 - 2 + r / RX;
      bool ok = true;
      if (x < 0 || x >= nx) {
        if (g.periodic[0]) x = ((x % nx) + nx) % nx;
        else ok = false;
      }
      if (y < 0 || y >= ny) {
        if (g.periodic[1]) y = ((y % ny) + ny) % ny;
        else ok = false;
      }
      if (ok) rcol[i] = (int64_t)x + (int64_t)nx * y;
    }
  }
  // plane p (any integer) present, and its index in the grid
  auto plane_ok = [&](int p) { return pz || (p >= 0 && p < nz); };
  auto plane_of = [&](int p) { return (p < 0 || p >= nz) ? ((p % nz) + nz) % nz : p; };

  RingPf<STAGE> pf;
  uint8_t pff[PER];
  auto issue = [&](int p) {  // the loads of ring plane p
    const bool okp = plane_ok(p);
    const int64_t base = (int64_t)plane_of(p) * sm;
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      pff[i] = 0;
      if (okp && rcol[i] >= 0) {
        const int64_t idx = rcol[i] + base;
#pragma unroll
        for (int k = 0; k < NF; ++k) {
          if (STAGE == 2) {
            pf.a[i][k] = in.U[k * fs + idx];
            pf.b[i][k] = in.R[k * fs + idx];
          } else {
            pf.a[i][k] = in.X[k * fs + idx];
          }
        }
        pff[i] = (uint8_t)(PRESENT | (MASK ? g.flags[idx] : (uint8_t)(CELL_FLUID | CELL_SOLVED)));
      }
    }
  };
  auto store = [&](int p) {  // the issued plane into its buffer
    const int buf = ((p % NBUF) + NBUF) % NBUF;
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      if (rr[i] < RP) {
        FL[buf][rr[i]] = pff[i];
        if (pff[i]) {
#pragma unroll
          for (int k = 0; k < NF; ++k)
            S[buf][k][rr[i]] = (STAGE == 2) ? pf.a[i][k] + div_vol(gp, dt * pf.b[i][k]) : pf.a[i][k];  // main.cpp:421
        }
      }
    }
  };
  auto ld = [&](int p, int pos, double *x) {
    const int buf = ((p % NBUF) + NBUF) % NBUF;
#pragma unroll
    for (int k = 0; k < NF; ++k) x[k] = S[buf][k][pos];
  };
  auto fl = [&](int p, int pos) { return FL[((p % NBUF) + NBUF) % NBUF][pos]; };

  const double A = g.A, halfA = g.halfA;
  double lam_max = 0.0;
  NegCount neg;

  dirichlet_tuples_o2(g, bct, tid);
  // ---- prologue: planes m0 .. m0 + 2 into LDS, the loads of m0 + 3 ----
  for (int p = m0; p < m0 + 3; ++p) {
    issue(p);
    store(p);
  }
  issue(m0 + 3);
  __syncthreads();

  // the column's cell below plane m0 (into its slot of plane m0 - 1's
  // buffer, which the passes read) and the face below plane m0
  double AFlo[NF] = {0., 0., 0., 0., 0.};
  if (valid) {
    double Xb[NF] = {0., 0., 0., 0., 0.};
    uint8_t fb = 0;  // its flags with PRESENT
    double Xbb[NF] = {0., 0., 0., 0., 0.};
    uint8_t fbb = 0;
    if (plane_ok(m0 - 1)) {
      const int64_t i1 = colidx + (int64_t)plane_of(m0 - 1) * sm;
      in.load(gp, i1, Xb);
      fb = (uint8_t)(PRESENT | cell_flags<MASK>(g, i1));
      if (plane_ok(m0 - 2)) {
        const int64_t i2 = colidx + (int64_t)plane_of(m0 - 2) * sm;
        in.load(gp, i2, Xbb);
        fbb = (uint8_t)(PRESENT | cell_flags<MASK>(g, i2));
      }
    }
    double Xc[NF], Xn[NF];
    ld(m0, cpos, Xc);
    ld(m0 + 1, cpos, Xn);
    const uint8_t fc = fl(m0, cpos), fn = fl(m0 + 1, cpos);
    const int64_t q[3] = {gx, gy, m0};
    double l = 0.0;
    bool have = false;
    if (!(fb & PRESENT)) {  // the grid's lower z border
      if (fc & CELL_SOLVED) {
        double sc[NF], Rm[NF];
        slope(Xb, Xc, Xn, false, (fn & PRESENT) && (fn & CELL_FLUID), sc);
        Tup TR, TL;
        recon_tup(Xc, sc, false, Rm, TR);
        border_tup(g, bct, 4, 2, Rm, TR, TL);
        l = face_flux_area(TL, TR, 2, P2, A, halfA, AFlo);
        neg.add(TL, TR);
        have = true;
      }
    } else if ((fb | fc) & CELL_SOLVED) {
      double sb[NF], sc[NF], Lv[NF], Rv[NF];
      Tup TL, TR;
      if (MASK && ((fb ^ fc) & CELL_FLUID)) {  // a wall: the fluid side's value mirrored
        if (fc & CELL_FLUID) {
          slope(Xb, Xc, Xn, false, (fn & PRESENT) && (fn & CELL_FLUID), sc);
          recon_tup(Xc, sc, false, Rv, TR);
          reflect_axis(Rv, 2, Lv);
          make_tup(Lv, TL);
        } else {
          slope(Xbb, Xb, Xc, (fbb & PRESENT) && (fbb & CELL_FLUID), false, sb);
          recon_tup(Xb, sb, true, Lv, TL);
          reflect_axis(Lv, 2, Rv);
          make_tup(Rv, TR);
        }
      } else {
        slope(Xbb, Xb, Xc, (fbb & PRESENT) && (fbb & CELL_FLUID), (fc & CELL_FLUID) != 0, sb);
        slope(Xb, Xc, Xn, (fb & CELL_FLUID) != 0, (fn & PRESENT) && (fn & CELL_FLUID), sc);
        recon_tup(Xb, sb, true, Lv, TL);
        recon_tup(Xc, sc, false, Rv, TR);
      }
      l = face_flux_area(TL, TR, 2, P2, A, halfA, AFlo);
      // counted here: the block below does not count its chunk's top face
      neg.add(TL, TR);
      have = true;
    }
    if (have) lam_max = lam_acc(lam_max, l);
    const int bb = ((m0 - 1) % NBUF + NBUF) % NBUF;
#pragma unroll
    for (int k = 0; k < NF; ++k) S[bb][k][cpos] = Xb[k];
    FL[bb][cpos] = fb;
  }

  for (int m = m0; m < m1; ++m) {
    const bool in_chunk = m + 1 < m1;
    const int64_t idx = colidx + (int64_t)m * sm;
    double Xc[NF];
    uint8_t fc = 0;
    double AFhi[NF] = {0., 0., 0., 0., 0.};
    double acc[NF] = {0., 0., 0., 0., 0.};  // euler.cpp:142-148
    if (valid) {
      ld(m, cpos, Xc);
      fc = fl(m, cpos);
      const int64_t q[3] = {gx, gy, m};
      if (fc & CELL_SOLVED) {
        // ---- x and y faces (the gather kernel's per-axis body, one face
        // at a time, neighbours read from the LDS ring as needed) ----
#pragma unroll
        for (int a = 0; a < 2; ++a) {
          const int st = (a == 0) ? 1 : RX;
          const uint8_t fm1 = fl(m, cpos - st), fp1 = fl(m, cpos + st);
          const bool em1 = fm1 & PRESENT, ep1 = fp1 & PRESENT;
          const bool wall_m = MASK && em1 && ((fm1 ^ fc) & CELL_FLUID);
          const bool wall_p = MASK && ep1 && ((fp1 ^ fc) & CELL_FLUID);
          const bool cfluid = fc & CELL_FLUID;
          double sc[NF];  // this cell's slope
          {
            double Xm1[NF], Xp1[NF];
            ld(m, cpos - st, Xm1);  // every ring slot is in range; an absent cell's difference is selected away
            ld(m, cpos + st, Xp1);
            slope(Xm1, Xc, Xp1, em1 && (fm1 & CELL_FLUID), ep1 && (fp1 & CELL_FLUID), sc);
          }
          double AFm[NF], AFp[NF];
          __builtin_amdgcn_sched_barrier(0);
          {  // the lower face
            double Rm[NF], Lm[NF];
            Tup TLm, TRm;
            recon_tup_r(Xc, sc, false, Rm, TRm, [&](double *v) { ld_fence(); ld(m, cpos, v); });
            if (!em1) {
              border_tup(g, bct, 2 * a, a, Rm, TRm, TLm);
            } else if (wall_m) {
              reflect_axis(Rm, a, Lm);
              make_tup(Lm, TLm);
            } else {
              const uint8_t fm2 = fl(m, cpos - 2 * st);
              const bool em2 = fm2 & PRESENT;
              double Xm1[NF], Xm2[NF], smv[NF];
              ld(m, cpos - st, Xm1);
              ld(m, cpos - 2 * st, Xm2);
              slope(Xm2, Xm1, Xc, em2 && (fm2 & CELL_FLUID), cfluid, smv);
              recon_tup_r(Xm1, smv, true, Lm, TLm, [&](double *v) { ld_fence(); ld(m, cpos - st, v); });
            }
            const double lm = face_flux_area(TLm, TRm, a, P2, A, halfA, AFm);
            lam_max = lam_acc(lam_max, lm);
            // positivity counts once per face (the gather kernel's rule)
            const bool other_solved_m = em1 && !wall_m && (fm1 & CELL_SOLVED);
            if (!other_solved_m) neg.add(TLm, TRm);
            if (a == 0) {
              // the first contribution (from zero: a wrap's order does not
              // change the bits of the first two)
#pragma unroll
              for (int k = 0; k < NF; ++k) acc[k] += AFm[k];
            }
          }
          __builtin_amdgcn_sched_barrier(0);
          {  // the upper face
            double Lp[NF], Rp[NF];
            Tup TLp, TRp;
            recon_tup_r(Xc, sc, true, Lp, TLp, [&](double *v) { ld_fence(); ld(m, cpos, v); });
            if (!ep1) {
              border_tup(g, bct, 2 * a + 1, a, Lp, TLp, TRp);
            } else if (wall_p) {
              reflect_axis(Lp, a, Rp);
              make_tup(Rp, TRp);
            } else {
              const uint8_t fp2 = fl(m, cpos + 2 * st);
              const bool ep2 = fp2 & PRESENT;
              double Xp1[NF], Xp2[NF], spv[NF];
              ld(m, cpos + st, Xp1);
              ld(m, cpos + 2 * st, Xp2);
              slope(Xc, Xp1, Xp2, cfluid, ep2 && (fp2 & CELL_FLUID), spv);
              recon_tup_r(Xp1, spv, false, Rp, TRp, [&](double *v) { ld_fence(); ld(m, cpos + st, v); });
            }
            const double lp = face_flux_area(TLp, TRp, a, P2, A, halfA, AFp);
            lam_max = lam_acc(lam_max, lp);
            neg.add(TLp, TRp);
          }
          __builtin_amdgcn_sched_barrier(0);  // keep the two faces' live ranges apart
          // periodic wrap: the wrap face is first in the reference-order face list
          const bool wrap_first = g.periodic[a] && (q[a] == g.n[a] - 1);
          if (a == 0) {
#pragma unroll
            for (int k = 0; k < NF; ++k) acc[k] -= AFp[k];
          } else if (wrap_first) {
#pragma unroll
            for (int k = 0; k < NF; ++k) { acc[k] -= AFp[k]; acc[k] += AFm[k]; }
          } else {
#pragma unroll
            for (int k = 0; k < NF; ++k) { acc[k] += AFm[k]; acc[k] -= AFp[k]; }
          }
          ld_fence();
          ld(m, cpos, Xc);  // re-read: nothing of the cell stays live across the sections
        }
      }

      __builtin_amdgcn_sched_barrier(0);
      // ---- the z face above this cell (each z face once) ----
      const uint8_t fn = fl(m + 1, cpos);
      const uint8_t fb = fl(m - 1, cpos);  // the cell below (its own slot)
      double Xb[NF] = {0., 0., 0., 0., 0.};
      ld(m - 1, cpos, Xb);
      if (!(fn & PRESENT)) {  // the grid's upper z border
        if (fc & CELL_SOLVED) {
          double Xn[NF] = {0., 0., 0., 0., 0.}, sc[NF], Lp[NF], Rp[NF];
          slope(Xb, Xc, Xn, (fb & PRESENT) && (fb & CELL_FLUID), false, sc);
          Tup TL, TR;
          recon_tup_r(Xc, sc, true, Lp, TL, [&](double *v) { ld_fence(); ld(m, cpos, v); });
          border_tup(g, bct, 5, 2, Lp, TL, TR);
          const double l = face_flux_area(TL, TR, 2, P2, A, halfA, AFhi);
          lam_max = lam_acc(lam_max, l);
          neg.add(TL, TR);
        }
      } else if ((fc | fn) & CELL_SOLVED) {
        Tup TL, TR;
        if (MASK && ((fc ^ fn) & CELL_FLUID)) {  // a wall: the fluid side's value mirrored
          double Xn[NF], s_[NF], Lv[NF], Rv[NF];
          ld(m + 1, cpos, Xn);
          if (fc & CELL_FLUID) {
            slope(Xb, Xc, Xn, (fb & PRESENT) && (fb & CELL_FLUID), false, s_);
            recon_tup_r(Xc, s_, true, Lv, TL, [&](double *v) { ld_fence(); ld(m, cpos, v); });
            reflect_axis(Lv, 2, Rv);
            make_tup(Rv, TR);
          } else {
            const uint8_t fnn = fl(m + 2, cpos);
            double Xnn[NF] = {0., 0., 0., 0., 0.};
            ld(m + 2, cpos, Xnn);
            slope(Xc, Xn, Xnn, false, (fnn & PRESENT) && (fnn & CELL_FLUID), s_);
            recon_tup_r(Xn, s_, false, Rv, TR, [&](double *v) { ld_fence(); ld(m + 1, cpos, v); });
            reflect_axis(Rv, 2, Lv);
            make_tup(Lv, TL);
          }
        } else {
          {  // this cell's upper value
            double Xn[NF], s_[NF], Lv[NF];
            ld(m + 1, cpos, Xn);
            slope(Xb, Xc, Xn, (fb & PRESENT) && (fb & CELL_FLUID), (fn & CELL_FLUID) != 0, s_);
            recon_tup_r(Xc, s_, true, Lv, TL, [&](double *v) { ld_fence(); ld(m, cpos, v); });
          }
          {  // the upper cell's lower value
            double Xn[NF], Xnn[NF] = {0., 0., 0., 0., 0.}, s_[NF], Rv[NF];
            const uint8_t fnn = fl(m + 2, cpos);
            ld(m + 1, cpos, Xn);
            ld(m + 2, cpos, Xnn);
            slope(Xc, Xn, Xnn, (fc & CELL_FLUID) != 0, (fnn & PRESENT) && (fnn & CELL_FLUID), s_);
            recon_tup_r(Xn, s_, false, Rv, TR, [&](double *v) { ld_fence(); ld(m + 1, cpos, v); });
          }
        }
        const double l = face_flux_area(TL, TR, 2, P2, A, halfA, AFhi);
        lam_max = lam_acc(lam_max, l);
        // the chunk's top face and the periodic wrap face: counted by the block above
        if (in_chunk) neg.add(TL, TR);
      }

      __builtin_amdgcn_sched_barrier(0);
      // ---- this cell is complete: its RK stage update ----
      if (fc & CELL_SOLVED) {
        // the periodic top cell: its upper face (the wrap face) precedes its lower face
        const bool wrap_top = pz && m == nz - 1;
#pragma unroll
        for (int k = 0; k < NF; ++k) {
          const double rhs = wrap_top ? (acc[k] - AFhi[k]) + AFlo[k] : (acc[k] + AFlo[k]) - AFhi[k];
          const double u = (STAGE == 1) ? 0.0 : U[k * fs + idx];
          out[k * fs + idx] = rk_update<STAGE>(gp, u, Xc[k], rhs, dt);
        }
      } else if (STAGE == 1) {
#pragma unroll
        for (int k = 0; k < NF; ++k) out[k * fs + idx] = 0.0;
      } else if (STAGE == 2) {
#pragma unroll
        for (int k = 0; k < NF; ++k) out[k * fs + idx] = U[k * fs + idx];
      }
    }
    // plane m + 3 into the buffer plane m - 1 leaves (each thread's own slot
    // of it read above by the thread itself)
#pragma unroll
    for (int k = 0; k < NF; ++k) AFlo[k] = AFhi[k];
    store(m + 3);
    if (m + 1 < m1) issue(m + 4);
    __syncthreads();
  }
  neg.flush(ctrl, tid);
  block_max_to<NT>(ctrl, slot_eig, lam_max, tid);
  // stage 3 of a step: the last block advances the time (no separate launch)
  if (STAGE == 3 && mode == MODE_STEP) close_step<false>(ctrl, tid);
}

template <int STAGE, bool MASK, bool P2>
static hipError_t launch_o2_t(const GridDev &g, int chunk, int mode, const double *U, const double *R,
                              const double *W, double *oR, double *oW, double *oU, Ctrl *ctrl, hipStream_t s) {
  StageIn<STAGE> in;
  in.fs = g.fs;
  in.dt = 0.0;
  in.U = U;
  in.R = R;
  in.X = (STAGE == 3) ? W : U;
  double *out = (STAGE == 1) ? oR : (STAGE == 2) ? oW : oU;
  const int ntx = (int)((g.n[0] + OX - 1) / OX);
  const int nty = (int)((g.n[1] + OY - 1) / OY);
  const int nch = (int)((g.n[2] + chunk - 1) / chunk);
  hipLaunchKernelGGL((k_march_o2<STAGE, MASK, P2>), dim3((unsigned)(ntx * nty * nch)), dim3(OX, OY), 0, s, g, in, U,
                     out, ctrl, mode, chunk, ntx, nty);
  return hipGetLastError();
}

template <bool MASK, bool P2>
static hipError_t launch_o2_s(const GridDev &g, int chunk, int stage, int mode, const double *U, const double *R,
                              const double *W, double *oR, double *oW, double *oU, Ctrl *ctrl, hipStream_t s) {
  switch (stage) {
    case 1: return launch_o2_t<1, MASK, P2>(g, chunk, mode, U, R, W, oR, oW, oU, ctrl, s);
    case 2: return launch_o2_t<2, MASK, P2>(g, chunk, mode, U, R, W, oR, oW, oU, ctrl, s);
    default: return launch_o2_t<3, MASK, P2>(g, chunk, mode, U, R, W, oR, oW, oU, ctrl, s);
  }
}

hipError_t launch_stage_march3d_o2(const GridDev &g, int chunk, int stage, int mode, const double *U, const double *R,
                                   const double *W, double *oR, double *oW, double *oU, Ctrl *ctrl, hipStream_t s) {
  const bool mask = g.flags != nullptr;
  const bool p2 = g.pow2A && g.pow2V;
  if (mask) {
    return p2 ? launch_o2_s<true, true>(g, chunk, stage, mode, U, R, W, oR, oW, oU, ctrl, s)
              : launch_o2_s<true, false>(g, chunk, stage, mode, U, R, W, oR, oW, oU, ctrl, s);
  }
  return p2 ? launch_o2_s<false, true>(g, chunk, stage, mode, U, R, W, oR, oW, oU, ctrl, s)
            : launch_o2_s<false, false>(g, chunk, stage, mode, U, R, W, oR, oW, oU, ctrl, s);
}

}  // namespace mmf
