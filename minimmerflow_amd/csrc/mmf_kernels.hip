// mmf_kernels.hip -- HIP kernels of the finite-volume step for gfx950.
//
// Stage kernel (k_stage): one thread per cell.  Each thread gathers the
// fluxes of its 2*dim faces in the order the reference's interface loop
// reaches them (all x faces, then y, then z; lower face before upper face),
// so the residual of every cell is accumulated in a fixed order without any
// float atomic: results are bitwise reproducible and independent of the
// launch shape.  The RK stage update (main.cpp:409-495) is fused:
//   stage 1: R  = RHS(U)                               (euler.cpp:127-249)
//   stage 2: W1 = U + dt R / V  (rebuilt on the fly per stencil cell)
//            W  = 0.75 U + 0.25 (W1 + dt RHS(W1) / V)
//   stage 3: U  = (1./3) U + (2./3) (W + dt RHS(W) / V)
// dt = 0.9 cfl h / maxEig(stage 1) is evaluated on the device from the
// stage-1 maximum, so no host round trip sits inside a step.
#include <hip/hip_runtime.h>

#include "mmf_device.hpp"
#include "mmf_internal.hpp"
#include "mmf_stage.hpp"

namespace mmf {

constexpr int BLOCK = 256;

// ---------------------------------------------------------------------------
// stage kernel
// ---------------------------------------------------------------------------
template <int DIM, int ORDER, int STAGE, bool MASK>
__global__ __launch_bounds__(BLOCK) void k_stage(GridDev g, StageIn<STAGE> in, const double *U, double *out,
                                                 Ctrl *ctrl, int mode) {
  const int slot = (mode == MODE_RHS_ONLY) ? 3 : STAGE - 1;
  if (mode == MODE_STEP && !(ctrl->t < ctrl->tmax)) return;  // main.cpp:377 loop guard
  if (STAGE >= 2) in.dt = step_dt(ctrl);

  const int64_t idx = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
  const bool live = idx < g.N;
  double lam_max = 0.0;

  if (live) {
    const int64_t q[3] = {idx % g.n[0], (idx / g.n[0]) % g.n[1], idx / (g.n[0] * g.n[1])};
    const uint8_t fc = cell_flags<MASK>(g, idx);
    const bool solved = fc & CELL_SOLVED;
    double Xc[NF];
    in.load(g, idx, Xc);

    if (solved) {
      double rhs[NF] = {0., 0., 0., 0., 0.};  // euler.cpp:142-148
      Tup Tc;
      if (ORDER == 1) make_tup(Xc, Tc);

#pragma unroll
      for (int a = 0; a < DIM; ++a) {
        // neighbours along axis a: -2, -1, +1, +2
        int64_t nm1 = 0, np1 = 0, nm2 = 0, np2 = 0;
        bool em1 = neighbour(g, q, a, -1, nm1);
        bool ep1 = neighbour(g, q, a, +1, np1);
        uint8_t fm1 = em1 ? cell_flags<MASK>(g, nm1) : 0;
        uint8_t fp1 = ep1 ? cell_flags<MASK>(g, np1) : 0;
        // BC_WALL where fluid flags differ (main.cpp:271-275); border where absent
        const bool wall_m = em1 && ((fm1 ^ fc) & CELL_FLUID);
        const bool wall_p = ep1 && ((fp1 ^ fc) & CELL_FLUID);

        double Lm[NF], Rm[NF], Lp[NF], Rp[NF];  // face states: lower face (L|R), upper face (L|R)
        Tup TLm2, TRm2, TLp2, TRp2;             // order 2: their tuples
        if (ORDER == 1) {
          // lower face: R = this cell
#pragma unroll
          for (int k = 0; k < NF; ++k) Rm[k] = Xc[k], Lp[k] = Xc[k];
          if (!em1) border_state(g, 2 * a, a, q, Xc, Lm);
          else if (wall_m) reflect_axis(Xc, a, Lm);
          else in.load(g, nm1, Lm);
          if (!ep1) border_state(g, 2 * a + 1, a, q, Xc, Rp);
          else if (wall_p) reflect_axis(Xc, a, Rp);
          else in.load(g, np1, Rp);
        } else {
          // limited linear: slopes of cells -1, 0, +1 along a
          double Xm1[NF], Xp1[NF], Xm2[NF], Xp2[NF];
          if (em1) in.load(g, nm1, Xm1);
          if (ep1) in.load(g, np1, Xp1);
          int64_t q1[3] = {q[0], q[1], q[2]};
          bool em2 = false, ep2 = false;
          uint8_t fm2 = 0, fp2 = 0;
          if (em1) {
            q1[a] = (q[a] - 1 + g.n[a]) % g.n[a];
            em2 = neighbour(g, q1, a, -1, nm2);
            if (em2) { fm2 = cell_flags<MASK>(g, nm2); in.load(g, nm2, Xm2); }
          }
          if (ep1) {
            q1[a] = (q[a] + 1) % g.n[a];
            ep2 = neighbour(g, q1, a, +1, np2);
            if (ep2) { fp2 = cell_flags<MASK>(g, np2); in.load(g, np2, Xp2); }
          }
          const bool okc_m = em1 && (fm1 & CELL_FLUID);  // this cell's lower difference
          const bool okc_p = ep1 && (fp1 & CELL_FLUID);
          const bool okm_m = em2 && (fm2 & CELL_FLUID);  // cell -1: its lower difference
          const bool okp_p = ep2 && (fp2 & CELL_FLUID);  // cell +1: its upper difference
          const bool cfluid = fc & CELL_FLUID;
          // reconstructed values; positivity fallback per face side
          double sc[NF], sm[NF], sp[NF];
#pragma unroll
          for (int k = 0; k < NF; ++k) {
            double dm = okc_m ? Xc[k] - Xm1[k] : 0.0;
            double dp = okc_p ? Xp1[k] - Xc[k] : 0.0;
            sc[k] = minmod(dp, dm);
            if (em1) {
              double dmm = okm_m ? Xm1[k] - Xm2[k] : 0.0;
              double dpm = cfluid ? Xc[k] - Xm1[k] : 0.0;
              sm[k] = minmod(dpm, dmm);
            }
            if (ep1) {
              double dmp = cfluid ? Xp1[k] - Xc[k] : 0.0;
              double dpp = okp_p ? Xp2[k] - Xp1[k] : 0.0;
              sp[k] = minmod(dpp, dmp);
            }
          }
          // reconstruction.cpp:67-97 extension: mean +/- slope/2, fallback to
          // mean; the tuple built from the primitives of the positivity check
          auto recon = [&](const double *mean, const double *s, double sgn, double *v, Tup &T) {
#pragma unroll
            for (int k = 0; k < NF; ++k) v[k] = (sgn > 0) ? mean[k] + 0.5 * s[k] : mean[k] - 0.5 * s[k];
            double pr[NF];
            cons2prim(v, pr);
            if (!(v[FID_RHO] > 0.0) || !(pr[FID_P] > 0.0)) {
#pragma unroll
              for (int k = 0; k < NF; ++k) v[k] = mean[k];
              cons2prim(v, pr);
            }
            make_tup_prim(v, pr, T);
          };
          recon(Xc, sc, -1.0, Rm, TRm2);
          recon(Xc, sc, +1.0, Lp, TLp2);
          if (!em1) border_state(g, 2 * a, a, q, Rm, Lm);
          else if (wall_m) reflect_axis(Rm, a, Lm);
          else recon(Xm1, sm, +1.0, Lm, TLm2);
          if (!ep1) border_state(g, 2 * a + 1, a, q, Lp, Rp);
          else if (wall_p) reflect_axis(Lp, a, Rp);
          else recon(Xp1, sp, -1.0, Rp, TRp2);
          if (!em1 || wall_m) make_tup(Lm, TLm2);
          if (!ep1 || wall_p) make_tup(Rp, TRp2);
        }

        Tup TLm, TRm, TLp, TRp;
        if (ORDER == 1) {
          TRm = Tc;
          TLp = Tc;
          // free-flow border and copy states reuse the computed tuple
          make_tup(Lm, TLm);
          make_tup(Rp, TRp);
        } else {
          TRm = TRm2;
          TLp = TLp2;
          TLm = TLm2;
          TRp = TRp2;
        }

        double Fm[NF], Fp[NF];
        double lm = face_flux(TLm, TRm, a, Fm);
        double lp = face_flux(TLp, TRp, a, Fp);
        lam_max = lam_acc(lam_max, lm);
        lam_max = lam_acc(lam_max, lp);

        // positivity counts once per face (the reference evaluates each face once)
        if (mode == MODE_STEP || mode == MODE_RHS_ONLY) {
          const bool other_solved_m = em1 && !wall_m && (fm1 & CELL_SOLVED);
          if (!other_solved_m) count_neg(TLm, TRm, ctrl);
          count_neg(TLp, TRp, ctrl);
        }

        const double A = g.A;
        // periodic wrap: the wrap face is first in the reference-order face list
        const bool wrap_first = g.periodic[a] && (q[a] == g.n[a] - 1);
        if (wrap_first) {
#pragma unroll
          for (int k = 0; k < NF; ++k) { rhs[k] -= A * Fp[k]; rhs[k] += A * Fm[k]; }
        } else {
#pragma unroll
          for (int k = 0; k < NF; ++k) { rhs[k] += A * Fm[k]; rhs[k] -= A * Fp[k]; }
        }
      }

      const int64_t fs = g.fs;
#pragma unroll
      for (int k = 0; k < NF; ++k) {
        double u = (STAGE == 1) ? 0.0 : U[k * fs + idx];
        out[k * fs + idx] = rk_update<STAGE>(g, u, Xc[k], rhs[k], in.dt);
      }
    } else {
      const int64_t fs = g.fs;
      if (STAGE == 1) {
#pragma unroll
        for (int k = 0; k < NF; ++k) out[k * fs + idx] = 0.0;
      } else if (STAGE == 2) {
#pragma unroll
        for (int k = 0; k < NF; ++k) out[k * fs + idx] = U[k * fs + idx];
      }
    }
  }
  block_max_to<BLOCK>(ctrl, slot, lam_max, threadIdx.x);
}

// t += dt after stage 3 (main.cpp:505-506); resets the stage maxima.
__global__ void k_advance(Ctrl *c) { advance_ctrl_wave(c, threadIdx.x); }

// The early maximum of state U: what the stage-1 faces of a step starting
// from U give on a grid whose unsolved cells are all solid (every face with a
// solved side takes the larger of its two sides' |u_a| + c, euler.cpp:232-234):
// over the solved cells, every axis (each has faces along every axis); along
// an axis with a wall face (a non-fluid neighbour) or a reflecting / wall
// border, the reflected virtual state's (euler.cpp:318-356); at a Dirichlet
// border, the border state's (euler.cpp:366-376).  Ctrl::lam_now is zero
// beforehand.
__global__ __launch_bounds__(256) void k_lam_init(GridDev g, const double *U, Ctrl *ctrl) {
  double m = 0.0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < g.N; i += stride) {
    if (g.flags && !(g.flags[i] & CELL_SOLVED)) continue;
    double c[NF];
#pragma unroll
    for (int k = 0; k < NF; ++k) c[k] = U[k * g.fs + i];
    Tup T;
    make_tup(c, T);
    const int64_t q[3] = {i % g.n[0], (i / g.n[0]) % g.n[1], i / (g.n[0] * g.n[1])};
    for (int a = 0; a < g.dim; ++a) {
      m = lam_acc(m, fabs(T.vel[a]) + T.a);
      bool refl = false;
      for (int d = 0; d < 2; ++d) {
        const int side = 2 * a + d;
        int64_t nb;
        if (neighbour(g, q, a, d ? 1 : -1, nb)) {
          refl = refl || (g.flags && !(g.flags[nb] & CELL_FLUID));
        } else if (g.bc[side] == BC_REFLECTING || g.bc[side] == BC_WALL) {
          refl = true;
        } else if (g.bc[side] == BC_DIRICHLET) {
          double v[NF];
          Tup V;
          prim2cons(g.info[side], v);
          make_tup(v, V);
          m = lam_acc(m, fabs(V.vel[a]) + V.a);
        }
      }
      if (refl) {
        double v[NF];
        Tup V;
        reflect_axis(c, a, v);
        make_tup(v, V);
        m = lam_acc(m, fabs(V.vel[a]) + V.a);
      }
    }
  }
  __shared__ unsigned long long red[4];
  unsigned long long b = dbits(m);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned long long o = __shfl_xor(b, off, 64);
    b = (o > b) ? o : b;
  }
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = b;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long r = red[0];
    for (int w = 1; w < 4; ++w) r = (red[w] > r) ? red[w] : r;
    if (r) atomicMax(&ctrl->lam_now, r);
  }
}

__global__ void k_zero_slot(Ctrl *c, int slot) {
  if (threadIdx.x < EIG_SPLIT) c->eig[slot][threadIdx.x][0] = 0ull;
}

template <int DIM, int ORDER, int STAGE, bool MASK>
static hipError_t launch_stage_t(const GridDev &g, int mode, const double *U, const double *R, const double *W,
                                 double *out_R, double *out_W, double *out_U, Ctrl *ctrl, hipStream_t stream) {
  StageIn<STAGE> in;
  in.fs = g.fs;
  in.dt = 0.0;
  in.U = U;
  in.R = R;
  in.X = (STAGE == 3) ? W : U;
  double *out = (STAGE == 1) ? out_R : (STAGE == 2) ? out_W : out_U;
  dim3 grid((unsigned)((g.N + BLOCK - 1) / BLOCK));
  hipLaunchKernelGGL((k_stage<DIM, ORDER, STAGE, MASK>), grid, dim3(BLOCK), 0, stream, g, in, U, out, ctrl, mode);
  return hipGetLastError();
}

template <int DIM, int ORDER, bool MASK>
static hipError_t launch_stage_s(const GridDev &g, int stage, int mode, const double *U, const double *R,
                                 const double *W, double *oR, double *oW, double *oU, Ctrl *ctrl, hipStream_t s) {
  switch (stage) {
    case 1: return launch_stage_t<DIM, ORDER, 1, MASK>(g, mode, U, R, W, oR, oW, oU, ctrl, s);
    case 2: return launch_stage_t<DIM, ORDER, 2, MASK>(g, mode, U, R, W, oR, oW, oU, ctrl, s);
    default: return launch_stage_t<DIM, ORDER, 3, MASK>(g, mode, U, R, W, oR, oW, oU, ctrl, s);
  }
}

hipError_t launch_stage(const GridDev &g, int stage, int mode, const double *U, const double *R, const double *W,
                        double *oR, double *oW, double *oU, Ctrl *ctrl, hipStream_t s) {
  const bool mask = g.flags != nullptr;
  if (g.dim == 2) {
    if (g.order == 1)
      return mask ? launch_stage_s<2, 1, true>(g, stage, mode, U, R, W, oR, oW, oU, ctrl, s)
                  : launch_stage_s<2, 1, false>(g, stage, mode, U, R, W, oR, oW, oU, ctrl, s);
    return mask ? launch_stage_s<2, 2, true>(g, stage, mode, U, R, W, oR, oW, oU, ctrl, s)
                : launch_stage_s<2, 2, false>(g, stage, mode, U, R, W, oR, oW, oU, ctrl, s);
  }
  if (g.order == 1)
    return mask ? launch_stage_s<3, 1, true>(g, stage, mode, U, R, W, oR, oW, oU, ctrl, s)
                : launch_stage_s<3, 1, false>(g, stage, mode, U, R, W, oR, oW, oU, ctrl, s);
  return mask ? launch_stage_s<3, 2, true>(g, stage, mode, U, R, W, oR, oW, oU, ctrl, s)
              : launch_stage_s<3, 2, false>(g, stage, mode, U, R, W, oR, oW, oU, ctrl, s);
}

hipError_t launch_lam_init(const GridDev &g, const double *U, Ctrl *ctrl, hipStream_t s) {
  const int64_t want = (g.N + 255) / 256;
  const unsigned nb = (unsigned)(want < 2048 ? (want > 0 ? want : 1) : 2048);
  hipLaunchKernelGGL(k_lam_init, dim3(nb), dim3(256), 0, s, g, U, ctrl);
  return hipGetLastError();
}

hipError_t launch_advance(Ctrl *ctrl, hipStream_t s) {
  hipLaunchKernelGGL(k_advance, dim3(1), dim3(64), 0, s, ctrl);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// output fields (solver_writer.cpp:56-119; main.cpp:511-517 converts every
// cell with utils::conservative2primitive first): one thread per cell
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_output_fields(GridDev g, const double *U, const double *R,
                                                       const uint8_t *flags, char *stage, OutputLayout L) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= g.N) return;
  const int64_t fs = g.fs;
  double c[NF], p[NF];
#pragma unroll
  for (int k = 0; k < NF; ++k) c[k] = U[k * fs + i];
  cons2prim(c, p);
  reinterpret_cast<int32_t *>(stage + L.solved)[i] = (flags == nullptr || (flags[i] & CELL_SOLVED)) ? 1 : 0;
  double *vel = reinterpret_cast<double *>(stage + L.vel);
  vel[3 * i] = p[FID_U];
  vel[3 * i + 1] = p[FID_V];
  vel[3 * i + 2] = p[FID_W];
  reinterpret_cast<double *>(stage + L.p)[i] = p[FID_P];
  reinterpret_cast<double *>(stage + L.T)[i] = p[FID_T];
  reinterpret_cast<double *>(stage + L.rho)[i] = c[FID_RHO];
#pragma unroll
  for (int k = 0; k < NF; ++k) reinterpret_cast<double *>(stage + L.res[k])[i] = R[k * fs + i];
}

hipError_t launch_output_fields(const GridDev &g, const double *U, const double *R, const uint8_t *flags, char *stage,
                                const OutputLayout &L, hipStream_t s) {
  if (g.N <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_output_fields, dim3((unsigned)((g.N + 255) / 256)), dim3(256), 0, s, g, U, R, flags, stage, L);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// host layout <-> device layout (storage.hpp:31-41): AoS rows of 5 doubles in
// raw order (row r holds cell cell_of_raw[r]; identity when null) and the
// SoA planes of the state
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_aos_to_soa(const double *aos, const int64_t *cell_of_raw, int64_t N,
                                                    int64_t fs, double *soa) {
  const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (r >= N) return;
  const int64_t c = cell_of_raw ? cell_of_raw[r] : r;
#pragma unroll
  for (int k = 0; k < NF; ++k) soa[k * fs + c] = aos[NF * r + k];
}

__global__ __launch_bounds__(256) void k_soa_to_aos(const double *soa, const int64_t *cell_of_raw, int64_t N,
                                                    int64_t fs, double *aos) {
  const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (r >= N) return;
  const int64_t c = cell_of_raw ? cell_of_raw[r] : r;
#pragma unroll
  for (int k = 0; k < NF; ++k) aos[NF * r + k] = soa[k * fs + c];
}

hipError_t launch_aos_to_soa(const double *aos, const int64_t *cell_of_raw, int64_t N, int64_t fs, double *soa,
                             hipStream_t s) {
  if (N <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_aos_to_soa, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, s, aos, cell_of_raw, N, fs, soa);
  return hipGetLastError();
}

hipError_t launch_soa_to_aos(const double *soa, const int64_t *cell_of_raw, int64_t N, int64_t fs, double *aos,
                             hipStream_t s) {
  if (N <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_soa_to_aos, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, s, soa, cell_of_raw, N, fs, aos);
  return hipGetLastError();
}

hipError_t launch_zero_slot(Ctrl *ctrl, int slot, hipStream_t s) {
  hipLaunchKernelGGL(k_zero_slot, dim3(1), dim3(64), 0, s, ctrl, slot);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// error norms (main.cpp:550-561): deterministic two-level tree sum
// ---------------------------------------------------------------------------
constexpr int ERR_BLOCK = 256;
constexpr int ERR_PER_THREAD = 16;

int64_t error_norm_blocks(int64_t N) {
  const int64_t per_block = (int64_t)ERR_BLOCK * ERR_PER_THREAD;
  return (N + per_block - 1) / per_block;
}

__global__ __launch_bounds__(ERR_BLOCK) void k_err_partial(const double *rho, const double *rho_ex, int64_t N,
                                                           double V, double *partials) {
  __shared__ double s1[ERR_BLOCK], s2[ERR_BLOCK];
  const int64_t base = (int64_t)blockIdx.x * ERR_BLOCK * ERR_PER_THREAD;
  double a1 = 0.0, a2 = 0.0;
  for (int r = 0; r < ERR_PER_THREAD; ++r) {
    int64_t i = base + (int64_t)r * ERR_BLOCK + threadIdx.x;
    if (i < N) {
      double d = rho[i] - rho_ex[i];
      a1 += fabs(d) * V;
      a2 += d * d * V;
    }
  }
  s1[threadIdx.x] = a1;
  s2[threadIdx.x] = a2;
  __syncthreads();
  for (int w = ERR_BLOCK / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      s1[threadIdx.x] += s1[threadIdx.x + w];
      s2[threadIdx.x] += s2[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    partials[2 * blockIdx.x] = s1[0];
    partials[2 * blockIdx.x + 1] = s2[0];
  }
}

__global__ __launch_bounds__(ERR_BLOCK) void k_err_final(const double *partials, int64_t nb, double *out2) {
  __shared__ double s1[ERR_BLOCK], s2[ERR_BLOCK];
  double a1 = 0.0, a2 = 0.0;
  for (int64_t i = threadIdx.x; i < nb; i += ERR_BLOCK) {
    a1 += partials[2 * i];
    a2 += partials[2 * i + 1];
  }
  s1[threadIdx.x] = a1;
  s2[threadIdx.x] = a2;
  __syncthreads();
  for (int w = ERR_BLOCK / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      s1[threadIdx.x] += s1[threadIdx.x + w];
      s2[threadIdx.x] += s2[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out2[0] = s1[0];
    out2[1] = s2[0];
  }
}

hipError_t launch_error_norms(const GridDev &g, const double *U, const double *rho_ex, double *partials,
                              int64_t nblocks, double *out2, hipStream_t s) {
  hipLaunchKernelGGL(k_err_partial, dim3((unsigned)nblocks), dim3(ERR_BLOCK), 0, s, U, rho_ex, g.N, g.V, partials);
  hipLaunchKernelGGL(k_err_final, dim3(1), dim3(ERR_BLOCK), 0, s, partials, nblocks, out2);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// body marking (body.cpp:80-95 boxes, inclusive; sphere level set)
// ---------------------------------------------------------------------------
__global__ void k_mark_bodies(GridDev g, double ox, double oy, double oz, const double *boxes, int nb,
                              const double *spheres, int ns, uint8_t *flags) {
  int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= g.N) return;
  int64_t i = idx % g.n[0], j = (idx / g.n[0]) % g.n[1], k = idx / (g.n[0] * g.n[1]);
  double p[3];
  p[0] = ox + (i + 0.5) * g.h;
  p[1] = oy + (j + 0.5) * g.h;
  p[2] = (g.dim == 3) ? oz + (k + 0.5) * g.h : oz;
  bool fluid = true;
  for (int b = 0; b < nb && fluid; ++b) {
    const double *x = boxes + 6 * b;
    bool out = (p[0] < x[0] || p[0] > x[3]) || (p[1] < x[1] || p[1] > x[4]) || (p[2] < x[2] || p[2] > x[5]);
    if (!out) fluid = false;
  }
  for (int s = 0; s < ns && fluid; ++s) {
    const double *c = spheres + 4 * s;
    double dx = p[0] - c[0], dy = p[1] - c[1], dz = p[2] - c[2];
    double phi = sqrt(dx * dx + dy * dy + dz * dz) - c[3];
    if (!(phi >= 0.0)) fluid = false;
  }
  flags[idx] = fluid ? (uint8_t)(CELL_FLUID | CELL_SOLVED) : (uint8_t)0;
}

hipError_t launch_mark_bodies(const GridDev &g, const double *origin, const double *boxes, int nboxes,
                              const double *spheres, int nspheres, uint8_t *flags, hipStream_t s) {
  dim3 grid((unsigned)((g.N + 255) / 256));
  hipLaunchKernelGGL(k_mark_bodies, grid, dim3(256), 0, s, g, origin[0], origin[1], origin[2], boxes, nboxes,
                     spheres, nspheres, flags);
  return hipGetLastError();
}

__global__ void k_fill(double *p, int64_t n, double v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

hipError_t launch_fill_double(double *p, int64_t n, double v, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n, v);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// one RK3 stage at a time: the update loops of main.cpp:409-422 (stage 1),
// 445-458 (stage 2) and 481-494 (stage 3) as their own launch, the same
// expressions as the fused stage kernels (rk_update, StageIn<2>)
// ---------------------------------------------------------------------------
template <int STAGE>
__global__ __launch_bounds__(256) void k_rk_update(GridDev g, double *U, const double *R, double *W, double dt) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= g.N) return;
  const bool solved = g.flags == nullptr || (g.flags[i] & CELL_SOLVED);
  const int64_t fs = g.fs;
#pragma unroll
  for (int k = 0; k < NF; ++k) {
    const int64_t o = k * fs + i;
    if (STAGE == 1) W[o] = solved ? U[o] + div_vol(g, dt * R[o]) : U[o];
    else if (STAGE == 2) W[o] = solved ? rk_update<2>(g, U[o], W[o], R[o], dt) : U[o];
    else if (solved) U[o] = rk_update<3>(g, U[o], W[o], R[o], dt);
  }
}

// main.cpp:505-506 after a stage-3 update: t += dt, the step counted, the
// three stage maxima kept as the step's
__global__ void k_rk_close(Ctrl *c, double dt, double e0, double e1, double e2) {
  if (threadIdx.x != 0) return;
  c->last_eig[0] = e0;
  c->last_eig[1] = e1;
  c->last_eig[2] = e2;
  c->dt = dt;
  c->t += dt;
  c->step += 1;
}

hipError_t launch_rk_update(const GridDev &g, int stage, double *U, const double *R, double *W, double dt,
                            hipStream_t s) {
  const dim3 grid((unsigned)((g.N + 255) / 256));
  if (stage == 1) hipLaunchKernelGGL(k_rk_update<1>, grid, dim3(256), 0, s, g, U, R, W, dt);
  else if (stage == 2) hipLaunchKernelGGL(k_rk_update<2>, grid, dim3(256), 0, s, g, U, R, W, dt);
  else hipLaunchKernelGGL(k_rk_update<3>, grid, dim3(256), 0, s, g, U, R, W, dt);
  return hipGetLastError();
}

hipError_t launch_rk_close(Ctrl *ctrl, double dt, const double *eig, hipStream_t s) {
  hipLaunchKernelGGL(k_rk_close, dim3(1), dim3(64), 0, s, ctrl, dt, eig[0], eig[1], eig[2]);
  return hipGetLastError();
}

// Ghost lists (communications.tpp:164-199 packs 5 doubles per listed cell):
// rows of the SoA state at cells[0..n) (cell order) <-> AoS buf [n][5].
__global__ __launch_bounds__(256) void k_pack_cells(const double *X, int64_t fs, const int64_t *cells, int64_t n,
                                                    double *buf) {
  const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (j >= n) return;
  const int64_t c = cells[j];
#pragma unroll
  for (int k = 0; k < NF; ++k) buf[NF * j + k] = X[k * fs + c];
}

__global__ __launch_bounds__(256) void k_unpack_cells(double *X, int64_t fs, const int64_t *cells, int64_t n,
                                                      const double *buf) {
  const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (j >= n) return;
  const int64_t c = cells[j];
#pragma unroll
  for (int k = 0; k < NF; ++k) X[k * fs + c] = buf[NF * j + k];
}

hipError_t launch_pack_cells(const double *X, int64_t fs, const int64_t *cells, int64_t n, double *buf,
                             hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_pack_cells, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, X, fs, cells, n, buf);
  return hipGetLastError();
}

hipError_t launch_unpack_cells(double *X, int64_t fs, const int64_t *cells, int64_t n, const double *buf,
                               hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_unpack_cells, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, X, fs, cells, n, buf);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// batched point functions (AoS [count][5] in and out)
// ---------------------------------------------------------------------------
__global__ void k_b_c2p(const double *c, int64_t n, double *p) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) cons2prim(c + NF * i, p + NF * i);
}
__global__ void k_b_p2c(const double *p, int64_t n, double *c) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) prim2cons(p + NF * i, c + NF * i);
}
__global__ void k_b_flux(const double *c, const double *p, const double *nrm, int64_t n, double *f) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int a = 0, b = 0;
  if (i < n) eval_fluxes(c + NF * i, p + NF * i, nrm + 3 * i, f + NF * i, &a, &b);
}
__global__ void k_b_split(const double *L, const double *R, const double *nrm, int64_t n, double *F, double *lam) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int a = 0, b = 0;
  if (i < n) eval_splitting(L + NF * i, R + NF * i, nrm + 3 * i, F + NF * i, lam + i, &a, &b);
}
__global__ void k_b_bc(int bc, const double *nrm, const double *info, const double *c, int64_t n, double *out) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double *ci = c + NF * i;
  double *o = out + NF * i;
  if (bc == BC_FREE_FLOW) {
    for (int k = 0; k < NF; ++k) o[k] = ci[k];
  } else if (bc == BC_REFLECTING || bc == BC_WALL) {
    reflect_general(nrm + 3 * i, ci, o);
  } else if (bc == BC_DIRICHLET) {
    prim2cons(info + NF * i, o);
  }
}

static dim3 grid_for(int64_t n) { return dim3((unsigned)((n + 255) / 256)); }

hipError_t launch_batch_cons2prim(const double *c, int64_t n, double *p, hipStream_t s) {
  hipLaunchKernelGGL(k_b_c2p, grid_for(n), dim3(256), 0, s, c, n, p);
  return hipGetLastError();
}
hipError_t launch_batch_prim2cons(const double *p, int64_t n, double *c, hipStream_t s) {
  hipLaunchKernelGGL(k_b_p2c, grid_for(n), dim3(256), 0, s, p, n, c);
  return hipGetLastError();
}
hipError_t launch_batch_eval_fluxes(const double *c, const double *p, const double *nrm, int64_t n, double *f,
                                    hipStream_t s) {
  hipLaunchKernelGGL(k_b_flux, grid_for(n), dim3(256), 0, s, c, p, nrm, n, f);
  return hipGetLastError();
}
hipError_t launch_batch_eval_splitting(const double *L, const double *R, const double *nrm, int64_t n, double *F,
                                       double *lam, hipStream_t s) {
  hipLaunchKernelGGL(k_b_split, grid_for(n), dim3(256), 0, s, L, R, nrm, n, F, lam);
  return hipGetLastError();
}
hipError_t launch_batch_bc_values(int bc, const double *nrm, const double *info, const double *c, int64_t n,
                                  double *out, hipStream_t s) {
  hipLaunchKernelGGL(k_b_bc, grid_for(n), dim3(256), 0, s, bc, nrm, info, c, n, out);
  return hipGetLastError();
}

// plain copy kernel for HBM byte-counter calibration
// one 16-byte element per thread: the most memory-level parallelism per byte
__global__ __launch_bounds__(256) void k_copy(const double2 *__restrict__ src, double2 *__restrict__ dst, int64_t n2) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n2) dst[i] = src[i];
}

hipError_t launch_copy(const double4 *src, double4 *dst, int64_t n4, hipStream_t s) {
  const int64_t n2 = 2 * n4;
  const int64_t blocks = (n2 + 255) / 256;
  if (blocks < 1) return hipSuccess;
  hipLaunchKernelGGL(k_copy, dim3((unsigned)blocks), dim3(256), 0, s, (const double2 *)src, (double2 *)dst, n2);
  return hipGetLastError();
}

}  // namespace mmf
