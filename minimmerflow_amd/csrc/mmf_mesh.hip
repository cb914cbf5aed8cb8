// mmf_mesh.hip -- the finite-volume step on an explicit interface list
// (the general MeshGeometricalInfo, mesh_info.hpp:31-105 / mesh_info.cpp:86-118):
// any conforming or hanging-face mesh given as owner / neighbour / area /
// normal per face and a volume per cell, order 1 (reconstruction.cpp:90-97).
//
// Two kernels per RK stage, no float atomics:
//   k_mesh_faces  one thread per interface, in the caller's interface order:
//                 coalesced SoA loads of the face's geometry and kind, the two
//                 side states gathered from the SoA state, the boundary or
//                 wall virtual state (euler.cpp:198-225), the Rusanov flux
//                 (euler.cpp:42-73), A*F written per face (SoA), the face
//                 eigenvalue into the stage maximum;
//   k_mesh_cells  one thread per cell: a segmented sum over the cell's
//                 incident faces in increasing face index (the order the
//                 reference's interface loop reaches them, euler.cpp:153-247:
//                 RHS -= A*F as owner, += A*F as neighbour), fused with the RK
//                 stage update (main.cpp:409-495, per-cell volume).
// The residual of every cell is summed in the reference's face order, so the
// results are bitwise reproducible and equal to a serial interface loop over
// the same face list.
#include <hip/hip_runtime.h>

#include "mmf_device.hpp"
#include "mmf_internal.hpp"
#include "mmf_mesh.hpp"
#include "mmf_stage.hpp"

namespace mmf {

constexpr int MB = 256;

__device__ __forceinline__ bool mesh_solved(const MeshDev &m, int64_t c) {
  return m.flags == nullptr || (m.flags[c] & CELL_SOLVED);
}

// euler.cpp:261-376: the virtual state beyond a border or wall face, from the
// fluid side's state and the normal pointing out of the fluid.
__device__ __forceinline__ void mesh_bc_state(const MeshDev &m, int32_t kind, const double *n, const double *fluid,
                                              double *out) {
  const int bc = (kind & MESH_KIND_BC_MASK) - 1;
  if (bc == BC_REFLECTING || bc == BC_WALL) {
    reflect_general(n, fluid, out);
  } else if (bc == BC_DIRICHLET) {
    prim2cons(m.info + NF * (kind >> MESH_KIND_SLOT_SHIFT), out);
  } else {
#pragma unroll
    for (int k = 0; k < NF; ++k) out[k] = fluid[k];
  }
}

__global__ __launch_bounds__(MB) void k_mesh_faces(MeshDev m, const double *__restrict__ X, double *__restrict__ AF,
                                                   Ctrl *ctrl, int slot, int mode) {
  if (mode == MODE_STEP && !(ctrl->t < ctrl->tmax)) return;  // main.cpp:377 loop guard
  const int64_t f = (int64_t)blockIdx.x * MB + threadIdx.x;
  double lam = 0.0;
  int np = 0, nr = 0;
  if (f < m.F) {
    const int64_t o = m.owner[f], nb = m.neigh[f];
    const bool os = mesh_solved(m, o);
    const bool ns = nb >= 0 && mesh_solved(m, nb);
    if (os || ns) {  // euler.cpp:181-183
      const int32_t kind = m.kind[f];
      const int32_t nc = (kind >> MESH_KIND_NRM_SHIFT) & 7;
      double n[3];
      if (nc == MESH_NRM_GENERAL) {
        n[0] = m.normal[f];
        n[1] = m.normal[m.F + f];
        n[2] = m.normal[2 * m.F + f];
      } else {  // +-e_a: the same doubles as the caller's (1 or -1, and +0)
        n[0] = n[1] = n[2] = 0.0;
        n[nc >> 1] = (nc & 1) ? -1.0 : 1.0;
      }
      double L[NF], R[NF];
      if ((kind & MESH_KIND_BC_MASK) == 0) {
#pragma unroll
        for (int k = 0; k < NF; ++k) {
          L[k] = X[k * m.fs + o];
          R[k] = X[k * m.fs + nb];
        }
      } else if (os) {
#pragma unroll
        for (int k = 0; k < NF; ++k) L[k] = X[k * m.fs + o];
        mesh_bc_state(m, kind, n, L, R);
      } else {
        // the fluid side is the neighbour: the boundary state sees the
        // flipped normal (euler.cpp:219-222); the flux keeps the owner's
        const double loc[3] = {-1. * n[0], -1. * n[1], -1. * n[2]};
#pragma unroll
        for (int k = 0; k < NF; ++k) R[k] = X[k * m.fs + nb];
        mesh_bc_state(m, kind, loc, R, L);
      }
      double Fx[NF];
      eval_splitting(L, R, n, Fx, &lam, &np, &nr);
      const int32_t as = (kind >> MESH_KIND_AREA_SHIFT) & 255;
      const double a = (as == MESH_AREA_GENERAL) ? m.area[f] : m.area_tab[as];
#pragma unroll
      for (int k = 0; k < NF; ++k) AF[k * m.F + f] = a * Fx[k];  // euler.cpp:239,245: interfaceArea * fluxes[k]
    }
  }
  // positivity counts: one atomic per wave
  unsigned int sp = np, sr = nr;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    sp += __shfl_xor(sp, off, 64);
    sr += __shfl_xor(sr, off, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    if (sp) atomicAdd(&ctrl->neg_p, (unsigned long long)sp);
    if (sr) atomicAdd(&ctrl->neg_rho, (unsigned long long)sr);
  }
  block_max_to<MB>(ctrl, slot, lam, threadIdx.x);
}

// One cell's residual: its incident faces in increasing face index.
__device__ __forceinline__ void mesh_gather(const MeshDev &m, int64_t c, const double *__restrict__ AF, double *rhs) {
#pragma unroll
  for (int k = 0; k < NF; ++k) rhs[k] = 0.;  // euler.cpp:142-148
  const int64_t e0 = m.off[c], e1 = m.off[c + 1];
  for (int64_t e = e0; e < e1; ++e) {
    const int64_t v = m.inc[e];
    const int64_t f = v >> 1;
    if (v & 1) {
#pragma unroll
      for (int k = 0; k < NF; ++k) rhs[k] += AF[k * m.F + f];
    } else {
#pragma unroll
      for (int k = 0; k < NF; ++k) rhs[k] -= AF[k * m.F + f];
    }
  }
}

template <int STAGE>
__global__ __launch_bounds__(MB) void k_mesh_cells(MeshDev m, const double *__restrict__ U, const double *W,
                                                   double *out, const double *__restrict__ AF, Ctrl *ctrl, int mode) {
  if (mode == MODE_STEP && !(ctrl->t < ctrl->tmax)) return;
  const int64_t c = (int64_t)blockIdx.x * MB + threadIdx.x;
  if (c >= m.N) return;
  const bool sv = mesh_solved(m, c);
  const int64_t fs = m.fs;
  if (mode == MODE_RHS_ONLY) {
    double rhs[NF] = {0., 0., 0., 0., 0.};
    if (sv) mesh_gather(m, c, AF, rhs);
#pragma unroll
    for (int k = 0; k < NF; ++k) out[k * fs + c] = rhs[k];
    return;
  }
  if (!sv) return;  // main.cpp:411-414: unsolved cells keep their state
  const double dt = step_dt(ctrl);
  double rhs[NF];
  mesh_gather(m, c, AF, rhs);
  const double V = m.vol[c];  // rawGetCellVolume
#pragma unroll
  for (int k = 0; k < NF; ++k) {
    const double u = U[k * fs + c];
    double r;
    if (STAGE == 1) r = u + dt * rhs[k] / V;                                     // main.cpp:421
    else if (STAGE == 2) r = 0.75 * u + 0.25 * (W[k * fs + c] + dt * rhs[k] / V);  // main.cpp:457
    else r = (1. / 3) * u + (2. / 3) * (W[k * fs + c] + dt * rhs[k] / V);          // main.cpp:493
    out[k * fs + c] = r;
  }
}

hipError_t launch_mesh_faces(const MeshDev &m, const double *X, double *AF, Ctrl *ctrl, int slot, int mode,
                             hipStream_t s) {
  const int64_t nb = (m.F + MB - 1) / MB;
  if (nb < 1) return hipSuccess;
  hipLaunchKernelGGL(k_mesh_faces, dim3((unsigned)nb), dim3(MB), 0, s, m, X, AF, ctrl, slot, mode);
  return hipGetLastError();
}

hipError_t launch_mesh_cells(const MeshDev &m, int stage, const double *U, const double *W, double *out,
                             const double *AF, Ctrl *ctrl, int mode, hipStream_t s) {
  const int64_t nb = (m.N + MB - 1) / MB;
  if (nb < 1) return hipSuccess;
  switch (stage) {
    case 1:
      hipLaunchKernelGGL(k_mesh_cells<1>, dim3((unsigned)nb), dim3(MB), 0, s, m, U, W, out, AF, ctrl, mode);
      break;
    case 2:
      hipLaunchKernelGGL(k_mesh_cells<2>, dim3((unsigned)nb), dim3(MB), 0, s, m, U, W, out, AF, ctrl, mode);
      break;
    default:
      hipLaunchKernelGGL(k_mesh_cells<3>, dim3((unsigned)nb), dim3(MB), 0, s, m, U, W, out, AF, ctrl, mode);
      break;
  }
  return hipGetLastError();
}

}  // namespace mmf
