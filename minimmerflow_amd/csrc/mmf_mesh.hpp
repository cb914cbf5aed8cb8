// mmf_mesh.hpp -- device view of an explicit interface list (mmf_mesh_*).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mmf_internal.hpp"

namespace mmf {

// Face kind: (BC code + 1) in the low bits (0 = interior face between two
// fluid cells, BC_NONE), the Dirichlet data slot above them.
constexpr int32_t MESH_KIND_BC_MASK = 15;
constexpr int MESH_KIND_SLOT_SHIFT = 4;

struct MeshDev {
  int64_t N;              // cells
  int64_t F;              // faces
  int64_t fs;             // field stride of the SoA state (>= N, multiple of 64)
  const int32_t *owner;   // [F] owner cell (Interface::getOwner)
  const int32_t *neigh;   // [F] neighbour cell, -1 on borders (Interface::getNeigh)
  const double *area;     // [F]
  const double *normal;   // [3][F] owner-outward unit normal, SoA
  const int32_t *kind;    // [F]
  const double *info;     // [slots][5] Dirichlet primitives
  const double *vol;      // [N]
  const uint8_t *flags;   // [N] MMF_CELL_* bits, or nullptr (every cell fluid and solved)
  const int32_t *off;     // [N+1] first incidence of each cell
  const int32_t *inc;     // incidences, per cell in increasing face index: face << 1 | (1 on the neighbour side)
};

hipError_t launch_mesh_faces(const MeshDev &m, const double *X, double *AF, Ctrl *ctrl, int slot, int mode,
                             hipStream_t s);
hipError_t launch_mesh_cells(const MeshDev &m, int stage, const double *U, const double *W, double *out,
                             const double *AF, Ctrl *ctrl, int mode, hipStream_t s);

}  // namespace mmf
