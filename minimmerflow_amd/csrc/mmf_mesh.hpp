// mmf_mesh.hpp -- device view of an explicit interface list (mmf_mesh_*).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mmf_internal.hpp"

namespace mmf {

// Face kind, 32 bits:
//   bits  0-3   BC code + 1 (0 = interior face between two fluid cells, BC_NONE)
//   bits  4-6   normal: 0-5 = +x, -x, +y, -y, +z, -z exactly (the octree's
//               axis-aligned faces), 7 = read it from the normal arrays
//   bits  7-14  area: slot 0-254 of the area table, 255 = read the area array
//   bits 15-30  Dirichlet data slot
// An octree mesh (bitpit's VolOctree) has axis-aligned faces and a handful of
// face areas, so its faces need 12 bytes of geometry (owner, neighbour, kind)
// instead of 44.
constexpr int32_t MESH_KIND_BC_MASK = 15;
constexpr int MESH_KIND_NRM_SHIFT = 4;
constexpr int32_t MESH_NRM_GENERAL = 7;
constexpr int MESH_KIND_AREA_SHIFT = 7;
constexpr int32_t MESH_AREA_GENERAL = 255;
constexpr int MESH_KIND_SLOT_SHIFT = 15;
constexpr int MESH_MAX_AREAS = 255;
constexpr int32_t MESH_MAX_SLOTS = 1 << 16;

struct MeshDev {
  int64_t N;              // cells
  int64_t F;              // faces
  int64_t fs;             // field stride of the SoA state (>= N, multiple of 64)
  const int32_t *owner;   // [F] owner cell (Interface::getOwner)
  const int32_t *neigh;   // [F] neighbour cell, -1 on borders (Interface::getNeigh)
  const double *area;     // [F] faces whose area is not in the table, else nullptr
  const double *normal;   // [3][F] owner-outward unit normal, SoA, for general faces, else nullptr
  const double *area_tab; // [<= 255] distinct face areas
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
