// mmf_internal.hpp -- structures shared by the C-ABI layer and the kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mmf.h"

namespace mmf {

// Uniform Cartesian grid as the kernels see it (the closed-form
// MeshGeometricalInfo, mesh_info.hpp:31-105, plus the border BC table).
struct GridDev {
  int64_t n[3];        // cells per axis (n[2] == 1 in 2D)
  int64_t N;           // cells
  int64_t fs;          // field stride of the SoA state (>= N, multiple of 64)
  int32_t dim;
  int32_t order;
  double h;            // cell size (main.cpp:351-356 minCellSize)
  double V;            // cell volume h^dim
  double A;            // face area h^(dim-1)
  int32_t bc[6];       // x-, x+, y-, y+, z-, z+
  int32_t periodic[3];
  double info[6][5];   // Dirichlet primitives per side
  const uint8_t *flags;  // MMF_CELL_* bits per cell, or nullptr (all fluid+solved)
  int32_t ghost_free;    // with flags: every fluid cell is solved (no ghost cells)
  double invV;         // 1/V, used only when V is a power of two (x/V == x*invV exactly)
  int32_t pow2V;
  int32_t pow2A;       // A is a power of two: A*(0.5*x) == (0.5*A)*x exactly
  double halfA;        // 0.5*A
  // 3D marching kernel on a grid with flags: per block, its tile, with
  // MARCH_TILE_MASKED set where the block must read the flags; or nullptr
  const int32_t *march_tiles;
  // Border conditions per border face (main.cpp:251-260 keeps one per
  // interface), or nullptr when bc / info above hold per side: the slot of
  // each border face, sides one after another from border_off[side], a
  // side's faces ordered by its two other axes, the lower one fastest; each
  // slot's BC code and Dirichlet primitives.  Only the per-cell gather kernel
  // reads it.
  const uint8_t *bslot;
  const int32_t *bslot_bc;
  const double *bslot_info;
  int64_t border_off[6];
};

// Border faces of side 'side' (product of the two other axes' counts).
inline int64_t border_faces_of_side(const int64_t *n, int side) {
  const int a = side / 2;
  return (a == 0 ? n[1] * n[2] : a == 1 ? n[0] * n[2] : n[0] * n[1]);
}

constexpr int32_t MARCH_TILE_MASKED = 1 << 30;

// Tile of block b of nblk: XCD-aware numbering (consecutive blocks go to the
// 8 XCDs in turn; each XCD takes a contiguous range of tiles, whose halos
// then meet in its L2).
__host__ __device__ inline int march_tile_of_block(int b, int nblk) {
  return (nblk % 8 == 0) ? (b % 8) * (nblk / 8) + b / 8 : b;
}

// Stage-kernel variants.
enum KernelKind { KERNEL_AUTO = 0, KERNEL_GATHER = 1, KERNEL_MARCH = 2, KERNEL_TILE = 3 };

// Launch shape of the marching kernel (tile rows and march chunk).
struct MarchShape {
  int ty;       // tile rows (3D): 4, 8 or 16; 1 in 2D
  int chunk;    // planes per block along the march axis
  int tx = 64;  // tile columns (3D): 64 (one wave per row) or 32
  // 3D order 1: > 0 spreads the planes of all tiles over this many blocks
  // in shares of equal cost (up to MARCH_SEGS segments each, from a segment
  // table) instead of one (tile, chunk) per block; 0 keeps the chunks
  int seg_blocks = 0;
};

// Segments per block in the 3D marching kernel's segment table.
constexpr int MARCH_SEGS = 4;

// Stage maxima and block counters are spread over EIG_SPLIT cache lines each
// (block b uses line b % EIG_SPLIT): device-scope atomics on one address
// from hundreds of blocks serialise at the point of coherence (about 10 ns
// each), on different lines they proceed in parallel.  Batched cases, whose
// blocks per case are few, use line 0 only.
constexpr int EIG_SPLIT = 16;
constexpr int LINE_U64 = 16;  // 128 bytes

// Time-control block resident in device memory.
struct Ctrl {
  double t, tmax, cfl, hmin;
  double dt;                     // dt of the last completed step
  double last_eig[3];            // stage maxima of the last completed step
  long long step;
  unsigned long long neg_p, neg_rho;
  // steps whose stage maxima (or early maximum) were not finite: the state
  // holds a NaN or infinity; such a step leaves t, dt and the step count
  // unchanged
  unsigned long long nonfinite;
  // Early dt (the 3D marching kernel on grids without a body or a reflected
  // border): the stage-1 maximum of a step, max over cells and axes of
  // |u_a| + c, known before the step from the state it starts from (bits).
  // The step's stage-3 launch gathers the next one into lam[k][0].
  unsigned long long lam_now;
  // running stage maxima (bits) in entry [s][k][0], s = 3: compute_rhs; the
  // stage maximum is the largest over k
  alignas(128) unsigned long long eig[4][EIG_SPLIT][LINE_U64];
  alignas(128) unsigned long long lam[EIG_SPLIT][LINE_U64];
  // blocks of a step-closing launch that have finished, per line
  // ([k][0]) and lines completed ([EIG_SPLIT][0]); 0 between launches
  unsigned long long done[EIG_SPLIT + 1][LINE_U64];
};

// Host view of stage maximum s (the largest entry over the lines).
inline double ctrl_stage_max(const Ctrl &c, int s) {
  unsigned long long m = 0;
  for (int k = 0; k < EIG_SPLIT; ++k) m = (c.eig[s][k][0] > m) ? c.eig[s][k][0] : m;
  double v;
  __builtin_memcpy(&v, &m, sizeof(v));
  return v;
}

enum StageMode { MODE_STEP = 0, MODE_RHS_ONLY = 1 };

// Byte offsets of the output fields in one staging buffer (the VTK appended
// arrays, solver_writer.cpp:56-119): solved (int32), velocity (3 doubles per
// cell), pressure, temperature, density, the five residual components.
struct OutputLayout {
  int64_t solved, vel, p, T, rho, res[5], total;  // five residual components
};

inline OutputLayout output_layout(int64_t N) {
  OutputLayout L;
  L.solved = 0;
  L.vel = (4 * N + 7) / 8 * 8;
  L.p = L.vel + 24 * N;
  L.T = L.p + 8 * N;
  L.rho = L.T + 8 * N;
  for (int k = 0; k < 5; ++k) L.res[k] = L.rho + 8 * N * (k + 1);
  L.total = L.res[4] + 8 * N;
  return L;
}

// The closed-form grid of a descriptor, validated as mmf_create documents
// (MMF_OK or the status mmf_create would return); no flags, field stride =
// cells rounded up to 64.
int grid_from_desc(const mmf_grid_desc *d, GridDev &g);

// Kernel launchers (mmf_kernels.hip).  All enqueue on 'stream'.
hipError_t launch_stage(const GridDev &g, int stage, int mode, const double *U, const double *R, const double *W,
                        double *out_R, double *out_W, double *out_U, Ctrl *ctrl, hipStream_t stream);
hipError_t launch_stage_march(const GridDev &g, const MarchShape &shape, int stage, int mode, const double *U,
                              const double *R, const double *W, double *out_R, double *out_W, double *out_U,
                              Ctrl *ctrl, hipStream_t stream);
bool march_shape_supported(const GridDev &g, const MarchShape &shape);
// 3D order 2: the LDS-staged marching kernel (64 x 8 tiles), bodies included.
hipError_t launch_stage_march3d_o2(const GridDev &g, int chunk, int stage, int mode, const double *U, const double *R,
                                   const double *W, double *oR, double *oW, double *oU, Ctrl *ctrl, hipStream_t s);
// Tile table of the 3D marching kernel for cell flags 'flags' (cell order):
// masked tiles first, the rest in the XCD-aware order.
void build_march_tiles(const GridDev &g, const MarchShape &shape, const uint8_t *flags, int32_t *tiles);
// The segment table of shape.seg_blocks blocks (3 * MARCH_SEGS entries
// each): the tile planes split into shares of equal cost (a plane that needs
// the masked code costs more), XCD-aware, each share cut into runs of one
// tile and one code; per block its runs as (tile | MARCH_TILE_MASKED, first
// plane, end plane), -1 after the last.  flags (cell order) or nullptr.
// False when a share would need more than MARCH_SEGS runs of different tiles.
bool build_march_segments(const GridDev &g, const MarchShape &shape, const uint8_t *flags, int32_t *table);
// Sets shape.seg_blocks where the chunks leave a partial round of blocks on
// num_cus CUs and every share stays within a few tiles.
void balance_march_shape(const GridDev &g, int num_cus, MarchShape &shape);
int64_t march_blocks(const GridDev &g, const MarchShape &shape);
// The marching kernel takes dt from Ctrl::lam_now (early dt) on this grid:
// 3D order 1 without flags, or with flags where every fluid cell is solved.
bool march_early(const GridDev &g);
// Ctrl::lam_now from state U (early dt before a step sequence).
hipError_t launch_lam_init(const GridDev &g, const double *U, Ctrl *ctrl, hipStream_t stream);
MarchShape default_march_shape(const GridDev &g, int num_cus);
hipError_t launch_advance(Ctrl *ctrl, hipStream_t stream);
// AoS rows (raw order via cell_of_raw, identity when null) <-> SoA planes.
hipError_t launch_aos_to_soa(const double *aos, const int64_t *cell_of_raw, int64_t N, int64_t fs, double *soa,
                             hipStream_t stream);
hipError_t launch_soa_to_aos(const double *soa, const int64_t *cell_of_raw, int64_t N, int64_t fs, double *aos,
                             hipStream_t stream);
// Output fields of state U and residual R into 'stage' (layout L).
hipError_t launch_output_fields(const GridDev &g, const double *U, const double *R, const uint8_t *flags, char *stage,
                                const OutputLayout &L, hipStream_t stream);
hipError_t launch_zero_slot(Ctrl *ctrl, int slot, hipStream_t stream);
hipError_t launch_error_norms(const GridDev &g, const double *U, const double *rho_ex, double *partials,
                              int64_t nblocks, double *out2, hipStream_t stream);
int64_t error_norm_blocks(int64_t N);
hipError_t launch_mark_bodies(const GridDev &g, const double *origin, const double *boxes, int nboxes,
                              const double *spheres, int nspheres, uint8_t *flags, hipStream_t stream);
hipError_t launch_fill_double(double *p, int64_t n, double v, hipStream_t stream);
// One RK3 stage's update (stage 1: W = U + dt R / V; 2: W = 3/4 U + 1/4 (W +
// dt R / V); 3: U = 1/3 U + 2/3 (W + dt R / V)) and the step's close.
hipError_t launch_rk_update(const GridDev &g, int stage, double *U, const double *R, double *W, double dt,
                            hipStream_t stream);
hipError_t launch_rk_close(Ctrl *ctrl, double dt, const double *eig, hipStream_t stream);
// Rows of a SoA state at a list of cells (cell order) to / from AoS [n][5].
hipError_t launch_pack_cells(const double *X, int64_t fs, const int64_t *cells, int64_t n, double *buf,
                             hipStream_t stream);
hipError_t launch_unpack_cells(double *X, int64_t fs, const int64_t *cells, int64_t n, const double *buf,
                               hipStream_t stream);

// Batched point functions.
hipError_t launch_batch_cons2prim(const double *c, int64_t n, double *p, hipStream_t s);
hipError_t launch_batch_prim2cons(const double *p, int64_t n, double *c, hipStream_t s);
hipError_t launch_batch_eval_fluxes(const double *c, const double *p, const double *nrm, int64_t n, double *f,
                                    hipStream_t s);
hipError_t launch_batch_eval_splitting(const double *L, const double *R, const double *nrm, int64_t n, double *F,
                                       double *lam, hipStream_t s);
hipError_t launch_batch_bc_values(int bc, const double *nrm, const double *info, const double *c, int64_t n,
                                  double *out, hipStream_t s);
hipError_t launch_copy(const double4 *src, double4 *dst, int64_t n4, hipStream_t s);

}  // namespace mmf
