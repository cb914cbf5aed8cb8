/*
 * mmf.h -- C-ABI of the MI355X finite-volume time step (minimmerflow_amd).
 *
 * Plain C types only: pointers, sizes, status codes.  No C++ exception and no
 * exit() crosses this boundary.  Every function returns MMF_OK (0) or an
 * MMF_ERR_* code; mmf_last_error(ctx) holds a message for the last failure on
 * that context.
 *
 * State layout at the boundary: host arrays of conservative variables in the
 * reference's AoS order [cell][5] = (rho, rho*u, rho*v, rho*w, rho*E)
 * (constants.hpp:43-47).  Cells are numbered i + nx*(j + ny*k) unless a raw
 * order is installed with mmf_set_raw_order().  On the device the state is
 * SoA (one plane per field) and stays resident between calls.
 *
 * One HIP stream per context.  Calls on one context must be serialised by
 * the caller; distinct contexts are independent.
 */
#ifndef MMF_H
#define MMF_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MMF_NF 5

/* status codes; MMF_ERR_UNSUPPORTED_ORDER replaces the reference's exit(2)
 * (reconstruction.cpp:75-76) */
enum {
    MMF_OK = 0,
    MMF_ERR_INVALID = 1,
    MMF_ERR_UNSUPPORTED_ORDER = 2,
    MMF_ERR_HIP = 3,
    MMF_ERR_NO_DEVICE = 4,
    MMF_ERR_OUT_OF_MEMORY = 5,
    MMF_ERR_STATE = 6
};

/* boundary condition codes: constants.hpp:58-62; PERIODIC is an extension */
enum {
    MMF_BC_NONE = -1,
    MMF_BC_FREE_FLOW = 0,
    MMF_BC_REFLECTING = 1,
    MMF_BC_WALL = 2,
    MMF_BC_DIRICHLET = 3,
    MMF_BC_PERIODIC = 4
};

/* cell flag bits (mmf_set_cell_flags) */
enum { MMF_CELL_FLUID = 1, MMF_CELL_SOLVED = 2 };

/*
 * Uniform Cartesian grid + boundary table: the closed-form replacement of
 * MeshGeometricalInfo (mesh_info.hpp:31-105) and of the per-face BC storage
 * built in computation() (main.cpp:251-277) with problem::getBorderBCType /
 * getBorderBCInfo (problem.cpp:404-477).
 *   cell volume = h^dim, face area = h^(dim-1), cell size = h,
 *   centroid = origin + (i+1/2) h, face normals = +-e_axis (owner-outward).
 * side order: x-, x+, y-, y+, z-, z+.  A periodic axis needs PERIODIC on
 * both of its sides.  side_info holds primitive (p,u,v,w,T) Dirichlet data.
 */
typedef struct mmf_grid_desc {
    int32_t dim;            /* 2 or 3 */
    int32_t order;          /* 1 (reference) or 2 (limited linear) */
    int64_t n[3];           /* cells per axis, 1 .. 2^30; n[2] ignored in 2D; at most 2^36 cells */
    double origin[3];
    double h;               /* cell edge length */
    int32_t side_bc[6];
    double side_info[6][MMF_NF];
} mmf_grid_desc;

typedef struct mmf_ctx mmf_ctx;

/* ---- lifetime (MeshGeometricalInfo ctor + storage allocation,
 *      main.cpp:197-219) -------------------------------------------------- */
int mmf_create(const mmf_grid_desc *desc, mmf_ctx **out);
int mmf_destroy(mmf_ctx *ctx);
const char *mmf_last_error(const mmf_ctx *ctx);
const char *mmf_status_string(int status);
/* cells; interfaces as the reference counts them (border faces included) */
int mmf_get_sizes(const mmf_ctx *ctx, int64_t *ncells, int64_t *ninterfaces);

/* ---- stage-kernel selection ------------------------------------------------ */
enum { MMF_KERNEL_AUTO = 0, MMF_KERNEL_GATHER = 1, MMF_KERNEL_MARCH = 2, MMF_KERNEL_TILE = 3 };
/* GATHER: per-cell kernel, any grid.  MARCH: the marching LDS-tile kernels
 * (order 1; order 2 in 2D without bodies).  TILE: the 8x8 LDS tile kernel of
 * 2D order-2 grids without bodies (each face once, the step's time advance
 * in its last stage; also the learned reconstruction's kernel).  AUTO: TILE
 * for 2D order-2 grids without bodies below 2^18 cells; MARCH where it
 * applies from 2^18 cells (2D order 1) or 2^20 cells (3D order 1, 2D order
 * 2); else GATHER.  mmf_set_kernel refuses MARCH or TILE on a grid they do not
 * apply to; bodies set afterwards make them fall back to GATHER.  tile_rows (3D: 4 or 8) and chunk (planes per block along the last axis)
 * shape the marching kernel; <= 0 keeps the default (3D order 1, 64 x 8
 * tiles: where the chunks would leave a partial round of blocks, the planes
 * of all tiles split over one block per CU in shares of equal cost, a plane
 * that needs the masked code near a body counting 2.5 times).  Every variant
 * gives bit-identical results. */
int mmf_set_kernel(mmf_ctx *ctx, int32_t kind, int32_t tile_rows, int32_t chunk);
/* 3D order 1, 64 x 8 tiles: the planes of all tiles split over 'blocks'
 * blocks of the marching kernel in shares of equal cost (between the tile
 * count and the tile planes; a block marches up to four segments, each of
 * one tile and one code, masked or not); 0 restores the default. */
int mmf_set_march_blocks(mmf_ctx *ctx, int32_t blocks);
/* kind actually used for the next launches, and the marching shape (chunk 0:
 * the planes split over the blocks in shares of equal cost) */
int mmf_get_kernel(const mmf_ctx *ctx, int32_t *kind, int32_t *tile_rows, int32_t *chunk);
/* The 3D marching kernel's tile (tile_x x tile_y cells: 64 x 4, 64 x 8,
 * 32 x 8 or 32 x 16) and chunk; <= 0 keeps the default of that field.  Does
 * not change the kernel kind.  Results do not depend on the shape. */
int mmf_set_march_shape(mmf_ctx *ctx, int32_t tile_x, int32_t tile_y, int32_t chunk);

/* ---- body / flags (body::isPointFluid + main.cpp:222-237, 251-277) ------- */
/* fluid: [ncells] 0/1 or NULL (all fluid); solved: NULL -> solved = fluid */
int mmf_set_cell_flags(mmf_ctx *ctx, const uint8_t *fluid, const uint8_t *solved);
/* Marks cells on the device from their centroids: boxes [nb][6]
 * (xmin,ymin,zmin,xmax,ymax,zmax; inclusive, body.cpp:80-95), spheres
 * [ns][4] (cx,cy,cz,r; solid where |x-c| - r < 0).  solved = fluid. */
int mmf_set_bodies(mmf_ctx *ctx, const double *boxes, int32_t nboxes,
                   const double *spheres, int32_t nspheres);
/* flags: [ncells] MMF_CELL_* bits */
int mmf_get_cell_flags(mmf_ctx *ctx, uint8_t *flags);

/* ---- border conditions per border face (main.cpp:251-260) ---------------- */
/* The reference keeps a BC code per border interface (interfaceBCs,
 * main.cpp:251-260, from problem::getBorderBCType at the face centroid) and
 * Dirichlet data per face (euler.cpp:261-268).  Entries side by side: x-, x+,
 * y-, y+ (, z-, z+); a side's faces ordered by its two other axes, the lower
 * one fastest (side y- of a 3D grid: i + n_x k).  mmf_border_face_count gives
 * the entry count of a descriptor's grid.  bc: MMF_BC_FREE_FLOW ..
 * MMF_BC_DIRICHLET; info: [count][5] Dirichlet primitives (p,u,v,w,T), read
 * for Dirichlet entries only, may be NULL when there is none.  Entries of
 * sides on periodic axes are ignored.  A side whose entries agree becomes a
 * per-side condition; with a side that does not (at most 256 distinct
 * conditions), the stage kernel is the per-cell gather kernel, the tile,
 * marching and learned kernels do not apply. */
int mmf_border_face_count(const mmf_grid_desc *desc, int64_t *count);
int mmf_set_border_bcs(mmf_ctx *ctx, const int32_t *bc, const double *info);

/* ---- state transfer (AoS host <-> SoA device) --------------------------- */
int mmf_set_state(mmf_ctx *ctx, const double *U_aos);
int mmf_get_state(mmf_ctx *ctx, double *U_aos);
/* Optional: cell_of_raw[r] = structured index of the reference's raw cell r
 * (bitpit PiercedStorage raw order).  After this, mmf_set_state /
 * mmf_get_state / mmf_compute_rhs read and write host arrays in raw order. */
int mmf_set_raw_order(mmf_ctx *ctx, const int64_t *cell_of_raw, int64_t nraw);

/* ---- euler::computeRHS (euler.cpp:127-249) ------------------------------- */
/* RHS of the current state (order-2 slopes included: the
 * reconstruction::computePolynomials hook, reconstruction.cpp:47-55).
 * rhs_aos: host [ncells][5] or NULL; max_eig: host scalar or NULL. */
int mmf_compute_rhs(mmf_ctx *ctx, double *rhs_aos, double *max_eig);
/* euler::computeRHS with the arguments the reference passes on every call
 * (euler.hpp:41-43): solved [ncells] 0/1 (host order; NULL keeps the
 * context's), order 1 or 2 (else MMF_ERR_UNSUPPORTED_ORDER, the reference's
 * exit(2), reconstruction.cpp:75-76), U_aos the state (NULL: the current
 * one).  Flags and order are re-sent only where they differ from the
 * context's, so a caller may change them between calls (ghost cells
 * unsolved, main.cpp:231-235).  The border conditions (interfaceBCs) are
 * the context's: mmf_set_border_bcs; fluid-solid faces follow the fluid
 * flags (main.cpp:271-275). */
int mmf_compute_rhs_ex(mmf_ctx *ctx, const uint8_t *solved, int32_t order, const double *U_aos,
                       double *rhs_aos, double *max_eig);

/* ---- one RK3 stage at a time (main.cpp:377-506) ----------------------------- */
/* For a caller that runs something between the stages of a step: the ghost
 * exchange of a partitioned mesh (main.cpp:425-430, 461-466, 497-502) or the
 * reduction of the eigenvalue over ranks (main.cpp:390-395).  One step is
 *   mmf_rk_rhs(ctx, MMF_STATE_U, &e);  [reduce e over ranks]
 *   mmf_rk_dt(ctx, e, &dt);             main.cpp:398-402
 *   mmf_rk_update(ctx, 1, dt);          W = U + dt R / V        [exchange W]
 *   mmf_rk_rhs(ctx, MMF_STATE_W, &e);
 *   mmf_rk_update(ctx, 2, dt);          W = 3/4 U + 1/4 (W + dt R / V)  [exchange W]
 *   mmf_rk_rhs(ctx, MMF_STATE_W, &e);
 *   mmf_rk_update(ctx, 3, dt);          U = 1/3 U + 2/3 (W + dt R / V), t += dt  [exchange U]
 * and gives the bits of mmf_advance(ctx, 1) on the same grid (the fused
 * kernels evaluate the same expressions).  Unsolved cells keep U; stages 1
 * and 2 copy it into W. */
enum { MMF_STATE_U = 0, MMF_STATE_W = 1, MMF_STATE_RHS = 2 };
/* R = RHS(state) (state U or W); max_eig: the stage maximum (host) */
int mmf_rk_rhs(mmf_ctx *ctx, int32_t state, double *max_eig);
/* dt of the step from max_eig (MMF_ERR_STATE unless finite and positive) */
int mmf_rk_dt(mmf_ctx *ctx, double max_eig, double *dt);
int mmf_rk_update(mmf_ctx *ctx, int32_t stage, double dt);
/* Ghost lists (communications.tpp:164-199): rows cells[n] (host order, as
 * mmf_set_state) of a state to / from host AoS buf [n][5]. */
int mmf_pack_cells(mmf_ctx *ctx, int32_t state, const int64_t *cells, int64_t n, double *buf);
int mmf_unpack_cells(mmf_ctx *ctx, int32_t state, const int64_t *cells, int64_t n, const double *buf);
/* Device SoA view of a state (field k at ptr + k * field_stride, cell order),
 * for exchanges that stay on the device. */
int mmf_get_device_state(mmf_ctx *ctx, int32_t state, double **ptr, int64_t *field_stride);

/* ---- the same RHS, allocating nothing -------------------------------------- */
/* No counterpart in the reference: a variant of mmf_compute_rhs for callers
 * that cannot allocate device memory (code under HIP graph capture, serving
 * loops, harnesses that forbid hipMalloc).  It creates no context and calls
 * no hipMalloc / hipFree: the stage-maximum lines and positivity counters
 * live in a library-owned device array of 12,672 bytes (its maximum size),
 * one per GPU, held until the process ends.
 *   desc      the grid (validated as in mmf_create; desc->order 1 or 2)
 *   flags     device [ncells] MMF_CELL_* bits, or NULL (every cell fluid and
 *             solved)
 *   U         device SoA state [5][field_stride] (conservative variables)
 *   rhs       device SoA residual [5][field_stride], written (not aliasing U)
 *   max_eig   device double or NULL: the maximum face eigenvalue
 *   counts    device int64[2] or NULL: negative pressure / density face sides
 *   stream    a hipStream_t, or NULL for the null stream
 * field_stride must be >= ncells and a multiple of 64.  Enqueues on 'stream'
 * and returns without synchronising; results are bitwise those of
 * mmf_compute_rhs on the same grid, flags and state.  Calls on one GPU share
 * the array: calls that overlap on one GPU (unordered streams) race.  Errors
 * (MMF_ERR_INVALID for a descriptor or stride beyond the limits) leave a
 * message for mmf_noalloc_last_error() in the calling thread. */
int mmf_compute_rhs_noalloc(const mmf_grid_desc *desc, const uint8_t *flags, const double *U,
                            int64_t field_stride, double *rhs, double *max_eig, int64_t *counts,
                            void *stream);
const char *mmf_noalloc_last_error(void);

/* ---- RK3 time loop (main.cpp:374-506) ------------------------------------ */
/* dt = 0.9 * cfl * h / maxEig(stage 1), clipped to tmax (main.cpp:398-402). */
int mmf_set_time(mmf_ctx *ctx, double t, double tmax, double cfl);
/* Enqueue nsteps SSP-RK3 steps; dt and t stay on the device and a step
 * whose start time is >= tmax does nothing.  Does not synchronise. */
int mmf_advance(mmf_ctx *ctx, int64_t nsteps);
/* Steps are replayed from HIP graphs captured on first use (re-captured
 * when the kernel choice, bodies or reconstruction change); on 0 they are
 * launched one by one.  Results are identical either way. */
int mmf_set_step_graphs(mmf_ctx *ctx, int32_t on);
/* Step until t >= tmax (or max_steps when >= 0); synchronises.
 * Non-finite states: a step whose stage maxima (the face eigenvalues, whose
 * running maximum keeps a NaN) are NaN or infinite leaves t, dt and the step
 * count unchanged and is counted; mmf_run then stops and returns
 * MMF_ERR_STATE (steps_done and t_final filled), as does mmf_get_time.  The
 * reference has no such check: it logs negative p / rho (euler.cpp:94-101)
 * and its t would become NaN (main.cpp:398-402, 505). */
int mmf_run(mmf_ctx *ctx, int64_t max_steps, int64_t *steps_done, double *t_final);
/* Synchronises and reads the time control block.  stage_eig: the three
 * stage maxima of the last step (main.cpp:389,433,469) or NULL.  Returns
 * MMF_ERR_STATE (outputs filled) once a step met a non-finite state. */
int mmf_get_time(mmf_ctx *ctx, double *t, double *dt, int64_t *step, double *stage_eig);
/* steps that met a non-finite state since mmf_set_time or
 * mmf_reset_positivity_counters (synchronises) */
int mmf_get_nonfinite_steps(mmf_ctx *ctx, int64_t *steps);
int mmf_synchronize(mmf_ctx *ctx);

/* ---- output (SolverWriter, solver_writer.cpp:56-119; the writes of
 *      main.cpp:284-298, 346, 509-540) ------------------------------------- */
/* VTK XML image data (.vti) of the current state: cell fields solved, velocity,
 * pressure, temperature, density and residualC/MX/MY/MZ/E, the reference's
 * names and order.  The conversion runs on the GPU; the residuals are the RHS
 * of the written state (the reference writes the last stage's).  Returns once
 * the file is written. */
int mmf_write_vtk(mmf_ctx *ctx, const char *path);
/* The same, returning at once: the fields are converted and copied on the
 * context's stream (later steps queue behind the copy) and a host thread
 * writes the file meanwhile.  Two writes are in flight at most; a third call
 * waits for the older one. */
int mmf_write_vtk_async(mmf_ctx *ctx, const char *path);
/* Waits for every pending write; reports the first failed one since the last
 * wait (MMF_ERR_STATE, the path in mmf_last_error). */
int mmf_wait_writes(mmf_ctx *ctx);

/* ---- diagnostics ---------------------------------------------------------- */
/* main.cpp:550-561: sum |rho - rho_exact| V over all cells; l2 =
 * sqrt(sum (rho - rho_exact)^2 V) (extension).  rho_exact: host [ncells]. */
int mmf_error_norms(mmf_ctx *ctx, const double *rho_exact, double *l1, double *l2);
/* euler.cpp:94-101 logs negative pressure / density per flux evaluation;
 * here they are counted (per interface side, as the reference evaluates). */
int mmf_get_positivity_counters(mmf_ctx *ctx, int64_t *neg_p, int64_t *neg_rho);
int mmf_reset_positivity_counters(mmf_ctx *ctx);
/* Kernel timing with HIP events on the context's stream.  on = 1: every
 * stage-kernel launch is bracketed by an event pair (steps launched one by
 * one); on = 2: one event pair brackets each mmf_advance call, whose steps
 * are still replayed from graphs, and counts its three stage launches per
 * step (the time then includes the gaps between them and, for kernels whose
 * step ends with a separate time update, that small kernel); 0: off. */
int mmf_enable_timing(mmf_ctx *ctx, int32_t on);
/* total device milliseconds and stage-launch count since timing was enabled
 * (synchronises) */
int mmf_read_timing(mmf_ctx *ctx, double *stage_ms, int64_t *stage_launches);

/* ---- point functions, batched on the device (host arrays in/out) --------- */
/* utils::conservative2primitive / primitive2conservative (utils.cpp:48-83) */
int mmf_batch_cons2prim(const double *cons, int64_t count, double *prim);
int mmf_batch_prim2cons(const double *prim, int64_t count, double *cons);
/* euler::evalFluxes (euler.cpp:83-113): normals [count][3] */
int mmf_batch_eval_fluxes(const double *cons, const double *prim, const double *normals,
                          int64_t count, double *flux);
/* euler::evalSplitting (euler.cpp:42-73) */
int mmf_batch_eval_splitting(const double *consL, const double *consR, const double *normals,
                             int64_t count, double *flux, double *lambda);
/* euler::evalInterfaceBCValues (euler.cpp:261-376); info [count][5] or NULL */
int mmf_batch_eval_bc_values(int32_t bc, const double *normals, const double *info,
                             const double *cons, int64_t count, double *cons_bc);

/* ---- problem catalogue (problem.cpp), host-side setup -------------------- */
/* problem.hpp:33-42; FREESTREAM (uniform state, Dirichlet inflow at x-) is an
 * extension for the immersed-body case */
enum {
    MMF_PROBLEM_VORTEX_XY = 0,
    MMF_PROBLEM_VORTEX_ZX = 1,
    MMF_PROBLEM_VORTEX_YZ = 2,
    MMF_PROBLEM_RADSOD = 3,
    MMF_PROBLEM_SOD_X = 4,
    MMF_PROBLEM_SOD_Y = 5,
    MMF_PROBLEM_SOD_Z = 6,
    MMF_PROBLEM_FFSTEP = 7,
    MMF_PROBLEM_FREESTREAM = 8
};

/* Defaults are the reference's hard-wired values (problem.cpp:70-149,
 * 257-262, 336-342, 383-387); the other fields parametrise the variants. */
typedef struct mmf_problem {
    int32_t type;
    int32_t dim;
    double vortex_beta;          /* problem.cpp:257 */
    double vortex_center[3];
    int32_t vortex_periodic;     /* exact solution wrapped on a periodic box */
    double domain_origin[3];
    double domain_length;
    double sod_r0;               /* problem.cpp:336 */
    double sod_p_in, sod_t_in;
    double sod_p_out, sod_t_out;
    double fs_prim[MMF_NF];      /* free stream / Dirichlet primitives */
} mmf_problem;

int mmf_problem_defaults(mmf_problem *pb, int32_t type, int32_t dim);
/* problem.cpp:175-211 */
int mmf_problem_default_end_time(const mmf_problem *pb, double *tend);
/* problem::evalExactConservatives (problem.cpp:247-392) at count points */
int mmf_problem_exact(const mmf_problem *pb, const double *points, int64_t count, double t,
                      double *cons);
/* problem::getBorderBCType / getBorderBCInfo (problem.cpp:404-477) */
int mmf_problem_border_bc(const mmf_problem *pb, const double *face_centroid, int32_t *bc,
                          double *info);
/* side_bc / side_info of desc from the problem's border rule, evaluated at
 * one face per side (the side's centre; FFSTEP, whose rule depends on x only:
 * a face of the first column).  Exact for every problem whose rule is uniform
 * per side; FFSTEP's y and z sides change condition at x = 3.2 (problem.cpp:
 * 422-432), which mmf_problem_border_bcs + mmf_set_border_bcs represent. */
int mmf_problem_side_bcs(const mmf_problem *pb, mmf_grid_desc *desc);
/* problem.cpp:404-477 at the centroid of every border face of desc's grid
 * (main.cpp:251-260), in mmf_set_border_bcs's layout; info may be NULL */
int mmf_problem_border_bcs(const mmf_problem *pb, const mmf_grid_desc *desc, int32_t *bc,
                           double *info);
/* main.cpp:335-344: upload the exact solution at t=0 sampled at centroids */
int mmf_set_initial_state(mmf_ctx *ctx, const mmf_problem *pb);
/* main.cpp:550-561 against the problem's exact density at time t */
int mmf_error_vs_exact(mmf_ctx *ctx, const mmf_problem *pb, double t, double *l1, double *l2);
/* cell centroids [ncells][3] (origin + (i+1/2) h; z = origin z in 2D) in the
 * current host order */
int mmf_get_centroids(const mmf_ctx *ctx, double *xyz);

/* ---- batched mode: many independent 2D order-1 grids ----------------------- */
/* The reference runs one grid per process; a batch advances ncases uneven 2D
 * grids together, each with the reference's own time loop (main.cpp:386-512:
 * its own dt from its own stage-1 maximum eigenvalue, its own end time).
 * One step is three persistent stage launches plus the per-case time update,
 * captured once into a HIP graph and replayed.  Cases are order 1, with no
 * immersed bodies; border BCs as in the grid descriptor, periodic allowed.  Per
 * case the results are bitwise those of a single-grid mmf_ctx. */
typedef struct mmf_batch mmf_batch;

int mmf_batch_create(const mmf_grid_desc *cases, int32_t ncases, mmf_batch **out);
int mmf_batch_destroy(mmf_batch *b);
const char *mmf_batch_last_error(const mmf_batch *b);
/* total cells over all cases and the number of (case, strip, row chunk) work items */
int mmf_batch_sizes(const mmf_batch *b, int64_t *total_cells, int32_t *nitems);
/* conservative state of case ci, host AoS [ncells][5], cells x-fastest */
int mmf_batch_set_state(mmf_batch *b, int32_t ci, const double *cons);
int mmf_batch_get_state(mmf_batch *b, int32_t ci, double *cons);
/* problems[ncases]: exact solution at t = 0 at the cell centroids (main.cpp:335-344) */
int mmf_batch_set_initial_states(mmf_batch *b, const mmf_problem *problems);
/* per-case start time, end time and CFL (arrays of ncases) */
int mmf_batch_set_time(mmf_batch *b, const double *t0, const double *tmax, const double *cfl);
/* rows per work item and persistent grid size of stages 2 and 3 (<= 0 keeps
 * the current value; stage 1, resident at three waves per SIMD where they
 * are at two, runs proportionally more blocks).  The items are cut into one
 * queue per XCD, runs of neighbouring strips and row bands of equal cost; a
 * block takes items from its XCD's queue, then from the others. */
int mmf_batch_set_shape(mmf_batch *b, int32_t chunk, int32_t blocks);
int mmf_batch_get_shape(const mmf_batch *b, int32_t *chunk, int32_t *blocks);
/* rows of the small items that end each stage: the last 'chunk' rows of
 * every strip split into items of this many rows, queued after the full
 * items (0: none).  Results do not depend on it. */
int mmf_batch_set_tail(mmf_batch *b, int32_t rows);
int mmf_batch_get_tail(const mmf_batch *b, int32_t *rows);
/* enqueue nsteps RK3 steps (graph replays); cases past their end time idle */
int mmf_batch_advance(mmf_batch *b, int64_t nsteps);
/* steps until every case reaches its end time (max_steps < 0: no limit) */
int mmf_batch_run(mmf_batch *b, int64_t max_steps, int64_t *steps_done);
/* per-case time and completed step count (synchronises); either may be NULL.
 * mmf_batch_run and mmf_batch_get_times return MMF_ERR_STATE once a case met
 * a non-finite state (its time then stays where it was; see mmf_run). */
int mmf_batch_get_times(mmf_batch *b, double *t, int64_t *steps);
/* main.cpp:550-561 per case against problems[i] at the case's current time */
int mmf_batch_error_vs_exact(mmf_batch *b, const mmf_problem *problems, double *l1, double *l2);
/* graph-replay timing with HIP events (whole step: three stages + update) */
int mmf_batch_enable_timing(mmf_batch *b, int32_t on);
int mmf_batch_read_timing(mmf_batch *b, double *ms, int64_t *steps);
/* Launch-shape tuner: times candidate (chunk, tail, blocks) triples over
 * steps_per_trial steps each on a snapshot of the state, restores the state
 * and keeps the fastest.  Results do not depend on the shape. */
int mmf_batch_autotune(mmf_batch *b, int32_t steps_per_trial, int32_t *best_chunk, int32_t *best_blocks,
                       double *best_ms_per_step);

/* ---- learned face reconstruction (2D, order 2) ----------------------------- */
/* A second form beside the plain drop-in, with no counterpart in the
 * reference: the face values of the limited linear polynomial
 * (reconstruction.cpp:67-97) are replaced by the output of a small network
 * (5x5 stencil of rho, rho u, rho v, rho E, normalised per cell -> 64 -> 64
 * (GELU, tanh form) -> 16 outputs z), run in bf16 on the matrix cores with fp32
 * accumulation inside the reconstruction kernel.  The network is a learned
 * limiter: per face and component the value is P + sigmoid(z) (Q - P), P the
 * minmod face value and Q the fifth-order upwind-biased linear value along
 * the face's axis, so every learned value is limited and upwind.  The plain
 * face values remain wherever the stencil crosses a non-periodic border, the
 * shock sensor (normalised second difference of rho or p) exceeds its
 * threshold, or a learned face state has non-positive density or pressure.
 * Training (fp32 forward, backward and Adam kernels) uses rows built on the
 * device from fine runs restricted by conservative averaging; targets are
 * the fine plain face values averaged over each coarse face. */
#define MMF_LEARN_NPARAMS 11664
#define MMF_LEARN_NIN 100
#define MMF_LEARN_NOUT 16
typedef struct mmf_learn mmf_learn;

/* network with U(-1/sqrt(fan_in), 1/sqrt(fan_in)) weights from SplitMix64(seed),
 * zero biases, and room for 'capacity' training rows */
int mmf_learn_create(uint64_t seed, int64_t capacity, mmf_learn **out);
int mmf_learn_destroy(mmf_learn *net);
const char *mmf_learn_last_error(const mmf_learn *net);
/* parameters W1[64][100], b1[64], W2[64][64], b2[64], W3[16][64], b3[16] (fp32);
 * setting them resets the Adam moments */
int mmf_learn_get_params(const mmf_learn *net, float *params);
int mmf_learn_set_params(mmf_learn *net, const float *params);
/* appends 8 rows (the dihedral images) per listed coarse cell of the fine
 * context's current state restricted by 'factor' (1..8); rows whose stencil
 * crosses a non-periodic border or whose sensor exceeds the threshold get
 * weight 0 */
int mmf_learn_collect(mmf_learn *net, mmf_ctx *fine, int32_t factor, const int64_t *coarse_cells, int32_t ncells,
                      double sensor_threshold);
int mmf_learn_num_samples(const mmf_learn *net, int64_t *count);
/* rows x[count][100], targets t[count][16], weights w[count]; any may be NULL on get */
int mmf_learn_get_samples(const mmf_learn *net, int64_t first, int64_t count, float *x, float *t, float *w);
int mmf_learn_set_samples(mmf_learn *net, const float *x, const float *t, const float *w, int64_t count);
/* loss sum w (yhat - t)^2 / (16 count), yhat = P + sigmoid(z) (Q - P) in
 * normalised units, and its gradient on the given rows (the training kernels,
 * no update) */
int mmf_learn_loss_grad(mmf_learn *net, const float *x, const float *t, const float *w, int64_t count,
                        double *loss, float *grad);
/* one Adam update (beta 0.9 / 0.999, eps 1e-8, bias-corrected) with a given gradient */
int mmf_learn_adam_step(mmf_learn *net, const float *grad, double lr);
/* 'steps' Adam steps on batches of 'batch' rows drawn from the training set,
 * learning rate cosine-annealed from lr0 to lr1; final_loss = mean of the
 * last 50 batch losses */
int mmf_learn_train(mmf_learn *net, int32_t steps, int32_t batch, double lr0, double lr1, uint64_t seed,
                    double *final_loss);
/* raw network outputs z[count][16]: fp32 SIMT (use_mfma 0) or the bf16
 * matrix-core path */
int mmf_learn_forward(mmf_learn *net, const float *x, int64_t count, int32_t use_mfma, float *y);
/* switch a 2D order-2 context without bodies to the learned face values
 * (enable 1) or back (enable 0).  net NULL keeps the plain face values
 * everywhere through the same face-state kernels.  The network must outlive
 * the context's use of it. */
int mmf_set_learned_reconstruction(mmf_ctx *ctx, const mmf_learn *net, double sensor_threshold, int32_t enable);
/* optional safeguards on the learned face values, applied before the
 * positivity check: clip (default 0) bounds each component between the means
 * of the two cells sharing the face; blend (default 1) returns plain + blend
 * (learned - plain) */
int mmf_set_learned_options(mmf_ctx *ctx, int32_t clip, double blend);
/* cells reconstructed by the network / on the fallback, summed over stage-1
 * launches since the last reset */
int mmf_get_learned_counts(mmf_ctx *ctx, int64_t *learned_cells, int64_t *fallback_cells);
int mmf_reset_learned_counts(mmf_ctx *ctx);
/* face states [ncells][4][5] (faces x-, x+, y-, y+; cells x-fastest) of the
 * current state, as the stage-1 reconstruction produces them */
int mmf_get_face_states(mmf_ctx *ctx, double *faces);
/* conservative average of the current 2D state over factor x factor blocks,
 * host AoS [ncells / factor^2][5] */
int mmf_restrict_state(mmf_ctx *ctx, int32_t factor, double *coarse);

/* ---- general mesh: an explicit interface list ---------------------------- */
/* The general MeshGeometricalInfo (mesh_info.hpp:31-105, cached by _extract,
 * mesh_info.cpp:86-118) handed over as arrays, for meshes the uniform
 * descriptor cannot express (bitpit's octree with hanging faces, any
 * polyhedral cell list).  Cells and faces are numbered by the caller (raw
 * order); the interface loop of euler::computeRHS (euler.cpp:153-247) runs in
 * the given face order, so each cell's residual is summed in that order and
 * the results equal a serial interface loop over the same list bit for bit.
 * Order 1 (reconstruction.cpp:90-97).  Borders: faces with neighbour -1 take
 * face_bc (FREE_FLOW, REFLECTING, WALL or DIRICHLET with face_info
 * primitives: problem::getBorderBCType / getBorderBCInfo evaluated by the
 * caller); interior faces become walls where exactly one side is fluid
 * (main.cpp:251-277).  dt = 0.9 cfl min_cell_size / maxEig (main.cpp:351-356,
 * 398); the RK update divides by each cell's volume (main.cpp:409-495). */
typedef struct mmf_mesh_desc {
    int32_t dim;                 /* 2 or 3 (informational: normals carry 3 components) */
    int64_t ncells;              /* < 2^31 */
    int64_t nfaces;              /* < 2^30 */
    const double *cell_volume;   /* [ncells] rawGetCellVolume */
    double min_cell_size;        /* min over cells of rawGetCellSize */
    const int64_t *face_owner;   /* [nfaces] Interface::getOwner as a cell index */
    const int64_t *face_neigh;   /* [nfaces] Interface::getNeigh as a cell index, -1 on borders */
    const double *face_area;     /* [nfaces] rawGetInterfaceArea */
    const double *face_normal;   /* [nfaces][3] rawGetInterfaceNormal (owner-outward, unit) */
    const int32_t *face_bc;      /* [nfaces] border faces' BC codes, or NULL (free flow) */
    const double *face_info;     /* [nfaces][5] Dirichlet primitives (p,u,v,w,T), or NULL */
} mmf_mesh_desc;

typedef struct mmf_mesh mmf_mesh;

int mmf_mesh_create(const mmf_mesh_desc *desc, mmf_mesh **out);
int mmf_mesh_destroy(mmf_mesh *m);
const char *mmf_mesh_last_error(const mmf_mesh *m);
int mmf_mesh_sizes(const mmf_mesh *m, int64_t *ncells, int64_t *nfaces);
/* fluid: [ncells] 0/1 or NULL; solved: NULL -> solved = fluid (main.cpp:222-237) */
int mmf_mesh_set_cell_flags(mmf_mesh *m, const uint8_t *fluid, const uint8_t *solved);
/* host AoS [ncells][5] conservative states */
int mmf_mesh_set_state(mmf_mesh *m, const double *U_aos);
int mmf_mesh_get_state(mmf_mesh *m, double *U_aos);
/* euler::computeRHS (euler.cpp:127-249) of the current state */
int mmf_mesh_compute_rhs(mmf_mesh *m, double *rhs_aos, double *max_eig);
/* the RK3 time loop, as mmf_set_time / mmf_advance / mmf_run / mmf_get_time */
int mmf_mesh_set_time(mmf_mesh *m, double t, double tmax, double cfl);
int mmf_mesh_advance(mmf_mesh *m, int64_t nsteps);
int mmf_mesh_set_step_graphs(mmf_mesh *m, int32_t on);
int mmf_mesh_run(mmf_mesh *m, int64_t max_steps, int64_t *steps_done, double *t_final);
int mmf_mesh_get_time(mmf_mesh *m, double *t, double *dt, int64_t *step, double *stage_eig);
/* negative p / rho face sides (euler.cpp:94-101) and steps that met a
 * non-finite state */
int mmf_mesh_get_counters(mmf_mesh *m, int64_t *neg_p, int64_t *neg_rho, int64_t *nonfinite_steps);
/* HIP-event timing of the face and cell kernels of every stage */
int mmf_mesh_enable_timing(mmf_mesh *m, int32_t on);
int mmf_mesh_read_timing(mmf_mesh *m, double *face_ms, double *cell_ms, int64_t *stages);

/* ---- bandwidth calibration ------------------------------------------------ */
/* Copies nbytes device-to-device with a plain vector-load/store kernel reps
 * times; returns the average kernel milliseconds (HIP events). */
int mmf_calibration_copy(int64_t nbytes, int32_t reps, double *ms_per_copy);

#ifdef __cplusplus
}
#endif
#endif /* MMF_H */
