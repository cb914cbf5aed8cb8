/*
 * mmf_oracle.h -- CPU restatement (FP64, plain C99) of minimmerflow's
 * finite-volume time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker the HIP path is compared
 * against.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it; the product library never links or calls it.
 *
 * It follows the reference's own formulation: a generic interface loop over
 * an explicit face list (owner / neighbour / normal / area), conservative
 * AoS state [cell][5], the RK3 loop written inline in the driver.  Each
 * function cites the reference file:line it restates.  The bitpit VolOctree
 * geometry is replaced by closed forms for a uniform Cartesian grid.
 *
 * Compile with -ffp-contract=off so that every multiply and add rounds the
 * way the reference's source says.
 */
#ifndef MMF_ORACLE_H
#define MMF_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_NF 5

/* constants.hpp:37-47 -- primitive and conservative field order */
enum { ORC_P = 0, ORC_U = 1, ORC_V = 2, ORC_W = 3, ORC_T = 4 };
enum { ORC_RHO = 0, ORC_RHO_U = 1, ORC_RHO_V = 2, ORC_RHO_W = 3, ORC_RHO_E = 4 };

/* constants.hpp:58-62 -- boundary condition codes; PERIODIC is an extension */
enum {
    ORC_BC_NONE = -1,
    ORC_BC_FREE_FLOW = 0,
    ORC_BC_REFLECTING = 1,
    ORC_BC_WALL = 2,
    ORC_BC_DIRICHLET = 3,
    ORC_BC_PERIODIC = 4
};

/* problem.hpp:33-42 */
enum {
    ORC_PROBLEM_VORTEX_XY = 0,
    ORC_PROBLEM_VORTEX_ZX = 1,
    ORC_PROBLEM_VORTEX_YZ = 2,
    ORC_PROBLEM_RADSOD = 3,
    ORC_PROBLEM_SOD_X = 4,
    ORC_PROBLEM_SOD_Y = 5,
    ORC_PROBLEM_SOD_Z = 6,
    ORC_PROBLEM_FFSTEP = 7
};

/* Problem parameters.  Defaults (orc_problem_defaults) are the reference's
 * hard-wired values (problem.cpp:257-262, 336-342); the other fields are the
 * parametrised variants used by the larger workloads. */
typedef struct orc_problem {
    int type;
    int dim;
    /* isentropic vortex */
    double vortex_beta;
    double vortex_center[3];
    int vortex_periodic;          /* wrap the exact solution on a periodic box */
    double domain_origin[3];
    double domain_length;
    /* radial Sod */
    double sod_r0;
    double sod_p_in, sod_t_in;
    double sod_p_out, sod_t_out;
    /* uniform free stream (sphere body case) */
    double fs_prim[5];
} orc_problem;

void orc_problem_defaults(orc_problem *pb, int type, int dim);

/* ---- utils (utils.cpp) ---------------------------------------------- */
double orc_normal_velocity(const double *prim, const double *n);
void orc_cons2prim(const double *c, double *p);
void orc_prim2cons(const double *p, double *c);

/* ---- euler point functions (euler.cpp) -------------------------------- */
void orc_eval_fluxes(const double *cons, const double *prim, const double *n,
                     double *flux, int64_t *neg_p, int64_t *neg_rho);
void orc_eval_splitting(const double *consL, const double *consR, const double *n,
                        double *flux, double *lambda,
                        int64_t *neg_p, int64_t *neg_rho);
void orc_eval_bc_values(int bc, const double *normal, const double *info,
                        const double *cons, double *cons_bc);

/* ---- problem (problem.cpp) -------------------------------------------- */
void orc_exact_conservatives(const orc_problem *pb, const double *point, double t,
                             double *cons);
int orc_border_bc_type(const orc_problem *pb, const double *face_centroid);
void orc_border_bc_info(const orc_problem *pb, int bc, double *info);

/* ---- mesh (mesh_info.cpp closed forms + explicit face list) ----------- */
typedef struct orc_mesh {
    int dim;
    int64_t n[3];
    double origin[3];
    double h;
    int periodic[3];
    int64_t ncells;
    int64_t nfaces;
    double volume;               /* every cell */
    double area;                 /* every face */
    double *cell_centroid;       /* [ncells][3] */
    int64_t *face_owner;         /* [nfaces] */
    int64_t *face_neigh;         /* [nfaces], -1 on borders */
    double *face_normal;         /* [nfaces][3], owner-outward */
    double *face_centroid;       /* [nfaces][3] */
    int *face_axis;              /* [nfaces] */
    int *face_side;              /* [nfaces]: border side 0..5 (x-,x+,y-,y+,z-,z+), -1 interior */
    double *face_area;           /* [nfaces] on an explicit mesh, else NULL (uniform 'area') */
    double *cell_volume;         /* [ncells] on an explicit mesh, else NULL (uniform 'volume') */
} orc_mesh;

int orc_mesh_build(orc_mesh *m, int dim, const int64_t *n, const double *origin, double h,
                   const int *periodic);
void orc_mesh_free(orc_mesh *m);
int64_t orc_cell_index(const orc_mesh *m, int64_t i, int64_t j, int64_t k);

/* ---- solver state: flags, BCs, reconstruction, RHS, RK -------------- */
typedef struct orc_solver {
    orc_mesh mesh;
    orc_problem pb;
    int order;
    uint8_t *fluid;              /* [ncells] */
    uint8_t *solved;             /* [ncells] */
    int *face_bc;                /* [nfaces] */
    double *face_info;           /* [nfaces][5] Dirichlet primitive data */
    double *slopes;              /* [ncells][3][5], order 2 only */
    int64_t neg_p, neg_rho;      /* positivity violations seen in fluxes */
} orc_solver;

/* side_bc: per border side BC code (x-,x+,y-,y+,z-,z+); NULL -> problem's
 * getBorderBCType.  side_info: [6][5] Dirichlet primitives, NULL -> problem's
 * getBorderBCInfo.  fluid: [ncells] or NULL (all fluid). */
int orc_solver_init(orc_solver *s, const orc_problem *pb, int order,
                    const int64_t *n, const double *origin, double h,
                    const int *side_bc, const double *side_info, const uint8_t *fluid);
void orc_solver_free(orc_solver *s);

/* A solver on an explicit interface list (the general MeshGeometricalInfo,
 * mesh_info.cpp:86-118): per face owner, neighbour (-1 on borders), area,
 * owner-outward normal [3]; per cell volume (and centroid [3] or NULL);
 * border faces' BC codes and Dirichlet primitives [5] (NULL: free flow / no
 * data); min_cell_size for the dt rule.  Order 1; the face loop runs in the
 * given order. */
int orc_solver_init_faces(orc_solver *s, const orc_problem *pb, int dim, int64_t ncells, int64_t nfaces,
                          const double *cell_volume, const double *cell_centroid, double min_cell_size,
                          const int64_t *owner, const int64_t *neigh, const double *area, const double *normal,
                          const int *border_bc, const double *border_info, const uint8_t *fluid);

void orc_initial_state(const orc_solver *s, double *U);
void orc_compute_polynomials(orc_solver *s, const double *U);
void orc_reconstruct(const orc_solver *s, int64_t cell, const double *outward_normal,
                     const double *U, double *values);
int orc_compute_rhs(orc_solver *s, const double *U, double *RHS, double *max_eig);

/* One full SSP-RK3 step; returns dt used.  work: [ncells][5]. */
double orc_rk3_step(orc_solver *s, double *U, double *work, double *RHS,
                    double cfl, double t, double tmax, double *stage_eig);

/* Threaded forms (OpenMP), bit for bit the serial ones: the CPU baseline on
 * all host cores.  The face loop runs axis by axis over groups of faces whose
 * cells are disjoint, each group in increasing position, so every cell adds
 * its faces in the serial order. */
int orc_compute_rhs_threads(orc_solver *s, const double *U, double *RHS, double *max_eig, int nthreads);
double orc_rk3_step_threads(orc_solver *s, double *U, double *work, double *RHS,
                            double cfl, double t, double tmax, double *stage_eig, int nthreads);
/* Face-list solvers (orc_solver_init_faces): every face's A*F in parallel,
 * then every cell's sum in increasing face index in parallel (the serial
 * loop's order).  orc_rk3_step_threads uses it for them. */
int orc_compute_rhs_faces_threads(orc_solver *s, const double *U, double *RHS, double *max_eig, int nthreads);
/* Loop until t >= tmax; returns number of steps, final t in *t_out. */
int64_t orc_run(orc_solver *s, double *U, double cfl, double t0, double tmax,
                int64_t max_steps, double *t_out);

/* main.cpp:550-561 -- L1 density error (and the L2 extension). */
void orc_error(const orc_solver *s, const double *U, double t, double *l1, double *l2);

/* body.cpp:80-95 -- inclusive boxes; sphere level set is an extension. */
int orc_point_in_box(const double *box /*xmin,ymin,zmin,xmax,ymax,zmax*/, const double *p);
void orc_mark_fluid(const orc_mesh *m, const double *boxes, int nboxes,
                    const double *spheres /*cx,cy,cz,r*/, int nspheres, uint8_t *fluid);

/* ---- flat helpers for ctypes callers (tests, bench) ------------------ */
orc_solver *orc_solver_alloc(void);
void orc_solver_dealloc(orc_solver *s);
int64_t orc_solver_ncells(const orc_solver *s);
int64_t orc_solver_nfaces(const orc_solver *s);
void orc_solver_counters(const orc_solver *s, int64_t *neg_p, int64_t *neg_rho);
void orc_solver_cell_centroids(const orc_solver *s, double *out);
int orc_solver_set_solved(orc_solver *s, const uint8_t *solved);
void orc_solver_flags(const orc_solver *s, uint8_t *fluid, uint8_t *solved);
void orc_solver_face_bc(const orc_solver *s, int *out);
void orc_solver_face_list(const orc_solver *s, int64_t *owner, int64_t *neigh, double *normal, double *area);
void orc_solver_cell_volumes(const orc_solver *s, double *vol);

#ifdef __cplusplus
}
#endif
#endif
