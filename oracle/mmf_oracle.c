/*
 * mmf_oracle.c -- CPU restatement of minimmerflow's finite-volume time step.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Never linked into the product.
 *
 * Every expression keeps the reference's operand order; build with
 * -ffp-contract=off so no multiply-add is fused.  Citations are to the
 * reference tree (src/...).
 *
 * Parity status of each part (see tests/test_oracle_*.py):
 *   utils, evalFluxes, evalSplitting, BC values, computeRHS, RK3, dt rule,
 *   L1 error, ICs: pinned end-to-end by the reference's five known-answer
 *   tests (test/<case>/CMakeLists.txt:33) and unit-pinned by closed-form
 *   invariants (round trips, consistency, antisymmetry, mirror symmetry,
 *   uniform-state zero residual, conservation).
 *   Order-2 limited reconstruction, periodic wrap, sphere mask, L2 error and
 *   the free-stream problem are extensions with no reference counterpart:
 *   pinned by invariants only (order 2 reduces to order 1 under a flat
 *   limiter, is exact on linear data, converges at ~2nd order).
 */
#include "mmf_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* constants.hpp:55 */
static const double ORC_GAMMA = 1.4;

/* main.cpp:417 rawGetCellVolume: per cell on an explicit mesh, else uniform */
#define ORC_VOL(m, c) ((m)->cell_volume ? (m)->cell_volume[(c)] : (m)->volume)

/* ------------------------------------------------------------------ */
/* utils                                                               */
/* ------------------------------------------------------------------ */

/* utils.cpp:37-40 */
double orc_normal_velocity(const double *p, const double *n)
{
    return (p[ORC_U] * n[0] + p[ORC_V] * n[1] + p[ORC_W] * n[2]);
}

/* utils.cpp:48-63 */
void orc_cons2prim(const double *c, double *p)
{
    double K = (c[ORC_RHO_U] * c[ORC_RHO_U] + c[ORC_RHO_V] * c[ORC_RHO_V] +
                c[ORC_RHO_W] * c[ORC_RHO_W]) / (c[ORC_RHO] * c[ORC_RHO]);
    p[ORC_T] = (2.0 * c[ORC_RHO_E] / c[ORC_RHO] - K) / (2.0 / (ORC_GAMMA - 1.0));
    p[ORC_U] = c[ORC_RHO_U] / c[ORC_RHO];
    p[ORC_V] = c[ORC_RHO_V] / c[ORC_RHO];
    p[ORC_W] = c[ORC_RHO_W] / c[ORC_RHO];
    p[ORC_P] = c[ORC_RHO] * p[ORC_T];
}

/* utils.cpp:71-83 */
void orc_prim2cons(const double *p, double *c)
{
    c[ORC_RHO] = p[ORC_P] / p[ORC_T];
    c[ORC_RHO_U] = c[ORC_RHO] * p[ORC_U];
    c[ORC_RHO_V] = c[ORC_RHO] * p[ORC_V];
    c[ORC_RHO_W] = c[ORC_RHO] * p[ORC_W];
    c[ORC_RHO_E] = c[ORC_RHO] * p[ORC_T] / (ORC_GAMMA - 1.0) +
                   0.5 * c[ORC_RHO] * (p[ORC_U] * p[ORC_U] + p[ORC_V] * p[ORC_V] + p[ORC_W] * p[ORC_W]);
}

/* ------------------------------------------------------------------ */
/* euler point functions                                               */
/* ------------------------------------------------------------------ */

/* euler.cpp:83-113.  The reference logs negative p / rho and continues
 * (euler.cpp:94-101); here they are counted. */
void orc_eval_fluxes(const double *cons, const double *prim, const double *n,
                     double *f, int64_t *neg_p, int64_t *neg_rho)
{
    double u = prim[ORC_U];
    double v = prim[ORC_V];
    double w = prim[ORC_W];

    double vel2 = u * u + v * v + w * w;
    double un = orc_normal_velocity(prim, n);

    double p = prim[ORC_P];
    if (p < 0. && neg_p) (*neg_p)++;

    double rho = cons[ORC_RHO];
    if (rho < 0. && neg_rho) (*neg_rho)++;

    double eto = p / (ORC_GAMMA - 1.) + 0.5 * rho * vel2;

    double massFlux = rho * un;

    f[0] = massFlux;
    f[1] = massFlux * u + p * n[0];
    f[2] = massFlux * v + p * n[1];
    f[3] = massFlux * w + p * n[2];
    f[4] = un * (eto + p);
}

/* euler.cpp:42-73 (local Lax-Friedrichs / Rusanov) */
void orc_eval_splitting(const double *L, const double *R, const double *n,
                        double *F, double *lambda, int64_t *neg_p, int64_t *neg_rho)
{
    double pL[ORC_NF], pR[ORC_NF];
    orc_cons2prim(L, pL);
    orc_cons2prim(R, pR);

    double fL[ORC_NF], fR[ORC_NF];
    orc_eval_fluxes(L, pL, n, fL, neg_p, neg_rho);
    orc_eval_fluxes(R, pR, n, fR, neg_p, neg_rho);

    double unL = orc_normal_velocity(pL, n);
    double aL = sqrt(ORC_GAMMA * pL[ORC_T]);
    double lambdaL = fabs(unL) + aL;

    double unR = orc_normal_velocity(pR, n);
    double aR = sqrt(ORC_GAMMA * pR[ORC_T]);
    double lambdaR = fabs(unR) + aR;

    /* std::max(lambdaR, lambdaL) */
    *lambda = (lambdaR < lambdaL) ? lambdaL : lambdaR;

    for (int k = 0; k < ORC_NF; ++k)
        F[k] = 0.5 * ((fR[k] + fL[k]) - (*lambda) * (R[k] - L[k]));
}

/* euler.cpp:318-337 (reflecting) and 347-356 (wall == reflecting) */
static void orc_reflecting(const double *normal, const double *cons, double *cons_bc)
{
    double prim[ORC_NF];
    orc_cons2prim(cons, prim);

    double u[3] = {prim[ORC_U], prim[ORC_V], prim[ORC_W]};
    double un = orc_normal_velocity(prim, normal);
    double u_n[3] = {un * normal[0], un * normal[1], un * normal[2]};

#ifdef ORC_PLANT_REFLECT /* planted mistake (tests only): tangential instead of normal mirror */
    prim[ORC_U] = u[0] - 2 * (u[0] - u_n[0]);
    prim[ORC_V] = u[1] - 2 * (u[1] - u_n[1]);
    prim[ORC_W] = u[2] - 2 * (u[2] - u_n[2]);
#else
    prim[ORC_U] = u[0] - 2 * u_n[0];
    prim[ORC_V] = u[1] - 2 * u_n[1];
    prim[ORC_W] = u[2] - 2 * u_n[2];
#endif

    orc_prim2cons(prim, cons_bc);
}

/* euler.cpp:261-288 dispatch; 298-308 free flow; 366-376 Dirichlet */
void orc_eval_bc_values(int bc, const double *normal, const double *info,
                        const double *cons, double *cons_bc)
{
    switch (bc) {
    case ORC_BC_FREE_FLOW:
        memcpy(cons_bc, cons, sizeof(double) * ORC_NF);
        break;
    case ORC_BC_REFLECTING:
    case ORC_BC_WALL:
        orc_reflecting(normal, cons, cons_bc);
        break;
    case ORC_BC_DIRICHLET:
        orc_prim2cons(info, cons_bc);
        break;
    default:
        break;
    }
}

/* ------------------------------------------------------------------ */
/* problem                                                             */
/* ------------------------------------------------------------------ */

void orc_problem_defaults(orc_problem *pb, int type, int dim)
{
    memset(pb, 0, sizeof(*pb));
    pb->type = type;
    pb->dim = dim;
    /* problem.cpp:257-262 */
    pb->vortex_beta = 5.0;
    pb->vortex_periodic = 0;
    /* problem.cpp:70-149 domain defaults */
    switch (type) {
    case ORC_PROBLEM_VORTEX_XY:
    case ORC_PROBLEM_VORTEX_ZX:
    case ORC_PROBLEM_VORTEX_YZ:
        pb->domain_origin[0] = pb->domain_origin[1] = pb->domain_origin[2] = -5.0;
        pb->domain_length = 10.0;
        break;
    case ORC_PROBLEM_SOD_X:
    case ORC_PROBLEM_SOD_Y:
    case ORC_PROBLEM_SOD_Z:
        pb->domain_origin[0] = pb->domain_origin[1] = pb->domain_origin[2] = -1.0;
        pb->domain_length = 2.0;
        break;
    case ORC_PROBLEM_RADSOD:
        pb->domain_length = 8.0;
        break;
    default:
        pb->domain_length = 0.0; /* mandatory in the reference (problem.cpp:145) */
        break;
    }
    /* problem.cpp:336-342 */
    pb->sod_r0 = 1.0;
    pb->sod_p_in = 1.0;
    pb->sod_t_in = 1.0 / 1.0;
    pb->sod_p_out = 0.1;
    pb->sod_t_out = 0.1 / 0.125;
    /* problem.cpp:383-387 (FFSTEP state); the free-stream extension reuses it */
    pb->fs_prim[ORC_P] = 1.0;
    pb->fs_prim[ORC_U] = 3.0;
    pb->fs_prim[ORC_V] = 0.0;
    pb->fs_prim[ORC_W] = 0.0;
    pb->fs_prim[ORC_T] = 1.0 / 1.4;
}

/* problem.cpp:247-392.  Extensions: vortex strength/centre, periodic image
 * of the convected vortex, radial-Sod radius and states, free stream. */
void orc_exact_conservatives(const orc_problem *pb, const double *point_in, double t,
                             double *cons)
{
    double point[3] = {point_in[0], point_in[1], point_in[2]};
    double prim[ORC_NF];
    int type = pb->type;

    switch (type) {
    case ORC_PROBLEM_VORTEX_XY:
    case ORC_PROBLEM_VORTEX_ZX:
    case ORC_PROBLEM_VORTEX_YZ: {
        const double VORTEX_BETA = pb->vortex_beta;
        const double VORTEX_P_INF = 1.0;
        const double VORTEX_T_INF = 1.0;
        double VORTEX_U_INF = 1.0;
        double VORTEX_V_INF = 1.0;
        double VORTEX_W_INF = 1.0;

        if (type == ORC_PROBLEM_VORTEX_XY) VORTEX_W_INF = 0.;
        else if (type == ORC_PROBLEM_VORTEX_ZX) VORTEX_V_INF = 0.;
        else VORTEX_U_INF = 0.;

        point[0] -= t * VORTEX_U_INF;
        point[1] -= t * VORTEX_V_INF;
        point[2] -= t * VORTEX_W_INF;

        /* extension: vortex centre and nearest periodic image */
        for (int d = 0; d < 3; ++d) {
            point[d] -= pb->vortex_center[d];
            if (pb->vortex_periodic && d < pb->dim) {
                double L = pb->domain_length;
                point[d] -= L * floor(point[d] / L + 0.5);
            }
        }

        double csi = 0., eta = 0.;
        if (type == ORC_PROBLEM_VORTEX_XY) { csi = point[0]; eta = point[1]; }
        else if (type == ORC_PROBLEM_VORTEX_ZX) { csi = point[2]; eta = point[0]; }
        else { csi = point[1]; eta = point[2]; }

        double r = sqrt(csi * csi + eta * eta);
        double shape = VORTEX_BETA / (2 * M_PI) * exp(0.5 * (1 - r * r));

        double du = 0., dv = 0., dw = 0.;
        if (type == ORC_PROBLEM_VORTEX_XY) { du = -eta * shape; dv = csi * shape; dw = 0.; }
        else if (type == ORC_PROBLEM_VORTEX_ZX) { dw = -eta * shape; du = csi * shape; dv = 0.; }
        else { dv = -eta * shape; dw = csi * shape; du = 0.; }

        const double dT = -(ORC_GAMMA - 1.0) / (2 * ORC_GAMMA) * shape * shape;
        const double dp = pow(VORTEX_T_INF + dT, ORC_GAMMA / (ORC_GAMMA - 1.0)) - 1.0;

        prim[ORC_P] = VORTEX_P_INF + dp;
        prim[ORC_U] = VORTEX_U_INF + du;
        prim[ORC_V] = VORTEX_V_INF + dv;
        prim[ORC_W] = (pb->dim == 3) ? VORTEX_W_INF + dw : 0;
        prim[ORC_T] = VORTEX_T_INF + dT;
        break;
    }
    case ORC_PROBLEM_RADSOD: {
        double r = sqrt(point[0] * point[0] + point[1] * point[1] + point[2] * point[2]);
        prim[ORC_U] = 0.0;
        prim[ORC_V] = 0.0;
        prim[ORC_W] = 0.0;
        if (r < pb->sod_r0) {
            prim[ORC_P] = pb->sod_p_in;
            prim[ORC_T] = pb->sod_t_in;
        } else {
            prim[ORC_P] = pb->sod_p_out;
            prim[ORC_T] = pb->sod_t_out;
        }
        break;
    }
    case ORC_PROBLEM_SOD_X:
    case ORC_PROBLEM_SOD_Y:
    case ORC_PROBLEM_SOD_Z: {
        prim[ORC_U] = 0.0;
        prim[ORC_V] = 0.0;
        prim[ORC_W] = 0.0;
        double r = (type == ORC_PROBLEM_SOD_X) ? point[0] : (type == ORC_PROBLEM_SOD_Y) ? point[1] : point[2];
        if (r < 0.) {
            prim[ORC_P] = 1.0;
            prim[ORC_T] = 1.0 / 1.0;
        } else {
            prim[ORC_P] = 0.1;
            prim[ORC_T] = 0.1 / 0.125;
        }
        break;
    }
    default: /* FFSTEP (problem.cpp:381-388) and the free-stream extension */
        for (int k = 0; k < ORC_NF; ++k) prim[k] = pb->fs_prim[k];
        break;
    }

    orc_prim2cons(prim, cons);
}

/* problem.cpp:404-438 */
int orc_border_bc_type(const orc_problem *pb, const double *fc)
{
    switch (pb->type) {
    case ORC_PROBLEM_VORTEX_XY:
    case ORC_PROBLEM_VORTEX_ZX:
    case ORC_PROBLEM_VORTEX_YZ:
    case ORC_PROBLEM_SOD_X:
    case ORC_PROBLEM_SOD_Y:
    case ORC_PROBLEM_SOD_Z:
        return ORC_BC_FREE_FLOW;
    case ORC_PROBLEM_RADSOD:
        return ORC_BC_REFLECTING;
    case ORC_PROBLEM_FFSTEP:
        if (fc[0] < 1e-10) return ORC_BC_DIRICHLET;
        else if (fc[0] > 3.2 - 1e-10) return ORC_BC_FREE_FLOW;
        else return ORC_BC_REFLECTING;
    default:
        return ORC_BC_FREE_FLOW;
    }
}

/* problem.cpp:450-477 */
void orc_border_bc_info(const orc_problem *pb, int bc, double *info)
{
    for (int k = 0; k < ORC_NF; ++k) info[k] = 0.0;
    if (bc == ORC_BC_DIRICHLET)
        for (int k = 0; k < ORC_NF; ++k) info[k] = pb->fs_prim[k];
}

/* ------------------------------------------------------------------ */
/* mesh                                                                */
/* ------------------------------------------------------------------ */

int64_t orc_cell_index(const orc_mesh *m, int64_t i, int64_t j, int64_t k)
{
    return i + m->n[0] * (j + m->n[1] * k);
}

/* Closed-form equivalent of MeshGeometricalInfo::_extract (mesh_info.cpp:86-118)
 * for a uniform Cartesian grid.  Face list order: all x-faces, then y, then z;
 * within an axis k-major, then j, then i.  Interior faces: owner = lower
 * cell, normal +e_axis.  Border faces: owner = the cell, neighbour -1,
 * normal owner-outward.  Periodic axes: the wrap face comes first in its
 * row, owner = last cell, neighbour = first cell, normal +e_axis. */
int orc_mesh_build(orc_mesh *m, int dim, const int64_t *n, const double *origin, double h,
                   const int *periodic)
{
    memset(m, 0, sizeof(*m));
    m->dim = dim;
    for (int d = 0; d < 3; ++d) {
        m->n[d] = (d < dim) ? n[d] : 1;
        m->origin[d] = origin[d];
        m->periodic[d] = (periodic && d < dim) ? periodic[d] : 0;
    }
    m->h = h;
    m->ncells = m->n[0] * m->n[1] * m->n[2];
    m->volume = (dim == 3) ? h * h * h : h * h;
    m->area = (dim == 3) ? h * h : h;

    int64_t nf = 0;
    for (int a = 0; a < dim; ++a) {
        int64_t per_row = m->periodic[a] ? m->n[a] : m->n[a] + 1;
        nf += per_row * (m->ncells / m->n[a]);
    }
    m->nfaces = nf;

    m->cell_centroid = (double *)malloc(sizeof(double) * 3 * m->ncells);
    m->face_owner = (int64_t *)malloc(sizeof(int64_t) * nf);
    m->face_neigh = (int64_t *)malloc(sizeof(int64_t) * nf);
    m->face_normal = (double *)malloc(sizeof(double) * 3 * nf);
    m->face_centroid = (double *)malloc(sizeof(double) * 3 * nf);
    m->face_axis = (int *)malloc(sizeof(int) * nf);
    m->face_side = (int *)malloc(sizeof(int) * nf);
    if (!m->cell_centroid || !m->face_owner || !m->face_neigh || !m->face_normal ||
        !m->face_centroid || !m->face_axis || !m->face_side) {
        orc_mesh_free(m);
        return -1;
    }

    for (int64_t k = 0; k < m->n[2]; ++k)
        for (int64_t j = 0; j < m->n[1]; ++j)
            for (int64_t i = 0; i < m->n[0]; ++i) {
                int64_t c = orc_cell_index(m, i, j, k);
                int64_t ijk[3] = {i, j, k};
                for (int d = 0; d < 3; ++d)
                    m->cell_centroid[3 * c + d] = (d < dim) ? origin[d] + (ijk[d] + 0.5) * h : origin[d];
            }

    int64_t f = 0;
    for (int a = 0; a < dim; ++a) {
        int64_t lim[3] = {m->n[0], m->n[1], m->n[2]};
        lim[a] = m->periodic[a] ? m->n[a] : m->n[a] + 1;
        for (int64_t k = 0; k < lim[2]; ++k)
            for (int64_t j = 0; j < lim[1]; ++j)
                for (int64_t i = 0; i < lim[0]; ++i) {
                    int64_t ijk[3] = {i, j, k};
                    int64_t p = ijk[a]; /* face p sits below cell p */
                    int64_t lo[3] = {i, j, k}, hi[3] = {i, j, k};
                    lo[a] = p - 1;
                    hi[a] = p;
                    double *nrm = &m->face_normal[3 * f];
                    double *fc = &m->face_centroid[3 * f];
                    nrm[0] = nrm[1] = nrm[2] = 0.0;
                    for (int d = 0; d < 3; ++d)
                        fc[d] = (d < dim) ? origin[d] + (ijk[d] + 0.5) * h : origin[d];
                    fc[a] = origin[a] + p * h;
                    m->face_axis[f] = a;
                    m->face_side[f] = -1;
                    if (m->periodic[a]) {
                        if (p == 0) lo[a] = m->n[a] - 1;
                        m->face_owner[f] = orc_cell_index(m, lo[0], lo[1], lo[2]);
                        m->face_neigh[f] = orc_cell_index(m, hi[0], hi[1], hi[2]);
                        nrm[a] = 1.0;
                    } else if (p == 0) {
                        m->face_owner[f] = orc_cell_index(m, hi[0], hi[1], hi[2]);
                        m->face_neigh[f] = -1;
                        nrm[a] = -1.0;
                        m->face_side[f] = 2 * a;
                    } else if (p == m->n[a]) {
                        m->face_owner[f] = orc_cell_index(m, lo[0], lo[1], lo[2]);
                        m->face_neigh[f] = -1;
                        nrm[a] = 1.0;
                        m->face_side[f] = 2 * a + 1;
                    } else {
                        m->face_owner[f] = orc_cell_index(m, lo[0], lo[1], lo[2]);
                        m->face_neigh[f] = orc_cell_index(m, hi[0], hi[1], hi[2]);
                        nrm[a] = 1.0;
                    }
                    ++f;
                }
    }
    return 0;
}

void orc_mesh_free(orc_mesh *m)
{
    free(m->cell_centroid);
    free(m->face_owner);
    free(m->face_neigh);
    free(m->face_normal);
    free(m->face_centroid);
    free(m->face_axis);
    free(m->face_side);
    free(m->face_area);
    free(m->cell_volume);
    memset(m, 0, sizeof(*m));
}

/* ------------------------------------------------------------------ */
/* body                                                                */
/* ------------------------------------------------------------------ */

/* body.cpp:80-95: a point on a box face is solid (inclusive bounds). */
int orc_point_in_box(const double *b, const double *p)
{
    if (p[0] < b[0] || p[0] > b[3]) return 0;
    if (p[1] < b[1] || p[1] > b[4]) return 0;
    if (p[2] < b[2] || p[2] > b[5]) return 0;
    return 1;
}

/* main.cpp:222-237 with body::isPointFluid; spheres are a level-set
 * extension: fluid iff |x - c| - r >= 0 at the centroid. */
void orc_mark_fluid(const orc_mesh *m, const double *boxes, int nboxes,
                    const double *spheres, int nspheres, uint8_t *fluid)
{
    for (int64_t c = 0; c < m->ncells; ++c) {
        const double *x = &m->cell_centroid[3 * c];
        int is_fluid = 1;
        for (int b = 0; b < nboxes && is_fluid; ++b)
            if (orc_point_in_box(&boxes[6 * b], x)) is_fluid = 0;
        for (int s = 0; s < nspheres && is_fluid; ++s) {
            const double *sp = &spheres[4 * s];
            double dx = x[0] - sp[0], dy = x[1] - sp[1], dz = x[2] - sp[2];
            double phi = sqrt(dx * dx + dy * dy + dz * dz) - sp[3];
            if (!(phi >= 0.0)) is_fluid = 0;
        }
        fluid[c] = (uint8_t)is_fluid;
    }
}

/* ------------------------------------------------------------------ */
/* solver                                                              */
/* ------------------------------------------------------------------ */

int orc_solver_init(orc_solver *s, const orc_problem *pb, int order,
                    const int64_t *n, const double *origin, double h,
                    const int *side_bc, const double *side_info, const uint8_t *fluid)
{
    memset(s, 0, sizeof(*s));
    if (order != 1 && order != 2) return -2; /* reconstruction.cpp:75-76 exit(2) */
    s->pb = *pb;
    s->order = order;

    int periodic[3] = {0, 0, 0};
    if (side_bc)
        for (int a = 0; a < pb->dim; ++a)
            periodic[a] = (side_bc[2 * a] == ORC_BC_PERIODIC);
    if (orc_mesh_build(&s->mesh, pb->dim, n, origin, h, periodic) != 0) return -1;
    orc_mesh *m = &s->mesh;

    s->fluid = (uint8_t *)malloc(m->ncells);
    s->solved = (uint8_t *)malloc(m->ncells);
    s->face_bc = (int *)malloc(sizeof(int) * m->nfaces);
    s->face_info = (double *)calloc((size_t)m->nfaces * ORC_NF, sizeof(double));
    s->slopes = (order == 2) ? (double *)calloc((size_t)m->ncells * 3 * ORC_NF, sizeof(double)) : NULL;
    if (!s->fluid || !s->solved || !s->face_bc || !s->face_info || (order == 2 && !s->slopes)) {
        orc_solver_free(s);
        return -1;
    }

    /* main.cpp:222-237: solved == fluid in a single-domain run */
    for (int64_t c = 0; c < m->ncells; ++c) {
        s->fluid[c] = fluid ? fluid[c] : 1;
        s->solved[c] = s->fluid[c];
    }

    /* main.cpp:251-277 */
    for (int64_t f = 0; f < m->nfaces; ++f) {
        int side = m->face_side[f];
        if (side >= 0) {
            int bc = side_bc ? side_bc[side] : orc_border_bc_type(pb, &m->face_centroid[3 * f]);
            s->face_bc[f] = bc;
            if (side_info) memcpy(&s->face_info[ORC_NF * f], &side_info[ORC_NF * side], sizeof(double) * ORC_NF);
            else orc_border_bc_info(pb, bc, &s->face_info[ORC_NF * f]);
        } else {
            int of = s->fluid[m->face_owner[f]];
            int nf = s->fluid[m->face_neigh[f]];
#ifdef ORC_PLANT_WALL_DROP /* planted mistake (tests only): fluid|solid faces left without a wall */
            s->face_bc[f] = ORC_BC_NONE + 0 * (of ^ nf);
#else
            s->face_bc[f] = (of ^ nf) ? ORC_BC_WALL : ORC_BC_NONE;
#endif
        }
    }
    return 0;
}

/* A solver on an explicit interface list, the general MeshGeometricalInfo
 * (mesh_info.cpp:86-118): owner / neighbour (-1 on borders) / area /
 * owner-outward normal per face, volume (and centroid) per cell, the border
 * faces' BC codes and Dirichlet data, min_cell_size (main.cpp:351-356).
 * Interior faces become walls where exactly one side is fluid
 * (main.cpp:251-277).  Order 1 (the reference's reconstruction).  The face
 * loop runs in the given face order. */
int orc_solver_init_faces(orc_solver *s, const orc_problem *pb, int dim, int64_t ncells, int64_t nfaces,
                          const double *cell_volume, const double *cell_centroid, double min_cell_size,
                          const int64_t *owner, const int64_t *neigh, const double *area, const double *normal,
                          const int *border_bc, const double *border_info, const uint8_t *fluid)
{
    memset(s, 0, sizeof(*s));
    s->pb = *pb;
    s->order = 1;
    orc_mesh *m = &s->mesh;
    m->dim = dim;
    m->n[0] = ncells;
    m->n[1] = m->n[2] = 1;
    m->h = min_cell_size;
    m->ncells = ncells;
    m->nfaces = nfaces;
    m->cell_centroid = (double *)calloc((size_t)ncells * 3, sizeof(double));
    m->cell_volume = (double *)malloc(sizeof(double) * ncells);
    m->face_owner = (int64_t *)malloc(sizeof(int64_t) * nfaces);
    m->face_neigh = (int64_t *)malloc(sizeof(int64_t) * nfaces);
    m->face_normal = (double *)malloc(sizeof(double) * 3 * nfaces);
    m->face_centroid = (double *)calloc((size_t)nfaces * 3, sizeof(double));
    m->face_axis = (int *)calloc((size_t)nfaces, sizeof(int));
    m->face_side = (int *)malloc(sizeof(int) * nfaces);
    m->face_area = (double *)malloc(sizeof(double) * nfaces);
    s->fluid = (uint8_t *)malloc(ncells);
    s->solved = (uint8_t *)malloc(ncells);
    s->face_bc = (int *)malloc(sizeof(int) * nfaces);
    s->face_info = (double *)calloc((size_t)nfaces * ORC_NF, sizeof(double));
    if (!m->cell_centroid || !m->cell_volume || !m->face_owner || !m->face_neigh || !m->face_normal ||
        !m->face_centroid || !m->face_axis || !m->face_side || !m->face_area || !s->fluid || !s->solved ||
        !s->face_bc || !s->face_info) {
        orc_solver_free(s);
        return -1;
    }
    memcpy(m->cell_volume, cell_volume, sizeof(double) * ncells);
    if (cell_centroid) memcpy(m->cell_centroid, cell_centroid, sizeof(double) * 3 * ncells);
    memcpy(m->face_owner, owner, sizeof(int64_t) * nfaces);
    memcpy(m->face_neigh, neigh, sizeof(int64_t) * nfaces);
    memcpy(m->face_area, area, sizeof(double) * nfaces);
    memcpy(m->face_normal, normal, sizeof(double) * 3 * nfaces);
    for (int64_t c = 0; c < ncells; ++c) {
        s->fluid[c] = fluid ? fluid[c] : 1;
        s->solved[c] = s->fluid[c];
    }
    for (int64_t f = 0; f < nfaces; ++f) {
        m->face_side[f] = -1;
        if (neigh[f] < 0) {
            s->face_bc[f] = border_bc ? border_bc[f] : ORC_BC_FREE_FLOW;
            if (border_info) memcpy(&s->face_info[ORC_NF * f], &border_info[ORC_NF * f], sizeof(double) * ORC_NF);
        } else {
            s->face_bc[f] = (s->fluid[owner[f]] ^ s->fluid[neigh[f]]) ? ORC_BC_WALL : ORC_BC_NONE;
        }
    }
    return 0;
}

void orc_solver_free(orc_solver *s)
{
    orc_mesh_free(&s->mesh);
    free(s->fluid);
    free(s->solved);
    free(s->face_bc);
    free(s->face_info);
    free(s->slopes);
    memset(s, 0, sizeof(*s));
}

/* main.cpp:335-344 -> problem.cpp:220-223 (exact solution at t=0, centroid) */
void orc_initial_state(const orc_solver *s, double *U)
{
    for (int64_t c = 0; c < s->mesh.ncells; ++c)
        orc_exact_conservatives(&s->pb, &s->mesh.cell_centroid[3 * c], 0.0, &U[ORC_NF * c]);
}

/* minmod of two undivided differences */
static double orc_minmod(double a, double b)
{
    if (a > 0.0 && b > 0.0) return (a < b) ? a : b;
    if (a < 0.0 && b < 0.0) return (a > b) ? a : b;
    return 0.0;
}

/* Neighbour of cell (ijk) along axis a, direction dir (+1/-1); -1 when it
 * lies outside a non-periodic border. */
static int64_t orc_neighbour(const orc_mesh *m, int64_t c, int a, int dir)
{
    int64_t ijk[3];
    ijk[0] = c % m->n[0];
    ijk[1] = (c / m->n[0]) % m->n[1];
    ijk[2] = c / (m->n[0] * m->n[1]);
    int64_t q = ijk[a] + dir;
    if (q < 0 || q >= m->n[a]) {
        if (!m->periodic[a]) return -1;
        q = (q + m->n[a]) % m->n[a];
    }
    ijk[a] = q;
    return orc_cell_index(m, ijk[0], ijk[1], ijk[2]);
}

/* reconstruction.cpp:47-55 is an empty hook at order 1.  Order 2 (an
 * extension): per cell and axis, minmod-limited undivided differences of the
 * conservative means; a neighbour across a border or inside a body gives a
 * zero one-sided difference. */
static void orc_polys(orc_solver *s, const double *U, int nthreads)
{
    if (s->order != 2) return;
    const orc_mesh *m = &s->mesh;
#pragma omp parallel for schedule(static) num_threads(nthreads) if (nthreads > 1)
    for (int64_t c = 0; c < m->ncells; ++c) {
        for (int a = 0; a < 3; ++a) {
            double *sl = &s->slopes[(3 * c + a) * ORC_NF];
            if (a >= m->dim) {
                for (int k = 0; k < ORC_NF; ++k) sl[k] = 0.0;
                continue;
            }
            int64_t cm = orc_neighbour(m, c, a, -1);
            int64_t cp = orc_neighbour(m, c, a, +1);
            int okm = (cm >= 0) && s->fluid[cm];
            int okp = (cp >= 0) && s->fluid[cp];
            for (int k = 0; k < ORC_NF; ++k) {
                double dm = okm ? U[ORC_NF * c + k] - U[ORC_NF * cm + k] : 0.0;
                double dp = okp ? U[ORC_NF * cp + k] - U[ORC_NF * c + k] : 0.0;
                sl[k] = orc_minmod(dp, dm);
            }
        }
    }
}

void orc_compute_polynomials(orc_solver *s, const double *U) { orc_polys(s, U, 1); }

/* reconstruction.cpp:67-97.  'outward' is the unit normal of the face
 * pointing out of this cell.  Order 1 copies the mean.  Order 2 evaluates
 * the limited linear polynomial at the face (mean +/- slope/2) and falls back
 * to the mean when the reconstructed density or pressure is not positive. */
void orc_reconstruct(const orc_solver *s, int64_t cell, const double *outward,
                     const double *U, double *values)
{
    const double *mean = &U[ORC_NF * cell];
    if (s->order == 1) {
        memcpy(values, mean, sizeof(double) * ORC_NF);
        return;
    }
    int a = 0;
    for (int d = 1; d < 3; ++d)
        if (fabs(outward[d]) > fabs(outward[a])) a = d;
    const double *sl = &s->slopes[(3 * cell + a) * ORC_NF];
    if (outward[a] > 0.0)
        for (int k = 0; k < ORC_NF; ++k) values[k] = mean[k] + 0.5 * sl[k];
    else
        for (int k = 0; k < ORC_NF; ++k) values[k] = mean[k] - 0.5 * sl[k];

    double prim[ORC_NF];
    orc_cons2prim(values, prim);
    if (!(values[ORC_RHO] > 0.0) || !(prim[ORC_P] > 0.0))
        memcpy(values, mean, sizeof(double) * ORC_NF);
}

/* One interface of euler.cpp:153-247 (the body of the face loop) up to the
 * sums: AF = interfaceArea * fluxes, the eigenvalue into *max_eig, negative
 * p / rho sides counted.  Returns which sides take AF: bit 0 the owner
 * (RHS -= AF), bit 1 the neighbour (RHS += AF); 0 for a skipped face. */
static int orc_face_af(const orc_solver *s, const double *U, int64_t f, double *AF, double *max_eig,
                       int64_t *neg_p, int64_t *neg_rho)
{
    const orc_mesh *m = &s->mesh;
    int64_t owner = m->face_owner[f];
    int64_t neigh = m->face_neigh[f];
    int ownerSolved = s->solved[owner];
    int neighSolved = (neigh >= 0) ? s->solved[neigh] : 0;

#ifndef ORC_PLANT_NO_SKIP /* planted mistake (tests only): faces of unsolved cells evaluated */
    if (!ownerSolved && !neighSolved) return 0;
#else
    if (neigh >= 0 && s->face_bc[f] == ORC_BC_NONE && !ownerSolved && !neighSolved) {
        double Fs[ORC_NF], es;
        orc_eval_splitting(&U[ORC_NF * owner], &U[ORC_NF * neigh], &m->face_normal[3 * f], Fs, &es, NULL, NULL);
        *max_eig = (es < *max_eig) ? *max_eig : es;
        return 0;
    }
    if (!ownerSolved && !neighSolved) return 0;
#endif

    int bc = s->face_bc[f];
    const double area = m->face_area ? m->face_area[f] : m->area; /* rawGetInterfaceArea */
    const double *normal = &m->face_normal[3 * f];

    double ownerRec[ORC_NF], neighRec[ORC_NF];
    if (bc == ORC_BC_NONE) {
        double flipped[3] = {-normal[0], -normal[1], -normal[2]};
        orc_reconstruct(s, owner, normal, U, ownerRec);
        orc_reconstruct(s, neigh, flipped, U, neighRec);
    } else {
        /* euler.cpp:198-225: the fluid side's normal points out of it */
        double local[3] = {normal[0], normal[1], normal[2]};
        int64_t fluidCell;
        double *fluidRec, *virtRec;
        if (ownerSolved) {
            fluidCell = owner;
            fluidRec = ownerRec;
            virtRec = neighRec;
        } else {
            fluidCell = neigh;
            fluidRec = neighRec;
            virtRec = ownerRec;
            local[0] = -1. * local[0];
            local[1] = -1. * local[1];
            local[2] = -1. * local[2];
        }
#ifdef ORC_PLANT_WALL_SLOT /* planted mistake (tests only): the virtual state in the fluid cell's slot */
        {
            double *tmp = fluidRec;
            fluidRec = virtRec;
            virtRec = tmp;
        }
#endif
        orc_reconstruct(s, fluidCell, local, U, fluidRec);
        orc_eval_bc_values(bc, local, &s->face_info[ORC_NF * f], fluidRec, virtRec);
    }

    double F[ORC_NF];
    double faceMaxEig;
    /* euler.cpp:232: the flux uses the owner-outward normal */
    orc_eval_splitting(ownerRec, neighRec, normal, F, &faceMaxEig, neg_p, neg_rho);

    *max_eig = (faceMaxEig < *max_eig) ? *max_eig : faceMaxEig; /* std::max(faceMaxEig, *maxEig) */

#ifdef ORC_PLANT_SOLID_UPDATE /* planted mistake (tests only): residuals of unsolved cells accumulated */
    ownerSolved = 1;
    neighSolved = neigh >= 0;
#endif
    for (int k = 0; k < ORC_NF; ++k) AF[k] = area * F[k];
    return (ownerSolved ? 1 : 0) | (neighSolved ? 2 : 0);
}

/* One interface with its sums: RHS_owner -= AF, RHS_neigh += AF
 * (euler.cpp:237-247). */
static void orc_face(const orc_solver *s, const double *U, double *RHS, int64_t f, double *max_eig,
                     int64_t *neg_p, int64_t *neg_rho)
{
    double AF[ORC_NF];
    const int sides = orc_face_af(s, U, f, AF, max_eig, neg_p, neg_rho);
    if (sides & 1)
        for (int k = 0; k < ORC_NF; ++k) RHS[ORC_NF * s->mesh.face_owner[f] + k] -= AF[k];
    if (sides & 2)
        for (int k = 0; k < ORC_NF; ++k) RHS[ORC_NF * s->mesh.face_neigh[f] + k] += AF[k];
}

/* euler.cpp:127-249 */
int orc_compute_rhs(orc_solver *s, const double *U, double *RHS, double *max_eig)
{
    const orc_mesh *m = &s->mesh;

    /* euler.cpp:142-148: reset over all cells */
    for (int64_t i = 0; i < m->ncells * ORC_NF; ++i) RHS[i] = 0.;

    *max_eig = 0.0;
    for (int64_t f = 0; f < m->nfaces; ++f)
        orc_face(s, U, RHS, f, max_eig, &s->neg_p, &s->neg_rho);
    return 0;
}

/* The same residual with the face loop spread over threads, bit for bit the
 * serial one: the faces of one axis are split into groups whose cells are
 * disjoint (x: rows; y: (plane, 64-column chunk); z: (row, 64-column chunk)),
 * each group runs its faces in increasing position along the axis, and the
 * axes run one after another, so every cell adds its faces in the serial
 * order (all x, then y, then z; lower before upper).  The eigenvalue maximum
 * and the counters are combined per thread (a maximum of finite values does
 * not depend on the order). */
int orc_compute_rhs_threads(orc_solver *s, const double *U, double *RHS, double *max_eig, int nthreads)
{
    const orc_mesh *m = &s->mesh;
    const int64_t n0 = m->n[0], n1 = m->n[1], n2 = m->n[2];
    const int64_t CH = 64;
    const int64_t nch = (n0 + CH - 1) / CH;
    double gmax = 0.0;
    int64_t gp = 0, gr = 0;
#pragma omp parallel num_threads(nthreads) if (nthreads > 1)
    {
        double lmax = 0.0;
        int64_t lp = 0, lr = 0;
#pragma omp for schedule(static)
        for (int64_t i = 0; i < m->ncells * ORC_NF; ++i) RHS[i] = 0.;
        int64_t base = 0;
        for (int a = 0; a < m->dim; ++a) {
            const int64_t lim = m->periodic[a] ? m->n[a] : m->n[a] + 1;
            if (a == 0) {
#pragma omp for schedule(static)
                for (int64_t r = 0; r < n1 * n2; ++r)
                    for (int64_t p = 0; p < lim; ++p) orc_face(s, U, RHS, base + p + lim * r, &lmax, &lp, &lr);
            } else if (a == 1) {
#pragma omp for schedule(static)
                for (int64_t t = 0; t < n2 * nch; ++t) {
                    const int64_t k = t / nch, i0 = (t % nch) * CH, i1 = (i0 + CH < n0) ? i0 + CH : n0;
                    for (int64_t p = 0; p < lim; ++p)
                        for (int64_t i = i0; i < i1; ++i)
                            orc_face(s, U, RHS, base + i + n0 * (p + lim * k), &lmax, &lp, &lr);
                }
            } else {
#pragma omp for schedule(static)
                for (int64_t t = 0; t < n1 * nch; ++t) {
                    const int64_t j = t / nch, i0 = (t % nch) * CH, i1 = (i0 + CH < n0) ? i0 + CH : n0;
                    for (int64_t p = 0; p < lim; ++p)
                        for (int64_t i = i0; i < i1; ++i)
                            orc_face(s, U, RHS, base + i + n0 * (j + n1 * p), &lmax, &lp, &lr);
                }
            }
            base += lim * (m->ncells / m->n[a]);
        }
#pragma omp critical
        {
            gmax = (lmax < gmax) ? gmax : lmax;
            gp += lp;
            gr += lr;
        }
    }
    s->neg_p += gp;
    s->neg_rho += gr;
    *max_eig = gmax;
    return 0;
}

/* main.cpp:381-506: one SSP-RK3 step written inline in the reference. */
static double orc_rk3(orc_solver *s, double *U, double *W, double *RHS, double cfl, double t, double tmax,
                      double *stage_eig, int nthreads)
{
    const orc_mesh *m = &s->mesh;
    const double minCellSize = m->h; /* main.cpp:351-356 */
    double maxEig;

    orc_polys(s, U, nthreads);
    if (nthreads > 1 && !m->face_area) orc_compute_rhs_threads(s, U, RHS, &maxEig, nthreads);
    else if (nthreads > 1) orc_compute_rhs_faces_threads(s, U, RHS, &maxEig, nthreads);
    else orc_compute_rhs(s, U, RHS, &maxEig);
    if (stage_eig) stage_eig[0] = maxEig;

#ifdef ORC_PLANT_DT /* planted mistake (tests only): the 0.9 safety factor dropped */
    double dt = cfl * minCellSize / maxEig;
#else
    double dt = 0.9 * cfl * minCellSize / maxEig; /* main.cpp:398 */
#endif
    if (t + dt > tmax) dt = tmax - t;             /* main.cpp:400-402 */

#pragma omp parallel for schedule(static) num_threads(nthreads) if (nthreads > 1)
    for (int64_t c = 0; c < m->ncells; ++c) {    /* main.cpp:409-423 */
#ifndef ORC_PLANT_SOLID_UPDATE
        if (!s->solved[c]) continue;
#endif
        for (int k = 0; k < ORC_NF; ++k)
            W[ORC_NF * c + k] = U[ORC_NF * c + k] + dt * RHS[ORC_NF * c + k] / ORC_VOL(m, c);
    }

    orc_polys(s, W, nthreads);
    if (nthreads > 1 && !m->face_area) orc_compute_rhs_threads(s, W, RHS, &maxEig, nthreads);
    else if (nthreads > 1) orc_compute_rhs_faces_threads(s, W, RHS, &maxEig, nthreads);
    else orc_compute_rhs(s, W, RHS, &maxEig);
    if (stage_eig) stage_eig[1] = maxEig;

#pragma omp parallel for schedule(static) num_threads(nthreads) if (nthreads > 1)
    for (int64_t c = 0; c < m->ncells; ++c) {    /* main.cpp:445-459 */
#ifndef ORC_PLANT_SOLID_UPDATE
        if (!s->solved[c]) continue;
#endif
        for (int k = 0; k < ORC_NF; ++k)
#ifdef ORC_PLANT_RK /* planted mistake (tests only): stage-2 weights of a different scheme */
            W[ORC_NF * c + k] = 0.5 * U[ORC_NF * c + k] + 0.5 * (W[ORC_NF * c + k] + dt * RHS[ORC_NF * c + k] / ORC_VOL(m, c));
#else
            W[ORC_NF * c + k] = 0.75 * U[ORC_NF * c + k] + 0.25 * (W[ORC_NF * c + k] + dt * RHS[ORC_NF * c + k] / ORC_VOL(m, c));
#endif
    }

    orc_polys(s, W, nthreads);
    if (nthreads > 1 && !m->face_area) orc_compute_rhs_threads(s, W, RHS, &maxEig, nthreads);
    else if (nthreads > 1) orc_compute_rhs_faces_threads(s, W, RHS, &maxEig, nthreads);
    else orc_compute_rhs(s, W, RHS, &maxEig);
    if (stage_eig) stage_eig[2] = maxEig;

#pragma omp parallel for schedule(static) num_threads(nthreads) if (nthreads > 1)
    for (int64_t c = 0; c < m->ncells; ++c) {    /* main.cpp:481-495 */
#ifndef ORC_PLANT_SOLID_UPDATE
        if (!s->solved[c]) continue;
#endif
        for (int k = 0; k < ORC_NF; ++k)
            U[ORC_NF * c + k] = (1. / 3) * U[ORC_NF * c + k] + (2. / 3) * (W[ORC_NF * c + k] + dt * RHS[ORC_NF * c + k] / ORC_VOL(m, c));
    }
    return dt;
}

double orc_rk3_step(orc_solver *s, double *U, double *W, double *RHS,
                    double cfl, double t, double tmax, double *stage_eig)
{
    return orc_rk3(s, U, W, RHS, cfl, t, tmax, stage_eig, 1);
}

/* The same step with the loops on nthreads threads (bit for bit the serial
 * step; the CPU baseline on all host cores). */
double orc_rk3_step_threads(orc_solver *s, double *U, double *W, double *RHS,
                            double cfl, double t, double tmax, double *stage_eig, int nthreads)
{
    return orc_rk3(s, U, W, RHS, cfl, t, tmax, stage_eig, nthreads < 1 ? 1 : nthreads);
}

/* main.cpp:374-506 time loop */
int64_t orc_run(orc_solver *s, double *U, double cfl, double t0, double tmax,
                int64_t max_steps, double *t_out)
{
    const int64_t nc = s->mesh.ncells;
    double *W = (double *)malloc(sizeof(double) * ORC_NF * nc);
    double *RHS = (double *)malloc(sizeof(double) * ORC_NF * nc);
    if (!W || !RHS) {
        free(W);
        free(RHS);
        return -1;
    }
    /* the work storage starts as a copy so unsolved cells hold finite values */
    memcpy(W, U, sizeof(double) * ORC_NF * nc);
    double t = t0;
    int64_t step = 0;
    while (t < tmax && (max_steps < 0 || step < max_steps)) {
        double dt = orc_rk3_step(s, U, W, RHS, cfl, t, tmax, NULL);
        t += dt;
        ++step;
    }
    free(W);
    free(RHS);
    if (t_out) *t_out = t;
    return step;
}

/* main.cpp:550-561 (FID_P == FID_RHO == 0: density).  L2 is an extension:
 * sqrt(sum (rho - rho_ex)^2 V). */
void orc_error(const orc_solver *s, const double *U, double t, double *l1, double *l2)
{
    double e1 = 0., e2 = 0.;
    double ex[ORC_NF];
    for (int64_t c = 0; c < s->mesh.ncells; ++c) {
        orc_exact_conservatives(&s->pb, &s->mesh.cell_centroid[3 * c], t, ex);
        double d = U[ORC_NF * c + 0] - ex[0];
        e1 += fabs(d) * ORC_VOL(&s->mesh, c);
        e2 += d * d * ORC_VOL(&s->mesh, c);
    }
    if (l1) *l1 = e1;
    if (l2) *l2 = sqrt(e2);
}

/* ------------------------------------------------------------------ */
/* flat helpers for ctypes callers                                     */
/* ------------------------------------------------------------------ */

orc_solver *orc_solver_alloc(void) { return (orc_solver *)calloc(1, sizeof(orc_solver)); }
void orc_solver_dealloc(orc_solver *s)
{
    if (!s) return;
    orc_solver_free(s);
    free(s);
}
int64_t orc_solver_ncells(const orc_solver *s) { return s->mesh.ncells; }
int64_t orc_solver_nfaces(const orc_solver *s) { return s->mesh.nfaces; }
void orc_solver_counters(const orc_solver *s, int64_t *neg_p, int64_t *neg_rho)
{
    *neg_p = s->neg_p;
    *neg_rho = s->neg_rho;
}
void orc_solver_cell_centroids(const orc_solver *s, double *out)
{
    memcpy(out, s->mesh.cell_centroid, sizeof(double) * 3 * s->mesh.ncells);
}
/* main.cpp:231-235: cells solved by another rank (ghosts) stay fluid but are
 * not solved here.  A solved cell must be fluid; returns -1 otherwise. */
int orc_solver_set_solved(orc_solver *s, const uint8_t *solved)
{
    for (int64_t c = 0; c < s->mesh.ncells; ++c)
        if (solved[c] && !s->fluid[c]) return -1;
    for (int64_t c = 0; c < s->mesh.ncells; ++c) s->solved[c] = solved[c] ? 1 : 0;
    return 0;
}

void orc_solver_flags(const orc_solver *s, uint8_t *fluid, uint8_t *solved)
{
    memcpy(fluid, s->fluid, s->mesh.ncells);
    memcpy(solved, s->solved, s->mesh.ncells);
}
void orc_solver_face_bc(const orc_solver *s, int *out)
{
    memcpy(out, s->face_bc, sizeof(int) * s->mesh.nfaces);
}

/* The residual of a face-list solver on nthreads threads, bit for bit the
 * serial interface loop: every face's A*F and eigenvalue first (faces in
 * parallel), then every cell sums its faces in increasing face index (cells
 * in parallel), the order the serial loop adds them. */
int orc_compute_rhs_faces_threads(orc_solver *s, const double *U, double *RHS, double *max_eig, int nthreads)
{
    const orc_mesh *m = &s->mesh;
    const int64_t F = m->nfaces, N = m->ncells;
    double *AF = (double *)malloc(sizeof(double) * ORC_NF * F);
    uint8_t *live = (uint8_t *)malloc(F);
    int64_t *off = (int64_t *)calloc((size_t)N + 1, sizeof(int64_t));
    int64_t *inc = (int64_t *)malloc(sizeof(int64_t) * 2 * F);
    if (!AF || !live || !off || !inc) {
        free(AF); free(live); free(off); free(inc);
        return -1;
    }
    for (int64_t f = 0; f < F; ++f) {
        off[m->face_owner[f] + 1] += 1;
        if (m->face_neigh[f] >= 0) off[m->face_neigh[f] + 1] += 1;
    }
    for (int64_t c = 0; c < N; ++c) off[c + 1] += off[c];
    {
        int64_t *pos = (int64_t *)malloc(sizeof(int64_t) * N);
        if (!pos) { free(AF); free(live); free(off); free(inc); return -1; }
        memcpy(pos, off, sizeof(int64_t) * N);
        for (int64_t f = 0; f < F; ++f) {
            inc[pos[m->face_owner[f]]++] = 2 * f;
            if (m->face_neigh[f] >= 0) inc[pos[m->face_neigh[f]]++] = 2 * f + 1;
        }
        free(pos);
    }
    double gmax = 0.0;
    int64_t gp = 0, gr = 0;
#pragma omp parallel num_threads(nthreads) if (nthreads > 1)
    {
        double lmax = 0.0;
        int64_t lp = 0, lr = 0;
#pragma omp for schedule(static)
        for (int64_t f = 0; f < F; ++f)
            live[f] = (uint8_t)orc_face_af(s, U, f, &AF[ORC_NF * f], &lmax, &lp, &lr);
#pragma omp for schedule(static)
        for (int64_t c = 0; c < N; ++c) {
            for (int k = 0; k < ORC_NF; ++k) RHS[ORC_NF * c + k] = 0.;
            for (int64_t e = off[c]; e < off[c + 1]; ++e) {
                const int64_t f = inc[e] >> 1;
                if (!(live[f] & ((inc[e] & 1) ? 2 : 1))) continue;
                if (inc[e] & 1)
                    for (int k = 0; k < ORC_NF; ++k) RHS[ORC_NF * c + k] += AF[ORC_NF * f + k];
                else
                    for (int k = 0; k < ORC_NF; ++k) RHS[ORC_NF * c + k] -= AF[ORC_NF * f + k];
            }
        }
#pragma omp critical
        {
            gmax = (lmax < gmax) ? gmax : lmax;
            gp += lp;
            gr += lr;
        }
    }
    s->neg_p += gp;
    s->neg_rho += gr;
    *max_eig = gmax;
    free(AF); free(live); free(off); free(inc);
    return 0;
}

/* The solver's face list (owner, neighbour, owner-outward normal, area) and
 * cell volumes: what an explicit-mesh caller passes for the same grid. */
void orc_solver_face_list(const orc_solver *s, int64_t *owner, int64_t *neigh, double *normal, double *area)
{
    const orc_mesh *m = &s->mesh;
    memcpy(owner, m->face_owner, sizeof(int64_t) * m->nfaces);
    memcpy(neigh, m->face_neigh, sizeof(int64_t) * m->nfaces);
    memcpy(normal, m->face_normal, sizeof(double) * 3 * m->nfaces);
    for (int64_t f = 0; f < m->nfaces; ++f) area[f] = m->face_area ? m->face_area[f] : m->area;
}
void orc_solver_cell_volumes(const orc_solver *s, double *vol)
{
    for (int64_t c = 0; c < s->mesh.ncells; ++c) vol[c] = ORC_VOL(&s->mesh, c);
}
