// mmf_problem.cpp -- problem catalogue on the host (problem.cpp of the
// reference): initial conditions and exact solutions sampled at cell
// centroids, border BC types and Dirichlet data, domain defaults.
//
// Host-only setup code: it runs once per case, outside the time loop, so the
// FP64 exp/pow of the vortex solution are the host libm's (the reference
// evaluates them on the host too, problem.cpp:293,314).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "mmf.h"
#include "mmf_internal.hpp"

namespace {

constexpr double GAMMA = 1.4;  // constants.hpp:55

// utils.cpp:71-83
void prim2cons(const double *p, double *c) {
  c[0] = p[0] / p[4];
  c[1] = c[0] * p[1];
  c[2] = c[0] * p[2];
  c[3] = c[0] * p[3];
  c[4] = c[0] * p[4] / (GAMMA - 1.0) + 0.5 * c[0] * (p[1] * p[1] + p[2] * p[2] + p[3] * p[3]);
}

bool is_vortex(int t) { return t == MMF_PROBLEM_VORTEX_XY || t == MMF_PROBLEM_VORTEX_ZX || t == MMF_PROBLEM_VORTEX_YZ; }
bool is_sod(int t) { return t == MMF_PROBLEM_SOD_X || t == MMF_PROBLEM_SOD_Y || t == MMF_PROBLEM_SOD_Z; }

// problem.cpp:247-392 (+ strength/centre/periodic image, Sod radius/states)
void exact_point(const mmf_problem &pb, const double *pin, double t, double *cons) {
  double point[3] = {pin[0], pin[1], pin[2]};
  double prim[5];
  const int type = pb.type;
  if (is_vortex(type)) {
    const double BETA = pb.vortex_beta, P_INF = 1.0, T_INF = 1.0;
    double U_INF = 1.0, V_INF = 1.0, W_INF = 1.0;
    if (type == MMF_PROBLEM_VORTEX_XY) W_INF = 0.;
    else if (type == MMF_PROBLEM_VORTEX_ZX) V_INF = 0.;
    else U_INF = 0.;
    point[0] -= t * U_INF;  // problem.cpp:273-275
    point[1] -= t * V_INF;
    point[2] -= t * W_INF;
    for (int d = 0; d < 3; ++d) {
      point[d] -= pb.vortex_center[d];
      if (pb.vortex_periodic && d < pb.dim) {
        const double L = pb.domain_length;
        point[d] -= L * std::floor(point[d] / L + 0.5);
      }
    }
    double csi, eta;
    if (type == MMF_PROBLEM_VORTEX_XY) { csi = point[0]; eta = point[1]; }
    else if (type == MMF_PROBLEM_VORTEX_ZX) { csi = point[2]; eta = point[0]; }
    else { csi = point[1]; eta = point[2]; }
    double r = std::sqrt(csi * csi + eta * eta);
    double shape = BETA / (2 * M_PI) * std::exp(0.5 * (1 - r * r));  // problem.cpp:293
    double du = 0., dv = 0., dw = 0.;
    if (type == MMF_PROBLEM_VORTEX_XY) { du = -eta * shape; dv = csi * shape; }
    else if (type == MMF_PROBLEM_VORTEX_ZX) { dw = -eta * shape; du = csi * shape; }
    else { dv = -eta * shape; dw = csi * shape; }
    const double dT = -(GAMMA - 1.0) / (2 * GAMMA) * shape * shape;          // problem.cpp:312
    const double dp = std::pow(T_INF + dT, GAMMA / (GAMMA - 1.0)) - 1.0;     // problem.cpp:314
    prim[0] = P_INF + dp;
    prim[1] = U_INF + du;
    prim[2] = V_INF + dv;
    prim[3] = (pb.dim == 3) ? W_INF + dw : 0;
    prim[4] = T_INF + dT;
  } else if (type == MMF_PROBLEM_RADSOD) {  // problem.cpp:329-345
    double r = std::sqrt(point[0] * point[0] + point[1] * point[1] + point[2] * point[2]);
    prim[1] = prim[2] = prim[3] = 0.0;
    if (r < pb.sod_r0) { prim[0] = pb.sod_p_in; prim[4] = pb.sod_t_in; }
    else { prim[0] = pb.sod_p_out; prim[4] = pb.sod_t_out; }
  } else if (is_sod(type)) {  // problem.cpp:347-380
    prim[1] = prim[2] = prim[3] = 0.0;
    double r = (type == MMF_PROBLEM_SOD_X) ? point[0] : (type == MMF_PROBLEM_SOD_Y) ? point[1] : point[2];
    if (r < 0.) { prim[0] = 1.0; prim[4] = 1.0 / 1.0; }
    else { prim[0] = 0.1; prim[4] = 0.1 / 0.125; }
  } else {  // FFSTEP (problem.cpp:381-388) and free stream
    for (int k = 0; k < 5; ++k) prim[k] = pb.fs_prim[k];
  }
  prim2cons(prim, cons);
}

}  // namespace

extern "C" {

int mmf_problem_defaults(mmf_problem *pb, int32_t type, int32_t dim) {
  if (!pb || type < 0 || type > MMF_PROBLEM_FREESTREAM || (dim != 2 && dim != 3)) return MMF_ERR_INVALID;
  std::memset(pb, 0, sizeof(*pb));
  pb->type = type;
  // problem.cpp:75-84: the zx / yz vortices are always 3D
  pb->dim = (type == MMF_PROBLEM_VORTEX_ZX || type == MMF_PROBLEM_VORTEX_YZ) ? 3 : dim;
  pb->vortex_beta = 5.0;  // problem.cpp:257
  if (is_vortex(type)) {  // problem.cpp:99-103,128-131
    pb->domain_origin[0] = pb->domain_origin[1] = pb->domain_origin[2] = -5.0;
    pb->domain_length = 10.0;
  } else if (is_sod(type)) {  // problem.cpp:105-109,134-137
    pb->domain_origin[0] = pb->domain_origin[1] = pb->domain_origin[2] = -1.0;
    pb->domain_length = 2.0;
  } else if (type == MMF_PROBLEM_RADSOD) {  // problem.cpp:140-141
    pb->domain_length = 8.0;
  }
  pb->sod_r0 = 1.0;  // problem.cpp:336-342
  pb->sod_p_in = 1.0;
  pb->sod_t_in = 1.0 / 1.0;
  pb->sod_p_out = 0.1;
  pb->sod_t_out = 0.1 / 0.125;
  pb->fs_prim[0] = 1.0;  // problem.cpp:383-387
  pb->fs_prim[1] = 3.0;
  pb->fs_prim[2] = 0.0;
  pb->fs_prim[3] = 0.0;
  pb->fs_prim[4] = 1.0 / 1.4;
  return MMF_OK;
}

// problem.cpp:175-211
int mmf_problem_default_end_time(const mmf_problem *pb, double *tend) {
  if (!pb || !tend) return MMF_ERR_INVALID;
  if (is_vortex(pb->type)) *tend = (pb->dim == 2) ? 2. : 1.;
  else if (is_sod(pb->type)) *tend = 0.3;
  else *tend = 4.;
  return MMF_OK;
}

int mmf_problem_exact(const mmf_problem *pb, const double *points, int64_t count, double t, double *cons) {
  if (!pb || (count && (!points || !cons)) || count < 0) return MMF_ERR_INVALID;
  const unsigned hw = std::thread::hardware_concurrency();
  const int64_t nthreads = (count < 65536 || hw < 2) ? 1 : (int64_t)std::min(hw, 16u);
  auto work = [&](int64_t b, int64_t e) {
    for (int64_t i = b; i < e; ++i) exact_point(*pb, points + 3 * i, t, cons + 5 * i);
  };
  if (nthreads == 1) {
    work(0, count);
    return MMF_OK;
  }
  std::vector<std::thread> th;
  const int64_t chunk = (count + nthreads - 1) / nthreads;
  for (int64_t w = 0; w < nthreads; ++w) {
    int64_t b = w * chunk, e = std::min(count, b + chunk);
    if (b < e) th.emplace_back(work, b, e);
  }
  for (auto &x : th) x.join();
  return MMF_OK;
}

// problem.cpp:404-438 (type) and 450-477 (info) for one border face
int mmf_problem_border_bc(const mmf_problem *pb, const double *face_centroid, int32_t *bc, double *info) {
  if (!pb || !face_centroid || !bc) return MMF_ERR_INVALID;
  int code = MMF_BC_FREE_FLOW;
  if (pb->type == MMF_PROBLEM_RADSOD) code = MMF_BC_REFLECTING;
  else if (pb->type == MMF_PROBLEM_FFSTEP) {
    if (face_centroid[0] < 1e-10) code = MMF_BC_DIRICHLET;
    else if (face_centroid[0] > 3.2 - 1e-10) code = MMF_BC_FREE_FLOW;
    else code = MMF_BC_REFLECTING;
  }
  *bc = code;
  if (info)
    for (int k = 0; k < 5; ++k) info[k] = (code == MMF_BC_DIRICHLET) ? pb->fs_prim[k] : 0.0;
  return MMF_OK;
}

// Fills desc->side_bc / side_info from the problem: the per-face rule of
// problem.cpp:404-477 evaluated at one face of each side (the centre of the
// side; for FFSTEP, whose rule depends on x only, a face of the first
// column).  A side whose faces do not share one condition (FFSTEP's y and z
// sides on a domain reaching past x = 3.2) takes that face's here;
// mmf_problem_border_bcs gives every face its own.
int mmf_problem_side_bcs(const mmf_problem *pb, mmf_grid_desc *desc) {
  if (!pb || !desc) return MMF_ERR_INVALID;
  for (int s = 0; s < 6; ++s) {
    const int a = s / 2;
    double fc[3];
    for (int d = 0; d < 3; ++d) fc[d] = desc->origin[d] + 0.5 * desc->h * (double)desc->n[d];
    fc[a] = desc->origin[a] + ((s & 1) ? desc->h * (double)desc->n[a] : 0.0);
    if (pb->type == MMF_PROBLEM_FFSTEP && a != 0) fc[0] = desc->origin[0] + 0.5 * desc->h;
    int32_t bc = 0;
    int rc = mmf_problem_border_bc(pb, fc, &bc, desc->side_info[s]);
    if (rc != MMF_OK) return rc;
    desc->side_bc[s] = bc;
  }
  return MMF_OK;
}

// problem.cpp:404-477 at the centroid of every border face of desc's grid
// (main.cpp:251-260), in mmf_set_border_bcs's layout: sides x-, x+, y-, y+
// (, z-, z+), a side's faces ordered by its two other axes, the lower one
// fastest; face centroid = origin + (i + 1/2) h along those axes, the origin
// or origin + n h across (z = origin z in 2D).
int mmf_problem_border_bcs(const mmf_problem *pb, const mmf_grid_desc *desc, int32_t *bc, double *info) {
  if (!pb || !bc) return MMF_ERR_INVALID;
  mmf::GridDev g;
  const int rcg = mmf::grid_from_desc(desc, g);
  if (rcg != MMF_OK) return rcg;
  int64_t e = 0;
  for (int s = 0; s < 2 * g.dim; ++s) {
    const int a = s / 2;
    const int b0 = (a == 0) ? 1 : 0, b1 = (a == 2) ? 1 : 2;
    for (int64_t j = 0; j < g.n[b1]; ++j)
      for (int64_t i = 0; i < g.n[b0]; ++i, ++e) {
        double fc[3];
        for (int d = 0; d < 3; ++d) fc[d] = desc->origin[d];
        if (b0 < g.dim) fc[b0] = desc->origin[b0] + ((double)i + 0.5) * desc->h;
        if (b1 < g.dim) fc[b1] = desc->origin[b1] + ((double)j + 0.5) * desc->h;
        fc[a] = desc->origin[a] + ((s & 1) ? (double)g.n[a] : 0.0) * desc->h;
        const int rc = mmf_problem_border_bc(pb, fc, &bc[e], info ? info + 5 * e : nullptr);
        if (rc != MMF_OK) return rc;
      }
  }
  return MMF_OK;
}

// main.cpp:335-344: state = exact solution at t = 0 at every centroid
int mmf_set_initial_state(mmf_ctx *ctx, const mmf_problem *pb) {
  if (!ctx || !pb) return MMF_ERR_INVALID;
  int64_t N = 0;
  int rc = mmf_get_sizes(ctx, &N, nullptr);
  if (rc != MMF_OK) return rc;
  std::vector<double> pts((size_t)(3 * N)), U((size_t)(5 * N));
  rc = mmf_get_centroids(ctx, pts.data());
  if (rc != MMF_OK) return rc;
  rc = mmf_problem_exact(pb, pts.data(), N, 0.0, U.data());
  if (rc != MMF_OK) return rc;
  return mmf_set_state(ctx, U.data());
}

// main.cpp:550-561 with the exact density at time t
int mmf_error_vs_exact(mmf_ctx *ctx, const mmf_problem *pb, double t, double *l1, double *l2) {
  if (!ctx || !pb) return MMF_ERR_INVALID;
  int64_t N = 0;
  int rc = mmf_get_sizes(ctx, &N, nullptr);
  if (rc != MMF_OK) return rc;
  std::vector<double> pts((size_t)(3 * N)), U((size_t)(5 * N)), rho((size_t)N);
  rc = mmf_get_centroids(ctx, pts.data());
  if (rc != MMF_OK) return rc;
  rc = mmf_problem_exact(pb, pts.data(), N, t, U.data());
  if (rc != MMF_OK) return rc;
  for (int64_t i = 0; i < N; ++i) rho[(size_t)i] = U[(size_t)(5 * i)];
  return mmf_error_norms(ctx, rho.data(), l1, l2);
}

}  // extern "C"
