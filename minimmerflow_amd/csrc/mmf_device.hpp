// mmf_device.hpp -- device-side physics of the finite-volume step (gfx950).
//
// Every expression keeps the operand order of the reference source it cites,
// and the library is compiled with -ffp-contract=off, so FP64 results are
// the correctly rounded values of the same operation sequence the reference
// executes.  FP64 division and sqrt on CDNA are correctly rounded
// (IEEE-754) in the default (non fast-math) lowering.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mmf {

constexpr int NF = 5;
constexpr double GAMMA = 1.4;  // constants.hpp:55

// constants.hpp:37-47
enum { FID_P = 0, FID_U = 1, FID_V = 2, FID_W = 3, FID_T = 4 };
enum { FID_RHO = 0, FID_RHO_U = 1, FID_RHO_V = 2, FID_RHO_W = 3, FID_RHO_E = 4 };

// constants.hpp:58-62 (+ periodic extension); identical to mmf.h
enum { BC_NONE = -1, BC_FREE_FLOW = 0, BC_REFLECTING = 1, BC_WALL = 2, BC_DIRICHLET = 3, BC_PERIODIC = 4 };
enum { CELL_FLUID = 1, CELL_SOLVED = 2 };

// utils.cpp:48-63
__device__ __forceinline__ void cons2prim(const double *c, double *p) {
  double K = (c[FID_RHO_U] * c[FID_RHO_U] + c[FID_RHO_V] * c[FID_RHO_V] + c[FID_RHO_W] * c[FID_RHO_W]) /
             (c[FID_RHO] * c[FID_RHO]);
  p[FID_T] = (2.0 * c[FID_RHO_E] / c[FID_RHO] - K) / (2.0 / (GAMMA - 1.0));
  p[FID_U] = c[FID_RHO_U] / c[FID_RHO];
  p[FID_V] = c[FID_RHO_V] / c[FID_RHO];
  p[FID_W] = c[FID_RHO_W] / c[FID_RHO];
  p[FID_P] = c[FID_RHO] * p[FID_T];
}

// utils.cpp:71-83
__device__ __forceinline__ void prim2cons(const double *p, double *c) {
  c[FID_RHO] = p[FID_P] / p[FID_T];
  c[FID_RHO_U] = c[FID_RHO] * p[FID_U];
  c[FID_RHO_V] = c[FID_RHO] * p[FID_V];
  c[FID_RHO_W] = c[FID_RHO] * p[FID_W];
  c[FID_RHO_E] = c[FID_RHO] * p[FID_T] / (GAMMA - 1.0) +
                 0.5 * c[FID_RHO] * (p[FID_U] * p[FID_U] + p[FID_V] * p[FID_V] + p[FID_W] * p[FID_W]);
}

// utils.cpp:37-40
__device__ __forceinline__ double normal_velocity(const double *p, const double *n) {
  return (p[FID_U] * n[0] + p[FID_V] * n[1] + p[FID_W] * n[2]);
}

// euler.cpp:83-113 for an arbitrary normal (batched point function)
__device__ __forceinline__ void eval_fluxes(const double *cons, const double *prim, const double *n, double *f,
                                            int *neg_p, int *neg_rho) {
  double u = prim[FID_U], v = prim[FID_V], w = prim[FID_W];
  double vel2 = u * u + v * v + w * w;
  double un = normal_velocity(prim, n);
  double p = prim[FID_P];
  if (p < 0.) *neg_p += 1;
  double rho = cons[FID_RHO];
  if (rho < 0.) *neg_rho += 1;
  double eto = p / (GAMMA - 1.) + 0.5 * rho * vel2;
  double massFlux = rho * un;
  f[0] = massFlux;
  f[1] = massFlux * u + p * n[0];
  f[2] = massFlux * v + p * n[1];
  f[3] = massFlux * w + p * n[2];
  f[4] = un * (eto + p);
}

// euler.cpp:42-73 for an arbitrary normal (batched point function)
__device__ __forceinline__ void eval_splitting(const double *L, const double *R, const double *n, double *F,
                                               double *lambda, int *neg_p, int *neg_rho) {
  double pL[NF], pR[NF], fL[NF], fR[NF];
  cons2prim(L, pL);
  cons2prim(R, pR);
  eval_fluxes(L, pL, n, fL, neg_p, neg_rho);
  eval_fluxes(R, pR, n, fR, neg_p, neg_rho);
  double lambdaL = fabs(normal_velocity(pL, n)) + sqrt(GAMMA * pL[FID_T]);
  double lambdaR = fabs(normal_velocity(pR, n)) + sqrt(GAMMA * pR[FID_T]);
  double lam = (lambdaR < lambdaL) ? lambdaL : lambdaR;  // std::max(lambdaR, lambdaL)
  *lambda = lam;
#pragma unroll
  for (int k = 0; k < NF; ++k) F[k] = 0.5 * ((fR[k] + fL[k]) - lam * (R[k] - L[k]));
}

// euler.cpp:318-337 for an arbitrary normal
__device__ __forceinline__ void reflect_general(const double *n, const double *cons, double *out) {
  double prim[NF];
  cons2prim(cons, prim);
  double u0 = prim[FID_U], u1 = prim[FID_V], u2 = prim[FID_W];
  double un = normal_velocity(prim, n);
  double un0 = un * n[0], un1 = un * n[1], un2 = un * n[2];
  prim[FID_U] = u0 - 2 * un0;
  prim[FID_V] = u1 - 2 * un1;
  prim[FID_W] = u2 - 2 * un2;
  prim2cons(prim, out);
}

// ---------------------------------------------------------------------------
// Axis-aligned form used by the stage kernels.
//
// On a Cartesian grid every normal is +-e_a.  The flux is evaluated with the
// normal +e_a from the lower (L) to the upper (R) cell.  Because the Rusanov
// flux is exactly antisymmetric, F(L,R,n) == -F(R,L,-n) bit for bit, and
// RHS -= A*F on the owner / += A*F on the neighbour yields the same bits
// whichever side bitpit names the owner.  For n = +e_a, u.n == u_a exactly and
// p*n_b == 0 for b != a, so the per-side quantities below are the reference's
// own intermediate values (up to the sign of zero).
// ---------------------------------------------------------------------------

// A cell's state with everything the flux needs, computed once per state.
struct Tup {
  double c[NF];     // conservative
  double vel[3];    // u, v, w
  double p;         // pressure
  double a;         // sqrt(gamma T)
  double H;         // eto + p, eto = p/(gamma-1) + 0.5 rho |u|^2 (euler.cpp:103)
};

// The tuple of state c from its primitives prim = cons2prim(c).
__device__ __forceinline__ void make_tup_prim(const double *c, const double *prim, Tup &t) {
#pragma unroll
  for (int k = 0; k < NF; ++k) t.c[k] = c[k];
  t.vel[0] = prim[FID_U];
  t.vel[1] = prim[FID_V];
  t.vel[2] = prim[FID_W];
  t.p = prim[FID_P];
  t.a = sqrt(GAMMA * prim[FID_T]);                                    // euler.cpp:60
  double vel2 = t.vel[0] * t.vel[0] + t.vel[1] * t.vel[1] + t.vel[2] * t.vel[2];  // euler.cpp:90
  double eto = t.p / (GAMMA - 1.) + 0.5 * c[FID_RHO] * vel2;         // euler.cpp:103
  t.H = eto + t.p;
}

__device__ __forceinline__ void make_tup(const double *c, Tup &t) {
  double prim[NF];
  cons2prim(c, prim);
  make_tup_prim(c, prim, t);
}

// Mirror of the velocity component along axis a (euler.cpp:318-337 with
// n = +-e_a: u - 2 (u.n) n), reflecting and wall BC (euler.cpp:347-356).
__device__ __forceinline__ void reflect_axis(const double *cons, int a, double *out) {
  double prim[NF];
  cons2prim(cons, prim);
  double ua = prim[FID_U + a];
  prim[FID_U + a] = ua - 2 * ua;
  prim2cons(prim, out);
}

// Physical flux along +e_a of one side (euler.cpp:83-113).
__device__ __forceinline__ void side_flux(const Tup &t, int a, double *f) {
  double un = t.vel[a];
  double massFlux = t.c[FID_RHO] * un;
  f[0] = massFlux;
  f[1] = massFlux * t.vel[0];
  f[2] = massFlux * t.vel[1];
  f[3] = massFlux * t.vel[2];
  f[1 + a] = f[1 + a] + t.p;
  f[4] = un * t.H;
}

// euler.cpp:42-73 along +e_a.  The per-side fluxes (side_flux) are formed
// component by component so that few of them are live at a time; the
// operations and their order are those of side_flux + the splitting sum.
__device__ __forceinline__ double face_flux(const Tup &L, const Tup &R, int a, double *F) {
  const double unL = L.vel[a], unR = R.vel[a];
  const double lambdaL = fabs(unL) + L.a;
  const double lambdaR = fabs(unR) + R.a;
  const double lam = (lambdaR < lambdaL) ? lambdaL : lambdaR;
  const double mfL = L.c[FID_RHO] * unL;
  const double mfR = R.c[FID_RHO] * unR;
  F[0] = 0.5 * ((mfR + mfL) - lam * (R.c[0] - L.c[0]));
#pragma unroll
  for (int b = 0; b < 3; ++b) {
    double fLb = mfL * L.vel[b];
    double fRb = mfR * R.vel[b];
    if (b == a) {
      fLb = fLb + L.p;
      fRb = fRb + R.p;
    }
    F[1 + b] = 0.5 * ((fRb + fLb) - lam * (R.c[1 + b] - L.c[1 + b]));
  }
  F[4] = 0.5 * ((unR * R.H + unL * L.H) - lam * (R.c[4] - L.c[4]));
  return lam;
}

// The residual contribution A*F of a face (euler.cpp:239,245), F as in
// face_flux.  With A a power of two, A*(0.5*x) is computed as (0.5*A)*x: the
// same double (both are exact scalings), one multiply fewer per component.
__device__ __forceinline__ double face_flux_area(const Tup &L, const Tup &R, int a, bool pow2A, double A,
                                                 double halfA, double *AF) {
  const double unL = L.vel[a], unR = R.vel[a];
  const double lambdaL = fabs(unL) + L.a;
  const double lambdaR = fabs(unR) + R.a;
  const double lam = (lambdaR < lambdaL) ? lambdaL : lambdaR;
  const double mfL = L.c[FID_RHO] * unL;
  const double mfR = R.c[FID_RHO] * unR;
  double X[NF];
  X[0] = (mfR + mfL) - lam * (R.c[0] - L.c[0]);
#pragma unroll
  for (int b = 0; b < 3; ++b) {
    double fLb = mfL * L.vel[b];
    double fRb = mfR * R.vel[b];
    if (b == a) {
      fLb = fLb + L.p;
      fRb = fRb + R.p;
    }
    X[1 + b] = (fRb + fLb) - lam * (R.c[1 + b] - L.c[1 + b]);
  }
  X[4] = (unR * R.H + unL * L.H) - lam * (R.c[4] - L.c[4]);
#pragma unroll
  for (int k = 0; k < NF; ++k) AF[k] = pow2A ? halfA * X[k] : A * (0.5 * X[k]);
  return lam;
}

// true when the state has a negative pressure or density (euler.cpp:94-101)
__device__ __forceinline__ bool tup_negative(const Tup &t) { return (t.p < 0.) | (t.c[FID_RHO] < 0.); }

// Positive doubles order like their bit patterns: max-reduce through u64.
__device__ __forceinline__ unsigned long long dbits(double x) { return (unsigned long long)__double_as_longlong(x); }

// Running maximum of face eigenvalues (non-negative doubles, euler.cpp:234's
// std::max) compared as u64 bits: for finite values the same double as
// (l < m) ? m : l, and a NaN (whose bits lie above every finite double and
// +inf) is kept instead of being replaced by the next face's value, so a
// non-finite state reaches the stage maximum and the step's time control.
__device__ __forceinline__ double lam_acc(double m, double l) { return (dbits(l) > dbits(m)) ? l : m; }

// Bits at and above those of +inf: the stage maximum is not finite.
constexpr unsigned long long NONFINITE_BITS = 0x7ff0000000000000ull;

}  // namespace mmf
