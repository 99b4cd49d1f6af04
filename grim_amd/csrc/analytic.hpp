// Analytic Jacobian of the per-zone residual (ideal MHD) and the fast FP64
// residual of the mixed-precision Newton path.
//
// Identities used (normal-observer split, beta^i = alpha^2 g^0i,
// beta_j = g_0j, gamma_ij = g_ij):
//   u^0 = gamma/alpha, u_i = gamma_ij utilde^j, u_0 = -alpha gamma + g_0j utilde^j
//   b^0 = B^i u_i,  b_mu = (B_mu + b^0 u_mu)/u^0  with B_mu = g_mu,j B^j
//   b^mu b_mu = (B^i B_i + (b^0)^2)/(u^0)^2
// so the t-row of the stress tensor needs no contravariant spatial vector.
//
// Tangent (forward mode, by hand).  For R = s (U(P) - rhs), s = mask/dt, the
// magnetic field held fixed, the columns of dU/dP are:
//   rho:  dT^0_nu = u^0 u_nu,                        dU_rho = g u^0
//   u:    dT^0_nu = Gamma u^0 u_nu + (Gamma-1) d_0nu,  dU_rho = 0
//   utilde^i, with gamma_i = utilde_i / gamma:
//     du^0  = gamma_i / alpha
//     du_nu = g_nu,i - alpha gamma_i d_0nu
//     db^0  = B_i
//     d(b^2)= 2 (b^0 (b_i - alpha gamma_i b^0) - (b.b) du^0) / u^0
//     dT^0_nu = c1 u_nu + c2 du_nu + c3 b_nu + 1/2 d(b^2) d_0nu
//       c1 = d(b^2) u^0 + w du^0 - b^0 db^0 / u^0
//       c2 = w u^0 - (b^0)^2 / u^0
//       c3 = b^0 du^0 / u^0 - db^0
//     dU_rho = g rho du^0
//   and dU_U = g dT^0_0 + dU_rho, dU_Ui = g dT^0_i.
// (Expanding db_nu = (db^0 u_nu + b^0 du_nu - b_nu du^0)/u^0 in
//  dT^0_nu = d(b^2) u^0 u_nu + w du^0 u_nu + w u^0 du_nu - db^0 b_nu - b^0 db_nu
//  + 1/2 d(b^2) d_0nu gives the c1, c2, c3 grouping.)
// Floors: set() uses max(P, floor); the derivative follows the one-sided
// forward difference of grim's FD rule (1 when P >= floor, else 0).
//
// The primal quantities are recomputed in the tangent's precision T (float
// for the mixed path, double for the analytic FP64 path).
#pragma once
#include <hip/hip_runtime.h>

#include "zone_physics.hpp"

namespace grim {

template <class T, class L>
__device__ __forceinline__ void analyticJacobianIdeal(const double* Pd, const ZoneGeom& Gd,
                                                      const DevParams& prm, double scale,
                                                      T (&A)[L::NF][L::NF]) {
  static_assert(!L::EMHD, "ideal-MHD tangent");
  T gc[4][4];
#pragma unroll
  for (int m = 0; m < 4; m++)
#pragma unroll
    for (int k = 0; k < 4; k++) gc[m][k] = (T)Gd.cov(m, k);
  const T alpha = (T)Gd.alpha, g = (T)Gd.g;
  const T sRho = (Pd[RHO] >= prm.rhoFloor) ? T(1) : T(0);
  const T sU = (Pd[UU] >= prm.uFloor) ? T(1) : T(0);
  const T rho = (T)smax(Pd[RHO], prm.rhoFloor);
  const T u = (T)smax(Pd[UU], prm.uFloor);
  const T gam = (T)prm.gam;
  const T p = (gam - T(1)) * u;
  const T ut[3] = {(T)Pd[U1], (T)Pd[U2], (T)Pd[U3]};
  const T B[3] = {(T)Pd[L::B1], (T)Pd[L::B2], (T)Pd[L::B3]};

  // primal
  T utLow[3], BLow[3];
#pragma unroll
  for (int i = 0; i < 3; i++) {
    utLow[i] = gc[i + 1][1] * ut[0] + gc[i + 1][2] * ut[1] + gc[i + 1][3] * ut[2];
    BLow[i] = gc[i + 1][1] * B[0] + gc[i + 1][2] * B[1] + gc[i + 1][3] * B[2];
  }
  const T lor = sqrt(T(1) + ut[0] * utLow[0] + ut[1] * utLow[1] + ut[2] * utLow[2]);
  const T invLor = T(1) / lor;
  const T invAlpha = T(1) / alpha;
  const T u0 = lor * invAlpha;
  const T invU0 = alpha * invLor;
  T uCov[4];
  uCov[0] = -alpha * lor + gc[0][1] * ut[0] + gc[0][2] * ut[1] + gc[0][3] * ut[2];
#pragma unroll
  for (int i = 0; i < 3; i++) uCov[i + 1] = utLow[i];
  const T b0 = B[0] * utLow[0] + B[1] * utLow[1] + B[2] * utLow[2];
  T bCov[4];
  bCov[0] = (gc[0][1] * B[0] + gc[0][2] * B[1] + gc[0][3] * B[2] + b0 * uCov[0]) * invU0;
#pragma unroll
  for (int i = 0; i < 3; i++) bCov[i + 1] = (BLow[i] + b0 * utLow[i]) * invU0;
  const T BB = B[0] * BLow[0] + B[1] * BLow[1] + B[2] * BLow[2];
  const T bb = (BB + b0 * b0) * invU0 * invU0;
  const T bsq = bb + (T)prm.bSqrFloor;
  const T w = rho + u + p + bsq;
  const T s = (T)scale;
  const T gs = g * s;

  // rho and u columns
  {
    const T a0 = sRho * u0;
    A[RHO][RHO] = gs * a0;
    A[UU][RHO] = gs * (a0 * uCov[0] + a0);
#pragma unroll
    for (int i = 0; i < 3; i++) A[U1 + i][RHO] = gs * (a0 * uCov[i + 1]);
    const T a1 = gam * sU * u0;
    A[RHO][UU] = T(0);
    A[UU][UU] = gs * (a1 * uCov[0] + (gam - T(1)) * sU);
#pragma unroll
    for (int i = 0; i < 3; i++) A[U1 + i][UU] = gs * (a1 * uCov[i + 1]);
  }
  // velocity columns
  const T c2 = w * u0 - b0 * b0 * invU0;
  const T gRho = g * rho;
#pragma unroll
  for (int i = 0; i < 3; i++) {
    const T gi = utLow[i] * invLor;
    const T du0 = gi * invAlpha;
    const T db0 = BLow[i];
    const T bdu = bCov[i + 1] - alpha * gi * b0;
    const T dbsq = T(2) * (b0 * bdu - bb * du0) * invU0;
    const T c1 = dbsq * u0 + w * du0 - b0 * db0 * invU0;
    const T c3 = b0 * du0 * invU0 - db0;
    T dT[4];
#pragma unroll
    for (int nu = 0; nu < 4; nu++) dT[nu] = c1 * uCov[nu] + c2 * gc[nu][i + 1] + c3 * bCov[nu];
    dT[0] = dT[0] - c2 * alpha * gi + T(0.5) * dbsq;
    const T dUrho = gRho * du0;
    A[RHO][U1 + i] = s * dUrho;
    A[UU][U1 + i] = s * (g * dT[0] + dUrho);
#pragma unroll
    for (int k = 0; k < 3; k++) A[U1 + k][U1 + i] = gs * dT[k + 1];
  }
}

// Residual of the mixed-precision path (ideal MHD): FP64 throughout, the same
// mathematics as residual() evaluated through the identities above, with one
// FP64 division (1/gamma) and one square root per evaluation instead of nine
// divisions; 1/alpha and mask/dt come in per zone.  Rounding differs from the
// grim-order residual at the 1e-16 level, far below the 1e-10 tolerance.
template <class L>
__device__ __forceinline__ void residualFastIdeal(const double* P, const ZoneGeom& G,
                                                  const ZoneConst<L>& c, const DevParams& prm,
                                                  double invAlpha, double scale, double* R) {
  static_assert(!L::EMHD, "ideal-MHD residual");
  const double rho = smax(P[RHO], prm.rhoFloor);
  const double u = smax(P[UU], prm.uFloor);
  const double ut[3] = {P[U1], P[U2], P[U3]};
  const double B[3] = {P[L::B1], P[L::B2], P[L::B3]};
  const double p = (prm.gam - 1.) * u;
  double utLow[3], BLow[3];
#pragma unroll
  for (int i = 0; i < 3; i++) {
    utLow[i] = G.cov(i + 1, 1) * ut[0] + G.cov(i + 1, 2) * ut[1] + G.cov(i + 1, 3) * ut[2];
    BLow[i] = G.cov(i + 1, 1) * B[0] + G.cov(i + 1, 2) * B[1] + G.cov(i + 1, 3) * B[2];
  }
  const double lor = sqrt(1. + ut[0] * utLow[0] + ut[1] * utLow[1] + ut[2] * utLow[2]);
  const double u0 = lor * invAlpha;
  const double invU0 = G.alpha / lor;
  const double uCov0 = -G.alpha * lor + G.cov(0, 1) * ut[0] + G.cov(0, 2) * ut[1] + G.cov(0, 3) * ut[2];
  const double b0 = B[0] * utLow[0] + B[1] * utLow[1] + B[2] * utLow[2];
  const double bCov0 = (G.cov(0, 1) * B[0] + G.cov(0, 2) * B[1] + G.cov(0, 3) * B[2] + b0 * uCov0) * invU0;
  const double BB = B[0] * BLow[0] + B[1] * BLow[1] + B[2] * BLow[2];
  const double bsq = (BB + b0 * b0) * invU0 * invU0 + prm.bSqrFloor;
  const double wu0 = (rho + u + p + bsq) * u0;
  const double Urho = G.g * (rho * u0);
  const double T00 = wu0 * uCov0 - b0 * bCov0 + (p + 0.5 * bsq);
  R[RHO] = (Urho - c.rhs[RHO]) * scale;
  R[UU] = (G.g * T00 + Urho - c.rhs[UU]) * scale;
#pragma unroll
  for (int i = 0; i < 3; i++) {
    const double bCovi = (BLow[i] + b0 * utLow[i]) * invU0;
    R[U1 + i] = (G.g * (wu0 * utLow[i] - b0 * bCovi) - c.rhs[U1 + i]) * scale;
  }
}

// Row and column equilibration around the FP32 LU: solve (Dr A Dc) y = Dr b,
// x = Dc y, with Dc = |P_k| (the natural scale of each primitive) and Dr the
// inverse of each scaled row's largest entry.
template <class T, int N>
__device__ __forceinline__ void equilibratedSolve(T (&A)[N][N], T (&b)[N], const double* P) {
  T cs[N];
#pragma unroll
  for (int k = 0; k < N; k++) {
    const T a = (T)fabs(P[k]);
    cs[k] = (a > T(0) && a < T(1e30)) ? a : T(1);
  }
#pragma unroll
  for (int i = 0; i < N; i++) {
    T m = T(0);
#pragma unroll
    for (int k = 0; k < N; k++) {
      A[i][k] = A[i][k] * cs[k];
      m = fmax(m, fabs(A[i][k]));
    }
    const T r = (m > T(0)) ? T(1) / m : T(1);
#pragma unroll
    for (int k = 0; k < N; k++) A[i][k] = A[i][k] * r;
    b[i] = b[i] * r;
  }
  luSolve<T, N>(A, b);
#pragma unroll
  for (int k = 0; k < N; k++) b[k] = b[k] * cs[k];
}

}  // namespace grim
