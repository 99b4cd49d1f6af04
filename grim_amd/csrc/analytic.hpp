// Analytic Jacobian of the per-zone residual (ideal MHD) and the fast FP64
// residual of the mixed-precision Newton path.
//
// Tangent (forward-mode, by hand).  For R = s (U(P) - rhs), s = mask/dt, with
// the magnetic field held fixed, the columns of dU/dP are:
//   rho:  dT^0_nu = u^0 u_nu,                       dU_rho = g u^0
//   u:    dT^0_nu = Gamma u^0 u_nu + (Gamma-1) d_0nu, dU_rho = 0
//   utilde^i (i = 1..3), with gamma_i = d gamma/d utilde^i = utilde_i / gamma
//   (utilde_i lowered with the spatial metric):
//     du^0     = gamma_i / alpha
//     du^j     = delta_ij - gamma_i alpha g^0j
//     du_mu    = g_mu,i - alpha gamma_i delta_mu0
//                (lowering du^nu = gamma_i (-alpha g^0nu) + delta^nu_i and
//                 g_mu,nu g^nu,0 = delta_mu0)
//     db^0     = B_i = g_ij B^j
//     db^mu    = (db^0 u^mu + b^0 du^mu - b^mu du^0) / u^0   (same for b_mu)
//     d(b^2)   = 2 b_mu db^mu
//     dT^0_nu  = d(b^2) u^0 u_nu + w du^0 u_nu + w u^0 du_nu
//                + 1/2 d(b^2) d_0nu - db^0 b_nu - b^0 db_nu
//     dU_rho   = g rho du^0
//   and dU_U = g dT^0_0 + dU_rho, dU_Ui = g dT^0_i.
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
  const T gcon0[4] = {(T)Gd.gcon0[0], (T)Gd.gcon0[1], (T)Gd.gcon0[2], (T)Gd.gcon0[3]};
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
  T uCon[4], uCov[4], bCon[4], bCov[4];
  uCon[0] = lor / alpha;
#pragma unroll
  for (int j = 0; j < 3; j++) uCon[j + 1] = ut[j] - lor * gcon0[j + 1] * alpha;
#pragma unroll
  for (int m = 0; m < 4; m++)
    uCov[m] = gc[m][0] * uCon[0] + gc[m][1] * uCon[1] + gc[m][2] * uCon[2] + gc[m][3] * uCon[3];
  const T invU0 = T(1) / uCon[0];
  bCon[0] = B[0] * uCov[1] + B[1] * uCov[2] + B[2] * uCov[3];
#pragma unroll
  for (int j = 0; j < 3; j++) bCon[j + 1] = (B[j] + bCon[0] * uCon[j + 1]) * invU0;
#pragma unroll
  for (int m = 0; m < 4; m++)
    bCov[m] = gc[m][0] * bCon[0] + gc[m][1] * bCon[1] + gc[m][2] * bCon[2] + gc[m][3] * bCon[3];
  const T bsq = bCon[0] * bCov[0] + bCon[1] * bCov[1] + bCon[2] * bCov[2] + bCon[3] * bCov[3] +
                (T)prm.bSqrFloor;
  const T w = rho + u + p + bsq;
  const T s = (T)scale;
  const T gs = g * s;

  // rho and u columns
  {
    const T a0 = sRho * uCon[0];
    A[RHO][RHO] = gs * a0;
    A[UU][RHO] = gs * (a0 * uCov[0] + a0);
#pragma unroll
    for (int i = 0; i < 3; i++) A[U1 + i][RHO] = gs * (a0 * uCov[i + 1]);
    const T wu = gam * sU;
    const T a1 = wu * uCon[0];
    A[RHO][UU] = T(0);
    A[UU][UU] = gs * (a1 * uCov[0] + (gam - T(1)) * sU);
#pragma unroll
    for (int i = 0; i < 3; i++) A[U1 + i][UU] = gs * (a1 * uCov[i + 1]);
  }
  // velocity columns
  const T wu0 = w * uCon[0];
#pragma unroll
  for (int i = 0; i < 3; i++) {
    const T gi = utLow[i] * invLor;
    const T du0 = gi / alpha;
    T duCon[4], duCov[4], dbCon[4], dbCov[4];
    duCon[0] = du0;
#pragma unroll
    for (int j = 0; j < 3; j++) duCon[j + 1] = ((i == j) ? T(1) : T(0)) - gi * alpha * gcon0[j + 1];
#pragma unroll
    for (int m = 0; m < 4; m++) duCov[m] = gc[m][i + 1];
    duCov[0] = duCov[0] - alpha * gi;
    const T db0 = BLow[i];
#pragma unroll
    for (int m = 0; m < 4; m++) {
      dbCon[m] = (db0 * uCon[m] + bCon[0] * duCon[m] - bCon[m] * du0) * invU0;
      dbCov[m] = (db0 * uCov[m] + bCon[0] * duCov[m] - bCov[m] * du0) * invU0;
    }
    dbCon[0] = db0;
    dbCov[0] = (db0 * uCov[0] + bCon[0] * duCov[0] - bCov[0] * du0) * invU0;
    const T dbsq = T(2) * (bCov[0] * dbCon[0] + bCov[1] * dbCon[1] + bCov[2] * dbCon[2] +
                           bCov[3] * dbCon[3]);
    T dT[4];
#pragma unroll
    for (int nu = 0; nu < 4; nu++)
      dT[nu] = (dbsq * uCon[0] + w * du0) * uCov[nu] + wu0 * duCov[nu] - db0 * bCov[nu] -
               bCon[0] * dbCov[nu];
    dT[0] = dT[0] + T(0.5) * dbsq;
    const T dUrho = g * rho * du0;
    A[RHO][U1 + i] = s * dUrho;
    A[UU][U1 + i] = s * (g * dT[0] + dUrho);
#pragma unroll
    for (int k = 0; k < 3; k++) A[U1 + k][U1 + i] = gs * dT[k + 1];
  }
}

// Residual of the mixed-precision path (ideal MHD), FP64 throughout but with
// the divisions of set() and of (U - rhs)/dt replaced by multiplications with
// reciprocals formed once per evaluation (1/alpha and mask/dt once per zone):
// 2 FP64 divisions instead of 9.  Same mathematics as residual(); rounding
// differs at the 1e-16 level, far below the 1e-10 relative tolerance.
template <class L>
__device__ __forceinline__ void residualFastIdeal(const double* P, const ZoneGeom& G,
                                                  const ZoneConst<L>& c, const DevParams& prm,
                                                  double invAlpha, double scale, double* R) {
  static_assert(!L::EMHD, "ideal-MHD residual");
  const double rho = smax(P[RHO], prm.rhoFloor);
  const double u = smax(P[UU], prm.uFloor);
  const double u1 = P[U1], u2 = P[U2], u3 = P[U3];
  const double B1 = P[L::B1], B2 = P[L::B2], B3 = P[L::B3];
  const double p = (prm.gam - 1.) * u;
  const double lor = sqrt(1. + G.cov(1, 1) * u1 * u1 + G.cov(2, 2) * u2 * u2 + G.cov(3, 3) * u3 * u3 +
                          2. * (G.cov(1, 2) * u1 * u2 + G.cov(1, 3) * u1 * u3 + G.cov(2, 3) * u2 * u3));
  double uCon[4], uCov[4], bCon[4], bCov[4];
  uCon[0] = lor * invAlpha;
  uCon[1] = u1 - lor * G.gcon0[1] * G.alpha;
  uCon[2] = u2 - lor * G.gcon0[2] * G.alpha;
  uCon[3] = u3 - lor * G.gcon0[3] * G.alpha;
#pragma unroll
  for (int mu = 0; mu < 4; mu++)
    uCov[mu] = G.cov(mu, 0) * uCon[0] + G.cov(mu, 1) * uCon[1] + G.cov(mu, 2) * uCon[2] +
               G.cov(mu, 3) * uCon[3];
  const double invU0 = G.alpha / lor;
  bCon[0] = B1 * uCov[1] + B2 * uCov[2] + B3 * uCov[3];
  bCon[1] = (B1 + bCon[0] * uCon[1]) * invU0;
  bCon[2] = (B2 + bCon[0] * uCon[2]) * invU0;
  bCon[3] = (B3 + bCon[0] * uCon[3]) * invU0;
#pragma unroll
  for (int mu = 0; mu < 4; mu++)
    bCov[mu] = G.cov(mu, 0) * bCon[0] + G.cov(mu, 1) * bCon[1] + G.cov(mu, 2) * bCon[2] +
               G.cov(mu, 3) * bCon[3];
  const double bsq = bCon[0] * bCov[0] + bCon[1] * bCov[1] + bCon[2] * bCov[2] + bCon[3] * bCov[3] +
                     prm.bSqrFloor;
  const double wu0 = (rho + u + p + bsq) * uCon[0];
  double T0[4];
#pragma unroll
  for (int nu = 0; nu < 4; nu++) T0[nu] = wu0 * uCov[nu] - bCon[0] * bCov[nu];
  T0[0] = T0[0] + (p + 0.5 * bsq);
  const double Urho = G.g * (rho * uCon[0]);
  R[RHO] = (Urho - c.rhs[RHO]) * scale;
  R[UU] = (G.g * T0[0] + Urho - c.rhs[UU]) * scale;
  R[U1] = (G.g * T0[1] - c.rhs[U1]) * scale;
  R[U2] = (G.g * T0[2] - c.rhs[U2]) * scale;
  R[U3] = (G.g * T0[3] - c.rhs[U3]) * scale;
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
