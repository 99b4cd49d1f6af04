// Analytic Jacobian of the per-zone residual (ideal MHD and EMHD) and the
// fast FP64 residual of the mixed-precision Newton path.
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
// EMHD adds to T^0_nu the terms E1_nu = q/|b| (u^0 b_nu + b^0 u_nu) and
// E2_nu = -dP (b^0 b_nu / b^2 - (d_0nu + u^0 u_nu)/3), with q = qTilde K_q
// and dP = dPTilde K_d depending on (rho, u) only:
//   K_q = T sqrt(rho alpha_c c_s^2),  K_d = sqrt(T rho alpha_v c_s^2),
//   c_s^2 = Gamma(Gamma-1) u/(rho + Gamma u),  T = (Gamma-1) u/rho,
//   dln K_q/drho = -1/rho + (1/rho - 1/(rho+Gamma u))/2,
//   dln K_q/du   =  1/u + rho/(2 u (rho+Gamma u)),
//   dln K_d/drho = (-1/rho + 1/rho - 1/(rho+Gamma u))/2,
//   dln K_d/du   = (1/u + rho/(u (rho+Gamma u)))/2
// (the 1/rho and 1/u terms of T drop out where T sits on its floor).  The
// guess element's own closure (tau, chi, nu) never enters the residual:
// only the coefficient element's does, and it is fixed during the solve.
//
// The primal quantities are recomputed in the tangent's precision T (float
// for the mixed path, double for the analytic FP64 path).
#pragma once
#include <hip/hip_runtime.h>

#include "zone_physics.hpp"

namespace grim {

// Per-zone constant weights of the EMHD time-derivative sources.  With the
// coefficient element c (elemOld on the half step, elemHalfStep on the full
// step) fixed during the solve, computeTimeDerivSources (physics.cpp:219-318)
// is affine in the new element's u_mu and T:
//   dS_DP/du_mu = [-(g/tau_c) K_DP (-rho_c nu_c g^0mu + 3 rho_c nu_c b^0_c b^mu_c / b^2_c)
//                  - 1/2 g dPTilde_c g^0mu] / dt
//   dS_Q /du_nu = [-(g/tau_c) K_Q (-rho_c chi_c T_c b^nu_c u^0_c / |b|_c)
//                  - 1/2 g qTilde_c g^0nu] / dt
//   dS_Q /dT    = [-(g/tau_c) K_Q (-rho_c chi_c b^0_c / |b|_c)] / dt
// with K_DP = sqrt(tau/rho/nu/T)_c and K_Q = sqrt(tau/rho/chi)_c / T_c (high-
// order terms; 1 otherwise).  Rows Q and DP also carry the normalisation
// (residual.cpp:61-93) and the mask.
struct EmhdWeights {
  double wQ[4], wQT, wDP[4];
  double sQ, sDP;      // mask * normQ, mask * normDP
  double gOverTau;     // g / tau_c
};

template <class L>
__device__ __forceinline__ void emhdWeights(const ZoneGeom& G, const ZoneConst<L>& c,
                                            const DevParams& prm, double dt, EmhdWeights& w) {
  const double invDt = 1. / dt;
  w.gOverTau = G.g / c.cTau;
  w.sQ = c.mask * c.normQ;
  w.sDP = c.mask * c.normDP;
  if constexpr (L::VISC_ON) {
    const double K = prm.highOrderVisc ? sqrt(c.cTau / c.cRho / c.cNu / c.cT) : 1.;
#pragma unroll
    for (int mu = 0; mu < 4; mu++) {
      double a = -w.gOverTau * K *
                 (-c.cRho * c.cNu * G.gcon0[mu] + 3. * c.cRho * c.cNu * c.cBCon[0] * c.cBCon[mu] / c.cBSqr);
      if (prm.highOrderVisc) a -= 0.5 * G.g * c.cDPTilde * G.gcon0[mu];
      w.wDP[mu] = a * invDt;
    }
  }
  if constexpr (L::COND_ON) {
    const double K = prm.highOrderCond ? sqrt(c.cTau / c.cRho / c.cChi) / c.cT : 1.;
#pragma unroll
    for (int nu = 0; nu < 4; nu++) {
      double a = -w.gOverTau * K * (-c.cRho * c.cChi * c.cT * c.cBCon[nu] * c.cUCon0 / c.cBNorm);
      if (prm.highOrderCond) a -= 0.5 * G.g * c.cQTilde * G.gcon0[nu];
      w.wQ[nu] = a * invDt;
    }
    w.wQT = -w.gOverTau * K * (-c.cRho * c.cChi * c.cBCon[0] / c.cBNorm) * invDt;
  }
}

template <class T, class L>
__device__ __forceinline__ void analyticJacobian(const double* Pd, const ZoneGeom& Gd,
                                                 const DevParams& prm, double scale,
                                                 const ZoneConst<L>& c, const EmhdWeights& ew,
                                                 double dt, T (&A)[L::NF][L::NF]) {
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

  // EMHD primal: q = qTilde K_q(rho,u), dP = dPTilde K_d(rho,u) (set(),
  // physics.cpp:118-148) and the logarithmic derivatives of K_q, K_d
  T qt = T(0), dpt = T(0), Kq = T(1), Kd = T(1), q = T(0), dP = T(0);
  T dlnKq_rho = T(0), dlnKq_u = T(0), dlnKd_rho = T(0), dlnKd_u = T(0);
  T ib = T(0), temp = T(0), sT = T(0);
  if constexpr (L::EMHD) {
    const T rawT = p / rho;
    sT = (rawT > (T)prm.tFloor) ? T(1) : T(0);
    temp = (rawT > (T)prm.tFloor) ? rawT : (T)prm.tFloor;
    const T cs2 = gam * p / (rho + gam * u);
    const T invRho = T(1) / rho;
    const T invRG = T(1) / (rho + gam * u);
    const T invU = T(1) / u;
    ib = T(1) / sqrt(bsq);
    if constexpr (L::COND_ON) {
      const T aC = (T)((prm.closure == GRIM_CLOSURE_PER_ZONE) ? c.alphaC : prm.condAlpha);
      qt = (T)Pd[L::Q];
      if (prm.highOrderCond) {
        Kq = temp * sqrt(rho * aC * cs2);
        dlnKq_rho = -sT * invRho + T(0.5) * (invRho - invRG);
        dlnKq_u = sT * invU + T(0.5) * rho * invU * invRG;
      }
      q = qt * Kq;
    }
    if constexpr (L::VISC_ON) {
      const T aV = (T)((prm.closure == GRIM_CLOSURE_PER_ZONE) ? c.alphaV : prm.viscAlpha);
      dpt = (T)Pd[L::DP];
      if (prm.highOrderVisc) {
        Kd = sqrt(temp * rho * aV * cs2);
        dlnKd_rho = T(0.5) * (-sT * invRho + invRho - invRG);
        dlnKd_u = T(0.5) * (sT * invU + rho * invU * invRG);
      }
      dP = dpt * Kd;
    }
  }
  // shapes of the EMHD stress terms: E1_nu = q ib (u^0 b_nu + b^0 u_nu),
  // E2_nu = -dP (b^0 b_nu / b^2 - (d_0nu + u^0 u_nu)/3)
  T f1[4], f2[4];
#pragma unroll
  for (int nu = 0; nu < 4; nu++) {
    f1[nu] = ib * (u0 * bCov[nu] + b0 * uCov[nu]);
    f2[nu] = -(b0 * bCov[nu] / bsq - T(1. / 3.) * ((nu == 0 ? T(1) : T(0)) + u0 * uCov[nu]));
  }

  // rho and u columns
  {
    const T a0 = sRho * u0;
    T dT[4];
#pragma unroll
    for (int nu = 0; nu < 4; nu++) {
      dT[nu] = a0 * uCov[nu];
      if constexpr (L::EMHD) dT[nu] += sRho * (q * dlnKq_rho * f1[nu] + dP * dlnKd_rho * f2[nu]);
    }
    A[RHO][RHO] = gs * a0;
    A[UU][RHO] = gs * (dT[0] + a0);
#pragma unroll
    for (int i = 0; i < 3; i++) A[U1 + i][RHO] = gs * dT[i + 1];
    const T a1 = gam * sU * u0;
#pragma unroll
    for (int nu = 0; nu < 4; nu++) {
      dT[nu] = a1 * uCov[nu];
      if constexpr (L::EMHD) dT[nu] += sU * (q * dlnKq_u * f1[nu] + dP * dlnKd_u * f2[nu]);
    }
    dT[0] = dT[0] + (gam - T(1)) * sU;
    A[RHO][UU] = T(0);
    A[UU][UU] = gs * dT[0];
#pragma unroll
    for (int i = 0; i < 3; i++) A[U1 + i][UU] = gs * dT[i + 1];
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
    if constexpr (L::EMHD) {
      // d(E1 + E2) with du_nu = g_nu,i - alpha gamma_i d_0nu and
      // db_nu = (db^0 u_nu + b^0 du_nu - b_nu du^0) / u^0
      const T dib = T(-0.5) * dbsq * ib * ib * ib;
#pragma unroll
      for (int nu = 0; nu < 4; nu++) {
        const T duCov = gc[nu][i + 1] - (nu == 0 ? alpha * gi : T(0));
        const T dbCov = (db0 * uCov[nu] + b0 * duCov - bCov[nu] * du0) * invU0;
        const T dE1 = q * (dib * (u0 * bCov[nu] + b0 * uCov[nu]) +
                           ib * (du0 * bCov[nu] + u0 * dbCov + db0 * uCov[nu] + b0 * duCov));
        const T dE2 = -dP * ((db0 * bCov[nu] + b0 * dbCov) / bsq - b0 * bCov[nu] * dbsq / (bsq * bsq) -
                             T(1. / 3.) * (du0 * uCov[nu] + u0 * duCov));
        dT[nu] += dE1 + dE2;
      }
    }
    const T dUrho = gRho * du0;
    A[RHO][U1 + i] = s * dUrho;
    A[UU][U1 + i] = s * (g * dT[0] + dUrho);
#pragma unroll
    for (int k = 0; k < 3; k++) A[U1 + k][U1 + i] = gs * dT[k + 1];
    if constexpr (L::COND_ON) {
      const T dS = (T)ew.wQ[0] * (gc[0][i + 1] - alpha * gi) + (T)ew.wQ[1] * gc[1][i + 1] +
                   (T)ew.wQ[2] * gc[2][i + 1] + (T)ew.wQ[3] * gc[3][i + 1];
      A[L::Q][U1 + i] = (T)ew.sQ * (g * qt * du0 / (T)dt + dS);
    }
    if constexpr (L::VISC_ON) {
      const T dS = (T)ew.wDP[0] * (gc[0][i + 1] - alpha * gi) + (T)ew.wDP[1] * gc[1][i + 1] +
                   (T)ew.wDP[2] * gc[2][i + 1] + (T)ew.wDP[3] * gc[3][i + 1];
      A[L::DP][U1 + i] = (T)ew.sDP * (g * dpt * du0 / (T)dt + dS);
    }
  }
  if constexpr (L::EMHD) {
    // EMHD columns of the ideal rows: dT^0_nu/dqTilde = K_q f1_nu,
    // dT^0_nu/ddPTilde = K_d f2_nu
    if constexpr (L::COND_ON) {
      A[RHO][L::Q] = T(0);
      A[UU][L::Q] = gs * Kq * f1[0];
#pragma unroll
      for (int k = 0; k < 3; k++) A[U1 + k][L::Q] = gs * Kq * f1[k + 1];
    }
    if constexpr (L::VISC_ON) {
      A[RHO][L::DP] = T(0);
      A[UU][L::DP] = gs * Kd * f2[0];
#pragma unroll
      for (int k = 0; k < 3; k++) A[U1 + k][L::DP] = gs * Kd * f2[k + 1];
    }
    // EMHD rows: U_x = g u^0 x, S_impl = g x / tau_c, S_td affine in (u_mu, T)
    const T invDt = T(1) / (T)dt;
    const T diag = g * u0 * invDt + T(0.5) * (T)ew.gOverTau;
    if constexpr (L::COND_ON) {
      const T sQ = (T)ew.sQ;
      const T dTdrho = -sT * temp / rho * sRho;
      const T dTdu = sT * temp / u * sU;
      A[L::Q][RHO] = sQ * (T)ew.wQT * dTdrho;
      A[L::Q][UU] = sQ * (T)ew.wQT * dTdu;
      A[L::Q][L::Q] = sQ * diag;
      if constexpr (L::VISC_ON) A[L::Q][L::DP] = T(0);
    }
    if constexpr (L::VISC_ON) {
      const T sDP = (T)ew.sDP;
      A[L::DP][RHO] = T(0);
      A[L::DP][UU] = T(0);
      A[L::DP][L::DP] = sDP * diag;
      if constexpr (L::COND_ON) A[L::DP][L::Q] = T(0);
    }
  }
}

// Adjoint of the same residual (reverse mode, by hand): g = J^T w for one
// zone without forming J.  With W_nu = (w_U, w_U1, w_U2, w_U3) the weighted
// sum of the residual rows is
//   w.R = gs [ (w_rho + w_U) rho u^0 + W^nu T^0_nu ] + w_Q R_Q + w_DP R_DP,
//   W^nu T^0_nu = wt u^0 (W.u) + (p + b^2/2) W_0 - b^0 (W.b)
//               + q ib (u^0 (W.b) + b^0 (W.u)) - dP (b^0 (W.b)/b^2 - (W_0 + u^0 (W.u))/3)
// (wt = rho + u + p + b^2, W.u = W^nu u_nu, W.b = W^nu b_nu), and
//   R_Q  = sQ  [g u^0 qTilde/dt + g qTilde/(2 tau_c) + wQT T + wQ^mu u_mu] + const,
//   R_DP = sDP [g u^0 dPTilde/dt + g dPTilde/(2 tau_c) + wDP^mu u_mu] + const.
// The sweep below runs the primal of analyticJacobian forward, then carries
// the adjoints (the _b names) of every intermediate back to the primitives
// in reverse order of their computation; floors follow the tangent's
// one-sided rule.  <J v, w> = <v, J^T w> to rounding, for any v and w.
template <class T, class L>
__device__ __forceinline__ void analyticAdjoint(const double* Pd, const ZoneGeom& Gd,
                                                const DevParams& prm, double scale,
                                                const ZoneConst<L>& c, const EmhdWeights& ew,
                                                double dt, const T (&wr)[L::NF], T (&out)[L::NF]) {
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

  // ---- primal (as analyticJacobian)
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
  const T wt = rho + u + p + bsq;
  const T gs = g * (T)scale;

  T qt = T(0), dpt = T(0), Kq = T(1), Kd = T(1), q = T(0), dP = T(0);
  T dlnKq_rho = T(0), dlnKq_u = T(0), dlnKd_rho = T(0), dlnKd_u = T(0);
  T ib = T(0), temp = T(0), sT = T(0);
  if constexpr (L::EMHD) {
    const T rawT = p / rho;
    sT = (rawT > (T)prm.tFloor) ? T(1) : T(0);
    temp = (rawT > (T)prm.tFloor) ? rawT : (T)prm.tFloor;
    const T cs2 = gam * p / (rho + gam * u);
    const T invRho = T(1) / rho;
    const T invRG = T(1) / (rho + gam * u);
    const T invU = T(1) / u;
    ib = T(1) / sqrt(bsq);
    if constexpr (L::COND_ON) {
      const T aC = (T)((prm.closure == GRIM_CLOSURE_PER_ZONE) ? c.alphaC : prm.condAlpha);
      qt = (T)Pd[L::Q];
      if (prm.highOrderCond) {
        Kq = temp * sqrt(rho * aC * cs2);
        dlnKq_rho = -sT * invRho + T(0.5) * (invRho - invRG);
        dlnKq_u = sT * invU + T(0.5) * rho * invU * invRG;
      }
      q = qt * Kq;
    }
    if constexpr (L::VISC_ON) {
      const T aV = (T)((prm.closure == GRIM_CLOSURE_PER_ZONE) ? c.alphaV : prm.viscAlpha);
      dpt = (T)Pd[L::DP];
      if (prm.highOrderVisc) {
        Kd = sqrt(temp * rho * aV * cs2);
        dlnKd_rho = T(0.5) * (-sT * invRho + invRho - invRG);
        dlnKd_u = T(0.5) * (sT * invU + rho * invU * invRG);
      }
      dP = dpt * Kd;
    }
  }
  const T W[4] = {wr[UU], wr[U1], wr[U2], wr[U3]};
  const T Wu = W[0] * uCov[0] + W[1] * uCov[1] + W[2] * uCov[2] + W[3] * uCov[3];
  const T Wb = W[0] * bCov[0] + W[1] * bCov[1] + W[2] * bCov[2] + W[3] * bCov[3];

  // ---- reverse sweep
  T rho_b = T(0), u_b = T(0), u0_b = T(0), Wu_b = T(0), Wb_b = T(0), bsq_b = T(0), b0_b = T(0);
  T qt_b = T(0), dpt_b = T(0);
  T uCov_b[4] = {T(0), T(0), T(0), T(0)};
  // ideal stress rows, weighted by gs
  const T wRhoU = wr[RHO] + wr[UU];
  rho_b += gs * wRhoU * u0;
  u0_b += gs * (wRhoU * rho + wt * Wu);
  const T wt_b = gs * u0 * Wu;
  Wu_b += gs * wt * u0;
  T p_b = gs * W[0];
  bsq_b += gs * T(0.5) * W[0];
  b0_b -= gs * Wb;
  Wb_b -= gs * b0;
  if constexpr (L::EMHD) {
    // E1 = q ib (u0 Wb + b0 Wu)
    const T e1 = u0 * Wb + b0 * Wu;
    const T q_b = gs * ib * e1;
    const T ib_b = gs * q * e1;
    u0_b += gs * q * ib * Wb;
    Wb_b += gs * q * ib * u0;
    b0_b += gs * q * ib * Wu;
    Wu_b += gs * q * ib * b0;
    // E2 = -dP (b0 Wb / bsq - (W0 + u0 Wu)/3)
    const T invBsq = T(1) / bsq;
    const T dP_b = -gs * (b0 * Wb * invBsq - T(1. / 3.) * (W[0] + u0 * Wu));
    b0_b -= gs * dP * Wb * invBsq;
    Wb_b -= gs * dP * b0 * invBsq;
    bsq_b += gs * dP * b0 * Wb * invBsq * invBsq;
    u0_b += gs * dP * Wu * T(1. / 3.);
    Wu_b += gs * dP * u0 * T(1. / 3.);
    // EMHD rows
    const T invDt = T(1) / (T)dt;
    const T diag = g * u0 * invDt + T(0.5) * (T)ew.gOverTau;
    T temp_b = T(0);
    if constexpr (L::COND_ON) {
      const T cQ = wr[L::Q] * (T)ew.sQ;
      qt_b += cQ * diag;
      u0_b += cQ * g * qt * invDt;
      temp_b += cQ * (T)ew.wQT;
#pragma unroll
      for (int mu = 0; mu < 4; mu++) uCov_b[mu] += cQ * (T)ew.wQ[mu];
    }
    if constexpr (L::VISC_ON) {
      const T cD = wr[L::DP] * (T)ew.sDP;
      dpt_b += cD * diag;
      u0_b += cD * g * dpt * invDt;
#pragma unroll
      for (int mu = 0; mu < 4; mu++) uCov_b[mu] += cD * (T)ew.wDP[mu];
    }
    // q = qTilde K_q(rho, u), dP = dPTilde K_d(rho, u), T = max(p/rho, floor)
    if constexpr (L::COND_ON) {
      qt_b += q_b * Kq;
      rho_b += q_b * q * dlnKq_rho;
      u_b += q_b * q * dlnKq_u;
    }
    if constexpr (L::VISC_ON) {
      dpt_b += dP_b * Kd;
      rho_b += dP_b * dP * dlnKd_rho;
      u_b += dP_b * dP * dlnKd_u;
    }
    rho_b -= temp_b * sT * temp / rho;
    u_b += temp_b * sT * temp / u;
    // ib = bsq^(-1/2)
    bsq_b += ib_b * T(-0.5) * ib * ib * ib;
  }
  // wt = rho + u + p + bsq, p = (Gamma - 1) u
  rho_b += wt_b;
  u_b += wt_b;
  p_b += wt_b;
  bsq_b += wt_b;
  u_b += p_b * (gam - T(1));
  // W.u and W.b
  T bCov_b[4];
#pragma unroll
  for (int nu = 0; nu < 4; nu++) {
    uCov_b[nu] += Wu_b * W[nu];
    bCov_b[nu] = Wb_b * W[nu];
  }
  // bsq = (BB + b0^2) / u0^2 + floor
  b0_b += bsq_b * T(2) * b0 * invU0 * invU0;
  u0_b -= bsq_b * T(2) * bb * invU0;
  // b_nu = (B_nu + b0 u_nu) / u0
  T utLow_b[3] = {T(0), T(0), T(0)};
  b0_b += bCov_b[0] * uCov[0] * invU0;
  uCov_b[0] += bCov_b[0] * b0 * invU0;
  u0_b -= bCov_b[0] * bCov[0] * invU0;
#pragma unroll
  for (int i = 0; i < 3; i++) {
    b0_b += bCov_b[i + 1] * utLow[i] * invU0;
    utLow_b[i] += bCov_b[i + 1] * b0 * invU0;
    u0_b -= bCov_b[i + 1] * bCov[i + 1] * invU0;
  }
  // b0 = B^i utLow_i
#pragma unroll
  for (int i = 0; i < 3; i++) utLow_b[i] += b0_b * B[i];
  // u_0 = -alpha lor + g_0j utilde^j,  u_i = utLow_i,  u^0 = lor / alpha
  T lor_b = u0_b * invAlpha - alpha * uCov_b[0];
  T ut_b[3];
#pragma unroll
  for (int j = 0; j < 3; j++) {
    utLow_b[j] += uCov_b[j + 1];
    ut_b[j] = uCov_b[0] * gc[0][j + 1];
  }
  // lor = sqrt(1 + utilde^i gamma_ij utilde^j): dlor/dutilde^j = utLow_j / lor
#pragma unroll
  for (int j = 0; j < 3; j++) ut_b[j] += lor_b * utLow[j] * invLor;
  // utLow_i = gamma_ij utilde^j
#pragma unroll
  for (int j = 0; j < 3; j++)
    ut_b[j] += utLow_b[0] * gc[1][j + 1] + utLow_b[1] * gc[2][j + 1] + utLow_b[2] * gc[3][j + 1];

  out[RHO] = sRho * rho_b;
  out[UU] = sU * u_b;
#pragma unroll
  for (int j = 0; j < 3; j++) out[U1 + j] = ut_b[j];
  if constexpr (L::COND_ON) out[L::Q] = qt_b;
  if constexpr (L::VISC_ON) out[L::DP] = dpt_b;
}

// Ideal-MHD entry kept for the kernels that carry no EMHD constants.
template <class T, class L>
__device__ __forceinline__ void analyticJacobianIdeal(const double* Pd, const ZoneGeom& Gd,
                                                      const DevParams& prm, double scale,
                                                      T (&A)[L::NF][L::NF]) {
  static_assert(!L::EMHD, "ideal-MHD tangent");
  ZoneConst<L> c;
  EmhdWeights ew;
  analyticJacobian<T, L>(Pd, Gd, prm, scale, c, ew, 1., A);
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

// The EMHD residual of the fused paths: the same rows as residual<L>, with
// the per-zone constants folded once per zone.  The time-derivative sources
// are affine in the new element's u_mu and T with the weights of
// emhdWeights (S_td = sum_mu w_mu (u_mu - u_mu,old) + w_T (T - T_old), the
// divisions by dt, tau_c, |b|_c and b^2_c taken once), the implicit source is
// g qTilde / tau_c, and the rows are scaled by 1/dt, the normalisation and
// the mask by products.  Equal to residual<L> up to rounding.
template <class L>
__device__ __forceinline__ void residualFastEmhd(const double* P, const ZoneGeom& G, const ZoneConst<L>& c,
                                                 const DevParams& prm, const EmhdWeights& ew, double invDt,
                                                 double* R) {
  static_assert(L::EMHD, "EMHD residual");
  Kinematics k;
  kinematics<L>(P, G, prm, k);
  Closure cl;
  closure<L>(P, k, prm, c.alphaC, c.alphaV, c.radius, cl);
  const double w = k.rho + k.u + k.pressure + k.bSqr;
  const double qOverB = L::COND_ON ? cl.q / cl.bNorm : 0.;
  const double dPOverB2 = L::VISC_ON ? cl.deltaP / k.bSqr : 0.;
  double T0[4];
#pragma unroll
  for (int nu = 0; nu < 4; nu++) {
    double t = w * k.uCon[0] * k.uCov[nu] - k.bCon[0] * k.bCov[nu];
    if (nu == 0) t += k.pressure + 0.5 * k.bSqr;
    if constexpr (L::COND_ON) t += qOverB * (k.uCon[0] * k.bCov[nu] + k.bCon[0] * k.uCov[nu]);
    if constexpr (L::VISC_ON)
      t -= dPOverB2 * k.bCon[0] * k.bCov[nu] - (cl.deltaP / 3.) * ((nu == 0 ? 1. : 0.) + k.uCon[0] * k.uCov[nu]);
    T0[nu] = t;
  }
  const double s = c.mask * invDt;
  const double Urho = G.g * (k.rho * k.uCon[0]);
  R[RHO] = (Urho - c.rhs[RHO]) * s;
  R[UU] = (G.g * T0[0] + Urho - c.rhs[UU]) * s;
#pragma unroll
  for (int i = 1; i < 4; i++) R[UU + i] = (G.g * T0[i] - c.rhs[UU + i]) * s;
  double du[4];
#pragma unroll
  for (int mu = 0; mu < 4; mu++) du[mu] = k.uCov[mu] - c.oldUCov[mu];
  if constexpr (L::COND_ON) {
    double std = ew.wQT * (k.temperature - c.oldT);
#pragma unroll
    for (int mu = 0; mu < 4; mu++) std += ew.wQ[mu] * du[mu];
    const double U = G.g * (k.uCon[0] * P[L::Q]);
    R[L::Q] = ((U - c.rhs[L::Q]) * invDt + 0.5 * ew.gOverTau * P[L::Q] + std) * ew.sQ;
  }
  if constexpr (L::VISC_ON) {
    double std = 0.;
#pragma unroll
    for (int mu = 0; mu < 4; mu++) std += ew.wDP[mu] * du[mu];
    const double U = G.g * (k.uCon[0] * P[L::DP]);
    R[L::DP] = ((U - c.rhs[L::DP]) * invDt + 0.5 * ew.gOverTau * P[L::DP] + std) * ew.sDP;
  }
}

// Row and column equilibration around the FP32 LU: solve (Dr A Dc) y = Dr b,
// x = Dc y, with Dc = |P_k| (the natural scale of each primitive) and Dr the
// inverse of each scaled row's largest entry.  The equilibrated matrix is
// factorised without pivoting: the row swaps of partial pivoting are selects
// on every entry (a fifth of the issued instructions of the ideal-MHD solve),
// and on the 2^22-zone Kerr-Schild batch, the EMHD batch and the torus they
// changed neither the iteration counts, nor the unconverged zones, nor the
// error against the known roots.  An FP32 step that goes wrong is rejected
// by the FP64 line search, and the escalating solve finishes such zones with
// the pivoted FP64 factorisation.
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
    const T r = (m > T(0)) ? lowRecip(m) : T(1);
#pragma unroll
    for (int k = 0; k < N; k++) A[i][k] = A[i][k] * r;
    b[i] = b[i] * r;
  }
  luSolve<T, N, false>(A, b);
#pragma unroll
  for (int k = 0; k < N; k++) b[k] = b[k] * cs[k];
}

}  // namespace grim
