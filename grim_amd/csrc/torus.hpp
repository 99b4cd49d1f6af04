// Device code of the torus problem (src/problem/torus/problem.cpp): the
// Fishbone-Moncrief equilibrium at one zone, applyFloor at one zone.
// Expressions follow the reference term by term (see the cited lines).
#pragma once
#include <hip/hip_runtime.h>

#include "grim_hip.h"
#include "step_physics.hpp"
#include "zone_physics.hpp"

namespace grim {

// problem.cpp:56-143 (M = 1)
__host__ __device__ inline double lFishboneMoncrief(double a, double r) {
  const double M = 1.;
  return sqrt(M / pow(r, 3.)) *
         (pow(r, 4.) + r * r * a * a - 2. * M * r * a * a - a * sqrt(M * r) * (r * r - a * a)) /
         (r * r - 3 * M * r + 2. * a * sqrt(M * r));
}
__host__ __device__ inline double computeDelta(double a, double r) { return r * r - 2 * 1. * r + a * a; }
__host__ __device__ inline double computeSigma(double a, double r, double theta) {
  return r * r + a * a * cos(theta) * cos(theta);
}
__host__ __device__ inline double computeA(double a, double r, double theta) {
  const double Delta = computeDelta(a, r);
  return pow(r * r + a * a, 2.) - Delta * a * a * sin(theta) * sin(theta);
}
__host__ __device__ inline double computeLnOfh(const grim_torus_params& t, double a, double r, double theta) {
  const double l = lFishboneMoncrief(a, t.PressureMaxRadius);
  auto term1 = [&](double rr, double th) {
    const double Delta = computeDelta(a, rr), Sigma = computeSigma(a, rr, th), A = computeA(a, rr, th);
    return 0.5 * log((1. + sqrt(1. + (4. * l * l * Sigma * Sigma * Delta) / (A * sin(th) * A * sin(th)))) /
                     (Sigma * Delta / A));
  };
  auto term2 = [&](double rr, double th) {
    const double Delta = computeDelta(a, rr), Sigma = computeSigma(a, rr, th), A = computeA(a, rr, th);
    return -0.5 * sqrt(1. + (4. * l * l * Sigma * Sigma * Delta) / (A * A * sin(th) * sin(th)));
  };
  auto term3 = [&](double rr, double th) {
    const double A = computeA(a, rr, th);
    const double M = 1.;
    return -2 * a * M * rr * l / A;
  };
  const double t1 = term1(r, theta), t2 = term2(r, theta), t3 = term3(r, theta);
  const double t1i = term1(t.InnerEdgeRadius, M_PI / 2.), t2i = term2(t.InnerEdgeRadius, M_PI / 2.),
               t3i = term3(t.InnerEdgeRadius, M_PI / 2.);
  return t1 + t2 + t3 - t1i - t2i - t3i;
}

// Fluid state of one zone of the torus initial conditions (problem.cpp:175-322).
__host__ __device__ inline void torusZone(const grim_torus_params& t, double aBH, double hSlope, double gam,
                                 double rhoFloor, double uFloor, double r, double theta, double X2,
                                 double lapse, const double* beta, double xi, double* P) {
  double lnOfh = 1.;
  if (r >= t.InnerEdgeRadius) lnOfh = computeLnOfh(t, aBH, r, theta);
  if (lnOfh < 0. || r < t.InnerEdgeRadius) {
    P[RHO] = rhoFloor;
    P[UU] = uFloor;
    P[U1] = 0.;
    P[U2] = 0.;
    P[U3] = 0.;
    return;
  }
  const double h = exp(lnOfh);
  const double Gamma = gam, Kappa = t.Adiabat;
  P[RHO] = pow((h - 1) * (Gamma - 1.) / (Kappa * Gamma), 1. / (Gamma - 1.));
  P[UU] = Kappa * pow(P[RHO], Gamma) / (Gamma - 1.) * (1. + t.InitialPerturbationAmplitude * (xi - 0.5));
  const double A = computeA(aBH, r, theta), Sigma = computeSigma(aBH, r, theta), Delta = computeDelta(aBH, r);
  const double l = lFishboneMoncrief(aBH, t.PressureMaxRadius);
  const double expOfMinus2Chi = Sigma * Sigma * Delta / (A * A * sin(theta) * sin(theta));
  const double uCovPhiBL = sqrt((-1. + sqrt(1. + 4 * l * l * expOfMinus2Chi)) / 2.);
  const double uConPhiBL = 2. * aBH * r * sqrt(1. + uCovPhiBL * uCovPhiBL) / sqrt(A * Sigma * Delta) +
                           sqrt(Sigma / A) * uCovPhiBL / sin(theta);
  double uConBL[4] = {0., 0., 0., uConPhiBL};
  double gCovBL[4][4], T[4][4];
  for (int a = 0; a < 4; a++)
    for (int b = 0; b < 4; b++) gCovBL[a][b] = T[a][b] = 0.;
  const double mu = 1 + aBH * aBH * cos(theta) * cos(theta) / (r * r);
  gCovBL[0][0] = -(1. - 2. / (r * mu));
  gCovBL[0][3] = -2. * aBH * sin(theta) * sin(theta) / (r * mu);
  gCovBL[3][0] = gCovBL[0][3];
  gCovBL[1][1] = mu * r * r / Delta;
  gCovBL[2][2] = r * r * mu;
  gCovBL[3][3] = r * r * sin(theta) * sin(theta) *
                 (1. + aBH * aBH / (r * r) + 2. * aBH * aBH * sin(theta) * sin(theta) / (r * r * r * mu));
  T[0][0] = T[1][1] = T[2][2] = T[3][3] = 1.;
  T[0][1] = 2. * r / Delta;
  T[3][1] = aBH / Delta;
  const double AA = gCovBL[0][0];
  const double BB = 2. * (gCovBL[0][1] * uConBL[1] + gCovBL[0][2] * uConBL[2] + gCovBL[0][3] * uConBL[3]);
  const double CC = 1. + gCovBL[1][1] * uConBL[1] * uConBL[1] + gCovBL[2][2] * uConBL[2] * uConBL[2] +
                    gCovBL[3][3] * uConBL[3] * uConBL[3] +
                    2. * (gCovBL[1][2] * uConBL[1] * uConBL[2] + gCovBL[1][3] * uConBL[1] * uConBL[3] +
                          gCovBL[2][3] * uConBL[2] * uConBL[3]);
  const double discriminent = BB * BB - 4. * AA * CC;
  uConBL[0] = -(BB + sqrt(discriminent)) / (2. * AA);
  double uConKS[4];
  for (int a = 0; a < 4; a++) {
    uConKS[a] = 0.;
    for (int b = 0; b < 4; b++) uConKS[a] += T[a][b] * uConBL[b];
  }
  const double rFactor = r;
  const double hFactor = M_PI + (1. - hSlope) * M_PI * cos(2. * M_PI * X2);
  const double uConMKS[4] = {uConKS[0], uConKS[1] / rFactor, uConKS[2] / hFactor, uConKS[3]};
  P[U1] = uConMKS[1] + pow(lapse, 2.) * beta[1] * uConMKS[0];
  P[U2] = uConMKS[2] + pow(lapse, 2.) * beta[2] * uConMKS[0];
  P[U3] = uConMKS[3] + pow(lapse, 2.) * beta[3] * uConMKS[0];
}

// applyFloor (problem.cpp:542-712) for one zone, given its floor values.
// The maximum-Lorentz-factor rescaling is applied to every zone: grim
// applies it when any zone exceeds the limit, and the factor of every other
// zone is exactly 1.
template <class L>
__host__ __device__ inline void applyFloorZone(const DevParams& prm, const grim_torus_params& t, const GeomFull& G,
                                      double minRho, double minU, double* P) {
  const ZoneGeom& zg = G.zg;
  const double rho_prefloor = P[RHO] * 1.;
  const double u_prefloor = P[UU] * 1.;
  double condition = P[RHO] < minRho ? 1. : 0.;
  double usefloor = condition;
  P[RHO] = condition * minRho + (1. - condition) * P[RHO];
  condition = P[UU] < minU ? 1. : 0.;
  usefloor = stdmax(condition, usefloor);
  P[UU] = condition * minU + (1. - condition) * P[UU];
  Kinematics k;
  Closure cc;
  kinematics<L>(P, zg, prm, k);
  double bSqr = k.bSqr;
  condition = bSqr > t.BsqrOverRhoMax * P[RHO] ? 1. : 0.;
  usefloor = stdmax(condition, usefloor);
  P[RHO] = P[RHO] * (1. - condition) + condition * bSqr / t.BsqrOverRhoMax;
  condition = bSqr > t.BsqrOverUMax * P[UU] ? 1. : 0.;
  usefloor = stdmax(condition, usefloor);
  P[UU] = P[UU] * (1. - condition) + condition * bSqr / t.BsqrOverUMax;
  const double zero = 0.;
  const double trans =
      stdmax(stdmin(zero + 1., (bSqr - 0.1 * P[RHO]) / P[RHO]), zero) * (usefloor > zero ? 1. : 0.);
  const double betapar = -k.bCon[0] / bSqr / k.uCon[0];
  double betasqr = betapar * betapar * bSqr;
  const double betasqrmax = 1. - 1. / t.MaxLorentzFactor / t.MaxLorentzFactor;
  betasqr = stdmin(betasqr, betasqrmax);
  const double gamma = 1. / sqrt(1. - betasqr);
  double ucondr[4];
  for (int m = 0; m < 4; m++) ucondr[m] = gamma * (k.uCon[m] + betapar * k.bCon[m]);
  const double Bcon[4] = {zero, P[L::B1], P[L::B2], P[L::B3]};
  double Bcov[4];
  for (int m = 0; m < 4; m++) {
    Bcov[m] = zero;
    for (int n = 0; n < 4; n++) Bcov[m] += zg.cov(n, m) * Bcon[n];
  }
  double udotB = zero, bSqr_fl = zero;
  for (int n = 0; n < 4; n++) {
    udotB += Bcov[n] * k.uCon[n];
    bSqr_fl += Bcov[n] * Bcon[n];
  }
  double bNorm = sqrt(bSqr_fl);
  const double bMin = sqrt(prm.bSqrFloor);
  bNorm = stdmax(bNorm, bMin);
  const double wold = rho_prefloor + u_prefloor * prm.gam;
  const double QdotB = udotB * wold * k.uCon[0];
  const double wnew = P[RHO] + P[UU] * prm.gam;
  const double x = 2. * QdotB / (bNorm * wnew * ucondr[0]);
  const double vpar = x / (ucondr[0] * (1. + sqrt(1. + x * x)));
  const double one_over_ucondr_t = 1. / ucondr[0];
  double vcon[4];
  vcon[0] = zero + 1.;
  for (int m = 1; m < 4; m++) vcon[m] = vpar * Bcon[m] / bNorm + ucondr[m] * one_over_ucondr_t;
  double vsqr = zero;
  for (int m = 0; m < 4; m++)
    for (int n = 0; n < 4; n++) vsqr += zg.cov(m, n) * vcon[m] * vcon[n];
  const double gcon00 = zg.gcon0[0];
  const double condition_v = (vsqr >= 0. || vsqr < 1. / gcon00) ? 1. : 0.;
  vsqr = (1. - condition_v) * vsqr + condition_v / gcon00;
  const double ut = sqrt(-1. / vsqr);
  double utcon[4];
  utcon[0] = zero;
  for (int m = 1; m < 4; m++) utcon[m] = ut * (vcon[m] - zg.gcon0[m] / gcon00);
  P[U1] = P[U1] * (1. - trans) + trans * utcon[1];
  P[U2] = P[U2] * (1. - trans) + trans * utcon[2];
  P[U3] = P[U3] * (1. - trans) + trans * utcon[3];
  kinematics<L>(P, zg, prm, k);
  const double lorentzFactorSqr = k.gammaLF * k.gammaLF;
  double MultFac = sqrt(stdmax((lorentzFactorSqr - 1.) / (t.MaxLorentzFactor * t.MaxLorentzFactor - 1.), 1.));
  MultFac = 1. / MultFac;
  P[U1] *= MultFac;
  P[U2] *= MultFac;
  P[U3] *= MultFac;
  if constexpr (L::EMHD) {
    kinematics<L>(P, zg, prm, k);
    closure<L>(P, k, prm, 1., 1., G.r, cc);
    if constexpr (L::COND_ON) {
      const double Qmax = 1.07 * prm.condClosureFactor * P[RHO] * pow(cc.soundSpeed, 3.);
      const double LimFac = stdmax(fabs(cc.q) / Qmax, 1.);
      P[L::Q] = P[L::Q] / LimFac;
    }
    if constexpr (L::VISC_ON) {
      const double pressure = k.pressure, deltaP = cc.deltaP;
      bSqr = k.bSqr;
      const double dPmod = stdmax(pressure - 2. / 3. * deltaP, 0.01 * prm.bSqrFloor) /
                           stdmax(pressure + 1. / 3. * deltaP, prm.bSqrFloor);
      const double dPmaxPlus = stdmin(1.07 * prm.viscClosureFactor * bSqr * 0.5 * dPmod, 1.49 * pressure);
      const double dPmaxMinus = stdmax(-1.07 * prm.viscClosureFactor * bSqr, -2.99 * pressure);
      const double cond = deltaP > 0. ? 1. : 0.;
      P[L::DP] = P[L::DP] * (cond / stdmax(deltaP / dPmaxPlus, 1.) + (1. - cond) / stdmax(deltaP / dPmaxMinus, 1.));
    }
  }
}

}  // namespace grim
