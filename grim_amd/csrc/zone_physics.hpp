// Per-zone physics of grim's implicit solve, as gfx950 device code.
//
// One thread owns one zone.  Everything the residual of a zone needs lives in
// registers: the 16 geometry doubles, the guess, the folded right-hand side
// and (EMHD) the handful of coefficient-element scalars that grim reads from
// elemOld / elemHalfStep.  Only the t-row of the stress tensor T^0_nu is
// formed, because computeFluxes(0) is the only consumer inside the solve
// (grim's fluidElement::set materialises all 16 components).
//
// Evaluation order follows grim's expressions term by term (fluidElement::set,
// src/physics/physics.cpp:53-183, and the residual, src/timestepper/
// residual.cpp:3-180), so that with -ffp-contract=off the FP64 results agree
// with a CPU FP64 evaluation to the last bit wherever division and square
// root are correctly rounded.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "grim_hip.h"

namespace grim {

enum { RHO = 0, UU = 1, U1 = 2, U2 = 3, U3 = 4 };

// gCov plane index of the symmetric (mu, nu) component: upper-triangle
// packing 00 01 02 03 11 12 13 22 23 33.  Closed form (no recursion or
// branches) so that it folds to a constant in every unrolled loop; a
// runtime index would demote the register array to scratch memory.
__host__ __device__ constexpr int gcovIdx(int mu, int nu) {
  return (mu < nu ? mu : nu) * 4 - ((mu < nu ? mu : nu) * ((mu < nu ? mu : nu) + 1)) / 2 +
         (mu < nu ? nu : mu);
}
static_assert(gcovIdx(0, 0) == 0 && gcovIdx(3, 0) == 3 && gcovIdx(1, 1) == 4 &&
                  gcovIdx(3, 1) == 6 && gcovIdx(2, 2) == 7 && gcovIdx(2, 3) == 8 &&
                  gcovIdx(3, 3) == 9,
              "gCov packing");
enum { GEOM_GCON0 = 10, GEOM_ALPHA = 14, GEOM_G = 15 };

// Scalars of grim_params the device code reads, plus the closure selector.
struct DevParams {
  double gam;  // adiabaticIndex
  double rhoFloor, uFloor, bSqrFloor, tFloor;
  double condAlpha, viscAlpha;
  double tau0;
  double condClosureFactor, viscClosureFactor;
  double atol, jacEps, lsFloor;
  int closure;
  int highOrderCond, highOrderVisc;
  int maxIter, maxLS;
};

inline DevParams makeDevParams(const grim_params& p) {
  DevParams d;
  d.gam = p.adiabaticIndex;
  d.rhoFloor = p.rhoFloorInFluidElement;
  d.uFloor = p.uFloorInFluidElement;
  d.bSqrFloor = p.bSqrFloorInFluidElement;
  d.tFloor = p.temperatureFloorInFluidElement;
  d.condAlpha = p.ConductionAlpha;
  d.viscAlpha = p.ViscosityAlpha;
  d.tau0 = p.tau0;
  d.condClosureFactor = p.ConductionClosureFactor;
  d.viscClosureFactor = p.ViscosityClosureFactor;
  d.atol = p.nonlinearsolve_atol;
  d.jacEps = p.JacobianAssembleEpsilon;
  d.lsFloor = p.linesearchfloor;
  d.closure = p.closure;
  d.highOrderCond = p.highOrderTermsConduction;
  d.highOrderVisc = p.highOrderTermsViscosity;
  d.maxIter = p.maxNonLinearIter;
  d.maxLS = p.maxLineSearchIters;
  return d;
}

// af::max / af::min against a scalar: plain selects (no NaN special-casing,
// exactly `a > b ? a : b`).
__device__ __forceinline__ double smax(double a, double b) { return a > b ? a : b; }
__device__ __forceinline__ double smin(double a, double b) { return a < b ? a : b; }

struct ZoneGeom {
  double gcov[10];
  double gcon0[4];
  double alpha, g;
  __device__ __forceinline__ double cov(int mu, int nu) const { return gcov[gcovIdx(mu, nu)]; }
};

__device__ __forceinline__ void loadGeom(const double* __restrict__ geom, int64_t n,
                                         int64_t z, ZoneGeom& G) {
#pragma unroll
  for (int k = 0; k < 10; k++) G.gcov[k] = geom[k * n + z];
#pragma unroll
  for (int k = 0; k < 4; k++) G.gcon0[k] = geom[(GEOM_GCON0 + k) * n + z];
  G.alpha = geom[GEOM_ALPHA * n + z];
  G.g = geom[GEOM_G * n + z];
}

// Compile-time variable layout (src/params.hpp:8-18 with the index rule of
// src/problem/linear_modes/params.cpp:84-92).
template <bool COND, bool VISC>
struct Layout {
  static constexpr int NF = 5 + COND + VISC;
  static constexpr int DOF = 8 + COND + VISC;
  static constexpr int Q = 5;
  static constexpr int DP = 5 + COND;
  static constexpr int B1 = NF, B2 = NF + 1, B3 = NF + 2;
  static constexpr bool COND_ON = COND, VISC_ON = VISC;
  static constexpr bool EMHD = COND || VISC;
};

// Kinematic part of fluidElement::set shared by every consumer: 4-velocity,
// magnetic 4-vector and b^2 (physics.cpp:60-116).
struct Kinematics {
  double rho, u, pressure, temperature;
  double gammaLF;  // Lorentz factor (fluidElement::gammaLorentzFactor)
  double uCon[4], uCov[4], bCon[4], bCov[4];
  double bSqr;
};

template <class L>
__device__ __forceinline__ void kinematics(const double* P, const ZoneGeom& G,
                                           const DevParams& prm, Kinematics& k) {
  k.rho = smax(P[RHO], prm.rhoFloor);
  k.u = smax(P[UU], prm.uFloor);
  const double u1 = P[U1], u2 = P[U2], u3 = P[U3];
  const double B1 = P[L::B1], B2 = P[L::B2], B3 = P[L::B3];
  k.pressure = (prm.gam - 1.) * k.u;
  k.temperature = smax(k.pressure / k.rho, prm.tFloor);
  const double gamma = sqrt(1. + G.cov(1, 1) * u1 * u1 + G.cov(2, 2) * u2 * u2 +
                            G.cov(3, 3) * u3 * u3 +
                            2. * (G.cov(1, 2) * u1 * u2 + G.cov(1, 3) * u1 * u3 +
                                  G.cov(2, 3) * u2 * u3));
  k.gammaLF = gamma;
  k.uCon[0] = gamma / G.alpha;
  k.uCon[1] = u1 - gamma * G.gcon0[1] * G.alpha;
  k.uCon[2] = u2 - gamma * G.gcon0[2] * G.alpha;
  k.uCon[3] = u3 - gamma * G.gcon0[3] * G.alpha;
#pragma unroll
  for (int mu = 0; mu < 4; mu++)
    k.uCov[mu] = G.cov(mu, 0) * k.uCon[0] + G.cov(mu, 1) * k.uCon[1] +
                 G.cov(mu, 2) * k.uCon[2] + G.cov(mu, 3) * k.uCon[3];
  k.bCon[0] = B1 * k.uCov[1] + B2 * k.uCov[2] + B3 * k.uCov[3];
  k.bCon[1] = (B1 + k.bCon[0] * k.uCon[1]) / k.uCon[0];
  k.bCon[2] = (B2 + k.bCon[0] * k.uCon[2]) / k.uCon[0];
  k.bCon[3] = (B3 + k.bCon[0] * k.uCon[3]) / k.uCon[0];
#pragma unroll
  for (int mu = 0; mu < 4; mu++)
    k.bCov[mu] = G.cov(mu, 0) * k.bCon[0] + G.cov(mu, 1) * k.bCon[1] +
                 G.cov(mu, 2) * k.bCon[2] + G.cov(mu, 3) * k.bCon[3];
  k.bSqr = k.bCon[0] * k.bCov[0] + k.bCon[1] * k.bCov[1] + k.bCon[2] * k.bCov[2] +
           k.bCon[3] * k.bCov[3] + prm.bSqrFloor;
}

// EMHD closure of one element (setFluidElementParameters; the torus form is
// src/problem/torus/problem.cpp:7-55) together with the rescaled heat flux q
// and pressure anisotropy deltaP of set() (physics.cpp:118-151).
struct Closure {
  double soundSpeed;
  double bNorm;
  double q, deltaP;
  double tau, chi, nu;
};

template <class L>
__device__ __forceinline__ void closure(const double* P, const Kinematics& k,
                                        const DevParams& prm, double alphaC,
                                        double alphaV, double radius, Closure& c) {
  c.soundSpeed = sqrt(prm.gam * k.pressure / (k.rho + prm.gam * k.u));
  c.bNorm = sqrt(k.bSqr);
  const double cs = c.soundSpeed;
  const double aC = (prm.closure == GRIM_CLOSURE_PER_ZONE) ? alphaC : prm.condAlpha;
  const double aV = (prm.closure == GRIM_CLOSURE_PER_ZONE) ? alphaV : prm.viscAlpha;
  c.q = 0.;
  c.deltaP = 0.;
  if constexpr (L::COND_ON) {
    const double qTilde = P[L::Q];
    c.q = prm.highOrderCond ? qTilde * k.temperature * sqrt(k.rho * aC * cs * cs) : qTilde;
  }
  if constexpr (L::VISC_ON) {
    const double dPTilde = P[L::DP];
    c.deltaP = prm.highOrderVisc ? dPTilde * sqrt(k.temperature * k.rho * aV * cs * cs)
                                 : dPTilde;
  }
  if (prm.closure == GRIM_CLOSURE_LINEAR_MODES) {
    c.tau = 1.0;
    c.chi = cs * cs * c.tau;
    c.nu = cs * cs * c.tau;
  } else if (prm.closure == GRIM_CLOSURE_PER_ZONE) {
    c.tau = prm.tau0;
    c.chi = alphaC * cs * cs * c.tau;
    c.nu = alphaV * cs * cs * c.tau;
  } else if (prm.closure == GRIM_CLOSURE_SHOCK_TESTS) {
    // src/problem/shock_tests/problem.cpp:13-18
    c.tau = 0.1;
    c.chi = prm.condAlpha * cs * cs * c.tau;
    c.nu = prm.viscAlpha * cs * cs * c.tau;
  } else if (prm.closure == GRIM_CLOSURE_BUOYANCY) {
    // src/problem/buoyancy_instabilities/problem.cpp:11-21; pow(r, 0.5) as
    // sqrt(r), correctly rounded on host and device
    c.chi = sqrt(radius);
    c.tau = 1.2 * c.chi / (cs * cs);
    c.nu = 0.;
  } else {
    // pow(r, 1.5) and pow(c_s, 3) of problem.cpp:7-55 as r sqrt(r) and
    // c_s c_s c_s: equal to the last bit, reproducible on host and device,
    // and a fraction of the cost of two library pow calls per element.
    const double dyn = radius * sqrt(radius);
    double tau = dyn;
    if constexpr (L::COND_ON) {
      const double Qmax = prm.condClosureFactor * k.rho * (cs * cs * cs);
      double y = fabs(c.q) / Qmax;
      y = exp(-(y - 1.) / 0.01);
      const double fd = y / (y + 1.) + 1.e-5;
      tau = smin(tau, dyn * fd);
    }
    if constexpr (L::VISC_ON) {
      const double p = k.pressure, dP = c.deltaP, bs = k.bSqr;
      const double dPmod = smax(p - 2. / 3. * dP, prm.bSqrFloor) /
                           smax(p + 1. / 3. * dP, prm.bSqrFloor);
      const double dPmaxPlus = smin(prm.viscClosureFactor * bs * 0.5 * dPmod, 1.49 * p / 1.07);
      const double dPmaxMinus = smax(-prm.viscClosureFactor * bs, -2.99 * p / 1.07);
      const double dPmax = (dP > 0.) ? dPmaxPlus : dPmaxMinus;
      double y = fabs(dP) / (fabs(dPmax) + prm.bSqrFloor);
      y = exp(-(y - 1.) / 0.01);
      const double fd = y / (y + 1.) + 1.e-5;
      tau = smin(tau, dyn * fd);
    }
    c.tau = tau;
    const double maxAlpha = (1. - cs * cs) / (2. * cs * cs + 1.e-12);
    c.chi = smin(maxAlpha, prm.condAlpha) * cs * cs * c.tau;
    c.nu = smin(maxAlpha, prm.viscAlpha) * cs * cs * c.tau;
  }
}

// Per-zone constants of the residual that do not depend on the guess.
template <class L>
struct ZoneConst {
  double rhs[L::NF];
  double mask;
  double alphaC, alphaV, radius;
  // EMHD: elemOld (u_mu, T) and the coefficient element (elemOld on the half
  // step, elemHalfStep on the full step).
  double oldUCov[4], oldT;
  double cRho, cT, cTau, cChi, cNu, cBCon[4], cBSqr, cBNorm, cUCon0, cQTilde, cDPTilde;
  double normQ, normDP;  // residual row normalisation, residual.cpp:61-93
};

// Geometry plane size and the geometry entry of zone z of a batch: one per
// zone, or one per column when geom_columns > 0 (grim_zone_batch).
__device__ __forceinline__ int64_t batchGeomPlane(const grim_zone_batch& b) {
  return b.geom_columns > 0 ? b.geom_columns : b.n;
}
__device__ __forceinline__ int64_t batchGeomIndex(const grim_zone_batch& b, int64_t z) {
  return b.geom_columns > 0 ? (int64_t)((uint32_t)z % (uint32_t)b.geom_columns) : z;
}

// The zone's geometry from the batch's planes.
__device__ __forceinline__ void loadBatchGeom(const grim_zone_batch& b, int64_t z, ZoneGeom& G) {
  loadGeom(b.geom, batchGeomPlane(b), batchGeomIndex(b, z), G);
}

// RHS_IN_MEMORY false: the caller sets rhs and mask itself.
template <class L, bool RHS_IN_MEMORY = true>
__device__ __forceinline__ void loadZoneConst(const grim_zone_batch& b, const DevParams& prm,
                                              const ZoneGeom& G, int64_t z, ZoneConst<L>& c) {
  const int64_t n = b.n;
  if constexpr (RHS_IN_MEMORY) {
#pragma unroll
    for (int v = 0; v < L::NF; v++) c.rhs[v] = b.rhs[v * n + z];
    c.mask = b.residual_mask ? b.residual_mask[z] : 1.;
  }
  c.alphaC = b.closure ? b.closure[z] : 1.;
  c.alphaV = b.closure ? b.closure[n + z] : 1.;
  c.radius = b.radius ? b.radius[batchGeomIndex(b, z)] : 0.;
  if constexpr (L::EMHD) {
    double P[L::DOF];
    Kinematics k;
    Closure cl;
#pragma unroll
    for (int v = 0; v < L::DOF; v++) P[v] = b.prim_old[v * n + z];
    kinematics<L>(P, G, prm, k);
#pragma unroll
    for (int mu = 0; mu < 4; mu++) c.oldUCov[mu] = k.uCov[mu];
    c.oldT = k.temperature;
    if (b.prim_coef) {
#pragma unroll
      for (int v = 0; v < L::DOF; v++) P[v] = b.prim_coef[v * n + z];
      kinematics<L>(P, G, prm, k);
    }
    closure<L>(P, k, prm, c.alphaC, c.alphaV, c.radius, cl);
    c.cRho = k.rho;
    c.cT = k.temperature;
    c.cTau = cl.tau;
    c.cChi = cl.chi;
    c.cNu = cl.nu;
#pragma unroll
    for (int mu = 0; mu < 4; mu++) c.cBCon[mu] = k.bCon[mu];
    c.cBSqr = k.bSqr;
    c.cBNorm = cl.bNorm;
    c.cUCon0 = k.uCon[0];
    c.cQTilde = L::COND_ON ? P[L::Q] : 0.;
    c.cDPTilde = L::VISC_ON ? P[L::DP] : 0.;
    c.normQ = prm.highOrderCond ? c.cT * sqrt(c.cRho * c.cChi * c.cTau) : c.cTau;
    c.normDP = prm.highOrderVisc ? sqrt(c.cRho * c.cNu * c.cT * c.cTau) : c.cTau;
  }
}

// The pieces of one zone's residual at guess P: conserved variables U
// (computeFluxes(0)), implicit sources Si and time-derivative sources Std
// (EMHD rows only; the ideal rows of Si and Std are zero).
template <class L>
__device__ __forceinline__ void residualPieces(const double* P, const ZoneGeom& G,
                                               const ZoneConst<L>& c, const DevParams& prm,
                                               double dt, double* U, double* Si, double* Std) {
  Kinematics k;
  kinematics<L>(P, G, prm, k);
  double q = 0., deltaP = 0., bNorm = 0.;
  if constexpr (L::EMHD) {
    Closure cl;
    closure<L>(P, k, prm, c.alphaC, c.alphaV, c.radius, cl);
    q = cl.q;
    deltaP = cl.deltaP;
    bNorm = cl.bNorm;
  }
  // t-row of T^mu_nu (physics.cpp:153-178) and U = computeFluxes(0)
  // (physics.cpp:185-217).
  double T0[4];
  const double w = k.rho + k.u + k.pressure + k.bSqr;
#pragma unroll
  for (int nu = 0; nu < 4; nu++) {
    double t = w * k.uCon[0] * k.uCov[nu];
    if (nu == 0) t = t + (k.pressure + 0.5 * k.bSqr);
    t = t - k.bCon[0] * k.bCov[nu];
    if constexpr (L::COND_ON) t += q / bNorm * (k.uCon[0] * k.bCov[nu] + k.bCon[0] * k.uCov[nu]);
    if constexpr (L::VISC_ON)
      t += (-deltaP) * (k.bCon[0] * k.bCov[nu] / k.bSqr -
                        (1. / 3.) * ((nu == 0 ? 1. : 0.) + k.uCon[0] * k.uCov[nu]));
    T0[nu] = t;
  }
  U[RHO] = G.g * (k.rho * k.uCon[0]);
  U[UU] = G.g * T0[0] + U[RHO];
  U[U1] = G.g * T0[1];
  U[U2] = G.g * T0[2];
  U[U3] = G.g * T0[3];
  if constexpr (L::COND_ON) U[L::Q] = G.g * (k.uCon[0] * P[L::Q]);
  if constexpr (L::VISC_ON) U[L::DP] = G.g * (k.uCon[0] * P[L::DP]);
  if constexpr (L::EMHD) {
    // computeTimeDerivSources on the coefficient element
    // (physics.cpp:219-318) and computeImplicitSources (physics.cpp:320-358).
    double dtuCov[4];
#pragma unroll
    for (int mu = 0; mu < 4; mu++) dtuCov[mu] = (k.uCov[mu] - c.oldUCov[mu]) / dt;
    double divuCov = 0.;
#pragma unroll
    for (int mu = 0; mu < 4; mu++) divuCov += G.gcon0[mu] * dtuCov[mu];
    if constexpr (L::VISC_ON) {
      double dP0 = -divuCov * c.cRho * c.cNu;
#pragma unroll
      for (int mu = 0; mu < 4; mu++)
        dP0 += 3. * c.cRho * c.cNu * c.cBCon[0] * c.cBCon[mu] / c.cBSqr * dtuCov[mu];
      if (prm.highOrderVisc) dP0 *= sqrt(c.cTau / c.cRho / c.cNu / c.cT);
      double st = -G.g * (dP0) / c.cTau;
      if (prm.highOrderVisc) st -= 0.5 * G.g * divuCov * c.cDPTilde;
      Std[L::DP] = st;
      Si[L::DP] = G.g * (P[L::DP]) / c.cTau;
    }
    if constexpr (L::COND_ON) {
      double q0 = -c.cRho * c.cChi * c.cBCon[0] / c.cBNorm * (k.temperature - c.oldT) / dt;
#pragma unroll
      for (int nu = 0; nu < 4; nu++)
        q0 -= c.cRho * c.cChi * c.cT * c.cBCon[nu] / c.cBNorm * c.cUCon0 * dtuCov[nu];
      if (prm.highOrderCond) q0 *= sqrt(c.cTau / c.cRho / c.cChi) / c.cT;
      double st = -G.g * (q0) / c.cTau;
      if (prm.highOrderCond) st -= 0.5 * G.g * divuCov * c.cQTilde;
      Std[L::Q] = st;
      Si[L::Q] = G.g * (P[L::Q]) / c.cTau;
    }
  }
}

// Residual rows of one zone at guess P (residual.cpp:37-45 with the EMHD row
// normalisation of :61-93 and the mask of :173-177).
template <class L>
__device__ __forceinline__ void residual(const double* P, const ZoneGeom& G,
                                         const ZoneConst<L>& c, const DevParams& prm,
                                         double dt, double* R) {
  double U[L::NF], Si[L::NF], Std[L::NF];
  residualPieces<L>(P, G, c, prm, dt, U, Si, Std);
#pragma unroll
  for (int v = 0; v < 5; v++) R[v] = (U[v] - c.rhs[v]) / dt;
  if constexpr (L::COND_ON) R[L::Q] = ((U[L::Q] - c.rhs[L::Q]) / dt + 0.5 * Si[L::Q] + Std[L::Q]) * c.normQ;
  if constexpr (L::VISC_ON) R[L::DP] = ((U[L::DP] - c.rhs[L::DP]) / dt + 0.5 * Si[L::DP] + Std[L::DP]) * c.normDP;
#pragma unroll
  for (int v = 0; v < L::NF; v++) R[v] *= c.mask;
}

template <int N>
__device__ __forceinline__ double sumSq(const double* R) {
  double s = R[0] * R[0];
#pragma unroll
  for (int v = 1; v < N; v++) s = s + R[v] * R[v];
  return s;
}

// Dense LU with partial pivoting (pivot: first row with the largest |A(i,k)|),
// reciprocal-of-pivot multipliers, in place on register arrays.  Solves
// A x = b, x returned in b.  Real is double or float.
//
// The pivot row is brought up by compare-and-swap: row k trades places with
// row i whenever |A(i,k)| exceeds the current |A(k,k)|.  That selects the
// same pivot as LAPACK's search-then-swap, and every remaining row is
// eliminated by the same operations wherever it sits, so the solution is
// bit-identical to the search-then-swap order (exact ties aside).  Every
// access stays at a compile-time index: a swap with a computed row index
// would force the matrix out of registers into scratch memory.  PIVOT false
// factorises in the given row order (the mixed path's equilibrated FP32
// systems).

// 1/x: FP64 correctly rounded; FP32 the hardware reciprocal (within one ulp;
// the FP32 factorisation is an approximate solve refined by the FP64 Newton
// updates around it).
__device__ __forceinline__ double lowRecip(double x) { return 1. / x; }
__device__ __forceinline__ float lowRecip(float x) { return __builtin_amdgcn_rcpf(x); }

template <class Real, int N, bool PIVOT = true>
__device__ __forceinline__ void luSolve(Real (&A)[N][N], Real (&b)[N]) {
  Real rinv[N];
#pragma unroll
  for (int k = 0; k < N; k++) {
#pragma unroll
    for (int i = k + 1; i < N && PIVOT; i++) {
      const bool swap = fabs(A[i][k]) > fabs(A[k][k]);
#pragma unroll
      for (int j = k; j < N; j++) {
        const Real hi = A[k][j], lo = A[i][j];
        A[k][j] = swap ? lo : hi;
        A[i][j] = swap ? hi : lo;
      }
      const Real hi = b[k], lo = b[i];
      b[k] = swap ? lo : hi;
      b[i] = swap ? hi : lo;
    }
    rinv[k] = lowRecip(A[k][k]);
#pragma unroll
    for (int i = k + 1; i < N; i++) {
      const Real m = A[i][k] * rinv[k];
      A[i][k] = m;
#pragma unroll
      for (int j = k + 1; j < N; j++) A[i][j] = A[i][j] - m * A[k][j];
      b[i] = b[i] - m * b[k];
    }
  }
#pragma unroll
  for (int i = N - 1; i >= 0; i--) {
    Real s = b[i];
#pragma unroll
    for (int j = i + 1; j < N; j++) s = s - A[i][j] * b[j];
    b[i] = s * rinv[i];
  }
}

}  // namespace grim
