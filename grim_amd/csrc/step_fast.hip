// The fast explicit stage: the two heaviest kernels of grim's stage, the face
// fluxes (riemannSolver::solve with reconstruction, riemannsolver.cpp:
// 202-328) and the explicit sources (fluidElement::computeExplicitSources
// with computeEMHDGradients, physics.cpp:360-586), compiled with
// -ffp-contract=fast-honor-pragmas (fused multiply-adds).
//
// Same mathematics as the exact kernels of step.hip, reorganised for fewer
// instructions:
//  * WENO5 (weno5.cpp:111-257) forms each face value with one division: the
//    nonlinear weights c_k / beta_k^2 and their sum are scaled by
//    16 beta_1^2 beta_2^2 beta_3^2, which leaves products of the betas;
//  * each quotient family with a common divisor (1/u^0 in b^i, 1/|b| and
//    1/b^2 in the stress rows, 1/tau and 1/(rho + Gamma u) in the speeds,
//    1/(2A) for the two roots) takes one reciprocal;
//  * the characteristic speeds read the inverse metric entries the unit
//    covectors pick instead of contracting with zeros;
//  * the sources read the 40 symmetric connection planes once, in one pass
//    over the upper index that feeds both the geometric source and the
//    covariant derivative of u_mu;
//  * a*b + c contracts to a fused multiply-add.
// Rounding differs from grim's order at the 1e-16 relative level; the
// results are tested against the exact kernels and the oracle to 1e-12.
#include <hip/hip_runtime.h>

#include "fast_ops.hpp"
#include "grim_hip.h"
#include "runtime.hpp"
#include "step_fast.hpp"
#include "step_physics.hpp"
#include "zone_physics.hpp"

namespace grim {
namespace {

// WENO5 face value of one zone from y[0..4] = in[i-2 .. i+2]: SIDE 0 the
// value at i + 1/2 (grim's primRight), SIDE 1 at i - 1/2 (primLeft).
// grim's smoothness indicators (weno5.cpp:165-193) are the quadratic forms
// beta_k = 13/12 (second difference)^2 + 1/4 (first-difference combination)^2;
// they are formed here as 12 beta_k (eps2 scaled with them), a factor that
// cancels in the normalised weights.  The weights c_k / beta_k^2 and their
// sum are scaled by beta_1^2 beta_2^2 beta_3^2 and the linear weights folded
// into the candidate stencils, which leaves one quotient.
template <int SIDE>
__device__ __forceinline__ double weno5Fast(const double* y) {
  const double eps = 12. * 1.0e-17;
  const double y0 = y[0], y1 = y[1], y2 = y[2], y3 = y[3], y4 = y[4];
  const double a12 = fabs(y1) + fabs(y2), a23 = fabs(y2) + fabs(y3);
  const double d1 = y0 - 2. * y1 + y2, e1 = y0 - 4. * y1 + 3. * y2;
  const double d2 = y1 - 2. * y2 + y3, e2 = y1 - y3;
  const double d3 = y2 - 2. * y3 + y4, e3 = 3. * y2 - 4. * y3 + y4;
  const double beta1 = 13. * d1 * d1 + 3. * e1 * e1 + eps * (1. + fabs(y0) + a12);
  const double beta2 = 13. * d2 * d2 + 3. * e2 * e2 + eps * (1. + a12 + fabs(y3));
  const double beta3 = 13. * d3 * d3 + 3. * e3 * e3 + eps * (1. + a23 + fabs(y4));
  const double b1 = beta1 * beta1, b2 = beta2 * beta2, b3 = beta3 * beta3;
  const double p1 = b2 * b3, p2 = b1 * b3, p3 = b1 * b2;
  double num, den;
  if constexpr (SIDE == 0) {
    // linear weights (1, 10, 5) / 16 of the candidates at i + 1/2
    num = p1 * (0.375 * y0 - 1.25 * y1 + 1.875 * y2) + p2 * (-1.25 * y1 + 7.5 * y2 + 3.75 * y3) +
          p3 * (1.875 * y2 + 3.75 * y3 - 0.625 * y4);
    den = p1 + 10. * p2 + 5. * p3;
  } else {
    // (5, 10, 1) / 16 at i - 1/2
    num = p1 * (-0.625 * y0 + 3.75 * y1 + 1.875 * y2) + p2 * (3.75 * y1 + 7.5 * y2 - 1.25 * y3) +
          p3 * (1.875 * y2 - 1.25 * y3 + 0.375 * y4);
    den = 5. * p1 + 10. * p2 + p3;
  }
  return fastDiv(num, den);
}

template <int RECON, int SIDE>
__device__ __forceinline__ double faceValue(const double* y, double theta) {
  if constexpr (RECON == GRIM_RECON_WENO5) {
    return weno5Fast<SIDE>(y);
  } else {
    double l, r;
    reconstructZone<RECON>(y, theta, l, r);
    return SIDE ? l : r;
  }
}

// fluidElement::set's kinematics (physics.cpp:60-116) with one reciprocal of
// u^0 for the three b^i.
template <class L>
__device__ __forceinline__ void kinematicsFast(const double* P, const ZoneGeom& G, const DevParams& prm,
                                               Kinematics& k) {
  k.rho = smax(P[RHO], prm.rhoFloor);
  k.u = smax(P[UU], prm.uFloor);
  const double u1 = P[U1], u2 = P[U2], u3 = P[U3];
  const double B1 = P[L::B1], B2 = P[L::B2], B3 = P[L::B3];
  k.pressure = (prm.gam - 1.) * k.u;
  k.temperature = smax(fastDiv(k.pressure, k.rho), prm.tFloor);
  const double gamma = sqrt(1. + G.cov(1, 1) * u1 * u1 + G.cov(2, 2) * u2 * u2 + G.cov(3, 3) * u3 * u3 +
                            2. * (G.cov(1, 2) * u1 * u2 + G.cov(1, 3) * u1 * u3 + G.cov(2, 3) * u2 * u3));
  k.gammaLF = gamma;
  k.uCon[0] = fastDiv(gamma, G.alpha);
  const double invU0 = fastDiv(G.alpha, gamma);
  const double ga = gamma * G.alpha;
  k.uCon[1] = u1 - ga * G.gcon0[1];
  k.uCon[2] = u2 - ga * G.gcon0[2];
  k.uCon[3] = u3 - ga * G.gcon0[3];
#pragma unroll
  for (int mu = 0; mu < 4; mu++)
    k.uCov[mu] = G.cov(mu, 0) * k.uCon[0] + G.cov(mu, 1) * k.uCon[1] + G.cov(mu, 2) * k.uCon[2] +
                 G.cov(mu, 3) * k.uCon[3];
  k.bCon[0] = B1 * k.uCov[1] + B2 * k.uCov[2] + B3 * k.uCov[3];
  k.bCon[1] = (B1 + k.bCon[0] * k.uCon[1]) * invU0;
  k.bCon[2] = (B2 + k.bCon[0] * k.uCon[2]) * invU0;
  k.bCon[3] = (B3 + k.bCon[0] * k.uCon[3]) * invU0;
#pragma unroll
  for (int mu = 0; mu < 4; mu++)
    k.bCov[mu] = G.cov(mu, 0) * k.bCon[0] + G.cov(mu, 1) * k.bCon[1] + G.cov(mu, 2) * k.bCon[2] +
                 G.cov(mu, 3) * k.bCon[3];
  k.bSqr = k.bCon[0] * k.bCov[0] + k.bCon[1] * k.bCov[1] + k.bCon[2] * k.bCov[2] + k.bCon[3] * k.bCov[3] +
           prm.bSqrFloor;
}

// setFluidElementParameters and the rescaled q, deltaP of set() (physics.cpp:
// 118-151; torus closure torus/problem.cpp:7-55), the element's closure with
// unit alphas, as elementClosure.
template <class L>
__device__ __forceinline__ void closureFast(const double* P, const Kinematics& k, const DevParams& prm,
                                            double radius, Closure& c) {
  const double cs2 = fastDiv(prm.gam * k.pressure, k.rho + prm.gam * k.u);
  c.soundSpeed = sqrt(cs2);
  c.bNorm = sqrt(k.bSqr);
  c.q = 0.;
  c.deltaP = 0.;
  c.tau = c.chi = c.nu = 0.;
  if constexpr (L::EMHD) {
    const double aC = prm.condAlpha, aV = prm.viscAlpha;
    if constexpr (L::COND_ON)
      c.q = prm.highOrderCond ? P[L::Q] * k.temperature * sqrt(k.rho * aC * cs2) : P[L::Q];
    if constexpr (L::VISC_ON)
      c.deltaP = prm.highOrderVisc ? P[L::DP] * sqrt(k.temperature * k.rho * aV * cs2) : P[L::DP];
    if (prm.closure == GRIM_CLOSURE_LINEAR_MODES) {
      c.tau = 1.0;
      c.chi = cs2;
      c.nu = cs2;
    } else if (prm.closure == GRIM_CLOSURE_SHOCK_TESTS) {
      c.tau = 0.1;
      c.chi = prm.condAlpha * cs2 * c.tau;
      c.nu = prm.viscAlpha * cs2 * c.tau;
    } else if (prm.closure == GRIM_CLOSURE_BUOYANCY) {
      c.chi = sqrt(radius);
      c.tau = fastDiv(1.2 * c.chi, cs2);
    } else {
      const double dyn = radius * sqrt(radius);
      double tau = dyn;
      const double cs3 = cs2 * c.soundSpeed;
      if constexpr (L::COND_ON) {
        double y = fastDiv(fabs(c.q), prm.condClosureFactor * k.rho * cs3);
        y = exp((1. - y) * 100.);
        tau = smin(tau, dyn * (fastDiv(y, y + 1.) + 1.e-5));
      }
      if constexpr (L::VISC_ON) {
        const double p = k.pressure, dP = c.deltaP, bs = k.bSqr;
        const double dPmod = fastDiv(smax(p - (2. / 3.) * dP, prm.bSqrFloor), smax(p + (1. / 3.) * dP, prm.bSqrFloor));
        const double dPmaxPlus = smin(prm.viscClosureFactor * bs * 0.5 * dPmod, (1.49 / 1.07) * p);
        const double dPmaxMinus = smax(-prm.viscClosureFactor * bs, (-2.99 / 1.07) * p);
        const double dPmax = (dP > 0.) ? dPmaxPlus : dPmaxMinus;
        double y = fastDiv(fabs(dP), fabs(dPmax) + prm.bSqrFloor);
        y = exp((1. - y) * 100.);
        tau = smin(tau, dyn * (fastDiv(y, y + 1.) + 1.e-5));
      }
      c.tau = tau;
      const double maxAlpha = fastDiv(1. - cs2, 2. * cs2 + 1.e-12);
      c.chi = smin(maxAlpha, prm.condAlpha) * cs2 * c.tau;
      c.nu = smin(maxAlpha, prm.viscAlpha) * cs2 * c.tau;
    }
  }
}

// Reciprocals every stress row shares.
struct StressInv {
  double qOverB;   // q / |b|
  double dPOverB2; // deltaP / b^2
};

template <class L>
__device__ __forceinline__ StressInv stressInv(const Kinematics& k, const Closure& c) {
  StressInv s{0., 0.};
  if constexpr (L::COND_ON) s.qOverB = fastDiv(c.q, c.bNorm);
  if constexpr (L::VISC_ON) s.dPOverB2 = fastDiv(c.deltaP, k.bSqr);
  return s;
}

// Row MU of T^mu_nu (physics.cpp:153-178).
template <class L, int MU>
__device__ __forceinline__ void stressRowFast(const Kinematics& k, const Closure& c, const StressInv& si,
                                              double* T) {
  const double w = k.rho + k.u + k.pressure + k.bSqr;
  const double wu = w * k.uCon[MU];
#pragma unroll
  for (int nu = 0; nu < 4; nu++) {
    double t = wu * k.uCov[nu] - k.bCon[MU] * k.bCov[nu];
    if (MU == nu) t += k.pressure + 0.5 * k.bSqr;
    if constexpr (L::COND_ON) t += si.qOverB * (k.uCon[MU] * k.bCov[nu] + k.bCon[MU] * k.uCov[nu]);
    if constexpr (L::VISC_ON)
      t += -si.dPOverB2 * (k.bCon[MU] * k.bCov[nu]) +
           (c.deltaP * (1. / 3.)) * ((MU == nu ? 1. : 0.) + k.uCon[MU] * k.uCov[nu]);
    T[nu] = t;
  }
}

// fluidElement::computeFluxes(DIR) (physics.cpp:185-217); DIR = 0 gives U.
template <class L, int DIR>
__device__ __forceinline__ void fluxesFast(const double* P, const Kinematics& k, const Closure& c,
                                           const StressInv& si, double g, double* F) {
  double T[4];
  stressRowFast<L, DIR>(k, c, si, T);
  const double gu = g * k.uCon[DIR];
  F[RHO] = gu * k.rho;
  F[UU] = g * T[0] + F[RHO];
  F[U1] = g * T[1];
  F[U2] = g * T[2];
  F[U3] = g * T[3];
  // the component along DIR is b^d u^d - b^d u^d = 0 exactly, as grim
  // computes it (a contracted b^d u^d - b^d u^d would leave the product's
  // rounding error)
#pragma unroll
  for (int c = 1; c <= 3; c++)
    F[L::B1 + c - 1] = (c == DIR) ? 0. : g * (k.bCon[c] * k.uCon[DIR] - k.bCon[DIR] * k.uCon[c]);
  if constexpr (L::COND_ON) F[L::Q] = gu * P[L::Q];
  if constexpr (L::VISC_ON) F[L::DP] = gu * P[L::DP];
}

// computeMinMaxCharSpeeds (riemannsolver.cpp:54-200) along DIR = 0, 1, 2.
// The unit covectors A = dX^(DIR+1), B = dt pick single inverse-metric
// entries: A.A = g^{ss}, B.B = g^{00}, A.B = g^{s0}, A.u = u^s, B.u = u^0.
template <class L, int DIR>
__device__ __forceinline__ void charSpeedsFast(const Kinematics& k, const Closure& c, double gconSS, double gcon00,
                                               double gcon0S, const DevParams& prm, double& minSpeed,
                                               double& maxSpeed) {
  const double rhoGu = k.rho + prm.gam * k.u;
  const double cAlvenSqr = fastDiv(k.bSqr, rhoGu + k.bSqr);
  double csSqr = c.soundSpeed * c.soundSpeed;
  double cVisSqr = 0., cConSqr = 0.;
  if constexpr (L::EMHD) {
    const double invTau = fastRcp(c.tau);
    if constexpr (L::VISC_ON) cVisSqr = fastDiv((4. / 3.) * k.rho * c.nu * invTau, rhoGu);
    if constexpr (L::COND_ON) cConSqr = (prm.gam - 1.) * c.chi * invTau;
  }
  cConSqr = 0.5 * (csSqr + cConSqr + sqrt(csSqr * csSqr + cConSqr * cConSqr));
  csSqr = cConSqr + cVisSqr;
  csSqr = csSqr + cAlvenSqr - csSqr * cAlvenSqr;
  csSqr = csSqr > 1. ? 1. : csSqr;
  constexpr int sdir = DIR + 1;
  const double ASqr = gconSS, BSqr = gcon00, ADotB = gcon0S;
  const double ADotU = k.uCon[sdir], BDotU = k.uCon[0];
  const double A = (BDotU * BDotU) - (BSqr + BDotU * BDotU) * csSqr;
  const double B = 2. * (ADotU * BDotU - (ADotB + ADotU * BDotU) * csSqr);
  const double C = ADotU * ADotU - (ASqr + ADotU * ADotU) * csSqr;
  const double discr = sqrt(smax(B * B - 4. * A * C, 1.e-16));
  const double inv2A = 0.5 * fastRcp(A);
  double mn = (B - discr) * inv2A;
  double mx = (B + discr) * inv2A;
  mn = mn > -1.e-15 ? -1.e-15 : mn;
  mx = mx < 1.e-15 ? 1.e-15 : mx;
  minSpeed = mn;
  maxSpeed = mx;
}

// The geometry one side of a face needs: the zone geometry, the inverse
// metric entry g^{ss} of the face direction and r (the torus closure).
struct FaceGeom {
  ZoneGeom zg;
  double gconSS, r;
};

template <int DIR>
__device__ __forceinline__ void loadFaceGeom(const double* geom, int64_t n, uint32_t zb, FaceGeom& G) {
#pragma unroll
  for (int k = 0; k < 10; k++) G.zg.gcov[k] = ldOff(geom + k * n, zb);
#pragma unroll
  for (int k = 0; k < 4; k++) G.zg.gcon0[k] = ldOff(geom + (GEOM_GCON0 + k) * n, zb);
  G.zg.alpha = ldOff(geom + GEOM_ALPHA * n, zb);
  G.zg.g = ldOff(geom + GEOM_G * n, zb);
  // planes 16..21 hold g^{11} g^{12} g^{13} g^{22} g^{23} g^{33}
  constexpr int ss = DIR == 0 ? 0 : (DIR == 1 ? 3 : 5);
  G.gconSS = ldOff(geom + (16 + ss) * n, zb);
  G.r = ldOff(geom + LOC_R * n, zb);
}

// One side of a face (as faceState in step.hip): the face value of every
// variable, the element at the face geometry, its flux, conserved
// variables and speeds.  zb: the byte offset of the face's zone in a plane,
// zgb: that of its (X1, X2) column in a geometry plane.
template <class L, int DIR, int RECON, int SIDE>
__device__ __forceinline__ void faceStateFast(const StepArgs& a, uint32_t zb, uint32_t zgb,
                                              const double* __restrict__ prim,
                                              const double* __restrict__ geomFace, double (&F)[L::DOF],
                                              double (&U)[L::DOF], double& mn, double& mx) {
  const int64_t n = a.D.n;
  // the stencil's offsets along DIR, shared by every variable's plane
  uint32_t off[5];
#pragma unroll
  for (int q = 0; q < 5; q++) off[q] = zb + (uint32_t)((q - 3 + SIDE) * a.D.s[DIR] * 8);
  double P[L::DOF];
#pragma unroll
  for (int v = 0; v < L::DOF; v++) {
    double y[5];
#pragma unroll
    for (int q = 0; q < 5; q++) y[q] = ldOff(prim + v * n, off[q]);
    P[v] = faceValue<RECON, SIDE>(y, a.theta);
  }
  FaceGeom G;
  loadFaceGeom<DIR>(geomFace, a.D.ng, zgb, G);
  Kinematics kk;
  Closure cc;
  kinematicsFast<L>(P, G.zg, a.prm, kk);
  closureFast<L>(P, kk, a.prm, G.r, cc);
  const StressInv si = stressInv<L>(kk, cc);
  fluxesFast<L, DIR + 1>(P, kk, cc, si, G.zg.g, F);
  fluxesFast<L, 0>(P, kk, cc, si, G.zg.g, U);
  charSpeedsFast<L, DIR>(kk, cc, G.gconSS, G.zg.gcon0[0], G.zg.gcon0[DIR + 1], a.prm, mn, mx);
}

// The split face kernel of step.hip with the fast element: wavefronts 0-1
// the left states, 2-3 the right states, which meet in LDS.
template <class L, int DIR, int RECON>
__global__ __launch_bounds__(kBlock) void faceFluxFastKernel(StepArgs a, Box box, const double* __restrict__ prim,
                                                             const double* __restrict__ geomFace,
                                                             double* __restrict__ flux) {
  __shared__ double sh[2 * L::DOF + 2][kFacesPerBlock];
  const int side = threadIdx.x / kFacesPerBlock;
  const int lf = threadIdx.x % kFacesPerBlock;
  const int64_t f = (int64_t)blockIdx.x * kFacesPerBlock + lf;
  const bool active = f < box.count();
  const int64_t n = a.D.n;
  uint32_t zb = 0;
  double F[L::DOF], U[L::DOF], mn = 0., mx = 0.;
  if (active) {
    int i, j, k;
    box.at(f, i, j, k);
    const uint32_t zgb = ((uint32_t)i + (uint32_t)j * (uint32_t)a.D.s[1]) * 8u;
    zb = zgb + (uint32_t)k * (uint32_t)a.D.s[2] * 8u;
    if (side == 0) {
      faceStateFast<L, DIR, RECON, 0>(a, zb, zgb, prim, geomFace, F, U, mn, mx);
    } else {
      faceStateFast<L, DIR, RECON, 1>(a, zb, zgb, prim, geomFace, F, U, mn, mx);
#pragma unroll
      for (int v = 0; v < L::DOF; v++) {
        sh[v][lf] = F[v];
        sh[L::DOF + v][lf] = U[v];
      }
      sh[2 * L::DOF][lf] = mn;
      sh[2 * L::DOF + 1][lf] = mx;
    }
  }
  __syncthreads();
  if (!active || side != 0) return;
  const double minSpeed = stdmin(mn, sh[2 * L::DOF][lf]);
  const double maxSpeed = stdmax(mx, sh[2 * L::DOF + 1][lf]);
  if (a.riemann == GRIM_RIEMANN_HLL) {
    const double inv = fastRcp(maxSpeed - minSpeed);
    const double mm = minSpeed * maxSpeed;
#pragma unroll
    for (int v = 0; v < L::DOF; v++) {
      const double FR = sh[v][lf], UR = sh[L::DOF + v][lf];
      stOff(flux + v * n, zb, (maxSpeed * F[v] - minSpeed * FR + mm * (UR - U[v])) * inv);
    }
  } else {
    const double c = stdmax(maxSpeed, -minSpeed);
#pragma unroll
    for (int v = 0; v < L::DOF; v++) {
      const double FR = sh[v][lf], UR = sh[L::DOF + v][lf];
      stOff(flux + v * n, zb, 0.5 * (F[v] + FR - c * (UR - U[v])));
    }
  }
}

__host__ __device__ constexpr int sym4(int k, int m) {
  return (k < m ? k : m) * 4 - ((k < m ? k : m) * ((k < m ? k : m) + 1)) / 2 + (k < m ? m : k);
}

// computeExplicitSources (physics.cpp:360-586) for the bulk zones of box:
// one pass over the connection's upper index l, whose 10 symmetric planes
// feed the geometric source -g T^k_l Gamma^l_{k n} and the covariant
// derivative u_{mu;nu} = d_nu u_mu - Gamma^l_{nu mu} u_l; the old implicit
// sources with the stage element's tau; on the half stage consOld.
template <class L, int RECON>
__global__ __launch_bounds__(kBlock) void sourcesFastKernel(StepArgs a, Box box, const double* __restrict__ prim,
                                                            const double* __restrict__ primOld,
                                                            const double* __restrict__ geomC,
                                                            const double* __restrict__ gamma40,
                                                            const double* __restrict__ uT, double* __restrict__ sExp,
                                                            double* __restrict__ sImp, double* __restrict__ consOld) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= box.count()) return;
  int i, j, k;
  box.at(t, i, j, k);
  const Dims& D = a.D;
  const int64_t n = D.n, z = flatIndex(D, i, j, k);
  const int64_t ng = D.ng, zg = geomIndex(D, i, j);
  double P[L::DOF];
#pragma unroll
  for (int v = 0; v < L::DOF; v++) P[v] = prim[v * n + z];
  GeomFull G;
  loadGeomFull(geomC, ng, zg, G);
  Kinematics kk;
  Closure cc;
  kinematicsFast<L>(P, G.zg, a.prm, kk);
  closureFast<L>(P, kk, a.prm, G.r, cc);
  const StressInv si = stressInv<L>(kk, cc);
  const double g = G.zg.g;
  if (consOld) {
    double U[L::DOF];
    fluxesFast<L, 0>(P, kk, cc, si, g, U);
#pragma unroll
    for (int v = 0; v < L::DOF; v++) consOld[v * n + z] = U[v];
  }
  double src[L::NF];
#pragma unroll
  for (int v = 0; v < L::NF; v++) src[v] = 0.;
  const bool curved = a.metric != GRIM_METRIC_MINKOWSKI;
  // T^k_l, k-major
  double T[4][4];
  if (curved) {
    stressRowFast<L, 0>(kk, cc, si, T[0]);
    stressRowFast<L, 1>(kk, cc, si, T[1]);
    stressRowFast<L, 2>(kk, cc, si, T[2]);
    stressRowFast<L, 3>(kk, cc, si, T[3]);
  }
  double graduCov[4][4];
  if constexpr (L::EMHD) {
    auto slopeOf = [&](const double* arr, int d) {
      double y[5];
#pragma unroll
      for (int q = 0; q < 5; q++) y[q] = arr[z + (q - 2) * D.s[d]];
      return slopeZone<RECON>(y, a.theta, a.dX[d]);
    };
#pragma unroll
    for (int mu = 0; mu < 4; mu++) {
      const double* um = uT + mu * n;
      graduCov[0][mu] = 0.;
      graduCov[1][mu] = slopeOf(um, 0);
      graduCov[2][mu] = D.dim > 1 ? slopeOf(um, 1) : 0.;
      graduCov[3][mu] = D.dim > 2 ? slopeOf(um, 2) : 0.;
    }
  }
  if (curved || L::EMHD) {
#pragma unroll
    for (int l = 0; l < 4; l++) {
      double gam[10];
#pragma unroll
      for (int q = 0; q < 10; q++) gam[q] = gamma40[(l * 10 + q) * ng + zg];
      if (curved) {
#pragma unroll
        for (int nu = 0; nu < 4; nu++) {
          double s = 0.;
#pragma unroll
          for (int kap = 0; kap < 4; kap++) s += T[kap][l] * gam[sym4(kap, nu)];
          src[UU + nu] -= g * s;
        }
      }
      if constexpr (L::EMHD) {
        const double ul = kk.uCov[l];
#pragma unroll
        for (int nu = 0; nu < 4; nu++)
#pragma unroll
          for (int mu = 0; mu < 4; mu++) graduCov[nu][mu] -= gam[sym4(nu, mu)] * ul;
      }
    }
  }
  if constexpr (L::EMHD) {
    double divuCov = 0.;
#pragma unroll
    for (int mu = 0; mu < 4; mu++)
#pragma unroll
      for (int nu = 0; nu < 4; nu++) divuCov += G.con(mu, nu) * graduCov[mu][nu];
    const double invTau = 1. / cc.tau;
    if constexpr (L::VISC_ON) {
      const double rn = kk.rho * cc.nu;
      const double c3 = 3. * rn / kk.bSqr;
      double bb = 0.;
#pragma unroll
      for (int mu = 0; mu < 4; mu++)
#pragma unroll
        for (int nu = 0; nu < 4; nu++) bb += kk.bCon[mu] * kk.bCon[nu] * graduCov[mu][nu];
      double deltaP0 = -divuCov * rn + c3 * bb;
      if (a.prm.highOrderVisc == 1) deltaP0 = deltaP0 * sqrt(cc.tau / (rn * kk.temperature));
      double sd = -g * deltaP0 * invTau;
      if (a.prm.highOrderVisc == 1) sd -= 0.5 * g * divuCov * P[L::DP];
      src[L::DP] = sd;
    }
    if constexpr (L::COND_ON) {
      const double* Tp = uT + 4 * n;
      double gradT[4];
      gradT[0] = 0.;
      {
        double y[5];
#pragma unroll
        for (int q = 0; q < 5; q++) y[q] = Tp[z + (q - 2) * D.s[0]];
        gradT[1] = slopeZone<RECON>(y, a.theta, a.dX[0]);
      }
      gradT[2] = 0.;
      gradT[3] = 0.;
      if (D.dim > 1) {
        double y[5];
#pragma unroll
        for (int q = 0; q < 5; q++) y[q] = Tp[z + (q - 2) * D.s[1]];
        gradT[2] = slopeZone<RECON>(y, a.theta, a.dX[1]);
      }
      if (D.dim > 2) {
        double y[5];
#pragma unroll
        for (int q = 0; q < 5; q++) y[q] = Tp[z + (q - 2) * D.s[2]];
        gradT[3] = slopeZone<RECON>(y, a.theta, a.dX[2]);
      }
      // q0 = -rho chi / |b| (b^mu d_mu T + T b^nu u^mu u_{nu;mu})
      const double rcb = kk.rho * cc.chi / cc.bNorm;
      double bgT = 0., buu = 0.;
#pragma unroll
      for (int mu = 0; mu < 4; mu++) {
        bgT += kk.bCon[mu] * gradT[mu];
#pragma unroll
        for (int nu = 0; nu < 4; nu++) buu += kk.bCon[nu] * kk.uCon[mu] * graduCov[mu][nu];
      }
      double q0 = -rcb * (bgT + kk.temperature * buu);
      if (a.prm.highOrderCond == 1) q0 = q0 * sqrt(cc.tau / (kk.rho * cc.chi)) / kk.temperature;
      double sq = -g * q0 * invTau;
      if (a.prm.highOrderCond == 1) sq -= 0.5 * g * divuCov * P[L::Q];
      src[L::Q] = sq;
    }
    if (sImp) {
#pragma unroll
      for (int v = 0; v < L::NF; v++) sImp[v * n + z] = 0.;
      const double gt = g * invTau;
      if constexpr (L::VISC_ON) sImp[L::DP * n + z] = gt * primOld[L::DP * n + z];
      if constexpr (L::COND_ON) sImp[L::Q * n + z] = gt * primOld[L::Q * n + z];
    }
  }
#pragma unroll
  for (int v = 0; v < L::NF; v++) sExp[v * n + z] = src[v];
}

__global__ __launch_bounds__(kBlock) void gamma40Kernel(int64_t n, const double* __restrict__ g64,
                                                        double* __restrict__ g40) {
  const int64_t z = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (z >= n) return;
#pragma unroll
  for (int l = 0; l < 4; l++)
#pragma unroll
    for (int k = 0; k < 4; k++)
#pragma unroll
      for (int m = k; m < 4; m++) g40[(l * 10 + sym4(k, m)) * n + z] = g64[(l * 16 + k * 4 + m) * n + z];
}

inline unsigned blocks(int64_t count, int per) { return (unsigned)((count + per - 1) / per); }

}  // namespace

int launchFaceFast(const grim_params& p, const StepArgs& a, int dir, int recon, const Box& b, const double* prim,
                   const double* geomFace, double* flux, hipStream_t s) {
  GRIM_REQUIRE(planeOffsetsFit(a.D.n), "fast explicit stage: planes must hold fewer than 2^29 zones");
  const unsigned nb = blocks(b.count(), kFacesPerBlock);
  return dispatchLayout(p, [&]<class L>() -> int {
    auto go = [&]<int DIR>() {
      switch (recon) {
        case GRIM_RECON_WENO5:
          faceFluxFastKernel<L, DIR, GRIM_RECON_WENO5><<<nb, kBlock, 0, s>>>(a, b, prim, geomFace, flux);
          break;
        case GRIM_RECON_PPM:
          faceFluxFastKernel<L, DIR, GRIM_RECON_PPM><<<nb, kBlock, 0, s>>>(a, b, prim, geomFace, flux);
          break;
        default:
          faceFluxFastKernel<L, DIR, GRIM_RECON_MINMOD><<<nb, kBlock, 0, s>>>(a, b, prim, geomFace, flux);
      }
    };
    if (dir == 0) go.template operator()<0>();
    else if (dir == 1) go.template operator()<1>();
    else go.template operator()<2>();
    return checkLaunch("face fluxes (fast)");
  });
}

int launchSourcesFast(const grim_params& p, const StepArgs& a, int recon, const Box& b, const double* prim,
                      const double* primOld, const double* geomC, const double* gamma40, const double* uT,
                      double* sExp, double* sImp, double* consOld, hipStream_t s) {
  GRIM_REQUIRE(planeOffsetsFit(a.D.n), "fast explicit stage: planes must hold fewer than 2^29 zones");
  const unsigned nb = blocks(b.count(), kBlock);
  return dispatchLayout(p, [&]<class L>() -> int {
    switch (recon) {
      case GRIM_RECON_WENO5:
        sourcesFastKernel<L, GRIM_RECON_WENO5><<<nb, kBlock, 0, s>>>(a, b, prim, primOld, geomC, gamma40, uT, sExp,
                                                                     sImp, consOld);
        break;
      case GRIM_RECON_PPM:
        sourcesFastKernel<L, GRIM_RECON_PPM><<<nb, kBlock, 0, s>>>(a, b, prim, primOld, geomC, gamma40, uT, sExp,
                                                                   sImp, consOld);
        break;
      default:
        sourcesFastKernel<L, GRIM_RECON_MINMOD><<<nb, kBlock, 0, s>>>(a, b, prim, primOld, geomC, gamma40, uT,
                                                                      sExp, sImp, consOld);
    }
    return checkLaunch("sources (fast)");
  });
}

void launchGamma40(int64_t n, const double* gamma64, double* gamma40, hipStream_t s) {
  gamma40Kernel<<<blocks(n, kBlock), kBlock, 0, s>>>(n, gamma64, gamma40);
}

}  // namespace grim
