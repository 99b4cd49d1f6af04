// Device code of the explicit part of grim's time step: geometry of a grid
// location, reconstruction of one zone's face values, the flux of one
// element along a direction, characteristic speeds, and the stress tensor.
//
// Every expression follows the reference term by term (file:line cited at
// each routine) so that, built with -ffp-contract=off, results agree with a
// CPU FP64 evaluation of the same expressions to the last bit wherever the
// operations are correctly rounded (+, -, *, /, sqrt).  The geometry
// routines call transcendental functions and agree to a few ulp.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "grim_hip.h"
#include "zone_physics.hpp"

namespace grim {

// std::min / std::max semantics ((b < a) ? b : a), including which operand
// is returned on ties.
__host__ __device__ __forceinline__ double stdmin(double a, double b) { return (b < a) ? b : a; }
__host__ __device__ __forceinline__ double stdmax(double a, double b) { return (a < b) ? b : a; }

// ---------------------------------------------------------------------------
// Grid dimensions (grid.cpp:66-150 for one rank)
// ---------------------------------------------------------------------------
struct Dims {
  int N[3], g[3], NT[3], dim;
  int64_t n;
  int64_t s[3];  // plane strides of X1, X2, X3
  int64_t ng;    // values per geometry plane: one per (X1, X2) column
};

// Geometry planes (metric, connection, floors) hold one value per (X1, X2)
// column: the metrics this build supports (modified Kerr-Schild, Minkowski)
// do not depend on X3.  Zone (i, j, k) reads them at i + NT1 j, so a plane
// holds D.ng = NT1 NT2 values (the zone count itself below three dimensions).
__host__ __device__ __forceinline__ int64_t geomIndex(const Dims& D, int i, int j) {
  return i + (int64_t)D.NT[0] * j;
}
// The same from a flat zone index (32-bit: grids hold fewer than 2^31 zones).
__host__ __device__ __forceinline__ int64_t geomIndexOf(const Dims& D, int64_t z) {
  return (int64_t)((uint32_t)z % (uint32_t)D.ng);
}

inline Dims makeDims(const grim_grid_config& c) {
  Dims D;
  D.dim = c.dim;
  D.N[0] = c.N1;
  D.N[1] = c.N2;
  D.N[2] = c.N3;
  for (int d = 0; d < 3; d++) {
    D.g[d] = (d < c.dim) ? c.numGhost : 0;
    D.NT[d] = D.N[d] + 2 * D.g[d];
  }
  D.n = (int64_t)D.NT[0] * D.NT[1] * D.NT[2];
  D.s[0] = 1;
  D.s[1] = D.NT[0];
  D.s[2] = (int64_t)D.NT[0] * D.NT[1];
  D.ng = (int64_t)D.NT[0] * D.NT[1];
  return D;
}

// A box of zone indices [lo, hi) in each dimension, walked X1-fastest by a
// flat thread index.
struct Box {
  int lo[3], hi[3];
  __host__ __device__ int64_t count() const {
    return (int64_t)(hi[0] - lo[0]) * (hi[1] - lo[1]) * (hi[2] - lo[2]);
  }
  // 32-bit index arithmetic: grim_create admits grids of fewer than 2^31
  // zones, so every box and index fits (64-bit division is a long software
  // sequence on the GPU, 32-bit a few instructions)
  __device__ __forceinline__ void at(int64_t t, int& i, int& j, int& k) const {
    const unsigned c0 = (unsigned)(hi[0] - lo[0]), c1 = (unsigned)(hi[1] - lo[1]);
    const unsigned u = (unsigned)t;
    i = lo[0] + (int)(u % c0);
    const unsigned r = u / c0;
    j = lo[1] + (int)(r % c1);
    k = lo[2] + (int)(r / c1);
  }
};

// ---------------------------------------------------------------------------
// Geometry of one location (23 planes, see include/grim_hip.h)
// ---------------------------------------------------------------------------
enum { LOC_R = 22, LOC_THETA = 23 };

struct GeomFull {
  ZoneGeom zg;        // gCov, gCon row 0, alpha, g
  double gconS[6];    // gCon 11 12 13 22 23 33
  double r;
  __device__ __forceinline__ double con(int mu, int nu) const {
    if (mu == 0) return zg.gcon0[nu];
    if (nu == 0) return zg.gcon0[mu];
    const int a = (mu < nu ? mu : nu) - 1, b = (mu < nu ? nu : mu) - 1;
    return gconS[a * 3 - (a * (a + 1)) / 2 + b];
  }
};

__device__ __forceinline__ void loadGeomFull(const double* __restrict__ gl, int64_t n, int64_t z,
                                             GeomFull& G) {
  loadGeom(gl, n, z, G.zg);
#pragma unroll
  for (int q = 0; q < 6; q++) G.gconS[q] = gl[(16 + q) * n + z];
  G.r = gl[LOC_R * n + z];
}

constexpr int kBlock = 256;

// Everything a step kernel reads besides its arrays.
struct StepArgs {
  Dims D;
  DevParams prm;
  double dX[3];
  double theta;
  int riemann;
  int metric;
  int fluxFilter;  // torus: applyProblemSpecificFluxFilter
};

__device__ __forceinline__ int64_t flatIndex(const Dims& D, int i, int j, int k) {
  return i + (int64_t)D.NT[0] * (j + (int64_t)D.NT[1] * k);
}

// ---------------------------------------------------------------------------
// Element quantities beyond the kinematics: sound speed, |b|, EMHD closure
// (fluidElement::set, physics.cpp:53-183).
// ---------------------------------------------------------------------------
template <class L>
__device__ __forceinline__ void elementClosure(const double* P, const Kinematics& k,
                                               const DevParams& prm, double radius, Closure& c) {
  if constexpr (L::EMHD) {
    closure<L>(P, k, prm, 1., 1., radius, c);
  } else {
    c.soundSpeed = sqrt(prm.gam * k.pressure / (k.rho + prm.gam * k.u));
    c.bNorm = sqrt(k.bSqr);
    c.q = 0.;
    c.deltaP = 0.;
    c.tau = c.chi = c.nu = 0.;
  }
}

// Row MU of T^mu_nu (physics.cpp:153-178), in the reference's form with the
// Kronecker delta multiplied in.
template <class L, int MU>
__device__ __forceinline__ void stressRow(const Kinematics& k, const Closure& c, double* T) {
  const double w = k.rho + k.u + k.pressure + k.bSqr;
#pragma unroll
  for (int nu = 0; nu < 4; nu++) {
    const double delta = (MU == nu) ? 1. : 0.;
    double t = w * k.uCon[MU] * k.uCov[nu] + (k.pressure + 0.5 * k.bSqr) * delta -
               k.bCon[MU] * k.bCov[nu];
    if constexpr (L::COND_ON)
      t += c.q / c.bNorm * (k.uCon[MU] * k.bCov[nu] + k.bCon[MU] * k.uCov[nu]);
    if constexpr (L::VISC_ON)
      t += (-c.deltaP) * (k.bCon[MU] * k.bCov[nu] / k.bSqr -
                          (1. / 3.) * (delta + k.uCon[MU] * k.uCov[nu]));
    T[nu] = t;
  }
}

// fluidElement::computeFluxes(DIR) (physics.cpp:185-217); DIR = 0 gives U.
template <class L, int DIR>
__device__ __forceinline__ void fluxes(const double* P, const Kinematics& k, const Closure& c,
                                       double g, double* F) {
  double T[4];
  stressRow<L, DIR>(k, c, T);
  F[RHO] = g * (k.rho * k.uCon[DIR]);
  F[UU] = g * T[0] + F[RHO];
  F[U1] = g * T[1];
  F[U2] = g * T[2];
  F[U3] = g * T[3];
  F[L::B1] = g * (k.bCon[1] * k.uCon[DIR] - k.bCon[DIR] * k.uCon[1]);
  F[L::B2] = g * (k.bCon[2] * k.uCon[DIR] - k.bCon[DIR] * k.uCon[2]);
  F[L::B3] = g * (k.bCon[3] * k.uCon[DIR] - k.bCon[DIR] * k.uCon[3]);
  if constexpr (L::COND_ON) F[L::Q] = g * (k.uCon[DIR] * P[L::Q]);
  if constexpr (L::VISC_ON) F[L::DP] = g * (k.uCon[DIR] * P[L::DP]);
}

// fluidElement::computeMinMaxCharSpeeds (riemannsolver.cpp:54-200), DIR = 0,
// 1, 2 for X1, X2, X3.  The contractions with the unit covectors are written
// out as the reference's loops (products with zero included) so that signed
// zeros round identically.
template <class L, int DIR>
__device__ __forceinline__ void charSpeeds(const Kinematics& k, const Closure& c,
                                           const GeomFull& G, const DevParams& prm, double& minSpeed,
                                           double& maxSpeed) {
  const double cAlvenSqr = k.bSqr / (k.rho + prm.gam * k.u + k.bSqr);
  double csSqr = c.soundSpeed * c.soundSpeed;
  double cVisSqr = 0.;
  if constexpr (L::VISC_ON) cVisSqr = 4. / 3. / (k.rho + prm.gam * k.u) * k.rho * c.nu / c.tau;
  double cConSqr = 0.;
  if constexpr (L::COND_ON) cConSqr = (prm.gam - 1.) * c.chi / c.tau;
  cConSqr = 0.5 * (csSqr + cConSqr + sqrt(csSqr * csSqr + cConSqr * cConSqr));
  csSqr = cConSqr + cVisSqr;
  csSqr = csSqr + cAlvenSqr - csSqr * cAlvenSqr;
  {
    const double condition = csSqr > 1. ? 1. : 0.;
    csSqr = csSqr * (1. - condition) + condition;
  }
  constexpr int sdir = DIR + 1;
  double ACov[4], BCov[4], ACon[4], BCon[4];
#pragma unroll
  for (int mu = 0; mu < 4; mu++) {
    ACov[mu] = (mu == sdir) ? 1. : 0.;
    BCov[mu] = (mu == 0) ? 1. : 0.;
  }
#pragma unroll
  for (int mu = 0; mu < 4; mu++) {
    ACon[mu] = 0.;
    BCon[mu] = 0.;
#pragma unroll
    for (int nu = 0; nu < 4; nu++) {
      ACon[mu] += G.con(mu, nu) * ACov[nu];
      BCon[mu] += G.con(mu, nu) * BCov[nu];
    }
  }
  double ASqr = 0., BSqr = 0., ADotU = 0., BDotU = 0., ADotB = 0.;
#pragma unroll
  for (int mu = 0; mu < 4; mu++) {
    ASqr += ACov[mu] * ACon[mu];
    BSqr += BCov[mu] * BCon[mu];
    ADotU += ACov[mu] * k.uCon[mu];
    BDotU += BCov[mu] * k.uCon[mu];
    ADotB += ACov[mu] * BCon[mu];
  }
  const double A = (BDotU * BDotU) - (BSqr + BDotU * BDotU) * csSqr;
  const double B = 2. * (ADotU * BDotU - (ADotB + ADotU * BDotU) * csSqr);
  const double C = ADotU * ADotU - (ASqr + ADotU * ADotU) * csSqr;
  const double discr = sqrt(smax(B * B - 4. * A * C, 1.e-16));
  double mn = -(-B + discr) / 2. / A;
  double mx = -(-B - discr) / 2. / A;
  {
    const double condition = mn > -1.e-15 ? 1. : 0.;
    mn = mn * (1 - condition) - 1.e-15 * condition;
  }
  {
    const double condition = mx < 1.e-15 ? 1. : 0.;
    mx = mx * (1 - condition) + 1.e-15 * condition;
  }
  minSpeed = mn;
  maxSpeed = mx;
}

// ---------------------------------------------------------------------------
// Reconstruction of one zone from y[0..4] = in[i-2 .. i+2]
// ---------------------------------------------------------------------------
// reconstruction::minmod, minmod.cpp:3-19 (af::sign(x) = 1 for x < 0, else 0)
__device__ __forceinline__ double minmod(double x, double y, double z) {
  const double minOfAll = stdmin(stdmin(fabs(x), fabs(y)), fabs(z));
  const double signx = 1. - 2. * (x < 0 ? 1. : 0.);
  const double signy = 1. - 2. * (y < 0 ? 1. : 0.);
  const double signz = 1. - 2. * (z < 0 ? 1. : 0.);
  return 0.25 * fabs(signx + signy) * (signx + signz) * minOfAll;
}

// slopeMM, minmod.cpp:21-71
__device__ __forceinline__ double slopeMM(double theta, double dX, double ym, double y, double yp) {
  const double forwardDiff = (yp - y) / dX;
  const double backwardDiff = (y - ym) / dX;
  const double centralDiff = backwardDiff + forwardDiff;
  const double left = theta * backwardDiff;
  const double center = 0.5 * centralDiff;
  const double right = theta * forwardDiff;
  return minmod(left, center, right);
}

// reconstructWENO5, weno5.cpp:111-257
__device__ __forceinline__ void weno5(const double* y, double& left, double& right) {
  const double eps2 = 1.0e-17;
  const double y0 = y[0], y1 = y[1], y2 = y[2], y3 = y[3], y4 = y[4];
  const double beta1 = ((4.0 / 3.0) * y0 * y0 - (19.0 / 3.0) * y0 * y1 + (25.0 / 3.0) * y1 * y1 +
                        (11.0 / 3.0) * y0 * y2 - (31.0 / 3.0) * y1 * y2 + (10.0 / 3.0) * y2 * y2) +
                       eps2 * (1.0 + fabs(y0) + fabs(y1) + fabs(y2));
  const double beta2 = ((4.0 / 3.0) * y1 * y1 - (13.0 / 3.0) * y1 * y2 + (13.0 / 3.0) * y2 * y2 +
                        (5.0 / 3.0) * y1 * y3 - (13.0 / 3.0) * y2 * y3 + (4.0 / 3.0) * y3 * y3) +
                       eps2 * (1.0 + fabs(y1) + fabs(y2) + fabs(y3));
  const double beta3 = ((10.0 / 3.0) * y2 * y2 - (31.0 / 3.0) * y2 * y3 + (25.0 / 3.0) * y3 * y3 +
                        (11.0 / 3.0) * y2 * y4 - (19.0 / 3.0) * y3 * y4 + (4.0 / 3.0) * y4 * y4) +
                       eps2 * (1.0 + fabs(y2) + fabs(y3) + fabs(y4));
  const double w1r = 1.0 / (16.0 * beta1 * beta1);
  const double w2r = 5.0 / (8.0 * beta2 * beta2);
  const double w3r = 5.0 / (16.0 * beta3 * beta3);
  const double w1l = 5.0 / (16.0 * beta1 * beta1);
  const double w2l = 5.0 / (8.0 * beta2 * beta2);
  const double w3l = 1.0 / (16.0 * beta3 * beta3);
  const double denl = w1l + w2l + w3l;
  const double denr = w1r + w2r + w3r;
  const double u1r = 0.375 * y0 - 1.25 * y1 + 1.875 * y2;
  const double u2r = -0.125 * y1 + 0.75 * y2 + 0.375 * y3;
  const double u3r = 0.375 * y2 + 0.75 * y3 - 0.125 * y4;
  const double u1l = -0.125 * y0 + 0.75 * y1 + 0.375 * y2;
  const double u2l = 0.375 * y1 + 0.75 * y2 - 0.125 * y3;
  const double u3l = 1.875 * y2 - 1.25 * y3 + 0.375 * y4;
  left = (w1l * u1l + w2l * u2l + w3l * u3l) / denl;
  right = (w1r * u1r + w2r * u2r + w3r * u3r) / denr;
}

// reconstructPPM / slopePPM, ppm.cpp:3-164 (face values)
__device__ __forceinline__ void ppm(const double* y, double& leftOut, double& rightOut) {
  const double y0 = y[0], y1 = y[1], y2 = y[2], y3 = y[3], y4 = y[4];
  const double d0 = 2. * (y1 - y0);
  const double d1 = 2. * (y2 - y1);
  const double d2 = 2. * (y3 - y2);
  const double d3 = 2. * (y4 - y3);
  const double D1 = 0.5 * (y2 - y0);
  const double D2 = 0.5 * (y3 - y1);
  const double D3 = 0.5 * (y4 - y2);
  const double condZeroSlope1 = (d1 * d0 <= 0.) ? 1. : 0.;
  const double sign1 = ((D1 > 0.) ? 1. : 0.) * 2. - 1.;
  const double DQ1 = (1. - condZeroSlope1) * sign1 * stdmin(fabs(D1), stdmin(fabs(d0), fabs(d1)));
  const double condZeroSlope2 = (d2 * d1 <= 0.) ? 1. : 0.;
  const double sign2 = ((D2 > 0.) ? 1. : 0.) * 2. - 1.;
  const double DQ2 = (1. - condZeroSlope2) * sign2 * stdmin(fabs(D2), stdmin(fabs(d1), fabs(d2)));
  const double condZeroSlope3 = (d3 * d2 <= 0.) ? 1. : 0.;
  const double sign3 = ((D3 > 0.) ? 1. : 0.) * 2. - 1.;
  const double DQ3 = (1. - condZeroSlope3) * sign3 * stdmin(fabs(D3), stdmin(fabs(d2), fabs(d3)));
  double leftV = 0.5 * (y2 + y1) - 1. / 6. * (DQ2 - DQ1);
  double rightV = 0.5 * (y3 + y2) - 1. / 6. * (DQ3 - DQ2);
  const double corr1 = ((rightV - y2) * (y2 - leftV) <= 0.) ? 1. : 0.;
  const double qd = rightV - leftV;
  const double qe = 6. * (y2 - 0.5 * (rightV + leftV));
  const double corr2 = (qd * (qd - qe) < 0.) ? 1. : 0.;
  const double corr3 = (qd * (qd + qe) < 0.) ? 1. : 0.;
  leftV = leftV * (1. - corr1) + corr1 * y2;
  rightV = rightV * (1. - corr1) + corr1 * y2;
  leftV = leftV * (1. - corr2) + corr2 * (3. * y2 - 2. * rightV);
  rightV = rightV * corr2 + (1. - corr2) * rightV * (1. - corr3) +
           (1. - corr2) * corr3 * (3. * y2 - 2. * leftV);
  leftOut = leftV;
  rightOut = rightV;
}

// reconstruction::reconstruct (reconstruction.cpp:3-38): face values of one
// zone; MINMOD uses dX = 1 (minmod.cpp:84-91).
template <int RECON>
__device__ __forceinline__ void reconstructZone(const double* y, double theta, double& left,
                                                double& right) {
  if constexpr (RECON == GRIM_RECON_WENO5) {
    weno5(y, left, right);
  } else if constexpr (RECON == GRIM_RECON_PPM) {
    ppm(y, left, right);
  } else {
    const double slope = slopeMM(theta, 1., y[1], y[2], y[3]);
    left = y[2] - 0.5 * slope;
    right = y[2] + 0.5 * slope;
  }
}

// reconstruction::slope (reconstruction.cpp:40-64)
template <int RECON>
__device__ __forceinline__ double slopeZone(const double* y, double theta, double dX) {
  if constexpr (RECON == GRIM_RECON_WENO5) {
    return (-y[4] + 8. * y[3] - 8. * y[1] + y[0]) / 12. / dX;  // weno5.cpp:52-65
  } else if constexpr (RECON == GRIM_RECON_PPM) {
    double l, r;
    ppm(y, l, r);
    return (r - l) / dX;
  } else {
    return slopeMM(theta, dX, y[1], y[2], y[3]);
  }
}

// ---------------------------------------------------------------------------
// Coordinates and metric (CoordinateChangeFunctionsArray.hpp:1-195,
// geometry.cpp:255-424, 510-629), one point at a time.
// ---------------------------------------------------------------------------
struct CoordMap {
  int metric, derefine, cylindrify;
  double a, hSlope, X1Start, X1cyl, X2cyl;

  __host__ __device__ double GammieRadius(double X1) const { return exp(X1); }
  __host__ __device__ double GammieTheta(double X2) const {
    return M_PI * X2 + 0.5 * (1 - hSlope) * sin(2. * M_PI * X2);
  }
  __host__ __device__ double ThetaNoCyl(double X1, double X2) const {
    if (derefine) {
      const double cTh0 = cos(M_PI * X2);
      const double cTh1 = cos(GammieTheta(X2));
      const double cTh = cTh1 + exp(X1Start) / exp(X1) * (cTh0 - cTh1);
      return acos(cTh);
    }
    return GammieTheta(X2);
  }
  __host__ __device__ static double Fangle(double x) {
    const double condition1 = (x <= -1) ? 1. : 0.;
    const double condition2 = (x >= 1) ? 1. : 0.;
    return condition2 * x +
           (condition1 - 1) * (condition2 - 1) *
               (x + 1. +
                (-140 * sin(((x + 1) * M_PI) / 2.) + (10 * sin(((x + 1) * 3 * M_PI) / 2.)) / 3. +
                 (2 * sin(((x + 1) * 5 * M_PI) / 2.)) / 5.) /
                    (64. * M_PI)) /
               2.;
  }
  __host__ __device__ static double limlin(double x, double x0, double dx, double y0) {
    return y0 - dx * Fangle((x - x0) * (-1.) / dx);
  }
  __host__ __device__ static double mins(double f1, double f2, double df) { return limlin(f1, f2, df, f2); }
  __host__ __device__ static double maxs(double f1, double f2, double df) { return (mins(-f1, -f2, df) * (-1.)); }
  __host__ __device__ static double minmaxs(double f1, double f2, double df, double dir) {
    const double condition = (dir >= 0) ? 1. : 0.;
    return condition * maxs(f1, f2, df) + (1 - condition) * mins(f1, f2, df);
  }
  __host__ __device__ double sinth1in(const double* X0, const double* X) const {
    return GammieRadius(X0[0]) * sin(ThetaNoCyl(X0[0], X[1])) / GammieRadius(X[0]);
  }
  __host__ __device__ double th2in(const double* X0, const double* X) const {
    const double th0 =
        asin(GammieRadius(X0[0]) * sin(ThetaNoCyl(X0[0], X0[1])) / GammieRadius(X[0]));
    return (ThetaNoCyl(X[0], X[1]) - ThetaNoCyl(X[0], X0[1])) /
               (ThetaNoCyl(X[0], 0.5) - ThetaNoCyl(X[0], X0[1])) * (ThetaNoCyl(X[0], 0.5) - th0) +
           th0;
  }
  __host__ __device__ double func1(const double* X) const { return sin(ThetaNoCyl(X[0], X[1])); }
  __host__ __device__ double func2(const double* X0, const double* X) const {
    const double Xca[3] = {X[0], 0., X[2]};
    const double sth1in = sinth1in(X0, X);
    const double sth2in = sin(th2in(X0, X));
    const double sth1inaxis = sinth1in(X0, Xca);
    const double sth2inaxis = sin(th2in(X0, Xca));
    return minmaxs(sth1in, sth2in, fabs(sth2inaxis - sth1inaxis) + 1.e-20, X[0] - X0[0]);
  }

  __host__ __device__ void toX(const double* X, double* x) const {
    if (metric == GRIM_METRIC_MINKOWSKI) {
      x[0] = X[0];
      x[1] = X[1];
      x[2] = X[2];
      return;
    }
    x[0] = GammieRadius(X[0]);
    x[1] = ThetaNoCyl(X[0], X[1]);
    x[2] = X[2];
    if (!cylindrify) return;
    double Xgrid[3] = {X[0], X[1], X[2]};
    double xGam[3] = {x[0], x[1], x[2]};
    const double IsMirrored = (Xgrid[1] > 0.5) ? 1. : 0.;
    Xgrid[1] = (1 - IsMirrored) * Xgrid[1] + IsMirrored * (1. - Xgrid[1]);
    xGam[1] = (1 - IsMirrored) * xGam[1] + IsMirrored * (M_PI - xGam[1]);
    const double IsGhost = (Xgrid[1] < 0.) ? 1. : 0.;
    Xgrid[1] = (1 - IsGhost) * Xgrid[1] + IsGhost * (-(1.0) * Xgrid[1]);
    xGam[1] = (1 - IsGhost) * xGam[1] + IsGhost * (-(1.0) * xGam[1]);
    const double X0[3] = {X1cyl, X2cyl, 0.};
    const double R_cyl = GammieRadius(X0[0]);
    const double X1_tr = log(0.5 * (R_cyl + exp(X1Start)));
    const double Xtr[3] = {X1_tr, Xgrid[1], Xgrid[2]};
    const double f1 = func1(Xgrid);
    const double f2 = func2(X0, Xgrid);
    const double dftr = func2(X0, Xtr) - func1(Xtr);
    const double sinth =
        maxs(xGam[0] * f1, xGam[0] * f2, GammieRadius(Xtr[0]) * fabs(dftr) + 1.e-20) / xGam[0];
    double th = asin(sinth);
    th = th * (1 - IsGhost * 2);
    x[1] = IsMirrored * (M_PI - th) + (1 - IsMirrored) * th;
  }

  __host__ __device__ void gCov(const double* X, double G[4][4]) const {
#pragma unroll
    for (int m = 0; m < 4; m++)
#pragma unroll
      for (int k = 0; k < 4; k++) G[m][k] = 0.;
    if (metric == GRIM_METRIC_MINKOWSKI) {
      G[0][0] = -1.;
      G[1][1] = 1.;
      G[2][2] = 1.;
      G[3][3] = 1.;
      return;
    }
    const double GAMMA_EPS = 1.e-5;
    double x[3];
    toX(X, x);
    const double r = x[0], theta = x[1];
    double dr_dX1 = r, dr_dX2 = 0.;
    double dtheta_dX2 = M_PI + M_PI * (1 - hSlope) * cos(2 * M_PI * X[1]);
    double dtheta_dX1 = 0.;
    if (cylindrify) {
      double xp[3], xm[3];
      const double Xp[3] = {X[0] + GAMMA_EPS, X[1], X[2]}, Xm[3] = {X[0] - GAMMA_EPS, X[1], X[2]};
      toX(Xp, xp);
      toX(Xm, xm);
      dr_dX1 = (xp[0] - xm[0]) / (2. * GAMMA_EPS);
      dtheta_dX1 = (xp[1] - xm[1]) / (2. * GAMMA_EPS);
      const double Yp[3] = {X[0], X[1] + GAMMA_EPS, X[2]}, Ym[3] = {X[0], X[1] - GAMMA_EPS, X[2]};
      toX(Yp, xp);
      toX(Ym, xm);
      dr_dX2 = (xp[0] - xm[0]) / (2. * GAMMA_EPS);
      dtheta_dX2 = (xp[1] - xm[1]) / (2. * GAMMA_EPS);
    }
    const double sigma = r * r + pow(a * cos(theta), 2.);
    const double s2 = pow(sin(theta), 2.);
    G[0][0] = -(1. - 2. * r / sigma);
    G[0][1] = (2. * r / sigma) * dr_dX1;
    G[0][2] = (2. * r / sigma) * dr_dX2;
    G[0][3] = -(2. * a * r * s2 / sigma);
    G[1][0] = G[0][1];
    G[1][1] = (1. + 2 * r / sigma) * dr_dX1 * dr_dX1 + (sigma)*dtheta_dX1 * dtheta_dX1;
    G[1][2] = (1. + 2 * r / sigma) * dr_dX1 * dr_dX2 + (sigma)*dtheta_dX1 * dtheta_dX2;
    G[1][3] = -a * (1. + 2. * r / sigma) * s2 * dr_dX1;
    G[2][0] = G[0][2];
    G[2][1] = G[1][2];
    G[2][2] = sigma * dtheta_dX2 * dtheta_dX2 + (1. + 2 * r / sigma) * dr_dX2 * dr_dX2;
    G[2][3] = -a * (1. + 2. * r / sigma) * s2 * dr_dX2;
    G[3][0] = G[0][3];
    G[3][1] = G[1][3];
    G[3][2] = G[2][3];
    G[3][3] = (s2 * (sigma + pow(a * sin(theta), 2.) * (1. + 2 * r / sigma)));
  }
};

// geometry::setgDetAndgConFromgCov (geometry.cpp:123-253): determinant by the
// Leibniz expansion, inverse by cofactors.
__host__ __device__ __forceinline__ void gDetAndgCon(const double G[4][4], double& det, double C[4][4]) {
  const double d =
      G[0][0] * G[1][1] * G[2][2] * G[3][3] + G[0][0] * G[1][2] * G[2][3] * G[3][1] +
      G[0][0] * G[1][3] * G[2][1] * G[3][2] + G[0][1] * G[1][0] * G[2][3] * G[3][2] +
      G[0][1] * G[1][2] * G[2][0] * G[3][3] + G[0][1] * G[1][3] * G[2][2] * G[3][0] +
      G[0][2] * G[1][0] * G[2][1] * G[3][3] + G[0][2] * G[1][1] * G[2][3] * G[3][0] +
      G[0][2] * G[1][3] * G[2][0] * G[3][1] + G[0][3] * G[1][0] * G[2][2] * G[3][1] +
      G[0][3] * G[1][1] * G[2][0] * G[3][2] + G[0][3] * G[1][2] * G[2][1] * G[3][0] -
      G[0][0] * G[1][1] * G[2][3] * G[3][2] - G[0][0] * G[1][2] * G[2][1] * G[3][3] -
      G[0][0] * G[1][3] * G[2][2] * G[3][1] - G[0][1] * G[1][0] * G[2][2] * G[3][3] -
      G[0][1] * G[1][2] * G[2][3] * G[3][0] - G[0][1] * G[1][3] * G[2][0] * G[3][2] -
      G[0][2] * G[1][0] * G[2][3] * G[3][1] - G[0][2] * G[1][1] * G[2][0] * G[3][3] -
      G[0][2] * G[1][3] * G[2][1] * G[3][0] - G[0][3] * G[1][0] * G[2][1] * G[3][2] -
      G[0][3] * G[1][1] * G[2][2] * G[3][0] - G[0][3] * G[1][2] * G[2][0] * G[3][1];
  det = d;
#pragma unroll
  for (int i = 0; i < 4; i++) {
#pragma unroll
    for (int j = 0; j < 4; j++) {
      int ri[3], cj[3];
      for (int k = 0, m = 0; k < 4; k++)
        if (k != j) ri[m++] = k;
      for (int k = 0, m = 0; k < 4; k++)
        if (k != i) cj[m++] = k;
      const double mnr =
          G[ri[0]][cj[0]] * (G[ri[1]][cj[1]] * G[ri[2]][cj[2]] - G[ri[1]][cj[2]] * G[ri[2]][cj[1]]) -
          G[ri[0]][cj[1]] * (G[ri[1]][cj[0]] * G[ri[2]][cj[2]] - G[ri[1]][cj[2]] * G[ri[2]][cj[0]]) +
          G[ri[0]][cj[2]] * (G[ri[1]][cj[0]] * G[ri[2]][cj[1]] - G[ri[1]][cj[1]] * G[ri[2]][cj[0]]);
      C[i][j] = (((i + j) & 1) ? -mnr : mnr) / d;
    }
  }
}

}  // namespace grim
