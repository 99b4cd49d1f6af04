// grim oracle: the full time step (TEST INFRASTRUCTURE ONLY; see grim_oracle.h).
//
// CPU FP64 restatement of grim's explicit half of the step and of the
// control flow around the per-zone Newton solve in grim_oracle.hpp:
//   coordinates        src/grid/grid.cpp:163-249
//   geometry           src/geometry/geometry.cpp:4-121, 255-424, 510-680,
//                      src/geometry/CoordinateChangeFunctionsArray.hpp:1-195
//   boundaries         src/boundary/boundary.cpp:5-243, src/grid/grid.cpp:306-315
//   reconstruction     src/reconstruction/{reconstruction,minmod,weno5,ppm}.cpp
//   Riemann solver     src/physics/riemannsolver.cpp:54-328
//   divergence, CT     src/timestepper/fvmfluxes.cpp:3-184, constrainedtransport.cpp:3-83
//   sources            src/physics/physics.cpp:360-586
//   time step, dt      src/timestepper/timestep.cpp:3-400
// grim evaluates these on whole arrays with af::shift (circular); here every
// value is computed per zone with the same expression tree, for the zones
// whose stencil lies inside the grid (the only ones that reach the bulk).
//
// Deviations, each without effect on bulk values:
//  * ghost zones are not solved for (grim solves their zero systems and
//    leaves non-finite values that the boundary conditions overwrite);
//  * the frozen residual terms are folded into one right-hand side,
//    rhs = consOld - dt'((divF + S_expl) + 1/2 S_implOld), so that
//    R = (U(P) - rhs)/dt' + 1/2 S_impl(P) + S_td(P) (residual.cpp:39-45
//    reassociated; differences are rounding-level);
//  * the first dt of a run is computed from the initial state (grim's
//    constructor sets elemOld before the initial conditions).
#include <omp.h>

#include <cmath>
#include <vector>

#include "grim_oracle.h"
#include "grim_oracle.hpp"

namespace oracle {
namespace {

struct Dims {
  int N[3], g[3], NT[3], dim, ng;
  int64_t n;
  explicit Dims(const grim_oracle_grid& c) {
    dim = c.dim;
    ng = c.numGhost;
    N[0] = c.N1;
    N[1] = c.N2;
    N[2] = c.N3;
    for (int d = 0; d < 3; d++) {
      g[d] = (d < dim) ? ng : 0;
      NT[d] = N[d] + 2 * g[d];
    }
    n = (int64_t)NT[0] * NT[1] * NT[2];
  }
  int64_t idx(int i, int j, int k) const { return i + (int64_t)NT[0] * (j + (int64_t)NT[1] * k); }
  int64_t stride(int d) const { return d == 0 ? 1 : (d == 1 ? NT[0] : (int64_t)NT[0] * NT[1]); }
  bool bulk(int i, int j, int k) const {
    return i >= g[0] && i < g[0] + N[0] && j >= g[1] && j < g[1] + N[1] && k >= g[2] &&
           k < g[2] + N[2];
  }
};

// ---------------------------------------------------------------------------
// Coordinates and geometry
// ---------------------------------------------------------------------------
enum { LOC_CENTER = 0, LOC_LEFT = 1, LOC_BOTTOM = 2 };

// coordinatesGrid::setXCoords (grid.cpp:168-249) for one index triple.
void XCoordsAt(const grim_oracle_grid& c, const Dims& D, int i, int j, int k, int loc, double X[3]) {
  const double dX1 = (c.X1End - c.X1Start) / c.N1;
  const double dX2 = (c.X2End - c.X2Start) / c.N2;
  const double dX3 = (c.X3End - c.X3Start) / c.N3;
  const double i1 = (double)i - D.g[0], i2 = (double)j - D.g[1], i3 = (double)k - D.g[2];
  X[0] = c.X1Start + (loc == LOC_LEFT ? (i1)*dX1 : (i1 + 0.5) * dX1);
  X[1] = c.X2Start + (loc == LOC_BOTTOM ? (i2)*dX2 : (i2 + 0.5) * dX2);
  X[2] = c.X3Start + (i3 + 0.5) * dX3;
}

// CoordinateChangeFunctionsArray.hpp, one point at a time.
struct Coord {
  const grim_oracle_grid& c;
  double GammieRadius(double X1) const { return std::exp(X1); }
  double GammieTheta(double X2) const {
    return M_PI * X2 + 0.5 * (1 - c.hSlope) * std::sin(2. * M_PI * X2);
  }
  double ThetaNoCyl(double X1, double X2) const {
    if (c.DerefineThetaHorizon) {
      double cTh0 = std::cos(M_PI * X2);
      double cTh1 = std::cos(GammieTheta(X2));
      double cTh = cTh1 + std::exp(c.X1Start) / std::exp(X1) * (cTh0 - cTh1);
      return std::acos(cTh);
    }
    return GammieTheta(X2);
  }
  static double Ftr(double x) {
    double condition1 = (x <= 0) ? 1. : 0.;
    double condition2 = (x >= 1) ? 1. : 0.;
    double res = condition2;
    res = res + (condition1 - 1) * (condition2 - 1) *
                    (std::cos(x * 5 * M_PI) + 70 * std::sin(((2 * x - 1) * M_PI) / 2.) +
                     5 * std::sin(((2 * x - 1) * 3 * M_PI) / 2.) + 64.) /
                    128.;
    return res;
  }
  static double Fangle(double x) {
    double condition1 = (x <= -1) ? 1. : 0.;
    double condition2 = (x >= 1) ? 1. : 0.;
    return condition2 * x +
           (condition1 - 1) * (condition2 - 1) *
               (x + 1. +
                (-140 * std::sin(((x + 1) * M_PI) / 2.) +
                 (10 * std::sin(((x + 1) * 3 * M_PI) / 2.)) / 3. +
                 (2 * std::sin(((x + 1) * 5 * M_PI) / 2.)) / 5.) /
                    (64. * M_PI)) /
               2.;
  }
  static double limlin(double x, double x0, double dx, double y0) {
    return y0 - dx * Fangle((x - x0) * (-1.) / dx);
  }
  static double mins(double f1, double f2, double df) { return limlin(f1, f2, df, f2); }
  static double maxs(double f1, double f2, double df) { return (mins(-f1, -f2, df) * (-1.)); }
  static double minmaxs(double f1, double f2, double df, double dir) {
    double condition = (dir >= 0) ? 1. : 0.;
    return condition * maxs(f1, f2, df) + (1 - condition) * mins(f1, f2, df);
  }
  double sinth0(const double X0[3], const double X[3]) const {
    return GammieRadius(X0[0]) * std::sin(ThetaNoCyl(X0[0], X0[1])) / GammieRadius(X[0]);
  }
  double sinth1in(const double X0[3], const double X[3]) const {
    return GammieRadius(X0[0]) * std::sin(ThetaNoCyl(X0[0], X[1])) / GammieRadius(X[0]);
  }
  double th2in(const double X0[3], const double X[3]) const {
    double th0 = std::asin(sinth0(X0, X));
    return (ThetaNoCyl(X[0], X[1]) - ThetaNoCyl(X[0], X0[1])) /
               (ThetaNoCyl(X[0], 0.5) - ThetaNoCyl(X[0], X0[1])) *
               (ThetaNoCyl(X[0], 0.5) - th0) +
           th0;
  }
  double func1(const double*, const double X[3]) const { return std::sin(ThetaNoCyl(X[0], X[1])); }
  double func2(const double X0[3], const double X[3]) const {
    double Xca[3] = {X[0], 0., X[2]};
    double sth1in = sinth1in(X0, X);
    double sth2in = std::sin(th2in(X0, X));
    double sth1inaxis = sinth1in(X0, Xca);
    double sth2inaxis = std::sin(th2in(X0, Xca));
    return minmaxs(sth1in, sth2in, std::fabs(sth2inaxis - sth1inaxis) + 1.e-20, X[0] - X0[0]);
  }

  // geometry::XCoordsToxCoords, geometry.cpp:510-629
  void XCoordsToxCoords(const double X[3], double x[3]) const {
    if (c.metric == GRIM_ORACLE_METRIC_MINKOWSKI) {
      x[0] = X[0];
      x[1] = X[1];
      x[2] = X[2];
      return;
    }
    x[0] = GammieRadius(X[0]);
    x[1] = ThetaNoCyl(X[0], X[1]);
    x[2] = X[2];
    if (!c.DoCylindrify) return;
    double Xgrid[3] = {X[0], X[1], X[2]};
    double xGam[3] = {x[0], x[1], x[2]};
    double IsMirrored = (Xgrid[1] > 0.5) ? 1. : 0.;
    Xgrid[1] = (1 - IsMirrored) * Xgrid[1] + IsMirrored * (1. - Xgrid[1]);
    xGam[1] = (1 - IsMirrored) * xGam[1] + IsMirrored * (M_PI - xGam[1]);
    double IsGhost = (Xgrid[1] < 0.) ? 1. : 0.;
    Xgrid[1] = (1 - IsGhost) * Xgrid[1] + IsGhost * (-(1.0) * Xgrid[1]);
    xGam[1] = (1 - IsGhost) * xGam[1] + IsGhost * (-(1.0) * xGam[1]);
    double X0[3] = {c.X1cyl, c.X2cyl, 0.};
    double R_cyl = GammieRadius(X0[0]);
    double X1_tr = std::log(0.5 * (R_cyl + std::exp(c.X1Start)));
    double Xtr[3] = {X1_tr, Xgrid[1], Xgrid[2]};
    double f1 = func1(X0, Xgrid);
    double f2 = func2(X0, Xgrid);
    double dftr = func2(X0, Xtr) - func1(X0, Xtr);
    double sinth = maxs(xGam[0] * f1, xGam[0] * f2, GammieRadius(Xtr[0]) * std::fabs(dftr) + 1.e-20) /
                   xGam[0];
    double th = std::asin(sinth);
    th = th * (1 - IsGhost * 2);
    x[0] = xGam[0];
    x[1] = IsMirrored * (M_PI - th) + (1 - IsMirrored) * th;
    x[2] = xGam[2];
  }

  // geometry::setgCovInXCoords, geometry.cpp:255-424
  void gCovAt(const double X[3], double G[4][4]) const {
    for (int m = 0; m < 4; m++)
      for (int k = 0; k < 4; k++) G[m][k] = 0.;
    if (c.metric == GRIM_ORACLE_METRIC_MINKOWSKI) {
      G[0][0] = -1.;
      G[1][1] = 1.;
      G[2][2] = 1.;
      G[3][3] = 1.;
      return;
    }
    const double a = c.blackHoleSpin;
    const double GAMMA_EPS = 1.e-5;
    double x[3];
    XCoordsToxCoords(X, x);
    double r = x[0], theta = x[1];
    double dr_dX1 = r, dr_dX2 = 0.;
    double dtheta_dX2 = M_PI + M_PI * (1 - c.hSlope) * std::cos(2 * M_PI * X[1]);
    double dtheta_dX1 = 0.;
    if (c.DoCylindrify) {
      double Xp[3] = {X[0] + GAMMA_EPS, X[1], X[2]}, Xm[3] = {X[0] - GAMMA_EPS, X[1], X[2]};
      double xp[3], xm[3];
      XCoordsToxCoords(Xp, xp);
      XCoordsToxCoords(Xm, xm);
      dr_dX1 = (xp[0] - xm[0]) / (2. * GAMMA_EPS);
      dtheta_dX1 = (xp[1] - xm[1]) / (2. * GAMMA_EPS);
      double Yp[3] = {X[0], X[1] + GAMMA_EPS, X[2]}, Ym[3] = {X[0], X[1] - GAMMA_EPS, X[2]};
      XCoordsToxCoords(Yp, xp);
      XCoordsToxCoords(Ym, xm);
      dr_dX2 = (xp[0] - xm[0]) / (2. * GAMMA_EPS);
      dtheta_dX2 = (xp[1] - xm[1]) / (2. * GAMMA_EPS);
    }
    double sigma = r * r + std::pow(a * std::cos(theta), 2.);
    double s2 = std::pow(std::sin(theta), 2.);
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
    G[3][3] = (s2 * (sigma + std::pow(a * std::sin(theta), 2.) * (1. + 2 * r / sigma)));
  }
};

const int kCov10[10][2] = {{0, 0}, {0, 1}, {0, 2}, {0, 3}, {1, 1}, {1, 2}, {1, 3}, {2, 2}, {2, 3}, {3, 3}};
const int kCon6[6][2] = {{1, 1}, {1, 2}, {1, 3}, {2, 2}, {2, 3}, {3, 3}};

void storeLocation(const double G[4][4], const double C[4][4], double gDet, double r, double theta,
                   int64_t n, int64_t z, double* out) {
  for (int q = 0; q < 10; q++) out[q * n + z] = G[kCov10[q][0]][kCov10[q][1]];
  for (int q = 0; q < 4; q++) out[(10 + q) * n + z] = C[0][q];
  out[14 * n + z] = 1. / std::sqrt(-C[0][0]);
  out[15 * n + z] = std::sqrt(-gDet);
  for (int q = 0; q < 6; q++) out[(16 + q) * n + z] = C[kCon6[q][0]][kCon6[q][1]];
  out[22 * n + z] = r;
  out[23 * n + z] = theta;
}

template <class T>
void loadLocation(const double* gl, int64_t n, int64_t z, ZoneGeom<T>* G) {
  for (int q = 0; q < 10; q++) {
    const int a = kCov10[q][0], b = kCov10[q][1];
    G->gCov[a][b] = G->gCov[b][a] = T(gl[q * n + z]);
  }
  for (int q = 0; q < 4; q++) G->gCon[0][q] = G->gCon[q][0] = T(gl[(10 + q) * n + z]);
  for (int q = 0; q < 6; q++) {
    const int a = kCon6[q][0], b = kCon6[q][1];
    G->gCon[a][b] = G->gCon[b][a] = T(gl[(16 + q) * n + z]);
  }
  G->alpha = T(gl[14 * n + z]);
  G->g = T(gl[15 * n + z]);
  G->r = T(gl[22 * n + z]);
}

// ---------------------------------------------------------------------------
// Reconstruction (values at the two faces of one zone from its stencil)
// ---------------------------------------------------------------------------
// reconstruction::minmod, minmod.cpp:3-19 (af::sign(x) = 1 for x < 0, else 0)
double minmod(double x, double y, double z) {
  double minOfAll = std::min(std::min(std::fabs(x), std::fabs(y)), std::fabs(z));
  double signx = 1. - 2. * (x < 0 ? 1. : 0.);
  double signy = 1. - 2. * (y < 0 ? 1. : 0.);
  double signz = 1. - 2. * (z < 0 ? 1. : 0.);
  return 0.25 * std::fabs(signx + signy) * (signx + signz) * minOfAll;
}

// slopeMM, minmod.cpp:21-71: y[0..2] = in[i-1], in[i], in[i+1]
double slopeMM(double theta, double dX, double ym, double y, double yp) {
  double forwardDiff = (yp - y) / dX;
  double backwardDiff = (y - ym) / dX;
  double centralDiff = backwardDiff + forwardDiff;
  double left = theta * backwardDiff;
  double center = 0.5 * centralDiff;
  double right = theta * forwardDiff;
  return minmod(left, center, right);
}

// y[0..4] = in[i-2..i+2] (af::shift(+2) .. af::shift(-2))
void weno5(const double* y, double* left, double* right) {
  // reconstructWENO5, weno5.cpp:111-257
  const double eps2 = 1.0e-17;
  const double y0 = y[0], y1 = y[1], y2 = y[2], y3 = y[3], y4 = y[4];
  double beta1 = ((4.0 / 3.0) * y0 * y0 - (19.0 / 3.0) * y0 * y1 + (25.0 / 3.0) * y1 * y1 +
                  (11.0 / 3.0) * y0 * y2 - (31.0 / 3.0) * y1 * y2 + (10.0 / 3.0) * y2 * y2) +
                 eps2 * (1.0 + std::fabs(y0) + std::fabs(y1) + std::fabs(y2));
  double beta2 = ((4.0 / 3.0) * y1 * y1 - (13.0 / 3.0) * y1 * y2 + (13.0 / 3.0) * y2 * y2 +
                  (5.0 / 3.0) * y1 * y3 - (13.0 / 3.0) * y2 * y3 + (4.0 / 3.0) * y3 * y3) +
                 eps2 * (1.0 + std::fabs(y1) + std::fabs(y2) + std::fabs(y3));
  double beta3 = ((10.0 / 3.0) * y2 * y2 - (31.0 / 3.0) * y2 * y3 + (25.0 / 3.0) * y3 * y3 +
                  (11.0 / 3.0) * y2 * y4 - (19.0 / 3.0) * y3 * y4 + (4.0 / 3.0) * y4 * y4) +
                 eps2 * (1.0 + std::fabs(y2) + std::fabs(y3) + std::fabs(y4));
  double w1r = 1.0 / (16.0 * beta1 * beta1);
  double w2r = 5.0 / (8.0 * beta2 * beta2);
  double w3r = 5.0 / (16.0 * beta3 * beta3);
  double w1l = 5.0 / (16.0 * beta1 * beta1);
  double w2l = 5.0 / (8.0 * beta2 * beta2);
  double w3l = 1.0 / (16.0 * beta3 * beta3);
  double denl = w1l + w2l + w3l;
  double denr = w1r + w2r + w3r;
  double u1r = 0.375 * y0 - 1.25 * y1 + 1.875 * y2;
  double u2r = -0.125 * y1 + 0.75 * y2 + 0.375 * y3;
  double u3r = 0.375 * y2 + 0.75 * y3 - 0.125 * y4;
  double u1l = -0.125 * y0 + 0.75 * y1 + 0.375 * y2;
  double u2l = 0.375 * y1 + 0.75 * y2 - 0.125 * y3;
  double u3l = 1.875 * y2 - 1.25 * y3 + 0.375 * y4;
  *left = (w1l * u1l + w2l * u2l + w3l * u3l) / denl;
  *right = (w1r * u1r + w2r * u2r + w3r * u3r) / denr;
}

// reconstructPPM / slopePPM, ppm.cpp:3-164 (face values before the division by dX)
void ppm(const double* y, double* leftOut, double* rightOut) {
  const double y0 = y[0], y1 = y[1], y2 = y[2], y3 = y[3], y4 = y[4];
  double d0 = 2. * (y1 - y0);
  double d1 = 2. * (y2 - y1);
  double d2 = 2. * (y3 - y2);
  double d3 = 2. * (y4 - y3);
  double D1 = 0.5 * (y2 - y0);
  double D2 = 0.5 * (y3 - y1);
  double D3 = 0.5 * (y4 - y2);
  double condZeroSlope1 = (d1 * d0 <= 0.) ? 1. : 0.;
  double sign1 = ((D1 > 0.) ? 1. : 0.) * 2. - 1.;
  double DQ1 = (1. - condZeroSlope1) * sign1 *
               std::min(std::fabs(D1), std::min(std::fabs(d0), std::fabs(d1)));
  double condZeroSlope2 = (d2 * d1 <= 0.) ? 1. : 0.;
  double sign2 = ((D2 > 0.) ? 1. : 0.) * 2. - 1.;
  double DQ2 = (1. - condZeroSlope2) * sign2 *
               std::min(std::fabs(D2), std::min(std::fabs(d1), std::fabs(d2)));
  double condZeroSlope3 = (d3 * d2 <= 0.) ? 1. : 0.;
  double sign3 = ((D3 > 0.) ? 1. : 0.) * 2. - 1.;
  double DQ3 = (1. - condZeroSlope3) * sign3 *
               std::min(std::fabs(D3), std::min(std::fabs(d2), std::fabs(d3)));
  double leftV = 0.5 * (y2 + y1) - 1. / 6. * (DQ2 - DQ1);
  double rightV = 0.5 * (y3 + y2) - 1. / 6. * (DQ3 - DQ2);
  double corr1 = ((rightV - y2) * (y2 - leftV) <= 0.) ? 1. : 0.;
  double qd = rightV - leftV;
  double qe = 6. * (y2 - 0.5 * (rightV + leftV));
  double corr2 = (qd * (qd - qe) < 0.) ? 1. : 0.;
  double corr3 = (qd * (qd + qe) < 0.) ? 1. : 0.;
  leftV = leftV * (1. - corr1) + corr1 * y2;
  rightV = rightV * (1. - corr1) + corr1 * y2;
  leftV = leftV * (1. - corr2) + corr2 * (3. * y2 - 2. * rightV);
  rightV = rightV * corr2 + (1. - corr2) * rightV * (1. - corr3) +
           (1. - corr2) * corr3 * (3. * y2 - 2. * leftV);
  *leftOut = leftV;
  *rightOut = rightV;
}

// Face values of one zone from its 5-point stencil y[0..4] = in[i-2..i+2]
// (reconstruction::reconstruct, reconstruction.cpp:3-38).
void reconstructZone(const grim_oracle_grid& c, const double* y, double* left, double* right) {
  if (c.reconstruction == GRIM_ORACLE_RECON_WENO5) {
    weno5(y, left, right);
  } else if (c.reconstruction == GRIM_ORACLE_RECON_PPM) {
    ppm(y, left, right);
  } else {
    // reconstructMM, minmod.cpp:73-95 (dX = 1)
    double slope = slopeMM(c.slopeLimTheta, 1., y[1], y[2], y[3]);
    *left = y[2] - 0.5 * slope;
    *right = y[2] + 0.5 * slope;
  }
}

// reconstruction::slope (reconstruction.cpp:40-64): y[0..4] = in[i-2..i+2]
double slopeZone(const grim_oracle_grid& c, double dX, const double* y) {
  if (c.reconstruction == GRIM_ORACLE_RECON_WENO5) {
    // slopeWENO5, weno5.cpp:52-65
    return (-y[4] + 8. * y[3] - 8. * y[1] + y[0]) / 12. / dX;
  }
  if (c.reconstruction == GRIM_ORACLE_RECON_PPM) {
    double l, r;
    ppm(y, &l, &r);
    return (r - l) / dX;
  }
  return slopeMM(c.slopeLimTheta, dX, y[1], y[2], y[3]);
}

double dXof(const grim_oracle_grid& c, int d) {
  return d == 0 ? (c.X1End - c.X1Start) / c.N1
                : (d == 1 ? (c.X2End - c.X2Start) / c.N2 : (c.X3End - c.X3Start) / c.N3);
}

void loadPrim(const double* prim, int nv, int64_t n, int64_t z, double* P) {
  for (int v = 0; v < nv; v++) P[v] = prim[v * n + z];
}

// ---------------------------------------------------------------------------
// Per-face flux (reconstruction + riemannSolver::solve for one face)
// ---------------------------------------------------------------------------
// Flux at the lower face (i - 1/2 along d) of zone (i,j,k).
void faceFlux(const grim_oracle_grid& c, const grim_oracle_params& p, const Dims& D,
              const grim_oracle_geometry& gm, const double* prim, int d, int64_t z, double* F) {
  const Vars vars(p);
  const int nv = vars.dof;
  const int64_t n = D.n, s = D.stride(d);
  // left state: primRight of zone z - s; right state: primLeft of zone z.
  double PL[GRIM_ORACLE_MAX_VARS], PR[GRIM_ORACLE_MAX_VARS];
  for (int v = 0; v < nv; v++) {
    const double* pv = prim + v * n;
    double y[6];
    for (int q = 0; q < 6; q++) y[q] = pv[z + (q - 3) * s];
    double l, r;
    reconstructZone(c, y, &l, &r);  // zone z - s: y[0..4]
    PL[v] = r;
    reconstructZone(c, y + 1, &l, &r);  // zone z: y[1..5]
    PR[v] = l;
  }
  const double* gl = (d == 0) ? gm.faceX1 : (d == 1 ? gm.faceX2 : gm.center);
  ZoneGeom<double> G;
  loadLocation(gl, n, z, &G);
  ZoneClosure zc{1., 1.};
  FluidElement<double> eL, eR;
  eL.set(PL, G, p, zc);
  eR.set(PR, G, p, zc);
  double fluxL[GRIM_ORACLE_MAX_VARS], consL[GRIM_ORACLE_MAX_VARS];
  double fluxR[GRIM_ORACLE_MAX_VARS], consR[GRIM_ORACLE_MAX_VARS];
  eL.computeFluxes(d + 1, fluxL, p);
  eL.computeFluxes(0, consL, p);
  eR.computeFluxes(d + 1, fluxR, p);
  eR.computeFluxes(0, consR, p);
  double minL, maxL, minR, maxR;
  eL.computeMinMaxCharSpeeds(d, &minL, &maxL, p);
  eR.computeMinMaxCharSpeeds(d, &minR, &maxR, p);
  const double minSpeed = std::min(minL, minR);
  const double maxSpeed = std::max(maxL, maxR);
  for (int v = 0; v < nv; v++) {
    if (c.riemannSolver == GRIM_ORACLE_RIEMANN_HLL)
      F[v] = (maxSpeed * fluxL[v] - minSpeed * fluxR[v] +
              minSpeed * maxSpeed * (consR[v] - consL[v])) /
             (maxSpeed - minSpeed);
    else
      F[v] = 0.5 * (fluxL[v] + fluxR[v] - std::max(maxSpeed, -minSpeed) * (consR[v] - consL[v]));
  }
}

// Range of face indices along d whose stencil fits (i-3 .. i+2 inside).
void faceRange(const Dims& D, int d, int lo[3], int hi[3]) {
  for (int e = 0; e < 3; e++) {
    lo[e] = 0;
    hi[e] = D.NT[e];
  }
  lo[d] = 3;
  hi[d] = D.NT[d] - 2;
}

template <class F>
void forRange(const int lo[3], const int hi[3], F f) {
  const int64_t c0 = hi[0] - lo[0], c1 = hi[1] - lo[1], c2 = hi[2] - lo[2];
  const int64_t total = c0 * c1 * c2;
#pragma omp parallel for schedule(static)
  for (int64_t t = 0; t < total; t++) {
    int i = lo[0] + (int)(t % c0);
    int j = lo[1] + (int)((t / c0) % c1);
    int k = lo[2] + (int)(t / (c0 * c1));
    f(i, j, k);
  }
}

void riemannAll(const grim_oracle_grid& c, const grim_oracle_params& p, const Dims& D,
                const grim_oracle_geometry& gm, const double* prim, int d, double* flux) {
  const Vars vars(p);
  int lo[3], hi[3];
  faceRange(D, d, lo, hi);
  forRange(lo, hi, [&](int i, int j, int k) {
    const int64_t z = D.idx(i, j, k);
    double F[GRIM_ORACLE_MAX_VARS];
    faceFlux(c, p, D, gm, prim, d, z, F);
    for (int v = 0; v < vars.dof; v++) flux[v * D.n + z] = F[v];
  });
}

// timeStepper::computeDivOfFluxes + fluxCT (fvmfluxes.cpp, constrainedtransport.cpp)
void divFluxes(const grim_oracle_grid& c, const grim_oracle_params& p, const Dims& D,
               const grim_oracle_geometry& gm, const double* prim, double* divF,
               double* fluxesOut, bool torusFilter = false) {
  const Vars vars(p);
  const int nv = vars.dof;
  const int64_t n = D.n;
  std::vector<double> fl((size_t)3 * nv * n, 0.);
  double* F[3] = {fl.data(), fl.data() + (size_t)nv * n, fl.data() + (size_t)2 * nv * n};
  for (int d = 0; d < D.dim; d++) riemannAll(c, p, D, gm, prim, d, F[d]);

  if (D.dim >= 2) {
    // computeEMF + fluxCT.  af::shift(x, 0, 1, 0) reads x[j-1]; (0,-1,0) x[j+1].
    const int64_t s1 = D.stride(0), s2 = D.stride(1), s3 = D.stride(2);
    std::vector<double> emf3(n, 0.), emf1, emf2;
    const double* F1B2 = F[0] + vars.B2 * n;
    const double* F2B1 = F[1] + vars.B1 * n;
    int lo[3] = {1, 1, 0}, hi[3] = {D.NT[0], D.NT[1], D.NT[2]};
    forRange(lo, hi, [&](int i, int j, int k) {
      const int64_t z = D.idx(i, j, k);
      emf3[z] = 0.25 * (F1B2[z] + F1B2[z - s2] - F2B1[z] - F2B1[z - s1]);
    });
    if (D.dim == 3) {
      emf1.assign(n, 0.);
      emf2.assign(n, 0.);
      const double* F2B3 = F[1] + vars.B3 * n;
      const double* F3B2 = F[2] + vars.B2 * n;
      const double* F3B1 = F[2] + vars.B1 * n;
      const double* F1B3 = F[0] + vars.B3 * n;
      int lo3[3] = {1, 1, 1};
      forRange(lo3, hi, [&](int i, int j, int k) {
        const int64_t z = D.idx(i, j, k);
        emf1[z] = 0.25 * (F2B3[z] + F2B3[z - s3] - F3B2[z] - F3B2[z - s2]);
        emf2[z] = 0.25 * (F3B1[z] + F3B1[z - s1] - F1B3[z] - F1B3[z - s3]);
      });
    }
    int lo2[3] = {0, 0, 0}, hi2[3] = {D.NT[0] - 1, D.NT[1] - 1, D.dim == 3 ? D.NT[2] - 1 : D.NT[2]};
    forRange(lo2, hi2, [&](int i, int j, int k) {
      const int64_t z = D.idx(i, j, k);
      F[0][vars.B1 * n + z] = 0.;
      F[0][vars.B2 * n + z] = 0.5 * (emf3[z] + emf3[z + s2]);
      F[1][vars.B1 * n + z] = -0.5 * (emf3[z] + emf3[z + s1]);
      F[1][vars.B2 * n + z] = 0.;
      if (D.dim == 3) {
        F[0][vars.B3 * n + z] = -0.5 * (emf2[z] + emf2[z + s3]);
        F[1][vars.B3 * n + z] = 0.5 * (emf1[z] + emf1[z + s3]);
        F[2][vars.B1 * n + z] = 0.5 * (emf2[z] + emf2[z + s1]);
        F[2][vars.B2 * n + z] = -0.5 * (emf1[z] + emf1[z + s2]);
        F[2][vars.B3 * n + z] = 0.;
      }
    });
  }

  if (torusFilter) {
    // applyProblemSpecificFluxFilter, torus/problem.cpp:1243-1285: no mass
    // inflow through the radial boundary faces, no flux through the polar
    // faces (the F1_B2 ghost-row copies it also makes are never read).
    const int i0 = D.g[0], i1 = D.N[0] + D.g[0];
    const int j0 = D.g[1], j1 = D.N[1] + D.g[1];
    for (int k = 0; k < D.NT[2]; k++)
      for (int j = 0; j < D.NT[1]; j++) {
        double* f = F[0] + RHO * n;
        f[D.idx(i0, j, k)] = std::min(f[D.idx(i0, j, k)], 0.);
        f[D.idx(i1, j, k)] = std::max(f[D.idx(i1, j, k)], 0.);
      }
    if (D.dim >= 2)
      for (int k = 0; k < D.NT[2]; k++)
        for (int i = 0; i < D.NT[0]; i++)
          for (int v = 0; v < nv; v++) {
            F[1][v * n + D.idx(i, j0, k)] = 0.;
            F[1][v * n + D.idx(i, j1, k)] = 0.;
          }
  }

  // Divergence over the bulk.  dim < 3: af::convolve with {1,-1,0}/dX
  // (F[i+1]*(1/dX) + F[i]*(-1/dX)); dim 3: (shift(F,-1) - F)/dX.
  const double dX[3] = {dXof(c, 0), dXof(c, 1), dXof(c, 2)};
  int lo[3] = {D.g[0], D.g[1], D.g[2]};
  int hi[3] = {D.g[0] + D.N[0], D.g[1] + D.N[1], D.g[2] + D.N[2]};
  forRange(lo, hi, [&](int i, int j, int k) {
    const int64_t z = D.idx(i, j, k);
    for (int v = 0; v < nv; v++) {
      double term[3] = {0., 0., 0.};
      for (int d = 0; d < D.dim; d++) {
        const double* f = F[d] + v * n;
        const int64_t s = D.stride(d);
        if (D.dim < 3) {
          const double w = 1. / dX[d];
          term[d] = f[z + s] * w + f[z] * (-w);
        } else {
          term[d] = (f[z + s] - f[z]) / dX[d];
        }
      }
      double acc = term[0];
      for (int d = 1; d < D.dim; d++) acc = acc + term[d];
      divF[v * n + z] = acc;
    }
  });
  if (fluxesOut)
    for (int d = 0; d < D.dim; d++)
      std::copy(F[d], F[d] + (size_t)nv * n, fluxesOut + (size_t)d * nv * n);
}

// fluidElement::computeExplicitSources + computeEMHDGradients
// (physics.cpp:360-586) for the bulk zones of the element set at prim.
void explicitSources(const grim_oracle_grid& c, const grim_oracle_params& p, const Dims& D,
                     const grim_oracle_geometry& gm, const double* prim, double* S) {
  const Vars vars(p);
  const int nv = vars.dof;
  const int64_t n = D.n;
  const bool emhd = p.conduction || p.viscosity;
  // u_mu and T of every zone (the element's arrays that the slopes read).
  std::vector<double> uT;
  if (emhd) {
    uT.assign((size_t)5 * n, 0.);
    int lo[3] = {0, 0, 0}, hi[3] = {D.NT[0], D.NT[1], D.NT[2]};
    forRange(lo, hi, [&](int i, int j, int k) {
      const int64_t z = D.idx(i, j, k);
      double P[GRIM_ORACLE_MAX_VARS] = {};
      loadPrim(prim, nv, n, z, P);
      ZoneGeom<double> G;
      loadLocation(gm.center, n, z, &G);
      FluidElement<double> e;
      e.set(P, G, p, ZoneClosure{1., 1.});
      for (int m = 0; m < 4; m++) uT[m * n + z] = e.uCov[m];
      uT[4 * n + z] = e.temperature;
    });
  }
  const double dX[3] = {dXof(c, 0), dXof(c, 1), dXof(c, 2)};
  int lo[3] = {D.g[0], D.g[1], D.g[2]};
  int hi[3] = {D.g[0] + D.N[0], D.g[1] + D.N[1], D.g[2] + D.N[2]};
  forRange(lo, hi, [&](int i, int j, int k) {
    const int64_t z = D.idx(i, j, k);
    double P[GRIM_ORACLE_MAX_VARS];
    loadPrim(prim, nv, n, z, P);
    ZoneGeom<double> G;
    loadLocation(gm.center, n, z, &G);
    FluidElement<double> e;
    e.set(P, G, p, ZoneClosure{1., 1.});
    double src[GRIM_ORACLE_MAX_VARS];
    for (int v = 0; v < nv; v++) src[v] = 0.;
    auto gamma = [&](int l, int a, int b) {
      return gm.gamma ? gm.gamma[(l * 16 + a * 4 + b) * n + z] : 0.;
    };
    if (c.metric != GRIM_ORACLE_METRIC_MINKOWSKI) {
      for (int nu = 0; nu < 4; nu++)
        for (int kappa = 0; kappa < 4; kappa++)
          for (int lamda = 0; lamda < 4; lamda++)
            src[UU + nu] -= G.g * e.TUpDown[kappa][lamda] * gamma(lamda, kappa, nu);
    }
    if (emhd) {
      auto slopeOf = [&](const double* a, int d) {
        const int64_t s = D.stride(d);
        double y[5];
        for (int q = 0; q < 5; q++) y[q] = a[z + (q - 2) * s];
        return slopeZone(c, dX[d], y);
      };
      double graduCov[4][4], gradT[4];
      for (int mu = 0; mu < 4; mu++) {
        const double* um = uT.data() + mu * n;
        graduCov[0][mu] = 0.;
        graduCov[1][mu] = slopeOf(um, 0);
        graduCov[2][mu] = 0.;
        if (D.dim > 1) graduCov[2][mu] = slopeOf(um, 1);
        graduCov[3][mu] = 0.;
        if (D.dim > 2) graduCov[3][mu] = slopeOf(um, 2);
        for (int nu = 0; nu < 4; nu++)
          for (int lambda = 0; lambda < 4; lambda++)
            graduCov[nu][mu] -= gamma(lambda, nu, mu) * e.uCov[lambda];
      }
      if (p.conduction) {
        const double* Tp = uT.data() + 4 * n;
        gradT[0] = 0.;
        gradT[1] = slopeOf(Tp, 0);
        gradT[2] = 0.;
        if (D.dim > 1) gradT[2] = slopeOf(Tp, 1);
        gradT[3] = 0.;
        if (D.dim > 2) gradT[3] = slopeOf(Tp, 2);
      }
      double divuCov = 0.;
      for (int mu = 0; mu < 4; mu++)
        for (int nu = 0; nu < 4; nu++) divuCov += G.gCon[mu][nu] * graduCov[mu][nu];
      if (p.viscosity) {
        double deltaP0 = -divuCov * e.rho * e.nu_emhd;
        for (int mu = 0; mu < 4; mu++)
          for (int nu = 0; nu < 4; nu++)
            deltaP0 += 3. * e.rho * e.nu_emhd * e.bCon[mu] * e.bCon[nu] / e.bSqr * graduCov[mu][nu];
        double deltaP0Tilde = deltaP0;
        if (p.highOrderTermsViscosity == 1)
          deltaP0Tilde = deltaP0 * std::sqrt(e.tau / e.rho / e.nu_emhd / e.temperature);
        src[vars.DP] = -G.g * (deltaP0Tilde) / e.tau;
        if (p.highOrderTermsViscosity == 1) src[vars.DP] -= 0.5 * G.g * divuCov * e.deltaPTilde;
      }
      if (p.conduction) {
        double q0 = 0.;
        for (int mu = 0; mu < 4; mu++) {
          q0 -= e.rho * e.chi_emhd * e.bCon[mu] / e.bNorm * gradT[mu];
          for (int nu = 0; nu < 4; nu++)
            q0 -= e.rho * e.chi_emhd * e.temperature * e.bCon[nu] / e.bNorm * e.uCon[mu] *
                  graduCov[mu][nu];
        }
        double q0Tilde = q0;
        if (p.highOrderTermsConduction == 1)
          q0Tilde = q0 * std::sqrt(e.tau / e.rho / e.chi_emhd) / e.temperature;
        src[vars.Q] = -G.g * (q0Tilde) / e.tau;
        if (p.highOrderTermsConduction == 1) src[vars.Q] -= 0.5 * G.g * divuCov * e.qTilde;
      }
    }
    for (int v = 0; v < nv; v++) S[v * n + z] = src[v];
  });
}

// grid::communicate on one rank (grid.cpp:306-360): the local vector of a
// DMDA with a box stencil receives the periodic image in every ghost zone
// whose out-of-range coordinates all lie along periodic directions; the
// other ghost zones lie outside the global domain, are never written by
// DMGlobalToLocal and keep the local vector's initial zeros, which
// copyLocalVecToVars then writes back.
void communicate(const grim_oracle_grid& c, const Dims& D, int nv, double* prim) {
  const int bLo[3] = {c.boundaryLeft, c.boundaryBottom, c.boundaryBack};
  const int bHi[3] = {c.boundaryRight, c.boundaryTop, c.boundaryFront};
  int lo[3] = {0, 0, 0}, hi[3] = {D.NT[0], D.NT[1], D.NT[2]};
  forRange(lo, hi, [&](int i, int j, int k) {
    int I[3] = {i, j, k}, W[3];
    bool ghost = false, outside = false;
    for (int d = 0; d < 3; d++) {
      W[d] = I[d];
      if (D.g[d] == 0) continue;
      if (I[d] < D.g[d]) {
        if (bLo[d] != GRIM_ORACLE_BC_PERIODIC) outside = true;
        W[d] = I[d] + D.N[d];
        ghost = true;
      } else if (I[d] >= D.g[d] + D.N[d]) {
        if (bHi[d] != GRIM_ORACLE_BC_PERIODIC) outside = true;
        W[d] = I[d] - D.N[d];
        ghost = true;
      }
    }
    if (!ghost) return;
    const int64_t z = D.idx(i, j, k), w = D.idx(W[0], W[1], W[2]);
    for (int v = 0; v < nv; v++) prim[v * D.n + z] = outside ? 0. : prim[v * D.n + w];
  });
}

// boundaries::applyBoundaryConditions, boundary.cpp:5-243
void applyBCs(const grim_oracle_grid& c, const Dims& D, int nv, double* prim) {
  const int types[6] = {c.boundaryLeft, c.boundaryRight, c.boundaryBottom,
                        c.boundaryTop,  c.boundaryBack,  c.boundaryFront};
  for (int d = 0; d < D.dim; d++) {
    for (int side = 0; side < 2; side++) {
      const int type = types[2 * d + side];
      if (type != GRIM_ORACLE_BC_MIRROR && type != GRIM_ORACLE_BC_OUTFLOW) continue;
      const int g = D.g[d], N = D.N[d];
      int lo[3] = {0, 0, 0}, hi[3] = {D.NT[0], D.NT[1], D.NT[2]};
      lo[d] = side == 0 ? 0 : N + g;
      hi[d] = side == 0 ? g : N + 2 * g;
      forRange(lo, hi, [&](int i, int j, int k) {
        int I[3] = {i, j, k};
        const int a = I[d];
        int src;
        if (type == GRIM_ORACLE_BC_MIRROR)
          src = side == 0 ? 2 * g - 1 - a : 2 * (N + g) - 1 - a;
        else
          src = side == 0 ? g : N + g - 1;
        I[d] = src;
        const int64_t z = D.idx(i, j, k), w = D.idx(I[0], I[1], I[2]);
        for (int v = 0; v < nv; v++) prim[v * D.n + z] = prim[v * D.n + w];
      });
    }
  }
}

double computeDt(const grim_oracle_grid& c, const grim_oracle_params& p, const Dims& D,
                 const grim_oracle_geometry& gm, const double* prim, double dt) {
  const Vars vars(p);
  const int nv = vars.dof;
  const int64_t n = D.n;
  const double dX[3] = {dXof(c, 0), dXof(c, 1), dXof(c, 2)};
  double maxInvDt = -INFINITY;
#pragma omp parallel for schedule(static) reduction(max : maxInvDt)
  for (int64_t z = 0; z < n; z++) {
    double P[GRIM_ORACLE_MAX_VARS];
    loadPrim(prim, nv, n, z, P);
    ZoneGeom<double> G;
    loadLocation(gm.center, n, z, &G);
    FluidElement<double> e;
    e.set(P, G, p, ZoneClosure{1., 1.});
    double maxSpeed = 0.;
    for (int d = 0; d < D.dim; d++) {
      double mn, mx;
      e.computeMinMaxCharSpeeds(d, &mn, &mx, p);
      mn = mn / dX[d];
      mx = mx / dX[d];
      const double m = std::max(mx, std::fabs(mn));
      maxSpeed = (d == 0) ? m : maxSpeed + m;
    }
    maxInvDt = std::max(maxInvDt, maxSpeed);
  }
  double newDt = c.CourantFactor / maxInvDt;
  if (newDt > c.maxDtIncrement * dt) newDt = c.maxDtIncrement * dt;
  return newDt;
}

// ---------------------------------------------------------------------------
// Torus problem (src/problem/torus/problem.cpp)
// ---------------------------------------------------------------------------
// Counter-based uniform in [0, 1): splitmix64 (Steele, Lea & Flood 2014) of
// seed + G (32 z + 1), the same stream as the product's batch generator.
double uniformAt(uint64_t seed, uint64_t z) {
  uint64_t x = seed + 0x9E3779B97F4A7C15ull * (32ull * z + 1ull);
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  x = x ^ (x >> 31);
  return (double)(x >> 11) * (1.0 / 9007199254740992.0);
}

// problem.cpp:56-143 (M = 1)
double lFishboneMoncrief(double a, double r, double) {
  const double M = 1.;
  return std::sqrt(M / std::pow(r, 3.)) *
         (std::pow(r, 4.) + r * r * a * a - 2. * M * r * a * a - a * std::sqrt(M * r) * (r * r - a * a)) /
         (r * r - 3 * M * r + 2. * a * std::sqrt(M * r));
}
double computeDelta(double a, double r, double) { return r * r - 2 * 1. * r + a * a; }
double computeSigma(double a, double r, double theta) {
  return r * r + a * a * std::cos(theta) * std::cos(theta);
}
double computeA(double a, double r, double theta) {
  const double Delta = computeDelta(a, r, theta);
  return std::pow(r * r + a * a, 2.) - Delta * a * a * std::sin(theta) * std::sin(theta);
}
double lnOfhTerm1(double a, double r, double theta, double l) {
  const double Delta = computeDelta(a, r, theta), Sigma = computeSigma(a, r, theta),
               A = computeA(a, r, theta);
  return 0.5 * std::log((1. + std::sqrt(1. + (4. * l * l * Sigma * Sigma * Delta) /
                                                 (A * std::sin(theta) * A * std::sin(theta)))) /
                        (Sigma * Delta / A));
}
double lnOfhTerm2(double a, double r, double theta, double l) {
  const double Delta = computeDelta(a, r, theta), Sigma = computeSigma(a, r, theta),
               A = computeA(a, r, theta);
  return -0.5 * std::sqrt(1. + (4. * l * l * Sigma * Sigma * Delta) /
                                   (A * A * std::sin(theta) * std::sin(theta)));
}
double lnOfhTerm3(double a, double r, double theta, double l) {
  const double A = computeA(a, r, theta);
  const double M = 1.;
  return -2 * a * M * r * l / A;
}
double computeLnOfh(const grim_oracle_torus& t, double a, double r, double theta) {
  const double l = lFishboneMoncrief(a, t.PressureMaxRadius, M_PI / 2.);
  const double term1 = lnOfhTerm1(a, r, theta, l);
  const double term2 = lnOfhTerm2(a, r, theta, l);
  const double term3 = lnOfhTerm3(a, r, theta, l);
  const double term1InnerEdge = lnOfhTerm1(a, t.InnerEdgeRadius, M_PI / 2., l);
  const double term2InnerEdge = lnOfhTerm2(a, t.InnerEdgeRadius, M_PI / 2., l);
  const double term3InnerEdge = lnOfhTerm3(a, t.InnerEdgeRadius, M_PI / 2., l);
  return term1 + term2 + term3 - term1InnerEdge - term2InnerEdge - term3InnerEdge;
}

// The torus's fluid state at one zone (problem.cpp:175-322): fills rho, u,
// u^1..u^3 of P (B left to the caller).
void torusZone(const grim_oracle_grid& c, const grim_oracle_params& p, const grim_oracle_torus& t,
               double r, double theta, double X2, double lapse, const double beta[4], double xi,
               double* P) {
  const double aBH = c.blackHoleSpin;
  double lnOfh = 1.;
  if (r >= t.InnerEdgeRadius) lnOfh = computeLnOfh(t, aBH, r, theta);
  if (lnOfh < 0. || r < t.InnerEdgeRadius) {
    P[RHO] = p.rhoFloorInFluidElement;
    P[UU] = p.uFloorInFluidElement;
    P[U1] = 0.;
    P[U2] = 0.;
    P[U3] = 0.;
    return;
  }
  const double h = std::exp(lnOfh);
  const double Gamma = p.adiabaticIndex, Kappa = t.Adiabat;
  P[RHO] = std::pow((h - 1) * (Gamma - 1.) / (Kappa * Gamma), 1. / (Gamma - 1.));
  P[UU] = Kappa * std::pow(P[RHO], Gamma) / (Gamma - 1.) *
          (1. + t.InitialPerturbationAmplitude * (xi - 0.5));
  const double A = computeA(aBH, r, theta), Sigma = computeSigma(aBH, r, theta),
               Delta = computeDelta(aBH, r, theta);
  const double l = lFishboneMoncrief(aBH, t.PressureMaxRadius, M_PI / 2.);
  const double expOfMinus2Chi = Sigma * Sigma * Delta / (A * A * std::sin(theta) * std::sin(theta));
  const double uCovPhiBL = std::sqrt((-1. + std::sqrt(1. + 4 * l * l * expOfMinus2Chi)) / 2.);
  const double uConPhiBL = 2. * aBH * r * std::sqrt(1. + uCovPhiBL * uCovPhiBL) / std::sqrt(A * Sigma * Delta) +
                           std::sqrt(Sigma / A) * uCovPhiBL / std::sin(theta);
  double uConBL[4] = {0., 0., 0., uConPhiBL};
  double gCovBL[4][4], T[4][4];
  for (int a = 0; a < 4; a++)
    for (int b = 0; b < 4; b++) gCovBL[a][b] = T[a][b] = 0.;
  const double mu = 1 + aBH * aBH * std::cos(theta) * std::cos(theta) / (r * r);
  gCovBL[0][0] = -(1. - 2. / (r * mu));
  gCovBL[0][3] = -2. * aBH * std::sin(theta) * std::sin(theta) / (r * mu);
  gCovBL[3][0] = gCovBL[0][3];
  gCovBL[1][1] = mu * r * r / Delta;
  gCovBL[2][2] = r * r * mu;
  gCovBL[3][3] = r * r * std::sin(theta) * std::sin(theta) *
                 (1. + aBH * aBH / (r * r) + 2. * aBH * aBH * std::sin(theta) * std::sin(theta) / (r * r * r * mu));
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
  uConBL[0] = -(BB + std::sqrt(discriminent)) / (2. * AA);
  double uConKS[4];
  for (int a = 0; a < 4; a++) {
    uConKS[a] = 0.;
    for (int b = 0; b < 4; b++) uConKS[a] += T[a][b] * uConBL[b];
  }
  const double rFactor = r;
  const double hFactor = M_PI + (1. - c.hSlope) * M_PI * std::cos(2. * M_PI * X2);
  const double uConMKS[4] = {uConKS[0], uConKS[1] / rFactor, uConKS[2] / hFactor, uConKS[3]};
  P[U1] = uConMKS[1] + std::pow(lapse, 2.) * beta[1] * uConMKS[0];
  P[U2] = uConMKS[2] + std::pow(lapse, 2.) * beta[2] * uConMKS[0];
  P[U3] = uConMKS[3] + std::pow(lapse, 2.) * beta[3] * uConMKS[0];
}

// applyFloor, problem.cpp:542-712, one zone.  minRho/minU are the floor
// values pow(r, slope)*ampl of the zone.
void applyFloorZone(const grim_oracle_params& p, const grim_oracle_torus& t, const ZoneGeom<double>& G,
                    double minRho, double minU, double* P) {
  const Vars vars(p);
  const ZoneClosure zc{1., 1.};
  const double rho_prefloor = P[RHO] * 1.;
  const double u_prefloor = P[UU] * 1.;
  double condition = P[RHO] < minRho ? 1. : 0.;
  double usefloor = condition;
  P[RHO] = condition * minRho + (1. - condition) * P[RHO];
  condition = P[UU] < minU ? 1. : 0.;
  usefloor = std::max(condition, usefloor);
  P[UU] = condition * minU + (1. - condition) * P[UU];
  FluidElement<double> e;
  e.set(P, G, p, zc);
  double bSqr = e.bSqr;
  condition = bSqr > t.BsqrOverRhoMax * P[RHO] ? 1. : 0.;
  usefloor = std::max(condition, usefloor);
  P[RHO] = P[RHO] * (1. - condition) + condition * bSqr / t.BsqrOverRhoMax;
  condition = bSqr > t.BsqrOverUMax * P[UU] ? 1. : 0.;
  usefloor = std::max(condition, usefloor);
  P[UU] = P[UU] * (1. - condition) + condition * bSqr / t.BsqrOverUMax;
  const double zero = 0.;
  const double trans =
      std::max(std::min(zero + 1., (bSqr - 0.1 * P[RHO]) / P[RHO]), zero) * (usefloor > zero ? 1. : 0.);
  // Sasha's floor method
  const double betapar = -e.bCon[0] / bSqr / e.uCon[0];
  double betasqr = betapar * betapar * bSqr;
  const double betasqrmax = 1. - 1. / t.MaxLorentzFactor / t.MaxLorentzFactor;
  betasqr = std::min(betasqr, betasqrmax);
  const double gamma = 1. / std::sqrt(1. - betasqr);
  double ucondr[4];
  for (int m = 0; m < 4; m++) ucondr[m] = gamma * (e.uCon[m] + betapar * e.bCon[m]);
  const double Bcon[4] = {zero, P[vars.B1], P[vars.B2], P[vars.B3]};
  double Bcov[4];
  for (int m = 0; m < 4; m++) {
    Bcov[m] = zero;
    for (int n = 0; n < 4; n++) Bcov[m] += G.gCov[n][m] * Bcon[n];
  }
  double udotB = zero, bSqr_fl = zero;
  for (int n = 0; n < 4; n++) {
    udotB += Bcov[n] * e.uCon[n];
    bSqr_fl += Bcov[n] * Bcon[n];
  }
  double bNorm = std::sqrt(bSqr_fl);
  const double bMin = std::sqrt(p.bSqrFloorInFluidElement);
  bNorm = std::max(bNorm, bMin);
  const double wold = rho_prefloor + u_prefloor * p.adiabaticIndex;
  const double QdotB = udotB * wold * e.uCon[0];
  const double wnew = P[RHO] + P[UU] * p.adiabaticIndex;
  const double x = 2. * QdotB / (bNorm * wnew * ucondr[0]);
  const double vpar = x / (ucondr[0] * (1. + std::sqrt(1. + x * x)));
  const double one_over_ucondr_t = 1. / ucondr[0];
  double vcon[4];
  vcon[0] = zero + 1.;
  for (int m = 1; m < 4; m++) vcon[m] = vpar * Bcon[m] / bNorm + ucondr[m] * one_over_ucondr_t;
  double vsqr = zero;
  for (int m = 0; m < 4; m++)
    for (int n = 0; n < 4; n++) vsqr += G.gCov[m][n] * vcon[m] * vcon[n];
  const double condition_v = (vsqr >= 0. || vsqr < 1. / G.gCon[0][0]) ? 1. : 0.;
  vsqr = (1. - condition_v) * vsqr + condition_v / G.gCon[0][0];
  const double ut = std::sqrt(-1. / vsqr);
  double utcon[4];
  utcon[0] = zero;
  for (int m = 1; m < 4; m++) utcon[m] = ut * (vcon[m] - G.gCon[0][m] / G.gCon[0][0]);
  P[U1] = P[U1] * (1. - trans) + trans * utcon[1];
  P[U2] = P[U2] * (1. - trans) + trans * utcon[2];
  P[U3] = P[U3] * (1. - trans) + trans * utcon[3];
  // Maximum Lorentz factor.  grim multiplies only when some zone of the grid
  // exceeds it; every other zone's factor is exactly 1, so applying it
  // everywhere changes no value.
  e.set(P, G, p, zc);
  const double lorentzFactorSqr = e.gammaLorentzFactor * e.gammaLorentzFactor;
  double MultFac = std::sqrt(std::max((lorentzFactorSqr - 1.) /
                                          (t.MaxLorentzFactor * t.MaxLorentzFactor - 1.),
                                      1.));
  MultFac = 1. / MultFac;
  P[U1] *= MultFac;
  P[U2] *= MultFac;
  P[U3] *= MultFac;
  e.set(P, G, p, zc);
  if (p.conduction) {
    const double Qmax = 1.07 * p.ConductionClosureFactor * P[RHO] * std::pow(e.soundSpeed, 3.);
    const double LimFac = std::max(std::fabs(e.q) / Qmax, 1.);
    P[vars.Q] = P[vars.Q] / LimFac;
  }
  if (p.viscosity) {
    const double pressure = e.pressure, deltaP = e.deltaP;
    bSqr = e.bSqr;
    const double dPmod = std::max(pressure - 2. / 3. * deltaP, 0.01 * p.bSqrFloorInFluidElement) /
                         std::max(pressure + 1. / 3. * deltaP, p.bSqrFloorInFluidElement);
    const double dPmaxPlus = std::min(1.07 * p.ViscosityClosureFactor * bSqr * 0.5 * dPmod, 1.49 * pressure);
    const double dPmaxMinus = std::max(-1.07 * p.ViscosityClosureFactor * bSqr, -2.99 * pressure);
    const double cond = deltaP > 0. ? 1. : 0.;
    P[vars.DP] = P[vars.DP] * (cond / std::max(deltaP / dPmaxPlus, 1.) +
                               (1. - cond) / std::max(deltaP / dPmaxMinus, 1.));
  }
}

void applyFloorAll(const grim_oracle_grid& c, const grim_oracle_params& p, const grim_oracle_torus& t,
                   const Dims& D, const grim_oracle_geometry& gm, double* prim) {
  const Vars vars(p);
  const int64_t n = D.n;
  int lo[3] = {0, 0, 0}, hi[3] = {D.NT[0], D.NT[1], D.NT[2]};
  forRange(lo, hi, [&](int i, int j, int k) {
    const int64_t z = D.idx(i, j, k);
    ZoneGeom<double> G;
    loadLocation(gm.center, n, z, &G);
    const double Radius = gm.center[22 * n + z];
    const double minRho = t.minRho ? t.minRho[z] : std::pow(Radius, t.RhoFloorSlope) * t.RhoFloorAmpl;
    const double minU = t.minU ? t.minU[z] : std::pow(Radius, t.UFloorSlope) * t.UFloorAmpl;
    double P[GRIM_ORACLE_MAX_VARS];
    loadPrim(prim, vars.dof, n, z, P);
    applyFloorZone(p, t, G, minRho, minU, P);
    for (int v = 0; v < vars.dof; v++) prim[v * n + z] = P[v];
  });
}

// inflowCheck (problem.cpp:941-1044) then fixPoles (:1046-1215).
void torusBCs(const grim_oracle_grid& c, const grim_oracle_params& p, const grim_oracle_torus& t,
              const Dims& D, const grim_oracle_geometry& gm, double* prim) {
  const Vars vars(p);
  const int64_t n = D.n;
  const int g = D.g[0];
  auto at = [&](int v, int i, int j, int k) -> double& { return prim[v * n + D.idx(i, j, k)]; };
  auto geo = [&](int plane, int i, int j, int k) { return gm.center[plane * n + D.idx(i, j, k)]; };
  const double vSqrMax = 1. - 1. / t.MaxLorentzFactor / t.MaxLorentzFactor;
  const double vSqrMin = 1.e-13;
  for (int side = 0; side < 2; side++) {
    const int iLo = side == 0 ? 0 : D.N[0] + g, iHi = side == 0 ? g : D.N[0] + 2 * g;
    int lo[3] = {iLo, 0, 0}, hi[3] = {iHi, D.NT[1], D.NT[2]};
    forRange(lo, hi, [&](int i, int j, int k) {
      const int64_t z = D.idx(i, j, k);
      ZoneGeom<double> G;
      loadLocation(gm.center, n, z, &G);
      double P[GRIM_ORACLE_MAX_VARS];
      loadPrim(prim, vars.dof, n, z, P);
      FluidElement<double> e;
      e.set(P, G, p, ZoneClosure{1., 1.});
      double u[3] = {P[U1] / e.gammaLorentzFactor, P[U2] / e.gammaLorentzFactor,
                     P[U3] / e.gammaLorentzFactor};
      if (side == 1) {
        // minUr of the last bulk column, same (j, k)
        const int il = D.N[0] + g - 1;
        const double minUr = geo(10 + 1, il, j, k) * geo(14, il, j, k);
        u[0] = std::max(minUr, u[0]);
      } else {
        u[0] = std::min(G.gCon[0][1] * G.alpha, u[0]);
      }
      double vSqr = u[0] * 0.;
      for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) vSqr += G.gCov[a + 1][b + 1] * u[a] * u[b];
      vSqr = std::max(std::min(vSqr, vSqrMax), vSqrMin);
      const double newLorentzFactor = 1. / std::sqrt(1. - vSqr);
      for (int a = 0; a < 3; a++) prim[(U1 + a) * n + z] = u[a] * newLorentzFactor;
    });
  }
  if (D.dim < 2) return;
  // fixPoles
  const int g2 = D.g[1], N2 = D.N[1];
  for (int pole = 0; pole < 2; pole++) {
    const int idx0 = pole == 0 ? g2 : N2 + g2 - 1;
    const int idx1 = pole == 0 ? g2 + 1 : idx0 - 1;
    const int idx2 = pole == 0 ? g2 + 2 : idx0 - 2;
    for (int k = 0; k < D.NT[2]; k++)
      for (int i = 0; i < D.NT[0]; i++) {
        auto Theta = [&](int j) {
          const double th = geo(23, i, j, k);
          return pole == 0 ? th : th * (-1.) + M_PI;
        };
        for (int jj : {idx0, idx1}) {
          at(RHO, i, jj, k) = at(RHO, i, idx2, k);
          at(UU, i, jj, k) = at(UU, i, idx2, k);
          at(U1, i, jj, k) = at(U1, i, idx2, k);
          at(U2, i, jj, k) = at(U2, i, idx2, k) * Theta(jj) / Theta(idx2);
          at(U3, i, jj, k) = at(U3, i, idx2, k);
          if (p.conduction) at(vars.Q, i, jj, k) = at(vars.Q, i, idx2, k) * Theta(jj) / Theta(idx2);
          if (p.viscosity) at(vars.DP, i, jj, k) = at(vars.DP, i, idx2, k) * Theta(jj) / Theta(idx2);
        }
        for (int q = 0; q < g2; q++) {
          const int dst = pole == 0 ? g2 - 1 - q : N2 + g2 + q;
          const int src = pole == 0 ? g2 + q : N2 + g2 - 1 - q;
          at(RHO, i, dst, k) = at(RHO, i, src, k);
          at(UU, i, dst, k) = at(UU, i, src, k);
          at(U1, i, dst, k) = at(U1, i, src, k);
          at(U3, i, dst, k) = at(U3, i, src, k);
          at(U2, i, dst, k) = at(U2, i, src, k) * (-1.0);
          at(vars.B1, i, dst, k) = at(vars.B1, i, src, k);
          at(vars.B3, i, dst, k) = at(vars.B3, i, src, k);
          at(vars.B2, i, dst, k) = at(vars.B2, i, src, k) * (-1.0);
          if (p.conduction) at(vars.Q, i, dst, k) = at(vars.Q, i, src, k) * (-1.0);
          if (p.viscosity) at(vars.DP, i, dst, k) = at(vars.DP, i, src, k) * (-1.0);
        }
      }
  }
}

// fullStepDiagnostics' NaN scrub (problem.cpp:722-741): when the bulk holds a
// NaN, every zone holding one is reset to zero in all variables.
void scrubNaNs(const Dims& D, int nv, double* prim) {
  int64_t bulkNaN = 0;
  for (int k = D.g[2]; k < D.g[2] + D.N[2]; k++)
    for (int j = D.g[1]; j < D.g[1] + D.N[1]; j++)
      for (int i = D.g[0]; i < D.g[0] + D.N[0]; i++)
        for (int v = 0; v < nv; v++) bulkNaN += std::isnan(prim[v * D.n + D.idx(i, j, k)]) ? 1 : 0;
  if (bulkNaN == 0) return;
  for (int64_t z = 0; z < D.n; z++) {
    bool bad = false;
    for (int v = 0; v < nv; v++) bad = bad || std::isnan(prim[v * D.n + z]);
    if (bad)
      for (int v = 0; v < nv; v++) prim[v * D.n + z] = 0.;
  }
}

// One stage of timeStep (timestep.cpp:14-107 half, 183-255 full).
// setProblemSpecificBCs of bondi_inflow (bondi_inflow/problem.cpp:157-183):
// every ghost zone of primOld taken from the initial state after the
// boundary conditions (null: hook off).  grim calls the hook in both stages
// and it always writes primOld, so only the half stage changes anything.
const double* g_ghostState = nullptr;

void ghostStateBCs(const Dims& D, int nv, const double* ic, double* prim) {
  for (int k = 0; k < D.NT[2]; k++)
    for (int j = 0; j < D.NT[1]; j++)
      for (int i = 0; i < D.NT[0]; i++) {
        if (D.bulk(i, j, k)) continue;
        const int64_t z = D.idx(i, j, k);
        for (int v = 0; v < nv; v++) prim[v * D.n + z] = ic[v * D.n + z];
      }
}

// setProblemSpecificBCs of buoyancy_instabilities (problem.cpp:221-293) on
// the stage state, both stages (null: hook off).  af::seq(0, numGhost) is
// inclusive: the first bulk zone of each X1 line is reset too.
const double* g_sideState = nullptr;

void sideStateBCs(const Dims& D, const Vars& vars, bool conduction, const double* ic, double* prim) {
  const int64_t n = D.n;
  for (int k = 0; k < D.NT[2]; k++)
    for (int j = 0; j < D.NT[1]; j++)
      for (int i = 0; i < D.NT[0]; i++) {
        if (i > D.g[0] && i < D.N[0] + D.g[0]) continue;
        const int64_t z = D.idx(i, j, k);
        prim[RHO * n + z] = ic[z];
        prim[UU * n + z] = ic[n + z];
        prim[U1 * n + z] = ic[2 * n + z];
        prim[U2 * n + z] = 0.;
        prim[U3 * n + z] = 0.;
        for (int d = 0; d < 3; d++) prim[(vars.B1 + d) * n + z] = 0.;
        if (conduction) prim[vars.Q * n + z] = 0.;
      }
}

void stage(const grim_oracle_grid& c, const grim_oracle_params& p, const Dims& D,
           const grim_oracle_geometry& gm, const grim_oracle_torus* torus, int st, double dt,
           double* primOld, double* primHalf, double* consOld, double* primNew, int semantics,
           grim_oracle_step_stats* stats) {
  const Vars vars(p);
  const int nv = vars.dof, nf = vars.numFluidVars;
  const int64_t n = D.n;
  const bool emhd = p.conduction || p.viscosity;
  // The stage state was communicated at the end of the previous solve (or
  // holds the initial conditions everywhere); only the physical boundaries
  // and the problem's boundary hook are applied here, as timestep.cpp:17-22
  // / 186-191 do.
  double* primStage = (st == 0) ? primOld : primHalf;
  applyBCs(c, D, nv, primStage);
  if (torus) torusBCs(c, p, *torus, D, gm, primStage);
  if (g_ghostState && st == 0) ghostStateBCs(D, nv, g_ghostState, primStage);
  if (g_sideState) sideStateBCs(D, vars, p.conduction != 0, g_sideState, primStage);
  const double dtPrime = (st == 0) ? dt / 2. : dt;

  std::vector<double> Sexp((size_t)nv * n, 0.), divF((size_t)nv * n, 0.);
  explicitSources(c, p, D, gm, primStage, Sexp.data());
  divFluxes(c, p, D, gm, primStage, divF.data(), nullptr, torus != nullptr);

  std::vector<double> rhs((size_t)nf * n, 0.), mask(n, 0.);
  int lo[3] = {D.g[0], D.g[1], D.g[2]};
  int hi[3] = {D.g[0] + D.N[0], D.g[1] + D.N[1], D.g[2] + D.N[2]};
  // Guess: primOld (half) or the half-step solution (full) for the fluid
  // variables of every zone; B from the induction equation in the bulk.
  if (st == 0) std::copy(primOld, primOld + (size_t)nv * n, primNew);
  else std::copy(primHalf, primHalf + (size_t)nv * n, primNew);
  forRange(lo, hi, [&](int i, int j, int k) {
    const int64_t z = D.idx(i, j, k);
    mask[z] = 1.;
    ZoneGeom<double> G;
    loadLocation(gm.center, n, z, &G);
    ZoneClosure zc{1., 1.};
    double P[GRIM_ORACLE_MAX_VARS], U[GRIM_ORACLE_MAX_VARS];
    FluidElement<double> eOld;
    if (st == 0) {
      loadPrim(primOld, nv, n, z, P);
      eOld.set(P, G, p, zc);
      eOld.computeFluxes(0, U, p);
      for (int v = 0; v < nv; v++) consOld[v * n + z] = U[v];
    } else {
      for (int v = 0; v < nv; v++) U[v] = consOld[v * n + z];
    }
    double Si[GRIM_ORACLE_MAX_VARS];
    for (int v = 0; v < nv; v++) Si[v] = 0.;
    if (emhd) {
      if (st == 1) {
        loadPrim(primOld, nv, n, z, P);
        eOld.set(P, G, p, zc);
      }
      double tauDamp = eOld.tau;
      if (st == 1) {
        FluidElement<double> eHalf;
        loadPrim(primHalf, nv, n, z, P);
        eHalf.set(P, G, p, zc);
        tauDamp = eHalf.tau;
      }
      eOld.computeImplicitSources(Si, tauDamp, p);
    }
    double rmax = 0.;
    for (int v = 0; v < nf; v++) {
      rhs[v * n + z] = U[v] - dtPrime * ((divF[v * n + z] + Sexp[v * n + z]) + 0.5 * Si[v]);
      rmax = std::max(rmax, std::fabs(rhs[v * n + z]));
    }
    if (c.residualScaling) mask[z] = 1. / rmax;
    // induction equation, timestep.cpp:80-89 / 226-237
    for (int b = 0; b < 3; b++) {
      const int v = vars.B1 + b;
      const double consB = U[v] - dtPrime * divF[v * n + z];
      primNew[v * n + z] = consB / G.g;
    }
  });

  grim_oracle_batch b;
  b.n = n;
  b.prim = primNew;
  b.geom = gm.center;
  b.rhs = rhs.data();
  b.primOld = primOld;
  b.primCoef = (st == 0) ? primOld : primHalf;
  b.closure = nullptr;
  b.radius = gm.center + 22 * n;
  b.dtPrime = dtPrime;
  b.residualMask = mask.data();
  std::vector<int32_t> iters(n, 0);
  std::vector<double> l2(n, 0.);
  grim_oracle_newton(&p, &b, semantics, iters.data(), l2.data(), nullptr, 0);
  if (stats) {
    int64_t unconv = 0, nonfinite = 0;
    int32_t mx = 0;
    for (int64_t z = 0; z < n; z++) {
      if (mask[z] == 0.) continue;
      if (l2[z] > p.nonlinearsolve_atol) unconv++;
      if (!std::isfinite(l2[z])) nonfinite++;
      mx = std::max(mx, iters[z]);
    }
    stats->unconverged[st] = unconv;
    stats->maxIterations[st] = mx;
    stats->nonfinite[st] = nonfinite;
  }
  communicate(c, D, nv, primNew);
  // halfStepDiagnostics / fullStepDiagnostics (torus/problem.cpp:714-741)
  if (torus) {
    if (st == 1) scrubNaNs(D, nv, primNew);
    applyFloorAll(c, p, *torus, D, gm, primNew);
  }
}

// timeStepper::initialConditions of the torus (problem.cpp:145-540).
void torusIC(const grim_oracle_grid& c, const grim_oracle_params& p, const grim_oracle_torus& t,
             const Dims& D, const grim_oracle_geometry& gm, double* prim) {
  const Vars vars(p);
  const int nv = vars.dof;
  const int64_t n = D.n;
  int lo[3] = {0, 0, 0}, hi[3] = {D.NT[0], D.NT[1], D.NT[2]};
  const double dX2 = dXof(c, 1);
  forRange(lo, hi, [&](int i, int j, int k) {
    const int64_t z = D.idx(i, j, k);
    double P[GRIM_ORACLE_MAX_VARS];
    for (int v = 0; v < nv; v++) P[v] = 0.;
    const double X2 = c.X2Start + ((double)j - D.g[1] + 0.5) * dX2;
    const double beta[4] = {0., gm.center[11 * n + z], gm.center[12 * n + z], gm.center[13 * n + z]};
    torusZone(c, p, t, gm.center[22 * n + z], gm.center[23 * n + z], X2, gm.center[14 * n + z], beta,
              uniformAt(t.seed, (uint64_t)z), P);
    for (int v = 0; v < nv; v++) prim[v * n + z] = P[v];
  });
  double rhoMax = -INFINITY;
  for (int64_t z = 0; z < n; z++) rhoMax = std::max(rhoMax, prim[RHO * n + z]);
  for (int64_t z = 0; z < n; z++) {
    prim[RHO * n + z] = prim[RHO * n + z] / rhoMax;
    prim[UU * n + z] = prim[UU * n + z] / rhoMax;
  }
  // Vector potential from the neighbour-averaged density (af::shift is
  // circular), then B = curl A (problem.cpp:355-404).
  std::vector<double> Avec(n, 0.);
  const int64_t s1 = D.stride(0), s2 = D.stride(1), s3 = D.stride(2);
  auto wrap = [&](int64_t z, int d, int off) {
    int I[3] = {(int)(z % D.NT[0]), (int)((z / D.NT[0]) % D.NT[1]), (int)(z / ((int64_t)D.NT[0] * D.NT[1]))};
    I[d] = (I[d] + off + D.NT[d]) % D.NT[d];
    return D.idx(I[0], I[1], I[2]);
  };
  (void)s1;
  (void)s2;
  (void)s3;
  for (int64_t z = 0; z < n; z++) {
    const double* rho = prim + RHO * n;
    double rhoAvg = rho[wrap(z, 0, -1)] + rho[wrap(z, 0, 1)] + rho[wrap(z, 1, -1)] + rho[wrap(z, 1, 1)];
    if (D.dim > 2)
      rhoAvg = (rhoAvg + rho[wrap(z, 2, -1)] + rho[wrap(z, 2, 1)]) / 6.;
    else
      rhoAvg = rhoAvg / 4.;
    const double zero = rhoAvg * 0.;
    if (t.UseMADdisk) {
      const double Radius = gm.center[22 * n + z];
      Avec[z] = std::max(zero, rhoAvg * rhoAvg * std::pow(Radius, 4));
    } else {
      Avec[z] = std::max(rhoAvg - 0.2, zero) * std::cos(gm.center[23 * n + z] * (t.MagneticLoops - 1));
    }
  }
  const double dX1 = dXof(c, 0);
  for (int64_t z = 0; z < n; z++) {
    const double g = gm.center[15 * n + z];
    const double A00 = Avec[z], A01 = Avec[wrap(z, 1, 1)], A10 = Avec[wrap(z, 0, 1)];
    const double A11 = Avec[wrap(wrap(z, 0, 1), 1, 1)];
    double B1 = (A01 - A00 + A11 - A10) / (2. * dX2 * g);
    double B2 = (A00 - A10 + A01 - A11) / (2. * dX1 * g);
    int I[3] = {(int)(z % D.NT[0]), (int)((z / D.NT[0]) % D.NT[1]), (int)(z / ((int64_t)D.NT[0] * D.NT[1]))};
    const double mask = D.bulk(I[0], I[1], I[2]) ? 1. : 0.;
    prim[vars.B1 * n + z] = B1 * mask;
    prim[vars.B2 * n + z] = B2 * mask;
    prim[vars.B3 * n + z] = 0.;
  }
  communicate(c, D, nv, prim);
  double BFactor;
  if (t.UseMADdisk) {
    double PmagMax = -INFINITY, PgasMax = -INFINITY;
    for (int64_t z = 0; z < n; z++) {
      ZoneGeom<double> G;
      loadLocation(gm.center, n, z, &G);
      double P[GRIM_ORACLE_MAX_VARS];
      loadPrim(prim, nv, n, z, P);
      FluidElement<double> e;
      e.set(P, G, p, ZoneClosure{1., 1.});
      PmagMax = std::max(PmagMax, e.bSqr / 2.);
      PgasMax = std::max(PgasMax, e.pressure);
    }
    BFactor = std::sqrt(PgasMax / PmagMax / t.MinPlasmaBeta);
  } else {
    double BetaMin = INFINITY;
    for (int64_t z = 0; z < n; z++) {
      ZoneGeom<double> G;
      loadLocation(gm.center, n, z, &G);
      double P[GRIM_ORACLE_MAX_VARS];
      loadPrim(prim, nv, n, z, P);
      FluidElement<double> e;
      e.set(P, G, p, ZoneClosure{1., 1.});
      BetaMin = std::min(BetaMin, 2. * (e.pressure + 1.e-13) / (e.bSqr + 1.e-18));
    }
    BFactor = std::sqrt(BetaMin / t.MinPlasmaBeta);
  }
  for (int64_t z = 0; z < n; z++)
    for (int b = 0; b < 3; b++) prim[(vars.B1 + b) * n + z] *= BFactor;
  // Q and dP start at zero (set above); floors, then fullStepDiagnostics'
  // scrub and floors (problem.cpp:525-538).
  applyFloorAll(c, p, t, D, gm, prim);
  scrubNaNs(D, nv, prim);
  applyFloorAll(c, p, t, D, gm, prim);
}

}  // namespace
}  // namespace oracle

using namespace oracle;

extern "C" {

int64_t grim_oracle_grid_zones(const grim_oracle_grid* c) { return Dims(*c).n; }

void grim_oracle_grid_xcoords(const grim_oracle_grid* c, double* x3) {
  const Dims D(*c);
  const Coord C{*c};
  int lo[3] = {0, 0, 0}, hi[3] = {D.NT[0], D.NT[1], D.NT[2]};
  forRange(lo, hi, [&](int i, int j, int k) {
    double X[3], x[3];
    XCoordsAt(*c, D, i, j, k, LOC_CENTER, X);
    C.XCoordsToxCoords(X, x);
    const int64_t z = D.idx(i, j, k);
    for (int d = 0; d < 3; d++) x3[d * D.n + z] = x[d];
  });
}

void grim_oracle_build_geometry(const grim_oracle_grid* c, grim_oracle_geometry* gm) {
  const Dims D(*c);
  const Coord C{*c};
  const int64_t n = D.n;
  double* outs[3] = {gm->center, gm->faceX1, gm->faceX2};
  for (int loc = 0; loc < 3; loc++) {
    if (!outs[loc]) continue;
    int lo[3] = {0, 0, 0}, hi[3] = {D.NT[0], D.NT[1], D.NT[2]};
    forRange(lo, hi, [&](int i, int j, int k) {
      double X[3], x[3], G[4][4], Cn[4][4], det;
      XCoordsAt(*c, D, i, j, k, loc, X);
      C.XCoordsToxCoords(X, x);
      C.gCovAt(X, G);
      gDetAndgCon(G, &det, Cn);
      storeLocation(G, Cn, det, x[0], x[1], n, D.idx(i, j, k), outs[loc]);
    });
  }
  if (gm->gamma) {
    int lo[3] = {0, 0, 0}, hi[3] = {D.NT[0], D.NT[1], D.NT[2]};
    forRange(lo, hi, [&](int i, int j, int k) {
      // geometry::computeConnectionCoeffs, geometry.cpp:61-121
      const double EPS = 1.e-5;
      double X[3], G[4][4], Cn[4][4], det;
      XCoordsAt(*c, D, i, j, k, LOC_CENTER, X);
      C.gCovAt(X, G);
      gDetAndgCon(G, &det, Cn);
      double Gp[3][4][4], Gm[3][4][4];
      for (int d = 0; d < 3; d++) {
        double Xp[3] = {X[0], X[1], X[2]}, Xm[3] = {X[0], X[1], X[2]};
        Xp[d] += EPS;
        Xm[d] -= EPS;
        C.gCovAt(Xp, Gp[d]);
        C.gCovAt(Xm, Gm[d]);
      }
      double down[4][4][4];
      for (int eta = 0; eta < 4; eta++)
        for (int mu = 0; mu < 4; mu++)
          for (int nu = 0; nu < 4; nu++) {
            double out = 0.;
            if (nu > 0) {
              out += 0.5 * Gp[nu - 1][eta][mu] / (2. * EPS);
              out -= 0.5 * Gm[nu - 1][eta][mu] / (2. * EPS);
            }
            if (mu > 0) {
              out += 0.5 * Gp[mu - 1][eta][nu] / (2. * EPS);
              out -= 0.5 * Gm[mu - 1][eta][nu] / (2. * EPS);
            }
            if (eta > 0) {
              out -= 0.5 * Gp[eta - 1][mu][nu] / (2. * EPS);
              out += 0.5 * Gm[eta - 1][mu][nu] / (2. * EPS);
            }
            down[eta][mu][nu] = out;
          }
      const int64_t z = D.idx(i, j, k);
      for (int mu = 0; mu < 4; mu++)
        for (int nu = 0; nu < 4; nu++)
          for (int la = 0; la < 4; la++) {
            double s = 0.;
            for (int eta = 0; eta < 4; eta++) s += Cn[mu][eta] * down[eta][nu][la];
            gm->gamma[(mu * 16 + nu * 4 + la) * n + z] = s;
          }
    });
  }
}

void grim_oracle_apply_boundaries(const grim_oracle_grid* c, int nvars, double* prim) {
  const Dims D(*c);
  communicate(*c, D, nvars, prim);
  applyBCs(*c, D, nvars, prim);
}

void grim_oracle_reconstruct(const grim_oracle_grid* c, int nvars, const double* prim, int dir,
                             double* left, double* right) {
  const Dims D(*c);
  const int64_t s = D.stride(dir);
  int lo[3] = {0, 0, 0}, hi[3] = {D.NT[0], D.NT[1], D.NT[2]};
  lo[dir] = 2;
  hi[dir] = D.NT[dir] - 2;
  forRange(lo, hi, [&](int i, int j, int k) {
    const int64_t z = D.idx(i, j, k);
    for (int v = 0; v < nvars; v++) {
      double y[5];
      for (int q = 0; q < 5; q++) y[q] = prim[v * D.n + z + (q - 2) * s];
      reconstructZone(*c, y, &left[v * D.n + z], &right[v * D.n + z]);
    }
  });
}

void grim_oracle_riemann_flux(const grim_oracle_grid* c, const grim_oracle_params* p,
                              const grim_oracle_geometry* gm, const double* prim, int dir,
                              double* flux) {
  const Dims D(*c);
  riemannAll(*c, *p, D, *gm, prim, dir, flux);
}

void grim_oracle_div_fluxes(const grim_oracle_grid* c, const grim_oracle_params* p,
                            const grim_oracle_geometry* gm, const double* prim, double* divF,
                            double* fluxes, int flux_filter) {
  const Dims D(*c);
  divFluxes(*c, *p, D, *gm, prim, divF, fluxes, flux_filter != 0);
}

void grim_oracle_uniform(uint64_t seed, int64_t n, double* out) {
  for (int64_t z = 0; z < n; z++) out[z] = uniformAt(seed, (uint64_t)z);
}

void grim_oracle_explicit_sources(const grim_oracle_grid* c, const grim_oracle_params* p,
                                  const grim_oracle_geometry* gm, const double* prim,
                                  double* sources) {
  const Dims D(*c);
  explicitSources(*c, *p, D, *gm, prim, sources);
}

double grim_oracle_compute_dt(const grim_oracle_grid* c, const grim_oracle_params* p,
                              const grim_oracle_geometry* gm, const double* prim, double dt) {
  const Dims D(*c);
  return computeDt(*c, *p, D, *gm, prim, dt);
}

void grim_oracle_set_ghost_state(const double* prim_ic) { g_ghostState = prim_ic; }
void grim_oracle_set_x1_side_state(const double* prim_ic) { g_sideState = prim_ic; }

void grim_oracle_stage(const grim_oracle_grid* c, const grim_oracle_params* p,
                       const grim_oracle_geometry* gm, const grim_oracle_torus* t, int st,
                       double dt, double* primOld, double* primHalf, double* consOld,
                       double* primNew, int semantics, grim_oracle_step_stats* stats) {
  const Dims D(*c);
  stage(*c, *p, D, *gm, t, st, dt, primOld, primHalf, consOld, primNew, semantics, stats);
}

void grim_oracle_time_step(const grim_oracle_grid* c, const grim_oracle_params* p,
                           const grim_oracle_geometry* gm, const grim_oracle_torus* t,
                           double* primOld, double* primDt, double* dt, double* time,
                           int semantics, grim_oracle_step_stats* stats) {
  const Dims D(*c);
  const Vars vars(*p);
  const size_t sz = (size_t)vars.dof * D.n;
  *dt = computeDt(*c, *p, D, *gm, primDt, *dt);
  if (stats) stats->dt = *dt;
  std::vector<double> primHalf(sz), consOld(sz), primNew(sz);
  stage(*c, *p, D, *gm, t, 0, *dt, primOld, nullptr, consOld.data(), primHalf.data(), semantics,
        stats);
  stage(*c, *p, D, *gm, t, 1, *dt, primOld, primHalf.data(), consOld.data(), primNew.data(),
        semantics, stats);
  std::copy(primNew.begin(), primNew.end(), primOld);
  // elemOld of the next computeDt: the problem's fullStepDiagnostics set it
  // from this step's final primOld (linear_modes/problem.cpp:182,
  // torus/problem.cpp:743 via applyFloor).
  std::copy(primOld, primOld + sz, primDt);
  *time += *dt;
}

void grim_oracle_torus_ic(const grim_oracle_grid* c, const grim_oracle_params* p,
                          const grim_oracle_torus* t, const grim_oracle_geometry* gm,
                          double* prim) {
  const Dims D(*c);
  torusIC(*c, *p, *t, D, *gm, prim);
}

void grim_oracle_apply_floor(const grim_oracle_grid* c, const grim_oracle_params* p,
                             const grim_oracle_torus* t, const grim_oracle_geometry* gm,
                             double* prim) {
  const Dims D(*c);
  applyFloorAll(*c, *p, *t, D, *gm, prim);
}

void grim_oracle_full_step_diagnostics(const grim_oracle_grid* c, const grim_oracle_params* p,
                                       const grim_oracle_torus* t, const grim_oracle_geometry* gm,
                                       double* prim) {
  const Dims D(*c);
  const Vars vars(*p);
  scrubNaNs(D, vars.dof, prim);
  applyFloorAll(*c, *p, *t, D, *gm, prim);
}

void grim_oracle_torus_bcs(const grim_oracle_grid* c, const grim_oracle_params* p,
                           const grim_oracle_torus* t, const grim_oracle_geometry* gm,
                           double* prim) {
  const Dims D(*c);
  torusBCs(*c, *p, *t, D, *gm, prim);
}

// fullStepDiagnostics' WriteData variables (torus/problem.cpp:808-875) of the
// element set at prim, bulk zones in natural order (X1 fastest), planes
//   rho, u (floored), v^i = u^i/u^t and B^i = b^i u^t - b^t u^i in the
//   x coordinates, [q], [deltaP], b^2, the Lorentz factor u^t alpha_x
// (torus params.cpp:136-147), then 3 planes of xCoords (the VTS points).
// dxdX by geometry::computeTransformationMatrices (geometry.cpp:441-508):
// central differences of XCoordsToxCoords with GAMMA_EPS, dxdX[0][0] = 1 and
// the other time entries 0; conXTox (geometry.cpp:426-439) sums over all four.
void grim_oracle_dump_vars(const grim_oracle_grid* c, const grim_oracle_params* p,
                           const grim_oracle_geometry* gm, const double* prim, double* out) {
  const Dims D(*c);
  const Vars vars(*p);
  const Coord C{*c};
  const int64_t n = D.n, nb = (int64_t)D.N[0] * D.N[1] * D.N[2];
  const double eps = 1.e-5;  // GAMMA_EPS, geometry.cpp:11
  int64_t t = 0;
  for (int k = D.g[2]; k < D.g[2] + D.N[2]; k++)
    for (int j = D.g[1]; j < D.g[1] + D.N[1]; j++)
      for (int i = D.g[0]; i < D.g[0] + D.N[0]; i++, t++) {
        const int64_t z = D.idx(i, j, k);
        double X[3], x[3];
        XCoordsAt(*c, D, i, j, k, LOC_CENTER, X);
        C.XCoordsToxCoords(X, x);
        double dxdX[4][4] = {{1., 0., 0., 0.}, {0., 0., 0., 0.}, {0., 0., 0., 0.}, {0., 0., 0., 0.}};
        for (int a = 0; a < 3; a++)
          for (int b = 0; b < 3; b++) {
            double Xp[3] = {X[0], X[1], X[2]}, Xm[3] = {X[0], X[1], X[2]}, xp[3], xm[3];
            Xp[b] += eps;
            Xm[b] -= eps;
            C.XCoordsToxCoords(Xp, xp);
            C.XCoordsToxCoords(Xm, xm);
            dxdX[a + 1][b + 1] = (xp[a] - xm[a]) / (2. * eps);
          }
        double P[GRIM_ORACLE_MAX_VARS];
        loadPrim(prim, vars.dof, n, z, P);
        ZoneGeom<double> G;
        loadLocation(gm->center, n, z, &G);
        FluidElement<double> e;
        e.set(P, G, *p, ZoneClosure{1., 1.});
        double ux[4], bx[4];
        for (int mu = 0; mu < 4; mu++) {
          ux[mu] = 0.;
          bx[mu] = 0.;
          for (int nu = 0; nu < 4; nu++) {
            ux[mu] += dxdX[mu][nu] * e.uCon[nu];
            bx[mu] += dxdX[mu][nu] * e.bCon[nu];
          }
        }
        double gCon00x = 0.;
        for (int mu = 0; mu < 4; mu++)
          for (int nu = 0; nu < 4; nu++) gCon00x += dxdX[0][mu] * dxdX[0][nu] * G.gCon[mu][nu];
        const double alphax = 1. / std::sqrt(-gCon00x);
        std::vector<double> v;
        v.push_back(e.rho);
        v.push_back(e.u);
        for (int a = 1; a <= 3; a++) v.push_back(ux[a] / ux[0]);
        if (p->conduction) v.push_back(e.q);
        if (p->viscosity) v.push_back(e.deltaP);
        for (int a = 1; a <= 3; a++) v.push_back(bx[a] * ux[0] - bx[0] * ux[a]);
        v.push_back(e.bSqr);
        v.push_back(ux[0] * alphax);
        for (int d = 0; d < 3; d++) v.push_back(x[d]);
        for (size_t q = 0; q < v.size(); q++) out[(int64_t)q * nb + t] = v[q];
      }
}

// fullStepDiagnostics' observers (torus/problem.cpp:757-776) as written in
// src/problem/torus/ObserversTorus.hpp, on one rank: the element is set at
// prim (centre geometry); af::sum(af::flat(X(domain))) becomes a serial sum
// over the bulk in flat order (X1 fastest); af::max / af::min over the
// bulk become fmax / fmin from -inf / +inf.
void grim_oracle_torus_observers(const grim_oracle_grid* c, const grim_oracle_params* p,
                                 const grim_oracle_geometry* gm, const double* prim,
                                 double* out) {
  const Dims D(*c);
  const Vars vars(*p);
  const int64_t n = D.n;
  // volElem (problem.cpp:764-774)
  double volElem = dXof(*c, 0);
  if (c->dim > 1) volElem *= dXof(*c, 1);
  if (c->dim > 2) volElem *= dXof(*c, 2);
  double mass = 0., eMag = 0., eTh = 0., rhoMax = -INFINITY, betaMin = INFINITY;
  double mdotIn = 0., mdotOut = 0., unbound = 0., relativistic = 0.;
  const int iIn = D.g[0], iOut = D.g[0] + D.N[0] - 1;  // MassIntegrand(3, ...), (N1Local+2, ...)
  for (int k = D.g[2]; k < D.g[2] + D.N[2]; k++)
    for (int j = D.g[1]; j < D.g[1] + D.N[1]; j++)
      for (int i = D.g[0]; i < D.g[0] + D.N[0]; i++) {
        const int64_t z = D.idx(i, j, k);
        double P[GRIM_ORACLE_MAX_VARS];
        loadPrim(prim, vars.dof, n, z, P);
        ZoneGeom<double> G;
        loadLocation(gm->center, n, z, &G);
        FluidElement<double> e;
        e.set(P, G, *p, ZoneClosure{1., 1.});
        const double rho = P[0], u = P[1];
        // ComputeEnergyIntegrals (:17, :40, :62)
        mass += rho * volElem * G.g * e.gammaLorentzFactor / G.alpha;
        eMag += volElem * e.bSqr * G.g * e.gammaLorentzFactor / G.alpha;
        eTh += volElem * u * G.g * e.gammaLorentzFactor / G.alpha;
        // ComputeMinMaxVariables (:93, :115-118)
        rhoMax = std::fmax(rhoMax, rho);
        betaMin = std::fmin(betaMin, 2. * (e.pressure + 1.e-13) / (e.bSqr + 1.e-18));
        // ComputeBoundaryFluxes (:150-173)
        if (i == iIn || i == iOut) {
          const double mdot = rho * volElem * G.g * e.uCon[1];
          if (i == iIn) mdotIn += mdot;
          if (i == iOut) {
            const double bernoulli = (1. + u / rho * p->adiabaticIndex) * e.uCov[0];
            const double unb = mdot * (bernoulli < -1. ? 1. : 0.);
            mdotOut += mdot;
            unbound += unb;
            relativistic += unb * (e.gammaLorentzFactor > 2. ? 1. : 0.);
          }
        }
      }
  const double r[9] = {mass, eMag, eTh, mdotIn, mdotOut, unbound, relativistic, rhoMax, betaMin};
  std::copy(r, r + 9, out);
}

}  // extern "C"
