// grim's geometry (src/geometry/geometry.hpp:7-78) over the library's device
// arrays: the metric of one grid location as device planes of the grid's
// zones.
//
//   alpha, g             lapse and sqrt(-det g)
//   gCov[mu][nu]         g_{mu nu}          (symmetric; both index orders alias)
//   gCon[mu][nu]         g^{mu nu}
//   gammaUpDownDown[l][k][n]  Gamma^l_{k n} (zone centres only, else NULL)
//
// grim builds a geometry from a coordinates grid at one location
// (geometry.cpp:4-59); the standalone constructor does the same through a
// context of its own.  The library keeps one value per (X1, X2) column (the
// metric does not depend on X3); a geometry holds every zone's value, as
// grim's arrays do, in device planes of its own written from the context's
// (grim_get_geometry).  The RIGHT and TOP locations (the upper faces of each
// zone) are the LEFT and BOTTOM planes of the next zone along X1 / X2,
// shifted by one zone; the zones whose upper face does not exist in the
// array (the last ghost zone of each row, or of the last X2 row) hold a
// neighbour's face, which no bulk stencil reads.
#pragma once

#include <memory>

#include "grid.hpp"

namespace grim {

class geometry {
  grim_ctx* ownCtx = nullptr;  // standalone geometry: its own context
  double* own = nullptr;       // this location's planes, every zone
  double* conn = nullptr;      // CENTER: the connection's 64 planes, every zone
  int64_t n = 0;

  // upper-triangle packing of a symmetric 4x4 (00 01 02 03 11 12 13 22 23 33)
  // and 3x3 (11 12 13 22 23 33) index pair
  static int packed(int mu, int nu) {
    const int a = mu < nu ? mu : nu, b = mu < nu ? nu : mu;
    return a * 4 - (a * (a + 1)) / 2 + b;
  }
  static int packed3(int i, int j) {
    const int a = i < j ? i : j, b = i < j ? j : i;
    return a * 3 - (a * (a + 1)) / 2 + b;
  }

 public:
  int N1, N2, N3, dim, numGhost;
  int metric;
  double blackHoleSpin, hSlope;
  int location;
  grim_ctx* ctx;
  const double* planes;  // GRIM_LOC_PLANES planes of this location
  const double* alpha;
  const double* g;
  const double* gCov[4][4];
  const double* gCon[4][4];
  const double* gammaUpDownDown[4][4][4];

  // geometry::geometry (geometry.cpp:4-59) on its own context: the metric of
  // the coordinates grid's location, the connection at CENTER.
  geometry(const int metric, const double blackHoleSpin, const double hSlope, const coordinatesGrid& XCoordsGrid)
      : N1(XCoordsGrid.N1), N2(XCoordsGrid.N2), N3(XCoordsGrid.N3), dim(XCoordsGrid.dim),
        numGhost(XCoordsGrid.numGhost), metric(metric), blackHoleSpin(blackHoleSpin), hSlope(hSlope),
        location(XCoordsGrid.location), ctx(nullptr), planes(nullptr) {
    grim_params p;
    grim_default_params(&p, 0);
    grim_grid_config c{};
    c.N1 = N1;
    c.N2 = N2;
    c.N3 = N3;
    c.dim = dim;
    c.numGhost = numGhost;
    c.X1Start = XCoordsGrid.X1Start;
    c.X1End = XCoordsGrid.X1End;
    c.X2Start = XCoordsGrid.X2Start;
    c.X2End = XCoordsGrid.X2End;
    c.X3Start = XCoordsGrid.X3Start;
    c.X3End = XCoordsGrid.X3End;
    c.metric = metric;
    c.blackHoleSpin = blackHoleSpin;
    c.hSlope = hSlope;
    c.DerefineThetaHorizon = XCoordsGrid.DerefineThetaHorizon;
    c.DoCylindrify = XCoordsGrid.DoCylindrify;
    c.X1cyl = XCoordsGrid.X1cyl;
    c.X2cyl = XCoordsGrid.X2cyl;
    c.reconstruction = GRIM_RECON_WENO5;
    c.riemannSolver = GRIM_RIEMANN_LLF;
    c.slopeLimTheta = 2.;
    c.CourantFactor = 0.9;
    c.maxDtIncrement = 1.3;
    check(grim_create(&p, &c, nullptr, nullptr, &ownCtx), "geometry: grim_create");
    attach(ownCtx, location);
  }
  // The geometry a context holds at a location (timeStepper's geomLeft,
  // geomRight, geomBottom, geomTop, geomCenter; timestepper.cpp:236-280).
  geometry(grim_ctx* context, const grim_grid_config& c, int loc)
      : N1(c.N1), N2(c.N2), N3(c.N3), dim(c.dim), numGhost(c.numGhost), metric(c.metric),
        blackHoleSpin(c.blackHoleSpin), hSlope(c.hSlope), location(loc), ctx(nullptr), planes(nullptr) {
    attach(context, loc);
  }
  ~geometry() {
    if (own) (void)hipFree(own);
    if (conn) (void)hipFree(conn);
    if (ownCtx) grim_destroy(ownCtx);
  }
  geometry(const geometry&) = delete;
  geometry& operator=(const geometry&) = delete;

  // geometry::computeConnectionCoeffs (geometry.cpp:61-121): built with the
  // geometry; this only checks it exists at this location.
  void computeConnectionCoeffs() {
    if (!gammaUpDownDown[0][0][0]) throw std::runtime_error("geometry: the connection exists at CENTER only");
  }

  // Host copy of one plane (grim's gCovGrid / gConGrid / ... accessors of
  // the Cython layer, geometry.hpp:64-77).
  std::vector<double> hostPlane(const double* plane) const {
    std::vector<double> out((size_t)n);
    checkHip(hipMemcpy(out.data(), plane, sizeof(double) * (size_t)n, hipMemcpyDeviceToHost), "geometry: copy");
    return out;
  }

 private:
  void attach(grim_ctx* context, int loc) {
    ctx = context;
    grim_grid_config c{};
    c.N1 = N1;
    c.N2 = N2;
    c.N3 = N3;
    c.dim = dim;
    c.numGhost = numGhost;
    n = grim_grid_zones(&c);
    const int which = (loc == CENTER || loc == FRONT || loc == BACK) ? 0 : ((loc == LEFT || loc == RIGHT) ? 1 : 2);
    if (which == 2 && dim < 2) throw std::runtime_error("geometry: no X2 faces on a 1D grid");
    const size_t locBytes = sizeof(double) * GRIM_LOC_PLANES * (size_t)n;
    checkHip(hipMalloc((void**)&own, locBytes), "geometry: hipMalloc");
    check(grim_get_geometry(ctx, which, own, 1), "geometry: grim_get_geometry");
    if (loc == RIGHT || loc == TOP) {
      // face i+1/2 of zone i is face (i+1)-1/2 of zone i+1
      const int64_t s = loc == RIGHT ? 1 : (int64_t)(N1 + 2 * numGhost);
      double* every = own;
      checkHip(hipMalloc((void**)&own, locBytes), "geometry: hipMalloc");
      for (int q = 0; q < GRIM_LOC_PLANES; q++) {
        const double* src = every + (size_t)q * n;
        double* dst = own + (size_t)q * n;
        checkHip(hipMemcpy(dst, src + s, sizeof(double) * (size_t)(n - s), hipMemcpyDeviceToDevice),
                 "geometry: shift");
        checkHip(hipMemcpy(dst + (n - s), src + (n - s), sizeof(double) * (size_t)s, hipMemcpyDeviceToDevice),
                 "geometry: shift (last row)");
      }
      (void)hipFree(every);
    }
    planes = own;
    alpha = planes + (size_t)14 * n;
    g = planes + (size_t)15 * n;
    for (int mu = 0; mu < 4; mu++)
      for (int nu = 0; nu < 4; nu++) {
        gCov[mu][nu] = planes + (size_t)packed(mu, nu) * n;
        if (mu == 0 || nu == 0)
          gCon[mu][nu] = planes + (size_t)(10 + (mu == 0 ? nu : mu)) * n;
        else
          gCon[mu][nu] = planes + (size_t)(16 + packed3(mu - 1, nu - 1)) * n;
      }
    if (which == 0) {
      checkHip(hipMalloc((void**)&conn, sizeof(double) * 64 * (size_t)n), "geometry: hipMalloc");
      check(grim_get_geometry(ctx, 3, conn, 1), "geometry: grim_get_geometry (connection)");
    }
    for (int l = 0; l < 4; l++)
      for (int k = 0; k < 4; k++)
        for (int m = 0; m < 4; m++)
          gammaUpDownDown[l][k][m] = conn ? conn + (size_t)(l * 16 + k * 4 + m) * n : nullptr;
  }
};

}  // namespace grim
