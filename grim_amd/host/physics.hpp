// grim's fluidElement and riemannSolver (src/physics/physics.hpp:14-133) on
// device grids, over the C ABI of include/grim_hip.h.
//
// grim's fluidElement materialises ~38 ArrayFire arrays in set(); here set()
// records the state and geometry and materialises the element's arrays
// (rho ... bCov, tau, chi_emhd, nu_emhd) into one device buffer of
// GRIM_ELEMENT_FIELDS planes, exposed under grim's member names.  The
// compute* methods evaluate their outputs from the recorded state.  numReads
// and numWrites, grim's array-access counters for its bandwidth report, are
// set to 0 (the kernels' traffic is measured with rocprofv3).
#pragma once

#include "geometry.hpp"
#include "grid.hpp"

namespace grim {

class fluidElement {
  double* fieldBuf = nullptr;
  int64_t n = 0;

 public:
  const grid* prim = nullptr;  // the state the element was set from
  geometry* geom = nullptr;

  // the element's arrays (device planes of the grid's zones)
  const double *rho, *u, *pressure, *temperature, *gammaLorentzFactor, *soundSpeed, *bSqr, *bNorm;
  const double *q, *deltaP, *tau, *chi_emhd, *nu_emhd;
  const double *uCon[4], *uCov[4], *bCon[4], *bCov[4];

  fluidElement(const grid& primIn, geometry& geomIn, int& numReads, int& numWrites) {
    set(primIn, geomIn, numReads, numWrites);
  }
  ~fluidElement() {
    if (fieldBuf) (void)hipFree(fieldBuf);
  }
  fluidElement(const fluidElement&) = delete;
  fluidElement& operator=(const fluidElement&) = delete;

  // fluidElement::set (physics.cpp:53-183), with setFluidElementParameters
  // (the context's closure) inside.
  void set(const grid& primIn, geometry& geomIn, int& numReads, int& numWrites) {
    if (primIn.ctx != geomIn.ctx) throw std::runtime_error("fluidElement::set: grid and geometry of two contexts");
    prim = &primIn;
    geom = &geomIn;
    if (!fieldBuf || n != primIn.zones()) {
      if (fieldBuf) (void)hipFree(fieldBuf);
      n = primIn.zones();
      checkHip(hipMalloc((void**)&fieldBuf, sizeof(double) * GRIM_ELEMENT_FIELDS * (size_t)n),
               "fluidElement: hipMalloc");
    }
    check(grim_element_fields(primIn.ctx, primIn.vars, location(), fieldBuf), "fluidElement::set");
    const double* f = fieldBuf;
    auto plane = [&](int q) { return f + (size_t)q * (size_t)n; };
    rho = plane(0);
    u = plane(1);
    pressure = plane(2);
    temperature = plane(3);
    gammaLorentzFactor = plane(4);
    soundSpeed = plane(5);
    bSqr = plane(6);
    bNorm = plane(7);
    q = plane(8);
    deltaP = plane(9);
    tau = plane(10);
    chi_emhd = plane(11);
    nu_emhd = plane(12);
    for (int m = 0; m < 4; m++) {
      uCon[m] = plane(13 + m);
      uCov[m] = plane(17 + m);
      bCon[m] = plane(21 + m);
      bCov[m] = plane(25 + m);
    }
    numReads = 0;
    numWrites = 0;
  }

  // fluidElement::computeFluxes (physics.cpp:185-217): direction 0 gives the
  // conserved variables, 1..3 the fluxes along X1..X3.
  void computeFluxes(const int direction, grid& flux, int& numReads, int& numWrites) {
    check(grim_element_fluxes(prim->ctx, prim->vars, location(), direction, flux.vars),
          "fluidElement::computeFluxes");
    numReads = 0;
    numWrites = 0;
  }

  // fluidElement::computeMinMaxCharSpeeds (riemannsolver.cpp:54-200), dir
  // 0..2 for X1..X3, into two device planes of the grid's zones.
  void computeMinMaxCharSpeeds(const int dir, double* MinSpeed, double* MaxSpeed, int& numReads, int& numWrites) {
    check(grim_element_char_speeds(prim->ctx, prim->vars, location(), dir, MinSpeed, MaxSpeed),
          "fluidElement::computeMinMaxCharSpeeds");
    numReads = 0;
    numWrites = 0;
  }

  // fluidElement::computeTimeDerivSources (physics.cpp:219-318): this
  // element supplies the coefficients (tau, chi, nu, b, ...), elemOld and
  // elemNew the states whose u_mu and T are differenced over dt.
  void computeTimeDerivSources(const fluidElement& elemOld, const fluidElement& elemNew, const double dt,
                               grid& sources, int& numReads, int& numWrites) {
    check(grim_element_time_deriv_sources(prim->ctx, prim->vars, elemOld.prim->vars, elemNew.prim->vars, dt,
                                          sources.vars),
          "fluidElement::computeTimeDerivSources");
    numReads = 0;
    numWrites = 0;
  }

  // fluidElement::computeImplicitSources (physics.cpp:320-358) with the
  // damping time tauDamp (a device plane of the grid's zones).
  void computeImplicitSources(grid& sources, const double* tauDamp, int& numReads, int& numWrites) {
    check(grim_element_implicit_sources(prim->ctx, prim->vars, tauDamp, sources.vars),
          "fluidElement::computeImplicitSources");
    numReads = 0;
    numWrites = 0;
  }

  // fluidElement::computeExplicitSources (physics.cpp:360-586): geometric
  // sources and the EMHD gradient terms at the zone centres, over the bulk.
  // dX must be the context's zone widths.
  void computeExplicitSources(const double dX[3], grid& sources, int& numReads, int& numWrites) {
    (void)dX;
    check(grim_explicit_sources(prim->ctx, prim->vars, sources.vars), "fluidElement::computeExplicitSources");
    numReads = 0;
    numWrites = 0;
  }

 private:
  int location() const {
    const int loc = geom->location;
    if (loc == CENTER || loc == FRONT || loc == BACK) return 0;
    if (loc == LEFT) return 1;
    if (loc == BOTTOM) return 2;
    throw std::runtime_error("fluidElement: set at a lower face (LEFT, BOTTOM) or the centre");
  }
};

class riemannSolver {
 public:
  // riemannSolver::riemannSolver (riemannsolver.cpp:3-52): grim allocates
  // its face elements and flux grids; the kernels keep them in registers.
  riemannSolver(const grid& prim, geometry& geom) {
    (void)prim;
    (void)geom;
  }

  // riemannSolver::solve (riemannsolver.cpp:202-328): the flux at the lower
  // face of each zone along dir from the reconstructed states, with the face
  // geometry grim passes for that direction (geomLeft/geomRight along X1,
  // geomBottom/geomTop along X2, geomCenter twice along X3).
  void solve(const grid& primLeft, const grid& primRight, geometry& geomLeft, geometry& geomRight, const int dir,
             grid& flux, int& numReads, int& numWrites) {
    const int want[3][2] = {{LEFT, RIGHT}, {BOTTOM, TOP}, {CENTER, CENTER}};
    if (dir < 0 || dir > 2 || geomLeft.location != want[dir][0] || geomRight.location != want[dir][1])
      throw std::runtime_error("riemannSolver::solve: geometries do not match the direction");
    check(grim_riemann_solve(primLeft.ctx, primLeft.vars, primRight.vars, dir, flux.vars),
          "riemannSolver::solve");
    numReads = 0;
    numWrites = 0;
  }
};

}  // namespace grim
