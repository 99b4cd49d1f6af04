// grim oracle: CPU FP64 restatement of grim's per-zone implicit-step algorithm.
//
// TEST INFRASTRUCTURE ONLY.  This code is the checker for the HIP library in
// grim_amd/: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
// leg may load it.  The product path never calls it.
//
// Every routine restates the reference line by line (citations are
// file:line in the grim source tree) and keeps grim's evaluation order so
// that the CPU and GPU results can be compared to FP64 rounding.  The file is
// built with -ffp-contract=off so that no multiply-add is fused.
//
// All per-zone routines are templated on the scalar type T so the same code
// instantiates `double`, a flop-counting type and a forward-mode dual number
// (the latter gives exact Jacobians against which analytic tangents are
// pinned).
//
// Pinning status of each part is recorded in the headers of
// tests/test_oracle.py (per-zone routines) and tests/test_step_oracle.py
// (the step, problems and observers).  Parity unpinned (no check beyond the
// restatement itself): only the ArrayFire boundary -- af::solve's pivoting
// and the summation order of af::sum (the observers' sums; the LU here is
// LAPACK's partial pivoting) -- which no reference test exercises.
#pragma once
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>

#include "grim_oracle.h"

namespace oracle {

// ---------------------------------------------------------------------------
// Scalar helpers.  Overloads for plain double; the special scalar types in
// scalars.hpp provide their own overloads found by ADL.
// ---------------------------------------------------------------------------
inline double oval(double x) { return x; }
// af::max(a, b) / af::min(a, b) with a scalar b: elementwise select.
inline double amax(double a, double b) { return a > b ? a : b; }
inline double amin(double a, double b) { return a < b ? a : b; }

// Variable layout (src/params.hpp:8-18; src/problem/linear_modes/params.cpp:84-92).
struct Vars {
  int Q, DP, numFluidVars, B1, B2, B3, dof;
  explicit Vars(const grim_oracle_params& p) {
    Q = 5;
    DP = 5 + p.conduction;
    numFluidVars = 5 + p.conduction + p.viscosity;
    B1 = 5 + p.conduction + p.viscosity;
    B2 = 6 + p.conduction + p.viscosity;
    B3 = 7 + p.conduction + p.viscosity;
    dof = 8 + p.conduction + p.viscosity;
  }
};
enum { RHO = 0, UU = 1, U1 = 2, U2 = 3, U3 = 4 };

// Geometry of one zone at the zone centre (src/geometry/geometry.hpp:7-78):
// gCov, gCon, g = sqrt(-det g), alpha = 1/sqrt(-g^00), and the radius r that
// the torus closure reads through geom->getxCoords.
template <class T>
struct ZoneGeom {
  T gCov[4][4];
  T gCon[4][4];
  T g;
  T alpha;
  T r;
};

// Per-zone closure inputs that are not primitives (conduction / viscosity
// alpha when they vary per zone).
struct ZoneClosure {
  double conductionAlpha;
  double viscosityAlpha;
};

// ---------------------------------------------------------------------------
// fluidElement (src/physics/physics.cpp, src/physics/physics.hpp:14-108)
// ---------------------------------------------------------------------------
template <class T>
struct FluidElement {
  const ZoneGeom<T>* geom = nullptr;
  T rho, u, u1, u2, u3, B1, B2, B3;
  T pressure, temperature, soundSpeed, gammaLorentzFactor;
  T uCon[4], uCov[4], bCon[4], bCov[4];
  T bSqr, bNorm;
  T qTilde, q, deltaPTilde, deltaP;
  T tau, chi_emhd, nu_emhd;
  T NUp[4];
  T TUpDown[4][4];

  // Closure hook setFluidElementParameters (physics.hpp:58).
  //  LINEAR_MODES: src/problem/linear_modes/problem.cpp:3-8
  //  PER_ZONE:     same form with per-zone alphas (chi = a_c cs^2 tau)
  //  TORUS:        src/problem/torus/problem.cpp:7-55
  void setFluidElementParameters(const grim_oracle_params& p,
                                 const ZoneClosure& zc) {
    using std::exp;
    using std::pow;
    using std::abs;
    if (p.closure == GRIM_ORACLE_CLOSURE_LINEAR_MODES) {
      tau = T(1.0);
      chi_emhd = soundSpeed * soundSpeed * tau;
      nu_emhd = soundSpeed * soundSpeed * tau;
    } else if (p.closure == GRIM_ORACLE_CLOSURE_PER_ZONE) {
      tau = T(p.tau0);
      chi_emhd = zc.conductionAlpha * soundSpeed * soundSpeed * tau;
      nu_emhd = zc.viscosityAlpha * soundSpeed * soundSpeed * tau;
    } else if (p.closure == GRIM_ORACLE_CLOSURE_SHOCK_TESTS) {
      // src/problem/shock_tests/problem.cpp:13-18
      tau = T(0.1);
      chi_emhd = p.ConductionAlpha * soundSpeed * soundSpeed * tau;
      nu_emhd = p.ViscosityAlpha * soundSpeed * soundSpeed * tau;
    } else if (p.closure == GRIM_ORACLE_CLOSURE_BUOYANCY) {
      // src/problem/buoyancy_instabilities/problem.cpp:11-21: af::pow(r,
      // 0.5) evaluated as sqrt(r) (correctly rounded; a library pow need not
      // be); the problem sets no nu_emhd and runs without viscosity
      chi_emhd = sqrt(geom->r);
      tau = 1.2 * chi_emhd / (soundSpeed * soundSpeed);
      nu_emhd = T(0.);
    } else {
      // torus closure, src/problem/torus/problem.cpp:7-55.  pow(r, 1.5) and
      // pow(c_s, 3) are evaluated as r sqrt(r) and c_s c_s c_s (equal up to
      // the last bit; sqrt and products round identically on host and device,
      // a library pow does not).
      T Radius = geom->r;
      T DynamicalTimescale = Radius * sqrt(Radius);
      tau = DynamicalTimescale;
      if (p.conduction) {
        T Qmax = p.ConductionClosureFactor * rho * (soundSpeed * soundSpeed * soundSpeed);
        double lambda = 0.01;
        T yCon = abs(q) / Qmax;
        yCon = exp(-(yCon - 1.) / lambda);
        T fdCon = yCon / (yCon + 1.) + 1.e-5;
        tau = amin(tau, DynamicalTimescale * fdCon);
      }
      if (p.viscosity) {
        T dPmod = amax(pressure - 2. / 3. * deltaP, p.bSqrFloorInFluidElement) /
                  amax(pressure + 1. / 3. * deltaP, p.bSqrFloorInFluidElement);
        T dPmaxPlus = amin(p.ViscosityClosureFactor * bSqr * 0.5 * dPmod,
                           1.49 * pressure / 1.07);
        T dPmaxMinus = amax(-p.ViscosityClosureFactor * bSqr,
                            -2.99 * pressure / 1.07);
        T dPmax = (oval(deltaP) > 0.) ? dPmaxPlus : dPmaxMinus;
        double lambda = 0.01;
        T yVis = abs(deltaP) / (abs(dPmax) + p.bSqrFloorInFluidElement);
        yVis = exp(-(yVis - 1.) / lambda);
        T fdVis = yVis / (yVis + 1.) + 1.e-5;
        tau = amin(tau, DynamicalTimescale * fdVis);
      }
      T maxAlpha = (1. - soundSpeed * soundSpeed) /
                   (2. * soundSpeed * soundSpeed + 1.e-12);
      chi_emhd = amin(maxAlpha, p.ConductionAlpha) * soundSpeed * soundSpeed * tau;
      nu_emhd = amin(maxAlpha, p.ViscosityAlpha) * soundSpeed * soundSpeed * tau;
    }
  }

  // fluidElement::set, src/physics/physics.cpp:53-183
  void set(const T* prim, const ZoneGeom<T>& geom_, const grim_oracle_params& p,
           const ZoneClosure& zc) {
    using std::sqrt;
    const Vars vars(p);
    geom = &geom_;
    rho = amax(prim[RHO], p.rhoFloorInFluidElement);
    u = amax(prim[UU], p.uFloorInFluidElement);
    u1 = prim[U1];
    u2 = prim[U2];
    u3 = prim[U3];
    B1 = prim[vars.B1];
    B2 = prim[vars.B2];
    B3 = prim[vars.B3];

    pressure = (p.adiabaticIndex - 1.) * u;
    temperature = amax(pressure / rho, p.temperatureFloorInFluidElement);

    soundSpeed = sqrt(p.adiabaticIndex * pressure / (rho + p.adiabaticIndex * u));

    gammaLorentzFactor =
        sqrt(1. + geom->gCov[1][1] * u1 * u1 + geom->gCov[2][2] * u2 * u2 +
             geom->gCov[3][3] * u3 * u3 +
             2. * (geom->gCov[1][2] * u1 * u2 + geom->gCov[1][3] * u1 * u3 +
                   geom->gCov[2][3] * u2 * u3));

    uCon[0] = gammaLorentzFactor / geom->alpha;
    uCon[1] = u1 - gammaLorentzFactor * geom->gCon[0][1] * geom->alpha;
    uCon[2] = u2 - gammaLorentzFactor * geom->gCon[0][2] * geom->alpha;
    uCon[3] = u3 - gammaLorentzFactor * geom->gCon[0][3] * geom->alpha;

    for (int mu = 0; mu < 4; mu++) {
      uCov[mu] = geom->gCov[mu][0] * uCon[0] + geom->gCov[mu][1] * uCon[1] +
                 geom->gCov[mu][2] * uCon[2] + geom->gCov[mu][3] * uCon[3];
    }

    bCon[0] = B1 * uCov[1] + B2 * uCov[2] + B3 * uCov[3];
    bCon[1] = (B1 + bCon[0] * uCon[1]) / uCon[0];
    bCon[2] = (B2 + bCon[0] * uCon[2]) / uCon[0];
    bCon[3] = (B3 + bCon[0] * uCon[3]) / uCon[0];

    for (int mu = 0; mu < 4; mu++) {
      bCov[mu] = geom->gCov[mu][0] * bCon[0] + geom->gCov[mu][1] * bCon[1] +
                 geom->gCov[mu][2] * bCon[2] + geom->gCov[mu][3] * bCon[3];
    }

    bSqr = bCon[0] * bCov[0] + bCon[1] * bCov[1] + bCon[2] * bCov[2] +
           bCon[3] * bCov[3] + p.bSqrFloorInFluidElement;
    bNorm = sqrt(bSqr);

    const double condAlpha = (p.closure == GRIM_ORACLE_CLOSURE_PER_ZONE)
                                 ? zc.conductionAlpha
                                 : p.ConductionAlpha;
    const double viscAlpha = (p.closure == GRIM_ORACLE_CLOSURE_PER_ZONE)
                                 ? zc.viscosityAlpha
                                 : p.ViscosityAlpha;
    if (p.conduction == 1) {
      qTilde = prim[vars.Q];
      if (p.highOrderTermsConduction == 1)
        q = qTilde * temperature * sqrt(rho * condAlpha * soundSpeed * soundSpeed);
      else
        q = qTilde;
    } else {
      qTilde = T(0.);
      q = T(0.);
    }
    if (p.viscosity == 1) {
      deltaPTilde = prim[vars.DP];
      if (p.highOrderTermsViscosity == 1)
        deltaP = deltaPTilde *
                 sqrt(temperature * rho * viscAlpha * soundSpeed * soundSpeed);
      else
        deltaP = deltaPTilde;
    } else {
      deltaPTilde = T(0.);
      deltaP = T(0.);
    }

    setFluidElementParameters(p, zc);

    for (int mu = 0; mu < 4; mu++) {
      NUp[mu] = rho * uCon[mu];
      for (int nu = 0; nu < 4; nu++) {
        const double delta = (mu == nu) ? 1. : 0.;
        TUpDown[mu][nu] = (rho + u + pressure + bSqr) * uCon[mu] * uCov[nu] +
                          (pressure + 0.5 * bSqr) * delta - bCon[mu] * bCov[nu];
        if (p.conduction == 1)
          TUpDown[mu][nu] +=
              q / bNorm * (uCon[mu] * bCov[nu] + bCon[mu] * uCov[nu]);
        if (p.viscosity == 1)
          TUpDown[mu][nu] +=
              (-deltaP) * (bCon[mu] * bCov[nu] / bSqr -
                           (1. / 3.) * (delta + uCon[mu] * uCov[nu]));
      }
    }
  }

  // fluidElement::computeFluxes, src/physics/physics.cpp:185-217
  void computeFluxes(int dir, T* flux, const grim_oracle_params& p) const {
    const Vars vars(p);
    T g = geom->g;
    flux[RHO] = g * NUp[dir];
    flux[UU] = g * TUpDown[dir][0] + flux[RHO];
    flux[U1] = g * TUpDown[dir][1];
    flux[U2] = g * TUpDown[dir][2];
    flux[U3] = g * TUpDown[dir][3];
    flux[vars.B1] = g * (bCon[1] * uCon[dir] - bCon[dir] * uCon[1]);
    flux[vars.B2] = g * (bCon[2] * uCon[dir] - bCon[dir] * uCon[2]);
    flux[vars.B3] = g * (bCon[3] * uCon[dir] - bCon[dir] * uCon[3]);
    if (p.conduction) flux[vars.Q] = g * (uCon[dir] * qTilde);
    if (p.viscosity) flux[vars.DP] = g * (uCon[dir] * deltaPTilde);
  }

  // fluidElement::computeTimeDerivSources, src/physics/physics.cpp:219-318.
  // `this` supplies the coefficients (elemOld or elemHalfStep in
  // src/timestepper/residual.cpp:28,106); elemOld/elemNew supply u_mu and T.
  void computeTimeDerivSources(const FluidElement& elemOld,
                               const FluidElement& elemNew, double dt,
                               T* sources, const grim_oracle_params& p) const {
    using std::sqrt;
    const Vars vars(p);
    for (int var = 0; var < vars.dof; var++) sources[var] = T(0.);
    if (!(p.conduction || p.viscosity)) return;
    T dtuCov[4];
    for (int mu = 0; mu < 4; mu++)
      dtuCov[mu] = (elemNew.uCov[mu] - elemOld.uCov[mu]) / dt;
    T divuCov = T(0.);
    for (int mu = 0; mu < 4; mu++) divuCov += geom->gCon[0][mu] * dtuCov[mu];

    if (p.viscosity) {
      T deltaP0 = -divuCov * rho * nu_emhd;
      for (int mu = 0; mu < 4; mu++)
        deltaP0 += 3. * rho * nu_emhd * bCon[0] * bCon[mu] / bSqr * dtuCov[mu];
      if (p.highOrderTermsViscosity == 1)
        deltaP0 *= sqrt(tau / rho / nu_emhd / temperature);
      sources[vars.DP] = -geom->g * (deltaP0) / tau;
      if (p.highOrderTermsViscosity == 1)
        sources[vars.DP] -= 0.5 * geom->g * divuCov * deltaPTilde;
    }
    if (p.conduction) {
      T q0 = -rho * chi_emhd * bCon[0] / bNorm *
             (elemNew.temperature - elemOld.temperature) / dt;
      for (int nu = 0; nu < 4; nu++)
        q0 -= rho * chi_emhd * temperature * bCon[nu] / bNorm * uCon[0] *
              dtuCov[nu];
      if (p.highOrderTermsConduction == 1)
        q0 *= sqrt(tau / rho / chi_emhd) / temperature;
      sources[vars.Q] = -geom->g * (q0) / tau;
      if (p.highOrderTermsConduction == 1)
        sources[vars.Q] -= 0.5 * geom->g * divuCov * qTilde;
    }
  }

  // fluidElement::computeImplicitSources, src/physics/physics.cpp:320-358
  void computeImplicitSources(T* sources, const T& tauDamp,
                              const grim_oracle_params& p) const {
    const Vars vars(p);
    for (int var = 0; var < vars.dof; var++) sources[var] = T(0.);
    if (!(p.conduction || p.viscosity)) return;
    if (p.viscosity) sources[vars.DP] = geom->g * (deltaPTilde) / tauDamp;
    if (p.conduction) sources[vars.Q] = geom->g * (qTilde) / tauDamp;
  }

  // fluidElement::computeMinMaxCharSpeeds, src/physics/riemannsolver.cpp:54-200.
  // dir = 0, 1, 2 for X1, X2, X3.  Needs the full gCon.
  void computeMinMaxCharSpeeds(int dir, T* minSpeed, T* maxSpeed,
                               const grim_oracle_params& p) const {
    using std::sqrt;
    T cAlvenSqr = bSqr / (rho + p.adiabaticIndex * u + bSqr);
    T csSqr = soundSpeed * soundSpeed;
    T cVisSqr = T(0.);
    if (p.viscosity) cVisSqr = 4. / 3. / (rho + p.adiabaticIndex * u) * rho * nu_emhd / tau;
    T cConSqr = T(0.);
    if (p.conduction) cConSqr = (p.adiabaticIndex - 1.) * chi_emhd / tau;
    cConSqr = 0.5 * (csSqr + cConSqr + sqrt(csSqr * csSqr + cConSqr * cConSqr));
    csSqr = cConSqr + cVisSqr;
    csSqr = csSqr + cAlvenSqr - csSqr * cAlvenSqr;
    {
      const double condition = oval(csSqr) > 1. ? 1. : 0.;
      csSqr = csSqr * (1. - condition) + condition;
    }
    const int sdir = dir + 1;
    T ACov[4], ACon[4], BCov[4], BCon[4];
    for (int mu = 0; mu < 4; mu++) {
      ACov[mu] = T(0.);
      BCov[mu] = T(0.);
    }
    ACov[sdir] = T(1.);
    BCov[0] = T(1.);
    for (int mu = 0; mu < 4; mu++) {
      ACon[mu] = T(0.);
      BCon[mu] = T(0.);
      for (int nu = 0; nu < 4; nu++) {
        ACon[mu] += geom->gCon[mu][nu] * ACov[nu];
        BCon[mu] += geom->gCon[mu][nu] * BCov[nu];
      }
    }
    T ASqr = T(0.), BSqr = T(0.), ADotU = T(0.), BDotU = T(0.), ADotB = T(0.);
    for (int mu = 0; mu < 4; mu++) {
      ASqr += ACov[mu] * ACon[mu];
      BSqr += BCov[mu] * BCon[mu];
      ADotU += ACov[mu] * uCon[mu];
      BDotU += BCov[mu] * uCon[mu];
      ADotB += ACov[mu] * BCon[mu];
    }
    T A = (BDotU * BDotU) - (BSqr + BDotU * BDotU) * csSqr;
    T B = 2. * (ADotU * BDotU - (ADotB + ADotU * BDotU) * csSqr);
    T C = ADotU * ADotU - (ASqr + ADotU * ADotU) * csSqr;
    T discr = sqrt(amax(B * B - 4. * A * C, 1.e-16));
    T mn = -(-B + discr) / 2. / A;
    T mx = -(-B - discr) / 2. / A;
    {
      const double condition = oval(mn) > -1.e-15 ? 1. : 0.;
      mn = mn * (1 - condition) - 1.e-15 * condition;
    }
    {
      const double condition = oval(mx) < 1.e-15 ? 1. : 0.;
      mx = mx * (1 - condition) + 1.e-15 * condition;
    }
    *minSpeed = mn;
    *maxSpeed = mx;
  }
};

// ---------------------------------------------------------------------------
// Frozen per-zone pieces of the residual (everything computed once per stage
// before solve(), src/timestepper/timestep.cpp:28-66 and 183-220).
// ---------------------------------------------------------------------------
struct ZoneFrozen {
  double consOld[GRIM_ORACLE_MAX_VARS];
  double divFluxes[GRIM_ORACLE_MAX_VARS];
  double sourcesExplicit[GRIM_ORACLE_MAX_VARS];
  double sourcesImplicitOld[GRIM_ORACLE_MAX_VARS];
};

// Everything computeResidual reads for one zone, apart from the guess.
template <class T>
struct ResidualContext {
  const grim_oracle_params* p;
  const ZoneGeom<T>* geom;
  ZoneClosure closure;
  // Either the literal frozen pieces (frozen != nullptr) or the batched
  // form with one folded right-hand side per row (rhs != nullptr):
  //   R = (U(P) - rhs)/dtPrime + 1/2 S_impl(P) + S_td(P),
  // which equals the literal form with consOld = rhs and the other frozen
  // terms zero (adding an exact zero leaves every rounding unchanged).
  const ZoneFrozen* frozen;
  const double* rhs;
  double dtPrime;
  double residualMask;  // residual.cpp:174-177 (1 in the bulk, 0 in ghosts)  // dt/2 on the half step, dt on the full step
  // elemOld and the coefficient element (elemOld on the half step,
  // elemHalfStep on the full step), already set().
  const FluidElement<T>* elemOld;
  const FluidElement<T>* elemCoef;
};

// timeStepper::computeResidual for one zone, src/timestepper/residual.cpp:3-180.
// Writes numFluidVars residual rows.
template <class T>
void computeResidual(const T* primGuess, const ResidualContext<T>& c, T* R) {
  using std::sqrt;
  const grim_oracle_params& p = *c.p;
  const Vars vars(p);
  FluidElement<T> elem;
  elem.set(primGuess, *c.geom, p, c.closure);
  T cons[GRIM_ORACLE_MAX_VARS];
  elem.computeFluxes(0, cons, p);

  T sourcesImplicit[GRIM_ORACLE_MAX_VARS];
  T sourcesTimeDer[GRIM_ORACLE_MAX_VARS];
  const bool emhd = p.conduction || p.viscosity;
  if (emhd) {
    elem.computeImplicitSources(sourcesImplicit, c.elemCoef->tau, p);
    c.elemCoef->computeTimeDerivSources(*c.elemOld, elem, c.dtPrime,
                                        sourcesTimeDer, p);
  } else {
    for (int v = 0; v < vars.dof; v++) {
      sourcesImplicit[v] = T(0.);
      sourcesTimeDer[v] = T(0.);
    }
  }
  if (c.frozen) {
    const ZoneFrozen& f = *c.frozen;
    for (int var = 0; var < vars.numFluidVars; var++) {
      R[var] = (cons[var] - f.consOld[var]) / c.dtPrime + f.divFluxes[var] +
               f.sourcesExplicit[var] +
               0.5 * (f.sourcesImplicitOld[var] + sourcesImplicit[var]) +
               sourcesTimeDer[var];
    }
  } else if (emhd) {
    for (int var = 0; var < vars.numFluidVars; var++)
      R[var] = (cons[var] - c.rhs[var]) / c.dtPrime +
               0.5 * sourcesImplicit[var] + sourcesTimeDer[var];
  } else {
    for (int var = 0; var < vars.numFluidVars; var++)
      R[var] = (cons[var] - c.rhs[var]) / c.dtPrime;
  }
  const FluidElement<T>& e = *c.elemCoef;
  if (p.conduction) {
    if (p.highOrderTermsConduction)
      R[vars.Q] *= e.temperature * sqrt(e.rho * e.chi_emhd * e.tau);
    else
      R[vars.Q] *= e.tau;
  }
  if (p.viscosity) {
    if (p.highOrderTermsViscosity)
      R[vars.DP] *= sqrt(e.rho * e.nu_emhd * e.temperature * e.tau);
    else
      R[vars.DP] *= e.tau;
  }
  // Zero the residual in global ghost zones, residual.cpp:173-177.
  for (int var = 0; var < vars.numFluidVars; var++) R[var] *= c.residualMask;
}

// Sum of squared residual rows (src/timestepper/solve.cpp:36-37).
template <class T>
T residualL2(const T* R, int n) {
  T s = R[0] * R[0];
  for (int v = 1; v < n; v++) s = s + R[v] * R[v];
  return s;
}

// FD Jacobian assembly, src/timestepper/solve.cpp:82-109.
// A is row-major nf x nf with A[i][j] = dR_i / dP_j (grim stores it as
// jacobianSoA(column + nf*row), i.e. column-major with column = residual
// row, then solves with A(i,j) = dR_i/dP_j).
template <class T>
void fdJacobian(const T* primGuess, const T* R, const ResidualContext<T>& c,
                T* A) {
  const grim_oracle_params& p = *c.p;
  const Vars vars(p);
  const int nf = vars.numFluidVars;
  T primPlusEps[GRIM_ORACLE_MAX_VARS];
  T Rplus[GRIM_ORACLE_MAX_VARS];
  for (int v = 0; v < vars.dof; v++) primPlusEps[v] = primGuess[v];
  for (int row = 0; row < nf; row++) {
    const double epsilon = p.JacobianAssembleEpsilon;
    const bool small = std::fabs(oval(primGuess[row])) < .5 * epsilon;
    // primGuess + epsilon*primGuess*(1-small) + small*epsilon
    primPlusEps[row] = small ? primGuess[row] + epsilon
                             : primGuess[row] + epsilon * primGuess[row];
    computeResidual(primPlusEps, c, Rplus);
    for (int column = 0; column < nf; column++)
      A[column * nf + row] =
          (Rplus[column] - R[column]) / (primPlusEps[row] - primGuess[row]);
    primPlusEps[row] = primGuess[row];
  }
}

// Dense LU with partial pivoting, solving A x = b in place (A row-major n x n).
// Stands in for af::solve (src/timestepper/solve.cpp:203-230), which calls a
// LAPACK-style getrf/getrs.  Pivot = first row holding the largest |A(i,k)|
// (LAPACK's idamax rule); multipliers use the reciprocal of the pivot as
// dgetf2 does.  Returns 0, or k+1 when the k-th pivot is exactly zero.
template <class T>
int luSolve(T* A, T* b, int n) {
  using std::fabs;
  int info = 0;
  T rinv[GRIM_ORACLE_MAX_VARS];
  for (int k = 0; k < n; k++) {
    int piv = k;
    double maxv = fabs(oval(A[k * n + k]));
    for (int i = k + 1; i < n; i++) {
      double v = fabs(oval(A[i * n + k]));
      if (v > maxv) {
        maxv = v;
        piv = i;
      }
    }
    if (piv != k) {
      for (int j = 0; j < n; j++) std::swap(A[k * n + j], A[piv * n + j]);
      std::swap(b[k], b[piv]);
    }
    if (maxv == 0. && info == 0) info = k + 1;
    rinv[k] = 1. / A[k * n + k];
    for (int i = k + 1; i < n; i++) {
      T m = A[i * n + k] * rinv[k];
      A[i * n + k] = m;
      for (int j = k + 1; j < n; j++) A[i * n + j] = A[i * n + j] - m * A[k * n + j];
      b[i] = b[i] - m * b[k];
    }
  }
  for (int i = n - 1; i >= 0; i--) {
    T s = b[i];
    for (int j = i + 1; j < n; j++) s = s - A[i * n + j] * b[j];
    b[i] = s * rinv[i];
  }
  return info;
}

struct NewtonZoneStats {
  int iterations;       // Newton updates applied
  int residualEvals;    // residual evaluations performed
  int lineSearchEvals;  // of which inside the line search
  double finalL2;       // sum of squared residual rows at the returned state
  int converged;
};

// One Newton iteration's update step for one zone (everything in
// src/timestepper/solve.cpp:80-198 after the convergence test), given the
// residual R at primGuess and l2 = sum R^2.  Updates primGuess in place.
template <class T>
void newtonUpdate(T* primGuess, const T* R, const T& l2,
                  const ResidualContext<T>& c, NewtonZoneStats* st) {
  const grim_oracle_params& p = *c.p;
  const Vars vars(p);
  const int nf = vars.numFluidVars;
  T A[GRIM_ORACLE_MAX_VARS * GRIM_ORACLE_MAX_VARS];
  fdJacobian(primGuess, R, c, A);
  st->residualEvals += nf;
  T deltaPrim[GRIM_ORACLE_MAX_VARS];
  for (int v = 0; v < nf; v++) deltaPrim[v] = -R[v];
  luSolve(A, deltaPrim, nf);

  // Quadratic backtracking, solve.cpp:130-191
  T f0 = 0.5 * l2;
  T fPrime0 = -2. * f0;
  T stepLength = T(1.);
  T trial[GRIM_ORACLE_MAX_VARS];
  T Rtrial[GRIM_ORACLE_MAX_VARS];
  for (int v = 0; v < vars.dof; v++) trial[v] = primGuess[v];
  for (int ls = 0; ls < p.maxLineSearchIters; ls++) {
    for (int v = 0; v < nf; v++) trial[v] = primGuess[v] + stepLength * deltaPrim[v];
    computeResidual(trial, c, Rtrial);
    st->residualEvals += 1;
    st->lineSearchEvals += 1;
    T f1 = 0.5 * residualL2(Rtrial, nf);
    const double alpha = 1e-4;
    const double EPS = p.linesearchfloor;
    const bool condition = oval(f1) > oval(f0 * (1. - alpha * stepLength) + EPS);
    if (!condition) break;
    T denom = f1 - f0 - fPrime0 * stepLength;
    stepLength = -fPrime0 * stepLength * stepLength / denom / 2.;
  }
  for (int v = 0; v < nf; v++) primGuess[v] = primGuess[v] + stepLength * deltaPrim[v];
  st->iterations += 1;
}

// Floors applied at the top of every Newton iteration (solve.cpp:17-18).
template <class T>
void applySolveFloors(T* primGuess, const grim_oracle_params& p) {
  primGuess[RHO] = amax(primGuess[RHO], p.rhoFloorInFluidElement);
  primGuess[UU] = amax(primGuess[UU], p.uFloorInFluidElement);
}

// Per-zone Newton (each zone stops at its own convergence).  This is the
// per-zone reading of timeStepper::solve (src/timestepper/solve.cpp:3-201)
// used by the batched solver; the lock-step reading (every zone iterates
// while any zone is unconverged) is grim_oracle_newton_batch's GLOBAL mode.
template <class T>
void newtonZone(T* primGuess, const ResidualContext<T>& c, NewtonZoneStats* st) {
  const grim_oracle_params& p = *c.p;
  const Vars vars(p);
  const int nf = vars.numFluidVars;
  T R[GRIM_ORACLE_MAX_VARS];
  st->iterations = st->residualEvals = st->lineSearchEvals = 0;
  st->converged = 0;
  for (int it = 0; it < p.maxNonLinearIter; it++) {
    applySolveFloors(primGuess, p);
    computeResidual(primGuess, c, R);
    st->residualEvals += 1;
    T l2 = residualL2(R, nf);
    if (!(oval(l2) > p.nonlinearsolve_atol)) {
      st->converged = 1;
      st->finalL2 = oval(l2);
      return;
    }
    newtonUpdate(primGuess, R, l2, c, st);
  }
  computeResidual(primGuess, c, R);
  T l2 = residualL2(R, nf);
  st->finalL2 = oval(l2);
  st->converged = !(st->finalL2 > p.nonlinearsolve_atol);
}

// ---------------------------------------------------------------------------
// Geometry (src/geometry/geometry.cpp)
// ---------------------------------------------------------------------------

// geometry::setgCovInXCoords for MKS without cylindrification,
// src/geometry/geometry.cpp:277-420, with r = exp(X1) and
// theta = ThetaNoCyl(X1, X2) (CoordinateChangeFunctionsArray.hpp:4-46).
inline void mksTheta(double X1, double X2, double hSlope, int derefine,
                     double X1Start, double* theta) {
  const double gammieTheta =
      M_PI * X2 + 0.5 * (1 - hSlope) * std::sin(2. * M_PI * X2);
  if (derefine) {
    double cTh0 = std::cos(M_PI * X2);
    double cTh1 = std::cos(gammieTheta);
    double cTh = cTh1 + std::exp(X1Start) / std::exp(X1) * (cTh0 - cTh1);
    *theta = std::acos(cTh);
  } else {
    *theta = gammieTheta;
  }
}

inline void mksGCov(double X1, double X2, double a, double hSlope,
                    double gCov[4][4]) {
  double r = std::exp(X1);
  double theta;
  mksTheta(X1, X2, hSlope, 0, 0., &theta);
  double dr_dX1 = r;
  double dr_dX2 = 0.;
  double dtheta_dX2 = M_PI + M_PI * (1 - hSlope) * std::cos(2 * M_PI * X2);
  double dtheta_dX1 = 0.;
  double sigma = r * r + std::pow(a * std::cos(theta), 2.);
  double s2 = std::pow(std::sin(theta), 2.);
  gCov[0][0] = -(1. - 2. * r / sigma);
  gCov[0][1] = (2. * r / sigma) * dr_dX1;
  gCov[0][2] = (2. * r / sigma) * dr_dX2;
  gCov[0][3] = -(2. * a * r * s2 / sigma);
  gCov[1][0] = gCov[0][1];
  gCov[1][1] = (1. + 2 * r / sigma) * dr_dX1 * dr_dX1 + (sigma)*dtheta_dX1 * dtheta_dX1;
  gCov[1][2] = (1. + 2 * r / sigma) * dr_dX1 * dr_dX2 + (sigma)*dtheta_dX1 * dtheta_dX2;
  gCov[1][3] = -a * (1. + 2. * r / sigma) * s2 * dr_dX1;
  gCov[2][0] = gCov[0][2];
  gCov[2][1] = gCov[1][2];
  gCov[2][2] = sigma * dtheta_dX2 * dtheta_dX2 + (1. + 2 * r / sigma) * dr_dX2 * dr_dX2;
  gCov[2][3] = -a * (1. + 2. * r / sigma) * s2 * dr_dX2;
  gCov[3][0] = gCov[0][3];
  gCov[3][1] = gCov[1][3];
  gCov[3][2] = gCov[2][3];
  gCov[3][3] = (s2 * (sigma + std::pow(a * std::sin(theta), 2.) * (1. + 2 * r / sigma)));
}

// geometry::setgDetAndgConFromgCov, src/geometry/geometry.cpp:123-253: the
// determinant by its 24-term Leibniz expansion and g^{mu nu} by cofactors.
inline void gDetAndgCon(const double G[4][4], double* gDet, double C[4][4]) {
  double d =
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
  *gDet = d;
  // Cofactor of (i, j) for a symmetric matrix: the 3x3 determinant of the
  // minor, signed.  Written generically rather than term by term.
  for (int i = 0; i < 4; i++) {
    for (int j = 0; j < 4; j++) {
      int ri[3], cj[3];
      for (int k = 0, n = 0; k < 4; k++)
        if (k != j) ri[n++] = k;  // rows of the minor (transpose: adjugate)
      for (int k = 0, n = 0; k < 4; k++)
        if (k != i) cj[n++] = k;
      double m = G[ri[0]][cj[0]] * (G[ri[1]][cj[1]] * G[ri[2]][cj[2]] - G[ri[1]][cj[2]] * G[ri[2]][cj[1]]) -
                 G[ri[0]][cj[1]] * (G[ri[1]][cj[0]] * G[ri[2]][cj[2]] - G[ri[1]][cj[2]] * G[ri[2]][cj[0]]) +
                 G[ri[0]][cj[2]] * (G[ri[1]][cj[0]] * G[ri[2]][cj[1]] - G[ri[1]][cj[1]] * G[ri[2]][cj[0]]);
      C[i][j] = (((i + j) & 1) ? -m : m) / d;
    }
  }
}

}  // namespace oracle
