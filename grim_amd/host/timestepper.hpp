// C++ host interface of the gfx950 time step, with the names and argument
// meaning of grim's timeStepper (src/timestepper/timestepper.hpp:19-132) and
// the parameter names of its `namespace params` (src/params.hpp:96-210).
//
// grim calls timeStepper member functions linked statically; this header
// gives a grim driver (src/grim.cpp:10-84) the same calls on top of the C ABI
// of include/grim_hip.h.  The grid arrays live on the GPU inside the
// library's context; the host sees them through getPrim/setPrim (grim's
// gridPy setVars/getVars index order, src/grid/gridPy.pyx:185-230).
//
// Errors: grim's member functions return void and fail through NaNs or
// MPI_Abort; here a failing C call throws std::runtime_error with the
// library's message (exceptions never cross the C ABI itself).  Numerical
// non-convergence is not an error: it is reported in lastStats().
//
// The stage-level members grim's timeStep is built from are here too:
// solve(grid&), computeResidual, computeDivOfFluxes, batchLinearSolve, the
// grids prim / primHalfStep / primOld / consOld / sourcesExplicit /
// sourcesImplicitOld / divFluxes / fluxesX1..3 and the geometries
// geomLeft / geomRight / geomBottom / geomTop / geomCenter, all on the
// context's device arrays (grid.hpp, geometry.hpp, physics.hpp,
// reconstruction.hpp mirror the classes they take).
#pragma once

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "geometry.hpp"
#include "grid.hpp"
#include "grim_hip.h"
#include "physics.hpp"
#include "reconstruction.hpp"

namespace grim {

// The members of grim's `namespace params` that the time step reads.
struct Params {
  // grid and geometry (src/problem/*/params.cpp)
  int metric = GRIM_METRIC_MINKOWSKI;
  double blackHoleSpin = 0.9375, hSlope = 0.3;
  int DerefineThetaHorizon = 0, DoCylindrify = 0;
  double X1cyl = 0., X2cyl = 0.;
  // scheme
  int reconstruction = GRIM_RECON_WENO5, riemannSolver = GRIM_RIEMANN_LLF;
  double slopeLimTheta = 2., CourantFactor = 0.9, maxDtIncrement = 1.3;
  // physics (grim_params)
  grim_params physics{};
  // Newton path of the two solves and the relative residual test
  grim_newton_opts newton{};
  int residualScaling = 0;
  // the explicit stage with shared reciprocals and fused multiply-adds
  int explicitFast = 0;

  explicit Params(bool emhd = false) {
    grim_default_params(&physics, emhd ? 1 : 0);
    newton.jacobian = GRIM_JAC_FD;
    newton.lu_precision = GRIM_LU_FP64;
    newton.semantics = GRIM_NEWTON_PER_ZONE;
  }
};

class timeStepper {
  grim_ctx* ctx = nullptr;
  grim_step_stats stats{};
  grim_grid_config cfg{};
  grim_params phys{};
  bool isTorus = false;

 public:
  double dt = 0., time = 0.;
  int N1, N2, N3, numGhost, dim;
  int numVars;
  int boundaryLeft, boundaryRight;
  int boundaryTop, boundaryBottom;
  int boundaryFront, boundaryBack;

  // timeStepper::timeStepper (src/timestepper/timestepper.cpp:3-386): the
  // grid arrays and geometry (zone centres, lower faces, connection).  grim
  // passes top/bottom in that order; they map to X2's upper/lower faces.
  timeStepper(const int N1, const int N2, const int N3, const int dim, const int numVars,
              const int numGhost, const double time, const double dt, const int boundaryLeft,
              const int boundaryRight, const int boundaryTop, const int boundaryBottom,
              const int boundaryFront, const int boundaryBack, const int metric,
              const double blackHoleSpin, const double hSlope, const double X1Start,
              const double X1End, const double X2Start, const double X2End, const double X3Start,
              const double X3End, const Params& params = Params(), void* stream = nullptr)
      : dt(dt), time(time), N1(N1), N2(N2), N3(N3), numGhost(numGhost), dim(dim), numVars(numVars),
        boundaryLeft(boundaryLeft), boundaryRight(boundaryRight), boundaryTop(boundaryTop),
        boundaryBottom(boundaryBottom), boundaryFront(boundaryFront), boundaryBack(boundaryBack) {
    if (numVars != 8 + params.physics.conduction + params.physics.viscosity)
      throw std::runtime_error("timeStepper: numVars does not match the conduction/viscosity switches");
    phys = params.physics;
    cfg.N1 = N1;
    cfg.N2 = N2;
    cfg.N3 = N3;
    cfg.dim = dim;
    cfg.numGhost = numGhost;
    cfg.X1Start = X1Start;
    cfg.X1End = X1End;
    cfg.X2Start = X2Start;
    cfg.X2End = X2End;
    cfg.X3Start = X3Start;
    cfg.X3End = X3End;
    cfg.metric = metric;
    cfg.blackHoleSpin = blackHoleSpin;
    cfg.hSlope = hSlope;
    cfg.DerefineThetaHorizon = params.DerefineThetaHorizon;
    cfg.DoCylindrify = params.DoCylindrify;
    cfg.X1cyl = params.X1cyl;
    cfg.X2cyl = params.X2cyl;
    cfg.boundaryLeft = boundaryLeft;
    cfg.boundaryRight = boundaryRight;
    cfg.boundaryBottom = boundaryBottom;
    cfg.boundaryTop = boundaryTop;
    cfg.boundaryBack = boundaryBack;
    cfg.boundaryFront = boundaryFront;
    cfg.reconstruction = params.reconstruction;
    cfg.riemannSolver = params.riemannSolver;
    cfg.slopeLimTheta = params.slopeLimTheta;
    cfg.CourantFactor = params.CourantFactor;
    cfg.maxDtIncrement = params.maxDtIncrement;
    cfg.residualScaling = params.residualScaling;
    cfg.explicitFast = params.explicitFast;
    check(grim_create(&params.physics, &cfg, &params.newton, (grim_stream_t)stream, &ctx), "grim_create");
    attachMembers();
  }
  ~timeStepper() {
    // the views go before the context that holds their memory
    prim.reset();
    primHalfStep.reset();
    primOld.reset();
    consOld.reset();
    sourcesExplicit.reset();
    sourcesImplicitOld.reset();
    fluxesX1.reset();
    fluxesX2.reset();
    fluxesX3.reset();
    divFluxes.reset();
    geomLeft.reset();
    geomRight.reset();
    geomBottom.reset();
    geomTop.reset();
    geomCenter.reset();
    grim_destroy(ctx);
  }
  timeStepper(const timeStepper&) = delete;
  timeStepper& operator=(const timeStepper&) = delete;

  // Zones of one variable plane, ghosts included (N1T * N2T * N3T).
  int64_t zones() const { return grim_grid_zones(&cfg); }
  const grim_grid_config& config() const { return cfg; }
  const grim_params& physics() const { return phys; }
  // the torus problem's hooks are active (initialConditionsTorus)
  bool torus() const { return isTorus; }

  // timeStepper::initialConditions of the two problems this path serves
  // (src/problem/linear_modes/problem.cpp:10-111, src/problem/torus/problem.cpp:145-540);
  // a caller with its own initial state uses setPrim.  Either is followed by
  // the start time and dt (the constructor's time and dt by default).
  void initialConditionsLinearMode(int mode, double amplitude) {
    check(grim_init_linear_mode(ctx, mode, amplitude), "grim_init_linear_mode");
    setTime(time, dt);
  }
  void initialConditionsTorus(const grim_torus_params& torus) {
    check(grim_init_torus(ctx, &torus), "grim_init_torus");
    isTorus = true;
    setTime(time, dt);
  }

  // timeStepper::initialConditions of buoyancy_instabilities (src/problem/
  // buoyancy_instabilities/problem.cpp:22-216) and from then on its
  // setProblemSpecificBCs (:221-293, grim_set_x1_side_state):
  //  * rho, u and (conduction) q~ of X1 zone i, ghosts included, from grim's
  //    input files in dir (atmosphere_soln_rho.txt, _u.txt, _phi.txt,
  //    _rCoords.txt: N1 + 2 numGhost values, one per line), after checking
  //    |r(i) - rCoords[i]| <= 1e-10 on the first X2 / X3 line;
  //  * u~^1 of the static Boyer-Lindquist observer, (c u^r / r - sqrt(-a b^4
  //    - a b^2 c (u^r)^2 / r^2 - b^2 c)) / (a b^2 + c) with a = g_11,
  //    b = g^01, c = g^00, u^r = 0, times 1 + amplitude (U - 1/2) on the
  //    bulk zones (U: the library's counter-based stream in place of
  //    PETSc's RAND48); u~^2 = u~^3 = 0;
  //  * B^1 = B^2 = 0, B^3 = 1e-5 / sqrt(-g).
  // Errors carry grim's messages.
  void initialConditionsBuoyancy(const std::string& dir, double amplitude, uint64_t seed) {
    const int64_t n = zones();
    const int ng = cfg.numGhost;
    const int N1T = cfg.N1 + 2 * ng;
    const int N2T = cfg.dim > 1 ? cfg.N2 + 2 * ng : 1;
    const int N3T = cfg.dim > 2 ? cfg.N3 + 2 * ng : 1;
    static const char* names[4] = {"atmosphere_soln_rho.txt", "atmosphere_soln_u.txt", "atmosphere_soln_phi.txt",
                                   "atmosphere_soln_rCoords.txt"};
    std::vector<double> prof[4];
    for (int f = 0; f < 4; f++) {
      std::ifstream in(dir + "/" + names[f]);
      if (!in) throw std::runtime_error("Input data files not found!");
      std::string line;
      for (int i = 0; i < N1T; i++) {
        if (!std::getline(in, line))
          throw std::runtime_error(
              "Found the input data files but error in reading them! Check number of grid zones in python script");
        prof[f].push_back(std::atof(line.c_str()));
      }
    }
    std::vector<double> geom((size_t)GRIM_LOC_PLANES * n);
    check(grim_get_geometry(ctx, 0, geom.data(), 0), "grim_get_geometry");
    const double* r = geom.data() + 22 * n;
    for (int i = 0; i < N1T; i++)
      if (std::fabs(r[i] - prof[3][i]) > 1e-10) {
        char msg[200];
        std::snprintf(msg, sizeof msg,
                      "Mismatch in rCoords! Check r coords in python script. r = %.18f, rCoords = %.18f, i = %d", r[i],
                      prof[3][i], i);
        throw std::runtime_error(msg);
      }
    std::vector<double> prim((size_t)numVars * n, 0.), side((size_t)3 * n);
    const int B3 = numVars - 1;
    for (int k = 0; k < N3T; k++)
      for (int j = 0; j < N2T; j++)
        for (int i = 0; i < N1T; i++) {
          const int64_t z = i + (int64_t)N1T * (j + (int64_t)N2T * k);
          const double a = geom[4 * n + z], b = geom[11 * n + z], c = geom[10 * n + z], rz = r[z];
          const double uCon1 = 0. * prim[2 * n + z];
          const double vUpr =
              (c * uCon1 / rz - std::sqrt(-a * b * b * b * b - a * b * b * c * uCon1 * uCon1 / (rz * rz) - b * b * c)) /
              (a * b * b + c);
          const bool bulk = i >= ng && i < cfg.N1 + ng && (cfg.dim < 2 || (j >= ng && j < cfg.N2 + ng)) &&
                            (cfg.dim < 3 || (k >= ng && k < cfg.N3 + ng));
          prim[z] = prof[0][i];
          prim[n + z] = prof[1][i];
          if (phys.conduction) prim[5 * n + z] = prof[2][i];
          prim[2 * n + z] = bulk ? vUpr * (1. + amplitude * (uniform(seed, (uint64_t)z) - 0.5)) : vUpr;
          prim[B3 * n + z] = 1e-5 / geom[15 * n + z];
          for (int v = 0; v < 3; v++) side[v * n + z] = prim[v * n + z];
        }
    setPrim(prim);
    check(grim_set_x1_side_state(ctx, side.data(), 0), "grim_set_x1_side_state");
    setTime(time, dt);
  }
  // The library's counter-based uniform in [0, 1) of zone z: splitmix64 of
  // seed + G (32 z + 1), G = 0x9E3779B97F4A7C15 (the torus's noise stream).
  static double uniform(uint64_t seed, uint64_t z) {
    uint64_t x = seed + 0x9E3779B97F4A7C15ull * (32ull * z + 1ull);
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    x = x ^ (x >> 31);
    return (double)(x >> 11) * (1.0 / 9007199254740992.0);
  }

  // grid::setVars / getVars: numVars planes of zones() doubles on the host.
  void setPrim(const std::vector<double>& prim) {
    if ((int64_t)prim.size() != (int64_t)numVars * zones()) throw std::runtime_error("setPrim: wrong size");
    check(grim_set_prim(ctx, prim.data(), 0), "grim_set_prim");
  }
  std::vector<double> getPrim() const {
    std::vector<double> out((size_t)numVars * (size_t)zones());
    check(grim_get_prim(ctx, out.data(), 0), "grim_get_prim");
    return out;
  }
  // fullStepDiagnostics' dump variables of primOld (grim_dump_vars): numVars
  // + 5 planes of the N1*N2*N3 bulk zones on the host, the last three the
  // x coordinates of the zone centres.
  std::vector<double> dumpVars() const {
    const size_t nb = (size_t)cfg.N1 * cfg.N2 * cfg.N3;
    std::vector<double> out((size_t)(numVars + 5) * nb);
    check(grim_dump_vars(ctx, nullptr, out.data(), 0), "grim_dump_vars");
    return out;
  }
  // Device address of primOld (numVars planes), for on-device analysis.
  double* primOldDevice() const {
    double* p = nullptr;
    check(grim_prim(ctx, &p), "grim_prim");
    return p;
  }

  // grid::communicate on primOld (periodic ghost zones, one rank) and the
  // torus's applyFloor on primOld.
  void communicate() { check(grim_communicate(ctx, primOldDevice(), numVars), "grim_communicate"); }
  void applyFloor() { check(grim_apply_floor(ctx, primOldDevice()), "grim_apply_floor"); }

  void setTime(double t, double d) {
    check(grim_set_time(ctx, t, d), "grim_set_time");
    time = t;
    dt = d;
  }

  // timeStepper::timeStep (src/timestepper/timestep.cpp:3-334).  grim fills
  // numReads/numWrites with array-access counts for its bandwidth report;
  // kernel traffic here is measured with rocprofv3 instead, so both are 0.
  void timeStep(int& numReads, int& numWrites) {
    check(grim_time_step(ctx, &stats), "grim_time_step");
    check(grim_get_time(ctx, &time, &dt), "grim_get_time");
    numReads = 0;
    numWrites = 0;
  }

  // The same step without statistics or host synchronisation (dt and time
  // advance on the device; call syncTime() to read them).
  void timeStepAsync() { check(grim_time_step(ctx, nullptr), "grim_time_step"); }
  void syncTime() { check(grim_get_time(ctx, &time, &dt), "grim_get_time"); }

  // timeStepper::computeDt (timestep.cpp:336-400) from the current primOld:
  // the next dt, bounded by maxDtIncrement times the current one.
  void computeDt(int& numReads, int& numWrites) {
    double next = 0.;
    check(grim_compute_dt(ctx, primOldDevice(), dt, &next), "grim_compute_dt");
    numReads = 0;
    numWrites = 0;
    dt = next;
  }

  // The torus's on-the-fly observers of primOld (ObserversTorus.hpp:6-251),
  // all three from one pass over the grid; each prints grim's line.
  grim_torus_observables observe() {
    grim_torus_observables o{};
    check(grim_torus_observers(ctx, nullptr, &o), "grim_torus_observers");
    return o;
  }
  static void ComputeMinMaxVariables(const grim_torus_observables& o) {
    std::printf("rhoMax = %e; betaMin = %e;\n", o.rho_max, o.beta_min);
  }
  static void ComputeEnergyIntegrals(const grim_torus_observables& o) {
    std::printf("Baryon Mass = %e; Magnetic Energy = %e; Thermal Energy = %e;\n", o.baryon_mass,
                o.magnetic_energy, o.thermal_energy);
  }
  static void ComputeBoundaryFluxes(const grim_torus_observables& o) {
    std::printf("MdotIn = %e; MdotOut = %e; ( Unbound = %e; Relativistic = %e)\n", o.mdot_in, o.mdot_out,
                o.mdot_out_unbound, o.mdot_out_relativistic);
  }

  // ---- the stage-level members (timestepper.hpp:36-41, 58-90) -----------
  // grids on the context's arrays: prim (the full stage's solution),
  // primHalfStep, primOld (dof planes), consOld, sourcesExplicit,
  // sourcesImplicitOld (numFluidVars of dof planes used), the face fluxes
  // of the last stage and divFluxes (owned, written by computeDivOfFluxes)
  std::unique_ptr<grid> prim, primHalfStep, primOld, consOld, sourcesExplicit, sourcesImplicitOld;
  std::unique_ptr<grid> fluxesX1, fluxesX2, fluxesX3, divFluxes;
  std::unique_ptr<geometry> geomLeft, geomRight, geomBottom, geomTop, geomCenter;
  int currentStep = 0;  // timeStepperSwitches: 0 = HALF_STEP, 1 = FULL_STEP

  // timeStepper::computeDivOfFluxes (fvmfluxes.cpp:3-184) of prim into
  // divFluxes, with flux-CT and the problem's flux filter.
  void computeDivOfFluxes(const grid& primIn, int& numReads, int& numWrites) {
    check(grim_div_fluxes(ctx, primIn.vars, divFluxes->vars, nullptr), "timeStepper::computeDivOfFluxes");
    numReads = 0;
    numWrites = 0;
  }

  // The part of timeStep before solve in the current stage (timestep.cpp:
  // 15-89 / 184-237): boundaries, sources, divergence of fluxes, induction,
  // and the frozen terms folded into the residual's right-hand side; the
  // guess goes into primHalfStep (half stage) or prim (full stage).
  void foldFrozenTerms() { check(grim_fold_rhs(ctx, currentStep), "timeStepper: grim_fold_rhs"); }

  // timeStepper::solve (solve.cpp:3-201) of the current stage on primGuess.
  void solve(grid& primGuess) {
    check(grim_solve(ctx, currentStep, primGuess.vars, &stats), "timeStepper::solve");
  }

  // timeStepper::computeResidual (residual.cpp:3-180) of the current stage.
  void computeResidual(const grid& primIn, grid& residual, int& numReads, int& numWrites) {
    check(grim_stage_residual(ctx, currentStep, primIn.vars, residual.vars), "timeStepper::computeResidual");
    numReads = 0;
    numWrites = 0;
  }

  // timeStepper::batchLinearSolve (solve.cpp:203-230): A x = b for n zones of
  // nvar unknowns, device planes (A: nvar*nvar planes, entry (i, j) in plane
  // i*nvar + j; b, x: nvar planes), LU with partial pivoting in FP64.
  void batchLinearSolve(const double* A, const double* b, double* x, int nvar, int64_t n) {
    check(grim_batch_linear_solve(n, nvar, A, b, x, GRIM_LU_FP64, nullptr), "timeStepper::batchLinearSolve");
  }

  // timeStep (timestep.cpp:3-334) written out as grim's control flow over
  // the stage-level members: computeDt, then for each stage the frozen
  // terms, solve, communicate and the diagnostics hooks, then primOld =
  // prim and time += dt.  The same kernels as timeStep(), in the same order.
  void timeStepByStages(int& numReads, int& numWrites) {
    check(grim_step_begin(ctx), "timeStepper: grim_step_begin");
    currentStep = 0;
    foldFrozenTerms();
    solve(*primHalfStep);
    check(grim_stage_finish(ctx, 0), "timeStepper: grim_stage_finish");
    currentStep = 1;
    foldFrozenTerms();
    solve(*prim);
    check(grim_stage_finish(ctx, 1), "timeStepper: grim_stage_finish");
    check(grim_step_end(ctx), "timeStepper: grim_step_end");
    check(grim_get_time(ctx, &time, &dt), "grim_get_time");
    numReads = 0;
    numWrites = 0;
  }

  // Statistics of the last timeStep: dt, unconverged zones and largest
  // Newton iteration count per stage, zones with a non-finite residual.
  const grim_step_stats& lastStats() const { return stats; }
  grim_ctx* context() const { return ctx; }

 private:
  void attachMembers() {
    grim_stage_arrays a{};
    check(grim_get_stage_arrays(ctx, &a), "grim_get_stage_arrays");
    const int nf = 5 + phys.conduction + phys.viscosity;
    prim = std::make_unique<grid>(ctx, cfg, numVars, a.prim);
    primHalfStep = std::make_unique<grid>(ctx, cfg, numVars, a.primHalfStep);
    primOld = std::make_unique<grid>(ctx, cfg, numVars, a.primOld);
    consOld = std::make_unique<grid>(ctx, cfg, numVars, a.consOld);
    sourcesExplicit = std::make_unique<grid>(ctx, cfg, nf, a.sourcesExplicit);
    sourcesImplicitOld = std::make_unique<grid>(ctx, cfg, nf, a.sourcesImplicitOld);
    fluxesX1 = std::make_unique<grid>(ctx, cfg, numVars, a.fluxesX1);
    if (a.fluxesX2) fluxesX2 = std::make_unique<grid>(ctx, cfg, numVars, a.fluxesX2);
    if (a.fluxesX3) fluxesX3 = std::make_unique<grid>(ctx, cfg, numVars, a.fluxesX3);
    divFluxes = std::make_unique<grid>(ctx, cfg, numVars);
    geomCenter = std::make_unique<geometry>(ctx, cfg, CENTER);
    geomLeft = std::make_unique<geometry>(ctx, cfg, LEFT);
    geomRight = std::make_unique<geometry>(ctx, cfg, RIGHT);
    if (dim > 1) {
      geomBottom = std::make_unique<geometry>(ctx, cfg, BOTTOM);
      geomTop = std::make_unique<geometry>(ctx, cfg, TOP);
    }
  }
};

}  // namespace grim
