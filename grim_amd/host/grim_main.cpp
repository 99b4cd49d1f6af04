// A grim driver on the gfx950 time step: the loop of src/grim.cpp:10-84
// (construct the timeStepper, set the initial conditions, step, report zone
// cycles per second) for the two problems of this path, through the C++
// interface of timestepper.hpp.
//
//   grim_step [--problem fastmode|emhdmode|torus|buoyancy] [--N N1 N2 N3]
//             [--steps K] [--solver fd|fd_lockstep|mixed] [--maxiter M]
//             [--dump FILE] [--observe-every DT] [--write-every DT]
//             [--restart] [--max-wall-time SECONDS] [--dt0 DT]
//             [--atmosphere DIR]
//
// fastmode: 2D Minkowski ideal MHD, 64x64 periodic, WENO5 + LLF, the fast
//           magnetosonic eigenmode at amplitude 1e-4, dt0 = 0.5/N1;
// emhdmode: grim's EMHD linear mode (MINMOD + HLL, amplitude 1e-8);
// torus:    3D MKS torus (a = 0.9375, r_in = 6, r_max = 12, SANE beta 100),
//           EMHD with the torus closure, WENO5 + LLF, relative residual test;
// buoyancy: buoyancy_instabilities' 2D atmosphere (r in [200, 300], a = 0),
//           conduction with the problem's closure, MINMOD + LLF, from grim's
//           four atmosphere_soln_*.txt files in --atmosphere DIR (default
//           the working directory), 4 % noise in u~^1, the X1-side hook;
//           maxNonLinearIter 3 unless --maxiter.
// --dump writes primOld (numVars planes of N1T*N2T*N3T doubles, X1 fastest)
// after the last step.  --observe-every DT prints the torus observers
// whenever time crosses a multiple of DT (fullStepDiagnostics,
// torus/problem.cpp:752-776).  --write-every DT writes primOld to
// primVarsT%06d.h5 and the dump variables to dump%06d.vts whenever time
// crosses a multiple of DT (WriteData, :777-876), and at the start the
// metric files gCov.h5, gCon.h5, sqrtDetg.h5 and xCoords.h5 with the first
// dumps (the initial conditions' WriteIdx 0, :780-792).  A run restarts as grim's constructor does
// (timestepper.cpp:387-441) when restartFileName.txt and restartFileTime.txt
// exist in the working directory, or from restartFile.h5 with --restart;
// --max-wall-time stops after the first step that ends past the limit and
// writes the restart files (CheckWallClockTermination, :880-906).  --dt0
// replaces the problem's initial dt (params::InitialDt).
// Exit status 1 on any library error.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "checkpoint.hpp"
#include "timestepper.hpp"

namespace {

struct Options {
  std::string problem = "fastmode";
  int N[3] = {0, 0, 0};
  int steps = 10;
  std::string solver = "fd";
  int maxIter = 0;  // 0: the problem's maxNonLinearIter
  std::string dump;
  double observeEvery = 0.;
  double writeEvery = 0.;
  bool restart = false;
  double maxWallTime = -1.;
  double dt0 = 0.;
  std::string atmosphere = ".";
};

Options parse(int argc, char** argv) {
  Options o;
  for (int i = 1; i < argc; i++) {
    const std::string a = argv[i];
    auto next = [&]() -> const char* {
      if (i + 1 >= argc) throw std::runtime_error("missing value after " + a);
      return argv[++i];
    };
    if (a == "--problem") {
      o.problem = next();
    } else if (a == "--N") {
      for (int d = 0; d < 3; d++) o.N[d] = std::atoi(next());
    } else if (a == "--steps") {
      o.steps = std::atoi(next());
    } else if (a == "--solver") {
      o.solver = next();
    } else if (a == "--maxiter") {
      o.maxIter = std::atoi(next());
    } else if (a == "--dump") {
      o.dump = next();
    } else if (a == "--observe-every") {
      o.observeEvery = std::atof(next());
    } else if (a == "--write-every") {
      o.writeEvery = std::atof(next());
    } else if (a == "--restart") {
      o.restart = true;
    } else if (a == "--max-wall-time") {
      o.maxWallTime = std::atof(next());
    } else if (a == "--dt0") {
      o.dt0 = std::atof(next());
    } else if (a == "--atmosphere") {
      o.atmosphere = next();
    } else {
      throw std::runtime_error("unknown argument " + a);
    }
  }
  return o;
}

void setSolver(grim::Params& prm, const std::string& solver) {
  if (solver == "fd") {
    prm.newton.semantics = GRIM_NEWTON_PER_ZONE;
  } else if (solver == "fd_lockstep") {
    prm.newton.semantics = GRIM_NEWTON_GLOBAL;
  } else if (solver == "mixed") {
    prm.newton.jacobian = GRIM_JAC_ANALYTIC;
    prm.newton.lu_precision = GRIM_LU_FP32_FP64;
  } else {
    throw std::runtime_error("unknown solver " + solver);
  }
}

int run(const Options& o) {
  const bool torus = o.problem == "torus";
  const bool buoyancy = o.problem == "buoyancy";
  const bool emhd = torus || buoyancy || o.problem == "emhdmode";
  if (!torus && !buoyancy && o.problem != "fastmode" && o.problem != "emhdmode")
    throw std::runtime_error("unknown problem " + o.problem);
  grim::Params prm(emhd);
  setSolver(prm, o.solver);
  prm.physics.maxNonLinearIter = o.maxIter > 0 ? o.maxIter : (buoyancy ? 3 : 10);
  int N1, N2, N3, dim;
  double X1s = 0., X1e = 1., X2s = 0., X2e = 1., X3s = 0., X3e = 1., dt0;
  int bl = GRIM_BC_PERIODIC, br = GRIM_BC_PERIODIC, bb = GRIM_BC_PERIODIC, bt = GRIM_BC_PERIODIC,
      bk = GRIM_BC_PERIODIC, bf = GRIM_BC_PERIODIC;
  int metric = GRIM_METRIC_MINKOWSKI;
  double a = 0.9375, hSlope = 0.3;
  if (torus) {
    N1 = o.N[0] ? o.N[0] : 192;
    N2 = o.N[1] ? o.N[1] : 192;
    N3 = o.N[2] ? o.N[2] : 128;
    dim = 3;
    metric = GRIM_METRIC_MKS;
    const double Rin = 0.85 * (1. + std::sqrt(1. - a * a));  // torus/params.cpp:51
    X1s = std::log(Rin);
    X1e = std::log(3000.);
    X2s = 1e-8;
    X2e = 1. - 1e-8;
    X3s = 0.;
    X3e = 2. * M_PI;
    bl = br = GRIM_BC_OUTFLOW;
    bb = bt = GRIM_BC_MIRROR;
    prm.DerefineThetaHorizon = 1;
    prm.DoCylindrify = 1;
    prm.X1cyl = std::log(4. * Rin);
    prm.X2cyl = 1. / N2;
    prm.physics.adiabaticIndex = 5. / 3.;
    prm.physics.closure = GRIM_CLOSURE_TORUS;
    prm.residualScaling = 1;
    dt0 = 0.005;
  } else if (buoyancy) {
    // buoyancy_instabilities/params.cpp:8-96 (a = 0, one X3 zone of grim's
    // N3 = 128 over [0, pi])
    N1 = o.N[0] ? o.N[0] : 64;
    N2 = o.N[1] ? o.N[1] : 64;
    N3 = 1;
    dim = 2;
    metric = GRIM_METRIC_MKS;
    const double Rin = 200.;
    X1s = std::log(Rin);
    X1e = std::log(300.);
    X2s = 1e-8;
    X2e = 1. - 1e-8;
    X3s = 0.;
    X3e = M_PI / 128.;
    bl = br = GRIM_BC_OUTFLOW;
    bb = bt = GRIM_BC_MIRROR;
    a = 0.;
    hSlope = 1.;
    prm.DerefineThetaHorizon = 1;
    prm.DoCylindrify = 1;
    prm.X1cyl = std::log(4. * Rin);
    prm.X2cyl = 3. / N2;
    prm.reconstruction = GRIM_RECON_MINMOD;
    prm.riemannSolver = GRIM_RIEMANN_LLF;
    prm.slopeLimTheta = 2.;
    prm.CourantFactor = 0.5;
    grim_params& ph = prm.physics;
    ph.conduction = 1;
    ph.viscosity = 0;
    ph.highOrderTermsConduction = ph.highOrderTermsViscosity = 1;
    ph.ConductionAlpha = ph.ViscosityAlpha = 1.;
    ph.ConductionClosureFactor = ph.ViscosityClosureFactor = 1.;
    ph.adiabaticIndex = 5. / 3.;
    ph.rhoFloorInFluidElement = ph.uFloorInFluidElement = 1e-20;
    ph.bSqrFloorInFluidElement = ph.temperatureFloorInFluidElement = 1e-20;
    ph.closure = GRIM_CLOSURE_BUOYANCY;
    ph.maxLineSearchIters = 3;
    ph.nonlinearsolve_atol = 1e-6;
    ph.JacobianAssembleEpsilon = 4e-8;
    ph.linesearchfloor = 1e-24;
    dt0 = 0.01;
  } else {
    N1 = o.N[0] ? o.N[0] : 64;
    N2 = o.N[1] ? o.N[1] : 64;
    N3 = 1;
    dim = 2;
    if (o.problem == "emhdmode") {
      prm.reconstruction = GRIM_RECON_MINMOD;
      prm.riemannSolver = GRIM_RIEMANN_HLL;
    }
    dt0 = 0.5 / N1;
  }
  if (o.dt0 > 0.) {
    dt0 = o.dt0;
  }
  const int numVars = 8 + prm.physics.conduction + prm.physics.viscosity;
  // grim's constructor order: ..., boundaryTop, boundaryBottom, boundaryFront, boundaryBack
  grim::timeStepper ts(N1, N2, N3, dim, numVars, 3, 0., dt0, bl, br, bt, bb, bf, bk, metric, a, hSlope, X1s,
                       X1e, X2s, X2e, X3s, X3e, prm);
  if (torus) {
    grim_torus_params t{};
    t.InnerEdgeRadius = 6.;
    t.PressureMaxRadius = 12.;
    t.Adiabat = 1e-3;
    t.MinPlasmaBeta = 100.;
    t.MagneticLoops = 1;
    t.UseMADdisk = 0;
    t.InitialPerturbationAmplitude = 4e-2;
    t.MaxLorentzFactor = 10.;
    t.RhoFloorAmpl = 1e-3;
    t.UFloorAmpl = 1e-5;
    t.RhoFloorSlope = -2.5;
    t.UFloorSlope = -2.5;
    t.BsqrOverRhoMax = 10.;
    t.BsqrOverUMax = 500.;
    t.seed = 20261202ull;
    ts.initialConditionsTorus(t);
  } else if (buoyancy) {
    ts.initialConditionsBuoyancy(o.atmosphere, 4e-2, 20261203ull);
  } else {
    ts.initialConditionsLinearMode(o.problem == "emhdmode" ? 1 : 0, o.problem == "emhdmode" ? 1e-8 : 1e-4);
  }
  const bool restarted = grim::restart(ts, o.restart);
  // the initial conditions' fullStepDiagnostics (torus/problem.cpp:539):
  // WriteData at WriteIdx 0 writes the metric files and the first dumps
  if (o.writeEvery > 0. && !restarted) {
    grim::dumpMetric(ts);
    grim::dumpPrimitives(ts, grim::primVarsName(ts.time, o.writeEvery));
    grim::dumpVTS(ts, grim::dumpVTSName(ts.time, o.writeEvery));
  }
  const auto start = std::chrono::steady_clock::now();
  const double zones = (double)N1 * N2 * N3;
  double total = 0.;
  int done = 0;
  for (int n = 0; n < o.steps; n++) {
    int numReads = 0, numWrites = 0;
    const auto t0 = std::chrono::steady_clock::now();
    ts.timeStep(numReads, numWrites);
    const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    total += s;
    const grim_step_stats& st = ts.lastStats();
    std::printf("Time = %.6e, dt = %.6e, unconverged = %lld %lld, max iterations = %d %d, "
                "zone cycles/sec = %.4e\n",
                ts.time, st.dt, (long long)st.unconverged[0], (long long)st.unconverged[1],
                st.max_iterations[0], st.max_iterations[1], zones / s);
    // ObserveData (torus/problem.cpp:752): time crossed a multiple of the
    // observation interval during this step
    if (o.observeEvery > 0. &&
        std::floor(ts.time / o.observeEvery) != std::floor((ts.time - st.dt) / o.observeEvery)) {
      std::printf("Global quantities at t = %e\n", ts.time);
      const grim_torus_observables obs = ts.observe();
      grim::timeStepper::ComputeMinMaxVariables(obs);
      grim::timeStepper::ComputeEnergyIntegrals(obs);
      grim::timeStepper::ComputeBoundaryFluxes(obs);
    }
    // WriteData (:753, :777-876): the primitives and the dump variables
    if (o.writeEvery > 0. &&
        std::floor(ts.time / o.writeEvery) != std::floor((ts.time - st.dt) / o.writeEvery)) {
      grim::dumpPrimitives(ts, grim::primVarsName(ts.time, o.writeEvery));
      grim::dumpVTS(ts, grim::dumpVTSName(ts.time, o.writeEvery));
    }
    done++;
    // CheckWallClockTermination (torus/problem.cpp:880-906)
    if (o.maxWallTime >= 0. &&
        std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count() >= o.maxWallTime) {
      const std::string file = grim::writeRestart(ts);
      std::printf("\n Termination reason: WallClock (restart file %s)\n", file.c_str());
      break;
    }
  }
  std::printf("%d steps of %s (%s): %.4e zone cycles/sec\n", done, o.problem.c_str(), o.solver.c_str(),
              zones * done / total);
  if (!o.dump.empty()) {
    const std::vector<double> prim = ts.getPrim();
    FILE* f = std::fopen(o.dump.c_str(), "wb");
    if (!f || std::fwrite(prim.data(), sizeof(double), prim.size(), f) != prim.size())
      throw std::runtime_error("cannot write " + o.dump);
    std::fclose(f);
  }
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  try {
    return run(parse(argc, argv));
  } catch (const std::exception& e) {
    std::fprintf(stderr, "grim_step: %s\n", e.what());
    return 1;
  }
}
