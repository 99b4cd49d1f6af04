// Runs grim's C++ interfaces of this build (timestepper.hpp, grid.hpp,
// geometry.hpp, physics.hpp, reconstruction.hpp) on one small EMHD torus
// grid and writes every result as raw float64 files into an output
// directory, with a manifest (name, number of doubles) for the Python side
// to compare against its own path through the C ABI.
//
//   facade_check OUTDIR
//
// Grid: MKS a = 0.9375, 48 x 32 x 8 zones, Fishbone-Moncrief torus with the
// SANE loop, EMHD with the torus closure, WENO5 + LLF, relative residual.
// Exit status 1 on any error.
#include <cmath>
#include <cstdio>
#include <fstream>
#include <string>
#include <vector>

#include "timestepper.hpp"

namespace {

struct Writer {
  std::string dir;
  std::ofstream manifest;
  explicit Writer(const std::string& d) : dir(d), manifest(d + "/manifest.txt") {}
  void raw(const std::string& name, const std::vector<double>& v) {
    std::ofstream f(dir + "/" + name + ".f64", std::ios::binary);
    f.write(reinterpret_cast<const char*>(v.data()), (std::streamsize)(v.size() * sizeof(double)));
    manifest << name << " " << v.size() << "\n";
  }
  void device(const std::string& name, const double* p, size_t count) {
    std::vector<double> v(count);
    grim::checkHip(hipMemcpy(v.data(), p, count * sizeof(double), hipMemcpyDeviceToHost), "copy");
    raw(name, v);
  }
  void grid(const std::string& name, const grim::grid& g) { raw(name, g.getVars()); }
};

grim_torus_params torusParams() {
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
  t.seed = 20261202;
  return t;
}

std::unique_ptr<grim::timeStepper> torus(int N1, int N2, int N3) {
  grim::Params P(true);
  P.metric = GRIM_METRIC_MKS;
  P.blackHoleSpin = 0.9375;
  P.hSlope = 0.3;
  P.DerefineThetaHorizon = 1;
  P.DoCylindrify = 1;
  const double Rin = 0.85 * (1. + std::sqrt(1. - 0.9375 * 0.9375));
  P.X1cyl = std::log(4. * Rin);
  P.X2cyl = 1. / N2;
  P.reconstruction = GRIM_RECON_WENO5;
  P.riemannSolver = GRIM_RIEMANN_LLF;
  P.residualScaling = 1;
  P.physics.adiabaticIndex = 5. / 3.;
  P.physics.closure = GRIM_CLOSURE_TORUS;
  P.physics.maxNonLinearIter = 10;
  auto ts = std::make_unique<grim::timeStepper>(
      N1, N2, N3, 3, 10, 3, 0., 0.005, GRIM_BC_OUTFLOW, GRIM_BC_OUTFLOW, GRIM_BC_MIRROR, GRIM_BC_MIRROR,
      GRIM_BC_PERIODIC, GRIM_BC_PERIODIC, GRIM_METRIC_MKS, 0.9375, 0.3, std::log(Rin), std::log(3000.), 1e-8,
      1. - 1e-8, 0., 2. * M_PI, P);
  ts->initialConditionsTorus(torusParams());
  return ts;
}

int run(const std::string& out) {
  const int N1 = 48, N2 = 32, N3 = 8;
  Writer w(out);
  auto ts = torus(N1, N2, N3);
  const grim_grid_config& cfg = ts->config();
  grim_ctx* ctx = ts->context();
  const int dof = ts->numVars;
  const size_t n = (size_t)ts->zones();
  int nR = 0, nW = 0;
  w.grid("primOld", *ts->primOld);

  // reconstruction::reconstruct and riemannSolver::solve per direction
  grim::grid primLeft(ctx, cfg, dof), primRight(ctx, cfg, dof), flux(ctx, cfg, dof);
  grim::riemannSolver riemann(*ts->primOld, *ts->geomCenter);
  grim::geometry* gl[3] = {ts->geomLeft.get(), ts->geomBottom.get(), ts->geomCenter.get()};
  grim::geometry* gr[3] = {ts->geomRight.get(), ts->geomTop.get(), ts->geomCenter.get()};
  for (int d = 0; d < 3; d++) {
    grim::reconstruction::reconstruct(*ts->primOld, d, primLeft, primRight, nR, nW);
    w.grid("primLeft" + std::to_string(d), primLeft);
    w.grid("primRight" + std::to_string(d), primRight);
    riemann.solve(primLeft, primRight, *gl[d], *gr[d], d, flux, nR, nW);
    w.grid("riemann" + std::to_string(d), flux);
  }

  // fluidElement at the zone centres and at the X1 faces
  grim::fluidElement elem(*ts->primOld, *ts->geomCenter, nR, nW);
  w.device("elem_rho", elem.rho, n);
  w.device("elem_tau", elem.tau, n);
  w.device("elem_uCov2", elem.uCov[2], n);
  w.device("elem_bCon0", elem.bCon[0], n);
  w.device("elem_soundSpeed", elem.soundSpeed, n);
  for (int d = 0; d < 4; d++) {
    elem.computeFluxes(d, flux, nR, nW);
    w.grid("fluxes" + std::to_string(d), flux);
  }
  {
    double *mn = nullptr, *mx = nullptr;
    grim::checkHip(hipMalloc((void**)&mn, 2 * n * sizeof(double)), "hipMalloc");
    mx = mn + n;
    for (int d = 0; d < 3; d++) {
      elem.computeMinMaxCharSpeeds(d, mn, mx, nR, nW);
      w.device("minSpeed" + std::to_string(d), mn, n);
      w.device("maxSpeed" + std::to_string(d), mx, n);
    }
    (void)hipFree(mn);
  }
  grim::fluidElement elemFace(*ts->primOld, *ts->geomLeft, nR, nW);
  elemFace.computeFluxes(1, flux, nR, nW);
  w.grid("fluxesFaceX1", flux);

  // sources: implicit with the element's own tau, time-derivative between
  // primOld and a perturbed state, explicit
  grim::grid sources(ctx, cfg, dof);
  elem.computeImplicitSources(sources, elem.tau, nR, nW);
  w.grid("implicitSources", sources);
  grim::grid primNew(ctx, cfg, dof);
  {
    std::vector<double> p = ts->primOld->getVars();
    for (size_t z = 0; z < n; z++) {
      p[1 * n + z] *= 1.01;  // u
      p[2 * n + z] += 1e-3;  // u^1
    }
    primNew.copyHostPtrToVars(p.data());
  }
  grim::fluidElement elemNew(primNew, *ts->geomCenter, nR, nW);
  elem.computeTimeDerivSources(elem, elemNew, 0.005, sources, nR, nW);
  w.grid("timeDerivSources", sources);
  const double dX[3] = {0., 0., 0.};
  grim::grid sExp(ctx, cfg, 7);
  elem.computeExplicitSources(dX, sExp, nR, nW);
  w.grid("explicitSources", sExp);

  // timeStepper::computeDivOfFluxes
  ts->computeDivOfFluxes(*ts->primOld, nR, nW);
  w.grid("divFluxes", *ts->divFluxes);

  // a standalone geometry at CENTER on the same coordinates
  {
    grim::coordinatesGrid X(N1, N2, N3, 3, 3, cfg.X1Start, cfg.X1End, cfg.X2Start, cfg.X2End, cfg.X3Start,
                            cfg.X3End);
    X.setXCoords(grim::CENTER);
    X.DerefineThetaHorizon = cfg.DerefineThetaHorizon;
    X.DoCylindrify = cfg.DoCylindrify;
    X.X1cyl = cfg.X1cyl;
    X.X2cyl = cfg.X2cyl;
    grim::geometry geom(GRIM_METRIC_MKS, 0.9375, 0.3, X);
    geom.computeConnectionCoeffs();
    w.raw("geom_gCov11", geom.hostPlane(geom.gCov[1][1]));
    w.raw("geom_gCon23", geom.hostPlane(geom.gCon[2][3]));
    w.raw("geom_alpha", geom.hostPlane(geom.alpha));
    w.raw("geom_g", geom.hostPlane(geom.g));
    w.raw("geom_gamma123", geom.hostPlane(geom.gammaUpDownDown[1][2][3]));
    w.raw("ts_gCon23", ts->geomCenter->hostPlane(ts->geomCenter->gCon[2][3]));
    w.raw("ts_gammaRight_alpha", ts->geomRight->hostPlane(ts->geomRight->alpha));
  }

  // the step: timeStep against timeStepByStages on a second context
  auto tb = torus(N1, N2, N3);
  for (int s = 0; s < 2; s++) {
    ts->timeStep(nR, nW);
    tb->timeStepByStages(nR, nW);
  }
  w.grid("step_timeStep", *ts->primOld);
  w.grid("step_byStages", *tb->primOld);
  w.raw("step_dt", {ts->dt, tb->dt, ts->time, tb->time});

  // the half stage of the next step: frozen terms, residual at the guess,
  // solve, residual at the solution
  tb->currentStep = 0;
  grim::check(grim_step_begin(tb->context()), "grim_step_begin");
  tb->foldFrozenTerms();
  w.grid("halfStep_guess", *tb->primHalfStep);
  grim::grid residual(tb->context(), tb->config(), 7);
  tb->computeResidual(*tb->primHalfStep, residual, nR, nW);
  w.grid("residual_guess", residual);
  tb->solve(*tb->primHalfStep);
  tb->computeResidual(*tb->primHalfStep, residual, nR, nW);
  w.grid("residual_solution", residual);
  w.grid("halfStep_solution", *tb->primHalfStep);
  grim_stage_arrays a{};
  grim::check(grim_get_stage_arrays(tb->context(), &a), "grim_get_stage_arrays");
  w.device("rhs", a.rhs, 7 * n);
  w.device("residualMask", a.residualMask, n);

  // timeStepper::batchLinearSolve on random well-conditioned 5x5 systems
  {
    const int nv = 5;
    const int64_t m = 1000;
    std::vector<double> A((size_t)nv * nv * m), b((size_t)nv * m), x((size_t)nv * m);
    uint64_t state = 12345;
    auto next = [&]() {
      state = state * 6364136223846793005ull + 1442695040888963407ull;
      return (double)(state >> 11) * (1. / 9007199254740992.) - 0.5;
    };
    for (int64_t z = 0; z < m; z++)
      for (int i = 0; i < nv; i++) {
        for (int j = 0; j < nv; j++) A[(size_t)(i * nv + j) * m + z] = next() + (i == j ? 3. : 0.);
        b[(size_t)i * m + z] = next();
      }
    double* dA = nullptr;
    grim::checkHip(hipMalloc((void**)&dA, sizeof(double) * (A.size() + b.size() + x.size())), "hipMalloc");
    double* db = dA + A.size();
    double* dx = db + b.size();
    grim::checkHip(hipMemcpy(dA, A.data(), sizeof(double) * A.size(), hipMemcpyHostToDevice), "copy");
    grim::checkHip(hipMemcpy(db, b.data(), sizeof(double) * b.size(), hipMemcpyHostToDevice), "copy");
    ts->batchLinearSolve(dA, db, dx, nv, m);
    w.raw("lu_A", A);
    w.raw("lu_b", b);
    w.device("lu_x", dx, x.size());
    (void)hipFree(dA);
  }
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc != 2) {
    std::fprintf(stderr, "usage: facade_check OUTDIR\n");
    return 2;
  }
  try {
    return run(argv[1]);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "facade_check: %s\n", e.what());
    return 1;
  }
}
