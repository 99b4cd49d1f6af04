// grim oracle C interface (TEST INFRASTRUCTURE ONLY; see grim_oracle.h).
#include "grim_oracle.h"

#include <omp.h>

#include <cmath>
#include <limits>
#include <vector>

#include "grim_oracle.hpp"
#include "scalars.hpp"

namespace oracle {
thread_local OpCounts g_ops;
thread_local Tape g_tape;

namespace {

const double kNaN = std::numeric_limits<double>::quiet_NaN();

template <class T>
void loadGeom(const double* geom, const double* radius, int64_t n, int64_t z,
              ZoneGeom<T>* G) {
  auto at = [&](int plane) { return geom[plane * n + z]; };
  const int idx[4][4] = {{0, 1, 2, 3}, {1, 4, 5, 6}, {2, 5, 7, 8}, {3, 6, 8, 9}};
  for (int m = 0; m < 4; m++)
    for (int k = 0; k < 4; k++) {
      G->gCov[m][k] = T(at(idx[m][k]));
      G->gCon[m][k] = T(kNaN);
    }
  for (int k = 0; k < 4; k++) {
    G->gCon[0][k] = T(at(10 + k));
    G->gCon[k][0] = T(at(10 + k));
  }
  G->alpha = T(at(14));
  G->g = T(at(15));
  G->r = T(radius ? radius[z] : kNaN);
}

ZoneClosure loadClosure(const double* closure, int64_t n, int64_t z) {
  ZoneClosure c{1., 1.};
  if (closure) {
    c.conductionAlpha = closure[z];
    c.viscosityAlpha = closure[n + z];
  }
  return c;
}

// Everything one zone's solve needs, loaded from the batch.
template <class T>
struct ZoneState {
  ZoneGeom<T> geom;
  FluidElement<T> elemOld, elemCoef;
  double rhs[GRIM_ORACLE_MAX_VARS];
  T prim[GRIM_ORACLE_MAX_VARS];
  ResidualContext<T> ctx;

  void load(const grim_oracle_params& p, const grim_oracle_batch& b, int64_t z) {
    const Vars vars(p);
    const int64_t n = b.n;
    loadGeom(b.geom, b.radius, n, z, &geom);
    ZoneClosure zc = loadClosure(b.closure, n, z);
    for (int v = 0; v < vars.numFluidVars; v++) rhs[v] = b.rhs[v * n + z];
    for (int v = 0; v < vars.dof; v++) prim[v] = T(b.prim[v * n + z]);
    const bool emhd = p.conduction || p.viscosity;
    if (emhd) {
      T po[GRIM_ORACLE_MAX_VARS];
      for (int v = 0; v < vars.dof; v++) po[v] = T(b.primOld[v * n + z]);
      elemOld.set(po, geom, p, zc);
      if (b.primCoef) {
        for (int v = 0; v < vars.dof; v++) po[v] = T(b.primCoef[v * n + z]);
        elemCoef.set(po, geom, p, zc);
      } else {
        elemCoef = elemOld;
      }
      // copies keep pointing at this zone's geometry
      elemOld.geom = &geom;
      elemCoef.geom = &geom;
    }
    ctx.p = &p;
    ctx.geom = &geom;
    ctx.closure = zc;
    ctx.frozen = nullptr;
    ctx.rhs = rhs;
    ctx.dtPrime = b.dtPrime;
    ctx.residualMask = b.residualMask ? b.residualMask[z] : 1.;
    ctx.elemOld = &elemOld;
    ctx.elemCoef = &elemCoef;
  }
};

void setThreads(int nthreads) {
  if (nthreads > 0) omp_set_num_threads(nthreads);
}

}  // namespace
}  // namespace oracle

using namespace oracle;

extern "C" {

void grim_oracle_default_params(grim_oracle_params* p, int emhd) {
  // Defaults of src/problem/linear_modes/params.cpp:40-69 (floors, EMHD
  // switches, solver tolerances), adiabatic index 4/3.
  p->conduction = emhd ? 1 : 0;
  p->viscosity = emhd ? 1 : 0;
  p->highOrderTermsConduction = 1;
  p->highOrderTermsViscosity = 1;
  p->adiabaticIndex = 4. / 3;
  p->rhoFloorInFluidElement = 1e-20;
  p->uFloorInFluidElement = 1e-20;
  p->bSqrFloorInFluidElement = 1e-20;
  p->temperatureFloorInFluidElement = 1e-20;
  p->ConductionAlpha = 1.;
  p->ViscosityAlpha = 1.;
  p->closure = GRIM_ORACLE_CLOSURE_LINEAR_MODES;
  p->tau0 = 1.;
  p->ConductionClosureFactor = 1.;
  p->ViscosityClosureFactor = 1.;
  p->maxNonLinearIter = 3;
  p->maxLineSearchIters = 3;
  p->nonlinearsolve_atol = 1.e-20;
  p->JacobianAssembleEpsilon = 4.e-8;
  p->linesearchfloor = 1.e-24;
}

void grim_oracle_mks_geometry(int64_t n, const double* X1, const double* X2,
                              double a, double hSlope, double* gCov,
                              double* gCon, double* g, double* alpha,
                              double* gamma) {
#pragma omp parallel for schedule(static)
  for (int64_t z = 0; z < n; z++) {
    double G[4][4], C[4][4], det;
    mksGCov(X1[z], X2[z], a, hSlope, G);
    gDetAndgCon(G, &det, C);
    for (int m = 0; m < 4; m++)
      for (int k = 0; k < 4; k++) {
        if (gCov) gCov[(m * 4 + k) * n + z] = G[m][k];
        if (gCon) gCon[(m * 4 + k) * n + z] = C[m][k];
      }
    if (g) g[z] = std::sqrt(-det);
    if (alpha) alpha[z] = 1. / std::sqrt(-C[0][0]);
    if (gamma) {
      // geometry::computeConnectionCoeffs, src/geometry/geometry.cpp:61-121,
      // and computeGammaDownDownDown, :631-680.  GAMMA_EPS = 1e-5 (:11).
      const double EPS = 1.e-5;
      double Gp[3][4][4], Gm[3][4][4];
      for (int d = 0; d < 3; d++) {
        double Xp[3] = {X1[z], X2[z], 0.}, Xm[3] = {X1[z], X2[z], 0.};
        Xp[d] += EPS;
        Xm[d] -= EPS;
        mksGCov(Xp[0], Xp[1], a, hSlope, Gp[d]);
        mksGCov(Xm[0], Xm[1], a, hSlope, Gm[d]);
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
      for (int mu = 0; mu < 4; mu++)
        for (int nu = 0; nu < 4; nu++)
          for (int la = 0; la < 4; la++) {
            double s = 0.;
            for (int eta = 0; eta < 4; eta++) s += C[mu][eta] * down[eta][nu][la];
            gamma[(mu * 16 + nu * 4 + la) * n + z] = s;
          }
    }
  }
}

void grim_oracle_pack_geometry(int64_t n, const double* gCov,
                               const double* gCon, const double* g,
                               const double* alpha, double* geom16) {
  const int cov[10] = {0, 1, 2, 3, 5, 6, 7, 10, 11, 15};
#pragma omp parallel for schedule(static)
  for (int64_t z = 0; z < n; z++) {
    for (int k = 0; k < 10; k++) geom16[k * n + z] = gCov[cov[k] * n + z];
    for (int k = 0; k < 4; k++) geom16[(10 + k) * n + z] = gCon[k * n + z];
    geom16[14 * n + z] = alpha[z];
    geom16[15 * n + z] = g[z];
  }
}

void grim_oracle_conserved(const grim_oracle_params* p, int64_t n,
                           const double* prim, const double* geom,
                           const double* closure, const double* radius,
                           double* cons) {
  const Vars vars(*p);
#pragma omp parallel for schedule(static)
  for (int64_t z = 0; z < n; z++) {
    ZoneGeom<double> G;
    loadGeom(geom, radius, n, z, &G);
    double P[GRIM_ORACLE_MAX_VARS], U[GRIM_ORACLE_MAX_VARS];
    for (int v = 0; v < vars.dof; v++) P[v] = prim[v * n + z];
    FluidElement<double> e;
    e.set(P, G, *p, loadClosure(closure, n, z));
    e.computeFluxes(0, U, *p);
    for (int v = 0; v < vars.dof; v++) cons[v * n + z] = U[v];
  }
}

void grim_oracle_fluxes(const grim_oracle_params* p, int64_t n, const double* prim,
                        const double* geom, const double* closure, const double* radius,
                        int dir, double* flux) {
  const Vars vars(*p);
#pragma omp parallel for schedule(static)
  for (int64_t z = 0; z < n; z++) {
    ZoneGeom<double> G;
    loadGeom(geom, radius, n, z, &G);
    double P[GRIM_ORACLE_MAX_VARS], F[GRIM_ORACLE_MAX_VARS];
    for (int v = 0; v < vars.dof; v++) P[v] = prim[v * n + z];
    FluidElement<double> e;
    e.set(P, G, *p, loadClosure(closure, n, z));
    e.computeFluxes(dir, F, *p);
    for (int v = 0; v < vars.dof; v++) flux[v * n + z] = F[v];
  }
}

void grim_oracle_element(const grim_oracle_params* p, int64_t n, const double* prim,
                         const double* geom, const double* closure, const double* radius,
                         double* out) {
  const Vars vars(*p);
#pragma omp parallel for schedule(static)
  for (int64_t z = 0; z < n; z++) {
    ZoneGeom<double> G;
    loadGeom(geom, radius, n, z, &G);
    double P[GRIM_ORACLE_MAX_VARS];
    for (int v = 0; v < vars.dof; v++) P[v] = prim[v * n + z];
    FluidElement<double> e;
    e.set(P, G, *p, loadClosure(closure, n, z));
    double vals[GRIM_ORACLE_ELEMENT_PLANES];
    int k = 0;
    for (int m = 0; m < 4; m++) vals[k++] = e.uCon[m];
    for (int m = 0; m < 4; m++) vals[k++] = e.uCov[m];
    for (int m = 0; m < 4; m++) vals[k++] = e.bCon[m];
    for (int m = 0; m < 4; m++) vals[k++] = e.bCov[m];
    vals[k++] = e.bSqr;
    vals[k++] = e.q;
    vals[k++] = e.deltaP;
    vals[k++] = e.pressure;
    vals[k++] = e.rho;
    vals[k++] = e.u;
    vals[k++] = e.tau;
    vals[k++] = e.chi_emhd;
    vals[k++] = e.nu_emhd;
    for (int m = 0; m < 4; m++)
      for (int q = 0; q < 4; q++) vals[k++] = e.TUpDown[m][q];
    for (int j = 0; j < GRIM_ORACLE_ELEMENT_PLANES; j++) out[j * n + z] = vals[j];
  }
}

void grim_oracle_residual_pieces(const grim_oracle_params* p, const grim_oracle_batch* b,
                                 double* Uo, double* Sio, double* Stdo) {
  const Vars vars(*p);
  const int64_t n = b->n;
  const bool emhd = p->conduction || p->viscosity;
#pragma omp parallel for schedule(static)
  for (int64_t z = 0; z < n; z++) {
    ZoneState<double> s;
    s.load(*p, *b, z);
    FluidElement<double> elem;
    elem.set(s.prim, s.geom, *p, s.ctx.closure);
    double U[GRIM_ORACLE_MAX_VARS], Si[GRIM_ORACLE_MAX_VARS], Std[GRIM_ORACLE_MAX_VARS];
    elem.computeFluxes(0, U, *p);
    for (int v = 0; v < vars.dof; v++) Si[v] = Std[v] = 0.;
    if (emhd) {
      elem.computeImplicitSources(Si, s.elemCoef.tau, *p);
      s.elemCoef.computeTimeDerivSources(s.elemOld, elem, b->dtPrime, Std, *p);
    }
    for (int v = 0; v < vars.numFluidVars; v++) {
      Uo[v * n + z] = U[v];
      Sio[v * n + z] = Si[v];
      Stdo[v * n + z] = Std[v];
    }
  }
}

void grim_oracle_known_root_rhs(const grim_oracle_params* p,
                                const grim_oracle_batch* b, double* rhs) {
  const Vars vars(*p);
  const int64_t n = b->n;
  const bool emhd = p->conduction || p->viscosity;
  std::vector<double> zero((size_t)vars.numFluidVars * n, 0.);
  grim_oracle_batch bb = *b;
  bb.rhs = zero.data();
#pragma omp parallel for schedule(static)
  for (int64_t z = 0; z < n; z++) {
    ZoneState<double> s;
    s.load(*p, bb, z);
    FluidElement<double> elem;
    elem.set(s.prim, s.geom, *p, s.ctx.closure);
    double U[GRIM_ORACLE_MAX_VARS], Si[GRIM_ORACLE_MAX_VARS], Std[GRIM_ORACLE_MAX_VARS];
    elem.computeFluxes(0, U, *p);
    if (emhd) {
      elem.computeImplicitSources(Si, s.elemCoef.tau, *p);
      s.elemCoef.computeTimeDerivSources(s.elemOld, elem, b->dtPrime, Std, *p);
    }
    for (int v = 0; v < vars.numFluidVars; v++)
      rhs[v * n + z] = emhd ? U[v] + b->dtPrime * (0.5 * Si[v] + Std[v]) : U[v];
  }
}

void grim_oracle_residual(const grim_oracle_params* p,
                          const grim_oracle_batch* b, double* R) {
  const Vars vars(*p);
  const int64_t n = b->n;
#pragma omp parallel for schedule(static)
  for (int64_t z = 0; z < n; z++) {
    ZoneState<double> s;
    s.load(*p, *b, z);
    double r[GRIM_ORACLE_MAX_VARS];
    computeResidual(s.prim, s.ctx, r);
    for (int v = 0; v < vars.numFluidVars; v++) R[v * n + z] = r[v];
  }
}

void grim_oracle_fd_jacobian(const grim_oracle_params* p,
                             const grim_oracle_batch* b, double* J) {
  const Vars vars(*p);
  const int nf = vars.numFluidVars;
  const int64_t n = b->n;
#pragma omp parallel for schedule(static)
  for (int64_t z = 0; z < n; z++) {
    ZoneState<double> s;
    s.load(*p, *b, z);
    double r[GRIM_ORACLE_MAX_VARS], A[GRIM_ORACLE_MAX_VARS * GRIM_ORACLE_MAX_VARS];
    computeResidual(s.prim, s.ctx, r);
    fdJacobian(s.prim, r, s.ctx, A);
    for (int k = 0; k < nf * nf; k++) J[k * n + z] = A[k];
  }
}

void grim_oracle_dual_jacobian(const grim_oracle_params* p,
                               const grim_oracle_batch* b, double* J) {
  const Vars vars(*p);
  const int nf = vars.numFluidVars;
  const int64_t n = b->n;
#pragma omp parallel for schedule(static)
  for (int64_t z = 0; z < n; z++) {
    ZoneState<Dual> s;
    s.load(*p, *b, z);
    for (int v = 0; v < nf; v++) s.prim[v] = Dual::seed(s.prim[v].v, v);
    Dual r[GRIM_ORACLE_MAX_VARS];
    computeResidual(s.prim, s.ctx, r);
    for (int i = 0; i < nf; i++)
      for (int j = 0; j < nf; j++) J[(i * nf + j) * n + z] = r[i].d[j];
  }
}

int64_t grim_oracle_newton(const grim_oracle_params* p,
                           const grim_oracle_batch* b, int mode,
                           int32_t* iters, double* finalL2, int32_t* resEvals,
                           int nthreads) {
  setThreads(nthreads);
  const Vars vars(*p);
  const int nf = vars.numFluidVars;
  const int64_t n = b->n;
  int64_t unconverged = 0;
  if (mode == GRIM_ORACLE_NEWTON_PER_ZONE) {
#pragma omp parallel for schedule(static) reduction(+ : unconverged)
    for (int64_t z = 0; z < n; z++) {
      ZoneState<double> s;
      s.load(*p, *b, z);
      NewtonZoneStats st{};
      newtonZone(s.prim, s.ctx, &st);
      for (int v = 0; v < nf; v++) b->prim[v * n + z] = s.prim[v];
      if (iters) iters[z] = st.iterations;
      if (finalL2) finalL2[z] = st.finalL2;
      if (resEvals) resEvals[z] = st.residualEvals;
      unconverged += st.converged ? 0 : 1;
    }
    return unconverged;
  }
  // GLOBAL: grim's lock-step loop, src/timestepper/solve.cpp:10-200.  Every
  // zone receives an update in each iteration that runs; the loop ends when
  // no zone has sum R^2 > atol.
  std::vector<double> Rall((size_t)nf * n), l2all(n);
  std::vector<int32_t> evals(n, 0);
  int updates = 0;
  bool done = false;
  for (int it = 0; it < p->maxNonLinearIter; it++) {
    int64_t count = 0;
#pragma omp parallel for schedule(static) reduction(+ : count)
    for (int64_t z = 0; z < n; z++) {
      ZoneState<double> s;
      s.load(*p, *b, z);
      applySolveFloors(s.prim, *p);
      b->prim[0 * n + z] = s.prim[RHO];
      b->prim[1 * n + z] = s.prim[UU];
      double r[GRIM_ORACLE_MAX_VARS];
      computeResidual(s.prim, s.ctx, r);
      evals[z] += 1;
      for (int v = 0; v < nf; v++) Rall[v * n + z] = r[v];
      l2all[z] = residualL2(r, nf);
      count += (l2all[z] > p->nonlinearsolve_atol) ? 1 : 0;
    }
    if (count == 0) {
      done = true;
      break;
    }
#pragma omp parallel for schedule(static)
    for (int64_t z = 0; z < n; z++) {
      // Zones outside the bulk (mask 0: zero residual, zero Jacobian) take
      // no update; grim solves their singular systems and leaves non-finite
      // values there until the boundary conditions overwrite them.
      if (b->residualMask && b->residualMask[z] == 0.) continue;
      ZoneState<double> s;
      s.load(*p, *b, z);
      double r[GRIM_ORACLE_MAX_VARS];
      for (int v = 0; v < nf; v++) r[v] = Rall[v * n + z];
      NewtonZoneStats st{};
      newtonUpdate(s.prim, r, l2all[z], s.ctx, &st);
      evals[z] += st.residualEvals;
      for (int v = 0; v < nf; v++) b->prim[v * n + z] = s.prim[v];
    }
    updates++;
  }
  if (!done) {
#pragma omp parallel for schedule(static)
    for (int64_t z = 0; z < n; z++) {
      ZoneState<double> s;
      s.load(*p, *b, z);
      double r[GRIM_ORACLE_MAX_VARS];
      computeResidual(s.prim, s.ctx, r);
      l2all[z] = residualL2(r, nf);
    }
  }
  for (int64_t z = 0; z < n; z++) {
    if (iters) iters[z] = updates;
    if (finalL2) finalL2[z] = l2all[z];
    if (resEvals) resEvals[z] = evals[z];
    unconverged += (l2all[z] > p->nonlinearsolve_atol) ? 1 : 0;
  }
  return unconverged;
}

int grim_oracle_lu_solve(int n, double* A, double* b) { return luSolve(A, b, n); }

void grim_oracle_count_ops(const grim_oracle_params* p,
                           const grim_oracle_batch* b, int64_t* residualOps,
                           int64_t* luOps) {
  const Vars vars(*p);
  const int nf = vars.numFluidVars;
  ZoneState<Count> s;
  s.load(*p, *b, 0);
  Count r[GRIM_ORACLE_MAX_VARS];
  g_ops = OpCounts();
  computeResidual(s.prim, s.ctx, r);
  OpCounts c = g_ops;
  residualOps[0] = c.add;
  residualOps[1] = c.mul;
  residualOps[2] = c.div;
  residualOps[3] = c.sqrt;
  residualOps[4] = c.other;
  Count A[GRIM_ORACLE_MAX_VARS * GRIM_ORACLE_MAX_VARS];
  fdJacobian(s.prim, r, s.ctx, A);
  Count x[GRIM_ORACLE_MAX_VARS];
  for (int v = 0; v < nf; v++) x[v] = -r[v];
  g_ops = OpCounts();
  luSolve(A, x, nf);
  c = g_ops;
  luOps[0] = c.add;
  luOps[1] = c.mul;
  luOps[2] = c.div;
  luOps[3] = c.sqrt;
  luOps[4] = c.other;
}

void grim_oracle_set_threads(int nthreads) { setThreads(nthreads); }

void grim_oracle_count_live_ops(const grim_oracle_params* p, const grim_oracle_batch* b,
                                int64_t* residualOps) {
  const Vars vars(*p);
  ZoneState<Traced> s;
  s.load(*p, *b, 0);
  g_tape.reset();  // everything load() formed is a per-zone constant
  for (int v = 0; v < vars.dof; v++) s.prim[v] = Traced::variable(oval(s.prim[v]));
  Traced r[GRIM_ORACLE_MAX_VARS];
  computeResidual(s.prim, s.ctx, r);
  long long c[6];
  liveCounts(r, vars.numFluidVars, c);
  for (int k = 0; k < 6; k++) residualOps[k] = c[k];
  g_tape.reset();
  g_tape.nodes.shrink_to_fit();
}

}  // extern "C"
