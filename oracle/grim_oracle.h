/* grim oracle C interface.
 *
 * TEST INFRASTRUCTURE ONLY: the CPU FP64 restatement of grim's per-zone
 * implicit solve, used as the checker for the HIP library.  Loaded by tests/
 * (ctypes), __graft_entry__.smoke() and bench.py's cpu_baseline leg only.
 *
 * Array layout matches the product's batched-zone layout: each variable is a
 * contiguous plane of n doubles (structure of arrays, zone index fastest).
 * Geometry planes (16 per zone):
 *   0..9   gCov 00 01 02 03 11 12 13 22 23 33
 *   10..13 gCon 00 01 02 03
 *   14     alpha = 1/sqrt(-gCon00)
 *   15     g = sqrt(-det gCov)
 */
#ifndef GRIM_ORACLE_H_
#define GRIM_ORACLE_H_
#include <stdint.h>

#define GRIM_ORACLE_MAX_VARS 10
#define GRIM_ORACLE_GEOM_PLANES 16

#define GRIM_ORACLE_CLOSURE_LINEAR_MODES 0
#define GRIM_ORACLE_CLOSURE_PER_ZONE 1
#define GRIM_ORACLE_CLOSURE_TORUS 2
#define GRIM_ORACLE_CLOSURE_SHOCK_TESTS 3
#define GRIM_ORACLE_CLOSURE_BUOYANCY 4

#define GRIM_ORACLE_NEWTON_PER_ZONE 0
#define GRIM_ORACLE_NEWTON_GLOBAL 1

#ifdef __cplusplus
extern "C" {
#endif

/* The subset of grim's `namespace params` (src/params.hpp:96-210) that the
 * per-zone solve reads, plus the closure selector. */
typedef struct grim_oracle_params {
  int conduction;
  int viscosity;
  int highOrderTermsConduction;
  int highOrderTermsViscosity;
  double adiabaticIndex;
  double rhoFloorInFluidElement;
  double uFloorInFluidElement;
  double bSqrFloorInFluidElement;
  double temperatureFloorInFluidElement;
  double ConductionAlpha;
  double ViscosityAlpha;
  int closure;
  double tau0;
  double ConductionClosureFactor;
  double ViscosityClosureFactor;
  int maxNonLinearIter;
  int maxLineSearchIters;
  double nonlinearsolve_atol;
  double JacobianAssembleEpsilon;
  double linesearchfloor;
} grim_oracle_params;

/* One batch of independent zones.  prim/primOld/primCoef have dof planes,
 * rhs has numFluidVars planes: the frozen residual terms folded into
 * consOld (R = (U(P) - rhs)/dtPrime + 1/2 S_impl(P) + S_td(P)).
 * primOld/primCoef may be NULL for ideal MHD; primCoef NULL means primOld.
 * closure: 2 planes (conduction alpha, viscosity alpha) or NULL.
 * radius: 1 plane (torus closure) or NULL.
 * residualMask: 1 plane multiplying each zone's residual rows (grim's
 * residualMask, src/timestepper/residual.cpp:174-177, generalised from
 * {0,1} to a positive per-zone weight) or NULL for 1. */
typedef struct grim_oracle_batch {
  int64_t n;
  double* prim;
  const double* geom;
  const double* rhs;
  const double* primOld;
  const double* primCoef;
  const double* closure;
  const double* radius;
  double dtPrime;
  const double* residualMask;
} grim_oracle_batch;

void grim_oracle_default_params(grim_oracle_params* p, int emhd);

/* MKS geometry at (X1, X2) without cylindrification: gCov (16 planes,
 * row-major mu*4+nu), gCon (16 planes), g, alpha, and the 64 connection
 * coefficients Gamma^mu_{nu lambda} (plane mu*16+nu*4+lambda) by grim's
 * central differences with GAMMA_EPS = 1e-5.  Any output may be NULL. */
void grim_oracle_mks_geometry(int64_t n, const double* X1, const double* X2,
                              double a, double hSlope, double* gCov,
                              double* gCon, double* g, double* alpha,
                              double* gamma);

/* Pack the 16 batch geometry planes from a full gCov/gCon/g/alpha set. */
void grim_oracle_pack_geometry(int64_t n, const double* gCov,
                               const double* gCon, const double* g,
                               const double* alpha, double* geom16);

/* Conserved variables U = computeFluxes(0) for each zone (dof planes). */
void grim_oracle_conserved(const grim_oracle_params* p, int64_t n,
                           const double* prim, const double* geom,
                           const double* closure, const double* radius,
                           double* cons);

/* computeFluxes(dir) for each zone (dof planes); dir 0 gives U. */
void grim_oracle_fluxes(const grim_oracle_params* p, int64_t n, const double* prim,
                        const double* geom, const double* closure, const double* radius,
                        int dir, double* flux);

/* fluidElement::set internals of every zone (GRIM_ORACLE_ELEMENT_PLANES
 * planes): uCon[4], uCov[4], bCon[4], bCov[4], bSqr, q, deltaP, pressure,
 * rho, u, tau, chi, nu, TUpDown[4][4] (row-major). */
#define GRIM_ORACLE_ELEMENT_PLANES 41
void grim_oracle_element(const grim_oracle_params* p, int64_t n,
                         const double* prim, const double* geom,
                         const double* closure, const double* radius,
                         double* out);

/* The pieces of the batched residual at batch->prim: U (conserved, fluid
 * rows), Si (implicit sources) and Std (time-derivative sources), each
 * numFluidVars planes. */
void grim_oracle_residual_pieces(const grim_oracle_params* p,
                                 const grim_oracle_batch* b, double* U,
                                 double* Si, double* Std);

/* Folded right-hand side that makes batch->prim an exact root:
 * rhs = U(P) + dtPrime (1/2 S_impl(P) + S_td(P))  (numFluidVars planes). */
void grim_oracle_known_root_rhs(const grim_oracle_params* p,
                                const grim_oracle_batch* b, double* rhs);

/* Residual rows (numFluidVars planes) at batch->prim. */
void grim_oracle_residual(const grim_oracle_params* p,
                          const grim_oracle_batch* b, double* R);

/* grim's forward-difference Jacobian, planes i*nf+j = dR_i/dP_j. */
void grim_oracle_fd_jacobian(const grim_oracle_params* p,
                             const grim_oracle_batch* b, double* J);

/* Exact Jacobian of the residual by forward-mode dual numbers through the
 * same restatement (planes i*nf+j = dR_i/dP_j). */
void grim_oracle_dual_jacobian(const grim_oracle_params* p,
                               const grim_oracle_batch* b, double* J);

/* Newton solve of every zone in place (batch->prim).  mode: PER_ZONE or
 * GLOBAL (grim's lock-step loop).  iters/finalL2/resEvals may be NULL.
 * nthreads <= 0 uses the OpenMP default.  Returns the number of zones that
 * did not reach nonlinearsolve_atol. */
int64_t grim_oracle_newton(const grim_oracle_params* p,
                           const grim_oracle_batch* b, int mode,
                           int32_t* iters, double* finalL2, int32_t* resEvals,
                           int nthreads);

/* In-place dense solve of n x n row-major A x = b (partial pivoting). */
int grim_oracle_lu_solve(int n, double* A, double* b);

/* Operation counts of one residual evaluation and of one LU solve, from the
 * counting scalar type: out[0]=add/sub, out[1]=mul, out[2]=div, out[3]=sqrt,
 * out[4]=other (exp, pow, ...), for zone 0 of the batch. */
void grim_oracle_count_ops(const grim_oracle_params* p,
                           const grim_oracle_batch* b, int64_t* residualOps,
                           int64_t* luOps);

/* OpenMP threads of every later parallel loop of the oracle (n > 0). */
void grim_oracle_set_threads(int nthreads);

/* Live operations of one residual evaluation, from the tracing scalar type:
 * those the residual rows depend on and that depend on the guess (work a
 * compiled evaluation repeats per call; dead code and per-zone constants
 * excluded): out[0]=add/sub, out[1]=mul, out[2]=div, out[3]=sqrt,
 * out[4]=other, out[5]=max/min selects, for zone 0 of the batch. */
void grim_oracle_count_live_ops(const grim_oracle_params* p,
                                const grim_oracle_batch* b, int64_t* residualOps);

/* ------------------------------------------------------------------------
 * Full time step (grim_step.cpp).
 *
 * Grid arrays are grim's `grid::vars` planes: N1T*N2T*N3T doubles per
 * variable, index i + N1T*(j + N2T*k), with numGhost ghost zones on each
 * side of every active dimension (none in X3 when dim < 3, none in X2 when
 * dim < 2).
 *
 * Geometry at one location: GRIM_ORACLE_LOC_PLANES planes
 *   0..9   gCov 00 01 02 03 11 12 13 22 23 33
 *   10..13 gCon 00 01 02 03
 *   14     alpha
 *   15     g
 *   16..21 gCon 11 12 13 22 23 33
 *   22     r (xCoords[X1])
 *   23     theta (xCoords[X2])
 * (planes 0..15 are the batched-zone layout above).  The connection at the
 * zone centre has 64 planes, Gamma^l_{k n} at plane l*16 + k*4 + n.
 * ------------------------------------------------------------------------ */
#define GRIM_ORACLE_LOC_PLANES 24

#define GRIM_ORACLE_METRIC_MINKOWSKI 0
#define GRIM_ORACLE_METRIC_MKS 1
#define GRIM_ORACLE_BC_PERIODIC 0
#define GRIM_ORACLE_BC_OUTFLOW 1
#define GRIM_ORACLE_BC_MIRROR 2
#define GRIM_ORACLE_RECON_MINMOD 0
#define GRIM_ORACLE_RECON_WENO5 1
#define GRIM_ORACLE_RECON_PPM 2
#define GRIM_ORACLE_RIEMANN_HLL 0
#define GRIM_ORACLE_RIEMANN_LLF 1

typedef struct grim_oracle_grid {
  int N1, N2, N3, dim, numGhost;
  double X1Start, X1End, X2Start, X2End, X3Start, X3End;
  int metric;
  double blackHoleSpin, hSlope;
  int DerefineThetaHorizon, DoCylindrify;
  double X1cyl, X2cyl;
  /* boundaries (src/params.hpp:48-54 names) */
  int boundaryLeft, boundaryRight, boundaryBottom, boundaryTop,
      boundaryBack, boundaryFront;
  int reconstruction, riemannSolver;
  double slopeLimTheta;
  double CourantFactor, maxDtIncrement;
  int residualScaling; /* 0: mask 1 in the bulk; 1: 1/max_v |rhs_v| */
} grim_oracle_grid;

/* Geometry of every zone of the grid.  center, faceX1 (LEFT), faceX2
 * (BOTTOM) hold GRIM_ORACLE_LOC_PLANES planes each; gamma 64 planes (may be
 * NULL).  X3 faces use the centre geometry (grim passes geomCenter for X3,
 * src/timestepper/fvmfluxes.cpp:147-151). */
typedef struct grim_oracle_geometry {
  double* center;
  double* faceX1;
  double* faceX2;
  double* gamma;
} grim_oracle_geometry;

int64_t grim_oracle_grid_zones(const grim_oracle_grid* c);
/* xCoords (r, theta, phi for MKS) at the CENTER of every zone: 3 planes. */
void grim_oracle_grid_xcoords(const grim_oracle_grid* c, double* x3);
void grim_oracle_build_geometry(const grim_oracle_grid* c, grim_oracle_geometry* gm);

/* grid::communicate for one rank (periodic ghost fill) followed by
 * boundaries::applyBoundaryConditions, in place on nvars planes. */
void grim_oracle_apply_boundaries(const grim_oracle_grid* c, int nvars, double* prim);

/* reconstruction::reconstruct along dir (0,1,2) for nvars planes: left =
 * value at the i-1/2 face of zone i, right = at i+1/2.  Written for the
 * zones whose stencil lies inside the grid; other entries untouched. */
void grim_oracle_reconstruct(const grim_oracle_grid* c, int nvars, const double* prim,
                             int dir, double* left, double* right);

/* Reconstruction + riemannSolver::solve: flux planes (dof) at the lower
 * face of every zone along dir, for faces whose stencil lies in the grid. */
void grim_oracle_riemann_flux(const grim_oracle_grid* c, const grim_oracle_params* p,
                              const grim_oracle_geometry* gm, const double* prim,
                              int dir, double* flux);

/* computeDivOfFluxes: fluxes along each active direction, flux-CT for
 * dim >= 2, the torus's applyProblemSpecificFluxFilter when flux_filter != 0
 * (torus/problem.cpp:1243-1285), divergence (dof planes, bulk zones only).
 * fluxes (dim*dof planes, after CT and the filter) may be NULL. */
void grim_oracle_div_fluxes(const grim_oracle_grid* c, const grim_oracle_params* p,
                            const grim_oracle_geometry* gm, const double* prim,
                            double* divF, double* fluxes, int flux_filter);

/* The torus noise stream: uniform in [0, 1) of zone z = 0..n-1 (splitmix64
 * of seed + 0x9E3779B97F4A7C15 (32 z + 1), top 53 bits). */
void grim_oracle_uniform(uint64_t seed, int64_t n, double* out);

/* fluidElement::computeExplicitSources of the element set at prim (dof
 * planes, bulk zones only). */
void grim_oracle_explicit_sources(const grim_oracle_grid* c, const grim_oracle_params* p,
                                  const grim_oracle_geometry* gm, const double* prim,
                                  double* sources);

/* timeStepper::computeDt: returns min(Courant/maxInvDt, maxDtIncrement*dt)
 * with maxInvDt the maximum over every zone of the element set at prim. */
double grim_oracle_compute_dt(const grim_oracle_grid* c, const grim_oracle_params* p,
                              const grim_oracle_geometry* gm, const double* prim,
                              double dt);

/* Torus problem (src/problem/torus/params.cpp, problem.cpp): initial
 * conditions and the per-stage hooks.  minRho/minU: optional floor planes
 * pow(r, slope)*ampl (NULL: computed here). */
typedef struct grim_oracle_torus {
  double InnerEdgeRadius, PressureMaxRadius, Adiabat, MinPlasmaBeta;
  int MagneticLoops, UseMADdisk;
  double InitialPerturbationAmplitude;
  double MaxLorentzFactor;
  double RhoFloorAmpl, UFloorAmpl, RhoFloorSlope, UFloorSlope;
  double BsqrOverRhoMax, BsqrOverUMax;
  uint64_t seed; /* counter-based noise replacing PETSc's RAND48 stream */
  const double* minRho;
  const double* minU;
} grim_oracle_torus;

/* FM torus + SANE/MAD loop field + floors (problem.cpp:145-540), all zones. */
void grim_oracle_torus_ic(const grim_oracle_grid* c, const grim_oracle_params* p,
                          const grim_oracle_torus* t, const grim_oracle_geometry* gm,
                          double* prim);
/* applyFloor (problem.cpp:542-712), all zones, in place. */
void grim_oracle_apply_floor(const grim_oracle_grid* c, const grim_oracle_params* p,
                             const grim_oracle_torus* t, const grim_oracle_geometry* gm,
                             double* prim);
/* setProblemSpecificBCs: inflowCheck then fixPoles (problem.cpp:941-1240). */
/* fullStepDiagnostics' state hooks (problem.cpp:719-746): the NaN scrub,
 * then applyFloor. */
void grim_oracle_full_step_diagnostics(const grim_oracle_grid* c, const grim_oracle_params* p,
                                       const grim_oracle_torus* t, const grim_oracle_geometry* gm,
                                       double* prim);
void grim_oracle_torus_bcs(const grim_oracle_grid* c, const grim_oracle_params* p,
                           const grim_oracle_torus* t, const grim_oracle_geometry* gm,
                           double* prim);

/* The torus's observers (src/problem/torus/ObserversTorus.hpp:6-251) of the
 * element set at prim, over the bulk: out[9] = baryon mass, magnetic and
 * thermal energy, mdot in, out, out unbound, out relativistic, rho max,
 * beta min. */
void grim_oracle_torus_observers(const grim_oracle_grid* c, const grim_oracle_params* p,
                                 const grim_oracle_geometry* gm, const double* prim,
                                 double* out);

/* fullStepDiagnostics' WriteData variables (src/problem/torus/problem.cpp:
 * 808-875) of the element set at prim: (dof + 5) planes of N1*N2*N3 bulk
 * zones, X1 fastest: rho, u, v^1..3, [q], [deltaP], B^1..3, b^2, gamma
 * (params.cpp:136-147), then the x coordinates of the zone centres. */
void grim_oracle_dump_vars(const grim_oracle_grid* c, const grim_oracle_params* p,
                           const grim_oracle_geometry* gm, const double* prim, double* out);

typedef struct grim_oracle_step_stats {
  double dt;
  int64_t unconverged[2];   /* bulk zones above atol after each stage */
  int32_t maxIterations[2]; /* per stage */
  int64_t nonfinite[2];     /* bulk zones whose final residual is not finite */
} grim_oracle_step_stats;

/* timeStepper::timeStep (src/timestepper/timestep.cpp:3-334): computeDt from
 * primDt (the state elemOld was last set from: primOld at the end of the
 * previous step, or the initial state), then the half and full stages with
 * the torus hooks when t != NULL.  primOld (dof planes) is advanced in
 * place and copied to primDt; *dt and *time advance.
 * semantics: GRIM_ORACLE_NEWTON_PER_ZONE or GLOBAL. */
void grim_oracle_time_step(const grim_oracle_grid* c, const grim_oracle_params* p,
                           const grim_oracle_geometry* gm, const grim_oracle_torus* t,
                           double* primOld, double* primDt, double* dt, double* time,
                           int semantics, grim_oracle_step_stats* stats);

/* setProblemSpecificBCs of bondi_inflow (src/problem/bondi_inflow/
 * problem.cpp:157-183): every ghost zone of primOld set from prim_ic (dof
 * planes, kept by the caller) after the half stage's boundary conditions;
 * NULL switches the hook off.  Process-wide. */
void grim_oracle_set_ghost_state(const double* prim_ic);

/* setProblemSpecificBCs of buoyancy_instabilities (src/problem/
 * buoyancy_instabilities/problem.cpp:221-293): in both stages after the
 * boundary conditions, the stage state's zones of X1 index 0..numGhost and
 * N1+numGhost..N1+2 numGhost-1 take rho, u, u~^1 from prim_ic (3 planes,
 * kept by the caller), u~^2 = u~^3 = B^i = 0 and q~ = 0; NULL switches the
 * hook off.  Process-wide. */
void grim_oracle_set_x1_side_state(const double* prim_ic);

/* The two stages as separate calls, for stage-level parity tests.
 * stage 0 (half): boundaries on primOld, consOld, sources, divF, induction,
 * Newton; writes primHalf (dof planes) and consOld (dof planes).
 * stage 1 (full): boundaries on primHalf, sources, divF, induction, Newton
 * from the guess primHalf; writes primNew. */
void grim_oracle_stage(const grim_oracle_grid* c, const grim_oracle_params* p,
                       const grim_oracle_geometry* gm, const grim_oracle_torus* t, int stage,
                       double dt, double* primOld, double* primHalf, double* consOld,
                       double* primNew, int semantics, grim_oracle_step_stats* stats);

#ifdef __cplusplus
}
#endif
#endif
