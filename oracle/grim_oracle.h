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

#ifdef __cplusplus
}
#endif
#endif
