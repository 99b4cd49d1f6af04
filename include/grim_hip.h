/* grim_hip.h -- C ABI of the MI355X (gfx950) implementation of grim's
 * per-timestep implicit hot path.
 *
 * grim itself has no FFI: its hot path is C++ member functions on
 * timeStepper / fluidElement, linked statically.  Each entry point below
 * names the reference routine it replaces (file:line in the grim source
 * tree).  All device arrays are caller-owned, structure-of-arrays planes of
 * n doubles (zone index fastest), enqueued on the caller's HIP stream.
 *
 * Errors: every entry point returns GRIM_OK (0) or a negative grim_status;
 * grim_last_error() gives a message for the calling thread.  Numerical
 * non-convergence is not an error: it is reported in grim_newton_stats.
 */
#ifndef GRIM_HIP_H_
#define GRIM_HIP_H_
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ihipStream_t* grim_stream_t; /* = hipStream_t */

typedef enum grim_status {
  GRIM_OK = 0,
  GRIM_ERR_HIP = -1,        /* a HIP runtime call failed */
  GRIM_ERR_BAD_ARG = -2,    /* null pointer, bad size or unsupported option */
  GRIM_ERR_NO_DEVICE = -3,  /* no usable gfx950 device */
  GRIM_ERR_OOM = -4         /* device allocation failed */
} grim_status;

#define GRIM_MAX_VARS 10
#define GRIM_GEOM_PLANES 16

/* Closure hook setFluidElementParameters (src/physics/physics.hpp:58). */
typedef enum grim_closure {
  GRIM_CLOSURE_LINEAR_MODES = 0, /* src/problem/linear_modes/problem.cpp:3-8 */
  GRIM_CLOSURE_PER_ZONE = 1,     /* same form, per-zone conduction/viscosity alpha */
  GRIM_CLOSURE_TORUS = 2,        /* src/problem/torus/problem.cpp:7-55 */
  GRIM_CLOSURE_SHOCK_TESTS = 3,  /* src/problem/shock_tests/problem.cpp:13-18: tau = 0.1,
                                    chi (nu) = ConductionAlpha (ViscosityAlpha) c_s^2 tau */
  GRIM_CLOSURE_BUOYANCY = 4      /* src/problem/buoyancy_instabilities/problem.cpp:11-21:
                                    chi = sqrt(r), tau = 1.2 chi / c_s^2 (r: the radius
                                    of the zone's xCoords); sets no viscosity coefficient,
                                    so it takes conduction only */
} grim_closure;

/* The part of grim's `namespace params` (src/params.hpp:96-210) that the
 * per-zone implicit solve reads. */
typedef struct grim_params {
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
  int closure; /* grim_closure */
  double tau0; /* relaxation time of GRIM_CLOSURE_PER_ZONE */
  double ConductionClosureFactor;
  double ViscosityClosureFactor;
  int maxNonLinearIter;
  int maxLineSearchIters;
  double nonlinearsolve_atol;
  double JacobianAssembleEpsilon;
  double linesearchfloor;
} grim_params;

/* A batch of independent zones for the per-zone implicit solve.
 * Geometry planes (GRIM_GEOM_PLANES): gCov 00 01 02 03 11 12 13 22 23 33,
 * gCon 00 01 02 03, alpha, g=sqrt(-det gCov).
 * The residual of zone z is (src/timestepper/residual.cpp:3-180)
 *   R = mask_z * N[(U(P) - rhs)/dt_prime + 1/2 S_impl(P) + S_td(P)]
 * where rhs folds consOld and every term frozen during the solve
 * (rhs = consOld - dt_prime (divF + S_expl + 1/2 S_impl,old)), N is grim's
 * EMHD row normalisation and mask_z grim's residualMask. */
typedef struct grim_zone_batch {
  int64_t n;                  /* number of zones */
  double* prim;               /* dof planes, in/out (fluid rows updated) */
  const double* geom;         /* GRIM_GEOM_PLANES planes */
  const double* rhs;          /* numFluidVars planes */
  const double* prim_old;     /* dof planes (EMHD: elemOld) or NULL */
  const double* prim_coef;    /* dof planes (EMHD: elemHalfStep) or NULL = prim_old */
  const double* closure;      /* 2 planes (conduction, viscosity alpha) or NULL */
  const double* radius;       /* 1 plane (torus and buoyancy closures) or NULL */
  const double* residual_mask;/* 1 plane or NULL (= 1) */
  double dt_prime;            /* dt/2 (half step) or dt (full step) */
  const double* dt_prime_device; /* optional: device address holding dt_prime,
                                    read by the kernels instead of the value
                                    (steps whose dt is computed on the device,
                                    e.g. under HIP graph capture); NULL = use
                                    dt_prime */
  const double* guess;        /* optional: numFluidVars planes of the starting
                                 fluid primitives, read instead of prim's own
                                 fluid rows; the solve's result still goes to
                                 prim (B rows are always read from prim).  It
                                 folds into the solve the copy grim makes
                                 before it (prim = primOld, timestep.cpp:
                                 72-75), so a caller solving again from the
                                 same start needs no copy.  Read by
                                 grim_batched_newton(_noalloc) and the
                                 autotuner only; NULL = start from prim. */
  int64_t geom_columns;       /* 0: geom and radius hold one value per zone
                                 (planes of n).  > 0: they hold one value per
                                 column (planes of geom_columns values), zone
                                 z reading entry z % geom_columns: a grid
                                 whose metric does not depend on X3 keeps one
                                 value per (X1, X2) column, z % (NT1 NT2)
                                 (grim_geometry_columns).  Must divide n. */
} grim_zone_batch;

typedef enum grim_jacobian_kind { GRIM_JAC_FD = 0, GRIM_JAC_ANALYTIC = 1 } grim_jacobian_kind;
typedef enum grim_lu_precision {
  GRIM_LU_FP64 = 0,
  GRIM_LU_FP32 = 1,
  GRIM_LU_FP32_FP64 = 2 /* analytic Jacobian only: FP32 Jacobian and LU for the
                           first phase1_iters updates of each zone, FP64 for
                           the rest (always run as the compacting solve) */
} grim_lu_precision;
typedef enum grim_semantics {
  GRIM_NEWTON_PER_ZONE = 0, /* each zone stops at its own convergence */
  GRIM_NEWTON_GLOBAL = 1    /* grim's lock-step loop: all zones iterate while any is unconverged */
} grim_semantics;

typedef struct grim_newton_opts {
  int jacobian;      /* grim_jacobian_kind */
  int lu_precision;  /* grim_lu_precision */
  int semantics;     /* grim_semantics */
  int compaction;    /* 1: every zone gets phase1_iters updates, then the
                        still unconverged zones are compacted into a dense
                        work list that a persistent kernel drains (at most
                        2^30 zones); 2: one pass in which, after
                        phase1_iters updates, each workgroup gathers its
                        unconverged zones onto its first lanes (no work
                        list).  Results equal the one-pass solve's bit for
                        bit in every case. */
  int block_size;    /* threads per workgroup (0 = library default) */
  int fused;         /* FD path only: 0 = grim's operation order, no fused
                        multiply-add (bit-reproducible against a CPU FP64
                        evaluation); 1 = fused multiply-adds and reciprocal
                        forms.  The mixed path is always fused. */
  int phase1_iters;  /* compaction: Newton updates before the work list is
                        compacted (0 = library default) */
  int phase2_group;  /* compaction: chunks of 64 work-list zones a phase-2
                        wavefront takes per grab, 1..8 (0 = default) */
  void* workspace;   /* compaction: device scratch of at least
                        grim_newton_workspace_bytes(n) bytes, or NULL to let
                        the library allocate it on the stream */
  int64_t workspace_bytes;
} grim_newton_opts;

/* Device scratch the compacting solve needs for n zones. */
int64_t grim_newton_workspace_bytes(int64_t n);

/* Per-zone outputs are optional device arrays (NULL to skip). */
typedef struct grim_newton_stats {
  int32_t* iterations;   /* Newton updates applied, per zone */
  int32_t* residual_evals;/* residual evaluations, per zone */
  double* final_l2;      /* sum of squared residual rows at the returned state */
} grim_newton_stats;

/* Library defaults of grim's params (src/problem/linear_modes/params.cpp:40-69). */
void grim_default_params(grim_params* p, int emhd);

const char* grim_last_error(void);

/* 0 when a gfx950 device is visible; fills the device name. */
int grim_device_info(char* name, int name_len);

/* fluidElement::set + computeFluxes(0) (src/physics/physics.cpp:53-217):
 * conserved variables U (dof planes) of every zone. */
int grim_conserved(const grim_params* p, int64_t n, const double* prim,
                   const double* geom, const double* closure,
                   const double* radius, double* cons, grim_stream_t stream);

/* timeStepper::computeResidual (src/timestepper/residual.cpp:3-180):
 * residual rows (numFluidVars planes) at batch->prim. */
int grim_residual(const grim_params* p, const grim_zone_batch* b, double* R,
                  grim_stream_t stream);

/* Jacobian assembly of timeStepper::solve (src/timestepper/solve.cpp:82-109):
 * J planes i*nf+j = dR_i/dP_j, by grim's forward differences
 * (GRIM_JAC_FD) or by the analytic tangent (GRIM_JAC_ANALYTIC). */
int grim_jacobian(const grim_params* p, const grim_zone_batch* b, int jacobian,
                  double* J, grim_stream_t stream);

/* Matrix-free products with the analytic derivative J = dR/dP of the
 * per-zone residual at batch->prim (numFluidVars planes in and out), in
 * FP64 (precision = GRIM_LU_FP64) or FP32 (GRIM_LU_FP32) arithmetic.  They
 * replace grim's finite-difference Jacobian (src/timestepper/solve.cpp:82-109,
 * one residual evaluation of src/timestepper/residual.cpp:3-180 per
 * primitive):
 *  grim_residual_tangent: Jv_z = J_z v_z (the hand-written tangent);
 *  grim_residual_adjoint: JTw_z = J_z^T w_z (the hand-written reverse sweep,
 *    J never formed).  <J v, w> = <v, J^T w> to rounding. */
int grim_residual_tangent(const grim_params* p, const grim_zone_batch* b, const double* v, double* Jv,
                          int precision, grim_stream_t stream);
int grim_residual_adjoint(const grim_params* p, const grim_zone_batch* b, const double* w, double* JTw,
                          int precision, grim_stream_t stream);

/* timeStepper::batchLinearSolve (src/timestepper/solve.cpp:203-230): solves
 * A_z x_z = b_z for every zone, A planes i*nvar+j, partial pivoting, in FP64
 * or FP32 arithmetic. */
int grim_batch_linear_solve(int64_t n, int nvar, const double* A,
                            const double* b, double* x, int lu_precision,
                            grim_stream_t stream);

/* timeStepper::solve (src/timestepper/solve.cpp:3-201): Newton with
 * line search on every zone of the batch, in place. */
int grim_batched_newton(const grim_params* p, const grim_zone_batch* b,
                        const grim_newton_opts* opts, grim_newton_stats* stats,
                        grim_stream_t stream);

/* No counterpart in the reference.  grim_batched_newton for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc): it calls no hipMalloc, hipFree or
 * allocating library routine.  The compacting solve (compaction = 1, or
 * GRIM_LU_FP32_FP64) keeps its work list in a library-owned device array of
 * compaction workspace for GRIM_NOALLOC_MAX_ZONES zones (about 35 MB, one
 * copy per device, held until the process ends).  Same options and results
 * as grim_batched_newton, except: per-zone semantics only, opts->workspace
 * must be NULL, and at most GRIM_NOALLOC_MAX_ZONES zones (GRIM_ERR_BAD_ARG
 * otherwise).  Every call on one device shares that device's array: calls
 * on one stream are ordered, calls that overlap on one device (other
 * streams, other threads) race.  A call rewrites every part of the array it
 * reads before reading it, so nothing left by an earlier call matters. */
#define GRIM_NOALLOC_MAX_ZONES 8388608
int grim_batched_newton_noalloc(const grim_params* p, const grim_zone_batch* b,
                                const grim_newton_opts* opts, grim_newton_stats* stats,
                                grim_stream_t stream);

/* Project autotuner for the batched solve: times the candidate launch
 * configurations (workgroup size, compaction on/off, phase-1 budget,
 * phase-2 grid) on this batch -- on a scratch copy of batch->prim, so the
 * batch is left unchanged -- and writes the fastest into *opts (jacobian,
 * lu_precision and fused are taken from *opts).  Synchronises the stream. */
int grim_autotune_newton(const grim_params* p, const grim_zone_batch* b,
                         grim_newton_opts* opts, grim_stream_t stream);

/* Number of zones whose final_l2 exceeds tol (device reduction, synchronises
 * the stream). */
int grim_count_above(int64_t n, const double* final_l2, double tol,
                     int64_t* count, grim_stream_t stream);

/* Synthetic isolated-zone batch on modified Kerr-Schild geometry (no grim
 * counterpart: it generates the inputs the reference gets from a stage):
 * spin a, hSlope h, counter-based RNG with the given seed.  Fills geom (16),
 * prim_star (dof, the known root), guess (dof), and for EMHD prim_old (dof)
 * and closure (2).  See grim_amd/csrc/batchgen.hpp for the sampling rule. */
int grim_generate_kerr_schild_batch(const grim_params* p, int64_t n,
                                    uint64_t seed, double a, double h,
                                    double* geom, double* prim_star,
                                    double* guess, double* prim_old,
                                    double* closure, grim_stream_t stream);

/* rhs (numFluidVars planes) that makes batch->prim an exact root of the
 * batched residual, and optionally the residual mask 1/max_v|rhs_v| that
 * turns the absolute tolerance into a relative one. */
int grim_known_root_rhs(const grim_params* p, const grim_zone_batch* b,
                        double* rhs, double* mask, grim_stream_t stream);

/* ------------------------------------------------------------------------
 * The full time step on a grid (grim's timeStepper).
 *
 * A context owns every device array of one grid: the primitive variables
 * (grim's grid::vars planes, N1T*N2T*N3T doubles each, index
 * i + N1T*(j + N2T*k), numGhost ghost zones per side of each active
 * dimension), the geometry and the work arrays.  All work is enqueued on
 * the context's stream.
 *
 * Geometry of one location: GRIM_LOC_PLANES planes
 *   0..9   gCov 00 01 02 03 11 12 13 22 23 33
 *   10..13 gCon 00 01 02 03,  14 alpha,  15 g = sqrt(-det gCov)
 *   16..21 gCon 11 12 13 22 23 33,  22 r (xCoords[X1]),  23 theta (xCoords[X2])
 * (planes 0..15 are the batched-zone layout).  The connection at the zone
 * centre has 64 planes, Gamma^l_{k n} at plane l*16 + k*4 + n.
 * ------------------------------------------------------------------------ */
#define GRIM_LOC_PLANES 24

typedef enum grim_metric { GRIM_METRIC_MINKOWSKI = 0, GRIM_METRIC_MKS = 1 } grim_metric;
/* src/params.hpp:48-54 (DIRICHLET is not supported) */
typedef enum grim_boundary {
  GRIM_BC_PERIODIC = 0, GRIM_BC_OUTFLOW = 1, GRIM_BC_MIRROR = 2
} grim_boundary;
/* src/params.hpp:72-86 */
typedef enum grim_reconstruction {
  GRIM_RECON_MINMOD = 0, GRIM_RECON_WENO5 = 1, GRIM_RECON_PPM = 2
} grim_reconstruction;
typedef enum grim_riemann { GRIM_RIEMANN_HLL = 0, GRIM_RIEMANN_LLF = 1 } grim_riemann;

/* Grid, geometry, boundary and scheme members of grim's `namespace params`
 * (e.g. src/problem/torus/params.cpp:8-105). */
typedef struct grim_grid_config {
  int N1, N2, N3, dim, numGhost;
  double X1Start, X1End, X2Start, X2End, X3Start, X3End;
  int metric; /* grim_metric */
  double blackHoleSpin, hSlope;
  int DerefineThetaHorizon, DoCylindrify;
  double X1cyl, X2cyl;
  int boundaryLeft, boundaryRight, boundaryBottom, boundaryTop, boundaryBack,
      boundaryFront; /* grim_boundary */
  int reconstruction; /* grim_reconstruction */
  int riemannSolver;  /* grim_riemann */
  double slopeLimTheta;
  double CourantFactor, maxDtIncrement;
  /* 0: grim's residual mask (1 in the bulk): nonlinearsolve_atol bounds the
   * absolute sum of squared residual rows.  1: the bulk mask is
   * 1/max_v |rhs_v| per zone, so the tolerance bounds the residual relative
   * to the zone's conserved variables (no grim counterpart: an absolute
   * 1e-20 lies below the FP64 rounding floor of zones with large U). */
  int residualScaling;
  /* 0: the explicit stage (face fluxes, explicit sources) in grim's
   * expression order, one rounding per operation, bit-reproducible by a CPU
   * FP64 evaluation.  1: the same mathematics with shared reciprocals, fused
   * multiply-adds and the 40 symmetric connection coefficients (+40 planes
   * of memory); results agree with 0 to rounding (tested to 1e-12). */
  int explicitFast;
} grim_grid_config;

typedef struct grim_ctx grim_ctx;

typedef struct grim_step_stats {
  double dt;                  /* dt taken by the step */
  int64_t unconverged[2];     /* bulk zones above nonlinearsolve_atol after the half / full stage */
  int32_t max_iterations[2];  /* largest per-zone Newton iteration count of each stage */
  int64_t nonfinite[2];       /* bulk zones whose final residual is not finite (grim's
                                 test counts them as converged) */
} grim_step_stats;

/* Number of zones (N1T*N2T*N3T) of a grid; 0 when N1, N2 or N3 lies outside
 * [1, 2^30]. */
int64_t grim_grid_zones(const grim_grid_config* c);

/* timeStepper::timeStepper (src/timestepper/timestepper.cpp:3-386 without
 * the initial conditions): allocates the grid arrays, builds the geometry at
 * the zone centres and lower faces and the connection (geometry.cpp:4-121).
 * opts selects the Newton path of the two solves (NULL: grim's FD-FP64,
 * per-zone convergence).  The primitive variables start at zero. */
int grim_create(const grim_params* p, const grim_grid_config* c,
                const grim_newton_opts* opts, grim_stream_t stream, grim_ctx** ctx);
int grim_destroy(grim_ctx* ctx);

/* Device pointer of primOld (dof planes of grim_grid_zones doubles), and
 * of a geometry array: which = 0 centre, 1 X1 faces (LEFT), 2 X2 faces
 * (BOTTOM), 3 connection (64 planes).  The metrics supported (modified
 * Kerr-Schild, Minkowski) do not depend on X3, so a geometry plane holds
 * one value per (X1, X2) column: grim_geometry_columns(ctx) = NT1 NT2
 * values, zone (i, j, k) at i + NT1 j (one per zone below three
 * dimensions). */
int grim_prim(grim_ctx* ctx, double** prim_old);
int grim_geometry(grim_ctx* ctx, int which, double** geom);
int64_t grim_geometry_columns(const grim_ctx* ctx);
/* Copies of primOld and of a geometry array to / from host (on_device = 0)
 * or device (on_device = 1) memory; synchronise the stream (grim's
 * grid::setVars / getVars, src/grid/gridPy.pyx:185-230).  grim_get_geometry
 * writes every zone's value (planes of grim_grid_zones doubles, each
 * column's value repeated along X3), as grim's geometry arrays hold them. */
int grim_set_prim(grim_ctx* ctx, const double* src, int on_device);
int grim_get_prim(grim_ctx* ctx, double* dst, int on_device);
int grim_get_geometry(grim_ctx* ctx, int which, double* dst, int on_device);
/* time and dt of the timeStepper (dt is the previous step's, which bounds
 * the next one through maxDtIncrement).  Both live in device memory;
 * grim_get_time synchronises the stream to read them. */
int grim_set_time(grim_ctx* ctx, double time, double dt);
int grim_get_time(grim_ctx* ctx, double* time, double* dt);

/* timeStepper::timeStep (src/timestepper/timestep.cpp:3-334): computeDt,
 * then the half and full stages (boundaries, conserved variables, sources,
 * divergence of fluxes with flux-CT, induction, Newton solve).  dt and time
 * advance on the device.  With stats == NULL the call enqueues the whole
 * step without a host synchronisation or allocation, so it may be captured
 * into a HIP graph and replayed; with stats it synchronises to fill them. */
int grim_time_step(grim_ctx* ctx, grim_step_stats* stats);

/* The pieces of the step on caller-provided device arrays (dof planes each):
 *  boundaries::applyBoundaryConditions after grid::communicate
 *    (src/boundary/boundary.cpp:5-243, src/grid/grid.cpp:306-315);
 *  grid::communicate alone (periodic ghost fill);
 *  reconstruction::reconstruct (src/reconstruction/reconstruction.cpp:3-38);
 *  riemannSolver::solve with reconstruction (src/physics/riemannsolver.cpp:202-328):
 *    flux at the lower face of each zone along dir;
 *  timeStepper::computeDivOfFluxes with fluxCT (src/timestepper/fvmfluxes.cpp:3-184,
 *    constrainedtransport.cpp:3-83): divF over the bulk, fluxes (dim*dof planes) may be NULL;
 *  fluidElement::computeExplicitSources (src/physics/physics.cpp:360-586) over the bulk;
 *  timeStepper::computeDt (src/timestepper/timestep.cpp:336-400). */
int grim_apply_boundaries(grim_ctx* ctx, double* prim, int nvars);
int grim_communicate(grim_ctx* ctx, double* prim, int nvars);
int grim_reconstruct(grim_ctx* ctx, const double* prim, int nvars, int dir,
                     double* left, double* right);
int grim_riemann_flux(grim_ctx* ctx, const double* prim, int dir, double* flux);
int grim_div_fluxes(grim_ctx* ctx, const double* prim, double* div_f, double* fluxes);
int grim_explicit_sources(grim_ctx* ctx, const double* prim, double* sources);
int grim_compute_dt(grim_ctx* ctx, const double* prim, double dt, double* new_dt);

/* riemannSolver::solve (src/physics/riemannsolver.cpp:202-328) from
 * reconstructed states: prim_left / prim_right are the face values
 * grim_reconstruct writes (primLeft at the lower face of each zone,
 * primRight at the upper face); flux at the lower face of each zone along
 * dir, from primRight of zone z - 1 and primLeft of zone z at the face
 * geometry (geomLeft/geomRight, geomBottom/geomTop, geomCenter for X3, as
 * grim passes them, timestepper.cpp:236-280 and fvmfluxes.cpp:151). */
int grim_riemann_solve(grim_ctx* ctx, const double* prim_left, const double* prim_right, int dir,
                       double* flux);

/* fluidElement on a whole grid (src/physics/physics.hpp:14-108): the element
 * set from prim (dof planes) at the geometry of a location (0 = zone
 * centres, 1 = X1 lower faces, 2 = X2 lower faces), then
 *  grim_element_fluxes: computeFluxes(dir) (physics.cpp:185-217), dof planes;
 *    dir 0 gives the conserved variables, 1..3 the fluxes along X1..X3;
 *  grim_element_char_speeds: computeMinMaxCharSpeeds(dir) (riemannsolver.cpp:
 *    54-200), dir 0..2 = X1..X3, one plane each;
 *  grim_element_fields: the element's arrays, GRIM_ELEMENT_FIELDS planes in
 *    the order rho, u, pressure, temperature, gammaLorentzFactor, soundSpeed,
 *    bSqr, bNorm, q, deltaP, tau, chi_emhd, nu_emhd, uCon[4], uCov[4],
 *    bCon[4], bCov[4];
 *  grim_element_implicit_sources: computeImplicitSources(sources, tauDamp)
 *    (physics.cpp:320-358) at the zone centres, dof planes;
 *  grim_element_time_deriv_sources: computeTimeDerivSources(elemOld,
 *    elemNew, dt, sources) (physics.cpp:219-318) of the element set from prim
 *    (the coefficient element), elemOld / elemNew set from prim_old /
 *    prim_new, at the zone centres, dof planes.
 * Every zone of the grid is evaluated, ghosts included. */
#define GRIM_ELEMENT_FIELDS 29
int grim_element_fluxes(grim_ctx* ctx, const double* prim, int location, int dir, double* out);
int grim_element_char_speeds(grim_ctx* ctx, const double* prim, int location, int dir, double* min_speed,
                             double* max_speed);
int grim_element_fields(grim_ctx* ctx, const double* prim, int location, double* out);
int grim_element_implicit_sources(grim_ctx* ctx, const double* prim, const double* tau_damp,
                                  double* sources);
int grim_element_time_deriv_sources(grim_ctx* ctx, const double* prim, const double* prim_old,
                                    const double* prim_new, double dt, double* sources);

/* grim_time_step in its pieces, for a caller that runs timeStepper::timeStep
 * (src/timestepper/timestep.cpp:3-334) itself; stage 0 = HALF_STEP, 1 =
 * FULL_STEP.  dt lives in the context's device clock:
 *  grim_step_begin: computeDt (timestep.cpp:8, 336-400) from the state the
 *    last step left;
 *  grim_fold_rhs: everything the stage does before timeStepper::solve --
 *    boundary conditions and problem hooks on the stage state, explicit and
 *    old implicit sources, the divergence of the fluxes with flux-CT, the
 *    induction update of B, and the frozen right-hand side folded into one
 *    array, rhs = consOld - dt' (divF + S_expl + S_impl,old / 2), with the
 *    residual mask (residual.cpp:39-45, 173-177); the guess (the stage
 *    state) goes into the stage's output (primHalfStep, then prim);
 *  grim_solve: timeStepper::solve (solve.cpp:3-201) of the stage on prim
 *    (dof planes, fluid rows updated; NULL = the stage's output); stats, if
 *    given, receive the stage's counts (synchronises);
 *  grim_stage_residual: timeStepper::computeResidual (residual.cpp:3-180) of
 *    the stage at prim, numFluidVars planes;
 *  grim_stage_finish: communicate and the problem's half/full-step
 *    diagnostics on the stage's output;
 *  grim_step_end: primOld = prim, time += dt.
 * begin, fold/solve/finish for stage 0 then 1, and end are exactly
 * grim_time_step. */
int grim_step_begin(grim_ctx* ctx);
int grim_fold_rhs(grim_ctx* ctx, int stage);
int grim_solve(grim_ctx* ctx, int stage, double* prim, grim_step_stats* stats);
int grim_stage_residual(grim_ctx* ctx, int stage, const double* prim, double* residual);
int grim_stage_finish(grim_ctx* ctx, int stage);
int grim_step_end(grim_ctx* ctx);

/* Device arrays of the context under grim's timeStepper member names
 * (timestepper.hpp:58-80): dof planes each for the states and consOld,
 * numFluidVars planes for the sources, rhs and fluxes, one plane for the
 * residual mask; fluxesX2/X3 are NULL below 2 / 3 dimensions.  The fluxes
 * hold the last stage's (or grim_div_fluxes') face fluxes. */
typedef struct grim_stage_arrays {
  double *primOld, *primHalfStep, *prim;
  double *consOld, *sourcesExplicit, *sourcesImplicitOld, *rhs, *residualMask;
  double *fluxesX1, *fluxesX2, *fluxesX3;
} grim_stage_arrays;
int grim_get_stage_arrays(grim_ctx* ctx, grim_stage_arrays* out);

/* Per-zone Newton iteration counts and final sums of squared residual rows
 * of the context's last solve (the full stage after grim_time_step), all
 * zones, into host (on_device = 0, synchronises) or device memory; either
 * pointer may be NULL. */
int grim_last_solve_stats(grim_ctx* ctx, int32_t* iterations, double* final_l2, int on_device);

/* Initial conditions of the linear-mode problem
 * (src/problem/linear_modes/problem.cpp:10-111) into primOld:
 * mode 0 = ideal fast magnetosonic wave, k = (2 pi, 2 pi), on rho = 1,
 * u = 2, B = (0.1, 0.3, 0) (eigenvector derived from the linearised
 * equations; not in the reference); mode 1 = the reference's EMHD mode,
 * k = (2 pi, 4 pi).  amplitude multiplies the eigenvector. */
int grim_init_linear_mode(grim_ctx* ctx, int mode, double amplitude);

/* The torus problem's parameters (src/problem/torus/params.cpp:38-80). */
typedef struct grim_torus_params {
  double InnerEdgeRadius, PressureMaxRadius, Adiabat, MinPlasmaBeta;
  int MagneticLoops, UseMADdisk;
  double InitialPerturbationAmplitude;
  double MaxLorentzFactor;
  double RhoFloorAmpl, UFloorAmpl, RhoFloorSlope, UFloorSlope;
  double BsqrOverRhoMax, BsqrOverUMax;
  uint64_t seed; /* counter-based noise stream replacing PETSc's RAND48 */
} grim_torus_params;

/* timeStepper::initialConditions of the torus (src/problem/torus/
 * problem.cpp:145-540: Fishbone-Moncrief torus, loop field normalised to
 * the minimum plasma beta, floors) into primOld, and from then on the
 * torus's per-stage hooks in grim_time_step: setProblemSpecificBCs
 * (inflowCheck, fixPoles, :941-1240), the flux filter (:1243-1285) and the
 * half/full-step diagnostics (NaN scrub and applyFloor, :714-741).  Needs
 * an MKS grid.  grim_get_geometry(ctx, 4, ...) then returns the two floor
 * planes pow(r, slope) * ampl (rho, u). */
int grim_init_torus(grim_ctx* ctx, const grim_torus_params* t);
/* applyFloor (problem.cpp:542-712) on caller-provided primitive planes. */
int grim_apply_floor(grim_ctx* ctx, double* prim);
/* The state hooks of the torus's fullStepDiagnostics (problem.cpp:719-746)
 * on caller-provided primitive planes, as the full stage of grim_time_step
 * runs them: when a bulk zone holds a NaN in any variable, every zone
 * holding one (ghosts included) is set to zero in all variables; then
 * applyFloor.  The observers and dumps of the same hook are
 * grim_torus_observers and grim_dump_vars. */
int grim_full_step_diagnostics(grim_ctx* ctx, double* prim);
/* setProblemSpecificBCs of the torus on caller-provided primitive planes. */
int grim_torus_bcs(grim_ctx* ctx, double* prim);

/* The torus's on-the-fly observers (src/problem/torus/ObserversTorus.hpp,
 * called from fullStepDiagnostics, problem.cpp:757-776), of the element set
 * at the given primitives, over the bulk zones; volElem = dX1 (* dX2 when
 * dim > 1, * dX3 when dim > 2). */
typedef struct grim_torus_observables {
  /* ComputeEnergyIntegrals (:6-81): sums of rho, b^2 and u, each times
   * volElem * g * gamma / alpha */
  double baryon_mass, magnetic_energy, thermal_energy;
  /* ComputeBoundaryFluxes (:141-251): sums of rho * volElem * g * u^1 over
   * the first (in) and last (out) bulk X1 planes; of the outflow, the part
   * with (1 + u/rho * Gamma) u_0 < -1 (unbound) and of that the part with
   * gamma > 2 (relativistic) */
  double mdot_in, mdot_out, mdot_out_unbound, mdot_out_relativistic;
  /* ComputeMinMaxVariables (:83-139): max rho, min 2 (p + 1e-13) / (b^2 +
   * 1e-18); NaN zones are skipped (fmax / fmin) */
  double rho_max, beta_min;
} grim_torus_observables;

/* Replaces ComputeEnergyIntegrals, ComputeMinMaxVariables and
 * ComputeBoundaryFluxes (ObserversTorus.hpp:6-251) on one rank.  prim:
 * device primitive planes, or NULL for primOld.  Valid on every grid (the
 * reference calls it for the torus).  Sums are combined in a fixed order:
 * repeated calls on the same state return identical values.  Synchronises
 * the context's stream. */
int grim_torus_observers(grim_ctx* ctx, const double* prim, grim_torus_observables* out);

/* Replaces setProblemSpecificBCs of bondi_inflow (src/problem/bondi_inflow/
 * problem.cpp:157-183): from the next step on, every ghost zone of primOld
 * is set from prim_ic (dof planes of grim_grid_zones doubles; host memory,
 * or device memory when on_device != 0; copied, so the caller may free it)
 * after the half stage's boundary conditions.  grim calls the hook in both
 * stages, but it always writes primOld, so only the half stage changes
 * anything.  prim_ic = NULL switches the hook off.  The hook's state buffer
 * stays allocated until grim_destroy and the on/off switch is read on the
 * device, so a step captured into a HIP graph after the first non-NULL call
 * follows later switches and never reads freed memory; a step captured
 * before that call does not contain the hook.  Synchronises the context's
 * stream. */
int grim_set_ghost_state(grim_ctx* ctx, const double* prim_ic, int on_device);

/* Replaces setProblemSpecificBCs of buoyancy_instabilities (src/problem/
 * buoyancy_instabilities/problem.cpp:221-293): from the next step on, in
 * both stages after the boundary conditions, the stage state (primOld in
 * the half stage, primHalfStep in the full stage) takes on the X1 sides
 * rho, u and u~^1 from prim_ic, u~^2 = u~^3 = 0, B^1..3 = 0 and (conduction
 * on) q~ = 0.  The sides are the zones of X1 index 0..numGhost (the first
 * bulk zone included: af::seq(0, numGhost) is inclusive, :233) and
 * N1 + numGhost .. N1 + 2 numGhost - 1, every X2 and X3 index.  prim_ic: 3
 * planes (rho, u, u~^1) of grim_grid_zones doubles, host memory or device
 * memory when on_device != 0, copied; NULL switches the hook off.  The
 * state buffer and its on/off switch follow the rules of
 * grim_set_ghost_state.  Synchronises the context's stream. */
int grim_set_x1_side_state(grim_ctx* ctx, const double* prim_ic, int on_device);

/* Replaces the dumpVars fill of fullStepDiagnostics (torus/problem.cpp:
 * 808-875, layout params.cpp:136-147).  out: device buffer of (dof + 5)
 * planes of grim_grid_bulk zones each (N1*N2*N3, X1 fastest, no ghosts):
 * rho, u, v^1..3 = u^i/u^t and B^1..3 = b^i u^t - b^t u^i in the physical
 * x coordinates, q (conduction on), deltaP (viscosity on) placed between
 * v^3 and B^1, b^2, the Lorentz factor; then x^1..3 of the zone centres (the
 * points of the VTS grid).  dx/dX by central differences with step 1e-5
 * (geometry.cpp:441-508).  prim: device planes, or NULL for primOld.
 * on_device != 0: out is device memory and the call is asynchronous on the
 * context's stream; 0: out is host memory, staged through a device buffer
 * that the context allocates on the first such call, and the call
 * synchronises the stream. */
int grim_dump_vars(grim_ctx* ctx, const double* prim, double* out, int on_device);

#ifdef __cplusplus
}
#endif
#endif /* GRIM_HIP_H_ */
