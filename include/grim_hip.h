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
  GRIM_CLOSURE_TORUS = 2         /* src/problem/torus/problem.cpp:7-55 */
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
  const double* radius;       /* 1 plane (torus closure) or NULL */
  const double* residual_mask;/* 1 plane or NULL (= 1) */
  double dt_prime;            /* dt/2 (half step) or dt (full step) */
} grim_zone_batch;

typedef enum grim_jacobian_kind { GRIM_JAC_FD = 0, GRIM_JAC_ANALYTIC = 1 } grim_jacobian_kind;
typedef enum grim_lu_precision { GRIM_LU_FP64 = 0, GRIM_LU_FP32 = 1 } grim_lu_precision;
typedef enum grim_semantics {
  GRIM_NEWTON_PER_ZONE = 0, /* each zone stops at its own convergence */
  GRIM_NEWTON_GLOBAL = 1    /* grim's lock-step loop: all zones iterate while any is unconverged */
} grim_semantics;

typedef struct grim_newton_opts {
  int jacobian;      /* grim_jacobian_kind */
  int lu_precision;  /* grim_lu_precision */
  int semantics;     /* grim_semantics */
  int compaction;    /* 1: persistent kernel over a compacted work list */
  int block_size;    /* threads per workgroup (0 = library default) */
  int fused;         /* FD path only: 0 = grim's operation order, no fused
                        multiply-add (bit-reproducible against a CPU FP64
                        evaluation); 1 = fused multiply-adds and reciprocal
                        forms.  The mixed path is always fused. */
  int phase1_iters;  /* compaction: Newton updates before the work list is
                        compacted (0 = library default) */
  int phase2_group;  /* compaction: phase-1 wavefronts whose unconverged zones
                        one phase-2 wavefront takes over, 1..8 (0 = default) */
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

#ifdef __cplusplus
}
#endif
#endif /* GRIM_HIP_H_ */
