// Batched per-zone Newton solve of grim's implicit stage on gfx950.
//
// Replaces timeStepper::solve (src/timestepper/solve.cpp:3-201), its
// Jacobian assembly (:82-109), batchLinearSolve (:203-230) and
// computeResidual (src/timestepper/residual.cpp:3-180).
//
// Design: one thread per zone, the whole Newton iteration in registers.  The
// zone's Jacobian is 5x5 (ideal MHD) to 7x7 (EMHD): far too small for MFMA
// tiles and small enough to stay in VGPRs, so LU with partial pivoting runs
// fully unrolled on register arrays with selects in place of row swaps.  All
// planes are zone-major structure of arrays: lane i of a wavefront touches
// zone base+i of every plane, so every load and store is a coalesced 512-byte
// wavefront access.  FP64 arithmetic throughout this file (the reference path).
#include <hip/hip_runtime.h>

#include "analytic.hpp"
#include "grim_hip.h"
#include "runtime.hpp"
#include "zone_physics.hpp"

namespace grim {

constexpr int kDefaultBlock = 256;

// Jacobian / precision modes of the per-zone Newton kernel.
//  MODE_FD:          grim's residual and FD Jacobian, FP64 LU (reference path)
//  MODE_ANALYTIC64:  grim's residual, analytic Jacobian and LU in FP64
//  MODE_MIXED:       FP64 residual (reciprocal form) and update; analytic
//                    Jacobian and equilibrated LU in FP32
enum { MODE_FD = 0, MODE_ANALYTIC64 = 1, MODE_MIXED = 2 };

// Per-zone quantities of the residual formed once per zone.
struct ZoneScale {
  double invAlpha;  // 1/alpha
  double scale;     // mask/dt
};

template <class L, int MODE>
__device__ __forceinline__ void evalResidual(const double* P, const ZoneGeom& G,
                                             const ZoneConst<L>& c, const DevParams& prm,
                                             double dt, const ZoneScale& zs, double* R) {
  if constexpr (MODE == MODE_MIXED)
    residualFastIdeal<L>(P, G, c, prm, zs.invAlpha, zs.scale, R);
  else
    residual<L>(P, G, c, prm, dt, R);
}

// grim's forward-difference Jacobian (solve.cpp:82-109): A[i][j] = dR_i/dP_j
// with P_j perturbed by eps*P_j, or by eps when |P_j| < eps/2.
template <class L>
__device__ __forceinline__ void fdJacobian(const double* P, const double* R, const ZoneGeom& G,
                                           const ZoneConst<L>& c, const DevParams& prm,
                                           double dt, double (&A)[L::NF][L::NF]) {
  double Pp[L::DOF];
#pragma unroll
  for (int v = 0; v < L::DOF; v++) Pp[v] = P[v];
#pragma unroll
  for (int row = 0; row < L::NF; row++) {
    const double eps = prm.jacEps;
    const bool small = fabs(P[row]) < .5 * eps;
    Pp[row] = small ? P[row] + eps : P[row] + eps * P[row];
    double Rp[L::NF];
    residual<L>(Pp, G, c, prm, dt, Rp);
    const double dP = Pp[row] - P[row];
#pragma unroll
    for (int col = 0; col < L::NF; col++) A[col][row] = (Rp[col] - R[col]) / dP;
    Pp[row] = P[row];
  }
}

// One Newton update after the convergence test (solve.cpp:80-198): Jacobian,
// LU solve of J d = -R, quadratic backtracking line search, P += lambda d.
// Returns true when the last evaluated trial was accepted; Rt then holds the
// residual at the new P (bit-identical to a fresh evaluation, since the new P
// is computed by the same expression as the trial).
template <class L, int MODE>
__device__ __forceinline__ bool newtonUpdate(double* P, const double* R, double l2,
                                             const ZoneGeom& G, const ZoneConst<L>& c,
                                             const DevParams& prm, double dt, const ZoneScale& zs,
                                             double* Rt, int& evals) {
  double d[L::NF];
  if constexpr (MODE == MODE_FD) {
    double A[L::NF][L::NF];
    fdJacobian<L>(P, R, G, c, prm, dt, A);
    evals += L::NF;
#pragma unroll
    for (int v = 0; v < L::NF; v++) d[v] = -R[v];
    luSolve<double, L::NF>(A, d);
  } else if constexpr (MODE == MODE_ANALYTIC64) {
    double A[L::NF][L::NF];
    analyticJacobianIdeal<double, L>(P, G, prm, c.mask / dt, A);
#pragma unroll
    for (int v = 0; v < L::NF; v++) d[v] = -R[v];
    luSolve<double, L::NF>(A, d);
  } else {
    float A[L::NF][L::NF], df[L::NF];
    analyticJacobianIdeal<float, L>(P, G, prm, zs.scale, A);
#pragma unroll
    for (int v = 0; v < L::NF; v++) df[v] = (float)(-R[v]);
    equilibratedSolve<float, L::NF>(A, df, P);
#pragma unroll
    for (int v = 0; v < L::NF; v++) d[v] = (double)df[v];
  }

  const double f0 = 0.5 * l2;
  const double fPrime0 = -2. * f0;
  double lambda = 1.;
  double Pt[L::DOF];
#pragma unroll
  for (int v = 0; v < L::DOF; v++) Pt[v] = P[v];
  bool accepted = false;
  for (int ls = 0; ls < prm.maxLS; ls++) {
#pragma unroll
    for (int v = 0; v < L::NF; v++) Pt[v] = P[v] + lambda * d[v];
    evalResidual<L, MODE>(Pt, G, c, prm, dt, zs, Rt);
    evals += 1;
    const double f1 = 0.5 * sumSq<L::NF>(Rt);
    const bool condition = f1 > f0 * (1. - 1e-4 * lambda) + prm.lsFloor;
    if (!condition) {
      accepted = true;
      break;
    }
    const double denom = f1 - f0 - fPrime0 * lambda;
    lambda = -fPrime0 * lambda * lambda / denom / 2.;
  }
#pragma unroll
  for (int v = 0; v < L::NF; v++) P[v] = P[v] + lambda * d[v];
  return accepted;
}

// Per-zone semantics: each thread iterates its own zone until
// sum R^2 <= atol or maxNonLinearIter updates.
template <class L, int MODE>
__global__ __launch_bounds__(kDefaultBlock) void newtonPerZone(DevParams prm, grim_zone_batch b,
                                                               grim_newton_stats st) {
  const int64_t z = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (z >= b.n) return;
  const int64_t n = b.n;
  const double dt = b.dt_prime;
  ZoneGeom G;
  loadGeom(b.geom, n, z, G);
  ZoneConst<L> c;
  loadZoneConst<L>(b, prm, G, z, c);
  ZoneScale zs;
  if constexpr (MODE == MODE_MIXED) {
    zs.invAlpha = 1. / G.alpha;
    zs.scale = c.mask / dt;
  }
  double P[L::DOF];
#pragma unroll
  for (int v = 0; v < L::DOF; v++) P[v] = b.prim[v * n + z];

  double R[L::NF];
  bool haveR = false;
  bool converged = false;
  int iters = 0, evals = 0;
  double l2 = 0.;
  for (int it = 0; it < prm.maxIter; it++) {
    // floors at the top of every iteration (solve.cpp:17-18)
    const double p0 = smax(P[RHO], prm.rhoFloor);
    const double p1 = smax(P[UU], prm.uFloor);
    if (!(p0 == P[RHO] && p1 == P[UU])) haveR = false;
    P[RHO] = p0;
    P[UU] = p1;
    if (!haveR) {
      evalResidual<L, MODE>(P, G, c, prm, dt, zs, R);
      evals++;
    }
    l2 = sumSq<L::NF>(R);
    if (!(l2 > prm.atol)) {
      converged = true;
      break;
    }
    double Rt[L::NF];
    haveR = newtonUpdate<L, MODE>(P, R, l2, G, c, prm, dt, zs, Rt, evals);
    iters++;
    if (haveR) {
#pragma unroll
      for (int v = 0; v < L::NF; v++) R[v] = Rt[v];
    }
  }
  if (!converged) {
    if (!haveR) evalResidual<L, MODE>(P, G, c, prm, dt, zs, R);
    l2 = sumSq<L::NF>(R);
  }
#pragma unroll
  for (int v = 0; v < L::NF; v++) b.prim[v * n + z] = P[v];
  if (st.iterations) st.iterations[z] = iters;
  if (st.residual_evals) st.residual_evals[z] = evals;
  if (st.final_l2) st.final_l2[z] = l2;
}

// grim's lock-step semantics (solve.cpp:10-200): iteration k first tests
// every zone (floors, residual, count of zones above atol), then -- if the
// count is nonzero -- updates every zone.  Launched as 2*maxNonLinearIter
// kernels with no host synchronisation; a zero count makes every later
// launch return immediately.
template <class L>
__global__ __launch_bounds__(kDefaultBlock) void newtonGlobalCheck(DevParams prm, grim_zone_batch b,
                                                                  int k, int* counts, double* Rbuf,
                                                                  double* l2buf, int* evalsBuf) {
  if (k > 0 && counts[k - 1] == 0) return;
  const int64_t z = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  bool above = false;
  if (z < b.n) {
    const int64_t n = b.n;
    ZoneGeom G;
    loadGeom(b.geom, n, z, G);
    ZoneConst<L> c;
    loadZoneConst<L>(b, prm, G, z, c);
    double P[L::DOF];
#pragma unroll
    for (int v = 0; v < L::DOF; v++) P[v] = b.prim[v * n + z];
    P[RHO] = smax(P[RHO], prm.rhoFloor);
    P[UU] = smax(P[UU], prm.uFloor);
    b.prim[RHO * n + z] = P[RHO];
    b.prim[UU * n + z] = P[UU];
    double R[L::NF];
    residual<L>(P, G, c, prm, b.dt_prime, R);
    const double l2 = sumSq<L::NF>(R);
#pragma unroll
    for (int v = 0; v < L::NF; v++) Rbuf[v * n + z] = R[v];
    l2buf[z] = l2;
    evalsBuf[z] += 1;
    above = l2 > prm.atol;
  }
  // one atomic per wavefront
  const unsigned long long ballot = __ballot(above);
  if ((threadIdx.x & 63) == 0 && ballot) atomicAdd(&counts[k], (int)__popcll(ballot));
}

template <class L>
__global__ __launch_bounds__(kDefaultBlock) void newtonGlobalUpdate(DevParams prm,
                                                                   grim_zone_batch b, int k,
                                                                   const int* counts,
                                                                   const double* Rbuf,
                                                                   const double* l2buf,
                                                                   int* evalsBuf) {
  if (counts[k] == 0) return;
  const int64_t z = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (z >= b.n) return;
  const int64_t n = b.n;
  ZoneGeom G;
  loadGeom(b.geom, n, z, G);
  ZoneConst<L> c;
  loadZoneConst<L>(b, prm, G, z, c);
  double P[L::DOF], R[L::NF], Rt[L::NF];
#pragma unroll
  for (int v = 0; v < L::DOF; v++) P[v] = b.prim[v * n + z];
#pragma unroll
  for (int v = 0; v < L::NF; v++) R[v] = Rbuf[v * n + z];
  int evals = 0;
  ZoneScale zs{};
  newtonUpdate<L, MODE_FD>(P, R, l2buf[z], G, c, prm, b.dt_prime, zs, Rt, evals);
#pragma unroll
  for (int v = 0; v < L::NF; v++) b.prim[v * n + z] = P[v];
  evalsBuf[z] += evals;
}

// Stats of the lock-step solve; recomputes the residual when the loop ran
// out of iterations (the last test then predates the last update).
template <class L>
__global__ __launch_bounds__(kDefaultBlock) void newtonGlobalFinish(DevParams prm, grim_zone_batch b,
                                                                   const int* counts, int maxIter,
                                                                   double* l2buf, const int* evalsBuf,
                                                                   grim_newton_stats st) {
  const int64_t z = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (z >= b.n) return;
  int updates = 0;
  bool stopped = false;
  for (int k = 0; k < maxIter; k++) {
    if (counts[k] == 0) {
      stopped = true;
      break;
    }
    updates++;
  }
  double l2 = l2buf[z];
  if (!stopped) {
    const int64_t n = b.n;
    ZoneGeom G;
    loadGeom(b.geom, n, z, G);
    ZoneConst<L> c;
    loadZoneConst<L>(b, prm, G, z, c);
    double P[L::DOF], R[L::NF];
#pragma unroll
    for (int v = 0; v < L::DOF; v++) P[v] = b.prim[v * n + z];
    residual<L>(P, G, c, prm, b.dt_prime, R);
    l2 = sumSq<L::NF>(R);
  }
  if (st.iterations) st.iterations[z] = updates;
  if (st.residual_evals) st.residual_evals[z] = evalsBuf[z];
  if (st.final_l2) st.final_l2[z] = l2;
}

__global__ void fillInt(int* p, int64_t n, int v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

template <class L>
__global__ __launch_bounds__(kDefaultBlock) void residualKernel(DevParams prm, grim_zone_batch b,
                                                               double* Rout) {
  const int64_t z = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (z >= b.n) return;
  const int64_t n = b.n;
  ZoneGeom G;
  loadGeom(b.geom, n, z, G);
  ZoneConst<L> c;
  loadZoneConst<L>(b, prm, G, z, c);
  double P[L::DOF], R[L::NF];
#pragma unroll
  for (int v = 0; v < L::DOF; v++) P[v] = b.prim[v * n + z];
  residual<L>(P, G, c, prm, b.dt_prime, R);
#pragma unroll
  for (int v = 0; v < L::NF; v++) Rout[v * n + z] = R[v];
}

template <class L>
__global__ __launch_bounds__(kDefaultBlock) void fdJacobianKernel(DevParams prm, grim_zone_batch b,
                                                                 double* J) {
  const int64_t z = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (z >= b.n) return;
  const int64_t n = b.n;
  ZoneGeom G;
  loadGeom(b.geom, n, z, G);
  ZoneConst<L> c;
  loadZoneConst<L>(b, prm, G, z, c);
  double P[L::DOF], R[L::NF], A[L::NF][L::NF];
#pragma unroll
  for (int v = 0; v < L::DOF; v++) P[v] = b.prim[v * n + z];
  residual<L>(P, G, c, prm, b.dt_prime, R);
  fdJacobian<L>(P, R, G, c, prm, b.dt_prime, A);
#pragma unroll
  for (int i = 0; i < L::NF; i++)
#pragma unroll
    for (int j = 0; j < L::NF; j++) J[(i * L::NF + j) * n + z] = A[i][j];
}

template <class L>
__global__ __launch_bounds__(kDefaultBlock) void analyticJacobianKernel(DevParams prm,
                                                                       grim_zone_batch b, double* J) {
  const int64_t z = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (z >= b.n) return;
  const int64_t n = b.n;
  ZoneGeom G;
  loadGeom(b.geom, n, z, G);
  const double mask = b.residual_mask ? b.residual_mask[z] : 1.;
  double P[L::DOF], A[L::NF][L::NF];
#pragma unroll
  for (int v = 0; v < L::DOF; v++) P[v] = b.prim[v * n + z];
  analyticJacobianIdeal<double, L>(P, G, prm, mask / b.dt_prime, A);
#pragma unroll
  for (int i = 0; i < L::NF; i++)
#pragma unroll
    for (int j = 0; j < L::NF; j++) J[(i * L::NF + j) * n + z] = A[i][j];
}

template <class Real, int N>
__global__ __launch_bounds__(kDefaultBlock) void batchLinearSolveKernel(int64_t n, const double* A,
                                                                       const double* bvec,
                                                                       double* x) {
  const int64_t z = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (z >= n) return;
  Real M[N][N], r[N];
#pragma unroll
  for (int i = 0; i < N; i++) {
#pragma unroll
    for (int j = 0; j < N; j++) M[i][j] = (Real)A[(i * N + j) * n + z];
    r[i] = (Real)bvec[i * n + z];
  }
  luSolve<Real, N>(M, r);
#pragma unroll
  for (int i = 0; i < N; i++) x[i * n + z] = (double)r[i];
}

namespace {

int validBatch(const grim_params* p, const grim_zone_batch* b) {
  if (!validParams(p) || !b || b->n < 0) return 0;
  if (b->n == 0) return 1;
  if (!b->prim || !b->geom || !b->rhs) return 0;
  if ((p->conduction || p->viscosity) && !b->prim_old) return 0;
  if (p->closure == GRIM_CLOSURE_PER_ZONE && (p->conduction || p->viscosity) && !b->closure) return 0;
  if (p->closure == GRIM_CLOSURE_TORUS && (p->conduction || p->viscosity) && !b->radius) return 0;
  if (!(b->dt_prime > 0.)) return 0;
  return 1;
}

int launchGlobal(const grim_params& p, const grim_zone_batch& b, const grim_newton_stats& st,
                 hipStream_t s) {
  const DevParams prm = makeDevParams(p);
  const int64_t n = b.n;
  const int nf = 5 + p.conduction + p.viscosity;
  const int maxIter = p.maxNonLinearIter;
  void* mem = nullptr;
  const size_t bytes = sizeof(double) * (size_t)n * (nf + 1) + sizeof(int) * (size_t)n +
                       sizeof(int) * (size_t)(maxIter + 1);
  GRIM_HIP_TRY(hipMallocAsync(&mem, bytes, s));
  double* Rbuf = (double*)mem;
  double* l2buf = Rbuf + (size_t)n * nf;
  int* evalsBuf = (int*)(l2buf + n);
  int* counts = evalsBuf + n;
  const unsigned grid = gridFor(n, kDefaultBlock);
  fillInt<<<gridFor(n, 256), 256, 0, s>>>(evalsBuf, n, 0);
  fillInt<<<gridFor(maxIter + 1, 64), 64, 0, s>>>(counts, maxIter + 1, 0);
  int rc = dispatchLayout(p, [&]<class L>() -> int {
    for (int k = 0; k < maxIter; k++) {
      newtonGlobalCheck<L><<<grid, kDefaultBlock, 0, s>>>(prm, b, k, counts, Rbuf, l2buf, evalsBuf);
      newtonGlobalUpdate<L><<<grid, kDefaultBlock, 0, s>>>(prm, b, k, counts, Rbuf, l2buf, evalsBuf);
    }
    newtonGlobalFinish<L><<<grid, kDefaultBlock, 0, s>>>(prm, b, counts, maxIter, l2buf, evalsBuf, st);
    return checkLaunch("newton (global)");
  });
  hipError_t e = hipFreeAsync(mem, s);
  if (rc == GRIM_OK && e != hipSuccess) {
    setError(std::string("hipFreeAsync: ") + hipGetErrorString(e));
    return GRIM_ERR_HIP;
  }
  return rc;
}

}  // namespace
}  // namespace grim

using namespace grim;

extern "C" {

int grim_residual(const grim_params* p, const grim_zone_batch* b, double* R,
                  grim_stream_t stream) {
  GRIM_REQUIRE(validBatch(p, b) && (b->n == 0 || R), "grim_residual: bad arguments");
  if (b->n == 0) return GRIM_OK;
  if (int rc = checkDevice()) return rc;
  hipStream_t s = (hipStream_t)stream;
  const DevParams prm = makeDevParams(*p);
  const grim_zone_batch bb = *b;
  return dispatchLayout(*p, [&]<class L>() -> int {
    residualKernel<L><<<gridFor(bb.n, kDefaultBlock), kDefaultBlock, 0, s>>>(prm, bb, R);
    return checkLaunch("grim_residual");
  });
}

int grim_jacobian(const grim_params* p, const grim_zone_batch* b, int jacobian, double* J,
                  grim_stream_t stream) {
  GRIM_REQUIRE(validBatch(p, b) && (b->n == 0 || J), "grim_jacobian: bad arguments");
  GRIM_REQUIRE(jacobian == GRIM_JAC_FD || jacobian == GRIM_JAC_ANALYTIC,
               "grim_jacobian: bad Jacobian kind");
  GRIM_REQUIRE(jacobian == GRIM_JAC_FD || !(p->conduction || p->viscosity),
               "grim_jacobian: the analytic Jacobian is built for ideal MHD only");
  if (b->n == 0) return GRIM_OK;
  if (int rc = checkDevice()) return rc;
  hipStream_t s = (hipStream_t)stream;
  const DevParams prm = makeDevParams(*p);
  const grim_zone_batch bb = *b;
  return dispatchLayout(*p, [&]<class L>() -> int {
    if (jacobian == GRIM_JAC_FD) {
      fdJacobianKernel<L><<<gridFor(bb.n, kDefaultBlock), kDefaultBlock, 0, s>>>(prm, bb, J);
    } else {
      if constexpr (!L::EMHD)
        analyticJacobianKernel<L><<<gridFor(bb.n, kDefaultBlock), kDefaultBlock, 0, s>>>(prm, bb, J);
    }
    return checkLaunch("grim_jacobian");
  });
}

int grim_batch_linear_solve(int64_t n, int nvar, const double* A, const double* bvec, double* x,
                            int lu_precision, grim_stream_t stream) {
  GRIM_REQUIRE(n >= 0 && nvar >= 1 && nvar <= GRIM_MAX_VARS, "grim_batch_linear_solve: bad size");
  GRIM_REQUIRE(lu_precision == GRIM_LU_FP64 || lu_precision == GRIM_LU_FP32,
               "grim_batch_linear_solve: bad precision");
  if (n == 0) return GRIM_OK;
  GRIM_REQUIRE(A && bvec && x, "grim_batch_linear_solve: null pointer");
  if (int rc = checkDevice()) return rc;
  hipStream_t s = (hipStream_t)stream;
  const unsigned grid = gridFor(n, kDefaultBlock);
  auto go = [&]<int N>() {
    if (lu_precision == GRIM_LU_FP64)
      batchLinearSolveKernel<double, N><<<grid, kDefaultBlock, 0, s>>>(n, A, bvec, x);
    else
      batchLinearSolveKernel<float, N><<<grid, kDefaultBlock, 0, s>>>(n, A, bvec, x);
  };
  switch (nvar) {
    case 1: go.template operator()<1>(); break;
    case 2: go.template operator()<2>(); break;
    case 3: go.template operator()<3>(); break;
    case 4: go.template operator()<4>(); break;
    case 5: go.template operator()<5>(); break;
    case 6: go.template operator()<6>(); break;
    case 7: go.template operator()<7>(); break;
    case 8: go.template operator()<8>(); break;
    case 9: go.template operator()<9>(); break;
    default: go.template operator()<10>(); break;
  }
  return checkLaunch("grim_batch_linear_solve");
}

int grim_batched_newton(const grim_params* p, const grim_zone_batch* b,
                        const grim_newton_opts* opts, grim_newton_stats* stats,
                        grim_stream_t stream) {
  GRIM_REQUIRE(validBatch(p, b), "grim_batched_newton: bad params or batch");
  grim_newton_opts o = {GRIM_JAC_FD, GRIM_LU_FP64, GRIM_NEWTON_PER_ZONE, 0, 0};
  if (opts) o = *opts;
  grim_newton_stats st = {nullptr, nullptr, nullptr};
  if (stats) st = *stats;
  GRIM_REQUIRE(o.semantics == GRIM_NEWTON_PER_ZONE || o.semantics == GRIM_NEWTON_GLOBAL,
               "grim_batched_newton: bad semantics");
  GRIM_REQUIRE(!o.compaction, "grim_batched_newton: compaction is not built");
  int mode = -1;
  if (o.jacobian == GRIM_JAC_FD && o.lu_precision == GRIM_LU_FP64) mode = MODE_FD;
  if (o.jacobian == GRIM_JAC_ANALYTIC && o.lu_precision == GRIM_LU_FP64) mode = MODE_ANALYTIC64;
  if (o.jacobian == GRIM_JAC_ANALYTIC && o.lu_precision == GRIM_LU_FP32) mode = MODE_MIXED;
  GRIM_REQUIRE(mode >= 0, "grim_batched_newton: unsupported Jacobian / LU precision pair");
  GRIM_REQUIRE(mode == MODE_FD || !(p->conduction || p->viscosity),
               "grim_batched_newton: the analytic Jacobian is built for ideal MHD only");
  GRIM_REQUIRE(mode == MODE_FD || o.semantics == GRIM_NEWTON_PER_ZONE,
               "grim_batched_newton: lock-step semantics run grim's FD-FP64 algorithm only");
  if (b->n == 0) return GRIM_OK;
  if (int rc = checkDevice()) return rc;
  hipStream_t s = (hipStream_t)stream;
  if (o.semantics == GRIM_NEWTON_GLOBAL) return launchGlobal(*p, *b, st, s);
  const DevParams prm = makeDevParams(*p);
  const grim_zone_batch bb = *b;
  const unsigned grid = gridFor(bb.n, kDefaultBlock);
  return dispatchLayout(*p, [&]<class L>() -> int {
    if (mode == MODE_FD) {
      newtonPerZone<L, MODE_FD><<<grid, kDefaultBlock, 0, s>>>(prm, bb, st);
    } else if constexpr (!L::EMHD) {
      if (mode == MODE_ANALYTIC64)
        newtonPerZone<L, MODE_ANALYTIC64><<<grid, kDefaultBlock, 0, s>>>(prm, bb, st);
      else
        newtonPerZone<L, MODE_MIXED><<<grid, kDefaultBlock, 0, s>>>(prm, bb, st);
    }
    return checkLaunch("grim_batched_newton");
  });
}

}  // extern "C"
