// Batched per-zone Newton solve of grim's implicit stage on gfx950: the
// exact translation unit (compiled with -ffp-contract=off).
//
// Replaces timeStepper::solve (src/timestepper/solve.cpp:3-201), its
// Jacobian assembly (:82-109), batchLinearSolve (:203-230) and
// computeResidual (src/timestepper/residual.cpp:3-180).  Kernels here round
// every FP64 operation as written in grim's expressions: the FD reference
// path, the analytic FP64 path, grim's lock-step loop, and the residual /
// Jacobian / linear-solve entry points.  The fused-arithmetic modes live in
// newton_fast.hip.
#include <hip/hip_runtime.h>

#include <atomic>
#include <mutex>

#include "grim_hip.h"
#include "newton_kernels.hpp"
#include "runtime.hpp"

namespace grim {
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
    loadBatchGeom(b, z, G);
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
    residual<L>(P, G, c, prm, batchDt(b), R);
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
  // Zones outside the bulk (mask 0) have a zero residual and a zero
  // Jacobian: they take no update (grim would write the non-finite result
  // of the singular solve there, later overwritten by the boundaries).
  if (b.residual_mask && b.residual_mask[z] == 0.) return;
  const int64_t n = b.n;
  ZoneGeom G;
  loadBatchGeom(b, z, G);
  ZoneConst<L> c;
  loadZoneConst<L>(b, prm, G, z, c);
  double P[L::DOF], R[L::NF], Rt[L::NF];
#pragma unroll
  for (int v = 0; v < L::DOF; v++) P[v] = b.prim[v * n + z];
#pragma unroll
  for (int v = 0; v < L::NF; v++) R[v] = Rbuf[v * n + z];
  int evals = 0;
  ZoneScale zs{};
  EmhdWeights ew;
  newtonUpdate<L, MODE_FD>(P, R, l2buf[z], G, c, prm, batchDt(b), zs, ew, Rt, evals);
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
    loadBatchGeom(b, z, G);
    ZoneConst<L> c;
    loadZoneConst<L>(b, prm, G, z, c);
    double P[L::DOF], R[L::NF];
#pragma unroll
    for (int v = 0; v < L::DOF; v++) P[v] = b.prim[v * n + z];
    residual<L>(P, G, c, prm, batchDt(b), R);
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
  loadBatchGeom(b, z, G);
  ZoneConst<L> c;
  loadZoneConst<L>(b, prm, G, z, c);
  double P[L::DOF], R[L::NF];
#pragma unroll
  for (int v = 0; v < L::DOF; v++) P[v] = b.prim[v * n + z];
  residual<L>(P, G, c, prm, batchDt(b), R);
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
  loadBatchGeom(b, z, G);
  ZoneConst<L> c;
  loadZoneConst<L>(b, prm, G, z, c);
  double P[L::DOF], R[L::NF], A[L::NF][L::NF];
#pragma unroll
  for (int v = 0; v < L::DOF; v++) P[v] = b.prim[v * n + z];
  residual<L>(P, G, c, prm, batchDt(b), R);
  fdJacobian<L>(P, R, G, c, prm, batchDt(b), A);
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
  loadBatchGeom(b, z, G);
  ZoneConst<L> c;
  loadZoneConst<L>(b, prm, G, z, c);
  EmhdWeights ew;
  if constexpr (L::EMHD) emhdWeights<L>(G, c, prm, batchDt(b), ew);
  double P[L::DOF], A[L::NF][L::NF];
#pragma unroll
  for (int v = 0; v < L::DOF; v++) P[v] = b.prim[v * n + z];
  analyticJacobian<double, L>(P, G, prm, c.mask / batchDt(b), c, ew, batchDt(b), A);
#pragma unroll
  for (int i = 0; i < L::NF; i++)
#pragma unroll
    for (int j = 0; j < L::NF; j++) J[(i * L::NF + j) * n + z] = A[i][j];
}

// Matrix-free products of the analytic derivative of the residual, in the
// arithmetic of T: the tangent J v (the column-wise tangent contracted with
// v in registers) and the adjoint J^T w (the reverse sweep, J never formed).
template <class L, class T, bool ADJOINT>
__global__ __launch_bounds__(kDefaultBlock) void derivativeProductKernel(DevParams prm, grim_zone_batch b,
                                                                        const double* __restrict__ x,
                                                                        double* __restrict__ out) {
  const int64_t z = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (z >= b.n) return;
  const int64_t n = b.n;
  const double dt = batchDt(b);
  ZoneGeom G;
  loadBatchGeom(b, z, G);
  ZoneConst<L> c;
  loadZoneConst<L>(b, prm, G, z, c);
  EmhdWeights ew;
  if constexpr (L::EMHD) emhdWeights<L>(G, c, prm, dt, ew);
  double P[L::DOF];
#pragma unroll
  for (int v = 0; v < L::DOF; v++) P[v] = b.prim[v * n + z];
  T xs[L::NF], y[L::NF];
#pragma unroll
  for (int v = 0; v < L::NF; v++) xs[v] = (T)x[v * n + z];
  if constexpr (ADJOINT) {
    analyticAdjoint<T, L>(P, G, prm, c.mask / dt, c, ew, dt, xs, y);
  } else {
    T A[L::NF][L::NF];
    analyticJacobian<T, L>(P, G, prm, c.mask / dt, c, ew, dt, A);
#pragma unroll
    for (int i = 0; i < L::NF; i++) {
      T s = T(0);
#pragma unroll
      for (int j = 0; j < L::NF; j++) s += A[i][j] * xs[j];
      y[i] = s;
    }
  }
#pragma unroll
  for (int v = 0; v < L::NF; v++) out[v * n + z] = (double)y[v];
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
  if ((p->closure == GRIM_CLOSURE_TORUS || p->closure == GRIM_CLOSURE_BUOYANCY) && (p->conduction || p->viscosity) &&
      !b->radius)
    return 0;
  if (!b->dt_prime_device && !(b->dt_prime > 0.)) return 0;
  // column geometry: whole columns, 32-bit zone indices
  if (b->geom_columns < 0) return 0;
  if (b->geom_columns > 0 && (b->n % b->geom_columns != 0 || b->n > INT32_MAX)) return 0;
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

// Scratch of the allocation-free solve: the compaction workspace of up to
// GRIM_NOALLOC_MAX_ZONES zones, one copy per device for the life of the
// process.  In host code `noAllocScratch` only names the symbol; its device
// address comes from hipGetSymbolAddress, once per device.
constexpr int64_t kNoAllocScratchBytes = compactionBytes(GRIM_NOALLOC_MAX_ZONES);
static __device__ __align__(256) unsigned char noAllocScratch[kNoAllocScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratchMutex;
std::atomic<void*> g_scratchPtr[kMaxDevices];

// Device address of the current device's scratch (looked up on the first
// call per device, with that device current).
int noAllocScratchPtr(void** out) {
  int dev = -1;
  GRIM_HIP_TRY(hipGetDevice(&dev));
  GRIM_REQUIRE(dev >= 0 && dev < kMaxDevices, "grim_batched_newton_noalloc: device ordinal out of range");
  void* ptr = g_scratchPtr[dev].load(std::memory_order_acquire);
  if (!ptr) {
    std::lock_guard<std::mutex> lock(g_scratchMutex);
    ptr = g_scratchPtr[dev].load(std::memory_order_relaxed);
    if (!ptr) {
      GRIM_HIP_TRY(hipGetSymbolAddress(&ptr, HIP_SYMBOL(noAllocScratch)));
      GRIM_REQUIRE(ptr != nullptr, "grim_batched_newton_noalloc: no device address for the scratch array");
      g_scratchPtr[dev].store(ptr, std::memory_order_release);
    }
  }
  *out = ptr;
  return GRIM_OK;
}
}  // namespace
}  // namespace grim

using namespace grim;

namespace grim {

bool foldFusable(const grim_params& p, const grim_newton_opts& o, int dim) {
  const bool escalate = o.jacobian == GRIM_JAC_ANALYTIC && o.lu_precision == GRIM_LU_FP32_FP64;
  const bool mixed = o.jacobian == GRIM_JAC_ANALYTIC && o.lu_precision == GRIM_LU_FP32;
  const int budget1 = o.phase1_iters ? o.phase1_iters : 2;
  return dim == 3 && o.semantics == GRIM_NEWTON_PER_ZONE && (escalate || (mixed && o.compaction == 1)) &&
         budget1 < p.maxNonLinearIter;
}

int batchedNewton(const grim_params* p, const grim_zone_batch* b, const grim_newton_opts* opts,
                  grim_newton_stats* stats, hipStream_t stream, const FoldArgs* fold) {
  GRIM_REQUIRE(validBatch(p, b), "grim_batched_newton: bad params or batch");
  grim_newton_opts o = {};
  o.jacobian = GRIM_JAC_FD;
  o.lu_precision = GRIM_LU_FP64;
  o.semantics = GRIM_NEWTON_PER_ZONE;
  if (opts) o = *opts;
  grim_newton_stats st = {nullptr, nullptr, nullptr};
  if (stats) st = *stats;
  GRIM_REQUIRE(o.semantics == GRIM_NEWTON_PER_ZONE || o.semantics == GRIM_NEWTON_GLOBAL,
               "grim_batched_newton: bad semantics");
  int mode = -1;
  if (o.jacobian == GRIM_JAC_FD && o.lu_precision == GRIM_LU_FP64) mode = o.fused ? MODE_FD_FAST : MODE_FD;
  if (o.jacobian == GRIM_JAC_ANALYTIC && o.lu_precision == GRIM_LU_FP64) mode = MODE_ANALYTIC64;
  if (o.jacobian == GRIM_JAC_ANALYTIC && o.lu_precision == GRIM_LU_FP32) mode = MODE_MIXED;
  const bool escalate = o.jacobian == GRIM_JAC_ANALYTIC && o.lu_precision == GRIM_LU_FP32_FP64;
  if (escalate) mode = MODE_MIXED;
  GRIM_REQUIRE(mode >= 0, "grim_batched_newton: unsupported Jacobian / LU precision pair");
  GRIM_REQUIRE(mode == MODE_FD || o.semantics == GRIM_NEWTON_PER_ZONE,
               "grim_batched_newton: lock-step semantics run grim's exact FD-FP64 algorithm only");
  GRIM_REQUIRE(o.compaction >= 0 && o.compaction <= 2, "grim_batched_newton: compaction must be 0, 1 or 2");
  GRIM_REQUIRE(o.semantics == GRIM_NEWTON_PER_ZONE || !o.compaction,
               "grim_batched_newton: compaction applies to per-zone semantics");
  GRIM_REQUIRE(o.block_size == 0 || o.block_size == 64 || o.block_size == 128 || o.block_size == 256,
               "grim_batched_newton: block_size must be 64, 128 or 256");
  GRIM_REQUIRE(o.phase1_iters >= 0 && o.phase2_group >= 0 && o.phase2_group <= kMaxGroup,
               "grim_batched_newton: bad compaction options");
  if (b->n == 0) return GRIM_OK;
  if (int rc = checkDevice()) return rc;
  hipStream_t s = stream;
  GRIM_REQUIRE(!fold || foldFusable(*p, o, fold->D.dim), "stage solve: the fold runs inside compacting mixed solves only");
  if (o.semantics == GRIM_NEWTON_GLOBAL) {
    // grim's lock-step loop works in prim from the start: copy the guess in
    if (b->guess && b->guess != b->prim)
      GRIM_HIP_TRY(hipMemcpyAsync(b->prim, b->guess, sizeof(double) * (size_t)(5 + p->conduction + p->viscosity) *
                                                         (size_t)b->n,
                                  hipMemcpyDeviceToDevice, s));
    return launchGlobal(*p, *b, st, s);
  }

  LaunchCfg cfg;
  cfg.block = o.block_size ? o.block_size : kDefaultBlock;
  cfg.compaction = o.compaction == 1 || escalate;
  cfg.workgroupSplit = o.compaction == 2 && !escalate;
  cfg.escalate = escalate;
  cfg.budget1 = o.phase1_iters ? o.phase1_iters : 2;
  cfg.group = o.phase2_group ? o.phase2_group : 1;
  // persistent phase 2: 8 wavefronts per compute unit
  cfg.persistentWaves = 8 * deviceCUs();
  cfg.ws = CompactionWorkspace{};
  void* owned = nullptr;
  const int64_t need = grim_newton_workspace_bytes(b->n);
  if (cfg.compaction) {
    GRIM_REQUIRE(b->n <= (int64_t(1) << 30), "grim_batched_newton: compaction supports at most 2^30 zones");
    void* ws = o.workspace;
    if (ws) {
      GRIM_REQUIRE(o.workspace_bytes >= need, "grim_batched_newton: workspace too small");
    } else {
      GRIM_HIP_TRY(hipMallocAsync(&owned, (size_t)need, s));
      ws = owned;
    }
    cfg.ws = carveCompaction(ws, b->n);
  }
  const DevParams prm = makeDevParams(*p);
  const grim_zone_batch bb = *b;
  int rc;
  if (mode == MODE_MIXED || mode == MODE_FD_FAST) {
    rc = launchNewtonFast(mode, *p, bb, st, cfg, s, fold);
  } else {
    rc = dispatchLayout(*p, [&]<class L>() -> int {
      if (mode == MODE_FD)
        launchSolve<L, MODE_FD>(prm, bb, st, cfg, s);
      else
        launchSolve<L, MODE_ANALYTIC64>(prm, bb, st, cfg, s);
      return checkLaunch("grim_batched_newton");
    });
  }
  if (owned) {
    hipError_t e = hipFreeAsync(owned, s);
    if (rc == GRIM_OK && e != hipSuccess) {
      setError(std::string("hipFreeAsync: ") + hipGetErrorString(e));
      rc = GRIM_ERR_HIP;
    }
  }
  return rc;
}

}  // namespace grim

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
  if (b->n == 0) return GRIM_OK;
  if (int rc = checkDevice()) return rc;
  hipStream_t s = (hipStream_t)stream;
  const DevParams prm = makeDevParams(*p);
  const grim_zone_batch bb = *b;
  return dispatchLayout(*p, [&]<class L>() -> int {
    if (jacobian == GRIM_JAC_FD) {
      fdJacobianKernel<L><<<gridFor(bb.n, kDefaultBlock), kDefaultBlock, 0, s>>>(prm, bb, J);
    } else {
      analyticJacobianKernel<L><<<gridFor(bb.n, kDefaultBlock), kDefaultBlock, 0, s>>>(prm, bb, J);
    }
    return checkLaunch("grim_jacobian");
  });
}

namespace {
int derivativeProduct(const char* what, bool adjoint, const grim_params* p, const grim_zone_batch* b,
                      const double* x, double* out, int precision, grim_stream_t stream) {
  GRIM_REQUIRE(validBatch(p, b) && (b->n == 0 || (x && out)), std::string(what) + ": bad arguments");
  GRIM_REQUIRE(precision == GRIM_LU_FP64 || precision == GRIM_LU_FP32, std::string(what) + ": bad precision");
  if (b->n == 0) return GRIM_OK;
  if (int rc = checkDevice()) return rc;
  hipStream_t s = (hipStream_t)stream;
  const DevParams prm = makeDevParams(*p);
  const grim_zone_batch bb = *b;
  const unsigned grid = gridFor(bb.n, kDefaultBlock);
  return dispatchLayout(*p, [&]<class L>() -> int {
    if (adjoint && precision == GRIM_LU_FP64)
      derivativeProductKernel<L, double, true><<<grid, kDefaultBlock, 0, s>>>(prm, bb, x, out);
    else if (adjoint)
      derivativeProductKernel<L, float, true><<<grid, kDefaultBlock, 0, s>>>(prm, bb, x, out);
    else if (precision == GRIM_LU_FP64)
      derivativeProductKernel<L, double, false><<<grid, kDefaultBlock, 0, s>>>(prm, bb, x, out);
    else
      derivativeProductKernel<L, float, false><<<grid, kDefaultBlock, 0, s>>>(prm, bb, x, out);
    return checkLaunch(what);
  });
}
}  // namespace

int grim_residual_tangent(const grim_params* p, const grim_zone_batch* b, const double* v, double* Jv,
                          int precision, grim_stream_t stream) {
  return derivativeProduct("grim_residual_tangent", false, p, b, v, Jv, precision, stream);
}

int grim_residual_adjoint(const grim_params* p, const grim_zone_batch* b, const double* w, double* JTw,
                          int precision, grim_stream_t stream) {
  return derivativeProduct("grim_residual_adjoint", true, p, b, w, JTw, precision, stream);
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

int64_t grim_newton_workspace_bytes(int64_t n) {
  // phase-1 masks, their prefix sums, the work list and two counters
  return n < 0 ? 0 : compactionBytes(n);
}

int grim_batched_newton(const grim_params* p, const grim_zone_batch* b,
                        const grim_newton_opts* opts, grim_newton_stats* stats,
                        grim_stream_t stream) {
  return grim::batchedNewton(p, b, opts, stats, (hipStream_t)stream, nullptr);
}

int grim_batched_newton_noalloc(const grim_params* p, const grim_zone_batch* b,
                                const grim_newton_opts* opts, grim_newton_stats* stats,
                                grim_stream_t stream) {
  GRIM_REQUIRE(validBatch(p, b), "grim_batched_newton_noalloc: bad params or batch");
  GRIM_REQUIRE(b->n <= GRIM_NOALLOC_MAX_ZONES,
               "grim_batched_newton_noalloc: more than GRIM_NOALLOC_MAX_ZONES (2^23) zones");
  grim_newton_opts o = {};
  o.jacobian = GRIM_JAC_FD;
  o.lu_precision = GRIM_LU_FP64;
  o.semantics = GRIM_NEWTON_PER_ZONE;
  if (opts) o = *opts;
  GRIM_REQUIRE(o.semantics == GRIM_NEWTON_PER_ZONE,
               "grim_batched_newton_noalloc: per-zone semantics only (the lock-step loop needs "
               "per-zone state beyond the scratch array)");
  GRIM_REQUIRE(o.workspace == nullptr && o.workspace_bytes == 0,
               "grim_batched_newton_noalloc: takes no workspace (it uses the library's scratch array)");
  if (b->n == 0) return GRIM_OK;
  if (int rc = checkDevice()) return rc;
  void* ws = nullptr;
  if (int rc = noAllocScratchPtr(&ws)) return rc;
  o.workspace = ws;
  o.workspace_bytes = kNoAllocScratchBytes;
  // the per-zone solve allocates nothing once the compaction workspace is given
  return grim_batched_newton(p, b, &o, stats, stream);
}

}  // extern "C"
