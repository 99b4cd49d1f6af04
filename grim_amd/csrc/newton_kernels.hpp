// Per-zone Newton kernels shared by the exact and the fast translation units.
//
// newton.hip includes this header and compiles with -ffp-contract=off (the
// reference FD path and the analytic FP64 path round every operation as
// written, bit-identical to the oracle); newton_fast.hip compiles the same
// templates with -ffp-contract=fast-honor-pragmas (fused multiply-adds) for the mixed
// precision path and the fast FD variant.  Each translation unit
// instantiates different modes, so no kernel is compiled twice.
//
// Design: one thread per zone, the whole Newton iteration in registers.  The
// zone's Jacobian is 5x5 (ideal MHD) to 7x7 (EMHD): far too small for MFMA
// tiles and small enough to stay in VGPRs, so LU with partial pivoting runs
// fully unrolled on register arrays with selects in place of row swaps.  All
// planes are zone-major structure of arrays: lane i of a wavefront touches
// zone base+i of every plane, so every load and store is a coalesced 512-byte
// wavefront access.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>

#include "analytic.hpp"
#include "fold.hpp"
#include "grim_hip.h"
#include "zone_physics.hpp"

namespace grim {

constexpr int kDefaultBlock = 256;
constexpr int kMaxGroup = 8;  // phase-2 grab, in chunks of 64 list entries, at most

// dt' of the batch: from device memory when the caller computed it there.
__device__ __forceinline__ double batchDt(const grim_zone_batch& b) {
  return b.dt_prime_device ? *b.dt_prime_device : b.dt_prime;
}

// Jacobian / precision modes of the per-zone Newton kernel.
//  MODE_FD:          grim's residual and FD Jacobian, FP64 LU (reference path)
//  MODE_ANALYTIC64:  grim's residual, analytic Jacobian and LU in FP64
//  MODE_MIXED:       FP64 residual (reciprocal form) and update; analytic
//                    Jacobian and equilibrated LU in FP32
//  MODE_FD_FAST:     grim's FD Jacobian rule on the reciprocal-form FP64
//                    residual, one reciprocal per FD column, FP64 LU (the
//                    fastest FD-FP64 variant, for a like-for-like comparison)
//  MODE_ANALYTIC64F: the mixed path's FP64 escalation: MODE_MIXED's residual
//                    with the analytic Jacobian and LU in FP64 (phase 2 of
//                    GRIM_LU_FP32_FP64, compiled in the fused unit)
enum { MODE_FD = 0, MODE_ANALYTIC64 = 1, MODE_MIXED = 2, MODE_FD_FAST = 3, MODE_ANALYTIC64F = 4 };

template <int MODE>
struct ModeTraits {
  static constexpr bool fastResidual = (MODE == MODE_MIXED || MODE == MODE_FD_FAST || MODE == MODE_ANALYTIC64F);
  static constexpr bool analytic = (MODE == MODE_ANALYTIC64 || MODE == MODE_MIXED || MODE == MODE_ANALYTIC64F);
};

// The EMHD mixed solve keeps each zone's metric (16 doubles) and EMHD
// weights (12) in LDS slots instead of VGPRs: held in registers across the
// Newton loop they pushed it past 256 VGPRs into scratch (224 bytes per
// lane).  Slots are 17 and 13 doubles apart, so that the 64 lanes of a b64
// read cover every bank.  Every residual evaluation and Jacobian assembly
// starts behind a compiler fence, so the slots are read where they are used
// instead of being hoisted back into registers for the whole loop.
template <class L, int MODE>
constexpr bool kLdsState = L::EMHD && MODE == MODE_MIXED;
constexpr int kGeomSlot = 17, kWeightSlot = 13;
constexpr size_t kLdsStateBytes = (size_t)(kGeomSlot + kWeightSlot) * sizeof(double);  // per lane

template <class L, int MODE>
__device__ __forceinline__ void slotFence() {
  if constexpr (kLdsState<L, MODE>) asm volatile("" ::: "memory");
}

// Per-zone quantities of the residual formed once per zone.
struct ZoneScale {
  double invAlpha;  // 1/alpha
  double scale;     // mask/dt
  double invDt;     // 1/dt
};

template <class L, int MODE>
__device__ __forceinline__ void evalResidual(const double* P, const ZoneGeom& G,
                                             const ZoneConst<L>& c, const DevParams& prm,
                                             double dt, const ZoneScale& zs, const EmhdWeights& ew,
                                             double* R) {
  slotFence<L, MODE>();
  if constexpr (ModeTraits<MODE>::fastResidual && !L::EMHD)
    residualFastIdeal<L>(P, G, c, prm, zs.invAlpha, zs.scale, R);
  else if constexpr (ModeTraits<MODE>::fastResidual)
    residualFastEmhd<L>(P, G, c, prm, ew, zs.invDt, R);
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

// grim's FD rule with the residual of MODE and one reciprocal per column.
template <class L, int MODE>
__device__ __forceinline__ void fdJacobianFast(const double* P, const double* R, const ZoneGeom& G,
                                               const ZoneConst<L>& c, const DevParams& prm,
                                               double dt, const ZoneScale& zs, const EmhdWeights& ew,
                                               double (&A)[L::NF][L::NF]);

// Small batches: grim's FD solve with one zone per kWide lanes of a
// wavefront.  Lane r + 1 of a zone's group evaluates the residual with P_r
// perturbed (grim's rule), so the NF columns of the FD Jacobian come from
// one residual evaluation's latency instead of NF; every other step runs on
// all lanes of the group with identical operands.  The result is the
// one-lane solve's bit for bit.
constexpr int kWide = 8;
constexpr int64_t kWideMaxZones = 16384;  // below this a one-lane grid leaves most CUs idle

template <class L>
__device__ __forceinline__ void fdJacobianWide(const double* P, const double* R, const ZoneGeom& G,
                                               const ZoneConst<L>& c, const DevParams& prm,
                                               double dt, double (&A)[L::NF][L::NF]) {
  static_assert(L::NF < kWide, "one lane per FD column");
  const int lane = threadIdx.x & 63;
  const int sub = lane % kWide;
  const int first = lane - sub;
  // lanes 1..NF take columns 0..NF-1; the others repeat column 0 (unused)
  const int row = (sub >= 1 && sub <= L::NF) ? sub - 1 : 0;
  const double eps = prm.jacEps;
  double Pp[L::DOF];
#pragma unroll
  for (int v = 0; v < L::DOF; v++) Pp[v] = P[v];
  double dP = 0.;
#pragma unroll
  for (int v = 0; v < L::NF; v++) {
    const bool small = fabs(P[v]) < .5 * eps;
    const double pv = small ? P[v] + eps : P[v] + eps * P[v];
    Pp[v] = (v == row) ? pv : P[v];
    dP = (v == row) ? pv - P[v] : dP;
  }
  double Rp[L::NF];
  residual<L>(Pp, G, c, prm, dt, Rp);
  double colv[L::NF];
#pragma unroll
  for (int k = 0; k < L::NF; k++) colv[k] = (Rp[k] - R[k]) / dP;
#pragma unroll
  for (int r = 0; r < L::NF; r++)
#pragma unroll
    for (int k = 0; k < L::NF; k++) A[k][r] = __shfl(colv[k], first + 1 + r, 64);
}

// One Newton update after the convergence test (solve.cpp:80-198): Jacobian,
// LU solve of J d = -R, quadratic backtracking line search, P += lambda d.
// Returns true when the last evaluated trial was accepted; Rt then holds the
// residual at the new P (bit-identical to a fresh evaluation, since the new P
// is computed by the same expression as the trial).
template <class L, int MODE, bool WIDE = false>
__device__ __forceinline__ bool newtonUpdate(double* P, const double* R, double l2,
                                             const ZoneGeom& G, const ZoneConst<L>& c,
                                             const DevParams& prm, double dt, const ZoneScale& zs,
                                             const EmhdWeights& ew, double* Rt, int& evals) {
  double d[L::NF];
  if constexpr (MODE == MODE_FD) {
    double A[L::NF][L::NF];
    if constexpr (WIDE)
      fdJacobianWide<L>(P, R, G, c, prm, dt, A);
    else
      fdJacobian<L>(P, R, G, c, prm, dt, A);
    evals += L::NF;
#pragma unroll
    for (int v = 0; v < L::NF; v++) d[v] = -R[v];
    luSolve<double, L::NF>(A, d);
  } else if constexpr (MODE == MODE_FD_FAST) {
    double A[L::NF][L::NF];
    fdJacobianFast<L, MODE>(P, R, G, c, prm, dt, zs, ew, A);
    evals += L::NF;
#pragma unroll
    for (int v = 0; v < L::NF; v++) d[v] = -R[v];
    luSolve<double, L::NF>(A, d);
  } else if constexpr (MODE == MODE_ANALYTIC64 || MODE == MODE_ANALYTIC64F) {
    double A[L::NF][L::NF];
    analyticJacobian<double, L>(P, G, prm, c.mask / dt, c, ew, dt, A);
#pragma unroll
    for (int v = 0; v < L::NF; v++) d[v] = -R[v];
    luSolve<double, L::NF>(A, d);
  } else {
    float A[L::NF][L::NF], df[L::NF];
    slotFence<L, MODE>();
    analyticJacobian<float, L>(P, G, prm, c.mask / dt, c, ew, dt, A);
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
  bool finiteTrial = true;
  for (int ls = 0; ls < prm.maxLS; ls++) {
#pragma unroll
    for (int v = 0; v < L::NF; v++) Pt[v] = P[v] + lambda * d[v];
    evalResidual<L, MODE>(Pt, G, c, prm, dt, zs, ew, Rt);
    evals += 1;
    const double f1 = 0.5 * sumSq<L::NF>(Rt);
    if constexpr (MODE == MODE_FD) {
      // grim's test: a non-finite trial compares false and is accepted
      const bool condition = f1 > f0 * (1. - 1e-4 * lambda) + prm.lsFloor;
      if (!condition) {
        accepted = true;
        break;
      }
    } else {
      // this build's solvers reject a non-finite trial and halve the step
      // (an FP32 Jacobian of a near-singular zone can overshoot into
      // states where the residual overflows)
      finiteTrial = isfinite(f1);
      if (f1 <= f0 * (1. - 1e-4 * lambda) + prm.lsFloor) {
        accepted = true;
        break;
      }
      if (!finiteTrial) {
        lambda = 0.5 * lambda;
        continue;
      }
    }
    const double denom = f1 - f0 - fPrime0 * lambda;
    lambda = -fPrime0 * lambda * lambda / denom / 2.;
  }
  // grim applies the last step length even when no trial passed; after a
  // non-finite last trial this build's solvers keep the zone where it is
  if (MODE != MODE_FD && !accepted && !finiteTrial) return false;
#pragma unroll
  for (int v = 0; v < L::NF; v++) P[v] = P[v] + lambda * d[v];
  return accepted;
}

// Newton iterations of one zone, from its current primitives, for at most
// `budget` updates (solve.cpp:10-200 read per zone: the zone stops when its
// own sum R^2 <= atol).  Writes the primitives back; stats are stored, or
// added to the earlier phase's when `accumulate`.  Returns true when the
// zone ran out of budget unconverged (a later phase may continue it: every
// iteration is a function of P alone, so stopping and resuming changes no
// result).  WIDE: the zone's kWide lanes run it together; only the first
// writes.
template <class L, int MODE, bool WIDE>
__device__ __forceinline__ bool solveZoneCore(const DevParams& prm, const grim_zone_batch& b,
                                              const grim_newton_stats& st, int64_t z, int budget,
                                              bool accumulate, bool lastPhase, double dt, const ZoneGeom& G,
                                              const ZoneConst<L>& c, const ZoneScale& zs, const EmhdWeights& ew,
                                              const double* start, const double* foldB);

// FOLDED: the zone's right-hand side, field rows and residual mask come from
// the caller's registers (foldRhs, foldB, foldMask: the stage fold computed
// in the same kernel) instead of the batch's planes.
template <class L, int MODE, bool WIDE = false, bool FOLDED = false>
__device__ __forceinline__ bool solveZone(const DevParams& prm, const grim_zone_batch& b,
                                          const grim_newton_stats& st, int64_t z, int budget,
                                          bool accumulate, bool lastPhase, const double* foldRhs = nullptr,
                                          const double* foldB = nullptr, double foldMask = 0.) {
  const int64_t n = b.n;
  const bool lead = !WIDE || (threadIdx.x % kWide) == 0;
  // the starting fluid rows: the caller's guess planes on a first pass,
  // else prim (a later phase resumes from what the earlier one wrote)
  const double* start = (b.guess && !accumulate) ? b.guess : b.prim;
  const double maskZ = FOLDED ? foldMask : (b.residual_mask ? b.residual_mask[z] : 1.);
  if ((FOLDED || b.residual_mask) && maskZ == 0.) {
    // outside the residual mask (ghost zones of a grid) R = 0: the zone
    // takes its floors and stops at the first test, as in the full path
    if (!lead) return false;
    if (start != b.prim)
#pragma unroll
      for (int v = U1; v < L::NF; v++) b.prim[v * n + z] = start[v * n + z];
    b.prim[RHO * n + z] = smax(start[RHO * n + z], prm.rhoFloor);
    b.prim[UU * n + z] = smax(start[UU * n + z], prm.uFloor);
    if (st.iterations) st.iterations[z] = accumulate ? st.iterations[z] : 0;
    if (st.residual_evals) st.residual_evals[z] = (accumulate ? st.residual_evals[z] : 0) + 1;
    if (st.final_l2) st.final_l2[z] = 0.;
    return false;
  }
  const double dt = batchDt(b);
  ZoneGeom G;
  loadBatchGeom(b, z, G);
  ZoneConst<L> c;
  loadZoneConst<L, !FOLDED>(b, prm, G, z, c);
  if constexpr (FOLDED) {
#pragma unroll
    for (int v = 0; v < L::NF; v++) c.rhs[v] = foldRhs[v];
    c.mask = maskZ;
  }
  ZoneScale zs;
  if constexpr (ModeTraits<MODE>::fastResidual) {
    zs.invAlpha = 1. / G.alpha;
    zs.scale = c.mask / dt;
    zs.invDt = 1. / dt;
  }
  EmhdWeights ew;
  if constexpr (L::EMHD && (ModeTraits<MODE>::analytic || ModeTraits<MODE>::fastResidual))
    emhdWeights<L>(G, c, prm, dt, ew);
  if constexpr (kLdsState<L, MODE>) {
    static_assert(sizeof(ZoneGeom) <= kGeomSlot * sizeof(double) &&
                  sizeof(EmhdWeights) <= kWeightSlot * sizeof(double), "LDS slot sizes");
    // the launch gives kLdsStateBytes of dynamic LDS per lane
    extern __shared__ double zoneSlots[];
    ZoneGeom* gs = reinterpret_cast<ZoneGeom*>(zoneSlots + (size_t)kGeomSlot * threadIdx.x);
    EmhdWeights* ws =
        reinterpret_cast<EmhdWeights*>(zoneSlots + (size_t)kGeomSlot * blockDim.x + (size_t)kWeightSlot * threadIdx.x);
    *gs = G;
    *ws = ew;
    return solveZoneCore<L, MODE, WIDE>(prm, b, st, z, budget, accumulate, lastPhase, dt, *gs, c, zs, *ws, start,
                                        FOLDED ? foldB : nullptr);
  } else {
    return solveZoneCore<L, MODE, WIDE>(prm, b, st, z, budget, accumulate, lastPhase, dt, G, c, zs, ew, start,
                                        FOLDED ? foldB : nullptr);
  }
}

template <class L, int MODE, bool WIDE>
__device__ __forceinline__ bool solveZoneCore(const DevParams& prm, const grim_zone_batch& b,
                                              const grim_newton_stats& st, int64_t z, int budget,
                                              bool accumulate, bool lastPhase, double dt, const ZoneGeom& G,
                                              const ZoneConst<L>& c, const ZoneScale& zs, const EmhdWeights& ew,
                                              const double* start, const double* foldB) {
  const int64_t n = b.n;
  const bool lead = !WIDE || (threadIdx.x % kWide) == 0;
  double P[L::DOF];
#pragma unroll
  for (int v = 0; v < L::DOF; v++)
    P[v] = v < L::NF ? start[v * n + z] : (foldB ? foldB[v - L::NF] : b.prim[v * n + z]);

  double R[L::NF];
  bool haveR = false;
  bool converged = false;
  int iters = 0, evals = 0;
  double l2 = 0.;
  for (int it = 0; it < budget; it++) {
    // floors at the top of every iteration (solve.cpp:17-18)
    const double p0 = smax(P[RHO], prm.rhoFloor);
    const double p1 = smax(P[UU], prm.uFloor);
    if (!(p0 == P[RHO] && p1 == P[UU])) haveR = false;
    P[RHO] = p0;
    P[UU] = p1;
    if (!haveR) {
      evalResidual<L, MODE>(P, G, c, prm, dt, zs, ew, R);
      evals++;
    }
    l2 = sumSq<L::NF>(R);
    if (!(l2 > prm.atol)) {
      converged = true;
      break;
    }
    double Rt[L::NF];
    haveR = newtonUpdate<L, MODE, WIDE>(P, R, l2, G, c, prm, dt, zs, ew, Rt, evals);
    iters++;
    if (haveR) {
#pragma unroll
      for (int v = 0; v < L::NF; v++) R[v] = Rt[v];
    }
  }
  if (!converged) {
    if (!haveR) {
      evalResidual<L, MODE>(P, G, c, prm, dt, zs, ew, R);
      if (!lastPhase) evals++;
    }
    l2 = sumSq<L::NF>(R);
    if (!lastPhase && !(l2 > prm.atol)) {
      // the one-pass solve would floor P at the top of its next iteration
      // and then stop: do the same here
      P[RHO] = smax(P[RHO], prm.rhoFloor);
      P[UU] = smax(P[UU], prm.uFloor);
      converged = true;
    }
  }
  if (!lead) return !converged && !lastPhase;
#pragma unroll
  for (int v = 0; v < L::NF; v++) b.prim[v * n + z] = P[v];
  if (st.iterations) st.iterations[z] = (accumulate ? st.iterations[z] : 0) + iters;
  if (st.residual_evals) st.residual_evals[z] = (accumulate ? st.residual_evals[z] : 0) + evals;
  if (st.final_l2) st.final_l2[z] = l2;
  return !converged && !lastPhase;
}

// Per-zone semantics, one pass: each thread iterates its own zone until
// sum R^2 <= atol or maxNonLinearIter updates.
//
// The EMHD solves need more than 256 registers per lane; asking for two
// wavefronts per SIMD (instead of one with AGPR spills) measured 8-10 %
// faster on the 3D torus, for the one-pass and the phase-1 kernel alike.
template <class L, int MODE>
__global__ __launch_bounds__(kDefaultBlock) __attribute__((amdgpu_waves_per_eu(2))) void newtonPerZone(
    DevParams prm, grim_zone_batch b, grim_newton_stats st) {
  const int64_t z = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (z >= b.n) return;
  solveZone<L, MODE>(prm, b, st, z, prm.maxIter, false, true);
}

// One-pass per-zone solve with the stragglers gathered inside the workgroup:
// every zone first gets budget1 updates; the zones still unconverged are then
// packed, by index, onto the workgroup's first lanes, which continue them for
// the remaining updates.  A zone's iterations are a function of its own
// primitives, so the results are the one-pass kernel's bit for bit, but the
// wavefronts left without a straggler end instead of idling through the last
// updates of a few lanes (about a quarter of the zones of a Kerr-Schild batch
// need a third update).  The zones the continuing lanes take are the
// workgroup's own, so their inputs come back from the caches.  It pays where
// updates are costly and a tail of zones needs many (grim's FD solve of the
// EMHD batch: 20.7 -> 15.0 ms); where the extra updates are few and cheap
// the two barriers and the restart cost more than they save (the mixed
// solve; the FD solve on the torus, 4.55 -> 5.05 ms per stage), so it is an
// option (grim_newton_opts.compaction = 2) the autotuner weighs.
template <class L, int MODE>
__global__ __launch_bounds__(kDefaultBlock) __attribute__((amdgpu_waves_per_eu(2))) void newtonPerZoneSplit(
    DevParams prm, grim_zone_batch b, grim_newton_stats st, int budget1) {
  __shared__ int list[kDefaultBlock];
  __shared__ int waveCount[kDefaultBlock / 64];
  const int64_t z = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  bool more = false;
  if (z < b.n) more = solveZone<L, MODE>(prm, b, st, z, budget1, false, false);
  const unsigned long long ballot = __ballot(more);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) waveCount[wave] = __popcll(ballot);
  __syncthreads();
  int base = 0, total = 0;
  for (int w = 0; w < (int)(blockDim.x >> 6); w++) {
    base += w < wave ? waveCount[w] : 0;
    total += waveCount[w];
  }
  if (more) list[base + __popcll(ballot & ((1ull << lane) - 1ull))] = (int)z;
  __syncthreads();
  if ((int)threadIdx.x < total)
    solveZone<L, MODE>(prm, b, st, (int64_t)list[threadIdx.x], prm.maxIter - budget1, true, true);
}

// The one-pass FD solve of a small batch, kWide lanes per zone (every lane
// of a group shares z, so groups leave together).
template <class L>
__global__ __launch_bounds__(kDefaultBlock) void newtonPerZoneWide(DevParams prm, grim_zone_batch b,
                                                                  grim_newton_stats st) {
  const int64_t z = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWide;
  if (z >= b.n) return;
  solveZone<L, MODE_FD, true>(prm, b, st, z, prm.maxIter, false, true);
}

// Compacting solve, phase 1: at most `budget1` updates per zone.  Each
// wavefront stores the 64-bit ballot of its lanes left unconverged (no
// atomics: a single global counter serialises 65536 wavefronts).
template <class L, int MODE>
__global__ __launch_bounds__(kDefaultBlock) __attribute__((amdgpu_waves_per_eu(2))) void newtonPhase1(
    DevParams prm, grim_zone_batch b, grim_newton_stats st, int budget1, unsigned long long* __restrict__ masks) {
  const int64_t z = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  bool more = false;
  if (z < b.n) more = solveZone<L, MODE>(prm, b, st, z, budget1, false, false);
  const unsigned long long ballot = __ballot(more);
  // lane 0 of a wavefront holding at least one zone (z < n) owns mask z/64;
  // wavefronts wholly past the end of the batch own none
  if ((threadIdx.x & 63) == 0 && z < b.n) masks[z >> 6] = ballot;
}

// Phase 1 of a grid stage's compacting solve with the stage fold in front of
// each zone's Newton iterations: a bulk zone forms its right-hand side, its
// new field rows and its residual scale from the face fluxes, EMFs, consOld
// and sources (foldZone, the fold kernel's arithmetic bit for bit), writes
// the field rows and the mask plane, and keeps the right-hand side in
// registers; only the zones left unconverged store it for phase 2.  The
// fold's memory traffic overlaps the solves' arithmetic.
template <class L, int MODE, int DIM>
__global__ __launch_bounds__(kDefaultBlock) __attribute__((amdgpu_waves_per_eu(2))) void newtonPhase1Fold(
    DevParams prm, grim_zone_batch b, grim_newton_stats st, int budget1, unsigned long long* __restrict__ masks,
    FoldArgs f) {
  const int64_t z = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  bool more = false;
  if (z < b.n) {
    const int64_t n = b.n;
    int i, j, k;
    zoneIndex(f.D, z, i, j, k);
    const bool bulk = i >= f.bulk.lo[0] && i < f.bulk.hi[0] && j >= f.bulk.lo[1] && j < f.bulk.hi[1] &&
                      k >= f.bulk.lo[2] && k < f.bulk.hi[2];
    double rhs[L::NF], B[3], mask = 0.;
#pragma unroll
    for (int v = 0; v < L::NF; v++) rhs[v] = 0.;
    B[0] = B[1] = B[2] = 0.;
    if (bulk) {
      double rmax;
      foldZone<L, DIM>(f.D, f.dX, f.F, f.emf, f.geomC, f.consOld, f.sExp, f.sImp, batchDt(b), z, i, j,
                       f.fluxFilter, rhs, B, rmax);
#pragma unroll
      for (int c = 0; c < 3; c++) b.prim[(L::NF + c) * n + z] = B[c];
      if (f.scaling) {
        mask = 1. / rmax;
        f.mask[z] = mask;
      } else {
        mask = f.mask[z];
      }
    } else {
#pragma unroll
      for (int c = 0; c < 3; c++) b.prim[(L::NF + c) * n + z] = f.fieldIn[c * n + z];
    }
    more = solveZone<L, MODE, false, true>(prm, b, st, z, budget1, false, false, rhs, B, mask);
    if (more)
#pragma unroll
      for (int v = 0; v < L::NF; v++) f.rhs[v * n + z] = rhs[v];
  }
  const unsigned long long ballot = __ballot(more);
  if ((threadIdx.x & 63) == 0 && z < b.n) masks[z >> 6] = ballot;
}

// The work list between the phases.  Phase 1 leaves one 64-bit mask of
// unconverged lanes per wavefront; an exclusive prefix sum of their
// popcounts numbers the unconverged zones, which are then scattered, in
// ascending zone order, into a dense list of zone indices.
constexpr int kScanBlock = 256;  // masks per scan workgroup

// Prefix sums of the popcounts within each workgroup's kScanBlock masks, and
// each workgroup's total.
static __global__ __launch_bounds__(kScanBlock) void scanMasksKernel(int64_t nMasks,
                                                              const unsigned long long* __restrict__ masks,
                                                              int* __restrict__ maskOff,
                                                              int* __restrict__ blockSum) {
  __shared__ int sh[kScanBlock];
  const int64_t m = (int64_t)blockIdx.x * kScanBlock + threadIdx.x;
  const int c = m < nMasks ? __popcll(masks[m]) : 0;
  sh[threadIdx.x] = c;
  __syncthreads();
  // Hillis-Steele inclusive scan in LDS
  for (int off = 1; off < kScanBlock; off <<= 1) {
    const int v = (int)threadIdx.x >= off ? sh[threadIdx.x - off] : 0;
    __syncthreads();
    sh[threadIdx.x] += v;
    __syncthreads();
  }
  if (m < nMasks) maskOff[m] = sh[threadIdx.x] - c;
  if (threadIdx.x == kScanBlock - 1) blockSum[blockIdx.x] = sh[threadIdx.x];
}

// Exclusive prefix over the workgroup totals (one workgroup, in chunks of
// kScanBlock), the list length into ctrl[0], and the grab counter ctrl[1]
// cleared for the persistent phase.
static __global__ __launch_bounds__(kScanBlock) void scanBlocksKernel(int64_t nBlocks, int* __restrict__ blockSum,
                                                               int* __restrict__ ctrl) {
  __shared__ int sh[kScanBlock];
  int carry = 0;
  for (int64_t base = 0; base < nBlocks; base += kScanBlock) {
    const int64_t q = base + threadIdx.x;
    const int c = q < nBlocks ? blockSum[q] : 0;
    sh[threadIdx.x] = c;
    __syncthreads();
    for (int off = 1; off < kScanBlock; off <<= 1) {
      const int v = (int)threadIdx.x >= off ? sh[threadIdx.x - off] : 0;
      __syncthreads();
      sh[threadIdx.x] += v;
      __syncthreads();
    }
    if (q < nBlocks) blockSum[q] = carry + sh[threadIdx.x] - c;
    const int chunk = sh[kScanBlock - 1];
    __syncthreads();
    carry += chunk;
  }
  if (threadIdx.x == 0) {
    ctrl[0] = carry;
    ctrl[1] = 0;
  }
}

// Each mask's unconverged zones into the list at their global rank.
static __global__ __launch_bounds__(kScanBlock) void scatterListKernel(int64_t nMasks,
                                                                const unsigned long long* __restrict__ masks,
                                                                const int* __restrict__ maskOff,
                                                                const int* __restrict__ blockOff,
                                                                int* __restrict__ list) {
  const int64_t m = (int64_t)blockIdx.x * kScanBlock + threadIdx.x;
  if (m >= nMasks) return;
  unsigned long long bits = masks[m];
  int at = blockOff[m / kScanBlock] + maskOff[m];
  while (bits) {
    const int bit = __ffsll((long long)bits) - 1;
    list[at++] = (int)(m * 64 + bit);
    bits &= bits - 1ull;
  }
}

// Compacting solve, phase 2: a persistent grid drains the list.  Each
// wavefront takes `grab` chunks of 64 list entries at a time (one vector
// atomic on the grab counter, by lane 0) and continues those zones for the
// remaining maxNonLinearIter - budget1 updates; it exits once the counter
// has passed the list length, which every wavefront reaches because the
// counter only grows and the length is fixed.
template <class L, int MODE>
__global__ __launch_bounds__(kDefaultBlock) void newtonPhase2(DevParams prm, grim_zone_batch b,
                                                              grim_newton_stats st, int budget2,
                                                              const int* __restrict__ list,
                                                              int* __restrict__ ctrl, int grab) {
  const int lane = threadIdx.x & 63;
  const int total = ctrl[0];
  for (;;) {
    int base = 0;
    if (lane == 0) base = atomicAdd(&ctrl[1], 64 * grab);
    base = __shfl(base, 0, 64);
    if (base >= total) break;
    for (int j = 0; j < grab; j++) {
      const int idx = base + j * 64 + lane;
      if (idx < total) solveZone<L, MODE>(prm, b, st, (int64_t)list[idx], budget2, true, true);
    }
  }
}

template <class L, int MODE>
__device__ __forceinline__ void fdJacobianFast(const double* P, const double* R, const ZoneGeom& G,
                                               const ZoneConst<L>& c, const DevParams& prm,
                                               double dt, const ZoneScale& zs, const EmhdWeights& ew,
                                               double (&A)[L::NF][L::NF]) {
  double Pp[L::DOF];
#pragma unroll
  for (int v = 0; v < L::DOF; v++) Pp[v] = P[v];
#pragma unroll
  for (int row = 0; row < L::NF; row++) {
    const double eps = prm.jacEps;
    const bool small = fabs(P[row]) < .5 * eps;
    Pp[row] = small ? P[row] + eps : P[row] + eps * P[row];
    double Rp[L::NF];
    evalResidual<L, MODE>(Pp, G, c, prm, dt, zs, ew, Rp);
    const double inv = 1. / (Pp[row] - P[row]);
#pragma unroll
    for (int col = 0; col < L::NF; col++) A[col][row] = (Rp[col] - R[col]) * inv;
    Pp[row] = P[row];
  }
}

// Carve-up of the caller's (or the library's) compaction workspace.
struct CompactionWorkspace {
  unsigned long long* masks;  // one per phase-1 wavefront
  int* maskOff;               // popcount prefix within each scan workgroup
  int* blockSum;              // scan workgroup totals, then their prefix
  int* list;                  // unconverged zones, ascending
  int* ctrl;                  // [0] list length, [1] grab counter
};

constexpr int64_t align256(int64_t x) { return (x + 255) / 256 * 256; }

// Bytes of the workspace for n zones, and the carve-up of a buffer of that size.
constexpr int64_t compactionBytes(int64_t n) {
  const int64_t nMasks = (n + 63) / 64, nScan = (nMasks + kScanBlock - 1) / kScanBlock;
  return align256(nMasks * 8) + align256(nMasks * 4) + align256(nScan * 4) + align256(n * 4) + 256;
}

inline CompactionWorkspace carveCompaction(void* base, int64_t n) {
  const int64_t nMasks = (n + 63) / 64, nScan = (nMasks + kScanBlock - 1) / kScanBlock;
  char* p = (char*)base;
  CompactionWorkspace w;
  w.masks = (unsigned long long*)p;
  p += align256(nMasks * 8);
  w.maskOff = (int*)p;
  p += align256(nMasks * 4);
  w.blockSum = (int*)p;
  p += align256(nScan * 4);
  w.list = (int*)p;
  p += align256(n * 4);
  w.ctrl = (int*)p;
  return w;
}

// Launch configuration of the per-zone solve (chosen by the autotuner or the
// caller).
struct LaunchCfg {
  int block;                 // threads per workgroup: 64, 128 or 256
  bool compaction;           // two-phase compacting solve
  bool workgroupSplit;       // one pass, stragglers gathered inside each workgroup (newtonPerZoneSplit)
  int budget1;               // phase-1 update budget
  int group;                 // phase-2 grab: chunks of 64 list entries per atomic (1..kMaxGroup)
  int persistentWaves;       // phase-2 grid, in wavefronts
  bool escalate;             // mixed path: phase 2 in FP64 (GRIM_LU_FP32_FP64)
  CompactionWorkspace ws;    // masks, prefix sums, work list, counters
};

// Dynamic LDS of a per-zone solve workgroup of `block` lanes.
template <class L, int MODE>
constexpr size_t ldsBytes(int block) {
  return kLdsState<L, MODE> ? kLdsStateBytes * (size_t)block : 0;
}

// Enqueue the per-zone solve of mode MODE on stream s (phase 2 of the
// compacting solve in MODE2).
// fold: a grid stage's fold runs inside phase 1 (newtonPhase1Fold; the
// caller has checked that the solve compacts, see foldFusable).
template <class L, int MODE, int MODE2 = MODE>
void launchSolve(const DevParams& prm, const grim_zone_batch& b, const grim_newton_stats& st,
                 const LaunchCfg& cfg, hipStream_t s, const FoldArgs* fold = nullptr) {
  const unsigned grid = (unsigned)((b.n + cfg.block - 1) / cfg.block);
  if (!cfg.compaction || cfg.budget1 >= prm.maxIter) {
    if constexpr (MODE == MODE_FD) {
      if (b.n <= kWideMaxZones) {
        const unsigned wideGrid = (unsigned)((b.n * kWide + kDefaultBlock - 1) / kDefaultBlock);
        newtonPerZoneWide<L><<<wideGrid, kDefaultBlock, 0, s>>>(prm, b, st);
        return;
      }
    }
    if (cfg.workgroupSplit && cfg.budget1 < prm.maxIter)
      newtonPerZoneSplit<L, MODE><<<grid, cfg.block, ldsBytes<L, MODE>(cfg.block), s>>>(prm, b, st, cfg.budget1);
    else
      newtonPerZone<L, MODE><<<grid, cfg.block, ldsBytes<L, MODE>(cfg.block), s>>>(prm, b, st);
    return;
  }
  static_assert(MODE2 == MODE || (MODE == MODE_MIXED && MODE2 == MODE_ANALYTIC64F), "phase-2 mode");
  const CompactionWorkspace& w = cfg.ws;
  if constexpr (MODE == MODE_MIXED) {
    if (fold)
      newtonPhase1Fold<L, MODE, 3><<<grid, cfg.block, ldsBytes<L, MODE>(cfg.block), s>>>(prm, b, st, cfg.budget1,
                                                                                        w.masks, *fold);
    else
      newtonPhase1<L, MODE><<<grid, cfg.block, ldsBytes<L, MODE>(cfg.block), s>>>(prm, b, st, cfg.budget1,
                                                                                 w.masks);
  } else {
    newtonPhase1<L, MODE><<<grid, cfg.block, ldsBytes<L, MODE>(cfg.block), s>>>(prm, b, st, cfg.budget1, w.masks);
  }
  const int64_t nMasks = (b.n + 63) / 64;
  const int64_t nScan = (nMasks + kScanBlock - 1) / kScanBlock;
  scanMasksKernel<<<(unsigned)nScan, kScanBlock, 0, s>>>(nMasks, w.masks, w.maskOff, w.blockSum);
  scanBlocksKernel<<<1, kScanBlock, 0, s>>>(nScan, w.blockSum, w.ctrl);
  scatterListKernel<<<(unsigned)nScan, kScanBlock, 0, s>>>(nMasks, w.masks, w.maskOff, w.blockSum, w.list);
  const int64_t waves = std::min<int64_t>(cfg.persistentWaves, nMasks);
  const unsigned grid2 = (unsigned)((waves * 64 + cfg.block - 1) / cfg.block);
  newtonPhase2<L, MODE2><<<grid2, cfg.block, ldsBytes<L, MODE2>(cfg.block), s>>>(
      prm, b, st, prm.maxIter - cfg.budget1, w.list, w.ctrl, cfg.group);
}

// Host launcher of the fast translation unit (newton_fast.hip).
int launchNewtonFast(int mode, const grim_params& p, const grim_zone_batch& b,
                     const grim_newton_stats& st, const LaunchCfg& cfg, hipStream_t s,
                     const FoldArgs* fold = nullptr);

}  // namespace grim
