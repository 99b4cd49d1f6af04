// The divergence of the face fluxes with flux-CT and the stage's fold of
// the frozen terms, shared by the fold kernel of the grid step and the
// compacting Newton solve that folds each zone's right-hand side itself.
//
// The arithmetic is fixed with fused multiply-add contraction off, whatever
// the including translation unit's -ffp-contract: the fold gives the same
// bits in every kernel that evaluates it.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "grim_hip.h"
#include "step_physics.hpp"
#include "zone_physics.hpp"

namespace grim {

// (i, j, k) of zone z, in 32-bit arithmetic (grids hold fewer than 2^31 zones)
__device__ __forceinline__ void zoneIndex(const Dims& D, int64_t z, int& i, int& j, int& k) {
  const unsigned u = (unsigned)z, n0 = (unsigned)D.NT[0], n1 = (unsigned)D.NT[1];
  i = (int)(u % n0);
  const unsigned r = u / n0;
  j = (int)(r % n1);
  k = (int)(r / n1);
}

// Flux of variable v through the lower face of zone z along d, after
// timeStepper::fluxCT (constrainedtransport.cpp:3-47) for dim >= 2.
template <class L, int DIM = 0>
__device__ __forceinline__ double ctFlux(const Dims& D, const double* __restrict__ F,
                                         const double* __restrict__ emf, int d, int v, int64_t z) {
#pragma clang fp contract(off)
  const int dim = DIM ? DIM : D.dim;  // DIM 0: the grid's dimension read at run time
  const int64_t n = D.n;
  if (dim >= 2 && v >= L::B1) {
    const double* e1 = emf;
    const double* e2 = emf + n;
    const double* e3 = emf + 2 * n;
    const int64_t s1 = D.s[0], s2 = D.s[1], s3 = D.s[2];
    const int b = v - L::B1;
    if (d == 0) {
      if (b == 0) return 0.;
      if (b == 1) return 0.5 * (e3[z] + e3[z + s2]);
      if (dim == 3) return -0.5 * (e2[z] + e2[z + s3]);
    } else if (d == 1) {
      if (b == 0) return -0.5 * (e3[z] + e3[z + s1]);
      if (b == 1) return 0.;
      if (dim == 3) return 0.5 * (e1[z] + e1[z + s3]);
    } else {
      if (b == 0) return 0.5 * (e2[z] + e2[z + s1]);
      if (b == 1) return -0.5 * (e1[z] + e1[z + s2]);
      return 0.;
    }
  }
  return F[((size_t)d * L::DOF + v) * n + z];
}

// Divergence of the fluxes of zone z, variable v (fvmfluxes.cpp:36-46,
// 91-103, 164-180): dim < 3 is af::convolve with {1, -1, 0}/dX,
// dim 3 is (shift(F, -1) - F)/dX.  With the dimension and the variable known
// at compile time every branch folds away and the flux filter becomes
// selects, so the loads of all faces issue together.
template <class L, int DIM>
__device__ __forceinline__ double divergence(const Dims& D, const double* dX,
                                             const double* __restrict__ F,
                                             const double* __restrict__ emf, int v, int64_t z,
                                             int i, int j, int filter) {
#pragma clang fp contract(off)
  double acc = 0.;
#pragma unroll
  for (int d = 0; d < DIM; d++) {
    double f0 = ctFlux<L, DIM>(D, F, emf, d, v, z);
    double f1 = ctFlux<L, DIM>(D, F, emf, d, v, z + D.s[d]);
    // The torus's flux filter (problem.cpp:1243-1285), after flux-CT: no
    // mass inflow through the radial boundary faces, nothing through the
    // polar faces.
    if (d == 0 && v == RHO) {
      f0 = (filter && i == D.g[0]) ? stdmin(f0, 0.) : f0;
      f1 = (filter && i + 1 == D.N[0] + D.g[0]) ? stdmax(f1, 0.) : f1;
    }
    if (d == 1) {
      f0 = (filter && j == D.g[1]) ? 0. : f0;
      f1 = (filter && j + 1 == D.N[1] + D.g[1]) ? 0. : f1;
    }
    double term;
    if constexpr (DIM < 3) {
      const double w = 1. / dX[d];
      term = f1 * w + f0 * (-w);
    } else {
      term = (f1 - f0) / dX[d];
    }
    acc = (d == 0) ? term : acc + term;
  }
  return acc;
}

// The stage fold of bulk zone z = (i, j, .) (timestep.cpp:39-45, 80-89):
//   rhs = consOld - dt'((divF + S_expl) + 1/2 S_implOld)       (fluid rows)
//   B   = (consOld_B - dt' divF_B) / g                        (field rows)
// and rmax = max |rhs| (the residual scale is 1/rmax).
template <class L, int DIM>
__device__ __forceinline__ void foldZone(const Dims& D, const double* dX, const double* __restrict__ F,
                                         const double* __restrict__ emf, const double* __restrict__ geomC,
                                         const double* __restrict__ consOld, const double* __restrict__ sExp,
                                         const double* __restrict__ sImp, double dtPrime, int64_t z, int i, int j,
                                         int filter, double (&rhs)[L::NF], double (&B)[3], double& rmax) {
#pragma clang fp contract(off)
  const int64_t n = D.n;
  rmax = 0.;
#pragma unroll
  for (int v = 0; v < L::DOF; v++) {
    const double div = divergence<L, DIM>(D, dX, F, emf, v, z, i, j, filter);
    const double U = consOld[v * n + z];
    if (v < L::NF) {
      const double si = L::EMHD ? sImp[v * n + z] : 0.;
      const double r = U - dtPrime * ((div + sExp[v * n + z]) + 0.5 * si);
      rhs[v] = r;
      rmax = stdmax(rmax, fabs(r));
    } else {
      const double consB = U - dtPrime * div;
      B[v - L::NF] = consB / geomC[GEOM_G * D.ng + geomIndex(D, i, j)];
    }
  }
}

// What a compacting stage solve needs to fold each zone's right-hand side
// itself (the fold kernel's inputs and outputs).
struct FoldArgs {
  Dims D;
  Box bulk;
  double dX[3];
  int fluxFilter;
  int scaling;  // residual scale 1/max|rhs| into mask (else mask is read)
  const double* F;
  const double* emf;
  const double* geomC;
  const double* consOld;
  const double* sExp;
  const double* sImp;
  double* rhs;   // written for the zones left to the second phase
  double* mask;  // the residual mask plane
  // the stage state's field rows (3 planes): zones outside the bulk take
  // them into prim's field rows, so no copy of the rows precedes the solve
  const double* fieldIn;
};

// Host: whether a grid stage's solve with these options compacts with the
// mixed analytic Jacobian on a 3D grid, so that its phase 1 can fold the
// stage's right-hand side itself.
bool foldFusable(const grim_params& p, const grim_newton_opts& o, int dim);

// grim_batched_newton, with the stage fold inside phase 1 when fold is given
// (then foldFusable must hold).
int batchedNewton(const grim_params* p, const grim_zone_batch* b, const grim_newton_opts* opts,
                  grim_newton_stats* stats, hipStream_t stream, const FoldArgs* fold);

}  // namespace grim
