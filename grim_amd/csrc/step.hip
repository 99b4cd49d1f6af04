// The full time step on a grid: grim's timeStepper on one gfx950 device.
//
// One context owns the grid arrays.  A stage is a short chain of kernels on
// the context's stream:
//   boundaries   -> ghost slabs of the stage state
//   [EMHD] uT    -> u_mu and T of every zone (the arrays grim's slopes read)
//   sources      -> S_expl, S_implOld and (half stage) consOld, bulk zones
//   face fluxes  -> one kernel per direction: reconstruction of both sides,
//                   two elements, fluxes, conserved variables, speeds, HLL/LLF
//   EMF          -> edge EMFs for flux-CT (dim >= 2)
//   fold         -> divergence (CT fluxes formed on the fly), the folded
//                   right-hand side of the Newton residual and the induction
//                   update of B, bulk zones
//   Newton       -> grim_batched_newton on the whole grid (mask 0 outside the
//                   bulk: those zones are not solved)
//   communicate  -> periodic ghost fill of the result
// Every kernel evaluates grim's expressions in grim's order (built with
// -ffp-contract=off), so a stage agrees bit for bit with the CPU FP64
// restatement given the same geometry.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <new>
#include <vector>

#include "grim_hip.h"
#include "runtime.hpp"
#include "batchgen.hpp"
#include "step_fast.hpp"
#include "fold.hpp"
#include "step_physics.hpp"
#include "torus.hpp"
#include "zone_physics.hpp"

struct grim_ctx;

namespace grim {
namespace {


inline unsigned blocksFor(int64_t count) { return (unsigned)((count + kBlock - 1) / kBlock); }

// ---------------------------------------------------------------------------
// Geometry (set-up): coordinates (grid.cpp:168-249), metric, inverse, alpha,
// g and r at one location; the connection at the centre (geometry.cpp:61-121,
// 631-680).
// ---------------------------------------------------------------------------
struct CoordGrid {
  double start[3], dX[3];
};

__device__ __forceinline__ void XCoordsAt(const Dims& D, const CoordGrid& cg, int i, int j, int k,
                                          int loc, double* X) {
  const double i1 = (double)i - D.g[0], i2 = (double)j - D.g[1], i3 = (double)k - D.g[2];
  X[0] = cg.start[0] + (loc == 1 ? (i1)*cg.dX[0] : (i1 + 0.5) * cg.dX[0]);
  X[1] = cg.start[1] + (loc == 2 ? (i2)*cg.dX[1] : (i2 + 0.5) * cg.dX[1]);
  X[2] = cg.start[2] + (i3 + 0.5) * cg.dX[2];
}

// One thread per (X1, X2) column (geomIndex): the metric does not depend on
// X3, so the column's value is every zone's of that column.
__global__ __launch_bounds__(kBlock) void geometryKernel(Dims D, CoordMap cm, CoordGrid cg, int loc,
                                                         double* __restrict__ out) {
  const int64_t z = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (z >= D.ng) return;
  const int i = (int)(z % D.NT[0]), j = (int)(z / D.NT[0]), k = 0;
  double X[3], x[3], G[4][4], C[4][4], det;
  XCoordsAt(D, cg, i, j, k, loc, X);
  cm.toX(X, x);
  cm.gCov(X, G);
  gDetAndgCon(G, det, C);
  const int64_t n = D.ng;
  const int cov[10][2] = {{0, 0}, {0, 1}, {0, 2}, {0, 3}, {1, 1}, {1, 2}, {1, 3}, {2, 2}, {2, 3}, {3, 3}};
#pragma unroll
  for (int q = 0; q < 10; q++) out[q * n + z] = G[cov[q][0]][cov[q][1]];
#pragma unroll
  for (int q = 0; q < 4; q++) out[(10 + q) * n + z] = C[0][q];
  out[14 * n + z] = 1. / sqrt(-C[0][0]);
  out[15 * n + z] = sqrt(-det);
#pragma unroll
  for (int q = 0; q < 6; q++) out[(16 + q) * n + z] = C[cov[4 + q][0]][cov[4 + q][1]];
  out[LOC_R * n + z] = x[0];
  out[LOC_THETA * n + z] = x[1];
}

__global__ __launch_bounds__(kBlock) void connectionKernel(Dims D, CoordMap cm, CoordGrid cg,
                                                           double* __restrict__ gamma) {
  const int64_t z = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (z >= D.ng) return;
  const int i = (int)(z % D.NT[0]), j = (int)(z / D.NT[0]), k = 0;
  const double EPS = 1.e-5;
  double X[3], G[4][4], C[4][4], det;
  XCoordsAt(D, cg, i, j, k, 0, X);
  cm.gCov(X, G);
  gDetAndgCon(G, det, C);
  double Gp[3][4][4], Gm[3][4][4];
  for (int d = 0; d < 3; d++) {
    double Xp[3] = {X[0], X[1], X[2]}, Xm[3] = {X[0], X[1], X[2]};
    Xp[d] += EPS;
    Xm[d] -= EPS;
    cm.gCov(Xp, Gp[d]);
    cm.gCov(Xm, Gm[d]);
  }
  for (int mu = 0; mu < 4; mu++)
    for (int nu = 0; nu < 4; nu++)
      for (int la = 0; la < 4; la++) {
        double s = 0.;
        for (int eta = 0; eta < 4; eta++) {
          double out = 0.;
          if (la > 0) {
            out += 0.5 * Gp[la - 1][eta][nu] / (2. * EPS);
            out -= 0.5 * Gm[la - 1][eta][nu] / (2. * EPS);
          }
          if (nu > 0) {
            out += 0.5 * Gp[nu - 1][eta][la] / (2. * EPS);
            out -= 0.5 * Gm[nu - 1][eta][la] / (2. * EPS);
          }
          if (eta > 0) {
            out -= 0.5 * Gp[eta - 1][nu][la] / (2. * EPS);
            out += 0.5 * Gm[eta - 1][nu][la] / (2. * EPS);
          }
          s += C[mu][eta] * out;
        }
        gamma[(mu * 16 + nu * 4 + la) * D.ng + z] = s;
      }
}

// dst[q n + z] = src[q ng + z mod ng]: planes of one value per (X1, X2)
// column written out for every zone.
__global__ __launch_bounds__(kBlock) void expandColumnsKernel(int64_t count, int64_t n, int64_t ng,
                                                              const double* __restrict__ src,
                                                              double* __restrict__ dst) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= count) return;
  const int64_t q = t / n, z = t - q * n;
  dst[t] = src[q * ng + z % ng];
}

__global__ __launch_bounds__(kBlock) void maskKernel(Dims D, double* __restrict__ mask) {
  const int64_t z = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (z >= D.n) return;
  int i, j, k;
  zoneIndex(D, z, i, j, k);
  const bool bulk = i >= D.g[0] && i < D.g[0] + D.N[0] && j >= D.g[1] && j < D.g[1] + D.N[1] &&
                    k >= D.g[2] && k < D.g[2] + D.N[2];
  mask[z] = bulk ? 1. : 0.;
}

// ---------------------------------------------------------------------------
// Boundaries
// ---------------------------------------------------------------------------
struct BcTypes {
  int lo[3], hi[3];
};

// grid::communicate on one rank (grid.cpp:88-121, 306-360): ghost zones
// whose out-of-range coordinates all lie along periodic directions take the
// periodic image (DMDA box stencil); the others lie outside the global
// domain, which DMGlobalToLocal never writes, and read back the local
// vector's initial zeros.  Reads only bulk zones, so zones are independent.
__global__ __launch_bounds__(kBlock) void communicateKernel(Dims D, BcTypes bc, int nvars,
                                                            double* __restrict__ prim) {
  const int64_t z = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (z >= D.n) return;
  int I[3];
  zoneIndex(D, z, I[0], I[1], I[2]);
  bool ghost = false, outside = false;
#pragma unroll
  for (int d = 0; d < 3; d++) {
    if (D.g[d] == 0) continue;
    if (I[d] < D.g[d]) {
      if (bc.lo[d] != GRIM_BC_PERIODIC) outside = true;
      I[d] += D.N[d];
      ghost = true;
    } else if (I[d] >= D.g[d] + D.N[d]) {
      if (bc.hi[d] != GRIM_BC_PERIODIC) outside = true;
      I[d] -= D.N[d];
      ghost = true;
    }
  }
  if (!ghost) return;
  const int64_t w = flatIndex(D, I[0], I[1], I[2]);
  // every load first, then every store (w is a bulk zone: nothing here
  // writes it), so the loads of all variables are in flight together
  double val[GRIM_MAX_VARS];
#pragma unroll
  for (int v = 0; v < GRIM_MAX_VARS; v++)
    if (v < nvars) val[v] = outside ? 0. : prim[v * D.n + w];
#pragma unroll
  for (int v = 0; v < GRIM_MAX_VARS; v++)
    if (v < nvars) prim[v * D.n + z] = val[v];
}

// One ghost slab of boundaries::applyBoundaryConditions (boundary.cpp:32-239):
// outflow copies the first interior layer, mirror the reflected layer.
__global__ __launch_bounds__(kBlock) void boundaryKernel(Dims D, Box box, int d, int side, int type,
                                                         int nvars, double* __restrict__ prim) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= box.count()) return;
  int I[3];
  box.at(t, I[0], I[1], I[2]);
  const int g = D.g[d], N = D.N[d], a = I[d];
  const int64_t z = flatIndex(D, I[0], I[1], I[2]);
  if (type == GRIM_BC_MIRROR)
    I[d] = side == 0 ? 2 * g - 1 - a : 2 * (N + g) - 1 - a;
  else
    I[d] = side == 0 ? g : N + g - 1;
  const int64_t w = flatIndex(D, I[0], I[1], I[2]);
  // w lies inside along d, z in the ghost slab: loads first, then stores
  double val[GRIM_MAX_VARS];
#pragma unroll
  for (int v = 0; v < GRIM_MAX_VARS; v++)
    if (v < nvars) val[v] = prim[v * D.n + w];
#pragma unroll
  for (int v = 0; v < GRIM_MAX_VARS; v++)
    if (v < nvars) prim[v * D.n + z] = val[v];
}

// ---------------------------------------------------------------------------
// Reconstruction and face fluxes
// ---------------------------------------------------------------------------
template <int RECON>
__global__ __launch_bounds__(kBlock) void reconstructKernel(Dims D, Box box, int64_t s, double theta,
                                                            int nvars, const double* __restrict__ prim,
                                                            double* __restrict__ left,
                                                            double* __restrict__ right) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= box.count()) return;
  int i, j, k;
  box.at(t, i, j, k);
  const int64_t z = flatIndex(D, i, j, k);
  for (int v = 0; v < nvars; v++) {
    double y[5];
#pragma unroll
    for (int q = 0; q < 5; q++) y[q] = prim[v * D.n + z + (q - 2) * s];
    double l, r;
    reconstructZone<RECON>(y, theta, l, r);
    left[v * D.n + z] = l;
    right[v * D.n + z] = r;
  }
}

// Flux at the lower face (i - 1/2 along DIR) of zone z: riemannSolver::solve
// (riemannsolver.cpp:202-328) for one face.  The left state is the right
// face value of zone z - s, the right state the left face value of zone z;
// both sit at the face geometry.
//
// The two states of a face are built on different wavefronts: a 256-thread
// workgroup takes 128 faces; wavefronts 0-1 build the left
// states (right face value of zone z - s), wavefronts 2-3 the right states
// (left face value of zone z).  Each thread reconstructs only its own side
// and evaluates one element, so the branch is uniform per wavefront and a
// thread holds half the state a one-thread-per-face kernel would (which
// needs 256 VGPRs plus spills and runs one wavefront per SIMD).  The right
// states' fluxes, conserved variables and speeds meet the left states in
// LDS.  The arithmetic follows grim's expression order exactly.

// RECON = GIVEN_STATES: the face values come ready-made, as
// reconstruction::reconstruct leaves them (prim holds primRight on the left
// side, i.e. the right face value of zone z - s, and primLeft on the right
// side), which is riemannSolver::solve's own interface.
constexpr int GIVEN_STATES = -1;

template <class L, int DIR, int RECON, int SIDE>
__device__ __forceinline__ void faceState(const StepArgs& a, int64_t z, int64_t zg, const double* __restrict__ prim,
                                          const double* __restrict__ geomFace, double (&F)[L::DOF],
                                          double (&U)[L::DOF], double& mn, double& mx) {
  const int64_t n = a.D.n, s = a.D.s[DIR];
  double P[L::DOF];
#pragma unroll
  for (int v = 0; v < L::DOF; v++) {
    if constexpr (RECON == GIVEN_STATES) {
      P[v] = prim[v * n + z - (SIDE ? 0 : s)];
    } else {
      double y[5];
#pragma unroll
      for (int q = 0; q < 5; q++) y[q] = prim[v * n + z + (q - 3 + SIDE) * s];
      double l, r;
      reconstructZone<RECON>(y, a.theta, l, r);
      P[v] = SIDE ? l : r;
    }
  }
  GeomFull G;
  loadGeomFull(geomFace, a.D.ng, zg, G);
  Kinematics kk;
  Closure cc;
  kinematics<L>(P, G.zg, a.prm, kk);
  elementClosure<L>(P, kk, a.prm, G.r, cc);
  fluxes<L, DIR + 1>(P, kk, cc, G.zg.g, F);
  fluxes<L, 0>(P, kk, cc, G.zg.g, U);
  charSpeeds<L, DIR>(kk, cc, G, a.prm, mn, mx);
}

// Only the right states pass through LDS (22.5 KB per workgroup), so LDS
// does not cap the occupancy; the left-state wavefronts combine and write
// every variable of their faces.
template <class L, int DIR, int RECON>
__global__ __launch_bounds__(kBlock) void faceFluxSplitKernel(
    StepArgs a, Box box, const double* __restrict__ prim, const double* __restrict__ geomFace,
    double* __restrict__ flux, const double* __restrict__ primRight = nullptr) {
  __shared__ double sh[2 * L::DOF + 2][kFacesPerBlock];
  const int side = threadIdx.x / kFacesPerBlock;
  const int lf = threadIdx.x % kFacesPerBlock;
  const int64_t f = (int64_t)blockIdx.x * kFacesPerBlock + lf;
  const bool active = f < box.count();
  const Dims& D = a.D;
  const int64_t n = D.n;
  int64_t z = 0;
  double F[L::DOF], U[L::DOF], mn = 0., mx = 0.;
  if (active) {
    int i, j, k;
    box.at(f, i, j, k);
    z = flatIndex(D, i, j, k);
    const int64_t zg = geomIndex(D, i, j);
    if (side == 0) {
      faceState<L, DIR, RECON, 0>(a, z, zg, RECON == GIVEN_STATES ? primRight : prim, geomFace, F, U, mn, mx);
    } else {
      faceState<L, DIR, RECON, 1>(a, z, zg, prim, geomFace, F, U, mn, mx);
#pragma unroll
      for (int v = 0; v < L::DOF; v++) {
        sh[v][lf] = F[v];
        sh[L::DOF + v][lf] = U[v];
      }
      sh[2 * L::DOF][lf] = mn;
      sh[2 * L::DOF + 1][lf] = mx;
    }
  }
  __syncthreads();
  if (!active || side != 0) return;
  const double minSpeed = stdmin(mn, sh[2 * L::DOF][lf]);
  const double maxSpeed = stdmax(mx, sh[2 * L::DOF + 1][lf]);
  if (a.riemann == GRIM_RIEMANN_HLL) {
#pragma unroll
    for (int v = 0; v < L::DOF; v++) {
      const double FR = sh[v][lf], UR = sh[L::DOF + v][lf];
      flux[v * n + z] = (maxSpeed * F[v] - minSpeed * FR + minSpeed * maxSpeed * (UR - U[v])) / (maxSpeed - minSpeed);
    }
  } else {
    const double c = stdmax(maxSpeed, -minSpeed);
#pragma unroll
    for (int v = 0; v < L::DOF; v++) {
      const double FR = sh[v][lf], UR = sh[L::DOF + v][lf];
      flux[v * n + z] = 0.5 * (F[v] + FR - c * (UR - U[v]));
    }
  }
}

// timeStepper::computeEMF (constrainedtransport.cpp:49-83) on the edges the
// bulk divergence reads.
template <class L>
__global__ __launch_bounds__(kBlock) void emfKernel(Dims D, Box box, const double* __restrict__ F,
                                                    double* __restrict__ emf) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= box.count()) return;
  int i, j, k;
  box.at(t, i, j, k);
  const int64_t n = D.n, z = flatIndex(D, i, j, k), s1 = D.s[0], s2 = D.s[1], s3 = D.s[2];
  const double* F1 = F;
  const double* F2 = F + (size_t)L::DOF * n;
  const double* F3 = F + (size_t)2 * L::DOF * n;
  const double* F1B2 = F1 + L::B2 * n;
  const double* F2B1 = F2 + L::B1 * n;
  emf[2 * n + z] = 0.25 * (F1B2[z] + F1B2[z - s2] - F2B1[z] - F2B1[z - s1]);
  if (D.dim == 3) {
    const double* F2B3 = F2 + L::B3 * n;
    const double* F3B2 = F3 + L::B2 * n;
    const double* F3B1 = F3 + L::B1 * n;
    const double* F1B3 = F1 + L::B3 * n;
    emf[0 * n + z] = 0.25 * (F2B3[z] + F2B3[z - s3] - F3B2[z] - F3B2[z - s2]);
    emf[1 * n + z] = 0.25 * (F3B1[z] + F3B1[z - s1] - F1B3[z] - F1B3[z - s3]);
  }
}

// Bulk zones of a DIM-dimensional grid.  FOLD: the stage fold
//   rhs = consOld - dt'((divF + S_expl) + 1/2 S_implOld)      (fluid rows)
//   B   = (consOld_B - dt' divF_B) / g                        (timestep.cpp:80-89)
// and, when mask is given, the residual scale 1/max|rhs|; otherwise divF
// into divOut.
template <class L, int DIM, bool FOLD>
__global__ __launch_bounds__(kBlock) void foldKernel(StepArgs a, Box box, const double* __restrict__ F,
                                                     const double* __restrict__ emf,
                                                     const double* __restrict__ geomC,
                                                     const double* __restrict__ consOld,
                                                     const double* __restrict__ sExp,
                                                     const double* __restrict__ sImp,
                                                     const double* __restrict__ dtPrimeDev,
                                                     double* __restrict__ divOut,
                                                     double* __restrict__ rhs,
                                                     double* __restrict__ primNew,
                                                     double* __restrict__ mask) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= box.count()) return;
  int i, j, k;
  box.at(t, i, j, k);
  const Dims& D = a.D;
  const int64_t n = D.n, z = flatIndex(D, i, j, k);
  if constexpr (!FOLD) {
#pragma unroll
    for (int v = 0; v < L::DOF; v++) divOut[v * n + z] = divergence<L, DIM>(D, a.dX, F, emf, v, z, i, j, a.fluxFilter);
  } else {
    double r[L::NF], B[3], rmax;
    foldZone<L, DIM>(D, a.dX, F, emf, geomC, consOld, sExp, sImp, *dtPrimeDev, z, i, j, a.fluxFilter, r, B, rmax);
#pragma unroll
    for (int v = 0; v < L::NF; v++) rhs[v * n + z] = r[v];
#pragma unroll
    for (int c = 0; c < 3; c++) primNew[(L::NF + c) * n + z] = B[c];
    if (mask) mask[z] = 1. / rmax;
  }
}

// foldKernel for the grid's dimension.
template <class L, bool FOLD>
void launchFold(const StepArgs& a, const Box& b, hipStream_t s, const double* F, const double* emf,
                const double* geomC, const double* consOld, const double* sExp, const double* sImp,
                const double* dtPrimeDev, double* divOut, double* rhs, double* primNew, double* mask) {
  const unsigned nb = blocksFor(b.count());
  switch (a.D.dim) {
    case 1:
      foldKernel<L, 1, FOLD><<<nb, kBlock, 0, s>>>(a, b, F, emf, geomC, consOld, sExp, sImp, dtPrimeDev, divOut,
                                                   rhs, primNew, mask);
      break;
    case 2:
      foldKernel<L, 2, FOLD><<<nb, kBlock, 0, s>>>(a, b, F, emf, geomC, consOld, sExp, sImp, dtPrimeDev, divOut,
                                                   rhs, primNew, mask);
      break;
    default:
      foldKernel<L, 3, FOLD><<<nb, kBlock, 0, s>>>(a, b, F, emf, geomC, consOld, sExp, sImp, dtPrimeDev, divOut,
                                                   rhs, primNew, mask);
  }
}

// Post-CT flux arrays for the faces of the bulk zones (API output only).
template <class L>
__global__ __launch_bounds__(kBlock) void ctWriteKernel(Dims D, Box box, const double* __restrict__ F,
                                                        const double* __restrict__ emf,
                                                        double* __restrict__ out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= box.count()) return;
  int i, j, k;
  box.at(t, i, j, k);
  const int64_t z = flatIndex(D, i, j, k);
  for (int d = 0; d < D.dim; d++)
    for (int v = L::B1; v < L::DOF; v++) out[((size_t)d * L::DOF + v) * D.n + z] = ctFlux<L>(D, F, emf, d, v, z);
}

// ---------------------------------------------------------------------------
// Sources
// ---------------------------------------------------------------------------
// u_mu and T of every zone of the stage state (EMHD gradients read them).
template <class L>
__global__ __launch_bounds__(kBlock) void uTKernel(StepArgs a, const double* __restrict__ prim,
                                                   const double* __restrict__ geomC,
                                                   double* __restrict__ uT) {
  const int64_t z = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t n = a.D.n;
  if (z >= n) return;
  double P[L::DOF];
#pragma unroll
  for (int v = 0; v < L::DOF; v++) P[v] = prim[v * n + z];
  ZoneGeom G;
  loadGeom(geomC, a.D.ng, geomIndexOf(a.D, z), G);
  Kinematics kk;
  kinematics<L>(P, G, a.prm, kk);
#pragma unroll
  for (int m = 0; m < 4; m++) uT[m * n + z] = kk.uCov[m];
  uT[4 * n + z] = kk.temperature;
}

// Bulk zones: fluidElement::computeExplicitSources with computeEMHDGradients
// (physics.cpp:360-586) of the stage element; computeImplicitSources of
// elemOld with the stage element's tau (timestep.cpp:56-60, 211-215); on the
// half stage also consOld = computeFluxes(0) of elemOld (timestep.cpp:33-36).
template <class L, int RECON>
__global__ __launch_bounds__(kBlock) void sourcesKernel(StepArgs a, Box box,
                                                        const double* __restrict__ prim,
                                                        const double* __restrict__ primOld,
                                                        const double* __restrict__ geomC,
                                                        const double* __restrict__ gamma,
                                                        const double* __restrict__ uT,
                                                        double* __restrict__ sExp,
                                                        double* __restrict__ sImp,
                                                        double* __restrict__ consOld) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= box.count()) return;
  int i, j, k;
  box.at(t, i, j, k);
  const Dims& D = a.D;
  const int64_t n = D.n, z = flatIndex(D, i, j, k);
  const int64_t ng = D.ng, zg = geomIndex(D, i, j);
  double P[L::DOF];
#pragma unroll
  for (int v = 0; v < L::DOF; v++) P[v] = prim[v * n + z];
  GeomFull G;
  loadGeomFull(geomC, ng, zg, G);
  Kinematics kk;
  Closure cc;
  kinematics<L>(P, G.zg, a.prm, kk);
  elementClosure<L>(P, kk, a.prm, G.r, cc);
  const double g = G.zg.g;
  if (consOld) {
    double U[L::DOF];
    fluxes<L, 0>(P, kk, cc, g, U);
#pragma unroll
    for (int v = 0; v < L::DOF; v++) consOld[v * n + z] = U[v];
  }
  double src[L::NF];
#pragma unroll
  for (int v = 0; v < L::NF; v++) src[v] = 0.;
  if (a.metric != GRIM_METRIC_MINKOWSKI) {
    // One row of T^kappa_lambda at a time: every src[UU + nu] still receives
    // its 16 terms in grim's (kappa, lambda) order, while only four stress
    // values and 16 connection planes are live at once.
    auto contract = [&](const double (&Tk)[4], int kappa) {
#pragma unroll
      for (int lamda = 0; lamda < 4; lamda++)
#pragma unroll
        for (int nu = 0; nu < 4; nu++)
          src[UU + nu] -= g * Tk[lamda] * gamma[(lamda * 16 + kappa * 4 + nu) * ng + zg];
    };
    {
      double T[4];
      stressRow<L, 0>(kk, cc, T);
      contract(T, 0);
    }
    {
      double T[4];
      stressRow<L, 1>(kk, cc, T);
      contract(T, 1);
    }
    {
      double T[4];
      stressRow<L, 2>(kk, cc, T);
      contract(T, 2);
    }
    {
      double T[4];
      stressRow<L, 3>(kk, cc, T);
      contract(T, 3);
    }
  }
  if constexpr (L::EMHD) {
    auto gam = [&](int l, int p, int q) { return gamma[(l * 16 + p * 4 + q) * ng + zg]; };
    auto slopeOf = [&](const double* arr, int d) {
      double y[5];
#pragma unroll
      for (int q = 0; q < 5; q++) y[q] = arr[z + (q - 2) * D.s[d]];
      return slopeZone<RECON>(y, a.theta, a.dX[d]);
    };
    double graduCov[4][4];
#pragma unroll
    for (int mu = 0; mu < 4; mu++) {
      const double* um = uT + mu * n;
      graduCov[0][mu] = 0.;
      graduCov[1][mu] = slopeOf(um, 0);
      graduCov[2][mu] = 0.;
      if (D.dim > 1) graduCov[2][mu] = slopeOf(um, 1);
      graduCov[3][mu] = 0.;
      if (D.dim > 2) graduCov[3][mu] = slopeOf(um, 2);
#pragma unroll
      for (int nu = 0; nu < 4; nu++)
#pragma unroll
        for (int lambda = 0; lambda < 4; lambda++) graduCov[nu][mu] -= gam(lambda, nu, mu) * kk.uCov[lambda];
    }
    double divuCov = 0.;
#pragma unroll
    for (int mu = 0; mu < 4; mu++)
#pragma unroll
      for (int nu = 0; nu < 4; nu++) divuCov += G.con(mu, nu) * graduCov[mu][nu];
    if constexpr (L::VISC_ON) {
      double deltaP0 = -divuCov * kk.rho * cc.nu;
#pragma unroll
      for (int mu = 0; mu < 4; mu++)
#pragma unroll
        for (int nu = 0; nu < 4; nu++)
          deltaP0 += 3. * kk.rho * cc.nu * kk.bCon[mu] * kk.bCon[nu] / kk.bSqr * graduCov[mu][nu];
      double deltaP0Tilde = deltaP0;
      if (a.prm.highOrderVisc == 1)
        deltaP0Tilde = deltaP0 * sqrt(cc.tau / kk.rho / cc.nu / kk.temperature);
      double sd = -g * (deltaP0Tilde) / cc.tau;
      if (a.prm.highOrderVisc == 1) sd -= 0.5 * g * divuCov * P[L::DP];
      src[L::DP] = sd;
    }
    if constexpr (L::COND_ON) {
      const double* Tp = uT + 4 * n;
      double gradT[4];
      gradT[0] = 0.;
      gradT[1] = slopeOf(Tp, 0);
      gradT[2] = 0.;
      if (D.dim > 1) gradT[2] = slopeOf(Tp, 1);
      gradT[3] = 0.;
      if (D.dim > 2) gradT[3] = slopeOf(Tp, 2);
      double q0 = 0.;
#pragma unroll
      for (int mu = 0; mu < 4; mu++) {
        q0 -= kk.rho * cc.chi * kk.bCon[mu] / cc.bNorm * gradT[mu];
#pragma unroll
        for (int nu = 0; nu < 4; nu++)
          q0 -= kk.rho * cc.chi * kk.temperature * kk.bCon[nu] / cc.bNorm * kk.uCon[mu] * graduCov[mu][nu];
      }
      double q0Tilde = q0;
      if (a.prm.highOrderCond == 1) q0Tilde = q0 * sqrt(cc.tau / kk.rho / cc.chi) / kk.temperature;
      double sq = -g * (q0Tilde) / cc.tau;
      if (a.prm.highOrderCond == 1) sq -= 0.5 * g * divuCov * P[L::Q];
      src[L::Q] = sq;
    }
    if (sImp) {
      // computeImplicitSources of elemOld with tauDamp = the stage element's tau
#pragma unroll
      for (int v = 0; v < L::NF; v++) sImp[v * n + z] = 0.;
      if constexpr (L::VISC_ON) sImp[L::DP * n + z] = g * (primOld[L::DP * n + z]) / cc.tau;
      if constexpr (L::COND_ON) sImp[L::Q * n + z] = g * (primOld[L::Q * n + z]) / cc.tau;
    }
  }
#pragma unroll
  for (int v = 0; v < L::NF; v++) sExp[v * n + z] = src[v];
}

// ---------------------------------------------------------------------------
// fluidElement's own outputs on a whole grid (physics.cpp:53-358)
// ---------------------------------------------------------------------------
enum { ELEM_FLUXES = 0, ELEM_SPEEDS = 1, ELEM_FIELDS = 2, ELEM_IMPLICIT = 3, ELEM_TIME_DERIV = 4 };

// Every zone z of the grid, the element set from prim at the geometry of one
// location (fluidElement::set, physics.cpp:53-183), then by `what`:
//  ELEM_FLUXES      computeFluxes(dir) (physics.cpp:185-217), dof planes of out0
//  ELEM_SPEEDS      computeMinMaxCharSpeeds(dir) (riemannsolver.cpp:54-200): out0, out1
//  ELEM_FIELDS      the element's arrays (GRIM_ELEMENT_FIELDS planes of out0)
//  ELEM_IMPLICIT    computeImplicitSources(sources, tauDamp = aux0) (physics.cpp:320-358)
//  ELEM_TIME_DERIV  computeTimeDerivSources(elemOld = aux0, elemNew = aux1, dt)
//                   (physics.cpp:219-318), this element the coefficient element
template <class L>
__global__ __launch_bounds__(kBlock) void elementKernel(StepArgs a, int what, int dir,
                                                        const double* __restrict__ prim,
                                                        const double* __restrict__ geom,
                                                        const double* __restrict__ aux0,
                                                        const double* __restrict__ aux1, double dt,
                                                        double* __restrict__ out0, double* __restrict__ out1) {
  const int64_t z = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t n = a.D.n;
  if (z >= n) return;
  double P[L::DOF];
#pragma unroll
  for (int v = 0; v < L::DOF; v++) P[v] = prim[v * n + z];
  GeomFull G;
  loadGeomFull(geom, a.D.ng, geomIndexOf(a.D, z), G);
  Kinematics kk;
  Closure cc;
  kinematics<L>(P, G.zg, a.prm, kk);
  elementClosure<L>(P, kk, a.prm, G.r, cc);
  const double g = G.zg.g;
  if (what == ELEM_FLUXES) {
    double F[L::DOF];
    if (dir == 0) fluxes<L, 0>(P, kk, cc, g, F);
    else if (dir == 1) fluxes<L, 1>(P, kk, cc, g, F);
    else if (dir == 2) fluxes<L, 2>(P, kk, cc, g, F);
    else fluxes<L, 3>(P, kk, cc, g, F);
#pragma unroll
    for (int v = 0; v < L::DOF; v++) out0[v * n + z] = F[v];
  } else if (what == ELEM_SPEEDS) {
    double mn, mx;
    if (dir == 0) charSpeeds<L, 0>(kk, cc, G, a.prm, mn, mx);
    else if (dir == 1) charSpeeds<L, 1>(kk, cc, G, a.prm, mn, mx);
    else charSpeeds<L, 2>(kk, cc, G, a.prm, mn, mx);
    out0[z] = mn;
    out1[z] = mx;
  } else if (what == ELEM_FIELDS) {
    const double f[GRIM_ELEMENT_FIELDS] = {
        kk.rho, kk.u, kk.pressure, kk.temperature, kk.gammaLF, cc.soundSpeed, kk.bSqr, cc.bNorm,
        cc.q, cc.deltaP, cc.tau, cc.chi, cc.nu,
        kk.uCon[0], kk.uCon[1], kk.uCon[2], kk.uCon[3], kk.uCov[0], kk.uCov[1], kk.uCov[2], kk.uCov[3],
        kk.bCon[0], kk.bCon[1], kk.bCon[2], kk.bCon[3], kk.bCov[0], kk.bCov[1], kk.bCov[2], kk.bCov[3]};
#pragma unroll
    for (int q = 0; q < GRIM_ELEMENT_FIELDS; q++) out0[q * n + z] = f[q];
  } else {
    double S[L::DOF];
#pragma unroll
    for (int v = 0; v < L::DOF; v++) S[v] = 0.;
    if constexpr (L::EMHD) {
      if (what == ELEM_IMPLICIT) {
        if constexpr (L::VISC_ON) S[L::DP] = g * (P[L::DP]) / aux0[z];
        if constexpr (L::COND_ON) S[L::Q] = g * (P[L::Q]) / aux0[z];
      } else {
        // elemOld and elemNew at the same geometry
        double Po[L::DOF], Pn[L::DOF];
#pragma unroll
        for (int v = 0; v < L::DOF; v++) {
          Po[v] = aux0[v * n + z];
          Pn[v] = aux1[v * n + z];
        }
        Kinematics ko, kn;
        kinematics<L>(Po, G.zg, a.prm, ko);
        kinematics<L>(Pn, G.zg, a.prm, kn);
        double dtuCov[4];
#pragma unroll
        for (int mu = 0; mu < 4; mu++) dtuCov[mu] = (kn.uCov[mu] - ko.uCov[mu]) / dt;
        double divuCov = 0.;
#pragma unroll
        for (int mu = 0; mu < 4; mu++) divuCov += G.zg.gcon0[mu] * dtuCov[mu];
        if constexpr (L::VISC_ON) {
          double dP0 = -divuCov * kk.rho * cc.nu;
#pragma unroll
          for (int mu = 0; mu < 4; mu++) dP0 += 3. * kk.rho * cc.nu * kk.bCon[0] * kk.bCon[mu] / kk.bSqr * dtuCov[mu];
          if (a.prm.highOrderVisc) dP0 *= sqrt(cc.tau / kk.rho / cc.nu / kk.temperature);
          S[L::DP] = -g * (dP0) / cc.tau;
          if (a.prm.highOrderVisc) S[L::DP] -= 0.5 * g * divuCov * P[L::DP];
        }
        if constexpr (L::COND_ON) {
          double q0 = -kk.rho * cc.chi * kk.bCon[0] / cc.bNorm * (kn.temperature - ko.temperature) / dt;
#pragma unroll
          for (int nu = 0; nu < 4; nu++)
            q0 -= kk.rho * cc.chi * kk.temperature * kk.bCon[nu] / cc.bNorm * kk.uCon[0] * dtuCov[nu];
          if (a.prm.highOrderCond) q0 *= sqrt(cc.tau / kk.rho / cc.chi) / kk.temperature;
          S[L::Q] = -g * (q0) / cc.tau;
          if (a.prm.highOrderCond) S[L::Q] -= 0.5 * g * divuCov * P[L::Q];
        }
      }
    }
#pragma unroll
    for (int v = 0; v < L::DOF; v++) out0[v * n + z] = S[v];
  }
}

// ---------------------------------------------------------------------------
// dt and stats
// ---------------------------------------------------------------------------
// computeDt's per-zone term (timestep.cpp:336-391): the sum over the
// active directions of max(|min speed|, max speed) / dX.
template <class L>
__device__ __forceinline__ double zoneInvDt(const StepArgs& a, const double* P, const GeomFull& G) {
  Kinematics kk;
  Closure cc;
  kinematics<L>(P, G.zg, a.prm, kk);
  elementClosure<L>(P, kk, a.prm, G.r, cc);
  double maxSpeed = 0.;
  for (int d = 0; d < a.D.dim; d++) {
    double mn, mx;
    if (d == 0) charSpeeds<L, 0>(kk, cc, G, a.prm, mn, mx);
    else if (d == 1) charSpeeds<L, 1>(kk, cc, G, a.prm, mn, mx);
    else charSpeeds<L, 2>(kk, cc, G, a.prm, mn, mx);
    mn = mn / a.dX[d];
    mx = mx / a.dX[d];
    const double s = stdmax(mx, fabs(mn));
    maxSpeed = (d == 0) ? s : maxSpeed + s;
  }
  return maxSpeed;
}

// The workgroup's maximum of m (a fixed tree, so the result does not depend
// on timing) into blockMax[blockIdx.x].
__device__ __forceinline__ void blockMaxStore(double* red, double m, double* __restrict__ blockMax) {
  red[threadIdx.x] = m;
  __syncthreads();
  for (int w = kBlock / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] = stdmax(red[threadIdx.x], red[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) blockMax[blockIdx.x] = red[0];
}

template <class L>
__global__ __launch_bounds__(kBlock) void dtKernel(StepArgs a, const double* __restrict__ prim,
                                                   const double* __restrict__ geomC,
                                                   double* __restrict__ blockMax) {
  __shared__ double red[kBlock];
  const int64_t z = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t n = a.D.n;
  double m = -INFINITY;
  if (z < n) {
    double P[L::DOF];
#pragma unroll
    for (int v = 0; v < L::DOF; v++) P[v] = prim[v * n + z];
    GeomFull G;
    loadGeomFull(geomC, a.D.ng, geomIndexOf(a.D, z), G);
    m = zoneInvDt<L>(a, P, G);
  }
  blockMaxStore(red, m, blockMax);
}

// The step's clock lives in device memory so that a whole step can be
// enqueued (or captured in a HIP graph) without a host round trip:
// clock[0] = time, clock[1] = dt, clock[2] = dt/2 (the half stage's dt').
enum { CLOCK_TIME = 0, CLOCK_DT = 1, CLOCK_HALF_DT = 2, CLOCK_WORDS = 4 };

__global__ __launch_bounds__(kBlock) void maxReduceKernel(int64_t count, const double* __restrict__ in,
                                                          double* __restrict__ out,
                                                          double* __restrict__ clock = nullptr,
                                                          double courant = 0., double maxIncrement = 0.) {
  __shared__ double red[kBlock];
  double m = -INFINITY;
  for (int64_t t = threadIdx.x; t < count; t += kBlock) m = stdmax(m, in[t]);
  red[threadIdx.x] = m;
  __syncthreads();
  for (int w = kBlock / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] = stdmax(red[threadIdx.x], red[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = red[0];
    // with a clock: the end of timeStepper::computeDt (timestep.cpp:393-399)
    // from the maximum just reduced
    if (clock) {
      double newDt = courant / red[0];
      if (newDt > maxIncrement * clock[CLOCK_DT]) newDt = maxIncrement * clock[CLOCK_DT];
      clock[CLOCK_DT] = newDt;
      clock[CLOCK_HALF_DT] = newDt / 2.;
    }
  }
}

__global__ void setClockKernel(double* __restrict__ clock, double time, double dt) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  clock[CLOCK_TIME] = time;
  clock[CLOCK_DT] = dt;
  clock[CLOCK_HALF_DT] = dt / 2.;
}

// Clears counters with a kernel (a memset node in a captured graph is not
// safe to replay).
__global__ void zeroCountersKernel(unsigned long long* __restrict__ c, int count) {
  if (blockIdx.x == 0 && (int)threadIdx.x < count) c[threadIdx.x] = 0ull;
}

// Bulk-zone Newton statistics: count above atol and the largest iteration
// count. Each wavefront reduces in registers and issues one vector-memory
// atomic per counter.
__global__ __launch_bounds__(kBlock) void stageStatsKernel(int64_t n, const double* __restrict__ mask,
                                                           const double* __restrict__ l2,
                                                           const int* __restrict__ iters, double atol,
                                                           unsigned long long* __restrict__ out) {
  // grid-stride over the zones (a fixed, small grid: same-address atomics
  // serialise, so each workgroup issues only one per counter)
  unsigned long long count = 0, bad = 0;
  int it = 0;
  for (int64_t z = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; z < n; z += (int64_t)gridDim.x * blockDim.x) {
    if (mask[z] != 0.) {
      count += l2[z] > atol ? 1ull : 0ull;
      bad += isfinite(l2[z]) ? 0ull : 1ull;
      it = max(it, iters[z]);
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    count += __shfl_xor(count, off, 64);
    bad += __shfl_xor(bad, off, 64);
    it = max(it, __shfl_xor(it, off, 64));
  }
  __shared__ unsigned long long sc[kBlock / 64], sb[kBlock / 64];
  __shared__ int si[kBlock / 64];
  if ((threadIdx.x & 63) == 0) {
    sc[threadIdx.x / 64] = count;
    sb[threadIdx.x / 64] = bad;
    si[threadIdx.x / 64] = it;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < kBlock / 64; w++) {
      count += sc[w];
      bad += sb[w];
      it = max(it, si[w]);
    }
    if (count) atomicAdd(&out[0], count);
    if (it > 0) atomicMax(&out[1], (unsigned long long)it);
    if (bad) atomicAdd(&out[2], bad);
  }
}
constexpr unsigned kStatsBlocks = 512;

// ---------------------------------------------------------------------------
// Linear-mode initial conditions (src/problem/linear_modes/problem.cpp:10-111)
// ---------------------------------------------------------------------------
// Ideal fast magnetosonic eigenvector on rho = 1, u = 2, B = (0.1, 0.3, 0),
// Gamma = 4/3, k = (2 pi, 2 pi), unit norm; derived from the linearised
// equations (tests/golden/fast_mode.json holds the derivation's output).
__constant__ double kFastMode[8] = {0.3454194280099405,  0.9211184746935905,  0.12567391680338838,
                                    0.11701881004980452, 0.,                  -0.03700566889854888,
                                    0.037005668898594656, 0.};

template <class L>
__global__ __launch_bounds__(kBlock) void linearModeKernel(StepArgs a, int mode, double amp,
                                                           const double* __restrict__ geomC,
                                                           CoordGrid cg, double* __restrict__ prim) {
  const int64_t z = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const Dims& D = a.D;
  const int64_t n = D.n;
  if (z >= n) return;
  int i, j, k;
  zoneIndex(D, z, i, j, k);
  double X[3];
  XCoordsAt(D, cg, i, j, k, 0, X);  // Minkowski: xCoords = XCoords
  double P[L::DOF];
#pragma unroll
  for (int v = 0; v < L::DOF; v++) P[v] = 0.;
  if (mode == 0) {
    const double cphi = cos(2. * M_PI * X[0] + 2. * M_PI * X[1]);
    const double bg[8] = {1., 2., 0., 0., 0., 0.1, 0.3, 0.};
#pragma unroll
    for (int v = 0; v < 5; v++) P[v] = bg[v] + amp * kFastMode[v] * cphi;
#pragma unroll
    for (int b = 0; b < 3; b++) P[L::B1 + b] = bg[5 + b] + amp * kFastMode[5 + b] * cphi;
  } else {
    const double Aw = amp;
    const double k1 = 2. * M_PI, k2 = 4. * M_PI;
    const double cphi = cos(k1 * X[0] + k2 * X[1]);
    const double sphi = sin(k1 * X[0] + k2 * X[1]);
    P[RHO] = 1.;
    P[UU] = 2.;
    P[L::B1] = 0.1;
    P[L::B2] = 0.3;
    P[RHO] += Aw * cphi * (-0.518522524082246) + Aw * sphi * 0.1792647678001878;
    P[UU] += Aw * cphi * 0.5516170736393813;
    P[U1] += Aw * cphi * 0.008463122479547856 + Aw * sphi * (-0.011862022608466367);
    P[U2] += Aw * cphi * (-0.16175466371870734) + Aw * sphi * (0.034828080823603294);
    P[L::B1] += Aw * cphi * (-0.05973794979640743) + Aw * sphi * 0.03351707506150924;
    P[L::B2] += Aw * cphi * 0.02986897489820372 - Aw * sphi * 0.016758537530754618;
    if constexpr (L::EMHD) {
      double Q = Aw * cphi * 0.5233486841539436 - Aw * sphi * 0.04767672501939603;
      double DP = Aw * cphi * 0.2909106062057657 - Aw * sphi * 0.02159452055336572;
      const double gam = a.prm.gam;
      const double pressure = (gam - 1.) * P[UU];
      const double temperature = pressure / P[RHO];
      const double soundSpeed = sqrt(gam * pressure / (P[RHO] + gam * P[UU]));
      if (a.prm.highOrderCond)
        Q = Q / temperature / sqrt(P[RHO] * a.prm.condAlpha * soundSpeed * soundSpeed);
      if (a.prm.highOrderVisc)
        DP = DP / sqrt(temperature * P[RHO] * a.prm.viscAlpha * soundSpeed * soundSpeed);
      if constexpr (L::COND_ON) P[L::Q] = Q;
      if constexpr (L::VISC_ON) P[L::DP] = DP;
    }
  }
#pragma unroll
  for (int v = 0; v < L::DOF; v++) prim[v * n + z] = P[v];
}

// ---------------------------------------------------------------------------
// Torus problem (src/problem/torus/problem.cpp)
// ---------------------------------------------------------------------------
__device__ __forceinline__ void zoneIJK(const Dims& D, int64_t z, int& i, int& j, int& k) {
  i = (int)(z % D.NT[0]);
  j = (int)((z / D.NT[0]) % D.NT[1]);
  k = (int)(z / ((int64_t)D.NT[0] * D.NT[1]));
}

// Floor values pow(r, slope) * ampl of every zone (applyFloor, :548-549).
__global__ __launch_bounds__(kBlock) void floorPlanesKernel(Dims D, grim_torus_params t,
                                                            const double* __restrict__ geomC,
                                                            double* __restrict__ floors) {
  const int64_t z = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (z >= D.n) return;
  const double Radius = geomC[LOC_R * D.ng + geomIndexOf(D, z)];
  floors[z] = pow(Radius, t.RhoFloorSlope) * t.RhoFloorAmpl;
  floors[D.n + z] = pow(Radius, t.UFloorSlope) * t.UFloorAmpl;
}

// The fluid part of the initial conditions at every zone (:175-322).
template <class L>
__global__ __launch_bounds__(kBlock) void torusICKernel(StepArgs a, grim_torus_params t, double aBH,
                                                        double hSlope, CoordGrid cg,
                                                        const double* __restrict__ geomC,
                                                        double* __restrict__ prim) {
  const int64_t z = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const Dims& D = a.D;
  const int64_t n = D.n;
  if (z >= n) return;
  int i, j, k;
  zoneIJK(D, z, i, j, k);
  const double X2 = cg.start[1] + ((double)j - D.g[1] + 0.5) * cg.dX[1];
  const int64_t ng = D.ng, zg = geomIndex(D, i, j);
  const double beta[4] = {0., geomC[11 * ng + zg], geomC[12 * ng + zg], geomC[13 * ng + zg]};
  double P[L::DOF];
#pragma unroll
  for (int v = 0; v < L::DOF; v++) P[v] = 0.;
  torusZone(t, aBH, hSlope, a.prm.gam, a.prm.rhoFloor, a.prm.uFloor, geomC[LOC_R * ng + zg],
            geomC[LOC_THETA * ng + zg], X2, geomC[GEOM_ALPHA * ng + zg], beta, uniform01(t.seed, (uint64_t)z, 0),
            P);
#pragma unroll
  for (int v = 0; v < L::DOF; v++) prim[v * n + z] = P[v];
}

// Per-block maximum of sign * plane (a plane minimum with sign = -1).
__global__ __launch_bounds__(kBlock) void planeMaxKernel(int64_t n, const double* __restrict__ in, double sign,
                                                         double* __restrict__ blockMax) {
  __shared__ double red[kBlock];
  const int64_t z = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  red[threadIdx.x] = z < n ? sign * in[z] : -INFINITY;
  __syncthreads();
  for (int w = kBlock / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] = stdmax(red[threadIdx.x], red[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) blockMax[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(kBlock) void normalizeKernel(int64_t n, const double* __restrict__ rhoMax,
                                                          double* __restrict__ prim) {
  const int64_t z = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (z >= n) return;
  const double m = rhoMax[0];
  prim[RHO * n + z] = prim[RHO * n + z] / m;
  prim[UU * n + z] = prim[UU * n + z] / m;
}

// Vector potential from the neighbour-averaged density (af::shift wraps
// around the grid), :355-380.
__global__ __launch_bounds__(kBlock) void vectorPotentialKernel(Dims D, grim_torus_params t,
                                                                const double* __restrict__ geomC,
                                                                const double* __restrict__ rho,
                                                                double* __restrict__ Avec) {
  const int64_t z = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (z >= D.n) return;
  int I[3];
  zoneIJK(D, z, I[0], I[1], I[2]);
  auto at = [&](int d, int off) {
    int J[3] = {I[0], I[1], I[2]};
    J[d] = (J[d] + off + D.NT[d]) % D.NT[d];
    return rho[flatIndex(D, J[0], J[1], J[2])];
  };
  double rhoAvg = at(0, -1) + at(0, 1) + at(1, -1) + at(1, 1);
  if (D.dim > 2)
    rhoAvg = (rhoAvg + at(2, -1) + at(2, 1)) / 6.;
  else
    rhoAvg = rhoAvg / 4.;
  const double zero = rhoAvg * 0.;
  const int64_t zg = geomIndex(D, I[0], I[1]);
  if (t.UseMADdisk)
    Avec[z] = stdmax(zero, rhoAvg * rhoAvg * pow(geomC[LOC_R * D.ng + zg], 4));
  else
    Avec[z] = stdmax(rhoAvg - 0.2, zero) * cos(geomC[LOC_THETA * D.ng + zg] * (t.MagneticLoops - 1));
}

// B = curl A in the bulk, zero in ghost zones (:382-404).
template <class L>
__global__ __launch_bounds__(kBlock) void curlKernel(Dims D, double dX1, double dX2,
                                                     const double* __restrict__ geomC,
                                                     const double* __restrict__ Avec, double* __restrict__ prim) {
  const int64_t z = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t n = D.n;
  if (z >= n) return;
  int I[3];
  zoneIJK(D, z, I[0], I[1], I[2]);
  auto A = [&](int di, int dj) {
    const int ii = (I[0] + di) % D.NT[0], jj = (I[1] + dj) % D.NT[1];
    return Avec[flatIndex(D, ii, jj, I[2])];
  };
  const double g = geomC[GEOM_G * D.ng + geomIndex(D, I[0], I[1])];
  const double A00 = A(0, 0), A01 = A(0, 1), A10 = A(1, 0), A11 = A(1, 1);
  const double B1 = (A01 - A00 + A11 - A10) / (2. * dX2 * g);
  const double B2 = (A00 - A10 + A01 - A11) / (2. * dX1 * g);
  const bool bulk = I[0] >= D.g[0] && I[0] < D.g[0] + D.N[0] && I[1] >= D.g[1] && I[1] < D.g[1] + D.N[1] &&
                    I[2] >= D.g[2] && I[2] < D.g[2] + D.N[2];
  const double mask = bulk ? 1. : 0.;
  prim[L::B1 * n + z] = B1 * mask;
  prim[L::B2 * n + z] = B2 * mask;
  prim[L::B3 * n + z] = 0.;
}

// Plasma-beta reductions for the field normalisation (:406-504): SANE needs
// min beta = -max(-beta), MAD max Pmag and max Pgas.
template <class L>
__global__ __launch_bounds__(kBlock) void betaKernel(StepArgs a, int mad, const double* __restrict__ prim,
                                                     const double* __restrict__ geomC,
                                                     double* __restrict__ blockA, double* __restrict__ blockB) {
  __shared__ double ra[kBlock], rb[kBlock];
  const int64_t z = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t n = a.D.n;
  double va = -INFINITY, vb = -INFINITY;
  if (z < n) {
    double P[L::DOF];
#pragma unroll
    for (int v = 0; v < L::DOF; v++) P[v] = prim[v * n + z];
    ZoneGeom G;
    loadGeom(geomC, a.D.ng, geomIndexOf(a.D, z), G);
    Kinematics k;
    kinematics<L>(P, G, a.prm, k);
    if (mad) {
      va = k.bSqr / 2.;
      vb = k.pressure;
    } else {
      va = -(2. * (k.pressure + 1.e-13) / (k.bSqr + 1.e-18));
    }
  }
  ra[threadIdx.x] = va;
  rb[threadIdx.x] = vb;
  __syncthreads();
  for (int w = kBlock / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      ra[threadIdx.x] = stdmax(ra[threadIdx.x], ra[threadIdx.x + w]);
      rb[threadIdx.x] = stdmax(rb[threadIdx.x], rb[threadIdx.x + w]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    blockA[blockIdx.x] = ra[0];
    blockB[blockIdx.x] = rb[0];
  }
}

template <class L>
__global__ __launch_bounds__(kBlock) void scaleFieldKernel(int64_t n, grim_torus_params t,
                                                           const double* __restrict__ redA,
                                                           const double* __restrict__ redB,
                                                           double* __restrict__ prim) {
  const int64_t z = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (z >= n) return;
  const double BFactor = t.UseMADdisk ? sqrt(redB[0] / redA[0] / t.MinPlasmaBeta)
                                      : sqrt(-redA[0] / t.MinPlasmaBeta);
#pragma unroll
  for (int b = 0; b < 3; b++) prim[(L::B1 + b) * n + z] *= BFactor;
}

template <class L>
__global__ __launch_bounds__(kBlock) void floorKernel(StepArgs a, grim_torus_params t,
                                                      const double* __restrict__ geomC,
                                                      const double* __restrict__ floors,
                                                      double* __restrict__ prim) {
  const int64_t z = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t n = a.D.n;
  if (z >= n) return;
  double P[L::DOF];
#pragma unroll
  for (int v = 0; v < L::DOF; v++) P[v] = prim[v * n + z];
  GeomFull G;
  loadGeomFull(geomC, a.D.ng, geomIndexOf(a.D, z), G);
  applyFloorZone<L>(a.prm, t, G, floors[z], floors[n + z], P);
#pragma unroll
  for (int v = 0; v < L::DOF; v++) prim[v * n + z] = P[v];
}

// fullStepDiagnostics' NaN scrub (:722-741): count NaNs in the bulk, then
// zero every zone holding one when the count is positive.
__global__ __launch_bounds__(kBlock) void nanCountKernel(Dims D, Box box, int nvars, const double* __restrict__ prim,
                                                         unsigned long long* __restrict__ count) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= box.count()) return;
  int i, j, k;
  box.at(t, i, j, k);
  const int64_t z = flatIndex(D, i, j, k);
  unsigned long long c = 0;
  for (int v = 0; v < nvars; v++) c += isnan(prim[v * D.n + z]) ? 1ull : 0ull;
  if (c) atomicAdd(count, c);
}

__global__ __launch_bounds__(kBlock) void nanZeroKernel(int64_t n, int nvars, const unsigned long long* __restrict__ count,
                                                        double* __restrict__ prim) {
  const int64_t z = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (z >= n || count[0] == 0) return;
  bool bad = false;
  for (int v = 0; v < nvars; v++) bad = bad || isnan(prim[v * n + z]);
  if (bad)
    for (int v = 0; v < nvars; v++) prim[v * n + z] = 0.;
}

// The end of grim_time_step in one pass over every zone (workgroups and
// zones mapped as in nanZeroKernel, floorKernel and dtKernel): the torus's
// full-stage NaN scrub (when the bulk count is positive) and applyFloor
// (fullStepDiagnostics, torus/problem.cpp:719-746) written back into prim,
// the result into primOld and primDt (timestep.cpp:290-300), time += dt, and
// the next computeDt's per-workgroup maxima of the inverse time step of the
// result (timestep.cpp:336-391), which the next step's reduction reads.
// Each value is the separate kernels' bit for bit.
template <class L, bool TORUS>
__global__ __launch_bounds__(kBlock) void endStepKernel(StepArgs a, grim_torus_params t,
                                                        const double* __restrict__ geomC,
                                                        const double* __restrict__ floors,
                                                        const unsigned long long* __restrict__ nanCount,
                                                        double* __restrict__ prim, double* __restrict__ primOld,
                                                        double* __restrict__ primDt, double* __restrict__ clock,
                                                        double* __restrict__ blockMax) {
  __shared__ double red[kBlock];
  const int64_t z = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t n = a.D.n;
  if (z == 0) clock[CLOCK_TIME] += clock[CLOCK_DT];
  double m = -INFINITY;
  if (z < n) {
    double P[L::DOF];
#pragma unroll
    for (int v = 0; v < L::DOF; v++) P[v] = prim[v * n + z];
    GeomFull G;
    loadGeomFull(geomC, a.D.ng, geomIndexOf(a.D, z), G);
    if constexpr (TORUS) {
      if (nanCount[0] != 0) {
        bool bad = false;
#pragma unroll
        for (int v = 0; v < L::DOF; v++) bad = bad || isnan(P[v]);
        if (bad)
#pragma unroll
          for (int v = 0; v < L::DOF; v++) P[v] = 0.;
      }
      applyFloorZone<L>(a.prm, t, G, floors[z], floors[n + z], P);
    }
#pragma unroll
    for (int v = 0; v < L::DOF; v++) {
      if constexpr (TORUS) prim[v * n + z] = P[v];
      primOld[v * n + z] = P[v];
      primDt[v * n + z] = P[v];
    }
    m = zoneInvDt<L>(a, P, G);
  }
  blockMaxStore(red, m, blockMax);
}

// On-the-fly observers of fullStepDiagnostics (:757-776), defined in
// src/problem/torus/ObserversTorus.hpp, of the element set at prim:
//  ComputeEnergyIntegrals (:6-81)   sums over the bulk of rho, b^2 and u
//                                   times volElem*g*gamma/alpha
//  ComputeMinMaxVariables (:83-139) max rho, min 2(p+1e-13)/(b^2+1e-18)
//  ComputeBoundaryFluxes (:141-251) sums of rho*volElem*g*u^1 over the first
//                                   and last bulk X1 planes; on the last, the
//                                   unbound part ((1+u/rho*Gamma)u_0 < -1) and
//                                   its relativistic part (gamma > 2)
// Each workgroup strides over a fixed share of the bulk and stores its
// partial sums, max and min; obsFinishKernel combines the partials in a
// fixed order, so repeated calls give identical results.
enum {
  OBS_MASS,
  OBS_EMAG,
  OBS_ETH,
  OBS_MDOT_IN,
  OBS_MDOT_OUT,
  OBS_MDOT_UNBOUND,
  OBS_MDOT_REL,
  OBS_SUMS,
  OBS_RHO_MAX = OBS_SUMS,
  OBS_BETA_MIN,
  OBS_WORDS
};
constexpr unsigned kObsBlocks = 2048;

// word q of two partials combined: sums add, the maximum and minimum as such
__device__ __forceinline__ double obsCombine(int q, double a, double b) {
  return q < OBS_SUMS ? a + b : (q == OBS_RHO_MAX ? fmax(a, b) : fmin(a, b));
}

template <class L>
__global__ __launch_bounds__(kBlock) void observersKernel(StepArgs a, Box box, double volElem,
                                                          const double* __restrict__ prim,
                                                          const double* __restrict__ geomC,
                                                          double* __restrict__ partial) {
  const int64_t n = a.D.n;
  const int iIn = box.lo[0], iOut = box.hi[0] - 1;
  double s[OBS_SUMS];
#pragma unroll
  for (int q = 0; q < OBS_SUMS; q++) s[q] = 0.;
  double rhoMax = -INFINITY, betaMin = INFINITY;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < box.count();
       t += (int64_t)gridDim.x * blockDim.x) {
    int i, j, k;
    box.at(t, i, j, k);
    const int64_t z = flatIndex(a.D, i, j, k);
    double P[L::DOF];
#pragma unroll
    for (int v = 0; v < L::DOF; v++) P[v] = prim[v * n + z];
    ZoneGeom G;
    loadGeom(geomC, a.D.ng, geomIndex(a.D, i, j), G);
    Kinematics e;
    kinematics<L>(P, G, a.prm, e);
    const double rho = P[RHO], u = P[UU];
    s[OBS_MASS] += rho * volElem * G.g * e.gammaLF / G.alpha;
    s[OBS_EMAG] += volElem * e.bSqr * G.g * e.gammaLF / G.alpha;
    s[OBS_ETH] += volElem * u * G.g * e.gammaLF / G.alpha;
    rhoMax = fmax(rhoMax, rho);
    betaMin = fmin(betaMin, 2. * (e.pressure + 1.e-13) / (e.bSqr + 1.e-18));
    if (i == iIn || i == iOut) {
      const double mdot = rho * volElem * G.g * e.uCon[1];
      if (i == iIn) s[OBS_MDOT_IN] += mdot;
      if (i == iOut) {
        const double bernoulli = (1. + u / rho * a.prm.gam) * e.uCov[0];
        const double unbound = mdot * (bernoulli < -1. ? 1. : 0.);
        s[OBS_MDOT_OUT] += mdot;
        s[OBS_MDOT_UNBOUND] += unbound;
        s[OBS_MDOT_REL] += unbound * (e.gammaLF > 2. ? 1. : 0.);
      }
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
    for (int q = 0; q < OBS_SUMS; q++) s[q] += __shfl_xor(s[q], off, 64);
    rhoMax = fmax(rhoMax, __shfl_xor(rhoMax, off, 64));
    betaMin = fmin(betaMin, __shfl_xor(betaMin, off, 64));
  }
  __shared__ double sh[kBlock / 64][OBS_WORDS];
  const int w = threadIdx.x / 64;
  if ((threadIdx.x & 63) == 0) {
#pragma unroll
    for (int q = 0; q < OBS_SUMS; q++) sh[w][q] = s[q];
    sh[w][OBS_RHO_MAX] = rhoMax;
    sh[w][OBS_BETA_MIN] = betaMin;
  }
  __syncthreads();
  if (threadIdx.x < OBS_WORDS) {
    const int q = threadIdx.x;
    double r = sh[0][q];
    for (int v = 1; v < kBlock / 64; v++)
      r = obsCombine(q, r, sh[v][q]);
    partial[(int64_t)blockIdx.x * OBS_WORDS + q] = r;
  }
}

// One workgroup combines the kObsBlocks partials of every word in a fixed
// order (a strided pass with all words in flight, then a tree in LDS),
// writing out[q].
__global__ __launch_bounds__(kBlock) void obsFinishKernel(const double* __restrict__ partial,
                                                          double* __restrict__ out) {
  static_assert(kObsBlocks % kBlock == 0, "every thread starts from one partial");
  __shared__ double sh[OBS_WORDS][kBlock];
  const int t = threadIdx.x;
  double r[OBS_WORDS];
#pragma unroll
  for (int q = 0; q < OBS_WORDS; q++) r[q] = partial[(int64_t)t * OBS_WORDS + q];
  for (unsigned b = t + kBlock; b < kObsBlocks; b += kBlock) {
#pragma unroll
    for (int q = 0; q < OBS_WORDS; q++) r[q] = obsCombine(q, r[q], partial[(int64_t)b * OBS_WORDS + q]);
  }
#pragma unroll
  for (int q = 0; q < OBS_WORDS; q++) sh[q][t] = r[q];
  __syncthreads();
  for (int w = kBlock / 2; w > 0; w >>= 1) {
    if (t < w) {
#pragma unroll
      for (int q = 0; q < OBS_WORDS; q++) sh[q][t] = obsCombine(q, sh[q][t], sh[q][t + w]);
    }
    __syncthreads();
  }
  if (t < OBS_WORDS) out[t] = sh[t][0];
}

// inflowCheck (:941-1044) on the radial ghost slabs.
template <class L>
__global__ __launch_bounds__(kBlock) void inflowCheckKernel(StepArgs a, grim_torus_params t, Box box, int side,
                                                            const double* __restrict__ geomC,
                                                            double* __restrict__ prim) {
  const int64_t tt = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (tt >= box.count()) return;
  int i, j, k;
  box.at(tt, i, j, k);
  const Dims& D = a.D;
  const int64_t n = D.n, z = flatIndex(D, i, j, k);
  double P[L::DOF];
#pragma unroll
  for (int v = 0; v < L::DOF; v++) P[v] = prim[v * n + z];
  const int64_t ng = D.ng;
  ZoneGeom G;
  loadGeom(geomC, ng, geomIndex(D, i, j), G);
  Kinematics kk;
  kinematics<L>(P, G, a.prm, kk);
  double u[3] = {P[U1] / kk.gammaLF, P[U2] / kk.gammaLF, P[U3] / kk.gammaLF};
  if (side == 1) {
    const int64_t zl = geomIndex(D, D.N[0] + D.g[0] - 1, j);
    const double minUr = geomC[(GEOM_GCON0 + 1) * ng + zl] * geomC[GEOM_ALPHA * ng + zl];
    u[0] = stdmax(minUr, u[0]);
  } else {
    u[0] = stdmin(G.gcon0[1] * G.alpha, u[0]);
  }
  double vSqr = u[0] * 0.;
#pragma unroll
  for (int p = 0; p < 3; p++)
#pragma unroll
    for (int q = 0; q < 3; q++) vSqr += G.cov(p + 1, q + 1) * u[p] * u[q];
  const double vSqrMax = 1. - 1. / t.MaxLorentzFactor / t.MaxLorentzFactor;
  const double vSqrMin = 1.e-13;
  vSqr = stdmax(stdmin(vSqr, vSqrMax), vSqrMin);
  const double newLorentzFactor = 1. / sqrt(1. - vSqr);
#pragma unroll
  for (int p = 0; p < 3; p++) prim[(U1 + p) * n + z] = u[p] * newLorentzFactor;
}

// fixPoles (:1046-1215): one thread per (i, k) column and pole; the two
// active rows next to the pole first, then the mirrored ghost rows.
template <class L>
__global__ __launch_bounds__(kBlock) void fixPolesKernel(Dims D, const double* __restrict__ geomC,
                                                         double* __restrict__ prim) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t cols = (int64_t)D.NT[0] * D.NT[2];
  if (t >= 2 * cols) return;
  const int pole = (int)(t / cols);
  const int64_t c = t % cols;
  const int i = (int)(c % D.NT[0]), k = (int)(c / D.NT[0]);
  const int64_t n = D.n;
  const int g2 = D.g[1], N2 = D.N[1];
  const int idx0 = pole == 0 ? g2 : N2 + g2 - 1;
  const int idx1 = pole == 0 ? g2 + 1 : idx0 - 1;
  const int idx2 = pole == 0 ? g2 + 2 : idx0 - 2;
  auto at = [&](int v, int j) -> double& { return prim[v * n + flatIndex(D, i, j, k)]; };
  auto Theta = [&](int j) {
    const double th = geomC[LOC_THETA * D.ng + geomIndex(D, i, j)];
    return pole == 0 ? th : th * (-1.) + M_PI;
  };
  const int rows[2] = {idx0, idx1};
  for (int r = 0; r < 2; r++) {
    const int jj = rows[r];
    at(RHO, jj) = at(RHO, idx2);
    at(UU, jj) = at(UU, idx2);
    at(U1, jj) = at(U1, idx2);
    at(U2, jj) = at(U2, idx2) * Theta(jj) / Theta(idx2);
    at(U3, jj) = at(U3, idx2);
    if constexpr (L::COND_ON) at(L::Q, jj) = at(L::Q, idx2) * Theta(jj) / Theta(idx2);
    if constexpr (L::VISC_ON) at(L::DP, jj) = at(L::DP, idx2) * Theta(jj) / Theta(idx2);
  }
  for (int q = 0; q < g2; q++) {
    const int dst = pole == 0 ? g2 - 1 - q : N2 + g2 + q;
    const int src = pole == 0 ? g2 + q : N2 + g2 - 1 - q;
    at(RHO, dst) = at(RHO, src);
    at(UU, dst) = at(UU, src);
    at(U1, dst) = at(U1, src);
    at(U3, dst) = at(U3, src);
    at(U2, dst) = at(U2, src) * (-1.0);
    at(L::B1, dst) = at(L::B1, src);
    at(L::B3, dst) = at(L::B3, src);
    at(L::B2, dst) = at(L::B2, src) * (-1.0);
    if constexpr (L::COND_ON) at(L::Q, dst) = at(L::Q, src) * (-1.0);
    if constexpr (L::VISC_ON) at(L::DP, dst) = at(L::DP, src) * (-1.0);
  }
}

}  // namespace
}  // namespace grim

using namespace grim;

// ---------------------------------------------------------------------------
// Context
// ---------------------------------------------------------------------------
struct grim_ctx {
  grim_params p;
  grim_grid_config c;
  grim_newton_opts opts;
  hipStream_t s;
  Dims D;
  StepArgs a;
  CoordGrid cg;
  CoordMap cm;
  int dof, nf;
  bool clockSet;  // grim_set_time was called
  double* clock;  // device: time, dt, dt/2 (CLOCK_*)
  double *primOld, *primHalf, *primNew, *primDt, *consOld;
  double *fluxes, *emf, *sExp, *sImp, *uT, *rhs, *mask;
  double *geomC, *geomX1, *geomX2, *gamma;
  double* gamma40;  // explicitFast: the 40 symmetric connection planes
  double* red;
  // computeDt's per-workgroup maxima of primDt (blocksFor(n) + 1 doubles):
  // written by grim_set_time and by the end of every step, reduced at the
  // start of the next
  double* dtPart;
  double* obs;  // observers: per-workgroup partials, then the result
  int* iters;
  double* l2;
  unsigned long long* counters;
  void* newtonWs;   // compacting Newton work list (grim_newton_workspace_bytes)
  double* hostBuf;  // pinned: dt / stats readback
  // torus problem hooks (grim_init_torus)
  int torus;
  grim_torus_params tp;
  double* floors;  // 2 planes: rho and u floor values
  double* dump;    // dump variables staged for a host copy (first such call)
  double* ghostState;  // dof planes whose ghost zones the half stage imposes, or null
  double* sideState;   // rho, u, u~^1 planes the X1 sides take in both stages, or null
};

namespace grim {
namespace {

template <class F>
int withLayout(const grim_ctx* x, F&& f) {
  return dispatchLayout(x->p, std::forward<F>(f));
}

Box bulkBox(const Dims& D) {
  Box b;
  for (int d = 0; d < 3; d++) {
    b.lo[d] = D.g[d];
    b.hi[d] = D.g[d] + D.N[d];
  }
  return b;
}

// Faces whose 6-point stencil lies in the grid (the oracle's face range).
Box faceBox(const Dims& D, int dir) {
  Box b;
  for (int d = 0; d < 3; d++) {
    b.lo[d] = 0;
    b.hi[d] = D.NT[d];
  }
  b.lo[dir] = 3;
  b.hi[dir] = D.NT[dir] - 2;
  return b;
}

// Edges read by the bulk divergence: [g, N + g] in each active dimension.
Box edgeBox(const Dims& D) {
  Box b;
  for (int d = 0; d < 3; d++) {
    b.lo[d] = D.g[d];
    b.hi[d] = D.g[d] + D.N[d] + (d < D.dim ? 1 : 0);
  }
  return b;
}

int applyBoundaries(grim_ctx* x, double* prim, int nvars, bool periodicFill) {
  const Dims& D = x->D;
  BcTypes bc;
  const int lo[3] = {x->c.boundaryLeft, x->c.boundaryBottom, x->c.boundaryBack};
  const int hi[3] = {x->c.boundaryRight, x->c.boundaryTop, x->c.boundaryFront};
  for (int d = 0; d < 3; d++) {
    bc.lo[d] = lo[d];
    bc.hi[d] = hi[d];
  }
  if (periodicFill) communicateKernel<<<blocksFor(D.n), kBlock, 0, x->s>>>(D, bc, nvars, prim);
  const int types[6] = {lo[0], hi[0], lo[1], hi[1], lo[2], hi[2]};
  for (int d = 0; d < D.dim; d++)
    for (int side = 0; side < 2; side++) {
      const int type = types[2 * d + side];
      if (type != GRIM_BC_OUTFLOW && type != GRIM_BC_MIRROR) continue;
      Box b;
      for (int e = 0; e < 3; e++) {
        b.lo[e] = 0;
        b.hi[e] = D.NT[e];
      }
      b.lo[d] = side == 0 ? 0 : D.N[d] + D.g[d];
      b.hi[d] = side == 0 ? D.g[d] : D.N[d] + 2 * D.g[d];
      boundaryKernel<<<blocksFor(b.count()), kBlock, 0, x->s>>>(D, b, d, side, type, nvars, prim);
    }
  return checkLaunch("boundaries");
}

int communicate(grim_ctx* x, double* prim, int nvars) {
  BcTypes bc;
  const int lo[3] = {x->c.boundaryLeft, x->c.boundaryBottom, x->c.boundaryBack};
  const int hi[3] = {x->c.boundaryRight, x->c.boundaryTop, x->c.boundaryFront};
  for (int d = 0; d < 3; d++) {
    bc.lo[d] = lo[d];
    bc.hi[d] = hi[d];
  }
  communicateKernel<<<blocksFor(x->D.n), kBlock, 0, x->s>>>(x->D, bc, nvars, prim);
  return checkLaunch("communicate");
}

// Faces the stage reads: along DIR the faces of the bulk zones, across it
// one zone beyond the bulk (the flux-CT EMFs read the transverse
// neighbours' fluxes); the whole face range when the caller asks for the
// flux arrays themselves.
Box stageFaceBox(const Dims& D, int dir) {
  Box b = faceBox(D, dir);
  for (int d = 0; d < 3; d++) {
    if (d == dir || d >= D.dim) continue;
    b.lo[d] = D.g[d] - 1;
    b.hi[d] = D.g[d] + D.N[d] + 1;
  }
  return b;
}

template <class L, int DIR>
void launchFace(grim_ctx* x, const double* prim, double* flux, bool allFaces = true) {
  const Box b = allFaces ? faceBox(x->D, DIR) : stageFaceBox(x->D, DIR);
  const double* gf = DIR == 0 ? x->geomX1 : (DIR == 1 ? x->geomX2 : x->geomC);
  if (x->c.explicitFast) {
    (void)launchFaceFast(x->p, x->a, DIR, x->c.reconstruction, b, prim, gf, flux, x->s);
    return;
  }
  const unsigned nb = (unsigned)((b.count() + kFacesPerBlock - 1) / kFacesPerBlock);
  switch (x->c.reconstruction) {
    case GRIM_RECON_WENO5:
      faceFluxSplitKernel<L, DIR, GRIM_RECON_WENO5><<<nb, kBlock, 0, x->s>>>(x->a, b, prim, gf, flux);
      break;
    case GRIM_RECON_PPM:
      faceFluxSplitKernel<L, DIR, GRIM_RECON_PPM><<<nb, kBlock, 0, x->s>>>(x->a, b, prim, gf, flux);
      break;
    default:
      faceFluxSplitKernel<L, DIR, GRIM_RECON_MINMOD><<<nb, kBlock, 0, x->s>>>(x->a, b, prim, gf, flux);
  }
}

// riemannSolver::solve from reconstructed states (primLeft / primRight of
// reconstruction::reconstruct) on the faces the fused launch covers.
template <class L, int DIR>
void launchFaceGiven(grim_ctx* x, const double* primLeft, const double* primRight, double* flux) {
  const Box b = faceBox(x->D, DIR);
  const double* gf = DIR == 0 ? x->geomX1 : (DIR == 1 ? x->geomX2 : x->geomC);
  const unsigned nb = (unsigned)((b.count() + kFacesPerBlock - 1) / kFacesPerBlock);
  faceFluxSplitKernel<L, DIR, GIVEN_STATES><<<nb, kBlock, 0, x->s>>>(x->a, b, primLeft, gf, flux, primRight);
}

// Riemann fluxes of every active direction into x->fluxes, then the EMFs.
template <class L>
int fluxesAndEmf(grim_ctx* x, const double* prim, bool allFaces) {
  const size_t plane = (size_t)L::DOF * x->D.n;
  launchFace<L, 0>(x, prim, x->fluxes, allFaces);
  if (x->D.dim > 1) launchFace<L, 1>(x, prim, x->fluxes + plane, allFaces);
  if (x->D.dim > 2) launchFace<L, 2>(x, prim, x->fluxes + 2 * plane, allFaces);
  if (x->D.dim > 1) {
    const Box e = edgeBox(x->D);
    emfKernel<L><<<blocksFor(e.count()), kBlock, 0, x->s>>>(x->D, e, x->fluxes, x->emf);
  }
  return checkLaunch("face fluxes");
}

template <class L>
int explicitSources(grim_ctx* x, const double* prim, const double* primOld, double* sExp, double* sImp,
                    double* consOld) {
  const Box b = bulkBox(x->D);
  if constexpr (L::EMHD)
    uTKernel<L><<<blocksFor(x->D.n), kBlock, 0, x->s>>>(x->a, prim, x->geomC, x->uT);
  if (x->c.explicitFast)
    return launchSourcesFast(x->p, x->a, x->c.reconstruction, b, prim, primOld, x->geomC, x->gamma40, x->uT, sExp,
                             sImp, consOld, x->s);
  const unsigned nb = blocksFor(b.count());
  switch (x->c.reconstruction) {
    case GRIM_RECON_WENO5:
      sourcesKernel<L, GRIM_RECON_WENO5><<<nb, kBlock, 0, x->s>>>(x->a, b, prim, primOld, x->geomC, x->gamma,
                                                                x->uT, sExp, sImp, consOld);
      break;
    case GRIM_RECON_PPM:
      sourcesKernel<L, GRIM_RECON_PPM><<<nb, kBlock, 0, x->s>>>(x->a, b, prim, primOld, x->geomC, x->gamma,
                                                              x->uT, sExp, sImp, consOld);
      break;
    default:
      sourcesKernel<L, GRIM_RECON_MINMOD><<<nb, kBlock, 0, x->s>>>(x->a, b, prim, primOld, x->geomC, x->gamma,
                                                                 x->uT, sExp, sImp, consOld);
  }
  return checkLaunch("sources");
}

// max over zones of the inverse time step, into x->red[nb] (device); with
// clock, also the new dt into the device clock.
template <class L>
const double* reduceInvDt(grim_ctx* x, const double* prim, double* clock = nullptr) {
  const unsigned nb = blocksFor(x->D.n);
  dtKernel<L><<<nb, kBlock, 0, x->s>>>(x->a, prim, x->geomC, x->red);
  maxReduceKernel<<<1, kBlock, 0, x->s>>>(nb, x->red, x->red + nb, clock, x->c.CourantFactor,
                                          x->c.maxDtIncrement);
  return x->red + nb;
}

// timeStepper::computeDt with the previous dt given by the caller; returns
// the new dt to the host (synchronises the stream).
template <class L>
int computeDt(grim_ctx* x, const double* prim, double dt, double* out) {
  const double* maxInv = reduceInvDt<L>(x, prim);
  GRIM_HIP_TRY(hipMemcpyAsync(x->hostBuf, maxInv, sizeof(double), hipMemcpyDeviceToHost, x->s));
  GRIM_HIP_TRY(hipStreamSynchronize(x->s));
  const double maxInvDt = x->hostBuf[0];
  double newDt = x->c.CourantFactor / maxInvDt;
  if (newDt > x->c.maxDtIncrement * dt) newDt = x->c.maxDtIncrement * dt;
  *out = newDt;
  return GRIM_OK;
}

// The same on the device clock: clock[dt] becomes the new dt, no host read.
template <class L>
int computeDtOnDevice(grim_ctx* x, const double* prim) {
  reduceInvDt<L>(x, prim, x->clock);
  return checkLaunch("computeDt");
}

template <class L>
int applyFloor(grim_ctx* x, double* prim) {
  floorKernel<L><<<blocksFor(x->D.n), kBlock, 0, x->s>>>(x->a, x->tp, x->geomC, x->floors, prim);
  return checkLaunch("applyFloor");
}

int scrubNaNs(grim_ctx* x, double* prim, int nvars) {
  const Box b = bulkBox(x->D);
  zeroCountersKernel<<<1, 64, 0, x->s>>>(x->counters, 1);
  nanCountKernel<<<blocksFor(b.count()), kBlock, 0, x->s>>>(x->D, b, nvars, prim, x->counters);
  nanZeroKernel<<<blocksFor(x->D.n), kBlock, 0, x->s>>>(x->D.n, nvars, x->counters, prim);
  return checkLaunch("NaN scrub");
}

// The observers of prim into x->obs[kObsBlocks * OBS_WORDS + q] (device);
// volElem = dX1 (* dX2 from dim 2, * dX3 from dim 3) as in
// fullStepDiagnostics (torus/problem.cpp:764-774).
template <class L>
int observers(grim_ctx* x, const double* prim) {
  const Box b = bulkBox(x->D);
  double volElem = x->a.dX[0];
  if (x->c.dim > 1) volElem *= x->a.dX[1];
  if (x->c.dim > 2) volElem *= x->a.dX[2];
  observersKernel<L><<<kObsBlocks, kBlock, 0, x->s>>>(x->a, b, volElem, prim, x->geomC, x->obs);
  obsFinishKernel<<<1, kBlock, 0, x->s>>>(x->obs, x->obs + (size_t)kObsBlocks * OBS_WORDS);
  return checkLaunch("observers");
}

// fullStepDiagnostics' WriteData variables (torus/problem.cpp:808-875) of
// the bulk zones, X1 fastest, as planes out[q * bulk + t]: rho, u, v^i =
// u^i/u^t and B^i = b^i u^t - b^t u^i in the x coordinates, [q], [deltaP],
// b^2, the Lorentz factor u^t alpha_x (params.cpp:136-147), then the three
// xCoords of the zone centre.  dxdX is computeTransformationMatrices
// (geometry.cpp:441-508): central differences of the coordinate map with
// GAMMA_EPS = 1e-5, dxdX[0][0] = 1 and the other time entries 0.
template <class L>
__global__ __launch_bounds__(kBlock) void dumpVarsKernel(StepArgs a, Box box, CoordMap cm, CoordGrid cg,
                                                         const double* __restrict__ prim,
                                                         const double* __restrict__ geomC,
                                                         double* __restrict__ out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t nb = box.count();
  if (t >= nb) return;
  const int64_t n = a.D.n;
  int i, j, k;
  box.at(t, i, j, k);
  const int64_t z = flatIndex(a.D, i, j, k);
  double X[3], x[3];
  XCoordsAt(a.D, cg, i, j, k, 0, X);
  cm.toX(X, x);
  const double eps = 1.e-5;
  double J[3][3];
  for (int b = 0; b < 3; b++) {
    double Xp[3] = {X[0], X[1], X[2]}, Xm[3] = {X[0], X[1], X[2]}, xp[3], xm[3];
    Xp[b] += eps;
    Xm[b] -= eps;
    cm.toX(Xp, xp);
    cm.toX(Xm, xm);
#pragma unroll
    for (int r = 0; r < 3; r++) J[r][b] = (xp[r] - xm[r]) / (2. * eps);
  }
  double P[L::DOF];
#pragma unroll
  for (int v = 0; v < L::DOF; v++) P[v] = prim[v * n + z];
  const int64_t zg = geomIndex(a.D, i, j);
  ZoneGeom G;
  loadGeom(geomC, a.D.ng, zg, G);
  Kinematics e;
  kinematics<L>(P, G, a.prm, e);
  Closure c;
  closure<L>(P, e, a.prm, 1., 1., geomC[(int64_t)LOC_R * a.D.ng + zg], c);
  double ux[4], bx[4];
  ux[0] = e.uCon[0];
  bx[0] = e.bCon[0];
#pragma unroll
  for (int r = 0; r < 3; r++) {
    ux[r + 1] = 0.;
    bx[r + 1] = 0.;
#pragma unroll
    for (int b = 0; b < 3; b++) {
      ux[r + 1] += J[r][b] * e.uCon[b + 1];
      bx[r + 1] += J[r][b] * e.bCon[b + 1];
    }
  }
  const double alphax = 1. / sqrt(-G.gcon0[0]);
  double* o = out + t;
  int q = 0;
  o[(q++) * nb] = e.rho;
  o[(q++) * nb] = e.u;
#pragma unroll
  for (int r = 1; r <= 3; r++) o[(q++) * nb] = ux[r] / ux[0];
  if constexpr (L::COND_ON) o[(q++) * nb] = c.q;
  if constexpr (L::VISC_ON) o[(q++) * nb] = c.deltaP;
#pragma unroll
  for (int r = 1; r <= 3; r++) o[(q++) * nb] = bx[r] * ux[0] - bx[0] * ux[r];
  o[(q++) * nb] = e.bSqr;
  o[(q++) * nb] = ux[0] * alphax;
#pragma unroll
  for (int d = 0; d < 3; d++) o[(q++) * nb] = x[d];
}

template <class L>
int dumpVars(grim_ctx* x, const double* prim, double* out) {
  const Box b = bulkBox(x->D);
  dumpVarsKernel<L><<<blocksFor(b.count()), kBlock, 0, x->s>>>(x->a, b, x->cm, x->cg, prim, x->geomC, out);
  return checkLaunch("dump variables");
}

// setProblemSpecificBCs of bondi_inflow (bondi_inflow/problem.cpp:157-183):
// every ghost zone of prim from the stored initial state.
// The hook is switched by a device flag that follows the dof planes of ic, so
// a captured step picks up later switches and never reads freed memory.
__global__ __launch_bounds__(kBlock) void ghostStateKernel(Dims D, int nvars, const double* __restrict__ ic,
                                                           double* __restrict__ prim) {
  const int64_t z = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (z >= D.n) return;
  if (ic[(int64_t)nvars * D.n] == 0.) return;
  int i, j, k;
  zoneIndex(D, z, i, j, k);
  if (i >= D.g[0] && i < D.g[0] + D.N[0] && j >= D.g[1] && j < D.g[1] + D.N[1] && k >= D.g[2] &&
      k < D.g[2] + D.N[2])
    return;
  for (int v = 0; v < nvars; v++) prim[v * D.n + z] = ic[v * D.n + z];
}

// setProblemSpecificBCs of buoyancy_instabilities (problem.cpp:221-293):
// the zones of X1 index 0..g (inclusive, as af::seq(0, numGhost)) and
// N1+g..N1+2g-1 take rho, u, u~^1 of the stored state, zero transverse
// velocity, zero field and (conduction) zero q~.  One thread per zone of
// the two X1 slabs, (2g + 1) NT2 NT3 threads; the switch follows the three
// planes, as ghostStateKernel's.
template <class L>
__global__ __launch_bounds__(kBlock) void sideStateKernel(Dims D, const double* __restrict__ ic,
                                                          double* __restrict__ prim) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int w = 2 * D.g[0] + 1;
  if (t >= (int64_t)w * D.NT[1] * D.NT[2]) return;
  if (ic[3 * D.n] == 0.) return;
  const int ii = (int)(t % w);
  const int64_t jk = t / w;
  const int i = ii <= D.g[0] ? ii : D.N[0] + D.g[0] + (ii - D.g[0] - 1);
  const int64_t z = i + (int64_t)D.NT[0] * jk;
  const int64_t n = D.n;
  prim[RHO * n + z] = ic[z];
  prim[UU * n + z] = ic[n + z];
  prim[U1 * n + z] = ic[2 * n + z];
  prim[U2 * n + z] = 0.;
  prim[U3 * n + z] = 0.;
  prim[L::B1 * n + z] = 0.;
  prim[L::B2 * n + z] = 0.;
  prim[L::B3 * n + z] = 0.;
  if constexpr (L::COND_ON) prim[L::Q * n + z] = 0.;
}

// setProblemSpecificBCs of the torus: inflowCheck on both radial ghost
// slabs, then fixPoles (problem.cpp:1217-1241).
template <class L>
int torusBCs(grim_ctx* x, double* prim) {
  const Dims& D = x->D;
  for (int side = 0; side < 2; side++) {
    Box b;
    for (int d = 0; d < 3; d++) {
      b.lo[d] = 0;
      b.hi[d] = D.NT[d];
    }
    b.lo[0] = side == 0 ? 0 : D.N[0] + D.g[0];
    b.hi[0] = side == 0 ? D.g[0] : D.N[0] + 2 * D.g[0];
    inflowCheckKernel<L><<<blocksFor(b.count()), kBlock, 0, x->s>>>(x->a, x->tp, b, side, x->geomC, prim);
  }
  if (D.dim >= 2) {
    const int64_t cols = 2 * (int64_t)D.NT[0] * D.NT[2];
    fixPolesKernel<L><<<blocksFor(cols), kBlock, 0, x->s>>>(D, x->geomC, prim);
  }
  return checkLaunch("torus boundaries");
}

// One stage of timeStep (timestep.cpp:14-107 half, 183-255 full).
// The stage buffers have fixed roles (no pointer swaps), so that a captured
// step reads and writes the same addresses on every replay: the half stage
// advances primOld into primHalf, the full stage primHalf into primNew,
// which the end of the step copies back into primOld.
inline double* stageIn(grim_ctx* x, int st) { return st == 0 ? x->primOld : x->primHalf; }
inline double* stageOut(grim_ctx* x, int st) { return st == 0 ? x->primHalf : x->primNew; }
inline const double* stageDtPrime(grim_ctx* x, int st) {
  return x->clock + (st == 0 ? CLOCK_HALF_DT : CLOCK_DT);
}

// Everything a stage does before its solve (timestep.cpp:15-89 / 184-237):
// boundary conditions and the problem's hooks on the stage state, the
// explicit and old implicit sources, the divergence of the fluxes with
// flux-CT, the induction update of B, and the folded right-hand side
// rhs = consOld - dt' (divF + S_expl + S_impl,old / 2) with the residual
// mask.  The Newton guess (the stage state's fluid rows) goes into the
// stage's output buffer.
//
// guessInSolve: the stage's own solve reads the fluid rows straight from the
// stage state (grim_zone_batch.guess), so only the B rows are copied here;
// the stage-level entry points, whose caller may solve another buffer or
// inspect the guess, get the whole state copied.
// foldInSolve: the stage's compacting solve folds each zone's right-hand
// side itself (newtonPhase1Fold), so no fold kernel runs here.
template <class L>
int stagePrepare(grim_ctx* x, int st, bool guessInSolve = false, bool foldInSolve = false) {
  const Dims& D = x->D;
  const size_t bytes = sizeof(double) * (size_t)L::DOF * D.n;
  double* primStage = stageIn(x, st);
  double* out = stageOut(x, st);
  if (int rc = applyBoundaries(x, primStage, L::DOF, false)) return rc;
  if (x->torus)
    if (int rc = torusBCs<L>(x, primStage)) return rc;
  // grim calls the hook in both stages, but it writes primOld: only the half
  // stage changes anything
  if (x->ghostState && st == 0) {
    ghostStateKernel<<<blocksFor(D.n), kBlock, 0, x->s>>>(D, L::DOF, x->ghostState, primStage);
    if (int rc = checkLaunch("ghost state")) return rc;
  }
  // buoyancy_instabilities' hook writes the stage state in both stages
  if (x->sideState) {
    const int64_t count = (int64_t)(2 * D.g[0] + 1) * D.NT[1] * D.NT[2];
    sideStateKernel<L><<<blocksFor(count), kBlock, 0, x->s>>>(D, x->sideState, primStage);
    if (int rc = checkLaunch("X1 side state")) return rc;
  }
  if (int rc = explicitSources<L>(x, primStage, x->primOld, x->sExp, L::EMHD ? x->sImp : nullptr,
                                  st == 0 ? x->consOld : nullptr))
    return rc;
  if (int rc = fluxesAndEmf<L>(x, primStage, false)) return rc;
  // the solve starts from primStage's fluid rows (guessInSolve); the field
  // rows come from the fold in the bulk and from primStage elsewhere, which
  // a fold inside the solve copies itself
  if (foldInSolve) return GRIM_OK;
  if (guessInSolve) {
    const size_t fluid = (size_t)L::NF * D.n;
    GRIM_HIP_TRY(hipMemcpyAsync(out + fluid, primStage + fluid, bytes - sizeof(double) * fluid,
                                hipMemcpyDeviceToDevice, x->s));
  } else {
    GRIM_HIP_TRY(hipMemcpyAsync(out, primStage, bytes, hipMemcpyDeviceToDevice, x->s));
  }
  const Box b = bulkBox(D);
  launchFold<L, true>(x->a, b, x->s, x->fluxes, x->emf, x->geomC, x->consOld, x->sExp, x->sImp,
                      stageDtPrime(x, st), nullptr, x->rhs, out, x->c.residualScaling ? x->mask : nullptr);
  return checkLaunch("fold");
}

// The stage solve folds in phase 1 when its options compact with the mixed
// analytic Jacobian on a 3D grid (the grid-step path only: the stage-level
// entry points keep the separate fold, whose rhs the caller may read).
bool foldInSolve(const grim_ctx* x) { return foldFusable(x->p, x->opts, x->D.dim); }

// The stage's implicit system as a batch of independent zones (the frozen
// terms folded into rhs; elemOld = primOld, the coefficient element primOld
// on the half stage and primHalf on the full one).
grim_zone_batch stageBatch(grim_ctx* x, int st, double* prim) {
  grim_zone_batch zb = {};
  zb.n = x->D.n;
  zb.prim = prim;
  zb.geom = x->geomC;
  zb.rhs = x->rhs;
  zb.prim_old = x->primOld;
  zb.prim_coef = st == 0 ? x->primOld : x->primHalf;
  zb.closure = nullptr;
  zb.radius = x->geomC + (size_t)LOC_R * x->D.ng;
  zb.geom_columns = x->D.ng;
  zb.residual_mask = x->mask;
  zb.dt_prime_device = stageDtPrime(x, st);
  return zb;
}

// timeStepper::solve of the stage (solve.cpp:3-201) on prim; the per-zone
// iteration counts and final residuals stay in the context.
int stageSolve(grim_ctx* x, int st, double* prim, const double* guess = nullptr, bool fold = false) {
  grim_zone_batch zb = stageBatch(x, st, prim);
  zb.guess = guess;
  grim_newton_stats ns = {x->iters, nullptr, x->l2};
  if (!fold) return grim_batched_newton(&x->p, &zb, &x->opts, &ns, x->s);
  FoldArgs f;
  f.D = x->D;
  f.bulk = bulkBox(x->D);
  for (int d = 0; d < 3; d++) f.dX[d] = x->a.dX[d];
  f.fluxFilter = x->a.fluxFilter;
  f.scaling = x->c.residualScaling ? 1 : 0;
  f.F = x->fluxes;
  f.emf = x->emf;
  f.geomC = x->geomC;
  f.consOld = x->consOld;
  f.sExp = x->sExp;
  f.sImp = x->sImp;
  f.rhs = x->rhs;
  f.mask = x->mask;
  f.fieldIn = guess + (size_t)(x->dof - 3) * x->D.n;
  return batchedNewton(&x->p, &zb, &x->opts, &ns, x->s, &f);
}

// Unconverged / non-finite bulk zones and the largest iteration count of the
// last solve (synchronises the stream).
int stageStats(grim_ctx* x, int st, grim_step_stats* stats) {
  zeroCountersKernel<<<1, 64, 0, x->s>>>(x->counters, 3);
  stageStatsKernel<<<kStatsBlocks, kBlock, 0, x->s>>>(x->D.n, x->mask, x->l2, x->iters, x->p.nonlinearsolve_atol,
                                                        x->counters);
  GRIM_HIP_TRY(hipMemcpyAsync(x->hostBuf, x->counters, 3 * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                              x->s));
  GRIM_HIP_TRY(hipStreamSynchronize(x->s));
  unsigned long long c[3];
  std::memcpy(c, x->hostBuf, sizeof(c));
  stats->unconverged[st] = (int64_t)c[0];
  stats->max_iterations[st] = (int32_t)c[1];
  stats->nonfinite[st] = (int64_t)c[2];
  return GRIM_OK;
}

// After the solve: communicate, and the problem's half/full-step
// diagnostics hooks (torus: NaN scrub on the full stage, applyFloor;
// torus/problem.cpp:714-746).
template <class L>
int stageFinish(grim_ctx* x, int st) {
  double* out = stageOut(x, st);
  if (int rc = communicate(x, out, L::DOF)) return rc;
  if (x->torus) {
    if (st == 1)
      if (int rc = scrubNaNs(x, out, L::DOF)) return rc;
    if (int rc = applyFloor<L>(x, out)) return rc;
  }
  return GRIM_OK;
}

template <class L>
int stage(grim_ctx* x, int st, grim_step_stats* stats) {
  const bool fold = foldInSolve(x);
  if (int rc = stagePrepare<L>(x, st, true, fold)) return rc;
  if (int rc = stageSolve(x, st, stageOut(x, st), stageIn(x, st), fold)) return rc;
  if (stats)
    if (int rc = stageStats(x, st, stats)) return rc;
  return stageFinish<L>(x, st);
}

// The end of timeStepper::timeStep: primOld = the full stage's result, the
// state the next computeDt reads, and time += dt (timestep.cpp:290-300).
template <class L>
int stepEnd(grim_ctx* x) {
  endStepKernel<L, false><<<blocksFor(x->D.n), kBlock, 0, x->s>>>(x->a, x->tp, x->geomC, x->floors, x->counters,
                                                                  x->primNew, x->primOld, x->primDt, x->clock,
                                                                  x->dtPart);
  return checkLaunch("end of step");
}

// The full stage's finish and the end of the step in one pass (the torus's
// NaN scrub and floors, the copies, the clock and the next dt's partial
// maxima): stageFinish(1) followed by stepEnd.
template <class L>
int finishStep(grim_ctx* x) {
  if (int rc = communicate(x, x->primNew, L::DOF)) return rc;
  if (x->torus) {
    const Box b = bulkBox(x->D);
    zeroCountersKernel<<<1, 64, 0, x->s>>>(x->counters, 1);
    nanCountKernel<<<blocksFor(b.count()), kBlock, 0, x->s>>>(x->D, b, L::DOF, x->primNew, x->counters);
    endStepKernel<L, true><<<blocksFor(x->D.n), kBlock, 0, x->s>>>(x->a, x->tp, x->geomC, x->floors, x->counters,
                                                                   x->primNew, x->primOld, x->primDt, x->clock,
                                                                   x->dtPart);
    return checkLaunch("end of step");
  }
  return stepEnd<L>(x);
}

int validConfig(const grim_params* p, const grim_grid_config* c) {
  if (!validParams(p) || !c) return 0;
  if (c->dim < 1 || c->dim > 3 || c->numGhost != 3) return 0;
  if (c->N1 < 1 || c->N2 < 1 || c->N3 < 1) return 0;
  // makeDims adds the ghost zones in int: keep N + 2 numGhost representable
  constexpr int kMaxN = (1 << 30);
  if (c->N1 > kMaxN || c->N2 > kMaxN || c->N3 > kMaxN) return 0;
  if (c->dim < 2 && c->N2 != 1) return 0;
  if (c->dim < 3 && c->N3 != 1) return 0;
  if (c->metric != GRIM_METRIC_MINKOWSKI && c->metric != GRIM_METRIC_MKS) return 0;
  const int b[6] = {c->boundaryLeft, c->boundaryRight, c->boundaryBottom, c->boundaryTop, c->boundaryBack,
                    c->boundaryFront};
  for (int q = 0; q < 6; q++)
    if (b[q] < GRIM_BC_PERIODIC || b[q] > GRIM_BC_MIRROR) return 0;
  if (c->reconstruction < 0 || c->reconstruction > 2 || c->riemannSolver < 0 || c->riemannSolver > 1)
    return 0;
  if ((p->conduction || p->viscosity) && p->closure == GRIM_CLOSURE_PER_ZONE) return 0;
  if (c->explicitFast != 0 && c->explicitFast != 1) return 0;
  return 1;
}

}  // namespace
}  // namespace grim

extern "C" {

int64_t grim_grid_zones(const grim_grid_config* c) {
  if (!c || c->N1 < 1 || c->N2 < 1 || c->N3 < 1 || c->N1 > (1 << 30) || c->N2 > (1 << 30) || c->N3 > (1 << 30) ||
      c->numGhost < 0 || c->numGhost > 64)
    return 0;
  return makeDims(*c).n;
}

int grim_create(const grim_params* p, const grim_grid_config* c, const grim_newton_opts* opts,
                grim_stream_t stream, grim_ctx** out) {
  GRIM_REQUIRE(out, "grim_create: null output");
  *out = nullptr;
  GRIM_REQUIRE(validConfig(p, c), "grim_create: bad params or grid config");
  GRIM_REQUIRE(grim_grid_zones(c) < (int64_t(1) << 31), "grim_create: grids hold fewer than 2^31 zones");
  if (int rc = checkDevice()) return rc;
  grim_ctx* x = new (std::nothrow) grim_ctx();
  GRIM_REQUIRE(x, "grim_create: host allocation failed");
  x->p = *p;
  x->c = *c;
  x->opts = grim_newton_opts{};
  x->opts.jacobian = GRIM_JAC_FD;
  x->opts.lu_precision = GRIM_LU_FP64;
  x->opts.semantics = GRIM_NEWTON_PER_ZONE;
  if (opts) x->opts = *opts;
  x->opts.workspace = nullptr;
  x->opts.workspace_bytes = 0;
  x->s = (hipStream_t)stream;
  x->D = makeDims(*c);
  x->dof = 8 + p->conduction + p->viscosity;
  x->nf = 5 + p->conduction + p->viscosity;
  x->a.D = x->D;
  x->a.prm = makeDevParams(*p);
  x->a.dX[0] = (c->X1End - c->X1Start) / c->N1;
  x->a.dX[1] = (c->X2End - c->X2Start) / c->N2;
  x->a.dX[2] = (c->X3End - c->X3Start) / c->N3;
  x->a.theta = c->slopeLimTheta;
  x->a.riemann = c->riemannSolver;
  x->a.metric = c->metric;
  x->cg.start[0] = c->X1Start;
  x->cg.start[1] = c->X2Start;
  x->cg.start[2] = c->X3Start;
  for (int d = 0; d < 3; d++) x->cg.dX[d] = x->a.dX[d];
  x->cm.metric = c->metric;
  x->cm.derefine = c->DerefineThetaHorizon;
  x->cm.cylindrify = c->DoCylindrify;
  x->cm.a = c->blackHoleSpin;
  x->cm.hSlope = c->hSlope;
  x->cm.X1Start = c->X1Start;
  x->cm.X1cyl = c->X1cyl;
  x->cm.X2cyl = c->X2cyl;
  x->clockSet = false;
  const int64_t n = x->D.n;
  const size_t dofB = sizeof(double) * (size_t)x->dof * n;
  const size_t nfB = sizeof(double) * (size_t)x->nf * n;
  // geometry planes: one value per (X1, X2) column (D.ng)
  const int64_t ng = x->D.ng;
  const size_t locB = sizeof(double) * (size_t)GRIM_LOC_PLANES * ng;
  const unsigned nb = blocksFor(n);
  struct Req {
    void** ptr;
    size_t bytes;
  } reqs[] = {
      {(void**)&x->primOld, dofB},
      {(void**)&x->primHalf, dofB},
      {(void**)&x->primNew, dofB},
      {(void**)&x->primDt, dofB},
      {(void**)&x->consOld, dofB},
      {(void**)&x->fluxes, dofB * (size_t)c->dim},
      {(void**)&x->emf, sizeof(double) * 3 * (size_t)n},
      {(void**)&x->sExp, nfB},
      {(void**)&x->sImp, nfB},
      {(void**)&x->uT, sizeof(double) * 5 * (size_t)n},
      {(void**)&x->rhs, nfB},
      {(void**)&x->mask, sizeof(double) * (size_t)n},
      {(void**)&x->geomC, locB},
      {(void**)&x->geomX1, locB},
      {(void**)&x->geomX2, locB},
      {(void**)&x->gamma, sizeof(double) * 64 * (size_t)ng},
      {(void**)&x->red, sizeof(double) * ((size_t)nb + 1)},
      {(void**)&x->dtPart, sizeof(double) * ((size_t)nb + 1)},
      {(void**)&x->obs, sizeof(double) * ((size_t)kObsBlocks + 1) * OBS_WORDS},
      {(void**)&x->iters, sizeof(int) * (size_t)n},
      {(void**)&x->l2, sizeof(double) * (size_t)n},
      {(void**)&x->counters, 4 * sizeof(unsigned long long)},
      {(void**)&x->newtonWs, (size_t)grim_newton_workspace_bytes(n)},
      {(void**)&x->clock, sizeof(double) * CLOCK_WORDS},
  };
  for (auto& r : reqs) {
    hipError_t e = hipMalloc(r.ptr, r.bytes);
    if (e != hipSuccess) {
      setError(std::string("grim_create: hipMalloc: ") + hipGetErrorString(e));
      grim_destroy(x);
      return GRIM_ERR_OOM;
    }
  }
  x->opts.workspace = x->newtonWs;
  x->opts.workspace_bytes = grim_newton_workspace_bytes(n);
  if (hipHostMalloc((void**)&x->hostBuf, 128) != hipSuccess) {
    setError("grim_create: hipHostMalloc failed");
    grim_destroy(x);
    return GRIM_ERR_OOM;
  }
  const struct {
    void* ptr;
    size_t bytes;
  } zeros[] = {{x->primOld, dofB}, {x->fluxes, dofB * (size_t)c->dim}, {x->emf, sizeof(double) * 3 * (size_t)n},
               {x->sImp, nfB},     {x->uT, sizeof(double) * 5 * (size_t)n}, {x->rhs, nfB}};
  for (auto& zr : zeros) {
    hipError_t e = hipMemsetAsync(zr.ptr, 0, zr.bytes, x->s);
    if (e != hipSuccess) {
      setError(std::string("grim_create: hipMemsetAsync: ") + hipGetErrorString(e));
      grim_destroy(x);
      return GRIM_ERR_HIP;
    }
  }
  const unsigned nbCols = blocksFor(x->D.ng);
  geometryKernel<<<nbCols, kBlock, 0, x->s>>>(x->D, x->cm, x->cg, 0, x->geomC);
  geometryKernel<<<nbCols, kBlock, 0, x->s>>>(x->D, x->cm, x->cg, 1, x->geomX1);
  geometryKernel<<<nbCols, kBlock, 0, x->s>>>(x->D, x->cm, x->cg, 2, x->geomX2);
  connectionKernel<<<nbCols, kBlock, 0, x->s>>>(x->D, x->cm, x->cg, x->gamma);
  if (c->explicitFast) {
    hipError_t e = hipMalloc((void**)&x->gamma40, sizeof(double) * 40 * (size_t)ng);
    if (e != hipSuccess) {
      setError(std::string("grim_create: hipMalloc: ") + hipGetErrorString(e));
      grim_destroy(x);
      return GRIM_ERR_OOM;
    }
    launchGamma40(x->D.ng, x->gamma, x->gamma40, x->s);
  }
  maskKernel<<<nb, kBlock, 0, x->s>>>(x->D, x->mask);
  if (int rc = checkLaunch("grim_create")) {
    grim_destroy(x);
    return rc;
  }
  hipError_t e = hipStreamSynchronize(x->s);
  if (e != hipSuccess) {
    setError(std::string("grim_create: ") + hipGetErrorString(e));
    grim_destroy(x);
    return GRIM_ERR_HIP;
  }
  *out = x;
  return GRIM_OK;
}

int grim_destroy(grim_ctx* x) {
  if (!x) return GRIM_OK;
  if (x->s) (void)hipStreamSynchronize(x->s);
  void* ptrs[] = {x->primOld, x->primHalf, x->primNew, x->primDt, x->consOld, x->fluxes, x->emf,
                  x->sExp,    x->sImp,     x->uT,      x->rhs,    x->mask,    x->geomC,  x->geomX1,
                  x->geomX2,  x->gamma,    x->red,     x->dtPart, x->iters,  x->l2,      x->counters,
                  x->newtonWs, x->clock,  x->obs,  x->gamma40};
  for (void* q : ptrs)
    if (q) (void)hipFree(q);
  if (x->hostBuf) (void)hipHostFree(x->hostBuf);
  if (x->floors) (void)hipFree(x->floors);
  if (x->dump) (void)hipFree(x->dump);
  if (x->ghostState) (void)hipFree(x->ghostState);
  if (x->sideState) (void)hipFree(x->sideState);
  delete x;
  return GRIM_OK;
}

int grim_prim(grim_ctx* x, double** prim_old) {
  GRIM_REQUIRE(x && prim_old, "grim_prim: null argument");
  *prim_old = x->primOld;
  return GRIM_OK;
}

int grim_geometry(grim_ctx* x, int which, double** geom) {
  GRIM_REQUIRE(x && geom && which >= 0 && which <= 3, "grim_geometry: bad argument");
  double* g[4] = {x->geomC, x->geomX1, x->geomX2, x->gamma};
  *geom = g[which];
  return GRIM_OK;
}

int grim_set_prim(grim_ctx* x, const double* src, int on_device) {
  GRIM_REQUIRE(x && src, "grim_set_prim: null argument");
  GRIM_HIP_TRY(hipMemcpyAsync(x->primOld, src, sizeof(double) * (size_t)x->dof * x->D.n,
                              on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, x->s));
  GRIM_HIP_TRY(hipStreamSynchronize(x->s));
  return GRIM_OK;
}

int grim_get_prim(grim_ctx* x, double* dst, int on_device) {
  GRIM_REQUIRE(x && dst, "grim_get_prim: null argument");
  GRIM_HIP_TRY(hipMemcpyAsync(dst, x->primOld, sizeof(double) * (size_t)x->dof * x->D.n,
                              on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, x->s));
  GRIM_HIP_TRY(hipStreamSynchronize(x->s));
  return GRIM_OK;
}

int64_t grim_geometry_columns(const grim_ctx* x) { return x ? x->D.ng : 0; }

int grim_get_geometry(grim_ctx* x, int which, double* dst, int on_device) {
  GRIM_REQUIRE(x && dst && which >= 0 && which <= 4, "grim_get_geometry: bad argument");
  GRIM_REQUIRE(which != 4 || x->floors, "grim_get_geometry: no floors before grim_init_torus");
  double* g[5] = {x->geomC, x->geomX1, x->geomX2, x->gamma, x->floors};
  const size_t planes = which == 3 ? 64 : (which == 4 ? 2 : GRIM_LOC_PLANES);
  const int64_t n = x->D.n;
  if (which == 4) {
    // the floor planes hold every zone
    GRIM_HIP_TRY(hipMemcpyAsync(dst, g[which], sizeof(double) * planes * n,
                                on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, x->s));
    GRIM_HIP_TRY(hipStreamSynchronize(x->s));
    return GRIM_OK;
  }
  // every zone's value from its (X1, X2) column's
  const int64_t ng = x->D.ng;
  if (on_device) {
    const int64_t count = (int64_t)planes * n;
    expandColumnsKernel<<<blocksFor(count), kBlock, 0, x->s>>>(count, n, ng, g[which], dst);
    if (int rc = checkLaunch("grim_get_geometry")) return rc;
    GRIM_HIP_TRY(hipStreamSynchronize(x->s));
    return GRIM_OK;
  }
  std::vector<double> cols;
  try {
    cols.resize(planes * (size_t)ng);
  } catch (const std::bad_alloc&) {
    setError("grim_get_geometry: host allocation failed");
    return GRIM_ERR_OOM;
  }
  GRIM_HIP_TRY(hipMemcpyAsync(cols.data(), g[which], sizeof(double) * planes * ng, hipMemcpyDeviceToHost, x->s));
  GRIM_HIP_TRY(hipStreamSynchronize(x->s));
  for (size_t q = 0; q < planes; q++)
    for (int64_t k = 0; k < n / ng; k++)
      std::memcpy(dst + q * n + k * ng, cols.data() + q * ng, sizeof(double) * ng);
  return GRIM_OK;
}

int grim_set_time(grim_ctx* x, double time, double dt) {
  GRIM_REQUIRE(x && dt > 0., "grim_set_time: bad argument");
  setClockKernel<<<1, 64, 0, x->s>>>(x->clock, time, dt);
  if (int rc = checkLaunch("grim_set_time")) return rc;
  x->clockSet = true;
  // primDt = primOld: computeDt of the first step reads the initial state.
  GRIM_HIP_TRY(hipMemcpyAsync(x->primDt, x->primOld, sizeof(double) * (size_t)x->dof * x->D.n,
                              hipMemcpyDeviceToDevice, x->s));
  return withLayout(x, [&]<class L>() -> int {
    dtKernel<L><<<blocksFor(x->D.n), kBlock, 0, x->s>>>(x->a, x->primDt, x->geomC, x->dtPart);
    return checkLaunch("grim_set_time");
  });
}

int grim_get_time(grim_ctx* x, double* time, double* dt) {
  GRIM_REQUIRE(x && time && dt && x->clockSet, "grim_get_time: null argument or time not set");
  GRIM_HIP_TRY(hipMemcpyAsync(x->hostBuf + 4, x->clock, 2 * sizeof(double), hipMemcpyDeviceToHost, x->s));
  GRIM_HIP_TRY(hipStreamSynchronize(x->s));
  *time = x->hostBuf[4 + CLOCK_TIME];
  *dt = x->hostBuf[4 + CLOCK_DT];
  return GRIM_OK;
}

int grim_time_step(grim_ctx* x, grim_step_stats* stats) {
  GRIM_REQUIRE(x && x->clockSet, "grim_time_step: no context or dt not set (grim_set_time)");
  return withLayout(x, [&]<class L>() -> int {
    // computeDt of primDt from the partial maxima the previous step's end
    // (or grim_set_time) left
    const unsigned nb = blocksFor(x->D.n);
    maxReduceKernel<<<1, kBlock, 0, x->s>>>(nb, x->dtPart, x->dtPart + nb, x->clock, x->c.CourantFactor,
                                            x->c.maxDtIncrement);
    if (int rc = checkLaunch("computeDt")) return rc;
    if (int rc = stage<L>(x, 0, stats)) return rc;
    const bool fold = foldInSolve(x);
    if (int rc = stagePrepare<L>(x, 1, true, fold)) return rc;
    if (int rc = stageSolve(x, 1, x->primNew, x->primHalf, fold)) return rc;
    if (stats)
      if (int rc = stageStats(x, 1, stats)) return rc;
    // primOld = the full stage's result, and elemOld of the next computeDt
    // (the problem's fullStepDiagnostics set it from this step's primOld).
    if (int rc = finishStep<L>(x)) return rc;
    if (stats) {
      GRIM_HIP_TRY(hipMemcpyAsync(x->hostBuf + 4, x->clock, 2 * sizeof(double), hipMemcpyDeviceToHost, x->s));
      GRIM_HIP_TRY(hipStreamSynchronize(x->s));
      stats->dt = x->hostBuf[4 + CLOCK_DT];
    }
    return GRIM_OK;
  });
}

int grim_apply_boundaries(grim_ctx* x, double* prim, int nvars) {
  GRIM_REQUIRE(x && prim && nvars > 0, "grim_apply_boundaries: bad argument");
  return applyBoundaries(x, prim, nvars, true);
}

int grim_communicate(grim_ctx* x, double* prim, int nvars) {
  GRIM_REQUIRE(x && prim && nvars > 0, "grim_communicate: bad argument");
  return communicate(x, prim, nvars);
}

int grim_reconstruct(grim_ctx* x, const double* prim, int nvars, int dir, double* left, double* right) {
  GRIM_REQUIRE(x && prim && left && right && nvars > 0 && dir >= 0 && dir < x->D.dim,
               "grim_reconstruct: bad argument");
  Box b;
  for (int d = 0; d < 3; d++) {
    b.lo[d] = 0;
    b.hi[d] = x->D.NT[d];
  }
  b.lo[dir] = 2;
  b.hi[dir] = x->D.NT[dir] - 2;
  const unsigned nb = blocksFor(b.count());
  const int64_t s = x->D.s[dir];
  switch (x->c.reconstruction) {
    case GRIM_RECON_WENO5:
      reconstructKernel<GRIM_RECON_WENO5><<<nb, kBlock, 0, x->s>>>(x->D, b, s, x->a.theta, nvars, prim, left, right);
      break;
    case GRIM_RECON_PPM:
      reconstructKernel<GRIM_RECON_PPM><<<nb, kBlock, 0, x->s>>>(x->D, b, s, x->a.theta, nvars, prim, left, right);
      break;
    default:
      reconstructKernel<GRIM_RECON_MINMOD><<<nb, kBlock, 0, x->s>>>(x->D, b, s, x->a.theta, nvars, prim, left, right);
  }
  return checkLaunch("grim_reconstruct");
}

int grim_riemann_flux(grim_ctx* x, const double* prim, int dir, double* flux) {
  GRIM_REQUIRE(x && prim && flux && dir >= 0 && dir < x->D.dim, "grim_riemann_flux: bad argument");
  return withLayout(x, [&]<class L>() -> int {
    if (dir == 0) launchFace<L, 0>(x, prim, flux);
    else if (dir == 1) launchFace<L, 1>(x, prim, flux);
    else launchFace<L, 2>(x, prim, flux);
    return checkLaunch("grim_riemann_flux");
  });
}

int grim_div_fluxes(grim_ctx* x, const double* prim, double* div_f, double* fluxes) {
  GRIM_REQUIRE(x && prim && div_f, "grim_div_fluxes: bad argument");
  return withLayout(x, [&]<class L>() -> int {
    if (int rc = fluxesAndEmf<L>(x, prim, fluxes != nullptr)) return rc;
    const Box b = bulkBox(x->D);
    launchFold<L, false>(x->a, b, x->s, x->fluxes, x->emf, x->geomC, nullptr, nullptr, nullptr, nullptr, div_f,
                         nullptr, nullptr, nullptr);
    if (fluxes) {
      GRIM_HIP_TRY(hipMemcpyAsync(fluxes, x->fluxes, sizeof(double) * (size_t)L::DOF * x->D.dim * x->D.n,
                                  hipMemcpyDeviceToDevice, x->s));
      if (x->D.dim > 1) {
        const Box e = bulkBox(x->D);
        ctWriteKernel<L><<<blocksFor(e.count()), kBlock, 0, x->s>>>(x->D, e, x->fluxes, x->emf, fluxes);
      }
    }
    return checkLaunch("grim_div_fluxes");
  });
}

int grim_explicit_sources(grim_ctx* x, const double* prim, double* sources) {
  GRIM_REQUIRE(x && prim && sources, "grim_explicit_sources: bad argument");
  return withLayout(x, [&]<class L>() -> int {
    return explicitSources<L>(x, prim, prim, sources, nullptr, nullptr);
  });
}

int grim_riemann_solve(grim_ctx* x, const double* prim_left, const double* prim_right, int dir, double* flux) {
  GRIM_REQUIRE(x && prim_left && prim_right && flux && dir >= 0 && dir < x->D.dim,
               "grim_riemann_solve: bad argument");
  return withLayout(x, [&]<class L>() -> int {
    if (dir == 0) launchFaceGiven<L, 0>(x, prim_left, prim_right, flux);
    else if (dir == 1) launchFaceGiven<L, 1>(x, prim_left, prim_right, flux);
    else launchFaceGiven<L, 2>(x, prim_left, prim_right, flux);
    return checkLaunch("grim_riemann_solve");
  });
}

namespace {
const double* locationGeometry(grim_ctx* x, int location) {
  return location == 0 ? x->geomC : (location == 1 ? x->geomX1 : (location == 2 ? x->geomX2 : nullptr));
}

int element(grim_ctx* x, const char* what, int kind, int dir, const double* prim, int location,
            const double* aux0, const double* aux1, double dt, double* out0, double* out1) {
  GRIM_REQUIRE(x && prim && out0, std::string(what) + ": null argument");
  const double* geom = locationGeometry(x, location);
  GRIM_REQUIRE(geom, std::string(what) + ": location must be 0 (centre), 1 (X1 faces) or 2 (X2 faces)");
  GRIM_REQUIRE(location < 2 || x->D.dim > 1, std::string(what) + ": no X2 faces on a 1D grid");
  return withLayout(x, [&]<class L>() -> int {
    elementKernel<L><<<blocksFor(x->D.n), kBlock, 0, x->s>>>(x->a, kind, dir, prim, geom, aux0, aux1, dt, out0,
                                                                out1);
    return checkLaunch(what);
  });
}
}  // namespace

int grim_element_fluxes(grim_ctx* x, const double* prim, int location, int dir, double* out) {
  GRIM_REQUIRE(dir >= 0 && dir <= 3, "grim_element_fluxes: dir must be 0..3");
  return element(x, "grim_element_fluxes", ELEM_FLUXES, dir, prim, location, nullptr, nullptr, 0., out, nullptr);
}

int grim_element_char_speeds(grim_ctx* x, const double* prim, int location, int dir, double* min_speed,
                             double* max_speed) {
  GRIM_REQUIRE(x && dir >= 0 && dir < x->D.dim && max_speed, "grim_element_char_speeds: bad argument");
  return element(x, "grim_element_char_speeds", ELEM_SPEEDS, dir, prim, location, nullptr, nullptr, 0., min_speed,
                 max_speed);
}

int grim_element_fields(grim_ctx* x, const double* prim, int location, double* out) {
  return element(x, "grim_element_fields", ELEM_FIELDS, 0, prim, location, nullptr, nullptr, 0., out, nullptr);
}

int grim_element_implicit_sources(grim_ctx* x, const double* prim, const double* tau_damp, double* sources) {
  GRIM_REQUIRE(x && tau_damp, "grim_element_implicit_sources: null argument");
  return element(x, "grim_element_implicit_sources", ELEM_IMPLICIT, 0, prim, 0, tau_damp, nullptr, 0., sources,
                 nullptr);
}

int grim_element_time_deriv_sources(grim_ctx* x, const double* prim, const double* prim_old,
                                    const double* prim_new, double dt, double* sources) {
  GRIM_REQUIRE(x && prim_old && prim_new && dt > 0., "grim_element_time_deriv_sources: bad argument");
  return element(x, "grim_element_time_deriv_sources", ELEM_TIME_DERIV, 0, prim, 0, prim_old, prim_new, dt,
                 sources, nullptr);
}

int grim_step_begin(grim_ctx* x) {
  GRIM_REQUIRE(x && x->clockSet, "grim_step_begin: no context or dt not set (grim_set_time)");
  return withLayout(x, [&]<class L>() -> int { return computeDtOnDevice<L>(x, x->primDt); });
}

int grim_fold_rhs(grim_ctx* x, int stage) {
  GRIM_REQUIRE(x && x->clockSet && (stage == 0 || stage == 1), "grim_fold_rhs: bad argument or dt not set");
  return withLayout(x, [&]<class L>() -> int { return stagePrepare<L>(x, stage); });
}

int grim_solve(grim_ctx* x, int stage, double* prim, grim_step_stats* stats) {
  GRIM_REQUIRE(x && x->clockSet && (stage == 0 || stage == 1), "grim_solve: bad argument or dt not set");
  if (int rc = stageSolve(x, stage, prim ? prim : stageOut(x, stage))) return rc;
  return stats ? stageStats(x, stage, stats) : GRIM_OK;
}

int grim_stage_residual(grim_ctx* x, int stage, const double* prim, double* residual) {
  GRIM_REQUIRE(x && x->clockSet && (stage == 0 || stage == 1) && prim && residual,
               "grim_stage_residual: bad argument");
  const grim_zone_batch zb = stageBatch(x, stage, const_cast<double*>(prim));
  return grim_residual(&x->p, &zb, residual, x->s);
}

int grim_stage_finish(grim_ctx* x, int stage) {
  GRIM_REQUIRE(x && (stage == 0 || stage == 1), "grim_stage_finish: bad argument");
  return withLayout(x, [&]<class L>() -> int { return stageFinish<L>(x, stage); });
}

int grim_step_end(grim_ctx* x) {
  GRIM_REQUIRE(x && x->clockSet, "grim_step_end: no context or dt not set");
  return withLayout(x, [&]<class L>() -> int { return stepEnd<L>(x); });
}

int grim_get_stage_arrays(grim_ctx* x, grim_stage_arrays* out) {
  GRIM_REQUIRE(x && out, "grim_get_stage_arrays: null argument");
  const size_t plane = (size_t)x->dof * x->D.n;
  out->primOld = x->primOld;
  out->primHalfStep = x->primHalf;
  out->prim = x->primNew;
  out->consOld = x->consOld;
  out->sourcesExplicit = x->sExp;
  out->sourcesImplicitOld = x->sImp;
  out->rhs = x->rhs;
  out->residualMask = x->mask;
  out->fluxesX1 = x->fluxes;
  out->fluxesX2 = x->D.dim > 1 ? x->fluxes + plane : nullptr;
  out->fluxesX3 = x->D.dim > 2 ? x->fluxes + 2 * plane : nullptr;
  return GRIM_OK;
}

int grim_last_solve_stats(grim_ctx* x, int32_t* iterations, double* final_l2, int on_device) {
  GRIM_REQUIRE(x && (iterations || final_l2), "grim_last_solve_stats: null argument");
  const hipMemcpyKind kind = on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
  if (iterations) GRIM_HIP_TRY(hipMemcpyAsync(iterations, x->iters, sizeof(int32_t) * (size_t)x->D.n, kind, x->s));
  if (final_l2) GRIM_HIP_TRY(hipMemcpyAsync(final_l2, x->l2, sizeof(double) * (size_t)x->D.n, kind, x->s));
  if (!on_device) GRIM_HIP_TRY(hipStreamSynchronize(x->s));
  return GRIM_OK;
}

int grim_compute_dt(grim_ctx* x, const double* prim, double dt, double* new_dt) {
  GRIM_REQUIRE(x && prim && new_dt && dt > 0., "grim_compute_dt: bad argument");
  return withLayout(x, [&]<class L>() -> int { return computeDt<L>(x, prim, dt, new_dt); });
}

int grim_init_torus(grim_ctx* x, const grim_torus_params* t) {
  GRIM_REQUIRE(x && t, "grim_init_torus: null argument");
  GRIM_REQUIRE(x->c.metric == GRIM_METRIC_MKS && x->c.dim >= 2 && x->c.N2 >= 6,
               "grim_init_torus: needs an MKS grid with dim >= 2 and N2 >= 6");
  const Dims& D = x->D;
  const int64_t n = D.n;
  GRIM_REQUIRE(2 * (int64_t)blocksFor(n) <= n, "grim_init_torus: grid too small");
  if (!x->floors) GRIM_HIP_TRY(hipMalloc(&x->floors, sizeof(double) * 2 * (size_t)n));
  x->tp = *t;
  x->torus = 1;
  x->a.fluxFilter = 1;
  const unsigned nb = blocksFor(n);
  floorPlanesKernel<<<nb, kBlock, 0, x->s>>>(D, x->tp, x->geomC, x->floors);
  return withLayout(x, [&]<class L>() -> int {
    double* prim = x->primOld;
    double* Avec = x->uT;  // scratch plane
    torusICKernel<L><<<nb, kBlock, 0, x->s>>>(x->a, x->tp, x->c.blackHoleSpin, x->c.hSlope, x->cg, x->geomC, prim);
    planeMaxKernel<<<nb, kBlock, 0, x->s>>>(n, prim + RHO * n, 1., x->red);
    maxReduceKernel<<<1, kBlock, 0, x->s>>>(nb, x->red, x->red + nb);
    normalizeKernel<<<nb, kBlock, 0, x->s>>>(n, x->red + nb, prim);
    vectorPotentialKernel<<<nb, kBlock, 0, x->s>>>(D, x->tp, x->geomC, prim + RHO * n, Avec);
    curlKernel<L><<<nb, kBlock, 0, x->s>>>(D, x->a.dX[0], x->a.dX[1], x->geomC, Avec, prim);
    if (int rc = communicate(x, prim, L::DOF)) return rc;
    // the l2 and sExp work arrays hold the two per-block reductions
    double* blockA = x->l2;
    double* blockB = x->l2 + nb;
    betaKernel<L><<<nb, kBlock, 0, x->s>>>(x->a, x->tp.UseMADdisk, prim, x->geomC, blockA, blockB);
    maxReduceKernel<<<1, kBlock, 0, x->s>>>(nb, blockA, x->sExp);
    maxReduceKernel<<<1, kBlock, 0, x->s>>>(nb, blockB, x->sExp + 1);
    scaleFieldKernel<L><<<nb, kBlock, 0, x->s>>>(n, x->tp, x->sExp, x->sExp + 1, prim);
    if (int rc = applyFloor<L>(x, prim)) return rc;
    if (int rc = scrubNaNs(x, prim, L::DOF)) return rc;
    if (int rc = applyFloor<L>(x, prim)) return rc;
    GRIM_HIP_TRY(hipStreamSynchronize(x->s));
    return checkLaunch("grim_init_torus");
  });
}

int grim_apply_floor(grim_ctx* x, double* prim) {
  GRIM_REQUIRE(x && prim && x->torus, "grim_apply_floor: no torus (grim_init_torus) or null prim");
  return withLayout(x, [&]<class L>() -> int { return applyFloor<L>(x, prim); });
}

int grim_full_step_diagnostics(grim_ctx* x, double* prim) {
  GRIM_REQUIRE(x && prim && x->torus, "grim_full_step_diagnostics: no torus (grim_init_torus) or null prim");
  return withLayout(x, [&]<class L>() -> int {
    if (int rc = scrubNaNs(x, prim, L::DOF)) return rc;
    return applyFloor<L>(x, prim);
  });
}

int grim_torus_bcs(grim_ctx* x, double* prim) {
  GRIM_REQUIRE(x && prim && x->torus, "grim_torus_bcs: no torus (grim_init_torus) or null prim");
  return withLayout(x, [&]<class L>() -> int { return torusBCs<L>(x, prim); });
}

int grim_torus_observers(grim_ctx* x, const double* prim, grim_torus_observables* out) {
  GRIM_REQUIRE(x && out, "grim_torus_observers: null argument");
  if (!prim) prim = x->primOld;
  if (int rc = withLayout(x, [&]<class L>() -> int { return observers<L>(x, prim); })) return rc;
  GRIM_HIP_TRY(hipMemcpyAsync(x->hostBuf, x->obs + (size_t)kObsBlocks * OBS_WORDS,
                              sizeof(double) * OBS_WORDS, hipMemcpyDeviceToHost, x->s));
  GRIM_HIP_TRY(hipStreamSynchronize(x->s));
  const double* r = x->hostBuf;
  out->baryon_mass = r[OBS_MASS];
  out->magnetic_energy = r[OBS_EMAG];
  out->thermal_energy = r[OBS_ETH];
  out->mdot_in = r[OBS_MDOT_IN];
  out->mdot_out = r[OBS_MDOT_OUT];
  out->mdot_out_unbound = r[OBS_MDOT_UNBOUND];
  out->mdot_out_relativistic = r[OBS_MDOT_REL];
  out->rho_max = r[OBS_RHO_MAX];
  out->beta_min = r[OBS_BETA_MIN];
  return GRIM_OK;
}

int grim_dump_vars(grim_ctx* x, const double* prim, double* out, int on_device) {
  GRIM_REQUIRE(x && out, "grim_dump_vars: null argument");
  if (!prim) prim = x->primOld;
  const size_t bytes = sizeof(double) * (size_t)(x->dof + 5) * (size_t)bulkBox(x->D).count();
  if (!on_device && !x->dump) GRIM_HIP_TRY(hipMalloc(&x->dump, bytes));
  double* dst = on_device ? out : x->dump;
  if (int rc = withLayout(x, [&]<class L>() -> int { return dumpVars<L>(x, prim, dst); })) return rc;
  if (on_device) return GRIM_OK;
  GRIM_HIP_TRY(hipMemcpyAsync(out, x->dump, bytes, hipMemcpyDeviceToHost, x->s));
  GRIM_HIP_TRY(hipStreamSynchronize(x->s));
  return GRIM_OK;
}

int grim_set_ghost_state(grim_ctx* x, const double* prim_ic, int on_device) {
  GRIM_REQUIRE(x, "grim_set_ghost_state: null context");
  // dof planes of the state, then the on/off flag (one double); the buffer
  // stays allocated until grim_destroy, so steps captured while the hook was
  // on keep reading valid memory after it is switched off
  const size_t bytes = sizeof(double) * (size_t)x->dof * x->D.n;
  if (!prim_ic && !x->ghostState) return GRIM_OK;
  if (!x->ghostState) GRIM_HIP_TRY(hipMalloc(&x->ghostState, bytes + sizeof(double)));
  x->hostBuf[0] = prim_ic ? 1. : 0.;
  if (prim_ic)
    GRIM_HIP_TRY(hipMemcpyAsync(x->ghostState, prim_ic, bytes,
                                on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, x->s));
  GRIM_HIP_TRY(hipMemcpyAsync(x->ghostState + (size_t)x->dof * x->D.n, x->hostBuf, sizeof(double),
                              hipMemcpyHostToDevice, x->s));
  GRIM_HIP_TRY(hipStreamSynchronize(x->s));
  return GRIM_OK;
}

int grim_set_x1_side_state(grim_ctx* x, const double* prim_ic, int on_device) {
  GRIM_REQUIRE(x, "grim_set_x1_side_state: null context");
  // rho, u, u~^1 planes, then the on/off flag, kept until grim_destroy as
  // the ghost state's
  const size_t bytes = sizeof(double) * 3 * (size_t)x->D.n;
  if (!prim_ic && !x->sideState) return GRIM_OK;
  if (!x->sideState) GRIM_HIP_TRY(hipMalloc(&x->sideState, bytes + sizeof(double)));
  x->hostBuf[0] = prim_ic ? 1. : 0.;
  if (prim_ic)
    GRIM_HIP_TRY(hipMemcpyAsync(x->sideState, prim_ic, bytes,
                                on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, x->s));
  GRIM_HIP_TRY(hipMemcpyAsync(x->sideState + 3 * (size_t)x->D.n, x->hostBuf, sizeof(double),
                              hipMemcpyHostToDevice, x->s));
  GRIM_HIP_TRY(hipStreamSynchronize(x->s));
  return GRIM_OK;
}

int grim_init_linear_mode(grim_ctx* x, int mode, double amplitude) {
  GRIM_REQUIRE(x && (mode == 0 || mode == 1), "grim_init_linear_mode: bad argument");
  GRIM_REQUIRE(x->c.metric == GRIM_METRIC_MINKOWSKI, "grim_init_linear_mode: Minkowski grids only");
  GRIM_REQUIRE(mode == 1 || !(x->p.conduction || x->p.viscosity),
               "grim_init_linear_mode: the fast mode is an ideal-MHD eigenvector");
  return withLayout(x, [&]<class L>() -> int {
    linearModeKernel<L><<<blocksFor(x->D.n), kBlock, 0, x->s>>>(x->a, mode, amplitude, x->geomC, x->cg,
                                                                 x->primOld);
    return checkLaunch("grim_init_linear_mode");
  });
}

}  // extern "C"
