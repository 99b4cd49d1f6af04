// Zone-batch utilities: conserved variables, known-root right-hand sides,
// the synthetic Kerr-Schild batch generator and a convergence count.
#include <hip/hip_runtime.h>

#include "batchgen.hpp"
#include "grim_hip.h"
#include "runtime.hpp"
#include "zone_physics.hpp"

namespace grim {

constexpr int kBlock = 256;

// fluidElement::set + computeFluxes(0) for every zone (all dof rows, the B
// rows being g (b^i u^0 - b^0 u^i), physics.cpp:201-203).
template <class L>
__global__ __launch_bounds__(kBlock) void conservedKernel(DevParams prm, int64_t n,
                                                         const double* prim, const double* geom,
                                                         const double* closureA,
                                                         const double* radius, double* cons) {
  const int64_t z = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (z >= n) return;
  ZoneGeom G;
  loadGeom(geom, n, z, G);
  ZoneConst<L> c;  // only the closure inputs are read by residualPieces' U
  c.alphaC = closureA ? closureA[z] : 1.;
  c.alphaV = closureA ? closureA[n + z] : 1.;
  c.radius = radius ? radius[z] : 0.;
#pragma unroll
  for (int mu = 0; mu < 4; mu++) { c.oldUCov[mu] = 0.; c.cBCon[mu] = 0.; }
  c.oldT = c.cRho = c.cT = c.cTau = c.cChi = c.cNu = c.cBSqr = c.cBNorm = c.cUCon0 = 1.;
  c.cQTilde = c.cDPTilde = 0.;
  double P[L::DOF], U[L::NF], Si[L::NF], Std[L::NF];
#pragma unroll
  for (int v = 0; v < L::DOF; v++) P[v] = prim[v * n + z];
  residualPieces<L>(P, G, c, prm, 1., U, Si, Std);
#pragma unroll
  for (int v = 0; v < L::NF; v++) cons[v * n + z] = U[v];
  Kinematics k;
  kinematics<L>(P, G, prm, k);
  cons[L::B1 * n + z] = G.g * (k.bCon[1] * k.uCon[0] - k.bCon[0] * k.uCon[1]);
  cons[L::B2 * n + z] = G.g * (k.bCon[2] * k.uCon[0] - k.bCon[0] * k.uCon[2]);
  cons[L::B3 * n + z] = G.g * (k.bCon[3] * k.uCon[0] - k.bCon[0] * k.uCon[3]);
}

// rhs = U(P) + dt (1/2 Si + Std): the folded right-hand side for which P is
// an exact root of the batched residual; mask = 1/max_v |rhs_v|.
template <class L>
__global__ __launch_bounds__(kBlock) void knownRootKernel(DevParams prm, grim_zone_batch b,
                                                         double* rhs, double* mask) {
  const int64_t z = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (z >= b.n) return;
  const int64_t n = b.n;
  ZoneGeom G;
  loadBatchGeom(b, z, G);
  ZoneConst<L> c;
  loadZoneConst<L>(b, prm, G, z, c);
  double P[L::DOF], U[L::NF], Si[L::NF], Std[L::NF];
#pragma unroll
  for (int v = 0; v < L::DOF; v++) P[v] = b.prim[v * n + z];
  residualPieces<L>(P, G, c, prm, b.dt_prime, U, Si, Std);
  double m = 0.;
#pragma unroll
  for (int v = 0; v < L::NF; v++) {
    const bool emhdRow = (L::COND_ON && v == L::Q) || (L::VISC_ON && v == L::DP);
    const double r = emhdRow ? U[v] + b.dt_prime * (0.5 * Si[v] + Std[v]) : U[v];
    rhs[v * n + z] = r;
    m = fmax(m, fabs(r));
  }
  if (mask) mask[z] = 1. / m;
}

template <class L>
__global__ __launch_bounds__(kBlock) void generateKernel(DevParams prm, int64_t n, uint64_t seed,
                                                        double a, double h, double* geom,
                                                        double* star, double* guess,
                                                        double* old, double* closureA) {
  const int64_t z = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (z >= n) return;
  auto u = [&](int k) { return uniform01(seed, (uint64_t)z, k); };
  const double X1 = log(1.5) + u(0) * (log(40.) - log(1.5));
  const double theta = M_PI * (0.05 + 0.9 * u(1));
  double X2 = theta / M_PI;
  for (int i = 0; i < 12; i++) {
    const double f = M_PI * X2 + 0.5 * (1. - h) * sin(2. * M_PI * X2) - theta;
    const double df = M_PI + M_PI * (1. - h) * cos(2. * M_PI * X2);
    X2 = X2 - f / df;
  }
  double Gm[4][4];
  mksMetric(X1, X2, a, h, Gm);
  double det, c0[4];
  gconRow0(Gm, &det, c0);
  const double alpha = 1. / sqrt(-c0[0]);
  const double g = sqrt(-det);
  const int cov[10][2] = {{0, 0}, {0, 1}, {0, 2}, {0, 3}, {1, 1}, {1, 2}, {1, 3}, {2, 2}, {2, 3}, {3, 3}};
  for (int k = 0; k < 10; k++) geom[k * n + z] = Gm[cov[k][0]][cov[k][1]];
  for (int k = 0; k < 4; k++) geom[(GEOM_GCON0 + k) * n + z] = c0[k];
  geom[GEOM_ALPHA * n + z] = alpha;
  geom[GEOM_G * n + z] = g;

  double rho = pow(10., -5. + 5. * u(2));
  if (L::EMHD && u(21) < 0.1) rho = 1e-8 * (1. + u(2));
  const double uu = rho * pow(10., -3. + 4. * u(3));
  const double W = 1. + 9. * u(4);
  double nv[3], e[3], ut[3];
  unitVector(u(5), u(6), nv);
  frameVector(Gm, nv, e);
  const double sw = sqrt(W * W - 1.);
  for (int i = 0; i < 3; i++) ut[i] = sw * e[i];

  // b^2 of a unit-norm field direction: (B.B + (B^i u_i)^2)/(u^0)^2
  double lor2 = 1.;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) lor2 += Gm[i + 1][j + 1] * ut[i] * ut[j];
  const double lor = sqrt(lor2);
  double uCon[4];
  uCon[0] = lor / alpha;
  for (int i = 0; i < 3; i++) uCon[i + 1] = ut[i] - lor * c0[i + 1] * alpha;
  double uCov[4];
  for (int m = 0; m < 4; m++) {
    uCov[m] = 0.;
    for (int k = 0; k < 4; k++) uCov[m] += Gm[m][k] * uCon[k];
  }
  double mv[3], eB[3];
  unitVector(u(7), u(8), mv);
  frameVector(Gm, mv, eB);
  const double Bu = eB[0] * uCov[1] + eB[1] * uCov[2] + eB[2] * uCov[3];
  const double bsqUnit = (1. + Bu * Bu) / (uCon[0] * uCon[0]);
  const double sigma = pow(10., -3. + (log10(50.) + 3.) * u(9));
  const double s = sqrt(sigma * rho / bsqUnit);

  double P[L::DOF];
  P[RHO] = rho;
  P[UU] = uu;
  P[U1] = ut[0];
  P[U2] = ut[1];
  P[U3] = ut[2];
  if constexpr (L::COND_ON) P[L::Q] = 0.1 * (2. * u(17) - 1.);
  if constexpr (L::VISC_ON) P[L::DP] = 0.1 * (2. * u(18) - 1.);
  P[L::B1] = s * eB[0];
  P[L::B2] = s * eB[1];
  P[L::B3] = s * eB[2];
  for (int v = 0; v < L::DOF; v++) {
    star[v * n + z] = P[v];
    guess[v * n + z] = (v < L::NF) ? P[v] * (1. + 1e-3 * (2. * u(10 + v) - 1.)) : P[v];
  }
  if constexpr (L::EMHD) {
    for (int v = 0; v < L::DOF; v++)
      old[v * n + z] = (v < L::NF) ? P[v] * (1. + 1e-2 * (2. * u(22 + v) - 1.)) : P[v];
    closureA[z] = pow(10., -2. + 3. * u(19));
    closureA[n + z] = pow(10., -2. + 3. * u(20));
  }
}

__global__ void countAboveKernel(int64_t n, const double* l2, double tol,
                                 unsigned long long* out) {
  const int64_t z = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool above = (z < n) && !(l2[z] <= tol);
  const unsigned long long ballot = __ballot(above);
  if ((threadIdx.x & 63) == 0 && ballot) atomicAdd(out, (unsigned long long)__popcll(ballot));
}

__global__ void zeroU64(unsigned long long* p) { *p = 0ull; }

}  // namespace grim

using namespace grim;

extern "C" {

int grim_conserved(const grim_params* p, int64_t n, const double* prim, const double* geom,
                   const double* closure, const double* radius, double* cons,
                   grim_stream_t stream) {
  GRIM_REQUIRE(validParams(p) && n >= 0, "grim_conserved: bad arguments");
  if (n == 0) return GRIM_OK;
  GRIM_REQUIRE(prim && geom && cons, "grim_conserved: null pointer");
  if (int rc = checkDevice()) return rc;
  hipStream_t s = (hipStream_t)stream;
  const DevParams prm = makeDevParams(*p);
  return dispatchLayout(*p, [&]<class L>() -> int {
    conservedKernel<L><<<gridFor(n, kBlock), kBlock, 0, s>>>(prm, n, prim, geom, closure, radius, cons);
    return checkLaunch("grim_conserved");
  });
}

int grim_known_root_rhs(const grim_params* p, const grim_zone_batch* b, double* rhs, double* mask,
                        grim_stream_t stream) {
  GRIM_REQUIRE(validParams(p) && b && b->n >= 0, "grim_known_root_rhs: bad arguments");
  if (b->n == 0) return GRIM_OK;
  GRIM_REQUIRE(b->prim && b->geom && rhs, "grim_known_root_rhs: null pointer");
  GRIM_REQUIRE(!(p->conduction || p->viscosity) || b->prim_old,
               "grim_known_root_rhs: EMHD needs prim_old");
  if (int rc = checkDevice()) return rc;
  hipStream_t s = (hipStream_t)stream;
  const DevParams prm = makeDevParams(*p);
  grim_zone_batch bb = *b;
  bb.rhs = rhs;  // read by loadZoneConst; overwritten below
  bb.residual_mask = nullptr;
  return dispatchLayout(*p, [&]<class L>() -> int {
    knownRootKernel<L><<<gridFor(bb.n, kBlock), kBlock, 0, s>>>(prm, bb, rhs, mask);
    return checkLaunch("grim_known_root_rhs");
  });
}

int grim_generate_kerr_schild_batch(const grim_params* p, int64_t n, uint64_t seed, double a,
                                    double h, double* geom, double* prim_star, double* guess,
                                    double* prim_old, double* closure, grim_stream_t stream) {
  GRIM_REQUIRE(validParams(p) && n >= 0, "grim_generate_kerr_schild_batch: bad arguments");
  if (n == 0) return GRIM_OK;
  GRIM_REQUIRE(geom && prim_star && guess, "grim_generate_kerr_schild_batch: null pointer");
  GRIM_REQUIRE(!(p->conduction || p->viscosity) || (prim_old && closure),
               "grim_generate_kerr_schild_batch: EMHD needs prim_old and closure");
  GRIM_REQUIRE(fabs(a) < 1. && h > 0. && h <= 1., "grim_generate_kerr_schild_batch: bad a or h");
  if (int rc = checkDevice()) return rc;
  hipStream_t s = (hipStream_t)stream;
  const DevParams prm = makeDevParams(*p);
  return dispatchLayout(*p, [&]<class L>() -> int {
    generateKernel<L><<<gridFor(n, kBlock), kBlock, 0, s>>>(prm, n, seed, a, h, geom, prim_star,
                                                           guess, prim_old, closure);
    return checkLaunch("grim_generate_kerr_schild_batch");
  });
}

int grim_count_above(int64_t n, const double* final_l2, double tol, int64_t* count,
                     grim_stream_t stream) {
  GRIM_REQUIRE(n >= 0 && count, "grim_count_above: bad arguments");
  *count = 0;
  if (n == 0) return GRIM_OK;
  GRIM_REQUIRE(final_l2, "grim_count_above: null pointer");
  if (int rc = checkDevice()) return rc;
  hipStream_t s = (hipStream_t)stream;
  unsigned long long* d = nullptr;
  GRIM_HIP_TRY(hipMallocAsync((void**)&d, sizeof(unsigned long long), s));
  zeroU64<<<1, 1, 0, s>>>(d);
  countAboveKernel<<<gridFor(n, kBlock), kBlock, 0, s>>>(n, final_l2, tol, d);
  unsigned long long h = 0;
  hipError_t e = hipMemcpyAsync(&h, d, sizeof(h), hipMemcpyDeviceToHost, s);
  hipError_t e2 = hipFreeAsync(d, s);
  hipError_t e3 = hipStreamSynchronize(s);
  if (e != hipSuccess || e2 != hipSuccess || e3 != hipSuccess) {
    setError(std::string("grim_count_above: ") +
             hipGetErrorString(e != hipSuccess ? e : (e2 != hipSuccess ? e2 : e3)));
    return GRIM_ERR_HIP;
  }
  *count = (int64_t)h;
  return GRIM_OK;
}

}  // extern "C"
