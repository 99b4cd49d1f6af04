// Synthetic isolated-zone batches on modified Kerr-Schild geometry.
//
// A batch stands in for one stage of grim's implicit solve with every zone
// independent (the flux divergence and explicit sources are frozen during
// solve(), src/timestepper/timestep.cpp:47-66).  Sampling rule, per zone z,
// with u_k(z) the k-th uniform of the counter-based stream below:
//   r      = exp(X1), X1 uniform in [ln 1.5, ln 40]
//   theta  = pi (0.05 + 0.9 u_1), X2 = inverse of grim's
//            theta(X2) = pi X2 + (1-h)/2 sin(2 pi X2) (12 Newton steps)
//   rho    = 10^(-5 + 5 u_2)          (EMHD: 1e-8 (1 + u_2) when u_21 < 0.1)
//   u      = rho 10^(-3 + 4 u_3)
//   utilde = sqrt(W^2 - 1) L^-T n,  W = 1 + 9 u_4, n uniform on S^2 (u_5, u_6),
//            g_ij = L L^T (Cholesky of the spatial metric)
//   B      = s L^-T m, m uniform on S^2 (u_7, u_8), s such that
//            b^2/rho = 10^(-3 + (log10 50 + 3) u_9)
//   EMHD:  qTilde = 0.1 (2 u_17 - 1), dPTilde = 0.1 (2 u_18 - 1),
//          closure alphas 10^(-2 + 3 u_19), 10^(-2 + 3 u_20),
//          old state = root * (1 + 1e-2 (2 u_{22+v} - 1)) per fluid variable
//   guess  = root * (1 + 1e-3 (2 u_{10+v} - 1)) per fluid variable.
// RNG: u_k(z) = (splitmix64(seed + G (32 z + k + 1)) >> 11) * 2^-53 with
// G = 0x9E3779B97F4A7C15; integer-exact, so host and device draw the same
// uniforms.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace grim {

__host__ __device__ inline uint64_t splitmix64(uint64_t x) {
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

__host__ __device__ inline double uniform01(uint64_t seed, uint64_t z, int k) {
  const uint64_t x = seed + 0x9E3779B97F4A7C15ull * (32ull * z + (uint64_t)(k + 1));
  return (double)(splitmix64(x) >> 11) * (1.0 / 9007199254740992.0);
}

// grim's setgCovInXCoords for MKS with r = exp(X1), theta = GammieTheta(X2),
// no cylindrification (src/geometry/geometry.cpp:277-420).
__device__ inline void mksMetric(double X1, double X2, double a, double h, double G[4][4]) {
  const double r = exp(X1);
  const double theta = M_PI * X2 + 0.5 * (1. - h) * sin(2. * M_PI * X2);
  const double dr = r;
  const double dth = M_PI + M_PI * (1. - h) * cos(2. * M_PI * X2);
  const double ct = cos(theta), st = sin(theta);
  const double sigma = r * r + (a * ct) * (a * ct);
  const double s2 = st * st;
  for (int m = 0; m < 4; m++)
    for (int k = 0; k < 4; k++) G[m][k] = 0.;
  G[0][0] = -(1. - 2. * r / sigma);
  G[0][1] = G[1][0] = (2. * r / sigma) * dr;
  G[0][3] = G[3][0] = -(2. * a * r * s2 / sigma);
  G[1][1] = (1. + 2. * r / sigma) * dr * dr;
  G[1][3] = G[3][1] = -a * (1. + 2. * r / sigma) * s2 * dr;
  G[2][2] = sigma * dth * dth;
  G[3][3] = s2 * (sigma + (a * st) * (a * st) * (1. + 2. * r / sigma));
}

// Inverse (first row) and determinant of a symmetric 4x4 by cofactors.
__device__ inline void gconRow0(const double G[4][4], double* det, double c0[4]) {
  double cof[4];
  for (int j = 0; j < 4; j++) {
    int rr[3], cc[3];
    for (int k = 0, m = 0; k < 4; k++)
      if (k != j) rr[m++] = k;
    for (int k = 0, m = 0; k < 4; k++)
      if (k != 0) cc[m++] = k;
    const double m3 =
        G[rr[0]][cc[0]] * (G[rr[1]][cc[1]] * G[rr[2]][cc[2]] - G[rr[1]][cc[2]] * G[rr[2]][cc[1]]) -
        G[rr[0]][cc[1]] * (G[rr[1]][cc[0]] * G[rr[2]][cc[2]] - G[rr[1]][cc[2]] * G[rr[2]][cc[0]]) +
        G[rr[0]][cc[2]] * (G[rr[1]][cc[0]] * G[rr[2]][cc[1]] - G[rr[1]][cc[1]] * G[rr[2]][cc[0]]);
    cof[j] = (j & 1) ? -m3 : m3;
  }
  // expansion of det along row 0 (cofactor of (0,j) of a symmetric matrix)
  double d = 0.;
  for (int j = 0; j < 4; j++) d += G[0][j] * cof[j];
  *det = d;
  for (int j = 0; j < 4; j++) c0[j] = cof[j] / d;
}

// Rows of L^-T with g_ij = L L^T, i,j = 1..3, applied to a unit vector.
__device__ inline void frameVector(const double G[4][4], const double nvec[3], double out[3]) {
  const double l00 = sqrt(G[1][1]);
  const double l10 = G[2][1] / l00;
  const double l20 = G[3][1] / l00;
  const double l11 = sqrt(G[2][2] - l10 * l10);
  const double l21 = (G[3][2] - l20 * l10) / l11;
  const double l22 = sqrt(G[3][3] - l20 * l20 - l21 * l21);
  // solve L^T x = n (back substitution)
  out[2] = nvec[2] / l22;
  out[1] = (nvec[1] - l21 * out[2]) / l11;
  out[0] = (nvec[0] - l10 * out[1] - l20 * out[2]) / l00;
}

__device__ inline void unitVector(double uc, double up, double v[3]) {
  const double ct = 2. * uc - 1.;
  const double st = sqrt(fmax(0., 1. - ct * ct));
  const double ph = 2. * M_PI * up;
  v[0] = st * cos(ph);
  v[1] = st * sin(ph);
  v[2] = ct;
}

}  // namespace grim
