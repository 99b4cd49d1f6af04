// Special scalar types for the oracle restatement (TEST INFRASTRUCTURE ONLY).
//
//  Dual  - forward-mode dual number with GRIM_ORACLE_MAX_VARS tangents; the
//          restatement instantiated on it yields exact Jacobians.
//  Count - a double that tallies every arithmetic operation it performs
//          (add/sub, mul, div, sqrt, other transcendental) into thread-local
//          counters; gives the algorithmic operation counts of one residual
//          evaluation and of one LU solve.
#pragma once
#include <cmath>

#include "grim_oracle.h"

namespace oracle {

// ---------------------------------------------------------------- Dual ----
struct Dual {
  double v;
  double d[GRIM_ORACLE_MAX_VARS];
  Dual() : v(0.) { for (double& x : d) x = 0.; }
  Dual(double x) : v(x) { for (double& y : d) y = 0.; }  // NOLINT
  static Dual seed(double x, int k) {
    Dual r(x);
    r.d[k] = 1.;
    return r;
  }
};
inline double oval(const Dual& x) { return x.v; }
inline Dual operator+(const Dual& a, const Dual& b) {
  Dual r(a.v + b.v);
  for (int k = 0; k < GRIM_ORACLE_MAX_VARS; k++) r.d[k] = a.d[k] + b.d[k];
  return r;
}
inline Dual operator-(const Dual& a, const Dual& b) {
  Dual r(a.v - b.v);
  for (int k = 0; k < GRIM_ORACLE_MAX_VARS; k++) r.d[k] = a.d[k] - b.d[k];
  return r;
}
inline Dual operator-(const Dual& a) {
  Dual r(-a.v);
  for (int k = 0; k < GRIM_ORACLE_MAX_VARS; k++) r.d[k] = -a.d[k];
  return r;
}
inline Dual operator*(const Dual& a, const Dual& b) {
  Dual r(a.v * b.v);
  for (int k = 0; k < GRIM_ORACLE_MAX_VARS; k++) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
  return r;
}
inline Dual operator/(const Dual& a, const Dual& b) {
  Dual r(a.v / b.v);
  for (int k = 0; k < GRIM_ORACLE_MAX_VARS; k++)
    r.d[k] = (a.d[k] * b.v - a.v * b.d[k]) / (b.v * b.v);
  return r;
}
inline Dual operator+(const Dual& a, double b) { return a + Dual(b); }
inline Dual operator+(double a, const Dual& b) { return Dual(a) + b; }
inline Dual operator-(const Dual& a, double b) { return a - Dual(b); }
inline Dual operator-(double a, const Dual& b) { return Dual(a) - b; }
inline Dual operator*(const Dual& a, double b) { return a * Dual(b); }
inline Dual operator*(double a, const Dual& b) { return Dual(a) * b; }
inline Dual operator/(const Dual& a, double b) { return a / Dual(b); }
inline Dual operator/(double a, const Dual& b) { return Dual(a) / b; }
inline Dual& operator+=(Dual& a, const Dual& b) { return a = a + b; }
inline Dual& operator-=(Dual& a, const Dual& b) { return a = a - b; }
inline Dual& operator*=(Dual& a, const Dual& b) { return a = a * b; }
inline Dual sqrt(const Dual& a) {
  double s = std::sqrt(a.v);
  Dual r(s);
  for (int k = 0; k < GRIM_ORACLE_MAX_VARS; k++) r.d[k] = a.d[k] / (2. * s);
  return r;
}
inline Dual exp(const Dual& a) {
  double e = std::exp(a.v);
  Dual r(e);
  for (int k = 0; k < GRIM_ORACLE_MAX_VARS; k++) r.d[k] = a.d[k] * e;
  return r;
}
inline Dual pow(const Dual& a, double b) {
  double pv = std::pow(a.v, b);
  Dual r(pv);
  double dp = b * std::pow(a.v, b - 1.);
  for (int k = 0; k < GRIM_ORACLE_MAX_VARS; k++) r.d[k] = a.d[k] * dp;
  return r;
}
inline Dual abs(const Dual& a) { return a.v < 0. ? -a : a; }
// One-sided derivative semantics: the selected operand's tangent.
inline Dual amax(const Dual& a, double b) { return a.v > b ? a : Dual(b); }
inline Dual amin(const Dual& a, double b) { return a.v < b ? a : Dual(b); }
inline Dual amax(const Dual& a, const Dual& b) { return a.v > b.v ? a : b; }
inline Dual amin(const Dual& a, const Dual& b) { return a.v < b.v ? a : b; }

// --------------------------------------------------------------- Count ----
struct OpCounts {
  long long add = 0, mul = 0, div = 0, sqrt = 0, other = 0;
};
extern thread_local OpCounts g_ops;

struct Count {
  double v;
  Count() : v(0.) {}
  Count(double x) : v(x) {}  // NOLINT
};
inline double oval(const Count& x) { return x.v; }
inline Count operator+(const Count& a, const Count& b) { g_ops.add++; return Count(a.v + b.v); }
inline Count operator-(const Count& a, const Count& b) { g_ops.add++; return Count(a.v - b.v); }
inline Count operator-(const Count& a) { return Count(-a.v); }
inline Count operator*(const Count& a, const Count& b) { g_ops.mul++; return Count(a.v * b.v); }
inline Count operator/(const Count& a, const Count& b) { g_ops.div++; return Count(a.v / b.v); }
inline Count operator+(const Count& a, double b) { return a + Count(b); }
inline Count operator+(double a, const Count& b) { return Count(a) + b; }
inline Count operator-(const Count& a, double b) { return a - Count(b); }
inline Count operator-(double a, const Count& b) { return Count(a) - b; }
inline Count operator*(const Count& a, double b) { return a * Count(b); }
inline Count operator*(double a, const Count& b) { return Count(a) * b; }
inline Count operator/(const Count& a, double b) { return a / Count(b); }
inline Count operator/(double a, const Count& b) { return Count(a) / b; }
inline Count& operator+=(Count& a, const Count& b) { return a = a + b; }
inline Count& operator-=(Count& a, const Count& b) { return a = a - b; }
inline Count& operator*=(Count& a, const Count& b) { return a = a * b; }
inline Count sqrt(const Count& a) { g_ops.sqrt++; return Count(std::sqrt(a.v)); }
inline Count exp(const Count& a) { g_ops.other++; return Count(std::exp(a.v)); }
inline Count pow(const Count& a, double b) { g_ops.other++; return Count(std::pow(a.v, b)); }
inline Count abs(const Count& a) { return Count(std::fabs(a.v)); }
inline Count amax(const Count& a, double b) { return a.v > b ? a : Count(b); }
inline Count amin(const Count& a, double b) { return a.v < b ? a : Count(b); }
inline Count amax(const Count& a, const Count& b) { return a.v > b.v ? a : b; }
inline Count amin(const Count& a, const Count& b) { return a.v < b.v ? a : b; }

}  // namespace oracle
