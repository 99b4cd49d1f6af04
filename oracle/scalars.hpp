// Special scalar types for the oracle restatement (TEST INFRASTRUCTURE ONLY).
//
//  Dual  - forward-mode dual number with GRIM_ORACLE_MAX_VARS tangents; the
//          restatement instantiated on it yields exact Jacobians.
//  Count - a double that tallies every arithmetic operation it performs
//          (add/sub, mul, div, sqrt, other transcendental) into thread-local
//          counters; gives the algorithmic operation counts of one residual
//          evaluation and of one LU solve.
//  Traced - a double that records every operation on a thread-local tape
//          (kind and operands).  Counting only the operations that the
//          outputs depend on, and that depend on a variable input, gives the
//          work a compiled evaluation must repeat per call: operations whose
//          result nothing reads are dead code, and operations on per-zone
//          constants alone can be formed once per zone.
#pragma once
#include <cmath>
#include <vector>

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

// -------------------------------------------------------------- Traced ----
// Operation kinds of the tape; kFree (negation, |x|) costs nothing on the
// GPU (a source modifier) and is not counted.
enum TapeKind { kAdd = 0, kMul = 1, kDiv = 2, kSqrt = 3, kOther = 4, kSelect = 5, kFree = 6, kLeaf = 7 };
struct TapeNode {
  int kind;
  long long a, b;  // operand ids (-1: a constant)
  bool var;        // depends on a variable leaf
};
struct Tape {
  long long base = 0;  // ids below base belong to an earlier recording: constants
  std::vector<TapeNode> nodes;
  void reset() {
    base += (long long)nodes.size();
    nodes.clear();
  }
  const TapeNode* at(long long id) const {
    return (id >= base && id - base < (long long)nodes.size()) ? &nodes[id - base] : nullptr;
  }
};
extern thread_local Tape g_tape;

struct Traced {
  double v;
  long long id;
  Traced() : v(0.), id(-1) {}
  Traced(double x) : v(x), id(-1) {}  // NOLINT
  static Traced variable(double x) {
    Traced r(x);
    g_tape.nodes.push_back({kLeaf, -1, -1, true});
    r.id = g_tape.base + (long long)g_tape.nodes.size() - 1;
    return r;
  }
};
inline double oval(const Traced& x) { return x.v; }
inline Traced traced(int kind, double v, const Traced& a, const Traced& b) {
  const TapeNode* na = g_tape.at(a.id);
  const TapeNode* nb = g_tape.at(b.id);
  Traced r(v);
  g_tape.nodes.push_back({kind, na ? a.id : -1, nb ? b.id : -1, (na && na->var) || (nb && nb->var)});
  r.id = g_tape.base + (long long)g_tape.nodes.size() - 1;
  return r;
}
inline Traced operator+(const Traced& a, const Traced& b) { return traced(kAdd, a.v + b.v, a, b); }
inline Traced operator-(const Traced& a, const Traced& b) { return traced(kAdd, a.v - b.v, a, b); }
inline Traced operator-(const Traced& a) { return traced(kFree, -a.v, a, Traced()); }
inline Traced operator*(const Traced& a, const Traced& b) { return traced(kMul, a.v * b.v, a, b); }
inline Traced operator/(const Traced& a, const Traced& b) { return traced(kDiv, a.v / b.v, a, b); }
inline Traced operator+(const Traced& a, double b) { return a + Traced(b); }
inline Traced operator+(double a, const Traced& b) { return Traced(a) + b; }
inline Traced operator-(const Traced& a, double b) { return a - Traced(b); }
inline Traced operator-(double a, const Traced& b) { return Traced(a) - b; }
inline Traced operator*(const Traced& a, double b) { return a * Traced(b); }
inline Traced operator*(double a, const Traced& b) { return Traced(a) * b; }
inline Traced operator/(const Traced& a, double b) { return a / Traced(b); }
inline Traced operator/(double a, const Traced& b) { return Traced(a) / b; }
inline Traced& operator+=(Traced& a, const Traced& b) { return a = a + b; }
inline Traced& operator-=(Traced& a, const Traced& b) { return a = a - b; }
inline Traced& operator*=(Traced& a, const Traced& b) { return a = a * b; }
inline Traced sqrt(const Traced& a) { return traced(kSqrt, std::sqrt(a.v), a, Traced()); }
inline Traced exp(const Traced& a) { return traced(kOther, std::exp(a.v), a, Traced()); }
inline Traced pow(const Traced& a, double b) { return traced(kOther, std::pow(a.v, b), a, Traced()); }
inline Traced abs(const Traced& a) { return traced(kFree, std::fabs(a.v), a, Traced()); }
inline Traced amax(const Traced& a, const Traced& b) { return traced(kSelect, a.v > b.v ? a.v : b.v, a, b); }
inline Traced amin(const Traced& a, const Traced& b) { return traced(kSelect, a.v < b.v ? a.v : b.v, a, b); }
inline Traced amax(const Traced& a, double b) { return amax(a, Traced(b)); }
inline Traced amin(const Traced& a, double b) { return amin(a, Traced(b)); }

// Operations of the current recording that the given outputs depend on and
// that depend on a variable leaf, by kind (add, mul, div, sqrt, other,
// select).
inline void liveCounts(const Traced* out, int nout, long long counts[6]) {
  for (int k = 0; k < 6; k++) counts[k] = 0;
  std::vector<char> seen(g_tape.nodes.size(), 0);
  std::vector<long long> stack;
  for (int i = 0; i < nout; i++)
    if (g_tape.at(out[i].id)) stack.push_back(out[i].id);
  while (!stack.empty()) {
    const long long id = stack.back();
    stack.pop_back();
    const TapeNode* nd = g_tape.at(id);
    if (!nd || seen[id - g_tape.base]) continue;
    seen[id - g_tape.base] = 1;
    if (!nd->var) continue;  // a per-zone constant: formed once, not per evaluation
    if (nd->kind < 6) counts[nd->kind]++;
    if (nd->a >= 0) stack.push_back(nd->a);
    if (nd->b >= 0) stack.push_back(nd->b);
  }
}

}  // namespace oracle
