// Host-side runtime helpers shared by the C-ABI entry points: thread-local
// error text, HIP status checks, the gfx950 device check and the dispatch of
// the (conduction, viscosity) compile-time layouts.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdio>
#include <string>

#include "grim_hip.h"
#include "zone_physics.hpp"

namespace grim {

void setError(const std::string& msg);

// Returns GRIM_OK when the current device is a gfx950 part (checked once per
// device per process), GRIM_ERR_NO_DEVICE otherwise.
int checkDevice();

// Compute units of the current device (after checkDevice; 256 if unknown).
int deviceCUs();

#define GRIM_HIP_TRY(expr)                                                       \
  do {                                                                           \
    hipError_t e_ = (expr);                                                      \
    if (e_ != hipSuccess) {                                                      \
      ::grim::setError(std::string(#expr) + ": " + hipGetErrorString(e_));       \
      return GRIM_ERR_HIP;                                                       \
    }                                                                            \
  } while (0)

#define GRIM_REQUIRE(cond, msg)                      \
  do {                                               \
    if (!(cond)) {                                   \
      ::grim::setError(msg);                         \
      return GRIM_ERR_BAD_ARG;                       \
    }                                                \
  } while (0)

inline int checkLaunch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    setError(std::string(what) + ": " + hipGetErrorString(e));
    return GRIM_ERR_HIP;
  }
  return GRIM_OK;
}

inline int validParams(const grim_params* p) {
  if (!p) return 0;
  if (p->conduction < 0 || p->conduction > 1 || p->viscosity < 0 || p->viscosity > 1) return 0;
  if (p->closure < GRIM_CLOSURE_LINEAR_MODES || p->closure > GRIM_CLOSURE_BUOYANCY) return 0;
  if (p->closure == GRIM_CLOSURE_BUOYANCY && p->viscosity) return 0;
  if (p->maxNonLinearIter < 0 || p->maxLineSearchIters < 0) return 0;
  return 1;
}

// Calls f.template operator()<Layout<c, v>>() for the params' EMHD switches.
template <class F>
int dispatchLayout(const grim_params& p, F&& f) {
  if (!p.conduction && !p.viscosity) return f.template operator()<Layout<false, false>>();
  if (p.conduction && !p.viscosity) return f.template operator()<Layout<true, false>>();
  if (!p.conduction && p.viscosity) return f.template operator()<Layout<false, true>>();
  return f.template operator()<Layout<true, true>>();
}

inline unsigned gridFor(int64_t n, int block) { return (unsigned)((n + block - 1) / block); }

}  // namespace grim
