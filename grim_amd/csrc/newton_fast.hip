// Batched per-zone Newton solve: the fused translation unit, compiled with
// -ffp-contract=fast-honor-pragmas so that a*b+c becomes one fused multiply-add
// (outside the fold arithmetic, which turns contraction off for itself).
//
// Holds the modes whose results are checked against tolerances rather than
// bit for bit: the mixed-precision path (FP64 residual and update, FP32
// analytic Jacobian and LU) and the fast FD-FP64 variant.
#include <hip/hip_runtime.h>

#include "grim_hip.h"
#include "newton_kernels.hpp"
#include "runtime.hpp"

namespace grim {

int launchNewtonFast(int mode, const grim_params& p, const grim_zone_batch& b,
                     const grim_newton_stats& st, const LaunchCfg& cfg, hipStream_t s, const FoldArgs* fold) {
  const DevParams prm = makeDevParams(p);
  return dispatchLayout(p, [&]<class L>() -> int {
    if (mode == MODE_FD_FAST)
      launchSolve<L, MODE_FD_FAST>(prm, b, st, cfg, s);
    else if (cfg.escalate)
      launchSolve<L, MODE_MIXED, MODE_ANALYTIC64F>(prm, b, st, cfg, s, fold);
    else
      launchSolve<L, MODE_MIXED>(prm, b, st, cfg, s, fold);
    return checkLaunch("grim_batched_newton (fused)");
  });
}

}  // namespace grim
