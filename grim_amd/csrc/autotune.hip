// Project autotuner of the batched per-zone solve.
//
// Times each candidate launch configuration on the caller's batch (on a
// scratch copy of the primitives, restored before every run) with HIP
// events and keeps the fastest.  Candidates: workgroup size 64/128/256;
// one pass, or the two-phase compacting solve with a phase-1 budget of
// 1-3 Newton updates and 2, 4 or 8 phase-1 wavefronts per phase-2
// wavefront.  The timing is the median of three runs.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "grim_hip.h"
#include "runtime.hpp"

extern "C" int grim_autotune_newton(const grim_params* p, const grim_zone_batch* b,
                                    grim_newton_opts* opts, grim_stream_t stream) {
  GRIM_REQUIRE(p && b && opts && b->n > 0 && b->prim, "grim_autotune_newton: bad arguments");
  if (int rc = grim::checkDevice()) return rc;
  hipStream_t s = (hipStream_t)stream;
  const int dof = 8 + p->conduction + p->viscosity;
  const size_t primBytes = sizeof(double) * (size_t)dof * (size_t)b->n;
  const int64_t wsBytes = grim_newton_workspace_bytes(b->n);
  void* prim = nullptr;
  void* ws = nullptr;
  GRIM_HIP_TRY(hipMalloc(&prim, primBytes));
  if (hipMalloc(&ws, (size_t)wsBytes) != hipSuccess) {
    (void)hipFree(prim);
    grim::setError("grim_autotune_newton: hipMalloc failed");
    return GRIM_ERR_OOM;
  }
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);

  std::vector<grim_newton_opts> cands;
  for (int block : {64, 128, 256}) {
    grim_newton_opts o = *opts;
    o.block_size = block;
    o.compaction = 0;
    cands.push_back(o);
    for (int k1 : {1, 2, 3}) {
      if (k1 >= p->maxNonLinearIter) continue;
      o.compaction = 2;
      o.phase1_iters = k1;
      o.phase2_group = 0;
      cands.push_back(o);
      for (int g : {2, 4, 8}) {
        o.compaction = 1;
        o.phase1_iters = k1;
        o.phase2_group = g;
        o.workspace = ws;
        o.workspace_bytes = wsBytes;
        cands.push_back(o);
      }
    }
  }
  grim_zone_batch bb = *b;
  bb.prim = (double*)prim;
  float best = 1e30f;
  grim_newton_opts bestOpts = *opts;
  int rc = GRIM_OK;
  for (const auto& o : cands) {
    std::vector<float> t;
    for (int rep = 0; rep < 3 && rc == GRIM_OK; rep++) {
      if (hipMemcpyAsync(prim, b->prim, primBytes, hipMemcpyDeviceToDevice, s) != hipSuccess) {
        grim::setError("grim_autotune_newton: copy failed");
        rc = GRIM_ERR_HIP;
        break;
      }
      (void)hipEventRecord(e0, s);
      rc = grim_batched_newton(p, &bb, &o, nullptr, stream);
      (void)hipEventRecord(e1, s);
      if (hipEventSynchronize(e1) != hipSuccess) rc = GRIM_ERR_HIP;
      float ms = 0.f;
      (void)hipEventElapsedTime(&ms, e0, e1);
      t.push_back(ms);
    }
    if (rc != GRIM_OK) break;
    std::sort(t.begin(), t.end());
    if (t[1] < best) {
      best = t[1];
      bestOpts = o;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(prim);
  (void)hipFree(ws);
  if (rc != GRIM_OK) return rc;
  bestOpts.workspace = opts->workspace;
  bestOpts.workspace_bytes = opts->workspace_bytes;
  *opts = bestOpts;
  return GRIM_OK;
}
