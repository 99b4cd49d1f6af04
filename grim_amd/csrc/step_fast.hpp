// Host launchers of the fast explicit stage (step_fast.hip, compiled with
// -ffp-contract=fast-honor-pragmas): the face fluxes and the explicit sources of grim's
// stage evaluated with shared reciprocals, fused multiply-adds and the
// 40-plane symmetric connection.  The results agree with the exact kernels
// of step.hip (grim's expression order, one rounding per operation) to
// rounding, not bit for bit.
#pragma once
#include <hip/hip_runtime.h>

#include "grim_hip.h"
#include "step_physics.hpp"

namespace grim {

// Faces per workgroup of the split face kernels (two threads per face).
constexpr int kFacesPerBlock = kBlock / 2;

// Riemann fluxes at the lower faces (along dir) of every face of box b.
int launchFaceFast(const grim_params& p, const StepArgs& a, int dir, int recon, const Box& b, const double* prim,
                   const double* geomFace, double* flux, hipStream_t s);

// Explicit sources (geometric + EMHD gradients), the old implicit sources
// and (consOld != NULL) the conserved variables of the bulk box b.
int launchSourcesFast(const grim_params& p, const StepArgs& a, int recon, const Box& b, const double* prim,
                      const double* primOld, const double* geomC, const double* gamma40, const double* uT,
                      double* sExp, double* sImp, double* consOld, hipStream_t s);

// The symmetric connection: plane l*10 + (packed k <= n) of gamma40 takes
// Gamma^l_{k n} of the 64-plane connection (plane l*16 + k*4 + n).
void launchGamma40(int64_t n, const double* gamma64, double* gamma40, hipStream_t s);

}  // namespace grim
