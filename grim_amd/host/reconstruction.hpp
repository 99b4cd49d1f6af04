// grim's reconstruction namespace (src/reconstruction/reconstruction.hpp:7-68)
// on device grids: the face values of every zone along one direction with
// the context's scheme (params::reconstruction: MINMOD, WENO5 or PPM).
#pragma once

#include "grid.hpp"

namespace grim {
namespace reconstruction {

// reconstruction::reconstruct (reconstruction.cpp:3-38): primLeft holds the
// value at the lower face (i - 1/2) of each zone, primRight the value at the
// upper face (i + 1/2), for the zones whose stencil lies in the grid.  grim
// counts array reads and writes for its bandwidth report; here both are 0
// (the kernels' traffic is measured with rocprofv3).
inline void reconstruct(const grid& prim, const int dir, grid& primLeft, grid& primRight, int& numReads,
                        int& numWrites) {
  if (primLeft.numVars < prim.numVars || primRight.numVars < prim.numVars)
    throw std::runtime_error("reconstruct: output grids hold fewer variables than prim");
  check(grim_reconstruct(prim.ctx, prim.vars, prim.numVars, dir, primLeft.vars, primRight.vars),
        "reconstruction::reconstruct");
  numReads = 0;
  numWrites = 0;
}

}  // namespace reconstruction
}  // namespace grim
