// grim's grid (src/grid/grid.hpp:13-94) over device memory: numVars planes
// of N1Total * N2Total * N3Total doubles (X1 fastest, numGhost ghost zones
// per side of each active dimension), the layout of grid::vars.
//
// In grim a grid owns a PETSc DMDA and ArrayFire arrays; here it owns (or
// views) one HIP allocation and is tied to a library context, which plays the
// DMDA's part: it knows the global domain, the boundary types and how
// communicate fills ghost zones.  Errors throw std::runtime_error.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "grim_hip.h"

namespace grim {

inline void check(int rc, const char* what) {
  if (rc != GRIM_OK) throw std::runtime_error(std::string(what) + ": " + grim_last_error());
}

inline void checkHip(hipError_t e, const char* what) {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

class grid {
  bool owned = false;

 public:
  int numGhost, numVars, dim;
  int N1, N2, N3;
  int N1Total, N2Total, N3Total;
  grim_ctx* ctx;  // the context whose domain this grid covers
  double* vars;   // device: numVars planes of zones() doubles

  // A grid of numVars variables on the context's domain; zero-initialised.
  grid(grim_ctx* context, const grim_grid_config& c, int numVars)
      : numGhost(c.numGhost), numVars(numVars), dim(c.dim), N1(c.N1), N2(c.N2), N3(c.N3), ctx(context),
        vars(nullptr) {
    setTotals();
    checkHip(hipMalloc((void**)&vars, bytes()), "grid: hipMalloc");
    owned = true;
    checkHip(hipMemset(vars, 0, bytes()), "grid: hipMemset");
  }
  // A view of numVars planes the context already holds (not freed here).
  grid(grim_ctx* context, const grim_grid_config& c, int numVars, double* view)
      : numGhost(c.numGhost), numVars(numVars), dim(c.dim), N1(c.N1), N2(c.N2), N3(c.N3), ctx(context),
        vars(view) {
    setTotals();
  }
  ~grid() {
    if (owned && vars) (void)hipFree(vars);
  }
  grid(const grid&) = delete;
  grid& operator=(const grid&) = delete;

  int64_t zones() const { return (int64_t)N1Total * N2Total * N3Total; }
  size_t bytes() const { return sizeof(double) * (size_t)numVars * (size_t)zones(); }
  double* var(int v) const { return vars + (size_t)v * (size_t)zones(); }

  // grid::communicate (grid.cpp:306-360): periodic ghost zones on one rank,
  // zeros outside the global domain.
  void communicate() { check(grim_communicate(ctx, vars, numVars), "grid::communicate"); }

  // grid::setVars / getVars of the Cython layer (gridPy.pyx:185-230): host
  // copies in the same plane order.
  void copyHostPtrToVars(const double* hostPtr) {
    checkHip(hipMemcpy(vars, hostPtr, bytes(), hipMemcpyHostToDevice), "grid: copy to device");
  }
  std::vector<double> getVars() const {
    std::vector<double> out((size_t)numVars * (size_t)zones());
    checkHip(hipMemcpy(out.data(), vars, bytes(), hipMemcpyDeviceToHost), "grid: copy to host");
    return out;
  }
  void copyFrom(const grid& other) {
    if (other.numVars != numVars || other.zones() != zones()) throw std::runtime_error("grid::copyFrom: shape");
    checkHip(hipMemcpy(vars, other.vars, bytes(), hipMemcpyDeviceToDevice), "grid: copy");
  }

 private:
  void setTotals() {
    N1Total = N1 + 2 * numGhost;
    N2Total = N2 + (dim > 1 ? 2 * numGhost : 0);
    N3Total = N3 + (dim > 2 ? 2 * numGhost : 0);
  }
};

// The coordinates grid of grim (grid.hpp:73-94): the domain and the location
// (locations enum of src/params.hpp) whose X coordinates it holds.
enum locations { CENTER = 0, LEFT = 1, RIGHT = 2, TOP = 3, BOTTOM = 4, FRONT = 5, BACK = 6 };

struct coordinatesGrid {
  int N1, N2, N3, dim, numGhost;
  double X1Start, X1End, X2Start, X2End, X3Start, X3End;
  double dX1, dX2, dX3;
  int location = CENTER;
  // the coordinate-map switches grim's geometry reads from namespace params
  // (DerefineThetaHorizon, DoCylindrify, X1cyl, X2cyl; geometry.cpp:255-424)
  int DerefineThetaHorizon = 0, DoCylindrify = 0;
  double X1cyl = 0., X2cyl = 0.;

  coordinatesGrid(const int N1, const int N2, const int N3, const int dim, const int numGhost,
                  const double X1Start, const double X1End, const double X2Start, const double X2End,
                  const double X3Start, const double X3End)
      : N1(N1), N2(N2), N3(N3), dim(dim), numGhost(numGhost), X1Start(X1Start), X1End(X1End),
        X2Start(X2Start), X2End(X2End), X3Start(X3Start), X3End(X3End), dX1((X1End - X1Start) / N1),
        dX2((X2End - X2Start) / N2), dX3((X3End - X3Start) / N3) {}

  // coordinatesGrid::setXCoords (grid.cpp:168-249): which location the
  // geometry built from this grid sits at.
  void setXCoords(const int loc) { location = loc; }
};

}  // namespace grim
