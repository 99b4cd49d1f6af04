// grim's HDF5 output and restart protocol on the C++ timeStepper interface
// (timestepper.hpp), over include/grim_io.h (link with -lgrim_io).
//
//  dumpPrimitives   grid::dump of primOld as "primitives" (the torus's
//                   WriteData dumps, torus/problem.cpp:778-840)
//  dumpVTS          the dump variables of primOld as dump%06d.vts
//                   (torus/problem.cpp:808-876)
//  dumpMetric       the metric files of the first WriteData (gCov.h5, gCon.h5,
//                   sqrtDetg.h5, xCoords.h5; torus/problem.cpp:780-792)
//  loadPrimitives   grid::load into primOld (src/grid/grid.cpp:482-499)
//  writeRestart     CheckWallClockTermination's restart dump and the two
//                   text files naming it (torus/problem.cpp:880-906)
//  restart          the timeStepper constructor's restart branch
//                   (src/timestepper/timestepper.cpp:387-441)
#pragma once

#include <sys/stat.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <string>
#include <vector>

#include "grim_io.h"
#include "timestepper.hpp"

namespace grim {

inline void checkIo(int rc, const char* what) {
  if (rc != 0) throw std::runtime_error(std::string(what) + ": " + grim_io_last_error());
}

inline bool fileExists(const std::string& f) {
  struct stat info;
  return stat(f.c_str(), &info) == 0;
}

inline void dumpPrimitives(const timeStepper& ts, const std::string& file,
                           const std::string& dataset = "primitives") {
  const std::vector<double> prim = ts.getPrim();
  checkIo(grim_io_dump(&ts.config(), ts.numVars, prim.data(), dataset.c_str(), file.c_str()), "grim_io_dump");
}

// The bulk from the file; ghost zones keep their values except the periodic
// ones, which communicate refills (grid::load's DMGlobalToLocal).
inline void loadPrimitives(timeStepper& ts, const std::string& file, const std::string& dataset = "primitives") {
  std::vector<double> prim = ts.getPrim();
  checkIo(grim_io_load(&ts.config(), ts.numVars, prim.data(), dataset.c_str(), file.c_str()), "grim_io_load");
  ts.setPrim(prim);
  ts.communicate();
}

// primVarsT%06d.h5, WriteIdx = floor(time / WriteDataEveryDt)
inline std::string primVarsName(double time, double writeDataEveryDt) {
  const std::string idx = std::to_string((long long)std::floor(time / writeDataEveryDt));
  return "primVarsT" + std::string(idx.size() < 6 ? 6 - idx.size() : 0, '0') + idx + ".h5";
}

// dump%06d.vts, WriteIdx = floor(time / WriteDataEveryDt)
inline std::string dumpVTSName(double time, double writeDataEveryDt) {
  const std::string idx = std::to_string((long long)std::floor(time / writeDataEveryDt));
  return "dump" + std::string(idx.size() < 6 ? 6 - idx.size() : 0, '0') + idx + ".vts";
}

// The variable names of torus/problem.cpp:808-825 in the plane order of
// grim_dump_vars (params.cpp:136-147), then the VTS file.
inline void dumpVTS(const timeStepper& ts, const std::string& file) {
  std::vector<const char*> names = {"rho", "u", "u1", "u2", "u3"};
  if (ts.physics().conduction) names.push_back("q");
  if (ts.physics().viscosity) names.push_back("dP");
  for (const char* n : {"B1", "B2", "B3", "bSqr", "gamma"}) names.push_back(n);
  if ((int)names.size() != ts.numVars + 2) throw std::runtime_error("dumpVTS: variable count");
  const std::vector<double> vars = ts.dumpVars();
  checkIo(grim_io_dump_vts(&ts.config(), (int)names.size(), names.data(), vars.data(), file.c_str()),
          "grim_io_dump_vts");
}

// The metric files the first WriteData writes (torus/problem.cpp:780-792):
// geomCenter's gCovGrid, gConGrid (16 variables each, index nu + 4 mu;
// geometry.cpp:686-714), gGrid (sqrt(-g)) and xCoordsGrid (the zone centres'
// x coordinates; geometry.cpp:753-764) as datasets "gCov", "gCon",
// "sqrtDetg", "xCoords" of gCov.h5, gCon.h5, sqrtDetg.h5 and xCoords.h5.
inline void dumpMetric(const timeStepper& ts, const std::string& dir = ".") {
  const int64_t n = ts.zones();
  std::vector<double> geom((size_t)GRIM_LOC_PLANES * (size_t)n);
  check(grim_get_geometry(ts.context(), 0, geom.data(), 0), "grim_get_geometry");
  auto packed = [](int a, int b, int m) {
    const int lo = a < b ? a : b, hi = a < b ? b : a;
    return lo * m - (lo * (lo + 1)) / 2 + hi;
  };
  std::vector<double> gCov((size_t)16 * n), gCon((size_t)16 * n);
  for (int mu = 0; mu < 4; mu++)
    for (int nu = 0; nu < 4; nu++) {
      const int cov = packed(mu, nu, 4);
      const int con = (mu == 0 || nu == 0) ? 10 + (mu == 0 ? nu : mu) : 16 + packed(mu - 1, nu - 1, 3);
      std::copy(geom.begin() + (size_t)cov * n, geom.begin() + (size_t)(cov + 1) * n,
                gCov.begin() + (size_t)(nu + 4 * mu) * n);
      std::copy(geom.begin() + (size_t)con * n, geom.begin() + (size_t)(con + 1) * n,
                gCon.begin() + (size_t)(nu + 4 * mu) * n);
    }
  const std::vector<double> g(geom.begin() + (size_t)15 * n, geom.begin() + (size_t)16 * n);
  // the x coordinates of the bulk zones, as grim_dump_vars gives them,
  // placed into ghosted planes (grid::dump writes the bulk only)
  const grim_grid_config& c = ts.config();
  const std::vector<double> vars = ts.dumpVars();
  const size_t nb = (size_t)c.N1 * c.N2 * c.N3;
  std::vector<double> x((size_t)3 * n, 0.);
  const int g1 = c.numGhost, g2 = c.dim > 1 ? c.numGhost : 0, g3 = c.dim > 2 ? c.numGhost : 0;
  const int64_t N1T = c.N1 + 2 * g1, N2T = c.N2 + 2 * g2;
  const double dX3 = (c.X3End - c.X3Start) / c.N3;
  for (int d = 0; d < 3; d++)
    for (int k = 0; k < c.N3; k++)
      for (int j = 0; j < c.N2; j++)
        for (int i = 0; i < c.N1; i++) {
          // x^3 = X^3 in both metrics (phi = X3); the dump variables' third
          // point coordinate is 0 below three dimensions, as PETSc's VTS
          // points are, so it is formed here from the zone centre
          const size_t b = (size_t)(i + (int64_t)c.N1 * (j + (int64_t)c.N2 * k));
          x[(size_t)d * n + (size_t)((i + g1) + N1T * ((j + g2) + N2T * (k + g3)))] =
              d < 2 ? vars[(size_t)(ts.numVars + 2 + d) * nb + b] : c.X3Start + (k + 0.5) * dX3;
        }
  checkIo(grim_io_dump(&c, 16, gCov.data(), "gCov", (dir + "/gCov.h5").c_str()), "grim_io_dump");
  checkIo(grim_io_dump(&c, 16, gCon.data(), "gCon", (dir + "/gCon.h5").c_str()), "grim_io_dump");
  checkIo(grim_io_dump(&c, 1, g.data(), "sqrtDetg", (dir + "/sqrtDetg.h5").c_str()), "grim_io_dump");
  checkIo(grim_io_dump(&c, 3, x.data(), "xCoords", (dir + "/xCoords.h5").c_str()), "grim_io_dump");
}

// restartVarsT%08d.h5 with index floor(100 time), then the file name and
// the time (default stream precision, as grim writes it) in the two text
// files.  Returns the dump's name.
inline std::string writeRestart(const timeStepper& ts, const std::string& restartFileName = "restartFileName.txt",
                                const std::string& restartFileTime = "restartFileTime.txt") {
  const std::string idx = std::to_string((long long)std::floor(ts.time * 100));
  const std::string file = "restartVarsT" + std::string(idx.size() < 8 ? 8 - idx.size() : 0, '0') + idx + ".h5";
  dumpPrimitives(ts, file);
  std::ofstream fName(restartFileName);
  fName << file << std::endl;
  std::ofstream fTime(restartFileTime);
  fTime << ts.time << std::endl;
  if (!fName || !fTime) throw std::runtime_error("writeRestart: cannot write " + restartFileName);
  return file;
}

// When both text files exist: the time and the dump they name; otherwise,
// when `force` (params::restart): restartFile.  Then the torus's floors
// (fullStepDiagnostics "to reset the time step").  Returns whether a state
// was loaded.
inline bool restart(timeStepper& ts, bool force = false, const std::string& restartFile = "restartFile.h5",
                    const std::string& restartFileName = "restartFileName.txt",
                    const std::string& restartFileTime = "restartFileTime.txt") {
  const bool haveFiles = fileExists(restartFileName) && fileExists(restartFileTime);
  if (!force && !haveFiles) return false;
  if (haveFiles) {
    double restartTime = 0.;
    std::ifstream fTime(restartFileTime);
    fTime >> restartTime;
    std::string file;
    std::ifstream fName(restartFileName);
    fName >> file;
    if (!fTime || !fName) throw std::runtime_error("restart: cannot read " + restartFileName);
    std::printf("\n Restarting from file %s at time %e\n\n", restartFileName.c_str(), restartTime);
    loadPrimitives(ts, file);
    ts.setTime(restartTime, ts.dt);
  } else {
    if (!fileExists(restartFile)) throw std::runtime_error("Restart file " + restartFile + " does not exist");
    std::printf("\nFound restart file: %s\n\n", restartFile.c_str());
    loadPrimitives(ts, restartFile);
  }
  if (ts.torus()) ts.applyFloor();
  return true;
}

}  // namespace grim
