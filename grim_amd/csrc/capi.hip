// C-ABI plumbing: parameter defaults, error text, device checks.
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>
#include <string>

#include "grim_hip.h"
#include "runtime.hpp"

namespace grim {

namespace {
thread_local std::string g_lastError;
std::mutex g_devMutex;
int g_devChecked[64] = {0};  // 0 unknown, 1 ok, -1 not gfx950
int g_devCUs[64] = {0};      // compute units of each checked device
}  // namespace

void setError(const std::string& msg) { g_lastError = msg; }

int checkDevice() {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess || dev < 0) {
    setError(std::string("no HIP device: ") + hipGetErrorString(e));
    return GRIM_ERR_NO_DEVICE;
  }
  if (dev < 64 && g_devChecked[dev] == 1) return GRIM_OK;
  std::lock_guard<std::mutex> lock(g_devMutex);
  hipDeviceProp_t prop;
  e = hipGetDeviceProperties(&prop, dev);
  if (e != hipSuccess) {
    setError(std::string("hipGetDeviceProperties: ") + hipGetErrorString(e));
    return GRIM_ERR_NO_DEVICE;
  }
  const bool ok = std::strncmp(prop.gcnArchName, "gfx950", 6) == 0;
  if (dev < 64) {
    g_devCUs[dev] = prop.multiProcessorCount;
    g_devChecked[dev] = ok ? 1 : -1;
  }
  if (!ok) {
    setError(std::string("device is ") + prop.gcnArchName + ", this library is built for gfx950");
    return GRIM_ERR_NO_DEVICE;
  }
  return GRIM_OK;
}

int deviceCUs() {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64 || g_devCUs[dev] <= 0) return 256;
  return g_devCUs[dev];
}

}  // namespace grim

extern "C" {

void grim_default_params(grim_params* p, int emhd) {
  if (!p) return;
  // src/problem/linear_modes/params.cpp:40-69: floors 1e-20, EMHD switches
  // with high-order terms, unit alphas, maxNonLinearIter = maxLineSearchIters
  // = 3, atol 1e-20, Jacobian epsilon 4e-8, line-search floor 1e-24.
  p->conduction = emhd ? 1 : 0;
  p->viscosity = emhd ? 1 : 0;
  p->highOrderTermsConduction = 1;
  p->highOrderTermsViscosity = 1;
  p->adiabaticIndex = 4. / 3;
  p->rhoFloorInFluidElement = 1e-20;
  p->uFloorInFluidElement = 1e-20;
  p->bSqrFloorInFluidElement = 1e-20;
  p->temperatureFloorInFluidElement = 1e-20;
  p->ConductionAlpha = 1.;
  p->ViscosityAlpha = 1.;
  p->closure = GRIM_CLOSURE_LINEAR_MODES;
  p->tau0 = 1.;
  p->ConductionClosureFactor = 1.;
  p->ViscosityClosureFactor = 1.;
  p->maxNonLinearIter = 3;
  p->maxLineSearchIters = 3;
  p->nonlinearsolve_atol = 1.e-20;
  p->JacobianAssembleEpsilon = 4.e-8;
  p->linesearchfloor = 1.e-24;
}

const char* grim_last_error(void) { return grim::g_lastError.c_str(); }

int grim_device_info(char* name, int name_len) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0) {
    grim::setError("no HIP device");
    return GRIM_ERR_NO_DEVICE;
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, dev) != hipSuccess) {
    grim::setError("hipGetDeviceProperties failed");
    return GRIM_ERR_NO_DEVICE;
  }
  if (name && name_len > 0) {
    std::strncpy(name, prop.gcnArchName, (size_t)name_len - 1);
    name[name_len - 1] = '\0';
  }
  return grim::checkDevice();
}

}  // extern "C"
