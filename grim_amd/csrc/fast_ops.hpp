// Device helpers of the fast translation units (*_fast.hip).
//
//  * ldOff / stOff: a plane element at a 32-bit byte offset from a base
//    that is uniform across the wavefront.  The base is pinned to scalar
//    registers, so the access is a global_load/store with an SGPR base and
//    one VGPR offset shared by every plane: no per-access 64-bit vector
//    address arithmetic.  Callers keep every plane under 4 GiB
//    (n * 8 < 2^32 bytes; planeOffsetsFit).
//  * fastRcp / fastDiv: FP64 reciprocal and quotient from the hardware
//    reciprocal estimate refined by two Newton-Raphson steps, and one
//    residual correction of the quotient (within about one ulp of the
//    correctly rounded result for finite, nonzero, normal divisors; a zero
//    or infinite divisor gives NaN where IEEE division gives inf or 0).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace grim {

typedef const double __attribute__((address_space(1)))* GlobalConstPtr;
typedef double __attribute__((address_space(1)))* GlobalPtr;

__device__ __forceinline__ double ldOff(const double* base, uint32_t off) {
  GlobalConstPtr b = (GlobalConstPtr)base;
  asm volatile("" : "+s"(b));
  return *(GlobalConstPtr)((const char __attribute__((address_space(1)))*)b + off);
}

__device__ __forceinline__ void stOff(double* base, uint32_t off, double v) {
  GlobalPtr b = (GlobalPtr)base;
  asm volatile("" : "+s"(b));
  *(GlobalPtr)((char __attribute__((address_space(1)))*)b + off) = v;
}

// Byte offsets of a zone within a plane of n doubles fit 32 bits.
inline bool planeOffsetsFit(int64_t n) { return n > 0 && n <= (int64_t)(UINT32_MAX / 8); }

__device__ __forceinline__ double fastRcp(double b) {
  double r = __builtin_amdgcn_rcp(b);
  double e = __builtin_fma(-b, r, 1.0);
  r = __builtin_fma(r, e, r);
  e = __builtin_fma(-b, r, 1.0);
  return __builtin_fma(r, e, r);
}

__device__ __forceinline__ double fastDiv(double a, double b) {
  const double r = fastRcp(b);
  const double q = a * r;
  return __builtin_fma(__builtin_fma(-b, q, a), r, q);
}

}  // namespace grim
