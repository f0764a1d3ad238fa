// Internal helpers shared by the pb3d HIP translation units (gfx950 / CDNA4 only).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <string>
#include "../../include/pb3d.h"

namespace pb3d {

// ------------------------------------------------------------------ error state (thread-local)
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);

#define PB3D_CHECK_HIP(expr)                                                              \
  do {                                                                                    \
    hipError_t _e = (expr);                                                               \
    if (_e != hipSuccess)                                                                 \
      return ::pb3d::fail(PB3D_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e)); \
  } while (0)

#define PB3D_CHECK_LAUNCH() PB3D_CHECK_HIP(hipGetLastError())

// ------------------------------------------------------------------ complex helpers
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }

// signed frequency index of position i on an axis of length n, following the reference grid
// convention fftshift(arange(-n//2, n//2)) (PolyBin3D/base.py:79,87): i for i < n//2, else i - n.
__host__ __device__ __forceinline__ int freq_index(int i, int n) { return i < n / 2 ? i : i - n; }

// Legendre L_l(mu) for l = 0..6 (PolyBin3D/cython/utils.pyx:385-400), evaluated in FP32.
__device__ __forceinline__ float legendre_f(int l, float mu) {
  float m2 = mu * mu;
  switch (l) {
    case 0: return 1.f;
    case 1: return mu;
    case 2: return 0.5f * (3.f * m2 - 1.f);
    case 3: return 0.5f * mu * (5.f * m2 - 3.f);
    case 4: return (3.f + m2 * (-30.f + 35.f * m2)) * 0.125f;
    case 5: return mu * (15.f + m2 * (-70.f + 63.f * m2)) * 0.125f;
    case 6: return (-5.f + m2 * (105.f + m2 * (-315.f + 231.f * m2))) * 0.0625f;
    default: return 0.f;
  }
}

// Real spherical harmonic Y_lm(n) for unit n=(x,y,z), normalised so that sum_m Y_lm(a)Y_lm(b) =
// L_l(a.b) (the convention of compute_real_harmonics, PolyBin3D/cython/utils.pyx:518-609).  The m
// order is this library's own; every caller sums over m.  n = 0 evaluates the same polynomials at
// the origin (Y_40(0) = 3/8), as the reference does for a zero coordinate vector.
__host__ __device__ __forceinline__ float ylm_f(int l, int m, float x, float y, float z) {
  const float s3 = 1.7320508075688772f, s5 = 2.2360679774997897f, s15 = 3.872983346207417f;
  const float s35 = 5.916079783099616f;
  float x2 = x * x, y2 = y * y, z2 = z * z;
  if (l == 1) return m == 0 ? x : (m == 1 ? y : z);
  if (l == 2) {
    switch (m) {
      case 0: return s3 * x * y;
      case 1: return s3 * y * z;
      case 2: return z2 - 0.5f * (x2 + y2);
      case 3: return s3 * x * z;
      default: return 0.5f * s3 * (x2 - y2);
    }
  }
  if (l == 3) {
    switch (m) {
      case 0: return 0.7905694150420949f * y * (3.f * x2 - y2);   // sqrt(5/8)
      case 1: return s15 * x * y * z;
      case 2: return 0.6123724356957945f * y * (5.f * z2 - 1.f);   // sqrt(3/8)
      case 3: return 0.5f * z * (5.f * z2 - 3.f);
      case 4: return 0.6123724356957945f * x * (5.f * z2 - 1.f);
      case 5: return 0.5f * s15 * z * (x2 - y2);
      default: return 0.7905694150420949f * x * (x2 - 3.f * y2);
    }
  }
  if (l == 4) {
    switch (m) {
      case 0: return 0.5f * s35 * x * y * (x2 - y2);
      case 1: return 2.0916500663351889f * y * z * (3.f * x2 - y2);  // sqrt(35/8)
      case 2: return 0.5f * s5 * x * y * (7.f * z2 - 1.f);
      case 3: return 0.7905694150420949f * y * z * (7.f * z2 - 3.f);  // sqrt(5/8)
      case 4: return (35.f * z2 * z2 - 30.f * z2 + 3.f) * 0.125f;
      case 5: return 0.7905694150420949f * x * z * (7.f * z2 - 3.f);
      case 6: return 0.25f * s5 * (x2 - y2) * (7.f * z2 - 1.f);
      case 7: return 2.0916500663351889f * x * z * (x2 - 3.f * y2);
      default: return 0.125f * s35 * (x2 * x2 - 6.f * x2 * y2 + y2 * y2);
    }
  }
  return l == 0 ? 1.f : 0.f;
}

}  // namespace pb3d
