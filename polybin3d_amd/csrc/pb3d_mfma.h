// Split-BF16 helpers for the gfx950 matrix cores (v_mfma_f32_32x32x16_bf16, FP32 accumulation),
// shared by the triplet and Fisher contractions.
//
// Split BF16: an FP32 operand x = hi + mid (+ lo), each part bf16.  NP = 2 parts: hi and mid rounded
// to nearest, 3-term product hi*hi + hi*mid + mid*hi (~2^-16 relative per product).  NP = 3 parts:
// truncating split, hi + mid + lo == x exactly, 6-term product adding hi*lo + lo*hi + mid*mid
// (~2^-24, FP32-like).
//
// MFMA 32x32x16 lane layout (checked on the device): A[i][k] from lane l: i = l%32,
// k = 8*(l/32) + r (r = 0..7 of an 8 x bf16 operand); B[k][j]: j = l%32, same k; accumulator
// register v of lane l holds C[8*(v/4) + 4*(l/32) + v%4][l%32].
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace pb3d {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float f32x16;


typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// x = p[0] + p[1] (+ p[2]) in bf16 parts; NP = 2 or 3
template <int NP>
__device__ __forceinline__ void split8(const float* x, bf16x8* p) {
  if constexpr (NP == 3) {
    // truncating split, exact: hi = the top 8 significand bits of x, mid = the top 8 of x - hi,
    // lo = x - hi - mid, which has at most 8 significant bits (a bf16 is the high half of a float);
    // masks, subtractions and one byte permute per pair of values
    uint32_t w[3][8];
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      const uint32_t h = __float_as_uint(x[r]) & 0xffff0000u;
      const float r1 = x[r] - __uint_as_float(h);
      const uint32_t m = __float_as_uint(r1) & 0xffff0000u;
      w[0][r] = h;
      w[1][r] = m;
      w[2][r] = __float_as_uint(r1 - __uint_as_float(m));
    }
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      u32x4 v;
#pragma unroll
      for (int j = 0; j < 4; ++j) v[j] = __builtin_amdgcn_perm(w[q][2 * j + 1], w[q][2 * j], 0x07060302u);
      p[q] = __builtin_bit_cast(bf16x8, v);
    }
    return;
  }
#pragma unroll
  for (int r = 0; r < 8; ++r) {
    float v = x[r];
    const __bf16 h = (__bf16)v;
    v -= (float)h;
    const __bf16 m = (__bf16)v;
    p[0][r] = h;
    p[1][r] = m;
    if constexpr (NP == 3) {
      v -= (float)m;
      p[2][r] = (__bf16)v;
    }
  }
}

template <int NP>
__device__ __forceinline__ f32x16 mma_split(const bf16x8* a, const bf16x8* b, f32x16 c) {
  if constexpr (NP == 3) {  // small terms first
    c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[1], b[1], c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[0], b[2], c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[2], b[0], c, 0, 0, 0);
  }
  c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[0], b[1], c, 0, 0, 0);
  c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[1], b[0], c, 0, 0, 0);
  c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[0], b[0], c, 0, 0, 0);
  return c;
}

}  // namespace pb3d
