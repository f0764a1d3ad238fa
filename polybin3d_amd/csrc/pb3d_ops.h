// Device-side geometry and the fused voxel / mode operations applied in FFT prologues and
// epilogues (and by the stand-alone mode-map kernel).
#pragma once
#include "pb3d_internal.h"

namespace pb3d {

// Geometry in FP32 for in-kernel mu, Y_lm and coordinates; bin membership never uses it (that
// comes from the FP64-built bin table).
struct Geo {
  int n0, n1, n2, n2h;
  float kF0, kF1, kF2;
  float dx0, dx1, dx2;
  float c0, c1, c2;
  float s0, s1, s2;
  long long H;  // n0*n1*n2h
  long long N;  // n0*n1*n2
};

inline Geo make_geo(const pb3d_grid& g) {
  Geo o;
  o.n0 = g.n[0]; o.n1 = g.n[1]; o.n2 = g.n[2]; o.n2h = g.n[2] / 2 + 1;
  const double twopi = 6.283185307179586476925286766559;
  o.kF0 = (float)(twopi / g.boxsize[0]); o.kF1 = (float)(twopi / g.boxsize[1]);
  o.kF2 = (float)(twopi / g.boxsize[2]);
  o.dx0 = (float)(g.boxsize[0] / g.n[0]); o.dx1 = (float)(g.boxsize[1] / g.n[1]);
  o.dx2 = (float)(g.boxsize[2] / g.n[2]);
  o.c0 = (float)g.boxcenter[0]; o.c1 = (float)g.boxcenter[1]; o.c2 = (float)g.boxcenter[2];
  o.s0 = (float)g.sight[0]; o.s1 = (float)g.sight[1]; o.s2 = (float)g.sight[2];
  o.H = (long long)g.n[0] * g.n[1] * o.n2h;
  o.N = (long long)g.n[0] * g.n[1] * g.n[2];
  return o;
}

// Plain-old-data copies of pb3d_rop / pb3d_kop passed by value to kernels.
struct ROp {
  int kind, ell, m, accumulate;
  const float* a; const float* b; const float* c; const float* d;
  long long sa, sb, sc, sd;
  float scale;
};
struct KOp {
  int kind, ell, ell2, m, accumulate;
  const uint8_t* bin_table;
  const int32_t* bins;
  const float* w;
  long long x_stride;
  int x_group;  // batch entries per input step (>= 1)
  float scale;
  const int32_t* pack_index;
  long long packed_len;
  const int32_t* trip;
  const int32_t* pair_index;
  const float2* ft;
  long long ft_lstride;
  int nk, terms, trip0, leg3_legendre;
};

inline ROp to_rop(const pb3d_rop* r) {
  ROp o{};
  if (!r) { o.kind = 0; o.scale = 1.f; return o; }
  o.kind = r->kind; o.ell = r->ell; o.m = r->m; o.accumulate = r->accumulate;
  o.a = r->a; o.b = r->b; o.c = r->c; o.d = r->d;
  o.sa = r->batch_stride_a; o.sb = r->batch_stride_b; o.sc = r->batch_stride_c;
  o.sd = r->batch_stride_d; o.scale = r->scale;
  return o;
}
inline KOp to_kop(const pb3d_kop* k) {
  KOp o{};
  if (!k) { o.kind = 0; o.scale = 1.f; o.x_group = 1; return o; }
  o.kind = k->kind; o.ell = k->ell; o.ell2 = k->ell2; o.m = k->m; o.accumulate = k->accumulate;
  o.bin_table = k->bin_table; o.bins = k->bins; o.w = k->w; o.x_stride = k->x_batch_stride;
  o.x_group = k->x_batch_group > 1 ? k->x_batch_group : 1;
  o.scale = k->scale;
  o.pack_index = k->pack_index; o.packed_len = k->packed_len; o.trip = k->trip;
  o.pair_index = k->pair_index; o.ft = (const float2*)k->ft; o.ft_lstride = k->ft_lstride;
  o.nk = k->nk; o.terms = k->terms; o.trip0 = k->trip0; o.leg3_legendre = k->leg3_legendre;
  return o;
}

// unit position vector of voxel (i0,i1,i2): r = fs(i)*L/n + boxcenter (PolyBin3D/base.py:77-80)
__device__ __forceinline__ void unit_r(const Geo& g, int i0, int i1, int i2, float& x, float& y,
                                       float& z) {
  x = freq_index(i0, g.n0) * g.dx0 + g.c0;
  y = freq_index(i1, g.n1) * g.dx1 + g.c1;
  z = freq_index(i2, g.n2) * g.dx2 + g.c2;
  float r2 = x * x + y * y + z * z;
  float inv = r2 > 0.f ? rsqrtf(r2) : 0.f;
  x *= inv; y *= inv; z *= inv;
}

// wavevector of half-grid mode (i,j,l) (PolyBin3D/base.py:86-92)
__device__ __forceinline__ void kvec(const Geo& g, int i, int j, int l, float& kx, float& ky,
                                     float& kz) {
  kx = freq_index(i, g.n0) * g.kF0;
  ky = freq_index(j, g.n1) * g.kF1;
  kz = l * g.kF2;
}

// Real-space voxel op on the value pair at (i0,i1,2m),(i0,i1,2m+1) of batch entry `bt`.
// For the R2C prologue `src` is ignored and a is read; returns the op result.
__device__ __forceinline__ float rop_value(const ROp& op, const Geo& g, long long bt,
                                           long long vox, int i0, int i1, int i2, float tval) {
  switch (op.kind) {
    case 0: return tval;
    case 1: return tval * op.b[bt * op.sb + vox];
    case 2: return tval * op.b[bt * op.sb + vox] - op.c[bt * op.sc + vox] * op.d[bt * op.sd + vox];
    case 3: {
      float x, y, z; unit_r(g, i0, i1, i2, x, y, z);
      return tval * ylm_f(op.ell, op.m, x, y, z);
    }
    case 4: {
      float x, y, z; unit_r(g, i0, i1, i2, x, y, z);
      return tval * op.b[bt * op.sb + vox] * ylm_f(op.ell, op.m, x, y, z);
    }
    case 5: return tval * tval;
    default: return tval;
  }
}

// Mode op on value v at half-grid mode (i,j,l), linear index `mode`, batch entry `bt`.
__device__ __forceinline__ float2 kop_value(const KOp& op, const Geo& g, long long bt,
                                            long long mode, int i, int j, int l, float2 v) {
  switch (op.kind) {
    case 0: return v;
    case 1: {
      int bin = op.bin_table[mode];
      int want = op.bins ? op.bins[bt] : -1;
      if (want >= 0 && bin != want) return make_float2(0.f, 0.f);
      if (want < 0 && bin == PB3D_NO_BIN) return make_float2(0.f, 0.f);
      float f = op.w ? op.w[mode] : 1.f;
      if (op.ell != 0 || op.ell2 != 0) {
        float kx, ky, kz; kvec(g, i, j, l, kx, ky, kz);
        float k2 = kx * kx + ky * ky + kz * kz;
        float mu = k2 > 0.f ? (g.s0 * kx + g.s1 * ky + g.s2 * kz) * rsqrtf(k2) : 0.f;
        f *= legendre_f(op.ell, mu) * legendre_f(op.ell2, mu);
      }
      return cscale(v, f);
    }
    case 2: {
      if (op.bin_table) {
        int bin = op.bin_table[mode];
        int want = op.bins ? op.bins[bt] : -1;
        if (want >= 0 ? bin != want : bin == PB3D_NO_BIN) return make_float2(0.f, 0.f);
      }
      float kx, ky, kz; kvec(g, i, j, l, kx, ky, kz);
      float k2 = kx * kx + ky * ky + kz * kz;
      float inv = k2 > 0.f ? rsqrtf(k2) : 0.f;
      float f = ylm_f(op.ell, op.m, kx * inv, ky * inv, kz * inv);
      if (op.w) f *= op.w[mode];
      return cscale(v, f);
    }
    case 3: return cscale(v, op.w[mode]);
    case 4: {
      // Q map of one triplet from packed FT[g_A g_B] rows (bspec.py:615-640)
      const int bin = op.bin_table[mode];
      const int p = op.pack_index[mode];
      if (bin == PB3D_NO_BIN || p < 0) return make_float2(0.f, 0.f);
      const int* tr = op.trip + 4 * (long long)(op.trip0 + bt);
      const int b1 = tr[0], b2 = tr[1], b3 = tr[2], ell = tr[3];
      float2 acc = make_float2(0.f, 0.f);
      if ((op.terms & 5) && bin == b3) {
        const int pr = op.pair_index[b1 * op.nk + b2];
        if (pr >= 0) {
          float2 f = op.ft[(long long)pr * op.packed_len + p];
          float kx, ky, kz; kvec(g, i, j, l, kx, ky, kz);
          const float k2 = kx * kx + ky * ky + kz * kz;
          const float inv = k2 > 0.f ? rsqrtf(k2) : 0.f;
          if (op.terms & 1) {
            float fl = 1.f;
            if (op.leg3_legendre && ell != 0) fl = legendre_f(ell, (g.s0 * kx + g.s1 * ky + g.s2 * kz) * inv);
            acc = cadd(acc, cscale(f, fl));
          }
          if (op.terms & 4) acc = cadd(acc, cscale(f, ylm_f(op.ell, op.m, kx * inv, ky * inv, kz * inv)));
        }
      }
      if (op.terms & 2) {
        const float2* fl = op.ft + (long long)(ell / 2) * op.ft_lstride;
        if (bin == b1) {
          const int pr = op.pair_index[b2 * op.nk + b3];
          if (pr >= 0) acc = cadd(acc, fl[(long long)pr * op.packed_len + p]);
        }
        if (bin == b2) {
          const int pr = op.pair_index[b1 * op.nk + b3];
          if (pr >= 0) acc = cadd(acc, fl[(long long)pr * op.packed_len + p]);
        }
      }
      return acc;
    }
    case 6: {
      const int bin = op.bin_table[mode];
      if (bin == PB3D_NO_BIN) return make_float2(0.f, 0.f);
      const float* coef = reinterpret_cast<const float*>(op.ft);
      float f = coef[bin];
      if (op.nk > 1) {
        float kx, ky, kz; kvec(g, i, j, l, kx, ky, kz);
        const float k2 = kx * kx + ky * ky + kz * kz;
        const float mu = k2 > 0.f ? (g.s0 * kx + g.s1 * ky + g.s2 * kz) * rsqrtf(k2) : 0.f;
        for (int q = 1; q < op.nk; ++q) f += coef[q * op.terms + bin] * legendre_f(2 * q, mu);
      }
      if (op.w) f *= op.w[mode];
      if ((g.n2 & 1) == 0 && l == g.n2h - 1) f *= 2.f;
      return cscale(v, f);
    }
    default: return v;
  }
}

}  // namespace pb3d
