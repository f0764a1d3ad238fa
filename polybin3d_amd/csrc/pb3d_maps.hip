// Mode-space elementwise maps (fourier_filter family, prod_fourier) and the Gaussian random field
// Hermitian projection.
#include "pb3d_ops.h"

namespace pb3d {

int check_kop(const pb3d_kop* k);

__global__ void kmap_kernel(const float2* __restrict__ x, float2* __restrict__ out, Geo g, KOp op) {
  const long long bt = blockIdx.y;
  const long long n12 = (long long)g.n1 * g.n2h;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < g.H;
       e += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(e / n12);
    const long long r = e % n12;
    const int j = (int)(r / g.n2h), l = (int)(r % g.n2h);
    float2 v = kop_value(op, g, bt, e, i, j, l, x[(bt / op.x_group) * op.x_stride + e]);
    v = cscale(v, op.scale);
    float2* o = out + bt * g.H + e;
    if (op.accumulate) v = cadd(v, *o);
    *o = v;
  }
}

// a(k) = 0.5 (r(k) + conj r(-k)) * sqrtPw(k) * norm, r = re + i im on the full grid, r(0)=0
// (PolyBin3D/base.py:332-344).  FP64 symmetrisation, FP32 output.
__global__ void grf_kernel(const double* __restrict__ re, const double* __restrict__ im,
                           const float* __restrict__ sqrtPw, float norm, float2* __restrict__ out, Geo g) {
  const long long n12 = (long long)g.n1 * g.n2h;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < g.H;
       e += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(e / n12);
    const long long r = e % n12;
    const int j = (int)(r / g.n2h), l = (int)(r % g.n2h);
    const int im_i = (g.n0 - i) % g.n0, im_j = (g.n1 - j) % g.n1, im_l = (g.n2 - l) % g.n2;
    const long long p = ((long long)i * g.n1 + j) * g.n2 + l;
    const long long q = ((long long)im_i * g.n1 + im_j) * g.n2 + im_l;
    double ar = re[p], ai = im[p], br = re[q], bi = im[q];
    if (p == 0) { ar = 0.; ai = 0.; }
    if (q == 0) { br = 0.; bi = 0.; }
    double vr = 0.5 * (ar + br), vi = 0.5 * (ai - bi);
    float f = norm * (sqrtPw ? sqrtPw[e] : 1.f);
    out[e] = make_float2((float)vr * f, (float)vi * f);
  }
}

// Counter-based normal pair for full-grid index q: a 64-bit SplitMix-style mix of (seed, q) gives
// two uniform words, Box-Muller gives (re, im) ~ N(0,1) each (the distribution of the reference's
// randn draws; the stream itself differs).
__device__ __forceinline__ unsigned long long mix64(unsigned long long z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
__device__ __forceinline__ double2 normal_pair(unsigned long long seed, long long q) {
  const unsigned long long h1 = mix64(seed * 0x632BE59BD9B4E019ull + 2ull * (unsigned long long)q);
  const unsigned long long h2 = mix64(h1 ^ (seed + 0x85157AF5ull));
  const double u1 = ((h1 >> 11) + 1.0) * (1.0 / 9007199254740993.0);  // (0, 1]
  const double u2 = (h2 >> 11) * (1.0 / 9007199254740992.0);          // [0, 1)
  const double rr = sqrt(-2.0 * log(u1));
  double sn, cs;
  sincospi(2.0 * u2, &sn, &cs);
  return make_double2(rr * cs, rr * sn);
}

// seed_at != nullptr: the seed is *seed_at + seed, read on the device (graph replays with new seeds)
__global__ void grf_device_kernel(unsigned long long seed, const unsigned long long* __restrict__ seed_at,
                                  const float* __restrict__ sqrtPw, float norm, float2* __restrict__ out, Geo g) {
  if (seed_at) seed += *seed_at;
  const long long n12 = (long long)g.n1 * g.n2h;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < g.H;
       e += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(e / n12);
    const long long r = e % n12;
    const int j = (int)(r / g.n2h), l = (int)(r % g.n2h);
    const int im_i = (g.n0 - i) % g.n0, im_j = (g.n1 - j) % g.n1, im_l = (g.n2 - l) % g.n2;
    const long long p = ((long long)i * g.n1 + j) * g.n2 + l;
    const long long q = ((long long)im_i * g.n1 + im_j) * g.n2 + im_l;
    double2 a = p == 0 ? make_double2(0., 0.) : normal_pair(seed, p);
    double2 c = q == 0 ? make_double2(0., 0.) : normal_pair(seed, q);
    const double vr = 0.5 * (a.x + c.x), vi = 0.5 * (a.y - c.y);
    const float f = norm * (sqrtPw ? sqrtPw[e] : 1.f);
    out[e] = make_float2((float)vr * f, (float)vi * f);
  }
}

// out[bt][p] (+)= scale * kop(mode = modes[p])  (x ignored for BKQ)
__global__ void kmap_packed_kernel(const float2* __restrict__ x, const int32_t* __restrict__ modes, long long M,
                                   float2* __restrict__ out, Geo g, KOp op) {
  const long long bt = blockIdx.y;
  const long long n12 = (long long)g.n1 * g.n2h;
  for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < M; p += (long long)gridDim.x * blockDim.x) {
    const long long e = modes[p];
    if (e < 0 || e >= g.H) continue;
    const int i = (int)(e / n12);
    const long long r = e % n12;
    const int j = (int)(r / g.n2h), l = (int)(r % g.n2h);
    const float2 xv = x ? x[(bt / op.x_group) * op.x_stride + e] : make_float2(0.f, 0.f);
    float2 v = cscale(kop_value(op, g, bt, e, i, j, l, xv), op.scale);
    float2* o = out + bt * M + p;
    if (op.accumulate) v = cadd(v, *o);
    *o = v;
  }
}

// out[r][q] = src[rows[r]*stride + off + q] * w[off + q]
__global__ void gather_rows_kernel(const float2* __restrict__ src, long long stride, int nsrc,
                                   const int32_t* __restrict__ rows, long long off, long long len,
                                   const float* __restrict__ w, float2* __restrict__ out) {
  const long long r = blockIdx.y;
  const int sr = rows[r];
  float2* o = out + r * len;
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < len; q += (long long)gridDim.x * blockDim.x) {
    float2 v = make_float2(__builtin_nanf(""), __builtin_nanf(""));
    if (sr >= 0 && sr < nsrc) {
      v = src[(long long)sr * stride + off + q];
      if (w) v = cscale(v, w[off + q]);
    }
    o[q] = v;
  }
}

static unsigned grid_for(long long n, int bs) {
  long long b = (n + bs - 1) / bs;
  if (b > 65536) b = 65536;
  if (b < 1) b = 1;
  return (unsigned)b;
}

static int check_grid(const pb3d_grid* g) {
  if (!g) return fail(PB3D_ERR_ARG, "null grid");
  for (int i = 0; i < 3; ++i)
    if (g->n[i] < 1 || !(g->boxsize[i] > 0)) return fail(PB3D_ERR_ARG, "invalid grid");
  if (g->n[2] % 2) return fail(PB3D_ERR_ARG, "n2 must be even for real FFT layouts");
  return 0;
}

}  // namespace pb3d

using namespace pb3d;

extern "C" int pb3d_kmap(const pb3d_grid* grid, const void* x, void* out, int batch, const pb3d_kop* kop,
                         void* stream) {
  if (int e = check_grid(grid)) return e;
  if (!x || !out || !kop || batch < 1) return fail(PB3D_ERR_ARG, "kmap: null pointer or batch < 1");
  if (int e = check_kop(kop)) return e;
  if (batch > 65535) return fail(PB3D_ERR_ARG, "kmap: batch too large");
  Geo g = make_geo(*grid);
  KOp op = to_kop(kop);
  dim3 gr(grid_for(g.H, 256), batch);
  kmap_kernel<<<gr, 256, 0, (hipStream_t)stream>>>((const float2*)x, (float2*)out, g, op);
  PB3D_CHECK_LAUNCH();
  return 0;
}

extern "C" int pb3d_grf_device(const pb3d_grid* grid, uint64_t seed, const float* sqrtP_w, float norm, void* out,
                               void* stream) {
  if (int e = check_grid(grid)) return e;
  if (!out) return fail(PB3D_ERR_ARG, "grf_device: null output");
  Geo g = make_geo(*grid);
  grf_device_kernel<<<grid_for(g.H, 256), 256, 0, (hipStream_t)stream>>>(seed, nullptr, sqrtP_w, norm, (float2*)out,
                                                                         g);
  PB3D_CHECK_LAUNCH();
  return 0;
}

extern "C" int pb3d_grf_device_at(const pb3d_grid* grid, const uint64_t* seed_dev, uint64_t offset,
                                  const float* sqrtP_w, float norm, void* out, void* stream) {
  if (int e = check_grid(grid)) return e;
  if (!out || !seed_dev) return fail(PB3D_ERR_ARG, "grf_device_at: null output or seed pointer");
  Geo g = make_geo(*grid);
  grf_device_kernel<<<grid_for(g.H, 256), 256, 0, (hipStream_t)stream>>>(
      offset, (const unsigned long long*)seed_dev, sqrtP_w, norm, (float2*)out, g);
  PB3D_CHECK_LAUNCH();
  return 0;
}

extern "C" int pb3d_kmap_packed(const pb3d_grid* grid, const void* x, const int32_t* modes, int64_t packed_len,
                                void* out, int batch, const pb3d_kop* kop, void* stream) {
  if (int e = check_grid(grid)) return e;
  if (!modes || !out || !kop || batch < 1 || batch > 65535 || packed_len < 1)
    return fail(PB3D_ERR_ARG, "kmap_packed: null pointer or bad sizes");
  if (int e = check_kop(kop)) return e;
  if (kop->kind == 5) return fail(PB3D_ERR_ARG, "kmap_packed: PACK is an R2C epilogue");
  if (kop->kind != 4 && !x) return fail(PB3D_ERR_ARG, "kmap_packed: input required");
  Geo g = make_geo(*grid);
  KOp op = to_kop(kop);
  dim3 gr(grid_for(packed_len, 256), batch);
  kmap_packed_kernel<<<gr, 256, 0, (hipStream_t)stream>>>((const float2*)x, modes, packed_len, (float2*)out, g, op);
  PB3D_CHECK_LAUNCH();
  return 0;
}

extern "C" int pb3d_gather_rows(const void* src, int64_t src_stride, int nsrc, const int32_t* rows, int nrows,
                                int64_t offset, int64_t len, const float* w, void* out, void* stream) {
  if (!src || !rows || !out || nrows < 1 || nrows > 65535 || len < 1 || offset < 0 || nsrc < 1 ||
      offset + len > src_stride)
    return fail(PB3D_ERR_ARG, "gather_rows: null pointer or bad sizes");
  dim3 gr(grid_for(len, 256), nrows);
  gather_rows_kernel<<<gr, 256, 0, (hipStream_t)stream>>>((const float2*)src, src_stride, nsrc, rows, offset, len, w,
                                                          (float2*)out);
  PB3D_CHECK_LAUNCH();
  return 0;
}

extern "C" int pb3d_grf_hermitian(const pb3d_grid* grid, const double* re, const double* im,
                                  const float* sqrtP_w, float norm, void* out, void* stream) {
  if (int e = check_grid(grid)) return e;
  if (!re || !im || !out) return fail(PB3D_ERR_ARG, "grf: null pointer");
  Geo g = make_geo(*grid);
  grf_kernel<<<grid_for(g.H, 256), 256, 0, (hipStream_t)stream>>>(re, im, sqrtP_w, norm, (float2*)out, g);
  PB3D_CHECK_LAUNCH();
  return 0;
}
