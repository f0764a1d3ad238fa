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
    float2 v = kop_value(op, g, bt, e, i, j, l, x[bt * op.x_stride + e]);
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

extern "C" int pb3d_grf_hermitian(const pb3d_grid* grid, const double* re, const double* im,
                                  const float* sqrtP_w, float norm, void* out, void* stream) {
  if (int e = check_grid(grid)) return e;
  if (!re || !im || !out) return fail(PB3D_ERR_ARG, "grf: null pointer");
  Geo g = make_geo(*grid);
  grf_kernel<<<grid_for(g.H, 256), 256, 0, (hipStream_t)stream>>>(re, im, sqrtP_w, norm, (float2*)out, g);
  PB3D_CHECK_LAUNCH();
  return 0;
}
