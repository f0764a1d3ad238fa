// Mixed-radix line DFTs in LDS for grids outside the power-of-two fast path: column passes with
// the mode prologue / epilogue, real rows with the voxel prologue / epilogue, and the fused
// pointing rows.  The stage algebra is stated in pb3d_fftgen.h.
#include "pb3d_fftgen.h"

namespace pb3d {

bool gen_factor(int L, GenAxis& ax) {
  ax.L = L;
  ax.nr = 0;
  if (L < 1) return false;
  int n = L;
  while (n % 4 == 0) {
    if (ax.nr >= kGenMaxRadices) return false;
    ax.r[ax.nr++] = 4;
    n /= 4;
  }
  for (int p = 2; n > 1; ++p)
    while (n % p == 0) {
      if (ax.nr >= kGenMaxRadices) return false;
      ax.r[ax.nr++] = p;
      n /= p;
    }
  return true;
}

static int cols_per_block(int L) {
  const int l = 2048 / L;
  return l < 1 ? 1 : (l > 16 ? 16 : l);
}
int gen_rows_per_block(int M) {
  const int l = 1024 / M;
  return l < 1 ? 1 : (l > 128 ? 128 : l);
}

__device__ __forceinline__ bool gen_sup(int p, int L, int rad) { return rad < 0 || p <= rad || p >= L - rad; }

// nl lines of length ax.L at a + c*L: all stages, ping-ponging between a and b (block-wide, every
// stage ends with a barrier; the caller's writes to a must be complete).  Returns the buffer that
// holds the result.
template <int DIR>
__device__ float2* line_dft(float2* a, float2* b, const float2* T, const GenAxis& ax, int nl) {
  const int L = ax.L;
  const int tid = threadIdx.x;
  int Ns = 1;
  for (int s = 0; s < ax.nr; ++s) {
    const int R = ax.r[s];
    const int LR = L / R, NsR = Ns * R, ts = L / NsR;
    for (int f = tid; f < nl * L; f += kGenThreads) {
      const int c = f / L, o = f - c * L;
      const int blk = o / NsR, rem = o - blk * NsR;
      const int q = rem / Ns, k = rem - q * Ns;
      // inputs j + r*L/R with j = blk*Ns + k < L/R; twiddle exponent r*step mod L, step < L
      const float2* src = a + c * L + blk * Ns + k;
      const int step = k * ts + q * LR;
      float2 acc = src[0];
      int e = 0;
      for (int r = 1; r < R; ++r) {
        e += step;
        if (e >= L) e -= L;
        float2 w = T[e];
        if (DIR > 0) w.y = -w.y;
        const float2 v = src[r * LR];
        acc.x = fmaf(v.x, w.x, fmaf(-v.y, w.y, acc.x));
        acc.y = fmaf(v.x, w.y, fmaf(v.y, w.x, acc.y));
      }
      b[f] = acc;
    }
    __syncthreads();
    float2* t = a;
    a = b;
    b = t;
    Ns = NsR;
  }
  return a;
}

// X_k (k = 0..M) of a real row from Z = DFT_M(x_even + i x_odd): E = (Z_k + conj Z_{M-k}) / 2,
// O = (Z_k - conj Z_{M-k}) / 2i, X_k = E + e^{-2 pi i k / 2M} O
__device__ __forceinline__ float2 gen_split(const float2* Z, int k, int M, const float2* W) {
  const float2 zk = Z[k == M ? 0 : k];
  const float2 zc = cconj(Z[(k == 0 || k == M) ? 0 : M - k]);
  const float2 e = make_float2(0.5f * (zk.x + zc.x), 0.5f * (zk.y + zc.y));
  const float2 o = make_float2(0.5f * (zk.y - zc.y), -0.5f * (zk.x - zc.x));
  return cadd(e, cmul(W[k], o));
}

// 2 * Z_k of a half-spectrum row (the inverse of gen_split): entries kz >= lz are zero, Im of
// kz = 0 and kz = M are dropped
__device__ __forceinline__ float2 gen_merge(const float2* __restrict__ src, int k, int M, int lz,
                                            const float2* W) {
  float2 a = k < lz ? src[k] : make_float2(0.f, 0.f);
  float2 c = M - k < lz ? src[M - k] : make_float2(0.f, 0.f);
  if (k == 0) { a.y = 0.f; c.y = 0.f; }
  const float2 sm = cadd(a, cconj(c));
  const float2 wd = cmul(cconj(W[k]), csub(a, cconj(c)));
  return make_float2(sm.x - wd.y, sm.y + wd.x);
}

template <int DIR, int MODE>
__global__ void __launch_bounds__(kGenThreads)
gen_cols_kernel(const float2* in, float2* out, GenCols cm, long long in_bs, long long out_bs,
                const float2* __restrict__ tw, GenAxis ax, Geo geo, KOp kop, int lpb) {
  extern __shared__ float2 gsm[];
  const int L = ax.L, tid = threadIdx.x;
  float2* T = gsm;
  float2* a = T + L;
  float2* b = a + lpb * L;
  for (int i = tid; i < L; i += kGenThreads) T[i] = tw[i];
  const long long bt = blockIdx.y;
  const long long c0 = (long long)blockIdx.x * lpb;
  const float2* src = in ? in + (MODE == 1 ? bt / kop.x_group : bt) * in_bs : nullptr;
  // element p of the lpb columns side by side: consecutive lanes read consecutive columns
  for (int f = tid; f < lpb * L; f += kGenThreads) {
    const int p = f / lpb, cl = f - p * lpb;
    const long long c = c0 + cl;
    float2 v = make_float2(0.f, 0.f);
    if (c < cm.ncols) {
      const int q = (int)(c / cm.cpp), l = (int)(c % cm.cpp);
      const bool zrow = MODE == 0 && cm.row_zero && !cm.row_zero[(long long)q * cm.row_n1 + p];
      if (!zrow && gen_sup(p, L, cm.rad_in) && gen_sup(q, cm.qlen, cm.qrad_in)) {
        if (MODE != 1 || kop.kind != 4) v = src[(long long)q * cm.in_plane + l + (long long)p * cm.in_S];
        if (MODE == 1) v = kop_value(kop, geo, bt, (long long)q * cm.plane + l + (long long)p * cm.S, p, q, l, v);
      }
    }
    a[cl * L + p] = v;
  }
  __syncthreads();
  const float2* res = line_dft<DIR>(a, b, T, ax, lpb);
  float2* dst = out + bt * out_bs;
  for (int f = tid; f < lpb * L; f += kGenThreads) {
    const int p = f / lpb, cl = f - p * lpb;
    const long long c = c0 + cl;
    if (c >= cm.ncols) continue;
    const int q = (int)(c / cm.cpp), l = (int)(c % cm.cpp);
    if (!gen_sup(p, L, cm.rad_out) || !gen_sup(q, cm.qlen, cm.qrad_out)) continue;
    float2 v = res[cl * L + p];
    const long long od = (long long)q * cm.out_plane + l + (long long)p * cm.out_S;
    if (MODE == 2) {
      const long long md = (long long)q * cm.plane + l + (long long)p * cm.S;
      if (kop.kind == 5) {  // packed output (in-bin modes only)
        const int pi = kop.pack_index[md];
        if (pi < 0) continue;
        v = cscale(v, kop.scale);
        if (kop.accumulate) v = cadd(v, dst[pi]);
        dst[pi] = v;
        continue;
      }
      v = cscale(kop_value(kop, geo, bt, md, p, q, l, v), kop.scale);
      if (kop.accumulate) v = cadd(v, dst[od]);
    }
    if (MODE == 0 && cm.row_on && !cm.row_on[(long long)q * cm.row_n1 + p]) continue;
    dst[od] = v;
  }
}

__global__ void __launch_bounds__(kGenThreads)
gen_r2c_rows_kernel(const float* in, float2* __restrict__ out, long long nrows, int rpb,
                    const float2* __restrict__ tw, const float2* __restrict__ twN, GenAxis ax, Geo geo, ROp rop,
                    int lz_out, int pitch, int lpb) {
  extern __shared__ float2 gsm[];
  const int M = ax.L, tid = threadIdx.x;
  float2* T = gsm;
  float2* W = T + M;
  float2* a = W + M + 1;
  float2* b = a + lpb * M;
  for (int i = tid; i < M; i += kGenThreads) T[i] = tw[i];
  for (int i = tid; i <= M; i += kGenThreads) W[i] = twN[i];
  const long long row0 = (long long)blockIdx.x * lpb;
  for (int f = tid; f < lpb * M; f += kGenThreads) {
    const int r = f / M, m = f - r * M;
    const long long row = row0 + r;
    float2 v = make_float2(0.f, 0.f);
    if (row < nrows) {
      const long long bt = row / rpb, rr = row - bt * rpb;
      const long long vox = rr * (2 * M) + 2 * m;
      const float* A = rop.a ? rop.a + bt * rop.sa : in + bt * (long long)rpb * 2 * M;
      float x0 = A[vox], x1 = A[vox + 1];
      if (rop.kind != 0) {
        const int i0 = (int)(rr / geo.n1), i1 = (int)(rr % geo.n1);
        x0 = rop_value(rop, geo, bt, vox, i0, i1, 2 * m, x0);
        x1 = rop_value(rop, geo, bt, vox + 1, i0, i1, 2 * m + 1, x1);
      }
      v = make_float2(x0 * rop.scale, x1 * rop.scale);
    }
    a[f] = v;
  }
  __syncthreads();
  const float2* Z = line_dft<-1>(a, b, T, ax, lpb);
  const int K = M + 1;
  for (int f = tid; f < lpb * K; f += kGenThreads) {
    const int r = f / K, k = f - r * K;
    const long long row = row0 + r;
    if (row >= nrows || k >= lz_out) continue;
    out[row * pitch + k] = gen_split(Z + r * M, k, M, W);
  }
}

__global__ void __launch_bounds__(kGenThreads)
gen_c2r_rows_kernel(const float2* __restrict__ in, float* out, long long nrows, int rpb,
                    const float2* __restrict__ tw, const float2* __restrict__ twN, GenAxis ax, Geo geo, ROp rop,
                    float norm, int lz_in, int pitch, int lpb) {
  extern __shared__ float2 gsm[];
  const int M = ax.L, tid = threadIdx.x;
  float2* T = gsm;
  float2* W = T + M;
  float2* a = W + M + 1;
  float2* b = a + lpb * M;
  for (int i = tid; i < M; i += kGenThreads) T[i] = tw[i];
  for (int i = tid; i <= M; i += kGenThreads) W[i] = twN[i];
  __syncthreads();  // W is read by the merge below
  const long long row0 = (long long)blockIdx.x * lpb;
  for (int f = tid; f < lpb * M; f += kGenThreads) {
    const int r = f / M, k = f - r * M;
    const long long row = row0 + r;
    a[f] = row < nrows ? gen_merge(in + row * pitch, k, M, lz_in, W) : make_float2(0.f, 0.f);
  }
  __syncthreads();
  const float2* Z = line_dft<+1>(a, b, T, ax, lpb);
  for (int f = tid; f < lpb * M; f += kGenThreads) {
    const int r = f / M, m = f - r * M;
    const long long row = row0 + r;
    if (row >= nrows) continue;
    const long long bt = row / rpb, rr = row - bt * rpb;
    const long long vox = rr * (2 * M) + 2 * m;
    float x0 = Z[f].x * norm, x1 = Z[f].y * norm;
    if (rop.kind != 0) {
      const int i0 = (int)(rr / geo.n1), i1 = (int)(rr % geo.n1);
      x0 = rop_value(rop, geo, bt, vox, i0, i1, 2 * m, x0);
      x1 = rop_value(rop, geo, bt, vox + 1, i0, i1, 2 * m + 1, x1);
    }
    x0 *= rop.scale;
    x1 *= rop.scale;
    float* o = out + bt * (long long)rpb * 2 * M;
    if (rop.accumulate) { x0 += o[vox]; x1 += o[vox + 1]; }
    o[vox] = x0;
    o[vox + 1] = x1;
  }
}

__global__ void __launch_bounds__(kGenThreads)
gen_pointing_kernel(const float2* __restrict__ work, long long slot_in, int pitch_in, float2* __restrict__ out,
                    long long slot_out, int pitch_out, long long rpm, int batch, GenSubs gs, GenIc ic,
                    const float* __restrict__ mask, float scale, int lz_in, int lz_out,
                    const float2* __restrict__ tw, const float2* __restrict__ twN, GenAxis ax, Geo geo, int lpb) {
  extern __shared__ float2 gsm[];
  __shared__ float icl[kGenThreads / 64][kGenMaxIcBins];
  const int M = ax.L, tid = threadIdx.x;
  float2* T = gsm;
  float2* W = T + M;
  float2* a = W + M + 1;
  float2* b = a + lpb * M;
  float* acc = reinterpret_cast<float*>(b + lpb * M);  // lpb * 2M voxels
  for (int i = tid; i < M; i += kGenThreads) T[i] = tw[i];
  for (int i = tid; i <= M; i += kGenThreads) W[i] = twN[i];
  const long long t = blockIdx.y;
  const long long row0 = (long long)blockIdx.x * lpb;
  if (gs.row_on) {
    // rows whose mask (and IC weight) is zero everywhere: a block with no other row leaves at once
    // (its IC partials are zero); the inverse y pass did not store them and the forward one reads
    // them as zeros.  In a block that runs, their inputs are taken as zeros below.
    bool any = false;
    for (int r = tid; r < lpb; r += kGenThreads) any = any || (row0 + r < rpm && gs.row_on[row0 + r] != 0);
    if (!__syncthreads_or(any)) {
      if (ic.weight && tid < ic.nbins) ic.partial[(t * gridDim.x + blockIdx.x) * ic.nbins + tid] = 0.;
      return;
    }
  }
  for (int i = tid; i < lpb * 2 * M; i += kGenThreads) acc[i] = 0.f;
  for (int g = 0; g < gs.n; ++g) {
    __syncthreads();  // tables staged; the previous sub-map's reads of a and b are done
    const float2* src = work + ((long long)g * batch + t) * slot_in;
    for (int f = tid; f < lpb * M; f += kGenThreads) {
      const int r = f / M, k = f - r * M;
      const long long row = row0 + r;
      const bool on = row < rpm && (!gs.row_on || gs.row_on[row] != 0);
      a[f] = on ? gen_merge(src + row * pitch_in, k, M, lz_in, W) : make_float2(0.f, 0.f);
    }
    __syncthreads();
    const float2* Z = line_dft<+1>(a, b, T, ax, lpb);
    const int ell = gs.ell[g], mm = gs.m[g];
    const float cf = gs.coef[g];
    for (int f = tid; f < lpb * M; f += kGenThreads) {
      const int r = f / M, m = f - r * M;
      const long long row = row0 + r;
      if (row >= rpm) continue;
      const float2 z = Z[f];
      float w0 = cf, w1 = cf;
      if (ell > 0) {
        const int i0 = (int)(row / geo.n1), i1 = (int)(row % geo.n1);
        float x, y, zz;
        unit_r(geo, i0, i1, 2 * m, x, y, zz);
        w0 = cf * ylm_f(ell, mm, x, y, zz);
        unit_r(geo, i0, i1, 2 * m + 1, x, y, zz);
        w1 = cf * ylm_f(ell, mm, x, y, zz);
      }
      acc[2 * f] = fmaf(w0, z.x, acc[2 * f]);
      acc[2 * f + 1] = fmaf(w1, z.y, acc[2 * f + 1]);
    }
  }
  __syncthreads();  // every sub-map's reads of a and b are done
  const float norm = scale / (float)geo.N;
  const int nv = lpb * 2 * M;
  if (ic.weight) {
    // per-bin sums of acc * norm * weight: per voxel slot a wave-wide sum per bin present (fixed
    // order), added by lane 0 to the wave's FP32 sums; the block's FP64 sum goes to its own slot
    const int wave = tid / 64, lane = tid % 64;
    if (lane < ic.nbins) icl[wave][lane] = 0.f;
    __syncthreads();
    for (int f0 = 0; f0 < nv; f0 += kGenThreads) {
      const int f = f0 + tid;
      int z = kGenMaxIcBins;
      float c = 0.f;
      if (f < nv) {
        const int r = f / (2 * M);
        const long long row = row0 + r;
        if (row < rpm) {
          const long long v0 = row * (2 * M) + (f - r * 2 * M);
          z = ic.bin ? ic.bin[v0] : 0;
          c = acc[f] * norm * ic.weight[v0];
        }
      }
      unsigned long long act = __ballot(z < ic.nbins);
      while (act) {
        const int k = __shfl(z, __ffsll((long long)act) - 1);
        float v = z == k ? c : 0.f;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
        if (lane == 0) icl[wave][k] += v;
        act &= ~__ballot(z == k);
      }
    }
    __syncthreads();
    if (tid < ic.nbins) {
      double s = 0.;
      for (int w = 0; w < kGenThreads / 64; ++w) s += (double)icl[w][tid];
      ic.partial[(t * gridDim.x + blockIdx.x) * ic.nbins + tid] = s;
    }
  }
  for (int f = tid; f < lpb * M; f += kGenThreads) {
    const int r = f / M, m = f - r * M;
    const long long row = row0 + r;
    float2 v = make_float2(0.f, 0.f);
    if (row < rpm) {
      float s0 = norm, s1 = norm;
      if (mask) {
        s0 *= mask[row * (2 * M) + 2 * m];
        s1 *= mask[row * (2 * M) + 2 * m + 1];
      }
      v = make_float2(acc[2 * f] * s0, acc[2 * f + 1] * s1);
    }
    a[f] = v;
  }
  __syncthreads();
  const float2* Z = line_dft<-1>(a, b, T, ax, lpb);
  float2* o = out + t * slot_out;
  const int K = M + 1;
  for (int f = tid; f < lpb * K; f += kGenThreads) {
    const int r = f / K, k = f - r * K;
    const long long row = row0 + r;
    if (row >= rpm || k >= lz_out) continue;
    o[row * pitch_out + k] = gen_split(Z + r * M, k, M, W);
  }
}

// sums[t * nbins + z] += the nwm block partials of map t, in block order
__global__ void gen_ic_finalize(const double* __restrict__ partial, long long nwm, int batch, int nbins,
                                double* __restrict__ sums) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long long)batch * nbins) return;
  const long long t = i / nbins;
  const int z = (int)(i % nbins);
  double s = 0.;
  for (long long w = 0; w < nwm; ++w) s += partial[(t * nwm + w) * nbins + z];
  sums[i] += s;
}

hipError_t gen_cols(int dir, int mode, const GenAxis& ax, const float2* in, float2* out, const GenCols& cm,
                    long long in_bs, long long out_bs, int batch, const float2* tw, const Geo& g, const KOp& kop,
                    hipStream_t s) {
  if (cm.ncols <= 0) return hipSuccess;
  if (batch < 1 || batch > 65535 || cm.cpp < 1) return hipErrorInvalidValue;
  const int lpb = cols_per_block(ax.L);
  const long long nb = (cm.ncols + lpb - 1) / lpb;
  if (nb > 0x7fffffffLL) return hipErrorInvalidValue;
  const size_t lds = sizeof(float2) * ((size_t)ax.L + 2 * (size_t)lpb * ax.L);
  dim3 grid((unsigned)nb, batch);
#define PB3D_GCOLS(D, MD) \
  gen_cols_kernel<D, MD><<<grid, kGenThreads, lds, s>>>(in, out, cm, in_bs, out_bs, tw, ax, g, kop, lpb)
  if (dir < 0) {
    if (mode == 0) PB3D_GCOLS(-1, 0);
    else if (mode == 2) PB3D_GCOLS(-1, 2);
    else return hipErrorInvalidValue;
  } else {
    if (mode == 0) PB3D_GCOLS(+1, 0);
    else if (mode == 1) PB3D_GCOLS(+1, 1);
    else return hipErrorInvalidValue;
  }
#undef PB3D_GCOLS
  return hipGetLastError();
}

hipError_t gen_r2c_rows(const GenAxis& axM, const float* in, float2* out, long long nrows, int rpb,
                        const float2* tw, const float2* twN, const Geo& g, const ROp& rop, int lz_out, int pitch,
                        hipStream_t s) {
  if (nrows <= 0) return hipSuccess;
  const int M = axM.L, lpb = gen_rows_per_block(M);
  const long long nb = (nrows + lpb - 1) / lpb;
  if (nb > 0x7fffffffLL) return hipErrorInvalidValue;
  const size_t lds = sizeof(float2) * ((size_t)2 * M + 1 + 2 * (size_t)lpb * M);
  gen_r2c_rows_kernel<<<(unsigned)nb, kGenThreads, lds, s>>>(in, out, nrows, rpb, tw, twN, axM, g, rop, lz_out,
                                                             pitch, lpb);
  return hipGetLastError();
}

hipError_t gen_c2r_rows(const GenAxis& axM, const float2* in, float* out, long long nrows, int rpb,
                        const float2* tw, const float2* twN, const Geo& g, const ROp& rop, float norm, int lz_in,
                        int pitch, hipStream_t s) {
  if (nrows <= 0) return hipSuccess;
  const int M = axM.L, lpb = gen_rows_per_block(M);
  const long long nb = (nrows + lpb - 1) / lpb;
  if (nb > 0x7fffffffLL) return hipErrorInvalidValue;
  const size_t lds = sizeof(float2) * ((size_t)2 * M + 1 + 2 * (size_t)lpb * M);
  gen_c2r_rows_kernel<<<(unsigned)nb, kGenThreads, lds, s>>>(in, out, nrows, rpb, tw, twN, axM, g, rop, norm, lz_in,
                                                             pitch, lpb);
  return hipGetLastError();
}

hipError_t gen_pointing_rows(const GenAxis& axM, const float2* work, long long slot_in, int pitch_in, float2* out,
                             long long slot_out, int pitch_out, long long rpm, int batch, const GenSubs& gs,
                             const GenIc& ic, const float* mask, float scale, int lz_in, int lz_out,
                             const float2* tw, const float2* twN, const Geo& g, hipStream_t s) {
  if (rpm <= 0) return hipSuccess;
  if (batch < 1 || batch > 65535 || gs.n < 1 || gs.n > kGenMaxSub) return hipErrorInvalidValue;
  if (ic.weight && (ic.nbins < 1 || ic.nbins > kGenMaxIcBins)) return hipErrorInvalidValue;
  const int M = axM.L, lpb = gen_rows_per_block(M);
  const long long nb = (rpm + lpb - 1) / lpb;
  if (nb > 0x7fffffffLL) return hipErrorInvalidValue;
  // tables (2M + 1), two line buffers (2 lpb M) and the voxel accumulators (lpb 2M floats)
  const size_t lds = sizeof(float2) * ((size_t)2 * M + 1 + 3 * (size_t)lpb * M);
  dim3 grid((unsigned)nb, batch);
  gen_pointing_kernel<<<grid, kGenThreads, lds, s>>>(work, slot_in, pitch_in, out, slot_out, pitch_out, rpm, batch, gs,
                                                     ic, mask, scale, lz_in, lz_out, tw, twN, axM, g, lpb);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || !ic.weight) return e;
  const long long nf = (long long)batch * ic.nbins;
  gen_ic_finalize<<<(unsigned)((nf + 255) / 256), 256, 0, s>>>(ic.partial, nb, batch, ic.nbins, ic.sums);
  return hipGetLastError();
}

}  // namespace pb3d
