// k-bin reductions: one coalesced pass over the half grid, wavefront-shuffle segmented sums keyed
// by (row, bin), FP64 LDS accumulation per one-wave block, fixed-order FP64 reduction across blocks.
// Replaces IntegrationUtils.integrate / cross_integrate / real_integrate_double
// (PolyBin3D/cython/utils.pyx:51-134), which sweep the whole grid once per bin.
//
// Along a row (fixed kx, ky) |k| grows with kz, so the bin index is non-decreasing along the
// contiguous axis: equal-bin modes form contiguous runs and a 64-lane segmented scan reduces each
// run to one value.  The run totals of a wave step are added to the block's LDS histogram one run
// at a time in lane order (a ballot over the run tails), by the lanes o < nout in parallel: every
// sum is formed in the same order on every call, so the results are bitwise reproducible.
#include "pb3d_ops.h"

#include <mutex>

namespace pb3d {

constexpr int kMaxOut = 16;
struct Ells { int n; int l1[kMaxOut]; int l2[kMaxOut]; };

constexpr int kBinThreads = 64;  // one wave per block

static long long bin_blocks(long long H, int nq) {
  long long gx = 8192 / (nq > 0 ? nq : 1);
  if (gx < 1) gx = 1;
  long long maxb = (H + kBinThreads - 1) / kBinThreads;
  if (gx > maxb) gx = maxb;
  return gx;
}

template <int MODE>
__global__ void __launch_bounds__(kBinThreads)
bin_reduce_kernel(const void* __restrict__ xv, long long xs, const float2* __restrict__ y, long long yq,
                  long long yo, Ells el, const uint8_t* __restrict__ table, int nbins, Geo g, long long chunk,
                  const int32_t* __restrict__ modes, long long count, double* __restrict__ partial) {
  extern __shared__ double hist[];
  const int nout = el.n;
  const int nh = nout * nbins;
  for (int i = threadIdx.x; i < nh; i += blockDim.x) hist[i] = 0.;
  __syncthreads();
  const long long q = blockIdx.y;
  const long long start = (long long)blockIdx.x * chunk;
  long long end = start + chunk;
  if (end > count) end = count;  // count = H, or the length of the mode list
  const int lane = threadIdx.x & 63;
  const long long n12 = (long long)g.n1 * g.n2h;
  bool need_mu = false;
  for (int o = 0; o < nout; ++o) need_mu |= (el.l1[o] != 0 || el.l2[o] != 0);
  const float2* x = MODE == 0 ? (const float2*)xv + q * xs : nullptr;
  const float* xr = MODE == 1 ? (const float*)xv + q * xs : nullptr;
  const float2* yb = y ? y + q * yq : nullptr;
  for (long long base = start; base < end; base += blockDim.x) {
    const long long idx = base + threadIdx.x;
    const bool valid = idx < end;
    // with a mode list the lanes walk the listed (ascending) modes: runs of one (row, bin) stay
    // contiguous, so the segmented scan below is unchanged
    const long long e = modes ? (valid ? (long long)modes[idx] : 0) : idx;
    const int key = valid ? table[e] : PB3D_NO_BIN;
    // Most of the half grid lies outside every bin (the bins end well below the Nyquist modes): a
    // wave whose 64 modes are all outside skips the loads and the scan (wave-uniform branch; the
    // loop holds no block barrier).
    if (!__any(key != PB3D_NO_BIN)) continue;
    const int row = valid ? (int)(e / g.n2h) : -1;
    float v[kMaxOut];
#pragma unroll
    for (int o = 0; o < kMaxOut; ++o) v[o] = 0.f;
    if (key != PB3D_NO_BIN) {
      const int l = (int)(e % g.n2h);
      const float degk = l > 0 ? 2.f : 1.f;
      float mu = 0.f;
      if (need_mu) {
        const int i = (int)(e / n12), j = (int)((e % n12) / g.n2h);
        float kx, ky, kz; kvec(g, i, j, l, kx, ky, kz);
        const float k2 = kx * kx + ky * ky + kz * kz;
        mu = k2 > 0.f ? (g.s0 * kx + g.s1 * ky + g.s2 * kz) * rsqrtf(k2) : 0.f;
      }
      float2 a = make_float2(0.f, 0.f);
      float ar = 0.f;
      if (MODE == 0) a = x[e]; else ar = xr[e];
#pragma unroll
      for (int o = 0; o < kMaxOut; ++o) {
        if (o < nout) {
          float s;
          if (MODE == 0) {
            const float2 b = yb ? yb[o * yo + e] : a;
            s = degk * (a.x * b.x + a.y * b.y);
          } else {
            s = degk * ar * ar;
          }
          if (el.l1[o] != 0 || el.l2[o] != 0) s *= legendre_f(el.l1[o], mu) * legendre_f(el.l2[o], mu);
          v[o] = s;
        }
      }
    }
    // segmented inclusive scan over the wavefront, segments = equal (row, bin); row < 2^23
    const int seg = ((row + 1) << 8) | key;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const int so = __shfl_up(seg, off);
      const bool take = lane >= off && so == seg;
#pragma unroll
      for (int o = 0; o < kMaxOut; ++o) {
        if (o < nout) {
          const float vo = __shfl_up(v[o], off);
          if (take) v[o] += vo;
        }
      }
    }
    const int sn = __shfl_down(seg, 1);
    const bool tail = (lane == 63) || sn != seg;
    // run totals into the histogram in lane order: lane o adds output o of each run in turn
    unsigned long long runs = __ballot(tail && key != PB3D_NO_BIN);
    while (runs) {
      const int src = __ffsll((long long)runs) - 1;
      runs &= runs - 1;
      const int k = __shfl(key, src);
      float mine = 0.f;
#pragma unroll
      for (int o = 0; o < kMaxOut; ++o) {
        if (o < nout) {
          const float vo = __shfl(v[o], src);
          if (lane == o) mine = vo;
        }
      }
      if (lane < nout) hist[lane * nbins + k] += (double)mine;
    }
  }
  __syncthreads();
  double* p = partial + ((long long)blockIdx.x * gridDim.y + q) * nh;
  for (int i = threadIdx.x; i < nh; i += blockDim.x) p[i] = hist[i];
}

__global__ void bin_finalize(const double* __restrict__ partial, long long gx, long long nq, int nh,
                             double* __restrict__ out) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nq * nh) return;
  const long long q = t / nh;
  const int i = (int)(t % nh);
  double s = 0.;
  for (long long b = 0; b < gx; ++b) s += partial[(b * nq + q) * nh + i];
  out[t] = s;
}

// Fixed device scratch of pb3d_bin_reduce_modes_noalloc (the block partials of one call): 16 MiB per
// GPU, held until the process ends.  Host code reaches it only through the per-device address from
// hipGetSymbolAddress, never through the symbol itself.
constexpr long long kNoallocDoubles = 2LL << 20;
static __device__ __align__(256) double scratch[kNoallocDoubles];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kMaxDevices] = {};

// the current device's address of `scratch`, looked up on the first call on that device
static int scratch_address(double** ptr) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices)
    return fail(PB3D_ERR_HIP, "bin_reduce_modes_noalloc: no current device (or device index >= 64)");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(PB3D_ERR_HIP, std::string("bin_reduce_modes_noalloc: hipGetSymbolAddress: ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = p;
  }
  *ptr = (double*)g_scratch_ptr[dev];
  return 0;
}

}  // namespace pb3d

using namespace pb3d;

extern "C" size_t pb3d_bin_reduce_work_size(const pb3d_grid* grid, int nq, int nout, int nbins) {
  if (!grid || nq < 1 || nout < 1 || nbins < 1) return 0;
  Geo g = make_geo(*grid);
  return (size_t)bin_blocks(g.H, nq) * nq * nout * nbins * sizeof(double);
}

static int bin_reduce_impl(const pb3d_grid* grid, const void* x, int64_t x_stride, const void* y,
                           int64_t y_qstride, int64_t y_ostride, int nq, int nout, const int32_t* ells,
                           const int32_t* ells2, int mode, const uint8_t* bin_table, int nbins,
                           const int32_t* modes, int64_t nmodes, double* out, void* work, void* stream) {
  if (!grid || !x || !bin_table || !out || !work) return fail(PB3D_ERR_ARG, "bin_reduce: null pointer");
  if (nq < 1 || nq > 65535) return fail(PB3D_ERR_ARG, "bin_reduce: nq out of range [1, 65535]");
  if (nout < 1 || nout > kMaxOut) return fail(PB3D_ERR_ARG, "bin_reduce: nout out of range [1, 16]");
  if (nbins < 1 || nbins > 254) return fail(PB3D_ERR_ARG, "bin_reduce: nbins out of range [1, 254]");
  if (mode != 0 && mode != 1) return fail(PB3D_ERR_ARG, "bin_reduce: mode must be 0 or 1");
  if (mode == 1 && y) return fail(PB3D_ERR_ARG, "bin_reduce: real mode takes no y");
  Ells el{};
  el.n = nout;
  for (int o = 0; o < nout; ++o) {
    el.l1[o] = ells ? ells[o] : 0;
    el.l2[o] = ells2 ? ells2[o] : 0;
    if (el.l1[o] < 0 || el.l1[o] > 6 || el.l2[o] < 0 || el.l2[o] > 6)
      return fail(PB3D_ERR_ARG, "bin_reduce: ell out of range [0, 6]");
  }
  Geo g = make_geo(*grid);
  if (modes && (nmodes < 0 || nmodes > g.H)) return fail(PB3D_ERR_ARG, "bin_reduce: nmodes out of range [0, H]");
  const long long count = modes ? (long long)nmodes : g.H;
  const long long gx = bin_blocks(count > 0 ? count : 1, nq);
  long long chunk = (count + gx - 1) / gx;
  chunk = (chunk + kBinThreads - 1) / kBinThreads * kBinThreads;
  const int nh = nout * nbins;
  dim3 gr((unsigned)gx, (unsigned)nq);
  hipStream_t s = (hipStream_t)stream;
  const size_t shm = (size_t)nh * sizeof(double);
  if (mode == 0)
    bin_reduce_kernel<0><<<gr, kBinThreads, shm, s>>>(x, x_stride, (const float2*)y, y_qstride, y_ostride, el, bin_table,
                                              nbins, g, chunk, modes, count, (double*)work);
  else
    bin_reduce_kernel<1><<<gr, kBinThreads, shm, s>>>(x, x_stride, nullptr, 0, 0, el, bin_table, nbins, g, chunk,
                                              modes, count, (double*)work);
  PB3D_CHECK_LAUNCH();
  const long long tot = (long long)nq * nh;
  bin_finalize<<<(unsigned)((tot + 255) / 256), 256, 0, s>>>((const double*)work, gx, nq, nh, out);
  PB3D_CHECK_LAUNCH();
  return 0;
}

extern "C" int pb3d_bin_reduce(const pb3d_grid* grid, const void* x, int64_t x_stride, const void* y,
                               int64_t y_qstride, int64_t y_ostride, int nq, int nout, const int32_t* ells,
                               const int32_t* ells2, int mode, const uint8_t* bin_table, int nbins,
                               double* out, void* work, void* stream) {
  return bin_reduce_impl(grid, x, x_stride, y, y_qstride, y_ostride, nq, nout, ells, ells2, mode, bin_table, nbins,
                         nullptr, 0, out, work, stream);
}

extern "C" int pb3d_bin_reduce_modes(const pb3d_grid* grid, const void* x, int64_t x_stride, const void* y,
                                     int64_t y_qstride, int64_t y_ostride, int nq, int nout, const int32_t* ells,
                                     const int32_t* ells2, int mode, const uint8_t* bin_table, int nbins,
                                     const int32_t* modes, int64_t nmodes, double* out, void* work, void* stream) {
  if (!modes) return fail(PB3D_ERR_ARG, "bin_reduce_modes: null mode list");
  return bin_reduce_impl(grid, x, x_stride, y, y_qstride, y_ostride, nq, nout, ells, ells2, mode, bin_table, nbins,
                         modes, nmodes, out, work, stream);
}

extern "C" int pb3d_bin_reduce_modes_noalloc(const pb3d_grid* grid, const void* x, int64_t x_stride, const void* y,
                                             int64_t y_qstride, int64_t y_ostride, int nq, int nout,
                                             const int32_t* ells, const int32_t* ells2, int mode,
                                             const uint8_t* bin_table, int nbins, const int32_t* modes, int64_t nmodes,
                                             double* out, void* stream) {
  if (!grid || !modes) return fail(PB3D_ERR_ARG, "bin_reduce_modes_noalloc: null grid or mode list");
  if (nq < 1 || nq > 65535 || nout < 1 || nout > kMaxOut || nbins < 1 || nbins > 254)
    return fail(PB3D_ERR_ARG, "bin_reduce_modes_noalloc: nq, nout or nbins out of range");
  const Geo g = make_geo(*grid);
  if (nmodes < 0 || nmodes > g.H) return fail(PB3D_ERR_ARG, "bin_reduce_modes_noalloc: nmodes out of range [0, H]");
  const long long gx = bin_blocks(nmodes > 0 ? nmodes : 1, nq);
  const long long need = gx * nq * (long long)nout * nbins;
  if (need > kNoallocDoubles)
    return fail(PB3D_ERR_ARG, "bin_reduce_modes_noalloc: this call needs " + std::to_string(need) +
                                  " doubles of scratch; the fixed buffer holds " + std::to_string(kNoallocDoubles) +
                                  " (use pb3d_bin_reduce_modes with a workspace)");
  double* ptr = nullptr;
  if (int e = scratch_address(&ptr)) return e;
  return bin_reduce_impl(grid, x, x_stride, y, y_qstride, y_ostride, nq, nout, ells, ells2, mode, bin_table, nbins,
                         modes, nmodes, out, ptr, stream);
}
