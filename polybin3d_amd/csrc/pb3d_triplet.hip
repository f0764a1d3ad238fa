// Triplet sums T[t] = sum_x g_a(x) g_b(x) g_c(x) over a stack of real maps (assemble_b3 /
// sum_triplet, PolyBin3D/cython/utils.pyx:357-366, 614-636).
//
// precision 0: FP32 products, FP64 accumulation (vector ALU).  Each block owns a group of
// triplets and a voxel chunk; consecutive blocks share the voxel chunk so the maps are reused from
// L2; per-chunk partials are reduced in a fixed order.
#include "pb3d_ops.h"

namespace pb3d {

constexpr int kTG = 16;  // triplets per block

__global__ void __launch_bounds__(256)
triplet_vec_kernel(const float* __restrict__ g, int nmaps, long long nvox, const int* __restrict__ i1,
                   const int* __restrict__ i2, const int* __restrict__ i3, int ntrip, long long chunk,
                   double* __restrict__ partial) {
  __shared__ double red[kTG][256 / 64];
  const int t0 = blockIdx.x * kTG;
  const long long start = (long long)blockIdx.y * chunk;
  long long end = start + chunk;
  if (end > nvox) end = nvox;
  const float* pa[kTG]; const float* pb[kTG]; const float* pc[kTG];
  bool bad[kTG];
#pragma unroll
  for (int k = 0; k < kTG; ++k) {
    const int t = t0 + k < ntrip ? t0 + k : ntrip - 1;
    int a = i1[t], b = i2[t], c = i3[t];
    bad[k] = a < 0 || a >= nmaps || b < 0 || b >= nmaps || c < 0 || c >= nmaps;
    if (bad[k]) { a = 0; b = 0; c = 0; }  // out-of-range index: result is NaN, no stray read
    pa[k] = g + (long long)a * nvox;
    pb[k] = g + (long long)b * nvox;
    pc[k] = g + (long long)c * nvox;
  }
  double acc[kTG];
#pragma unroll
  for (int k = 0; k < kTG; ++k) acc[k] = 0.;
  for (long long x = start + threadIdx.x; x < end; x += blockDim.x) {
#pragma unroll
    for (int k = 0; k < kTG; ++k) acc[k] += (double)(pa[k][x] * pb[k][x] * pc[k][x]);
  }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < kTG; ++k) {
    double v = acc[k];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off);
    if (lane == 0) red[k][w] = v;
  }
  __syncthreads();
  if (threadIdx.x < kTG && t0 + threadIdx.x < ntrip) {
    double s = 0.;
    for (int i = 0; i < 256 / 64; ++i) s += red[threadIdx.x][i];
    bool isbad = false;
#pragma unroll
    for (int k = 0; k < kTG; ++k) if (k == (int)threadIdx.x) isbad = bad[k];
    if (isbad) s = __builtin_nan("");
    partial[(long long)blockIdx.y * ntrip + t0 + threadIdx.x] = s;
  }
}

__global__ void triplet_finalize(const double* __restrict__ partial, int nchunks, int ntrip,
                                 double* __restrict__ out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ntrip) return;
  double s = 0.;
  for (int c = 0; c < nchunks; ++c) s += partial[(long long)c * ntrip + t];
  out[t] = s;
}

static long long trip_chunks(long long nvox) {
  long long c = (nvox + 65535) / 65536;  // ~64k voxels per chunk
  if (c < 1) c = 1;
  if (c > 4096) c = 4096;
  return c;
}

}  // namespace pb3d

using namespace pb3d;

extern "C" size_t pb3d_triplet_work_size(int64_t nvox, int ntrip, int precision) {
  (void)precision;
  if (nvox < 1 || ntrip < 1) return 0;
  return (size_t)trip_chunks(nvox) * ntrip * sizeof(double);
}

extern "C" int pb3d_triplet_sum(const float* g, int nmaps, int64_t nvox, const int32_t* i1, const int32_t* i2,
                                const int32_t* i3, int ntrip, int precision, double* out, void* work,
                                void* stream) {
  if (!g || !i1 || !i2 || !i3 || !out || !work) return fail(PB3D_ERR_ARG, "triplet_sum: null pointer");
  if (nmaps < 1 || nvox < 1 || ntrip < 1) return fail(PB3D_ERR_ARG, "triplet_sum: empty input");
  if (precision != 0) return fail(PB3D_ERR_UNSUPPORTED, "triplet_sum: precision mode not built");
  hipStream_t s = (hipStream_t)stream;
  const long long nc = trip_chunks(nvox);
  long long chunk = (nvox + nc - 1) / nc;
  dim3 gr((unsigned)((ntrip + kTG - 1) / kTG), (unsigned)nc);
  triplet_vec_kernel<<<gr, 256, 0, s>>>(g, nmaps, nvox, i1, i2, i3, ntrip, chunk, (double*)work);
  PB3D_CHECK_LAUNCH();
  triplet_finalize<<<(ntrip + 255) / 256, 256, 0, s>>>((const double*)work, (int)nc, ntrip, out);
  PB3D_CHECK_LAUNCH();
  return 0;
}
