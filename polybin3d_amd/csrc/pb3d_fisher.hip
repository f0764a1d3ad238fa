// Bispectrum Monte Carlo Fisher contraction on the gfx950 matrix cores.
//
// The reference forms each Fisher row with integrate_row (PolyBin3D/cython/utils.pyx:162-179, called
// at PolyBin3D/bspec.py:597): for a left map m of triangle i with middle bin b2,
//   fish[i][j] = sum_{k in b2} degk Re[conj(m(k)) Q_j(k)]   over every Q map j.
// With the in-bin modes stored bin-major (one contiguous segment per bin) and the left maps of all
// rows of bin b2 gathered into a [rows][2 M_b2] real block (re, im interleaved, degk folded in), each
// segment is one real GEMM  C = L . Q_seg^T  with K = 2 M_b2.  This kernel runs all segments of a
// realization in one persistent launch over a list of 128 x 128 output tiles (largest K first), with
// split-BF16 operands (NP = 2: 3 products, NP = 3: 6 products, FP32-like) and two-level FP32
// accumulation (the MFMA accumulator over 512 reals, then a second accumulator), and adds
//   out[row_map[r]][col_map[n]] += alpha * scale[row_map[r]] * scale[col_map[n]] * C[r][n]
// in FP64, so that the column permutation of the Q maps, the normalisation and the accumulation
// over realizations are fused into the store.  Each output is written by one tile: no atomics, and
// the summation order is fixed (bit-reproducible).
//
// Tile: 256 threads (4 waves as 2 x 2), each wave 64 x 64 outputs (2 x 2 MFMA tiles of 32 x 32).
// Stage: 32 reals of K.  Each load instruction of a wave reads 4 rows x 128 bytes (16 lanes per row,
// float2 per lane: rows are only 8-byte aligned); every thread splits its reals into NP bf16 planes
// and stores them in LDS rows padded to 80 bytes (the 32 rows a k-step reads 16 bytes from then fall
// on distinct banks).  The next stage's global loads are issued before the current stage's MFMAs.
#include "pb3d_internal.h"
#include "pb3d_mfma.h"

#include <algorithm>
#include <vector>

namespace pb3d {

constexpr int kFm = 128;         // output rows per tile
constexpr int kFn = 128;         // output columns per tile
constexpr int kFk = 32;          // reals of K per stage
constexpr int kFthreads = 256;
constexpr int kFpitch = 40;      // bf16 per LDS row (32 + 8 pad)
constexpr int kFflush = 16;      // stages per first-level accumulation (512 reals)

// the NP bf16 parts of two consecutive reals, one 4-byte LDS store per plane (planes `plane` apart)
template <int NP>
__device__ __forceinline__ void store_split(__bf16* dst, int plane, float2 v) {
  float x[8] = {v.x, v.y, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  bf16x8 p[NP];
  split8<NP>(x, p);
#pragma unroll
  for (int q = 0; q < NP; ++q) {
    const uint32_t lohi = __builtin_bit_cast(uint32_t, __builtin_shufflevector(p[q], p[q], 0, 1));
    *(uint32_t*)(dst + q * plane) = lohi;
  }
}

template <int NP>
__global__ __launch_bounds__(kFthreads, 2) void fisher_gemm_kernel(
    const float* __restrict__ left, const float* __restrict__ q, long long q_ld, int nq,
    const pb3d_fisher_seg* __restrict__ segs, int nseg, const int32_t* __restrict__ tiles, int ntiles,
    const int32_t* __restrict__ row_map, const int32_t* __restrict__ col_map, const double* __restrict__ scale,
    double alpha, double* __restrict__ out, long long out_ld) {
  __shared__ __attribute__((aligned(16))) __bf16 sA[NP][kFm][kFpitch];
  __shared__ __attribute__((aligned(16))) __bf16 sB[NP][kFn][kFpitch];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wm = w >> 1, wn = w & 1;
  // loads: instruction j of wave w reads rows w*32 + 4j + lane/16, reals 2*(lane%16) and +1, so that
  // 16 consecutive lanes cover one row's 128 bytes of the stage
  const int lrow = w * 32 + (lane >> 4), lcol = 2 * (lane & 15);
  const int fr = lane & 31, fk = 8 * (lane >> 5);
  for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const int si = tiles[3 * t], m0 = tiles[3 * t + 1], n0 = tiles[3 * t + 2];
    if (si < 0 || si >= nseg) continue;
    const pb3d_fisher_seg sg = segs[si];
    const long long K = sg.k;
    f32x16 acc[2][2], acc2[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b)
#pragma unroll
        for (int v = 0; v < 16; ++v) acc[a][b][v] = acc2[a][b][v] = 0.f;
    float2 ra[8], rb[8];
    auto load = [&](long long k0) {
      const long long kk = k0 + lcol;
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int r = lrow + 4 * j;
        const bool va = m0 + r < sg.rows && kk < K, vb = n0 + r < nq && kk < K;
        ra[j] = va ? *(const float2*)(left + sg.a_off + (long long)(m0 + r) * K + kk) : make_float2(0.f, 0.f);
        rb[j] = vb ? *(const float2*)(q + (long long)(n0 + r) * q_ld + sg.q_off + kk) : make_float2(0.f, 0.f);
      }
    };
    const int nst = (int)((K + kFk - 1) / kFk);
    load(0);
    for (int st = 0; st < nst; ++st) {
      __syncthreads();  // the previous stage's LDS reads are done
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const int r = lrow + 4 * j;
        store_split<NP>(&sA[0][r][lcol], kFm * kFpitch, ra[j]);
        store_split<NP>(&sB[0][r][lcol], kFn * kFpitch, rb[j]);
      }
      __syncthreads();
      if (st + 1 < nst) load((long long)(st + 1) * kFk);
#pragma unroll
      for (int kk = 0; kk < kFk; kk += 16) {
        bf16x8 fa[2][NP], fb[2][NP];
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
          for (int p = 0; p < NP; ++p) fa[a][p] = *(const bf16x8*)&sA[p][wm * 64 + a * 32 + fr][kk + fk];
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
          for (int p = 0; p < NP; ++p) fb[b][p] = *(const bf16x8*)&sB[p][wn * 64 + b * 32 + fr][kk + fk];
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
          for (int b = 0; b < 2; ++b) acc[a][b] = mma_split<NP>(fa[a], fb[b], acc[a][b]);
      }
      if ((st + 1) % kFflush == 0 || st + 1 == nst) {
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
          for (int b = 0; b < 2; ++b) {
            acc2[a][b] += acc[a][b];
#pragma unroll
            for (int v = 0; v < 16; ++v) acc[a][b][v] = 0.f;
          }
      }
    }
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        const int col = n0 + wn * 64 + b * 32 + fr;
        if (col >= nq) continue;
        const int j = col_map[col];
        const double sj = scale ? alpha * scale[j] : alpha;
#pragma unroll
        for (int v = 0; v < 16; ++v) {
          const int row = m0 + wm * 64 + a * 32 + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
          if (row < sg.rows) {
            const int i = row_map[sg.row0 + row];
            const double s = scale ? sj * scale[i] : sj;
            out[(long long)i * out_ld + j] += s * (double)acc2[a][b][v];
          }
        }
      }
  }
}

// compute units of the current device, queried once per device (no runtime query inside a graph capture
// after the first eager call)
static int device_cus() {
  static int cached[64] = {};
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 256;
  if (dev >= 0 && dev < 64 && cached[dev] > 0) return cached[dev];
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1) cus = 256;
  if (dev >= 0 && dev < 64) cached[dev] = cus;
  return cus;
}

}  // namespace pb3d

using namespace pb3d;

extern "C" int pb3d_fisher_tiles(const pb3d_fisher_seg* segs, int nseg, int nq, int32_t* tiles, int max_tiles) {
  if (!segs || nseg < 1 || nq < 1) return fail(PB3D_ERR_ARG, "fisher_tiles: null segments or nseg/nq < 1");
  std::vector<int> order;
  long long count = 0;
  for (int s = 0; s < nseg; ++s) {
    if (segs[s].rows < 0 || segs[s].k < 0 || (segs[s].k & 1) || segs[s].a_off < 0 || segs[s].q_off < 0 ||
        segs[s].row0 < 0)
      return fail(PB3D_ERR_ARG, "fisher_tiles: negative size or offset, or odd k");
    if (segs[s].rows == 0 || segs[s].k == 0) continue;
    order.push_back(s);
    count += (long long)((segs[s].rows + kFm - 1) / kFm) * ((nq + kFn - 1) / kFn);
  }
  if (count > 0x7fffffff) return fail(PB3D_ERR_ARG, "fisher_tiles: too many tiles");
  if (!tiles) return (int)count;
  if (count > max_tiles) return fail(PB3D_ERR_ARG, "fisher_tiles: tile buffer too small");
  // largest K first (the persistent grid then ends on short tiles); within a segment the row tiles
  // of one column tile are consecutive, so the blocks that share a Q tile run together
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return segs[a].k > segs[b].k; });
  long long t = 0;
  for (int s : order) {
    const int mt = (segs[s].rows + kFm - 1) / kFm, nt = (nq + kFn - 1) / kFn;
    for (int n = 0; n < nt; ++n)
      for (int m = 0; m < mt; ++m) {
        tiles[3 * t] = s;
        tiles[3 * t + 1] = m * kFm;
        tiles[3 * t + 2] = n * kFn;
        ++t;
      }
  }
  return (int)count;
}

extern "C" int pb3d_fisher_contract(const float* left, const float* q, int64_t q_ld, int nq,
                                    const pb3d_fisher_seg* segs, int nseg, const int32_t* tiles, int ntiles,
                                    const int32_t* row_map, const int32_t* col_map, const double* scale,
                                    double alpha, int precision, double* out, int64_t out_ld, void* stream) {
  if (!left || !q || !segs || !tiles || !row_map || !col_map || !out)
    return fail(PB3D_ERR_ARG, "fisher_contract: null pointer");
  if (nq < 1 || nseg < 1 || ntiles < 1 || q_ld < 2 || (q_ld & 1) || out_ld < 1)
    return fail(PB3D_ERR_ARG, "fisher_contract: bad sizes");
  if (precision != 1 && precision != 2)
    return fail(PB3D_ERR_UNSUPPORTED, "fisher_contract: precision must be 1 (bf16x3) or 2 (bf16x6)");
  const int grid = (int)std::min<long long>(ntiles, 2LL * device_cus());
  hipStream_t st = (hipStream_t)stream;
  if (precision == 2)
    fisher_gemm_kernel<3><<<grid, kFthreads, 0, st>>>(left, q, q_ld, nq, segs, nseg, tiles, ntiles, row_map, col_map,
                                                      scale, alpha, out, out_ld);
  else
    fisher_gemm_kernel<2><<<grid, kFthreads, 0, st>>>(left, q, q_ld, nq, segs, nseg, tiles, ntiles, row_map, col_map,
                                                      scale, alpha, out, out_ld);
  PB3D_CHECK_LAUNCH();
  return 0;
}
