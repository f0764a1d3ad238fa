// Triplet contractions as split-BF16 GEMMs on the gfx950 matrix cores (v_mfma_f32_32x32x16_bf16,
// FP32 accumulation).
//
// Forward (assemble_b3 / sum_triplet, PolyBin3D/cython/utils.pyx:357-366, 614-636):
//   T[a][p] = sum_x g[amap[a]](x) * g[pb[p]](x) * g[pc[p]](x)
// i.e. the GEMM T = G_A . P^T with K = voxels and the pair products P_p = g_b g_c formed on the fly;
// every triplet sum is one entry of T.  Adjoint (no counterpart in the reference):
//   out[a][x] = sum_p U[a][p] * g[pb[p]](x) * g[pc[p]](x)
// the GEMM U . P with voxels as the N axis.
//
// Split BF16: an FP32 operand x = hi + mid (+ lo), each part bf16 (x - hi - mid exact in FP32).
// precision 1 = 3-term product (hi*hi + hi*mid + mid*hi: ~2^-16 relative per product; parts rounded
// to nearest), precision 2 = 6-term product (adds hi*lo + lo*hi + mid*mid: ~2^-24, FP32-like; parts
// truncated, which makes hi + mid + lo exactly x).
//
// MFMA 32x32x16 lane layout (checked on the device): A[i][k] from lane l: i = l%32,
// k = 8*(l/32) + r (r = 0..7 of an 8 x bf16 operand); B[k][j]: j = l%32, same k; accumulator
// register v of lane l holds C[8*(v/4) + 4*(l/32) + v%4][l%32].
#include "pb3d_ops.h"
#include "pb3d_mfma.h"

namespace pb3d {

constexpr int kPgKS = 16;         // voxels (forward) / pairs (adjoint) per MFMA k-step
constexpr int kPgWaves = 4;       // waves per block
constexpr int kPgThreads = 64 * kPgWaves;
constexpr int kPgMaxM = 128;      // A rows (maps) per call: up to 4 M-tiles of 32
constexpr int kPgMaxMaps = 256;   // maps in the stack the forward kernel stages

// ------------------------------------------------------------------------------------------
// Forward.  Block (blockIdx.x = group of nw N-tiles of 32 pairs, one per wave, blockIdx.y = voxel
// chunk; nw = kPgWaves, or 4..8 waves for a single M-tile so that up to 256 pairs share one staging):
// stages kPgStage = ST (32, 64 or 128) voxels of every map of the stack in LDS (for stacks up to 96
// maps the next stage's loads go into registers before the current stage is used; larger stacks load
// at the start of each stage), splits the A rows once into bf16 parts shared by
// the waves, then each wave forms the products of its 32 pairs and runs NP-part MFMAs for every
// M-tile, kPgStage/kPgKS k-steps per stage.  Partial T of the chunk -> part[chunk][a][p] (FP64),
// accumulated in two FP32 levels (MFMA accumulator over 1024 voxels, then a second accumulator).
// LDS (dynamic): sg[nmaps + 2][kPgStage + 4] floats, then sa[NP][MT*32][kPgStage/8 + SAP] bf16x8.  The
// row pads put consecutive rows 16 bytes apart in the bank space: the lanes of a k-step read 8
// consecutive floats of 32 different rows (the pair products, consecutive pairs share b and take
// consecutive c) and 16 bytes of 32 consecutive A rows, which an unpadded power-of-two row length
// maps onto the same banks.
// ------------------------------------------------------------------------------------------
constexpr int kPgFlushVox = 1024;                     // voxels per first-level accumulation
constexpr size_t kPgLdsBudget = 64 * 1024 - 512;      // dynamic LDS per block (+ <= 384 B static: two blocks per CU)
constexpr int kPgSgPad = 4;                           // floats of pad per staged map row

static size_t pg_fwd_lds(int nmaps, int np_parts, int mt, int stage, int sap) {
  return (size_t)(nmaps + 2) * (stage + kPgSgPad) * sizeof(float) + (size_t)np_parts * mt * 32 * (stage / 8 + sap) * 16;
}

// ST = voxels per LDS stage (a stage costs three block barriers, its global loads and the A split,
// so small stacks take long stages); PF = prefetch registers per thread: nmaps * ST <= PF * kPgThreads
// (and blocks have at least kPgThreads threads);
// SAP = pad (bf16x8) per split A row (0 or 1)
template <int NP, int MT, int ST, int PF, int SAP>
__global__ void __launch_bounds__(MT == 1 && PF <= 24 ? 2 * kPgThreads : kPgThreads)  // kWide below
pair_gemm_fwd(const float* __restrict__ g, int nmaps, long long nvox, const int* __restrict__ amap, int na,
              const int* __restrict__ pb, const int* __restrict__ pc, int np, long long chunk,
              double* __restrict__ part) {
  constexpr int kPgStage = ST;
  constexpr int kPgFlush = kPgFlushVox / ST;  // stages per flush of the first-level accumulator
  extern __shared__ __align__(16) char smem[];
  constexpr int SGS = kPgStage + kPgSgPad;                                             // sg row stride
  float* sg = reinterpret_cast<float*>(smem);                                         // [nmaps + 2][SGS]
  bf16x8* sa = reinterpret_cast<bf16x8*>(smem + (size_t)(nmaps + 2) * SGS * sizeof(float));
  constexpr int K8 = kPgStage / 8;                                                     // bf16x8 per row
  constexpr int K8S = K8 + SAP;                                                        // sa row stride
  // kPgWaves waves per block (a compile-time stride), or 4..8 for a single M-tile over a stack of
  // at most 96 maps (one pair group holds every pair)
  constexpr bool kWide = MT == 1 && PF <= 24;
  const int tid = threadIdx.x, lane = tid % 64, wave = tid / 64;
  const int nthr = kWide ? (int)blockDim.x : kPgThreads, nw = nthr / 64;
  const long long x0 = (long long)blockIdx.y * chunk;
  long long x1 = x0 + chunk;
  if (x1 > nvox) x1 = nvox;
  // this lane's pair (column j = lane%32 of the wave's N-tile)
  const int p = (blockIdx.x * nw + wave) * 32 + lane % 32;
  const bool pvalid = p < np;
  const bool wave_on = (blockIdx.x * nw + wave) * 32 < np;  // wave-uniform
  int mb = pvalid ? pb[p] : 0, mc = pvalid ? pc[p] : 0;
  // two constant rows follow the staged maps in sg: row nmaps is zero (the products of a pair
  // beyond np), row nmaps + 1 is NaN (an index outside the stack gives NaN for that pair, with no
  // stray LDS read), so the products need no selects
  if (mb < 0 || mb >= nmaps || mc < 0 || mc >= nmaps) { mb = nmaps + 1; mc = nmaps + 1; }
  if (!pvalid) { mb = nmaps; mc = nmaps; }
  for (int k = tid; k < 2 * SGS; k += nthr) sg[nmaps * SGS + k] = k < SGS ? 0.f : __builtin_nanf("");
  const int kh = 8 * (lane / 32);  // this lane's k offset inside a k-step
  // two-level FP32 accumulation: the MFMA accumulator acc covers at most kPgFlush stages (1024
  // voxels), then is added into acc2 (rounding grows like sqrt(steps per level), not sqrt(steps))
  f32x16 acc[MT], acc2[MT];
#pragma unroll
  for (int t = 0; t < MT; ++t)
#pragma unroll
    for (int v = 0; v < 16; ++v) { acc[t][v] = 0.f; acc2[t][v] = 0.f; }
  // sg offset of each A row's map, read from LDS in every stage's split (-1: row beyond na, zero;
  // -2: map index outside the stack, NaN); visible after the first barrier of the stage loop
  __shared__ int s_aoff[MT * 32];
  for (int r = tid; r < MT * 32; r += nthr) {
    int o = -1;
    if (r < na) {
      const int am = amap[r];
      o = am < 0 || am >= nmaps ? -2 : am * SGS;
    }
    s_aoff[r] = o;
  }
  // staged element e = m * kPgStage + k, k = e % kPgStage: thread tid loads e = tid + i * nthr
  const int nel = nmaps * kPgStage;
  float pf[PF];
  auto fetch = [&](long long xs) {
#pragma unroll
    for (int i = 0; i < PF; ++i) {
      const int e = tid + i * nthr;
      float v = 0.f;
      if (e < nel) {
        const long long x = xs + (e % kPgStage);
        if (x < x1) v = g[(long long)(e / kPgStage) * nvox + x];
      }
      pf[i] = v;
    }
  };
  // PF <= 24: the next stage's loads are in flight in registers while this one computes; the largest
  // prefetch (32 registers) is loaded at the start of its own stage instead, so that those registers
  // are free during the MFMA loop (another block on the CU covers the load latency)
  constexpr bool kPre = PF <= 24;
  if (kPre && x0 < x1) fetch(x0);
  int nstage = 0;
  for (long long xs = x0; xs < x1; xs += kPgStage) {
    __syncthreads();  // previous stage's readers of sg / sa are done
    if constexpr (!kPre) fetch(xs);
#pragma unroll
    for (int i = 0; i < PF; ++i) {
      const int e = tid + i * nthr;
      if (e < nel) sg[(e / kPgStage) * SGS + e % kPgStage] = pf[i];
    }
    if (kPre && xs + kPgStage < x1) fetch(xs + kPgStage);  // next stage in flight while this one computes
    __syncthreads();
    // split the A rows (rows beyond na are zero)
    for (int e = tid; e < MT * 32 * K8; e += nthr) {
      const int row = e / K8, kb = e % K8;
      const int o = s_aoff[row];
      float v[8];
      if (o >= 0) {  // two 16-byte LDS reads (offsets are multiples of 4 floats)
        const float4 u0 = *reinterpret_cast<const float4*>(sg + o + kb * 8);
        const float4 u1 = *reinterpret_cast<const float4*>(sg + o + kb * 8 + 4);
        v[0] = u0.x; v[1] = u0.y; v[2] = u0.z; v[3] = u0.w; v[4] = u1.x; v[5] = u1.y; v[6] = u1.z; v[7] = u1.w;
      } else {
#pragma unroll
        for (int r = 0; r < 8; ++r) v[r] = o == -1 ? 0.f : __builtin_nanf("");
      }
      bf16x8 parts[NP];
      split8<NP>(v, parts);
#pragma unroll
      for (int q = 0; q < NP; ++q) sa[(q * MT * 32 + row) * K8S + kb] = parts[q];
    }
    __syncthreads();  // sa ready
    // a wave whose 32 pairs all lie beyond np only helps stage (uniform skip of its MFMAs)
#pragma unroll 1  // (unrolled k-steps hold more registers: an occupancy step for the 3-M-tile kernels)
    for (int ks = 0; ks < (wave_on ? kPgStage / kPgKS : 0); ++ks) {
      // B operand: this lane's 8 pair products of k-step ks
      float pr[8];
      const int k0 = ks * kPgKS + kh;
      {  // two 16-byte LDS reads per operand (rows and k0 are multiples of 4 floats)
        const float4* rb = reinterpret_cast<const float4*>(sg + mb * SGS + k0);
        const float4* rc = reinterpret_cast<const float4*>(sg + mc * SGS + k0);
        const float4 b0 = rb[0], b1 = rb[1], c0 = rc[0], c1 = rc[1];
        const float bv[8] = {b0.x, b0.y, b0.z, b0.w, b1.x, b1.y, b1.z, b1.w};
        const float cv[8] = {c0.x, c0.y, c0.z, c0.w, c1.x, c1.y, c1.z, c1.w};
#pragma unroll
        for (int r = 0; r < 8; ++r) pr[r] = bv[r] * cv[r];
      }
      bf16x8 bparts[NP];
      split8<NP>(pr, bparts);
#pragma unroll
      for (int t = 0; t < MT; ++t) {
        bf16x8 aparts[NP];
#pragma unroll
        for (int q = 0; q < NP; ++q) aparts[q] = sa[(q * MT * 32 + t * 32 + lane % 32) * K8S + k0 / 8];
        acc[t] = mma_split<NP>(aparts, bparts, acc[t]);
      }
    }
    if (++nstage == kPgFlush) {  // bound the first-level accumulation length (uniform)
      nstage = 0;
#pragma unroll
      for (int t = 0; t < MT; ++t)
#pragma unroll
        for (int v = 0; v < 16; ++v) { acc2[t][v] += acc[t][v]; acc[t][v] = 0.f; }
    }
  }
  // partial results: C[row][col] with col = lane%32 (pair), row = 8*(v/4) + 4*(lane/32) + v%4
  double* out = part + (long long)blockIdx.y * na * np;
  if (pvalid) {
#pragma unroll
    for (int t = 0; t < MT; ++t)
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const int row = t * 32 + 8 * (v / 4) + 4 * (lane / 32) + v % 4;
        if (row < na) out[(long long)row * np + p] = (double)acc2[t][v] + (double)acc[t][v];
      }
  }
}

// out[i] = sum_c part[c][i].  A block takes 16 consecutive outputs with 16 chunk lanes each (lane q
// sums chunks q, q+16, ...), then adds the 16 lane sums in a fixed order: deterministic, and enough
// blocks and independent loads in flight for a few outputs over a thousand chunks.
constexpr int kRedOut = 16, kRedLanes = 16;
__global__ void __launch_bounds__(kRedOut * kRedLanes)
pair_gemm_reduce(const double* __restrict__ part, int nchunks, long long n, double* __restrict__ out) {
  __shared__ double red[kRedLanes][kRedOut + 1];
  const int j = threadIdx.x % kRedOut, q = threadIdx.x / kRedOut;
  const long long i = (long long)blockIdx.x * kRedOut + j;
  double s = 0.;
  if (i < n) {
#pragma unroll 4
    for (int c = q; c < nchunks; c += kRedLanes) s += part[(long long)c * n + i];
  }
  red[q][j] = s;
  __syncthreads();
  if (q == 0 && i < n) {
    double t = 0.;
#pragma unroll
    for (int l = 0; l < kRedLanes; ++l) t += red[l][j];
    out[i] = t;
  }
}

// ------------------------------------------------------------------------------------------
// Adjoint.  Block = kPgWaves N-tiles of 32 voxels; the maps of the block's voxels are staged in
// LDS once; the K loop runs over pairs (kPgKS per step) with the A operand = split U rows read
// from global memory (small, cache resident) and the B operand = pair products of this lane's
// voxel.  Output rows a < na of out[a][x] (FP32).
// ------------------------------------------------------------------------------------------
constexpr int kAdjVox = 32 * kPgWaves;  // voxels per block
constexpr int kAdjMaxMaps = 128;

template <int NP, int MT>
__global__ void __launch_bounds__(kPgThreads)
pair_adjoint_kernel(const float* __restrict__ g, int nmaps, long long nvox, const int* __restrict__ pb,
                    const int* __restrict__ pc, int np_pad, const bf16x8* __restrict__ usplit, int na,
                    float* __restrict__ out) {
  __shared__ float sg[kAdjMaxMaps][kAdjVox];
  const int tid = threadIdx.x, lane = tid % 64, wave = tid / 64;
  const long long xb = (long long)blockIdx.x * kAdjVox;
  for (int e = tid; e < nmaps * kAdjVox; e += kPgThreads) {
    const int m = e / kAdjVox, k = e % kAdjVox;
    const long long x = xb + k;
    sg[m][k] = x < nvox ? g[(long long)m * nvox + x] : 0.f;
  }
  __syncthreads();
  const int col = wave * 32 + lane % 32;  // this lane's voxel inside the block
  const int kh = 8 * (lane / 32);
  f32x16 acc[MT];
#pragma unroll
  for (int t = 0; t < MT; ++t)
#pragma unroll
    for (int v = 0; v < 16; ++v) acc[t][v] = 0.f;
  const int ksteps = np_pad / kPgKS;
  for (int s = 0; s < ksteps; ++s) {
    const int p0 = s * kPgKS + kh;
    float pr[8];
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      const int b = pb[p0 + r], c = pc[p0 + r];
      const bool ok = b >= 0 && b < nmaps && c >= 0 && c < nmaps;  // else NaN, no stray LDS read
      pr[r] = ok ? sg[b][col] * sg[c][col] : __builtin_nanf("");
    }
    bf16x8 bparts[NP];
    split8<NP>(pr, bparts);
#pragma unroll
    for (int t = 0; t < MT; ++t) {
      bf16x8 aparts[NP];
      // usplit layout: [part][row (MT*32 padded)][np_pad / 8] bf16x8
#pragma unroll
      for (int q = 0; q < NP; ++q)
        aparts[q] = usplit[((long long)q * (MT * 32) + t * 32 + lane % 32) * (np_pad / 8) + (p0 / 8)];
      acc[t] = mma_split<NP>(aparts, bparts, acc[t]);
    }
  }
  const long long x = xb + col;
  if (x < nvox) {
#pragma unroll
    for (int t = 0; t < MT; ++t)
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const int row = t * 32 + 8 * (v / 4) + 4 * (lane / 32) + v % 4;
        if (row < na) out[(long long)row * nvox + x] = acc[t][v];
      }
  }
}

// U [na][np] FP32 -> split parts [NP][rows][np_pad/8] bf16x8 (rows = MT*32, zero padded)
template <int NP>
__global__ void split_rows_kernel(const float* __restrict__ U, int na, int np, int rows, int np_pad,
                                  bf16x8* __restrict__ out) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;  // (row, k8)
  const long long n = (long long)rows * (np_pad / 8);
  if (i >= n) return;
  const int row = (int)(i / (np_pad / 8)), k8 = (int)(i % (np_pad / 8));
  float v[8];
#pragma unroll
  for (int r = 0; r < 8; ++r) {
    const int p = k8 * 8 + r;
    v[r] = (row < na && p < np) ? U[(long long)row * np + p] : 0.f;
  }
  bf16x8 parts[NP];
  split8<NP>(v, parts);
#pragma unroll
  for (int q = 0; q < NP; ++q) out[(long long)q * n + i] = parts[q];
}

// pair list padded to a multiple of kPgKS with pair (0, 0) and zero U columns
__global__ void pad_pairs_kernel(const int* __restrict__ pb, const int* __restrict__ pc, int np, int np_pad,
                                 int* __restrict__ ob, int* __restrict__ oc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np_pad) return;
  ob[i] = i < np ? pb[i] : 0;
  oc[i] = i < np ? pc[i] : 0;
}

static int pg_mtiles(int na) { return (na + 31) / 32; }
static long long pg_chunks(long long nvox, int ngroups) {
  // enough blocks to fill the device several times over, chunks of at least 4096 voxels
  long long c = (2048 + ngroups - 1) / ngroups;
  const long long maxc = (nvox + 4095) / 4096;
  if (c > maxc) c = maxc;
  if (c < 1) c = 1;
  return c;
}
static int pad16(int n) { return (n + kPgKS - 1) / kPgKS * kPgKS; }
// waves per forward block: up to 256 pairs of a single M-tile (<= 32 A rows) over a stack of at
// most 96 maps in one pair group of 4..8 waves (one staging of the stack for all of them; those
// kernels are built for 512 threads), else kPgWaves
static int pg_waves(int nmaps, int na, int np) {
  if (nmaps > 96 || na > 32 || np > 8 * 32) return kPgWaves;
  const int w = (np + 31) / 32;
  return w < kPgWaves ? kPgWaves : w;
}

}  // namespace pb3d

using namespace pb3d;

extern "C" size_t pb3d_pair_gemm_work_size(int64_t nvox, int na, int np) {
  if (nvox < 1 || na < 1 || np < 1) return 0;
  // the widest block any stack allows: fewest pair groups, most chunks (an upper bound for every nmaps)
  const int nw = pg_waves(1, na, np), ngroups = (np + 32 * nw - 1) / (32 * nw);
  return (size_t)pg_chunks(nvox, ngroups) * na * np * sizeof(double);
}

extern "C" int pb3d_pair_gemm_cfg(const float* g, int nmaps, int64_t nvox, const int32_t* amap, int na, const int32_t* pb,
                                  const int32_t* pc, int np, int precision, int stage_cfg, int rows_cfg, double* out,
                                  void* work, void* stream) {
  if (!g || !amap || !pb || !pc || !out || !work) return fail(PB3D_ERR_ARG, "pair_gemm: null pointer");
  if (nmaps < 1 || nmaps > kPgMaxMaps) return fail(PB3D_ERR_ARG, "pair_gemm: nmaps out of [1, 256]");
  if (nvox < 1 || na < 1 || na > kPgMaxM || np < 1) return fail(PB3D_ERR_ARG, "pair_gemm: na out of [1, 128] or empty input");
  if (precision != 1 && precision != 2) return fail(PB3D_ERR_ARG, "pair_gemm: precision must be 1 (bf16 x3) or 2 (bf16 x6)");
  hipStream_t s = (hipStream_t)stream;
  const int nw = pg_waves(nmaps, na, np), ngroups = (np + 32 * nw - 1) / (32 * nw);  // nc <= the work size's
  const long long nc = pg_chunks(nvox, ngroups);
  // stack-size bucket (maps <= 32, 64, 96, 128, 256) and stage length: the long stage of a bucket
  // (128 voxels for <= 32 maps, 64 for <= 128) when its LDS fits kPgLdsBudget, else 32 voxels;
  // prefetch registers = bucket maps * stage / threads.  The largest stacks keep the A rows unpadded.
  const int pfb = nmaps <= 32 ? 0 : nmaps <= 64 ? 1 : nmaps <= 96 ? 2 : nmaps <= 128 ? 3 : 4;
  const int nparts = precision == 2 ? 3 : 2, sap = pfb == 4 ? 0 : 1;
  // More than 96 A rows run as launches of 64 rows (2 M-tiles each): a 4-M-tile wave holds 128
  // accumulator registers and runs at one wave per SIMD, and two 2-M-tile launches, which form the
  // pair products twice, took 20 % less time (128 rows x 2080 pairs at 256^3).
  const int long_stage = pfb == 0 ? 128 : pfb <= 3 ? 64 : 32;
  if (stage_cfg != 0 && stage_cfg != 32 && stage_cfg != long_stage)
    return fail(PB3D_ERR_ARG, "pair_gemm: stage must be 0 (auto), 32 or " + std::to_string(long_stage) +
                                  " for a stack of this size");
  if (rows_cfg != 0 && rows_cfg != 32 && rows_cfg != 64 && rows_cfg != 96)
    return fail(PB3D_ERR_ARG, "pair_gemm: rows per launch must be 0 (auto), 32, 64 or 96");
  const int rows_per_launch = rows_cfg ? rows_cfg : (na > 3 * 32 ? 2 * 32 : na);
  for (int r0 = 0; r0 < na; r0 += rows_per_launch) {
    const int nar = na - r0 < rows_per_launch ? na - r0 : rows_per_launch;
    const int mt = pg_mtiles(nar);  // 1..3
    int stage = stage_cfg ? stage_cfg : long_stage;
    if (!stage_cfg && pg_fwd_lds(nmaps, nparts, mt, stage, sap) > kPgLdsBudget) stage = 32;
    const size_t lds = pg_fwd_lds(nmaps, nparts, mt, stage, sap);
    if (lds > kPgLdsBudget) return fail(PB3D_ERR_ARG, "pair_gemm: stack does not fit in LDS");
    long long chunk = (nvox + nc - 1) / nc;
    chunk = (chunk + stage - 1) / stage * stage;
    dim3 grid((unsigned)ngroups, (unsigned)nc);
    const int32_t* am = amap + r0;
    double* part = (double*)work + nc * r0 * np;  // [nc][nar][np] region of this row block
#define PB3D_PG3(NP, MT, ST, PF, SAP) \
  pair_gemm_fwd<NP, MT, ST, PF, SAP><<<grid, 64 * nw, lds, s>>>(g, nmaps, nvox, am, nar, pb, pc, np, chunk, part)
#define PB3D_PG(NP, MT)                                                       \
    switch (pfb * 2 + (stage == 32)) {                                        \
      case 0: PB3D_PG3(NP, MT, 128, 32 * 128 / kPgThreads, 1); break;         \
      case 1: PB3D_PG3(NP, MT, 32, 32 * 32 / kPgThreads, 1); break;           \
      case 2: PB3D_PG3(NP, MT, 64, 64 * 64 / kPgThreads, 1); break;           \
      case 3: PB3D_PG3(NP, MT, 32, 64 * 32 / kPgThreads, 1); break;           \
      case 4: PB3D_PG3(NP, MT, 64, 96 * 64 / kPgThreads, 1); break;           \
      case 5: PB3D_PG3(NP, MT, 32, 96 * 32 / kPgThreads, 1); break;           \
      case 6: PB3D_PG3(NP, MT, 64, 128 * 64 / kPgThreads, 1); break;          \
      case 7: PB3D_PG3(NP, MT, 32, 128 * 32 / kPgThreads, 1); break;          \
      default: PB3D_PG3(NP, MT, 32, 256 * 32 / kPgThreads, 0); break;         \
    }
    if (precision == 1) {
      switch (mt) { case 1: PB3D_PG(2, 1); break; case 2: PB3D_PG(2, 2); break; default: PB3D_PG(2, 3); }
    } else {
      switch (mt) { case 1: PB3D_PG(3, 1); break; case 2: PB3D_PG(3, 2); break; default: PB3D_PG(3, 3); }
    }
#undef PB3D_PG
#undef PB3D_PG3
    PB3D_CHECK_LAUNCH();
    const long long n = (long long)nar * np;
    pair_gemm_reduce<<<(unsigned)((n + kRedOut - 1) / kRedOut), kRedOut * kRedLanes, 0, s>>>(part, (int)nc, n,
                                                                                            out + (long long)r0 * np);
    PB3D_CHECK_LAUNCH();
  }
  return 0;
}

extern "C" int pb3d_pair_gemm(const float* g, int nmaps, int64_t nvox, const int32_t* amap, int na, const int32_t* pb,
                              const int32_t* pc, int np, int precision, double* out, void* work, void* stream) {
  return pb3d_pair_gemm_cfg(g, nmaps, nvox, amap, na, pb, pc, np, precision, 0, 0, out, work, stream);
}

extern "C" size_t pb3d_pair_adjoint_work_size(int na, int np, int precision) {
  if (na < 1 || np < 1) return 0;
  const int nparts = precision == 2 ? 3 : 2;
  const size_t rows = (size_t)pg_mtiles(na) * 32;
  const size_t npp = (size_t)pad16(np);
  return nparts * rows * npp * 2 + 2 * npp * sizeof(int32_t) + 256;
}

extern "C" int pb3d_pair_adjoint(const float* g, int nmaps, int64_t nvox, const int32_t* pb, const int32_t* pc, int np,
                                 const float* U, int na, int precision, float* out, void* work, void* stream) {
  if (!g || !pb || !pc || !U || !out || !work) return fail(PB3D_ERR_ARG, "pair_adjoint: null pointer");
  if (nmaps < 1 || nmaps > kAdjMaxMaps) return fail(PB3D_ERR_ARG, "pair_adjoint: nmaps out of [1, 128]");
  if (nvox < 1 || na < 1 || na > kPgMaxM || np < 1) return fail(PB3D_ERR_ARG, "pair_adjoint: na out of [1, 128] or empty input");
  if (precision != 1 && precision != 2) return fail(PB3D_ERR_ARG, "pair_adjoint: precision must be 1 (bf16 x3) or 2 (bf16 x6)");
  hipStream_t s = (hipStream_t)stream;
  const int mt = pg_mtiles(na), rows = mt * 32, npp = pad16(np);
  const int nparts = precision == 2 ? 3 : 2;
  bf16x8* us = (bf16x8*)work;
  const size_t ubytes = (size_t)nparts * rows * npp * 2;
  int* ob = (int*)((char*)work + ((ubytes + 255) / 256) * 256);
  int* oc = ob + npp;
  const long long nsplit = (long long)rows * (npp / 8);
  if (precision == 1) split_rows_kernel<2><<<(unsigned)((nsplit + 255) / 256), 256, 0, s>>>(U, na, np, rows, npp, us);
  else split_rows_kernel<3><<<(unsigned)((nsplit + 255) / 256), 256, 0, s>>>(U, na, np, rows, npp, us);
  PB3D_CHECK_LAUNCH();
  pad_pairs_kernel<<<(npp + 255) / 256, 256, 0, s>>>(pb, pc, np, npp, ob, oc);
  PB3D_CHECK_LAUNCH();
  const unsigned blocks = (unsigned)((nvox + kAdjVox - 1) / kAdjVox);
#define PB3D_PA(NP, MT) pair_adjoint_kernel<NP, MT><<<blocks, kPgThreads, 0, s>>>(g, nmaps, nvox, ob, oc, npp, us, na, out)
  if (precision == 1) {
    switch (mt) { case 1: PB3D_PA(2, 1); break; case 2: PB3D_PA(2, 2); break; case 3: PB3D_PA(2, 3); break; default: PB3D_PA(2, 4); }
  } else {
    switch (mt) { case 1: PB3D_PA(3, 1); break; case 2: PB3D_PA(3, 2); break; case 3: PB3D_PA(3, 3); break; default: PB3D_PA(3, 4); }
  }
#undef PB3D_PA
  PB3D_CHECK_LAUNCH();
  return 0;
}
