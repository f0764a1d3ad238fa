// 3D real FFTs for gfx950: Stockham radix-16/8/4/2 butterflies in registers, LDS exchange
// between stages, one pass per axis.  R2C = z-rows (real->half complex) -> y-columns ->
// x-columns; C2R = x-columns -> y-columns -> z-rows.  The voxel/mode operations of pb3d_ops.h are
// fused into the first load and last store (prologue/epilogue), so filters, masks, Y_lm weights
// and products never take a separate pass over HBM.
//
// Replaces the reference's FFT dispatch to_fourier / to_real / _fft_setup
// (PolyBin3D/base.py:176-229, 256-302); numpy conventions (forward unnormalised, inverse 1/N).
#include "pb3d_ops.h"
#include "pb3d_fftgen.h"
#include <vector>
#include <cmath>
#include <atomic>
#include <type_traits>

namespace pb3d {

// ------------------------------------------------------------------------------------------
// Register DFTs (decimation in frequency, natural-order output).  DIR=-1 forward, +1 inverse.
// ------------------------------------------------------------------------------------------
// w16(j) = exp(DIR * 2*pi*i*j/16) for compile-time j after unrolling.
template <int DIR>
__device__ __forceinline__ float2 w16(int j) {
  constexpr float c1 = 0.92387953251128674f, s1 = 0.38268343236508978f, h = 0.70710678118654752f;
  float c, s;
  switch (j & 15) {
    case 0: c = 1.f; s = 0.f; break;
    case 1: c = c1; s = s1; break;
    case 2: c = h; s = h; break;
    case 3: c = s1; s = c1; break;
    case 4: c = 0.f; s = 1.f; break;
    case 5: c = -s1; s = c1; break;
    case 6: c = -h; s = h; break;
    case 7: c = -c1; s = s1; break;
    default: c = -1.f; s = 0.f; break;
  }
  return make_float2(c, DIR * s);
}

template <int R, int DIR>
struct DFT {
  __device__ __forceinline__ static void run(float2* v) {
    constexpr int H = R / 2;
    float2 b[H], c[H];
#pragma unroll
    for (int k = 0; k < H; ++k) {
      b[k] = cadd(v[k], v[k + H]);
      float2 d = csub(v[k], v[k + H]);
      const int j = k * (16 / R);  // exp(DIR 2 pi i k / R)
      if (j == 0) {
        c[k] = d;
      } else if (j == 4) {  // DIR * i
        c[k] = DIR < 0 ? make_float2(d.y, -d.x) : make_float2(-d.y, d.x);
      } else {
        c[k] = cmul(d, w16<DIR>(j));
      }
    }
    DFT<H, DIR>::run(b);
    DFT<H, DIR>::run(c);
#pragma unroll
    for (int m = 0; m < H; ++m) {
      v[2 * m] = b[m];
      v[2 * m + 1] = c[m];
    }
  }
};
template <int DIR>
struct DFT<1, DIR> {
  __device__ __forceinline__ static void run(float2*) {}
};

// ------------------------------------------------------------------------------------------
// Compile-time decomposition: radices R0,R1,R2 (1 = unused) and threads per pencil P.
// ------------------------------------------------------------------------------------------
template <int L> struct Decomp;
template <> struct Decomp<2>    { static constexpr int R0 = 2,  R1 = 1,  R2 = 1, P = 1; };
template <> struct Decomp<4>    { static constexpr int R0 = 4,  R1 = 1,  R2 = 1, P = 1; };
template <> struct Decomp<8>    { static constexpr int R0 = 8,  R1 = 1,  R2 = 1, P = 1; };
template <> struct Decomp<16>   { static constexpr int R0 = 16, R1 = 1,  R2 = 1, P = 1; };
template <> struct Decomp<32>   { static constexpr int R0 = 8,  R1 = 4,  R2 = 1, P = 4; };
template <> struct Decomp<64>   { static constexpr int R0 = 8,  R1 = 8,  R2 = 1, P = 8; };
template <> struct Decomp<128>  { static constexpr int R0 = 16, R1 = 8,  R2 = 1, P = 8; };
template <> struct Decomp<256>  { static constexpr int R0 = 16, R1 = 16, R2 = 1, P = 16; };
template <> struct Decomp<512>  { static constexpr int R0 = 8,  R1 = 8,  R2 = 8, P = 64; };

// LDS ordering between the lanes of one wave: a wave's LDS instructions execute in issue order,
// so a compiler-level barrier is enough (no s_barrier).  Block variant: __syncthreads.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
template <bool WAVE>
__device__ __forceinline__ void lds_sync() {
  if constexpr (WAVE) wave_lds_sync(); else __syncthreads();
}

// One Stockham stage: butterflies j = t + b*P, inputs src(j + k*L/R), twiddles from the length-L
// table tw[m] = exp(-2 pi i m / L) (conjugated for DIR=+1), outputs dst((j/NS)*NS*R + j%NS + k*NS).
// If `barrier` the block (WAVE: the wave) synchronises between the reads and the writes (src/dst
// share LDS).
template <int L, int R, int NS, int P, int DIR, bool WAVE = false, typename Src, typename Dst>
__device__ __forceinline__ void stage(int t, Src src, Dst dst, const float2* __restrict__ tw,
                                      bool barrier) {
  constexpr int NB = L / R;
  constexpr int B = NB / P;
  static_assert(B >= 1 && NB % P == 0, "bad decomposition");
  float2 v[B][R];
#pragma unroll
  for (int b = 0; b < B; ++b) {
    const int j = t + b * P;
#pragma unroll
    for (int k = 0; k < R; ++k) v[b][k] = src(j + k * NB);
    if (NS > 1) {
      const int jm = j % NS;
#pragma unroll
      for (int k = 1; k < R; ++k) {
        float2 w = tw[jm * k * (L / (NS * R))];
        if (DIR > 0) w.y = -w.y;
        v[b][k] = cmul(v[b][k], w);
      }
    }
    DFT<R, DIR>::run(v[b]);
  }
  if (barrier) lds_sync<WAVE>();
#pragma unroll
  for (int b = 0; b < B; ++b) {
    const int j = t + b * P;
    const int base = (j / NS) * NS * R + (j % NS);
#pragma unroll
    for (int k = 0; k < R; ++k) dst(base + k * NS, v[b][k]);
  }
}

// Full pencil FFT: first stage reads `src0`, last stage writes `dstN`, intermediate data lives in
// the LDS accessor `lds`.  All threads of the block must call this (it contains barriers).
template <int L, int DIR, typename Src0, typename Lds, typename DstN>
__device__ __forceinline__ void pencil_fft(int t, Src0 src0, Lds lds, DstN dstN,
                                           const float2* __restrict__ tw, bool src_is_lds) {
  using D = Decomp<L>;
  constexpr int P = D::P;
  if constexpr (D::R1 == 1) {
    stage<L, D::R0, 1, P, DIR>(t, src0, dstN, tw, src_is_lds);
  } else if constexpr (D::R2 == 1) {
    stage<L, D::R0, 1, P, DIR>(t, src0, lds, tw, src_is_lds);
    __syncthreads();
    stage<L, D::R1, D::R0, P, DIR>(t, lds, dstN, tw, true);
  } else {
    stage<L, D::R0, 1, P, DIR>(t, src0, lds, tw, src_is_lds);
    __syncthreads();
    stage<L, D::R1, D::R0, P, DIR>(t, lds, lds, tw, true);
    __syncthreads();
    stage<L, D::R2, D::R0 * D::R1, P, DIR>(t, lds, dstN, tw, true);
  }
}

#ifndef PB3D_COL_THREADS
#define PB3D_COL_THREADS 1024
#endif
// Streaming stores of whole-map pass outputs (each far larger than L2 and the MALL; the next pass
// reads them from HBM): 1 = non-temporal stores.  Used by the plain and first column passes and the
// FT-table rows, whose stores cover whole rows.  Not by the scattered stores of the packed epilogue,
// which non-temporal stores slow down, nor by the fused rows, whose rows of lz_out entries start off
// the 64-byte boundaries: there the non-temporal stores wrote 1.3 times the bytes for the same time.
#ifndef PB3D_NT_STORE
#define PB3D_NT_STORE 1
#endif
__device__ __forceinline__ void stream_store(float2* p, float2 v) {
#if PB3D_NT_STORE
  typedef float f2v __attribute__((ext_vector_type(2)));
  f2v w = {v.x, v.y};
  __builtin_nontemporal_store(w, reinterpret_cast<f2v*>(p));
#else
  *p = v;
#endif
}
// row passes over several batch entries: XCD-aware block order (see r2c_rows)
#ifndef PB3D_ROW_XCD
#define PB3D_ROW_XCD 1
#endif
// plane-aligned y-pass tiles (see launch_cols)
#ifndef PB3D_COL_ALIGN
#define PB3D_COL_ALIGN 1
#endif
// minimum waves per SIMD the fused pointing rows are compiled for (4: at most 128 VGPRs)
#ifndef PB3D_POINT_WAVES
#define PB3D_POINT_WAVES 4
#endif
// the same for the combined-row instances, which hold the Lmax+1 combined rows next to the
// accumulators: 3 waves per SIMD (168 VGPRs) for Lmax = 2, 2 for Lmax = 4 (no spills)
#ifndef PB3D_POINT_WAVES_COMB
#define PB3D_POINT_WAVES_COMB 3
#endif
#ifndef PB3D_POINT_WAVES_COMB5
#define PB3D_POINT_WAVES_COMB5 2
#endif
// Last Stockham stage with its outputs delivered to registers: for the last stage NB = L/R = NS,
// so butterfly j = t + b*P < NS writes positions j + k*NS = t + P*(b + k*NS/P); the sink gets the
// compile-time slot e = b + k*NS/P (slot e of thread t holds position t + e*P).
template <int L, int R, int NS, int P, int DIR, typename Src, typename Sink>
__device__ __forceinline__ void last_stage_to_regs(int t, Src src, Sink sink, const float2* tw) {
  constexpr int NB = L / R;
  constexpr int B = NB / P;
  static_assert(NB == NS && NB % P == 0, "last stage shape");
#pragma unroll
  for (int b = 0; b < B; ++b) {
    const int j = t + b * P;
    float2 v[R];
#pragma unroll
    for (int k = 0; k < R; ++k) v[k] = src(j + k * NB);
    if (NS > 1) {
#pragma unroll
      for (int k = 1; k < R; ++k) {
        float2 w = tw[j * k * (L / (NS * R))];
        if (DIR > 0) w.y = -w.y;
        v[k] = cmul(v[k], w);
      }
    }
    DFT<R, DIR>::run(v);
#pragma unroll
    for (int k = 0; k < R; ++k) sink(b + k * (NS / P), j + k * NS, v[k]);
  }
}

// First Stockham stage with inputs from registers: butterfly j = t + b*P reads positions
// j + k*NB = t + P*(b + k*NB/P), i.e. register slot e = b + k*NB/P of the same layout.
template <int L, int R, int P, int DIR, typename Get, typename Dst>
__device__ __forceinline__ void first_stage_from_regs(int t, Get get, Dst dst) {
  constexpr int NB = L / R;
  constexpr int B = NB / P;
  static_assert(NB % P == 0, "first stage shape");
#pragma unroll
  for (int b = 0; b < B; ++b) {
    const int j = t + b * P;
    float2 v[R];
#pragma unroll
    for (int k = 0; k < R; ++k) v[k] = get(b + k * (NB / P));
    DFT<R, DIR>::run(v);
#pragma unroll
    for (int k = 0; k < R; ++k) dst(j * R + k, v[k]);
  }
}

// Inverse pencil FFT whose final stage accumulates into registers (LDS source, all stages but the
// last in LDS).  Contains barriers: every thread of the block calls it.
// Split in two so that callers can specialise the sink per call: head = every stage but the last
// (in LDS, ends synchronised), tail = the last stage into registers.
template <int L, bool WAVE, typename Lds>
__device__ __forceinline__ void pencil_inverse_head(int t, Lds lds, const float2* tw) {
  using D = Decomp<L>;
  if constexpr (D::R1 != 1) {
    static_assert(D::R2 == 1, "row transforms use at most two stages");
    stage<L, D::R0, 1, D::P, +1, WAVE>(t, lds, lds, tw, true);
    lds_sync<WAVE>();
  }
}
template <int L, typename Lds, typename Sink>
__device__ __forceinline__ void pencil_inverse_tail(int t, Lds lds, Sink sink, const float2* tw) {
  using D = Decomp<L>;
  if constexpr (D::R1 == 1) last_stage_to_regs<L, D::R0, 1, D::P, +1>(t, lds, sink, tw);
  else last_stage_to_regs<L, D::R1, D::R0, D::P, +1>(t, lds, sink, tw);
}

// Forward pencil FFT whose first stage reads registers; result left in LDS (natural order).
template <int L, bool WAVE, typename Get, typename Lds>
__device__ __forceinline__ void pencil_forward_from_regs(int t, Get get, Lds lds, const float2* tw) {
  using D = Decomp<L>;
  constexpr int P = D::P;
  first_stage_from_regs<L, D::R0, P, -1>(t, get, lds);
  if constexpr (D::R1 != 1) {
    static_assert(D::R2 == 1, "row transforms use at most two stages");
    lds_sync<WAVE>();
    stage<L, D::R1, D::R0, P, -1, WAVE>(t, lds, lds, tw, true);
  }
}

template <int L> struct ColCfg {
  static constexpr int T = (PB3D_COL_THREADS / Decomp<L>::P) > 8 ? PB3D_COL_THREADS / Decomp<L>::P : 8;
};
template <> struct ColCfg<2>    { static constexpr int T = 256; };
template <> struct ColCfg<4>    { static constexpr int T = 256; };
template <> struct ColCfg<8>    { static constexpr int T = 128; };
template <> struct ColCfg<16>   { static constexpr int T = 128; };
// Columns per tile by pass: the first C2R pass (mode-op prologue) and the last R2C pass (mode-op
// epilogue) gather from mode tables and are latency bound; half-width tiles halve their LDS, so two
// blocks share a CU and one's table gathers overlap the other's butterflies (measured at 256^3:
// 8 % faster for both).  The plain passes stream whole planes and keep the wide tiles (32-column
// tiles there cost 35 %).
#ifndef PB3D_COL_NARROW
#define PB3D_COL_NARROW 1
#endif
template <int L, int MODE> struct ColTile {
  static constexpr int T = (PB3D_COL_NARROW && MODE != 0 && ColCfg<L>::T / 2 >= 16) ? ColCfg<L>::T / 2 : ColCfg<L>::T;
};

// ------------------------------------------------------------------------------------------
// Column pass along axis 0 or 1 of the half-complex grid, optionally pruned to a band limit.
//   active column c in [0, ncols): outer(q) = q < split ? q : q + gap,
//   base(c) = outer(c / cpp) * plane + (c % cpp); element p of the pencil at base + p*S in
//   half-complex mode coordinates (what the KOp sees).  The input and output arrays are addressed
//   the same way with their own (plane, S): the half-complex layout itself, or a compact layout
//   that keeps only the kz < pitch planes of a band-limited map (pitch along z instead of n2/2+1).
//   Elements with p outside the radius rad_in (|fs(p)| > rad_in) are known zeros (not read);
//   outputs outside rad_out are not needed (not written).  rad < 0 = whole pencil.
//   MODE 0: plain.  MODE 1: first C2R pass, input through KOp (broadcast by x_stride).
//   MODE 2: last R2C pass, output through KOp (may accumulate / pack).
// ------------------------------------------------------------------------------------------
//   tpp > 0: plane-aligned tiles, tpp tiles per plane (tile b covers columns (b % tpp)*T + col of
//   plane b / tpp), so that no tile straddles two planes, whose data lie far apart in the y pass.
struct ColMap {
  const unsigned char* row_on;     // MODE 0: (plane, position) rows not stored where 0 (NULL: all)
  const unsigned char* row_zero;   // MODE 0: (plane, position) rows read as zeros where 0 (NULL: none)
  int row_n1;                      // positions per plane in row_on / row_zero
  int ncols, cpp, split, gap, tpp;
  long long plane, S;          // mode coordinates
  long long in_plane, in_S;    // input addressing
  long long out_plane, out_S;  // output addressing
  int rad_in, rad_out;
};

__device__ __forceinline__ bool in_radius(int p, int L, int rad) {
  return rad < 0 || p <= rad || p >= L - rad;
}

template <int L, int R, int P, int DIR, typename Get, typename Dst>
__device__ __forceinline__ void first_stage_from_regs(int t, Get get, Dst dst);
template <int L, int R, int NS, int P, int DIR, typename Src, typename Sink>
__device__ __forceinline__ void last_stage_to_regs(int t, Src src, Sink sink, const float2* tw);

#ifndef PB3D_ROW_GROUP
#define PB3D_ROW_GROUP 16
#endif
#ifndef PB3D_BKQ_BATCH
#define PB3D_BKQ_BATCH 1
#endif
// positions per batch of table reads in the Q-map prologue: 4 keeps the pass at 121 VGPRs (two
// 512-thread blocks per CU); 8 and 16 need 129 and 143 and measured 30 % slower at 256^3
#ifndef PB3D_BKQ_GROUP
#define PB3D_BKQ_GROUP 4
#endif

// The BKQ prologue of one pencil (kop_value case 4 for the E = L/P positions p = t + e*P of this
// thread) with its table reads batched: per group of 8 positions, first the bin and pack-index
// loads, then the packed-row gathers, then the weights, so that the dependent loads of the
// positions overlap instead of running one chain at a time.
template <int L, int E>
__device__ __forceinline__ void bkq_gather(const KOp& op, const Geo& g, long long bt, long long base, long long S,
                                           int t, bool valid, int rad_in, int mj, int ml, float2 (&vals)[E]) {
  constexpr int P = L / E;
  constexpr int CH = E < PB3D_BKQ_GROUP ? E : PB3D_BKQ_GROUP;
  const int* tr = op.trip + 4 * (long long)(op.trip0 + bt);
  const int b1 = tr[0], b2 = tr[1], b3 = tr[2], ell = tr[3];
  const int pr3 = (op.terms & 5) ? op.pair_index[b1 * op.nk + b2] : -1;
  const int pr1 = (op.terms & 2) ? op.pair_index[b2 * op.nk + b3] : -1;
  const int pr2 = (op.terms & 2) ? op.pair_index[b1 * op.nk + b3] : -1;
  const float2* fl = op.ft + (long long)(ell / 2) * op.ft_lstride;
#pragma unroll
  for (int c0 = 0; c0 < E; c0 += CH) {
    int bn[CH], pk[CH];
#pragma unroll
    for (int u = 0; u < CH; ++u) {
      const int p = t + (c0 + u) * P;
      const bool ok = valid && in_radius(p, L, rad_in);
      const long long md = base + (long long)p * S;
      bn[u] = ok ? op.bin_table[md] : PB3D_NO_BIN;
      pk[u] = ok ? op.pack_index[md] : -1;
    }
    float2 f3[CH], f1[CH], f2[CH];
#pragma unroll
    for (int u = 0; u < CH; ++u) {
      const bool in = bn[u] != PB3D_NO_BIN && pk[u] >= 0;
      f3[u] = (in && bn[u] == b3 && pr3 >= 0) ? op.ft[(long long)pr3 * op.packed_len + pk[u]] : make_float2(0.f, 0.f);
      f1[u] = (in && bn[u] == b1 && pr1 >= 0) ? fl[(long long)pr1 * op.packed_len + pk[u]] : make_float2(0.f, 0.f);
      f2[u] = (in && bn[u] == b2 && pr2 >= 0) ? fl[(long long)pr2 * op.packed_len + pk[u]] : make_float2(0.f, 0.f);
    }
#pragma unroll
    for (int u = 0; u < CH; ++u) {
      const int p = t + (c0 + u) * P;
      float2 acc = make_float2(0.f, 0.f);
      if (bn[u] == b3 && pr3 >= 0 && pk[u] >= 0) {
        float kx, ky, kz; kvec(g, p, mj, ml, kx, ky, kz);
        const float k2 = kx * kx + ky * ky + kz * kz;
        const float inv = k2 > 0.f ? rsqrtf(k2) : 0.f;
        if (op.terms & 1) {
          float w = 1.f;
          if (op.leg3_legendre && ell != 0) w = legendre_f(ell, (g.s0 * kx + g.s1 * ky + g.s2 * kz) * inv);
          acc = cadd(acc, cscale(f3[u], w));
        }
        if (op.terms & 4) acc = cadd(acc, cscale(f3[u], ylm_f(op.ell, op.m, kx * inv, ky * inv, kz * inv)));
      }
      acc = cadd(cadd(acc, f1[u]), f2[u]);
      vals[c0 + u] = acc;
    }
  }
}

template <int L, int DIR, int MODE>
__global__ void __launch_bounds__((Decomp<L>::P * ColTile<L, MODE>::T))
fft_cols(const float2* in, float2* out, ColMap cm, long long in_bstride, long long out_bstride,
         const float2* __restrict__ tw, Geo geo, KOp kop) {
  constexpr int T = ColTile<L, MODE>::T;
  constexpr int LD = T + 1;
  __shared__ float2 lds[(Decomp<L>::R1 == 1) ? 1 : L * LD];
  __shared__ float2 stw[L];
  const int tid = threadIdx.x;
  // twiddles to LDS; first used after the barrier that follows stage 0 (stage 0 needs none)
  for (int i = tid; i < L; i += Decomp<L>::P * T) stw[i] = tw[i];
  const int col = tid % T;
  const int t = tid / T;
  const long long bt = blockIdx.y;
  int c;
  bool valid;
  if (cm.tpp > 0) {
    const int lc = (blockIdx.x % cm.tpp) * T + col;
    valid = lc < cm.cpp;
    c = (blockIdx.x / cm.tpp) * cm.cpp + lc;
  } else {
    c = blockIdx.x * T + col;
    valid = c < cm.ncols;
  }
  long long base = 0, bin = 0, bout = 0;
  int qplane = 0;
  if (valid) {
    int q = c / cm.cpp;
    const int l = c % cm.cpp;
    q = q < cm.split ? q : q + cm.gap;
    qplane = q;
    base = (long long)q * cm.plane + l;
    bin = (long long)q * cm.in_plane + l;
    bout = (long long)q * cm.out_plane + l;
  }
  // the input of a C2R first pass steps once per kop.x_group batch entries
  const float2* src = in + (MODE == 1 ? bt / kop.x_group : bt) * in_bstride;
  float2* dst = out + bt * out_bstride;
  // mode coordinates for KOp (x-column pass only: base = j*n2h + l)
  int mj = 0, ml = 0;
  if (MODE != 0) { mj = (int)(base / geo.n2h); ml = (int)(base % geo.n2h); }

  auto load = [&](int p) -> float2 {
    if (!valid || !in_radius(p, L, cm.rad_in)) return make_float2(0.f, 0.f);
    if (MODE == 1 && kop.kind == 4)  // BKQ assembles the map from packed rows; no input read
      return kop_value(kop, geo, bt, base + p * cm.S, p, mj, ml, make_float2(0.f, 0.f));
    if (MODE == 0 && cm.row_zero && !cm.row_zero[(long long)qplane * cm.row_n1 + p]) return make_float2(0.f, 0.f);
    float2 v = src[bin + p * cm.in_S];
    if (MODE == 1) v = kop_value(kop, geo, bt, base + p * cm.S, p, mj, ml, v);
    return v;
  };
  auto store = [&](int p, float2 v) {
    if (!valid || !in_radius(p, L, cm.rad_out)) return;
    if (MODE == 2) {
      long long md = base + p * cm.S;
      if (kop.kind == 5) {  // packed output (in-bin modes only)
        const int pi = kop.pack_index[md];
        if (pi < 0) return;
        v = cscale(v, kop.scale);
        if (kop.accumulate) v = cadd(v, dst[pi]);
        dst[pi] = v;
        return;
      }
      v = kop_value(kop, geo, bt, md, p, mj, ml, v);
      v = cscale(v, kop.scale);
      const long long od = bout + p * cm.out_S;
      if (kop.accumulate) v = cadd(v, dst[od]);
      dst[od] = v;
    } else {
      if (MODE == 0 && cm.row_on && !cm.row_on[(long long)qplane * cm.row_n1 + p]) return;
      stream_store(dst + bout + p * cm.out_S, v);
    }
  };
  struct LdsRW {
    float2* l; int LDv; int colv;
    __device__ float2 operator()(int p) const { return l[p * LDv + colv]; }
    __device__ void operator()(int p, float2 v) const { l[p * LDv + colv] = v; }
  } lrw{lds, LD, col};
  if constexpr (MODE == 1 && PB3D_BKQ_BATCH) {
    if (kop.kind == 4) {
      using D = Decomp<L>;
      constexpr int E = L / D::P;
      float2 vals[E];
      bkq_gather<L, E>(kop, geo, bt, base, cm.S, t, valid, cm.rad_in, mj, ml, vals);
      auto get = [&](int e) { return vals[e]; };
      // pencil_fft with the first stage fed from registers (slot e holds position t + e*P)
      if constexpr (D::R1 == 1) {
        first_stage_from_regs<L, D::R0, D::P, DIR>(t, get, store);
      } else if constexpr (D::R2 == 1) {
        first_stage_from_regs<L, D::R0, D::P, DIR>(t, get, lrw);
        __syncthreads();
        stage<L, D::R1, D::R0, D::P, DIR>(t, lrw, store, stw, true);
      } else {
        first_stage_from_regs<L, D::R0, D::P, DIR>(t, get, lrw);
        __syncthreads();
        stage<L, D::R1, D::R0, D::P, DIR>(t, lrw, lrw, stw, true);
        __syncthreads();
        stage<L, D::R2, D::R0 * D::R1, D::P, DIR>(t, lrw, store, stw, true);
      }
      return;
    }
  }
  if constexpr (MODE == 2 && PB3D_BKQ_BATCH) {
    if (kop.kind == 5) {
      // packed output: the pack indices of this thread's outputs do not depend on the data, so
      // they are loaded up front and the last stage stores straight from registers
      using D = Decomp<L>;
      constexpr int E = L / D::P;
      int pk[E];
#pragma unroll
      for (int e = 0; e < E; ++e) {
        const int p = t + e * D::P;
        pk[e] = (valid && in_radius(p, L, cm.rad_out)) ? kop.pack_index[base + (long long)p * cm.S] : -1;
      }
      auto sink = [&](int e, int, float2 v) {
        if (pk[e] < 0) return;
        v = cscale(v, kop.scale);
        if (kop.accumulate) v = cadd(v, dst[pk[e]]);
        dst[pk[e]] = v;
      };
      if constexpr (D::R1 == 1) {
        last_stage_to_regs<L, D::R0, 1, D::P, DIR>(t, load, sink, stw);
      } else if constexpr (D::R2 == 1) {
        stage<L, D::R0, 1, D::P, DIR>(t, load, lrw, stw, false);
        __syncthreads();
        last_stage_to_regs<L, D::R1, D::R0, D::P, DIR>(t, lrw, sink, stw);
      } else {
        stage<L, D::R0, 1, D::P, DIR>(t, load, lrw, stw, false);
        __syncthreads();
        stage<L, D::R1, D::R0, D::P, DIR>(t, lrw, lrw, stw, true);
        __syncthreads();
        last_stage_to_regs<L, D::R2, D::R0 * D::R1, D::P, DIR>(t, lrw, sink, stw);
      }
      return;
    }
  }
  pencil_fft<L, DIR>(t, load, lrw, store, stw, false);
}

// ------------------------------------------------------------------------------------------
// Row pass (axis 2).  R2C: real rows of n2 = 2M floats packed as M complex, FFT_M, then split
// into the n2/2+1 half spectrum (only kz < lz_out written).  C2R: the inverse (numpy/pocketfft
// semantics: Im of the kz=0 and kz=M entries ignored; entries kz >= lz_in known zero), with 1/N
// and the real-space epilogue fused into the store.
// LDS holds a tile of T rows transposed, [position][row] with the row index fastest, and the
// lanes of a wave walk the rows: every butterfly read and write is then bank-conflict free (the
// layout of the column kernel).  Global rows are contiguous, so loads and stores stage through
// the tile with coalesced flat indexing.  Twiddles are copied to LDS once per block.
// ------------------------------------------------------------------------------------------
template <int M> struct RowCfg { static constexpr int T = (256 / Decomp<M>::P) > 8 ? 256 / Decomp<M>::P : 8; };
template <> struct RowCfg<2>  { static constexpr int T = 128; };
template <> struct RowCfg<4>  { static constexpr int T = 128; };
template <> struct RowCfg<8>  { static constexpr int T = 64; };
template <> struct RowCfg<16> { static constexpr int T = 64; };

// accessor for row r of a transposed tile: element p at l[p * LDT + r + PADQ * (p / 16)] (PADQ > 0:
// every 16 positions shifted by PADQ entries, so that a radix-16 stage's writes to positions 16j + k
// of different j fall on different LDS banks)
template <int PADQ>
__device__ __forceinline__ int tile_at(int p, int LDT, int r) { return p * LDT + r + PADQ * (p >> 4); }
template <int PADQ = 0>
struct RowTp {
  float2* l; int LDT; int r;
  __device__ float2 operator()(int p) const { return l[tile_at<PADQ>(p, LDT, r)]; }
  __device__ void operator()(int p, float2 v) const { l[tile_at<PADQ>(p, LDT, r)] = v; }
};
using RowT = RowTp<0>;

// LDS pitch of the transposed tile: T rows (+1 to spread the flat staging accesses)
template <int M> struct RowTile {
  static constexpr int T = RowCfg<M>::T;
  static constexpr int P = Decomp<M>::P;
  static constexpr int NT = P * T;
  static constexpr int LDT = T + 1;
  static constexpr int SIZE = (M + 1) * LDT;
};

template <int M>
__device__ __forceinline__ void stage_twiddles(float2* stw, float2* stwN, const float2* __restrict__ tw,
                                               const float2* __restrict__ twN, int tid, int nt) {
  for (int i = tid; i < M; i += nt) stw[i] = tw[i];
  for (int i = tid; i <= M; i += nt) stwN[i] = twN[i];
}

// X[k] = E[k] + W^k O[k], E = (Z[k]+conj Z[M-k])/2, O = (Z[k]-conj Z[M-k])/(2i)   (row r of a tile)
template <int M, int PADQ = 0>
__device__ __forceinline__ float2 r2c_split(const float2* b, int LDT, int r, int k, const float2* twN) {
  const float2 zk = b[tile_at<PADQ>(k == M ? 0 : k, LDT, r)];
  const float2 zm = cconj(b[tile_at<PADQ>(k == 0 ? 0 : M - k, LDT, r)]);
  const float2 e = make_float2(0.5f * (zk.x + zm.x), 0.5f * (zk.y + zm.y));
  const float2 d = make_float2(0.5f * (zk.x - zm.x), 0.5f * (zk.y - zm.y));
  const float2 od = make_float2(d.y, -d.x);
  return cadd(e, cmul(twN[k], od));
}

// Load a tile of T half-spectrum rows from global memory and apply the C2R pre-pass on the fly
// (pairs (k, M-k) read together, Z' written to the transposed LDS tile): one staging step, no
// separate LDS pass.  Entries kz >= lz_in are known zeros; Im of kz=0 and kz=M is ignored.  The
// trip count is a compile-time constant and every load is issued before the first one is used, so
// a thread waits for HBM once per tile rather than once per element.  Row rr of the source starts
// at src + rr * pitch (pitch >= lz_in).
template <int M, int T, int NT>
__device__ __forceinline__ void load_prepass(float2* buf, int LDT, const float2* __restrict__ src, int pitch,
                                             long long rows_left, int lz_in, int tid, const float2* twN) {
  constexpr int K = M / 2 + 1;
  constexpr int TOT = T * K;
  constexpr int IT = (TOT + NT - 1) / NT;
  float2 xk[IT], xm[IT];
#pragma unroll
  for (int it = 0; it < IT; ++it) {
    const int f = tid + it * NT;
    const int rr = f / K, k = f % K;
    xk[it] = make_float2(0.f, 0.f);
    xm[it] = make_float2(0.f, 0.f);
    if ((IT * NT == TOT || f < TOT) && rr < rows_left) {
      if (k < lz_in) xk[it] = src[(long long)rr * pitch + k];
      if (M - k < lz_in) xm[it] = src[(long long)rr * pitch + M - k];
    }
  }
#pragma unroll
  for (int it = 0; it < IT; ++it) {
    const int f = tid + it * NT;
    if (IT * NT != TOT && f >= TOT) continue;
    const int rr = f / K, k = f % K;
    float2 a = xk[it], c = xm[it];
    if (k == 0) { a.y = 0.f; c.y = 0.f; }   // (k, M) pair: both self-conjugate entries
    if (2 * k == M) c = a;
    {
      const float2 sm = cadd(a, cconj(c));
      const float2 d = csub(a, cconj(c));
      const float2 wd = cmul(cconj(twN[k]), d);
      buf[k * LDT + rr] = make_float2(sm.x - wd.y, sm.y + wd.x);
    }
    if (k != 0 && 2 * k != M) {
      const float2 sm = cadd(c, cconj(a));
      const float2 d = csub(c, cconj(a));
      const float2 wd = cmul(cconj(twN[M - k]), d);
      buf[(M - k) * LDT + rr] = make_float2(sm.x - wd.y, sm.y + wd.x);
    }
  }
}

// The same staging with lanes walking the rows: lane (rr = tid % T, kb = tid / T) handles
// k = kb + it * (NT / T), so every address offset is a compile-time immediate and no division is
// needed.  Used where a tile has few rows (one wave's rows).  Needs rows_left >= 1: rows past the
// end are read as row 0 and replaced by zeros.  Two uniform regimes: lz_in <= M/2+1 (the right
// half X[M-k] is a known zero apart from k = M/2, so each pair needs one load and one complex
// multiply per output; loads are unpredicated at clamped addresses), lz_in > M/2+1 (both halves
// read, general pair formula).  Row rr starts at src + rr * pitch (pitch >= lz_in).
template <int M, int T, int NT, int PADQ = 0>
__device__ __forceinline__ void load_prepass_rows(float2* buf, int LDT, const float2* __restrict__ src, int pitch,
                                                  long long rows_left, int lz_in, int tid, const float2* twN) {
  constexpr int K = M / 2 + 1;
  constexpr int KS = NT / T;
  constexpr int IT = (K + KS - 1) / KS;
  static_assert(NT % T == 0, "tile shape");
  const int rr = tid % T, kb = tid / T;
  const bool rv = rr < rows_left;
  const float2* rowp = src + (rv ? rr : 0) * pitch;
  const float2 zero = make_float2(0.f, 0.f);
  if (lz_in <= K) {
    // With X[M-k] = 0 (k != M/2): Z'[k] = a (1 + i conj W^k) and Z'[M-k] = conj(a) (1 - i conj W^(M-k)),
    // W^k = twN[k]; at k = M/2 the pair is (a, a): Z' = (2 Re a, -2 Im a).
    float2 xa[IT];
#pragma unroll
    for (int it = 0; it < IT; ++it) {
      const int k = kb + it * KS;
      const float2 v = rowp[k < lz_in ? k : lz_in - 1];
      xa[it] = (rv && k < lz_in) ? v : zero;
    }
#pragma unroll
    for (int it = 0; it < IT; ++it) {
      const int k = kb + it * KS;
      if (IT * KS != K && k >= K) continue;
      float2 a = xa[it];
      if (k == 0) a.y = 0.f;  // Im of the DC entry is ignored
      const float2 w = twN[k];
      const float p = 1.f + w.y;
      float2 z1 = make_float2(fmaf(a.x, p, -a.y * w.x), fmaf(a.x, w.x, a.y * p));
      if (2 * k == M) z1 = make_float2(2.f * a.x, -2.f * a.y);
      buf[tile_at<PADQ>(k, LDT, rr)] = z1;
      if (k != 0 && 2 * k != M) {
        const float2 w2 = twN[M - k];
        const float q = 1.f - w2.y;
        buf[tile_at<PADQ>(M - k, LDT, rr)] = make_float2(fmaf(a.x, q, -a.y * w2.x), -fmaf(a.x, w2.x, a.y * q));
      }
    }
    return;
  }
  float2 xk[IT], xm[IT];
  {
#pragma unroll
    for (int it = 0; it < IT; ++it) {
      const int k = kb + it * KS;
      xk[it] = zero;
      xm[it] = zero;
      if (rv && (IT * KS == K || k < K)) {
        xk[it] = rowp[k];
        if (M - k < lz_in) xm[it] = rowp[M - k];
      }
    }
  }
#pragma unroll
  for (int it = 0; it < IT; ++it) {
    const int k = kb + it * KS;
    if (IT * KS != K && k >= K) continue;
    float2 a = xk[it], c = xm[it];
    if (k == 0) { a.y = 0.f; c.y = 0.f; }   // (k, M) pair: both self-conjugate entries
    if (2 * k == M) c = a;
    {
      const float2 sm = cadd(a, cconj(c));
      const float2 d = csub(a, cconj(c));
      const float2 wd = cmul(cconj(twN[k]), d);
      buf[tile_at<PADQ>(k, LDT, rr)] = make_float2(sm.x - wd.y, sm.y + wd.x);
    }
    if (k != 0 && 2 * k != M) {
      const float2 sm = cadd(c, cconj(a));
      const float2 d = csub(c, cconj(a));
      const float2 wd = cmul(cconj(twN[M - k]), d);
      buf[tile_at<PADQ>(M - k, LDT, rr)] = make_float2(sm.x - wd.y, sm.y + wd.x);
    }
  }
}

// linear block b is dispatched to XCD b % 8 (round robin), so row blocks are dealt out in groups of 8
// with all maps of one row block on one XCD, whose L2 then serves the rows every map reads (the mask
// rows of the fused rows, the fixed operand of the FT-table rows) to all of them.  (map t, row block
// rb) from b, a bijection onto [0, batch) x [0, nrb) whatever the dispatch order.
__device__ __forceinline__ void xcd_rows_block(long long b, long long nrb, long long batch, long long& t,
                                               long long& rb) {
  const long long full = (nrb / 8) * 8 * batch;
  if (b < full) {
    const long long j = b / 8;
    t = j % batch;
    rb = (j / batch) * 8 + b % 8;
  } else {
    const long long rem = nrb % 8, q = b - full;
    t = q / rem;
    rb = (nrb / 8) * 8 + q % rem;
  }
}

template <int M>
__global__ void __launch_bounds__(RowTile<M>::NT)
r2c_rows(const float* __restrict__ in, float2* __restrict__ out, long long nrows, int rows_per_batch,
         const float2* __restrict__ tw, const float2* __restrict__ twN, Geo geo, ROp rop, int lz_out, int pitch) {
  using RT = RowTile<M>;
  constexpr int T = RT::T, NT = RT::NT, LDT = RT::LDT;
  __shared__ float2 buf[RT::SIZE];
  __shared__ float2 stw[M], stwN[M + 1];
  const int tid = threadIdx.x;
  stage_twiddles<M>(stw, stwN, tw, twN, tid, NT);
  long long row0 = (long long)blockIdx.x * T;
  if (PB3D_ROW_XCD && rows_per_batch % T == 0 && nrows > rows_per_batch) {
    // every batch entry of one row block on one XCD, whose L2 then serves the rows of an operand the
    // entries share (stride 0: the fixed map A of the FT-table rows) to all of them
    long long t, rb;
    xcd_rows_block(blockIdx.x, rows_per_batch / T, nrows / rows_per_batch, t, rb);
    row0 = t * rows_per_batch + rb * T;
  }
  // 1. coalesced load of T rows (T*M float2) with the voxel prologue (T*M is a multiple of NT)
  static_assert((T * M) % NT == 0, "row tile");
  constexpr int IT = (T * M) / NT;
  const float* abase = rop.a ? rop.a : in;
  const long long astride = rop.a ? rop.sa : (long long)rows_per_batch * 2 * M;
  const bool prod = (rop.kind == 1 || rop.kind == 2) &&
                    ((reinterpret_cast<unsigned long long>(abase) | reinterpret_cast<unsigned long long>(rop.b) |
                      (rop.kind == 2 ? (reinterpret_cast<unsigned long long>(rop.c) |
                                        reinterpret_cast<unsigned long long>(rop.d)) : 0ull)) & 7ull) == 0 &&
                    ((astride | rop.sb | (rop.kind == 2 ? (rop.sc | rop.sd) : 0ll)) & 1ll) == 0;
  if (prod) {
    // Products of maps (a*b, a*b - c*d): every operand read as a float2, and the loads of
    // PB3D_ROW_GROUP iterations (by default all of them) issued before any is used, so that they
    // overlap instead of waiting one by one (measured 0.98 -> 0.71 ms per launch for the 15-entry
    // FT-table rows at 256^3).
    constexpr int G4 = IT < PB3D_ROW_GROUP ? IT : PB3D_ROW_GROUP;
#pragma unroll
    for (int g0 = 0; g0 < IT; g0 += G4) {
      float2 va[G4], vb[G4], vc[G4], vd[G4];
#pragma unroll
      for (int u = 0; u < G4; ++u) {
        const int f = tid + (g0 + u) * NT;
        const int r = f / M, m = f % M;
        const long long row = row0 + r;
        const bool ok = row < nrows;
        const long long rowc = ok ? row : 0;
        const long long bt = rowc / rows_per_batch, rr = rowc % rows_per_batch;
        const long long vox = rr * (2 * M) + 2 * m;
        const float2 z = make_float2(0.f, 0.f);
        va[u] = ok ? *reinterpret_cast<const float2*>(abase + bt * astride + vox) : z;
        vb[u] = ok ? *reinterpret_cast<const float2*>(rop.b + bt * rop.sb + vox) : z;
        vc[u] = (ok && rop.kind == 2) ? *reinterpret_cast<const float2*>(rop.c + bt * rop.sc + vox) : z;
        vd[u] = (ok && rop.kind == 2) ? *reinterpret_cast<const float2*>(rop.d + bt * rop.sd + vox) : z;
      }
#pragma unroll
      for (int u = 0; u < G4; ++u) {
        const int f = tid + (g0 + u) * NT;
        const int r = f / M, m = f % M;
        const float x0 = va[u].x * vb[u].x - vc[u].x * vd[u].x;
        const float x1 = va[u].y * vb[u].y - vc[u].y * vd[u].y;
        buf[m * LDT + r] = make_float2(x0 * rop.scale, x1 * rop.scale);
      }
    }
  }
#pragma unroll
  for (int it = 0; it < (prod ? 0 : IT); ++it) {
    const int f = tid + it * NT;
    const int r = f / M, m = f % M;
    const long long row = row0 + r;
    float2 v = make_float2(0.f, 0.f);
    if (row < nrows) {
      const long long bt = row / rows_per_batch;
      const long long rr = row % rows_per_batch;
      const long long vox = rr * (2 * M) + 2 * m;
      const float* a = rop.a ? rop.a + bt * rop.sa : in + bt * (long long)rows_per_batch * 2 * M;
      float x0 = a[vox], x1 = a[vox + 1];
      if (rop.kind != 0) {
        const int i0 = (int)(rr / geo.n1), i1 = (int)(rr % geo.n1);
        x0 = rop_value(rop, geo, bt, vox, i0, i1, 2 * m, x0);
        x1 = rop_value(rop, geo, bt, vox + 1, i0, i1, 2 * m + 1, x1);
      }
      v = make_float2(x0 * rop.scale, x1 * rop.scale);
    }
    buf[m * LDT + r] = v;
  }
  __syncthreads();
  // 2. complex FFT of each row in LDS (lanes walk the rows)
  const int r = tid % T, t = tid / T;
  RowT rw{buf, LDT, r};
  pencil_fft<M, -1>(t, rw, rw, rw, stw, true);
  __syncthreads();
  // 3. split into the half spectrum: kz < lz_out of row r at out + r * pitch
  float2* o = out + row0 * pitch;
  constexpr int ITS = (T * (M + 1) + NT - 1) / NT;
#pragma unroll
  for (int it = 0; it < ITS; ++it) {
    const int f = tid + it * NT;
    const int rr = f / (M + 1), k = f % (M + 1);
    if (f >= T * (M + 1) || row0 + rr >= nrows || k >= lz_out) continue;
    stream_store(o + (long long)rr * pitch + k, r2c_split<M>(buf, LDT, rr, k, stwN));
  }
}

template <int M>
__global__ void __launch_bounds__(RowTile<M>::NT)
c2r_rows(const float2* __restrict__ in, float* __restrict__ out, long long nrows, int rows_per_batch,
         const float2* __restrict__ tw, const float2* __restrict__ twN, Geo geo, ROp rop, float norm, int lz_in,
         int pitch) {
  using RT = RowTile<M>;
  constexpr int T = RT::T, NT = RT::NT, LDT = RT::LDT;
  __shared__ float2 buf[RT::SIZE];
  __shared__ float2 stw[M], stwN[M + 1];
  const int tid = threadIdx.x;
  stage_twiddles<M>(stw, stwN, tw, twN, tid, NT);
  const long long row0 = (long long)blockIdx.x * T;
  const float2* src = in + row0 * pitch;
  __syncthreads();  // twiddles staged
  load_prepass<M, T, NT>(buf, LDT, src, pitch, nrows - row0, lz_in, tid, stwN);
  __syncthreads();
  const int r = tid % T, t = tid / T;
  RowT rw{buf, LDT, r};
  pencil_fft<M, +1>(t, rw, rw, rw, stw, true);
  __syncthreads();
#pragma unroll
  for (int it = 0; it < (T * M) / NT; ++it) {
    const int f = tid + it * NT;
    const int rr = f / M, m = f % M;
    const long long row = row0 + rr;
    if (row >= nrows) continue;
    const long long bt = row / rows_per_batch;
    const long long rw_ = row % rows_per_batch;
    const long long vox = rw_ * (2 * M) + 2 * m;
    const float2 z = buf[m * LDT + rr];
    float x0 = z.x * norm, x1 = z.y * norm;
    float* o = out + bt * (long long)rows_per_batch * 2 * M;
    if (rop.kind != 0) {
      const int i0 = (int)(rw_ / geo.n1), i1 = (int)(rw_ % geo.n1);
      x0 = rop_value(rop, geo, bt, vox, i0, i1, 2 * m, x0);
      x1 = rop_value(rop, geo, bt, vox + 1, i0, i1, 2 * m + 1, x1);
    }
    x0 *= rop.scale; x1 *= rop.scale;
    if (rop.accumulate) {
      float2 prev = *reinterpret_cast<const float2*>(o + vox);
      x0 += prev.x; x1 += prev.y;
    }
    *reinterpret_cast<float2*>(o + vox) = make_float2(x0, x1);
  }
}

// ------------------------------------------------------------------------------------------
// Fused pointing row kernel: for output map t and row (i0,i1): invert the rows of all nsub
// sub-maps (work map g*batch + t), accumulate coef_g * Y_g(x_hat) * z in registers, multiply by the
// mask and scale, forward-transform the row and write kz < lz_out into output map t.  Work map j
// starts at work + j * slot_in, output map t at out + t * slot_out; row r of a map at + r * pitch
// (compact maps: pitch_in >= lz_in, pitch_out >= lz_out).
// Each wave works on its own RW = 64/P rows with its own transposed LDS tile, so the exchanges
// between FFT stages need only wave-level ordering: the sub-map loop has no block barrier.
// ------------------------------------------------------------------------------------------
constexpr int kMaxSub = 16;
constexpr int kMaxMono = 15;  // monomials of a homogeneous polynomial of degree <= 4
// Real-space weight of sub-map g: coef * Y_{ell,m}(r_hat) written as the homogeneous polynomial
// sum_{k<=ell} sum_{a<=ell-k} poly[idx] rx^a ry^(ell-k-a) rz^k / |r|^ell (idx runs over (k, a) in
// that order); y0 = Y(0) at r = 0, where the reference's zero-vector convention applies.
// The first nfact sub-maps are the ell = 2 ones: their weights are accumulated without the common
// 1/|r|^2, which multiplies the sum once before the remaining sub-maps are added.
// Combined form (nu > 0, every ell even, Lmax = nu - 1): w_g = cpoly_g(r) / |r|^Lmax with cpoly_g the
// sub-map's polynomial times coef * |r|^(Lmax - ell) (degree Lmax, same index order).  Along a row
// (rx, ry fixed) sum_g w_g z_g = |r|^-Lmax sum_u rz^u IFFT_z(sum_g D_ug S_g) with row constants D_ug,
// so the rows need nu inverse transforms instead of n.
struct SubW {
  const unsigned char* row_on;  // per (x, y) row: 0 = the mask (and IC weight) is zero on its group
  int n, nfact, nu;
  int ell[kMaxSub];
  float coef[kMaxSub];
  float y0[kMaxSub];
  float poly[kMaxSub][kMaxMono];
  float cpoly[kMaxSub][kMaxMono];
};

// Per-bin weighted sums of the combined real map (pb3d_icsum), by value for the kernel: each wave
// stores its FP64 partial sums in partial[(map * nwm + wave) * nbins + bin] (nwm = waves per map),
// and ic_finalize adds them to sums in wave order.
constexpr int kMaxIcBins = 64;
struct IcSum {
  const float* weight;  // nullptr: no sums
  const uint8_t* bin;   // nullptr: every voxel in bin 0
  int nbins;
  double* sums;
  double* partial;
  long long nwm;
};

template <int M> struct PointCfg {
  static constexpr int P = Decomp<M>::P;      // lanes per row
  static constexpr int RW = 64 / P;           // rows per wave
  static constexpr int WAVES = 4;
  static constexpr int NT = 64 * WAVES;
  static constexpr int RB = RW * WAVES;       // rows per block
  static constexpr int LDT = RW + 1;          // tile pitch: element p of row r at p * LDT + r (+ pad)
  static constexpr int PADQ = 0;              // bank padding per 16 positions (see RowTp): none
                                               // (4 measured no fewer conflict cycles at M = 128)
  static constexpr int SIZE = (M + 1) * LDT + PADQ * (M / 16 + 1);  // one wave's tile
  static constexpr int E = M / P;             // values per lane
  static_assert(64 % P == 0, "lanes per row must divide the wave");
};

// Band-limited C2R pre-pass (lz_in <= M/2, so X[k] = 0 for k >= M/2) from registers: slot it of
// lane (row r, tp) holds X[k], k = tp + P*it; writes Z'[k], Z'[M-k] of row r and Z'[M/2] = 0.
template <int M, int P, int IV, int PADQ>
__device__ __forceinline__ void prepass_regs(float2* buf, int LDT, int r, int tp, const float2 (&x)[IV],
                                             const float2* twN) {
#pragma unroll
  for (int it = 0; it < IV; ++it) {
    const int k = tp + P * it;
    float2 a = x[it];
    if (it == 0 && tp == 0) a.y = 0.f;  // Im of the DC entry is ignored
    const float2 w = twN[k];
    const float p = 1.f + w.y;
    buf[tile_at<PADQ>(k, LDT, r)] = make_float2(fmaf(a.x, p, -a.y * w.x), fmaf(a.x, w.x, a.y * p));
    if (!(it == 0 && tp == 0)) {
      const float2 w2 = twN[M - k];
      const float q = 1.f - w2.y;
      buf[tile_at<PADQ>(M - k, LDT, r)] = make_float2(fmaf(a.x, q, -a.y * w2.x), -fmaf(a.x, w2.x, a.y * q));
    }
  }
  if (tp == 0) buf[tile_at<PADQ>(M / 2, LDT, r)] = make_float2(0.f, 0.f);
}

// Block order of the fused pointing rows, whose blocks of different maps read the same mask rows:

template <int M, bool IC, int NU>
__global__ void __launch_bounds__(PointCfg<M>::NT, NU == 0 ? PB3D_POINT_WAVES : NU == 3 ? PB3D_POINT_WAVES_COMB : PB3D_POINT_WAVES_COMB5)
pointing_rows(const float2* work, long long slot_in, int pitch_in, float2* out, long long slot_out, int pitch_out,
              long long rows_per_map, int batch, SubW sw, IcSum ic,
              const float* __restrict__ mask, float scale, int lz_in, int lz_out, const float2* __restrict__ tw,
              const float2* __restrict__ twN, Geo geo) {
  using PC = PointCfg<M>;
  constexpr int RW = PC::RW, LDT = PC::LDT, E = PC::E, P = PC::P, PADQ = PC::PADQ;
  __shared__ float2 tiles[PC::WAVES * PC::SIZE];
  __shared__ float2 stw[M], stwN[M + 1];
  __shared__ float icl[IC ? PC::WAVES : 1][IC ? kMaxIcBins : 1];
  const int tid = threadIdx.x;
  stage_twiddles<M>(stw, stwN, tw, twN, tid, PC::NT);
  __syncthreads();  // twiddles staged: the only block barrier
  const int wave = __builtin_amdgcn_readfirstlane(tid / 64), lane = tid % 64;  // wave index: scalar
  float2* buf = tiles + wave * PC::SIZE;
  long long t, rb;
  xcd_rows_block(blockIdx.x, (rows_per_map + PC::RB - 1) / PC::RB, batch, t, rb);
  const long long row0 = rb * PC::RB + wave * RW;  // this wave's first row
  if (row0 >= rows_per_map) return;  // no rows for this wave (no block barrier follows)
  const int r = lane % RW, tp = lane / RW;
  if (sw.row_on) {
    // rows whose mask (and IC weight) is zero everywhere are zero rows of the output: a wave with no
    // other row leaves at once (its IC partials are zero); the C2R y pass did not store these rows and
    // the R2C y pass reads them as zeros, so they are not written either
    const long long rr = row0 + r;
    if (__ballot(rr < rows_per_map && sw.row_on[rr] != 0) == 0ull) {
      if constexpr (IC) {
        if (lane < ic.nbins) ic.partial[(t * ic.nwm + row0 / RW) * ic.nbins + lane] = 0.0;
      }
      return;
    }
  }
  RowTp<PADQ> rw{buf, LDT, r};
  // this lane's voxels after the inverse transforms: positions m = tp + e*P of row r
  const long long row = row0 + r;
  // grid indices of this lane's row (rows past the end take the last row): the wave's first row is
  // divided once (scalar), the lane offset (< RW) carried into i0
  const int dr = (int)((row < rows_per_map ? row : rows_per_map - 1) - row0);
  int i0 = (int)(row0 / geo.n1), i1 = (int)(row0 % geo.n1) + dr;
  while (i1 >= geo.n1) { i1 -= geo.n1; ++i0; }
  const float rx = freq_index(i0, geo.n0) * geo.dx0 + geo.c0;
  const float ry = freq_index(i1, geo.n1) * geo.dx1 + geo.c1;
  const float rxy2 = rx * rx + ry * ry;
  const float rzb = (float)(2 * tp) * geo.dx2 + geo.c2;  // z coordinate of position 2*tp
  static_assert((M / 2) % P == 0, "negative-frequency slots must be whole");
  float acc[E][2];  // set by the first transform (combined rows) or zeroed for the per-sub-map sums
  // acc *= 1/|r|^2 (the ell = 2 weights accumulated so far carry only their numerators); at r = 0
  // those numerators vanish exactly, as Y_2m(0) = 0 does
  auto rescale = [&](float rzb_v) {
    const float rzn = rzb_v - (float)geo.n2 * geo.dx2;
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const float rz0 = fmaf((float)(2 * e * P), geo.dx2, 2 * e >= E ? rzn : rzb_v), rz1 = rz0 + geo.dx2;
      acc[e][0] *= __builtin_amdgcn_rcpf(fmaxf(fmaf(rz0, rz0, rxy2), 1e-30f));
      acc[e][1] *= __builtin_amdgcn_rcpf(fmaxf(fmaf(rz1, rz1, rxy2), 1e-30f));
    }
  };
  bool combined = false;
  if constexpr (NU > 0) {
    // Combined rows: V_u = sum_g D_ug S_g in kz space (row constants D_ug = sum_a cpoly rx^a ry^b),
    // then acc = IFFT(V_{NU-1}) rz^(NU-1) + ... + IFFT(V_0) by Horner, divided by |r|^(NU-1).  Waves
    // holding a row with rx = ry = 0 (where r = 0 can occur) take the per-sub-map path.
    constexpr int IV = (M / 2) / P;  // kz slots per lane: k = tp + P*it < M/2
    constexpr int LM = NU - 1;
    if (sw.nu == NU && lz_in >= 1 && lz_in <= M / 2 && __ballot(rxy2 < 1e-20f) == 0ull) {
      combined = true;
      const bool rv = row < rows_per_map;
      // monomials rx^a ry^b, a + b = LM - u, in the index order of cpoly
      float mono[(LM + 1) * (LM + 2) / 2];
      {
        float px[LM + 1], py[LM + 1];
        px[0] = 1.f; py[0] = 1.f;
#pragma unroll
        for (int q = 1; q <= LM; ++q) { px[q] = px[q - 1] * rx; py[q] = py[q - 1] * ry; }
        int idx = 0;
#pragma unroll
        for (int u = 0; u <= LM; ++u)
#pragma unroll
          for (int a = 0; a <= LM - u; ++a) mono[idx++] = px[a] * py[LM - u - a];
      }
      float2 V[NU][IV];
      const int kl = lz_in - 1;
      // every row of the wave present and lz_in = M/2: no clamped addresses or zero selects
      const bool whole = lz_in == M / 2 && __ballot(!rv) == 0ull;
      auto load = [&](int g, float2 (&s)[IV]) {
        const float2* rowp = work + ((long long)g * batch + t) * slot_in + (row0 + (rv ? r : 0)) * pitch_in;
        if (whole) {
#pragma unroll
          for (int it = 0; it < IV; ++it) s[it] = rowp[tp + P * it];
          return;
        }
#pragma unroll
        for (int it = 0; it < IV; ++it) {
          const int k = tp + P * it;
          s[it] = rowp[k < lz_in ? k : kl];
        }
#pragma unroll
        for (int it = 0; it < IV; ++it) {
          const int k = tp + P * it;
          if (!(rv && k < lz_in)) s[it] = make_float2(0.f, 0.f);
        }
      };
      // V_u += D_ug S_g (the first sub-map sets V)
      auto combine = [&](int g, auto first) {
        float2 s[IV];
        load(g, s);
        const float* pc = sw.cpoly[g];
        int idx = 0;
#pragma unroll
        for (int u = 0; u <= LM; ++u) {
          float d = 0.f;
#pragma unroll
          for (int a = 0; a <= LM - u; ++a, ++idx) d = fmaf(pc[idx], mono[idx], d);
#pragma unroll
          for (int it = 0; it < IV; ++it) {
            if constexpr (decltype(first)::value) {
              V[u][it] = make_float2(d * s[it].x, d * s[it].y);
            } else {
              V[u][it].x = fmaf(d, s[it].x, V[u][it].x);
              V[u][it].y = fmaf(d, s[it].y, V[u][it].y);
            }
          }
        }
      };
      combine(0, std::true_type{});
      for (int g = 1; g < sw.n; ++g) combine(g, std::false_type{});
      const float rzb_neg = rzb - (float)geo.n2 * geo.dx2;
      auto rz_of = [&](int e) { return fmaf((float)(2 * e * P), geo.dx2, 2 * e >= E ? rzb_neg : rzb); };
#pragma unroll
      for (int u = LM; u >= 0; --u) {
        // opaque copies of the lane indices per transform: the twiddle reads derived from them are
        // repeated per transform instead of being hoisted and held in registers across all of them;
        // the masks restore their ranges, so index products stay full-rate 24-bit multiplies
        int tp_u = tp, r_u = r;
        asm volatile("" : "+v"(tp_u), "+v"(r_u));
        tp_u &= P - 1;
        r_u &= RW - 1;
        const RowTp<PADQ> rw_u{buf, LDT, r_u};
        wave_lds_sync();  // the previous transform's reads of the tile are issued
        prepass_regs<M, P, IV, PADQ>(buf, LDT, r_u, tp_u, V[u], stwN);
        wave_lds_sync();
        pencil_inverse_head<M, true>(tp_u, rw_u, stw);
        if (u == LM) {
          auto sink = [&](int e, int, float2 z) { acc[e][0] = z.x; acc[e][1] = z.y; };
          pencil_inverse_tail<M>(tp_u, rw_u, sink, stw);
        } else {
          auto sink = [&](int e, int, float2 z) {
            const float rz0 = rz_of(e);
            acc[e][0] = fmaf(acc[e][0], rz0, z.x);
            acc[e][1] = fmaf(acc[e][1], rz0 + geo.dx2, z.y);
          };
          pencil_inverse_tail<M>(tp_u, rw_u, sink, stw);
        }
      }
      // divide by |r|^LM (r > 0 on every row of this wave)
#pragma unroll
      for (int e = 0; e < E; ++e) {
        const float rz0 = rz_of(e), rz1 = rz0 + geo.dx2;
        float i0 = __builtin_amdgcn_rcpf(fmaf(rz0, rz0, rxy2)), i1 = __builtin_amdgcn_rcpf(fmaf(rz1, rz1, rxy2));
        if (LM == 4) { i0 *= i0; i1 *= i1; }
        acc[e][0] *= i0;
        acc[e][1] *= i1;
      }
    }
  }
  if (!combined) {
#pragma unroll
    for (int e = 0; e < E; ++e) { acc[e][0] = 0.f; acc[e][1] = 0.f; }
  }
  for (int g = 0; g < (combined ? 0 : sw.n); ++g) {
    const float2* src = work + ((long long)g * batch + t) * slot_in + row0 * pitch_in;
    // Opaque copies of the lane indices: index arithmetic derived from them is recomputed per
    // sub-map instead of being hoisted out of the loop and held in (spilled) registers; the masks
    // restore their ranges, so index products stay full-rate 24-bit multiplies.
    int lane_g = lane, tp_g = tp;
    float rzb_g = rzb;
    asm volatile("" : "+v"(lane_g), "+v"(tp_g), "+v"(rzb_g));
    lane_g &= 63;
    tp_g &= P - 1;
    if (g == sw.nfact && g > 0) rescale(rzb_g);
    wave_lds_sync();  // the previous sub-map's reads of the tile are issued
    load_prepass_rows<M, RW, 64, PADQ>(buf, LDT, src, pitch_in, rows_per_map - row0, lz_in, lane_g, stwN);
    wave_lds_sync();
    const int ell = sw.ell[g];
    const float cf = sw.coef[g];
    const float* pc = sw.poly[g];
    pencil_inverse_head<M, true>(tp_g, rw, stw);
    // Last inverse stage: adds w(x) * z straight into the accumulators.  w = coef Y(r_hat) =
    // D(rz) / r^ell with the z-polynomial coefficients of this row D[k] = coef sum_a poly rx^a
    // ry^(ell-k-a), in a form specialised to ell (uniform per sub-map).  Slot e of a lane holds z
    // position m = tp + e*P, and m >= M/2 (negative frequency) exactly when e >= E/2, so the
    // z coordinate of slot e is one FMA from the lane base rzb.
    const float rzb_neg = rzb_g - (float)geo.n2 * geo.dx2;
    auto rz_of = [&](int e) { return fmaf((float)(2 * e * P), geo.dx2, 2 * e >= E ? rzb_neg : rzb_g); };
    switch (ell) {
      case 0: {
        auto sink = [&](int e, int, float2 z) { acc[e][0] += z.x * cf; acc[e][1] += z.y * cf; };
        pencil_inverse_tail<M>(tp_g, rw, sink, stw);
        break;
      }
      case 2: {
        const float D0 = cf * (pc[0] * ry * ry + pc[1] * rx * ry + pc[2] * rx * rx);
        const float D1 = cf * (pc[3] * ry + pc[4] * rx);
        const float D2 = cf * pc[5];
        // numerator only (the ell = 2 sub-maps come first; rescale() divides by |r|^2)
        auto sink = [&](int e, int, float2 z) {
          const float rz0 = rz_of(e), rz1 = rz0 + geo.dx2;
          acc[e][0] = fmaf(z.x, fmaf(fmaf(D2, rz0, D1), rz0, D0), acc[e][0]);
          acc[e][1] = fmaf(z.y, fmaf(fmaf(D2, rz1, D1), rz1, D0), acc[e][1]);
        };
        pencil_inverse_tail<M>(tp_g, rw, sink, stw);
        break;
      }
      case 4: {
        const float rx2 = rx * rx, ry2 = ry * ry;
        const float D0 = cf * (pc[0] * ry2 * ry2 + pc[1] * rx * ry2 * ry + pc[2] * rx2 * ry2 + pc[3] * rx2 * rx * ry +
                               pc[4] * rx2 * rx2);
        const float D1 = cf * (pc[5] * ry2 * ry + pc[6] * rx * ry2 + pc[7] * rx2 * ry + pc[8] * rx2 * rx);
        const float D2 = cf * (pc[9] * ry2 + pc[10] * rx * ry + pc[11] * rx2);
        const float D3 = cf * (pc[12] * ry + pc[13] * rx);
        const float D4 = cf * pc[14];
        const float y0c = sw.y0[g] * cf;  // Y_4m(0) is not zero for every m
        auto sink = [&](int e, int, float2 z) {
          float w[2];
#pragma unroll
          for (int h = 0; h < 2; ++h) {
            const float rz = rz_of(e) + h * geo.dx2;
            const float r2 = fmaf(rz, rz, rxy2);
            const float pz = fmaf(fmaf(fmaf(fmaf(D4, rz, D3), rz, D2), rz, D1), rz, D0);
            const float ri2 = __builtin_amdgcn_rcpf(r2);
            w[h] = r2 > 0.f ? pz * ri2 * ri2 : y0c;
          }
          acc[e][0] = fmaf(z.x, w[0], acc[e][0]);
          acc[e][1] = fmaf(z.y, w[1], acc[e][1]);
        };
        pencil_inverse_tail<M>(tp_g, rw, sink, stw);
        break;
      }
      default: {  // odd ell (1, 3): Y(0) = 0, and pz = 0 exactly at r = 0
        float D[4] = {0.f, 0.f, 0.f, 0.f};
        int idx = 0;
        for (int k = 0; k <= ell; ++k) {
          float sacc = 0.f;
          for (int a = 0; a <= ell - k; ++a) {
            float mono = 1.f;
            for (int q = 0; q < a; ++q) mono *= rx;
            for (int q = 0; q < ell - k - a; ++q) mono *= ry;
            sacc += pc[idx++] * mono;
          }
          sacc *= cf;
          if (k == 0) D[0] = sacc; else if (k == 1) D[1] = sacc; else if (k == 2) D[2] = sacc; else D[3] = sacc;
        }
        auto sink = [&](int e, int, float2 z) {
          float w[2];
#pragma unroll
          for (int h = 0; h < 2; ++h) {
            const float rz = rz_of(e) + h * geo.dx2;
            const float ri = rsqrtf(fmaxf(fmaf(rz, rz, rxy2), 1e-30f));
            const float pz = fmaf(fmaf(fmaf(D[3], rz, D[2]), rz, D[1]), rz, D[0]);
            w[h] = pz * (ell == 1 ? ri : ri * ri * ri);
          }
          acc[e][0] = fmaf(z.x, w[0], acc[e][0]);
          acc[e][1] = fmaf(z.y, w[1], acc[e][1]);
        };
        pencil_inverse_tail<M>(tp_g, rw, sink, stw);
        break;
      }
    }
  }
  if (!combined && sw.nfact == sw.n) rescale(rzb);
  wave_lds_sync();  // all reads of the tile by the last inverse stage are issued
  // mask and scale in registers, then the forward row transform (first stage from registers)
  const float norm = scale / (float)geo.N;
  if constexpr (IC) {
    // per-bin sums of m * weight over this wave's voxels, in a fixed order: for each voxel slot and
    // each bin present in it, a butterfly sum over the wave, added by lane 0 to the wave's FP32 LDS
    // sums; then the wave's FP64 partials go to their own slot (no atomics)
    if (lane < ic.nbins) icl[wave][lane] = 0.f;
    wave_lds_sync();
    const bool rv = row < rows_per_map;
    auto add_keyed = [&](int z, float c) {
      unsigned long long act = __ballot(z < ic.nbins);
      while (act) {
        const int k = __shfl(z, __ffsll((long long)act) - 1);
        float v = z == k ? c : 0.f;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
        if (lane == 0) icl[wave][k] += v;
        act &= ~__ballot(z == k);
      }
    };
    // one voxel pair at a time (this path runs only for integral constraints; keeping the weights
    // of all slots live next to the accumulators would spill)
#pragma unroll
    for (int e = 0; e < E; ++e) {
      float c0 = 0.f, c1 = 0.f;
      int z0 = kMaxIcBins, z1 = kMaxIcBins;
      if (rv) {
        const long long v0 = row * (2 * M) + 2 * (tp + e * P);
        const float2 wv = *reinterpret_cast<const float2*>(ic.weight + v0);
        z0 = ic.bin ? ic.bin[v0] : 0;
        z1 = ic.bin ? ic.bin[v0 + 1] : 0;
        c0 = acc[e][0] * norm * wv.x;
        c1 = acc[e][1] * norm * wv.y;
      }
      add_keyed(z0, c0);
      add_keyed(z1, c1);
    }
    wave_lds_sync();
    if (lane < ic.nbins) ic.partial[(t * ic.nwm + row0 / RW) * ic.nbins + lane] = (double)icl[wave][lane];
  }
  if (mask && row < rows_per_map) {
    float2 mk[E];  // all mask loads in flight together
#pragma unroll
    for (int e = 0; e < E; ++e)
      mk[e] = *reinterpret_cast<const float2*>(mask + row * (2 * M) + 2 * (tp + e * P));
#pragma unroll
    for (int e = 0; e < E; ++e) { acc[e][0] *= norm * mk[e].x; acc[e][1] *= norm * mk[e].y; }
  } else {
#pragma unroll
    for (int e = 0; e < E; ++e) { acc[e][0] *= norm; acc[e][1] *= norm; }
  }
  auto get = [&](int e) { return make_float2(acc[e][0], acc[e][1]); };
  pencil_forward_from_regs<M, true>(tp, get, rw, stw);
  wave_lds_sync();
  // kz < lz_out of this lane's row: positions tp, tp + P, ... (each store instruction writes P
  // consecutive entries of every row of the wave)
  if (row < rows_per_map) {
    float2* o = out + t * slot_out + row * pitch_out;
    for (int k = tp; k < lz_out; k += P) o[k] = r2c_split<M, PADQ>(buf, LDT, r, k, stwN);
  }
}

// sums[t * nbins + z] += sum over the waves of map t of their partials, in wave order
__global__ void ic_finalize(const double* __restrict__ partial, long long nwm, int batch, int nbins,
                            double* __restrict__ sums) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long long)batch * nbins) return;
  const long long t = i / nbins;
  const int z = (int)(i % nbins);
  double s = 0.;
  for (long long w = 0; w < nwm; ++w) s += partial[(t * nwm + w) * nbins + z];
  sums[i] += s;
}

// ------------------------------------------------------------------------------------------
// host side
// ------------------------------------------------------------------------------------------
static std::atomic<long long> g_fwd{0}, g_rev{0};

// ---- sampled launch timing (roofline reporting): every `period`-th launch of a pass kind is
// bracketed by hipEvents on its own stream; read() synchronises and sums the durations.
// 0 r2c rows, 1 c2r rows, 2 plain C2R cols, 3 C2R first cols, 4 R2C last cols, 5 pointing rows,
// 6 plain R2C cols: one kernel template per kind
constexpr int kProfKinds = 7;
struct ProfSample { int kind; hipEvent_t a, b; double bytes; };
static int g_prof_period = 0;
static long long g_prof_count[kProfKinds] = {};
static std::vector<ProfSample> g_prof_pending;
static std::vector<hipEvent_t> g_prof_free;
static long long g_prof_n[kProfKinds] = {};
static double g_prof_ms[kProfKinds] = {};
static double g_prof_bytes[kProfKinds] = {};

static hipEvent_t prof_event() {
  if (!g_prof_free.empty()) { hipEvent_t e = g_prof_free.back(); g_prof_free.pop_back(); return e; }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

// folds the completed samples at the front of the pending list into the totals (non-blocking)
static void prof_drain() {
  size_t done = 0;
  for (; done < g_prof_pending.size(); ++done) {
    ProfSample& smp = g_prof_pending[done];
    if (hipEventQuery(smp.b) != hipSuccess) break;
    float t = 0.f;
    if (hipEventElapsedTime(&t, smp.a, smp.b) == hipSuccess) {
      g_prof_n[smp.kind] += 1; g_prof_ms[smp.kind] += t; g_prof_bytes[smp.kind] += smp.bytes;
    }
    g_prof_free.push_back(smp.a); g_prof_free.push_back(smp.b);
  }
  if (done) g_prof_pending.erase(g_prof_pending.begin(), g_prof_pending.begin() + done);
}

// returns true and records the start event when this launch is sampled (never while the stream is
// being captured into a graph: the events would only be recorded by the replays)
static bool prof_begin(int kind, hipStream_t s, ProfSample& smp, double bytes) {
  if (g_prof_period <= 0) return false;
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(s, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) return false;
  if ((g_prof_count[kind]++ % g_prof_period) != 0) return false;
  if (g_prof_pending.size() >= 512) prof_drain();           // keeps the whole timed region sampled
  if (g_prof_pending.size() > (1u << 16)) return false;     // the host runs that far ahead: skip
  smp.kind = kind; smp.bytes = bytes; smp.a = prof_event(); smp.b = prof_event();
  if (!smp.a || !smp.b) return false;
  (void)hipEventRecord(smp.a, s);
  return true;
}
static void prof_end(bool on, ProfSample& smp, hipStream_t s) {
  if (!on) return;
  (void)hipEventRecord(smp.b, s);
  g_prof_pending.push_back(smp);
}

static bool pow2_in(int n, int lo, int hi) { return n >= lo && n <= hi && (n & (n - 1)) == 0; }
// power-of-two fast path: n0, n1 in [2, 512], n2 in [4, 512]; other sizes take the mixed-radix passes
static bool fast_grid(int n0, int n1, int n2) { return pow2_in(n0, 2, 512) && pow2_in(n1, 2, 512) && pow2_in(n2, 4, 512); }
static bool generic_grid(int n0, int n1, int n2) {
  return n0 >= 2 && n0 <= 1024 && n1 >= 2 && n1 <= 1024 && n2 >= 4 && n2 <= 2048 && n2 % 2 == 0;
}

}  // namespace pb3d

struct pb3d_plan {
  pb3d_grid grid;
  pb3d::Geo geo;
  int L[3];         // n0, n1, n2/2
  float2* tw[3];    // device twiddles exp(-2 pi i m / L), m < L
  float2* twN;      // exp(-2 pi i k / n2), k = 0..n2/2
  int generic;      // 1: mixed-radix passes (sizes outside the power-of-two fast path)
  pb3d::GenAxis ax[3];  // radices of n0, n1, n2/2 (generic plans)
};

namespace pb3d {

template <int DIR, int MODE>
static hipError_t launch_cols(int L, const float2* in, float2* out, const ColMap& cm, long long in_bs, long long out_bs,
                              int batch, const float2* tw, const Geo& geo, const KOp& kop, hipStream_t s) {
  if (cm.ncols <= 0) return hipSuccess;
  // Plane-aligned tiles for the y pass when a plane's columns leave at least half a tile over:
  // a straddling tile would read a few columns of the next plane at a stride.
#define PB3D_COLS(LL)                                                                          \
  case LL: {                                                                                   \
    constexpr int T = ColTile<LL, MODE>::T;                                                    \
    ColMap c2 = cm;                                                                            \
    c2.tpp = 0;                                                                                \
    if (PB3D_COL_ALIGN && cm.tpp && cm.cpp % T >= T / 2) c2.tpp = (cm.cpp + T - 1) / T;       \
    const long long nt = c2.tpp ? (long long)(cm.ncols / cm.cpp) * c2.tpp : (cm.ncols + T - 1) / T; \
    dim3 grid((unsigned)nt, batch);                                                            \
    fft_cols<LL, DIR, MODE><<<grid, Decomp<LL>::P * T, 0, s>>>(in, out, c2, in_bs, out_bs, tw, geo, kop); \
    return hipGetLastError();                                                                  \
  }
  switch (L) {
    PB3D_COLS(2) PB3D_COLS(4) PB3D_COLS(8) PB3D_COLS(16) PB3D_COLS(32) PB3D_COLS(64)
    PB3D_COLS(128) PB3D_COLS(256) PB3D_COLS(512)
    default: return hipErrorInvalidValue;
  }
#undef PB3D_COLS
}

static hipError_t launch_r2c_rows(int M, const float* in, float2* out, long long nrows, int rpb,
                                  const float2* tw, const float2* twN, const Geo& geo, const ROp& rop, int lz_out,
                                  int pitch, hipStream_t s) {
#define PB3D_ROWS(MM)                                                                            \
  case MM: {                                                                                     \
    constexpr int T = RowCfg<MM>::T;                                                             \
    r2c_rows<MM><<<(unsigned)((nrows + T - 1) / T), RowTile<MM>::NT, 0, s>>>(in, out, nrows, rpb, tw, \
                                                                              twN, geo, rop, lz_out, pitch); \
    return hipGetLastError();                                                                    \
  }
  switch (M) {
    PB3D_ROWS(2) PB3D_ROWS(4) PB3D_ROWS(8) PB3D_ROWS(16) PB3D_ROWS(32) PB3D_ROWS(64)
    PB3D_ROWS(128) PB3D_ROWS(256)
    default: return hipErrorInvalidValue;
  }
#undef PB3D_ROWS
}

static hipError_t launch_c2r_rows(int M, const float2* in, float* out, long long nrows, int rpb,
                                  const float2* tw, const float2* twN, const Geo& geo, const ROp& rop,
                                  float norm, int lz_in, int pitch, hipStream_t s) {
#define PB3D_ROWS(MM)                                                                            \
  case MM: {                                                                                     \
    constexpr int T = RowCfg<MM>::T;                                                             \
    c2r_rows<MM><<<(unsigned)((nrows + T - 1) / T), RowTile<MM>::NT, 0, s>>>(in, out, nrows, rpb, tw, \
                                                                              twN, geo, rop, norm, lz_in, pitch); \
    return hipGetLastError();                                                                    \
  }
  switch (M) {
    PB3D_ROWS(2) PB3D_ROWS(4) PB3D_ROWS(8) PB3D_ROWS(16) PB3D_ROWS(32) PB3D_ROWS(64)
    PB3D_ROWS(128) PB3D_ROWS(256)
    default: return hipErrorInvalidValue;
  }
#undef PB3D_ROWS
}

template <int MM, int NU>
static hipError_t launch_prows(const float2* work, long long slot_in, int pitch_in, float2* out, long long slot_out,
                               int pitch_out, long long rpm, int batch, const SubW& sw, const IcSum& ic,
                               const float* mask, float scale, int lz_in, int lz_out, const float2* tw,
                               const float2* twN, const Geo& geo, hipStream_t s) {
  constexpr int RB = PointCfg<MM>::RB;
  const long long nblk = (rpm + RB - 1) / RB * batch;  // one linear grid, see xcd_rows_block
  if (nblk > 0x7fffffffLL) return hipErrorInvalidValue;
  dim3 grid((unsigned)nblk);
  if (ic.weight) {
    IcSum icw = ic;
    icw.nwm = (rpm + PointCfg<MM>::RW - 1) / PointCfg<MM>::RW;
    pointing_rows<MM, true, NU><<<grid, PointCfg<MM>::NT, 0, s>>>(work, slot_in, pitch_in, out, slot_out, pitch_out,
                                                                  rpm, batch, sw, icw, mask, scale, lz_in, lz_out, tw,
                                                                  twN, geo);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const long long nf = (long long)batch * ic.nbins;
    ic_finalize<<<(unsigned)((nf + 255) / 256), 256, 0, s>>>(icw.partial, icw.nwm, batch, ic.nbins, ic.sums);
  } else {
    pointing_rows<MM, false, NU><<<grid, PointCfg<MM>::NT, 0, s>>>(work, slot_in, pitch_in, out, slot_out, pitch_out,
                                                                   rpm, batch, sw, ic, mask, scale, lz_in, lz_out, tw,
                                                                   twN, geo);
  }
  return hipGetLastError();
}

// row lengths with a combined-row instance (sw.nu > 0 is set only for these)
constexpr bool combined_rows_len(int M) { return M == 64 || M == 128 || M == 256; }

template <int MM>
static hipError_t launch_prows_m(const float2* work, long long slot_in, int pitch_in, float2* out, long long slot_out,
                                 int pitch_out, long long rpm, int batch, const SubW& sw, const IcSum& ic,
                                 const float* mask, float scale, int lz_in, int lz_out, const float2* tw,
                                 const float2* twN, const Geo& geo, hipStream_t s) {
  if constexpr (combined_rows_len(MM)) {
    if (sw.nu == 3)
      return launch_prows<MM, 3>(work, slot_in, pitch_in, out, slot_out, pitch_out, rpm, batch, sw, ic, mask, scale,
                                 lz_in, lz_out, tw, twN, geo, s);
    if (sw.nu == 5)
      return launch_prows<MM, 5>(work, slot_in, pitch_in, out, slot_out, pitch_out, rpm, batch, sw, ic, mask, scale,
                                 lz_in, lz_out, tw, twN, geo, s);
  }
  return launch_prows<MM, 0>(work, slot_in, pitch_in, out, slot_out, pitch_out, rpm, batch, sw, ic, mask, scale,
                             lz_in, lz_out, tw, twN, geo, s);
}

static hipError_t launch_pointing_rows(int M, const float2* work, long long slot_in, int pitch_in, float2* out,
                                       long long slot_out, int pitch_out, long long rpm, int batch, const SubW& sw,
                                       const IcSum& ic,
                                       const float* mask, float scale, int lz_in, int lz_out, const float2* tw,
                                       const float2* twN, const Geo& geo, hipStream_t s) {
#define PB3D_PROWS(MM)                                                                                     \
  case MM:                                                                                                 \
    return launch_prows_m<MM>(work, slot_in, pitch_in, out, slot_out, pitch_out, rpm, batch, sw, ic, mask, \
                              scale, lz_in, lz_out, tw, twN, geo, s);
  switch (M) {
    PB3D_PROWS(2) PB3D_PROWS(4) PB3D_PROWS(8) PB3D_PROWS(16) PB3D_PROWS(32) PB3D_PROWS(64)
    PB3D_PROWS(128) PB3D_PROWS(256)
    default: return hipErrorInvalidValue;
  }
#undef PB3D_PROWS
}

// c * p(x, y, z) * (x^2 + y^2 + z^2)^((L - ell) / 2) for p homogeneous of degree ell given in the
// (k, a) index order (x^a y^(ell-k-a) z^k; p = 1 for ell = 0), written in the same order for degree L.
static void raise_poly(int ell, const float* p, int L, float c, float* out) {
  double A[5][5][5] = {};  // [x power][y power][z power]
  int idx = 0;
  for (int k = 0; k <= ell; ++k)
    for (int a = 0; a <= ell - k; ++a, ++idx) A[a][ell - k - a][k] = ell > 0 ? (double)p[idx] : 1.0;
  for (int d = ell; d < L; d += 2) {
    double B[5][5][5] = {};
    for (int a = 0; a <= d; ++a)
      for (int b = 0; a + b <= d; ++b) {
        const int k = d - a - b;
        const double v = A[a][b][k];
        B[a + 2][b][k] += v; B[a][b + 2][k] += v; B[a][b][k + 2] += v;
      }
    for (int a = 0; a < 5; ++a)
      for (int b = 0; b < 5; ++b)
        for (int k = 0; k < 5; ++k) A[a][b][k] = B[a][b][k];
  }
  idx = 0;
  for (int k = 0; k <= L; ++k)
    for (int a = 0; a <= L - k; ++a, ++idx) out[idx] = (float)(c * A[a][L - k - a][k]);
}

// Homogeneous-polynomial form of Y_{ell,m} (degree ell <= 4) on the unit sphere, fitted by least
// squares on fixed sample directions in FP64 (exact up to rounding: the homogeneous polynomials of
// degree ell restricted to the sphere are linearly independent).  Coefficient order: z power k
// outer, x power a inner, y power ell-k-a.
static bool ylm_poly(int ell, int m, float* coef) {
  int K = 0;
  int pa[kMaxMono], pb[kMaxMono], pk[kMaxMono];
  for (int k = 0; k <= ell; ++k)
    for (int a = 0; a <= ell - k; ++a) { pa[K] = a; pb[K] = ell - k - a; pk[K] = k; ++K; }
  double A[kMaxMono][kMaxMono] = {}, rhs[kMaxMono] = {};
  const int ns = 97;
  for (int s = 0; s < ns; ++s) {
    // deterministic spread of directions (spherical Fibonacci lattice)
    const double zz = 1.0 - (2.0 * s + 1.0) / ns, rr = std::sqrt(1.0 - zz * zz);
    const double ph = 2.399963229728653 * s;
    const double x = rr * std::cos(ph), y = rr * std::sin(ph), z = zz;
    double mono[kMaxMono];
    for (int i = 0; i < K; ++i) mono[i] = std::pow(x, pa[i]) * std::pow(y, pb[i]) * std::pow(z, pk[i]);
    const double yv = (double)ylm_f(ell, m, (float)x, (float)y, (float)z);
    for (int i = 0; i < K; ++i) {
      rhs[i] += mono[i] * yv;
      for (int j = 0; j < K; ++j) A[i][j] += mono[i] * mono[j];
    }
  }
  // Gaussian elimination with partial pivoting
  for (int c = 0; c < K; ++c) {
    int piv = c;
    for (int r = c + 1; r < K; ++r) if (std::fabs(A[r][c]) > std::fabs(A[piv][c])) piv = r;
    if (std::fabs(A[piv][c]) < 1e-12) return false;
    for (int j = 0; j < K; ++j) std::swap(A[c][j], A[piv][j]);
    std::swap(rhs[c], rhs[piv]);
    for (int r = 0; r < K; ++r) {
      if (r == c) continue;
      const double f = A[r][c] / A[c][c];
      for (int j = 0; j < K; ++j) A[r][j] -= f * A[c][j];
      rhs[r] -= f * rhs[c];
    }
  }
  for (int i = 0; i < K; ++i) coef[i] = (float)(rhs[i] / A[i][i]);
  return true;
}

// Band-limit bookkeeping: radii r0, r1 on the two complex axes (-1 = whole axis) and the number of
// kz planes lz in the support.
struct Prune { int r0, r1, lz; };
static Prune make_prune(const Geo& g, const pb3d_kop* k) {
  Prune pr{-1, -1, g.n2h};
  if (!k) return pr;
  if (k->support[0] >= 0 && 2 * k->support[0] < g.n0) pr.r0 = k->support[0];
  if (k->support[1] >= 0 && 2 * k->support[1] < g.n1) pr.r1 = k->support[1];
  if (k->support[2] >= 0 && k->support[2] + 1 < g.n2h) pr.lz = k->support[2] + 1;
  return pr;
}
static Prune merge_prune(const Geo& g, Prune a, Prune b) {
  Prune o;
  o.r0 = (a.r0 < 0 || b.r0 < 0) ? -1 : (a.r0 > b.r0 ? a.r0 : b.r0);
  o.r1 = (a.r1 < 0 || b.r1 < 0) ? -1 : (a.r1 > b.r1 ? a.r1 : b.r1);
  o.lz = a.lz > b.lz ? a.lz : b.lz;
  (void)g;
  return o;
}
// Row pitch of a compact work map holding the z planes kz < lz (unpadded: the column passes read
// and write whole compact planes, measured faster than padded pitches).
static int compact_pitch(const Geo& g, int lz) { return lz < g.n2h ? lz : g.n2h; }
// Band limits of a pointing chain: inverse side (all sub-maps) and forward side (kop_out).
static void pointing_prunes(const Geo& g, int nsub, const pb3d_submap* subs, const pb3d_kop* kop_out, Prune& pin,
                            Prune& pout) {
  pin = make_prune(g, &subs[0].kop);
  for (int k = 1; k < nsub; ++k) pin = merge_prune(g, pin, make_prune(g, &subs[k].kop));
  pout = make_prune(g, kop_out);
}
// Column-pass array addressing.  pitch 0: the half-complex layout; pitch > 0: the compact layout
// of a band-limited map, z planes kz < pitch only (row (i, j) at (i * n1 + j) * pitch).
static void x_side(const Geo& g, const ColMap& cm, int pitch, long long& plane, long long& S) {
  if (pitch <= 0 || cm.plane == 0) { plane = cm.plane; S = cm.S; return; }  // whole planes: pitch == n2h
  plane = pitch; S = (long long)g.n1 * pitch;
}
static void y_side(const Geo& g, const ColMap& cm, int pitch, long long& plane, long long& S) {
  if (pitch <= 0) { plane = cm.plane; S = cm.S; return; }
  plane = (long long)g.n1 * pitch; S = pitch;
}
// x-column pass (axis 0) over the active (j, l) columns; pitch_in / pitch_out select the layout of
// the input / output array (0 = half-complex)
static ColMap xcols(const Geo& g, const Prune& pr, bool c2r, int pitch_in = 0, int pitch_out = 0) {
  ColMap cm{};
  cm.S = (long long)g.n1 * g.n2h;
  if (pr.r1 < 0 && pr.lz == g.n2h) {
    cm.ncols = g.n1 * g.n2h; cm.cpp = g.n1 * g.n2h; cm.plane = 0; cm.split = 1 << 30; cm.gap = 0;
  } else {
    const int J = pr.r1 < 0 ? g.n1 : 2 * pr.r1 + 1;
    cm.ncols = J * pr.lz; cm.cpp = pr.lz; cm.plane = g.n2h;
    cm.split = pr.r1 < 0 ? (1 << 30) : pr.r1 + 1;
    cm.gap = pr.r1 < 0 ? 0 : g.n1 - J;
  }
  cm.rad_in = c2r ? pr.r0 : -1;
  cm.rad_out = c2r ? -1 : pr.r0;
  x_side(g, cm, pitch_in, cm.in_plane, cm.in_S);
  x_side(g, cm, pitch_out, cm.out_plane, cm.out_S);
  return cm;
}
// y-column pass (axis 1) over the active (i, l) columns
static ColMap ycols(const Geo& g, const Prune& pr, bool c2r, int pitch_in = 0, int pitch_out = 0) {
  ColMap cm{};
  cm.S = g.n2h;
  cm.ncols = g.n0 * pr.lz; cm.cpp = pr.lz; cm.plane = (long long)g.n1 * g.n2h; cm.split = 1 << 30; cm.gap = 0;
  cm.tpp = 1;  // plane-aligned tiles allowed (the launcher decides)
  cm.rad_in = c2r ? pr.r1 : -1;
  cm.rad_out = c2r ? -1 : pr.r1;
  y_side(g, cm, pitch_in, cm.in_plane, cm.in_S);
  y_side(g, cm, pitch_out, cm.out_plane, cm.out_S);
  return cm;
}

// Algorithmic HBM bytes of one launch (what the pass must read and write, after pruning): 8 B per
// complex element read or written; k-space prologue/epilogue tables (bin tables, packed-row
// indices) are not counted.
static int npos(int rad, int L) { return rad < 0 ? L : (2 * rad + 1 < L ? 2 * rad + 1 : L); }
static double col_bytes(const ColMap& cm, int L, int batch) {
  return 8.0 * batch * cm.ncols * (double)(npos(cm.rad_in, L) + npos(cm.rad_out, L));
}
// Mixed-radix column maps over the (j, l < lz) columns of the x pass and the (i, l < lz) columns of
// the y pass; pitch 0 < pitch <= n2h selects the row pitch of the input / output array.
static GenCols gen_xcols(const Geo& g, int lz, int pitch_in, int pitch_out, int rad_in, int rad_out, int qrad_in,
                         int qrad_out) {
  GenCols cm{};
  cm.ncols = (long long)g.n1 * lz; cm.cpp = lz;
  cm.plane = g.n2h; cm.S = (long long)g.n1 * g.n2h;
  cm.in_plane = pitch_in; cm.in_S = (long long)g.n1 * pitch_in;
  cm.out_plane = pitch_out; cm.out_S = (long long)g.n1 * pitch_out;
  cm.rad_in = rad_in; cm.rad_out = rad_out; cm.qlen = g.n1; cm.qrad_in = qrad_in; cm.qrad_out = qrad_out;
  return cm;
}
static GenCols gen_ycols(const Geo& g, int lz, int pitch_in, int pitch_out, int rad_in) {
  GenCols cm{};
  cm.ncols = (long long)g.n0 * lz; cm.cpp = lz;
  cm.plane = (long long)g.n1 * g.n2h; cm.S = g.n2h;
  cm.in_plane = (long long)g.n1 * pitch_in; cm.in_S = pitch_in;
  cm.out_plane = (long long)g.n1 * pitch_out; cm.out_S = pitch_out;
  cm.rad_in = rad_in; cm.rad_out = -1; cm.qlen = g.n0; cm.qrad_in = -1; cm.qrad_out = -1;
  return cm;
}
static double gen_col_bytes(const GenCols& cm, int L, int batch) { return 16.0 * batch * (double)cm.ncols * L; }

static int rop_operands(const pb3d_rop* r) {
  if (!r) return 1;
  return 1 + (r->b ? 1 : 0) + (r->c ? 1 : 0) + (r->d ? 1 : 0);
}

static float2* upload_twiddles(int L, int count) {
  std::vector<float2> h(count);
  const double twopi = 6.283185307179586476925286766559;
  for (int m = 0; m < count; ++m) {
    double a = -twopi * (double)m / (double)L;
    h[m] = make_float2((float)std::cos(a), (float)std::sin(a));
  }
  float2* d = nullptr;
  if (hipMalloc(&d, sizeof(float2) * count) != hipSuccess) return nullptr;
  if (hipMemcpy(d, h.data(), sizeof(float2) * count, hipMemcpyHostToDevice) != hipSuccess) {
    (void)hipFree(d);
    return nullptr;
  }
  return d;
}

int check_rop(const pb3d_rop* r, bool input_side) {
  if (!r) return 0;
  if (r->kind < 0 || r->kind > 5) return fail(PB3D_ERR_ARG, "rop.kind out of range");
  if (input_side && r->kind == 5) return fail(PB3D_ERR_ARG, "rop SQUARE is an output-side op");
  if ((r->kind == 1 || r->kind == 2 || r->kind == 4) && !r->b) return fail(PB3D_ERR_ARG, "rop.b required");
  if (r->kind == 2 && (!r->c || !r->d)) return fail(PB3D_ERR_ARG, "rop DIFF needs c and d");
  if ((r->kind == 3 || r->kind == 4) && (r->ell < 0 || r->ell > 4 || r->m < 0 || r->m > 2 * r->ell))
    return fail(PB3D_ERR_ARG, "rop Y_lm index out of range (l<=4)");
  return 0;
}

int check_kop(const pb3d_kop* k) {
  if (!k) return 0;
  if (k->x_batch_group < 0) return fail(PB3D_ERR_ARG, "kop.x_batch_group must be >= 0");
  if (k->kind < 0 || k->kind > 6) return fail(PB3D_ERR_ARG, "kop.kind out of range");
  if (k->kind == 6 && (!k->bin_table || !k->ft || k->nk < 1 || k->nk > 4 || k->terms < 1 || k->terms > 254))
    return fail(PB3D_ERR_ARG, "kop BINPOLY needs bin_table, coef (ft), 1<=nk<=4 and 1<=terms<=254");
  if (k->kind == 4 && (!k->bin_table || !k->pack_index || !k->trip || !k->pair_index || !k->ft ||
                       k->packed_len < 1 || k->nk < 1 || k->nk > 254 || k->trip0 < 0 ||
                       ((k->terms & 4) && (k->ell < 0 || k->ell > 4 || k->m < 0 || k->m > 2 * k->ell))))
    return fail(PB3D_ERR_ARG, "kop BKQ: missing tables or bad fields");
  if (k->kind == 5 && (!k->pack_index || k->packed_len < 1))
    return fail(PB3D_ERR_ARG, "kop PACK needs pack_index and packed_len");
  if (k->kind == 1 && (!k->bin_table || k->ell < 0 || k->ell > 6 || k->ell2 < 0 || k->ell2 > 6))
    return fail(PB3D_ERR_ARG, "kop FILTER needs bin_table and 0<=ell,ell2<=6");
  if (k->kind == 2 && (k->ell < 0 || k->ell > 4 || k->m < 0 || k->m > 2 * k->ell))
    return fail(PB3D_ERR_ARG, "kop Y_lm index out of range (l<=4)");
  if (k->kind == 3 && !k->w) return fail(PB3D_ERR_ARG, "kop WEIGHT needs w");
  return 0;
}

}  // namespace pb3d

using namespace pb3d;

extern "C" int64_t pb3d_fft_counter(int dir) {
  if (dir < 0) { g_fwd = 0; g_rev = 0; return 0; }
  return dir == 0 ? (int64_t)g_fwd.load() : (int64_t)g_rev.load();
}

extern "C" int pb3d_prof_enable(int period) {
  if (period < 0) return fail(PB3D_ERR_ARG, "prof_enable: period must be >= 0");
  g_prof_period = period;
  for (int k = 0; k < kProfKinds; ++k) { g_prof_count[k] = 0; g_prof_n[k] = 0; g_prof_ms[k] = 0; g_prof_bytes[k] = 0; }
  return 0;
}

extern "C" int pb3d_prof_read(int kind, int64_t* sampled, double* ms, double* bytes) {
  if (kind < 0 || kind >= kProfKinds || !sampled || !ms || !bytes) return fail(PB3D_ERR_ARG, "prof_read: bad args");
  for (auto& smp : g_prof_pending) {
    if (hipEventSynchronize(smp.b) != hipSuccess) return fail(PB3D_ERR_HIP, "prof_read: event sync failed");
    float t = 0.f;
    if (hipEventElapsedTime(&t, smp.a, smp.b) == hipSuccess) {
      g_prof_n[smp.kind] += 1; g_prof_ms[smp.kind] += t; g_prof_bytes[smp.kind] += smp.bytes;
    }
    g_prof_free.push_back(smp.a); g_prof_free.push_back(smp.b);
  }
  g_prof_pending.clear();
  *sampled = g_prof_n[kind]; *ms = g_prof_ms[kind]; *bytes = g_prof_bytes[kind];
  return 0;
}

extern "C" int pb3d_plan_create(pb3d_plan** plan, const pb3d_grid* grid) {
  if (!plan || !grid) return fail(PB3D_ERR_ARG, "null plan/grid pointer");
  *plan = nullptr;
  const int n0 = grid->n[0], n1 = grid->n[1], n2 = grid->n[2];
  const bool fast = fast_grid(n0, n1, n2);
  if (!fast && !generic_grid(n0, n1, n2))
    return fail(PB3D_ERR_UNSUPPORTED, "grid sizes must be n0, n1 in [2, 1024] and even n2 in [4, 2048]");
  for (int i = 0; i < 3; ++i)
    if (!(grid->boxsize[i] > 0)) return fail(PB3D_ERR_ARG, "boxsize must be positive");
  pb3d_plan* p = new pb3d_plan();
  p->grid = *grid;
  p->geo = make_geo(*grid);
  p->L[0] = n0; p->L[1] = n1; p->L[2] = n2 / 2;
  p->generic = fast ? 0 : 1;
  for (int a = 0; a < 3; ++a)
    if (!gen_factor(p->L[a], p->ax[a])) { pb3d_plan_destroy(p); return fail(PB3D_ERR_UNSUPPORTED, "axis length not factorable"); }
  for (int a = 0; a < 3; ++a) {
    p->tw[a] = upload_twiddles(p->L[a], p->L[a]);
    if (!p->tw[a]) { pb3d_plan_destroy(p); return fail(PB3D_ERR_ALLOC, "twiddle upload failed"); }
  }
  p->twN = upload_twiddles(n2, n2 / 2 + 1);
  if (!p->twN) { pb3d_plan_destroy(p); return fail(PB3D_ERR_ALLOC, "twiddle upload failed"); }
  *plan = p;
  return 0;
}

extern "C" int pb3d_plan_destroy(pb3d_plan* p) {
  if (!p) return 0;
  for (int a = 0; a < 3; ++a) if (p->tw[a]) (void)hipFree(p->tw[a]);
  if (p->twN) (void)hipFree(p->twN);
  delete p;
  return 0;
}

extern "C" int pb3d_fft_r2c(const pb3d_plan* p, const float* in, void* out_, void* work_, int batch,
                            const pb3d_rop* rop_, const pb3d_kop* kop_, void* stream) {
  if (!p || !out_ || batch < 1 || batch > 65535) return fail(PB3D_ERR_ARG, "r2c: null plan/output or batch out of [1, 65535]");
  if (!in && !(rop_ && rop_->a)) return fail(PB3D_ERR_ARG, "r2c: no input");
  if (int e = check_rop(rop_, true)) return e;
  if (int e = check_kop(kop_)) return e;
  if (kop_ && kop_->accumulate && !work_) return fail(PB3D_ERR_ARG, "r2c: accumulating output needs work");
  if (kop_ && kop_->kind == 5 && !work_) return fail(PB3D_ERR_ARG, "r2c: packed output needs work");
  if (kop_ && kop_->kind == 4) return fail(PB3D_ERR_ARG, "r2c: BKQ is a C2R prologue");
  hipStream_t s = (hipStream_t)stream;
  const Geo& g = p->geo;
  float2* out = (float2*)out_;
  float2* work = work_ ? (float2*)work_ : out;
  ROp rop = to_rop(rop_);
  KOp kop = to_kop(kop_);
  KOp plain = to_kop(nullptr);
  const Prune pr = make_prune(g, kop_);
  const long long rows = (long long)batch * g.n0 * g.n1;
  // a separate work array holds the band-limited intermediate compactly (z planes kz < lz only);
  // without one the output array is the work array and keeps the half-complex layout
  const int pitch = work == out ? 0 : compact_pitch(g, pr.lz);
  const long long wbs = pitch ? (long long)g.n0 * g.n1 * pitch : g.H;
  const ColMap cy = ycols(g, pr, false, pitch, pitch), cx = xcols(g, pr, false, pitch, 0);
  const double br = (double)batch * (g.N * 4.0 * rop_operands(rop_) + (double)g.n0 * g.n1 * pr.lz * 8.0);
  ProfSample sm{};
  const long long out_bs = (kop_ && kop_->kind == 5) ? kop_->packed_len : g.H;
  if (p->generic) {
    const int P = pitch ? pitch : g.n2h;
    bool on = prof_begin(0, s, sm, br);
    PB3D_CHECK_HIP(gen_r2c_rows(p->ax[2], in, work, rows, g.n0 * g.n1, p->tw[2], p->twN, g, rop, pr.lz, P, s));
    prof_end(on, sm, s);
    const GenCols gy = gen_ycols(g, pr.lz, P, P, -1);
    on = prof_begin(6, s, sm, gen_col_bytes(gy, g.n1, batch));
    PB3D_CHECK_HIP(gen_cols(-1, 0, p->ax[1], work, work, gy, wbs, wbs, batch, p->tw[1], g, plain, s));
    prof_end(on, sm, s);
    const GenCols gx = gen_xcols(g, pr.lz, P, g.n2h, -1, pr.r0, -1, pr.r1);
    on = prof_begin(4, s, sm, gen_col_bytes(gx, g.n0, batch));
    PB3D_CHECK_HIP(gen_cols(-1, 2, p->ax[0], work, out, gx, wbs, out_bs, batch, p->tw[0], g, kop, s));
    prof_end(on, sm, s);
    g_fwd += batch;
    return 0;
  }
  bool on = prof_begin(0, s, sm, br);
  PB3D_CHECK_HIP(launch_r2c_rows(p->L[2], in, work, rows, g.n0 * g.n1, p->tw[2], p->twN, g, rop, pr.lz,
                                 pitch ? pitch : g.n2h, s));
  prof_end(on, sm, s);
  on = prof_begin(6, s, sm, col_bytes(cy, p->L[1], batch));
  PB3D_CHECK_HIP((launch_cols<-1, 0>(p->L[1], work, work, cy, wbs, wbs, batch, p->tw[1], g, plain, s)));
  prof_end(on, sm, s);
  on = prof_begin(4, s, sm, col_bytes(cx, p->L[0], batch));
  PB3D_CHECK_HIP((launch_cols<-1, 2>(p->L[0], work, out, cx, wbs, out_bs, batch, p->tw[0], g, kop, s)));
  prof_end(on, sm, s);
  g_fwd += batch;
  return 0;
}

extern "C" int pb3d_fft_c2r(const pb3d_plan* p, const void* in_, float* out, void* work_, int batch,
                            const pb3d_kop* kop_, const pb3d_rop* rop_, void* stream) {
  if (!p || !out || !work_ || batch < 1 || batch > 65535) return fail(PB3D_ERR_ARG, "c2r: null pointer or batch out of [1, 65535]");
  if (!in_ && !(kop_ && kop_->kind == 4)) return fail(PB3D_ERR_ARG, "c2r: null input");
  if (in_ == work_) return fail(PB3D_ERR_ARG, "c2r: work must not alias the input");
  if (kop_ && kop_->kind == 5) return fail(PB3D_ERR_ARG, "c2r: PACK is an R2C epilogue");
  if (int e = check_rop(rop_, false)) return e;
  if (int e = check_kop(kop_)) return e;
  hipStream_t s = (hipStream_t)stream;
  const Geo& g = p->geo;
  const float2* in = (const float2*)in_;
  float2* work = (float2*)work_;
  ROp rop = to_rop(rop_);
  KOp kop = to_kop(kop_);
  KOp plain = to_kop(nullptr);
  const Prune pr = make_prune(g, kop_);
  const long long in_bs = kop_ ? kop_->x_batch_stride : g.H;
  // the band-limited intermediate is kept compactly in work (z planes kz < lz only)
  const int pitch = compact_pitch(g, pr.lz);
  const long long wbs = (long long)g.n0 * g.n1 * pitch;
  const ColMap cx = xcols(g, pr, true, 0, pitch), cy = ycols(g, pr, true, pitch, pitch);
  const double br = (double)batch * ((double)g.n0 * g.n1 * pr.lz * 8.0 +
                                     g.N * 4.0 * (rop_operands(rop_) + (rop_ && rop_->accumulate ? 1 : 0)));
  ProfSample sm{};
  if (p->generic) {
    const GenCols gx = gen_xcols(g, pr.lz, g.n2h, pitch, pr.r0, -1, pr.r1, -1);
    bool on = prof_begin(3, s, sm, gen_col_bytes(gx, g.n0, batch));
    PB3D_CHECK_HIP(gen_cols(+1, 1, p->ax[0], in, work, gx, in_bs, wbs, batch, p->tw[0], g, kop, s));
    prof_end(on, sm, s);
    const GenCols gy = gen_ycols(g, pr.lz, pitch, pitch, pr.r1);
    on = prof_begin(2, s, sm, gen_col_bytes(gy, g.n1, batch));
    PB3D_CHECK_HIP(gen_cols(+1, 0, p->ax[1], work, work, gy, wbs, wbs, batch, p->tw[1], g, plain, s));
    prof_end(on, sm, s);
    on = prof_begin(1, s, sm, br);
    PB3D_CHECK_HIP(gen_c2r_rows(p->ax[2], work, out, (long long)batch * g.n0 * g.n1, g.n0 * g.n1, p->tw[2], p->twN,
                                g, rop, (float)(1.0 / (double)g.N), pr.lz, pitch, s));
    prof_end(on, sm, s);
    g_rev += batch;
    return 0;
  }
  bool on = prof_begin(3, s, sm, col_bytes(cx, p->L[0], batch));
  PB3D_CHECK_HIP((launch_cols<+1, 1>(p->L[0], in, work, cx, in_bs, wbs, batch, p->tw[0], g, kop, s)));
  prof_end(on, sm, s);
  on = prof_begin(2, s, sm, col_bytes(cy, p->L[1], batch));
  PB3D_CHECK_HIP((launch_cols<+1, 0>(p->L[1], work, work, cy, wbs, wbs, batch, p->tw[1], g, plain, s)));
  prof_end(on, sm, s);
  const long long rows = (long long)batch * g.n0 * g.n1;
  on = prof_begin(1, s, sm, br);
  PB3D_CHECK_HIP(launch_c2r_rows(p->L[2], work, out, rows, g.n0 * g.n1, p->tw[2], p->twN, g, rop,
                                 (float)(1.0 / (double)g.N), pr.lz, pitch, s));
  prof_end(on, sm, s);
  g_rev += batch;
  return 0;
}

extern "C" int64_t pb3d_fft_pointing_work_size(const pb3d_grid* grid, int batch, int nsub, const pb3d_submap* subs,
                                               const pb3d_kop* kop_out) {
  if (!grid || !subs || batch < 1 || nsub < 1 || nsub > kMaxSub) {
    fail(PB3D_ERR_ARG, "pointing_work_size: null pointer, batch < 1 or nsub out of [1, 16]");
    return -1;
  }
  const Geo g = make_geo(*grid);
  Prune pin, pout;
  pointing_prunes(g, nsub, subs, kop_out, pin, pout);
  return (int64_t)batch * g.n0 * g.n1 * ((int64_t)nsub * compact_pitch(g, pin.lz) + compact_pitch(g, pout.lz));
}

extern "C" int64_t pb3d_icsum_partial_size(const pb3d_grid* grid, int batch, int nbins) {
  if (!grid || batch < 1 || nbins < 1 || nbins > kMaxIcBins) {
    fail(PB3D_ERR_ARG, "icsum_partial_size: null grid, batch < 1 or nbins out of [1, 64]");
    return -1;
  }
  const Geo g = make_geo(*grid);
  const long long rpm = (long long)g.n0 * g.n1;
  if (!fast_grid(g.n0, g.n1, g.n2)) {  // mixed-radix pointing rows: one partial per block of rows
    const long long rb = gen_rows_per_block(g.n2 / 2);
    return (int64_t)batch * ((rpm + rb - 1) / rb) * nbins;
  }
  long long rw = 64;
  switch (g.n2 / 2) {
    case 2: rw = PointCfg<2>::RW; break;
    case 4: rw = PointCfg<4>::RW; break;
    case 8: rw = PointCfg<8>::RW; break;
    case 16: rw = PointCfg<16>::RW; break;
    case 32: rw = PointCfg<32>::RW; break;
    case 64: rw = PointCfg<64>::RW; break;
    case 128: rw = PointCfg<128>::RW; break;
    case 256: rw = PointCfg<256>::RW; break;
    default: rw = 1; break;
  }
  return (int64_t)batch * ((rpm + rw - 1) / rw) * nbins;
}

static std::atomic<int> g_point_mode{1};

extern "C" int pb3d_pointing_rows_mode(int mode) {
  if (mode != 0 && mode != 1) {
    fail(PB3D_ERR_ARG, "pointing_rows_mode: mode must be 0 or 1");
    return -1;
  }
  return g_point_mode.exchange(mode);
}

extern "C" int pb3d_fft_pointing(const pb3d_plan* p, int batch, int nsub, const pb3d_submap* subs, const float* mask,
                                 float scale, const pb3d_kop* kop_out, const pb3d_icsum* ic_,
                                 const unsigned char* row_on, void* out_, void* work_, void* stream) {
  if (!p || !subs || !out_ || !work_ || batch < 1 || batch > 65535)
    return fail(PB3D_ERR_ARG, "pointing: null pointer or batch out of [1, 65535]");
  if (nsub < 1 || nsub > kMaxSub) return fail(PB3D_ERR_ARG, "pointing: nsub out of range [1, 16]");
  if (kop_out && kop_out->kind == 4) return fail(PB3D_ERR_ARG, "pointing: BKQ is a prologue");
  if (int e = check_kop(kop_out)) return e;
  IcSum ic{nullptr, nullptr, 1, nullptr, nullptr, 0};
  if (ic_) {
    if (!ic_->weight || !ic_->sums || !ic_->partial || ic_->nbins < 1 || ic_->nbins > kMaxIcBins)
      return fail(PB3D_ERR_ARG, "pointing: ic needs weight, sums, partial and 1 <= nbins <= 64");
    ic = IcSum{ic_->weight, ic_->bin, ic_->nbins, ic_->sums, ic_->partial, 0};
  }
  hipStream_t s = (hipStream_t)stream;
  const Geo& g = p->geo;
  float2* work = (float2*)work_;
  SubW sw{};
  sw.n = nsub;
  // row skipping needs a mask (rows are off only where it is zero)
  sw.row_on = mask ? row_on : nullptr;
  // sub-map order in the fused rows (and work slots): the ell = 2 sub-maps first
  int order[kMaxSub];
  int no = 0;
  for (int k = 0; k < nsub; ++k) if (subs[k].ell == 2) order[no++] = k;
  sw.nfact = no;
  for (int k = 0; k < nsub; ++k) if (subs[k].ell != 2) order[no++] = k;
  for (int k = 0; k < nsub; ++k) {
    const pb3d_submap& sb = subs[order[k]];
    if (int e = check_kop(&sb.kop)) return e;
    if (sb.kop.kind == 5) return fail(PB3D_ERR_ARG, "pointing: PACK is an R2C epilogue");
    if (!sb.in && sb.kop.kind != 4) return fail(PB3D_ERR_ARG, "pointing: sub-map input missing");
    if (sb.ell < 0 || sb.ell > 4 || (sb.ell > 0 && (sb.m < 0 || sb.m > 2 * sb.ell)))
      return fail(PB3D_ERR_ARG, "pointing: sub-map Y_lm index out of range (l<=4)");
    sw.ell[k] = sb.ell; sw.coef[k] = sb.coef;
    sw.y0[k] = sb.ell > 0 ? ylm_f(sb.ell, sb.m, 0.f, 0.f, 0.f) : 1.f;
    if (sb.ell > 0 && !ylm_poly(sb.ell, sb.m, sw.poly[k]))
      return fail(PB3D_ERR_ARG, "pointing: Y_lm polynomial fit failed");
  }
  // combined rows: every ell even, Lmax 2 or 4, fewer row transforms (Lmax + 1) than sub-maps
  {
    int lm = 0;
    bool even = true;
    for (int k = 0; k < nsub; ++k) { lm = sw.ell[k] > lm ? sw.ell[k] : lm; even = even && sw.ell[k] % 2 == 0; }
    sw.nu = 0;
    if (g_point_mode.load() == 1 && !p->generic && even && (lm == 2 || lm == 4) && lm + 1 < nsub && combined_rows_len(p->L[2])) {
      for (int k = 0; k < nsub; ++k) raise_poly(sw.ell[k], sw.poly[k], lm, sw.coef[k], sw.cpoly[k]);
      sw.nu = lm + 1;
    }
  }
  Prune pin, pout;
  pointing_prunes(g, nsub, subs, kop_out, pin, pout);
  KOp plain = to_kop(nullptr);
  ProfSample sm{};
  // Work maps use the compact layout: the nsub*batch sub-maps keep the z planes kz < pzi of the
  // inverse-side band limit, one map every wsi entries; the batch output maps after them keep
  // kz < pzo of the forward-side limit, one every wso entries.
  const int pzi = compact_pitch(g, pin.lz), pzo = compact_pitch(g, pout.lz);
  const long long wsi = (long long)g.n0 * g.n1 * pzi, wso = (long long)g.n0 * g.n1 * pzo;
  float2* outw = work + (long long)nsub * batch * wsi;
  if (p->generic) {
    GenSubs gs{};
    gs.n = nsub;
    gs.row_on = mask ? row_on : nullptr;
    for (int k = 0; k < nsub; ++k) {
      gs.ell[k] = subs[order[k]].ell; gs.m[k] = subs[order[k]].m; gs.coef[k] = subs[order[k]].coef;
    }
    for (int k = 0; k < nsub; ++k) {
      const pb3d_submap& sb = subs[order[k]];
      const GenCols gx = gen_xcols(g, pin.lz, g.n2h, pzi, pin.r0, -1, pin.r1, -1);
      bool on = prof_begin(3, s, sm, gen_col_bytes(gx, g.n0, batch));
      PB3D_CHECK_HIP(gen_cols(+1, 1, p->ax[0], (const float2*)sb.in, work + (long long)k * batch * wsi, gx,
                              sb.kop.x_batch_stride, wsi, batch, p->tw[0], g, to_kop(&sb.kop), s));
      prof_end(on, sm, s);
    }
    GenCols gyi = gen_ycols(g, pin.lz, pzi, pzi, pin.r1);
    gyi.row_on = gs.row_on;  // rows the row kernel skips are not stored
    gyi.row_n1 = g.n1;
    bool on = prof_begin(2, s, sm, gen_col_bytes(gyi, g.n1, batch * nsub));
    PB3D_CHECK_HIP(gen_cols(+1, 0, p->ax[1], work, work, gyi, wsi, wsi, batch * nsub, p->tw[1], g, plain, s));
    prof_end(on, sm, s);
    const double rows1 = (double)g.n0 * g.n1;
    on = prof_begin(5, s, sm, (double)batch * (rows1 * 8.0 * (pin.lz * (double)nsub + pout.lz) + (mask ? g.N * 4.0 : 0.0)));
    const GenIc gic{ic.weight, ic.bin, ic.nbins, ic.sums, ic.partial};
    PB3D_CHECK_HIP(gen_pointing_rows(p->ax[2], work, wsi, pzi, outw, wso, pzo, (long long)g.n0 * g.n1, batch, gs, gic,
                                     mask, scale, pin.lz, pout.lz, p->tw[2], p->twN, g, s));
    prof_end(on, sm, s);
    GenCols gyo = gen_ycols(g, pout.lz, pzo, pzo, -1);
    gyo.row_zero = gs.row_on;  // and read as zeros
    gyo.row_n1 = g.n1;
    on = prof_begin(6, s, sm, gen_col_bytes(gyo, g.n1, batch));
    PB3D_CHECK_HIP(gen_cols(-1, 0, p->ax[1], outw, outw, gyo, wso, wso, batch, p->tw[1], g, plain, s));
    prof_end(on, sm, s);
    const GenCols gxo = gen_xcols(g, pout.lz, pzo, g.n2h, -1, pout.r0, -1, pout.r1);
    const long long out_bs = (kop_out && kop_out->kind == 5) ? kop_out->packed_len : g.H;
    on = prof_begin(4, s, sm, gen_col_bytes(gxo, g.n0, batch));
    PB3D_CHECK_HIP(gen_cols(-1, 2, p->ax[0], outw, (float2*)out_, gxo, wso, out_bs, batch, p->tw[0], g, to_kop(kop_out),
                            s));
    prof_end(on, sm, s);
    g_rev += (long long)batch * nsub;
    g_fwd += batch;
    return 0;
  }
  // 1. pruned x-column C2R pass of every sub-map into work[k*batch + b], over the common support
  //    (so that the y pass reads no entry that was not written)
  for (int k = 0; k < nsub; ++k) {
    const pb3d_submap& sb = subs[order[k]];
    KOp kk = to_kop(&sb.kop);
    const ColMap cx = xcols(g, pin, true, 0, pzi);
    bool on = prof_begin(3, s, sm, col_bytes(cx, p->L[0], batch));
    PB3D_CHECK_HIP((launch_cols<+1, 1>(p->L[0], (const float2*)sb.in, work + (long long)k * batch * wsi, cx,
                                       sb.kop.x_batch_stride, wsi, batch, p->tw[0], g, kk, s)));
    prof_end(on, sm, s);
  }
  // 2. pruned y-column C2R pass over all sub-maps (rows the row kernel skips are not stored)
  ColMap cyi = ycols(g, pin, true, pzi, pzi);
  cyi.row_on = sw.row_on;
  cyi.row_n1 = g.n1;
  bool on = prof_begin(2, s, sm, col_bytes(cyi, p->L[1], batch * nsub));
  PB3D_CHECK_HIP((launch_cols<+1, 0>(p->L[1], work, work, cyi, wsi, wsi, batch * nsub, p->tw[1], g, plain, s)));
  prof_end(on, sm, s);
  // 3. fused rows: inverse rows of every sub-map, weights, mask, forward row -> output map b
  const double rows1 = (double)g.n0 * g.n1;
  on = prof_begin(5, s, sm, (double)batch * (rows1 * 8.0 * (pin.lz * (double)nsub + pout.lz) + (mask ? g.N * 4.0 : 0.0)));
  PB3D_CHECK_HIP(launch_pointing_rows(p->L[2], work, wsi, pzi, outw, wso, pzo, (long long)g.n0 * g.n1, batch, sw, ic,
                                      mask, scale, pin.lz, pout.lz, p->tw[2], p->twN, g, s));
  prof_end(on, sm, s);
  // 4. pruned forward y and x column passes, epilogue into out
  ColMap cyo = ycols(g, pout, false, pzo, pzo);
  const ColMap cxo = xcols(g, pout, false, pzo, 0);
  cyo.row_zero = sw.row_on;  // the rows skipped above are zero rows: not read again
  cyo.row_n1 = g.n1;
  on = prof_begin(6, s, sm, col_bytes(cyo, p->L[1], batch));
  PB3D_CHECK_HIP((launch_cols<-1, 0>(p->L[1], outw, outw, cyo, wso, wso, batch, p->tw[1], g, plain, s)));
  prof_end(on, sm, s);
  KOp ko = to_kop(kop_out);
  const long long out_bs = (kop_out && kop_out->kind == 5) ? kop_out->packed_len : g.H;
  on = prof_begin(4, s, sm, col_bytes(cxo, p->L[0], batch));
  PB3D_CHECK_HIP((launch_cols<-1, 2>(p->L[0], outw, (float2*)out_, cxo, wso, out_bs, batch, p->tw[0], g, ko, s)));
  prof_end(on, sm, s);
  g_rev += (long long)batch * nsub;
  g_fwd += batch;
  return 0;
}
