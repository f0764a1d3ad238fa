// Mixed-radix transforms for grid sizes outside the power-of-two fast path (any n0, n1 in
// [2, 1024], even n2 in [4, 2048]; PolyBin3D/base.py:50-54 only requires an even last axis).
//
// Every pass is a Stockham autosort line DFT in LDS: a block holds `lpb` lines of length L, and
// a stage of radix R (L = R_0 R_1 ... R_{S-1}) writes output o = blk*Ns*R + q*Ns + k from
// the R inputs j + r*L/R (j = blk*Ns + k) as sum_r x * w^(r*(k*L/(Ns*R) + q*L/R)), w = e^{-+2 pi i/L},
// from one table of e^{-2 pi i m/L} per axis.  A prime factor p costs p terms per output, so the
// work per line is L * sum R_s.  The real axis is transformed as M = n2/2 complex points (even and
// odd voxels packed into one complex value), split / merged with e^{-2 pi i k/n2}.
//
// Pass semantics are those of the power-of-two passes (supports, compact work pitches, voxel ops
// on the rows, mode ops on the x columns), so the entry points can route either way.
#pragma once
#include "pb3d_ops.h"

namespace pb3d {

constexpr int kGenMaxRadices = 24;
constexpr int kGenMaxSub = 16;
constexpr int kGenMaxIcBins = 64;
constexpr int kGenThreads = 256;

struct GenAxis {
  int L, nr;
  int r[kGenMaxRadices];
};
// radices of L (factors 4 first, then primes in increasing order); false if L < 1
bool gen_factor(int L, GenAxis& ax);

// Column pass along an axis of length L over columns c < ncols: q = c / cpp, l = c % cpp;
// element p of column c sits at mode index q*plane + l + p*S, input q*in_plane + l + p*in_S,
// output q*out_plane + l + p*out_S.  Inputs with p outside rad_in or q outside qrad_in are known
// zeros (not read); outputs with p outside rad_out or q outside qrad_out are not written
// (rad < 0: whole axis).  qlen is the length of the q axis.
struct GenCols {
  const unsigned char* row_on;    // MODE 0: (q, p) rows not stored where 0 (NULL: all)
  const unsigned char* row_zero;  // MODE 0: (q, p) rows read as zeros where 0 (NULL: none)
  int row_n1;                     // p positions per q in the row flags
  long long ncols;
  int cpp;
  long long plane, S, in_plane, in_S, out_plane, out_S;
  int rad_in, rad_out, qlen, qrad_in, qrad_out;
};

// Pointing sub-maps: weight coef * Y_{ell,m}(r_hat) (coef alone for ell = 0).
struct GenSubs {
  const unsigned char* row_on;  // per row: 0 = the mask (and IC weight) is zero on it (NULL: all on)
  int n;
  int ell[kGenMaxSub], m[kGenMaxSub];
  float coef[kGenMaxSub];
};
struct GenIc {
  const float* weight;  // nullptr: no sums
  const uint8_t* bin;   // nullptr: every voxel in bin 0
  int nbins;
  double* sums;
  double* partial;
};

// MODE 0 plain, 1 first C2R pass (input through kop, broadcast by x_group), 2 last R2C pass
// (output through kop: scale, accumulate, packed rows).  DIR -1 forward, +1 inverse (unnormalised).
hipError_t gen_cols(int dir, int mode, const GenAxis& ax, const float2* in, float2* out, const GenCols& cm,
                    long long in_bs, long long out_bs, int batch, const float2* tw, const Geo& g, const KOp& kop,
                    hipStream_t s);
// Real rows of n2 = 2M voxels -> kz < lz_out of the half spectrum at out + row*pitch (rows of all
// batch entries back to back, rpb per entry); the voxels are read through rop.
hipError_t gen_r2c_rows(const GenAxis& axM, const float* in, float2* out, long long nrows, int rpb,
                        const float2* tw, const float2* twN, const Geo& g, const ROp& rop, int lz_out, int pitch,
                        hipStream_t s);
// Half-spectrum rows (kz < lz_in at in + row*pitch; Im of kz = 0 and kz = n2/2 ignored) -> real
// voxels * norm, through rop, into out.
hipError_t gen_c2r_rows(const GenAxis& axM, const float2* in, float* out, long long nrows, int rpb,
                        const float2* tw, const float2* twN, const Geo& g, const ROp& rop, float norm, int lz_in,
                        int pitch, hipStream_t s);
// Fused pointing rows: per output map t and row, the inverse rows of every sub-map (work map
// g*batch + t), sum of coef_g Y_g * z, optional per-bin sums of the result * weight, mask * scale/N,
// forward row -> kz < lz_out of output map t.
hipError_t gen_pointing_rows(const GenAxis& axM, const float2* work, long long slot_in, int pitch_in, float2* out,
                             long long slot_out, int pitch_out, long long rpm, int batch, const GenSubs& gs,
                             const GenIc& ic, const float* mask, float scale, int lz_in, int lz_out,
                             const float2* tw, const float2* twN, const Geo& g, hipStream_t s);
// rows per block of the pointing rows (the number of per-bin partial slots per map is
// ceil(rows / this))
int gen_rows_per_block(int M);

}  // namespace pb3d
