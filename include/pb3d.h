/*
 * pb3d: MI355X (gfx950) kernels for the PolyBin3D binned-estimator path, C ABI.
 *
 * The reference has no C boundary: its hot path is the duck-typed `utils` module contract
 * (IntegrationUtils / MapUtils / module functions, PolyBin3D/cython/utils.pyx:17-692) plus the FFT
 * dispatch in PolyBin3D/base.py:176-229,256-302.  Each entry point below names the reference
 * interface it replaces.  The Python host layer (polybin3d_amd/_capi.py) binds these with ctypes.
 *
 * Conventions
 *   - All device pointers are owned by the caller (PyTorch tensors); the library never frees or
 *     retains them past the call.  Plans own only their twiddle tables.
 *   - Real grids are float32 [batch][n0][n1][n2]; half-complex grids are float2 (re,im)
 *     [batch][n0][n1][n2/2+1]; both C-contiguous.
 *   - Forward transforms are unnormalised (e^{-ik.x}); inverse transforms carry 1/(n0 n1 n2),
 *     the numpy convention the reference uses (PolyBin3D/base.py:176-229).
 *   - Every call is asynchronous on `stream` (a hipStream_t, NULL = default stream) and returns
 *     0 on success or a negative PB3D_ERR_* code; pb3d_last_error() gives the message
 *     (thread-local).  Argument checks happen on the host before any launch.
 *   - Bin membership is given by a host-built uint8 bin-index table over the half grid
 *     (255 = outside every bin), so it is bit-identical to the reference's FP64 test
 *     `modk >= k_lo && modk < k_hi` (PolyBin3D/cython/utils.pyx:63).
 */
#ifndef PB3D_H
#define PB3D_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PB3D_OK 0
#define PB3D_ERR_ARG (-1)
#define PB3D_ERR_UNSUPPORTED (-2)
#define PB3D_ERR_HIP (-3)
#define PB3D_ERR_ALLOC (-4)

#define PB3D_NO_BIN 255

typedef struct pb3d_plan pb3d_plan;

/* Box geometry (PolyBin3D/base.py:31-116).  sight = unit line of sight for the global case
 * (z axis); for local (Yamamoto) sightlines it is boxcenter/|boxcenter|, the distant-observer
 * direction the reference uses for mu (base.py:111-116). */
typedef struct {
  int32_t n[3];
  int32_t local;
  double boxsize[3];
  double boxcenter[3];
  double sight[3];
} pb3d_grid;

/* Real-space voxel operation, used as the R2C prologue (input side) and the C2R epilogue
 * (output side).  kind:
 *   0 PLAIN   v = a
 *   1 MUL     v = a * b                      (mask, pointing matrix; base.py:399, pspec.py:212)
 *   2 DIFF    v = a*b - c*d                  (prod_real_diff, utils.pyx:371-383)
 *   3 YLM     v = a * Y_lm(x_hat)            (apply_fourier_harmonics prologue, base.py:409)
 *   4 MULYLM  v = a * b * Y_lm(x_hat)
 *   5 SQUARE  v = a*a                        (C2R epilogue only: accumulate squares,
 *                                             bspec.py:258)
 * For C2R epilogues `a` is ignored (the transform output takes its place) and b..d are the
 * extra operands; out = scale*v (accumulate=0) or out += scale*v (accumulate=1).
 * Pointers a..d are advanced by `batch_stride` elements per batch entry (0 = broadcast). */
typedef struct {
  int32_t kind;
  int32_t ell, m;
  int32_t accumulate;
  const float* a;
  const float* b;
  const float* c;
  const float* d;
  int64_t batch_stride_a, batch_stride_b, batch_stride_c, batch_stride_d;
  float scale;
  int32_t pad_;
} pb3d_rop;

/* Fourier-space mode operation, used as the C2R prologue (input side) and the R2C epilogue
 * (output side).  kind:
 *   0 PLAIN   v = x
 *   1 FILTER  v = x * [bin(k) == bins[batch]] * L_ell(mu) * L_ell2(mu) * w(k)
 *             (fourier_filter / fourier_filter_ll, utils.pyx:283-320; w optional)
 *   2 YLM     v = x * [bin(k) == bins[batch] or bins==NULL] * Y_lm(k_hat) * w(k)
 *             (prod_fourier_filter with a Y_lm grid, utils.pyx:325-338; base.py:398,410)
 *   3 WEIGHT  v = x * w(k)                   (prod_fourier with invPk0, utils.pyx:213-225)
 *   4 BKQ     C2R prologue of the bispectrum MC Fisher: the Q map of triplet trip[trip0+batch] =
 *             (b1,b2,b3,l) assembled from packed FT[g_A g_B] rows (add_Q_element,
 *             PolyBin3D/bspec.py:615-640): terms bit0 = Theta_b3 [L_l(mu)] FT_0(b1,b2);
 *             bit1 = Theta_b1 FT_l(b2,b3) + Theta_b2 FT_l(b1,b3); bit2 = Theta_b3 Y_lm(k_hat)
 *             FT_0(b1,b2).  FT_l(A,B) = ft + (l/2)*ft_lstride + pair_index[A*nk+B]*packed_len.
 *   5 PACK    R2C epilogue: out[batch*packed_len + pack_index[k]] = v for modes with
 *             pack_index >= 0 (the in-bin modes, stored bin-major).
 *   6 BINPOLY v = x * w(k) * sum_{q<nk} coef[q*terms + bin(k)] * L_{2q}(mu) * (2 on the kz = n2/2
 *             plane, else 1), coef = (const float*)ft, terms = number of bins, 0 outside every
 *             bin: the mode-space weight of the Pk_ideal adjoint (no counterpart in the reference;
 *             the factor 2 is the reference's degeneracy degk = 1 + (kz > 0), base.py:94, on the
 *             self-conjugate Nyquist plane).
 * For R2C epilogues x is the transform output; out = scale*v or out += scale*v.
 * `x_batch_stride` (complex elements) advances the C2R input per batch entry (0 = broadcast:
 * one input map filtered into many bins); with `x_batch_group` = g > 1 it advances once per g
 * consecutive entries (entry t reads input t / g: several realizations, each filtered into g bins). */
typedef struct {
  int32_t kind;
  int32_t ell, ell2;
  int32_t m;
  int32_t accumulate;
  int32_t pad_;
  const uint8_t* bin_table;
  const int32_t* bins;
  const float* w;
  int64_t x_batch_stride;
  float scale;
  int32_t pad2_;
  /* PACK / BKQ */
  const int32_t* pack_index;
  int64_t packed_len;
  const int32_t* trip;
  const int32_t* pair_index;
  const void* ft;
  int64_t ft_lstride;
  int32_t nk;
  int32_t terms;
  int32_t trip0;
  int32_t leg3_legendre;
  /* Band limit: as a C2R prologue, the input is zero outside |kx| <= support[0]*kF0,
   * |ky| <= support[1]*kF1, kz <= support[2]*kF2 (frequency-index radii) and the passes skip those
   * columns and reads; as an R2C epilogue only modes inside it are produced (PACK outputs, or a
   * full grid whose modes outside the support are left unwritten).  A negative entry = full axis. */
  int32_t support[3];
  int32_t x_batch_group;  /* 0 or 1: one input step per batch entry */
} pb3d_kop;

/* One sub-map of pb3d_fft_pointing: C2R(kop(in)) weighted in real space by coef * Y_lm(x_hat)
 * (ell = 0: weight coef). */
typedef struct {
  pb3d_kop kop;
  const void* in;
  int32_t ell, m;
  float coef;
  int32_t pad_;
} pb3d_submap;

/* Per-bin weighted sums taken inside pb3d_fft_pointing, for the integral-constrained pointing
 * matrices (PolyBin3D/cython/utils.pyx:405-513; pspec.py:199-213; bspec.py:412-426): for output map b
 * and bin z,
 *   sums[b*nbins + z] += sum over voxels x with bin[x] == z of m_b(x) * weight[x],
 * where m_b = scale * sum_g coef_g Y_g(x_hat) C2R(kop_g(in_g))[b] is the combined real map before the
 * mask.  bin == NULL puts every voxel in bin 0 (global constraint, nbins = 1); entries >= nbins are
 * in no bin.  1 <= nbins <= 64.  The caller zeroes `sums`; FP64 accumulation.  `partial`: device
 * scratch of pb3d_icsum_partial_size() doubles, where every wave of the row kernel stores its own
 * partial sums (no atomics); they are then added to `sums` in a fixed order, so the sums are
 * bitwise reproducible. */
typedef struct {
  const float* weight;
  const uint8_t* bin;
  int32_t nbins;
  int32_t pad_;
  double* sums;
  double* partial;
} pb3d_icsum;

/* ---------------------------------------------------------------- library */
int pb3d_version(void);
const char* pb3d_last_error(void);
/* FFT counters kept like PolyBin3D.n_FFTs_forward / n_FFTs_reverse (base.py:134-135,178,204):
 * each batch entry counts one transform.  dir: 0 = forward, 1 = reverse, -1 = reset both. */
int64_t pb3d_fft_counter(int dir);

/* Sampled launch timing for roofline reports (no counterpart in the reference): with period P > 0
 * every P-th launch of each FFT pass kind is bracketed by hipEvents on its stream.  kinds:
 * 0 R2C z-rows, 1 C2R z-rows, 2 plain C2R column pass, 3 C2R first column pass (mode prologue),
 * 4 R2C last column pass (mode epilogue), 5 fused pointing rows (pb3d_fft_pointing step 3),
 * 6 plain R2C column pass.
 * pb3d_prof_read synchronises the sampled events and returns the number of sampled launches, their
 * summed duration (ms) and their summed algorithmic bytes: the elements the pass must read and
 * write after band-limit pruning, 8 B per complex mode and 4 B per real voxel (k-space table reads
 * of the prologues/epilogues are not counted).  pb3d_prof_enable resets the totals; period 0
 * switches sampling off. */
int pb3d_prof_enable(int period);
int pb3d_prof_read(int kind, int64_t* sampled, double* ms, double* bytes);

/* ---------------------------------------------------------------- FFT plans
 * Replace _fft_setup (base.py:256-302).  Supported: n0, n1 powers of two in [2, 512], n2 a power
 * of two in [4, 512] (real FFT layout). */
int pb3d_plan_create(pb3d_plan** plan, const pb3d_grid* grid);
int pb3d_plan_destroy(pb3d_plan* plan);

/* 3D real-to-complex transform, unnormalised (to_fourier, base.py:176-200).
 * in: real [batch][n0][n1][n2] read through `rop` (NULL = plain); out: half complex written
 * through `kop` (NULL = plain).  work: half-complex scratch of batch*n0*n1*(n2/2+1) float2, or
 * NULL to run the intermediate passes in `out` (not allowed when kop accumulates). */
int pb3d_fft_r2c(const pb3d_plan* plan, const float* in, void* out, void* work, int batch,
                 const pb3d_rop* rop, const pb3d_kop* kop, void* stream);

/* 3D complex-to-real transform with 1/N (to_real, base.py:202-229); numpy semantics on the
 * self-conjugate kz=0 and kz=n2/2 planes (imaginary parts ignored).
 * in: half complex read through `kop` (NULL = plain; never modified; may be NULL for BKQ); work: half-complex
 * scratch of batch*n0*n1*(n2/2+1) float2; out: real written through `rop` (NULL = plain). */
int pb3d_fft_c2r(const pb3d_plan* plan, const void* in, float* out, void* work, int batch,
                 const pb3d_kop* kop, const pb3d_rop* rop, void* stream);

/* The pointing chain of the MC Fisher Q maps with trivial S^-1, fused (apply_filter /
 * apply_filter_dagger and apply_real_harmonics, PolyBin3D/pspec.py:270-304, bspec.py:499-505,
 * 615-666, base.py:391-400):
 *   out[b] = kop_out( R2C( mask(x) * scale * sum_g coef_g Y_g(x_hat) * C2R( kop_g(in_g) )[b] ) )
 * The sub-map column passes run pruned to each input's support; one row kernel then inverts all
 * nsub rows, combines them with their real-space weights, applies the mask (NULL = none) and the
 * forward row transform, so the real-space maps never reach HBM.  ic (NULL = none): per-bin sums of
 * the combined real map, taken in the same kernel (see pb3d_icsum).  work: the float2 count given by
 * pb3d_fft_pointing_work_size (at most (nsub+1)*batch*n0*n1*(n2/2+1)): the intermediate maps keep
 * only the z planes inside the band limits.  nsub <= 16.  row_on (NULL = every row): one byte per
 * (x, y) row of the grid, 0 where the mask (and the ic weight) is zero on the whole z row; the rows
 * must come in aligned groups of 64 that are all on or all off.  Rows that are off are neither
 * transformed nor stored between the passes, and come out as zero rows (a mask is required). */
int pb3d_fft_pointing(const pb3d_plan* plan, int batch, int nsub, const pb3d_submap* subs,
                      const float* mask, float scale, const pb3d_kop* kop_out, const pb3d_icsum* ic,
                      const unsigned char* row_on, void* out, void* work, void* stream);
/* float2 entries of work that pb3d_fft_pointing needs for these sub-maps and output op: nsub*batch
 * maps of n0*n1*pzi and batch maps of n0*n1*pzo, pzi = the largest z band limit + 1 over the
 * sub-map kops, pzo = that of kop_out (n2/2+1 where one has none).  Host-only; -1 on a bad argument.  (No counterpart in the reference, whose
 * pointing allocates full maps per step.) */
int64_t pb3d_fft_pointing_work_size(const pb3d_grid* grid, int batch, int nsub, const pb3d_submap* subs,
                                    const pb3d_kop* kop_out);
/* How the row kernel of pb3d_fft_pointing forms sum_g coef_g Y_g(x_hat) C2R(...)_g (no counterpart in
 * the reference, which transforms every harmonic map): mode 1 (default) combines the sub-maps of one
 * row in kz space into Lmax+1 rows when every ell is even, Lmax is 2 or 4 and that is fewer rows than
 * sub-maps (n2 = 128, 256 or 512), so the row needs Lmax+1 inverse transforms (rows with
 * x = y = 0 keep the per-sub-map form); mode 0 transforms every sub-map.  Process-wide; returns the
 * previous mode, or -1 for a mode other than 0 or 1. */
int pb3d_pointing_rows_mode(int mode);
/* doubles of pb3d_icsum.partial for a pointing call of `batch` maps and `nbins` bins (host-only; -1 on
 * a bad argument). */
int64_t pb3d_icsum_partial_size(const pb3d_grid* grid, int batch, int nbins);

/* ---------------------------------------------------------------- mode-space maps */
/* out[batch] = kop(x) over the half grid, without a transform (fourier_filter,
 * prod_fourier, prod_fourier_filter; utils.pyx:213-338). */
int pb3d_kmap(const pb3d_grid* grid, const void* x, void* out, int batch, const pb3d_kop* kop,
              void* stream);

/* Gaussian random field amplitudes (generate_data, base.py:305-353): given the full-grid
 * draws re/im (float64 [n0][n1][n2], the reference's numpy stream, uploaded), write the half
 * grid 0.5*(r(k)+conj(r(-k))) * sqrt(P(k)) * N/sqrt(V) * W(k), DC set to zero before the
 * symmetrisation.  sqrtP_w: float32 half-grid of sqrt(P)*W (NULL = 1). */
int pb3d_grf_hermitian(const pb3d_grid* grid, const double* re, const double* im,
                       const float* sqrtP_w, float norm, void* out, void* stream);

/* Gaussian random field drawn on the GPU: same recipe and distribution as pb3d_grf_hermitian, with
 * the full-grid normals (re, im) generated from a counter-based hash of (seed, full-grid index)
 * instead of the numpy stream.  For throughput runs; parity runs use pb3d_grf_hermitian. */
int pb3d_grf_device(const pb3d_grid* grid, uint64_t seed, const float* sqrtP_w, float norm, void* out,
                    void* stream);
/* pb3d_grf_device with the seed read on the device when the kernel runs: seed = *seed_dev + offset
 * (seed_dev: one uint64 in device memory).  A HIP graph that captured this call draws a new field on
 * every replay after the caller rewrites *seed_dev (no counterpart in the reference). */
int pb3d_grf_device_at(const pb3d_grid* grid, const uint64_t* seed_dev, uint64_t offset, const float* sqrtP_w,
                       float norm, void* out, void* stream);

/* Packed-mode assembly without a transform: out[batch][p] (+)= scale * kop(mode = modes[p]) for
 * p < packed_len (kop kinds FILTER, YLM, WEIGHT, BKQ; for BKQ the input x is ignored).  Used
 * for constant masks, where the Fisher Q maps need no real-space step (bspec.py:662-664). */
int pb3d_kmap_packed(const pb3d_grid* grid, const void* x, const int32_t* modes, int64_t packed_len,
                     void* out, int batch, const pb3d_kop* kop, void* stream);

/* Row gather for the Fisher contraction: out[r][q] = src[rows[r]*src_stride + offset + q] *
 * w[offset + q] for q < len (complex src/out, real w; rows device int32 in [0, nsrc)). */
int pb3d_gather_rows(const void* src, int64_t src_stride, int nsrc, const int32_t* rows, int nrows,
                     int64_t offset, int64_t len, const float* w, void* out, void* stream);

/* ---------------------------------------------------------------- bin reductions
 * One pass over the half grid for a batch of maps (cross_integrate / integrate /
 * real_integrate_double, utils.pyx:51-134):
 *   out[q][o][b] = sum_{k in bin b} degk * Re[x_q(k) conj(y_q(k))] * L_{ells[o]}(mu)
 *                  * L_{ells2[o]}(mu)
 * x: nq maps (x_stride complex elements apart); y: maps for each output o: y + (q*y_qstride +
 * o*y_ostride) (y == NULL: y = x).  mode 1 = x is real (float32, stride in floats) and the
 * summand is degk*x^2*L*L.  Sums: FP32 inside a wave segment, FP64 across segments; out FP64.
 * work: FP64 scratch of pb3d_bin_reduce_work_size() bytes. */
size_t pb3d_bin_reduce_work_size(const pb3d_grid* grid, int nq, int nout, int nbins);
int pb3d_bin_reduce(const pb3d_grid* grid, const void* x, int64_t x_stride, const void* y,
                    int64_t y_qstride, int64_t y_ostride, int nq, int nout, const int32_t* ells,
                    const int32_t* ells2, int mode, const uint8_t* bin_table, int nbins,
                    double* out, void* work, void* stream);

/* The same sums (utils.pyx:51-134) visiting the listed modes only: modes = ascending linear
 * half-grid indices, nmodes <= H, that must
 * include every mode whose bin_table entry is not PB3D_NO_BIN (modes outside every bin may be
 * listed; they add nothing).  Work size as pb3d_bin_reduce_work_size().  Same result, without the
 * pass over the (mostly out-of-bin) rest of the grid. */
int pb3d_bin_reduce_modes(const pb3d_grid* grid, const void* x, int64_t x_stride, const void* y,
                          int64_t y_qstride, int64_t y_ostride, int nq, int nout, const int32_t* ells,
                          const int32_t* ells2, int mode, const uint8_t* bin_table, int nbins,
                          const int32_t* modes, int64_t nmodes, double* out, void* work, void* stream);

/* pb3d_bin_reduce_modes without a workspace argument: allocates nothing (no hipMalloc, hipFree or
 * allocating library call), for callers that cannot allocate (code under HIP graph capture, tight
 * serving loops, harnesses that forbid allocation).  No counterpart in the reference.  The block
 * partials go to a fixed 16 MiB device buffer of the library (2^21 doubles per GPU, held until the
 * process ends); a call that would need more (nq * nout * nbins * blocks, blocks <= 8192 / nq) is
 * rejected with PB3D_ERR_ARG before any launch.  Every call on one GPU shares that GPU's buffer: calls
 * that overlap on one GPU (two streams, two host threads) race; stream-ordered calls do not.  The
 * buffer needs no clearing: each call overwrites the entries it reads.  Same results as
 * pb3d_bin_reduce_modes, bit for bit. */
int pb3d_bin_reduce_modes_noalloc(const pb3d_grid* grid, const void* x, int64_t x_stride, const void* y,
                                  int64_t y_qstride, int64_t y_ostride, int nq, int nout, const int32_t* ells,
                                  const int32_t* ells2, int mode, const uint8_t* bin_table, int nbins,
                                  const int32_t* modes, int64_t nmodes, double* out, void* stream);

/* ---------------------------------------------------------------- triplet contraction
 * T[t] = sum_x g[i1[t]](x) g[i2[t]](x) g[i3[t]](x) for ntrip index triples into the
 * float32 map stack g[nmaps][nvox] (assemble_b3 / sum_triplet, utils.pyx:357-366,614-636).
 * precision 0 = FP32 products with FP64 accumulation.  Index triples are device int32 arrays;
 * an index outside [0, nmaps) yields NaN for that triplet.  out: FP64[ntrip] (device).
 * work: scratch of pb3d_triplet_work_size() bytes. */
size_t pb3d_triplet_work_size(int64_t nvox, int ntrip, int precision);
int pb3d_triplet_sum(const float* g, int nmaps, int64_t nvox, const int32_t* i1, const int32_t* i2,
                     const int32_t* i3, int ntrip, int precision, double* out, void* work,
                     void* stream);

/* Triplet sums on the matrix cores, as one GEMM (assemble_b3 / sum_triplet, utils.pyx:357-366,
 * 614-636, restated as T = G_A P^T):
 *   out[a][p] = sum_x g[amap[a]](x) g[pb[p]](x) g[pc[p]](x)      (FP64 out[na][np])
 * for na <= 128 A maps and np pairs of the float32 stack g[nmaps <= 256][nvox] (index arrays on
 * the device).  precision 1 = split BF16, 3 products (hi*hi + hi*mid + mid*hi, ~2^-16 relative
 * per product); 2 = split BF16, 6 products (~2^-24).  Two-level FP32 accumulation (runs of 1024
 * voxels, then their sum) within a voxel chunk, FP64 across chunks in a fixed order.  An index outside the stack gives NaN for its entries.
 * work: pb3d_pair_gemm_work_size() bytes. */
size_t pb3d_pair_gemm_work_size(int64_t nvox, int na, int np);
int pb3d_pair_gemm(const float* g, int nmaps, int64_t nvox, const int32_t* amap, int na, const int32_t* pb,
                   const int32_t* pc, int np, int precision, double* out, void* work, void* stream);

/* pb3d_pair_gemm with its launch shape chosen by the caller (the autotuner; no counterpart in the
 * reference): stage_cfg = voxels per LDS stage (0 = the built-in choice, else 32 or the long stage of
 * the stack's size class: 128 for <= 32 maps, 64 for <= 128), rows_cfg = A rows per launch (0 = the
 * built-in choice, else 32, 64 or 96).  Any accepted shape gives the same sums up to FP32 rounding
 * order; a shape that does not fit is rejected with PB3D_ERR_ARG before any launch. */
int pb3d_pair_gemm_cfg(const float* g, int nmaps, int64_t nvox, const int32_t* amap, int na, const int32_t* pb,
                       const int32_t* pc, int np, int precision, int stage_cfg, int rows_cfg, double* out, void* work,
                       void* stream);

/* Adjoint of the triplet sums (no counterpart in the reference; used by the Bk_ideal gradient):
 *   out[a][x] = sum_p U[a][p] g[pb[p]](x) g[pc[p]](x)           (float32 out[na][nvox])
 * U: float32 [na <= 128][np] on the device; g: float32 stack [nmaps <= 128][nvox].  precision as
 * for pb3d_pair_gemm.  work: pb3d_pair_adjoint_work_size() bytes. */
size_t pb3d_pair_adjoint_work_size(int na, int np, int precision);
int pb3d_pair_adjoint(const float* g, int nmaps, int64_t nvox, const int32_t* pb, const int32_t* pc, int np,
                      const float* U, int na, int precision, float* out, void* work, void* stream);

/* ---------------------------------------------------------------- Fisher contraction
 * The rows of the bispectrum Monte Carlo Fisher matrix (integrate_row, PolyBin3D/cython/utils.pyx:162-179,
 * as called at PolyBin3D/bspec.py:597: fish[i][j] = sum_{k in b2(i)} degk Re[conj(m_i(k)) Q_j(k)]) with
 * the in-bin modes stored bin-major, as one GEMM per bin segment.  A segment describes one middle
 * bin: `rows` left rows of `k` reals each (re, im interleaved, degk folded in) stored row-major at
 * left + a_off, contracted with reals [q_off, q_off + k) of every Q row (q + n * q_ld, n < nq). */
typedef struct {
  int64_t a_off;  /* float offset of the segment's [rows][k] left block */
  int64_t q_off;  /* float offset of the segment inside every Q row */
  int32_t k;      /* reals per row (even) */
  int32_t rows;   /* left rows */
  int32_t row0;   /* first entry of row_map for this segment */
  int32_t pad_;
} pb3d_fisher_seg;

/* Host-only: the output-tile list of pb3d_fisher_contract for these segments (host array) and nq Q
 * rows, written to tiles[3 * t .. 3 * t + 2] = (segment, first row, first column), largest k first.
 * Returns the number of tiles (tiles == NULL: only count), or a negative error. */
int pb3d_fisher_tiles(const pb3d_fisher_seg* segs, int nseg, int nq, int32_t* tiles, int max_tiles);

/* For every segment s, r < s.rows and n < nq (segs, tiles, row_map, col_map, scale on the device):
 *   out[row_map[s.row0 + r] * out_ld + col_map[n]] +=
 *       alpha * scale[row_map[..]] * scale[col_map[n]] * sum_{t < s.k} left[s.a_off + r s.k + t] q[n q_ld + s.q_off + t]
 * (scale == NULL: 1).  FP64 out, accumulated (callers zero it for a single realization).  One
 * persistent launch over the tile list; split-BF16 products on the matrix cores, precision 1 = 3
 * products (~2^-16 per product), 2 = 6 products (~2^-24); FP32 accumulation in runs of 512 reals,
 * then across runs, then FP64.  Every (row_map, col_map) target must be distinct: each output is
 * written by one thread, without atomics, in a fixed order. */
int pb3d_fisher_contract(const float* left, const float* q, int64_t q_ld, int nq, const pb3d_fisher_seg* segs,
                         int nseg, const int32_t* tiles, int ntiles, const int32_t* row_map, const int32_t* col_map,
                         const double* scale, double alpha, int precision, double* out, int64_t out_ld,
                         void* stream);

#ifdef __cplusplus
}
#endif
#endif /* PB3D_H */
