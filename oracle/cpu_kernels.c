/*
 * CPU restatement of the reference's bin-integral loops, kept only to time the reference's cost
 * model on the host (bench.py cpu_baseline leg).  TEST / BASELINE INFRASTRUCTURE: never linked into
 * the product.
 *
 * Loop structure follows PolyBin3D/cython/utils.pyx: each output (bin or map) sweeps the whole
 * half grid and tests bin membership per mode, parallelised over outputs with OpenMP (the Cython
 * `prange(..., schedule='static')`).
 *   cross_integrate: utils.pyx:51-75     integrate_row: utils.pyx:162-179
 */
#include <complex.h>
#include <stdint.h>

/* out[b] = sum_{edges[b] <= modk < edges[b+1]} degk * Re[s1 conj(s2)]  (l = 0) */
void cpu_cross_integrate(const double* modk, const double* degk, const double complex* s1,
                         const double complex* s2, int64_t n, const double* edges, int nbin, double* out) {
#pragma omp parallel for schedule(static)
  for (int b = 0; b < nbin; ++b) {
    double t = 0.;
    for (int64_t i = 0; i < n; ++i)
      if (modk[i] >= edges[b] && modk[i] < edges[b + 1]) t += degk[i] * creal(s1[i] * conj(s2[i]));
    out[b] = t;
  }
}

/* out[j] = sum_{kmin <= modk < kmax} degk * Re[conj(m) arr_j]  for j < nout */
void cpu_integrate_row(const double* modk, const double* degk, const double complex* m,
                       const double complex* arr, int64_t n, int nout, double kmin, double kmax, double* out) {
#pragma omp parallel for schedule(static)
  for (int j = 0; j < nout; ++j) {
    double t = 0.;
    const double complex* a = arr + (int64_t)j * n;
    for (int64_t i = 0; i < n; ++i)
      if (modk[i] >= kmin && modk[i] < kmax) t += degk[i] * creal(conj(m[i]) * a[i]);
    out[j] = t;
  }
}
