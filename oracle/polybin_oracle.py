"""FP64 CPU oracle for the PolyBin3D binned-estimator path.

TEST INFRASTRUCTURE ONLY.  This module is the checker: it may be imported by ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` and by nothing else.  The
product package ``polybin3d_amd`` never imports it and has no CPU fallback.

It restates, in plain numpy/scipy FP64, the algorithm of the reference estimators so that the HIP
path can be compared against it on identical seeded inputs.  Every function cites the reference
file:line it follows (paths relative to the reference repository root).  The FFTs are
``scipy.fft.rfftn/irfftn`` (numpy conventions: forward unnormalised, inverse x 1/N), standing in for
the reference's FFTW/MKL calls; FFT counters are kept exactly as the reference keeps them
(PolyBin3D/base.py:176-229) because the published notebook counters pin the call graph.

How each part is pinned (tests/test_oracle_*.py):
  * grids, bin membership, pixel windows: closed forms and the notebook's printed kF/kNy
    (Tutorial 1 cell 8) plus mode counting by brute force;
  * FFT wrappers: brute-force DFT on small grids, Parseval, plane wave <-> delta;
  * utils kernels (integrate, cross_integrate, real_integrate_double, integrate_row, filters,
    sum_pair/sum_triplet, assemble_b3/b1/bshot, the integral-constrained pointing_GIC/RIC and their
    transposes): explicit per-bin loops on tiny grids, transposes as adjoints, and the compiled
    reference kernels in oracle/_ref when that build is present;
  * real spherical harmonics: the addition theorem sum_m Y_lm(a) Y_lm(b) = L_l(a.b);
  * Pk_ideal: plane-wave known answer P = A^2 V / (2 N_modes);
  * Bk numerator: exact Fourier-space triple sum by brute force on 8^3 grids;
  * FFT counts of the whole call graphs: the notebook counters 1/0, 15/1, 2/96 and 452/682;
  * Pk MC Fisher (global, constant mask): per-seed closed form from the GRF amplitudes; local and
    masked branches: dense-operator restatement on 6^3-8^3 grids via the addition theorem, also with
    odd_l (odd-l legs and their -i / i phases) for the MC Fisher and the numerator;
  * Bk MC Fisher: in expectation over 200 seeds (unit mask, ideal weights) equal to the ideal Fisher
    with the discreteness correction, a separate code path, which it thereby also pins for l > 0;
    per seed, with a varying mask (global and local sightline) and with the global integral
    constraint, against a dense-operator restatement (explicit filter matrices, the cubic form's
    gradient through their transposes, the pointing matrix as a matrix; tests/test_oracle_pins_bk.py),
    which rejects planted mistakes (a dropped Legendre leg, a missing or transposed pointing matrix, a
    conjugated row, an unweighted A^-1, transposed harmonics);
  * Bk symmetry factors for l > 0: lattice triangle sums with the addition theorem (brute force);
  * local-sightline Bk numerator: dense DFT-matrix restatement with L_l(k_hat . x_hat);
  * apply_xi (local): the dense symmetric operator sum_k e^{ik(x-y)} [P_0 + P_2 (L_2(k.x)+L_2(k.y))/2];
  * Pk covariance contribution: per-seed closed form (global, constant mask, filter m^2 P + m) and the
    dense restatement with the filter P xi P^T + diag(mask_shot) (local, varying mask);
  * Pk theory binning matrix: equal to the Fisher matrix on the data bins, its l = 4 block equal to
    the lmax = 4 Fisher matrix's, refined bins summing to coarse ones; the wide-angle terms by a dense
    restatement; derivative_coefficient on hand-evaluated edges; cross_integrate_wcoeff against
    the compiled reference kernel.
  * PSpec.compute_shot_contribution: per-seed closed form (global, constant mask) and a dense
    restatement (local, varying and shot-noise masks).
  * BSpec.compute_shot_contribution: per seed as exact Fourier-space lattice sums with the
    real-space products as circular convolutions (global, constant masks), against dense local
    filter matrices (local, varying mask and shot-noise mask), and the cubic term's expectation
    against its closed form c T_i / (V sym_i); a dropped leg permutation or a Legendre factor on the
    wrong leg fails them.
Every function above has an independent pin; none is left as restatement only.
"""
import numpy as np
import scipy.fft
import scipy.interpolate
import os


# ----------------------------------------------------------------------------------------------
# Small helpers (PolyBin3D/cython/utils.pyx)
# ----------------------------------------------------------------------------------------------

def legendre(l, mu):
    """Explicit Legendre polynomials L_0..L_6 (PolyBin3D/cython/utils.pyx:385-400)."""
    mu = np.asarray(mu, dtype=np.float64)
    if l == 0:
        return np.ones_like(mu)
    if l == 1:
        return mu
    if l == 2:
        return 0.5 * (3 * mu ** 2 - 1)
    if l == 3:
        return (-3 * mu + 5 * mu ** 3) / 2.
    if l == 4:
        return (3 - 30 * mu ** 2 + 35 * mu ** 4) / 8.
    if l == 5:
        return (15 * mu - 70 * mu ** 3 + 63 * mu ** 5) / 8.
    if l == 6:
        return (-5 + 105 * mu ** 2 - 315 * mu ** 4 + 231 * mu ** 6) / 16.
    raise ValueError("Legendre polynomial only defined up to l=6")


def real_harmonics(coords, lmax, odd_l):
    """Real spherical harmonics Y_lm(n) for l=1..lmax, normalised so that
    sum_m Y_lm(a) Y_lm(b) = L_l(a.b) (the convention of compute_real_harmonics,
    PolyBin3D/cython/utils.pyx:518-609).  Odd l only when odd_l.  A zero vector is normalised to
    (0, 0, 0) (utils.pyx:534-538) and the polynomials are evaluated there: every member vanishes
    except the l=4 one (35 z^4 - 30 z^2 + 3)/8 = 3/8 (utils.pyx:599).

    Returns {l: array[2l+1, ...]}.  The basis is built from the standard tesseral harmonics
    (Condon-Shortley phases dropped) times sqrt(4 pi / (2l+1)); any orthonormal real basis gives
    identical estimator outputs because only sum_m Y_lm Y_lm enters them.
    """
    c = np.asarray(coords, dtype=np.float64)
    r = np.sqrt(c[0] ** 2 + c[1] ** 2 + c[2] ** 2)
    safe = np.where(r == 0, 1.0, r)
    x = np.where(r == 0, 0.0, c[0] / safe)
    y = np.where(r == 0, 0.0, c[1] / safe)
    z = np.where(r == 0, 0.0, c[2] / safe)
    out = {}
    if odd_l and lmax >= 1:
        out[1] = np.stack([x, y, z])
    if lmax >= 2:
        s3 = np.sqrt(3.)
        out[2] = np.stack([s3 * x * y, s3 * y * z, z * z - 0.5 * (x * x + y * y), s3 * x * z,
                           0.5 * s3 * (x * x - y * y)])
    if odd_l and lmax >= 3:
        out[3] = np.stack([
            np.sqrt(5. / 8.) * y * (3 * x * x - y * y),
            np.sqrt(15.) * x * y * z,
            np.sqrt(3. / 8.) * y * (5 * z * z - 1),
            0.5 * z * (5 * z * z - 3),
            np.sqrt(3. / 8.) * x * (5 * z * z - 1),
            0.5 * np.sqrt(15.) * z * (x * x - y * y),
            np.sqrt(5. / 8.) * x * (x * x - 3 * y * y)])
    if lmax >= 4:
        out[4] = np.stack([
            0.5 * np.sqrt(35.) * x * y * (x * x - y * y),
            np.sqrt(35. / 8.) * y * z * (3 * x * x - y * y),
            0.5 * np.sqrt(5.) * x * y * (7 * z * z - 1),
            np.sqrt(5. / 8.) * y * z * (7 * z * z - 3),
            (35 * z ** 4 - 30 * z * z + 3) / 8.,
            np.sqrt(5. / 8.) * x * z * (7 * z * z - 3),
            0.25 * np.sqrt(5.) * (x * x - y * y) * (7 * z * z - 1),
            np.sqrt(35. / 8.) * x * z * (x * x - 3 * y * y),
            np.sqrt(35.) / 8. * (x ** 4 - 6 * x * x * y * y + y ** 4)])
    return out


def bin_index(modk, edges):
    """Index b with edges[b] <= modk < edges[b+1], or -1 (half-open bins as in
    PolyBin3D/cython/utils.pyx:63,92,128,173,292,313,336).  Requires increasing edges."""
    edges = np.asarray(edges, dtype=np.float64)
    assert np.all(np.diff(edges) > 0), "bin edges must be strictly increasing"
    idx = np.searchsorted(edges, modk, side='right') - 1
    idx[(idx < 0) | (idx >= len(edges) - 1)] = -1
    return idx


# ----------------------------------------------------------------------------------------------
# Integration / map utilities (utils contract, PolyBin3D/cython/utils.pyx:17-383)
# ----------------------------------------------------------------------------------------------

class Integrator:
    """IntegrationUtils restated (PolyBin3D/cython/utils.pyx:17-179) for real FFT grids."""

    def __init__(self, modk, degk, muk):
        self.modk, self.degk, self.muk = modk, degk, muk
        self._cache = {}

    def _idx(self, edges):
        key = tuple(np.asarray(edges, dtype=np.float64).tolist())
        if key not in self._cache:
            self._cache[key] = bin_index(self.modk, edges)
        return self._cache[key]

    def _binsum(self, w, edges):
        idx = self._idx(edges)
        nb = len(edges) - 1
        sel = idx >= 0
        return np.bincount(idx[sel], weights=w[sel], minlength=nb)[:nb]

    def cross_integrate(self, s1, s2, edges, l):
        """sum_{k in b} degk Re[s1 conj(s2)] L_l(mu) (utils.pyx:51-75)."""
        w = self.degk * np.real(s1 * np.conj(s2)) * (legendre(l, self.muk) if l != 0 else 1.)
        return self._binsum(w, edges)

    def cross_integrate_wcoeff(self, coeff, s1, s2, edges):
        """sum_{k in b} degk Re[coeff_b s1 conj(s2)], one complex coefficient per bin (utils.pyx:139-157)."""
        coeff = np.asarray(coeff, dtype=np.complex128)
        prod = self.degk * s1 * np.conj(s2)
        return coeff.real * self._binsum(prod.real, edges) - coeff.imag * self._binsum(prod.imag, edges)

    def integrate(self, s, edges, l):
        """sum_{k in b} degk |s|^2 L_l(mu) (utils.pyx:80-104)."""
        w = self.degk * np.real(s * np.conj(s)) * (legendre(l, self.muk) if l != 0 else 1.)
        return self._binsum(w, edges)

    def real_integrate_double(self, s, l1, l2, edges):
        """sum_{k in b} degk s^2 L_l1 L_l2 for real s (utils.pyx:109-134)."""
        leg = 1. if (l1 == 0 and l2 == 0) else legendre(l1, self.muk) * legendre(l2, self.muk)
        return self._binsum(self.degk * s ** 2 * leg, edges)

    def integrate_row(self, m, arr, kmin, kmax):
        """out[j] = sum_{kmin<=k<kmax} degk Re[conj(m) arr_j] (utils.pyx:162-179)."""
        sel = (self.modk >= kmin) & (self.modk < kmax)
        w = (self.degk[sel] * np.conj(m[sel]))[None, :]
        return np.real(np.sum(w * arr[:, sel], axis=1))


class MapUtils:
    """MapUtils restated (PolyBin3D/cython/utils.pyx:184-383)."""

    def __init__(self, modk, muk):
        self.modk, self.muk = modk, muk

    def fourier_filter(self, kmap, l, kmin, kmax):
        """Theta_[kmin,kmax)(|k|) L_l(mu) kmap (utils.pyx:283-299)."""
        sel = (self.modk >= kmin) & (self.modk < kmax)
        leg = legendre(l, self.muk) if l != 0 else 1.
        return np.where(sel, kmap * leg, 0.).astype(np.complex128)

    def fourier_filter_ll(self, kmap, l1, l2, kmin, kmax):
        """Theta L_l1 L_l2 kmap (utils.pyx:304-320)."""
        sel = (self.modk >= kmin) & (self.modk < kmax)
        leg = 1. if (l1 == 0 and l2 == 0) else legendre(l1, self.muk) * legendre(l2, self.muk)
        return np.where(sel, kmap * leg, 0.).astype(np.complex128)

    def prod_fourier_filter(self, m1, m2, kmin, kmax):
        """Theta m1 m2 (utils.pyx:325-338)."""
        sel = (self.modk >= kmin) & (self.modk < kmax)
        return np.where(sel, m1 * m2, 0.).astype(np.complex128)

    @staticmethod
    def div_fourier(m1, m2):
        """m1/m2 with 0 where m2==0 (utils.pyx:230-243)."""
        out = np.zeros(m1.shape, dtype=np.complex128)
        nz = m2 != 0
        out[nz] = m1[nz] / m2[nz]
        return out


def sum_pair(a, b):
    """sum_x a b (utils.pyx:343-352)."""
    return float(np.sum(a * b))


def sum_triplet(a, b, c):
    """sum_x a b c (utils.pyx:357-366)."""
    return float(np.sum(a * b * c))


def assemble_b3(g, all_bins):
    """B3[i] = sum_x g[0][b1] g[0][b2] g[l/2][b3] (utils.pyx:614-636)."""
    out = np.zeros(len(all_bins))
    for i, (b1, b2, b3, l) in enumerate(all_bins):
        out[i] = np.sum(g[0][b1] * g[0][b2] * g[l // 2][b3])
    return out


def pointing_GIC(m, mask, mask_IC):
    """mask (m - <m mask_IC> / <mask_IC>): pointing with the global integral constraint
    (utils.pyx:405-423)."""
    return mask * (m - np.sum(m * mask_IC) / np.sum(mask_IC))


def pointing_GIC_transpose(m, mask, mask_IC):
    """mask m - mask_IC <m mask> / <mask_IC> (utils.pyx:428-446)."""
    return mask * m - mask_IC * (np.sum(m * mask) / np.sum(mask_IC))


def _radial_bin(modr, edges):
    """Half-open radial bin index per voxel, -1 outside every bin (utils.pyx:462-463)."""
    z = np.full(modr.shape, -1, dtype=np.int64)
    for iz in range(len(edges) - 1):
        z[(modr >= edges[iz]) & (modr < edges[iz + 1])] = iz
    return z


def pointing_RIC(m, mask, mask_IC, edges, modr):
    """mask (m - av_z) in radial bin z, av_z = sum_z m mask_IC / sum_z mask_IC; 0 outside every bin
    (utils.pyx:451-481)."""
    z = _radial_bin(modr, edges)
    out = np.zeros(m.shape)
    for iz in range(len(edges) - 1):
        s = z == iz
        out[s] = mask[s] * (m[s] - np.sum(m[s] * mask_IC[s]) / np.sum(mask_IC[s]))
    return out


def pointing_RIC_transpose(m, mask, mask_IC, edges, modr):
    """mask m - mask_IC av_z in radial bin z, av_z = sum_z m mask / sum_z mask_IC; 0 outside every bin
    (utils.pyx:486-513)."""
    z = _radial_bin(modr, edges)
    out = np.zeros(m.shape)
    for iz in range(len(edges) - 1):
        s = z == iz
        out[s] = mask[s] * m[s] - mask_IC[s] * (np.sum(m[s] * mask[s]) / np.sum(mask_IC[s]))
    return out


def apply_pointing_ic(spec, m, transpose):
    """PSpec/BSpec.apply_pointing (pspec.py:199-213, bspec.py:412-426): mask multiply, or the
    integral-constrained pointing matrix."""
    if spec.add_RIC:
        f = pointing_RIC_transpose if transpose else pointing_RIC
        return f(m, spec.mask, spec.mask_IC, spec.radial_bins_RIC, spec.base.modr_grid)
    if spec.add_GIC:
        f = pointing_GIC_transpose if transpose else pointing_GIC
        return f(m, spec.mask, spec.mask_IC)
    return m * spec.mask


def _setup_ic(spec, mask, add_GIC, mask_IC, add_RIC, radial_bins_RIC):
    """Integral-constraint options (pspec.py:60-72, 101-109; bspec.py:86-112)."""
    spec.add_GIC, spec.add_RIC = add_GIC, add_RIC
    if mask is None:
        assert not add_GIC and not add_RIC
    if mask_IC is not None:
        spec.mask_IC = np.asarray(mask_IC, dtype=np.float64)
    if add_GIC:
        assert hasattr(spec, 'mask_IC')
    if add_RIC:
        assert not add_GIC and hasattr(spec, 'mask_IC') and len(radial_bins_RIC) > 0
        b = spec.base
        b.modr_grid = np.sqrt(b.r_grids[0] ** 2. + b.r_grids[1] ** 2. + b.r_grids[2] ** 2.)
        spec.radial_bins_RIC = np.asarray(radial_bins_RIC, dtype=np.float64)


def assemble_bshot(gA, gB, gC, all_bins):
    """Shot-noise triple sums: the six assignments of the map sets (A, B, C) to the legs (b1, b2,
    b3 with its multipole) (utils.pyx:641-663)."""
    out = np.zeros(len(all_bins))
    for i, (b1, b2, b3, l) in enumerate(all_bins):
        L = l // 2
        out[i] = np.sum(gA[0][b1] * gB[0][b2] * gC[L][b3] + gA[0][b1] * gC[0][b2] * gB[L][b3]
                        + gB[0][b1] * gA[0][b2] * gC[L][b3] + gB[0][b1] * gC[0][b2] * gA[L][b3]
                        + gC[0][b1] * gA[0][b2] * gB[L][b3] + gC[0][b1] * gB[0][b2] * gA[L][b3])
    return out


def assemble_b1(g, s, all_bins):
    """Linear-term triple sums with simulation maps s (utils.pyx:668-692)."""
    out = np.zeros(len(all_bins))
    for i, (b1, b2, b3, l) in enumerate(all_bins):
        out[i] = np.sum(g[0][b1] * s[0][b2] * s[l // 2][b3] + s[0][b1] * g[0][b2] * s[l // 2][b3]
                        + s[0][b1] * s[0][b2] * g[l // 2][b3])
    return out


# ----------------------------------------------------------------------------------------------
# Base class (PolyBin3D/base.py)
# ----------------------------------------------------------------------------------------------

PIXEL_WINDOWS = ['none', 'cic', 'tsc', 'pcs', 'interlaced-cic', 'interlaced-tsc', 'interlaced-pcs']


def pixel_window_1d(kind, x):
    """1D voxel windows at x = pi m / n (PolyBin3D/base.py:231-254)."""
    if kind == 'cic':
        s = np.sin(x) ** 2
        return np.sqrt(1. - 2. / 3. * s)
    if kind == 'tsc':
        s = np.sin(x) ** 2
        return np.sqrt(1. - s + 2. / 15 * s ** 2)
    if kind == 'pcs':
        s = np.sin(x) ** 2
        return np.sqrt(1. - 4. / 3. * s + 2. / 5. * s ** 2 - 4. / 315. * s ** 3)
    if kind == 'interlaced-cic':
        s = np.sinc(x / np.pi)
        v = np.ones_like(s)
        v[s != 0.] = s[s != 0.] ** 2
        return v
    if kind == 'interlaced-tsc':
        return np.sinc(x / np.pi) ** 3
    if kind == 'interlaced-pcs':
        return np.sinc(x / np.pi) ** 4
    raise ValueError("Unknown window function '%s' supplied!" % kind)


class PolyBin3D:
    """FP64 restatement of PolyBin3D.PolyBin3D (PolyBin3D/base.py:6-437), real_fft=True only."""

    def __init__(self, boxsize, gridsize, Pk=None, boxcenter=None, pixel_window='none',
                 sightline='global', nthreads=None):
        # base.py:31-60
        self.boxsize = np.asarray([boxsize] * 3 if np.ndim(boxsize) == 0 else boxsize, dtype=np.float64)
        self.boxcenter = (np.zeros(3) if boxcenter is None else
                          np.asarray([boxcenter] * 3 if np.ndim(boxcenter) == 0 else boxcenter, dtype=np.float64))
        self.gridsize = np.asarray([gridsize] * 3 if np.ndim(gridsize) == 0 else gridsize, dtype=int)
        assert len(self.boxsize) == 3 and len(self.gridsize) == 3 and len(self.boxcenter) == 3
        if self.gridsize[2] % 2 != 0:
            raise Exception("Last dimension of grid must be even for real-to-complex FFTs.")
        assert sightline in ['local', 'global']
        if sightline == 'global':
            assert np.abs(self.boxcenter).sum() == 0.
        else:
            assert np.abs(self.boxcenter).sum() != 0.
        self.sightline = sightline
        self.pixel_window = pixel_window
        self.nthreads = nthreads or os.cpu_count()
        # base.py:63-74
        self.volume = self.boxsize.prod()
        self.kF = 2. * np.pi / self.boxsize
        self.kNy = self.gridsize * self.kF * 0.5
        if Pk is None:
            k_tmp = np.arange(0., np.max(self.kNy) * 2, np.min(self.kF) / 10.)
            self.Pfid = np.asarray([k_tmp, 1. + 0. * k_tmp])
        else:
            assert len(Pk) == 2
            self.Pfid = np.asarray(Pk, dtype=np.float64)
        # base.py:77-80 (no half-cell offset)
        if sightline == 'local':
            r_arrs = [np.fft.fftshift(np.arange(-self.gridsize[i] // 2, self.gridsize[i] // 2))
                      * self.boxsize[i] / self.gridsize[i] + self.boxcenter[i] for i in range(3)]
            self.r_grids = np.meshgrid(*r_arrs, indexing='ij')
        # base.py:86-94
        k_arrs = [np.fft.fftshift(np.arange(-self.gridsize[i] // 2, self.gridsize[i] // 2)) * self.kF[i]
                  for i in range(2)]
        k_arrs += [np.arange(0, self.gridsize[2] // 2 + 1) * self.kF[2]]
        self.k_arrs = k_arrs
        self.k_grids = np.meshgrid(*k_arrs, indexing='ij')
        self.modk_grid = np.sqrt(self.k_grids[0] ** 2 + self.k_grids[1] ** 2 + self.k_grids[2] ** 2)
        self.degeneracy_factor = np.asarray(1. + (self.k_grids[2] > 0), dtype=np.float64)
        # base.py:103-108
        assert pixel_window in PIXEL_WINDOWS
        if pixel_window != 'none':
            w1 = [pixel_window_1d(pixel_window, np.pi / self.gridsize[i] * k_arrs[i] / self.kF[i])
                  for i in range(3)]
            self.pixel_window_grid = np.asarray(np.meshgrid(*w1, indexing='ij')).prod(axis=0)
        # base.py:111-116
        sv = np.asarray([0., 0., 1.]) if sightline == 'global' else self.boxcenter / np.sqrt(np.sum(self.boxcenter ** 2))
        self.sight_vector = sv
        self.muk_grid = np.zeros(self.modk_grid.shape)
        nz = self.modk_grid != 0
        self.muk_grid[nz] = np.sum(sv[:, None] * np.asarray(self.k_grids)[:, nz], axis=0) / self.modk_grid[nz]
        # base.py:122-131
        self.Pk0_grid = scipy.interpolate.interp1d(self.Pfid[0], self.Pfid[1], bounds_error=False,
                                                   fill_value=0.)(self.modk_grid)
        self.invPk0_grid = np.zeros(self.modk_grid.shape)
        good = (self.Pk0_grid != 0) & np.isfinite(self.Pk0_grid)
        self.invPk0_grid[good] = 1. / self.Pk0_grid[good]
        self.Pk0_grid[~np.isfinite(self.Pk0_grid)] = 0.
        self.invPk0_grid[self.modk_grid == 0] = 1.
        self.Pk0_grid[self.modk_grid == 0] = 1.
        self.n_FFTs_forward = 0
        self.n_FFTs_reverse = 0
        self.integrator = Integrator(self.modk_grid, self.degeneracy_factor, self.muk_grid)
        self.map_utils = MapUtils(self.modk_grid, self.muk_grid)

    def real_zeros(self, size=None):
        shape = list(self.gridsize) if size is None else (
            [size] if np.ndim(size) == 0 else list(size)) + list(self.gridsize)
        return np.zeros(shape)

    def to_fourier(self, m):
        """Unnormalised R2C (base.py:176-200)."""
        self.n_FFTs_forward += 1
        return scipy.fft.rfftn(np.asarray(m, dtype=np.float64), axes=(0, 1, 2), workers=self.nthreads)

    def to_real(self, m):
        """C2R with 1/N (base.py:202-229)."""
        self.n_FFTs_reverse += 1
        return scipy.fft.irfftn(m, s=tuple(self.gridsize), axes=(0, 1, 2), workers=self.nthreads)

    def generate_data(self, seed=None, Pk_input=(), output_type='real', include_pixel_window=True):
        """Gaussian random field with the reference's numpy stream (base.py:305-353)."""
        rng = np.random.RandomState(seed) if seed is not None else np.random
        if len(Pk_input) == 0:
            Pk_grid = self.Pk0_grid
        else:
            Pk_grid = scipy.interpolate.interp1d(Pk_input[0], Pk_input[1], bounds_error=False,
                                                 fill_value=0.)(self.modk_grid)
            if len(Pk_input) > 2:
                Pk_grid = Pk_grid + legendre(2, self.muk_grid) * scipy.interpolate.interp1d(
                    Pk_input[0], Pk_input[2], bounds_error=False, fill_value=0.)(self.modk_grid)
            if len(Pk_input) > 3:
                Pk_grid = Pk_grid + legendre(4, self.muk_grid) * scipy.interpolate.interp1d(
                    Pk_input[0], Pk_input[3], bounds_error=False, fill_value=0.)(self.modk_grid)
        n = self.gridsize
        r = rng.randn(*n) + 1.0j * rng.randn(*n)
        r[0, 0, 0] = 0.
        ii = (-np.arange(n[0])) % n[0]
        jj = (-np.arange(n[1])) % n[1]
        kk = (-np.arange(n[2])) % n[2]
        r = 0.5 * (r + np.conj(r[ii][:, jj][:, :, kk]))
        r = np.ascontiguousarray(r[:, :, :n[2] // 2 + 1])
        r = r * np.sqrt(Pk_grid) * (n.prod() / np.sqrt(self.volume))
        if self.pixel_window != 'none' and include_pixel_window:
            r = r * self.pixel_window_grid
        return r if output_type == 'fourier' else self.to_real(r)

    def applySinv_trivial(self, d, input_type='real', output_type='real'):
        """base.py:355-362."""
        if input_type == output_type:
            return d
        if input_type == 'fourier':
            return self.to_real(d)
        return self.to_fourier(d)

    def applyAinv(self, d, invPk0_grid=None, input_type='real', output_type='real'):
        """base.py:364-389."""
        invPk0 = self.invPk0_grid if invPk0_grid is None else invPk0_grid
        f = self.to_fourier(d) if input_type == 'real' else d.copy()
        s = f * invPk0
        if self.pixel_window != 'none':
            s = MapUtils.div_fourier(MapUtils.div_fourier(s, self.pixel_window_grid), self.pixel_window_grid)
        return self.to_real(s) if output_type == 'real' else s

    def apply_real_harmonics(self, a_fourier, Ylm_real, Ylm_fourier):
        """sum_m IFFT[a Y_lm(k)] Y_lm(x) (base.py:391-400)."""
        out = self.real_zeros()
        for m in range(len(Ylm_real)):
            out += self.to_real(a_fourier * Ylm_fourier[m]) * Ylm_real[m]
        return out

    def apply_fourier_harmonics(self, a_real, Ylm_real, Ylm_fourier):
        """sum_m FFT[a Y_lm(x)] Y_lm(k) (base.py:402-411)."""
        out = np.zeros(self.modk_grid.shape, dtype=np.complex128)
        for m in range(len(Ylm_real)):
            out += self.to_fourier(a_real * Ylm_real[m]) * Ylm_fourier[m]
        return out

    def apply_xi(self, input_fourier, input_real=None, Ylm_real=None, Ylm_fourier=None, Pk_grid=None,
                 output_type='real'):
        """The correlation function xi(x, y) applied symmetrically (base.py:413-437).  Pk_grid: {0: P(k, mu)}
        (global) or {0: P_0(k), 2: P_2(k), 4: P_4(k)} (local, multipoles beyond the monopole optional).
        With input_real None and local multipoles the reference fails; here it is the C2R of the input."""
        if self.sightline == 'global':
            P = input_fourier * Pk_grid[0]
            return P if output_type == 'fourier' else self.to_real(P)
        P = input_fourier * Pk_grid[0]
        ells = [l for l in Pk_grid if l != 0]
        if not ells:
            return P if output_type == 'fourier' else self.to_real(P)
        if input_real is None:
            input_real = self.to_real(input_fourier)
        P_real = self.real_zeros()
        for l in ells:
            P = P + 0.5 * self.apply_fourier_harmonics(input_real, Ylm_real[l], Ylm_fourier[l]) * Pk_grid[l]
            P_real = P_real + 0.5 * self.apply_real_harmonics(input_fourier * Pk_grid[l], Ylm_real[l], Ylm_fourier[l])
        if output_type == 'fourier':
            return P + self.to_fourier(P_real)
        return self.to_real(P) + P_real


# ----------------------------------------------------------------------------------------------
# Power spectrum (PolyBin3D/pspec.py)
# ----------------------------------------------------------------------------------------------

def derivative_coefficient(k_bins_theory):
    """Per-bin weights of the wide-angle k dP/dk term (pspec.py:369-379): the column sums of a
    central-difference operator on the bin edges, with one-sided end rows."""
    kb = np.asarray(k_bins_theory, dtype=np.float64)
    n = len(kb) - 1
    kfac = np.zeros(n)
    for ki in range(n):
        if 0 < ki - 1 < n - 1:
            kfac[ki] += kb[ki - 1] / (kb[ki] - kb[ki - 2])
        if 0 < ki + 1 < n - 1:
            kfac[ki] -= kb[ki + 1] / (kb[ki + 2] - kb[ki])
        kfac[ki] += kb[0] / (kb[1] - kb[0]) * ((ki == 1) - (ki == 0))
        kfac[ki] -= kb[-1] / (kb[-1] - kb[-2]) * ((ki == n - 1) - (ki == n - 2))
    return kfac


def covariance_filter(spec, m, input_type, Pk_grid, noise_mask, Ylm_real, Ylm_fourier):
    """S^-1 [P xi P^T + N] S^-dag of a map, the covariance counterpart of S^-1 P (pspec.py:278-289,
    312-325; bspec.py:507-519): mask_mean^2 xi + mask_mean for a constant mask (Fourier output), else
    the pointing matrix around xi plus the noise mask N, with xi = apply_xi."""
    b = spec.base
    if spec.const_mask:
        Sid = spec.applySinv_transpose(m, input_type=input_type, output_type='fourier')
        CSid = spec.mask_mean ** 2 * b.apply_xi(Sid, None, Ylm_real, Ylm_fourier, Pk_grid, 'fourier') \
            + spec.mask_mean * Sid
        return spec.applySinv(CSid, input_type='fourier', output_type='fourier')
    Sid = spec.applySinv_transpose(m, input_type=input_type, output_type='real')
    PdSid_real = spec.apply_pointing(Sid, transpose=True)
    CPdSid = b.apply_xi(b.to_fourier(PdSid_real), PdSid_real, Ylm_real, Ylm_fourier, Pk_grid, 'real')
    Cov = spec.apply_pointing(CPdSid) + Sid * noise_mask
    return spec.applySinv(Cov, input_type='real', output_type='fourier')


def cov_pk_grid(base, Pk_cov):
    """Gridded theory spectrum of the covariance (pspec.py:237-251; bspec.py:484-497): global
    {0: P_0 + L_2(mu) P_2 + L_4(mu) P_4}, local {0: P_0, 2: P_2, 4: P_4}, zero outside the table."""
    b = base
    interp = lambda col: scipy.interpolate.interp1d(Pk_cov[0], Pk_cov[col], bounds_error=False,  # noqa: E731
                                                    fill_value=0.)(b.modk_grid)
    if b.sightline == 'global':
        P = interp(1)
        for col, l in ((2, 2), (3, 4)):
            if len(Pk_cov) > col:
                P = P + legendre(l, b.muk_grid) * interp(col)
        return {0: P}
    out = {0: interp(1)}
    for col, l in ((2, 2), (3, 4)):
        if len(Pk_cov) > col:
            out[l] = interp(col)
    return out


class PSpec:
    """FP64 restatement of PolyBin3D.PSpec (PolyBin3D/pspec.py:10-784)."""

    def __init__(self, base, k_bins, lmax=2, mask=None, applySinv=None, odd_l=False,
                 applySinv_transpose=None, mask_shot=None, add_GIC=False, mask_IC=None, add_RIC=False,
                 radial_bins_RIC=()):
        # pspec.py:27-120
        self.base = base
        self.k_bins = np.asarray(k_bins, dtype=np.float64)
        self.lmax = lmax
        self.Nk = len(k_bins) - 1
        self.odd_l = odd_l
        self.Nl = lmax + 1 if odd_l else lmax // 2 + 1
        self.Nl_even = lmax // 2 + 1
        self.N_bins = self.Nk * self.Nl
        assert np.max(self.k_bins) < np.min(base.kNy), "k_max must be less than k_Nyquist!"
        assert np.min(self.k_bins) >= np.max(base.kF), "k_min must be at least the k_fundamental!"
        assert np.max(self.k_bins) <= base.Pfid[0][-1]
        assert np.min(self.k_bins) >= base.Pfid[0][0]
        if odd_l:
            assert lmax > 0 and base.sightline == 'local'
        assert lmax <= 4
        if mask is None:
            self.const_mask, self.mask_mean, self.sq_mask_mean = True, 1., 1.
        else:
            self.mask = np.asarray(mask, dtype=np.float64)
            if np.std(self.mask) < 1e-12:
                self.const_mask, self.mask_mean = True, np.mean(self.mask)
            else:
                self.const_mask = False
            self.sq_mask_mean = np.mean(self.mask ** 2)
        if mask_shot is not None:
            self.mask_shot = np.asarray(mask_shot, dtype=np.float64)
        _setup_ic(self, mask, add_GIC, mask_IC, add_RIC, radial_bins_RIC)
        if applySinv is None:
            self.applySinv = base.applySinv_trivial
            self.applySinv_transpose = base.applySinv_trivial if applySinv_transpose is None else applySinv_transpose
        else:
            self.applySinv = applySinv
            self.applySinv_transpose = applySinv if applySinv_transpose is None else applySinv_transpose
        if base.sightline == 'local' and lmax > 0:
            self.Ylm_real = real_harmonics(np.asarray(base.r_grids), lmax, odd_l)
            self.Ylm_fourier = real_harmonics(np.asarray(base.k_grids), lmax, odd_l)
        else:
            self.Ylm_real = self.Ylm_fourier = None

    def get_ks(self):
        """pspec.py:122-128."""
        return 0.5 * (self.k_bins[1:] + self.k_bins[:-1])

    def apply_pointing(self, m, transpose=False):
        """pspec.py:199-213."""
        return apply_pointing_ic(self, m, transpose)

    # --- numerator ------------------------------------------------------------------------
    def _compute_pk_numerator(self, data, filtering='ideal'):
        """pspec.py:613-686."""
        b = self.base
        data = np.asarray(data, dtype=np.float64)
        if b.pixel_window != 'none':
            df = MapUtils.div_fourier(b.to_fourier(data), b.pixel_window_grid)
            if filtering == 'ideal':
                S = df * b.invPk0_grid / np.sqrt(self.sq_mask_mean)
            else:
                S = self.applySinv(df, input_type='fourier', output_type='fourier')
        else:
            if filtering == 'ideal':
                S = b.to_fourier(data) * b.invPk0_grid / np.sqrt(self.sq_mask_mean)
            else:
                S = self.applySinv(data, input_type='real', output_type='fourier')
        if b.sightline == 'local' and self.lmax > 0:
            S_real = b.to_real(S)
        Nl, lf = (self.Nl, 2 - self.odd_l) if filtering == 'Sinv' else (self.Nl_even, 2)
        out = np.zeros(Nl * self.Nk)
        out[:self.Nk] = 0.5 * b.integrator.integrate(S, self.k_bins, 0)
        for li in range(1, Nl if self.lmax > 0 else 1):
            if b.sightline == 'global':
                out[li * self.Nk:(li + 1) * self.Nk] = 0.5 * b.integrator.integrate(S, self.k_bins, 2 * li)
            else:
                lm = b.apply_fourier_harmonics(S_real, self.Ylm_real[lf * li], self.Ylm_fourier[lf * li])
                if filtering == 'Sinv' and self.odd_l and li % 2 == 1:
                    lm = lm * (-1.0j)
                out[li * self.Nk:(li + 1) * self.Nk] = 0.5 * b.integrator.cross_integrate(S, lm, self.k_bins, 0)
        return out * b.volume / b.gridsize.prod() ** 2

    def Pk_numerator_ideal(self, data):
        return self._compute_pk_numerator(data, 'ideal')

    def Pk_numerator(self, data):
        return self._compute_pk_numerator(data, 'Sinv')

    # --- ideal Fisher ---------------------------------------------------------------------
    def compute_fisher_ideal(self, discreteness_correction=True):
        """pspec.py:702-752 (Pk_fid='default')."""
        b = self.base
        Nk, Nl = self.Nk, self.Nl_even
        fish = np.zeros((Nl * Nk, Nl * Nk))
        inv = b.invPk0_grid
        if discreteness_correction:
            for la in range(Nl):
                for lb in range(Nl):
                    fish[la * Nk:(la + 1) * Nk, lb * Nk:(lb + 1) * Nk] = np.diag(
                        0.5 * b.integrator.real_integrate_double(inv, 2 * la, 2 * lb, self.k_bins))
        else:
            d = 0.5 * b.integrator.real_integrate_double(inv, 0, 0, self.k_bins)
            for la in range(Nl):
                fish[la * Nk:(la + 1) * Nk, la * Nk:(la + 1) * Nk] = np.diag(d) / (4. * la + 1.)
        self.fish_ideal = fish
        self.inv_fish_ideal = np.linalg.inv(fish)
        return fish

    def Pk_ideal(self, data, fish_ideal=(), discreteness_correction=True):
        """pspec.py:754-784."""
        if len(fish_ideal) != 0:
            self.fish_ideal = np.asarray(fish_ideal)
            self.inv_fish_ideal = np.linalg.inv(self.fish_ideal)
        if not hasattr(self, 'inv_fish_ideal'):
            self.compute_fisher_ideal(discreteness_correction)
        out = self.inv_fish_ideal @ self.Pk_numerator_ideal(data)
        return {'p%d' % l: out[i * self.Nk:(i + 1) * self.Nk] for i, l in enumerate(range(0, self.lmax + 1, 2))}

    # --- MC Fisher ------------------------------------------------------------------------
    def compute_fisher_contribution(self, seed, verb=False):
        """pspec.py:141-147."""
        return self._compute_fisher(seed)

    def compute_covariance_contribution(self, seed, Pk_cov, verb=False):
        """One realization of the Gaussian covariance for a theory spectrum Pk_cov = [k, P_0 (, P_2, P_4)]
        without shot noise (pspec.py:149-163)."""
        if not self.const_mask:
            assert hasattr(self, 'mask_shot'), "Must supply mask_shot to compute shot-noise contribution"
        assert len(Pk_cov) != 0, "Must provide fiducial power spectrum if computing covariance!"
        Pk_cov = np.asarray(Pk_cov, dtype=np.float64)
        assert Pk_cov.ndim == 2 and len(Pk_cov) in [2, 3, 4], "Pk should contain k and P_0, (and optionally P_2, P_4) columns"
        assert (len(Pk_cov) - 2) * 2 <= self.lmax, "Can't use higher multipoles than lmax"
        return self._compute_fisher(seed, compute_cov=True, Pk_cov=Pk_cov)

    def compute_theory_contribution(self, seed, k_bins_theory, lmax_theory=None, include_wideangle=False,
                                    verb=False):
        """(Fisher contribution, rectangular theory binning matrix) from one realization (pspec.py:165-197)."""
        b = self.base
        k_bins_theory = np.asarray(k_bins_theory, dtype=np.float64)
        assert np.max(k_bins_theory) < np.min(b.kNy), "k_max must be less than k_Nyquist!"
        assert np.min(k_bins_theory) >= np.max(b.kF), "k_min must be at least the k_fundamental!"
        assert np.max(k_bins_theory) <= b.Pfid[0][-1], "Fiducial power spectrum should extend up to k_max!"
        assert np.min(k_bins_theory) >= b.Pfid[0][0], "Fiducial power spectrum should extend down to k_min!"
        if lmax_theory is None:
            lmax_theory = self.lmax
        # the reference looks up the data's multipoles among the theory legs (KeyError otherwise)
        assert lmax_theory >= self.lmax, "Can't use larger lmax for data than for theory!"
        if include_wideangle:
            assert b.sightline != 'global', "Wide-angle effects require local lines-of-sight!"
            assert not self.odd_l, "Cannot include wide-angle effects for odd l!"
            assert lmax_theory > 1, "There are no wide-angle effects for the monopole!"
            assert lmax_theory <= 4, "Higher-order effects not yet implemented!"
            if not hasattr(b, 'modr_grid'):
                b.modr_grid = np.sqrt(b.r_grids[0] ** 2. + b.r_grids[1] ** 2. + b.r_grids[2] ** 2.)
        elif not self.odd_l:
            assert lmax_theory % 2 == 0, "Must use even lmax if not including wide-angle effects!"
        return self._compute_fisher(seed, compute_theory=True, k_bins_theory=k_bins_theory,
                                    lmax_theory=lmax_theory, include_wideangle=include_wideangle)

    def _cov_pk_grid(self, Pk_cov):
        """pspec.py:237-251."""
        return cov_pk_grid(self.base, Pk_cov)

    def _compute_fisher(self, seed, compute_cov=False, Pk_cov=(), compute_theory=False, k_bins_theory=(),
                        lmax_theory=None, include_wideangle=False):
        """pspec.py:215-493: the MC Fisher contribution, its covariance counterpart (the filter
        S^-1 [P xi P^T + N] S^-dag in place of S^-1 P) or the theory binning matrix."""
        b = self.base
        Nk = self.Nk
        F = np.zeros((self.N_bins, self.N_bins))
        lf = 2 - self.odd_l
        if compute_theory:                                                          # 221-231
            odd_th = True if include_wideangle else self.odd_l
            Nl_th = lmax_theory + 1 if (odd_th and not include_wideangle) else lmax_theory // 2 + 1
            lf_th = 2 - odd_th
            Nk_th = len(k_bins_theory) - 1
            Bm = np.zeros((self.N_bins, Nl_th * Nk_th))
        if compute_cov:                                                             # 234-251
            assert not compute_theory and not include_wideangle
            Pk_grid = self._cov_pk_grid(Pk_cov)
        Ylm_real, Ylm_fourier = self.Ylm_real, self.Ylm_fourier
        if compute_theory and b.sightline == 'local':                               # 254-264
            if lmax_theory > self.lmax or odd_th != self.odd_l:
                Yr_th = real_harmonics(np.asarray(b.r_grids), lmax_theory, odd_th)
                Yf_th = real_harmonics(np.asarray(b.k_grids), lmax_theory, odd_th)
            else:
                Yr_th, Yf_th = Ylm_real, Ylm_fourier
        a = b.generate_data(seed=seed + int(1e7), output_type='fourier')            # 266
        if not self.const_mask:
            a_real = b.to_real(a)                                                   # 268

        def cov_filter(m, it):                                                      # 278-289, 312-325
            return covariance_filter(self, m, it, Pk_grid, getattr(self, 'mask_shot', None), Ylm_real, Ylm_fourier)

        def apply_filter(m):                                                        # 270-289
            if compute_cov:
                return cov_filter(m, 'fourier' if self.const_mask else 'real')
            if self.const_mask:
                return self.mask_mean * self.applySinv(m, input_type='fourier', output_type='fourier')
            return self.applySinv(self.apply_pointing(m), input_type='real', output_type='fourier')

        def apply_filter_dagger(m, real=False):                                     # 291-326
            it = 'real' if real else 'fourier'
            if compute_cov:
                return cov_filter(m, it)
            if self.const_mask:
                return self.mask_mean * self.applySinv_transpose(m, input_type=it, output_type='fourier')
            return b.to_fourier(self.apply_pointing(self.applySinv_transpose(m, input_type=it, output_type='real'),
                                                    transpose=True))

        Ainv = b.applyAinv(a, input_type='fourier', output_type='fourier')          # 328
        SinvP = apply_filter(a if self.const_mask else a_real)                     # 329-332
        if b.sightline == 'local' and self.lmax > 0:
            SinvP_real = b.to_real(SinvP)                                           # 335-336
        if (b.sightline == 'local' and self.lmax > 0) or (compute_theory and b.sightline == 'local' and lmax_theory > 0):
            Ainv_real = b.to_real(Ainv)                                             # 337-343
        leg_maps = {}
        if not compute_theory:
            if b.sightline == 'local':                                              # 347-352
                for li in range(1, self.Nl):
                    leg_maps[li * lf] = b.apply_fourier_harmonics(Ainv_real, Ylm_real[lf * li], Ylm_fourier[lf * li])
        elif b.sightline == 'local' and not include_wideangle:                      # 354-357
            for li in range(1, Nl_th):
                leg_maps[li * lf_th] = b.apply_fourier_harmonics(Ainv_real, Yr_th[li * lf_th], Yf_th[li * lf_th])
        elif b.sightline == 'local':                                                # 359-367
            inv_r_Ainv = Ainv_real / b.modr_grid
            for l in range(1, lmax_theory + 1):
                src = Ainv_real if l % 2 == 0 else inv_r_Ainv
                leg_maps[l] = b.apply_fourier_harmonics(src, Yr_th[l], Yf_th[l])
        if compute_theory and include_wideangle:
            kfac = derivative_coefficient(k_bins_theory)                            # 369-379
        scaling = 0.5 * b.volume / b.gridsize.prod() ** 2                          # 390

        def add_to_matrix(row, Q):                                                  # 381-436
            F[row, :Nk] += b.integrator.cross_integrate(Ainv, Q, self.k_bins, 0) * scaling
            if compute_theory:
                Bm[row, :Nk_th] += b.integrator.cross_integrate(Ainv, Q, k_bins_theory, 0) * scaling
            if b.sightline == 'global':
                for li in range(1, self.Nl):
                    F[row, li * Nk:(li + 1) * Nk] += b.integrator.cross_integrate(Ainv, Q, self.k_bins, 2 * li) * scaling
                if compute_theory:
                    for li in range(1, Nl_th):
                        Bm[row, li * Nk_th:(li + 1) * Nk_th] += b.integrator.cross_integrate(
                            Ainv, Q, k_bins_theory, 2 * li) * scaling
                return
            for li in range(1, self.Nl):
                Qx = 1.0j * Q if (self.odd_l and li % 2 == 1) else Q
                F[row, li * Nk:(li + 1) * Nk] += b.integrator.cross_integrate(leg_maps[li * lf], Qx, self.k_bins, 0) * scaling
            if not compute_theory:
                return
            if not include_wideangle:
                for li in range(1, Nl_th):
                    Qx = 1.0j * Q if (odd_th and li % 2 == 1) else Q
                    Bm[row, li * Nk_th:(li + 1) * Nk_th] += b.integrator.cross_integrate(
                        leg_maps[li * lf_th], Qx, k_bins_theory, 0) * scaling
                return
            wc = b.integrator.cross_integrate_wcoeff
            for l in range(1, lmax_theory + 1):
                if l % 2 == 0:
                    Bm[row, l // 2 * Nk_th:(l // 2 + 1) * Nk_th] += b.integrator.cross_integrate(
                        leg_maps[l], Q, k_bins_theory, 0) * scaling
                if l == 1:      # the l = 1 leg feeds l = 2
                    Bm[row, Nk_th:2 * Nk_th] += wc(3.0j / 5. * (3 + kfac), leg_maps[l], Q, k_bins_theory) * scaling
                if l == 3:      # the l = 3 leg feeds l = 2 and l = 4
                    Bm[row, Nk_th:2 * Nk_th] += wc(3.0j / 5. * (2 - kfac), leg_maps[l], Q, k_bins_theory) * scaling
                    if lmax_theory >= 4:
                        Bm[row, 2 * Nk_th:3 * Nk_th] += wc(10.0j / 9. * (5 + kfac), leg_maps[l], Q,
                                                           k_bins_theory) * scaling

        for ki in range(Nk):                                                        # 439-442
            km = b.map_utils.fourier_filter(SinvP, 0, self.k_bins[ki], self.k_bins[ki + 1])
            add_to_matrix(ki, apply_filter_dagger(km))
        for li in range(1, self.Nl):                                                # 444-488
            if b.sightline == 'global':
                for ki in range(Nk):
                    km = b.map_utils.fourier_filter(SinvP, li * 2, self.k_bins[ki], self.k_bins[ki + 1])
                    add_to_matrix(li * Nk + ki, apply_filter_dagger(km))
            else:
                Yr, Yf = Ylm_real[lf * li], Ylm_fourier[lf * li]
                leg_map = b.apply_fourier_harmonics(SinvP_real, Yr, Yf)
                for ki in range(Nk):
                    km = b.map_utils.fourier_filter(leg_map, 0, self.k_bins[ki], self.k_bins[ki + 1])
                    if li % 2 == 1 and self.odd_l:
                        km = km * (-1.0j)
                    Q = 0.5 * apply_filter_dagger(km)
                    real_map = b.real_zeros()
                    for m in range(len(Yr)):
                        k_map = b.map_utils.prod_fourier_filter(SinvP, Yf[m], self.k_bins[ki], self.k_bins[ki + 1])
                        if li % 2 == 1 and self.odd_l:
                            k_map = k_map * (-1.0j)
                        real_map += b.to_real(k_map) * Yr[m]
                    Q = Q + 0.5 * apply_filter_dagger(real_map, real=True)
                    add_to_matrix(li * Nk + ki, Q)
        return (F, Bm) if compute_theory else F

    def compute_fisher(self, N_it, verb=False):
        """pspec.py:553-569."""
        fish = np.zeros((self.N_bins, self.N_bins))
        for seed in range(N_it):
            fish += self.compute_fisher_contribution(seed) / N_it
        self.fish = fish
        self.inv_fish = np.linalg.inv(fish)
        return fish

    def compute_shot_contribution(self, seed):
        """pspec.py:495-551."""
        b = self.base
        shot = np.zeros(self.N_bins)
        PkA = [b.Pfid[0], 1. / b.Pfid[1]]
        PkA_grid = scipy.interpolate.interp1d(PkA[0], PkA[1], bounds_error=False,
                                              fill_value=(PkA[1][0], PkA[1][-1]))(b.modk_grid)
        invPkA = np.zeros_like(PkA_grid)
        invPkA[b.modk_grid != 0] = 1. / PkA_grid[b.modk_grid != 0]
        a = b.generate_data(seed=seed + int(1e7), Pk_input=PkA, output_type='fourier')
        if self.const_mask:
            AiA = b.applyAinv(a, invPk0_grid=invPkA, input_type='fourier', output_type='fourier')
            SNA = self.mask_mean * self.applySinv(AiA, input_type='fourier', output_type='fourier')
        else:
            AiA = b.applyAinv(a, invPk0_grid=invPkA, input_type='fourier', output_type='real')
            SNA = self.applySinv(AiA * self.mask_shot, input_type='real', output_type='fourier')
        Sa = self.applySinv(a, input_type='fourier', output_type='fourier')
        if self.lmax > 0 and b.sightline == 'local':
            Sa_real = b.to_real(Sa)
        shot[:self.Nk] = 0.5 * b.integrator.cross_integrate(Sa, SNA, self.k_bins, 0)
        lf = 2 - self.odd_l
        for li in range(1, self.Nl):
            if b.sightline == 'global':
                shot[li * self.Nk:(li + 1) * self.Nk] = 0.5 * b.integrator.cross_integrate(Sa, SNA, self.k_bins, 2 * li)
            else:
                lm = b.apply_fourier_harmonics(Sa_real, self.Ylm_real[lf * li], self.Ylm_fourier[lf * li])
                if self.odd_l and li % 2 == 1:
                    lm = lm * (-1.0j)
                shot[li * self.Nk:(li + 1) * self.Nk] = 0.5 * b.integrator.cross_integrate(lm, SNA, self.k_bins, 0)
        return shot * b.volume / b.gridsize.prod() ** 2

    def Pk_unwindowed(self, data, fish=(), shot_num=(), subtract_shotnoise=False):
        """pspec.py:571-611."""
        if len(fish) != 0:
            self.fish = np.asarray(fish)
            self.inv_fish = np.linalg.inv(self.fish)
        if len(shot_num) != 0 or subtract_shotnoise:
            self.shot_num = np.asarray(shot_num)
            self.shot = self.inv_fish @ self.shot_num
        if not hasattr(self, 'inv_fish'):
            raise Exception("Need to compute Fisher matrix first!")
        out = self.inv_fish @ self.Pk_numerator(data)
        if subtract_shotnoise:
            out = out - self.shot
        res, idx = {}, 0
        for l in range(0, self.lmax + 1, 2 - self.odd_l):
            res['p%d' % l] = out[idx:idx + self.Nk]
            idx += self.Nk
        return res


# ----------------------------------------------------------------------------------------------
# Bispectrum (PolyBin3D/bspec.py)
# ----------------------------------------------------------------------------------------------

def triangle_bins(k_bins, k_bins_squeeze, lmax, include_partial_triangles=False):
    """all_bins = [(b1<=b2<=b3, l)], l outer (PolyBin3D/bspec.py:67-75, _check_bin 171-199)."""
    kb, ks = np.asarray(k_bins, dtype=np.float64), np.asarray(k_bins_squeeze, dtype=np.float64)
    Nk, Nks = len(kb) - 1, len(ks) - 1

    def check(b1, b2, b3):
        if include_partial_triangles:
            hi = ks[b1 + 1] + ks[b2 + 1]
            if b1 > b2:
                lo = ks[b1] - ks[b2 + 1]
            elif b1 == b2:
                lo = 0.
            else:
                lo = ks[b2] - ks[b1 + 1]
            return not (ks[b3] >= hi - 1e-10 or ks[b3 + 1] <= lo + 1e-10)
        k1 = 0.5 * (kb[b1] + kb[b1 + 1])
        k2 = 0.5 * (ks[b2] + ks[b2 + 1])
        k3 = 0.5 * (ks[b3] + ks[b3 + 1])
        return not (k3 < abs(k1 - k2) or k3 > k1 + k2)

    out = []
    for l in range(0, lmax + 1, 2):
        for b1 in range(Nk):
            for b2 in range(b1, Nks):
                for b3 in range(b2, Nks):
                    if check(b1, b2, b3):
                        out.append((b1, b2, b3, l))
    return np.asarray(out, dtype=np.int64).reshape(-1, 4)


class BSpec:
    """FP64 restatement of PolyBin3D.BSpec (PolyBin3D/bspec.py:12-1119), hot-path members only."""

    def __init__(self, base, k_bins, lmax=0, mask=None, applySinv=None, k_bins_squeeze=None,
                 include_partial_triangles=False, Pk_fid=None, applySinv_transpose=None, mask_shot=None,
                 add_GIC=False, mask_IC=None, add_RIC=False, radial_bins_RIC=()):
        b = self.base = base
        self.k_bins = np.asarray(k_bins, dtype=np.float64)
        self.lmax = lmax
        self.Nk = len(k_bins) - 1
        self.Nl = lmax // 2 + 1
        self.include_partial_triangles = include_partial_triangles
        assert lmax % 2 == 0 and lmax <= 4
        assert np.max(self.k_bins) < np.min(b.kNy) and np.min(self.k_bins) >= np.max(b.kF)
        if k_bins_squeeze is not None:
            self.k_bins_squeeze = np.asarray(k_bins_squeeze, dtype=np.float64)
            assert (self.k_bins_squeeze[:self.Nk + 1] == self.k_bins).all()
            assert np.max(self.k_bins_squeeze) < np.min(b.kNy)
        else:
            self.k_bins_squeeze = self.k_bins.copy()
        self.Nk_squeeze = len(self.k_bins_squeeze) - 1
        self.all_bins = triangle_bins(self.k_bins, self.k_bins_squeeze, lmax, include_partial_triangles)
        self.N_bins = len(self.all_bins)
        self.N3 = self.N_bins // self.Nl
        if mask is None:
            self.const_mask, self.mask_mean, self.cube_mask_mean = True, 1., 1.
        else:
            self.mask = np.asarray(mask, dtype=np.float64)
            if np.std(self.mask) < 1e-12:
                self.const_mask, self.mask_mean = True, np.mean(self.mask)
            else:
                self.const_mask = False
            self.cube_mask_mean = np.mean(self.mask ** 3)
        if applySinv is None:
            self.applySinv = b.applySinv_trivial
            self.applySinv_transpose = b.applySinv_trivial if applySinv_transpose is None else applySinv_transpose
        else:
            self.applySinv = applySinv
            self.applySinv_transpose = applySinv if applySinv_transpose is None else applySinv_transpose
        # bspec.py:138-156
        if isinstance(Pk_fid, str) and Pk_fid == 'default':
            self.invPk0 = b.invPk0_grid
        else:
            if Pk_fid is None:
                k_tmp = np.arange(np.min(b.kF) / 10., np.max(b.kNy) * 2, np.min(b.kF) / 10.)
                Pfid = np.asarray([k_tmp, 1. + 0. * k_tmp])
            else:
                Pfid = np.asarray(Pk_fid, dtype=np.float64)
            P = scipy.interpolate.interp1d(Pfid[0], Pfid[1], bounds_error=False, fill_value=0.)(b.modk_grid)
            self.invPk0 = np.zeros(P.shape)
            self.invPk0[P != 0] = 1. / P[P != 0]
        # bspec.py:158-169
        self.Ylm_real = real_harmonics(np.asarray(b.r_grids), lmax, False) if (b.sightline == 'local' and lmax > 0) else None
        self.Ylm_fourier = real_harmonics(np.asarray(b.k_grids), lmax, False) if lmax > 0 else None
        if mask_shot is not None:  # bspec.py:113-114
            self.mask_shot = np.asarray(mask_shot, dtype=np.float64)
        _setup_ic(self, mask, add_GIC, mask_IC, add_RIC, radial_bins_RIC)

    def get_ks(self):
        """bspec.py:201-220."""
        kc = 0.5 * (self.k_bins[1:] + self.k_bins[:-1])
        ksc = 0.5 * (self.k_bins_squeeze[1:] + self.k_bins_squeeze[:-1])
        t = self.all_bins[:self.N3]
        return np.asarray([kc[t[:, 0]], ksc[t[:, 1]], ksc[t[:, 2]]])

    def _filt(self, m, l, b):
        ks = self.k_bins_squeeze
        return self.base.map_utils.fourier_filter(m, l, ks[b], ks[b + 1])

    def _compute_symmetry_factor(self):
        """bspec.py:222-279."""
        b = self.base
        t = self.all_bins[:self.N3]
        sym = []
        for (b1, b2, b3, _) in t:
            sym.append(6. if (b1 == b2 == b3) else (2. if (b1 == b2 or b1 == b3 or b2 == b3) else 1.))
        if self.lmax > 0:
            inv = self.invPk0.astype(np.complex128)
            bins = np.asarray([b.to_real(self._filt(inv, 0, q)) for q in range(self.Nk_squeeze)])
            for l in range(2, self.lmax + 1, 2):
                b23 = b.real_zeros(self.Nk_squeeze)
                ks = self.k_bins_squeeze
                for q in range(self.Nk_squeeze):
                    for m in range(len(self.Ylm_fourier[l])):
                        b23[q] += b.to_real(b.map_utils.prod_fourier_filter(inv, self.Ylm_fourier[l][m], ks[q], ks[q + 1])) ** 2
                for (b1, b2, b3, _) in t:
                    if b1 != b2 and b2 != b3:
                        sym.append(1.)
                    elif b1 == b2 and b2 != b3:
                        sym.append(2.)
                    elif b1 != b2 and b2 == b3:
                        sym.append(1. + sum_pair(bins[b1], b23[b3]) / sum_triplet(bins[b1], bins[b3], bins[b3]))
                    else:
                        sym.append(2. * (1. + 2. * sum_pair(bins[b3], b23[b3]) / sum_triplet(bins[b3], bins[b3], bins[b3])))
        self.sym_factor = np.asarray(sym)
        assert len(self.sym_factor) == self.N_bins

    def _g_maps(self, S_fourier):
        """g_{l,b}(x) maps (bspec.py:931-956)."""
        b = self.base
        g = b.real_zeros((self.Nl, self.Nk_squeeze))
        if b.sightline == 'local' and self.lmax > 0:
            S_real = b.to_real(S_fourier)
        for l in range(0, self.lmax + 1, 2):
            if l == 0:
                for q in range(self.Nk_squeeze):
                    g[0, q] = b.to_real(self._filt(S_fourier, 0, q))
            elif b.sightline == 'global':
                for q in range(self.Nk_squeeze):
                    g[l // 2, q] = b.to_real(self._filt(S_fourier, l, q))
            else:
                leg = b.apply_fourier_harmonics(S_real, self.Ylm_real[l], self.Ylm_fourier[l])
                for q in range(self.Nk_squeeze):
                    g[l // 2, q] = b.to_real(self._filt(leg, 0, q))
        return g

    def _compute_bk_numerator(self, data, include_linear_term=True, filtering='ideal', return_scale=False):
        """bspec.py:895-983.  return_scale (for tests): also the same sums over |g| maps, the
        magnitude that FP32 rounding acts on (the cubic and linear terms cancel by design, and each
        grid sum can cancel between voxels)."""
        b = self.base
        if include_linear_term and not hasattr(self, 'preload'):
            raise Exception("Need to generate or specify bias simulations!")
        data = np.asarray(data, dtype=np.float64)
        if not hasattr(self, 'sym_factor'):
            self._compute_symmetry_factor()
        if b.pixel_window != 'none':
            df = MapUtils.div_fourier(b.to_fourier(data), b.pixel_window_grid)
            S = (df * self.invPk0 / self.cube_mask_mean ** (1. / 3.) if filtering == 'ideal'
                 else self.applySinv(df, input_type='fourier', output_type='fourier'))
        else:
            S = (b.to_fourier(data) * self.invPk0 / self.cube_mask_mean ** (1. / 3.) if filtering == 'ideal'
                 else self.applySinv(data, input_type='real', output_type='fourier'))
        g = self._g_maps(S)
        B3 = assemble_b3(g, self.all_bins)
        B1 = np.zeros(self.N_bins)
        if include_linear_term:
            for ii in range(self.N_it):
                B1 += -1. / self.N_it * assemble_b1(g, self.g_bl_maps[ii], self.all_bins)
        norm = b.volume / b.gridsize.prod() / self.sym_factor
        if not return_scale:
            return (B3 + B1) * norm
        ga = np.abs(g)
        mag = assemble_b3(ga, self.all_bins)
        if include_linear_term:
            for ii in range(self.N_it):
                mag += 1. / self.N_it * assemble_b1(ga, np.abs(self.g_bl_maps[ii]), self.all_bins)
        return (B3 + B1) * norm, mag * np.abs(norm)

    def Bk_numerator_ideal(self, data):
        return self._compute_bk_numerator(data, include_linear_term=False, filtering='ideal')

    def compute_shot_contribution(self, seed, data=None, data2=None, cubic_only=False, return_scale=False):
        """bspec.py:695-861: shot-noise numerator from a GRF pair drawn with 1/P_fid (seeds seed+1e7,
        seed+2e7) and weighted by A^-1 = P_fid.  Returns quadratic - 2 cubic, where the quadratic
        term also uses the data's g-maps and data2 (the doubly weighted data, default data); or the
        cubic term alone.  return_scale (for tests): also the same sums over |g| maps and |terms|,
        the magnitude that FP32 rounding acts on."""
        b = self.base
        if not cubic_only:
            assert data is not None
            data = np.asarray(data, dtype=np.float64)
            data2 = data if data2 is None else np.asarray(data2, dtype=np.float64)
        assert hasattr(self, 'mask_shot')
        if not hasattr(self, 'sym_factor'):
            self._compute_symmetry_factor()
        if not cubic_only:
            if b.pixel_window != 'none':  # bspec.py:739-744
                df = MapUtils.div_fourier(b.to_fourier(data), b.pixel_window_grid)
                Sd = self.applySinv(df, input_type='fourier', output_type='fourier')
            else:
                Sd = self.applySinv(data, input_type='real', output_type='fourier')
            g = self._g_maps(Sd)
        PkA = [b.Pfid[0], 1. / b.Pfid[1]]  # bspec.py:777-779
        a1 = b.generate_data(seed=seed + int(1e7), Pk_input=PkA, output_type='fourier')
        a2 = b.generate_data(seed=seed + int(2e7), Pk_input=PkA, output_type='fourier')
        A1 = b.applyAinv(a1, invPk0_grid=b.Pk0_grid, input_type='fourier', output_type='real')
        A2 = b.applyAinv(a2, invPk0_grid=b.Pk0_grid, input_type='fourier', output_type='real')
        gC = self._g_maps(self.applySinv(A1 * A2 * self.mask_shot, input_type='real', output_type='fourier'))
        if not cubic_only:
            gD = self._g_maps(self.applySinv(A1 * data2, input_type='real', output_type='fourier'))
            gE = self._g_maps(self.applySinv(A2 * data2, input_type='real', output_type='fourier'))
        gA = self._g_maps(self.applySinv(a1, input_type='fourier', output_type='fourier'))
        gB = self._g_maps(self.applySinv(a2, input_type='fourier', output_type='fourier'))
        norm = b.volume / b.gridsize.prod() / self.sym_factor
        ab = self.all_bins
        cubic = assemble_bshot(gA, gB, gC, ab) / 6. * norm  # bspec.py:846-847
        scale = np.abs(assemble_bshot(np.abs(gA), np.abs(gB), np.abs(gC), ab) / 6. * norm)
        if cubic_only:
            return (cubic, scale) if return_scale else cubic
        quad = 0.5 * assemble_bshot(g, gA, gD, ab) + 0.5 * assemble_bshot(g, gB, gE, ab)
        quad *= 0.5 * norm  # bspec.py:851-854
        scale = 2. * scale + np.abs(0.25 * norm * (assemble_bshot(np.abs(g), np.abs(gA), np.abs(gD), ab)
                                                   + assemble_bshot(np.abs(g), np.abs(gB), np.abs(gE), ab)))
        return (quad - 2. * cubic, scale) if return_scale else quad - 2. * cubic

    def Bk_numerator(self, data, include_linear_term=True):
        return self._compute_bk_numerator(data, include_linear_term=include_linear_term, filtering='Sinv')

    def compute_fisher_ideal(self, discreteness_correction=True):
        """bspec.py:999-1083."""
        b = self.base
        if not hasattr(self, 'sym_factor'):
            self._compute_symmetry_factor()
        inv = self.invPk0.astype(np.complex128)
        ks = self.k_bins_squeeze
        N3 = self.N3
        if discreteness_correction:
            fish = np.zeros((self.N_bins, self.N_bins))
            bins_l = np.asarray([[b.to_real(self._filt(inv, l, q)) for q in range(self.Nk_squeeze)]
                                 for l in range(0, self.lmax + 1, 2)])
            for l in range(0, self.lmax + 1, 2):
                for lp in range(l, self.lmax + 1, 2):
                    bllp = np.asarray([b.to_real(b.map_utils.fourier_filter_ll(inv, l, lp, ks[q], ks[q + 1]))
                                       for q in range(self.Nk_squeeze)])
                    for i in range(N3):
                        b1, b2, b3 = self.all_bins[i, :3]
                        r, c = l // 2 * N3 + i, lp // 2 * N3 + i
                        if b1 != b2 and b2 != b3:
                            fish[r, c] = sum_triplet(bins_l[0][b1], bins_l[0][b2], bllp[b3])
                        elif b1 == b2 and b2 != b3:
                            fish[r, c] = 2. * sum_triplet(bins_l[0][b1], bins_l[0][b2], bllp[b3])
                        elif b1 != b2 and b2 == b3:
                            fish[r, c] = (sum_triplet(bins_l[0][b1], bins_l[l // 2][b2], bins_l[lp // 2][b3])
                                          + sum_triplet(bins_l[0][b1], bins_l[0][b2], bllp[b3]))
                        else:
                            fish[r, c] = (4. * sum_triplet(bins_l[0][b1], bins_l[l // 2][b2], bins_l[lp // 2][b3])
                                          + 2. * sum_triplet(bins_l[0][b1], bins_l[0][b2], bllp[b3]))
            fish = fish + fish.T - np.diag(np.diag(fish))
            fish = fish / np.outer(self.sym_factor, self.sym_factor) * b.gridsize.prod() ** 2 / b.volume
        else:
            bins = np.asarray([b.to_real(self._filt(inv, 0, q)) for q in range(self.Nk_squeeze)])
            Ntri = np.asarray([sum_triplet(bins[b1], bins[b2], bins[b3]) for (b1, b2, b3, _) in self.all_bins[:N3]])
            d = np.concatenate([Ntri / (2. * l + 1.) for l in range(0, self.lmax + 1, 2)])
            fish = np.diag(d / self.sym_factor) * b.gridsize.prod() ** 2 / b.volume
        self.fish_ideal = fish
        self.inv_fish_ideal = np.linalg.inv(fish)
        self.discreteness_correction = discreteness_correction
        return fish

    def Bk_ideal(self, data, fish_ideal=(), discreteness_correction=True):
        """bspec.py:1085-1119."""
        if len(fish_ideal) != 0:
            self.fish_ideal = np.asarray(fish_ideal)
            self.inv_fish_ideal = np.linalg.inv(self.fish_ideal)
            self.discreteness_correction = discreteness_correction
        if not hasattr(self, 'inv_fish_ideal') or self.discreteness_correction != discreteness_correction:
            self.compute_fisher_ideal(discreteness_correction)
        out = self.inv_fish_ideal @ self.Bk_numerator_ideal(data)
        return {'b%d' % l: out[i * self.N3:(i + 1) * self.N3] for i, l in enumerate(range(0, self.lmax + 1, 2))}

    # --- linear term simulations ------------------------------------------------------------
    def apply_pointing(self, m, transpose=False):
        """bspec.py:412-426."""
        return apply_pointing_ic(self, m, transpose)

    def _process_sim(self, sim, input_type='real'):
        """bspec.py:281-326."""
        b = self.base
        if b.pixel_window != 'none':
            sf = b.to_fourier(sim) if input_type == 'real' else sim.copy()
            sf = sf / b.pixel_window_grid
            S = self.applySinv(sf, input_type='fourier', output_type='fourier')
        else:
            S = self.applySinv(sim, input_type=input_type, output_type='fourier')
        return self._g_maps(S)

    def generate_sims(self, N_sim, Pk_input=()):
        """bspec.py:362-410 (preload=True)."""
        b = self.base
        self.N_it = N_sim
        self.preload = True
        maps = []
        for ii in range(N_sim):
            if self.const_mask:
                sf = self.mask_mean * b.generate_data(int(1e5) + ii, Pk_input=Pk_input, output_type='fourier')
                maps.append(self._process_sim(sf, input_type='fourier'))
            else:
                sr = self.apply_pointing(b.generate_data(int(1e5) + ii, Pk_input=Pk_input, output_type='real',
                                                         include_pixel_window=False))
                if b.pixel_window != 'none':
                    maps.append(self._process_sim(b.to_fourier(sr) * b.pixel_window_grid, input_type='fourier'))
                else:
                    maps.append(self._process_sim(sr, input_type='real'))
        self.g_bl_maps = np.asarray(maps)

    # --- MC Fisher ------------------------------------------------------------------------
    def compute_fisher_contribution(self, seed, verb=False):
        """bspec.py:441-445."""
        return self._compute_fisher(seed)

    def compute_covariance_contribution(self, seed, Pk_cov, verb=False, mask_shot_2pt=None):
        """One realization of the Gaussian covariance for Pk_cov = [k, P_0 (, P_2, P_4)] without shot
        noise (bspec.py:447-471); mask_shot_2pt defaults to mask_shot.  (The reference fails for a
        constant mask without mask_shot_2pt, which the constant-mask filter does not use.)"""
        b = self.base
        if not self.const_mask and mask_shot_2pt is None:
            assert getattr(self, 'mask_shot', None) is not None, \
                "Must supply a shot-noise mask to compute the shot-noise contribution!"
            noise = self.mask_shot
        else:
            noise = None if mask_shot_2pt is None else np.asarray(mask_shot_2pt, dtype=np.float64)
        assert len(Pk_cov) != 0, "Must provide fiducial power spectrum if computing covariance!"
        Pk_cov = np.asarray(Pk_cov, dtype=np.float64)
        assert Pk_cov.ndim == 2 and len(Pk_cov) in [2, 3, 4], "Pk should contain k and P_0 (and optionally P_2, P_4) columns"
        Yr, Yf = self.Ylm_real, self.Ylm_fourier
        lcov = (len(Pk_cov) - 2) * 2
        if b.sightline == 'local' and lcov > self.lmax:                            # 461-467
            Yr = real_harmonics(np.asarray(b.r_grids), lcov, False)
            Yf = real_harmonics(np.asarray(b.k_grids), lcov, False)
        return self._compute_fisher(seed, cov=(cov_pk_grid(b, Pk_cov), noise, Yr, Yf))

    def _compute_fisher(self, seed, cov=None):
        """bspec.py:473-676; cov = (gridded spectrum, noise mask, Ylm tables) selects the covariance
        filter S^-1 [P xi P^T + N] S^-dag in place of S^-1 P."""
        b = self.base
        if not hasattr(self, 'sym_factor'):
            self._compute_symmetry_factor()
        fish = np.zeros((self.N_bins, self.N_bins))
        ks = self.k_bins_squeeze
        ab = self.all_bins

        def apply_filter(m):                                                       # 499-519
            it = 'fourier' if self.const_mask else 'real'
            if cov is not None:
                return covariance_filter(self, m, it, cov[0], cov[1], cov[2], cov[3])
            if self.const_mask:
                return self.mask_mean * self.applySinv(m, input_type='fourier', output_type='fourier')
            return self.applySinv(self.apply_pointing(m), input_type='real', output_type='fourier')

        Q_Ainv = np.zeros((self.N_bins,) + b.modk_grid.shape, dtype=np.complex128)   # 643

        def compute_Q(weighting, af1, ar1, af2, ar2):                              # 521-640
            if weighting == 'Sinv':
                w1 = apply_filter(af1 if self.const_mask else ar1)
                w2 = apply_filter(af2 if self.const_mask else ar2)
            else:
                w1 = b.applyAinv(af1, input_type='fourier', output_type='fourier')
                w2 = b.applyAinv(af2, input_type='fourier', output_type='fourier')
            local = b.sightline == 'local' and self.lmax > 0
            if local:
                wr1, wr2 = b.to_real(w1), b.to_real(w2)
            g1 = np.asarray([b.to_real(self._filt(w1, 0, q)) for q in range(self.Nk_squeeze)])
            g2 = np.asarray([b.to_real(self._filt(w2, 0, q)) for q in range(self.Nk_squeeze)])
            if local:
                leg1 = [b.apply_fourier_harmonics(wr1, self.Ylm_real[l], self.Ylm_fourier[l]) for l in range(2, self.lmax + 1, 2)]
                leg2 = [b.apply_fourier_harmonics(wr2, self.Ylm_real[l], self.Ylm_fourier[l]) for l in range(2, self.lmax + 1, 2)]
            for B in range(self.Nk_squeeze):
                if self.lmax > 0:
                    gB1, gB2 = [], []
                    for l in range(2, self.lmax + 1, 2):
                        if b.sightline == 'global':
                            gB1.append(b.to_real(self._filt(w1, l, B)))
                            gB2.append(b.to_real(self._filt(w2, l, B)))
                        else:
                            gB1.append(b.to_real(self._filt(leg1[l // 2 - 1], 0, B)))
                            gB2.append(b.to_real(self._filt(leg2[l // 2 - 1], 0, B)))
                for A in range(B + 1):
                    if weighting == 'Sinv':
                        ind = np.where((ab[:, 2] == B) & (ab[:, 0] == A))[0]
                        if len(ind) == 0:
                            continue
                        for l in range(0, self.lmax + 1, 2):
                            if l == 0:
                                ft = b.to_fourier(g1[A] * g1[B] - g2[A] * g2[B])
                            else:
                                ft = b.to_fourier(g1[A] * gB1[l // 2 - 1] - g2[A] * gB2[l // 2 - 1])
                            for ii in ind[ab[ind, 3] == l]:
                                q2 = ab[ii, 1]
                                fish[ii] = b.integrator.integrate_row(ft, Q_Ainv, ks[q2], ks[q2 + 1])
                    else:
                        i1 = np.where((ab[:, 0] == A) & (ab[:, 1] == B))[0]
                        i2 = np.where((ab[:, 1] == A) & (ab[:, 2] == B))[0]
                        i3 = np.where((ab[:, 2] == B) & (ab[:, 0] == A))[0]
                        if len(i1) + len(i2) + len(i3) == 0:
                            continue
                        ft = [b.to_fourier(g1[A] * g1[B] - g2[A] * g2[B])]
                        for l in range(2, self.lmax + 1, 2):
                            ft.append(b.to_fourier(g1[A] * gB1[l // 2 - 1] - g2[A] * gB2[l // 2 - 1]))

                        def add_Q(ci, inds, largest):
                            for ii in inds:
                                C = ab[ii, ci]
                                l = ab[ii, 3]
                                if l == 0 or not largest:
                                    Q_Ainv[ii] += self._filt(ft[l // 2], 0, C)
                                elif b.sightline == 'global':
                                    Q_Ainv[ii] += self._filt(ft[0], l, C)
                                else:
                                    fC = self._filt(ft[0], 0, C)
                                    Q_Ainv[ii] += b.to_fourier(b.apply_real_harmonics(fC, self.Ylm_real[l], self.Ylm_fourier[l]))
                        add_Q(2, i1, True)
                        add_Q(0, i2, False)
                        add_Q(1, i3, False)

        a1 = b.generate_data(seed=seed + int(1e7), output_type='fourier')            # 648-649
        a2 = b.generate_data(seed=seed + int(2e7), output_type='fourier')
        ar1 = ar2 = None
        if not self.const_mask:
            ar1, ar2 = b.to_real(a1), b.to_real(a2)
        compute_Q('Ainv', a1, ar1, a2, ar2)                                         # 659
        for i in range(self.N_bins):                                                # 662-666
            Q_Ainv[i] = apply_filter(Q_Ainv[i]) if self.const_mask else apply_filter(b.to_real(Q_Ainv[i]))
        compute_Q('Sinv', a1, ar1, a2, ar2)                                         # 671
        return 0.5 * fish * b.volume / b.gridsize.prod() ** 2 / np.outer(self.sym_factor, self.sym_factor)

    def fisher_subblock(self, seed, rows, cols, return_scale=False):
        """fish[rows][:, cols] of _compute_fisher(seed) (bspec.py:473-676, trivial or user S^-1, no
        covariance), computing only the Q[A^-1 a] maps of the listed columns and the S^-1-weighted
        left maps of the listed rows, so that it runs on grids where the full [N_bins, H] Q-map array
        cannot be held.  Per triplet (b1, b2, b3, l) the Q map is the sum of the three leg terms the
        reference places through add_Q_element (bspec.py:615-640): Theta_b3 W_l FT[g_0,b1 g_0,b2]
        (W_l = L_l(mu), or the Y_lm sum of apply_real_harmonics for the local sightline),
        Theta_b1 FT[g_0,b2 g_l,b3] and Theta_b2 FT[g_0,b1 g_l,b3], each with the GRF-pair difference
        (prod_real_diff), then S^-1 P (bspec.py:662-666); a row is integrate_row of
        FT[g_0,b1 g_l,b3] (S^-1 P weighting) over bin b2 (bspec.py:579-597).  return_scale (for tests):
        also the same sums over |FT| and |Q|, sum_{k in b2} degk |m(k)| |Q_j(k)| with the normalisation,
        the magnitude that FP32 rounding of the maps acts on (an entry that cancels to far below it is
        ill-conditioned in FP32)."""
        b = self.base
        if not hasattr(self, 'sym_factor'):
            self._compute_symmetry_factor()
        ab = self.all_bins
        ks = self.k_bins_squeeze
        rows, cols = np.asarray(rows, dtype=int), np.asarray(cols, dtype=int)
        local = b.sightline == 'local' and self.lmax > 0

        def apply_filter(m):                                                       # 499-519
            if self.const_mask:
                return self.mask_mean * self.applySinv(m, input_type='fourier', output_type='fourier')
            return self.applySinv(self.apply_pointing(m), input_type='real', output_type='fourier')

        a1 = b.generate_data(seed=seed + int(1e7), output_type='fourier')            # 648-649
        a2 = b.generate_data(seed=seed + int(2e7), output_type='fourier')

        def gmaps(w1, w2, need):
            """{(l, q): (g1, g2)} for the (multipole, bin) pairs in `need` (bspec.py:546-576)."""
            out, legs = {}, {}
            for l, q in sorted(need):
                if l == 0 or b.sightline == 'global':
                    out[l, q] = tuple(b.to_real(self._filt(w, l, q)) for w in (w1, w2))
                else:
                    if l not in legs:
                        legs[l] = tuple(b.apply_fourier_harmonics(b.to_real(w), self.Ylm_real[l], self.Ylm_fourier[l])
                                        for w in (w1, w2))
                    out[l, q] = tuple(b.to_real(self._filt(lg, 0, q)) for lg in legs[l])
            return out

        def diff(g, A, B):
            return g[A][0] * g[B][0] - g[A][1] * g[B][1]

        # columns: Q[A^-1 a] (compute_Q('Ainv'), bspec.py:659)
        wA = [b.applyAinv(a, input_type='fourier', output_type='fourier') for a in (a1, a2)]
        needA = set()
        for b1, b2, b3, l in ab[cols]:
            needA |= {(0, b1), (0, b2), (0, b3), (l, b3)}
        gA = gmaps(wA[0], wA[1], needA)
        Q = np.zeros((len(cols),) + b.modk_grid.shape, dtype=np.complex128)
        for c, (b1, b2, b3, l) in enumerate(ab[cols]):
            ft0 = b.to_fourier(diff(gA, (0, b1), (0, b2)))
            if l == 0:
                Q[c] += self._filt(ft0, 0, b3)
            elif b.sightline == 'global':
                Q[c] += self._filt(ft0, l, b3)
            else:
                Q[c] += b.to_fourier(b.apply_real_harmonics(self._filt(ft0, 0, b3), self.Ylm_real[l], self.Ylm_fourier[l]))
            Q[c] += self._filt(b.to_fourier(diff(gA, (0, b2), (l, b3))), 0, b1)
            Q[c] += self._filt(b.to_fourier(diff(gA, (0, b1), (l, b3))), 0, b2)
            Q[c] = apply_filter(Q[c]) if self.const_mask else apply_filter(b.to_real(Q[c]))
        del gA
        # rows: S^-1 P weighting (compute_Q('Sinv'), bspec.py:671)
        if self.const_mask:
            wS = [apply_filter(a) for a in (a1, a2)]
        else:
            wS = [apply_filter(b.to_real(a)) for a in (a1, a2)]
        needS = set()
        for b1, b2, b3, l in ab[rows]:
            needS |= {(0, b1), (l, b3)}
        gS = gmaps(wS[0], wS[1], needS)
        fish = np.zeros((len(rows), len(cols)))
        mag = np.zeros((len(rows), len(cols)))
        Qabs = np.abs(Q) if return_scale else None
        for r, (b1, b2, b3, l) in enumerate(ab[rows]):
            ft = b.to_fourier(diff(gS, (0, b1), (l, b3)))
            fish[r] = b.integrator.integrate_row(ft, Q, ks[b2], ks[b2 + 1])
            if return_scale:
                mag[r] = b.integrator.integrate_row(np.abs(ft), Qabs, ks[b2], ks[b2 + 1])
        sym = self.sym_factor
        norm = 0.5 * b.volume / b.gridsize.prod() ** 2 / np.outer(sym[rows], sym[cols])
        return (fish * norm, np.abs(mag * norm)) if return_scale else fish * norm

    def compute_fisher(self, N_it, verb=False):
        """bspec.py:678-693."""
        fish = np.zeros((self.N_bins, self.N_bins))
        for seed in range(N_it):
            fish += self.compute_fisher_contribution(seed) / N_it
        self.fish = fish
        self.inv_fish = np.linalg.inv(fish)
        return fish

    def Bk_unwindowed(self, data, fish=(), include_linear_term=True):
        """bspec.py:863-893."""
        if len(fish) != 0:
            self.fish = np.asarray(fish)
            self.inv_fish = np.linalg.inv(self.fish)
        if not hasattr(self, 'inv_fish'):
            raise Exception("Need to compute Fisher matrix first!")
        out = self.inv_fish @ self.Bk_numerator(data, include_linear_term=include_linear_term)
        return {'b%d' % l: out[i * self.N3:(i + 1) * self.N3] for i, l in enumerate(range(0, self.lmax + 1, 2))}


# ----------------------------------------------------------------------------------------------
# Structural FFT counts of the reference call graphs (for extrapolating the CPU baseline)
# ----------------------------------------------------------------------------------------------

def bk_fisher_fft_count(all_bins, Nk_squeeze, lmax, local, const_mask):
    """(forward, reverse) FFTs of one BSpec._compute_fisher call with trivial S^-1
    (PolyBin3D/bspec.py:473-676), counted from the loop structure without transforming anything."""
    ab = np.asarray(all_bins)
    Nl = lmax // 2 + 1
    loc = local and lmax > 0
    harm = sum(2 * l + 1 for l in range(2, lmax + 1, 2))
    fwd = rev = 0
    if not const_mask:
        rev += 2                                                    # a_real1, a_real2 (652-653)
    for weighting in ('Ainv', 'Sinv'):
        if weighting == 'Sinv' and not const_mask:
            fwd += 2                                                # apply_filter of both GRFs
        if loc:
            rev += 2                                                # weighted_map_real (543-544)
            fwd += 2 * harm                                         # leg maps (558-560)
        rev += 2 * Nk_squeeze                                       # g_{b,0} maps (550-552)
        for B in range(Nk_squeeze):
            if lmax > 0:
                rev += 2 * (Nl - 1)                                 # g_{B,l} maps (570-576)
            for A in range(B + 1):
                if weighting == 'Sinv':
                    if np.any((ab[:, 2] == B) & (ab[:, 0] == A)):
                        fwd += Nl                                   # ft_ABl (588-593)
                else:
                    i1 = np.where((ab[:, 0] == A) & (ab[:, 1] == B))[0]
                    i2 = np.where((ab[:, 1] == A) & (ab[:, 2] == B))[0]
                    i3 = np.where((ab[:, 2] == B) & (ab[:, 0] == A))[0]
                    if len(i1) + len(i2) + len(i3) == 0:
                        continue
                    fwd += Nl                                       # ft_ABl (609-613)
                    if loc:
                        for ii in i1:
                            l = ab[ii, 3]
                            if l > 0:
                                rev += 2 * l + 1                    # apply_real_harmonics (632)
                                fwd += 1                            # to_fourier (634/636)
        if weighting == 'Ainv' and not const_mask:
            rev += len(ab)                                          # to_real(Q_Ainv[i]) (666)
            fwd += len(ab)                                          # apply_filter (666)
    return fwd, rev


def pk_fisher_fft_count(Nk, lmax, local, const_mask):
    """(forward, reverse) FFTs of one PSpec._compute_fisher call with trivial S^-1, even l
    (PolyBin3D/pspec.py:215-493): the formula the Tutorial 3 counter (452/682) pins."""
    Nl = lmax // 2 + 1
    fwd = rev = 0
    if not const_mask:
        rev += 1                                                    # a_real (268)
        fwd += 1                                                    # apply_filter (276)
    loc = local and lmax > 0
    if loc:
        rev += 2                                                    # SinvP_real, Ainv_real (336, 343)
        fwd += sum(2 * (2 * li) + 1 for li in range(1, Nl))         # leg_maps (352)
    per_dagger = (1, 1) if not const_mask else (0, 0)               # (fwd, rev) of apply_filter_dagger
    fwd += Nk * per_dagger[0]
    rev += Nk * per_dagger[1]                                       # l = 0 rows (444-446)
    for li in range(1, Nl):
        ell = 2 * li
        if not loc:
            fwd += Nk * per_dagger[0]
            rev += Nk * per_dagger[1]
        else:
            fwd += 2 * ell + 1                                      # leg_map (462)
            rev += Nk * (2 * ell + 1)                               # real_map pieces (482)
            fwd += Nk * (per_dagger[0] + (1 if not const_mask else 1))  # two daggers (471, 485)
            rev += Nk * per_dagger[1]
    return fwd, rev
