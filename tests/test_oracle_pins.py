"""Pins for the FP64 oracle (CPU only): every check here is independent of the oracle's own code
path (notebook-printed counters and sizes, closed forms, brute force, the addition theorem), so a
dropped term, a wrong sign or index, or a transposed operand in oracle/polybin_oracle.py fails one.
"""
import itertools

import numpy as np
import pytest

from oracle import polybin_oracle as O


def field(n, seed=1):
    return np.random.RandomState(seed).randn(*n)


# ---------------------------------------------------------------- printed metadata / counters
def test_grid_metadata_tutorial1():
    """T1 cell 8 prints kF = 0.0063 and kNy = 0.402 for 128^3, L = 1000."""
    b = O.PolyBin3D(1000., [16, 16, 16])
    assert "%.4f" % (2 * np.pi / 1000.) == "0.0063"
    assert "%.3f" % (128 * 2 * np.pi / 1000. * 0.5) == "0.402"
    assert np.allclose(b.kF, 2 * np.pi / 1000.)


def test_fft_counter_pk_ideal_global():
    """T1 cell 10: Pk_ideal, global, lmax=4 -> 1 forward, 0 reverse."""
    b = O.PolyBin3D(1000., [32, 32, 32])
    ps = O.PSpec(b, np.arange(0.01, 0.1, 0.01), lmax=4)
    ps.Pk_ideal(field((32, 32, 32)))
    assert (b.n_FFTs_forward, b.n_FFTs_reverse) == (1, 0)


def test_fft_counter_pk_ideal_local():
    """T1 cell 21: Pk_ideal, local, lmax=4 -> 15 forward, 1 reverse."""
    b = O.PolyBin3D(1000., [32, 32, 32], boxcenter=[0., 1170., 826.], sightline='local')
    ps = O.PSpec(b, np.arange(0.01, 0.1, 0.01), lmax=4)
    ps.Pk_ideal(field((32, 32, 32)))
    assert (b.n_FFTs_forward, b.n_FFTs_reverse) == (15, 1)


def test_fft_counter_bk_ideal_and_nbins():
    """T4 cells 10/12: 4 bins + 6 squeezed bins, lmax=2 -> N_bins 78; Bk_ideal with and without
    the discreteness correction -> 2 forward, 96 reverse."""
    b = O.PolyBin3D(1000., [48, 48, 48])
    bs = O.BSpec(b, np.arange(0.01, 0.10, 0.02), lmax=2, k_bins_squeeze=np.arange(0.01, 0.15, 0.02),
                 Pk_fid='default')
    assert bs.N_bins == 78
    d = field((48, 48, 48))
    bs.Bk_ideal(d, discreteness_correction=True)
    bs.Bk_ideal(d, discreteness_correction=False)
    assert (b.n_FFTs_forward, b.n_FFTs_reverse) == (2, 96)


def test_fft_counter_pk_mc_fisher_local_masked():
    """T3 cells 16-26: one Pk_unwindowed, 10 Fisher + 10 shot realizations, one Pk_unwindowed
    (local, lmax=2, 9 bins, spatially varying mask) -> 452 forward, 682 reverse."""
    n = (16, 16, 16)
    b = O.PolyBin3D(1000., list(n), boxcenter=[700., 10., 200.], sightline='local')
    mask = 1 + 0.1 * np.random.RandomState(2).rand(*n)
    ps = O.PSpec(b, np.linspace(0.0063, 0.049, 10), lmax=2, mask=mask, mask_shot=mask)
    d = field(n)
    ps.Pk_unwindowed(d, fish=np.eye(ps.N_bins))
    F = [ps.compute_fisher_contribution(s) for s in range(10)]
    S = [ps.compute_shot_contribution(100 + s) for s in range(10)]
    ps.Pk_unwindowed(d, fish=np.mean(F, 0), shot_num=np.mean(S, 0))
    assert (b.n_FFTs_forward, b.n_FFTs_reverse) == (452, 682)


def test_nbins_and_allocation_tutorials():
    """T1 cell 10 '38 bins'; T5 cell 12 'N_bins: 56'; Q-map allocation prints 1.24 GB (T4 cell 25,
    128^3, 78 bins) and 0.40 GB (T5 cell 19, 78x146x82, 56 bins) = N_bins * n0 n1 (n2/2+1) * 16 B."""
    assert len(np.arange(0.01, 0.4, 0.01)) - 1 == 38
    e5 = np.linspace(0.02, 0.10, 6)
    t5 = O.triangle_bins(e5, e5, 2)
    assert len(t5) == 56
    t4 = O.triangle_bins(np.arange(0.01, 0.10, 0.02), np.arange(0.01, 0.15, 0.02), 2)
    assert "%.2f" % (len(t4) * 128 * 128 * 65 * 16 / 1024. ** 3) == "1.24"
    assert "%.2f" % (len(t5) * 78 * 146 * 42 * 16 / 1024. ** 3) == "0.40"


# ---------------------------------------------------------------- harmonics, FFT conventions
def test_real_harmonics_addition_theorem():
    rng = np.random.default_rng(0)
    a, c = rng.normal(size=(3, 500)), rng.normal(size=(3, 500))
    an, cn = a / np.linalg.norm(a, axis=0), c / np.linalg.norm(c, axis=0)
    Ya, Yc = O.real_harmonics(a, 4, True), O.real_harmonics(c, 4, True)
    for l in (1, 2, 3, 4):
        assert Ya[l].shape[0] == 2 * l + 1
        np.testing.assert_allclose((Ya[l] * Yc[l]).sum(0), O.legendre(l, (an * cn).sum(0)), atol=1e-13)
        # orthogonality on the sphere: <Y_m Y_m'> = delta / (2l+1)
        v = rng.normal(size=(3, 200000))
        Y = O.real_harmonics(v, 4, True)[l]
        np.testing.assert_allclose(Y @ Y.T / v.shape[1], np.eye(2 * l + 1) / (2 * l + 1), atol=1e-2)


def test_legendre_closed_forms():
    mu = np.linspace(-1, 1, 41)
    for l in range(7):
        ref = np.polynomial.legendre.legval(mu, [0] * l + [1])
        np.testing.assert_allclose(O.legendre(l, mu), ref, atol=1e-13)


def test_fft_bruteforce_dft():
    n = (4, 6, 8)
    b = O.PolyBin3D(100., list(n))
    x = field(n, 3)
    X = b.to_fourier(x)
    i, j, l = np.meshgrid(*[np.arange(m) for m in n], indexing='ij')
    ref = np.zeros(b.modk_grid.shape, dtype=complex)
    for p in range(n[0]):
        for q in range(n[1]):
            for r in range(n[2] // 2 + 1):
                ph = np.exp(-2j * np.pi * (p * i / n[0] + q * j / n[1] + r * l / n[2]))
                ref[p, q, r] = np.sum(x * ph)
    np.testing.assert_allclose(X, ref, atol=1e-10)
    np.testing.assert_allclose(b.to_real(X), x, atol=1e-12)
    # Parseval with the degeneracy weights: sum_x x^2 = (1/N) sum_k degk |X|^2 (no Nyquist double count)
    deg = np.ones(X.shape)
    deg[:, :, 1:n[2] // 2] = 2
    assert np.isclose(np.sum(x ** 2), np.sum(deg * np.abs(X) ** 2) / np.prod(n))


def test_grid_conventions():
    """k ordering fftshift(arange(-n//2, n//2)) (also for odd n), half axis, degk, mu."""
    b = O.PolyBin3D(10., [5, 4, 6])
    assert np.allclose(b.k_arrs[0] / b.kF[0], [0, 1, -3, -2, -1])
    assert np.allclose(b.k_arrs[1] / b.kF[1], [0, 1, -2, -1])
    assert np.allclose(b.k_arrs[2] / b.kF[2], [0, 1, 2, 3])
    assert np.all(b.degeneracy_factor[:, :, 0] == 1) and np.all(b.degeneracy_factor[:, :, 1:] == 2)
    kz = b.k_grids[2]
    nz = b.modk_grid > 0
    np.testing.assert_allclose(b.muk_grid[nz], kz[nz] / b.modk_grid[nz])
    assert b.muk_grid[0, 0, 0] == 0


# ---------------------------------------------------------------- bin integrals
def test_integrators_bruteforce_loops():
    """Per-bin full sweeps, written as the reference's explicit loop nest (utils.pyx:51-134)."""
    n = (6, 6, 8)
    b = O.PolyBin3D(100., list(n))
    rng = np.random.RandomState(4)
    s1 = rng.randn(*b.modk_grid.shape) + 1j * rng.randn(*b.modk_grid.shape)
    s2 = rng.randn(*b.modk_grid.shape) + 1j * rng.randn(*b.modk_grid.shape)
    r = rng.rand(*b.modk_grid.shape)
    edges = np.array([0.07, 0.13, 0.2, 0.25])
    for l in (0, 2, 4):
        ci = np.zeros(3)
        it = np.zeros(3)
        rd = np.zeros(3)
        for bb in range(3):
            for idx in itertools.product(*[range(m) for m in b.modk_grid.shape]):
                k = b.modk_grid[idx]
                if edges[bb] <= k < edges[bb + 1]:
                    L = O.legendre(l, b.muk_grid[idx])
                    g = b.degeneracy_factor[idx]
                    ci[bb] += g * (s1[idx] * np.conj(s2[idx])).real * L
                    it[bb] += g * abs(s1[idx]) ** 2 * L
                    rd[bb] += g * r[idx] ** 2 * L * O.legendre(2, b.muk_grid[idx])
        np.testing.assert_allclose(b.integrator.cross_integrate(s1, s2, edges, l), ci, rtol=1e-12)
        np.testing.assert_allclose(b.integrator.integrate(s1, edges, l), it, rtol=1e-12)
        np.testing.assert_allclose(b.integrator.real_integrate_double(r, l, 2, edges), rd, rtol=1e-12)


def test_mode_count_invariant():
    """real_integrate_double(1, 0, 0)[b] = number of full-grid lattice modes in shell b."""
    n = 12
    b = O.PolyBin3D(1000., [n, n, n])
    edges = np.array([1.0, 2.0, 3.5, 5.5]) * b.kF[0]
    got = b.integrator.real_integrate_double(np.ones(b.modk_grid.shape), 0, 0, edges)
    cnt = np.zeros(3)
    for v in itertools.product(range(-n // 2 + 1, n // 2), repeat=3):
        k = np.sqrt(sum((c * b.kF[0]) ** 2 for c in v))
        for bb in range(3):
            if edges[bb] <= k < edges[bb + 1]:
                cnt[bb] += 1
    np.testing.assert_array_equal(got, cnt)


# ---------------------------------------------------------------- estimators: known answers
def test_pk_ideal_plane_wave():
    """delta = A cos(k0.x), unit P_fid, global -> P0[b] = A^2 V / (2 N_modes(b)), other bins 0."""
    n, L, A = 16, 400., 2.5
    b = O.PolyBin3D(L, [n, n, n])
    kF = b.kF[0]
    edges = np.array([1.0, 2.0, 3.0, 4.0]) * kF + 1e-9
    ps = O.PSpec(b, edges, lmax=0)
    x = np.meshgrid(*[np.arange(n) * L / n] * 3, indexing='ij')
    k0 = np.array([1, 2, 0]) * kF  # |k0| = sqrt(5) kF in bin 1
    d = A * np.cos(k0[0] * x[0] + k0[1] * x[1] + k0[2] * x[2])
    p0 = ps.Pk_ideal(d)['p0']
    nmodes = sum(1 for v in itertools.product(range(-n // 2 + 1, n // 2), repeat=3)
                 if edges[1] <= np.sqrt(sum(c * c for c in v)) * kF < edges[2])
    np.testing.assert_allclose(p0, [0., A ** 2 * b.volume / (2 * nmodes), 0.], atol=1e-9)


def test_bk_numerator_fourier_bruteforce():
    """sum_x g_a g_b g_c = N^-2 sum_{k1+k2+k3=0} g_a(k1) g_b(k2) g_c(k3) on an 8^3 grid: the oracle's
    g-maps + assemble_b3 + V/N/sym normalisation against a direct Fourier-space triangle sum."""
    n = 8
    b = O.PolyBin3D(100., [n, n, n])
    kF = b.kF[0]
    edges = np.array([1.0, 1.5, 2.5, 3.5]) * kF
    bs = O.BSpec(b, edges, lmax=0, Pk_fid='default')
    d = field((n, n, n), 5)
    d = d + 0.3 * (d ** 2 - np.mean(d ** 2))
    num = bs.Bk_numerator_ideal(d)
    dk = np.fft.fftn(d)
    f = np.fft.fftfreq(n, 1. / n)
    kk = np.sqrt(f[:, None, None] ** 2 + f[None, :, None] ** 2 + f[None, None, :] ** 2) * kF
    masks = [(kk >= edges[i]) & (kk < edges[i + 1]) for i in range(len(edges) - 1)]
    idx = np.arange(n)
    ref = []
    for (b1, b2, b3, _), sym in zip(bs.all_bins, bs.sym_factor):
        ga, gb, gc = dk * masks[b1], dk * masks[b2], dk * masks[b3]
        tot = 0.
        for p in itertools.product(range(n), repeat=3):
            if ga[p] == 0:
                continue
            # sum over k2 of gb(k2) gc(-k1-k2)
            mi = (-p[0] - idx[:, None, None]) % n
            mj = (-p[1] - idx[None, :, None]) % n
            mk = (-p[2] - idx[None, None, :]) % n
            tot += ga[p] * np.sum(gb * gc[mi, mj, mk])
        ref.append((tot / n ** 6).real * b.volume / n ** 3 / sym)
    np.testing.assert_allclose(num, ref, rtol=1e-9, atol=1e-12 * np.max(np.abs(ref)))


def test_bk_ideal_fisher_triangle_count():
    """compute_fisher_ideal(discreteness_correction=False), l=0, unit weights: diagonal entries are
    #{k1+k2+k3=0, ki in bin bi} / sym / V (closed form by lattice enumeration)."""
    n = 8
    b = O.PolyBin3D(100., [n, n, n])
    kF = b.kF[0]
    edges = np.array([1.0, 1.5, 2.5, 3.5]) * kF
    bs = O.BSpec(b, edges, lmax=0, Pk_fid='default')
    F = bs.compute_fisher_ideal(discreteness_correction=False)
    f = np.fft.fftfreq(n, 1. / n)
    kk = np.sqrt(f[:, None, None] ** 2 + f[None, :, None] ** 2 + f[None, None, :] ** 2) * kF
    binof = np.full(kk.shape, -1)
    for i in range(len(edges) - 1):
        binof[(kk >= edges[i]) & (kk < edges[i + 1])] = i
    idx = np.arange(n)
    for t, (b1, b2, b3, _) in enumerate(bs.all_bins):
        cnt = 0
        for p in itertools.product(range(n), repeat=3):
            if binof[p] != b1:
                continue
            mi = (-p[0] - idx[:, None, None]) % n
            mj = (-p[1] - idx[None, :, None]) % n
            mk = (-p[2] - idx[None, None, :]) % n
            cnt += int(np.sum((binof == b2) & (binof[mi, mj, mk] == b3)))
        sym = 6. if b1 == b2 == b3 else (2. if (b1 == b2 or b2 == b3) else 1.)
        assert bs.sym_factor[t] == sym
        np.testing.assert_allclose(F[t, t], cnt / sym / b.volume, rtol=1e-10)


def test_grf_statistics_and_hermiticity():
    """generate_data: E|a_k|^2 = P N^2 / V; real output; reproducible per seed."""
    n = 16
    b = O.PolyBin3D(500., [n, n, n])
    a = [b.generate_data(seed=s, output_type='fourier') for s in range(40)]
    p = np.mean([np.abs(x[b.modk_grid > 0]) ** 2 for x in a])
    assert abs(p / (n ** 6 / b.volume) - 1) < 0.02
    np.testing.assert_array_equal(a[3], b.generate_data(seed=3, output_type='fourier'))
    r = b.to_real(a[0])
    np.testing.assert_allclose(b.to_fourier(r), a[0], atol=1e-9)


def test_pk_mc_fisher_global_closed_form():
    """Global sightline, constant mask, trivial S^-1: per seed
    F[(l,b),(l',b')] = delta_bb' m^2 V/(2N^2) sum_{k in b} degk invPk0 |a_k|^2 L_l L_l'."""
    n = 16
    Pk = [np.linspace(0, 2, 400), 1e3 * (1 + np.linspace(0, 2, 400))]
    b = O.PolyBin3D(600., [n, n, n], Pk=Pk)
    edges = np.linspace(0.012, 0.07, 5)
    m = 0.7
    ps = O.PSpec(b, edges, lmax=4, mask=m * np.ones((n, n, n)))
    F = ps.compute_fisher_contribution(3)
    a = b.generate_data(seed=3 + int(1e7), output_type='fourier')
    ref = np.zeros_like(F)
    Nk = ps.Nk
    idx = O.bin_index(b.modk_grid, edges)
    for l1 in range(3):
        for l2 in range(3):
            for k in range(Nk):
                sel = idx == k
                w = b.degeneracy_factor[sel] * b.invPk0_grid[sel] * np.abs(a[sel]) ** 2
                w = w * O.legendre(2 * l1, b.muk_grid[sel]) * O.legendre(2 * l2, b.muk_grid[sel])
                ref[l1 * Nk + k, l2 * Nk + k] = m ** 2 * b.volume / (2 * n ** 6) * w.sum()
    np.testing.assert_allclose(F, ref, atol=1e-12 * np.abs(ref).max())


def _dense_ops(b):
    """Full-grid DFT matrix and the addition-theorem kernels L_l(k_hat . x_hat) e^{-ik.x}."""
    n = b.gridsize
    N = int(np.prod(n))
    idx = np.stack(np.meshgrid(*[np.arange(m) for m in n], indexing='ij'), -1).reshape(N, 3)
    fr = np.stack(np.meshgrid(*[np.fft.fftfreq(m, 1. / m) for m in n], indexing='ij'), -1).reshape(N, 3)
    ph = np.exp(-2j * np.pi * (fr / n) @ idx.T)                      # [k, x]
    kv = fr * b.kF
    xs = np.stack(np.meshgrid(*[np.fft.fftshift(np.arange(-m // 2, m // 2)) * b.boxsize[i] / m + b.boxcenter[i]
                                for i, m in enumerate(n)], indexing='ij'), -1).reshape(N, 3)
    kh = kv / np.maximum(np.linalg.norm(kv, axis=1), 1e-300)[:, None]
    xh = xs / np.maximum(np.linalg.norm(xs, axis=1), 1e-300)[:, None]
    return ph, kv, kh @ xh.T, fr


def test_pk_mc_fisher_local_masked_dense():
    """Local sightline, varying mask, l<=2: the per-seed MC Fisher against a dense-matrix restatement
    (explicit DFT matrices and addition-theorem kernels instead of Y_lm sums and FFTs)."""
    n = 8
    b = O.PolyBin3D(200., [n, n, n], boxcenter=[300., 140., 90.], sightline='local')
    mask = 0.5 + np.random.RandomState(6).rand(n, n, n)
    edges = np.array([1.0, 2.0, 3.0]) * b.kF[0] + 1e-9
    ps = O.PSpec(b, edges, lmax=2, mask=mask)
    F = ps.compute_fisher_contribution(2)
    N = n ** 3
    ph, kv, cosang, fr = _dense_ops(b)
    # full-grid Fourier amplitudes of the same GRF
    ah = b.generate_data(seed=2 + int(1e7), output_type='fourier')
    a = np.fft.fftn(b.to_real(ah)).reshape(N)
    ar = np.fft.ifftn(a.reshape(n, n, n)).real.reshape(N)
    modk = np.linalg.norm(kv, axis=1)
    Pk0 = np.ones(N)
    inv = 1. / Pk0
    mk = mask.reshape(N)
    SinvP = ph @ (mk * ar)
    Ainv = inv * a
    Ainv_r = (ph.conj().T @ Ainv).real / N
    Lk = {0: np.ones((N, N)), 2: O.legendre(2, cosang)}
    lefts = {0: Ainv, 2: (ph * Lk[2]) @ Ainv_r}
    mu = np.where(modk > 0, kv[:, 2] * 0, 0)  # unused for local
    Nk = 2
    ref = np.zeros((2 * Nk, 2 * Nk))
    SinvP_r = (ph.conj().T @ SinvP).real / N
    for li, ell in enumerate((0, 2)):
        for kb in range(Nk):
            th = (modk >= edges[kb]) & (modk < edges[kb + 1])
            if ell == 0:
                km = th * SinvP
                Q = ph @ (mk * (ph.conj().T @ km).real / N)
            else:
                leg = (ph * Lk[2]) @ SinvP_r
                t1 = ph @ (mk * (ph.conj().T @ (th * leg)).real / N)
                real_map = ((ph.conj() * Lk[2]).T @ (th * SinvP)).real / N
                t2 = ph @ (mk * real_map)
                Q = 0.5 * t1 + 0.5 * t2
            for lj, ell2 in enumerate((0, 2)):
                for kb2 in range(Nk):
                    th2 = (modk >= edges[kb2]) & (modk < edges[kb2 + 1])
                    ref[li * Nk + kb, lj * Nk + kb2] = 0.5 * b.volume / N ** 2 * np.sum(
                        th2 * (lefts[ell2] * np.conj(Q)).real)
    np.testing.assert_allclose(F, ref, atol=1e-10 * np.abs(ref).max(), rtol=1e-8)


def test_pk_unwindowed_equals_ideal_with_ideal_fisher():
    """Pk_unwindowed with F = F_ideal, S^-1 = invPk0 and no mask reproduces Pk_ideal
    (pspec.py:631-640 apply the same filter)."""
    n = 16
    Pk = [np.linspace(0, 2, 300), 2e3 / (1 + 10 * np.linspace(0, 2, 300))]
    b = O.PolyBin3D(800., [n, n, n], Pk=Pk)
    edges = np.linspace(0.01, 0.05, 6)
    ps = O.PSpec(b, edges, lmax=2, applySinv=lambda m, input_type='real', output_type='fourier':
                 b.to_fourier(m) * b.invPk0_grid)
    d = field((n, n, n), 9)
    ideal = ps.Pk_ideal(d)
    unw = ps.Pk_unwindowed(d, fish=ps.fish_ideal)
    for k in ideal:
        np.testing.assert_allclose(unw[k], ideal[k], rtol=1e-10)


def test_bk_mc_fisher_expectation_equals_ideal():
    """BSpec._compute_fisher (bspec.py:473-676) in expectation: with unit mask, global sightline and
    S^-1 = A^-1 = invPk0 (the ideal weighting), the Monte Carlo Fisher built from pairs of GRFs is
    unbiased for the exact Gaussian Fisher, which is the ideal Fisher with the discreteness
    correction (bspec.py:1025-1054; a separate code path: triplet sums of filtered invPk0 maps).
    200 seeds on a 16^3 grid, l <= 2: every entry of the mean within 5 standard errors, diagonal
    entries within 10%.  A dropped permutation term, a factor 2 in the normalisation or a transposed
    index moves entries by tens of standard errors."""
    L, n = 1000., 16
    kk = np.linspace(0, 3, 6000)
    P = np.zeros_like(kk)
    P[1:] = 2e4 * (kk[1:] / 0.1) ** -1.5
    b = O.PolyBin3D(L, [n] * 3, Pk=[kk, P])

    def Sinv(m, input_type='real', output_type='real'):
        f = b.to_fourier(m) if input_type == 'real' else m
        f = f * b.invPk0_grid
        return b.to_real(f) if output_type == 'real' else f

    edges = (2 * np.pi / L) * np.array([1., 2., 3., 4., 5.])
    bs = O.BSpec(b, edges, lmax=2, Pk_fid='default', applySinv=Sinv)
    Fid = bs.compute_fisher_ideal(discreteness_correction=True)
    Fs = np.array([bs.compute_fisher_contribution(s) for s in range(200)])
    m, se = Fs.mean(0), Fs.std(0) / np.sqrt(len(Fs))
    d = np.diag(Fid)
    scale = np.sqrt(np.outer(d, d))
    zero = se <= 1e-12 * scale  # entries that vanish in every realization
    assert np.all(np.abs(m - Fid)[zero] <= 1e-10 * scale[zero])
    z = np.abs(m - Fid)[~zero] / se[~zero]
    assert z.max() < 5., "max deviation %.1f standard errors" % z.max()
    assert np.max(np.abs(np.diag(m) - d) / d) < 0.1


def test_bk_symmetry_factor_l2_lattice_sum():
    """_compute_symmetry_factor for l > 0 (bspec.py:245-276): for b2 == b3 the FFT ratio
    sum_x u_b1 sum_m v_m^2 / sum_x u_b1 u_b3 u_b3 (u = IFFT Theta w, v_m = IFFT Theta Y_lm w,
    w = invPk0) equals, by the convolution and addition theorems, the lattice triangle sum
    sum_{k1+k2+k3 = 0 mod n} w1 w2 w3 L_l(k2_hat . k3_hat) / sum w1 w2 w3 over k1 in b1,
    k2, k3 in b3 (brute-force enumeration on a 12^3 grid)."""
    L, n = 1000., 12
    kk = np.linspace(0, 3, 6000)
    P = np.zeros_like(kk)
    P[1:] = 2e4 * (kk[1:] / 0.1) ** -1.5
    b = O.PolyBin3D(L, [n] * 3, Pk=[kk, P])
    kF = 2 * np.pi / L
    edges = kF * np.array([1., 2., 3., 4.])
    bs = O.BSpec(b, edges, lmax=2, Pk_fid='default')
    bs._compute_symmetry_factor()
    f = np.fft.fftfreq(n, 1. / n).astype(int)
    K = np.stack(np.meshgrid(f, f, f, indexing='ij'), -1).reshape(-1, 3)
    modk = np.sqrt((K ** 2).sum(1)) * kF
    binof = np.searchsorted(edges, modk, side='right') - 1
    binof[(modk < edges[0]) | (modk >= edges[-1])] = -1
    # invPk0 on the full grid: the half-grid value at k or -k
    half = np.asarray(b.invPk0_grid)

    def w_of(k):
        k = np.where(k[..., 2:3] < 0, -k, k)
        k = np.where((k[..., 2:3] == 0) & (k[..., 1:2] < 0), -k, k)  # kz = 0 plane: either half works
        return half[k[..., 0] % n, k[..., 1] % n, k[..., 2]]

    def wrap(k):
        return (k + n // 2) % n - n // 2

    t = bs.all_bins[bs.N3:]  # the l = 2 block
    checked = set()
    for i, (b1, b2, b3, ell) in enumerate(t):
        if b2 != b3:
            continue
        checked.add(b1 == b2)
        k2 = K[binof == b3]
        k3 = K[binof == b3]
        kk2 = np.repeat(k2, len(k3), 0)
        kk3 = np.tile(k3, (len(k2), 1))
        k1 = wrap(-(kk2 + kk3))
        m1 = np.sqrt((k1 ** 2).sum(1)) * kF
        in1 = (m1 >= edges[b1]) & (m1 < edges[b1 + 1])
        w = w_of(k1[in1]) * w_of(kk2[in1]) * w_of(kk3[in1])
        c = (kk2[in1] * kk3[in1]).sum(1) / np.sqrt((kk2[in1] ** 2).sum(1) * (kk3[in1] ** 2).sum(1))
        ratio = np.sum(w * O.legendre(2, c)) / np.sum(w)
        want = 1. + ratio if b1 != b2 else 2. * (1. + 2. * ratio)
        np.testing.assert_allclose(bs.sym_factor[bs.N3 + i], want, rtol=1e-10)
    assert checked == {True, False}  # both the b1 == b2 == b3 and the b1 < b2 == b3 forms


def test_bk_numerator_local_dense():
    """Local-sightline Bk numerator (bspec.py:931-956: g_{l,b} from the Y_lm sums of
    apply_fourier_harmonics): by the addition theorem the l-leg is the dense operator
    g_{l,b}(x) = N^-1 sum_k Theta_b(k) e^{ik.x} sum_x' L_l(k_hat . x_hat') e^{-ik.x'} S(x'), built here
    with explicit DFT matrices on an 8^3 grid (no voxel at the origin), then the triplet sums and
    V/N/sym normalisation."""
    n, Lb = 8, 1000.
    center = [400., 250., -300.]
    b = O.PolyBin3D(Lb, [n] * 3, boxcenter=center, sightline='local')
    kF = b.kF[0]
    edges = np.array([1.0, 1.5, 2.5, 3.5]) * kF
    bs = O.BSpec(b, edges, lmax=2, Pk_fid='default')
    d = field((n, n, n), 6)
    d = d + 0.3 * (d ** 2 - np.mean(d ** 2))
    num = bs.Bk_numerator_ideal(d)
    # full-grid k vectors and positions (fft order, as the FFTs use them)
    f = np.fft.fftfreq(n, 1. / n)
    kv = np.stack(np.meshgrid(f, f, f, indexing='ij'), -1).reshape(-1, 3) * kF
    r = np.asarray(b.r_grids).reshape(3, -1).T
    assert np.min(np.sum(r ** 2, 1)) > 0
    modk = np.sqrt(np.sum(kv ** 2, 1))
    khat = kv / np.where(modk > 0, modk, 1.)[:, None]
    rhat = r / np.sqrt(np.sum(r ** 2, 1))[:, None]
    pos = np.stack(np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing='ij'), -1).reshape(-1, 3)
    E = np.exp(-2j * np.pi * (pos @ np.stack(np.meshgrid(f, f, f, indexing='ij'), -1).reshape(-1, 3).T) / n)  # [x, k]
    # S(x) = IFFT(invPk0 * FT d) / <mask^3>^(1/3) with unit mask: the ideal filter on the full grid
    half = np.asarray(b.invPk0_grid)
    ki = np.stack(np.meshgrid(f, f, f, indexing='ij'), -1).reshape(-1, 3).astype(int)
    kk = np.where(ki[:, 2:3] < 0, -ki, ki)
    inv_full = half[kk[:, 0] % n, kk[:, 1] % n, kk[:, 2] % n]
    dk = (E.T @ d.reshape(-1))  # sum_x d(x) e^{-ik.x}
    Sx = (E.conj() @ (inv_full * dk)).real / n ** 3
    g = {}
    for ell in (0, 2):
        Lmat = O.legendre(ell, khat @ rhat.T)  # [k, x']
        lk = (Lmat * E.T) @ Sx if ell else dk * inv_full  # sum_x' L_l(k.x') e^{-ik.x'} S(x')
        for q in range(len(edges) - 1):
            th = (modk >= edges[q]) & (modk < edges[q + 1])
            g[ell, q] = (E.conj() @ (th * lk)).real / n ** 3
    ref = []
    for (b1, b2, b3, ell), sym in zip(bs.all_bins, bs.sym_factor):
        ref.append(np.sum(g[0, b1] * g[0, b2] * g[ell, b3]) * b.volume / n ** 3 / sym)
    np.testing.assert_allclose(num, ref, rtol=1e-9, atol=1e-12 * np.max(np.abs(ref)))


def test_integral_constraint_transposes_are_adjoints():
    """<P m1, m2> = <m1, P^T m2> for the global and radial integral-constraint pointing matrices
    (the transposes are derived independently of the forward maps)."""
    rng = np.random.RandomState(9)
    n = (6, 5, 4)
    m1, m2, mask, mic = rng.randn(*n), rng.randn(*n), rng.rand(*n), rng.rand(*n)
    modr = rng.uniform(0., 10., n)
    edges = np.array([1., 3.5, 6., 9.])
    lhs = np.sum(O.pointing_GIC(m1, mask, mic) * m2)
    assert abs(lhs - np.sum(m1 * O.pointing_GIC_transpose(m2, mask, mic))) < 1e-12 * abs(lhs) + 1e-12
    lhs = np.sum(O.pointing_RIC(m1, mask, mic, edges, modr) * m2)
    rhs = np.sum(m1 * O.pointing_RIC_transpose(m2, mask, mic, edges, modr))
    assert abs(lhs - rhs) < 1e-12 * abs(lhs) + 1e-12
    # a constant map is annihilated by the constraint
    assert np.max(np.abs(O.pointing_GIC(np.ones(n), mask, mic))) < 1e-14


def test_pk_local_odd_l_dense():
    """odd_l (l = 0, 1, 2), local sightline, varying mask: the per-seed MC Fisher and the numerator
    against a dense-matrix restatement.  Odd-l legs carry the reference's factors: -i on the
    filtered leg maps and on the Y_lm-weighted k-space maps (pspec.py:469-481), i on Q in the
    cross-integral (pspec.py:410-411), -i on the numerator's leg map (pspec.py:677-678)."""
    n = 8
    b = O.PolyBin3D(200., [n, n, n], boxcenter=[300., 140., 90.], sightline='local')
    mask = 0.5 + np.random.RandomState(6).rand(n, n, n)
    edges = np.array([1.0, 2.0, 3.0]) * b.kF[0] + 1e-9
    ps = O.PSpec(b, edges, lmax=2, mask=mask, odd_l=True)
    F = ps.compute_fisher_contribution(2)
    N = n ** 3
    ph, kv, cosang, fr = _dense_ops(b)
    ah = b.generate_data(seed=2 + int(1e7), output_type='fourier')
    a = np.fft.fftn(b.to_real(ah)).reshape(N)
    ar = np.fft.ifftn(a.reshape(n, n, n)).real.reshape(N)
    modk = np.linalg.norm(kv, axis=1)
    mk = mask.reshape(N)
    SinvP = ph @ (mk * ar)
    Ainv = a.copy()
    Ainv_r = (ph.conj().T @ Ainv).real / N
    ells = (0, 1, 2)
    Lk = {l: O.legendre(l, cosang) for l in ells}
    lefts = {0: Ainv, 1: (ph * Lk[1]) @ Ainv_r, 2: (ph * Lk[2]) @ Ainv_r}
    Nk = 2
    ref = np.zeros((3 * Nk, 3 * Nk))
    SinvP_r = (ph.conj().T @ SinvP).real / N
    for li, ell in enumerate(ells):
        ph_odd = -1j if ell % 2 == 1 else 1.
        for kb in range(Nk):
            th = (modk >= edges[kb]) & (modk < edges[kb + 1])
            if ell == 0:
                Q = ph @ (mk * (ph.conj().T @ (th * SinvP)).real / N)
            else:
                leg = (ph * Lk[ell]) @ SinvP_r
                t1 = ph @ (mk * (ph.conj().T @ (ph_odd * th * leg)).real / N)
                real_map = ((ph.conj() * Lk[ell]).T @ (ph_odd * th * SinvP)).real / N
                Q = 0.5 * t1 + 0.5 * (ph @ (mk * real_map))
            for lj, ell2 in enumerate(ells):
                Qx = 1j * Q if ell2 % 2 == 1 else Q
                for kb2 in range(Nk):
                    th2 = (modk >= edges[kb2]) & (modk < edges[kb2 + 1])
                    ref[li * Nk + kb, lj * Nk + kb2] = 0.5 * b.volume / N ** 2 * np.sum(
                        th2 * (lefts[ell2] * np.conj(Qx)).real)
    np.testing.assert_allclose(F, ref, atol=1e-10 * np.abs(ref).max(), rtol=1e-8)
    # numerator (S^-1 filtering: S = FFT[d]): 0.5 sum degk Re[S conj(-i lm)] over the half grid
    d = np.random.RandomState(3).randn(n, n, n)
    num = ps.Pk_numerator(d)
    S = np.fft.fftn(d).reshape(N)
    S_r = d.reshape(N)
    half = fr[:, 2] >= 0
    nyq = fr[:, 2] == -n // 2  # fftfreq puts the z Nyquist plane at -n/2: it belongs to the half grid
    deg = np.where(fr[:, 2] > 0, 2., 1.)
    deg = np.where(nyq, 2., deg)
    sel = half | nyq
    refn = np.zeros(3 * Nk)
    for li, ell in enumerate(ells):
        lm = S if ell == 0 else (ph * Lk[ell]) @ S_r
        if ell % 2 == 1:
            lm = -1j * lm
        for kb in range(Nk):
            th = (modk >= edges[kb]) & (modk < edges[kb + 1]) & sel
            refn[li * Nk + kb] = 0.5 * b.volume / N ** 2 * np.sum(deg[th] * (S[th] * np.conj(lm[th])).real)
    np.testing.assert_allclose(num, refn, atol=1e-10 * np.abs(refn).max(), rtol=1e-8)


# ---------------------------------------------------------------- covariance and theory matrices
def _dense_q_rows(b, Cr, edges, seed, ells):
    """Dense restatement of the MC-Fisher Q maps (pspec.py:439-488) for trivial S^-1 and a real-space
    filter matrix Cr in place of the pointing (S^-1 P -> FFT Cr IFFT): full-grid DFT matrices and
    addition-theorem kernels.  Returns the Q maps in row order and the A^-1 a maps."""
    n = b.gridsize
    N = int(np.prod(n))
    ph, kv, cosang, fr = _dense_ops(b)
    ah = b.generate_data(seed=seed + int(1e7), output_type='fourier')
    a = np.fft.fftn(b.to_real(ah)).reshape(N)
    ar = np.fft.ifftn(a.reshape(*n)).real.reshape(N)
    modk = np.linalg.norm(kv, axis=1)
    SinvP = ph @ (Cr @ ar)
    SinvP_r = (ph.conj().T @ SinvP).real / N
    rows = []
    for ell in ells:
        Lk = O.legendre(ell, cosang)
        for kb in range(len(edges) - 1):
            th = (modk >= edges[kb]) & (modk < edges[kb + 1])
            if ell == 0:
                rows.append(ph @ (Cr @ ((ph.conj().T @ (th * SinvP)).real / N)))
            else:
                leg = (ph * Lk) @ SinvP_r
                t1 = ph @ (Cr @ ((ph.conj().T @ (th * leg)).real / N))
                real_map = ((ph.conj() * Lk).T @ (th * SinvP)).real / N
                rows.append(0.5 * t1 + 0.5 * (ph @ (Cr @ real_map)))
    Ainv = a.copy()  # unit fiducial spectrum
    Ainv_r = (ph.conj().T @ Ainv).real / N
    return rows, Ainv, Ainv_r, modk, ph, cosang


def test_derivative_coefficient_known_answers():
    """The wide-angle k dP/dk weights (pspec.py:369-379) on hand-evaluated edges, and -1 in the
    interior of a linear binning (the column sum of a central difference of k P(k))."""
    np.testing.assert_allclose(O.derivative_coefficient([1., 2., 3., 4., 5.]), [-2., -0.5, 6., -3.5])
    kf = O.derivative_coefficient(0.01 + 0.005 * np.arange(12))
    np.testing.assert_allclose(kf[2:-3], -1., rtol=1e-12)


def test_apply_xi_local_dense():
    """apply_xi (base.py:413-437), local sightline with P_0 and P_2: against the dense operator
    xi(x, y) = (1/N) sum_k e^{ik(x-y)} [P_0(k) + P_2(k) (L_2(k.x) + L_2(k.y)) / 2]; both output types,
    and the operator is symmetric."""
    n = 8
    N = n ** 3
    b = O.PolyBin3D(200., [n, n, n], boxcenter=[300., 140., 90.], sightline='local')
    Yr = O.real_harmonics(np.asarray(b.r_grids), 2, False)
    Yf = O.real_harmonics(np.asarray(b.k_grids), 2, False)
    kk = np.linspace(0.5 * b.kF[0], 0.85 * b.kNy[0], 60)  # vanishes at the Nyquist modes and k = 0
    P0f = lambda k: 1. + 30. * k  # noqa: E731
    P2f = lambda k: 0.7 - 11. * k ** 2  # noqa: E731
    Pk_grid = {0: np.interp(b.modk_grid, kk, P0f(kk), left=0., right=0.) * (b.modk_grid <= kk[-1]) * (b.modk_grid >= kk[0]),
               2: np.interp(b.modk_grid, kk, P2f(kk), left=0., right=0.) * (b.modk_grid <= kk[-1]) * (b.modk_grid >= kk[0])}
    ph, kv, cosang, fr = _dense_ops(b)
    modk = np.linalg.norm(kv, axis=1)
    inr = (modk >= kk[0]) & (modk <= kk[-1])
    P0 = np.where(inr, np.interp(modk, kk, P0f(kk)), 0.)
    P2 = np.where(inr, np.interp(modk, kk, P2f(kk)), 0.)
    L2 = O.legendre(2, cosang)
    Xi = (ph.conj().T @ (P0[:, None] * ph) + 0.5 * ph.conj().T @ (P2[:, None] * (ph * L2))
          + 0.5 * (ph * L2).conj().T @ (P2[:, None] * ph)) / N
    assert np.abs(Xi.imag).max() < 1e-9 * np.abs(Xi).max()
    Xi = Xi.real
    np.testing.assert_allclose(Xi, Xi.T, atol=1e-10 * np.abs(Xi).max())
    u = np.random.RandomState(4).randn(n, n, n)
    uf = b.to_fourier(u)
    out_r = b.apply_xi(uf, u, Yr, Yf, Pk_grid, 'real')
    np.testing.assert_allclose(out_r.reshape(N), Xi @ u.reshape(N), atol=1e-10 * np.abs(out_r).max())
    out_f = b.apply_xi(uf, u, Yr, Yf, Pk_grid, 'fourier')
    np.testing.assert_allclose(out_f, b.to_fourier(out_r), atol=1e-10 * np.abs(out_f).max())


def test_pk_covariance_global_closed_form():
    """Covariance contribution, global sightline, constant mask m, trivial S^-1 (pspec.py:278-289): the
    filter is c(k) = m^2 P(k, mu) + m, so per seed
    C[(l,b),(l',b')] = delta_bb' V/(2N^2) sum_{k in b} degk c^2 invPk0 |a_k|^2 L_l L_l'."""
    n = 16
    Pk = [np.linspace(0, 2, 400), 1e3 * (1 + np.linspace(0, 2, 400))]
    b = O.PolyBin3D(600., [n, n, n], Pk=Pk)
    edges = np.linspace(0.012, 0.07, 5)
    m = 0.7
    ps = O.PSpec(b, edges, lmax=4, mask=m * np.ones((n, n, n)))
    kc = np.linspace(0., 0.08, 50)
    Pc = [kc, 2e3 / (1 + 50 * kc), 300. - 2e3 * kc, 40. + kc]
    C = ps.compute_covariance_contribution(3, Pc)
    a = b.generate_data(seed=3 + int(1e7), output_type='fourier')
    k, mu = b.modk_grid, b.muk_grid
    Pkmu = (np.interp(k, kc, Pc[1], right=0.) + O.legendre(2, mu) * np.interp(k, kc, Pc[2], right=0.)
            + O.legendre(4, mu) * np.interp(k, kc, Pc[3], right=0.))
    c = m ** 2 * Pkmu + m
    ref = np.zeros_like(C)
    Nk = ps.Nk
    idx = O.bin_index(k, edges)
    for l1 in range(3):
        for l2 in range(3):
            for kb in range(Nk):
                sel = idx == kb
                w = b.degeneracy_factor[sel] * b.invPk0_grid[sel] * np.abs(a[sel]) ** 2 * c[sel] ** 2
                w = w * O.legendre(2 * l1, mu[sel]) * O.legendre(2 * l2, mu[sel])
                ref[l1 * Nk + kb, l2 * Nk + kb] = b.volume / (2 * n ** 6) * w.sum()
    np.testing.assert_allclose(C, ref, atol=1e-12 * np.abs(ref).max())


def test_pk_covariance_local_masked_dense():
    """Covariance contribution, local sightline, varying mask with a shot-noise mask (pspec.py:278-325):
    the MC-Fisher rows with the real-space filter P xi P^T + diag(mask_shot), xi the dense local
    correlation operator."""
    n = 8
    N = n ** 3
    b = O.PolyBin3D(200., [n, n, n], boxcenter=[300., 140., 90.], sightline='local')
    rng = np.random.RandomState(6)
    mask = 0.5 + rng.rand(n, n, n)
    mshot = 0.2 + rng.rand(n, n, n)
    edges = np.array([1.0, 2.0, 3.0]) * b.kF[0] + 1e-9
    ps = O.PSpec(b, edges, lmax=2, mask=mask, mask_shot=mshot)
    kk = np.linspace(0.5 * b.kF[0], 0.85 * b.kNy[0], 60)
    Pc = [kk, 1. + 30. * kk, 0.7 - 11. * kk ** 2]
    C = ps.compute_covariance_contribution(2, Pc)
    ph, kv, cosang, fr = _dense_ops(b)
    modk = np.linalg.norm(kv, axis=1)
    inr = (modk >= kk[0]) & (modk <= kk[-1])
    P0 = np.where(inr, np.interp(modk, kk, Pc[1]), 0.)
    P2 = np.where(inr, np.interp(modk, kk, Pc[2]), 0.)
    L2 = O.legendre(2, cosang)
    Xi = ((ph.conj().T @ (P0[:, None] * ph) + 0.5 * ph.conj().T @ (P2[:, None] * (ph * L2))
           + 0.5 * (ph * L2).conj().T @ (P2[:, None] * ph)) / N).real
    mk = mask.reshape(N)
    Cr = mk[:, None] * Xi * mk[None, :] + np.diag(mshot.reshape(N))
    rows, Ainv, Ainv_r, modk, ph, cosang = _dense_q_rows(b, Cr, edges, 2, (0, 2))
    lefts = [Ainv, (ph * L2) @ Ainv_r]
    Nk = 2
    ref = np.zeros((2 * Nk, 2 * Nk))
    for r, Q in enumerate(rows):
        for lj in range(2):
            for kb2 in range(Nk):
                th2 = (modk >= edges[kb2]) & (modk < edges[kb2 + 1])
                ref[r, lj * Nk + kb2] = 0.5 * b.volume / N ** 2 * np.sum(th2 * (lefts[lj] * np.conj(Q)).real)
    np.testing.assert_allclose(C, ref, atol=1e-10 * np.abs(ref).max(), rtol=1e-8)


def test_pk_theory_matrix_cross_pins():
    """Theory binning matrix (pspec.py:165-197, 381-436), local sightline, varying mask: with theory bins
    equal to the data bins it is the Fisher matrix; with lmax_theory = 4 its l = 4 block is the l = 4
    column block of the lmax = 4 Fisher matrix's l <= 2 rows (a separate code path); with refined
    theory bins that extend past the data's k-range, the sums of the refined bins are the coarse ones."""
    n = 8
    b = O.PolyBin3D(200., [n, n, n], boxcenter=[300., 140., 90.], sightline='local')
    mask = 0.5 + np.random.RandomState(6).rand(n, n, n)
    kF = b.kF[0]
    edges = np.array([1.0, 2.0, 3.0]) * kF + 1e-9
    ps2 = O.PSpec(b, edges, lmax=2, mask=mask)
    ps4 = O.PSpec(b, edges, lmax=4, mask=mask)
    F2 = ps2.compute_fisher_contribution(5)
    F4 = ps4.compute_fisher_contribution(5)
    Ft, Bm = ps2.compute_theory_contribution(5, edges, lmax_theory=4)
    np.testing.assert_array_equal(Ft, F2)
    np.testing.assert_allclose(Bm, F4[:4, :], atol=1e-12 * np.abs(F4).max())
    fine = np.array([1.0, 1.5, 2.0, 2.4, 3.0, 3.5]) * kF + 1e-9
    _, Bf = ps2.compute_theory_contribution(5, fine, lmax_theory=2)
    Nf = len(fine) - 1
    for blk in range(2):
        coarse = np.stack([Bf[:, blk * Nf:blk * Nf + 2].sum(1), Bf[:, blk * Nf + 2:blk * Nf + 4].sum(1)], 1)
        np.testing.assert_allclose(coarse, F2[:, blk * 2:blk * 2 + 2], atol=1e-12 * np.abs(F2).max())
    # the bin past the data's k range sees Q maps spread by the mask
    assert np.abs(Bf[:, Nf - 1]).max() > 1e-3 * np.abs(F2).max()


def test_pk_theory_wideangle_dense():
    """Wide-angle theory matrix (lmax_theory = 4, pspec.py:359-367, 414-436): the odd legs are the
    harmonic transforms of A^-1 a / |r|, entering the l = 2 and l = 4 blocks with per-bin weights
    3i/5 (3 + kfac), 3i/5 (2 - kfac) and 10i/9 (5 + kfac); dense restatement."""
    n = 8
    N = n ** 3
    b = O.PolyBin3D(200., [n, n, n], boxcenter=[300., 140., 90.], sightline='local')
    mask = 0.5 + np.random.RandomState(6).rand(n, n, n)
    kF = b.kF[0]
    edges = np.array([1.0, 2.0, 3.0]) * kF + 1e-9
    th_edges = np.array([1.0, 1.6, 2.2, 3.1]) * kF + 1e-9
    ps = O.PSpec(b, edges, lmax=2, mask=mask)
    F, Bm = ps.compute_theory_contribution(4, th_edges, lmax_theory=4, include_wideangle=True)
    rows, Ainv, Ainv_r, modk, ph, cosang = _dense_q_rows(b, np.diag(mask.reshape(N)), edges, 4, (0, 2))
    xs = np.stack([g.reshape(N) for g in b.r_grids], -1)
    inv_r = 1. / np.linalg.norm(xs, axis=1)
    leg = {l: (ph * O.legendre(l, cosang)) @ (Ainv_r * (inv_r if l % 2 else 1.)) for l in range(1, 5)}
    kf = O.derivative_coefficient(th_edges)
    Nt = len(th_edges) - 1
    s = 0.5 * b.volume / N ** 2
    ref = np.zeros_like(Bm)
    for r, Q in enumerate(rows):
        for kb in range(Nt):
            th = (modk >= th_edges[kb]) & (modk < th_edges[kb + 1])
            X = lambda y: np.sum(th * y * np.conj(Q))  # noqa: E731
            ref[r, kb] = s * X(Ainv).real
            ref[r, Nt + kb] = s * (X(leg[2]).real + (3j / 5. * (3 + kf[kb]) * X(leg[1])).real
                                   + (3j / 5. * (2 - kf[kb]) * X(leg[3])).real)
            ref[r, 2 * Nt + kb] = s * (X(leg[4]).real + (10j / 9. * (5 + kf[kb]) * X(leg[3])).real)
    np.testing.assert_allclose(Bm, ref, atol=1e-10 * np.abs(ref).max(), rtol=1e-8)


@pytest.mark.parametrize("sightline", ["global", "local"])
def test_bk_covariance_equals_fisher_with_filter(sightline):
    """BSpec covariance contribution (bspec.py:447-519) with a constant mask m and trivial S^-1: its
    filter S^-1 [m^2 xi + m] S^-dag is the mode-space multiplier c = m^2 P(k, mu) + m, so it equals the
    MC Fisher contribution (unmasked) with the user filter S^-1 = c: the covariance code path against
    the Fisher one, in both weightings (the Q[A^-1 a] maps and the S^-1-weighted g maps).  Global:
    P_0, P_2 and P_4 (a mu-dependent c); local: P_0 (c is isotropic, the Yamamoto legs are exercised
    by lmax = 2)."""
    L, n = 1000., 16
    kw = {} if sightline == 'global' else dict(boxcenter=[300., 200., 500.], sightline='local')
    b = O.PolyBin3D(L, [n] * 3, **kw)
    edges = (2 * np.pi / L) * np.array([1., 2., 3., 4., 5.])
    m = 0.6
    kc = np.linspace(0., 0.2, 100)
    Pc = [kc, 3e3 / (1 + 40 * kc), 500. - 2e3 * kc, 80. + kc]
    if sightline == 'local':
        Pc = Pc[:2]
    bc = O.BSpec(b, edges, lmax=2, mask=m * np.ones((n, n, n)))
    C = bc.compute_covariance_contribution(2, Pc, mask_shot_2pt=np.ones((n, n, n)))
    P = np.interp(b.modk_grid, kc, Pc[1], right=0.)
    if sightline == 'global':
        mu = b.muk_grid
        P = P + O.legendre(2, mu) * np.interp(b.modk_grid, kc, Pc[2], right=0.) \
            + O.legendre(4, mu) * np.interp(b.modk_grid, kc, Pc[3], right=0.)
    c = m ** 2 * P + m

    def Sinv(x, input_type='real', output_type='real'):
        f = b.to_fourier(x) if input_type == 'real' else x
        f = f * c
        return b.to_real(f) if output_type == 'real' else f
    bf = O.BSpec(b, edges, lmax=2, applySinv=Sinv)
    F = bf.compute_fisher_contribution(2)
    np.testing.assert_allclose(C, F, atol=1e-11 * np.abs(F).max(), rtol=1e-9)
    assert np.abs(F).max() > 0


def test_pk_shot_contribution_closed_form_and_dense():
    """PSpec.compute_shot_contribution (pspec.py:495-551).  Global sightline, constant mask m, trivial
    S^-1: the GRF a is drawn with P = 1/P_fid and the numerator is
      shot[(l, b)] = V/(2N^2) sum_{k in b} degk m |a_k|^2 / P_A(k) L_l(mu),  P_A = 1/P_fid (interpolated,
    held at the table ends);
    local sightline with a shot-noise mask: the dense restatement
      shot[(l, b)] = V/(2N^2) sum_{k in b} Re[leg_l(k) conj(FFT[N IFFT(a / P_A)](k))],
    leg_0 = a, leg_l = sum_x e^{-ik.x} L_l(k.x) IFFT(a)(x)."""
    n = 16
    kk = np.linspace(0, 2, 400)
    Pk = [kk, 1e3 * (1 + kk)]
    b = O.PolyBin3D(600., [n, n, n], Pk=Pk)
    edges = np.linspace(0.012, 0.07, 5)
    m = 0.7
    ps = O.PSpec(b, edges, lmax=4, mask=m * np.ones((n, n, n)))
    shot = ps.compute_shot_contribution(4)
    PkA = [b.Pfid[0], 1. / b.Pfid[1]]
    a = b.generate_data(seed=4 + int(1e7), Pk_input=PkA, output_type='fourier')
    PA = np.interp(b.modk_grid, PkA[0], PkA[1])
    idx = O.bin_index(b.modk_grid, edges)
    ref = np.zeros_like(shot)
    for li in range(3):
        for kb in range(ps.Nk):
            sel = idx == kb
            w = b.degeneracy_factor[sel] * m * np.abs(a[sel]) ** 2 / PA[sel] * O.legendre(2 * li, b.muk_grid[sel])
            ref[li * ps.Nk + kb] = b.volume / (2 * n ** 6) * w.sum()
    np.testing.assert_allclose(shot, ref, rtol=1e-10, atol=1e-12 * np.abs(ref).max())
    # local sightline, varying mask and shot-noise mask: dense restatement
    n = 8
    N = n ** 3
    bl = O.PolyBin3D(200., [n, n, n], boxcenter=[300., 140., 90.], sightline='local')
    rng = np.random.RandomState(8)
    mask, mshot = 0.5 + rng.rand(n, n, n), 0.2 + rng.rand(n, n, n)
    e2 = np.array([1.0, 2.0, 3.0]) * bl.kF[0] + 1e-9
    psl = O.PSpec(bl, e2, lmax=2, mask=mask, mask_shot=mshot)
    shot = psl.compute_shot_contribution(3)
    ph, kv, cosang, fr = _dense_ops(bl)
    modk = np.linalg.norm(kv, axis=1)
    PkA = [bl.Pfid[0], 1. / bl.Pfid[1]]
    ah = bl.generate_data(seed=3 + int(1e7), Pk_input=PkA, output_type='fourier')
    a = np.fft.fftn(bl.to_real(ah)).reshape(N)
    ar = (ph.conj().T @ a).real / N
    PA = np.interp(modk, PkA[0], PkA[1])
    AiA_r = (ph.conj().T @ (a / PA)).real / N
    SNA = ph @ (mshot.reshape(N) * AiA_r)
    legs = {0: a, 2: (ph * O.legendre(2, cosang)) @ ar}
    ref = np.zeros(2 * 2)
    for li, ell in enumerate((0, 2)):
        for kb in range(2):
            th = (modk >= e2[kb]) & (modk < e2[kb + 1])
            ref[li * 2 + kb] = 0.5 * bl.volume / N ** 2 * np.sum(th * (legs[ell] * np.conj(SNA)).real)
    np.testing.assert_allclose(shot, ref, rtol=1e-8, atol=1e-10 * np.abs(ref).max())
