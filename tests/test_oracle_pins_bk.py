"""Pins for the FP64 oracle's bispectrum Monte Carlo Fisher and shot noise (CPU only).

Both are checked per seed against restatements written from the estimator's mathematics, not from
the reference's loop structure:

  * BSpec._compute_fisher (PolyBin3D/bspec.py:473-676).  With g_{l,b} = K_{l,b} S the filtered maps of
    a real map S (K an explicit N x N matrix: DFT, bin indicator Theta_b, L_l(mu) for the global
    sightline or the addition-theorem kernel L_l(k_hat . x_hat') for the local one), the cubic form
    zeta_i(S) = sum_x g_{0,b1} g_{0,b2} g_{l,b3} has the gradient
        grad zeta_i = K_{0,b1}^T (g_{0,b2} g_{l,b3}) + K_{0,b2}^T (g_{0,b1} g_{l,b3}) + K_{l,b3}^T (g_{0,b1} g_{0,b2}).
    With two GRFs a1, a2 (seeds s+1e7, s+2e7), w^A = IFFT(A^-1 a) and w^S = P IFFT(a) (trivial S^-1,
    P the pointing matrix as a dense operator), the Monte Carlo row is
        F_ij = V / (2 N^2 sym_i sym_j) sum_k Theta_b2(i)(k) Re[ conj(FT h_i)(k) FT(P q_j)(k) ],
        h_i = [g_{0,b1} g_{l,b3}](w^S_1) - [same](w^S_2),   q_j = grad zeta_j(w^A_1) - grad zeta_j(w^A_2),
    summed over the full grid (no degeneracy factors).  Every product is an explicit matrix product,
    every transpose an explicit .T, so a dropped leg, a transposed operator, a missing mask or a wrong
    conjugation in the oracle changes the result.
  * BSpec.compute_shot_contribution (bspec.py:695-861).  Global sightline, constant masks: the cubic
    and quadratic terms as exact lattice sums N^-2 sum_{k1+k2+k3=0} F1(k1) F2(k2) F3(k3) over the six leg
    assignments, with the real-space products A1 A2 and A_i d2 as circular convolutions of their
    Fourier amplitudes.  Local sightline, varying mask and shot-noise mask: the same sums with the
    dense local filter matrices.  In expectation, for unit weights, the cubic term equals the closed
    form c T_i / (V sym_i), T_i the lattice triangle count of bin i weighted by L_l(mu_3).

Each pin is followed by planted mistakes in the oracle (monkeypatched, CPU only) that it must reject.
"""
import itertools

import numpy as np
import pytest

from oracle import polybin_oracle as O


def _L(ell, x):
    """Legendre polynomials written out here (independent of the oracle's table)."""
    if ell == 0:
        return np.ones_like(x)
    if ell == 2:
        return 1.5 * x ** 2 - 0.5
    if ell == 4:
        return (35. * x ** 4 - 30. * x ** 2 + 3.) / 8.
    raise ValueError(ell)


class Dense:
    """Full-grid DFT matrix E[k, x] = exp(-i k.x) (numpy fftn order, C-ordered voxels) and the
    geometry the filter matrices need."""

    def __init__(self, b):
        n = np.asarray(b.gridsize)
        self.n = n
        self.N = N = int(np.prod(n))
        fr = np.stack(np.meshgrid(*[np.fft.fftfreq(m, 1. / m) for m in n], indexing='ij'), -1).reshape(N, 3)
        idx = np.stack(np.meshgrid(*[np.arange(m) for m in n], indexing='ij'), -1).reshape(N, 3)
        self.E = np.exp(-2j * np.pi * (fr / n) @ idx.T)
        self.fr = fr.astype(int)
        kv = fr * np.asarray(b.kF)
        self.modk = np.linalg.norm(kv, axis=1)
        nz = self.modk > 0
        self.khat = np.zeros_like(kv)
        self.khat[nz] = kv[nz] / self.modk[nz, None]
        self.mu = self.khat @ np.asarray(b.sight_vector)
        self.local = b.sightline == 'local'
        if self.local:
            xs = np.stack(np.meshgrid(*[np.fft.fftshift(np.arange(-m // 2, m // 2)) * b.boxsize[i] / m
                                        + b.boxcenter[i] for i, m in enumerate(n)], indexing='ij'), -1).reshape(N, 3)
            self.xhat = xs / np.linalg.norm(xs, axis=1)[:, None]
            self.cos_kx = self.khat @ self.xhat.T  # [k, x']

    def full(self, half):
        """Full-grid values of a real, k -> -k symmetric half-grid table."""
        f = self.fr
        n = self.n
        neg = f[:, 2] < 0
        g = np.where(neg[:, None], -f, f)
        return np.asarray(half)[g[:, 0] % n[0], g[:, 1] % n[1], g[:, 2]]

    def fourier_full(self, half_map):
        """Full-grid Fourier amplitudes of a half-grid map (numpy's C2R projection, then fftn)."""
        r = np.fft.irfftn(half_map, s=tuple(self.n), axes=(0, 1, 2))
        return np.fft.fftn(r).reshape(self.N)

    def theta(self, lo, hi):
        return ((self.modk >= lo) & (self.modk < hi)).astype(np.float64)

    def K(self, ell, lo, hi):
        """Real N x N matrix of x <- y: (1/N) sum_k e^{ik.x} Theta(k) W(k, y) e^{-ik.y}, W = L_l(mu_k)
        (global) or L_l(k_hat . y_hat) (local); W = 1 for l = 0."""
        th = self.theta(lo, hi)
        if ell == 0:
            W = th[:, None] * self.E
        elif self.local:
            W = th[:, None] * _L(ell, self.cos_kx) * self.E
        else:
            W = (th * _L(ell, self.mu))[:, None] * self.E
        return (self.E.conj().T @ W).real / self.N


def bk_fisher_dense(bs, seed, P):
    """Per-seed Monte Carlo Fisher of the BSpec `bs` (trivial S^-1), P the dense pointing matrix."""
    b = bs.base
    if not hasattr(bs, 'sym_factor'):
        bs._compute_symmetry_factor()
    D = Dense(b)
    N = D.N
    ks = bs.k_bins_squeeze
    Nq = len(ks) - 1
    K = {(l, q): D.K(l, ks[q], ks[q + 1]) for l in range(0, bs.lmax + 1, 2) for q in range(Nq)}
    inv = D.full(b.invPk0_grid)
    ws, wa = [], []
    for s in (seed + int(1e7), seed + int(2e7)):
        a = D.fourier_full(b.generate_data(seed=s, output_type='fourier'))
        ar = (D.E.conj().T @ a).real / N
        wa.append((D.E.conj().T @ (inv * a)).real / N)   # A^-1 a (base.py:364-389)
        ws.append(P @ ar)                                 # S^-1 P a, trivial S^-1
    ga = [{key: Kq @ w for key, Kq in K.items()} for w in wa]
    gs = [{key: Kq @ w for key, Kq in K.items()} for w in ws]

    def grad(g, b1, b2, b3, l):
        return (K[0, b1].T @ (g[0, b2] * g[l, b3]) + K[0, b2].T @ (g[0, b1] * g[l, b3])
                + K[l, b3].T @ (g[0, b1] * g[0, b2]))

    ab = bs.all_bins
    Q = np.asarray([grad(ga[0], *t) - grad(ga[1], *t) for t in ab])          # [N_bins, N] real
    QP = (D.E @ (P @ Q.T)).T                                                  # FT(P q_j)
    H = np.asarray([gs[0][0, b1] * gs[0][l, b3] - gs[1][0, b1] * gs[1][l, b3] for b1, _, b3, l in ab])
    FH = (D.E @ H.T).T
    F = np.zeros((len(ab), len(ab)))
    for i, (_, b2, _, _) in enumerate(ab):
        th = D.theta(ks[b2], ks[b2 + 1])
        F[i] = (np.conj(FH[i])[None, :] * th[None, :] * QP).sum(1).real
    sym = bs.sym_factor
    return F * b.volume / (2. * N ** 2) / np.outer(sym, sym)


def _gic_matrix(mask, mask_ic):
    """Dense P of the global integral constraint (utils.pyx:405-423): P m = mask (m - <m mask_IC> / <mask_IC>)."""
    m, mi = mask.reshape(-1), mask_ic.reshape(-1)
    return np.diag(m) @ (np.eye(len(m)) - np.outer(np.ones(len(m)), mi) / mi.sum())


def _rel(a, b):
    d = np.sqrt(np.abs(np.outer(np.diag(b), np.diag(b))))
    return float(np.max(np.abs(a - b) / np.maximum(d, 1e-300)))


def _bk_case(kind):
    """(BSpec, dense P) for the three Fisher cases on an 8^3 grid, 3 bins, lmax = 2."""
    n = 8
    kk = np.linspace(0., 2., 2000)
    Pk = np.zeros_like(kk)
    Pk[1:] = 3e3 * (kk[1:] / 0.05) ** -1.2
    rng = np.random.RandomState(11)
    mask = 0.4 + rng.rand(n, n, n)
    if kind == 'global_mask':
        b = O.PolyBin3D(200., [n] * 3, Pk=[kk, Pk])
        bs = O.BSpec(b, np.array([1.0, 1.8, 2.6, 3.4]) * b.kF[0], lmax=2, mask=mask, Pk_fid='default')
        return bs, np.diag(mask.reshape(-1))
    if kind == 'local_mask':
        b = O.PolyBin3D(200., [n] * 3, Pk=[kk, Pk], boxcenter=[300., 140., 90.], sightline='local')
        bs = O.BSpec(b, np.array([1.0, 1.8, 2.6, 3.4]) * b.kF[0], lmax=2, mask=mask, Pk_fid='default')
        return bs, np.diag(mask.reshape(-1))
    if kind == 'local_gic':
        b = O.PolyBin3D(200., [n] * 3, boxcenter=[300., 140., 90.], sightline='local')
        mic = 0.2 + rng.rand(n, n, n)
        bs = O.BSpec(b, np.array([1.0, 1.8, 2.6, 3.4]) * b.kF[0], lmax=2, mask=mask, add_GIC=True, mask_IC=mic)
        return bs, _gic_matrix(mask, mic)
    raise ValueError(kind)


@pytest.mark.parametrize("kind", ["global_mask", "local_mask", "local_gic"])
def test_bk_mc_fisher_dense(kind):
    bs, P = _bk_case(kind)
    F = bs.compute_fisher_contribution(3)
    ref = bk_fisher_dense(bs, 3, P)
    err = _rel(F, ref)
    print("Bk MC Fisher (%s, %d bins): max |F - F_dense| / sqrt(F_ii F_jj) = %.2e" % (kind, bs.N_bins, err))
    assert bs.N_bins >= 12 and np.abs(ref).max() > 0
    assert err < 1e-9


def _plant_fisher(bs, which, monkeypatch):
    """A plausible mistake planted in the oracle object (never in product code)."""
    if which == 'dropped_legendre_leg':
        orig = O.BSpec._filt
        monkeypatch.setattr(bs, '_filt', lambda m, l, q: orig(bs, m, 0, q))
    elif which == 'no_mask':
        monkeypatch.setattr(bs, 'apply_pointing', lambda m, transpose=False: m)
    elif which == 'transposed_pointing':
        monkeypatch.setattr(bs, 'apply_pointing', lambda m, transpose=False: O.apply_pointing_ic(bs, m, True))
    elif which == 'conjugated_row':
        integ = bs.base.integrator
        monkeypatch.setattr(integ, 'integrate_row', lambda m, arr, lo, hi: O.Integrator.integrate_row(
            integ, np.conj(m) * 1j, arr, lo, hi))
    elif which == 'transposed_harmonics':
        b = bs.base
        monkeypatch.setattr(b, 'apply_real_harmonics', lambda a_f, Yr, Yf: b.to_real(
            O.PolyBin3D.apply_fourier_harmonics(b, b.to_real(a_f), Yr, Yf)))
    elif which == 'ainv_without_weight':
        monkeypatch.setattr(bs.base, 'applyAinv', lambda d, invPk0_grid=None, input_type='real',
                            output_type='real': d)


@pytest.mark.parametrize("kind,which", [("global_mask", "dropped_legendre_leg"), ("global_mask", "no_mask"),
                                        ("global_mask", "conjugated_row"), ("global_mask", "ainv_without_weight"),
                                        ("local_gic", "transposed_pointing"), ("local_mask", "transposed_harmonics")])
def test_bk_mc_fisher_dense_rejects_planted_mistakes(kind, which, monkeypatch):
    bs, P = _bk_case(kind)
    ref = bk_fisher_dense(bs, 3, P)
    _plant_fisher(bs, which, monkeypatch)
    F = bs.compute_fisher_contribution(3)
    err = _rel(F, ref)
    print("planted %s (%s): max relative deviation %.2e" % (which, kind, err))
    assert err > 1e-3


# ---------------------------------------------------------------- shot noise
def _perm_sum(all_bins, sets):
    """sum over the six leg assignments of N^-2 sum_{k1+k2+k3=0} F[s1][0,b1](k1) F[s2][0,b2](k2) F[s3][l,b3](k3)
    for three map sets (full-grid Fourier amplitudes of their filtered maps), per triangle bin."""
    A, B, Cm = sets
    out = np.zeros(len(all_bins))
    for i, (b1, b2, b3, l) in enumerate(all_bins):
        tot = 0.
        for x, y, z in itertools.permutations((A, B, Cm)):
            tot += _tri(x[0, b1], y[0, b2], z[l // 2, b3])
        out[i] = tot
    return out


def _tri(fa, fb, fc):
    """N^-2 sum_{k1+k2+k3 = 0 (mod n)} fa(k1) fb(k2) fc(k3) over full-grid arrays [n0, n1, n2]."""
    n = fa.shape
    N = fa.size
    tot = 0.
    i0, i1, i2 = (np.arange(m) for m in n)
    for p in zip(*np.nonzero(fa)):
        mi = (-p[0] - i0[:, None, None]) % n[0]
        mj = (-p[1] - i1[None, :, None]) % n[1]
        mk = (-p[2] - i2[None, None, :]) % n[2]
        tot += fa[p] * np.sum(fb * fc[mi, mj, mk])
    return (tot / N ** 2).real


def _conv(fa, fb):
    """Fourier amplitudes of the real-space product: (1/N) sum_q fa(q) fb(k - q), circular."""
    n = fa.shape
    N = fa.size
    out = np.zeros(n, dtype=np.complex128)
    i0, i1, i2 = (np.arange(m) for m in n)
    for p in zip(*np.nonzero(fa)):
        mi = (i0[:, None, None] - p[0]) % n[0]
        mj = (i1[None, :, None] - p[1]) % n[1]
        mk = (i2[None, None, :] - p[2]) % n[2]
        out += fa[p] * fb[mi, mj, mk]
    return out / N


def _shot_global_case():
    n = 8
    kk = np.linspace(0., 2., 2000)
    b = O.PolyBin3D(200., [n] * 3, Pk=[kk, 4e2 * (1. + 3. * kk)])
    edges = np.array([1.0, 1.8, 2.6, 3.4]) * b.kF[0]
    c_mask, c_shot = 0.8, 0.6
    bs = O.BSpec(b, edges, lmax=2, mask=c_mask * np.ones((n,) * 3), mask_shot=c_shot * np.ones((n,) * 3))
    d = np.random.RandomState(4).randn(n, n, n)
    d2 = 0.5 + np.random.RandomState(5).rand(n, n, n)
    return bs, d, d2, c_shot


def bk_shot_global_lattice(bs, seed, d, d2, c_shot):
    """Cubic and quadratic shot-noise terms (global sightline, constant masks, trivial S^-1) as
    Fourier-space lattice sums; returns (cubic, quadratic - 2 cubic)."""
    b = bs.base
    if not hasattr(bs, 'sym_factor'):
        bs._compute_symmetry_factor()
    n = tuple(int(v) for v in b.gridsize)
    N = int(np.prod(n))
    D = Dense(b)
    ks = bs.k_bins_squeeze
    Nq = len(ks) - 1
    filt = {(l, q): (D.theta(ks[q], ks[q + 1]) * _L(l, D.mu)).reshape(n)
            for l in range(0, bs.lmax + 1, 2) for q in range(Nq)}

    def gset(fk):
        out = np.zeros((bs.Nl, Nq) + n, dtype=np.complex128)
        for (l, q), w in filt.items():
            out[l // 2, q] = w * fk
        return out
    PkA = [b.Pfid[0], 1. / b.Pfid[1]]
    Pfid = D.full(b.Pk0_grid).reshape(n)
    a = [D.fourier_full(b.generate_data(seed=seed + int(s), Pk_input=PkA, output_type='fourier')).reshape(n)
         for s in (1e7, 2e7)]
    A = [Pfid * x for x in a]                                  # A^-1 a with A^-1 = P_fid (bspec.py:780-785)
    C = c_shot * _conv(A[0], A[1])
    fd = np.fft.fftn(d)
    fd2 = np.fft.fftn(d2)
    gA, gB, gC = gset(a[0]), gset(a[1]), gset(C)
    gD, gE, g = gset(_conv(A[0], fd2)), gset(_conv(A[1], fd2)), gset(fd)
    norm = b.volume / N / bs.sym_factor
    cubic = _perm_sum(bs.all_bins, (gA, gB, gC)) / 6. * norm
    quad = 0.25 * norm * (_perm_sum(bs.all_bins, (g, gA, gD)) + _perm_sum(bs.all_bins, (g, gB, gE)))
    return cubic, quad - 2. * cubic


def test_bk_shot_global_lattice_sums():
    bs, d, d2, c_shot = _shot_global_case()
    cub = bs.compute_shot_contribution(2, cubic_only=True)
    tot = bs.compute_shot_contribution(2, data=d, data2=d2)
    rc, rt = bk_shot_global_lattice(bs, 2, d, d2, c_shot)
    ec = np.max(np.abs(cub - rc)) / np.max(np.abs(rc))
    et = np.max(np.abs(tot - rt)) / np.max(np.abs(rt))
    print("Bk shot noise (global, constant masks): cubic %.2e, quadratic - 2 cubic %.2e (relative to max)" % (ec, et))
    assert ec < 1e-10 and et < 1e-10


def bk_shot_local_dense(bs, seed, d, d2):
    """Cubic and total shot-noise terms with the dense (local) filter matrices and real-space masks."""
    b = bs.base
    if not hasattr(bs, 'sym_factor'):
        bs._compute_symmetry_factor()
    D = Dense(b)
    N = D.N
    ks = bs.k_bins_squeeze
    Nq = len(ks) - 1
    K = {(l, q): D.K(l, ks[q], ks[q + 1]) for l in range(0, bs.lmax + 1, 2) for q in range(Nq)}

    def gset(x):
        out = np.zeros((bs.Nl, Nq, N))
        for (l, q), Kq in K.items():
            out[l // 2, q] = Kq @ x
        return out
    PkA = [b.Pfid[0], 1. / b.Pfid[1]]
    Pfid = D.full(b.Pk0_grid)
    a = [D.fourier_full(b.generate_data(seed=seed + int(s), Pk_input=PkA, output_type='fourier')) for s in (1e7, 2e7)]
    ar = [(D.E.conj().T @ x).real / N for x in a]
    A = [(D.E.conj().T @ (Pfid * x)).real / N for x in a]
    ms = bs.mask_shot.reshape(-1)
    dd, dd2 = d.reshape(-1), d2.reshape(-1)
    gA, gB, gC = gset(ar[0]), gset(ar[1]), gset(A[0] * A[1] * ms)
    gD, gE, g = gset(A[0] * dd2), gset(A[1] * dd2), gset(dd)

    def bshot(X, Y, Z):
        out = np.zeros(bs.N_bins)
        for i, (b1, b2, b3, l) in enumerate(bs.all_bins):
            out[i] = sum(np.sum(x[0, b1] * y[0, b2] * z[l // 2, b3]) for x, y, z in itertools.permutations((X, Y, Z)))
        return out
    norm = b.volume / N / bs.sym_factor
    cubic = bshot(gA, gB, gC) / 6. * norm
    quad = 0.25 * norm * (bshot(g, gA, gD) + bshot(g, gB, gE))
    return cubic, quad - 2. * cubic


def _shot_local_case():
    n = 8
    kk = np.linspace(0., 2., 2000)
    b = O.PolyBin3D(200., [n] * 3, Pk=[kk, 4e2 * (1. + 3. * kk)], boxcenter=[300., 140., 90.], sightline='local')
    rng = np.random.RandomState(9)
    mask, mshot = 0.4 + rng.rand(n, n, n), 0.2 + rng.rand(n, n, n)
    bs = O.BSpec(b, np.array([1.0, 1.8, 2.6, 3.4]) * b.kF[0], lmax=2, mask=mask, mask_shot=mshot)
    d = np.random.RandomState(4).randn(n, n, n) * mask
    d2 = 0.5 + np.random.RandomState(5).rand(n, n, n)
    return bs, d, d2


def test_bk_shot_local_dense():
    bs, d, d2 = _shot_local_case()
    cub = bs.compute_shot_contribution(3, cubic_only=True)
    tot = bs.compute_shot_contribution(3, data=d, data2=d2)
    rc, rt = bk_shot_local_dense(bs, 3, d, d2)
    ec = np.max(np.abs(cub - rc)) / np.max(np.abs(rc))
    et = np.max(np.abs(tot - rt)) / np.max(np.abs(rt))
    print("Bk shot noise (local, varying masks): cubic %.2e, quadratic - 2 cubic %.2e (relative to max)" % (ec, et))
    assert ec < 1e-9 and et < 1e-9


def test_bk_shot_cubic_expectation_closed_form():
    """Unit fiducial spectrum, constant shot mask c, global sightline: E[cubic_i] = c T_i / (V sym_i) with
    T_i = sum_{k1+k2+k3=0} Theta_b1 Theta_b2 Theta_b3 L_l(mu_3) (the GRFs have P_A = 1/P_fid = 1, so
    <g_A(x) A_1(y)> = V^-1 sum_k Theta e^{ik(x-y)}).  Mean of 150 seeds within 5 standard errors."""
    n = 8
    b = O.PolyBin3D(200., [n] * 3)
    c = 0.7
    bs = O.BSpec(b, np.array([1.0, 1.8, 2.6, 3.4]) * b.kF[0], lmax=2, mask_shot=c * np.ones((n,) * 3))
    bs._compute_symmetry_factor()
    D = Dense(b)
    ks = bs.k_bins_squeeze
    sh = tuple(int(v) for v in b.gridsize)
    want = []
    for b1, b2, b3, l in bs.all_bins:
        t1 = D.theta(ks[b1], ks[b1 + 1]).reshape(sh)
        t2 = D.theta(ks[b2], ks[b2 + 1]).reshape(sh)
        t3 = (D.theta(ks[b3], ks[b3 + 1]) * _L(l, D.mu)).reshape(sh)
        want.append(_tri(t1, t2, t3) * D.N ** 2)
    want = c * np.asarray(want) / b.volume / bs.sym_factor
    import contextlib
    import io
    with contextlib.redirect_stdout(io.StringIO()):
        s = np.asarray([bs.compute_shot_contribution(seed, cubic_only=True) for seed in range(150)])
    m, se = s.mean(0), s.std(0) / np.sqrt(len(s))
    z = np.abs(m - want) / se
    print("Bk shot cubic expectation: max deviation %.2f standard errors" % z.max())
    assert z.max() < 5.
    assert np.all(np.abs(want) > 0)


def _plant_shot(which, monkeypatch):
    if which == 'dropped_permutation':
        def bad(gA, gB, gC, all_bins):
            out = np.zeros(len(all_bins))
            for i, (b1, b2, b3, l) in enumerate(all_bins):
                out[i] = (np.sum(gA[0][b1] * gB[0][b2] * gC[l // 2][b3]) + np.sum(gA[0][b1] * gC[0][b2] * gB[l // 2][b3])
                          + np.sum(gB[0][b1] * gA[0][b2] * gC[l // 2][b3]) + np.sum(gB[0][b1] * gC[0][b2] * gA[l // 2][b3])
                          + np.sum(gC[0][b1] * gA[0][b2] * gB[l // 2][b3]))
            return out
        monkeypatch.setattr(O, 'assemble_bshot', bad)
    elif which == 'legendre_on_wrong_leg':
        def bad(gA, gB, gC, all_bins):
            out = np.zeros(len(all_bins))
            for i, (b1, b2, b3, l) in enumerate(all_bins):
                for x, y, z in itertools.permutations((gA, gB, gC)):
                    out[i] += np.sum(x[l // 2][b1] * y[0][b2] * z[0][b3])
            return out
        monkeypatch.setattr(O, 'assemble_bshot', bad)


@pytest.mark.parametrize("which", ["dropped_permutation", "legendre_on_wrong_leg"])
def test_bk_shot_pins_reject_planted_mistakes(which, monkeypatch):
    bs, d, d2 = _shot_local_case()
    rc, rt = bk_shot_local_dense(bs, 3, d, d2)
    bg, dg, d2g, c_shot = _shot_global_case()
    gc, gt = bk_shot_global_lattice(bg, 2, dg, d2g, c_shot)
    _plant_shot(which, monkeypatch)
    tot = bs.compute_shot_contribution(3, data=d, data2=d2)
    totg = bg.compute_shot_contribution(2, data=dg, data2=d2g)
    el = np.max(np.abs(tot - rt)) / np.max(np.abs(rt))
    eg = np.max(np.abs(totg - gt)) / np.max(np.abs(gt))
    print("planted %s: local %.2e, global %.2e" % (which, el, eg))
    assert el > 1e-3 and eg > 1e-3


@pytest.mark.parametrize("kind", ["global_mask", "local_mask", "local_gic"])
def test_bk_fisher_subblock_equals_full(kind):
    """BSpec.fisher_subblock (the restricted oracle used at 256^3, where the full Q-map array does not
    fit) equals the rows x columns of the full per-seed oracle Fisher, for scattered rows and columns
    of both multipoles."""
    bs, _ = _bk_case(kind)
    F = bs.compute_fisher_contribution(5)
    rng = np.random.RandomState(2)
    rows = np.sort(rng.choice(bs.N_bins, 7, replace=False))
    cols = np.sort(rng.choice(bs.N_bins, 9, replace=False))
    sub = bs.fisher_subblock(5, rows, cols)
    ref = F[np.ix_(rows, cols)]
    err = np.max(np.abs(sub - ref)) / np.max(np.abs(ref))
    print("Bk Fisher sub-block (%s): %.2e" % (kind, err))
    assert err < 1e-12


@pytest.mark.parametrize("kind", ["global_mask", "local_mask"])
def test_bk_numerator_magnitude_scale(kind):
    """The oracle's numerator with return_scale (the sums over |g| maps that GPU tests scale by)
    returns the plain call's values, and each |g| sum bounds its entry (triangle inequality), also
    with the linear term."""
    bs, _ = _bk_case(kind)
    bs.generate_sims(2)
    d = np.random.RandomState(8).randn(*bs.base.gridsize)
    for lin in (False, True):
        plain = bs._compute_bk_numerator(d, include_linear_term=lin, filtering='Sinv')
        val, mag = bs._compute_bk_numerator(d, include_linear_term=lin, filtering='Sinv', return_scale=True)
        assert np.array_equal(plain, val)
        assert np.all(mag >= np.abs(val) * (1 - 1e-12)) and np.all(mag > 0)
