"""GPU parity of the hand-written R2C/C2R FFTs and their fused prologues/epilogues against numpy FP64
(numpy/pocketfft conventions, the ones the reference's backends follow).  Tolerance: FP32 transforms,
relative L2 error < 2e-6 * log2(N) and max error < 1e-5 of the field's max modulus."""
import numpy as np
import pytest
import torch

from oracle import polybin_oracle as O

pytestmark = pytest.mark.gpu


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel())


SHAPES = [(2, 2, 4), (4, 8, 16), (8, 8, 8), (16, 32, 8), (32, 16, 64), (64, 64, 64), (128, 128, 128),
          (16, 256, 32), (256, 8, 512), (512, 4, 4)]


def make_base(n, **kw):
    import polybin3d_amd as pb
    return pb.PolyBin3D(1000., list(n), **kw)


@pytest.mark.parametrize("n", SHAPES, ids=lambda n: "x".join(map(str, n)))
def test_r2c_c2r_match_numpy(hip, n):
    b = make_base(n)
    x = np.random.RandomState(0).randn(*n)
    X = b.to_fourier(x).cpu().numpy()
    ref = np.fft.rfftn(x)
    tol = 2e-6 * np.log2(np.prod(n))
    assert rel(X, ref) < tol
    # C2R of a non-Hermitian half grid: numpy semantics (Im of kz=0 / Nyquist planes ignored)
    Y = (np.random.RandomState(1).randn(*ref.shape) + 1j * np.random.RandomState(2).randn(*ref.shape))
    y = b.to_real(Y).cpu().numpy()
    yr = np.fft.irfftn(Y, s=n, axes=(0, 1, 2))
    assert rel(y, yr) < tol
    assert b.n_FFTs_forward == 1 and b.n_FFTs_reverse == 1


def test_batched_and_counters(hip):
    n = (16, 16, 32)
    b = make_base(n)
    x = np.random.RandomState(3).randn(5, *n)
    X = b.r2c(b.as_real(x)).cpu().numpy()
    assert rel(X, np.fft.rfftn(x, axes=(1, 2, 3))) < 1e-5
    xb = b.c2r(torch.from_numpy(X).cuda()).cpu().numpy()
    assert rel(xb, x) < 1e-5
    assert (b.n_FFTs_forward, b.n_FFTs_reverse) == (5, 5)


def test_fused_ops(hip):
    from polybin3d_amd import _capi as C
    n = (16, 32, 16)
    b = make_base(n, boxcenter=[500., 300., 200.], sightline='local')
    o = O.PolyBin3D(1000., list(n), boxcenter=[500., 300., 200.], sightline='local')
    rng = np.random.RandomState(4)
    a, m, c, d = (rng.randn(*n) for _ in range(4))
    A, M, Cc, D = (b.as_real(v) for v in (a, m, c, d))
    # prologue MUL and DIFF
    got = b.r2c(None, rop=C.rop(C.ROP_MUL, a=A, b=M)).cpu().numpy()
    assert rel(got, np.fft.rfftn(a * m)) < 1e-5
    got = b.r2c(None, rop=C.rop(C.ROP_DIFF, a=A, b=M, c=Cc, d=D)).cpu().numpy()
    assert rel(got, np.fft.rfftn(a * m - c * d)) < 1e-5
    # Y_lm prologue + Y_lm accumulate epilogue = apply_fourier_harmonics
    Yr = O.real_harmonics(np.asarray(o.r_grids), 4, False)
    Yk = O.real_harmonics(np.asarray(o.k_grids), 4, False)
    for ell in (2, 4):
        got = b.apply_fourier_harmonics(A, ell, ell).cpu().numpy()
        ref = o.apply_fourier_harmonics(a, Yr[ell], Yk[ell])
        assert rel(got, ref) < 2e-5
        F = np.fft.rfftn(a)
        got = b.apply_real_harmonics(torch.from_numpy(F).cuda(), ell, ell).cpu().numpy()
        ref = o.apply_real_harmonics(F, Yr[ell], Yk[ell])
        assert rel(got, ref) < 2e-5
    # filter prologue (Theta_b L_l w) broadcast over a batch of bins, mask epilogue
    edges = np.linspace(0.01, 0.045, 6)
    tab = b.bin_table(edges)
    bins = torch.arange(5, dtype=torch.int32, device='cuda')
    F = np.fft.rfftn(a)
    Fd = torch.from_numpy(F.astype(np.complex64)).cuda()
    for ell in (0, 2):
        out = b.c2r(Fd, batch=5, kop=C.kop(C.KOP_FILTER, bin_table=tab, bins=bins, ell=ell, w=b.invPk0_grid,
                                            x_batch_stride=0),
                    rop=C.rop(C.ROP_MUL, b=M)).cpu().numpy()
        for q in range(5):
            ref = o.to_real(o.map_utils.fourier_filter(F, ell, edges[q], edges[q + 1])) * m
            assert rel(out[q], ref) < 2e-5
    # WEIGHT epilogue with scale, and accumulate
    w = b.invPk0_grid
    out = b.r2c(A, kop=C.kop(C.KOP_WEIGHT, w=w, scale=0.5))
    b.r2c(A, out=out, kop=C.kop(C.KOP_PLAIN, accumulate=1))
    assert rel(out.cpu().numpy(), 1.5 * np.fft.rfftn(a)) < 1e-5


def test_grf_matches_reference_stream(hip):
    n = (16, 16, 16)
    Pk = [np.linspace(0, 2, 500), 3e3 * np.exp(-np.linspace(0, 2, 500))]
    b = make_base(n, Pk=Pk, pixel_window='cic')
    o = O.PolyBin3D(1000., list(n), Pk=Pk, pixel_window='cic')
    for seed in (0, 17):
        got = b.generate_data(seed=seed, output_type='fourier').cpu().numpy()
        ref = o.generate_data(seed=seed, output_type='fourier')
        assert rel(got, ref) < 1e-6
    got = b.generate_data(seed=5, output_type='real').cpu().numpy()
    assert rel(got, o.generate_data(seed=5, output_type='real')) < 1e-5


def test_device_grf_statistics(hip):
    """rng='device': same recipe as the reference, counter-based normals: E|a_k|^2 = P N^2/V,
    a real field (C2R then R2C returns the input), DC zero, reproducible per seed."""
    n = 64
    b = make_base((n, n, n), rng='device')
    a = [b.generate_data(seed=s, output_type='fourier') for s in range(4)]
    p = torch.stack([x.abs() ** 2 for x in a]).mean().item()
    assert abs(p / (n ** 6 / b.volume) - 1) < 0.01
    assert a[0][0, 0, 0].abs().item() == 0
    r = b.to_real(a[1])
    assert rel(b.to_fourier(r).cpu().numpy(), a[1].cpu().numpy()) < 1e-5
    assert torch.equal(a[2], b.generate_data(seed=2, output_type='fourier'))
    assert not torch.equal(a[2], a[3])


# (shape, kmax): the second has 40 z planes in the band limit against 64-column y-pass tiles, so the
# y passes run plane-aligned tiles over compact work maps
@pytest.mark.parametrize("n,kmax", [((64, 32, 64), 0.09), ((64, 256, 128), 0.25)], ids=["small", "aligned_y_tiles"])
def test_pruned_transforms(hip, n, kmax):
    """Band-limited C2R (input zero outside the support) and R2C with packed in-support output match
    numpy; the pruned passes skip only known-zero or unused data."""
    from polybin3d_amd import _capi as C
    b = make_base(n)
    o = O.PolyBin3D(1000., list(n))
    sup = b.support(kmax)
    rng = np.random.RandomState(5)
    X = rng.randn(*o.modk_grid.shape) + 1j * rng.randn(*o.modk_grid.shape)
    X[o.modk_grid >= kmax] = 0.
    Xd = torch.from_numpy(X.astype(np.complex64)).cuda()
    got = b.c2r(Xd, kop=C.kop(C.KOP_PLAIN, support=sup, x_batch_stride=b.H)).cpu().numpy()
    assert rel(got, np.fft.irfftn(X, s=n, axes=(0, 1, 2))) < 1e-5
    x = rng.randn(*n)
    edges = np.array([0.01, 0.05, kmax])
    idx = b.bin_index_host(edges).reshape(-1)
    modes = np.nonzero(idx >= 0)[0]
    modes = modes[np.argsort(idx[modes], kind='stable')]
    pi = np.full(b.H, -1, dtype=np.int32)
    pi[modes] = np.arange(len(modes), dtype=np.int32)
    pid = torch.from_numpy(pi).cuda()
    out = torch.zeros(2 * len(modes), dtype=torch.complex64, device='cuda')
    xd = torch.from_numpy(np.stack([x, 2 * x]).astype(np.float32)).cuda()
    b.r2c(xd, out=out, kop=C.kop(C.KOP_PACK, pack_index=pid, packed_len=len(modes), support=sup))
    ref = np.fft.rfftn(x).reshape(-1)[modes]
    got = out.cpu().numpy().reshape(2, -1)
    assert rel(got[0], ref) < 1e-5 and rel(got[1], 2 * ref) < 1e-5


# (shape, l of the Y_lm sub-maps, box centre): row lengths n2/2 = 16, 128, 32, 64, 256 and every weight
# form (l = 1..4); (4, 4, 256) has fewer rows than one block holds; (16, 256, 64) runs the forward y pass
# on plane-aligned tiles (33 z planes against 64-column tiles).  Even l on n2 = 128, 256, 512 runs the
# combined rows (Lmax+1 row transforms) in mode 1.  A centre with x = y = 0 gives rows through the line
# of sight (they keep the per-sub-map form); z = 5 * 1000/256 also puts r = 0 exactly on a voxel (the
# reference refuses a zero centre with a local sightline).
FAR = (400., 250., -300.)
POINT_CASES = [((32, 32, 32), 2, FAR), ((4, 4, 256), 2, FAR), ((16, 8, 64), 2, FAR), ((8, 16, 128), 2, FAR),
               ((8, 8, 512), 2, FAR), ((32, 32, 32), 1, FAR), ((32, 32, 32), 3, FAR), ((32, 32, 32), 4, FAR),
               ((8, 16, 128), 4, FAR), ((16, 256, 64), 2, FAR), ((16, 8, 256), 2, FAR), ((8, 8, 256), 4, FAR),
               ((16, 16, 256), 2, (0., 0., 19.53125)), ((16, 16, 256), 4, (0., 0., 37.)), ((8, 8, 512), 4, FAR)]


def _case_id(c):
    if isinstance(c, tuple) and len(c) == 3 and isinstance(c[0], float):
        return "far" if c == FAR else "c%g_%g_%g" % c
    return "x".join(map(str, c)) if isinstance(c, tuple) else "l%d" % c


@pytest.mark.parametrize("mode", [1, 0], ids=["combined", "per_submap"])
@pytest.mark.parametrize("n,L,center", POINT_CASES, ids=_case_id)
def test_fused_pointing_chain(hip, n, L, center, mode):
    """pb3d_fft_pointing = R2C(mask * scale * sum_g coef_g Y_g(x) C2R(kop_g(in))) composed from numpy,
    in both row modes (pb3d_pointing_rows_mode)."""
    from polybin3d_amd import _capi as C
    center = list(center)
    b = make_base(n, boxcenter=center, sightline='local')
    o = O.PolyBin3D(1000., list(n), boxcenter=center, sightline='local')
    rng = np.random.RandomState(6)
    A = np.fft.rfftn(rng.randn(*n))
    mask = rng.rand(*n)
    edges = np.linspace(0.01, 0.08, 5)
    tab = b.bin_table(edges)
    bins = torch.arange(4, dtype=torch.int32, device='cuda')
    sup = b.support(edges[-1])
    Ad = torch.from_numpy(A.astype(np.complex64)).cuda()
    Md = b.as_real(mask)
    common = dict(bin_table=tab, bins=bins, x_batch_stride=0, support=sup)
    subs = [(C.kop(C.KOP_FILTER, ell=2, **common), Ad, 0, 0, 0.5)]
    subs += [(C.kop(C.KOP_YLM, ell=L, m=m, **common), Ad, L, m, 0.25) for m in range(2 * L + 1)]
    out = torch.empty((4,) + b.half_shape, dtype=torch.complex64, device='cuda')
    prev = C.lib().pb3d_pointing_rows_mode(mode)
    try:
        b.pointing(subs, 4, out, mask=Md, scale=1.5, kop_out=None)
        got = out.cpu().numpy()
    finally:
        C.lib().pb3d_pointing_rows_mode(prev)
    Yr = O.real_harmonics(np.asarray(o.r_grids), L, True)[L]
    Yk = O.real_harmonics(np.asarray(o.k_grids), L, True)[L]
    for q in range(4):
        r = 0.5 * o.to_real(o.map_utils.fourier_filter(A, 2, edges[q], edges[q + 1]))
        for m in range(2 * L + 1):
            r += 0.25 * Yr[m] * o.to_real(o.map_utils.prod_fourier_filter(A, Yk[m], edges[q], edges[q + 1]))
        ref = o.to_fourier(1.5 * mask * r)
        err = rel(got[q], ref)
        print("pointing %s l=%d centre=%s mode=%d bin %d: rel L2 %.2e" % (n, L, center, mode, q, err))
        assert err < 2e-5


def test_pointing_rows_mode_rejects_bad_values(hip):
    from polybin3d_amd import _capi as C
    assert C.lib().pb3d_pointing_rows_mode(2) == -1
    assert C.lib().pb3d_pointing_rows_mode(-1) == -1
    prev = C.lib().pb3d_pointing_rows_mode(1)
    assert prev in (0, 1)
    C.lib().pb3d_pointing_rows_mode(prev)


@pytest.mark.parametrize("n", [16, 32, 64, 128, 256, 512], ids=lambda n: "L%d" % n)
def test_bkq_prologue_and_packed_epilogue(hip, n):
    """The first C2R pass with the bispectrum Q-map prologue (table reads batched per thread) equals the
    element-wise evaluation of the same mode op (pb3d_kmap_packed) followed by a plain C2R, for every
    pencil length L = n0 of the x pass and every term set (the Theta terms of bspec.py:615-640 with
    and without the external-leg Legendre factor, and the Y_lm-weighted leg); the last R2C pass with
    packed output equals a plain R2C gathered at the in-bin modes."""
    import polybin3d_amd as pb
    from polybin3d_amd import _capi as C
    shape = [n, 8, 16] if n == 512 else [n] * 3
    L = 1000.
    b = pb.PolyBin3D(L, shape, boxcenter=[300., 200., 500.], sightline='local')
    kF = 2 * np.pi / L
    edges = kF * (np.array([1.0, 2.0, 3.0, 3.9]) if n == 512 else (n / 2) * np.array([0.15, 0.35, 0.6, 0.85]))
    bs = pb.BSpec(b, edges, lmax=2)
    pk = bs._packed_layout()
    M, Nq = pk['M'], bs.Nk_squeeze
    ab = bs.all_bins
    NB = len(ab)
    trip = torch.from_numpy(ab.astype(np.int32).reshape(-1)).cuda()
    pairs = sorted({(int(x), int(y)) for t in ab for x, y in ((t[0], t[1]), (t[1], t[2]), (t[0], t[2]))})
    tab = np.full(Nq * Nq, -1, dtype=np.int32)
    for r, (x, y) in enumerate(pairs):
        tab[x * Nq + y] = r
    tab = torch.from_numpy(tab).cuda()
    rng = np.random.RandomState(n)
    FT = torch.from_numpy((rng.randn(bs.Nl, len(pairs), M) + 1j * rng.randn(bs.Nl, len(pairs), M))
                          .astype(np.complex64)).cuda()
    sup = b.support(edges[-1])
    common = dict(bin_table=bs.table, pack_index=pk['pack_index'], packed_len=M, trip=trip, pair_index=tab,
                  ft=FT, ft_lstride=len(pairs) * M, nk=Nq, support=sup)
    modes = pk['modes'].long()
    ops = [dict(terms=3, leg3_legendre=1), dict(terms=3, leg3_legendre=0), dict(terms=2, leg3_legendre=0),
           dict(terms=4, ell=2, m=1), dict(terms=4, ell=2, m=4)]
    for op in ops:
        kop = C.kop(C.KOP_BKQ, trip0=0, **op, **common)
        got = b.c2r(None, batch=NB, kop=kop).reshape((NB,) + tuple(shape))
        packed = torch.empty((NB, M), dtype=torch.complex64, device='cuda')
        C.check(C.lib().pb3d_kmap_packed(C.ctypes.byref(b.cgrid), None, pk['modes'].data_ptr(), M, packed.data_ptr(),
                                         NB, C.ctypes.byref(kop), b.stream()))
        full = torch.zeros((NB, b.H), dtype=torch.complex64, device='cuda')
        full[:, modes] = packed
        ref = b.c2r(full.reshape((NB,) + b.half_shape), batch=NB).reshape((NB,) + tuple(shape))
        err = (torch.linalg.vector_norm((got - ref).reshape(NB, -1), dim=1) /
               torch.linalg.vector_norm(ref.reshape(NB, -1), dim=1).clamp_min(1e-30)).max().item()
        assert err < 1e-5, (op, err)
    # packed epilogue: R2C of real maps, kept at the in-bin modes only
    x = torch.from_numpy(rng.randn(3, *shape).astype(np.float32)).cuda()
    out = torch.empty((3, M), dtype=torch.complex64, device='cuda')
    b.r2c(x, out=out, kop=C.kop(C.KOP_PACK, pack_index=pk['pack_index'], packed_len=M, support=sup))
    ref = b.r2c(x).reshape(3, -1)[:, modes]
    assert rel(out.cpu().numpy(), ref.cpu().numpy()) < 1e-6
