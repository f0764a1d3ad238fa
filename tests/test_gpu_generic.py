"""Grids outside the power-of-two fast path (any n0, n1, even n2; PolyBin3D/base.py:50-54 only asks
for an even last axis), which run the mixed-radix passes: the transforms and their fused ops against
numpy FP64, and the estimators against the oracle, on the [78, 146, 82] grid of the reference's
survey tutorials (Tutorial 3 cell 15, Tutorial 5 cell 10) and on small grids with odd axes and large
prime factors.  Tolerances as for the power-of-two transforms (relative L2 error below 2e-6 log2 N or
1e-5) and estimators (1e-5 of the per-bin scale)."""
import inspect

import numpy as np
import pytest
import torch

from oracle import polybin_oracle as O

import parity
from parity import own_scale

pytestmark = pytest.mark.gpu

SURVEY = (78, 146, 82)
SURVEY_BOX = [1000., 1872., 1051.]


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel())


def close(got, ref, tol=1e-5, scale=None, what=None):
    """parity.check labelled by the calling test and line."""
    f = inspect.currentframe().f_back
    label = "%s:%d" % (f.f_code.co_name, f.f_lineno)
    parity.check(label if what is None else "%s %s" % (label, what), got, ref, scale=scale, tol=tol)


def fisher_scale(F):
    d = np.abs(np.diag(F))
    return np.sqrt(np.outer(d, d))


def survey_mask(n, box, center, seed=1, r_lo=350., r_hi=1100., nholes=120, smooth=15.):
    """Shell r in [r_lo, r_hi] around the observer at the origin minus random holes, Gaussian-smoothed,
    on a non-cubic grid (the coordinates of PolyBin3D/base.py:77-80)."""
    x = [np.fft.fftshift(np.arange(-n[i] // 2, n[i] // 2)) * box[i] / n[i] + center[i] for i in range(3)]
    X = np.meshgrid(*x, indexing='ij')
    r = np.sqrt(X[0] ** 2 + X[1] ** 2 + X[2] ** 2)
    m = ((r >= r_lo) & (r <= r_hi)).astype(np.float64)
    rng = np.random.RandomState(seed)
    for _ in range(nholes):
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        c = u * rng.uniform(r_lo, r_hi)
        rad = rng.uniform(20., 60.)
        m[(X[0] - c[0]) ** 2 + (X[1] - c[1]) ** 2 + (X[2] - c[2]) ** 2 < rad ** 2] = 0.
    k = [np.fft.fftfreq(n[i], box[i] / n[i]) * 2 * np.pi for i in range(2)]
    k += [np.fft.rfftfreq(n[2], box[2] / n[2]) * 2 * np.pi]
    K = np.meshgrid(*k, indexing='ij')
    g = np.exp(-0.5 * smooth ** 2 * (K[0] ** 2 + K[1] ** 2 + K[2] ** 2))
    return np.fft.irfftn(np.fft.rfftn(m) * g, s=n, axes=(0, 1, 2))


# odd axes (7, 9, 15, 21), prime factors 13, 41 and 73, n2/2 odd and even, a power of two above the
# fast path's 512 (1024 on axis 1; n2/2 = 1024 rows, one per block)
SHAPES = [SURVEY, (15, 21, 26), (7, 9, 10), (30, 6, 12), (5, 3, 4), (73, 41, 14), (6, 1024, 6), (9, 10, 2048)]


def test_plan_limits(hip):
    import polybin3d_amd as pb
    from polybin3d_amd import _capi as C
    for n in ([1, 8, 8], [8, 8, 2], [8, 1025, 8], [8, 8, 4098]):
        g = C.Grid((C.ctypes.c_int32 * 3)(*n), 0, (C.ctypes.c_double * 3)(1., 1., 1.),
                   (C.ctypes.c_double * 3)(0., 0., 0.), (C.ctypes.c_double * 3)(0., 0., 1.))
        plan = C.ctypes.c_void_p()
        assert C.lib().pb3d_plan_create(C.ctypes.byref(plan), C.ctypes.byref(g)) == -2  # PB3D_ERR_UNSUPPORTED
    with pytest.raises(Exception):
        pb.PolyBin3D(1000., [8, 8, 9])


@pytest.mark.parametrize("n", SHAPES, ids=lambda n: "x".join(map(str, n)))
def test_r2c_c2r_match_numpy(hip, n):
    import polybin3d_amd as pb
    print("[generic] r2c/c2r", n, flush=True)
    b = pb.PolyBin3D(1000., list(n))
    x = np.random.RandomState(0).randn(*n)
    X = b.to_fourier(x).cpu().numpy()
    ref = np.fft.rfftn(x)
    tol = 2e-6 * np.log2(np.prod(n))
    e = rel(X, ref)
    print("  r2c rel L2 %.2e (tol %.1e)" % (e, tol))
    assert e < tol
    # C2R of a non-Hermitian half grid: numpy semantics (Im of kz=0 / Nyquist planes ignored)
    Y = (np.random.RandomState(1).randn(*ref.shape) + 1j * np.random.RandomState(2).randn(*ref.shape))
    y = b.to_real(Y).cpu().numpy()
    e = rel(y, np.fft.irfftn(Y, s=n, axes=(0, 1, 2)))
    print("  c2r rel L2 %.2e" % e)
    assert e < tol
    # batched, and the round trip
    xb = np.random.RandomState(3).randn(3, *n)
    Xb = b.r2c(b.as_real(xb)).cpu().numpy()
    assert rel(Xb, np.fft.rfftn(xb, axes=(1, 2, 3))) < tol
    assert rel(b.c2r(torch.from_numpy(Xb).cuda(), batch=3).cpu().numpy().reshape(xb.shape), xb) < tol


def test_fused_ops(hip):
    """Voxel prologues (MUL, DIFF, Y_lm), mode prologues broadcast over a batch of bins, the mask
    epilogue and the weighted / accumulating epilogue on an odd, non-cubic grid."""
    from polybin3d_amd import _capi as C
    import polybin3d_amd as pb
    n = (15, 21, 26)
    box, center = [900., 1100., 1000.], [500., 300., 200.]
    b = pb.PolyBin3D(box, list(n), boxcenter=center, sightline='local')
    o = O.PolyBin3D(box, list(n), boxcenter=center, sightline='local')
    rng = np.random.RandomState(4)
    a, m, c, d = (rng.randn(*n) for _ in range(4))
    A, M, Cc, D = (b.as_real(v) for v in (a, m, c, d))
    got = b.r2c(None, rop=C.rop(C.ROP_MUL, a=A, b=M)).cpu().numpy()
    assert rel(got, np.fft.rfftn(a * m)) < 1e-5
    got = b.r2c(None, rop=C.rop(C.ROP_DIFF, a=A, b=M, c=Cc, d=D)).cpu().numpy()
    assert rel(got, np.fft.rfftn(a * m - c * d)) < 1e-5
    Yr = O.real_harmonics(np.asarray(o.r_grids), 4, False)
    Yk = O.real_harmonics(np.asarray(o.k_grids), 4, False)
    for ell in (2, 4):
        got = b.apply_fourier_harmonics(A, ell, ell).cpu().numpy()
        assert rel(got, o.apply_fourier_harmonics(a, Yr[ell], Yk[ell])) < 2e-5
        F = np.fft.rfftn(a)
        got = b.apply_real_harmonics(torch.from_numpy(F).cuda(), ell, ell).cpu().numpy()
        assert rel(got, o.apply_real_harmonics(F, Yr[ell], Yk[ell])) < 2e-5
    kF = 2 * np.pi / 1100.
    edges = kF * np.array([1., 2.5, 4., 5.5, 7.])
    tab = b.bin_table(edges)
    bins = torch.arange(4, dtype=torch.int32, device='cuda')
    F = np.fft.rfftn(a)
    Fd = torch.from_numpy(F.astype(np.complex64)).cuda()
    for ell in (0, 2):
        out = b.c2r(Fd, batch=4, kop=C.kop(C.KOP_FILTER, bin_table=tab, bins=bins, ell=ell, w=b.invPk0_grid,
                                            x_batch_stride=0),
                    rop=C.rop(C.ROP_MUL, b=M)).cpu().numpy()
        for q in range(4):
            ref = o.to_real(o.map_utils.fourier_filter(F, ell, edges[q], edges[q + 1])) * m
            assert rel(out[q], ref) < 2e-5
    out = b.r2c(A, kop=C.kop(C.KOP_WEIGHT, w=b.invPk0_grid, scale=0.5))
    b.r2c(A, out=out, kop=C.kop(C.KOP_PLAIN, accumulate=1))
    assert rel(out.cpu().numpy(), 1.5 * np.fft.rfftn(a)) < 1e-5


def outside_support(n, sup):
    """Half-grid modes outside the support box of the radii `sup` (|fs(i)| <= r0, |fs(j)| <= r1,
    kz <= r2; an axis whose radius covers it is not limited)."""
    def axis(m, r):
        p = np.arange(m)
        return np.zeros(m, bool) if 2 * r >= m else ~((p <= r) | (p >= m - r))
    lz = np.arange(n[2] // 2 + 1)
    o2 = (lz > sup[2]) if sup[2] + 1 < n[2] // 2 + 1 else np.zeros(len(lz), bool)
    return axis(n[0], sup[0])[:, None, None] | axis(n[1], sup[1])[None, :, None] | o2[None, None, :]


@pytest.mark.parametrize("n,kmax", [(SURVEY, 0.08), ((73, 41, 14), 0.2)], ids=["survey", "primes"])
def test_pruned_transforms(hip, n, kmax):
    """Band-limited C2R (support given: inputs outside it are not read) and R2C with packed in-support
    output."""
    from polybin3d_amd import _capi as C
    import polybin3d_amd as pb
    box = SURVEY_BOX if n == SURVEY else [1000.] * 3
    b = pb.PolyBin3D(box, list(n))
    o = O.PolyBin3D(box, list(n))
    sup = b.support(kmax)
    rng = np.random.RandomState(5)
    X = rng.randn(*o.modk_grid.shape) + 1j * rng.randn(*o.modk_grid.shape)
    X[o.modk_grid >= kmax] = 0.
    # NaN outside the support box: the pass must not read it
    Xg = X.copy()
    Xg[outside_support(n, sup)] = np.nan
    Xd = torch.from_numpy(Xg.astype(np.complex64)).cuda()
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


@pytest.mark.parametrize("n,L", [(SURVEY, 2), ((15, 21, 26), 1), ((15, 21, 26), 3), ((15, 21, 26), 4),
                                 ((7, 9, 10), 2)],
                         ids=lambda c: "x".join(map(str, c)) if isinstance(c, tuple) else "l%d" % c)
def test_fused_pointing_chain(hip, n, L):
    """pb3d_fft_pointing = R2C(mask * scale * sum_g coef_g Y_g(x) C2R(kop_g(in))) composed from numpy."""
    from polybin3d_amd import _capi as C
    import polybin3d_amd as pb
    center = [400., 250., -300.]
    box = SURVEY_BOX if n == SURVEY else [1000.] * 3
    b = pb.PolyBin3D(box, list(n), boxcenter=center, sightline='local')
    o = O.PolyBin3D(box, list(n), boxcenter=center, sightline='local')
    rng = np.random.RandomState(6)
    A = np.fft.rfftn(rng.randn(*n))
    mask = rng.rand(*n)
    kF = 2 * np.pi / 1000.
    edges = kF * np.linspace(1., 1. + 0.8 * min(n) / 2, 5)
    tab = b.bin_table(edges)
    bins = torch.arange(4, dtype=torch.int32, device='cuda')
    sup = b.support(edges[-1])
    Ad = torch.from_numpy(A.astype(np.complex64)).cuda()
    Md = b.as_real(mask)
    common = dict(bin_table=tab, bins=bins, x_batch_stride=0, support=sup)
    subs = [(C.kop(C.KOP_FILTER, ell=2, **common), Ad, 0, 0, 0.5)]
    subs += [(C.kop(C.KOP_YLM, ell=L, m=m, **common), Ad, L, m, 0.25) for m in range(2 * L + 1)]
    out = torch.empty((4,) + b.half_shape, dtype=torch.complex64, device='cuda')
    b.pointing(subs, 4, out, mask=Md, scale=1.5, kop_out=None)
    Yr = O.real_harmonics(np.asarray(o.r_grids), L, True)[L]
    Yk = O.real_harmonics(np.asarray(o.k_grids), L, True)[L]
    got = out.cpu().numpy()
    for q in range(4):
        r = 0.5 * o.to_real(o.map_utils.fourier_filter(A, 2, edges[q], edges[q + 1]))
        for m in range(2 * L + 1):
            r += 0.25 * Yr[m] * o.to_real(o.map_utils.prod_fourier_filter(A, Yk[m], edges[q], edges[q + 1]))
        ref = o.to_fourier(1.5 * mask * r)
        e = rel(got[q], ref)
        print("[generic] pointing", n, "l", L, "bin", q, "rel L2 %.2e" % e)
        assert e < 2e-5


PK = [np.linspace(0, 2, 500), 1e4 / (1 + (np.linspace(0, 2, 500) / 0.05) ** 2)]


def test_pk_masked_mc_fisher_survey_grid(hip):
    """Masked, local-sightline MC Fisher seeds of the power spectrum on the [78, 146, 82] grid (the
    reference's BOSS tutorial shape) against the oracle; the batched-seed pass; Pk_unwindowed."""
    import polybin3d_amd as pb
    n, box, center = SURVEY, SURVEY_BOX, [-200., 150., 400.]
    mask = survey_mask(n, box, center)
    kw = dict(boxcenter=center, sightline='local', Pk=PK)
    b, o = pb.PolyBin3D(box, list(n), **kw), O.PolyBin3D(box, list(n), **kw)
    edges = np.linspace(0.01, 0.1, 10)
    ps = pb.PSpec(b, edges, lmax=4, mask=mask, mask_shot=mask)
    po = O.PSpec(o, edges, lmax=4, mask=mask, mask_shot=mask)
    for seed in (0, 1):
        print("[generic] Pk MC Fisher seed", seed, flush=True)
        F1 = ps.compute_fisher_contribution(seed)
        F2 = po.compute_fisher_contribution(seed)
        close(F1, F2, scale=fisher_scale(F2))
    s1, s2 = ps.compute_shot_contribution(5), po.compute_shot_contribution(5)
    close(s1, s2, scale=np.abs(s2[:ps.Nk]).max())
    fish = po.compute_fisher(2)
    close(ps.compute_fisher(2, batch=2), fish, scale=fisher_scale(fish))
    kk = np.linspace(0, 2, 4000)
    P = np.zeros_like(kk)
    P[1:] = 2e4 * (kk[1:] / 0.1) ** -1.5
    d = O.PolyBin3D(box, list(n), Pk=[kk, P]).generate_data(seed=3, output_type='real') * mask
    got, ref = ps.Pk_unwindowed(d, fish=fish), po.Pk_unwindowed(d, fish=fish)
    parity.check_solve("Pk_unwindowed survey grid", got, ref, sorted(ref), fish, fish, po.Pk_numerator(d),
                       ps.Pk_numerator(d))


@pytest.mark.parametrize("case", ["gic_global", "ric_local"])
def test_pk_mc_fisher_integral_constraints(hip, case):
    """Integral-constrained pointing (per-bin sums inside the mixed-radix pointing rows) on an odd,
    non-cubic grid."""
    import polybin3d_amd as pb
    n, box = (27, 30, 22), [1000., 1100., 900.]
    if case == "gic_global":
        kw, ic, center = {}, dict(add_GIC=True), [0., 0., 0.]
    else:
        center = [10., 10., 10.]
        kw, ic = dict(boxcenter=center, sightline='local'), dict(add_RIC=True, radial_bins_RIC=[100., 250., 400., 600.])
    mask = survey_mask(n, box, center, r_lo=150., r_hi=480., nholes=40, smooth=30.)
    mask_IC = mask * (1.2 + 0.1 * np.cos(np.arange(n[2]) * 0.3))[None, None, :]
    b, o = pb.PolyBin3D(box, list(n), Pk=PK, **kw), O.PolyBin3D(box, list(n), Pk=PK, **kw)
    edges = np.linspace(0.01, 0.07, 7)
    ps = pb.PSpec(b, edges, lmax=2, mask=mask, mask_IC=mask_IC, **ic)
    po = O.PSpec(o, edges, lmax=2, mask=mask, mask_IC=mask_IC, **ic)
    for seed in (0, 1):
        close(ps.compute_fisher_contribution(seed), po.compute_fisher_contribution(seed),
              scale=fisher_scale(po.compute_fisher_contribution(seed)))


def ng_field(n, box, seed=0):
    kk = np.linspace(0, 3, 6000)
    P = np.zeros_like(kk)
    P[1:] = 2e4 * (kk[1:] / 0.1) ** -1.5
    d = O.PolyBin3D(box, list(n), Pk=[kk, P]).generate_data(seed=seed, output_type='real')
    return d + 0.5 * (d ** 2 - np.mean(d ** 2))


def test_bk_ideal_and_mc_fisher(hip):
    """Bk_ideal (triplet sums over the mixed-radix g-maps) and masked MC Fisher seeds of the bispectrum
    (Q maps assembled in the first C2R pass, fused pointing chains) on odd, non-cubic grids."""
    import polybin3d_amd as pb
    n, box = (27, 30, 22), [1000., 1100., 900.]
    Pk = [np.linspace(0, 2, 800), 5e3 / (1 + (np.linspace(0, 2, 800) / 0.05) ** 2)]
    center = [0., 1170., 826.]
    kw = dict(Pk=Pk, boxcenter=center, sightline='local')
    b, o = pb.PolyBin3D(box, list(n), **kw), O.PolyBin3D(box, list(n), **kw)
    edges = np.linspace(0.01, 0.07, 4)
    bs, bo = pb.BSpec(b, edges, lmax=2), O.BSpec(o, edges, lmax=2)
    assert bs.N_bins == bo.N_bins and np.array_equal(bs.all_bins, bo.all_bins)
    d = ng_field(n, box)
    got, ref = bs.Bk_ideal(d), bo.Bk_ideal(d)
    for k in ref:
        close(got[k], ref[k], scale=own_scale(ref[k]), what=k)
    n2, box2 = (15, 18, 14), [1000., 1200., 950.]
    mask = 0.2 + survey_mask(n2, box2, [10., 10., 10.], r_lo=150., r_hi=480., nholes=10, smooth=40.)
    kw2 = dict(Pk=Pk, boxcenter=[10., 10., 10.], sightline='local')
    b2, o2 = pb.PolyBin3D(box2, list(n2), **kw2), O.PolyBin3D(box2, list(n2), **kw2)
    e2 = np.linspace(0.008, 0.045, 4)
    bs2, bo2 = pb.BSpec(b2, e2, lmax=2, mask=mask), O.BSpec(o2, e2, lmax=2, mask=mask)
    for seed in (0, 3):
        print("[generic] Bk MC Fisher seed", seed, flush=True)
        close(bs2.compute_fisher_contribution(seed), bo2.compute_fisher_contribution(seed),
              scale=fisher_scale(bo2.compute_fisher_contribution(seed)))
