"""GPU parity of the bispectrum path against the FP64 oracle on identical seeded inputs.
Tolerance: 1e-5 of the per-bin scale (each multipole's own |B|, floored at 1e-3 of its largest bin;
sqrt(F_ii F_jj) for Fisher entries, which cross zero off the diagonal); every check prints its error."""
import inspect

import numpy as np
import pytest

from oracle import polybin_oracle as O

import parity
from parity import own_scale

pytestmark = pytest.mark.gpu

L = 1000.


def ng_field(n, seed=0):
    """delta_G + 0.5 (delta_G^2 - <delta_G^2>) with the power-law GRF (the Bk_ideal workload input)."""
    kk = np.linspace(0, 3, 6000)
    P = np.zeros_like(kk)
    P[1:] = 2e4 * (kk[1:] / 0.1) ** -1.5
    o = O.PolyBin3D(L, list(n), Pk=[kk, P])
    d = o.generate_data(seed=seed, output_type='real')
    return d + 0.5 * (d ** 2 - np.mean(d ** 2))


def close(got, ref, tol=1e-5, scale=None, what=None):
    """parity.check labelled by the calling test and line."""
    f = inspect.currentframe().f_back
    label = "%s:%d" % (f.f_code.co_name, f.f_lineno)
    parity.check(label if what is None else "%s %s" % (label, what), got, ref, scale=scale, tol=tol)


def fisher_scale(F):
    d = np.abs(np.diag(F))
    return np.sqrt(np.outer(d, d))


CASES = {
    "32_global_l0": dict(n=32, kw={}, lmax=0, edges=np.linspace(0.01, 0.09, 7), sq=None, Pk_fid=None),
    "32_global_l2_squeeze": dict(n=32, kw={}, lmax=2, edges=np.linspace(0.01, 0.05, 3),
                                 sq=np.linspace(0.01, 0.09, 5), Pk_fid='default'),
    "32_local_l2": dict(n=32, kw=dict(boxcenter=[0., 1170., 826.], sightline='local'), lmax=2,
                        edges=np.linspace(0.01, 0.07, 4), sq=None, Pk_fid=None),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_bk_ideal(hip, case):
    import polybin3d_amd as pb
    c = CASES[case]
    n = c['n']
    Pk = [np.linspace(0, 2, 800), 5e3 / (1 + (np.linspace(0, 2, 800) / 0.05) ** 2)]
    b, o = pb.PolyBin3D(L, [n] * 3, Pk=Pk, **c['kw']), O.PolyBin3D(L, [n] * 3, Pk=Pk, **c['kw'])
    kw = dict(lmax=c['lmax'], k_bins_squeeze=c['sq'], Pk_fid=c['Pk_fid'])
    bs, bo = pb.BSpec(b, c['edges'], **kw), O.BSpec(o, c['edges'], **kw)
    assert bs.N_bins == bo.N_bins and np.array_equal(bs.all_bins, bo.all_bins)
    d = ng_field((n, n, n))
    for disc in (True, False):
        got, ref = bs.Bk_ideal(d, discreteness_correction=disc), bo.Bk_ideal(d, discreteness_correction=disc)
        close(bs.sym_factor, bo.sym_factor, tol=2e-6)
        close(bs.fish_ideal, bo.fish_ideal, scale=fisher_scale(bo.fish_ideal))
        sc = np.abs(bo.Bk_numerator_ideal(d))
        close(bs.Bk_numerator_ideal(d), bo.Bk_numerator_ideal(d), scale=np.maximum(sc, sc.max() * 1e-3))
        for k in ref:
            close(got[k], ref[k], scale=own_scale(ref[k]), what=k)


def shell_mask(n, L, center):
    x = [np.fft.fftshift(np.arange(-n // 2, n // 2)) * L / n + center[i] for i in range(3)]
    X = np.meshgrid(*x, indexing='ij')
    r = np.sqrt(X[0] ** 2 + X[1] ** 2 + X[2] ** 2)
    return 0.2 + ((r >= 150.) & (r <= 480.)).astype(np.float64)


@pytest.mark.parametrize("case", ["global_const_l2", "global_mask_l0", "local_mask_l2"])
def test_bk_mc_fisher_per_seed(hip, case):
    import polybin3d_amd as pb
    n = 16
    edges = np.linspace(0.007, 0.05, 4)
    if case == "global_const_l2":
        kw, mask, lmax = {}, 0.7 * np.ones((n, n, n)), 2
    elif case == "global_mask_l0":
        kw, mask, lmax = {}, shell_mask(n, L, [0., 0., 0.]), 0
    else:
        kw, mask, lmax = dict(boxcenter=[10., 10., 10.], sightline='local'), shell_mask(n, L, [10., 10., 10.]), 2
    Pk = [np.linspace(0, 2, 500), 3e3 / (1 + (np.linspace(0, 2, 500) / 0.05) ** 2)]
    b, o = pb.PolyBin3D(L, [n] * 3, Pk=Pk, **kw), O.PolyBin3D(L, [n] * 3, Pk=Pk, **kw)
    bs, bo = pb.BSpec(b, edges, lmax=lmax, mask=mask), O.BSpec(o, edges, lmax=lmax, mask=mask)
    for seed in (0, 3):
        F1 = bs.compute_fisher_contribution(seed)
        F2 = bo.compute_fisher_contribution(seed)
        close(F1, F2, scale=fisher_scale(F2))


@pytest.mark.parametrize("case", ["global_l2", "local_l4"])
def test_bk_unwindowed_with_linear_term(hip, case):
    """Bk_unwindowed with the linear term (generated sims), its MC Fisher and numerator against the
    oracle; local l = 4 on 16^3 runs the l = 4 Q-map chains one sub-map at a time (rows too short to
    combine).  The numerator is the cubic term minus the linear term, which cancel by design, and
    each term is a grid sum whose voxels cancel (an entry of the local l = 4 case is 4e-5 of the
    largest).  FP32 rounding acts on the same sums over |g| maps, so each entry is held to 1e-5 of
    max(|entry|, 0.1 x its |g| sum).  Bk_unwindowed = inv(F) num goes through the solve bound (the
    same F on both sides, so only the measured numerator error propagates)."""
    import polybin3d_amd as pb
    n = 16
    edges = np.linspace(0.007, 0.05, 4)
    local = case.startswith("local")
    c = [10., 10., 10.] if local else [0., 0., 0.]
    kw, lmax = (dict(boxcenter=c, sightline='local'), 4) if local else ({}, 2)
    mask = shell_mask(n, L, c)
    Pk = [np.linspace(0, 2, 500), 3e3 / (1 + (np.linspace(0, 2, 500) / 0.05) ** 2)]
    b, o = pb.PolyBin3D(L, [n] * 3, Pk=Pk, **kw), O.PolyBin3D(L, [n] * 3, Pk=Pk, **kw)
    bs, bo = pb.BSpec(b, edges, lmax=lmax, mask=mask), O.BSpec(o, edges, lmax=lmax, mask=mask)
    bs.generate_sims(3)
    bo.generate_sims(3)
    fish = bo.compute_fisher(2)
    close(bs.compute_fisher(2), fish, scale=fisher_scale(fish))
    d = ng_field((n, n, n), 4) * mask
    n1 = np.asarray(bs.Bk_numerator(d))
    n2, mag = bo._compute_bk_numerator(d, include_linear_term=True, filtering='Sinv', return_scale=True)
    err = np.abs(n1 - n2)
    i = int(np.argmax(err / mag))
    print("numerator: worst entry %d at %.1e of the maximum, %.1e of its |g| sum; error %.2e of that sum"
          % (i, abs(n2[i]) / np.abs(n2).max(), abs(n2[i]) / mag[i], err[i] / mag[i]))
    close(n1, n2, scale=np.maximum(np.abs(n2), 0.1 * mag))
    got, ref = bs.Bk_unwindowed(d, fish=fish), bo.Bk_unwindowed(d, fish=fish)
    parity.check_solve("Bk_unwindowed %s" % case, got, ref, ['b%d' % l for l in range(0, lmax + 1, 2)], fish, fish,
                       n2, n1)


@pytest.mark.parametrize("case", ["global_mask_l2", "local_mask_l2"])
def test_bk_shot_contribution(hip, case):
    """BSpec.compute_shot_contribution (bspec.py:695-861): cubic term alone and quadratic - 2 cubic
    with singly and doubly weighted data, per seed against the oracle; scale = the same sums over
    |g| maps (the magnitude FP32 rounding acts on: single bins cancel to 1e-4 of it)."""
    import polybin3d_amd as pb
    n = 16
    edges = np.linspace(0.007, 0.05, 4)
    kw = {} if case.startswith("global") else dict(boxcenter=[10., 10., 10.], sightline='local')
    mask = shell_mask(n, L, kw.get('boxcenter', [0., 0., 0.]))
    mask_shot = mask ** 3
    Pk = [np.linspace(0, 2, 500), 3e3 / (1 + (np.linspace(0, 2, 500) / 0.05) ** 2)]
    b, o = pb.PolyBin3D(L, [n] * 3, Pk=Pk, **kw), O.PolyBin3D(L, [n] * 3, Pk=Pk, **kw)
    bs = pb.BSpec(b, edges, lmax=2, mask=mask, mask_shot=mask_shot)
    bo = O.BSpec(o, edges, lmax=2, mask=mask, mask_shot=mask_shot)
    d = ng_field((n, n, n), 5) * mask
    d2 = d * mask
    for seed in (0, 2):
        c1 = bs.compute_shot_contribution(seed, cubic_only=True)
        c2, sc = bo.compute_shot_contribution(seed, cubic_only=True, return_scale=True)
        close(c1, c2, scale=sc)
        s1 = bs.compute_shot_contribution(seed, data=d, data2=d2)
        s2, sc = bo.compute_shot_contribution(seed, data=d, data2=d2, return_scale=True)
        close(s1, s2, scale=sc)


@pytest.mark.parametrize("case", ["gic_global_l2", "ric_local_l2"])
def test_bk_integral_constraints(hip, case):
    """MC Fisher per seed and the linear-term numerator with the integral-constrained pointing
    matrix (bspec.py:412-426; utils.pyx:405-513) against the oracle."""
    import polybin3d_amd as pb
    n = 16
    edges = np.linspace(0.007, 0.05, 4)
    if case.startswith("gic"):
        kw, ic = {}, dict(add_GIC=True)
    else:
        kw, ic = dict(boxcenter=[10., 10., 10.], sightline='local'), dict(add_RIC=True,
                                                                          radial_bins_RIC=[100., 250., 400., 600.])
    mask = shell_mask(n, L, kw.get('boxcenter', [0., 0., 0.]))
    mask_IC = mask * (1.2 + 0.1 * np.cos(np.arange(n) * 0.3))[None, None, :]
    Pk = [np.linspace(0, 2, 500), 3e3 / (1 + (np.linspace(0, 2, 500) / 0.05) ** 2)]
    b, o = pb.PolyBin3D(L, [n] * 3, Pk=Pk, **kw), O.PolyBin3D(L, [n] * 3, Pk=Pk, **kw)
    bs = pb.BSpec(b, edges, lmax=2, mask=mask, mask_IC=mask_IC, **ic)
    bo = O.BSpec(o, edges, lmax=2, mask=mask, mask_IC=mask_IC, **ic)
    for seed in (0, 3):
        F1 = bs.compute_fisher_contribution(seed)
        F2 = bo.compute_fisher_contribution(seed)
        close(F1, F2, scale=fisher_scale(F2))
    bs.generate_sims(2)
    bo.generate_sims(2)
    d = ng_field((n, n, n), 4) * mask
    n1, n2 = bs.Bk_numerator(d), bo.Bk_numerator(d)
    # scale: the same triple sums over |g| maps (local l = 2 bins cancel to ~1e-2 of it)
    g = np.abs(bo._g_maps(o.to_fourier(np.asarray(d, dtype=np.float64))))
    sc = O.assemble_b3(g, bo.all_bins)
    for ii in range(bo.N_it):
        sc = sc + O.assemble_b1(g, np.abs(bo.g_bl_maps[ii]), bo.all_bins) / bo.N_it
    close(n1, n2, scale=sc * o.volume / o.gridsize.prod() / bo.sym_factor)


@pytest.mark.parametrize("case", ["global_mask_l2", "local_const_l2"])
def test_bk_mc_fisher_user_sinv(hip, case):
    """MC Fisher per seed with a user S^-1 (a mode-space weight given as a callable, bspec.py:499-505):
    the generic path (user callable on every Q map) against the oracle with the same S^-1."""
    import polybin3d_amd as pb
    import torch
    n = 16
    edges = np.linspace(0.007, 0.05, 4)
    if case.startswith("global"):
        kw, mask = {}, shell_mask(n, L, [0., 0., 0.])
    else:
        kw, mask = dict(boxcenter=[10., 10., 10.], sightline='local'), 0.6 * np.ones((n, n, n))
    Pk = [np.linspace(0, 2, 500), 3e3 / (1 + (np.linspace(0, 2, 500) / 0.05) ** 2)]
    b, o = pb.PolyBin3D(L, [n] * 3, Pk=Pk, **kw), O.PolyBin3D(L, [n] * 3, Pk=Pk, **kw)
    wk = 1. / (1. + (o.modk_grid / 0.02) ** 2)
    wk_d = torch.from_numpy(wk.astype(np.float32)).cuda()

    def sinv_o(m, input_type='real', output_type='fourier'):
        f = o.to_fourier(m) if input_type == 'real' else m
        f = f * wk
        return f if output_type == 'fourier' else o.to_real(f)

    def sinv_g(m, input_type='real', output_type='fourier'):
        f = b.to_fourier(m) if input_type == 'real' else m
        f = f * wk_d
        return f if output_type == 'fourier' else b.to_real(f)
    bs = pb.BSpec(b, edges, lmax=2, mask=mask, applySinv=sinv_g)
    bo = O.BSpec(o, edges, lmax=2, mask=mask, applySinv=sinv_o)
    for seed in (0, 3):
        F1 = bs.compute_fisher_contribution(seed)
        F2 = bo.compute_fisher_contribution(seed)
        close(F1, F2, scale=fisher_scale(F2))


@pytest.mark.parametrize("case", ["global_const_l2", "local_mask_l2", "local_gic_l2"])
def test_bk_covariance_contribution(hip, case):
    """compute_covariance_contribution (bspec.py:447-519) per seed against the oracle: the MC Fisher with
    S^-1 [P xi P^T + N] S^-dag on both sides; global P(k, mu) from P_0, P_2, P_4; local P_0 and P_2
    through the harmonic transforms, the shot-noise mask, and the integral-constrained pointing."""
    import polybin3d_amd as pb
    n = 16
    edges = np.linspace(0.007, 0.05, 4)
    kc = np.linspace(0, 2, 300)
    Pc = [kc, 3e3 / (1 + (kc / 0.05) ** 2), 800. / (1 + (kc / 0.05) ** 2), 200. / (1 + (kc / 0.04) ** 2)]
    ic, opt = {}, {}
    if case == "global_const_l2":
        kw, mask = {}, 0.7 * np.ones((n, n, n))
    else:
        kw, mask, Pc = dict(boxcenter=[10., 10., 10.], sightline='local'), shell_mask(n, L, [10., 10., 10.]), Pc[:3]
        opt = dict(mask_shot=0.5 * mask)
        if case == "local_gic_l2":
            ic = dict(add_GIC=True, mask_IC=1.1 * mask)
    Pk = [np.linspace(0, 2, 500), 3e3 / (1 + (np.linspace(0, 2, 500) / 0.05) ** 2)]
    b, o = pb.PolyBin3D(L, [n] * 3, Pk=Pk, **kw), O.PolyBin3D(L, [n] * 3, Pk=Pk, **kw)
    bs = pb.BSpec(b, edges, lmax=2, mask=mask, **opt, **ic)
    bo = O.BSpec(o, edges, lmax=2, mask=mask, **opt, **ic)
    for seed in (0, 3):
        ms2 = 0.3 * mask if seed == 3 and case != "global_const_l2" else None
        C1 = bs.compute_covariance_contribution(seed, Pc, mask_shot_2pt=ms2)
        C2 = bo.compute_covariance_contribution(seed, Pc, mask_shot_2pt=ms2)
        close(C1, C2, scale=fisher_scale(C2))


@pytest.mark.parametrize("window", ["tsc", "interlaced-cic"])
def test_bk_pixel_windows(hip, window):
    """Pixel windows through the bispectrum path: Bk_ideal and Bk_numerator (the data deconvolved by
    the window) and one masked MC Fisher realization against the oracle."""
    import polybin3d_amd as pb
    n = 16
    edges = np.linspace(0.007, 0.05, 4)
    kw = dict(boxcenter=[10., 10., 10.], sightline='local')
    Pk = [np.linspace(0, 2, 500), 3e3 / (1 + (np.linspace(0, 2, 500) / 0.05) ** 2)]
    b = pb.PolyBin3D(L, [n] * 3, Pk=Pk, pixel_window=window, **kw)
    o = O.PolyBin3D(L, [n] * 3, Pk=Pk, pixel_window=window, **kw)
    d = ng_field((n, n, n), seed=2)
    bs, bo = pb.BSpec(b, edges, lmax=2), O.BSpec(o, edges, lmax=2)
    got, ref = bs.Bk_ideal(d), bo.Bk_ideal(d)
    for k in ref:
        close(got[k], ref[k], scale=own_scale(ref[k]), what=k)
    n1, n2 = bs.Bk_numerator(d, include_linear_term=False), bo.Bk_numerator(d, include_linear_term=False)
    s = np.abs(n2)
    close(n1, n2, scale=np.maximum(s, 1e-3 * s.max()))
    mask = shell_mask(n, L, [10., 10., 10.])
    bsm, bom = pb.BSpec(b, edges, lmax=2, mask=mask), O.BSpec(o, edges, lmax=2, mask=mask)
    F1, F2 = bsm.compute_fisher_contribution(1), bom.compute_fisher_contribution(1)
    close(F1, F2, scale=fisher_scale(F2))


def test_bk_mc_fisher_squeezed_partial(hip):
    """MC Fisher per seed with squeezed bins (k_bins_squeeze) and partial triangles
    (bspec.py:31-170, 473-676), masked, global sightline, against the oracle."""
    import polybin3d_amd as pb
    n = 16
    edges = np.linspace(0.007, 0.035, 3)
    sq = np.append(edges, [0.042, 0.049])
    Pk = [np.linspace(0, 2, 500), 3e3 / (1 + (np.linspace(0, 2, 500) / 0.05) ** 2)]
    b, o = pb.PolyBin3D(L, [n] * 3, Pk=Pk), O.PolyBin3D(L, [n] * 3, Pk=Pk)
    mask = shell_mask(n, L, [0., 0., 0.])
    for partial in (False, True):
        bs = pb.BSpec(b, edges, lmax=2, mask=mask, k_bins_squeeze=sq, include_partial_triangles=partial)
        bo = O.BSpec(o, edges, lmax=2, mask=mask, k_bins_squeeze=sq, include_partial_triangles=partial)
        assert bs.N_bins == bo.N_bins and np.array_equal(bs.all_bins, bo.all_bins)
        F1, F2 = bs.compute_fisher_contribution(2), bo.compute_fisher_contribution(2)
        close(F1, F2, scale=fisher_scale(F2))


def test_non_cubic_grid_and_box(hip):
    """Anisotropic grid and box (kF and kNy differ per axis; base.py:63-94): Pk_ideal, a masked Pk MC
    Fisher realization (local sightline) and Bk_ideal against the oracle."""
    import polybin3d_amd as pb
    shape, box = [32, 64, 16], [1000., 2000., 500.]
    kw = dict(boxcenter=[10., 20., 30.], sightline='local')
    Pk = [np.linspace(0, 2, 500), 3e3 / (1 + (np.linspace(0, 2, 500) / 0.05) ** 2)]
    b, o = pb.PolyBin3D(box, shape, Pk=Pk, **kw), O.PolyBin3D(box, shape, Pk=Pk, **kw)
    rng = np.random.RandomState(3)
    d = rng.randn(*shape)
    d = d + 0.3 * (d ** 2 - 1)
    edges = np.linspace(0.013, 0.09, 6)
    pg, po = pb.PSpec(b, edges, lmax=2), O.PSpec(o, edges, lmax=2)
    got, ref = pg.Pk_ideal(d), po.Pk_ideal(d)
    parity.check_solve("anisotropic Pk_ideal", got, ref, ['p0', 'p2'], po.fish_ideal, pg.fish_ideal,
                       po.Pk_numerator_ideal(d), pg.Pk_numerator_ideal(d))
    mask = 0.3 + rng.rand(*shape)
    F1 = pb.PSpec(b, edges, lmax=2, mask=mask).compute_fisher_contribution(0)
    F2 = O.PSpec(o, edges, lmax=2, mask=mask).compute_fisher_contribution(0)
    close(F1, F2, scale=fisher_scale(F2))
    e3 = np.linspace(0.013, 0.09, 4)
    gb, rb = pb.BSpec(b, e3, lmax=2).Bk_ideal(d), O.BSpec(o, e3, lmax=2).Bk_ideal(d)
    for k in rb:
        close(gb[k], rb[k], scale=own_scale(rb[k]), what=k)


def test_bk_linear_term_sims(hip):
    """The linear term of the Bk numerator (bspec.py:281-410, 963-978), local sightline, mask, pixel
    window: sims generated with an input spectrum (streamed, preload=False) and sims handed in by a
    loader (load_sims, preloaded and streamed), against the oracle with the same sims."""
    import polybin3d_amd as pb
    n = 16
    edges = np.linspace(0.007, 0.05, 4)
    kw = dict(boxcenter=[10., 10., 10.], sightline='local', pixel_window='cic')
    Pk = [np.linspace(0, 2, 500), 3e3 / (1 + (np.linspace(0, 2, 500) / 0.05) ** 2)]
    PkS = [np.linspace(0, 2, 500), 2e3 / (1 + (np.linspace(0, 2, 500) / 0.04) ** 2)]
    b, o = pb.PolyBin3D(L, [n] * 3, Pk=Pk, **kw), O.PolyBin3D(L, [n] * 3, Pk=Pk, **kw)
    mask = shell_mask(n, L, [10., 10., 10.])
    d = ng_field((n, n, n), 5) * mask
    bs, bo = pb.BSpec(b, edges, lmax=2, mask=mask), O.BSpec(o, edges, lmax=2, mask=mask)
    bs.generate_sims(3, Pk_input=PkS, preload=False)
    bo.generate_sims(3, Pk_input=PkS)
    n1, n2 = bs.Bk_numerator(d), bo.Bk_numerator(d)
    close(n1, n2, scale=np.maximum(np.abs(n2), 1e-3 * np.abs(n2).max()))
    sims = [ng_field((n, n, n), 10 + i) * mask for i in range(3)]
    bo.g_bl_maps = np.asarray([bo._process_sim(s, input_type='real') for s in sims])
    bo.N_it = 3
    ref = bo.Bk_numerator(d)
    for preload in (True, False):
        bl = pb.BSpec(b, edges, lmax=2, mask=mask)
        bl.load_sims(lambda ii: sims[ii], 3, input_type='real', preload=preload)
        close(bl.Bk_numerator(d), ref, scale=np.maximum(np.abs(ref), 1e-3 * np.abs(ref).max()))


@pytest.mark.parametrize("shape", [(32, 32, 32), (24, 30, 36)])
def test_masked_rows_skipped(hip, shape):
    """A mask that is exactly zero on whole groups of (x, y) rows (here every row with x <= 0): the
    row pass skips those rows, the C2R y pass does not store them and the R2C y pass reads them as
    zeros (power-of-two and mixed-radix plans).  The Bk and the GIC-constrained Pk MC Fisher
    contributions equal, to the bit, the ones computed with every row transformed, and match the
    oracle."""
    import polybin3d_amd as pb
    n, c = shape, [10., 10., 10.]
    x = [np.fft.fftshift(np.arange(-n[i] // 2, n[i] // 2)) * L / n[i] + c[i] for i in range(3)]
    X = np.meshgrid(*x, indexing='ij')
    r = np.sqrt(X[0] ** 2 + X[1] ** 2 + X[2] ** 2)
    mask = ((r >= 150.) & (r <= 480.) & (X[0] > 0.)).astype(np.float64) * (0.5 + 0.5 * X[2] / 500.)
    kw = dict(boxcenter=c, sightline='local')
    Pk = [np.linspace(0, 2, 500), 3e3 / (1 + (np.linspace(0, 2, 500) / 0.05) ** 2)]
    b, o = pb.PolyBin3D(L, list(n), Pk=Pk, **kw), O.PolyBin3D(L, list(n), Pk=Pk, **kw)
    edges = np.linspace(0.007, 0.05, 4)
    bs, bo = pb.BSpec(b, edges, lmax=2, mask=mask), O.BSpec(o, edges, lmax=2, mask=mask)
    rows = b.mask_rows(bs._P.mask)
    assert rows is not None and 0 < int(rows.sum()) < rows.numel()
    F1 = bs.compute_fisher_contribution(3)
    b.mask_rows = lambda m, w=None: None  # every row transformed
    try:
        F2 = bs.compute_fisher_contribution(3)
        ps = pb.PSpec(b, np.linspace(0.01, 0.07, 7), lmax=2, mask=mask, add_GIC=True, mask_IC=mask)
        P2 = ps.compute_fisher_contribution(1)
    finally:
        del b.mask_rows
    assert np.array_equal(F1, F2)
    close(F1, bo.compute_fisher_contribution(3), scale=fisher_scale(F1))
    P1 = ps.compute_fisher_contribution(1)
    assert np.array_equal(P1, P2)
    po = O.PSpec(o, np.linspace(0.01, 0.07, 7), lmax=2, mask=mask, add_GIC=True, mask_IC=mask)
    close(P1, po.compute_fisher_contribution(1), scale=fisher_scale(P1))


@pytest.mark.parametrize("case", ["global_mask_l4", "local_mask_l4"])
def test_bk_lmax4(hip, case):
    """lmax = 4, the largest the reference takes (bspec.py:84): Bk_ideal (b0, b2, b4) and a masked MC
    Fisher realization against the oracle.  Global: the l = 4 Legendre weight on the third leg, 32^3.
    Local: the l = 4 Q-map chains carry ten sub-maps, combined into five row transforms on the
    16 x 16 x 128 grid (rows of 128 points) in a 500 x 500 x 1000 box."""
    import polybin3d_amd as pb
    local = case.startswith("local")
    shape, box = ([16, 16, 128], [500., 500., 1000.]) if local else ([32, 32, 32], [L] * 3)
    c = [10., 20., 30.] if local else [0., 0., 0.]
    kw = dict(boxcenter=c, sightline='local') if local else {}
    x = [np.fft.fftshift(np.arange(-shape[i] // 2, shape[i] // 2)) * box[i] / shape[i] + c[i] for i in range(3)]
    X = np.meshgrid(*x, indexing='ij')
    r = np.sqrt(X[0] ** 2 + X[1] ** 2 + X[2] ** 2)
    mask = 0.2 + ((r >= 150.) & (r <= 480.)).astype(np.float64)
    Pk = [np.linspace(0, 2, 500), 3e3 / (1 + (np.linspace(0, 2, 500) / 0.05) ** 2)]
    b, o = pb.PolyBin3D(box, shape, Pk=Pk, **kw), O.PolyBin3D(box, shape, Pk=Pk, **kw)
    edges = np.linspace(0.013, 0.07, 4)
    rng = np.random.RandomState(11)
    d = rng.randn(*shape)
    d = d + 0.3 * (d ** 2 - 1)
    bi, bio = pb.BSpec(b, edges, lmax=4), O.BSpec(o, edges, lmax=4)
    got, ref = bi.Bk_ideal(d), bio.Bk_ideal(d)
    assert sorted(got) == sorted(ref) == ['b0', 'b2', 'b4']
    # Bk_ideal = inv(F_ideal) num, with F_ideal coupling the multipoles: the numerator per
    # multipole, F_ideal on its diagonal, and the solve held to the error bound that the measured
    # numerator and Fisher errors give through |inv(F_ideal)|.  The numerator's entries are sums
    # over the grid of FP32 products, whose rounding is about 1e-7 of the multipole's largest
    # entry; where harmonic weights of both signs cancel (the local l = 4 entry of bins (0, 1, 2) is
    # 2e-3 of its multipole's maximum) that is above 1e-5 of the entry, so each entry is held to
    # 1e-5 of max(|entry|, 0.1 x the multipole's maximum).
    num, num_ref = np.asarray(bi.Bk_numerator_ideal(d)), np.asarray(bio.Bk_numerator_ideal(d))
    n3 = len(num_ref) // 3
    for j in range(3):
        blk = slice(j * n3, (j + 1) * n3)
        err, a = np.abs(num[blk] - num_ref[blk]), np.abs(num_ref[blk])
        i = int(np.argmax(err / a))
        print("numerator l=%d: worst entry %d at %.1e of the maximum, error %.2e of it, %.2e of the maximum"
              % (2 * j, i, a[i] / a.max(), err[i] / a[i], err[i] / a.max()))
        close(num[blk], num_ref[blk], scale=own_scale(num_ref[blk], floor=0.1), what="numerator l=%d" % (2 * j))
    close(bi.fish_ideal, bio.fish_ideal, scale=fisher_scale(bio.fish_ideal))
    parity.check_solve("Bk_ideal lmax 4 %s" % case, got, ref, ['b0', 'b2', 'b4'], bio.fish_ideal, bi.fish_ideal,
                       num_ref, num)
    bs, bo = pb.BSpec(b, edges, lmax=4, mask=mask), O.BSpec(o, edges, lmax=4, mask=mask)
    F1, F2 = bs.compute_fisher_contribution(2), bo.compute_fisher_contribution(2)
    close(F1, F2, scale=fisher_scale(F2))


def test_integral_constraints_on_combined_rows(hip):
    """Integral constraints where the row pass combines the harmonic sub-maps (rows of 128 points, a
    16 x 16 x 128 grid in a 500 x 500 x 1000 box, local sightline): a Bk MC Fisher seed with the
    radial constraint at lmax 2 (three combined row transforms) and a Pk MC Fisher seed with the global
    constraint at lmax 4 (five), against the oracle."""
    import polybin3d_amd as pb
    shape, box, c = [16, 16, 128], [500., 500., 1000.], [10., 20., 30.]
    kw = dict(boxcenter=c, sightline='local')
    x = [np.fft.fftshift(np.arange(-shape[i] // 2, shape[i] // 2)) * box[i] / shape[i] + c[i] for i in range(3)]
    X = np.meshgrid(*x, indexing='ij')
    r = np.sqrt(X[0] ** 2 + X[1] ** 2 + X[2] ** 2)
    mask = 0.2 + ((r >= 150.) & (r <= 480.)).astype(np.float64)
    mask_IC = mask * (1.2 + 0.1 * np.cos(np.arange(shape[2]) * 0.05))[None, None, :]
    Pk = [np.linspace(0, 2, 500), 3e3 / (1 + (np.linspace(0, 2, 500) / 0.05) ** 2)]
    b, o = pb.PolyBin3D(box, shape, Pk=Pk, **kw), O.PolyBin3D(box, shape, Pk=Pk, **kw)
    ric = dict(add_RIC=True, mask_IC=mask_IC, radial_bins_RIC=[100., 250., 400., 600.])
    e3 = np.linspace(0.013, 0.07, 4)
    F1 = pb.BSpec(b, e3, lmax=2, mask=mask, **ric).compute_fisher_contribution(1)
    F2 = O.BSpec(o, e3, lmax=2, mask=mask, **ric).compute_fisher_contribution(1)
    close(F1, F2, scale=fisher_scale(F2), what="Bk RIC lmax 2")
    gic = dict(add_GIC=True, mask_IC=mask_IC)
    e2 = np.linspace(0.013, 0.09, 6)
    P1 = pb.PSpec(b, e2, lmax=4, mask=mask, **gic).compute_fisher_contribution(2)
    P2 = O.PSpec(o, e2, lmax=4, mask=mask, **gic).compute_fisher_contribution(2)
    close(P1, P2, scale=fisher_scale(P2), what="Pk GIC lmax 4")
