"""real_fft=False: the reference's complex-to-complex conventions (PolyBin3D/base.py:18-27, 53, 86-96,
176-229) at the boundary, over the same half-spectrum transforms inside.

Pins: the public transforms against numpy's fftn / ifftn; generate_data's full grid is Hermitian and
equal to the real_fft=True half spectrum where they overlap; the estimators (ideal and MC Fisher, with
and without a user S^-1 that sees full grids) against the oracle run with real_fft=True, which the
reference's own full-grid sums equal for the Fourier transforms of real maps when every bin lies below
the kz Nyquist plane (each pair k, -k of the full grid is one half-spectrum mode of weight degk = 2);
MapUtils and IntegrationUtils against numpy sums over the full grid, including non-Hermitian inputs.
Tolerance 1e-5 relative to each quantity's scale, as for the other parity tests."""
import inspect

import numpy as np
import pytest
import torch

from oracle import polybin_oracle as O

import parity
from parity import own_scale

pytestmark = pytest.mark.gpu

L = 1000.


def close(got, ref, tol=1e-5, scale=None, what=None):
    """parity.check labelled by the calling test and line."""
    f = inspect.currentframe().f_back
    label = "%s:%d" % (f.f_code.co_name, f.f_lineno)
    got, ref = [x.cpu().numpy() if torch.is_tensor(x) else np.asarray(x) for x in (got, ref)]
    parity.check(label if what is None else "%s %s" % (label, what), got, ref, scale=scale, tol=tol)


def fisher_scale(F):
    d = np.abs(np.diag(F))
    return np.sqrt(np.outer(d, d))


def powerlaw_field(n, seed=0):
    kk = np.linspace(0, 2, 4000)
    P = np.zeros_like(kk)
    P[1:] = 2e4 * (kk[1:] / 0.1) ** -1.5
    return O.PolyBin3D(L, list(n), Pk=[kk, P]).generate_data(seed=seed, output_type='real')


def neg_k(x):
    """x(-k) on a full FFT grid."""
    return np.roll(np.flip(x, axis=(0, 1, 2)), 1, axis=(0, 1, 2))


def test_full_grid_transforms_and_grids(hip):
    import polybin3d_amd as pb
    n = (16, 24, 32)
    b = pb.PolyBin3D(L, list(n), real_fft=False)
    assert b.degeneracy_factor is None
    kg = np.meshgrid(*[np.fft.fftshift(np.arange(-m // 2, m // 2)) * (2 * np.pi / L) for m in n], indexing='ij')
    for i in range(3):
        assert np.allclose(b.k_grids[i], kg[i], rtol=1e-14, atol=0)
    assert np.allclose(b.modk_grid, np.sqrt(kg[0] ** 2 + kg[1] ** 2 + kg[2] ** 2))
    rng = np.random.RandomState(0)
    x = rng.randn(*n)
    k = b.to_fourier(x)
    assert tuple(k.shape) == n
    ref = np.fft.fftn(x)
    close(k, ref, scale=np.abs(ref).max())
    z = rng.randn(*n) + 1j * rng.randn(*n)  # not Hermitian: to_real keeps Re IFFT as the reference
    ref = np.fft.ifftn(z).real
    close(b.to_real(z), ref, scale=np.abs(ref).max())
    close(b.to_real(k), x, scale=np.abs(x).max())
    Pk = [np.linspace(0, 2, 500), 1e4 / (1 + (np.linspace(0, 2, 500) / 0.05) ** 2)]
    b = pb.PolyBin3D(L, list(n), Pk=Pk, real_fft=False)
    bt = pb.PolyBin3D(L, list(n), Pk=Pk)
    f = b.generate_data(seed=5, output_type='fourier').cpu().numpy()
    assert f.shape == n
    close(f, np.conj(neg_k(f)), scale=np.abs(f).max(), what="Hermitian")
    close(f[..., :n[2] // 2 + 1], bt.generate_data(seed=5, output_type='fourier'), scale=np.abs(f).max())
    close(b.to_real(f), b.generate_data(seed=5, output_type='real'), scale=np.abs(b.to_real(f).cpu().numpy()).max())
    with pytest.raises(NotImplementedError):
        pb.PolyBin3D(L, [16, 16, 15], real_fft=False)


def test_full_grid_map_and_integration_utils(hip):
    """MapUtils / IntegrationUtils on full grids (utils.pyx:51-178, 213-323 with real_fft=False) against
    numpy sums over every full-grid mode, for Hermitian and non-Hermitian inputs and odd and even l."""
    import polybin3d_amd as pb
    n = (16, 16, 16)
    b = pb.PolyBin3D(L, list(n), boxcenter=[300., 200., 800.], sightline='local', real_fft=False)
    modk, mu = b.modk_grid, b.muk_grid
    kg = b.k_grids
    los = np.asarray(b.boxcenter) / np.linalg.norm(b.boxcenter)
    ref_mu = np.where(modk > 0, (los[0] * kg[0] + los[1] * kg[1] + los[2] * kg[2]) / np.where(modk > 0, modk, 1.), 0.)
    assert np.allclose(mu, ref_mu)  # base.py:111-116 on the full grid
    rng = np.random.RandomState(1)
    herm = np.fft.fftn(rng.randn(*n))
    other = rng.randn(*n) + 1j * rng.randn(*n)
    edges = np.linspace(0.0, 0.045, 6)  # below the Nyquist planes (0.0503)

    def leg(l):
        return np.polynomial.legendre.legval(mu, [0] * l + [1])

    def shells(f, l):
        return np.array([np.sum((f * leg(l))[(modk >= lo) & (modk < hi)]) for lo, hi in zip(edges[:-1], edges[1:])])
    for s1, s2, what in ((herm, herm, "hh"), (herm, other, "ho"), (other, other, "oo")):
        for l in (0, 1, 2, 3):
            ref = shells(np.real(s1 * np.conj(s2)), l)
            close(b.integrator.cross_integrate(s1, s2, edges, l), ref, scale=np.abs(shells(np.abs(s1 * s2), 0)).max(),
                  what="%s l=%d" % (what, l))
        ref = shells(np.abs(s1) ** 2, 2)
        close(b.integrator.integrate(s1, edges, 2), ref, scale=shells(np.abs(s1) ** 2, 0).max(), what=what)
    Q = np.stack([herm, other, 2 * other])
    sel = (modk >= edges[1]) & (modk < edges[2])
    ref = np.array([np.sum(np.real(np.conj(herm) * q)[sel]) for q in Q])
    close(b.integrator.integrate_row(herm, Q, edges[1], edges[2]), ref, scale=np.abs(ref).max())
    for src in (herm, other):
        for l in (1, 2):
            sel = (modk >= edges[1]) & (modk < edges[4])
            ref = np.where(sel, src * leg(l), 0.)
            close(b.map_utils.fourier_filter(src, l, edges[1], edges[4]), ref, scale=np.abs(ref).max(), what="l=%d" % l)
        w = 1. / (1. + (modk / 0.03) ** 2)
        close(b.map_utils.prod_fourier(src, w), src * w, scale=np.abs(src * w).max())
        close(b.applyAinv(src, invPk0_grid=w, input_type='fourier', output_type='fourier'), src * w,
              scale=np.abs(src * w).max(), what="applyAinv")
        close(b.map_utils.div_fourier(src, w), src / w, scale=np.abs(src / w).max(), what="div_fourier")
        sel = (modk >= edges[1]) & (modk < edges[4])
        ref = np.where(sel, src * w, 0.)
        close(b.map_utils.prod_fourier_filter(src, w, edges[1], edges[4]), ref, scale=np.abs(ref).max(),
              what="prod_fourier_filter")
    coeff = np.array([1. + 2j, -0.5 + 1j, 3., -2j, 0.7 - 0.3j])
    for s1, s2, what in ((herm, herm, "hh"), (herm, other, "ho"), (other, other, "oo")):
        ref = coeff.real * shells(np.real(s1 * np.conj(s2)), 0) - coeff.imag * shells(np.imag(s1 * np.conj(s2)), 0)
        close(b.integrator.cross_integrate_wcoeff(coeff, s1, s2, edges), ref,
              scale=np.abs(coeff) * shells(np.abs(s1 * s2), 0).max(), what="wcoeff " + what)


@pytest.mark.parametrize("case", ["global_l4", "local_l2"])
def test_full_grid_pk_ideal_and_mc_fisher(hip, case):
    """Pk_ideal, compute_fisher_contribution and Pk_unwindowed with real_fft=False against the oracle
    (real_fft=True): the full grid's mode sums equal the half spectrum's with degk below Nyquist."""
    import polybin3d_amd as pb
    n = 32
    edges = np.linspace(0.01, 0.09, 9)
    kw = {} if case == "global_l4" else dict(boxcenter=[10., 10., 10.], sightline='local')
    lmax = 4 if case == "global_l4" else 2
    Pk = [np.linspace(0, 2, 500), 1e4 / (1 + (np.linspace(0, 2, 500) / 0.05) ** 2)]
    b = pb.PolyBin3D(L, [n] * 3, Pk=Pk, real_fft=False, **kw)
    o = O.PolyBin3D(L, [n] * 3, Pk=Pk, **kw)
    d = powerlaw_field((n, n, n), seed=3)
    ps, po = pb.PSpec(b, edges, lmax=lmax), O.PSpec(o, edges, lmax=lmax)
    got, ref = ps.Pk_ideal(d), po.Pk_ideal(d)
    close(ps.fish_ideal, po.fish_ideal, scale=fisher_scale(po.fish_ideal))
    parity.check_solve("Pk_ideal full grid %s" % case, got, ref, sorted(ref), po.fish_ideal, ps.fish_ideal,
                       po.Pk_numerator_ideal(d), ps.Pk_numerator_ideal(d))
    mask = 0.5 + 0.5 * np.cos(np.linspace(0, 2 * np.pi, n))[:, None, None] * np.ones((n, n, n))
    ps = pb.PSpec(b, edges, lmax=lmax, mask=mask)
    po = O.PSpec(o, edges, lmax=lmax, mask=mask)
    F1, F2 = ps.compute_fisher_contribution(1), po.compute_fisher_contribution(1)
    close(F1, F2, scale=fisher_scale(F2))
    fish = po.compute_fisher(2)
    got, ref = ps.Pk_unwindowed(d * mask, fish=fish), po.Pk_unwindowed(d * mask, fish=fish)
    parity.check_solve("Pk_unwindowed full grid %s" % case, got, ref, sorted(ref), fish, fish,
                       po.Pk_numerator(d * mask), ps.Pk_numerator(d * mask))


def test_full_grid_user_sinv(hip):
    """A user S^-1 written for real_fft=False (pspec.py:18-43): its Fourier inputs and outputs are full
    grids.  The MC Fisher and shot contributions equal the oracle's with the same filter on the half
    spectrum."""
    import polybin3d_amd as pb
    n = 32
    edges = np.linspace(0.01, 0.09, 9)
    kw = dict(boxcenter=[10., 10., 10.], sightline='local')
    Pk = [np.linspace(0, 2, 500), 1e4 / (1 + (np.linspace(0, 2, 500) / 0.05) ** 2)]
    b = pb.PolyBin3D(L, [n] * 3, Pk=Pk, real_fft=False, **kw)
    o = O.PolyBin3D(L, [n] * 3, Pk=Pk, **kw)
    mask = 0.4 + 0.6 * np.exp(-np.linspace(-1, 1, n) ** 2)[None, :, None] * np.ones((n, n, n))
    wk_o = 1. / (1. + (o.modk_grid / 0.03) ** 2)
    wk_full = torch.from_numpy((1. / (1. + (b.modk_grid / 0.03) ** 2)).astype(np.float32)).cuda()
    seen = set()

    def sinv_o(m, input_type='real', output_type='fourier'):
        f = (o.to_fourier(m) if input_type == 'real' else m) * wk_o
        return f if output_type == 'fourier' else o.to_real(f)

    def sinv_g(m, input_type='real', output_type='fourier'):
        if input_type == 'fourier':
            seen.add(tuple(m.shape[-3:]))
        f = (b.to_fourier(m) if input_type == 'real' else m) * wk_full
        assert tuple(f.shape[-3:]) == (n, n, n)
        return f if output_type == 'fourier' else b.to_real(f)
    ps = pb.PSpec(b, edges, lmax=2, mask=mask, mask_shot=mask, applySinv=sinv_g)
    po = O.PSpec(o, edges, lmax=2, mask=mask, mask_shot=mask, applySinv=sinv_o)
    F1, F2 = ps.compute_fisher_contribution(2), po.compute_fisher_contribution(2)
    close(F1, F2, scale=fisher_scale(F2))
    s1, s2 = ps.compute_shot_contribution(4), po.compute_shot_contribution(4)
    for li in range(ps.Nl):
        sl = slice(li * ps.Nk, (li + 1) * ps.Nk)
        close(s1[sl], s2[sl], scale=own_scale(s2[sl]), what="l=%d" % (2 * li))
    assert seen <= {(n, n, n)}


def test_full_grid_bispectrum(hip):
    """Bk_ideal and one Bk MC Fisher contribution with real_fft=False against the oracle (real_fft=True)."""
    import polybin3d_amd as pb
    n = 16
    edges = np.linspace(0.007, 0.05, 4)
    Pk = [np.linspace(0, 2, 500), 3e3 / (1 + (np.linspace(0, 2, 500) / 0.05) ** 2)]
    b = pb.PolyBin3D(L, [n] * 3, Pk=Pk, real_fft=False)
    o = O.PolyBin3D(L, [n] * 3, Pk=Pk)
    d = powerlaw_field((n, n, n), seed=2)
    d = d + 0.5 * (d ** 2 - np.mean(d ** 2))
    bs, bo = pb.BSpec(b, edges, lmax=2), O.BSpec(o, edges, lmax=2)
    got, ref = bs.Bk_ideal(d), bo.Bk_ideal(d)
    close(bs.fish_ideal, bo.fish_ideal, scale=fisher_scale(bo.fish_ideal))
    parity.check_solve("Bk_ideal full grid", got, ref, ['b0', 'b2'], bo.fish_ideal, bs.fish_ideal,
                       bo.Bk_numerator_ideal(d), bs.Bk_numerator_ideal(d))
    mask = 0.2 + np.linspace(0, 1, n)[:, None, None] * np.ones((n, n, n))
    bs, bo = pb.BSpec(b, edges, lmax=2, mask=mask), O.BSpec(o, edges, lmax=2, mask=mask)
    F1, F2 = bs.compute_fisher_contribution(0), bo.compute_fisher_contribution(0)
    close(F1, F2, scale=fisher_scale(F2))


def test_full_grid_odd_l(hip):
    """odd_l (l = 0..3, local sightline, masked) with real_fft=False: the numerator and an MC Fisher
    seed against the oracle (real_fft=True).  The odd-l weights are odd under k -> -k; they act inside
    the half-spectrum engine, and only the real map and the real-space mask cross the boundary."""
    import polybin3d_amd as pb
    n = 32
    edges = np.linspace(0.01, 0.09, 9)
    kw = dict(boxcenter=[10., 10., 10.], sightline='local')
    Pk = [np.linspace(0, 2, 500), 1e4 / (1 + (np.linspace(0, 2, 500) / 0.05) ** 2)]
    b = pb.PolyBin3D(L, [n] * 3, Pk=Pk, real_fft=False, **kw)
    o = O.PolyBin3D(L, [n] * 3, Pk=Pk, **kw)
    mask = 0.3 + np.linspace(0, 1, n)[None, None, :] * np.linspace(0.5, 1, n)[:, None, None] * np.ones((n, n, n))
    ps = pb.PSpec(b, edges, lmax=3, mask=mask, odd_l=True)
    po = O.PSpec(o, edges, lmax=3, mask=mask, odd_l=True)
    d = powerlaw_field((n, n, n), seed=4) * mask
    n1, n2 = ps.Pk_numerator(d), po.Pk_numerator(d)
    close(n1, n2, scale=np.abs(n2[:ps.Nk]).max())
    F1, F2 = ps.compute_fisher_contribution(1), po.compute_fisher_contribution(1)
    close(F1, F2, scale=fisher_scale(F2))


def test_full_grid_bk_user_sinv(hip):
    """BSpec MC Fisher with a user S^-1 written for full grids (the generic path: the callable runs on
    every Q map, bspec.py:499-505) and real_fft=False, against the oracle (real_fft=True) with the same
    filter on the half spectrum."""
    import polybin3d_amd as pb
    n = 16
    edges = np.linspace(0.007, 0.05, 4)
    kw = dict(boxcenter=[10., 10., 10.], sightline='local')
    Pk = [np.linspace(0, 2, 500), 3e3 / (1 + (np.linspace(0, 2, 500) / 0.05) ** 2)]
    b = pb.PolyBin3D(L, [n] * 3, Pk=Pk, real_fft=False, **kw)
    o = O.PolyBin3D(L, [n] * 3, Pk=Pk, **kw)
    mask = 0.4 + np.linspace(0, 1, n)[None, :, None] * np.ones((n, n, n))
    wk_o = 1. / (1. + (o.modk_grid / 0.02) ** 2)
    wk_full = torch.from_numpy((1. / (1. + (b.modk_grid / 0.02) ** 2)).astype(np.float32)).cuda()
    seen = set()

    def sinv_o(m, input_type='real', output_type='fourier'):
        f = (o.to_fourier(m) if input_type == 'real' else m) * wk_o
        return f if output_type == 'fourier' else o.to_real(f)

    def sinv_g(m, input_type='real', output_type='fourier'):
        if input_type == 'fourier':
            seen.add(tuple(m.shape[-3:]))
        f = (b.to_fourier(m) if input_type == 'real' else m) * wk_full
        return f if output_type == 'fourier' else b.to_real(f)
    bs = pb.BSpec(b, edges, lmax=2, mask=mask, applySinv=sinv_g)
    bo = O.BSpec(o, edges, lmax=2, mask=mask, applySinv=sinv_o)
    F1, F2 = bs.compute_fisher_contribution(3), bo.compute_fisher_contribution(3)
    close(F1, F2, scale=fisher_scale(F2))
    assert seen <= {(n, n, n)}
