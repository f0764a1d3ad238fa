"""GPU parity of the power-spectrum path against the FP64 oracle on identical seeded inputs.
Tolerances: spectra and Fisher entries within 1e-5 relative to the per-bin scale (each multipole's own
|value|, floored at 1e-3 of its largest bin, or, for Fisher entries that cross zero, sqrt(F_ii F_jj)),
the bound the FP32 kernels are held to; every check prints its error."""
import inspect

import numpy as np
import pytest
import torch

from oracle import polybin_oracle as O

import parity
from parity import own_scale

pytestmark = pytest.mark.gpu

L = 1000.


def powerlaw_field(n, seed=0):
    """float64 GRF with P(k) = 2e4 (k/0.1)^-1.5 (the Pk_ideal workload's input), via the oracle."""
    kk = np.linspace(0, 2, 4000)
    P = np.zeros_like(kk)
    P[1:] = 2e4 * (kk[1:] / 0.1) ** -1.5
    o = O.PolyBin3D(L, list(n), Pk=[kk, P])
    return o.generate_data(seed=seed, output_type='real')


def close(got, ref, tol=1e-5, scale=None, what=None):
    """parity.check labelled by the calling test and line."""
    f = inspect.currentframe().f_back
    label = "%s:%d" % (f.f_code.co_name, f.f_lineno)
    parity.check(label if what is None else "%s %s" % (label, what), got, ref, scale=scale, tol=tol)


def test_bin_reduce_matches_oracle(hip):
    import polybin3d_amd as pb
    n = (32, 32, 32)
    b = pb.PolyBin3D(L, list(n), boxcenter=[300., 200., 800.], sightline='local')
    o = O.PolyBin3D(L, list(n), boxcenter=[300., 200., 800.], sightline='local')
    rng = np.random.RandomState(0)
    s1 = (rng.randn(*o.modk_grid.shape) + 1j * rng.randn(*o.modk_grid.shape)).astype(np.complex64)
    s2 = (rng.randn(*o.modk_grid.shape) + 1j * rng.randn(*o.modk_grid.shape)).astype(np.complex64)
    r = rng.rand(*o.modk_grid.shape).astype(np.float32)
    edges = np.linspace(0.007, 0.1, 13)
    for l in (0, 2, 4):
        close(b.integrator.cross_integrate(s1, s2, edges, l), o.integrator.cross_integrate(s1, s2, edges, l),
              scale=o.integrator.integrate(s1, edges, 0))
        close(b.integrator.integrate(s1, edges, l), o.integrator.integrate(s1, edges, l),
              scale=o.integrator.integrate(s1, edges, 0))
        close(b.integrator.real_integrate_double(r, l, 2, edges), o.integrator.real_integrate_double(r, l, 2, edges),
              scale=o.integrator.real_integrate_double(r, 0, 0, edges))
    Q = np.stack([s1, s2, s1 * 2])
    close(b.integrator.integrate_row(s2, Q, edges[3], edges[4]), o.integrator.integrate_row(s2, Q, edges[3], edges[4]),
          scale=np.abs(o.integrator.integrate(s2, edges, 0)[3]))


def test_map_and_integration_utils_match_oracle(hip):
    """The rest of the reference's MapUtils / IntegrationUtils contract (utils.pyx:139-157, 213-383):
    div_fourier, prod_fourier_filter, div_real, prod_real_diff, sum_pair, sum_triplet and
    cross_integrate_wcoeff against the oracle (or numpy for the real-space ones) on the same inputs."""
    import polybin3d_amd as pb
    n = (32, 32, 32)
    b = pb.PolyBin3D(L, list(n), boxcenter=[300., 200., 800.], sightline='local')
    o = O.PolyBin3D(L, list(n), boxcenter=[300., 200., 800.], sightline='local')
    rng = np.random.RandomState(2)
    hshape = o.modk_grid.shape
    m1 = (rng.randn(*hshape) + 1j * rng.randn(*hshape)).astype(np.complex64)
    w = np.where(rng.rand(*hshape) < 0.2, 0., 0.5 + rng.rand(*hshape)).astype(np.float32)
    ref = o.map_utils.div_fourier(m1.astype(np.complex128), w.astype(np.float64))
    close(b.map_utils.div_fourier(m1, w).cpu().numpy(), ref, scale=np.abs(ref).max())
    assert np.all(b.map_utils.div_fourier(m1, w).cpu().numpy()[w == 0] == 0)
    ref = o.map_utils.prod_fourier_filter(m1.astype(np.complex128), w.astype(np.float64), 0.02, 0.08)
    close(b.map_utils.prod_fourier_filter(m1, w, 0.02, 0.08).cpu().numpy(), ref, scale=np.abs(ref).max())
    x = [rng.randn(*n).astype(np.float32) for _ in range(4)]
    x[1][rng.rand(*n) < 0.1] = 0.
    ref = np.where(x[1] != 0, x[0] / np.where(x[1] != 0, x[1], 1.), 0.)
    close(b.map_utils.div_real(x[0], x[1]).cpu().numpy(), ref, scale=np.abs(ref).max())
    ref = x[0].astype(np.float64) * x[1] - x[2].astype(np.float64) * x[3]
    close(b.map_utils.prod_real_diff(*x).cpu().numpy(), ref, scale=np.abs(ref).max())
    sp = b.map_utils.sum_pair(x[0], x[2])
    close(np.array([sp]), np.array([O.sum_pair(x[0].astype(np.float64), x[2])]), scale=np.sqrt(np.sum(x[0].astype(np.float64) ** 2 * x[2] ** 2)))
    st = b.map_utils.sum_triplet(x[0], x[2], x[3])
    close(np.array([st]), np.array([O.sum_triplet(x[0].astype(np.float64), x[2], x[3])]),
          scale=np.sqrt(np.sum((x[0].astype(np.float64) * x[2] * x[3]) ** 2)))
    s1 = np.fft.rfftn(x[0]).astype(np.complex64)
    s2 = np.fft.rfftn(x[2]).astype(np.complex64)
    edges = np.linspace(0.007, 0.1, 9)
    coeff = rng.randn(8) + 1j * rng.randn(8)
    ref = o.integrator.cross_integrate_wcoeff(coeff, s1.astype(np.complex128), s2.astype(np.complex128), edges)
    scale = np.abs(coeff) * o.integrator.integrate(s1.astype(np.complex128), edges, 0) ** 0.5 * \
        o.integrator.integrate(s2.astype(np.complex128), edges, 0) ** 0.5
    close(b.integrator.cross_integrate_wcoeff(coeff, s1, s2, edges), ref, scale=scale)


@pytest.mark.parametrize("case", ["64_global_l2", "64_local_l4", "128_global_l4"])
def test_pk_ideal(hip, case):
    import polybin3d_amd as pb
    if case == "64_global_l2":
        n, kw, lmax, edges = 64, {}, 2, np.arange(0.01, 0.2001, 0.01)
    elif case == "64_local_l4":
        n, kw, lmax, edges = 64, dict(boxcenter=[0., 1170., 826.], sightline='local'), 4, np.arange(0.01, 0.2001, 0.01)
    else:
        n, kw, lmax, edges = 128, {}, 4, np.arange(0.01, 0.3001, 0.01)
    d = powerlaw_field((n, n, n))
    b = pb.PolyBin3D(L, [n] * 3, **kw)
    o = O.PolyBin3D(L, [n] * 3, **kw)
    ps, po = pb.PSpec(b, edges, lmax=lmax), O.PSpec(o, edges, lmax=lmax)
    got, ref = ps.Pk_ideal(d), po.Pk_ideal(d)
    close(ps.fish_ideal, po.fish_ideal, scale=np.abs(np.diag(po.fish_ideal)).max())
    parity.check_solve("test_pk_ideal %s" % case, got, ref, sorted(ref), po.fish_ideal, ps.fish_ideal,
                       po.Pk_numerator_ideal(d), ps.Pk_numerator_ideal(d))
    assert (b.n_FFTs_forward, b.n_FFTs_reverse) == (o.n_FFTs_forward, o.n_FFTs_reverse)


def test_pk_fisher_ideal_no_discreteness(hip):
    import polybin3d_amd as pb
    n = 32
    edges = np.linspace(0.01, 0.09, 9)
    b, o = pb.PolyBin3D(L, [n] * 3), O.PolyBin3D(L, [n] * 3)
    f1 = pb.PSpec(b, edges, lmax=4).compute_fisher_ideal(discreteness_correction=False)
    f2 = O.PSpec(o, edges, lmax=4).compute_fisher_ideal(discreteness_correction=False)
    close(f1, f2, scale=np.abs(f2).max())


def shell_mask(n, L, center, seed=1, r_lo=300., r_hi=480., nholes=200, smooth=5.):
    """Shell r in [r_lo, r_hi] around the observer at the origin, minus random holes, Gaussian-smoothed."""
    x = [np.fft.fftshift(np.arange(-n // 2, n // 2)) * L / n + center[i] for i in range(3)]
    X = np.meshgrid(*x, indexing='ij')
    r = np.sqrt(X[0] ** 2 + X[1] ** 2 + X[2] ** 2)
    m = ((r >= r_lo) & (r <= r_hi)).astype(np.float64)
    rng = np.random.RandomState(seed)
    for _ in range(nholes):
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        c = u * rng.uniform(r_lo, r_hi)
        rad = rng.uniform(10., 30.)
        m[(X[0] - c[0]) ** 2 + (X[1] - c[1]) ** 2 + (X[2] - c[2]) ** 2 < rad ** 2] = 0.
    k = [np.fft.fftfreq(n, L / n) * 2 * np.pi] * 2 + [np.fft.rfftfreq(n, L / n) * 2 * np.pi]
    K = np.meshgrid(*k, indexing='ij')
    g = np.exp(-0.5 * smooth ** 2 * (K[0] ** 2 + K[1] ** 2 + K[2] ** 2))
    return np.fft.irfftn(np.fft.rfftn(m) * g, s=(n, n, n), axes=(0, 1, 2))


@pytest.mark.parametrize("case", ["global_const", "local_mask"])
def test_pk_mc_fisher_and_unwindowed(hip, case):
    import polybin3d_amd as pb
    n = 32
    edges = np.linspace(0.01, 0.09, 9)
    if case == "global_const":
        kw, mask, lmax = {}, 0.8 * np.ones((n, n, n)), 4
    else:
        kw = dict(boxcenter=[10., 10., 10.], sightline='local')
        mask, lmax = shell_mask(n, L, [10., 10., 10.]), 4
    Pk = [np.linspace(0, 2, 500), 1e4 / (1 + (np.linspace(0, 2, 500) / 0.05) ** 2)]
    b, o = pb.PolyBin3D(L, [n] * 3, Pk=Pk, **kw), O.PolyBin3D(L, [n] * 3, Pk=Pk, **kw)
    ps = pb.PSpec(b, edges, lmax=lmax, mask=mask, mask_shot=mask)
    po = O.PSpec(o, edges, lmax=lmax, mask=mask, mask_shot=mask)
    for seed in (0, 1):
        F1 = ps.compute_fisher_contribution(seed)
        F2 = po.compute_fisher_contribution(seed)
        sc = np.sqrt(np.outer(np.abs(np.diag(F2)), np.abs(np.diag(F2))))
        close(F1, F2, scale=sc)
    s1, s2 = ps.compute_shot_contribution(5), po.compute_shot_contribution(5)
    close(s1, s2, scale=np.abs(s2[:ps.Nk]).max())
    d = powerlaw_field((n, n, n), seed=3) * mask
    fish = po.compute_fisher(2)
    close(ps.compute_fisher(2), fish, scale=np.sqrt(np.outer(np.abs(np.diag(fish)), np.abs(np.diag(fish)))))
    got = ps.Pk_unwindowed(d, fish=fish)
    ref = po.Pk_unwindowed(d, fish=fish)
    parity.check_solve("Pk_unwindowed %s" % case, got, ref, sorted(ref), fish, fish, po.Pk_numerator(d),
                       ps.Pk_numerator(d))


@pytest.mark.parametrize("case", ["global_const", "local_mask"])
def test_pk_mc_fisher_batched_seeds(hip, case):
    """compute_fisher(N_it, batch=R): the R seeds of a batch share every pass (their GRFs, filters and
    Q maps are batch entries of the same launches, sub-map inputs grouped per seed); the result is the
    oracle's mean over the same seeds."""
    import polybin3d_amd as pb
    n = 32
    edges = np.linspace(0.01, 0.09, 9)
    if case == "global_const":
        kw, mask = {}, 0.8 * np.ones((n, n, n))
    else:
        kw = dict(boxcenter=[10., 10., 10.], sightline='local')
        mask = shell_mask(n, L, [10., 10., 10.])
    Pk = [np.linspace(0, 2, 500), 1e4 / (1 + (np.linspace(0, 2, 500) / 0.05) ** 2)]
    b, o = pb.PolyBin3D(L, [n] * 3, Pk=Pk, **kw), O.PolyBin3D(L, [n] * 3, Pk=Pk, **kw)
    ps = pb.PSpec(b, edges, lmax=4, mask=mask)
    po = O.PSpec(o, edges, lmax=4, mask=mask)
    fish = po.compute_fisher(3)
    sc = np.sqrt(np.outer(np.abs(np.diag(fish)), np.abs(np.diag(fish))))
    close(ps.compute_fisher(3, batch=3), fish, scale=sc)
    close(ps.compute_fisher(3, batch=2), fish, scale=sc)


@pytest.mark.parametrize("case", ["gic_global", "ric_local"])
def test_pk_mc_fisher_integral_constraints(hip, case):
    """MC Fisher with the global / radial integral-constrained pointing matrix (pspec.py:199-213;
    utils.pyx:405-513): per seed against the oracle, and the batched-seed mean."""
    import polybin3d_amd as pb
    n = 32
    edges = np.linspace(0.01, 0.09, 9)
    if case == "gic_global":
        kw, ic = {}, dict(add_GIC=True)
    else:
        kw, ic = dict(boxcenter=[10., 10., 10.], sightline='local'), dict(add_RIC=True,
                                                                          radial_bins_RIC=[100., 250., 400., 600.])
    mask = shell_mask(n, L, kw.get('boxcenter', [0., 0., 0.]))
    mask_IC = mask * (1.2 + 0.1 * np.cos(np.arange(n) * 0.3))[None, None, :]
    Pk = [np.linspace(0, 2, 500), 1e4 / (1 + (np.linspace(0, 2, 500) / 0.05) ** 2)]
    b, o = pb.PolyBin3D(L, [n] * 3, Pk=Pk, **kw), O.PolyBin3D(L, [n] * 3, Pk=Pk, **kw)
    ps = pb.PSpec(b, edges, lmax=4, mask=mask, mask_IC=mask_IC, **ic)
    po = O.PSpec(o, edges, lmax=4, mask=mask, mask_IC=mask_IC, **ic)
    for seed in (0, 1):
        F1 = ps.compute_fisher_contribution(seed)
        F2 = po.compute_fisher_contribution(seed)
        close(F1, F2, scale=np.sqrt(np.outer(np.abs(np.diag(F2)), np.abs(np.diag(F2)))))
    fish = po.compute_fisher(2)
    close(ps.compute_fisher(2, batch=2), fish, scale=np.sqrt(np.outer(np.abs(np.diag(fish)), np.abs(np.diag(fish)))))


def test_pk_odd_l(hip):
    """odd_l (l = 0..3, local sightline, varying mask): numerator, MC Fisher per seed (single and
    batched seeds), shot noise and Pk_unwindowed against the oracle (pspec.py:410-411, 469-481,
    539-543, 607, 677-678)."""
    import polybin3d_amd as pb
    n = 32
    edges = np.linspace(0.01, 0.09, 9)
    kw = dict(boxcenter=[10., 10., 10.], sightline='local')
    mask = shell_mask(n, L, [10., 10., 10.])
    Pk = [np.linspace(0, 2, 500), 1e4 / (1 + (np.linspace(0, 2, 500) / 0.05) ** 2)]
    b, o = pb.PolyBin3D(L, [n] * 3, Pk=Pk, **kw), O.PolyBin3D(L, [n] * 3, Pk=Pk, **kw)
    ps = pb.PSpec(b, edges, lmax=3, mask=mask, odd_l=True, mask_shot=mask)
    po = O.PSpec(o, edges, lmax=3, mask=mask, odd_l=True, mask_shot=mask)
    assert ps.N_bins == po.N_bins == 4 * 8
    d = powerlaw_field((n, n, n), seed=3) * mask
    n1, n2 = ps.Pk_numerator(d), po.Pk_numerator(d)
    close(n1, n2, scale=np.abs(n2[:ps.Nk]).max())
    for seed in (0, 1):
        F1, F2 = ps.compute_fisher_contribution(seed), po.compute_fisher_contribution(seed)
        close(F1, F2, scale=np.sqrt(np.outer(np.abs(np.diag(F2)), np.abs(np.diag(F2)))))
    s1, s2 = ps.compute_shot_contribution(5), po.compute_shot_contribution(5)
    close(s1, s2, scale=np.abs(s2[:ps.Nk]).max())
    fish = po.compute_fisher(2)
    sc = np.sqrt(np.outer(np.abs(np.diag(fish)), np.abs(np.diag(fish))))
    close(ps.compute_fisher(2, batch=2), fish, scale=sc)
    got, ref = ps.Pk_unwindowed(d, fish=fish), po.Pk_unwindowed(d, fish=fish)
    assert sorted(got) == sorted(ref) == ['p0', 'p1', 'p2', 'p3']
    parity.check_solve("Pk_unwindowed odd_l", got, ref, ['p0', 'p1', 'p2', 'p3'], fish, fish, n2, n1)


def _diag_scale(F):
    d = np.abs(np.diag(F))
    return np.sqrt(np.outer(d, d))


@pytest.mark.parametrize("case", ["global_const_l4", "local_mask_fine", "local_wideangle", "local_odd_l",
                                  "local_user_sinv"])
def test_pk_theory_contribution(hip, case):
    """compute_theory_contribution (pspec.py:165-197, 381-436): the Fisher contribution and the
    rectangular theory binning matrix against the oracle per seed; theory bins finer than the data's
    and extending past them (the Q maps are written up to the theory k_max), higher theory lmax,
    the wide-angle odd legs with their per-bin weights, odd_l phases, and a user S^-1 (the
    one-Q-map-at-a-time path)."""
    import polybin3d_amd as pb
    import torch
    n = 32
    edges = np.linspace(0.01, 0.09, 9)
    local = dict(boxcenter=[10., 10., 10.], sightline='local')
    kw, mask, lmax, opts, th = local, shell_mask(n, L, [10., 10., 10.]), 2, {}, {}
    tb = np.linspace(0.01, 0.1, 19)
    if case == "global_const_l4":
        kw, mask, th = {}, 0.8 * np.ones((n, n, n)), dict(lmax_theory=4)
    elif case == "local_mask_fine":
        th = dict(lmax_theory=4)
    elif case == "local_wideangle":
        th = dict(lmax_theory=4, include_wideangle=True)
    elif case == "local_odd_l":
        lmax, opts, th = 1, dict(odd_l=True), dict(lmax_theory=3)
    Pk = [np.linspace(0, 2, 500), 1e4 / (1 + (np.linspace(0, 2, 500) / 0.05) ** 2)]
    b, o = pb.PolyBin3D(L, [n] * 3, Pk=Pk, **kw), O.PolyBin3D(L, [n] * 3, Pk=Pk, **kw)
    if case == "local_user_sinv":
        wk = 1. / (1. + (o.modk_grid / 0.03) ** 2)
        wk_d = torch.from_numpy(wk.astype(np.float32)).cuda()

        def sinv_o(m, input_type='real', output_type='fourier'):
            f = (o.to_fourier(m) if input_type == 'real' else m) * wk
            return f if output_type == 'fourier' else o.to_real(f)

        def sinv_g(m, input_type='real', output_type='fourier'):
            f = (b.to_fourier(m) if input_type == 'real' else m) * wk_d
            return f if output_type == 'fourier' else b.to_real(f)
        ps = pb.PSpec(b, edges, lmax=lmax, mask=mask, applySinv=sinv_g)
        po = O.PSpec(o, edges, lmax=lmax, mask=mask, applySinv=sinv_o)
        th = dict(lmax_theory=2)
    else:
        ps = pb.PSpec(b, edges, lmax=lmax, mask=mask, **opts)
        po = O.PSpec(o, edges, lmax=lmax, mask=mask, **opts)
    for seed in (0, 2):
        F1, B1 = ps.compute_theory_contribution(seed, tb, **th)
        F2, B2 = po.compute_theory_contribution(seed, tb, **th)
        assert B1.shape == B2.shape
        sc = _diag_scale(F2)
        close(F1, F2, scale=sc)
        # per row: the row's diagonal Fisher scale (binning entries of a row share its Q map)
        close(B1, B2, scale=np.abs(np.diag(F2))[:, None] * np.ones_like(B2))


@pytest.mark.parametrize("case", ["global_const_l4", "local_mask_shot", "local_gic"])
def test_pk_covariance_contribution(hip, case):
    """compute_covariance_contribution (pspec.py:149-163, 278-325): the MC-Fisher rows with the filter
    S^-1 [P xi P^T + N] S^-dag, xi = apply_xi of the theory spectrum (global P(k, mu); local P_0 and
    P_2 through the fused harmonic transforms), N = mask_shot; per seed against the oracle."""
    import polybin3d_amd as pb
    n = 32
    edges = np.linspace(0.01, 0.09, 9)
    kc = np.linspace(0, 1.5, 300)
    Pc = [kc, 2e4 / (1 + (kc / 0.04) ** 2), 5e3 / (1 + (kc / 0.05) ** 2), 1e3 / (1 + (kc / 0.05) ** 2)]
    ic = {}
    if case == "global_const_l4":
        kw, mask, lmax = {}, 0.8 * np.ones((n, n, n)), 4
    else:
        kw, mask, lmax, Pc = dict(boxcenter=[10., 10., 10.], sightline='local'), None, 2, Pc[:3]
        mask = shell_mask(n, L, [10., 10., 10.])
        if case == "local_gic":
            ic = dict(add_GIC=True, mask_IC=mask * 1.1)
    Pk = [np.linspace(0, 2, 500), 1e4 / (1 + (np.linspace(0, 2, 500) / 0.05) ** 2)]
    b, o = pb.PolyBin3D(L, [n] * 3, Pk=Pk, **kw), O.PolyBin3D(L, [n] * 3, Pk=Pk, **kw)
    ps = pb.PSpec(b, edges, lmax=lmax, mask=mask, mask_shot=0.5 * mask, **ic)
    po = O.PSpec(o, edges, lmax=lmax, mask=mask, mask_shot=0.5 * mask, **ic)
    for seed in (0, 1):
        C1 = ps.compute_covariance_contribution(seed, Pc)
        C2 = po.compute_covariance_contribution(seed, Pc)
        close(C1, C2, scale=_diag_scale(C2))


def test_apply_xi_matches_oracle(hip):
    """base.apply_xi on device maps (base.py:413-437), local sightline with P_0, P_2, P_4, both output
    types, against the oracle."""
    import polybin3d_amd as pb
    n = 32
    kw = dict(boxcenter=[10., 10., 10.], sightline='local')
    b, o = pb.PolyBin3D(L, [n] * 3, **kw), O.PolyBin3D(L, [n] * 3, **kw)
    kc = np.linspace(0, 1.5, 300)
    Pc = [kc, 2e4 / (1 + (kc / 0.04) ** 2), 5e3 / (1 + (kc / 0.05) ** 2), 1e3 / (1 + (kc / 0.05) ** 2)]
    po = O.PSpec(o, np.linspace(0.01, 0.09, 9), lmax=4)
    grid_o = po._cov_pk_grid(np.asarray(Pc))
    grid_g = b.pk_grid(Pc)
    u = np.random.RandomState(2).randn(n, n, n)
    uf = o.to_fourier(u)
    for ot in ('real', 'fourier'):
        ref = o.apply_xi(uf, u, po.Ylm_real, po.Ylm_fourier, grid_o, ot)
        got = b.apply_xi(uf.astype(np.complex64), u.astype(np.float32), None, None, grid_g, ot).cpu().numpy()
        close(got, ref, scale=np.abs(ref).max())


def _cg_filters(base, smooth_mask, beta, Pkl, Yr, Yf, dot, n_it, rnd=lambda v: v):
    """The conjugate-gradient S^-1 and S^-dag of the reference's BOSS tutorial (Tutorial 3, cells 45 and
    49): solve [W xi + beta] x = u (and [xi W + beta] x = u, nulled outside the mask) with apply_xi of a
    multipole spectrum; a fixed number of unpreconditioned iterations (converged to ~1e-7 here).
    rnd rounds every operator output and iterate (FP32 rounding of the FP64 oracle's CG)."""
    def fwd(m):
        return rnd(smooth_mask * rnd(base.apply_xi(base.to_fourier(m), m, Yr, Yf, Pk_grid=Pkl, output_type='real'))
                   + beta * m)

    def tr(m):
        mm = rnd(smooth_mask * m)
        return rnd((smooth_mask != 0) * rnd(base.apply_xi(base.to_fourier(mm), mm, Yr, Yf, Pk_grid=Pkl,
                                                          output_type='real')) + beta * m)

    def cg(u, A):
        x = u * 1.
        r = rnd(u - A(x))
        p = r * 1.
        Cp = A(p)
        old = dot(r, r)
        alpha = old / dot(p, Cp)
        for _ in range(n_it):
            x = rnd(x + alpha * p)
            r = rnd(r - alpha * Cp)
            new = dot(r, r)
            p = rnd(r + (new / old) * p)
            old = new
            Cp = A(p)
            alpha = old / dot(p, Cp)
        return x

    def wrap(A):
        def f(d, input_type='real', output_type='real'):
            u = base.to_real(d) if input_type == 'fourier' else d
            x = cg(u, A)
            return x if output_type == 'real' else base.to_fourier(x)
        return f
    return wrap(fwd), wrap(tr)


def test_pk_fisher_with_cg_sinv(hip):
    """MC Fisher with the tutorial's conjugate-gradient S^-1 (a user callable built on base.apply_xi with
    numpy spectrum grids, as the reference's callers write it) on the GPU against the same CG on the
    oracle.  The CG amplifies rounding: the FP64 oracle's own CG with every operator output and iterate
    rounded to FP32 moves the entries by `spread` (measured here, in the same test).  The GPU is held
    to 4 x spread of the diagonal scale, and spread itself to 1e-3 (a wrong term in apply_xi or the
    filter path moves entries by O(1))."""
    import polybin3d_amd as pb
    import torch
    n = 32
    edges = np.linspace(0.01, 0.09, 9)
    kw = dict(boxcenter=[10., 10., 10.], sightline='local')
    mask = shell_mask(n, L, [10., 10., 10.])
    Pk = [np.linspace(0, 2, 500), 1e4 / (1 + (np.linspace(0, 2, 500) / 0.05) ** 2)]
    b, o = pb.PolyBin3D(L, [n] * 3, Pk=Pk, **kw), O.PolyBin3D(L, [n] * 3, Pk=Pk, **kw)
    kk = np.linspace(0., 1., 400)
    Pkl = {0: np.interp(o.modk_grid, kk, 8e3 / (1 + (kk / 0.05) ** 2)),
           2: np.interp(o.modk_grid, kk, 3e3 / (1 + (kk / 0.05) ** 2))}
    beta = 2e4
    po_tmp = O.PSpec(o, edges, lmax=2)
    so, sto = _cg_filters(o, mask, beta, Pkl, po_tmp.Ylm_real, po_tmp.Ylm_fourier,
                          lambda a, c: float(np.sum(a * c)), 10)
    mask_d = torch.from_numpy(mask.astype(np.float32)).cuda()
    sg, stg = _cg_filters(b, mask_d, beta, Pkl, None, None,
                          lambda a, c: float((a.double() * c.double()).sum()), 10)
    ps = pb.PSpec(b, edges, lmax=2, mask=mask, applySinv=sg, applySinv_transpose=stg)
    po = O.PSpec(o, edges, lmax=2, mask=mask, applySinv=so, applySinv_transpose=sto)
    F1, F2 = ps.compute_fisher_contribution(1), po.compute_fisher_contribution(1)
    r32 = lambda v: np.asarray(v, dtype=np.float32).astype(np.float64)  # noqa: E731
    so32, sto32 = _cg_filters(o, mask, beta, Pkl, po_tmp.Ylm_real, po_tmp.Ylm_fourier,
                              lambda a, c: float(np.sum(a * c)), 10, rnd=r32)
    po32 = O.PSpec(o, edges, lmax=2, mask=mask, applySinv=so32, applySinv_transpose=sto32)
    spread = parity.scaled_error(po32.compute_fisher_contribution(1), F2, _diag_scale(F2))
    print("[parity] oracle CG with FP32 rounding: spread %.2e of the diagonal scale" % spread)
    assert spread < 1e-3
    close(F1, F2, tol=max(parity.TOL, 4. * spread), scale=_diag_scale(F2), what="GPU CG vs FP64 oracle CG")


def test_pk_fisher_mixed_grids(hip):
    """compute_fisher_mixed: the MC Fishers of a 32^3 and a 64^3 spec, issued together on a stream
    each, equal each spec's own compute_fisher (same seeds; FP64 summation order aside)."""
    import polybin3d_amd as pb
    edges = np.linspace(0.01, 0.09, 9)
    kw = dict(boxcenter=[10., 10., 10.], sightline='local')
    Pk = [np.linspace(0, 2, 500), 1e4 / (1 + (np.linspace(0, 2, 500) / 0.05) ** 2)]
    specs = []
    for n in (32, 64):
        b = pb.PolyBin3D(L, [n] * 3, Pk=Pk, **kw)
        specs.append(pb.PSpec(b, edges, lmax=4, mask=shell_mask(n, L, [10., 10., 10.])))
    mixed = pb.compute_fisher_mixed(specs, 5, batch=2)
    for ps, fm in zip(specs, mixed):
        ref = ps.compute_fisher(5, batch=2)
        np.testing.assert_allclose(fm, ref, rtol=1e-9, atol=1e-12 * np.abs(ref).max())
    with pytest.raises(ValueError):
        pb.compute_fisher_mixed([specs[0], pb.PSpec(specs[0].base, edges, lmax=2)], 1)


def test_bin_reduce_full_grid_and_mode_list_agree(hip):
    """pb3d_bin_reduce (a pass over the whole half grid) and pb3d_bin_reduce_modes (the list of in-bin
    modes) give the same sums, complex and real modes, with Legendre weights.  Sums are FP32 within a
    wave's run of one bin and FP64 across runs; the two walks group the modes into different runs,
    so they agree to FP32 rounding (1e-6 of the largest sum)."""
    import polybin3d_amd as pb
    from polybin3d_amd import _capi as C
    n = 64
    b = pb.PolyBin3D(L, [n] * 3, boxcenter=[300., 200., 800.], sightline='local')
    rng = np.random.RandomState(1)
    x = torch.from_numpy((rng.randn(3, *b.half_shape) + 1j * rng.randn(3, *b.half_shape)).astype(np.complex64)).cuda()
    y = torch.from_numpy((rng.randn(2, *b.half_shape) + 1j * rng.randn(2, *b.half_shape)).astype(np.complex64)).cuda()
    xr = torch.from_numpy(rng.rand(3, *b.half_shape).astype(np.float32)).cuda()
    edges = np.linspace(0.01, 0.12, 12)
    table, modes = b.bin_table(edges), b.bin_modes(edges)
    nb = len(edges) - 1
    for real in (0, 1):
        xx = xr if real else x
        yy = None if real else y
        ells = C.int_array([0, 2]) if not real else C.int_array([2])
        nout = 1 if real else 2
        ws = C.lib().pb3d_bin_reduce_work_size(C.ctypes.byref(b.cgrid), 3, nout, nb)
        outs = []
        for use_list in (0, 1):
            work = torch.empty(ws // 8 + 1, dtype=torch.float64, device='cuda')
            out = torch.empty((3, nout, nb), dtype=torch.float64, device='cuda')
            args = [C.ctypes.byref(b.cgrid), xx.data_ptr(), b.H, C._ptr(yy), 0, b.H, 3, nout, ells, None, real,
                    table.data_ptr(), nb]
            if use_list:
                C.check(C.lib().pb3d_bin_reduce_modes(*args, modes.data_ptr(), int(modes.numel()), out.data_ptr(),
                                                      work.data_ptr(), b.stream()))
            else:
                C.check(C.lib().pb3d_bin_reduce(*args, out.data_ptr(), work.data_ptr(), b.stream()))
            outs.append(out.cpu().numpy())
        np.testing.assert_allclose(outs[1], outs[0], rtol=1e-6, atol=1e-6 * np.abs(outs[0]).max())


@pytest.mark.parametrize("window", ["cic", "tsc", "pcs", "interlaced-cic", "interlaced-tsc", "interlaced-pcs"])
def test_pk_pixel_windows(hip, window):
    """Pixel-window deconvolution (base.py:103-108, 231-254) through the power-spectrum path: Pk_ideal
    (global), the S^-1 numerator and one masked MC Fisher realization (local; the GRF carries the
    window and A^-1 divides it out twice) against the oracle."""
    import polybin3d_amd as pb
    n = 32
    edges = np.linspace(0.01, 0.09, 9)
    d = powerlaw_field((n, n, n), seed=4)
    b, o = pb.PolyBin3D(L, [n] * 3, pixel_window=window), O.PolyBin3D(L, [n] * 3, pixel_window=window)
    got, ref = pb.PSpec(b, edges, lmax=4).Pk_ideal(d), O.PSpec(o, edges, lmax=4).Pk_ideal(d)
    for k in ref:
        close(got[k], ref[k], scale=own_scale(ref[k]), what=k)
    kw = dict(boxcenter=[10., 10., 10.], sightline='local')
    Pk = [np.linspace(0, 2, 500), 1e4 / (1 + (np.linspace(0, 2, 500) / 0.05) ** 2)]
    bl = pb.PolyBin3D(L, [n] * 3, Pk=Pk, pixel_window=window, **kw)
    ol = O.PolyBin3D(L, [n] * 3, Pk=Pk, pixel_window=window, **kw)
    mask = shell_mask(n, L, [10., 10., 10.])
    ps, po = pb.PSpec(bl, edges, lmax=2, mask=mask), O.PSpec(ol, edges, lmax=2, mask=mask)
    n1, n2 = ps.Pk_numerator(d * mask), po.Pk_numerator(d * mask)
    close(n1, n2, scale=np.abs(n2[:ps.Nk]).max())
    F1, F2 = ps.compute_fisher_contribution(2), po.compute_fisher_contribution(2)
    close(F1, F2, scale=_diag_scale(F2))


def test_pk_unwindowed_shot_subtraction(hip):
    """Pk_unwindowed with a shot-noise numerator and subtract_shotnoise (pspec.py:571-611), masked,
    local sightline, against the oracle; without a Fisher matrix it raises as the reference does."""
    import polybin3d_amd as pb
    n = 32
    edges = np.linspace(0.01, 0.09, 9)
    kw = dict(boxcenter=[10., 10., 10.], sightline='local')
    Pk = [np.linspace(0, 2, 500), 1e4 / (1 + (np.linspace(0, 2, 500) / 0.05) ** 2)]
    b, o = pb.PolyBin3D(L, [n] * 3, Pk=Pk, **kw), O.PolyBin3D(L, [n] * 3, Pk=Pk, **kw)
    mask = shell_mask(n, L, [10., 10., 10.])
    ps = pb.PSpec(b, edges, lmax=2, mask=mask, mask_shot=mask)
    po = O.PSpec(o, edges, lmax=2, mask=mask, mask_shot=mask)
    d = powerlaw_field((n, n, n), seed=6) * mask
    with pytest.raises(Exception):
        ps.Pk_unwindowed(d)
    fish = po.compute_fisher(2)
    shot = po.compute_shot_contribution(7)
    got = ps.Pk_unwindowed(d, fish=fish, shot_num=shot, subtract_shotnoise=True)
    ref = po.Pk_unwindowed(d, fish=fish, shot_num=shot, subtract_shotnoise=True)
    n1, n2 = ps.Pk_numerator(d), po.Pk_numerator(d)
    parity.check_solve("Pk_unwindowed with shot subtraction", got, ref, ['p0', 'p2'], fish, fish, n2 - shot,
                       n1 - shot)
    # the shot term is stored and reused by a later call that only asks for the subtraction
    again = ps.Pk_unwindowed(d, subtract_shotnoise=False)
    plain = po.Pk_unwindowed(d, fish=fish)
    parity.check_solve("Pk_unwindowed without subtraction", again, plain, ['p0', 'p2'], fish, fish, n2, n1)


@pytest.mark.parametrize("case", ["global_const", "local_mask"])
def test_pk_shot_contribution_user_sinv(hip, case):
    """PSpec.compute_shot_contribution with a user S^-1 (pspec.py:514-522): the callable filters
    A^-1 a times the shot-noise mask (or the mask mean) and a itself; per seed against the oracle."""
    import polybin3d_amd as pb
    import torch
    n = 32
    edges = np.linspace(0.01, 0.09, 9)
    kw = {} if case == "global_const" else dict(boxcenter=[10., 10., 10.], sightline='local')
    mask = 0.7 * np.ones((n, n, n)) if case == "global_const" else shell_mask(n, L, [10., 10., 10.])
    Pk = [np.linspace(0, 2, 500), 1e4 / (1 + (np.linspace(0, 2, 500) / 0.05) ** 2)]
    b, o = pb.PolyBin3D(L, [n] * 3, Pk=Pk, **kw), O.PolyBin3D(L, [n] * 3, Pk=Pk, **kw)
    wk = 1. / (1. + (o.modk_grid / 0.03) ** 2)
    wk_d = torch.from_numpy(wk.astype(np.float32)).cuda()

    def sinv_o(m, input_type='real', output_type='fourier'):
        f = (o.to_fourier(m) if input_type == 'real' else m) * wk
        return f if output_type == 'fourier' else o.to_real(f)

    def sinv_g(m, input_type='real', output_type='fourier'):
        f = (b.to_fourier(m) if input_type == 'real' else m) * wk_d
        return f if output_type == 'fourier' else b.to_real(f)
    ps = pb.PSpec(b, edges, lmax=2, mask=mask, mask_shot=0.5 * mask, applySinv=sinv_g)
    po = O.PSpec(o, edges, lmax=2, mask=mask, mask_shot=0.5 * mask, applySinv=sinv_o)
    for seed in (0, 3):
        s1, s2 = ps.compute_shot_contribution(seed), po.compute_shot_contribution(seed)
        for li in range(ps.Nl):
            sl = slice(li * ps.Nk, (li + 1) * ps.Nk)
            close(s1[sl], s2[sl], scale=own_scale(s2[sl]), what="l=%d" % (2 * li))
