"""Gradients of the differentiable Pk_ideal / Bk_ideal (polybin3d_amd.autograd) against central
finite differences.

* 32^3, global sightline: the GPU adjoint against central differences of the FP64 oracle's own
  Pk_ideal / Bk_ideal on the same float32 field, on 16 random voxels, eps = 1e-3 sigma.  The chi^2
  loss is a polynomial of degree 4 (Pk) or 6 (Bk) in the field, so the difference quotient is exact
  up to O(eps^2) ~ 1e-6 relative; tolerance 1e-4 of the largest checked gradient.
* chi^2 weights: sigma of P_l(k) = 5% of |P_0(k)|, sigma of B = 5% of |B| + 1e-3 max|B|.
* 512^3 (chi^2 over P_0,2,4 in 30 bins + B_0 in 40 bins, quadratic non-Gaussian field): the
  single-voxel change of the loss (~1e-11 relative) is below float32 resolution, so the check is the
  directional derivative along v = field * (1 + sin(k.x)/2), Richardson-extrapolated central
  differences of the GPU loss at h = 5e-3 and 1e-2 (error O(h^4)) against <grad, v>; tolerance 1e-4
  relative.
"""
import numpy as np
import pytest
import torch

from oracle import polybin_oracle as O

pytestmark = pytest.mark.gpu

L = 1000.


def powerlaw():
    kk = np.linspace(0, 3, 6000)
    P = np.zeros_like(kk)
    P[1:] = 2e4 * (kk[1:] / 0.1) ** -1.5
    return [kk, P]


def ng_field(o, seed):
    d = o.generate_data(seed=seed, output_type='real')
    return d + 0.5 * (d ** 2 - np.mean(d ** 2))


def chi2_np(v, t, w):
    return float(np.sum((v - t) ** 2 * w))


def fd_oracle(loss, d64, vox, eps):
    out = []
    for v in vox:
        dd = d64.copy()
        dd.flat[v] += eps
        lp = loss(dd)
        dd.flat[v] -= 2 * eps
        lm = loss(dd)
        out.append((lp - lm) / (2 * eps))
    return np.asarray(out)


def test_pk_ideal_gradient_matches_oracle_fd(hip):
    import polybin3d_amd as pb
    from polybin3d_amd import autograd as AG
    n, edges, lmax = 32, np.linspace(0.01, 0.09, 9), 4
    o = O.PolyBin3D(L, [n] * 3, Pk=powerlaw())
    d32 = ng_field(o, 11).astype(np.float32)
    d64 = d32.astype(np.float64)
    po = O.PSpec(o, edges, lmax=lmax)
    ps = pb.PSpec(pb.PolyBin3D(L, [n] * 3, Pk=powerlaw()), edges, lmax=lmax)
    x = torch.tensor(d32, device='cuda', requires_grad=True)
    Pk = AG.pk_ideal(ps, x)
    P0 = Pk.detach().cpu().numpy()
    ref0 = np.concatenate([v for v in po.Pk_ideal(d64).values()])
    assert np.max(np.abs(P0 - ref0) / np.abs(ref0)) < 1e-5
    rng = np.random.RandomState(3)
    t = P0 * (1. + 0.2 * rng.randn(len(P0)))
    Nk = len(edges) - 1
    w = 1. / (0.05 * np.abs(np.tile(P0[:Nk], len(P0) // Nk))) ** 2  # sigma of P_l(k): 5% of |P_0(k)|
    loss = ((Pk - torch.from_numpy(t).cuda()) ** 2 * torch.from_numpy(w).cuda()).sum()
    loss.backward()
    grad = x.grad.detach().cpu().numpy().reshape(-1)
    vox = rng.choice(n ** 3, 16, replace=False)
    fd = fd_oracle(lambda dd: chi2_np(np.concatenate([v for v in po.Pk_ideal(dd).values()]), t, w),
                   d64, vox, 1e-3 * d64.std())
    err = np.max(np.abs(grad[vox] - fd)) / np.max(np.abs(fd))
    assert err < 1e-4, "max gradient error %.3e of max |grad|" % err


@pytest.mark.parametrize("precision,tol", [("bf16x6", 1e-4), ("bf16x3", 1e-3)])
def test_bk_ideal_gradient_matches_oracle_fd(hip, precision, tol):
    import polybin3d_amd as pb
    from polybin3d_amd import autograd as AG
    n, edges, lmax = 32, np.linspace(0.01, 0.09, 6), 2
    o = O.PolyBin3D(L, [n] * 3, Pk=powerlaw())
    d32 = ng_field(o, 12).astype(np.float32)
    d64 = d32.astype(np.float64)
    bo = O.BSpec(o, edges, lmax=lmax)
    bs = pb.BSpec(pb.PolyBin3D(L, [n] * 3, Pk=powerlaw()), edges, lmax=lmax, precision=precision)
    x = torch.tensor(d32, device='cuda', requires_grad=True)
    Bk = AG.bk_ideal(bs, x)
    B0 = Bk.detach().cpu().numpy()
    rng = np.random.RandomState(4)
    t = B0 * (1. + 0.2 * rng.randn(len(B0)))
    w = 1. / (0.05 * np.abs(B0) + 1e-3 * np.max(np.abs(B0))) ** 2
    loss = ((Bk - torch.from_numpy(t).cuda()) ** 2 * torch.from_numpy(w).cuda()).sum()
    loss.backward()
    grad = x.grad.detach().cpu().numpy().reshape(-1)
    vox = rng.choice(n ** 3, 16, replace=False)
    fd = fd_oracle(lambda dd: chi2_np(np.concatenate([v for v in bo.Bk_ideal(dd).values()]), t, w),
                   d64, vox, 1e-3 * d64.std())
    err = np.max(np.abs(grad[vox] - fd)) / np.max(np.abs(fd))
    assert err < tol, "max gradient error %.3e of max |grad|" % err


def test_chi2_gradient_512_directional(hip):
    """The differentiable-estimator workload at 512^3: chi^2 over P_0,2,4 (30 bins to 0.3) + B_0
    (40 bins to 0.4) of a float32 quadratic non-Gaussian field, global sightline."""
    import polybin3d_amd as pb
    from polybin3d_amd import autograd as AG
    n = 512
    b = pb.PolyBin3D(L, [n] * 3, Pk=powerlaw(), rng='device')
    ps = pb.PSpec(b, np.linspace(0.01, 0.30, 31), lmax=4)
    bs = pb.BSpec(b, np.linspace(0.01, 0.40, 41), lmax=0, precision='bf16x6')
    dG = b.generate_data(seed=2, output_type='real')
    field = (dG + 0.5 * (dG ** 2 - (dG ** 2).mean())).contiguous()
    P0 = AG.pk_ideal(ps, field).detach()
    B0 = AG.bk_ideal(bs, field).detach()
    gen = torch.Generator(device='cuda').manual_seed(5)
    tP = P0 * (1. + 0.2 * torch.randn(P0.shape, generator=gen, device='cuda', dtype=torch.float64))
    tB = B0 * (1. + 0.2 * torch.randn(B0.shape, generator=gen, device='cuda', dtype=torch.float64))
    # sigma of P_l(k): 5% of |P_0(k)| (near-zero higher multipoles would otherwise get weights that
    # turn the loss into a heavily cancelling sum, where float32 rounding dominates)
    wP = 1. / (0.05 * P0[:30].abs().repeat(3)) ** 2
    wB = 1. / (0.05 * B0.abs() + 1e-3 * B0.abs().max()) ** 2

    chi2P = lambda f: ((AG.pk_ideal(ps, f) - tP) ** 2 * wP).sum()
    chi2B = lambda f: ((AG.bk_ideal(bs, f) - tB) ** 2 * wB).sum()
    # direction: the field under a smooth modulation, so that the derivative is O(chi^2) (a random
    # direction is nearly orthogonal to a quadratic/cubic estimator's gradient and its difference
    # quotient drowns in float32 rounding of chi^2 itself)
    i = torch.arange(n, device='cuda', dtype=torch.float32)
    ph = (i[:, None, None] + 2 * i[None, :, None] + 3 * i[None, None, :]) * (2 * np.pi / n) + 0.3
    v = field * (1. + 0.5 * torch.sin(ph))
    del ph
    msgs = []
    for name, chi2 in (("P_l", chi2P), ("B_0", chi2B), ("P_l + B_0", lambda f: chi2P(f) + chi2B(f))):
        x = field.clone().requires_grad_(True)
        chi2(x).backward()
        dirder = float((x.grad.double() * v.double()).sum())
        with torch.no_grad():
            def fd(h):
                return (float(chi2(field + h * v)) - float(chi2(field - h * v))) / (2 * h)
            rich = (4. * fd(5e-3) - fd(1e-2)) / 3.
        err = abs(dirder - rich) / abs(rich)
        msgs.append("%s: <grad, v> %.9e vs finite differences %.9e (rel %.2e)" % (name, dirder, rich, err))
        assert err < 1e-4, "; ".join(msgs)


def test_adjoints_equal_the_reference_q_maps(hip):
    """The hand-written adjoints against the Q-derivative maps the reference forms for its MC Fisher
    (pspec.py:444-446: Q = (S^-1 P)^dag Theta_b L_l S^-1 P a; bspec.py:615-640: the three-leg
    placement of add_Q_element), built here with the oracle's filters for the ideal weighting
    (S^-1 = invPk0, P = 1, a = the data):
      d n_(l,b) / d d = (V/N) C2R(Q_(l,b)),   d num_i / d d = V/(N sym_i) C2R(invPk0 Q_i).
    The adjoint of grad . estimator with grad = F^T e_j is d n_j / d d (F the ideal Fisher).  Every
    power-spectrum bin and a set of bispectrum triangles of both multipoles, 32^3, global sightline;
    tolerance 1e-5 of the largest entry of the map."""
    import polybin3d_amd as pb
    n = 32
    o = O.PolyBin3D(L, [n] * 3, Pk=powerlaw())
    d = ng_field(o, 5).astype(np.float32)
    d64 = d.astype(np.float64)
    b = pb.PolyBin3D(L, [n] * 3, Pk=powerlaw())
    V, N = o.volume, float(n ** 3)
    D = o.to_fourier(d64)
    # power spectrum
    edges = np.linspace(0.01, 0.09, 9)
    ps = pb.PSpec(b, edges, lmax=2)
    ps.Pk_ideal(d)
    F = ps.fish_ideal
    w = o.invPk0_grid
    Nk = len(edges) - 1
    worst = 0.
    for j in range(2 * Nk):
        ell, kb = 2 * (j // Nk), j % Nk
        Q = w * o.map_utils.fourier_filter(w * D, ell, edges[kb], edges[kb + 1])
        ref = V / N * o.to_real(Q)
        got = ps._pk_ideal_adjoint(d, torch.from_numpy(F[j, :].copy()).cuda()).cpu().numpy()
        worst = max(worst, np.max(np.abs(got - ref)) / np.max(np.abs(ref)))
    print("[adjoint] Pk_ideal vs (V/N) C2R(Q): max error %.2e over %d bins" % (worst, 2 * Nk))
    assert worst < 1e-5
    # bispectrum
    e3 = np.linspace(0.01, 0.07, 5)
    bs = pb.BSpec(b, e3, lmax=2)
    bs.Bk_ideal(d)
    bo = O.BSpec(o, e3, lmax=2)
    bo._compute_symmetry_factor()
    g = bo._g_maps(bo.invPk0 * D)
    Fb = bs.fish_ideal
    sel = [i for i in range(bs.N_bins) if i % 3 == 0] + [i for i, t in enumerate(bs.all_bins) if t[0] == t[2]]
    worst = 0.
    for j in sorted(set(sel)):
        b1, b2, b3, ell = bs.all_bins[j]
        Q = bo._filt(o.to_fourier(g[0, b1] * g[0, b2]), ell, b3)
        Q = Q + bo._filt(o.to_fourier(g[0, b2] * g[ell // 2, b3]), 0, b1)
        Q = Q + bo._filt(o.to_fourier(g[0, b1] * g[ell // 2, b3]), 0, b2)
        ref = V / N / bo.sym_factor[j] * o.to_real(bo.invPk0 * Q)
        got = bs._bk_ideal_adjoint(d, torch.from_numpy(Fb[j, :].copy()).cuda()).cpu().numpy()
        worst = max(worst, np.max(np.abs(got - ref)) / np.max(np.abs(ref)))
    print("[adjoint] Bk_ideal vs V/(N sym) C2R(invPk0 Q): max error %.2e over %d triangles" % (worst, len(set(sel))))
    assert worst < 1e-5
