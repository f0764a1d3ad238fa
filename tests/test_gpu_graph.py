"""HIP-graph capture of MC Fisher passes (FisherGraph): every replay, with new seeds and with the
program's own eager work and synchronisations between replays, adds exactly what the eager path adds
for the same seeds (power spectrum with batched seeds, bispectrum), and the seed buffer drives the
device GRF like the eager seed argument."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

L = 1000.


def shell(n, c):
    x = [np.fft.fftshift(np.arange(-n // 2, n // 2)) * L / n + c[i] for i in range(3)]
    X = np.meshgrid(*x, indexing='ij')
    r = np.sqrt(X[0] ** 2 + X[1] ** 2 + X[2] ** 2)
    return 0.2 + ((r >= 150.) & (r <= 480.)).astype(np.float64)


def replay_matches_eager(spec, R, seeds):
    import polybin3d_amd as pb
    N = spec.N_bins
    out = torch.zeros((N, N), dtype=torch.float64, device='cuda')
    g = pb.FisherGraph(spec, out, seeds_per_replay=R)
    errs = []
    for s0 in seeds:
        before = out.clone()
        g.replay(s0)
        torch.cuda.synchronize()
        inc = (out - before).cpu().numpy()
        # eager work between replays: the same seeds through the uncaptured path
        if hasattr(spec, 'all_bins'):
            ref = sum(spec._fisher_contribution_dev(s).cpu().numpy() for s in range(s0, s0 + R))
        elif R > 1:
            ref = spec._fisher_contributions_dev(list(range(s0, s0 + R))).cpu().numpy()
        else:
            ref = spec._fisher_contribution_dev(s0).cpu().numpy()
        torch.cuda.synchronize()
        d = np.sqrt(np.abs(np.outer(np.diag(ref), np.diag(ref))))
        err = float(np.max(np.abs(inc - ref) / d))
        errs.append(err)
        print("[graph] %s replay seeds %d..%d: max |graph - eager| / sqrt(F_ii F_jj) = %.2e"
              % (type(spec).__name__, s0, s0 + R - 1, err))
        assert err < 1e-12
    return errs


def test_pk_fisher_graph_replays(hip):
    import polybin3d_amd as pb
    n = 64
    c = [10., 10., 10.]
    b = pb.PolyBin3D(L, [n] * 3, boxcenter=c, sightline='local', rng='device')
    ps = pb.PSpec(b, np.linspace(0.01, 0.15, 8), lmax=2, mask=shell(n, c))
    replay_matches_eager(ps, 2, [0, 7, 3])
    ps1 = pb.PSpec(b, np.linspace(0.01, 0.15, 8), lmax=2, mask=shell(n, c))
    replay_matches_eager(ps1, 1, [4, 11])


def test_bk_fisher_graph_replays(hip):
    import polybin3d_amd as pb
    n = 32
    c = [10., 10., 10.]
    b = pb.PolyBin3D(L, [n] * 3, boxcenter=c, sightline='local', rng='device')
    bs = pb.BSpec(b, np.linspace(0.01, 0.07, 5), lmax=2, mask=shell(n, c))
    replay_matches_eager(bs, 1, [0, 5, 2])


def test_graph_rejects_host_paths(hip):
    import polybin3d_amd as pb
    b = pb.PolyBin3D(L, [32] * 3, rng='numpy')
    ps = pb.PSpec(b, np.linspace(0.01, 0.09, 5), lmax=0)
    with pytest.raises(ValueError):
        pb.FisherGraph(ps, torch.zeros((ps.N_bins,) * 2, dtype=torch.float64, device='cuda'))
