"""Bitwise reproducibility: the reductions are formed in a fixed order (bin sums: run totals added in
lane order per one-wave block, block partials summed in block order; integral-constraint sums: per-wave
butterfly sums stored per wave and added in wave order; Fisher contraction: one writer per output), so
a realization computed twice gives identical bits."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

L = 1000.


def shell(n, c):
    x = [np.fft.fftshift(np.arange(-n // 2, n // 2)) * L / n + c[i] for i in range(3)]
    X = np.meshgrid(*x, indexing='ij')
    r = np.sqrt(X[0] ** 2 + X[1] ** 2 + X[2] ** 2)
    return 0.2 + ((r >= 150.) & (r <= 480.)).astype(np.float64)


@pytest.mark.parametrize("ic", ["none", "gic", "ric"])
def test_pk_fisher_realization_bitwise(hip, ic):
    import polybin3d_amd as pb
    n, c = 64, [10., 10., 10.]
    b = pb.PolyBin3D(L, [n] * 3, boxcenter=c, sightline='local')
    mask = shell(n, c)
    kw = {} if ic == "none" else (dict(add_GIC=True, mask_IC=mask * 1.1) if ic == "gic" else
                                   dict(add_RIC=True, mask_IC=mask * 1.1, radial_bins_RIC=[100., 250., 400., 600.]))
    ps = pb.PSpec(b, np.linspace(0.01, 0.15, 11), lmax=4, mask=mask, **kw)
    F1 = ps.compute_fisher_contribution(3)
    F2 = ps.compute_fisher_contribution(3)
    assert np.abs(F1).max() > 0
    assert np.array_equal(F1, F2)


@pytest.mark.parametrize("ic", ["none", "gic"])
def test_bk_fisher_realization_bitwise(hip, ic):
    import polybin3d_amd as pb
    n, c = 32, [10., 10., 10.]
    b = pb.PolyBin3D(L, [n] * 3, boxcenter=c, sightline='local')
    mask = shell(n, c)
    kw = {} if ic == "none" else dict(add_GIC=True, mask_IC=mask * 1.1)
    bs = pb.BSpec(b, np.linspace(0.01, 0.07, 5), lmax=2, mask=mask, **kw)
    F1 = bs.compute_fisher_contribution(2)
    F2 = bs.compute_fisher_contribution(2)
    assert np.abs(F1).max() > 0
    assert np.array_equal(F1, F2)


@pytest.mark.parametrize("streams", [1, 4])
def test_bk_fisher_streams_and_deferred_contractions_bitwise(hip, streams):
    """The Q-map chains dealt to several streams, and the contractions of consecutive realizations
    deferred to a stream of their own (each next realization's Q maps waiting for the previous
    contraction), give the bits of the serial one-stream sum."""
    import polybin3d_amd as pb
    import torch
    n, c = 32, [10., 10., 10.]
    b = pb.PolyBin3D(L, [n] * 3, boxcenter=c, sightline='local', rng='device')
    mask = shell(n, c)
    edges = np.linspace(0.01, 0.09, 9)  # 8 bins: several chains per multipole
    serial = pb.BSpec(b, edges, lmax=2, mask=mask, chain_streams=1)
    ref = torch.zeros((serial.N_bins,) * 2, dtype=torch.float64, device='cuda')
    for sd in (4, 9, 2):
        serial._fisher_contribution_dev(sd, out=ref)
    bs = pb.BSpec(b, edges, lmax=2, mask=mask, chain_streams=streams)
    assert bs.N_bins > 2 * 64
    got = torch.zeros_like(ref)
    bs._fisher_contributions_dev([4, 9, 2], got)
    torch.cuda.synchronize()
    assert ref.abs().max().item() > 0
    assert torch.equal(got, ref)


def test_bin_sums_bitwise(hip):
    import polybin3d_amd as pb
    import torch
    n = 64
    b = pb.PolyBin3D(L, [n] * 3)
    rng = np.random.RandomState(5)
    x = torch.from_numpy((rng.randn(4, *b.half_shape) + 1j * rng.randn(4, *b.half_shape)).astype(np.complex64)).cuda()
    edges = np.linspace(0.01, 0.19, 19)
    r1 = b.bin_reduce(x, edges, [0, 2, 4], nq=4).cpu().numpy()
    r2 = b.bin_reduce(x, edges, [0, 2, 4], nq=4).cpu().numpy()
    assert np.array_equal(r1, r2)


def test_bk_fisher_streams_fall_back_when_memory_is_short(hip, monkeypatch):
    """With no device memory beyond the 2 GiB kept free, the side streams whose work maps already
    hold the largest chain keep running; without them, and with room for one more stream's work
    maps, the Q-map chains go to two streams instead of four (streams_used).  The Fisher contribution
    keeps the bits of the one-stream path throughout."""
    import polybin3d_amd as pb
    import torch
    n, c = 32, [10., 10., 10.]
    b = pb.PolyBin3D(L, [n] * 3, boxcenter=c, sightline='local', rng='device')
    mask = shell(n, c)
    edges = np.linspace(0.01, 0.09, 9)
    ref = pb.BSpec(b, edges, lmax=2, mask=mask, chain_streams=1)._fisher_contribution_dev(4).clone()
    bs = pb.BSpec(b, edges, lmax=2, mask=mask, chain_streams=4)
    first = bs._fisher_contribution_dev(4).clone()
    assert bs.streams_used == 4 and torch.equal(first, ref)
    lanes = [('fft_point', lane) for lane in (1, 2, 3)]
    need = max(b._bufs[k].numel() for k in lanes + ['fft_point'])  # the largest chain's work maps
    held = sum(b._bufs[k].numel() >= need for k in lanes)
    real, extra = torch.cuda.mem_get_info, [0]

    def short(dev=None):
        free, total = real(dev)
        slack = torch.cuda.memory_reserved(dev) - torch.cuda.memory_allocated(dev)
        return (2 << 30) + extra[0] - slack, total
    monkeypatch.setattr(torch.cuda, "mem_get_info", short)
    got = bs._fisher_contribution_dev(4).clone()
    print("streams used with no spare memory: %d (side streams holding the largest chain: %d)"
          % (bs.streams_used, held))
    assert bs.streams_used == 1 + held and torch.equal(got, ref)
    for k in lanes:
        del b._bufs[k]
    extra[0] = need + need // 2
    got = bs._fisher_contribution_dev(4)
    torch.cuda.synchronize()
    print("streams used with memory for one more:", bs.streams_used)
    assert bs.streams_used == 2 and torch.equal(got, ref)
