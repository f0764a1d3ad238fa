"""Host-side logic that needs no GPU: the masked-row flags handed to pb3d_fft_pointing, the number of
chain streams that fit in device memory, and the full-grid helpers of real_fft=False (Hermitian
completion and the Hermitian / anti-Hermitian split), on CPU tensors against numpy."""
import types

import numpy as np
import pytest
import torch

from polybin3d_amd.base import PolyBin3D
from polybin3d_amd.bspec import BSpec


def fake_base(n):
    """The attributes the host helpers read, without a plan or a device."""
    b = types.SimpleNamespace()
    b.gridsize = np.asarray(n)
    b.full_shape = tuple(n)
    b.half_shape = (n[0], n[1], n[2] // 2 + 1)
    b._bufs = {}
    b.sightline = 'local'
    b.device = torch.device('cpu')
    return b


def test_mask_rows_groups_of_64():
    n = (8, 24, 16)  # 192 rows: three groups of 64
    b = fake_base(n)
    mask = torch.zeros(n)
    mask[0, 5, 3] = 1.  # row 5: group 0 on
    w = torch.zeros(n)
    w[7, 23, 0] = 2.  # row 191: group 2 on through the IC weight only
    flags = PolyBin3D.mask_rows(b, mask, w)
    assert flags.dtype == torch.uint8 and flags.numel() == 192
    assert flags[:64].tolist() == [1] * 64 and flags[64:128].tolist() == [0] * 64 and flags[128:].tolist() == [1] * 64
    assert PolyBin3D.mask_rows(b, mask, w) is flags  # cached for the same tensors and versions
    mask[3, 0, 0] = 1.  # in place: a new version, flags rebuilt (row 72: group 1 on)
    assert PolyBin3D.mask_rows(b, mask, w) is None  # every group on: no flags
    # a partial last group (rows not a multiple of 64)
    b2 = fake_base((5, 20, 8))  # 100 rows
    m2 = torch.zeros((5, 20, 8))
    m2[4, 19, 7] = 1.  # row 99, group 1 (rows 64..99)
    f2 = PolyBin3D.mask_rows(b2, m2)
    assert f2.numel() == 100 and f2[:64].sum().item() == 0 and f2[64:].sum().item() == 36


@pytest.mark.parametrize("free_gib,held,expect", [(200, 0, 4), (20, 0, 3), (10, 0, 2), (3, 0, 1), (3, 3, 4)])
def test_chain_streams_that_fit(monkeypatch, free_gib, held, expect):
    n = (256, 256, 256)
    b = fake_base(n)
    b.kF = np.full(3, 2 * np.pi / 1000.)
    b.support = types.MethodType(PolyBin3D.support, b)
    # bin edges to 0.4: input and output band limits of 63 frequency indices (64 kz planes)
    spec = types.SimpleNamespace(base=b, _sup=[63, 63, 63], k_bins_squeeze=np.array([0.01, 0.4]))
    # two chains: l = 0 (one sub-map, 64 maps) and l = 2 (six sub-maps, 32 maps), local sightline;
    # the l = 2 chain's work maps are the larger: 32 maps x (6 sub-maps + the output) x 256^2 x 64
    ab = np.zeros((96, 4), dtype=np.int64)
    ab[64:, 3] = 2
    chunks = [(0, 64), (64, 96)]
    need = (6 + 1) * 8 * 256 * 256 * 64 * 32
    for lane in range(1, held + 1):
        b._bufs[('fft_point', lane)] = torch.empty(need, dtype=torch.uint8, device='meta')
    monkeypatch.setattr(torch.cuda, "mem_get_info", lambda dev=None: (free_gib << 30, 288 << 30))
    monkeypatch.setattr(torch.cuda, "memory_reserved", lambda dev=None: 0)
    monkeypatch.setattr(torch.cuda, "memory_allocated", lambda dev=None: 0)
    assert BSpec._streams_that_fit(spec, ab, chunks, 4) == expect


def test_full_grid_completion_and_split():
    n = (6, 8, 10)
    b = fake_base(n)
    rng = np.random.RandomState(0)
    x = rng.randn(*n)
    full = np.fft.fftn(x)
    half = torch.from_numpy(full[..., :n[2] // 2 + 1].astype(np.complex64))
    got = PolyBin3D.to_full(b, half).numpy()
    assert np.allclose(got, full, rtol=0, atol=1e-4 * np.abs(full).max())
    # anti-Hermitian completion of i * (a Hermitian map)
    got = PolyBin3D.to_full(b, 1j * half, parity=-1).numpy()
    assert np.allclose(got, 1j * full, rtol=0, atol=1e-4 * np.abs(full).max())
    z = rng.randn(*n) + 1j * rng.randn(*n)
    zneg = np.conj(np.roll(np.flip(z, axis=(0, 1, 2)), 1, axis=(0, 1, 2)))
    h, a = PolyBin3D.full_parts(b, torch.from_numpy(z))
    assert np.allclose(h.numpy(), (0.5 * (z + zneg))[..., :n[2] // 2 + 1])
    assert np.allclose(a.numpy(), (0.5 * (z - zneg))[..., :n[2] // 2 + 1])
    # the parts rebuild the map
    back = PolyBin3D.to_full(b, h) + PolyBin3D.to_full(b, a, parity=-1)
    assert np.allclose(back.numpy(), z)
