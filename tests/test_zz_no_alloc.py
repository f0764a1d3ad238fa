"""pb3d_bin_reduce_modes_noalloc (the allocation-free variant, fixed 16 MiB device scratch) through the
C ABI against pb3d_bin_reduce_modes with a caller workspace on the same inputs: bit-identical sums
(both run the same fixed-order kernels) on two calls in a row, the second smaller, and an oversized
request rejected before any launch."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _call(C, b, fn, x, y, nq, nout, ells, edges, real, work=None):
    table, modes = b.bin_table(edges), b.bin_modes(edges)
    nb = len(edges) - 1
    out = torch.full((nq, nout, nb), np.nan, dtype=torch.float64, device='cuda')
    args = [C.ctypes.byref(b.cgrid), x.data_ptr(), b.H, C._ptr(y), 0, b.H, nq, nout, C.int_array(ells), None,
            1 if real else 0, table.data_ptr(), nb, modes.data_ptr(), int(modes.numel()), out.data_ptr()]
    if work is not None:
        args.append(work.data_ptr())
    rc = fn(*args, b.stream())
    torch.cuda.synchronize()
    return rc, out


def test_noalloc_bin_reduce_matches_workspace_variant(hip):
    import polybin3d_amd as pb
    C = hip
    b = pb.PolyBin3D(1000., [64] * 3, boxcenter=[300., 200., 800.], sightline='local')
    rng = np.random.RandomState(7)
    x = torch.from_numpy((rng.randn(6, *b.half_shape) + 1j * rng.randn(6, *b.half_shape)).astype(np.complex64)).cuda()
    y = torch.from_numpy((rng.randn(3, *b.half_shape) + 1j * rng.randn(3, *b.half_shape)).astype(np.complex64)).cuda()
    cases = [(6, 3, [0, 2, 4], np.linspace(0.01, 0.19, 25), False, y),   # first call
             (2, 1, [2], np.linspace(0.02, 0.11, 7), False, None)]        # second call, smaller
    for nq, nout, ells, edges, real, yy in cases:
        ws = C.lib().pb3d_bin_reduce_work_size(C.ctypes.byref(b.cgrid), nq, nout, len(edges) - 1)
        work = torch.empty(ws // 8 + 1, dtype=torch.float64, device='cuda')
        rc1, ref = _call(C, b, C.lib().pb3d_bin_reduce_modes, x, yy, nq, nout, ells, edges, real, work)
        rc2, got = _call(C, b, C.lib().pb3d_bin_reduce_modes_noalloc, x, yy, nq, nout, ells, edges, real)
        assert rc1 == 0 and rc2 == 0, C.lib().pb3d_last_error()
        assert torch.isfinite(ref).all() and torch.equal(got, ref)
        print("[noalloc] nq %d nout %d bins %d: bit-identical to the workspace variant" % (nq, nout, len(edges) - 1))
    # nq * nout * nbins * blocks above the fixed buffer: rejected, nothing launched
    edges = np.linspace(0.01, 0.19, 255)
    rc, out = _call(C, b, C.lib().pb3d_bin_reduce_modes_noalloc, x, None, 1, 16, [0] * 16, edges, False)
    assert rc == -1 and b"fixed buffer" in C.lib().pb3d_last_error()
    assert torch.isnan(out).all()
