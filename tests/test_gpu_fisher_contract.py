"""The split-BF16 MFMA Fisher contraction (pb3d_fisher_contract, the integrate_row sums of
PolyBin3D/cython/utils.pyx:162-179 as segment GEMMs) against an FP64 numpy GEMM of the same float32
operands, through the C ABI: ragged segments (rows and reals not multiples of the 128 x 128 x 32
tile), permuted row and column maps, per-index scales, accumulation into a non-zero output and
bit-reproducibility.  Tolerance relative to alpha * scale_i * scale_j * sum_t |l_t q_t| (the scale a
rounding error is proportional to): 6-product split 1e-6, 3-product split 1e-4."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

TOL = {1: 1e-4, 2: 1e-6}


def run(C, left, Q, segs, row_map, col_map, scale, alpha, prec, out):
    arr = (C.FisherSeg * len(segs))(*segs)
    nt = C.lib().pb3d_fisher_tiles(arr, len(segs), Q.shape[0], None, 0)
    assert nt > 0
    tiles = np.zeros(3 * nt, dtype=np.int32)
    assert C.lib().pb3d_fisher_tiles(arr, len(segs), Q.shape[0], tiles.ctypes.data, nt) == nt
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    segs_d = dev(np.frombuffer(bytes(arr), dtype=np.uint8).copy())
    keep = [dev(left), dev(Q), dev(tiles), dev(row_map), dev(col_map), dev(scale)]
    C.check(C.lib().pb3d_fisher_contract(keep[0].data_ptr(), keep[1].data_ptr(), Q.shape[1], Q.shape[0],
                                         segs_d.data_ptr(), len(segs), keep[2].data_ptr(), nt, keep[3].data_ptr(),
                                         keep[4].data_ptr(), keep[5].data_ptr(), alpha, prec, out.data_ptr(),
                                         out.shape[1], None))
    torch.cuda.synchronize()


@pytest.mark.parametrize("prec", [1, 2])
@pytest.mark.parametrize("shape", [((5, 2), (130, 62), (257, 1034)), ((1, 32),), ((129, 4100), (40, 78002))],
                         ids=["ragged3", "tiny", "long"])
def test_fisher_contract_matches_fp64(hip, prec, shape):
    C = hip
    rng = np.random.RandomState(3)
    nq = 300 if len(shape) > 1 else 7
    kq = sum(k for _, k in shape) + 6
    Q = (rng.randn(nq, kq) * rng.uniform(0.3, 3., (nq, 1))).astype(np.float32)
    segs, blocks, rows_all = [], [], []
    a_off, q_off = 0, 4
    for rows, k in shape:
        L = (rng.randn(rows, k) * rng.uniform(0.3, 3., (rows, 1))).astype(np.float32)
        s = C.FisherSeg()
        s.a_off, s.q_off, s.k, s.rows, s.row0 = a_off, q_off, k, rows, len(rows_all)
        segs.append(s)
        blocks.append((L, q_off, k))
        rows_all.extend(range(len(rows_all), len(rows_all) + rows))
        a_off += rows * k
        q_off += k
    left = np.concatenate([b[0].reshape(-1) for b in blocks])
    nrows = len(rows_all)
    row_map = rng.permutation(nrows).astype(np.int32)
    col_map = rng.permutation(nq).astype(np.int32)
    nout = max(nrows, nq)
    scale = rng.uniform(0.5, 2., nout)
    alpha = 0.37
    base = rng.randn(nout, nout)
    out = torch.from_numpy(base.copy()).cuda()
    run(C, left, Q, segs, row_map, col_map, scale, alpha, prec, out)
    ref, mag = base.copy(), np.zeros_like(base)
    r0 = 0
    for L, qo, k in blocks:
        L64, Q64 = L.astype(np.float64), Q[:, qo:qo + k].astype(np.float64)
        i = row_map[r0:r0 + len(L)]
        w = alpha * np.outer(scale[i], scale[col_map])
        ref[np.ix_(i, col_map)] += w * (L64 @ Q64.T)
        mag[np.ix_(i, col_map)] += w * (np.abs(L64) @ np.abs(Q64).T)
        r0 += len(L)
    got = out.cpu().numpy()
    touched = mag > 0
    assert np.array_equal(got[~touched], base[~touched])
    err = np.max(np.abs(got - ref)[touched] / mag[touched])
    print("[parity] fisher_contract prec %d %s: max error %.2e of sum |terms|" % (prec, "x".join(
        "%dx%d" % s for s in shape), err))
    assert err < TOL[prec]
    again = torch.from_numpy(base.copy()).cuda()
    run(C, left, Q, segs, row_map, col_map, scale, alpha, prec, again)
    assert torch.equal(out, again)


def test_fisher_contract_rejects_bad_arguments(hip):
    C = hip
    s = C.FisherSeg()
    s.k, s.rows = 3, 2
    arr = (C.FisherSeg * 1)(s)
    assert C.lib().pb3d_fisher_tiles(arr, 1, 10, None, 0) == -1
    assert C.lib().pb3d_fisher_contract(None, None, 2, 1, None, 1, None, 1, None, None, None, 1., 2, None, 1,
                                        None) == -1
