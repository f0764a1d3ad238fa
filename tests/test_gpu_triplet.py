"""Split-BF16 MFMA triplet contraction (pb3d_pair_gemm) and its adjoint (pb3d_pair_adjoint) against
FP64 numpy on the same float32 inputs, and the Bk_ideal numerator at each precision against the
FP64 oracle.  Tolerances (relative to the sum of |terms| of each entry, the scale a rounding error
is proportional to): 6-term split 1e-6, 3-term split 1e-4."""
import ctypes

import numpy as np
import pytest
import torch

from oracle import polybin_oracle as O

pytestmark = pytest.mark.gpu

TOL = {1: 1e-4, 2: 1e-6}


def stack(nm, nvox, seed):
    rng = np.random.RandomState(seed)
    return (rng.randn(nm, nvox) * rng.uniform(0.5, 2., (nm, 1))).astype(np.float32)


@pytest.mark.parametrize("prec", [1, 2])
@pytest.mark.parametrize("nm,nvox,na,npair", [(6, 5000, 3, 7), (40, 40000, 40, 300), (130, 4099, 100, 50),
                                             (128, 70001, 128, 260), (90, 30000, 120, 140), (24, 9000, 30, 70),
                                             (16, 30001, 16, 136), (40, 50000, 20, 200), (96, 20000, 32, 256)])
def test_pair_gemm_matches_fp64(hip, prec, nm, nvox, na, npair):
    C = hip
    g = stack(nm, nvox, 1)
    rng = np.random.RandomState(2)
    amap = rng.randint(0, nm, na).astype(np.int32)
    pb, pc = rng.randint(0, nm, npair).astype(np.int32), rng.randint(0, nm, npair).astype(np.int32)
    gd = torch.from_numpy(g).cuda()
    dev = lambda a: torch.from_numpy(a).cuda()
    out = torch.empty((na, npair), dtype=torch.float64, device='cuda')
    ws = C.lib().pb3d_pair_gemm_work_size(nvox, na, npair)
    work = torch.empty(max(ws, 1), dtype=torch.uint8, device='cuda')
    ad, bd, cd = dev(amap), dev(pb), dev(pc)
    C.check(C.lib().pb3d_pair_gemm(gd.data_ptr(), nm, nvox, ad.data_ptr(), na, bd.data_ptr(), cd.data_ptr(), npair,
                                   prec, out.data_ptr(), work.data_ptr(), None))
    g64 = g.astype(np.float64)
    P = g64[pb] * g64[pc]
    ref = g64[amap] @ P.T
    scale = np.abs(g64[amap]) @ np.abs(P).T
    err = np.max(np.abs(out.cpu().numpy() - ref) / scale)
    assert err < TOL[prec], "max scaled error %.3e" % err


def test_pair_gemm_bad_index_is_nan(hip):
    C = hip
    g = torch.from_numpy(stack(4, 1000, 3)).cuda()
    amap = torch.tensor([0, 1], dtype=torch.int32, device='cuda')
    pb = torch.tensor([0, 7], dtype=torch.int32, device='cuda')
    pc = torch.tensor([1, 2], dtype=torch.int32, device='cuda')
    out = torch.empty((2, 2), dtype=torch.float64, device='cuda')
    work = torch.empty(C.lib().pb3d_pair_gemm_work_size(1000, 2, 2), dtype=torch.uint8, device='cuda')
    C.check(C.lib().pb3d_pair_gemm(g.data_ptr(), 4, 1000, amap.data_ptr(), 2, pb.data_ptr(), pc.data_ptr(), 2, 2,
                                   out.data_ptr(), work.data_ptr(), None))
    o = out.cpu().numpy()
    assert np.isfinite(o[:, 0]).all() and np.isnan(o[:, 1]).all()
    with pytest.raises(ValueError):
        C.check(C.lib().pb3d_pair_gemm(g.data_ptr(), 4, 1000, amap.data_ptr(), 2, pb.data_ptr(), pc.data_ptr(), 2, 0,
                                       out.data_ptr(), work.data_ptr(), None))


@pytest.mark.parametrize("prec", [1, 2])
@pytest.mark.parametrize("nm,nvox,na,npair", [(5, 3000, 5, 9), (40, 20000, 40, 820), (80, 1000, 80, 33)])
def test_pair_adjoint_matches_fp64(hip, prec, nm, nvox, na, npair):
    C = hip
    g = stack(nm, nvox, 4)
    rng = np.random.RandomState(5)
    pb, pc = rng.randint(0, nm, npair).astype(np.int32), rng.randint(0, nm, npair).astype(np.int32)
    U = rng.randn(na, npair).astype(np.float32)
    gd, Ud = torch.from_numpy(g).cuda(), torch.from_numpy(U).cuda()
    bd, cd = torch.from_numpy(pb).cuda(), torch.from_numpy(pc).cuda()
    out = torch.empty((na, nvox), dtype=torch.float32, device='cuda')
    ws = C.lib().pb3d_pair_adjoint_work_size(na, npair, prec)
    work = torch.empty(ws, dtype=torch.uint8, device='cuda')
    C.check(C.lib().pb3d_pair_adjoint(gd.data_ptr(), nm, nvox, bd.data_ptr(), cd.data_ptr(), npair, Ud.data_ptr(), na,
                                      prec, out.data_ptr(), work.data_ptr(), None))
    g64 = g.astype(np.float64)
    P = g64[pb] * g64[pc]
    ref = U.astype(np.float64) @ P
    scale = np.abs(U.astype(np.float64)) @ np.abs(P)
    err = np.max(np.abs(out.cpu().numpy() - ref) / scale)
    # FP32 output: the stored value itself carries 2^-24 relative rounding
    assert err < TOL[prec] + 1e-6, "max scaled error %.3e" % err


@pytest.mark.parametrize("precision,tol", [("bf16x6", 1e-5), ("bf16x3", 2e-4), ("fp32", 1e-5)])
def test_bk_numerator_precisions(hip, precision, tol):
    """Bk_ideal numerator (global, l=2) at each triplet precision vs the FP64 oracle; the scale is
    sum over triplets of |g g g| (the magnitude the rounding acts on)."""
    import polybin3d_amd as pb
    n, edges = 32, np.linspace(0.01, 0.09, 6)
    kk = np.linspace(0, 3, 6000)
    P = np.zeros_like(kk)
    P[1:] = 2e4 * (kk[1:] / 0.1) ** -1.5
    o = O.PolyBin3D(1000., [n] * 3, Pk=[kk, P])
    d = o.generate_data(seed=3, output_type='real')
    d = d + 0.5 * (d ** 2 - np.mean(d ** 2))
    d32 = d.astype(np.float32)
    b = pb.PolyBin3D(1000., [n] * 3, Pk=[kk, P])
    bs = pb.BSpec(b, edges, lmax=2, precision=precision)
    bo = O.BSpec(o, edges, lmax=2)
    got = bs.Bk_numerator_ideal(d32)
    ref = bo.Bk_numerator_ideal(d32.astype(np.float64))
    # magnitude scale: the same triple sums of |g| (the oracle's own g-maps)
    S = o.to_fourier(d32.astype(np.float64)) * bo.invPk0 / bo.cube_mask_mean ** (1. / 3.)
    g = np.abs(bo._g_maps(S))
    scale = O.assemble_b3(g, bo.all_bins) * o.volume / o.gridsize.prod() / bo.sym_factor
    err = np.max(np.abs(got - ref) / scale)
    assert err < tol, "max scaled error %.3e" % err
