"""GPU parity at the benchmarked workloads' own shapes (the HIP path against the FP64 oracle on the
same seeded inputs, rng='numpy'), and size-independent properties where the oracle cannot run:

  * Bk_ideal at 128^3, 16 bins 0.01-0.17 h/Mpc, lmax 0 (508 triangles), on the quadratic non-Gaussian
    field: symmetry factors, ideal Fisher, numerator and b0 against the oracle;
  * the Pk MC Fisher at 128^3 (local sightline, lmax 4, 30 bins to 0.30, shell mask with 200 holes):
    one seed against the oracle, and Pk_unwindowed through that Fisher with its condition-number bound;
    compute_fisher_mixed over a 128^3 and a 256^3 spec with batched seeds against each spec's serial
    compute_fisher;
  * the Bk MC Fisher at 256^3 (40 bins to 0.40, lmax 2, 12 700 triangle bins, local, shell mask): a
    24 x 24 block spanning both multipoles and Q-map chunks of small and large b3 against the oracle
    restricted to those rows and columns (BSpec.fisher_subblock); this is the shape whose packed Q maps
    exceed 2^31 elements and whose contraction runs over segments of up to 78 000 reals;
  * at the same shape with a unit mask, global sightline and S^-1 = A^-1 = 1: one realization's
    diagonal against the ideal Fisher with the discreteness correction (its expectation), and its
    off-diagonal entries between different triangles (zero in expectation).
Every check prints its error family and size."""
import time

import numpy as np
import pytest
import torch

from oracle import polybin_oracle as O

import parity

pytestmark = pytest.mark.gpu

L = 1000.
CENTER = [10., 10., 10.]


def log(msg):
    print("[configs] %s" % msg, flush=True)


def powerlaw(kmax=3.):
    k = np.linspace(0., kmax, 6000)
    P = np.zeros_like(k)
    P[1:] = 2e4 * (k[1:] / 0.1) ** -1.5
    return [k, P]


def shell_mask(n, center, seed=1, r_lo=300., r_hi=480., nholes=200, smooth=5.):
    """1 inside r_lo <= r <= r_hi (observer at the origin) minus nholes spheres of radius U[10, 30]
    centred in the shell, Gaussian-smoothed (the benchmark's survey mask)."""
    ax = [np.fft.fftshift(np.arange(-n // 2, n // 2)) * L / n + center[i] for i in range(3)]
    X = np.meshgrid(*ax, indexing='ij', sparse=True)
    r2 = X[0] ** 2 + X[1] ** 2 + X[2] ** 2
    m = ((r2 >= r_lo ** 2) & (r2 <= r_hi ** 2)).astype(np.float64)
    rng = np.random.RandomState(seed)
    for _ in range(nholes):
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        c = u * rng.uniform(r_lo, r_hi)
        rad = rng.uniform(10., 30.)
        sl = [np.nonzero(np.abs(ax[i] - c[i]) <= rad + L / n)[0] for i in range(3)]
        if any(len(s) == 0 for s in sl):
            continue
        sub = np.ix_(*sl)
        d2 = (ax[0][sl[0]][:, None, None] - c[0]) ** 2 + (ax[1][sl[1]][None, :, None] - c[1]) ** 2 + \
             (ax[2][sl[2]][None, None, :] - c[2]) ** 2
        blk = m[sub]
        blk[d2 < rad ** 2] = 0.
        m[sub] = blk
    k = [np.fft.fftfreq(n, L / n) * 2 * np.pi] * 2 + [np.fft.rfftfreq(n, L / n) * 2 * np.pi]
    K = np.meshgrid(*k, indexing='ij', sparse=True)
    g = np.exp(-0.5 * smooth ** 2 * (K[0] ** 2 + K[1] ** 2 + K[2] ** 2))
    return np.fft.irfftn(np.fft.rfftn(m) * g, s=(n, n, n), axes=(0, 1, 2))


def test_cfg_bk_ideal_128(hip):
    """Bk_ideal at 128^3, lmax 0, bins np.arange(0.01, 0.1701, 0.01): N3 = 508 triangles."""
    import polybin3d_amd as pb
    n = 128
    edges = np.arange(0.01, 0.1701, 0.01)
    o = O.PolyBin3D(L, [n] * 3, Pk=powerlaw())
    dG = o.generate_data(seed=0, output_type='real')
    d = dG + 0.5 * (dG ** 2 - np.mean(dG ** 2))
    b = pb.PolyBin3D(L, [n] * 3)
    oo = O.PolyBin3D(L, [n] * 3)
    bs, bo = pb.BSpec(b, edges, lmax=0), O.BSpec(oo, edges, lmax=0)
    assert bs.N_bins == bo.N_bins == 508 and np.array_equal(bs.all_bins, bo.all_bins)
    t = time.time()
    got = bs.Bk_ideal(d)
    log("GPU Bk_ideal 128^3: %.2f s (first call, with the ideal Fisher)" % (time.time() - t))
    t = time.time()
    ref = bo.Bk_ideal(d)
    log("oracle Bk_ideal 128^3: %.1f s" % (time.time() - t))
    parity.check("cfg2 symmetry factors", bs.sym_factor, bo.sym_factor)
    parity.check_fisher("cfg2 ideal Fisher (508 x 508)", bs.fish_ideal, bo.fish_ideal)
    n1, n2 = bs.Bk_numerator_ideal(d), bo.Bk_numerator_ideal(d)
    parity.check("cfg2 numerator", n1, n2, scale=parity.own_scale(n2))
    parity.check_solve("cfg2 Bk_ideal", got, ref, ['b0'], bo.fish_ideal, bs.fish_ideal, n2, n1)


def test_cfg_pk_mc_fisher_128_seed(hip):
    """One seed of the Pk MC Fisher at 128^3: local, lmax 4, 30 bins 0.01-0.30, shell mask (90 x 90)."""
    import polybin3d_amd as pb
    n = 128
    edges = np.linspace(0.01, 0.30, 31)
    kw = dict(boxcenter=CENTER, sightline='local')
    Pk = powerlaw()
    mask = shell_mask(n, CENTER)
    b, o = pb.PolyBin3D(L, [n] * 3, Pk=Pk, **kw), O.PolyBin3D(L, [n] * 3, Pk=Pk, **kw)
    ps = pb.PSpec(b, edges, lmax=4, mask=mask)
    po = O.PSpec(o, edges, lmax=4, mask=mask)
    t = time.time()
    F1 = ps.compute_fisher_contribution(0)
    log("GPU Pk MC Fisher seed at 128^3: %.2f s" % (time.time() - t))
    t = time.time()
    F2 = po.compute_fisher_contribution(0)
    log("oracle Pk MC Fisher seed at 128^3: %.1f s" % (time.time() - t))
    assert F1.shape == (90, 90)
    parity.check_fisher("cfg3 Pk MC Fisher seed 0 (90 x 90)", F1, F2)
    # Pk_unwindowed through this Fisher: numerator errors amplified by cond(F)
    dG = O.PolyBin3D(L, [n] * 3, Pk=Pk).generate_data(seed=7, output_type='real')
    d = dG * mask
    n1, n2 = ps.Pk_numerator(d), po.Pk_numerator(d)
    parity.check("cfg3 Pk numerator", n1, n2, scale=np.concatenate(
        [parity.own_scale(n2[i * 30:(i + 1) * 30]) for i in range(3)]))
    got, ref = ps.Pk_unwindowed(d, fish=F2), po.Pk_unwindowed(d, fish=F2)
    parity.check_solve("cfg3 Pk_unwindowed", got, ref, ['p0', 'p2', 'p4'], F2, F2, n2, n1)


def test_cfg_pk_mixed_batched_equals_serial(hip):
    """compute_fisher_mixed over a 128^3 and a 256^3 spec, 2 seeds per pass, against each spec's serial
    compute_fisher over the same 4 seeds (device GRFs)."""
    import polybin3d_amd as pb
    edges = np.linspace(0.01, 0.30, 31)
    specs = []
    for n in (128, 256):
        b = pb.PolyBin3D(L, [n] * 3, Pk=powerlaw(), boxcenter=CENTER, sightline='local', rng='device')
        specs.append(pb.PSpec(b, edges, lmax=4, mask=shell_mask(n, CENTER)))
    mixed = pb.compute_fisher_mixed(specs, 4, batch=2)
    for ps, fm in zip(specs, mixed):
        ref = ps.compute_fisher(4)
        err = np.max(np.abs(fm - ref)) / np.abs(ref).max()
        log("cfg3 mixed batched vs serial at %d^3: %.2e" % (int(ps.base.gridsize[0]), err))
        assert err < 1e-12


def _subblock_bins(ab, nb, count=24):
    """A spread of triangle bins for the restricted comparison: both multipoles, b3 from the smallest
    to the largest bin, each with three legs in a small set of bins."""
    legs = {0, 1, 4, 9, 15, 22, 30, nb - 1}
    cand = [i for i, (b1, b2, b3, _) in enumerate(ab) if {b1, b2, b3} <= legs]
    cand = np.asarray(cand)
    order = np.lexsort((ab[cand, 2], ab[cand, 3]))
    pick = cand[order][np.linspace(0, len(cand) - 1, count).round().astype(int)]
    return np.unique(pick)


def test_cfg_bk_mc_fisher_256_subblock(hip):
    """A block of one 256^3 Bk MC Fisher realization (N_bins = 12 700) against the restricted oracle."""
    import polybin3d_amd as pb
    n = 256
    edges = np.linspace(0.01, 0.40, 41)
    kw = dict(boxcenter=CENTER, sightline='local')
    Pk = powerlaw()
    mask = shell_mask(n, CENTER)
    b = pb.PolyBin3D(L, [n] * 3, Pk=Pk, **kw)
    bs = pb.BSpec(b, edges, lmax=2, mask=mask)
    assert bs.N_bins == 12700
    t = time.time()
    F = bs._fisher_contribution_dev(0)
    torch.cuda.synchronize()
    log("GPU Bk MC Fisher realization at 256^3: %.2f s" % (time.time() - t))
    sel = _subblock_bins(bs.all_bins, bs.Nk)
    assert len(sel) >= 20 and len(set(bs.all_bins[sel, 3])) == 2
    idx = torch.from_numpy(sel).to(F.device)
    got = F[idx][:, idx].cpu().numpy()
    Fd = torch.diagonal(F).cpu().numpy()
    del F
    o = O.PolyBin3D(L, [n] * 3, Pk=Pk, **kw)
    bo = O.BSpec(o, edges, lmax=2, mask=mask)
    t = time.time()
    ref, mag = bo.fisher_subblock(0, sel, sel, return_scale=True)
    log("oracle restricted to %d x %d at 256^3: %.1f s" % (len(sel), len(sel), time.time() - t))
    parity.check("cfg4 symmetry factors", bs.sym_factor, bo.sym_factor)
    assert np.all(np.diag(ref) > 0)
    # Each entry is a sum over the middle bin's modes of products of four GRF-pair fields; one seed's
    # entry can cancel to far below the magnitude sum |terms| (mag) that the FP32 rounding of the maps
    # acts on.  The tolerance is 1e-5 of sqrt(F_ii F_jj), or of mag where that is larger: FP32 maps
    # cannot resolve an entry better than ~1e-6 of mag.
    fs = parity.fisher_scale(ref)
    cond = mag / fs
    log("cfg4 block: sum|terms| / sqrt(F_ii F_jj) median %.1f, max %.1f; %d of %d entries above 1"
        % (np.median(cond), cond.max(), int((cond > 1).sum()), cond.size))
    parity.check("cfg4 Bk MC Fisher seed 0, %d x %d block (scale max(sqrt(F_ii F_jj), sum|terms|))"
                 % (len(sel), len(sel)), got, ref, scale=np.maximum(fs, mag))
    parity.check("cfg4 Bk MC Fisher block, relative to sum|terms|", got, ref, scale=mag)
    parity.check("cfg4 Fisher diagonal of the block", Fd[sel], np.diag(ref), scale=np.maximum(np.diag(ref),
                                                                                             np.diag(mag)))


def test_cfg_bk_mc_fisher_256_against_ideal(hip):
    """Unit mask, global sightline, unit power (S^-1 = A^-1 = 1) at 256^3, 40 bins, lmax 2: one MC
    realization is unbiased for the ideal Fisher with the discreteness correction (pinned on the
    oracle over 200 seeds).  For l = 0 triangles whose three bins each hold more than 1e4 modes, the
    diagonal ratio F_MC / F_ideal is held within 0.2 of 1 per bin and its mean within 0.02 of 1; the
    entries between two different l = 0 triangles of that set (zero in expectation) are held to an rms
    of 0.05 of sqrt(F_ii F_jj).  A column taken from the wrong Q map, a wrong segment offset or an
    overflowed index moves these by O(1)."""
    import polybin3d_amd as pb
    n = 256
    edges = np.linspace(0.01, 0.40, 41)
    b = pb.PolyBin3D(L, [n] * 3, rng='device')
    bs = pb.BSpec(b, edges, lmax=2)
    t = time.time()
    F = bs._fisher_contribution_dev(3)
    torch.cuda.synchronize()
    log("GPU Bk MC Fisher realization (unit mask, global) at 256^3: %.2f s" % (time.time() - t))
    Fid = bs.compute_fisher_ideal(discreteness_correction=True)
    counts = np.bincount(b.bin_index_host(edges).reshape(-1)[b.bin_index_host(edges).reshape(-1) >= 0],
                         minlength=bs.Nk)
    ab = bs.all_bins
    big = np.nonzero((ab[:, 3] == 0) & (counts[ab[:, 0]] > 1e4) & (counts[ab[:, 1]] > 1e4)
                     & (counts[ab[:, 2]] > 1e4))[0]
    assert len(big) > 100
    Fd = torch.diagonal(F).cpu().numpy()
    r = Fd[big] / np.diag(Fid)[big]
    idx = torch.from_numpy(big).to(F.device)
    blk = F[idx][:, idx].cpu().numpy()
    del F
    off = blk / np.sqrt(np.outer(Fd[big], Fd[big]))
    off = off[~np.eye(len(big), dtype=bool)]
    log("cfg4 MC/ideal diagonal over %d triangles: mean %.4f, min %.3f, max %.3f; off-diagonal rms %.4f"
        % (len(big), r.mean(), r.min(), r.max(), np.sqrt(np.mean(off ** 2))))
    assert np.max(np.abs(r - 1.)) < 0.2
    assert abs(r.mean() - 1.) < 0.02
    assert np.sqrt(np.mean(off ** 2)) < 0.05
