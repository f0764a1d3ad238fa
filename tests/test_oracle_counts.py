"""The structural FFT counts used to extrapolate the CPU baseline equal the oracle's live counters
(which the notebook counters pin) on small configurations of every branch."""
import numpy as np
import pytest

from oracle import polybin_oracle as O


@pytest.mark.parametrize("local", [False, True])
@pytest.mark.parametrize("const_mask", [True, False])
@pytest.mark.parametrize("lmax", [0, 2])
def test_fft_counts_match_live_counters(local, const_mask, lmax):
    n = 16
    kw = dict(boxcenter=[300., 100., 50.], sightline='local') if local else {}
    b = O.PolyBin3D(1000., [n] * 3, **kw)
    mask = None if const_mask else 0.5 + np.random.RandomState(0).rand(n, n, n)
    bs = O.BSpec(b, np.linspace(0.007, 0.05, 4), lmax=lmax, mask=mask)
    bs._compute_symmetry_factor()
    b.n_FFTs_forward = b.n_FFTs_reverse = 0
    bs.compute_fisher_contribution(0)
    assert (b.n_FFTs_forward, b.n_FFTs_reverse) == O.bk_fisher_fft_count(bs.all_bins, bs.Nk_squeeze, lmax, local,
                                                                          bs.const_mask)
    ps = O.PSpec(b, np.linspace(0.007, 0.05, 5), lmax=2 * lmax, mask=mask)
    b.n_FFTs_forward = b.n_FFTs_reverse = 0
    ps.compute_fisher_contribution(0)
    assert (b.n_FFTs_forward, b.n_FFTs_reverse) == O.pk_fisher_fft_count(ps.Nk, 2 * lmax, local, ps.const_mask)


def test_metric_configuration_counts():
    """256^3 Pk Fisher (30 bins, lmax=4, local, shell mask): 179 + 513 FFTs; Bk Fisher (40 bins to
    0.4, lmax=2): N_bins 12700 and 67128 FFTs local (29004 global) per realization."""
    assert O.pk_fisher_fft_count(30, 4, True, False) == (179, 513)
    e = np.linspace(0.01, 0.40, 41)
    ab = O.triangle_bins(e, e, 2)
    assert len(ab) == 12700
    assert sum(O.bk_fisher_fft_count(ab, 40, 2, True, False)) == 67128
    assert sum(O.bk_fisher_fft_count(ab, 40, 2, False, False)) == 29004
