"""The autotuner (polybin3d_amd.tune): with tuning on, the first call of each knob and shape times its
candidates and writes the choice to the cache file; the tuned calls give the same results as the
built-in shapes (the triplet GEMM up to FP32 rounding order, the Q-map chunking and the seed batching
exactly); a later call reads the cache instead of timing again."""
import json

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

L = 1000.


@pytest.fixture
def tuner(tmp_path, monkeypatch):
    from polybin3d_amd import tune
    monkeypatch.setenv("PB3D_TUNE_CACHE", str(tmp_path / "tune.json"))
    tune.enable(True)
    yield tune
    tune.enable(None)


def test_tuned_shapes_give_the_same_results(hip, tuner):
    import polybin3d_amd as pb
    n, c = 32, [10., 10., 10.]
    b = pb.PolyBin3D(L, [n] * 3, boxcenter=c, sightline='local', rng='device')
    x = [np.fft.fftshift(np.arange(-n // 2, n // 2)) * L / n + c[i] for i in range(3)]
    X = np.meshgrid(*x, indexing='ij')
    mask = 0.2 + (np.sqrt(X[0] ** 2 + X[1] ** 2 + X[2] ** 2) < 450.)
    d = np.random.RandomState(3).randn(n, n, n)
    edges = np.linspace(0.01, 0.07, 5)
    tuner.enable(False)
    ref_bk = pb.BSpec(b, edges, lmax=2).Bk_ideal(d)
    ref_F = pb.BSpec(b, edges, lmax=2, mask=mask).compute_fisher_contribution(1)
    ps = pb.PSpec(b, np.linspace(0.01, 0.09, 9), lmax=4, mask=mask)
    ref_P = ps.compute_fisher(4)
    tuner.enable(True)
    got_bk = pb.BSpec(b, edges, lmax=2).Bk_ideal(d)
    got_F = pb.BSpec(b, edges, lmax=2, mask=mask).compute_fisher_contribution(1)
    got_P = ps.compute_fisher(4, batch='auto')
    cache = json.load(open(tuner.cache_path()))
    knobs = sorted({k.split('|')[3] for k in cache})
    print("[tune] cached knobs:", knobs, "choices:", {k.split('|')[3]: v['choice'] for k, v in cache.items()})
    assert {'pair_gemm', 'bk_qmap_chunk', 'pk_batch'} <= set(knobs)
    for k in ref_bk:
        s = np.abs(ref_bk[k]).max()
        assert np.max(np.abs(got_bk[k] - ref_bk[k])) < 1e-5 * s
    assert np.array_equal(got_F, ref_F)
    np.testing.assert_allclose(got_P, ref_P, rtol=1e-12, atol=1e-14 * np.abs(ref_P).max())
    # the cache now answers: no candidate is run again
    from polybin3d_amd import tune
    calls = []
    orig = tune.choose

    def spy(knob, shape, candidates, run, device, default, work=None, reps=3):
        def counted(cand):
            calls.append(knob)
            return run(cand)
        return orig(knob, shape, candidates, counted, device, default, work, reps)
    tune.choose = spy
    try:
        pb.BSpec(b, edges, lmax=2).Bk_ideal(d)
    finally:
        tune.choose = orig
    assert calls == []


def test_pair_gemm_cfg_rejects_bad_shapes(hip):
    C = hip
    assert C.lib().pb3d_pair_gemm_cfg(None, 4, 100, None, 4, None, None, 4, 2, 17, 0, None, None, None) == -1
