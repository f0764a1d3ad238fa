"""Cross-check of the oracle's utils restatement against the reference's own utils.pyx, compiled
from its source into oracle/_ref/ by oracle/build_ref.py (present only where the reference tree is;
skipped elsewhere).  CPU only."""
import glob
import importlib.util
import os

import numpy as np
import pytest

from oracle import polybin_oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def ref_utils():
    so = glob.glob(os.path.join(ROOT, "oracle", "_ref", "utils*.so"))
    if not so:
        pytest.skip("oracle/_ref not built (needs the reference tree)")
    spec = importlib.util.spec_from_file_location("utils", so[0])
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


@pytest.fixture(scope="module")
def setup():
    U = ref_utils()
    b = O.PolyBin3D(500., [8, 12, 10], boxcenter=[200., -50., 120.], sightline='local')
    I = U.IntegrationUtils(b.modk_grid, True, 2, degk=b.degeneracy_factor, muk=b.muk_grid)
    M = U.MapUtils(b.modk_grid, b.gridsize, 2, muk=b.muk_grid)
    return U, b, I, M


def test_integrals_match_reference(setup):
    U, b, I, M = setup
    rng = np.random.RandomState(0)
    sh = b.modk_grid.shape
    s1 = rng.randn(*sh) + 1j * rng.randn(*sh)
    s2 = rng.randn(*sh) + 1j * rng.randn(*sh)
    r = rng.rand(*sh)
    edges = np.array([0.013, 0.03, 0.05, 0.07])
    for l in (0, 2, 4):
        np.testing.assert_allclose(b.integrator.cross_integrate(s1, s2, edges, l), I.cross_integrate(s1, s2, edges, l),
                                   rtol=1e-12, atol=1e-12)
        np.testing.assert_allclose(b.integrator.integrate(s1, edges, l), I.integrate(s1, edges, l), rtol=1e-12)
        for l2 in (0, 2):
            np.testing.assert_allclose(b.integrator.real_integrate_double(r, l, l2, edges),
                                       I.real_integrate_double(r, l, l2, edges), rtol=1e-12, atol=1e-14)
    arr = np.stack([s1, s2, s1 - s2])
    np.testing.assert_allclose(b.integrator.integrate_row(s2, arr, 0.03, 0.05), I.integrate_row(s2, arr, 0.03, 0.05),
                               rtol=1e-12)
    coeff = rng.randn(3) + 1j * rng.randn(3)
    np.testing.assert_allclose(b.integrator.cross_integrate_wcoeff(coeff, s1, s2, edges),
                               I.cross_integrate_wcoeff(coeff, s1, s2, edges), rtol=1e-12, atol=1e-12)


def test_maps_match_reference(setup):
    U, b, I, M = setup
    rng = np.random.RandomState(1)
    sh = b.modk_grid.shape
    s = rng.randn(*sh) + 1j * rng.randn(*sh)
    w = rng.rand(*sh)
    for l in (0, 2, 4):
        np.testing.assert_allclose(b.map_utils.fourier_filter(s, l, 0.02, 0.06), M.fourier_filter(s, l, 0.02, 0.06),
                                   rtol=1e-14, atol=1e-15)
        np.testing.assert_allclose(b.map_utils.fourier_filter_ll(s, l, 2, 0.02, 0.06),
                                   M.fourier_filter_ll(s, l, 2, 0.02, 0.06), rtol=1e-14)
    np.testing.assert_array_equal(b.map_utils.prod_fourier_filter(s, w, 0.02, 0.06), M.prod_fourier_filter(s, w, 0.02, 0.06))
    x = [rng.randn(*b.gridsize) for _ in range(4)]
    assert np.isclose(O.sum_pair(x[0], x[1]), M.sum_pair(x[0], x[1]), rtol=1e-12)
    assert np.isclose(O.sum_triplet(x[0], x[1], x[2]), M.sum_triplet(x[0], x[1], x[2]), rtol=1e-12)


def test_harmonics_match_reference(setup):
    """Same addition-theorem kernel: sum_m Y_lm(a) Y_lm(b) equal for the reference's basis and ours."""
    U, b, I, M = setup
    rg = np.ascontiguousarray(np.asarray(b.r_grids))
    kg = np.ascontiguousarray(np.asarray(b.k_grids))
    R = U.compute_real_harmonics(rg, 4, True, 2)
    K = U.compute_real_harmonics(kg, 4, True, 2)
    Ro, Ko = O.real_harmonics(rg, 4, True), O.real_harmonics(kg, 4, True)
    for l in (1, 2, 3, 4):
        a = np.asarray(R[l]).reshape(2 * l + 1, -1)[:, :50]
        c = np.asarray(K[l]).reshape(2 * l + 1, -1)[:, :60]
        ao = Ro[l].reshape(2 * l + 1, -1)[:, :50]
        co = Ko[l].reshape(2 * l + 1, -1)[:, :60]
        np.testing.assert_allclose(a.T @ c, ao.T @ co, atol=1e-12)


def test_triplet_assembly_matches_reference(setup):
    U, b, I, M = setup
    rng = np.random.RandomState(2)
    g = rng.randn(2, 3, *b.gridsize)
    s = rng.randn(2, 3, *b.gridsize)
    bins = O.triangle_bins([0.01, 0.02, 0.03, 0.04], [0.01, 0.02, 0.03, 0.04], 2)
    np.testing.assert_allclose(O.assemble_b3(g, bins), U.assemble_b3(g, bins, 2), rtol=1e-12)
    np.testing.assert_allclose(O.assemble_b1(g, s, bins), U.assemble_b1(g, s, bins, 2), rtol=1e-12)
    c = rng.randn(2, 3, *b.gridsize)
    np.testing.assert_allclose(O.assemble_bshot(g, s, c, bins), U.assemble_bshot(g, s, c, bins, 2), rtol=1e-12)


def test_integral_constraint_pointing_matches_reference(setup):
    """pointing_GIC/RIC and their transposes (utils.pyx:405-513) against the compiled kernels."""
    U, b, I, M = setup
    rng = np.random.RandomState(4)
    n = tuple(int(v) for v in b.gridsize)
    m, mask, mic = rng.randn(*n), rng.rand(*n), rng.rand(*n)
    modr = rng.uniform(0., 10., n)
    edges = np.array([1., 3.5, 6., 9.])
    np.testing.assert_allclose(O.pointing_GIC(m, mask, mic), U.pointing_GIC(m, mask, mic, 2), rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(O.pointing_GIC_transpose(m, mask, mic), U.pointing_GIC_transpose(m, mask, mic, 2),
                               rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(O.pointing_RIC(m, mask, mic, edges, modr),
                               U.pointing_RIC(m, mask, mic, edges, modr, 2), rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(O.pointing_RIC_transpose(m, mask, mic, edges, modr),
                               U.pointing_RIC_transpose(m, mask, mic, edges, modr, 2), rtol=1e-12, atol=1e-14)
