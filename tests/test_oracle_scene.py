"""Pins of the seeded height-field oracle (oracle/scene.py, restating
compare/haynes.py:94-112) and of the product's host generators against it."""

import numpy as np
import pytest
from scipy import stats

from oracle import scene as osc
from radar_sounder_simulator_amd import roughness as R
from radar_sounder_simulator_amd import scene as S


def _closed_form_kernel(n, m, dx, corr_len):
    """exp(-2 rho^2 / l^2) with rho the periodic (minimal-image) distance,
    written as an explicit double loop."""
    out = np.empty((n, m))
    for i in range(n):
        for j in range(m):
            dy = min(i, n - i) * dx
            dxx = min(j, m - j) * dx
            out[i, j] = np.exp(-2.0 * (dy * dy + dxx * dxx) / corr_len ** 2)
    return out


@pytest.mark.parametrize("shape,pos", [((16, 24), (0, 0)), ((17, 9), (5, 3))])
def test_impulse_response_is_normalised_kernel(shape, pos):
    n, m = shape
    dx, l = 10.0, 35.0
    w = np.zeros(shape)
    w[pos] = 1.0
    z = osc.gaussian_field(w, dx, l)
    g = _closed_form_kernel(n, m, dx, l)
    want = np.roll(np.roll(g, pos[0], axis=0), pos[1], axis=1) / np.sqrt((g * g).sum())
    np.testing.assert_allclose(z, want, rtol=0, atol=1e-14)


def test_fft_filter_equals_direct_circular_sum():
    rng = np.random.default_rng(0)
    w = rng.standard_normal((5, 7))
    np.testing.assert_allclose(osc.gaussian_field(w, 20.0, 45.0),
                               osc.gaussian_field_brute(w, 20.0, 45.0), rtol=0, atol=1e-13)


@pytest.mark.parametrize("l_cells", [4.0, 9.5])
def test_field_autocorrelation_is_gaussian_acf(l_cells):
    """The field's exact autocorrelation (the impulse response correlated
    with itself) against the continuous law exp(-rho^2 / l^2): unit variance
    at zero lag, the 1/e point at rho = l.  A kernel width a = l (not
    l/sqrt 2) or a missing normalisation fails by O(1)."""
    n, dx = 128, 5.0
    l = l_cells * dx
    w = np.zeros((n, n))
    w[0, 0] = 1.0
    h = osc.gaussian_field(w, dx, l)
    H = np.fft.fft2(h)
    acf = np.real(np.fft.ifft2(H * np.conj(H)))
    lag = np.arange(n // 4)
    np.testing.assert_allclose(acf[0, lag], np.exp(-(lag * dx) ** 2 / l ** 2), rtol=0, atol=1e-9)
    np.testing.assert_allclose(acf[lag, lag], np.exp(-2 * (lag * dx) ** 2 / l ** 2), rtol=0,
                               atol=1e-9)


def test_keyed_noise_product_equals_scalar_restatement():
    for seed in (0, 3, (4, 17, 2)):
        np.testing.assert_allclose(R.keyed_normals(6, 9, seed), osc.keyed_noise(6, 9, seed),
                                   rtol=1e-15, atol=1e-15)


def test_keyed_noise_is_standard_normal_and_crop_invariant():
    z = R.keyed_normals(400, 500, (9, 1))
    n = z.size
    assert abs(z.mean()) < 5.0 / np.sqrt(n)
    assert abs(z.var() - 1.0) < 5.0 * np.sqrt(2.0 / n)
    assert stats.kstest(z.ravel(), "norm").pvalue > 1e-4
    # neighbouring values uncorrelated
    assert abs(np.mean(z[:, 1:] * z[:, :-1])) < 5.0 / np.sqrt(n)
    # keyed by vertex: a smaller lattice is a crop of a larger one
    np.testing.assert_array_equal(R.keyed_normals(37, 41, (9, 1)), z[:37, :41])
    # distinct seeds give distinct noise
    assert not np.allclose(R.keyed_normals(8, 8, (9, 2)), z[:8, :8])


def test_product_generators_equal_oracle():
    rng = np.random.default_rng(1)
    w = rng.standard_normal((33, 20))
    np.testing.assert_allclose(S.gaussian_filter(w, 25.0, 120.0),
                               osc.gaussian_field(w, 25.0, 120.0), rtol=0, atol=1e-14)
    np.testing.assert_allclose(S.gaussian_surface(33, 25.0, 120.0, np.random.default_rng(1),
                                                  m=20),
                               osc.gaussian_field(w, 25.0, 120.0), rtol=0, atol=1e-14)
    want = osc.gaussian_field(osc.keyed_noise(12, 15, (2, 5)), 30.0, 90.0)
    np.testing.assert_allclose(S.gaussian_surface_keyed(12, 30.0, 90.0, (2, 5), m=15), want,
                               rtol=0, atol=1e-13)
    lat = S.SeededLattice(11, 14, 30.0, sigma=4.0, corr_len=90.0, seed=(2, 5), mean=-50.0)
    np.testing.assert_allclose(lat.heights, -50.0 + 4.0 * want, rtol=0, atol=1e-12)


def test_seeded_lattice_model():
    lat = S.SeededLattice(6, 9, 25.0, sigma=0.0, mean=-120.0)
    assert lat.horizontal() and lat.level() == -120.0 and len(lat) == 54
    assert lat.grid_shape == (6, 9)
    np.testing.assert_array_equal(lat.heights, np.full((7, 10), -120.0))
    f = S.LatticeFacets(lat)
    assert len(f) == 54 and f._f is None          # length without tessellating
    np.testing.assert_array_equal(f.centers[:, 2], -120.0)
    rough = S.SeededLattice(6, 9, 25.0, sigma=1.0, corr_len=80.0, seed=1)
    assert not rough.horizontal()
    with pytest.raises(ValueError):
        rough.heights = np.zeros((6, 9))
    rough.heights = np.ones((7, 10))
    np.testing.assert_array_equal(rough.heights, 1.0)
    assert "SeededLattice" in repr(rough) and rough != S.SeededLattice(6, 9, 25.0)


def test_keyed_phasor_spec_equals_keyed_generator():
    from oracle import roughness as orr
    kp = R.KeyedPhasors((5, 7), (3, 2), key_origin=(10, -4))
    keys = kp.keys()
    assert keys.shape == (35, 2) and tuple(keys[8]) == (11, -3)
    got = np.asarray(kp)
    assert got.dtype == np.complex64
    np.testing.assert_array_equal(got, R.speckle_phasors_for_keys(keys, (3, 2)))
    np.testing.assert_allclose(got, orr.speckle_phasors_for_keys(keys, (3, 2)), rtol=1e-6,
                               atol=1e-7)
