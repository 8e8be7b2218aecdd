"""The pyramid oracle (oracle/pyramid.py) pinned by mathematics.

SimpleITK, which the reference's drawing.scale_image calls, is not installed,
so its recursive Gaussian is restated from Deriche (1993) and checked here
against what the mathematics fixes:

* the recursion reproduces the closed-form impulse response it is derived
  from (h(|n|) normalised), to rounding: a wrong pole, zero or sign in the
  causal/anti-causal split fails this;
* that impulse response approximates the sampled Gaussian to within
  5e-4 of its peak (the fourth-order approximation's own error, measured on
  the continuous h(x) against exp(-x^2/2)), and its second moment is sigma^2
  within 0.5 % for sigma >= 1;
* unit DC gain (a constant stays constant, borders included), symmetry, and a
  linear ramp left unchanged away from the borders;
* ITK's resampling: a linear ramp is resampled exactly where the neighbours
  are not clamped, nearest neighbour picks floor(c + 0.5), and voxels whose
  continuous index leaves [-0.5, n - 0.5) are 0.
"""

import numpy as np
import pytest

from oracle import pyramid as P


def _closed_form(sigma, k):
    c = P.DERICHE_GAUSSIAN
    t = np.abs(k) / sigma
    return (c["a0"] * np.cos(c["w0"] * t) + c["a1"] * np.sin(c["w0"] * t)) * np.exp(-c["b0"] * t) \
        + (c["c0"] * np.cos(c["w1"] * t) + c["c1"] * np.sin(c["w1"] * t)) * np.exp(-c["b1"] * t)


def _impulse(sigma):
    L = int(40 * sigma) + 41
    c = L // 2
    x = np.zeros(L)
    x[c] = 1.0
    return P.recursive_gaussian_1d(x, sigma), np.arange(L) - c


def test_continuous_approximation_error():
    x = np.linspace(0.0, 8.0, 8001)
    err = np.abs(_closed_form(1.0, x) - np.exp(-0.5 * x * x)).max()
    assert err < 5.3e-4


@pytest.mark.parametrize("sigma", [0.5, 1.0, 2.0, 3.5, 8.0])
def test_recursion_reproduces_closed_form(sigma):
    r, k = _impulse(sigma)
    h = _closed_form(sigma, k)
    np.testing.assert_allclose(r, h / h.sum(), rtol=0, atol=1e-13)


@pytest.mark.parametrize("sigma", [1.0, 2.0, 3.5, 8.0])
def test_impulse_response_is_gaussian(sigma):
    r, k = _impulse(sigma)
    g = np.exp(-0.5 * (k / sigma) ** 2)
    g /= g.sum()
    assert np.abs(r - g).max() <= 5e-4 * g.max()
    assert abs((k * k * r).sum() - sigma * sigma) <= 5e-3 * sigma * sigma


@pytest.mark.parametrize("sigma", [0.5, 2.0, 8.0])
def test_unit_dc_gain_and_symmetry(sigma):
    r, _ = _impulse(sigma)
    assert abs(r.sum() - 1.0) < 1e-12
    np.testing.assert_allclose(r, r[::-1], rtol=0, atol=1e-15)
    x = np.full((3, 37), 3.7)
    np.testing.assert_allclose(P.recursive_gaussian_1d(x, sigma), x, rtol=1e-13)


def test_linear_ramp_unchanged_in_the_interior():
    x = 2.0 + 0.3 * np.arange(200)
    y = P.recursive_gaussian_1d(x, 2.0)
    np.testing.assert_allclose(y[40:160], x[40:160], rtol=0, atol=1e-11)
    # the borders continue the signal as constants, which bends the ramp there
    assert abs(y[0] - x[0]) > 0.1


def test_smooth_is_separable_and_per_axis_sigma():
    rng = np.random.default_rng(3)
    v = rng.normal(size=(1, 9, 10, 11))
    sp = (1.0, 2.0, 0.5)
    a = P.smooth(v, sp, sigma=2.0)
    b = P.recursive_gaussian_1d(P.recursive_gaussian_1d(P.recursive_gaussian_1d(v, 2.0, axis=3), 1.0, axis=2),
                                4.0, axis=1)
    np.testing.assert_allclose(a, b, rtol=0, atol=1e-14)


def test_resample_linear_ramp_is_exact():
    nz, ny, nx = 12, 10, 14
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    v = (1.5 + 0.2 * x - 0.3 * y + 0.05 * z)[None]
    size = (9, 7, 5)
    out = P.resample(v, size)
    cx = np.arange(size[0]) * nx / size[0]
    cy = np.arange(size[1]) * ny / size[1]
    cz = np.arange(size[2]) * nz / size[2]
    want = 1.5 + 0.2 * cx[None, None, :] - 0.3 * cy[None, :, None] + 0.05 * cz[:, None, None]
    ok = (cx <= nx - 1)[None, None, :] & (cy <= ny - 1)[None, :, None] & (cz <= nz - 1)[:, None, None]
    np.testing.assert_allclose(out[0][ok], np.broadcast_to(want, ok.shape)[ok], rtol=0, atol=1e-12)


def test_resample_nearest_and_outside():
    v = np.arange(10, dtype=np.float64).reshape(1, 1, 1, 10)
    # upsampling 10 -> 25 along x: c = 0.4 i, nearest = floor(c + 0.5); the
    # last index, c = 9.6, lies past n - 0.5 = 9.5 and reads 0
    out = P.resample(v, (25, 1, 1), nearest=True)
    c = np.arange(25) * 0.4
    np.testing.assert_array_equal(out[0, 0, 0], np.where(c < 9.5, np.floor(c + 0.5), 0.0))
    assert out[0, 0, 0, -1] == 0.0
    # downsampling maps every output index inside
    w = np.ones((1, 3, 3, 3))
    out = P.resample(w, (2, 2, 2))
    assert (out == 1.0).all()


def test_scale_image_constant_volume():
    v = np.full((2, 16, 12, 10), 0.75)
    out, sp = P.scale_image(v, (1.0, 1.0, 1.5), (5, 6, 8))
    np.testing.assert_allclose(out, 0.75, rtol=1e-13)
    assert sp == pytest.approx((2.0, 2.0, 3.0))


def test_pyramid_sizes_halve():
    assert P.pyramid_sizes((130, 66, 41), 3) == [(130, 66, 41), (65, 33, 20), (32, 16, 10)]
