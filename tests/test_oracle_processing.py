"""Pins of oracle/processing.py (CPU): the boxcar conventions of presum /
unfocused SAR / multilook and the backprojection of focused SAR, against
identities, closed-form gains, speckle statistics and a simulated point
target whose focused resolution is the textbook lambda r / (2 L)."""

import numpy as np
import pytest

from oracle import processing as op

C = 299792458.0


def test_presum_is_decimated_unfocused_and_linear():
    rng = np.random.default_rng(0)
    T, n = 103, 7
    x = rng.standard_normal((T, 40)) + 1j * rng.standard_normal((T, 40))
    Wp = op.boxcar_weights(T, n, n)
    Wu = op.boxcar_weights(T, n, 1)
    assert Wp.shape == (T // n, T) and Wu.shape == (T - n + 1, T)
    p = op.apply_weights(Wp, x)
    u = op.apply_weights(Wu, x)
    np.testing.assert_array_equal(p, u[::n][:T // n])
    y = rng.standard_normal((T, 40))
    np.testing.assert_allclose(op.apply_weights(Wp, 2 * x + 3 * y),
                               2 * p + 3 * op.apply_weights(Wp, y), rtol=1e-13, atol=1e-12)


def test_coherent_gain_and_noise_growth():
    """A constant field gains n in amplitude; independent unit noise grows
    as sqrt(n) (SNR gain sqrt(n), processing.py:153-156)."""
    rng = np.random.default_rng(1)
    T, n = 4000, 16
    W = op.boxcar_weights(T, n, n)
    np.testing.assert_allclose(op.apply_weights(W, np.full((T, 3), 2.0 - 1.0j)),
                               n * (2.0 - 1.0j), rtol=1e-14)
    z = (rng.standard_normal((T, 64)) + 1j * rng.standard_normal((T, 64))) / np.sqrt(2)
    s = op.apply_weights(W, z)
    assert abs(np.mean(np.abs(s) ** 2) / n - 1.0) < 0.05


def test_multilook_speckle_contrast():
    """Fully developed speckle: exponential intensities, contrast 1 ->
    1/sqrt(n) after an n-look average (processing.py:224-228)."""
    rng = np.random.default_rng(2)
    T, n = 8192, 16
    I = rng.exponential(1.0, (T, 64))
    m = op.apply_weights(op.boxcar_weights(T, n, n), I, average=True)
    contrast = m.std() / m.mean()
    assert abs(contrast * np.sqrt(n) - 1.0) < 0.03


def _point_target(T=401, dx=1.0, H=1000.0, f0=150e6, B=15e6, nb=256, over=4):
    """Straight track along x at height H over a point target at the origin
    below the middle trace: F[j, b] = sinc(B (twtt_b - 2 R_j / c))
    exp(-2 j k R_j), bins aligned so that bin 80 is the nadir delay."""
    lam = C / f0
    k = 2 * np.pi / lam
    dt = 1.0 / (over * B)
    t0 = 2 * H / C - 80 * dt
    x = (np.arange(T) - T // 2) * dx
    R = np.sqrt(H * H + x * x)
    twtt = t0 + dt * np.arange(nb)
    F = np.sinc(B * (twtt[None, :] - 2 * R[:, None] / C)) * np.exp(-2j * k * R)[:, None]
    pos = np.column_stack([x, np.zeros(T), np.full(T, H)])
    return F, op.along_track(pos), dict(twtt=twtt, c=C, k=k), lam


def _width_3db(a, dx):
    """-3 dB width (in the sample unit dx) of a peaked profile |a|, by
    linear interpolation on both flanks."""
    p = np.abs(a) ** 2
    i = int(np.argmax(p))
    h = p[i] / 2
    l = i
    while p[l] > h:
        l -= 1
    r = i
    while p[r] > h:
        r += 1
    xl = l + (h - p[l]) / (p[l + 1] - p[l])
    xr = r - 1 + (p[r - 1] - h) / (p[r - 1] - p[r])
    return (xr - xl) * dx


@pytest.mark.parametrize("window,bins", [("none", 0.886), ("hann", 1.44)])
def test_point_target_focuses_to_the_textbook_resolution(window, bins):
    F, s, kw, lam = _point_target()
    H, L = 1000.0, 200.0
    out = op.focused_sar(F, s, aperture_m=L, window=window, **kw)
    i_t, b_t = np.unravel_index(np.argmax(np.abs(out)), out.shape)
    assert (i_t, b_t) == (200, 80)
    d = s - s[200]
    J = np.abs(d) <= L / 2
    wsum = (np.ones(J.sum()) if window == "none"
            else 0.5 * (1 + np.cos(np.pi * d[J] / (L / 2)))).sum()
    gain = np.abs(out[200, 80]) / wsum
    assert 0.95 < gain <= 1.0 + 1e-9, gain
    # azimuth resolution: the window's -3 dB width in bins (Harris 1978:
    # 0.89 rectangular, 1.44 hann) times lambda r / (2 L), measured on the
    # -3 dB points of the focused profile
    res = _width_3db(out[100:301, 80], 1.0)
    want = bins * lam * H / (2 * L)
    assert abs(res / want - 1.0) < 0.06, (res, want)
    # the unfocused boxcar of the same n-trace history sums a quadratic
    # phase 2 pi x^2 / (lambda r): its gain is the Fresnel integral's
    # |int exp(-2 pi j x^2 / (lambda r)) dx| = sqrt(lambda r) (per unit
    # spacing; ~45 traces here, processing.py:194-198), not n = 201
    n = int(J.sum())
    u = op.apply_weights(op.boxcar_weights(F.shape[0], n, 1), F)
    lu = np.sqrt(lam * H) / 1.0
    assert 0.7 * lu < np.abs(u).max() < 1.3 * lu < 0.35 * n


def test_focused_sar_window_edges_and_validity():
    """Traces outside the aperture contribute nothing: a one-trace
    aperture reproduces the field at zero migration (the last bin, whose
    right neighbour is missing, hangs on the rounding of (twtt[-1] - t0) /
    dt); delays migrated past the window read zero."""
    F, s, kw, _ = _point_target(T=21, nb=64)
    one = op.focused_sar(F, s, aperture_m=0.5, window="none", **kw)
    np.testing.assert_allclose(one[:, :-1], F[:, :-1], rtol=0, atol=1e-12)
    ones = np.ones_like(F)
    far = op.focused_sar(ones, s * 200.0, aperture_m=2000.0, window="none", **kw)
    # traces 200 m apart at ~1 km: the neighbours' delays migrate >= 20 m,
    # >= 8 bins, so the last bins of a trace collect only its own sample
    # while earlier bins sum several traces
    assert np.abs(far[10, -8:-1] - 1.0).max() < 1e-12
    assert np.abs(far[10, :40]).max() > 4.0
    with pytest.raises(ValueError):
        op.focused_sar(F, s, aperture_m=10.0, window="kaiser", **kw)
