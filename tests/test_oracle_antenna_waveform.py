"""Pins of the antenna and waveform oracles -- CPU.

Antenna (tests/test_antenna.py:119-294 of the reference): half-wave dipole
closed form, finite dipole L = lambda/2 equals the dipole, uniform-array
factor vs the explicit element phasor sum, equal-weight tapered array equals
the uniform array, a fine tabulated dipole reproduces the dipole, and the
array directivity scale for half-wave-spaced 3 dBi elements (D = 2N).

Waveform (tests/test_waveform.py:81-161, 317-327, tests/test_waveform_chirp
.py:23-79): Harris (1978) peak sidelobe levels and -3 dB widths, p(0) = 1,
FFT 'same' convolution vs np.convolve, chirp matched filter vs the analytic
mainlobe, Tukey limits, and product pulses equal to the oracle's.
"""

import numpy as np
import pytest

from oracle import antenna as oa
from oracle import waveform as ow
from radar_sounder_simulator_amd import antenna as pa
from radar_sounder_simulator_amd import waveform as pw


def _dirs(n=400, seed=0):
    rng = np.random.default_rng(seed)
    d = rng.standard_normal((n, 3))
    return d / np.linalg.norm(d, axis=1, keepdims=True)


def test_dipole_closed_form():
    d = _dirs()
    axis = np.array([1.0, 0.0, 0.0])
    g = oa.gain("dipole", d, axis, 0.0, 0.0)
    psi = np.arccos(np.clip(d @ axis, -1, 1))
    np.testing.assert_allclose(g, np.cos(np.pi / 2 * np.cos(psi)) / np.sin(psi), rtol=1e-9)
    assert oa.gain("dipole", np.array([[0.0, 1.0, 0.0]]), axis, 0, 0)[0] == pytest.approx(1.0)
    fd = oa.gain("finite_dipole", d, axis, 0.5, 0.0)
    np.testing.assert_allclose(fd, np.abs(g), rtol=1e-9)


def test_array_factor_vs_phasor_sum():
    d = _dirs()
    axis, bore = np.array([0.0, 1.0, 0.0]), np.array([0.0, 0.0, -1.0])
    N, s = 6, 0.5
    g = oa.gain("array", d, np.stack([axis, bore]), float(N), [s, 1.0, 0.0, 0.0])
    u = d @ axis
    m = np.arange(N) - (N - 1) / 2
    af = np.abs(np.exp(2j * np.pi * s * u[:, None] * m).sum(1)) / N
    np.testing.assert_allclose(g, af, rtol=1e-7, atol=1e-12)
    gt = oa.gain("array_tapered", d, np.stack([axis, bore]), np.ones((2, N)),
                 [s, 1.0, 0.0, 0.0])
    np.testing.assert_allclose(gt, g, rtol=1e-9, atol=1e-12)
    q = 2.0
    ge = oa.gain("array", d, np.stack([axis, bore]), float(N), [s, 1.0, q, 1.0])
    cb = d @ bore
    ef = np.where(cb >= 0, np.clip(cb, 0, 1) ** (q / 2), 0.0)
    np.testing.assert_allclose(ge, af * ef, rtol=1e-7, atol=1e-12)


def test_tabulated_reproduces_dipole():
    th = np.linspace(0, np.pi, 20001)
    gains = np.cos(np.pi / 2 * np.cos(th)) / np.maximum(np.sin(th), 1e-12)
    gains[0] = gains[-1] = 0.0
    d = _dirs()
    axis = np.array([0.0, 0.0, -1.0])
    g = oa.gain("tabulated", d, axis, th, gains)
    ref = oa.gain("dipole", d, axis, 0.0, 0.0)
    np.testing.assert_allclose(g, ref, atol=1e-6)


def test_array_directivity_scale_half_wave():
    ant = pa.Antenna(kind="array", n_elements=4, spacing_lam=0.5, element_directivity_db=3.0103)
    assert pa.array_field_scale(ant) == pytest.approx(np.sqrt(2 * 4), rel=2e-3)


def test_pattern_args_frames():
    u_at = np.tile([1.0, 0.0, 0.0], (3, 1))
    u_ct = np.tile([0.0, -1.0, 0.0], (3, 1))
    k, pv, a, b = pa.pattern_args(pa.Antenna(kind="dipole"), u_at, u_ct)
    np.testing.assert_array_equal(pv, u_at)
    k, pv, a, b = pa.pattern_args(pa.Antenna(kind="array", n_elements=5), u_at, u_ct)
    assert pv.shape == (3, 2, 3)
    np.testing.assert_allclose(pv[0, 1], [0.0, 0.0, -1.0])    # boresight nadir
    roll = np.array([0.0, 0.1, -0.2])
    k, pv, a, b = pa.pattern_args(pa.Antenna(kind="tabulated", theta_deg=[0, 90],
                                             gain=[1, 0], roll_source="nav"), u_at, u_ct, roll)
    # positive roll (right wing down) tilts the nadir boresight to the left
    # of travel: +y for a +x heading
    np.testing.assert_allclose(pv[1], [0.0, np.sin(0.1), -np.cos(0.1)], atol=1e-12)
    assert pa.pattern_args(pa.Antenna(), u_at, u_ct) is None


def _psl_width(window, B=15e6):
    dt = 1.0 / (B * 40)
    p, m = ow.compressed_pulse(B, 6e-6, dt, window)
    a = np.abs(p)
    i = m
    while a[i + 1] < a[i]:
        i += 1
    psl = 20 * np.log10(a[i:].max())
    half = np.where(a >= 10 ** (-3 / 20))[0]
    width = (half[-1] - half[0] + 1) * dt * B
    return psl, width


@pytest.mark.parametrize("window,psl,width", [("none", -13.26, 0.886),
                                              ("hann", -31.47, 1.44),
                                              ("hamming", -42.7, 1.30)])
def test_harris_figures(window, psl, width):
    got_psl, got_w = _psl_width(window)
    assert got_psl == pytest.approx(psl, abs=0.6)
    assert got_w == pytest.approx(width, rel=0.04)


def test_pulse_peak_symmetry_and_products():
    for window in ("none", "hann", "hamming", "tukey"):
        p, m = ow.compressed_pulse(15e6, 20e-6, 1 / 48e6, window)
        assert p[m] == pytest.approx(1.0)
        np.testing.assert_allclose(p, p[::-1], atol=1e-12)
        q, mq = pw.compressed_pulse(15e6, 20e-6, 1 / 48e6, window)
        assert mq == m
        np.testing.assert_allclose(q, p, atol=1e-12)
    p, m = ow.compressed_pulse(15e6, 20e-6, 1 / 48e6, "hann", "chirp")
    q, _ = pw.compressed_pulse(15e6, 20e-6, 1 / 48e6, "hann", "chirp")
    np.testing.assert_allclose(q, p, atol=1e-9)


def test_tukey_limits():
    B, dt = 15e6, 1 / 48e6
    rect, m = ow.compressed_pulse(B, 5e-6, dt, "none")
    t0 = ow.tukey_pulse(B, dt, m, alpha=1e-9)
    np.testing.assert_allclose(t0, rect, atol=2e-3)
    hann, _ = ow.compressed_pulse(B, 5e-6, dt, "hann")
    t1 = ow.tukey_pulse(B, dt, m, alpha=1.0)
    np.testing.assert_allclose(t1, hann, atol=2e-3)


def test_fft_convolution_vs_direct():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((3, 300)) + 1j * rng.standard_normal((3, 300))
    p, m = ow.compressed_pulse(15e6, 1e-6, 1 / 48e6, "hann", "chirp")
    y = ow.convolve_fast_time(x, p, m)
    for r in range(3):
        ref = np.convolve(x[r], p)[m:m + 300]
        np.testing.assert_allclose(y[r], ref, atol=1e-10)


def test_chirp_mainlobe_matches_analytic():
    B, T, dt = 15e6, 20e-6, 1 / 48e6
    a, m = ow.compressed_pulse(B, T, dt, "hann")
    c, _ = ow.compressed_pulse(B, T, dt, "hann", "chirp")
    sl = slice(m - 3, m + 4)
    np.testing.assert_allclose(np.abs(c[sl]), np.abs(a[sl]), atol=0.02)
    assert np.abs(c[m]) == 1.0
