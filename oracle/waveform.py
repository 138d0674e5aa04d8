"""Range compression (FP64 oracle), restating ``waveform.py:76-160``.

``compressed_pulse``: analytic raised-cosine-windowed sinc
[a sinc(x) + (1-a)/2 (sinc(x-1) + sinc(x+1))]/a, x = B dt m, m in [-M, M],
M = ceil(T/dt) (``:79-101``), or the explicit LFM matched filter built at an
internal rate >= 2B, correlated against its windowed conjugate, peak-
normalised and decimated to dt (``:104-123``).  ``window="tukey"`` (alpha =
0.2, a Tukey pedestal across the compressed band) is an extension the
reference does not have; its analytic form is the inverse Fourier transform
of the Tukey weighting sampled numerically (see ``tukey_pulse``).
``convolve_fast_time``: complex128 FFT "same" convolution along axis 1
(``:126-138``).

Pinned by tests/test_oracle_waveform.py: Harris (1978) peak sidelobe and
-3 dB width figures for none/hann/hamming, p(0) = 1 and symmetry, FFT
convolution vs direct ``np.convolve``, chirp vs analytic mainlobe agreement,
Tukey(alpha -> 0) = rectangular and Tukey(alpha = 1) = Hann.
"""

import numpy as np

WINDOW_A = {"none": 1.0, "hann": 0.5, "hamming": 0.54}


def tukey_weight(f, bandwidth, alpha):
    """Tukey(alpha) weighting across |f| <= B/2 (flat centre, cosine tapers
    over the outer alpha/2 fraction on each side)."""
    u = np.abs(np.asarray(f, np.float64)) / (bandwidth / 2.0)  # 0..1 in band
    w = np.ones_like(u)
    if alpha > 0:
        edge = u > 1.0 - alpha
        w[edge] = 0.5 * (1.0 + np.cos(np.pi * (u[edge] - 1.0 + alpha) / alpha))
    w[u > 1.0] = 0.0
    return w


def tukey_pulse(bandwidth, dt, m, alpha=0.2, n_quad=4097):
    """Inverse Fourier transform of the Tukey band weighting at tau = m dt,
    peak-normalised (midpoint quadrature over the band, float64)."""
    f = (np.arange(n_quad) + 0.5) / n_quad * bandwidth - bandwidth / 2.0
    w = tukey_weight(f, bandwidth, alpha)
    tau = np.arange(-m, m + 1) * dt
    p = (w[None, :] * np.cos(2.0 * np.pi * f[None, :] * tau[:, None])).sum(1)
    return p / p[m]


def compressed_pulse(bandwidth, pulse_length, dt, window="hann",
                     construction="analytic", tukey_alpha=0.2):
    m = int(np.ceil(pulse_length / dt))
    if construction == "chirp":
        return chirp_pulse(bandwidth, pulse_length, dt, window, tukey_alpha)
    if window == "tukey":
        return tukey_pulse(bandwidth, dt, m, tukey_alpha), m
    a = WINDOW_A[window]
    x = bandwidth * dt * np.arange(-m, m + 1)
    p = (a * np.sinc(x) + 0.5 * (1.0 - a) * (np.sinc(x - 1.0)
                                             + np.sinc(x + 1.0))) / a
    return p, m


def chirp_pulse(bandwidth, pulse_length, dt, window, tukey_alpha=0.2):
    """Explicit matched filter (``waveform.py:104-123``), direct correlation."""
    m = int(np.ceil(pulse_length / dt))
    over = max(1, int(np.ceil(2.0 * bandwidth * dt)))
    dts = dt / over
    n = int(np.ceil(pulse_length / dts)) | 1
    t = (np.arange(n) - (n - 1) / 2) * dts
    s = np.exp(1j * np.pi * (bandwidth / pulse_length) * t * t)
    if window == "tukey":
        # Tukey taper in time maps onto the band for a large-TB chirp
        w = tukey_weight(t * (bandwidth / pulse_length), bandwidth, tukey_alpha)
    else:
        a = WINDOW_A[window]
        w = a + (1.0 - a) * np.cos(2.0 * np.pi * t / pulse_length)
    h = np.conj(s[::-1]) * w
    y = np.convolve(s, h)                       # full, lag 0 at n - 1
    y = y / y[n - 1]
    p = np.zeros(2 * m + 1, np.complex128)
    j = np.arange(-m, m + 1)
    idx = n - 1 + j * over
    ok = (idx >= 0) & (idx < len(y)) & (np.abs(j * dt) <= pulse_length)
    p[ok] = y[idx[ok]]
    p[m] = 1.0
    return p, m


def convolve_fast_time(arr, kernel, center):
    """'Same'-mode FFT convolution along axis 1 (``waveform.py:126-138``)."""
    arr = np.asarray(arr)
    n, m = arr.shape[1], len(kernel)
    nfft = 1 << (n + m - 2).bit_length()
    kf = np.fft.fft(kernel, nfft).reshape((1, nfft) + (1,) * (arr.ndim - 2))
    y = np.fft.ifft(np.fft.fft(arr, nfft, axis=1) * kf, axis=1)
    return y[:, center:center + n]


def apply_waveform(out, bandwidth, pulse_length, dt, window="hann",
                   construction="analytic", kind="chirp"):
    """Coherent-mode waveform application, complex128 result."""
    if kind == "delta":
        return np.asarray(out, np.complex128)
    p, m = compressed_pulse(bandwidth, pulse_length, dt, window, construction)
    return convolve_fast_time(np.asarray(out, np.complex128), p, m)
