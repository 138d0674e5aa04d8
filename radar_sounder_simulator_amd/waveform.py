"""Pulse compression: host pulse construction, device fast-time convolution.

``compressed_pulse`` builds the sampled compressed pulse p(m dt), m in
[-M, M], M = ceil(T/dt) (``waveform.py:79-123``):

* ``construction="analytic"``: raised-cosine-weighted sinc
  [a sinc(x) + (1 - a)/2 (sinc(x - 1) + sinc(x + 1))]/a, x = B m dt, with
  a = 1 / 0.5 / 0.54 for none / hann / hamming; ``window="tukey"`` (an
  extension: Tukey(alpha) band weighting, alpha = 0.2 by default) is
  evaluated as the inverse Fourier integral of the weighting.
* ``construction="chirp"``: the explicit LFM matched filter at an internal
  rate >= 2B, correlated with its windowed conjugate replica, peak-
  normalised to exactly 1 and decimated onto the dt lattice.

``apply_waveform`` convolves the field along fast time on the GPU (direct
float64 summation, csrc compress_kernel) -- the reference's complex128 FFT
'same' convolution (``waveform.py:126-160``).
"""

from dataclasses import dataclass

import numpy as np
import torch

from . import device as _dv

WINDOW_A = {"none": 1.0, "hann": 0.5, "hamming": 0.54}


@dataclass
class Waveform:
    kind: str = "delta"            # "delta" | "chirp"
    bandwidth: float = None
    pulse_length: float = None
    window: str = "hann"           # none | hann | hamming | tukey
    construction: str = "analytic"
    interp_bins: bool = False
    incoherent_envelope: bool = False
    tukey_alpha: float = 0.2


def _tukey_band(f, bandwidth, alpha):
    u = np.abs(f) / (bandwidth / 2.0)
    w = np.ones_like(u)
    if alpha > 0:
        e = u > 1.0 - alpha
        w[e] = 0.5 * (1.0 + np.cos(np.pi * (u[e] - 1.0 + alpha) / alpha))
    w[u > 1.0] = 0.0
    return w


def compressed_pulse(bandwidth, pulse_length, dt, window="hann",
                     construction="analytic", tukey_alpha=0.2):
    """Sampled compressed pulse and its half-length M (peak p[M] = 1)."""
    m = int(np.ceil(pulse_length / dt))
    if construction == "chirp":
        return _chirp_pulse(bandwidth, pulse_length, dt, window, tukey_alpha), m
    if window == "tukey":
        nq = 4097
        f = (np.arange(nq) + 0.5) / nq * bandwidth - bandwidth / 2.0
        w = _tukey_band(f, bandwidth, tukey_alpha)
        tau = np.arange(-m, m + 1) * dt
        p = np.cos(2.0 * np.pi * np.outer(tau, f)) @ w
        return p / p[m], m
    a = WINDOW_A[window]
    x = bandwidth * dt * np.arange(-m, m + 1)
    p = (a * np.sinc(x) + 0.5 * (1.0 - a) * (np.sinc(x - 1.0) + np.sinc(x + 1.0))) / a
    return p, m


def _chirp_pulse(bandwidth, pulse_length, dt, window, tukey_alpha):
    m = int(np.ceil(pulse_length / dt))
    over = max(1, int(np.ceil(2.0 * bandwidth * dt)))
    dts = dt / over
    n = int(np.ceil(pulse_length / dts)) | 1
    t = (np.arange(n) - (n - 1) / 2) * dts
    s = np.exp(1j * np.pi * (bandwidth / pulse_length) * t * t)
    if window == "tukey":
        w = _tukey_band(t * (bandwidth / pulse_length), bandwidth, tukey_alpha)
    else:
        a = WINDOW_A[window]
        w = a + (1.0 - a) * np.cos(2.0 * np.pi * t / pulse_length)
    # correlation y[l] = sum_i s[i + l] conj(s[i]) w[i] via FFT, lag 0 at n-1
    h = np.conj(s[::-1]) * w
    nfft = 1 << (2 * n - 2).bit_length()
    y = np.fft.ifft(np.fft.fft(s, nfft) * np.fft.fft(h, nfft))[:2 * n - 1]
    y = y / y[n - 1]
    p = np.zeros(2 * m + 1, np.complex128)
    j = np.arange(-m, m + 1)
    idx = n - 1 + j * over
    ok = (idx >= 0) & (idx < len(y)) & (np.abs(j * dt) <= pulse_length)
    p[ok] = y[idx[ok]]
    p[m] = 1.0
    return p


def apply_waveform_device(field, waveform, dt, stream=None):
    """Convolve a (..., n, 2) float64 device field with the pulse (coherent)."""
    if waveform is None or waveform.kind == "delta":
        return field
    p, m = compressed_pulse(waveform.bandwidth, waveform.pulse_length, dt,
                            waveform.window, waveform.construction,
                            waveform.tukey_alpha)
    out, _ = _dv.range_compress(field, p, m, stream=stream)
    return out


def apply_waveform(out, radar, mode):
    """Reference-signature ``apply_waveform(out, radar, mode)``
    (``waveform.py:141-160``): ``radar`` has ``.waveform`` and ``.dt``.

    Coherent: the complex field (T, n[, ...]) convolved along fast time on
    the device in float64.  The result keeps the input's precision: a
    complex64 field (the reference's kernel output, and the mixed tier's)
    comes back complex64 as in the reference; a complex128 field (the FP64
    tier's) comes back complex128, so a float64 run stays float64 end to end.
    Incoherent: the |p|^2 power envelope only when
    ``waveform.incoherent_envelope`` is set (else unchanged), clamped at 0,
    float32 (float64 for a float64 input).  Delta waveforms return ``out``
    itself."""
    wf = radar.waveform
    if wf.kind == "delta":
        return out
    arr = np.asarray(out)
    if mode != "coherent" and not wf.incoherent_envelope:
        return out
    wide = arr.dtype in (np.complex128, np.float64)
    p, m = compressed_pulse(wf.bandwidth, wf.pulse_length, radar.dt, wf.window,
                            wf.construction, getattr(wf, "tukey_alpha", 0.2))
    if mode != "coherent":
        p = np.abs(p) ** 2
    # fast time (axis 1) to the last axis: rows = (T, trailing...)
    x = np.moveaxis(np.asarray(arr, np.complex128), 1, -1)
    shp = x.shape
    flat = np.ascontiguousarray(x.reshape(-1, shp[-1]))
    dev = torch.device("cuda", torch.cuda.current_device())
    t = torch.as_tensor(np.stack([flat.real, flat.imag], -1), device=dev)
    y, _ = _dv.range_compress(t, p, m)
    torch.cuda.synchronize(dev)
    yh = y.cpu().numpy()
    yc = (yh[..., 0] + 1j * yh[..., 1]).reshape(shp)
    yc = np.moveaxis(yc, -1, 1)
    if mode == "coherent":
        return yc if wide else yc.astype(np.complex64)
    y = np.maximum(yc.real, 0.0)
    return y if wide else y.astype(np.float32)
