"""Along-track post-processing -- FP64 oracle, restating ``processing.py:143-348``.

* ``boxcar_weights(T, n, stride)``: the (n_out, T) 0/1 weight matrix of a
  boxcar of ``n`` traces advanced by ``stride`` -- presum / multilook
  (stride n, trailing remainder dropped, ``:162-165, 236-238``) and the
  unfocused moving window (stride 1, ``T - n + 1`` outputs, ``:208-210``);
* ``apply_weights``: ``tensordot(W, x)`` along slow time, the field summed,
  every other slow-time quantity averaged with W row-normalised
  (``:88-142``);
* ``focused_sar(field, s, twtt, c, k, aperture_m, window)``: straight-track
  time-domain backprojection through air (``:270-348``): for output trace i
  and bin b (r0 = c twtt / 2) the traces j with |s_j - s_i| <= L/2 are read
  at the migrated delay twtt + 2 dr / c, dr = sqrt(r0^2 + d^2) - r0, by
  linear complex interpolation (zero outside [0, nb - 1)), phase-corrected
  by exp(+2j k dr), weighted by the window (hann 0.5 (1 + cos(pi d / half))
  or rectangular) and summed.

Pinned by tests/test_oracle_processing.py: presum == every n-th unfocused
output; linearity; the coherent gain n of a constant field and the sqrt(n)
growth of independent noise; the multilook speckle contrast 1/sqrt(n) of
exponential intensities; a simulated point target (exact migrated phase
history, band-limited range pulse) focuses at its own trace and range with
coherent gain within 5 % of sum(window) and an azimuth -3 dB width of
0.886 lambda r / (2 L) (rectangular) / 1.44 lambda r / (2 L) (hann; the
windows' -3 dB widths in bins, Harris 1978), while the unfocused sum of the
same history does not.
"""

import numpy as np


def boxcar_weights(T, n, stride):
    n_out = (T // n) if stride == n else (T - n + 1)
    W = np.zeros((n_out, T))
    for k in range(n_out):
        W[k, k * stride:k * stride + n] = 1.0
    return W


def apply_weights(W, x, average=False):
    """Sum (or, with ``average``, the row-normalised mean) of ``x`` along its
    first (slow-time) axis with the weight matrix W."""
    w = W / W.sum(1, keepdims=True) if average else W
    return np.tensordot(w, np.asarray(x), axes=([1], [0]))


def along_track(positions):
    """Arc length along the track, s[0] = 0 (``processing.py:63-67``)."""
    pos = np.asarray(positions, np.float64)
    seg = np.linalg.norm(np.diff(pos, axis=0), axis=1)
    return np.concatenate([[0.0], np.cumsum(seg)])


def focused_sar(field, s, twtt, *, c, k, aperture_m, window="hann"):
    """(T, nb) complex128 field, along-track arc length s (T,), bin delays
    twtt (nb,) -> (T, nb) complex128 focused field; t0 = twtt[0] and dt =
    twtt[1] - twtt[0] as the reference takes them (``:303-306``)."""
    F = np.asarray(field, np.complex128)
    T, nb = F.shape
    twtt = np.asarray(twtt, np.float64)
    dt = float(twtt[1] - twtt[0])
    t0 = float(twtt[0])
    r0 = c * twtt / 2.0
    half = aperture_m / 2.0
    if window == "hann":
        def win(d):
            return 0.5 * (1.0 + np.cos(np.pi * d / half))
    elif window in ("none", "rect", None):
        def win(d):
            return np.ones_like(d)
    else:
        raise ValueError(f"unknown window {window!r}")
    out = np.zeros((T, nb), np.complex128)
    for i in range(T):
        d = s - s[i]
        J = np.nonzero(np.abs(d) <= half)[0]
        if J.size == 0:
            continue
        dJ = d[J][:, None]
        dr = np.sqrt(r0[None, :] ** 2 + dJ ** 2) - r0[None, :]
        tau = twtt[None, :] + 2.0 * dr / c
        x = (tau - t0) / dt
        i0 = np.floor(x).astype(np.int64)
        frac = x - i0
        valid = (i0 >= 0) & (i0 < nb - 1)
        i0c = np.clip(i0, 0, nb - 2)
        rows = J[:, None]
        samp = F[rows, i0c] * (1.0 - frac) + F[rows, i0c + 1] * frac
        samp = np.where(valid, samp, 0.0)
        w = win(d[J])[:, None]
        out[i] = np.sum(w * samp * np.exp(2j * k * dr), axis=0)
    return out
