"""Seeded synthetic height fields (FP64), restating ``compare/haynes.py:94-112``.

TEST INFRASTRUCTURE ONLY: tests/ import this as the checker of the product's
host generators (``scene.gaussian_filter`` / ``gaussian_surface_keyed``) and
of the device generator (``ss_gaussian_heights``); nothing in the package
imports it.

``gaussian_field(w, dx, corr_len)``: the reference's Gaussian-ACF filter --
circular FFT convolution of white noise w with g(rho) = exp(-rho^2/a^2),
a = corr_len/sqrt(2), rho the minimal-image (periodic) distance on a grid of
spacing dx, divided by sqrt(sum g^2) so the periodic field has unit
variance and correlation exp(-rho^2/corr_len^2).  The reference draws w with
``rng.standard_normal``; ``keyed_noise`` gives the keyed white noise the
device draws instead: the real part x sqrt 2 of the keyed complex normal of
``roughness.py:207-233`` with key (i, j) (``oracle.roughness``'s scalar
integer splitmix64).

Pinned by tests/test_oracle_scene.py: the response to a unit impulse is the
normalised kernel g/sqrt(sum g^2) itself (normalisation, distance
convention, a = l/sqrt 2); a brute-force circular double sum on small
grids; the field's exact autocorrelation (from the impulse response)
against the continuous closed form exp(-rho^2/l^2); keyed noise against the
scalar splitmix64 restatement, unit variance and normality.
"""

import math

import numpy as np

from .roughness import _splitmix_scalar


def gaussian_kernel(n, m, dx, corr_len):
    """(n, m) g = exp(-rho^2/a^2) on minimal-image distances."""
    cy = np.arange(n, dtype=np.float64)
    cy = np.minimum(cy, n - cy) * dx
    cx = np.arange(m, dtype=np.float64)
    cx = np.minimum(cx, m - cx) * dx
    a = corr_len / np.sqrt(2.0)
    return np.exp(-(cy[:, None] ** 2 + cx[None, :] ** 2) / (a * a))


def gaussian_field(w, dx, corr_len):
    """Unit-variance Gaussian-ACF field of white noise ``w`` (n, m)."""
    w = np.asarray(w, np.float64)
    n, m = w.shape
    g = gaussian_kernel(n, m, dx, corr_len)
    z = np.fft.irfft2(np.fft.rfft2(w) * np.fft.rfft2(g), s=(n, m))
    return z / np.sqrt((g * g).sum())


def gaussian_field_brute(w, dx, corr_len):
    """The same circular convolution as a direct double sum (small grids)."""
    w = np.asarray(w, np.float64)
    n, m = w.shape
    g = gaussian_kernel(n, m, dx, corr_len)
    z = np.zeros((n, m))
    for i in range(n):
        for j in range(m):
            z += w[i, j] * np.roll(np.roll(g, i, axis=0), j, axis=1)
    return z / np.sqrt((g * g).sum())


def keyed_noise(rows, cols, seed):
    """(rows, cols) keyed standard normals, scalar integer arithmetic (small
    grids only)."""
    base = int(np.random.SeedSequence(seed).generate_state(1, dtype=np.uint64)[0])
    mask = (1 << 64) - 1
    out = np.empty((rows, cols))
    for i in range(rows):
        for j in range(cols):
            keyed = base ^ (i & mask) ^ (((j & mask) * 0xD2B74407B1CE6E93) & mask)
            u1 = (float(_splitmix_scalar(keyed) >> 11) + 0.5) * (1.0 / (1 << 53))
            u2 = (float(_splitmix_scalar(keyed ^ 0xCA5A826395121157) >> 11) + 0.5) * (
                1.0 / (1 << 53))
            out[i, j] = math.sqrt(-2.0 * math.log(u1)) * math.cos(2.0 * math.pi * u2)
    return out
