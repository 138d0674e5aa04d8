"""In-kernel antenna gain (FP64), restating ``antenna.py:138-258``.

``gain(kind, dhat, pv, pa, pb)`` is the one-way FIELD gain at departure
direction ``dhat`` (antenna -> facet, (..., 3)); kernels weight fields by
g**2.  ``pv`` is the trace's pattern vector ((3,) or (2, 3) for arrays).

Pinned by tests/test_oracle_antenna.py: half-wave dipole closed form and its
null/broadside values, finite dipole with L = 0.5 lambda equal to the dipole,
uniform-array factor vs the explicit phasor sum over elements, tapered
array with equal weights equal to the uniform array, tabulated table of the
dipole reproducing the dipole (``tests/test_antenna.py:119-294``).
"""

import numpy as np


def element_field_pattern(cos_theta, power_exponent, enabled=True):
    """Forward-hemisphere cosine-power element (``antenna.py:138-143``)."""
    c = np.clip(cos_theta, 0.0, 1.0)
    forward = np.where(cos_theta >= 0.0, c ** (power_exponent / 2.0), 0.0)
    return np.where(enabled, forward, np.ones_like(c))


def _dot(a, b):
    return (a[..., 0] * b[..., 0] + a[..., 1] * b[..., 1]) + a[..., 2] * b[..., 2]


def gain(kind, dhat, pv, pa, pb):
    dhat = np.asarray(dhat, np.float64)
    pv = np.asarray(pv, np.float64)
    if kind == "isotropic":
        return np.ones(dhat.shape[:-1])
    if kind == "dipole":
        ca = _dot(dhat, pv)
        s = np.sqrt(np.maximum(1.0 - ca * ca, 1e-12))
        return np.cos((np.pi / 2.0) * ca) / s
    if kind == "finite_dipole":
        kh = np.pi * float(pa)
        ca = _dot(dhat, pv)
        s = np.sqrt(np.maximum(1.0 - ca * ca, 1e-12))
        return np.abs(np.cos(kh * ca) - np.cos(kh)) / (s * (1.0 - np.cos(kh)))
    if kind == "array":
        pa = float(pa)
        pb = np.asarray(pb, np.float64)
        u = _dot(dhat, pv[0])
        ct = _dot(dhat, pv[1])
        x = np.pi * pb[0] * u
        sx = np.sin(x)
        near = np.abs(sx) < 1e-5
        with np.errstate(divide="ignore", invalid="ignore"):
            af = np.where(near, np.cos(pa * x) / np.cos(x),
                          np.sin(pa * x) / (pa * np.where(near, 1.0, sx)))
        ef = element_field_pattern(ct, pb[2], bool(pb[3]))
        return pb[1] * ef * np.abs(af)
    if kind == "array_tapered":
        pa = np.asarray(pa, np.float64)
        pb = np.asarray(pb, np.float64)
        n = pa.shape[1]
        m = np.arange(n) - (n - 1) / 2.0
        u = _dot(dhat, pv[0])
        ct = _dot(dhat, pv[1])
        ph = (2.0 * np.pi * pb[0]) * u[..., None] * m
        c, s = np.cos(ph), np.sin(ph)

        def af(w):
            return np.sqrt((c @ w) ** 2 + (s @ w) ** 2) / np.sum(w)

        ef = element_field_pattern(ct, pb[2], bool(pb[3]))
        return pb[1] * ef * np.sqrt(af(pa[0]) * af(pa[1]))
    if kind == "tabulated":
        ct = np.clip(_dot(dhat, pv), -1.0, 1.0)
        return np.interp(np.arccos(ct), np.asarray(pa, np.float64),
                         np.asarray(pb, np.float64))
    raise ValueError(f"unknown pattern kind {kind!r}")
