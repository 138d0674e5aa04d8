"""Brute-force physical-optics referees (FP64) used to PIN the oracle.

Restates ``compare/brute_force.py:36-127`` (direct sub-wavelength summation of
(k/2pi) gamma cos dA e^{-2jkr}/r^2, image-method plate Gamma e^{-2jkh}/(2h))
and the two-media referee of ``compare/brute_force_layered.py:149-182``
(exact flat-interface Fermat crossing per sample, same amplitude convention
as the refracted kernel).
"""

import numpy as np

from .physics import fresnel_te, refraction_spreading
from .refraction import snell_crossing

TWO_PI = 2.0 * np.pi


def contributions(p, pts, nrm, dA, k, gamma):
    d = np.asarray(p, np.float64) - np.asarray(pts, np.float64)
    r = np.sqrt(np.sum(d * d, axis=-1))
    cos = np.sum(d * nrm, axis=-1) / r
    amp = (k / TWO_PI) * gamma * cos * dA / (r * r)
    return 1j * amp * np.exp(-2j * k * r), r


def flat_plate_field(k, h, gamma):
    return gamma * np.exp(-2j * k * h) / (2.0 * h)


def facet_samples(center, e1, e2, spacing):
    """Cell-centred sub-samples of one parallelogram facet."""
    e1, e2 = np.asarray(e1, float), np.asarray(e2, float)
    n1 = max(1, int(np.ceil(np.linalg.norm(e1) / spacing)))
    n2 = max(1, int(np.ceil(np.linalg.norm(e2) / spacing)))
    u = (np.arange(n1) + 0.5) / n1 - 0.5
    v = (np.arange(n2) + 0.5) / n2 - 0.5
    U, V = np.meshgrid(u, v, indexing="ij")
    off = U.ravel()[:, None] * e1 + V.ravel()[:, None] * e2
    raw = np.cross(e1, e2)
    area = np.linalg.norm(raw)
    return np.asarray(center, float)[None, :] + off, raw / area, area / (n1 * n2)


def facet_field(p, center, e1, e2, k, gamma, spacing):
    pts, nrm, dA = facet_samples(center, e1, e2, spacing)
    c, _ = contributions(p, pts, np.tile(nrm, (len(pts), 1)),
                         np.full(len(pts), dA), k, gamma)
    return c.sum()


def two_media_field(p, pts, dA, n1, n2, eps1, eps2, k0, gamma_b, att=None):
    """Field at p from flat-bed samples below a flat z=0 interface: exact
    two-point Fermat crossing per sample, TE two-way transmission, ray-tube
    spreading with the flux factor, in-medium k_j prefactor."""
    x, th1, th2, s1, s2, res, ok = snell_crossing(
        p, pts, np.zeros(3), np.array([0.0, 0.0, 1.0]), n1, n2)
    c1, c2 = np.cos(th1), np.cos(th2)
    g = fresnel_te(eps1, eps2, c1)[0]
    tau2 = 1.0 - g * g
    l_par, l_perp = refraction_spreading(np.stack([s1, s2], -1),
                                         np.stack([c1, c2], -1), [n1, n2])
    flux = (n1 * c1) / (n2 * c2)
    opl = n1 * s1 + n2 * s2
    amp = ((k0 * n2 / TWO_PI) * gamma_b * c2 * dA * tau2 * flux
           / (l_par * l_perp))
    if att is not None:
        amp = amp * 10.0 ** (-(s2 * att / 1000.0) / 10.0)
    return np.sum(np.where(ok, 1j * amp * np.exp(-2j * k0 * opl), 0.0)), opl
