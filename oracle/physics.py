"""Scalar wave physics (FP64), restating ``physics.py``.

Pinned by tests/test_oracle_physics.py: normal-incidence reduction, R + T = 1,
Stokes reciprocity, TIR flag (``tests/test_fresnel_spreading.py:26-91``), and
finite-difference ray-tube widths for the spreading lengths (``:96-120``).
"""

import numpy as np

C = 299792458.0


def fresnel_normal(eps1, eps2):
    """(n1 - n2)/(n1 + n2) (``physics.py:16-21``)."""
    n1, n2 = np.sqrt(eps1), np.sqrt(eps2)
    return (n1 - n2) / (n1 + n2)


def fresnel_te(eps1, eps2, c1):
    """TE gamma, tau, cos(theta2), tir (``physics.py:33-67``).

    Evaluation order is the one the kernels use: sin2^2 = (eps1/eps2) *
    (1 - c1*c1); c2 = sqrt(max(1 - sin2^2, 0)); gamma = (n1 c1 - n2 c2) /
    (n1 c1 + n2 c2)."""
    eps1 = np.asarray(eps1, np.float64)
    eps2 = np.asarray(eps2, np.float64)
    c1 = np.asarray(c1, np.float64)
    n1, n2 = np.sqrt(eps1), np.sqrt(eps2)
    sin2_sq = (eps1 / eps2) * (1.0 - c1 * c1)
    tir = sin2_sq > 1.0
    c2 = np.sqrt(np.maximum(1.0 - sin2_sq, 0.0))
    gamma = (n1 * c1 - n2 * c2) / (n1 * c1 + n2 * c2)
    return gamma, 1.0 + gamma, c2, tir


def refraction_spreading(lengths, cos_leg, n):
    """(L_par, L_perp) of an astigmatic ray tube (``physics.py:109-150``)."""
    s, c, n = np.broadcast_arrays(np.asarray(lengths, np.float64),
                                  np.asarray(cos_leg, np.float64),
                                  np.asarray(n, np.float64))
    n0, c0 = n[..., 0], c[..., 0]
    l_perp = n0 * np.sum(s / n, axis=-1)
    l_par = n0 * c0 * c0 * np.sum(s / (n * c * c), axis=-1)
    return l_par, l_perp
