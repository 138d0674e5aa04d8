"""Host-side antenna pattern set-up (per-trace frames and kernel parameters).

Conventions follow ``antenna.py:1-32`` of the reference: ``g`` is the one-way
FIELD gain at the departure direction; coherent kernels weight fields by
g**2.  ``pattern_args`` returns ``None`` (isotropic) or ``(kind, pv, pa, pb)``
with the same meaning as ``antenna.py:76-118``:

  dipole / finite_dipole: pv = dipole axis (T, 3); pa = length_lam (finite)
  array:         pv = (T, 2, 3) [cross-track axis, boresight]; pa = n_elements;
                 pb = [spacing_lam, field_scale, q, element_enabled]
  array_tapered: pa = (2, n) [tx; rx] weights, pb as for arrays
  tabulated:     pv = nadir boresight; pa = theta samples (rad, ascending),
                 pb = field gains

The gain itself is evaluated on the device (csrc/ss_math.h antenna_gain).
"""

import functools
from dataclasses import dataclass, field

import numpy as np

NADIR = np.array([0.0, 0.0, -1.0])


@dataclass
class Antenna:
    """Antenna description (the fields of the reference's AntennaConfig)."""

    kind: str = "isotropic"
    axis: str = "along_track"
    n_elements: int = 5
    spacing_lam: float = 0.5
    element_directivity_db: float = None
    tx_weights: list = None
    rx_weights: list = None
    length_lam: float = 0.5
    theta_deg: list = None
    gain: list = None
    roll_source: str = "none"


def rodrigues(v, k, ang):
    """Right-handed rotation of rows of ``v`` by ``ang`` about unit rows ``k``."""
    ang = np.asarray(ang, np.float64)[..., None]
    c, s = np.cos(ang), np.sin(ang)
    return v * c + np.cross(k, v) * s + k * np.sum(k * v, -1, keepdims=True) * (1.0 - c)


def frame_vectors(ant, u_at, u_ct, roll=None):
    u_at = np.asarray(u_at, np.float64)
    u_ct = np.asarray(u_ct, np.float64)
    if ant.kind in ("dipole", "finite_dipole"):
        v = u_at if ant.axis == "along_track" else u_ct
    elif ant.kind in ("array", "array_tapered"):
        v = u_ct
    elif ant.kind == "tabulated":
        v = np.broadcast_to(NADIR, u_at.shape)
    else:
        raise ValueError(f"no pattern vector for kind {ant.kind!r}")
    if ant.roll_source == "nav" and roll is not None:
        phi = np.nan_to_num(np.asarray(roll, np.float64), nan=0.0)
        v = rodrigues(np.ascontiguousarray(v), u_at, phi)
    return np.ascontiguousarray(v)


def element_power_exponent(directivity_db):
    """q of a cos^q forward-hemisphere element with directivity 2(q + 1)."""
    if directivity_db is None:
        return 0.0
    return max(10.0 ** (directivity_db / 10.0) / 2.0 - 1.0, 0.0)


def _line_array_power(u, cos_bore, w, spacing_lam, q):
    m = np.arange(len(w)) - (len(w) - 1) / 2.0
    ph = (2.0 * np.pi * spacing_lam) * u[..., None] * m
    af = np.hypot(np.cos(ph) @ w, np.sin(ph) @ w) / w.sum()
    c = np.clip(cos_bore, 0.0, 1.0)
    ef = np.where(cos_bore >= 0.0, c ** (q / 2.0), 0.0)
    return (ef * af) ** 2


def _side_directivity(w, spacing_lam, q, n=720):
    """Peak directivity of one side's pattern, midpoint rule on the sphere."""
    th = (np.arange(n) + 0.5) * (np.pi / n)
    ph = (np.arange(n) + 0.5) * (2.0 * np.pi / n)
    st, ct = np.sin(th)[:, None], np.cos(th)[:, None]
    u = st * np.cos(ph)[None, :]
    P = _line_array_power(u, np.broadcast_to(ct, u.shape), np.asarray(w, np.float64),
                          spacing_lam, q)
    dom = st * (np.pi / n) * (2.0 * np.pi / n)
    return 4.0 * np.pi * P.max() / np.sum(P * dom)


@functools.lru_cache(maxsize=64)
def _directivity_scale(tx, rx, spacing_lam, directivity_db):
    q = element_power_exponent(directivity_db)
    d_tx = _side_directivity(tx, spacing_lam, q)
    d_rx = d_tx if rx == tx else _side_directivity(rx, spacing_lam, q)
    return float((d_tx * d_rx) ** 0.25)


def array_field_scale(ant):
    if ant.element_directivity_db is None:
        return 1.0
    if ant.kind == "array":
        tx = rx = (1.0,) * ant.n_elements
    else:
        tx = tuple(float(v) for v in ant.tx_weights)
        rx = tuple(float(v) for v in ant.rx_weights)
    return _directivity_scale(tx, rx, float(ant.spacing_lam),
                              float(ant.element_directivity_db))


def pattern_args(ant, u_at, u_ct, roll=None):
    if ant is None or ant.kind == "isotropic":
        return None
    pv = frame_vectors(ant, u_at, u_ct, roll)
    ctl = [ant.spacing_lam, None, element_power_exponent(ant.element_directivity_db),
           float(ant.element_directivity_db is not None)]
    if ant.kind in ("array", "array_tapered"):
        bore = np.cross(np.asarray(u_at, np.float64), pv)
        pv = np.stack([pv, bore], axis=1)
        ctl[1] = array_field_scale(ant)
        pb = np.asarray(ctl, np.float64)
        if ant.kind == "array":
            pa = np.float64(ant.n_elements)
        else:
            pa = np.asarray([ant.tx_weights, ant.rx_weights], np.float64)
    elif ant.kind == "finite_dipole":
        pa, pb = np.float64(ant.length_lam), np.float64(0.0)
    elif ant.kind == "tabulated":
        pa = np.deg2rad(np.asarray(ant.theta_deg, np.float64))
        pb = np.asarray(ant.gain, np.float64)
    else:
        pa = pb = np.float64(0.0)
    return ant.kind, pv, pa, pb


def cos_power_table(power_exponent=2.0, n=181):
    """Tabulated pattern whose POWER is cos^p(theta) off nadir (field gain
    cos^(p/2), zero beyond 90 deg): the 'cos^2' pattern as a tabulated kind."""
    th = np.linspace(0.0, 90.0, n)
    g = np.cos(np.deg2rad(th)) ** (power_exponent / 2.0)
    g[-1] = 0.0
    th = np.concatenate([th, [180.0]])
    g = np.concatenate([g, [0.0]])
    return Antenna(kind="tabulated", theta_deg=list(th), gain=list(g))
