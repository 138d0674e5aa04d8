"""Host-side roughness set-up: series length and frozen speckle phasors.

* ``n_terms_for`` -- ``roughness.py:189-196``: ceil(x + 6 sqrt(x) + 5)
  clipped to [10, 300] for the run's largest (sigma K)^2.
* ``speckle_phasors`` -- ``roughness.py:199-204``: (e1 + i e2)/sqrt(2) from
  numpy ``default_rng(seed)`` (two standard-normal draws of length n).
* ``speckle_phasors_for_keys`` -- ``roughness.py:207-233``: stateless keyed
  splitmix64 + Box-Muller so overlapping crops share phasors.  The hash
  constants are the published splitmix64 finaliser (Steele, Lea & Flood
  2014) and must match the reference for identical realisations.

The D_Phi series itself runs on the device (csrc/ss_math.h).
"""

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_MIX1 = np.uint64(0xBF58476D1CE4E5B9)
_MIX2 = np.uint64(0x94D049BB133111EB)
_KEY_MUL = np.uint64(0xD2B74407B1CE6E93)
_STREAM2 = np.uint64(0xCA5A826395121157)


def n_terms_for(x_max):
    x = float(x_max)
    return int(np.clip(np.ceil(x + 6.0 * np.sqrt(x) + 5.0), 10, 300))


def speckle_phasors(n, seed):
    """Complex64 (n,) unit-variance phasors from default_rng(seed)."""
    rng = np.random.default_rng(seed)
    re = rng.standard_normal(n)
    im = rng.standard_normal(n)
    return ((re + 1j * im) / np.sqrt(2.0)).astype(np.complex64)


def _splitmix(x):
    with np.errstate(over="ignore"):
        x = x + _GOLDEN
        x = (x ^ (x >> np.uint64(30))) * _MIX1
        x = (x ^ (x >> np.uint64(27))) * _MIX2
    return x ^ (x >> np.uint64(31))


def speckle_phasors_for_keys(keys, seed):
    """Complex64 phasors keyed by (n, 2) integer lattice cells."""
    keys = np.asarray(keys, dtype=np.int64)
    if keys.ndim != 2 or keys.shape[1] != 2:
        raise ValueError("speckle keys must have shape (n, 2)")
    base = np.random.SeedSequence(seed).generate_state(1, dtype=np.uint64)[0]
    with np.errstate(over="ignore"):
        keyed = (base ^ keys[:, 0].astype(np.uint64)
                 ^ (keys[:, 1].astype(np.uint64) * _KEY_MUL))
    u1 = ((_splitmix(keyed) >> np.uint64(11)).astype(np.float64) + 0.5) / float(1 << 53)
    u2 = ((_splitmix(keyed ^ _STREAM2) >> np.uint64(11)).astype(np.float64)
          + 0.5) / float(1 << 53)
    rad = np.sqrt(-2.0 * np.log(u1))
    ang = 2.0 * np.pi * u2
    return (rad * (np.cos(ang) + 1j * np.sin(ang)) / np.sqrt(2.0)).astype(np.complex64)
