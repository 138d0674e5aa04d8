"""Host-side roughness set-up: series length and frozen speckle phasors.

* ``n_terms_for`` -- ``roughness.py:189-196``: ceil(x + 6 sqrt(x) + 5)
  clipped to [10, 300] for the run's largest (sigma K)^2.
* ``speckle_phasors`` -- ``roughness.py:199-204``: (e1 + i e2)/sqrt(2) from
  numpy ``default_rng(seed)`` (two standard-normal draws of length n).
* ``speckle_phasors_for_keys`` -- ``roughness.py:207-233``: stateless keyed
  splitmix64 + Box-Muller so overlapping crops share phasors.  The hash
  constants are the published splitmix64 finaliser (Steele, Lea & Flood
  2014) and must match the reference for identical realisations.
* ``keyed_normals`` -- the same keying as white noise for the seeded
  Gaussian-ACF height generator (``scene.SeededLattice``), so the device
  can draw a lattice's noise from its seed alone.

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


def key_base(seed):
    """The 64-bit stream base of the keyed generators: numpy
    SeedSequence(seed).generate_state(1) (``roughness.py:211``)."""
    return np.random.SeedSequence(seed).generate_state(1, dtype=np.uint64)[0]


def _keyed_polar(kr, kc, base):
    """Box-Muller radius and angle of the keyed uniforms of lattice keys
    (kr, kc) (integer arrays of one shape) under stream ``base``."""
    with np.errstate(over="ignore"):
        keyed = (np.uint64(base) ^ np.asarray(kr, np.int64).astype(np.uint64)
                 ^ (np.asarray(kc, np.int64).astype(np.uint64) * _KEY_MUL))
    u1 = ((_splitmix(keyed) >> np.uint64(11)).astype(np.float64) + 0.5) / float(1 << 53)
    u2 = ((_splitmix(keyed ^ _STREAM2) >> np.uint64(11)).astype(np.float64)
          + 0.5) / float(1 << 53)
    return np.sqrt(-2.0 * np.log(u1)), 2.0 * np.pi * u2


def speckle_phasors_for_keys(keys, seed):
    """Complex64 phasors keyed by (n, 2) integer lattice cells."""
    keys = np.asarray(keys, dtype=np.int64)
    if keys.ndim != 2 or keys.shape[1] != 2:
        raise ValueError("speckle keys must have shape (n, 2)")
    rad, ang = _keyed_polar(keys[:, 0], keys[:, 1], key_base(seed))
    return (rad * (np.cos(ang) + 1j * np.sin(ang)) / np.sqrt(2.0)).astype(np.complex64)


def keyed_normals(rows, cols, seed):
    """(rows, cols) float64 standard normals keyed by lattice vertex (i, j):
    the real part x sqrt 2 of the keyed complex normal, rad cos(angle).  The
    white noise of the seeded Gaussian-ACF generator (``ss_gaussian_heights``
    draws the same values on the device)."""
    i, j = np.meshgrid(np.arange(rows, dtype=np.int64), np.arange(cols, dtype=np.int64),
                       indexing="ij")
    rad, ang = _keyed_polar(i, j, key_base(seed))
    return rad * np.cos(ang)


class KeyedPhasors:
    """The keyed speckle phasors of a lattice's (ny, nx) cells, cell (r, c)
    under key (key_origin[0] + r, key_origin[1] + c): what
    ``speckle_phasors_for_keys`` gives for a grid scene's ``phase_keys``,
    given by its seed so the device can draw it (``ss_keyed_phasors``).
    ``np.asarray`` materialises the host values in grid (row-major) order."""

    def __init__(self, grid_shape, seed, key_origin=(0, 0)):
        self.grid_shape = (int(grid_shape[0]), int(grid_shape[1]))
        self.seed = seed
        self.key_origin = (int(key_origin[0]), int(key_origin[1]))

    def __len__(self):
        return self.grid_shape[0] * self.grid_shape[1]

    def keys(self):
        ny, nx = self.grid_shape
        r, c = np.divmod(np.arange(ny * nx, dtype=np.int64), nx)
        return np.stack([r + self.key_origin[0], c + self.key_origin[1]], axis=1)

    def __array__(self, dtype=None, copy=None):
        ph = speckle_phasors_for_keys(self.keys(), self.seed)
        return ph if dtype is None else ph.astype(dtype)
