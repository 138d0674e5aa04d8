"""Scene data model on the host: facets, tracks and seeded synthetic surfaces.

Facets follow the reference's record (``scene.py:80-101``): per rectangular
mean-plane facet a centre, a unit +z-oriented normal, the parallelogram area
|e1 x e2| and the two mean edge vectors, plus the source grid cell and grid
shape (the refracted kernel's crossing lookup uses both).  Coordinates are a
local frame in metres (east, north, up): a planar lattice here, or the
WGS84 ENU frame of a projected DEM built by ``geodesy.build_facets``.

Corner math restates ``scene.py:158-176`` / ``compare/brute_force_layered.py
:44-71``: from the (ny+1, nx+1) vertex lattice V, e1 = ((v01 + v11) - (v00 +
v10)) / 2, e2 = ((v10 + v11) - (v00 + v01)) / 2, centre = mean of the four
corners, normal = (e1 x e2)/|e1 x e2| flipped to +z.

The Gaussian-ACF height field restates ``compare/haynes.py:94-112`` (circular
FFT convolution of white noise with exp(-rho^2/a^2), a = l/sqrt(2),
normalised by sqrt(sum g^2) so the periodic field has unit variance), with
white noise from a numpy Generator (as the reference) or keyed by vertex
(``SeededLattice``, drawn on the device by ``ss_gaussian_heights``).
"""

import warnings
from dataclasses import dataclass, field

import numpy as np

from .roughness import keyed_normals


@dataclass
class Facets:
    """Rectangular facets (float64, (N, 3) arrays as in the reference)."""

    centers: np.ndarray
    normals: np.ndarray
    areas: np.ndarray
    e1: np.ndarray
    e2: np.ndarray
    cell: np.ndarray = None          # (N, 2) int (row, col) source cell
    grid_shape: tuple = None         # (ny, nx) cells
    phase_keys: np.ndarray = None    # (N, 2) stable lattice keys

    def __len__(self):
        return int(self.centers.shape[0])

    def subset(self, idx):
        """Facets at integer/boolean index ``idx`` (grid metadata dropped)."""
        return Facets(self.centers[idx], self.normals[idx], self.areas[idx],
                      self.e1[idx], self.e2[idx])


@dataclass
class NavTrack:
    """Per-trace platform geometry (``nav.py:9-30``)."""

    positions: np.ndarray   # (T, 3)
    u_at: np.ndarray        # (T, 3) unit along-track
    u_ct: np.ndarray        # (T, 3) unit cross-track, right of travel
    roll: np.ndarray = None  # (T,) radians or None


def facets_from_vertices(V):
    """Facets of a (ny+1, nx+1, 3) vertex lattice (corner math above)."""
    V = np.asarray(V, dtype=np.float64)
    v00, v01 = V[:-1, :-1], V[:-1, 1:]
    v10, v11 = V[1:, :-1], V[1:, 1:]
    e1 = ((v01 + v11) - (v00 + v10)) / 2.0
    e2 = ((v10 + v11) - (v00 + v01)) / 2.0
    centers = ((v00 + v01 + v10 + v11) / 4.0).reshape(-1, 3)
    raw = np.cross(e1, e2).reshape(-1, 3)
    mag = np.linalg.norm(raw, axis=1)
    normals = raw / mag[:, None]
    normals *= np.where(normals[:, 2] < 0.0, -1.0, 1.0)[:, None]
    ny, nx = V.shape[0] - 1, V.shape[1] - 1
    ci, cj = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    cell = np.stack([ci.ravel(), cj.ravel()], axis=1)
    return Facets(centers, normals, mag, e1.reshape(-1, 3), e2.reshape(-1, 3),
                  cell, (ny, nx), cell.astype(np.int64).copy())


def check_facet_size(facets, wavelength, min_range, beta=0.5):
    """Fresnel-zone validity of the linear-phase approximation
    (``scene.py:179-195`` of the reference): warns when the longest facet edge
    exceeds beta sqrt(wavelength min_range) (half a Fresnel-zone radius by
    default); returns the largest edge / limit ratio."""
    edge = np.maximum(np.linalg.norm(facets.e1, axis=1), np.linalg.norm(facets.e2, axis=1))
    limit = beta * np.sqrt(wavelength * min_range)
    ratio = float(edge.max() / limit)
    if ratio > 1.0:
        warnings.warn(f"max facet edge {edge.max():.1f} m exceeds the Fresnel-zone limit "
                      f"{limit:.1f} m (ratio {ratio:.2f}): the linear-phase approximation "
                      "may be invalid; subdivide the scene")
    return ratio


def grid_facets(heights, spacing, origin=(0.0, 0.0)):
    """Facets of a height lattice ``heights`` (ny+1, nx+1) at ``spacing`` m,
    centred on ``origin`` (x = columns/east, y = rows/north)."""
    h = np.asarray(heights, dtype=np.float64)
    ny1, nx1 = h.shape
    xs = (np.arange(nx1) - (nx1 - 1) / 2.0) * spacing + origin[0]
    ys = (np.arange(ny1) - (ny1 - 1) / 2.0) * spacing + origin[1]
    X, Y = np.meshgrid(xs, ys)
    return facets_from_vertices(np.stack([X, Y, h], axis=-1))


@dataclass
class Lattice:
    """A regular height lattice: ``heights`` (ny+1, nx+1) vertex heights at
    ``spacing`` m centred on ``origin`` (the input of ``grid_facets``).  The
    device builds its facets itself (``device.build_lattices``); ``facets()``
    is the host tessellation of the same lattice."""

    heights: np.ndarray
    spacing: float
    origin: tuple = (0.0, 0.0)

    @property
    def grid_shape(self):
        return (self.heights.shape[0] - 1, self.heights.shape[1] - 1)

    def __len__(self):
        ny, nx = self.grid_shape
        return ny * nx

    def horizontal(self):
        return bool(np.all(self.heights == self.heights.flat[0]))

    def facets(self):
        return grid_facets(self.heights, self.spacing, self.origin)

    def level(self):
        """Height of a horizontal lattice."""
        return float(self.heights.flat[0])


class SeededLattice(Lattice):
    """A (ny+1, nx+1) height lattice given by its generator instead of its
    heights: mean + sigma x the seeded Gaussian-ACF field of correlation
    length ``corr_len`` (``gaussian_surface_keyed``; sigma 0 is a horizontal
    plane at ``mean``).  The device draws it from these few numbers
    (``ss_gaussian_heights``); ``heights`` materialises the host
    restatement (FFT filter, equal to the device's direct convolution to
    rounding) on first use, or holds heights read back from the device
    (``device.lattice_heights``)."""

    def __init__(self, ny, nx, spacing, sigma=0.0, corr_len=1.0, seed=0, mean=0.0,
                 origin=(0.0, 0.0)):
        self.ny, self.nx = int(ny), int(nx)
        self.spacing = float(spacing)
        self.sigma, self.corr_len, self.seed = float(sigma), float(corr_len), seed
        self.mean = float(mean)
        self.origin = tuple(origin)
        self._heights = None

    @property
    def heights(self):
        if self._heights is None:
            if self.sigma == 0.0:
                self._heights = np.full((self.ny + 1, self.nx + 1), self.mean)
            else:
                self._heights = self.mean + self.sigma * gaussian_surface_keyed(
                    self.ny + 1, self.spacing, self.corr_len, self.seed, m=self.nx + 1)
        return self._heights

    @heights.setter
    def heights(self, h):
        h = np.asarray(h, np.float64)
        if h.shape != (self.ny + 1, self.nx + 1):
            raise ValueError(f"heights must have shape {(self.ny + 1, self.nx + 1)}")
        self._heights = h

    @property
    def grid_shape(self):
        return (self.ny, self.nx)

    def __len__(self):
        return self.ny * self.nx

    def horizontal(self):
        return self.sigma == 0.0

    def level(self):
        return self.mean

    def __repr__(self):
        return (f"SeededLattice(ny={self.ny}, nx={self.nx}, spacing={self.spacing}, "
                f"sigma={self.sigma}, corr_len={self.corr_len}, seed={self.seed!r}, "
                f"mean={self.mean}, origin={self.origin})")

    __eq__ = object.__eq__
    __hash__ = object.__hash__


class LatticeFacets:
    """The facets of a lattice, tessellated on the host only when an
    attribute is read (``len`` needs none): a layer of a device-built scene
    carries one, so the host never builds facets it does not use."""

    def __init__(self, lattice):
        self.lattice = lattice
        self._f = None

    def __len__(self):
        return len(self.lattice)

    def __getattr__(self, name):
        if name.startswith("__"):
            raise AttributeError(name)
        if self._f is None:
            self._f = self.lattice.facets()
        return getattr(self._f, name)


def gaussian_filter(w, dx, corr_len):
    """The Gaussian-ACF filter of ``compare/haynes.py:94-112`` applied to a
    white-noise array ``w`` (n, m): circular FFT convolution with
    exp(-rho^2/a^2), a = l/sqrt(2), on minimal-image distances, divided by
    sqrt(sum g^2)."""
    n, m = w.shape
    cy = np.arange(n, dtype=np.float64)
    cy = np.minimum(cy, n - cy) * dx
    cx = np.arange(m, dtype=np.float64)
    cx = np.minimum(cx, m - cx) * dx
    r2 = cy[:, None] ** 2 + cx[None, :] ** 2
    a = corr_len / np.sqrt(2.0)
    g = np.exp(-r2 / (a * a))
    z = np.fft.irfft2(np.fft.rfft2(w) * np.fft.rfft2(g), s=(n, m))
    return z / np.sqrt((g * g).sum())


def gaussian_surface(n, dx, corr_len, rng, m=None):
    """(n, m) zero-mean unit-variance Gaussian-ACF field, C = exp(-rho^2/l^2)
    (``compare/haynes.py:94-112``, extended to rectangular grids): white
    noise from ``rng.standard_normal``."""
    m = n if m is None else m
    return gaussian_filter(rng.standard_normal((n, m)), dx, corr_len)


def gaussian_surface_keyed(n, dx, corr_len, seed, m=None):
    """As ``gaussian_surface`` with keyed white noise (``roughness.
    keyed_normals``): the realisation is a function of ``seed`` and the
    vertex indices alone, which is what ``ss_gaussian_heights`` draws on the
    device."""
    m = n if m is None else m
    return gaussian_filter(keyed_normals(n, m, seed), dx, corr_len)


def straight_track(n_traces, spacing, altitude, start=None, heading=(1.0, 0.0),
                   cross_offset=0.0):
    """Straight level track centred on the origin: (T, 3) positions, u_at
    from central differences and u_ct = u_at x up (``nav.py:21-30``)."""
    hd = np.asarray(heading, np.float64)
    hd = hd / np.linalg.norm(hd)
    s = (np.arange(n_traces) - (n_traces - 1) / 2.0) * spacing
    if start is not None:
        s = s - s[0] + start
    perp = np.array([hd[1], -hd[0]])
    xy = s[:, None] * hd[None, :] + cross_offset * perp[None, :]
    pos = np.column_stack([xy, np.full(n_traces, float(altitude))])
    if n_traces > 1:
        g = np.gradient(pos, axis=0)
    else:
        g = np.array([[hd[0], hd[1], 0.0]])
    u_at = g / np.linalg.norm(g, axis=1, keepdims=True)
    u_ct = np.cross(u_at, np.array([0.0, 0.0, 1.0]))
    u_ct = u_ct / np.linalg.norm(u_ct, axis=1, keepdims=True)
    return NavTrack(pos, u_at, u_ct)


def eps_kovacs(rho_kgm3):
    """Kovacs (1993) permittivity (1 + 0.845 rho[g/cc])^2 (``firn.py:19-21``)."""
    return (1.0 + 0.845 * np.asarray(rho_kgm3, np.float64) / 1000.0) ** 2


@dataclass
class LayeredScene:
    """Interfaces top-down plus media (one more medium than interfaces)."""

    interfaces: list
    eps: list
    att_db_per_km: list
    names: list = field(default_factory=list)
