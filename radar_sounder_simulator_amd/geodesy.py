"""WGS84 geodesy and projected-DEM scenes on the host (float64, NumPy).

The reference does every coordinate step with pyproj (``scene.py:1-176``,
``nav.py:21-30``, ``synthetic.py:1-268``); this module restates the
published formulas instead (no pyproj in this library):

* geodetic <-> ECEF on the WGS84 ellipsoid (a = 6378137 m, f =
  1/298.257223563): the closed form forward, and the inverse by fixed-point
  iteration on the latitude with the height taken as p cos(phi) + z sin(phi)
  - a sqrt(1 - e^2 sin^2 phi), which stays regular at every latitude;
* the NSIDC sea-ice polar stereographic north projection (EPSG:3413: WGS84,
  latitude of true scale 70 N, central meridian 45 W, no false origin) in
  Snyder (1987), *Map Projections -- A Working Manual*, eqs. 15-9, 14-15,
  21-33 to 21-38 (ellipsoidal polar stereographic with a standard
  parallel), the inverse latitude by iterating eq. 7-9, and the point scale
  factor k = rho / (a m) (eq. 21-32; the projection is conformal, so the
  meridional scale h equals k);
* the ENU ``LocalFrame`` anchored on the ellipsoid (``scene.py:25-78``);
* ``build_facets``: a projected DEM window tessellated into one mean-plane
  facet per cell, every vertex going DEM -> projected -> geodetic -> ECEF ->
  local ENU (``scene.py:126-176``);
* ``nav_to_frame`` (``nav.py:21-30``);
* the five synthetic verification scenes of ``synthetic.py:28-122`` (flat,
  tilted, hill, sinusoid, crater) and the multilayer ones of ``:124-252``,
  on the same grid, track and centre (75 N, 40 W).

The affine DEM transform is kept as its three numbers (x_min, y_max,
posting): pixel (col, row) -> (x_min + col posting, y_max - row posting).
"""

from dataclasses import dataclass, field

import numpy as np

from .scene import Facets, NavTrack, facets_from_vertices

A_WGS84 = 6378137.0
F_WGS84 = 1.0 / 298.257223563
E2_WGS84 = F_WGS84 * (2.0 - F_WGS84)
E_WGS84 = np.sqrt(E2_WGS84)

# EPSG:3413
PS_LAT_TS = 70.0
PS_LON0 = -45.0

CENTER_LATLON = (75.0, -40.0)


# ------------------------------------------------------------- ellipsoid
def geodetic_to_ecef(lat, lon, h):
    """(lat, lon) degrees, h metres above the ellipsoid -> ECEF x, y, z."""
    phi, lam = np.deg2rad(np.asarray(lat, np.float64)), np.deg2rad(np.asarray(lon, np.float64))
    h = np.asarray(h, np.float64)
    s = np.sin(phi)
    n = A_WGS84 / np.sqrt(1.0 - E2_WGS84 * s * s)
    cp = np.cos(phi)
    return ((n + h) * cp * np.cos(lam), (n + h) * cp * np.sin(lam),
            (n * (1.0 - E2_WGS84) + h) * s)


def ecef_to_geodetic(x, y, z, iters=8):
    """ECEF -> (lat deg, lon deg, h m)."""
    x, y, z = (np.asarray(v, np.float64) for v in (x, y, z))
    p = np.hypot(x, y)
    lam = np.arctan2(y, x)
    phi = np.arctan2(z, p * (1.0 - E2_WGS84))
    for _ in range(iters):
        s = np.sin(phi)
        n = A_WGS84 / np.sqrt(1.0 - E2_WGS84 * s * s)
        h = p * np.cos(phi) + z * s - A_WGS84 * np.sqrt(1.0 - E2_WGS84 * s * s)
        phi = np.arctan2(z, p * (1.0 - E2_WGS84 * n / (n + h)))
    s = np.sin(phi)
    h = p * np.cos(phi) + z * s - A_WGS84 * np.sqrt(1.0 - E2_WGS84 * s * s)
    return np.rad2deg(phi), np.rad2deg(lam), h


# ------------------------------------------------- polar stereographic (N)
def _t_of(phi):
    s = np.sin(phi)
    return np.tan(np.pi / 4.0 - phi / 2.0) / ((1.0 - E_WGS84 * s) / (1.0 + E_WGS84 * s)) ** (
        E_WGS84 / 2.0)


def _m_of(phi):
    s = np.sin(phi)
    return np.cos(phi) / np.sqrt(1.0 - E2_WGS84 * s * s)


_PHI_C = np.deg2rad(PS_LAT_TS)
_AMC_TC = A_WGS84 * _m_of(_PHI_C) / _t_of(_PHI_C)


def ps_forward(lon, lat):
    """EPSG:3413 (lon, lat) degrees -> projected (x, y) metres."""
    phi = np.deg2rad(np.asarray(lat, np.float64))
    dl = np.deg2rad(np.asarray(lon, np.float64) - PS_LON0)
    rho = _AMC_TC * _t_of(phi)
    return rho * np.sin(dl), -rho * np.cos(dl)


def ps_inverse(x, y, iters=12):
    """EPSG:3413 projected (x, y) metres -> (lon, lat) degrees."""
    x, y = np.asarray(x, np.float64), np.asarray(y, np.float64)
    rho = np.hypot(x, y)
    t = rho / _AMC_TC
    phi = np.pi / 2.0 - 2.0 * np.arctan(t)
    for _ in range(iters):
        s = np.sin(phi)
        phi = np.pi / 2.0 - 2.0 * np.arctan(
            t * ((1.0 - E_WGS84 * s) / (1.0 + E_WGS84 * s)) ** (E_WGS84 / 2.0))
    lon = PS_LON0 + np.rad2deg(np.arctan2(x, -y))
    return lon, np.rad2deg(phi)


def ps_scale(lon, lat):
    """Point scale factor of EPSG:3413 at (lon, lat) (meridional = parallel)."""
    phi = np.deg2rad(np.asarray(lat, np.float64))
    return _AMC_TC * _t_of(phi) / (A_WGS84 * _m_of(phi))


def projected_to_ecef(x, y, h):
    lon, lat = ps_inverse(x, y)
    return geodetic_to_ecef(lat, lon, h)


# ------------------------------------------------------------ local frame
class LocalFrame:
    """East-North-Up frame anchored at (lat0, lon0, h0) on the ellipsoid
    (``scene.py:25-78``): rows of R are the E, N, U unit vectors in ECEF and
    local = R (xyz - origin)."""

    def __init__(self, lat0, lon0, h0=0.0):
        self.lat0, self.lon0, self.h0 = float(lat0), float(lon0), float(h0)
        self.origin_ecef = np.array(geodetic_to_ecef(self.lat0, self.lon0, self.h0))
        la, lo = np.deg2rad(self.lat0), np.deg2rad(self.lon0)
        sl, cl, so, co = np.sin(la), np.cos(la), np.sin(lo), np.cos(lo)
        self.R = np.array([[-so, co, 0.0],
                           [-sl * co, -sl * so, cl],
                           [cl * co, cl * so, sl]])

    @classmethod
    def centered_on(cls, scene):
        """Anchored on the centre of the scene's DEM grid, h0 = 0."""
        ny, nx = scene.dem.shape
        x_min, y_max, post = scene.transform
        lon0, lat0 = ps_inverse(x_min + (nx / 2.0) * post, y_max - (ny / 2.0) * post)
        return cls(float(lat0), float(lon0), 0.0)

    def ecef_to_local(self, xyz):
        xyz = np.atleast_2d(np.asarray(xyz, np.float64))
        return (xyz - self.origin_ecef) @ self.R.T

    def local_to_ecef(self, enu):
        return np.atleast_2d(np.asarray(enu, np.float64)) @ self.R + self.origin_ecef

    def llh_to_local(self, llh):
        llh = np.atleast_2d(np.asarray(llh, np.float64))
        x, y, z = geodetic_to_ecef(llh[:, 0], llh[:, 1], llh[:, 2])
        return self.ecef_to_local(np.column_stack([x, y, z]))

    def local_to_llh(self, enu):
        xyz = self.local_to_ecef(enu)
        lat, lon, h = ecef_to_geodetic(xyz[:, 0], xyz[:, 1], xyz[:, 2])
        return np.column_stack([lat, lon, h])


# ---------------------------------------------------------------- scenes
@dataclass
class ProjectedScene:
    """One or more DEMs on one EPSG:3413 grid (``dems[0]`` the surface;
    heights above the WGS84 ellipsoid) plus a nav track (lat, lon, h)."""

    name: str
    dems: list               # (ny, nx) float32 arrays, top-down
    transform: tuple         # (x_min, y_max, posting)
    nav_llh: np.ndarray      # (T, 3) float64
    params: dict = field(default_factory=dict)
    nav_roll: np.ndarray = None
    eps: list = None         # multilayer media permittivities (len(dems) + 1)

    @property
    def dem(self):
        return self.dems[0]


def _grid(extent, posting):
    cx, cy = ps_forward(CENTER_LATLON[1], CENTER_LATLON[0])
    cx, cy = float(cx), float(cy)
    n = int(round(extent / posting))
    x_min, y_max = cx - extent / 2.0, cy + extent / 2.0
    xs = x_min + (np.arange(n) + 0.5) * posting
    ys = y_max - (np.arange(n) + 0.5) * posting
    X, Y = np.meshgrid(xs, ys)
    return cx, cy, (x_min, y_max, posting), X, Y


def _nav(cx, cy, n_traces, spacing, h):
    k = float(ps_scale(CENTER_LATLON[1], CENTER_LATLON[0]))
    offs = (np.arange(n_traces) - (n_traces - 1) / 2.0) * spacing * k
    lon, lat = ps_inverse(cx + offs, np.full(n_traces, cy))
    return np.column_stack([lat, lon, np.full(n_traces, h)]).astype(np.float64)


def _scene(name, z_funcs, params, eps=None, *, n_traces=20, altitude=1000.0,
           extent=8000.0, posting=50.0, spacing=100.0):
    cx, cy, tf, X, Y = _grid(extent, posting)
    dems = [zf(X, Y, cx, cy).astype(np.float32) for zf in z_funcs]
    nav = _nav(cx, cy, n_traces, spacing, float(dems[0].mean()) + altitude)
    p = {"n_traces": n_traces, "altitude": altitude, "extent": extent,
         "posting": posting, "spacing": spacing, **params}
    return ProjectedScene(name, dems, tf, nav, p, eps=eps)


def _flat(e):
    return lambda X, Y, cx, cy: np.full_like(X, e)


def flat_scene(elevation=500.0, **kw):
    return _scene("flat", [_flat(elevation)], {"elevation": elevation}, **kw)


def tilted_scene(slope_deg=5.0, elevation=500.0, **kw):
    m = np.tan(np.deg2rad(slope_deg))
    return _scene("tilted", [lambda X, Y, cx, cy: elevation + m * (Y - cy)],
                  {"slope_deg": slope_deg, "elevation": elevation}, **kw)


def hill_scene(height=200.0, sigma=800.0, offset=1500.0, elevation=500.0, **kw):
    def z(X, Y, cx, cy):
        return elevation + height * np.exp(
            -((X - cx) ** 2 + (Y - (cy + offset)) ** 2) / (2 * sigma ** 2))
    return _scene("hill", [z], {"height": height, "sigma": sigma, "offset": offset,
                                "elevation": elevation}, **kw)


def sinusoid_scene(amplitude=100.0, wavelength=2000.0, elevation=500.0, **kw):
    def z(X, Y, cx, cy):
        return elevation + amplitude * np.sin(2 * np.pi * (Y - cy) / wavelength)
    return _scene("sinusoid", [z], {"amplitude": amplitude, "wavelength": wavelength,
                                    "elevation": elevation}, **kw)


def crater_scene(depth=200.0, sigma=800.0, elevation=500.0, **kw):
    def z(X, Y, cx, cy):
        return elevation - depth * np.exp(-((X - cx) ** 2 + (Y - cy) ** 2) / (2 * sigma ** 2))
    return _scene("crater", [z], {"depth": depth, "sigma": sigma,
                                  "elevation": elevation}, **kw)


SCENES = {"flat": flat_scene, "tilted": tilted_scene, "hill": hill_scene,
          "sinusoid": sinusoid_scene, "crater": crater_scene}


def slab_scene(surface=500.0, depth=300.0, eps=(1.0, 3.17, 6.0), **kw):
    """Flat surface over a flat bed ``depth`` below (``synthetic.py:180-185``)."""
    return _scene("slab", [_flat(surface), _flat(surface - depth)],
                  {"surface": surface, "depth": depth}, eps=list(eps), **kw)


def tilted_bed_scene(surface=500.0, depth=300.0, slope_deg=5.0, eps=(1.0, 3.17, 6.0), **kw):
    m = np.tan(np.deg2rad(slope_deg))
    return _scene("tilted_bed",
                  [_flat(surface), lambda X, Y, cx, cy: (surface - depth) + m * (Y - cy)],
                  {"surface": surface, "depth": depth, "slope_deg": slope_deg},
                  eps=list(eps), **kw)


# ------------------------------------------------------------- facets/nav
def build_facets(dem, transform, frame):
    """One mean-plane facet per cell of the DEM's pixel-centre lattice in
    ``frame`` (native posting; ``scene.py:126-176`` with spacing=None):
    pixel centres -> projected (E, N) -> geodetic -> ECEF (with the DEM
    height) -> local ENU, then the corner math of ``facets_from_vertices``."""
    dem = np.asarray(dem, np.float64)
    ny, nx = dem.shape
    x_min, y_max, post = transform
    J, I = np.meshgrid(np.arange(nx) + 0.5, np.arange(ny) + 0.5)
    E = J * post + x_min
    N = I * (-post) + y_max
    x, y, z = projected_to_ecef(E.ravel(), N.ravel(), dem.ravel())
    V = frame.ecef_to_local(np.column_stack([x, y, z])).reshape(ny, nx, 3)
    return facets_from_vertices(V)


def nav_to_frame(nav_llh, frame, roll=None):
    """(T, 3) lat/lon/h -> NavTrack in ``frame``: u_at from central
    differences (one-sided at the ends), u_ct = unit(u_at x up)."""
    pos = frame.llh_to_local(nav_llh)
    g = np.gradient(pos, axis=0) if len(pos) > 1 else np.array([[1.0, 0.0, 0.0]])
    u_at = g / np.linalg.norm(g, axis=1, keepdims=True)
    u_ct = np.cross(u_at, np.array([0.0, 0.0, 1.0]))
    u_ct = u_ct / np.linalg.norm(u_ct, axis=1, keepdims=True)
    return NavTrack(pos, u_at, u_ct, roll)
