"""Pins of oracle/incoherent.py and of the geodesy that takes the reference's
verification scenes into the local frame (CPU).

* geodesy: the nav tracks of the five simc fixtures (``nav_ecef``, written
  by the reference through pyproj) reproduced from the scene parameters;
  ellipsoid and projection round trips; the projection's point scale factor
  against a finite-difference meridian arc;
* incoherent oracle: the simc golden fixtures under the reference's own
  parity metrics and thresholds (``tests/test_parity.py:94-99``); the
  closed-form integral of (A cos)^2 / r^4 over a plane under a nadir
  platform; the nadir leading-edge bin; left/right symmetry; power +
  dropped conservation under a window cut.
"""

import numpy as np
import pytest

import simc_metrics as M
from oracle import incoherent as oi
from radar_sounder_simulator_amd import geodesy as G
from radar_sounder_simulator_amd.scene import grid_facets, straight_track

C = 299792458.0


@pytest.mark.parametrize("name", M.NAMES)
def test_scene_nav_matches_fixture_ecef(name):
    """Projection (forward for the centre, scale factor, inverse for the
    track) and geodetic -> ECEF against the fixture's nav track: 0.2 mm
    (the tilted scene's nav height carries a float32 ulp of its DEM mean)."""
    fx = M.load(name)
    sc = G.SCENES[name](**fx["meta"]["scene"]["params"])
    x, y, z = G.geodetic_to_ecef(*sc.nav_llh.T)
    err = np.abs(np.column_stack([x, y, z]) - fx["nav_ecef"]).max()
    assert err < 2e-4, err


def test_ellipsoid_and_projection_round_trips():
    rng = np.random.default_rng(0)
    lat = rng.uniform(60.0, 89.0, 200)
    lon = rng.uniform(-180.0, 180.0, 200)
    h = rng.uniform(-500.0, 9000.0, 200)
    la, lo, hh = G.ecef_to_geodetic(*G.geodetic_to_ecef(lat, lon, h))
    assert np.abs(la - lat).max() < 1e-11 and np.abs(hh - h).max() < 1e-7
    assert np.abs(((lo - lon) + 180.0) % 360.0 - 180.0).max() < 1e-11
    x, y = G.ps_forward(lon, lat)
    lo2, la2 = G.ps_inverse(x, y)
    assert np.abs(la2 - lat).max() < 1e-11
    assert np.abs(((lo2 - lon) + 180.0) % 360.0 - 180.0).max() < 1e-11
    # true scale on the standard parallel
    assert abs(G.ps_scale(-45.0, 70.0) - 1.0) < 1e-14


def test_projection_scale_is_meridian_arc_ratio():
    """k(75 N) = d(rho)/d(meridian arc), the arc element a (1 - e^2) /
    (1 - e^2 sin^2 phi)^1.5 d phi (independent of the rho formula)."""
    phi, d = 75.0, 1e-5
    x1, y1 = G.ps_forward(-40.0, phi - d)
    x2, y2 = G.ps_forward(-40.0, phi + d)
    drho = np.hypot(x2 - x1, y2 - y1)
    s = np.sin(np.deg2rad(phi))
    arc = G.A_WGS84 * (1 - G.E2_WGS84) / (1 - G.E2_WGS84 * s * s) ** 1.5 * np.deg2rad(2 * d)
    assert abs(drho / arc - G.ps_scale(-40.0, phi)) < 1e-9
    assert abs(G.ps_scale(-40.0, 75.0) - 0.98666) < 1e-5


def _run_oracle(name):
    fx = M.load(name)
    rc = fx["meta"]["radar_config"]
    sc = G.SCENES[name](**fx["meta"]["scene"]["params"])
    fr = G.LocalFrame.centered_on(sc)
    f = G.build_facets(sc.dem, sc.transform, fr)
    tr = G.nav_to_frame(sc.nav_llh, fr)
    p, _ = oi.incoherent_cluttergram(tr.positions, tr.u_ct, f.centers, f.normals, f.areas,
                                     t0=rc["t0"], dt=rc["dt"], n_samples=rc["n_samples"],
                                     c=rc["c"], split_sides=True)
    r = np.linalg.norm(f.centers[None] - tr.positions[:, None], axis=-1)
    ft = 2.0 * r / rc["c"]
    ft = np.where((ft >= rc["t0"]) & (ft < rc["t0"] + rc["n_samples"] * rc["dt"]), ft, np.inf)
    return fx, p, ft.min(1)


@pytest.mark.parametrize("name", M.NAMES)
def test_oracle_passes_simc_parity(name):
    fx, p, frt = _run_oracle(name)
    m = M.compare(p.sum(-1), frt, fx, M.FRET_TOL[name])
    bad = {k: v for k, v in m.items() if not v["ok"]}
    assert not bad, bad
    # the constant between the two tools: 2 (1/k)^2 = 2.055 (one rectangle
    # of area (50/k)^2 per projected cell against simc's two triangles)
    assert 2.04 < m["power_ratio_cv"]["median_ratio"] < 2.07


def test_plane_integral_closed_form():
    """sum (A cos)^2 / r^4 over a plane under a nadir platform at height h
    = A * pi / (2 h^2) (the integral of h^2/r^6 over the plane), truncated
    at radius R by h^4 / (h^2 + R^2)^2."""
    h, dx, n = 100.0, 2.0, 1000
    f = grid_facets(np.zeros((n + 1, n + 1)), dx)
    tr = straight_track(1, 10.0, h)
    out, drop = oi.incoherent_cluttergram(tr.positions, tr.u_ct, f.centers, f.normals, f.areas,
                                          t0=0.0, dt=1e-9, n_samples=100000, c=C)
    total = out.sum() + drop.sum()
    half = n * dx / 2.0
    # the square [-half, half]^2 lies between the discs of radius half and
    # half * sqrt(2): bracket the truncation
    lo = dx * dx * np.pi / (2 * h * h) * (1 - h ** 4 / (h * h + half * half) ** 2)
    hi = dx * dx * np.pi / (2 * h * h) * (1 - h ** 4 / (h * h + 2 * half * half) ** 2)
    assert lo * (1 - 1e-3) < total < hi * (1 + 1e-3), (lo, total, hi)


def test_nadir_bin_symmetry_and_conservation():
    h, dx, n = 500.0, 10.0, 60
    f = grid_facets(np.zeros((n + 1, n + 1)), dx)
    tr = straight_track(3, 20.0, h)
    kw = dict(t0=2 * 400.0 / C, dt=1e-8, c=C, split_sides=True)
    full, dfull, bins = oi.incoherent_cluttergram(tr.positions, tr.u_ct, f.centers, f.normals,
                                                  f.areas, n_samples=4000, return_bins=True, **kw)
    # leading edge at the nearest facet's range (an even cell count: the
    # nearest centres sit half a cell off in x and y)
    first = np.floor((2 * np.sqrt(h * h + 2 * (dx / 2.0) ** 2) / C - kw["t0"]) / kw["dt"])
    nz = np.nonzero(full.sum(-1)[1])[0]
    assert nz[0] == first and bins.min() == first
    # a symmetric plane about the track: left == right
    l, r = full[..., 0].sum(1), full[..., 1].sum(1)
    np.testing.assert_allclose(l, r, rtol=1e-12)
    # a window cut moves power into dropped and conserves the total
    cut, dcut = oi.incoherent_cluttergram(tr.positions, tr.u_ct, f.centers, f.normals, f.areas,
                                          n_samples=40, **kw)
    np.testing.assert_allclose(cut.sum((1, 2)) + dcut, full.sum((1, 2)) + dfull, rtol=1e-12)
    assert (dcut > 0).all() and np.array_equal(cut, full[:, :40])
