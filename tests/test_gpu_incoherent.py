"""GPU: the incoherent (power-summing) surface kernel and the incoherent
simulate() path.

* kernel against oracle/incoherent.py: bins exact, power and dropped to
  1e-9 relative (float64 tier) / 1e-5 (mixed: float32 power and gain),
  with split sides, a window that cuts pairs, and antenna patterns;
* simulate() on the reference's five simc verification scenes (projected
  DEMs, WGS84 geodesy, lat/lon nav) against the simc golden fixtures under
  the reference's parity metrics and thresholds, and against the oracle
  run on the same facets;
* an incoherent two-interface (slab) run: layer 0 by the incoherent kernel,
  layer 1 by the refracted kernel in incoherent mode, each against its
  oracle.
"""

import numpy as np
import pytest

import simc_metrics as M
from oracle import incoherent as oi
from oracle import multilayer as om
from radar_sounder_simulator_amd import antenna as A
from radar_sounder_simulator_amd import config as cf
from radar_sounder_simulator_amd import geodesy as G
from radar_sounder_simulator_amd import kernels as K
from radar_sounder_simulator_amd import simulate as sm
from radar_sounder_simulator_amd.scene import gaussian_surface, grid_facets, straight_track

pytestmark = pytest.mark.gpu
C = 299792458.0

CASES = ["plain", "split", "cut", "tabulated", "dipole", "array"]


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_incoherent_kernel_vs_oracle(case, precision):
    print(f"case incoherent/{case}/{precision}", flush=True)
    rng = np.random.default_rng(21)
    f = grid_facets(25.0 * gaussian_surface(97, 20.0, 150.0, rng), 20.0)
    tr = straight_track(9, 50.0, 900.0)
    kw = dict(t0=2 * 850 / C, dt=1 / 48e6, n_samples=1024, c=C)
    if case == "split":
        kw.update(split_sides=True)
    elif case == "cut":
        kw.update(t0=2 * 930 / C, n_samples=60, split_sides=True)
    elif case == "tabulated":
        kw.update(pattern=A.pattern_args(A.cos_power_table(2.0), tr.u_at, tr.u_ct))
    elif case == "dipole":
        kw.update(pattern=A.pattern_args(A.Antenna(kind="dipole", axis="cross_track"),
                                         tr.u_at, tr.u_ct))
    elif case == "array":
        kw.update(pattern=A.pattern_args(A.Antenna(kind="array", n_elements=4,
                                                   axis="along_track"), tr.u_at, tr.u_ct))
    ref, dref, bref = oi.incoherent_cluttergram(tr.positions, tr.u_ct, f.centers, f.normals,
                                                f.areas, return_bins=True, **kw)
    out, drop, bins = K.incoherent_cluttergram(tr.positions, tr.u_ct, f.centers, f.normals,
                                               f.areas, precision=precision, return_bins=True,
                                               **kw)
    np.testing.assert_array_equal(bins, bref)
    tol = 1e-9 if precision == "fp64" else 1e-5
    assert np.linalg.norm(ref) > 0
    assert np.linalg.norm(out - ref) <= tol * np.linalg.norm(ref)
    np.testing.assert_allclose(drop, dref, rtol=tol, atol=1e-300)
    if case == "cut":
        assert (dref > 0).all()


def _simc_config(rc, precision):
    return cf.SimConfig(mode="incoherent", split_sides=True, precision=precision,
                        radar=cf.RadarConfig(dt=rc["dt"], n_samples=rc["n_samples"],
                                             t0=rc["t0"], c=rc["c"]))


@pytest.mark.parametrize("name", M.NAMES)
def test_simulate_simc_scene(name):
    """The reference's verification runs (``tests/test_parity.py:58-99``)
    through simulate() on the device, float64 and mixed."""
    print(f"case incoherent/simc/{name}", flush=True)
    fx = M.load(name)
    rc = fx["meta"]["radar_config"]
    scene = G.SCENES[name](**fx["meta"]["scene"]["params"])
    res = sm.simulate(scene, _simc_config(rc, "fp64"), device="cuda")
    assert res.power.shape == (20, rc["n_samples"], 2) and res.field is None
    m = M.compare(res.power.sum(-1), res.first_return_twtt, fx, M.FRET_TOL[name])
    bad = {k: v for k, v in m.items() if not v["ok"]}
    assert not bad, bad
    # the same facets through the oracle
    fr = G.LocalFrame.centered_on(scene)
    f = G.build_facets(scene.dem, scene.transform, fr)
    tr = G.nav_to_frame(scene.nav_llh, fr)
    ref, dref = oi.incoherent_cluttergram(tr.positions, tr.u_ct, f.centers, f.normals, f.areas,
                                          t0=rc["t0"], dt=rc["dt"], n_samples=rc["n_samples"],
                                          c=rc["c"], split_sides=True)
    assert np.linalg.norm(res.power - ref) <= 1e-6 * np.linalg.norm(ref)  # float32 output
    np.testing.assert_allclose(res.dropped_power, dref, rtol=1e-6, atol=1e-30)
    # first return in lat/lon near simc's ground point (simc regrids along
    # its own track and samples its DEM by truncation; the oracle's first
    # return lies 34-92 m from simc's on these scenes): within three 50 m
    # postings
    lat, lon = res.first_return_lat, res.first_return_lon
    x, y, z = G.geodetic_to_ecef(lat, lon, np.zeros_like(lat))
    ours = np.column_stack([x, y, z])
    simc = fx["fret_xyz"]
    u = ours / np.linalg.norm(ours, axis=1, keepdims=True)
    v = simc / np.linalg.norm(simc, axis=1, keepdims=True)
    loc = np.linalg.norm(u - v, axis=1) * np.linalg.norm(simc, axis=1)
    assert loc.max() < 150.0, loc.max()
    mixed = sm.simulate(scene, _simc_config(rc, "mixed"), device="cuda")
    assert np.linalg.norm(mixed.power - res.power) <= 1e-5 * np.linalg.norm(res.power)


def test_simulate_incoherent_slab_vs_oracle():
    print("case incoherent/slab", flush=True)
    scene = G.slab_scene(extent=3000.0, posting=50.0, n_traces=8)
    cfg = cf.SimConfig(mode="incoherent", precision="fp64",
                       radar=cf.RadarConfig(dt=1e-8, n_samples=2000, t0=0.0),
                       media=[cf.Medium(name="air", eps_r=1.0),
                              cf.Medium(name="ice", eps_r=3.17, attenuation_db_per_km=8.0),
                              cf.Medium(name="bed", eps_r=6.0)],
                       interfaces=[cf.DemInterface(name="surface"), cf.DemInterface(name="bed")])
    res = sm.simulate(scene, cfg, device="cuda")
    assert res.layer == ["surface", "bed"] and res.field is None
    fr = G.LocalFrame.centered_on(scene)
    surf = G.build_facets(scene.dems[0], scene.transform, fr)
    bed = G.build_facets(scene.dems[1], scene.transform, fr)
    tr = G.nav_to_frame(scene.nav_llh, fr)
    win = dict(t0=0.0, dt=1e-8, n_samples=2000, c=C)
    r0, _ = oi.incoherent_cluttergram(tr.positions, tr.u_ct, surf.centers, surf.normals,
                                      surf.areas, **win)
    r1, _ = om.refracted_cluttergram(tr.positions, tr.u_ct, bed, [surf], [1.0, 3.17],
                                     [0.0, 8.0], mode="incoherent", **win)
    for j, ref in enumerate((r0, r1)):
        out = res.power[..., j]
        assert np.linalg.norm(ref) > 0
        assert np.linalg.norm(out - ref) <= 1e-6 * np.linalg.norm(ref), j
