"""Host-side logic of the library -- CPU only (no kernel calls).

Config validation (the reference's config.py ValueError cases), interface
resolution, joint padding buckets, the device buffer layout built on the
host, Morton tiling compactness, shim argument checks raised before any
device work, the frozen flop model, and the workload builders.
"""

import numpy as np
import pytest
import torch

from radar_sounder_simulator_amd import config as cf
from radar_sounder_simulator_amd import device as dv
from radar_sounder_simulator_amd import kernels as K
from radar_sounder_simulator_amd import roofline, workloads
from radar_sounder_simulator_amd import simulate as sm
from radar_sounder_simulator_amd.scene import gaussian_surface, grid_facets, straight_track


def _radar(**kw):
    return cf.RadarConfig(dt=1e-8, n_samples=128, t0=0.0, f0=60e6, **kw)


def test_config_validation():
    with pytest.raises(ValueError):
        cf.WaveformConfig(kind="chirp", bandwidth=None, pulse_length=1e-6)
    with pytest.raises(ValueError):
        cf.AntennaConfig(kind="tabulated", theta_deg=[0, 10], gain=[1.0])
    with pytest.raises(ValueError):
        cf.AntennaConfig(kind="array", n_elements=1)
    with pytest.raises(ValueError):
        cf.SimConfig(mode="coherent", radar=_radar(), media=[cf.Medium(name="a", eps_r=1)])
    with pytest.raises(ValueError):
        cf.SimConfig(mode="coherent", radar=_radar(), interfaces=[
            cf.DemInterface(roughness=cf.RoughnessConfig(sigma_m=0.1, corr_length_m=1.0,
                                                         acf="exponential"))])
    with pytest.raises(ValueError):
        cf.SimConfig(mode="incoherent", radar=_radar(), grazing_fix=cf.GrazingFixConfig())
    with pytest.raises(ValueError):
        cf.SimConfig(mode="coherent", radar=_radar(),
                     media=[cf.Medium(name=m, eps_r=e) for m, e in (("a", 1), ("b", 3), ("c", 8))],
                     interfaces=[cf.DemInterface(name="s"),
                                 cf.OffsetInterface(name="o", reference="zz", offset=-5)])
    ok = cf.SimConfig(mode="coherent", radar=_radar(), grazing_fix=cf.GrazingFixConfig(),
                      interfaces=[cf.DemInterface(roughness=cf.RoughnessConfig(
                          sigma_m=0.1, corr_length_m=1.0, acf="exponential"))])
    assert cf.SimConfig.model_validate_json(ok.model_dump_json()) == ok


def test_interface_resolution_and_padding():
    rng = np.random.default_rng(0)
    dems = [gaussian_surface(9, 10.0, 30.0, rng), -200 + gaussian_surface(9, 10.0, 30.0, rng)]
    scene = sm.PlanarScene(dems=dems, spacing=10.0, track=straight_track(2, 10, 1000))
    c = cf.SimConfig(mode="coherent", radar=_radar(),
                     media=[cf.Medium(name=m, eps_r=e) for m, e in
                            (("air", 1), ("ice", 3.15), ("x", 4.0), ("rock", 8))],
                     interfaces=[cf.DemInterface(name="surface"), cf.DemInterface(name="bed"),
                                 cf.OffsetInterface(name="shadow", reference="bed",
                                                    offset=-5.0)])
    f, names = sm.build_interfaces(scene, c)
    assert names == ["surface", "bed", "shadow"]
    np.testing.assert_allclose(f[2].centers[:, 2], f[1].centers[:, 2] - 5.0)
    np.testing.assert_array_equal(f[2].normals, f[1].normals)
    assert [sm.joint_pad_to(j, 24) for j in (1, 2, 3, 5, 17, 24)] == [1, 2, 4, 8, 24, 24]


def test_device_buffer_layout_on_host():
    f = grid_facets(np.arange(25.0).reshape(5, 5), 2.0)
    g = np.linspace(0, 1, len(f))
    ph = np.exp(1j * np.arange(len(f)))
    d = dv.DeviceFacets.from_facets(f, gamma=g, phasors=ph, device="cpu")
    buf = d.buf.numpy()
    o = d.order
    np.testing.assert_array_equal(np.sort(o), np.arange(len(f)))
    np.testing.assert_array_equal(buf[0:3, :len(f)].T, f.centers[o])
    np.testing.assert_array_equal(buf[13, :len(f)], g[o])
    np.testing.assert_array_equal(buf[14, :len(f)] + 1j * buf[15, :len(f)], ph[o])
    pairs = np.arange(3 * len(f)).reshape(3, -1)
    np.testing.assert_array_equal(d.unpermute_pairs(pairs[:, ]) [:, o], pairs)


def test_morton_tiles_are_compact():
    f = grid_facets(np.zeros((129, 129)), 20.0)
    o = dv.morton_order(f.centers[:, :2])
    tiles = f.centers[o].reshape(-1, 256, 3)
    ext = tiles[..., :2].max(1) - tiles[..., :2].min(1)
    assert ext.max() <= 16 * 20.0          # every tile is a 16 x 16 patch


def test_interfaces_descriptor_padding():
    surf = grid_facets(np.zeros((5, 5)), 10.0)
    mid = grid_facets(np.full((5, 5), -30.0), 10.0)
    ifc = dv.DeviceInterfaces([surf, mid], device="cpu", pad_to=4, z_platform=1000.0)
    d = ifc.desc.numpy()
    assert ifc.k_pad == 2 and ifc.n == 4
    assert np.all(d[:2, 9] == 0) and np.all(np.diff(d[:2, 10]) < 0)   # pads top-down
    assert d[0, 10] < 1000.0 and d[1, 10] > 0.0
    np.testing.assert_allclose(d[2, 11:14], [0, 0, 1])
    with pytest.raises(ValueError):
        dv.DeviceInterfaces([surf], device="cpu", pad_to=2, z_platform=-5.0)


def test_shim_errors_before_device_work():
    f = grid_facets(np.zeros((3, 3)), 10.0)
    tr = straight_track(1, 1.0, 100.0)
    kw = dict(k=1.0, t0=0.0, dt=1e-8, n_samples=8, c=3e8)
    with pytest.raises(ValueError):
        K.coherent_cluttergram(tr.positions, tr.u_ct, f.centers, f.normals, f.areas, f.e1,
                               f.e2, gamma=np.zeros(7), **kw)
    with pytest.raises(ValueError):
        K.coherent_cluttergram(tr.positions, tr.u_ct, f.centers, f.normals, f.areas, f.e1,
                               f.e2, gamma=0.1, roughness=(0.1, 1, np.ones(4), 10,
                                                           "exponential"), **kw)
    kw2 = dict(t0=0.0, dt=1e-8, n_samples=8, c=3e8)
    with pytest.raises(ValueError):
        K.refracted_cluttergram(tr.positions, tr.u_ct, f, [f], [1, 3], [0, 0],
                                mode="incoherent", taper_s=0.1, **kw2)
    with pytest.raises(ValueError):
        K.refracted_cluttergram(tr.positions, tr.u_ct, f, [f], [1, 3, 4], [0, 0, 0],
                                mode="coherent", **kw2)
    with pytest.raises(ValueError):
        K.refracted_cluttergram(tr.positions, tr.u_ct, f, [f], [1, 3], [0, 0],
                                mode="coherent", refraction="bogus", **kw2)
    if not torch.cuda.is_available():
        with pytest.raises(RuntimeError):
            K.coherent_cluttergram(tr.positions, tr.u_ct, f.centers, f.normals, f.areas,
                                   f.e1, f.e2, gamma=0.1, **kw)


def test_flop_model_and_workloads():
    assert roofline.surface_pair_flops() == 59
    assert roofline.refracted_pair_flops(1) == 1240 + 42 + 27 + 15
    assert roofline.refracted_pair_flops(2) - roofline.refracted_pair_flops(1) == 1240
    wl = workloads.config2(n=16, n_traces=4, n_samples=256)
    assert [L.name for L in wl.layers] == ["surface", "bed"]
    assert wl.meta["n_terms"] == (10, 24)
    assert wl.pairs() == 2 * 16 * 16 * 4
    assert wl.layers[1].gamma == pytest.approx(-10 ** -0.5)
    wl1 = workloads.config1(n=8, n_traces=2)
    assert wl1.layers[0].gamma == pytest.approx((1 - np.sqrt(3.15)) / (1 + np.sqrt(3.15)))


def test_quantization_alias_warning():
    """The chirped-run alias check (reference simulate.py:71-95): 195 MHz at
    dt = 5 ns aliases to 5 MHz, inside B/2 = 7.5 MHz -> warns; 60 MHz at
    48 MHz sampling aliases to 12 MHz, outside -> silent; interp_bins or a
    delta waveform -> silent."""
    import warnings as w
    chirp = dict(kind="chirp", bandwidth=15e6, pulse_length=20e-6)
    rc = cf.RadarConfig(dt=5e-9, n_samples=64, t0=0.0, f0=195e6,
                        waveform=cf.WaveformConfig(**chirp))
    with pytest.warns(UserWarning, match="alias in band"):
        sm.warn_quantization_alias(rc)
    quiet = [cf.RadarConfig(dt=1 / 48e6, n_samples=64, t0=0.0, f0=60e6,
                            waveform=cf.WaveformConfig(**chirp)),
             cf.RadarConfig(dt=5e-9, n_samples=64, t0=0.0, f0=195e6,
                            waveform=cf.WaveformConfig(interp_bins=True, **chirp)),
             cf.RadarConfig(dt=5e-9, n_samples=64, t0=0.0, f0=195e6)]
    for q in quiet:
        with w.catch_warnings():
            w.simplefilter("error")
            sm.warn_quantization_alias(q)


def test_facet_size_and_corr_length_warnings():
    """Fresnel-zone facet check (reference scene.py:179-195): limit
    0.5 sqrt(lambda R); at 5 m and R = 1000 m the limit is 35.4 m, so 25 m
    facets pass (ratio 0.707) and 100 m facets warn (ratio 2.83).  The
    roughness check (simulate.py:157-162) warns when the correlation length
    exceeds the facet edge."""
    import warnings as w
    from radar_sounder_simulator_amd.scene import check_facet_size
    small = grid_facets(np.zeros((5, 5)), 25.0)
    big = grid_facets(np.zeros((5, 5)), 100.0)
    with w.catch_warnings():
        w.simplefilter("error")
        assert check_facet_size(small, 5.0, 1000.0) == pytest.approx(25.0 / (0.5 * np.sqrt(5e3)))
    with pytest.warns(UserWarning, match="Fresnel-zone limit"):
        assert check_facet_size(big, 5.0, 1000.0) == pytest.approx(100.0 / (0.5 * np.sqrt(5e3)))
    small.phase_keys = None
    rough = cf.DemInterface(name="s", roughness=cf.RoughnessConfig(sigma_m=0.1,
                                                                   corr_length_m=40.0))
    with pytest.warns(UserWarning, match="exceeds the facet size"):
        args = sm._roughness_args(rough, small, 1.0, (0, 0))
    assert args[3] == 10 and len(args[2]) == len(small)
    fine = cf.DemInterface(name="s", roughness=cf.RoughnessConfig(sigma_m=0.1,
                                                                  corr_length_m=10.0))
    with w.catch_warnings():
        w.simplefilter("error")
        sm._roughness_args(fine, small, 1.0, (0, 0))


def test_joint_form_default_is_reference():
    """refraction='joint' runs the reference's solver unless the flat-stack
    form is asked for (SimConfig.joint_form, flat_stack=...)."""
    cfg = cf.SimConfig(mode="coherent", radar=_radar())
    assert cfg.refraction == "joint" and cfg.joint_form == "reference"
    import inspect
    assert inspect.signature(K.refracted_cluttergram).parameters["flat_stack"].default is False
    assert inspect.signature(dv.refracted_field).parameters["flat_stack"].default is False
    assert all(not L.flat_stack for L in workloads.config3(nx=16, ny=8, n_traces=2,
                                                           flat_stack=False).layers)
