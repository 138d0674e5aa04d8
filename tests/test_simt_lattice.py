"""The device scene builder (ss_facets_from_heights) on the CPU emulation:
facets, lookup tables and plane statistics against the host tessellation
(scene.grid_facets, device.DeviceInterfaces), and a refracted field through
device-built facets and interfaces against the host-built path."""

import numpy as np
import pytest
import torch

import simt_harness as sh
from radar_sounder_simulator_amd import _lib
from radar_sounder_simulator_amd import device as dv
from radar_sounder_simulator_amd.scene import (Lattice, SeededLattice, gaussian_surface,
                                               grid_facets, straight_track)

C = 299792458.0


@pytest.fixture(scope="module")
def sim():
    return sh.lib()


@pytest.mark.parametrize("ny,nx", [(21, 37), (16, 16), (3, 40)])
def test_device_facets_match_host_tessellation(sim, ny, nx):
    rng = np.random.default_rng(ny * nx)
    h = 5.0 * gaussian_surface(ny + 1, 25.0, 100.0, rng, m=nx + 1)
    lat = Lattice(h, 25.0, (130.0, -70.0))
    f = lat.facets()
    with _lib.use(sim):
        built, table = dv.build_lattices([lat], {id(lat)}, device="cpu")
    b = built[id(lat)]
    o = b.fac.order
    buf = b.fac.buf.numpy()
    # tile order: a permutation, 16x16 patches
    assert sorted(o) == list(range(ny * nx))
    np.testing.assert_array_equal(buf[0:3].T, f.centers[o])
    np.testing.assert_array_equal(buf[3:6].T, f.normals[o])
    np.testing.assert_array_equal(buf[6], f.areas[o])
    np.testing.assert_array_equal(buf[7:10].T, f.e1[o])
    np.testing.assert_array_equal(buf[10:13].T, f.e2[o])
    np.testing.assert_array_equal(table.numpy()[:ny * nx, :3], f.centers)
    np.testing.assert_array_equal(table.numpy()[:ny * nx, 3:], f.normals)
    A = f.areas.sum()
    np.testing.assert_allclose(b.stats[0], A, rtol=1e-13)
    np.testing.assert_allclose(b.stats[1:4], (f.centers * f.areas[:, None]).sum(0), rtol=1e-12)
    np.testing.assert_allclose(b.center_mean, f.centers.mean(0), rtol=1e-12, atol=1e-9)
    # descriptor: the same affine cell map and mean plane as the host build
    host = dv.DeviceInterfaces([f], device="cpu").desc.numpy()[0]
    np.testing.assert_allclose(b.desc_row()[dv.IDESC["COEF"]:dv.IDESC["COEF"] + 6],
                               host[dv.IDESC["COEF"]:dv.IDESC["COEF"] + 6], atol=1e-9)
    np.testing.assert_allclose(b.desc_row()[dv.IDESC["MP"]:dv.IDESC["MN"] + 3],
                               host[dv.IDESC["MP"]:dv.IDESC["MN"] + 3], rtol=1e-12, atol=1e-12)


def test_refracted_field_through_device_built_scene(sim):
    """Surface + horizontal plane + rough bed as lattices, the joint solve
    padded: device-built scene vs the host-built one, float64."""
    rng = np.random.default_rng(2)
    n, dx = 20, 25.0
    surf = Lattice(3.0 * gaussian_surface(n + 1, dx, 100.0, rng), dx)
    plane = Lattice(np.full((n + 1, n + 1), -60.0), dx)
    bed = Lattice(-300.0 + 30.0 * gaussian_surface(n + 1, dx, 200.0, rng), dx)
    tr = straight_track(4, 40.0, 3000.0)
    eps, att = [1.0, 2.0, 3.15], [0.0, 5.0, 10.0]
    kw = dict(mode="coherent", t0=2 * 2900 / C, dt=1 / 48e6, n_samples=512, c=C,
              gamma=-0.3162, k0=2 * np.pi * 60e6 / C, refraction="joint", precision="fp64",
              pair_bins=True)
    zp = tr.positions[:, 2].min()
    with _lib.use(sim):
        built, table = dv.build_lattices([surf, plane, bed], {id(surf), id(plane)},
                                         device="cpu", n_pad_rows=8)
        ifc_d = dv.DeviceInterfaces.lattice_stacks([[surf, plane]], built, table,
                                                   pad_row0=table.shape[0] - 8, pad_tos=[4],
                                                   z_platform=zp, device="cpu")[0]
        trs = dv.DeviceTraces(tr.positions, tr.u_ct, None, None, device="cpu")
        od, dd, bd, _ = dv.refracted_field(built[id(bed)].fac, trs, ifc_d, eps, att, **kw)
        fb = bed.facets()
        fac_h = dv.DeviceFacets(fb.centers, fb.normals, fb.areas, fb.e1, fb.e2, device="cpu")
        ifc_h = dv.DeviceInterfaces([surf.facets(), plane.facets()], device="cpu", pad_to=4,
                                    z_platform=zp)
        oh, dh, bh, _ = dv.refracted_field(fac_h, trs, ifc_h, eps, att, **kw)
    assert ifc_d.flat_below and ifc_d.k_pad == 2
    np.testing.assert_array_equal(built[id(bed)].fac.unpermute_pairs(bd.numpy()),
                                  fac_h.unpermute_pairs(bh.numpy()))
    od, oh = od.numpy(), oh.numpy()
    assert np.linalg.norm(od - oh) <= 1e-12 * np.linalg.norm(oh)
    np.testing.assert_allclose(dd.numpy(), dh.numpy(), rtol=1e-12, atol=1e-30)


@pytest.mark.parametrize("built", [False, True])
def test_trace_diagnostics_match_oracle(sim, built):
    """ss_trace_diagnostics (nadir facet, first return) against the oracle,
    indices and values exact, on a grid with equidistant ties (the track runs
    along cell edges) in Morton and in tile order."""
    from oracle import output as oo
    rng = np.random.default_rng(7)
    lat = Lattice(2.0 * gaussian_surface(33, 20.0, 100.0, rng, m=41), 20.0)
    f = lat.facets()
    tr = straight_track(13, 20.0, 2000.0)                 # y = 0: on a row of cell edges
    t0, dt, n, c = 2 * 1980.0 / C, 1 / 48e6, 600, C
    with _lib.use(sim):
        if built:
            fac = dv.build_lattices([lat], set(), device="cpu")[0][id(lat)].fac
        else:
            fac = dv.DeviceFacets(f.centers, f.normals, f.areas, f.e1, f.e2, device="cpu")
        trs = dv.DeviceTraces(tr.positions, tr.u_ct, None, None, device="cpu")
        d = dv.trace_diagnostics(fac, trs, t0=t0, dt=dt, n_samples=n, c=c)
    ni, nz = oo.nadir(tr.positions, f.centers)
    fi, ftw, fxyz = oo.first_return(tr.positions, f.centers, t0, dt, n, c)
    np.testing.assert_array_equal(d["nadir_index"].numpy(), ni)
    np.testing.assert_array_equal(d["nadir_z"].numpy(), nz)
    np.testing.assert_array_equal(d["first_index"].numpy(), fi)
    np.testing.assert_array_equal(d["first_twtt"].numpy(), ftw)
    np.testing.assert_array_equal(d["first_xyz"].numpy(), fxyz)
    # a window that no facet reaches: -1 / +inf on the device (the reference
    # then reports facet 0)
    with _lib.use(sim):
        d = dv.trace_diagnostics(fac, trs, t0=1.0, dt=dt, n_samples=10, c=c)
    assert np.all(d["first_index"].numpy() == -1) and np.all(np.isinf(d["first_twtt"].numpy()))


@pytest.mark.parametrize("rows,cols,l", [(13, 21, 90.0), (40, 7, 300.0), (1, 5, 40.0)])
def test_gaussian_heights_match_oracle(sim, rows, cols, l):
    """ss_gaussian_heights (keyed noise, two direct separable circular
    passes) against the oracle's FFT filter of the scalar keyed noise."""
    from oracle import scene as osc
    lat = SeededLattice(rows - 1, cols - 1, 30.0, sigma=4.0, corr_len=l, seed=(5, rows),
                        mean=-75.0)
    with _lib.use(sim):
        h = dv.lattice_heights(lat, device="cpu").numpy()
    want = -75.0 + 4.0 * osc.gaussian_field(osc.keyed_noise(rows, cols, (5, rows)), 30.0, l)
    np.testing.assert_allclose(h, want, rtol=0, atol=1e-12)


def test_gaussian_heights_rejects_bad_arguments(sim):
    lib = sim
    ws = torch.zeros(8, dtype=torch.float64)
    out = torch.zeros(8, dtype=torch.float64)
    p = out.data_ptr()
    assert lib.ss_gaussian_heights(1, 4081, 2, 1.0, 5.0, 1.0, 0.0, p, ws.data_ptr(), 1 << 30,
                                   None) == -3
    assert lib.ss_gaussian_heights(1, 2, 2, 0.0, 5.0, 1.0, 0.0, p, ws.data_ptr(), 64,
                                   None) == -1
    assert lib.ss_gaussian_heights(1, 2, 2, 1.0, 5.0, 1.0, 0.0, p, ws.data_ptr(), 8,
                                   None) == -1          # workspace too small
    assert lib.ss_gaussian_heights_workspace_bytes(3, 5) == 120
    assert lib.ss_keyed_phasors(1, 0, 3, 0, 0, p, 8, None) == -1


def test_seeded_scene_built_on_device(sim):
    """A seeded surface + horizontal plane + seeded bed with keyed speckle
    phasors: heights drawn, tessellated and phasors keyed on the device,
    against the host materialisation of the same lattices (facets to
    rounding of the heights, phasors equal as complex64), and a refracted
    field through the device-built scene against the host-built one."""
    from radar_sounder_simulator_amd.roughness import KeyedPhasors
    n, dx = 20, 25.0
    surf = SeededLattice(n, n, dx, sigma=3.0, corr_len=100.0, seed=(1, 0))
    plane = SeededLattice(n, n, dx, mean=-60.0)
    bed = SeededLattice(n, n, dx, sigma=30.0, corr_len=200.0, seed=(1, 1), mean=-300.0)
    kp = KeyedPhasors((n, n), (1, 99), key_origin=(3, -5))
    with _lib.use(sim):
        built, table = dv.build_lattices([surf, plane, bed], {id(surf), id(plane)},
                                         device="cpu", phasors={id(bed): kp}, n_pad_rows=0)
    b = built[id(bed)]
    o = b.fac.order
    buf = b.fac.buf.numpy()
    ph = np.asarray(kp)[o]
    np.testing.assert_array_equal(buf[14], ph.real.astype(np.float64))
    np.testing.assert_array_equal(buf[15], ph.imag.astype(np.float64))
    for lat in (surf, plane, bed):
        f = lat.facets()                        # host heights (FFT filter)
        fb = built[id(lat)].fac.buf.numpy()
        o = built[id(lat)].fac.order
        np.testing.assert_allclose(fb[0:3].T, f.centers[o], rtol=0, atol=1e-11)
        np.testing.assert_allclose(fb[3:6].T, f.normals[o], rtol=0, atol=1e-12)
        np.testing.assert_array_equal(built[id(lat)].heights.numpy().shape, (n + 1, n + 1))
    assert built[id(plane)].horizontal and table.shape[0] == n * n + 1
    np.testing.assert_array_equal(built[id(plane)].fac.buf.numpy()[2], -60.0)
    # the highest surface facet centre, reduced on the device
    zs = surf.facets().centers[:, 2].max()
    assert abs(built[id(surf)].z_max() - zs) < 1e-11
