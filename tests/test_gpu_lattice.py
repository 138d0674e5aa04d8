"""GPU: the device scene builder (ss_facets_from_heights) against the host
tessellation, bit for bit, and a refracted field through a device-built
scene against the host-built one."""

import numpy as np
import pytest
import torch

from radar_sounder_simulator_amd import device as dv
from radar_sounder_simulator_amd.scene import Lattice, gaussian_surface, straight_track

pytestmark = pytest.mark.gpu
C = 299792458.0


@pytest.mark.parametrize("ny,nx", [(21, 37), (256, 300)])
def test_device_facets_match_host_tessellation(ny, nx):
    print(f"case lattice/facets/{ny}x{nx}", flush=True)
    rng = np.random.default_rng(ny * nx)
    lat = Lattice(5.0 * gaussian_surface(ny + 1, 25.0, 100.0, rng, m=nx + 1), 25.0, (130.0, -70.0))
    f = lat.facets()
    built, table = dv.build_lattices([lat], {id(lat)}, device="cuda")
    torch.cuda.synchronize()
    b = built[id(lat)]
    o = b.fac.order
    buf = b.fac.buf.cpu().numpy()
    np.testing.assert_array_equal(buf[0:3].T, f.centers[o])
    np.testing.assert_array_equal(buf[3:6].T, f.normals[o])
    np.testing.assert_array_equal(buf[6], f.areas[o])
    np.testing.assert_array_equal(buf[7:10].T, f.e1[o])
    np.testing.assert_array_equal(buf[10:13].T, f.e2[o])
    t = table.cpu().numpy()
    np.testing.assert_array_equal(t[:ny * nx, :3], f.centers)
    np.testing.assert_array_equal(t[:ny * nx, 3:], f.normals)
    np.testing.assert_allclose(b.stats[0], f.areas.sum(), rtol=1e-13)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_refracted_field_through_device_built_scene(precision):
    print(f"case lattice/refracted/{precision}", flush=True)
    rng = np.random.default_rng(2)
    n, dx = 48, 25.0
    surf = Lattice(3.0 * gaussian_surface(n + 1, dx, 100.0, rng), dx)
    plane = Lattice(np.full((n + 1, n + 1), -60.0), dx)
    bed = Lattice(-300.0 + 30.0 * gaussian_surface(n + 1, dx, 200.0, rng), dx)
    tr = straight_track(6, 40.0, 3000.0)
    eps, att = [1.0, 2.0, 3.15], [0.0, 5.0, 10.0]
    kw = dict(mode="coherent", t0=2 * 2900 / C, dt=1 / 48e6, n_samples=512, c=C,
              gamma=-0.3162, k0=2 * np.pi * 60e6 / C, refraction="joint", precision=precision,
              pair_bins=True)
    zp = tr.positions[:, 2].min()
    built, table = dv.build_lattices([surf, plane, bed], {id(surf), id(plane)}, device="cuda",
                                     n_pad_rows=8)
    ifc_d = dv.DeviceInterfaces.lattice_stacks([[surf, plane]], built, table,
                                               pad_row0=table.shape[0] - 8, pad_tos=[4],
                                               z_platform=zp, device="cuda")[0]
    trs = dv.DeviceTraces(tr.positions, tr.u_ct, None, None, device="cuda")
    od, dd, bd, _ = dv.refracted_field(built[id(bed)].fac, trs, ifc_d, eps, att, **kw)
    fb = bed.facets()
    fac_h = dv.DeviceFacets(fb.centers, fb.normals, fb.areas, fb.e1, fb.e2, device="cuda")
    ifc_h = dv.DeviceInterfaces([surf.facets(), plane.facets()], device="cuda", pad_to=4,
                                z_platform=zp)
    oh, dh, bh, _ = dv.refracted_field(fac_h, trs, ifc_h, eps, att, **kw)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(built[id(bed)].fac.unpermute_pairs(bd.cpu().numpy()),
                                  fac_h.unpermute_pairs(bh.cpu().numpy()))
    od, oh = od.cpu().numpy(), oh.cpu().numpy()
    tol = 1e-12 if precision == "fp64" else 1e-6
    assert np.linalg.norm(od - oh) <= tol * np.linalg.norm(oh)
