"""GPU: the allocation-free refracted entry point (ss_refracted_field_noalloc,
the library's per-device scratch array instead of a caller workspace)
against ss_refracted_field with a caller workspace, through the C ABI: two
calls in a row, the second on a smaller input, every pair through the
float64 redo record (SS_REDO_ALL) and in the plain mixed tier, and the
float64 tier.  Bins exact; fields to the summation-order tolerance of the
tier (float64 and forced redo 1e-12, mixed 1e-5)."""

import numpy as np
import pytest
import torch

from radar_sounder_simulator_amd import _lib
from radar_sounder_simulator_amd import device as dv
from radar_sounder_simulator_amd.scene import gaussian_surface, grid_facets, straight_track

pytestmark = pytest.mark.gpu
C = 299792458.0


def _scene(n, n_traces, seed):
    rng = np.random.default_rng(seed)
    dx = 25.0
    surf = grid_facets(3.0 * gaussian_surface(n + 1, dx, 100.0, rng), dx)
    bed = grid_facets(-300.0 + 30.0 * gaussian_surface(n + 1, dx, 200.0, rng), dx)
    tr = straight_track(n_traces, 40.0, 3000.0)
    return surf, bed, tr


@pytest.mark.parametrize("case", ["mixed", "redo_all", "fp64"])
def test_noalloc_matches_workspace_entry(case):
    kw = dict(mode="coherent", t0=2 * 2900 / C, dt=1 / 48e6, n_samples=512, c=C,
              gamma=-0.3162, k0=2 * np.pi * 60e6 / C)
    prec = "fp64" if case == "fp64" else "mixed"
    flags = _lib.SS_REDO_ALL if case == "redo_all" else 0
    tol = 1e-5 if case == "mixed" else 1e-12
    for n, nt, seed in ((40, 21, 3), (17, 5, 4)):    # second call smaller
        print(f"case noalloc/{case}/n{n}/traces{nt}", flush=True)
        surf, bed, tr = _scene(n, nt, seed)
        fac = dv.DeviceFacets(bed.centers, bed.normals, bed.areas, bed.e1, bed.e2,
                              device="cuda")
        ifc = dv.DeviceInterfaces([surf], device="cuda")
        trs = dv.DeviceTraces(tr.positions, tr.u_ct, None, None, device="cuda")
        assert dv.refracted_workspace_bytes(prec, fac.n, trs.n) <= _lib.NOALLOC_SCRATCH_BYTES
        res = {}
        for noalloc in (False, True):
            out, drop, bins, _ = dv.refracted_field(
                fac, trs, ifc, [1.0, 3.15], [0.0, 10.0], precision=prec, pair_bins=True,
                extra_flags=flags, noalloc=noalloc, **kw)
            torch.cuda.synchronize()
            res[noalloc] = (out.cpu().numpy(), drop.cpu().numpy(), bins.cpu().numpy())
        (oa, da, ba), (ob, db, bb) = res[False], res[True]
        np.testing.assert_array_equal(bb, ba)
        assert np.linalg.norm(oa) > 0
        assert np.linalg.norm(ob - oa) <= tol * np.linalg.norm(oa)
        np.testing.assert_allclose(db, da, rtol=tol, atol=1e-30)
