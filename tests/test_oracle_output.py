"""Pins of oracle/output.py (nadir and first-return diagnostics) by closed
forms: a horizontal plane under the platform, equidistant ties, a window
that excludes every facet, a two-layer slab's nadir delay."""

import numpy as np

from oracle import output as oo
from radar_sounder_simulator_amd.scene import grid_facets, straight_track

C = 299792458.0


def test_plane_nadir_and_first_return_closed_form():
    f = grid_facets(np.full((9, 9), -20.0), 10.0)       # centres at +-5, +-15, ...
    pos = np.array([[5.0, 5.0, 1000.0], [14.0, -6.0, 500.0]])
    idx, z = oo.nadir(pos, f.centers)
    np.testing.assert_array_equal(f.centers[idx, :2], [[5.0, 5.0], [15.0, -5.0]])
    np.testing.assert_array_equal(z, [-20.0, -20.0])
    fi, tw, xyz = oo.first_return(pos, f.centers, 0.0, 1e-8, 100000, C)
    np.testing.assert_array_equal(fi, idx)           # the nearest centre is the nearest point
    np.testing.assert_allclose(tw, 2 * np.linalg.norm(xyz - pos, axis=1) / C, rtol=0)
    np.testing.assert_allclose(tw[0], 2 * 1020.0 / C, rtol=1e-15)


def test_ties_take_the_lowest_index_and_empty_window_reports_facet_0():
    f = grid_facets(np.zeros((3, 3)), 10.0)           # centres (+-5, +-5)
    pos = np.array([[0.0, 0.0, 100.0]])               # equidistant to all four
    idx, _ = oo.nadir(pos, f.centers)
    assert idx[0] == 0
    fi, tw, _ = oo.first_return(pos, f.centers, 0.0, 1e-8, 100000, C)
    assert fi[0] == 0 and np.isfinite(tw[0])
    fi, tw, _ = oo.first_return(pos, f.centers, 1.0, 1e-8, 10, C)   # window far later
    assert fi[0] == 0 and tw[0] == np.inf


def test_slab_nadir_twtt_layers():
    tr = straight_track(3, 10.0, 800.0)
    top = grid_facets(np.full((5, 5), 0.0), 10.0)
    bot = grid_facets(np.full((5, 5), -100.0), 10.0)
    out = oo.nadir_twtt_layers(tr.positions, [top, bot], [1.0, 3.15], C)
    np.testing.assert_allclose(out[:, 0], 2 * 800.0 / C, rtol=1e-15)
    np.testing.assert_allclose(out[:, 1], 2 * (800.0 + np.sqrt(3.15) * 100.0) / C, rtol=1e-15)
