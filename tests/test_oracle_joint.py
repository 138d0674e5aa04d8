"""Pins of the joint N-interface solver oracle (oracle/joint.py) -- CPU.

Mirrors tests/test_refraction_joint.py:47-356 and tests/test_multilayer_joint
.py:212-241 of the reference.
"""

import numpy as np
import pytest
from scipy.optimize import brentq, minimize

from oracle import joint as oj
from oracle import multilayer as om
from oracle.refraction import snell_crossing
from radar_sounder_simulator_amd.scene import grid_facets, straight_track

C = 299792458.0


def _stack(zs, B):
    N = len(zs)
    o = np.zeros((N, B, 3))
    o[..., 2] = np.asarray(zs)[:, None]
    nrm = np.zeros((N, B, 3))
    nrm[..., 2] = 1.0
    return o, nrm


def test_n1_equals_two_point():
    rng = np.random.default_rng(0)
    B = 64
    p = np.array([0.0, 0.0, 3000.0])
    q = np.column_stack([rng.uniform(-3000, 3000, B), rng.uniform(-3000, 3000, B),
                         -rng.uniform(50, 2000, B)])
    o, nrm = _stack([0.0], B)
    r = oj.joint_crossings(p, q, o, nrm, np.array([1.0, 1.78]), n_newton=6, n_backtrack=4)
    x2 = snell_crossing(p, q, np.zeros(3), np.array([0.0, 0.0, 1.0]), 1.0, 1.78)[0]
    assert r["valid"].all()
    assert np.abs(r["x"][0] - x2).max() < 1e-8


def _ray_parameter_path(p, q, zs, n):
    """Analytic flat-stack path: horizontal slowness conserved."""
    zs = np.asarray(zs)
    levels = np.concatenate([[p[2]], zs, [q[2]]])
    dz = -np.diff(levels)                      # positive thicknesses
    D = np.hypot(*(q[:2] - p[:2]))
    f = lambda s: np.sum(dz * s / np.sqrt(n ** 2 - s ** 2)) - D
    s = brentq(f, 0.0, n.min() * (1 - 1e-15), xtol=1e-15, rtol=1e-15)
    hor = dz * s / np.sqrt(n ** 2 - s ** 2)
    u = (q[:2] - p[:2]) / D
    cum = np.cumsum(hor)[:-1]
    x = np.column_stack([p[0] + cum * u[0], p[1] + cum * u[1], zs])
    opl = np.sum(n * dz / np.sqrt(1 - (s / n) ** 2))
    return x, opl


@pytest.mark.parametrize("N", [2, 3, 6])
def test_flat_stack_equals_ray_parameter_solution(N):
    rng = np.random.default_rng(N)
    zs = -np.sort(rng.uniform(5, 150, N - 1))[::-1]
    zs = np.concatenate([[0.0], -np.sort(-zs)])[:N]
    n = np.concatenate([[1.0], np.linspace(1.3, 1.78, N)])
    p = np.array([0.0, 0.0, 2000.0])
    B = 16
    q = np.column_stack([rng.uniform(-1500, 1500, B), rng.uniform(-1500, 1500, B),
                         np.full(B, zs[-1] - 400.0)])
    o, nrm = _stack(zs, B)
    r = oj.joint_crossings(p, q, o, nrm, n, n_newton=20, n_backtrack=10)
    assert r["valid"].all()
    for b in range(B):
        x, opl = _ray_parameter_path(p, q[b], zs, n)
        assert np.abs(r["x"][:, b] - x).max() < 1e-7
        got = np.sum(n * r["s"][:, b])
        assert got == pytest.approx(opl, abs=1e-7)


def test_tilted_stack_equals_fermat_minimum():
    rng = np.random.default_rng(3)
    tilts = np.deg2rad([3.0, -4.0, 5.0])     # planes stay separated over the scene
    zs = [0.0, -150.0, -400.0]
    n = np.array([1.0, 1.4, 1.78, 2.2])
    nrm1 = np.stack([np.array([np.sin(t), 0.2 * np.sin(t), np.cos(t)]) for t in tilts])
    nrm1 /= np.linalg.norm(nrm1, axis=1, keepdims=True)
    p = np.array([100.0, -50.0, 1500.0])
    for _ in range(4):
        q = np.array([rng.uniform(-400, 400), rng.uniform(-400, 400), -700.0])
        o = np.array([[0.0, 0.0, z] for z in zs])[:, None, :]
        nr = nrm1[:, None, :]
        r = oj.joint_crossings(p, q[None], o, nr, n, n_newton=20, n_backtrack=10)
        assert r["valid"][0]
        basis = oj.plane_basis(nr)[:, 0]

        def opl(v):
            pts = [p] + [o[i, 0] + basis[i] @ v[2 * i:2 * i + 2] for i in range(3)] + [q]
            return sum(n[k] * np.linalg.norm(pts[k + 1] - pts[k]) for k in range(4))

        v0 = np.concatenate([basis[i].T @ (r["x"][i, 0] - o[i, 0]) for i in range(3)])
        res = minimize(opl, v0 + 3.0, method="BFGS", options=dict(gtol=1e-10))
        assert opl(v0) <= res.fun + 1e-8          # no lower optical path nearby
        h = 1e-3                                  # and the path is stationary
        grad = np.array([(opl(v0 + h * e) - opl(v0 - h * e)) / (2 * h) for e in np.eye(6)])
        assert np.abs(grad).max() < 1e-6


def _multilayer_case():
    rng = np.random.default_rng(8)
    surf = grid_facets(3 * rng.standard_normal((13, 13)), 25.0)
    mid = grid_facets(-60 + 2 * rng.standard_normal((13, 13)), 25.0)
    bed = grid_facets(-250 + 10 * rng.standard_normal((13, 13)), 25.0)
    tr = straight_track(2, 40.0, 2500.0)
    kw = dict(mode="coherent", t0=2 * 2400 / C, dt=1 / 48e6, n_samples=1024, c=C,
              gamma=-0.3, k0=2 * np.pi * 60e6 / C, refraction="joint")
    return surf, mid, bed, tr, kw


def test_budget_and_padding_invariance():
    surf, mid, bed, tr, kw = _multilayer_case()
    args = (tr.positions, tr.u_ct, bed, [surf, mid], [1.0, 1.6, 3.15], [0.0, 0.0, 5.0])
    a, _ = om.refracted_cluttergram(*args, **kw)
    b, _ = om.refracted_cluttergram(*args, joint_newton=12, joint_backtrack=8, **kw)
    c, _ = om.refracted_cluttergram(*args, pad_to=4, **kw)
    assert np.abs(a).max() > 0
    assert np.linalg.norm(a - b) / np.linalg.norm(a) < 1e-10
    assert np.linalg.norm(a - c) / np.linalg.norm(a) < 1e-10


def test_joint_single_crossing_equals_sequential():
    surf, mid, bed, tr, kw = _multilayer_case()
    args = (tr.positions, tr.u_ct, bed, [surf], [1.0, 3.15], [0.0, 5.0])
    a, _ = om.refracted_cluttergram(*args, **kw)
    kw["refraction"] = "sequential"
    b, _ = om.refracted_cluttergram(*args, **kw)
    assert np.linalg.norm(a - b) / np.linalg.norm(a) < 1e-9
