"""Pins of the two-point Snell oracle (oracle/refraction.py) and the scalar
physics (oracle/physics.py) -- CPU.

Mirrors tests/test_refraction.py:63-272 and tests/test_fresnel_spreading.py
:26-120 of the reference: closed forms (vertical incidence, equal index,
reciprocity), the Fermat optimum found by direct minimisation, a seeded
geometry sweep, the iteration budget, masking without NaN; TE Fresnel
identities and ray-tube spreading by finite-difference forward tracing.
"""

import numpy as np
import pytest
from scipy.optimize import minimize

from oracle.physics import fresnel_normal, fresnel_te, refraction_spreading
from oracle.refraction import snell_crossing

N_ICE = float(np.sqrt(3.17))
UP = np.array([0.0, 0.0, 1.0])


def test_vertical_incidence():
    x, th1, th2, s1, s2, res, ok = snell_crossing([3.0, -2.0, 500.0], [3.0, -2.0, -200.0],
                                                  np.zeros(3), UP, 1.0, N_ICE)
    np.testing.assert_allclose(x, [3.0, -2.0, 0.0], atol=1e-12)
    assert th1 == 0.0 and th2 == 0.0 and ok
    assert s1 == pytest.approx(500.0) and s2 == pytest.approx(200.0)


def test_equal_index_is_straight_line():
    p, q = np.array([0.0, 0.0, 300.0]), np.array([400.0, 120.0, -100.0])
    x = snell_crossing(p, q, np.zeros(3), UP, 1.7, 1.7)[0]
    t = p[2] / (p[2] - q[2])
    np.testing.assert_allclose(x, p + t * (q - p), atol=1e-9)


def test_reciprocity():
    rng = np.random.default_rng(1)
    for _ in range(20):
        p = np.array([*rng.uniform(-500, 500, 2), rng.uniform(100, 3000)])
        q = np.array([*rng.uniform(-2000, 2000, 2), -rng.uniform(10, 2000)])
        a = snell_crossing(p, q, np.zeros(3), UP, 1.0, N_ICE)
        b = snell_crossing(q, p, np.zeros(3), UP, N_ICE, 1.0)
        np.testing.assert_allclose(a[0], b[0], atol=1e-7)
        assert a[1] == pytest.approx(b[2], abs=1e-10)


@pytest.mark.parametrize("tilt_deg", [0.0, 12.0])
def test_fermat_optimum(tilt_deg):
    """The crossing minimises n1|p - x| + n2|x - q| over the (tilted) plane."""
    t = np.deg2rad(tilt_deg)
    nrm = np.array([np.sin(t), 0.0, np.cos(t)])
    u1 = np.array([np.cos(t), 0.0, -np.sin(t)])
    u2 = np.array([0.0, 1.0, 0.0])
    rng = np.random.default_rng(4)
    for _ in range(8):
        p = np.array([*rng.uniform(-300, 300, 2), rng.uniform(500, 2000)])
        q = np.array([*rng.uniform(-1500, 1500, 2), -rng.uniform(50, 800)])
        x = snell_crossing(p, q, np.zeros(3), nrm, 1.0, N_ICE)[0]

        def opl(ab):
            y = ab[0] * u1 + ab[1] * u2
            return np.linalg.norm(p - y) + N_ICE * np.linalg.norm(y - q)

        r = minimize(opl, [x @ u1 + 5.0, x @ u2 - 5.0], method="Nelder-Mead",
                     options=dict(xatol=1e-9, fatol=1e-12, maxiter=20000))
        y = r.x[0] * u1 + r.x[1] * u2
        assert np.linalg.norm(y - x) < 1e-4
        assert opl([x @ u1, x @ u2]) <= r.fun + 1e-9


def test_sweep_residual_and_budget():
    """4096 seeded stratospheric/airborne geometries: valid with residual
    <= 1e-9 in float64, and a 4x budget moves no crossing by > 1e-8 m."""
    rng = np.random.default_rng(7)
    n = 4096
    p = np.column_stack([np.zeros(n), np.zeros(n),
                         np.where(rng.random(n) < 0.5, rng.uniform(300, 3000, n),
                                  rng.uniform(14000, 20000, n))])
    q = np.column_stack([rng.uniform(-7000, 7000, n), rng.uniform(-7000, 7000, n),
                         -rng.uniform(10, 4000, n)])
    ang = np.deg2rad(rng.uniform(0, 20, n))
    az = rng.uniform(0, 2 * np.pi, n)
    nrm = np.column_stack([np.sin(ang) * np.cos(az), np.sin(ang) * np.sin(az), np.cos(ang)])
    a = snell_crossing(p, q, np.zeros(3), nrm, 1.0, N_ICE, n_iter=25)
    opposite = (p * nrm).sum(1) * (q * nrm).sum(1) < 0
    np.testing.assert_array_equal(a[6], opposite)   # same side <=> masked
    assert opposite.sum() > 3500 and np.abs(a[5][opposite]).max() <= 1e-9
    b = snell_crossing(p, q, np.zeros(3), nrm, 1.0, N_ICE, n_iter=100)
    assert np.abs(a[0] - b[0])[opposite].max() < 1e-8


def test_same_side_masked_not_nan():
    x, th1, th2, s1, s2, res, ok = snell_crossing([0.0, 0.0, 100.0], [50.0, 0.0, 20.0],
                                                  np.zeros(3), UP, 1.0, N_ICE)
    assert not ok
    assert np.all(np.isfinite(x)) and np.isfinite(s1) and np.isfinite(res)


def test_fresnel_te_identities():
    for e1, e2 in [(1.0, 3.17), (3.17, 6.0), (2.0, 1.5)]:
        g = fresnel_te(e1, e2, 1.0)[0]
        assert g == pytest.approx(fresnel_normal(e1, e2), abs=1e-15)
    th = np.linspace(0.0, np.deg2rad(89.0), 200)
    for e1, e2 in [(1.0, 3.17), (3.17, 6.0), (3.17, 1.0)]:
        c1 = np.cos(th)
        g, tau, c2, tir = fresnel_te(e1, e2, c1)
        t_pow = (np.sqrt(e2) * c2) / (np.sqrt(e1) * c1) * tau ** 2
        np.testing.assert_allclose((g ** 2 + t_pow)[~tir], 1.0, atol=1e-12)
    g, tau, c2, _ = fresnel_te(1.0, 3.17, np.cos(np.deg2rad(np.linspace(0, 80, 50))))
    gu, tau_u, c2u, _ = fresnel_te(3.17, 1.0, c2)
    np.testing.assert_allclose(tau * tau_u, 1.0 - g ** 2, atol=1e-12)
    np.testing.assert_allclose(gu, -g, atol=1e-12)
    crit = np.arcsin(1.0 / N_ICE)
    _, _, _, tir = fresnel_te(3.17, 1.0, np.cos([crit - 0.01, crit + 0.01]))
    assert list(tir) == [False, True]


def _forward(p, d, z_t, n1, n2):
    d = d / np.linalg.norm(d)
    t = -p[2] / d[2]
    x = p + t * d
    cos1 = -d[2]
    sv = d + cos1 * UP
    sin1 = np.linalg.norm(sv)
    sin2 = n1 * sin1 / n2
    cos2 = np.sqrt(1 - sin2 ** 2)
    d2 = sin2 * sv / sin1 - cos2 * UP
    q = x + (z_t - x[2]) / d2[2] * d2
    return x, q, t, np.linalg.norm(q - x), cos1, cos2, sin1


@pytest.mark.parametrize("theta_deg", [10.0, 30.0, 50.0, 65.0])
def test_spreading_vs_ray_tube(theta_deg):
    n1, n2 = 1.0, N_ICE
    p = np.array([0.0, 0.0, 700.0])
    th = np.deg2rad(theta_deg)
    d0 = np.array([np.sin(th), 0.0, -np.cos(th)])
    eps = 1e-6
    dth = np.array([np.sin(th + eps), 0.0, -np.cos(th + eps)])
    dph = np.array([d0[0] * np.cos(eps), d0[0] * np.sin(eps), d0[2]])
    x0, q0, s1, s2, c1, c2, sin1 = _forward(p, d0, -400.0, n1, n2)
    _, qa, *_ = _forward(p, dth, -400.0, n1, n2)
    _, qb, *_ = _forward(p, dph, -400.0, n1, n2)
    ray = (q0 - x0) / np.linalg.norm(q0 - x0)
    w_par = np.linalg.norm((qa - q0) - np.dot(qa - q0, ray) * ray)
    w_perp = np.linalg.norm((qb - q0) - np.dot(qb - q0, ray) * ray)
    l_par, l_perp = refraction_spreading([s1, s2], [c1, c2], [n1, n2])
    assert w_par / eps * (c1 / c2) == pytest.approx(l_par, rel=1e-4)
    assert w_perp / (sin1 * eps) == pytest.approx(l_perp, rel=1e-6)
    l1, l2 = refraction_spreading([1234.5], [0.7], [1.0])
    assert l1 == pytest.approx(1234.5) and l2 == pytest.approx(1234.5)
