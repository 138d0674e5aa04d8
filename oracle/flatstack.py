"""Flat-stack specialisation of the joint refraction solve -- test
infrastructure only (see oracle/__init__.py).

The joint solve (``refraction_joint.py:249-322``, ``kernels/multilayer.py:
408-457``) finds the stationary path p -> x_0 -> ... -> x_{N-1} -> q through
the local planes it anchored with a mean-plane chain.  When every crossed
interface below the surface is a horizontal plane (firn sublayers, internal
horizons), the anchoring of those planes cannot depend on where the chain
crossed them, and the stationary path is fixed by two numbers: the
horizontal slowness v = n_k sin(theta_k) u_h, conserved below the surface
(Snell at horizontal planes).  Given v the path is explicit, built upward
from q: c_k = sqrt(n_k^2 - |v|^2); the crossing of the first horizontal
plane is x_top = q - sum_{k>=2} d_k (v, -c_k)/c_k (d_k the layer thicknesses,
d_N = z_{N-1} - q_z); from there the ray climbs along u = (-v, c_1) to the
anchored surface plane: s = x_top + lam u, lam = n.(o - x_top)/(n.u).  Snell
at the surface closes the system:

    F(v) = B^T (n_0 (s - p)/|s - p| - (v, -c_1)) = 0,

B an orthonormal basis of the anchored plane, solved by damped Newton with
the joint solve's budgets from the slowness of the ray p -> s0 refracted at
the anchored plane (s0 the pass-1 crossing).  The Jacobian is analytic:
dx_top/dv = -sum_{k>=2} d_k (I/c_k + v v^T/c_k^3), du/dv = [-I; -v^T/c_1],
dlam/dv = -(n.dx_top/dv + lam n.du/dv)/(n.u).  The result is the joint
solve's stationary point (the reference pins that to ~1e-13 m against the
ray-parameter closed form for flat stacks, ``tests/test_refraction_joint.py
:134-168``); validity and sums are then taken from the polyline exactly as
the joint solve takes them.
"""

import numpy as np

from . import joint as oj
from .joint import plane_basis


def _flat_v(s, q, z_flat, n_seg, n_iter=12):
    """Horizontal slowness below a GIVEN surface crossing s reaching q
    through the flat planes (the stationarity check moves s and re-solves
    the rest).  s, q (B, 3); z_flat (N-1,) top-down; n_seg (N+1,) media.
    Solved along u = (q_h - s_h)/R in t = tan(theta) of the slowest layer
    m: g(t) = sum_k d_k n_m t / sqrt(n_k^2 (1 + t^2) - n_m^2 t^2) - R is
    increasing and concave in t (each term is), so Newton from t = 0
    converges monotonically from below.  Returns v (B, 2), c (N, B),
    d (N, B), M^-1 (B, 2, 2) with M = dH/dv."""
    N = len(n_seg) - 1
    zs = np.concatenate([s[None, :, 2], np.broadcast_to(z_flat[:, None], (N - 1, len(s))),
                         q[None, :, 2]])
    d = zs[:-1] - zs[1:]                       # (N, B) layer thicknesses
    nk = n_seg[1:, None]                       # (N, 1)
    nm = float(n_seg[1:].min())
    D = q[:, :2] - s[:, :2]
    R = np.hypot(D[:, 0], D[:, 1])
    t = np.zeros(len(s))
    for _ in range(n_iter):
        den = nk * nk * (1.0 + t * t) - nm * nm * t * t      # (N, B) > 0
        sq = np.sqrt(den)
        g = np.sum(d * nm * t / sq, axis=0) - R
        gp = np.sum(d * nm * nk * nk / (den * sq), axis=0)
        t = t - g / gp
    sig = t / np.sqrt(1.0 + t * t)
    pr = nm * sig
    u = np.where(R[:, None] > 0.0, D / np.where(R > 0.0, R, 1.0)[:, None], 0.0)
    v = pr[:, None] * u
    c = np.sqrt(nk * nk - (pr * pr)[None])
    w = np.sum(d / c ** 3, axis=0)
    M = (np.sum(d / c, axis=0)[:, None, None] * np.eye(2)[None]
         + w[:, None, None] * v[:, :, None] * v[:, None, :])
    return v, c, d, np.linalg.inv(M)


def _v_eval(v, P, Q, pt0, nr0, Bm, zf, nseg):
    """F(v), J = dF/dv, the surface crossing s, c (N, B) and feasibility
    (real c, lam > 0, the climbing ray meets the plane from below)."""
    n0, nl = nseg[0], nseg[1:]
    N, B = len(nl), Q.shape[0]
    vv = np.sum(v * v, 1)
    c = np.sqrt(nl[:, None] ** 2 - vv[None])
    if N > 1:
        d = np.empty((N - 1, B))
        for k in range(1, N):
            d[k - 1] = zf[k - 1] - (zf[k] if k < N - 1 else Q[:, 2])
        xt = np.column_stack([Q[:, :2] - np.sum((d / c[1:])[..., None] * v[None], 0),
                              np.full(B, zf[0])])
        M2 = (np.sum(d / c[1:], 0)[:, None, None] * np.eye(2)[None]
              + np.sum(d / c[1:] ** 3, 0)[:, None, None] * v[:, :, None] * v[:, None, :])
    else:
        xt = Q.copy()
        M2 = np.zeros((B, 2, 2))
    c1 = c[0]
    u = np.column_stack([-v, c1])
    nu = np.sum(nr0 * u, 1)
    lam = np.sum(nr0 * (pt0 - xt), 1) / nu
    s = xt + lam[:, None] * u
    ds = s - P[None]
    L = np.linalg.norm(ds, axis=1)
    dh = ds / L[:, None]
    F = np.einsum("bca,bc->ba", Bm, n0 * dh - np.column_stack([v, -c1]))
    dxt = np.zeros((B, 3, 2))
    dxt[:, :2, :] = -M2
    du = np.zeros((B, 3, 2))
    du[:, 0, 0] = du[:, 1, 1] = -1.0
    du[:, 2, :] = -v / c1[:, None]
    dlam = -(np.einsum("bc,bcj->bj", nr0, dxt)
             + lam[:, None] * np.einsum("bc,bcj->bj", nr0, du)) / nu[:, None]
    dsdv = dxt + u[:, :, None] * dlam[:, None, :] + lam[:, None, None] * du
    Pm = (np.eye(3)[None] - dh[:, :, None] * dh[:, None, :]) / L[:, None, None]
    dk1 = np.zeros((B, 3, 2))
    dk1[:, 0, 0] = dk1[:, 1, 1] = 1.0
    dk1[:, 2, :] = v / c1[:, None]
    J = np.einsum("bca,bcj->baj", Bm, n0 * np.einsum("bcd,bdj->bcj", Pm, dsdv) - dk1)
    ok = np.all(np.isfinite(c), 0) & (lam > 0) & (nu > 0)
    return F, J, s, c, ok


def start_slowness(p, s0, nr0, n0, n1):
    """Horizontal slowness of the ray p -> s0 refracted (n0 -> n1) at the
    plane with upward normal nr0."""
    i = s0 - p[None]
    i = i / np.linalg.norm(i, axis=1)[:, None]
    cosi = -np.sum(nr0 * i, 1)
    r = n0 / n1
    cost = np.sqrt(1.0 - r * r * (1.0 - cosi * cosi))
    kt = n0 * i + (n0 * cosi - n1 * cost)[:, None] * nr0
    return kt[:, :2].copy()


def flat_crossings(p, q, pt0, nr0, z_flat, n_seg, s0, n_newton=6, n_backtrack=4):
    """Stationary path p -> anchored surface plane (pt0, nr0) -> horizontal
    planes z_flat (top-down) -> q.  p (3,), q (B, 3), pt0/nr0/s0 (B, 3) (s0:
    the pass-1 crossing, where the start slowness is taken), z_flat (N-1,),
    n_seg (N+1,).  Damped Newton on v with the joint solve's budgets: of the
    steps lam = 1, 1/2, ... (n_backtrack of them) the first that is feasible
    and lowers |F|^2 is taken; none is taken when none is.  Returns x
    (N, B, 3)."""
    p = np.asarray(p, np.float64)
    q = np.asarray(q, np.float64)
    z_flat = np.asarray(z_flat, np.float64)
    n_seg = np.asarray(n_seg, np.float64)
    Bm = plane_basis(nr0[None])[0]
    N = len(n_seg) - 1
    s0 = s0 - np.sum((s0 - pt0) * nr0, axis=1)[:, None] * nr0     # onto the plane
    with np.errstate(all="ignore"):
        v = start_slowness(p, s0, nr0, n_seg[0], n_seg[1])
        for _ in range(n_newton):
            F, J, s, c, ok = _v_eval(v, p, q, pt0, nr0, Bm, z_flat, n_seg)
            m0 = np.sum(F * F, axis=1)
            det = J[:, 0, 0] * J[:, 1, 1] - J[:, 0, 1] * J[:, 1, 0]
            step = np.column_stack([J[:, 1, 1] * F[:, 0] - J[:, 0, 1] * F[:, 1],
                                    J[:, 0, 0] * F[:, 1] - J[:, 1, 0] * F[:, 0]]) / det[:, None]
            lam = np.ones(len(v))
            done = np.zeros(len(v), bool)
            v_new = v.copy()
            for _ in range(n_backtrack):
                trial = v - lam[:, None] * step
                Ft, _, _, _, okt = _v_eval(trial, p, q, pt0, nr0, Bm, z_flat, n_seg)
                acc = ~done & okt & np.isfinite(Ft).all(axis=1) & (np.sum(Ft * Ft, axis=1) < m0)
                v_new[acc] = trial[acc]
                done |= acc
                lam = np.where(done, lam, 0.5 * lam)
            v = v_new
        F, J, s, c, ok = _v_eval(v, p, q, pt0, nr0, Bm, z_flat, n_seg)
    x = np.empty((N, len(v), 3))
    x[0] = s
    for k in range(1, N):
        top = s[:, 2] if k == 1 else z_flat[k - 2]
        dk = top - z_flat[k - 1]
        x[k, :, :2] = x[k - 1, :, :2] + (dk / c[k - 1])[:, None] * v
        x[k, :, 2] = z_flat[k - 1]
    return x


def flat_path(p, fc, jc, n_leg, eps, att, k0, sig_c, n_newton, n_backtrack):
    """The joint path (``kernels/multilayer.py:408-457``) of a stack whose
    real interfaces below the surface are horizontal planes, with the
    crossings from ``flat_crossings``: the same pass-1 anchoring (only the
    surface's cell matters), the surface start taken from the pass-1 chain,
    the pads' crossings on the straight air leg, then the joint path's
    validity and sums.  Returns the tuple of ``multilayer.sequential_path``."""
    p = np.asarray(p, np.float64)
    k_pad = int(jc[-1])
    pt, nr, r1 = oj.anchor_planes(p, fc, jc, n_leg)
    z_flat = pt[k_pad + 1:, 0, 2]
    xr = flat_crossings(p, fc, pt[k_pad], nr[k_pad], z_flat, n_leg[k_pad:], r1["x"][k_pad],
                        n_newton=n_newton, n_backtrack=n_backtrack)
    s = xr[0]
    zp = pt[:k_pad, 0, 2]
    f = (p[2] - zp[:, None]) / (p[2] - s[None, :, 2])
    xp = p[None, None] + f[..., None] * (s[None] - p[None, None])
    xp[..., 2] = zp[:, None]
    x = np.concatenate([xp, xr], 0)
    r2 = oj.polyline(x, p, fc, nr, n_leg, 1e-9)
    return oj.path_sums(r2, k_pad, n_leg, eps, att, k0, sig_c)
