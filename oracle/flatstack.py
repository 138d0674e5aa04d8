"""Flat-stack specialisation of the joint refraction solve -- test
infrastructure only (see oracle/__init__.py).

The joint solve (``refraction_joint.py:249-322``, ``kernels/multilayer.py:
408-457``) finds the stationary path p -> x_0 -> ... -> x_{N-1} -> q through
the local planes it anchored with a mean-plane chain.  When every crossed
interface below the surface is a flat horizontal plane (firn sublayers,
internal horizons), the anchoring of those planes cannot depend on where
the chain crossed them, and the stationary path has two unknowns: the
surface crossing s on the surface's anchored plane.  Below s the horizontal
slowness v = n_k sin(theta_k) u_h is conserved (Snell at horizontal
planes); it solves

    H(v; s) = sum_k d_k v / c_k(|v|) - (q_h - s_h) = 0,   c_k = sqrt(n_k^2 - |v|^2),

with d_1 = s_z - z_1, d_k = z_{k-1} - z_k, d_N = z_{N-1} - q_z, and Snell at
the surface closes the system:

    F(s) = B^T (n_0 (s - p)/|s - p| - k_1) = 0,   k_1 = (v, -c_1),

B an orthonormal basis of the anchored plane.  Both are solved by Newton
(implicit derivative dv/ds = -M^-1 [I | v/c_1], M = dH/dv).  The result is
the same stationary point as the joint solve's (which the reference pins to
~1e-13 m against the ray-parameter closed form for flat stacks,
``tests/test_refraction_joint.py:134-168``); the validity and path sums are
then taken from the polyline exactly as the joint solve takes them.
"""

import numpy as np

from . import joint as oj
from .joint import plane_basis


def _flat_v(s, q, z_flat, n_seg, n_iter=12):
    """Horizontal slowness below s reaching q through the flat planes.
    s, q (B, 3); z_flat (N-1,) top-down; n_seg (N+1,) media indices.
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


def _residual(s, p, q, Bm, z_flat, n_seg):
    """F(s) (B, 2) and the pieces of its Jacobian."""
    v, c, d, Minv = _flat_v(s, q, z_flat, n_seg)
    ds = s - p[None]
    L = np.linalg.norm(ds, axis=1)
    dh = ds / L[:, None]
    k1 = np.column_stack([v, -c[0]])
    F = np.einsum("bca,bc->ba", Bm, n_seg[0] * dh - k1)
    return F, v, c, d, Minv, dh, L


def flat_crossings(p, q, pt0, nr0, z_flat, n_seg, s0, n_newton=6, n_backtrack=4):
    """Stationary path p -> anchored surface plane (pt0, nr0) -> horizontal
    planes z_flat -> q.  p (3,), q (B, 3), pt0/nr0/s0 (B, 3) (s0: start on or
    near the plane), z_flat (N-1,), n_seg (N+1,).  Damped Newton with the
    joint solve's budgets: of the steps lam = 1, 1/2, ... (n_backtrack of
    them) the first that lowers |F|^2 and keeps the surface crossing above
    the first flat plane is taken; none is taken when none does.  Returns
    x (N, B, 3)."""
    p = np.asarray(p, np.float64)
    q = np.asarray(q, np.float64)
    z_flat = np.asarray(z_flat, np.float64)
    n_seg = np.asarray(n_seg, np.float64)
    z_top = z_flat[0] if len(z_flat) else q[:, 2]
    Bm = plane_basis(nr0[None])[0]             # (B, 3, 2)
    s = s0 - np.sum((s0 - pt0) * nr0, axis=1)[:, None] * nr0
    with np.errstate(all="ignore"):
        for _ in range(n_newton):
            F, v, c, d, Minv, dh, L = _residual(s, p, q, Bm, z_flat, n_seg)
            m0 = np.sum(F * F, axis=1)
            dHds = np.zeros((len(s), 2, 3))
            dHds[:, 0, 0] = dHds[:, 1, 1] = 1.0
            dHds[:, :, 2] = v / c[0][:, None]
            dvds = -np.einsum("bij,bjk->bik", Minv, dHds)         # dv/ds
            dk1 = np.concatenate(
                [dvds, (np.einsum("bi,bik->bk", v, dvds) / c[0][:, None])[:, None]], axis=1)
            Pm = (np.eye(3)[None] - dh[:, :, None] * dh[:, None, :]) / L[:, None, None]
            J = np.einsum("bca,bcd,bde->bae", Bm, n_seg[0] * Pm - dk1, Bm)
            det = J[:, 0, 0] * J[:, 1, 1] - J[:, 0, 1] * J[:, 1, 0]
            step = np.column_stack([J[:, 1, 1] * F[:, 0] - J[:, 0, 1] * F[:, 1],
                                    J[:, 0, 0] * F[:, 1] - J[:, 1, 0] * F[:, 0]]) / det[:, None]
            lam = np.ones(len(s))
            done = np.zeros(len(s), bool)
            s_new = s.copy()
            for _ in range(n_backtrack):
                trial = s - lam[:, None] * np.einsum("bca,ba->bc", Bm, step)
                Ft = _residual(trial, p, q, Bm, z_flat, n_seg)[0]
                ok = (~done & np.isfinite(Ft).all(axis=1) & (np.sum(Ft * Ft, axis=1) < m0)
                      & (trial[:, 2] > z_top))
                s_new[ok] = trial[ok]
                done |= ok
                lam = np.where(done, lam, 0.5 * lam)
            s = s_new
        v, c, d, _ = _flat_v(s, q, z_flat, n_seg)
    N = len(n_seg) - 1
    x = np.empty((N, len(s), 3))
    x[0] = s
    off = np.zeros((len(s), 2))
    for k in range(1, N):
        off = off + (d[k - 1] / c[k - 1])[:, None] * v
        x[k, :, :2] = s[:, :2] + off
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
