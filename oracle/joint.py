"""Joint N-interface refraction solve (FP64 oracle), restating
``refraction_joint.py:109-322`` and the joint branch of
``kernels/multilayer.py:286-330, 408-457``.

Unknowns: two in-plane offsets per crossed local plane.  Residual
F_i = B_i^T (n_i d_i - n_{i+1} d_{i+1}) (tangential Snell + plane of
incidence); Jacobian block-tridiagonal with A_i = B_i^T (W_i + W_{i+1}) B_i
(+ Levenberg eps^0.75 I), C_i = -B_i^T W_{i+1} B_{i+1}, W_k = n_k (I - d_k
d_k^T)/s_k, solved per Newton step by block Thomas elimination; damped by
step halving on |F|^2 (first decreasing candidate of 2^-j, j < n_backtrack;
none -> step rejected); initialised from the sequential two-point chain;
coordinates scaled by |p - q|.  Validity: residual <= 1e-9, every plane
straddled by its neighbouring path points, no evanescent crossing, finite.

Batch layout: a leading interface axis (N, B, ...) with B targets, p (3,)
one platform position.

Pinned by tests/test_oracle_joint.py: N = 1 equals the two-point solve, flat
parallel stacks equal the analytic ray-parameter solution, tilted stacks
equal a direct Fermat minimisation, a doubled budget moves nothing, no-op
padding planes change nothing (``tests/test_refraction_joint.py:47-356``,
``tests/test_multilayer_joint.py:212-241``).
"""

import numpy as np

from .multilayer import C_MIN, grid_consts
from .physics import fresnel_te
from .refraction import snell_c

TINY = 1e-30
EPS = np.finfo(np.float64).eps


def _dot(a, b):
    return (a[..., 0] * b[..., 0] + a[..., 1] * b[..., 1]) + a[..., 2] * b[..., 2]


def chain(p, q, o, nrm, n_seg, n_iter):
    """Sequential two-point chain: x (N, B, 3), ok (B,)."""
    N = o.shape[0]
    cur = tuple(np.broadcast_to(p[..., i], q.shape[:-1]) for i in range(3))
    qc = (q[..., 0], q[..., 1], q[..., 2])
    xs, ok = [], np.ones(q.shape[:-1], bool)
    for i in range(N):
        x, _, _, _, _, _, v = snell_c(cur, qc, (o[i, ..., 0], o[i, ..., 1], o[i, ..., 2]),
                                      (nrm[i, ..., 0], nrm[i, ..., 1], nrm[i, ..., 2]),
                                      n_seg[i], n_seg[i + 1], n_iter)
        xs.append(np.stack(x, -1))
        ok = ok & v
        cur = x
    return np.stack(xs, 0), ok


def polyline(x, p, q, nrm, n_seg, tol, converged=None):
    """Angles, segments, recomputed residual and validity of p -> x -> q."""
    B = q.shape[:-1]
    P = np.concatenate([np.broadcast_to(p, B + (3,))[None], x, q[None]], 0)
    d = P[1:] - P[:-1]
    s = np.sqrt(np.sum(d * d, -1))
    dh = d / np.maximum(s, TINY)[..., None]
    din, dout = dh[:-1], dh[1:]
    ct1 = np.sum(nrm * din, -1)
    ct2 = np.sum(nrm * dout, -1)
    t1 = din - ct1[..., None] * nrm
    t2 = dout - ct2[..., None] * nrm
    st1 = np.sqrt(np.sum(t1 * t1, -1))
    st2 = np.sqrt(np.sum(t2 * t2, -1))
    th1 = np.arctan2(st1, np.abs(ct1))
    th2 = np.arctan2(st2, np.abs(ct2))
    ns = n_seg.reshape((-1,) + (1,) * len(B))
    dG = ns[:-1, ..., None] * din - ns[1:, ..., None] * dout
    rt = dG - np.sum(nrm * dG, -1, keepdims=True) * nrm
    residual = np.max(np.sqrt(np.sum(rt * rt, -1)), axis=0)
    hp = np.sum(nrm * (P[:-2] - x), -1)
    hn = np.sum(nrm * (P[2:] - x), -1)
    sides = np.all(hp * hn < 0, axis=0)
    evan = np.any(ns[:-1] * st1 > ns[1:] * (1.0 + 64.0 * EPS), axis=0)
    fin = (np.all(np.isfinite(x), axis=(0, -1)) & np.isfinite(residual)
           & np.all(np.isfinite(s), axis=0))
    if converged is None:
        converged = (residual <= tol) & fin
    valid = converged & sides & ~evan & fin
    return dict(x=x, theta1=th1, theta2=th2, s=s, residual=residual, valid=valid)


def plane_basis(nrm):
    """(N, B, 3) unit normals -> (N, B, 3, 2) orthonormal in-plane basis."""
    ez = np.zeros_like(nrm)
    ez[..., 2] = 1.0
    ex = np.zeros_like(nrm)
    ex[..., 0] = 1.0
    a = np.where((np.abs(nrm[..., 2]) < 0.9)[..., None], ez, ex)
    u = np.cross(a, nrm)
    u = u / np.maximum(np.sqrt(np.sum(u * u, -1, keepdims=True)), TINY)
    return np.stack([u, np.cross(nrm, u)], -1)


def inv2(m):
    a, b, c, d = m[..., 0, 0], m[..., 0, 1], m[..., 1, 0], m[..., 1, 1]
    det = a * d - b * c
    det = np.where(np.abs(det) < TINY, TINY, det)
    out = np.stack([np.stack([d, -b], -1), np.stack([-c, a], -1)], -2)
    return out / det[..., None, None]


def block_thomas(A, Cu, b):
    """Solve blocks A_i (N,B,2,2), upper Cu_i (N-1,B,2,2) (lower = Cu^T), rhs b."""
    N = A.shape[0]
    dinv = np.empty_like(A)
    y = np.empty_like(b)
    dprev = np.zeros_like(A[0])
    yprev = np.zeros_like(b[0])
    for i in range(N):
        cp = Cu[i - 1] if i > 0 else np.zeros_like(A[0])
        cpt = np.swapaxes(cp, -1, -2) @ dprev
        dinv[i] = inv2(A[i] - cpt @ cp)
        y[i] = b[i] - np.einsum("...ij,...j->...i", cpt, yprev)
        dprev, yprev = dinv[i], y[i]
    delta = np.empty_like(b)
    dn = np.zeros_like(b[0])
    for i in range(N - 1, -1, -1):
        cn = Cu[i] if i < N - 1 else np.zeros_like(A[0])
        rhs = y[i] - np.einsum("...ij,...j->...i", cn, dn)
        delta[i] = np.einsum("...ij,...j->...i", dinv[i], rhs)
        dn = delta[i]
    return delta


def joint_crossings(p, q, o, nrm, n_seg, n_newton=20, n_backtrack=10, n_iter_init=25,
                    tol=1e-9):
    """p (3,), q (B, 3), o/nrm (N, B, 3), n_seg (N+1,)."""
    x0, _ = chain(p, q, o, nrm, n_seg, n_iter_init)
    scale = np.maximum(np.sqrt(np.sum((p - q) ** 2, -1)), TINY)   # (B,)
    inv = (1.0 / scale)[..., None]
    ps, qs, os_ = p * inv, q * inv, o * inv
    basis = plane_basis(nrm)
    ab = np.einsum("...kj,...k->...j", basis, x0 * inv - os_)
    lam = 2.0 ** (-np.arange(n_backtrack, dtype=np.float64))
    mu = EPS ** 0.75
    i2 = np.eye(2)
    i3 = np.eye(3)
    ns = n_seg.reshape((-1,) + (1,) * (q.ndim - 1))

    def geom(ab):
        x = os_ + np.einsum("...kj,...j->...k", basis, ab)
        pl = np.concatenate([ps[None], x, qs[None]], 0)
        d = pl[1:] - pl[:-1]
        s = np.maximum(np.sqrt(np.sum(d * d, -1)), TINY)
        dh = d / s[..., None]
        g = ns[..., None] * dh
        f = np.einsum("...kj,...k->...j", basis, g[:-1] - g[1:])
        return x, s, dh, f

    for _ in range(n_newton):
        _, s, dh, f = geom(ab)
        w = (ns / s)[..., None, None] * (i3 - dh[..., :, None] * dh[..., None, :])
        a = np.einsum("...ki,...kl,...lj->...ij", basis, w[:-1] + w[1:], basis) + mu * i2
        c = -np.einsum("...ki,...kl,...lj->...ij", basis[:-1], w[1:-1], basis[1:])
        delta = block_thomas(a, c, -f)
        m0 = np.sum(f * f, axis=(0, -1))
        lam_sel = np.zeros_like(m0)
        chosen = np.zeros(m0.shape, bool)
        for lj in lam:
            with np.errstate(invalid="ignore", over="ignore"):
                mc = np.sum(geom(ab + lj * delta)[3] ** 2, axis=(0, -1))
                dec = (mc < m0) & ~chosen
            lam_sel = np.where(dec, lj, lam_sel)
            chosen = chosen | dec
        ab = ab + lam_sel[None, ..., None] * delta
    x = geom(ab)[0] * scale[..., None]
    return polyline(x, p, q, nrm, n_seg, tol)


def sequential_chain(p, q, o, nrm, n_seg, n_iter=25):
    x, ok = chain(p, q, o, nrm, n_seg, n_iter)
    return polyline(x, p, q, nrm, n_seg, None, converged=ok)


def joint_consts(crossed, pad_to, z_platform):
    """Stacked lookup constants with index-matched no-op air planes padded
    at the top (``kernels/multilayer.py:286-330``), float64."""
    per = [grid_consts(f) for f in crossed]
    k = pad_to - len(crossed)
    if k < 0:
        raise ValueError(f"pad_to={pad_to} < {len(crossed)} crossed interfaces")
    tabs, coefs, shps, mps, mns = [], [], [], [], []
    if k:
        z_surf = float(crossed[0].centers[:, 2].max())
        gap = float(z_platform) - z_surf
        if gap <= 0:
            raise ValueError("joint refraction padding needs the platform above the "
                             f"surface (gap {gap:.1f} m)")
        zs = z_surf + gap * (np.arange(k, 0, -1) / (k + 1.0))
        for z in zs:
            tabs.append(np.array([[[0.0, 0.0, z, 0.0, 0.0, 1.0]]]))
            coefs.append(np.zeros((3, 2)))
            shps.append((1, 1))
            mps.append(np.array([0.0, 0.0, z]))
            mns.append(np.array([0.0, 0.0, 1.0]))
    for (tab, coef, mp, mn), f in zip(per, crossed):
        tabs.append(tab)
        coefs.append(coef)
        shps.append(f.grid_shape)
        mps.append(mp)
        mns.append(mn)
    return tabs, coefs, shps, np.array(mps), np.array(mns), k


def anchor_planes(p, fc, jc, n_leg):
    """Pass 1 of the joint path (``kernels/multilayer.py:420-433``): the
    sequential chain against the mean planes (10 iterations), then the cell
    lookup of every crossing.  Returns the local planes pt, nr (N, B, 3) and
    the chain."""
    tabs, coefs, shps, mp, mn, k_pad = jc
    N = len(tabs)
    B = fc.shape[0]
    o1 = np.broadcast_to(mp[:, None, :], (N, B, 3))
    n1 = np.broadcast_to(mn[:, None, :], (N, B, 3))
    r1 = sequential_chain(p, fc, o1, n1, n_leg, n_iter=10)
    pt = np.empty((N, B, 3))
    nr = np.empty((N, B, 3))
    for i in range(N):
        xi = r1["x"][i]
        cf = coefs[i]
        rc0 = (xi[:, 0] * cf[0, 0] + xi[:, 1] * cf[1, 0]) + cf[2, 0]
        rc1 = (xi[:, 0] * cf[0, 1] + xi[:, 1] * cf[1, 1]) + cf[2, 1]
        row = np.clip(np.round(rc0), 0, shps[i][0] - 1).astype(np.int64)
        col = np.clip(np.round(rc1), 0, shps[i][1] - 1).astype(np.int64)
        g6 = tabs[i][row, col]
        pt[i] = g6[:, :3]
        v = g6[:, 3:]
        nr[i] = v / np.linalg.norm(v, axis=-1, keepdims=True)
    return pt, nr, r1


def path_sums(r2, k_pad, n_leg, eps, att, k0, sig_c):
    """Path sums of a solved polyline (``kernels/multilayer.py:436-457``):
    pads booked with the surface cosine; returns the tuple of
    ``multilayer.sequential_path``."""
    N = r2["x"].shape[0]
    s_up = r2["s"][:-1]
    c_inc = np.maximum(np.cos(r2["theta1"]), C_MIN)
    c_surf = c_inc[k_pad]
    is_pad = (np.arange(N) < k_pad)[:, None]
    g = fresnel_te(eps[:-1, None], eps[1:, None], c_inc)[0]
    tau2 = np.prod(1.0 - g * g, axis=0)
    if sig_c is not None:
        kt = k0 * (n_leg[:-1, None] * np.cos(r2["theta1"])
                   - n_leg[1:, None] * np.cos(r2["theta2"]))
        tau2 = tau2 * np.prod(np.exp(-2.0 * (sig_c[:, None] * kt) ** 2), axis=0)
    c_par = np.where(is_pad, c_surf[None], c_inc)
    nl = n_leg[:-1, None]
    opl = np.sum(nl * s_up, axis=0)
    loss = np.sum(s_up * (att[:-1, None] / 1000.0), axis=0)
    sum_perp = np.sum(s_up / nl, axis=0)
    sum_par = np.sum(s_up / (nl * c_par * c_par), axis=0)
    c_last = np.maximum(np.cos(r2["theta2"][-1]), C_MIN)
    x = r2["x"]
    comp = lambda v: (v[:, 0], v[:, 1], v[:, 2])
    return (comp(x[-1]), r2["valid"], opl, loss, sum_par, sum_perp, tau2, c_surf, c_last,
            comp(x[0]))


def joint_path(p, fc, jc, n_leg, eps, att, k0, sig_c, n_newton, n_backtrack):
    """Joint refracted path for one trace (``kernels/multilayer.py:408-457``);
    returns the same tuple as ``multilayer.sequential_path``."""
    p = np.asarray(p, np.float64)
    pt, nr, _ = anchor_planes(p, fc, jc, n_leg)
    r2 = joint_crossings(p, fc, pt, nr, n_leg, n_newton=n_newton,
                         n_backtrack=n_backtrack)
    return path_sums(r2, jc[-1], n_leg, eps, att, k0, sig_c)
