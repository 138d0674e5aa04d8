"""Two-point Snell crossing against a flat plane (FP64 oracle).

Restates the component form ``kernels/multilayer.py:196-266`` (identical
operations to ``refraction.py:97-161``): safeguarded Newton on sigma =
sin(theta_rare) over the non-dimensionalised horizontal-closure mismatch
F(sigma) = ar sigma/sqrt(1-sigma^2) + bd v/sqrt(1-v^2) - Lh, v = ratio*sigma,
with a fixed iteration count, bracket [lo, hi] and midpoint fallback; the
residual n1 t/r1 - n2 d2/r2 is recomputed from the crossing geometry and
valid = (h1 h2 < 0) & (|residual| <= 1e-9).

Evaluation order is explicit (no fused operations) so a float64 device
implementation reproduces crossing points, validity and optical paths bit
for bit when it uses correctly rounded sqrt and division.

Pinned by tests/test_oracle_refraction.py: vertical incidence, equal index
(straight line), reciprocity (swap endpoints and indices), Fermat optimum by
brute-force minimisation on flat and tilted planes, 4096-case residual sweep
and iteration-budget stability (``tests/test_refraction.py:63-249``).
"""

import numpy as np

TINY = 1e-30
EPS64 = np.finfo(np.float64).eps


def _dot(a, b):
    return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]


def snell_c(p, q, o, nrm, n1, n2, n_iter, tol=1e-9):
    """p/q/o/nrm are 3-tuples of broadcastable float64 arrays (nrm unit).
    Returns (x (3-tuple), theta1, theta2, s1, s2, residual, valid)."""
    h1 = _dot((p[0] - o[0], p[1] - o[1], p[2] - o[2]), nrm)
    h2 = _dot((q[0] - o[0], q[1] - o[1], q[2] - o[2]), nrm)
    a, b = np.abs(h1), np.abs(h2)
    fp = (p[0] - h1 * nrm[0], p[1] - h1 * nrm[1], p[2] - h1 * nrm[2])
    fq = (q[0] - h2 * nrm[0], q[1] - h2 * nrm[1], q[2] - h2 * nrm[2])
    w = (fq[0] - fp[0], fq[1] - fp[1], fq[2] - fp[2])
    L = np.sqrt(_dot(w, w))
    Lm = np.maximum(L, TINY)
    u = (w[0] / Lm, w[1] / Lm, w[2] / Lm)
    s = np.maximum((a + b) + L, TINY)
    inv_s = 1.0 / s
    ah, bh, Lh = a * inv_s, b * inv_s, L * inv_s
    swap = n1 > n2
    ratio = np.minimum(n1, n2) / np.maximum(n1, n2)
    ar = np.where(swap, bh, ah)
    bd = np.where(swap, ah, bh)
    hi0 = 1.0 - 8.0 * EPS64
    ab = ah + bh
    sig = np.minimum(Lh / np.maximum(np.sqrt(Lh * Lh + ab * ab), TINY), hi0)
    lo = np.zeros_like(Lh)
    hi = hi0 * np.ones_like(Lh)
    for _ in range(n_iter):
        v = ratio * sig
        c1 = np.maximum((1.0 - sig) * (1.0 + sig), TINY)
        c2 = np.maximum((1.0 - v) * (1.0 + v), TINY)
        F = (ar * sig * (1.0 / np.sqrt(c1)) + bd * v * (1.0 / np.sqrt(c2))) - Lh
        le = F <= 0
        lo = np.where(le, sig, lo)
        hi = np.where(le, hi, sig)
        Fp = ar / (c1 * np.sqrt(c1)) + (bd * ratio) / (c2 * np.sqrt(c2))
        with np.errstate(invalid="ignore", divide="ignore", over="ignore"):
            sn = sig - F / np.maximum(Fp, TINY)
        ok = (sn >= lo) & (sn <= hi) & np.isfinite(sn)
        sig = np.where(ok, sn, 0.5 * (lo + hi))
    c1 = np.maximum((1.0 - sig) * (1.0 + sig), TINY)
    tr = ar * sig * (1.0 / np.sqrt(c1))
    t = np.clip(np.where(swap, Lh - tr, tr), 0.0, Lh)
    d2 = Lh - t
    r1 = np.maximum(np.sqrt(ah * ah + t * t), TINY)
    r2 = np.maximum(np.sqrt(bh * bh + d2 * d2), TINY)
    residual = n1 * t / r1 - n2 * d2 / r2
    s1, s2 = s * r1, s * r2
    st = s * t
    x = (fp[0] + st * u[0], fp[1] + st * u[1], fp[2] + st * u[2])
    theta1 = np.arctan2(t, ah)
    theta2 = np.arctan2(d2, bh)
    valid = (h1 * h2 < 0) & (np.abs(residual) <= tol)
    return x, theta1, theta2, s1, s2, residual, valid


def snell_crossing(p, q, plane_point, plane_normal, n1, n2, n_iter=25):
    """Array-form wrapper ((..., 3) inputs) of ``snell_c``."""
    p, q, o, nr = (np.broadcast_to(np.asarray(v, np.float64), np.broadcast_shapes(
        np.shape(p), np.shape(q), np.shape(plane_point), np.shape(plane_normal)))
        for v in (p, q, plane_point, plane_normal))
    comp = lambda v: (v[..., 0], v[..., 1], v[..., 2])
    x, th1, th2, s1, s2, res, ok = snell_c(comp(p), comp(q), comp(o), comp(nr),
                                           np.float64(n1), np.float64(n2), n_iter)
    return np.stack(x, -1), th1, th2, s1, s2, res, ok
