"""Surface coherent LPA kernel (FP64 oracle), restating ``kernels/coherent.py``.

Per (trace p, facet): d = p - c, r = |d|, cos = d.n / r, rhat = d / r,
s_i = sinc((rhat . e_i) * k/pi) (normalised sinc), amp = (k/2pi) gamma cos A
s1 s2 / (r r) [x off-specular taper], phase = -2k (r - r_ref), contribution
j amp e^{j phase} (``:107-128``); the rough variant adds the incoherent
sqrt(D_Phi) phi_r term with the same non-phase factor and multiplies the
smooth term by exp(-sigma^2 K^2 / 2), K = 2k cos (``:131-168``); a
non-isotropic antenna multiplies by g^2 at the departure direction
(``:216-218``); twtt = 2r/c, bin = floor((twtt - t0)/dt), out-of-window parts
are dropped with their |.|^2 booked per trace (``:219-236``); the per-trace
reference range r_ref = |p - mean(centres)| is folded back as
exp(-j ((2k r_ref) mod 2pi)) (``:309-313, 359``).

Everything is float64 / complex128 (the reference kernel runs f32/c64 on
device; its own float64 truth is the same formula through ``xp=np``,
``tests/test_coherent_kernel.py:70-81``).  Evaluation order is explicit so a
float64 device implementation without FMA contraction matches bins and
masks bit for bit.  No facet culling or blocking: skipped blocks are silent
in the reference (``geometry.py:27-38``), so the all-facet sum is the
reference's ``field`` exactly; ``dropped`` here counts every facet (the
reference's culled ``dropped`` is a lower bound when culling is on).

Pinned by tests/test_oracle_coherent.py: small-facet limit vs the single
sample brute-force expression (1e-10), single facet vs sub-lambda brute-force
integration (LPA validity table), facet-binned flat plate vs brute force and
vs the continuous plate, energy bookkeeping (binned + dropped = total),
split sides sum to combined, interp bins preserve the complex sum, image-
method plate anchor, sigma = 0 rough == smooth exactly.
"""

import numpy as np

from . import antenna as _ant
from .roughness import d_phi, mean_attenuation

TWO_PI = 2.0 * np.pi


def dot3(a, b):
    """((ax bx + ay by) + az bz) on (..., 3) arrays (explicit order)."""
    return (a[..., 0] * b[..., 0] + a[..., 1] * b[..., 1]) + a[..., 2] * b[..., 2]


def sinc(x):
    """np.sinc (normalised): sin(pi x)/(pi x), 1 at x = 0."""
    return np.sinc(x)


def off_specular_taper(cos, taper_s):
    """exp(-(1 - c^2)/(c^2 (2 s^2))), c = clip(cos, 1e-9, 1) (``:94-104``)."""
    c = np.clip(cos, 1e-9, 1.0)
    c2 = c * c
    return np.exp(-(1.0 - c2) / (c2 * (2.0 * taper_s * taper_s)))


def twtt_bin(twtt, t0, dt):
    """floor((twtt - t0)/dt) as int64 (``geometry.py:18-24``)."""
    return np.floor((twtt - t0) / dt).astype(np.int64)


def reference_range(positions, centers):
    """Per-trace |p - mean(centres)| (``coherent.py:311-312``)."""
    m = np.asarray(centers, np.float64).mean(0)
    d = np.asarray(positions, np.float64) - m
    return np.sqrt(dot3(d, d))


def facet_terms(p, centers, normals, areas, e1, e2, k, gamma, r_ref,
                taper_s=None, roughness=None, d_phi_area=False):
    """Complex contributions (before antenna) and ranges for one trace."""
    d = p[None, :] - centers
    r = np.sqrt(dot3(d, d))
    cos = dot3(d, normals) / r
    rhat = d / r[:, None]
    d1 = dot3(rhat, e1)
    d2 = dot3(rhat, e2)
    s1 = sinc(d1 * (k / np.pi))
    s2 = sinc(d2 * (k / np.pi))
    amp = (k / TWO_PI) * gamma * cos * areas * s1 * s2 / (r * r)
    if taper_s is not None:
        amp = amp * off_specular_taper(cos, taper_s)
    phase = (-2.0 * k) * (r - r_ref)
    if roughness is None:
        return 1j * amp * np.exp(1j * phase), r, d
    sigma, lc, phasors, n_terms, acf = roughness
    l1 = np.sqrt(dot3(e1, e1))
    l2 = np.sqrt(dot3(e2, e2))
    kk = 2.0 * k * cos
    live = areas > 0
    l1s, l2s = np.where(live, l1, 1.0), np.where(live, l2, 1.0)
    dp = d_phi(sigma, lc, kk, 2.0 * k * d1 / l1s, 2.0 * k * d2 / l2s, l1s,
               l2s, n_terms=n_terms, area_only=d_phi_area, acf=acf)
    amp_i = np.where(live, (k / TWO_PI) * gamma * cos / (r * r) * np.sqrt(dp),
                     0.0)
    contrib = amp * mean_attenuation(sigma, kk) + amp_i * phasors
    return 1j * contrib * np.exp(1j * phase), r, d


def coherent_cluttergram(positions, u_ct, centers, normals, areas, e1, e2, *,
                         k, gamma, t0, dt, n_samples, c, split_sides=False,
                         interp_bins=False, pattern=None, roughness=None,
                         taper_s=None, d_phi_area=False, return_bins=False):
    """FP64 field (T, n[, 2]) complex128 and dropped (T,) float64.

    ``roughness``: None or (sigma, l, phasors (F,), n_terms[, acf]).
    ``pattern``: None or (kind, pv (T, ...), pa, pb).  ``return_bins`` also
    returns the per-pair (T, F) int64 bin (before side offset) and the
    validity mask (bin inside the window)."""
    pos = np.asarray(positions, np.float64)
    uct = np.asarray(u_ct, np.float64)
    C = np.asarray(centers, np.float64)
    N = np.asarray(normals, np.float64)
    A = np.asarray(areas, np.float64)
    E1 = np.asarray(e1, np.float64)
    E2 = np.asarray(e2, np.float64)
    nT, nF = pos.shape[0], C.shape[0]
    g = np.asarray(gamma, np.float64)
    if g.ndim > 0 and g.shape != (nF,):
        raise ValueError(f"per-facet gamma shape {g.shape} != ({nF},)")
    rough = None
    if roughness is not None:
        sigma, lc, phasors, n_terms, *rest = roughness
        acf = rest[0] if rest else "gaussian"
        if acf == "exponential" and not d_phi_area:
            raise ValueError("roughness acf='exponential' requires the "
                             "area-only D_Phi (d_phi_area / grazing fix)")
        rough = (float(sigma), float(lc),
                 np.asarray(phasors, np.complex128), int(n_terms), acf)
    n_seg = (2 if split_sides else 1) * n_samples
    hist = np.zeros((nT, n_seg), np.complex128)
    dropped = np.zeros(nT, np.float64)
    bins_out = np.zeros((nT, nF), np.int64)
    valid_out = np.zeros((nT, nF), bool)
    r_ref = reference_range(pos, C) if nF else np.zeros(nT)
    for t in range(nT):
        p = pos[t]
        contrib, r, d = facet_terms(p, C, N, A, E1, E2, k, g, r_ref[t],
                                    taper_s, rough, d_phi_area)
        if pattern is not None and pattern[0] != "isotropic":
            kind, pv, pa, pb = pattern
            gg = _ant.gain(kind, -d / r[:, None], np.asarray(pv)[t], pa, pb)
            contrib = contrib * (gg * gg)
        twtt = 2.0 * r / c
        b = twtt_bin(twtt, t0, dt)
        bins_out[t] = b
        valid_out[t] = (b >= 0) & (b < n_samples)
        if interp_bins:
            w = (twtt - t0) / dt - b
            parts = ((contrib * (1.0 - w), b), (contrib * w, b + 1))
        else:
            parts = ((contrib, b),)
        right = (dot3(-d, uct[t][None, :]) > 0).astype(np.int64)
        for pc, pb_ in parts:
            ok = (pb_ >= 0) & (pb_ < n_samples)
            seg = pb_ + right * n_samples if split_sides else pb_
            np.add.at(hist[t], seg[ok], pc[ok])
            dropped[t] += np.sum(np.abs(pc[~ok]) ** 2)
    phase_ref = np.exp(-1j * ((2.0 * k * r_ref) % TWO_PI))
    field = hist * phase_ref[:, None]
    if split_sides:
        field = field.reshape(nT, 2, n_samples).transpose(0, 2, 1)
    if return_bins:
        return field, dropped, bins_out, valid_out
    return field, dropped
