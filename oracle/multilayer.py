"""Refracted multilayer kernel (FP64 oracle), restating ``kernels/multilayer.py``.

Per (trace p, target facet q) the sequential chain over the crossed
interfaces (``:352-406``): two-point Snell against the interface's area-
weighted mean plane (10 iterations) -> affine (x, y, 1) -> (row, col) cell
lookup with round-half-even and clip -> two-point Snell against that cell's
local plane (25 iterations); TE Fresnel two-way transmission 1 - gamma^2 (and
the optional crossing roughness exp(-2 sigma^2 K_t^2)), optical path, one-way
loss, ray-tube sums.  The final leg, astigmatic spreading, flux factor,
attenuation, target LPA amplitude with k_j = k0 n_j, taper, sub-facet
roughness, diffuse channel, antenna gain at the air-leg departure direction,
contribution j amp e^{-2j k0 opl} and binning of 2 opl / c follow
``:500-583``.  ``mode="incoherent"`` books (A cos_t spread)^2 att_f^2 power.

The joint (block-tridiagonal Newton) crossing solver is in ``oracle.joint``.

All float64, lookup tables included (the reference stores its tables and
facet blocks as float32; its geometry runs in float64 under x64).
Evaluation order is explicit (see ``oracle.refraction``).

Pinned by tests/test_oracle_multilayer.py: eps -> 1 reduction to the surface
oracle, exact slab nadir delay bin, attenuation law 10^(-0.6), image-in-
dielectric slab closed form tau_d tau_u Gamma e^{-2jk0(h+nd)}/(2(h+d/n)) by
radial summation through the same amplitude convention, two-media brute-
force referee on a flat slab, dropped bookkeeping for shadowed geometry.
"""

import numpy as np

from . import antenna as _ant
from .coherent import TWO_PI, dot3, off_specular_taper, sinc, twtt_bin
from .physics import fresnel_te
from .refraction import snell_c
from .roughness import d_phi, mean_attenuation

C_MIN = 1e-9


def grid_consts(facets):
    """(tab (ny, nx, 6) [centre, normal] float64, coef (3, 2), mean point,
    unit mean normal) -- ``kernels/multilayer.py:269-283`` in float64."""
    ny, nx = facets.grid_shape
    tab = np.concatenate([facets.centers.reshape(ny, nx, 3),
                          facets.normals.reshape(ny, nx, 3)], axis=-1)
    a = np.column_stack([facets.centers[:, 0], facets.centers[:, 1],
                         np.ones(len(facets.centers))])
    coef, *_ = np.linalg.lstsq(a, facets.cell.astype(np.float64), rcond=None)
    w = (facets.areas / facets.areas.sum())[:, None]
    mp = (facets.centers * w).sum(0)
    mn = (facets.normals * w).sum(0)
    mn = mn / np.linalg.norm(mn)
    return np.ascontiguousarray(tab, np.float64), coef, mp, mn


def sequential_path(p, q, consts, n_leg, eps_leg, att, k0, sig_c):
    """Chained path platform -> target centres (``:352-406``), component form."""
    shape = q[0].shape
    cur = tuple(p[i] + np.zeros(shape) for i in range(3))
    valid = np.ones(shape, bool)
    opl = np.zeros(shape)
    loss_db = np.zeros(shape)
    sum_par = np.zeros(shape)
    sum_perp = np.zeros(shape)
    tau2 = np.ones(shape)
    c_first = c_last = x_first = None
    for i, (tab, coef, mp, mn) in enumerate(consts):
        x1 = snell_c(cur, q, (mp[0], mp[1], mp[2]), (mn[0], mn[1], mn[2]),
                     n_leg[i], n_leg[i + 1], 10)[0]
        rc0 = (x1[0] * coef[0, 0] + x1[1] * coef[1, 0]) + coef[2, 0]
        rc1 = (x1[0] * coef[0, 1] + x1[1] * coef[1, 1]) + coef[2, 1]
        row = np.clip(np.round(rc0), 0, tab.shape[0] - 1).astype(np.int64)
        col = np.clip(np.round(rc1), 0, tab.shape[1] - 1).astype(np.int64)
        g6 = tab[row, col]
        pt = (g6[..., 0], g6[..., 1], g6[..., 2])
        nr = (g6[..., 3], g6[..., 4], g6[..., 5])
        inn = 1.0 / np.sqrt((nr[0] * nr[0] + nr[1] * nr[1]) + nr[2] * nr[2])
        nr = (nr[0] * inn, nr[1] * inn, nr[2] * inn)
        x2, th1, th2, s1, _, _, ok = snell_c(cur, q, pt, nr, n_leg[i],
                                             n_leg[i + 1], 25)
        valid = valid & ok
        c_inc = np.maximum(np.cos(th1), C_MIN)
        if i == 0:
            c_first = c_inc
            x_first = x2
        g = fresnel_te(eps_leg[i], eps_leg[i + 1], c_inc)[0]
        tau2 = tau2 * (1.0 - g * g)
        if sig_c is not None:
            kt = k0 * (n_leg[i] * c_inc - n_leg[i + 1] * np.cos(th2))
            tau2 = tau2 * np.exp(-2.0 * (sig_c[i] * kt) ** 2)
        opl = opl + n_leg[i] * s1
        loss_db = loss_db + s1 * (att[i] / 1000.0)
        sum_perp = sum_perp + s1 / n_leg[i]
        sum_par = sum_par + s1 / (n_leg[i] * c_inc * c_inc)
        c_last = np.maximum(np.cos(th2), C_MIN)
        cur = x2
    return (cur, valid, opl, loss_db, sum_par, sum_perp, tau2, c_first,
            c_last, x_first)


def radiometry(p, q, fn, fa, f1, f2, path, *, n_leg, att, k0, gam, coherent,
               pattern_t=None, taper_s=None, rough=None, d_phi_area=False,
               diffuse=None):
    """Final leg + radiometry for one trace (``:500-573``): returns
    (contribution, power, 2 opl)."""
    (cur, valid, opl, loss_db, sum_par, sum_perp, tau2, c0, c_last,
     x_first) = path
    d = (cur[0] - q[0], cur[1] - q[1], cur[2] - q[2])
    s_j = np.sqrt((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2])
    inv = 1.0 / np.maximum(s_j, 1e-30)
    rhat = (d[0] * inv, d[1] * inv, d[2] * inv)
    nj = n_leg[-1]
    opl = opl + nj * s_j
    loss_db = loss_db + s_j * (att[-1] / 1000.0)
    sum_perp = sum_perp + s_j / nj
    sum_par = sum_par + s_j / (nj * c_last * c_last)
    l_perp = n_leg[0] * sum_perp
    l_par = n_leg[0] * c0 * c0 * sum_par
    flux = (n_leg[0] * c0) / (nj * c_last)
    att_f = 10.0 ** (-loss_db / 10.0)
    cos_t = (rhat[0] * fn[:, 0] + rhat[1] * fn[:, 1]) + rhat[2] * fn[:, 2]
    spread = tau2 * flux / (l_par * l_perp)
    if pattern_t is not None:
        kind, pv, pa, pb = pattern_t
        d0 = np.stack([x_first[0] - p[0], x_first[1] - p[1],
                       x_first[2] - p[2]], -1)
        dm = np.maximum(np.sqrt(dot3(d0, d0)), 1e-30)
        g = _ant.gain(kind, d0 / dm[:, None], pv, pa, pb)
        spread = spread * (g * g)
    if coherent:
        kj = k0 * nj
        rh = np.stack(rhat, -1)
        d1 = dot3(rh, f1)
        d2 = dot3(rh, f2)
        s1 = sinc(d1 * (kj / np.pi))
        s2 = sinc(d2 * (kj / np.pi))
        amp = (kj / TWO_PI) * gam * cos_t * fa * s1 * s2 * spread * att_f
        if taper_s is not None:
            ct2 = np.clip(cos_t, C_MIN, 1.0) ** 2
            amp = amp * np.exp(-(1.0 - ct2) / (ct2 * (2.0 * taper_s * taper_s)))
        if rough is not None:
            sig_t, l_t, fph, n_terms, acf = rough
            l1 = np.sqrt(dot3(f1, f1))
            l2 = np.sqrt(dot3(f2, f2))
            live = fa > 0
            l1s, l2s = np.where(live, l1, 1.0), np.where(live, l2, 1.0)
            kk = 2.0 * kj * cos_t
            dp = d_phi(sig_t, l_t, kk, 2.0 * kj * d1 / l1s, 2.0 * kj * d2 / l2s,
                       l1s, l2s, n_terms=n_terms, area_only=d_phi_area, acf=acf)
            amp = (amp * mean_attenuation(sig_t, kk)
                   + np.where(live, (kj / TWO_PI) * gam * cos_t * spread
                              * att_f * np.sqrt(dp) * fph, 0.0))
        if diffuse is not None:
            fdif, fdph, n_exp = diffuse
            ct = np.maximum(cos_t, 0.0)
            a_d = (np.sqrt(np.maximum(fa, 0.0) / TWO_PI) * fdif * ct
                   * ct ** (0.5 * n_exp) * spread * att_f)
            amp = amp + np.where(fa > 0, a_d * fdph, 0.0)
        contrib = 1j * amp * np.exp(-2j * k0 * opl)
        pwr = contrib.real ** 2 + contrib.imag ** 2
    else:
        a = fa * cos_t * spread
        contrib = a * a * att_f * att_f
        pwr = contrib
    return contrib, pwr, opl, valid


def refracted_cluttergram(positions, u_ct, target, crossed, eps_leg, att_leg, *,
                          mode, t0, dt, n_samples, c, gamma=0.0, k0=None,
                          split_sides=False, pattern=None, roughness=None,
                          crossed_sigma=None, diffuse=None, taper_s=None,
                          d_phi_area=False, refraction="sequential",
                          pad_to=None, joint_newton=6, joint_backtrack=4,
                          return_bins=False, flat_stack=False):
    """FP64 restatement of ``refracted_cluttergram`` (``:594-770``): out
    (T, n[, 2]) complex128 (coherent) / float64 (incoherent), dropped (T,).
    ``flat_stack`` (not in the reference; joint only): solve the crossings
    with ``flatstack.flat_path`` -- the oracle of the kernels' flat-stack
    form, for stacks whose interfaces below the surface are horizontal."""
    coherent = mode == "coherent"
    if (roughness is not None or crossed_sigma is not None) and not coherent:
        raise ValueError("roughness requires coherent mode")
    if (taper_s is not None or d_phi_area) and not coherent:
        raise ValueError("the grazing fix requires coherent mode")
    gamma_arr = np.asarray(gamma, np.float64)
    gamma_facet = gamma_arr.ndim > 0
    if gamma_facet and not coherent:
        raise ValueError("per-facet gamma requires coherent mode")
    eps = np.asarray(eps_leg, np.float64)
    n_leg = np.sqrt(eps)
    att = np.asarray(att_leg, np.float64)
    pos = np.asarray(positions, np.float64)
    uct = np.asarray(u_ct, np.float64)
    nT = pos.shape[0]
    n = len(target.centers)
    if gamma_facet and gamma_arr.shape != (n,):
        raise ValueError(f"per-facet gamma shape {gamma_arr.shape} != ({n},)")
    sig_c = None if crossed_sigma is None else np.asarray(crossed_sigma,
                                                          np.float64)
    rough = None
    if roughness is not None:
        sig_t, l_t, phasors, n_terms, *rest = roughness
        acf = rest[0] if rest else "gaussian"
        if acf == "exponential" and not d_phi_area:
            raise ValueError("roughness acf='exponential' requires the "
                             "area-only D_Phi (d_phi_area / grazing fix)")
        rough = (float(sig_t), float(l_t), np.asarray(phasors, np.complex128),
                 int(n_terms), acf)
    dif = None
    if diffuse is not None:
        if not coherent:
            raise ValueError("diffuse channel requires coherent mode")
        d_amp, d_ph, n_exp = diffuse
        d_amp = np.asarray(d_amp, np.float64)
        if d_amp.shape != (n,):
            raise ValueError(f"diffuse amp shape {d_amp.shape} != ({n},)")
        dif = (d_amp, np.asarray(d_ph, np.complex128), float(n_exp))
    joint = refraction == "joint"
    if joint:
        from .joint import joint_consts, joint_path
        from .flatstack import flat_path
        jc = joint_consts(crossed, pad_to or len(crossed), pos[:, 2].min())
        kp = int(jc[-1])
        n_leg = np.concatenate([np.full(kp, n_leg[0]), n_leg])
        eps = np.concatenate([np.full(kp, eps[0]), eps])
        att = np.concatenate([np.full(kp, att[0]), att])
        if sig_c is not None:
            sig_c = np.concatenate([np.zeros(kp), sig_c])
    else:
        consts = [grid_consts(f) for f in crossed]
    fc = np.asarray(target.centers, np.float64)
    fn = np.asarray(target.normals, np.float64)
    fa = np.asarray(target.areas, np.float64)
    f1 = np.asarray(target.e1, np.float64)
    f2 = np.asarray(target.e2, np.float64)
    q = (fc[:, 0], fc[:, 1], fc[:, 2])
    n_seg = (2 if split_sides else 1) * n_samples
    hist = np.zeros((nT, n_seg), np.complex128 if coherent else np.float64)
    dropped = np.zeros(nT)
    bins_out = np.zeros((nT, n), np.int64)
    valid_out = np.zeros((nT, n), bool)
    gam = gamma_arr if gamma_facet else float(gamma)
    for t in range(nT):
        p = (pos[t, 0], pos[t, 1], pos[t, 2])
        if joint:
            path = (flat_path if flat_stack else joint_path)(
                p, fc, jc, n_leg, eps, att, k0, sig_c, joint_newton, joint_backtrack)
        else:
            path = sequential_path(p, q, consts, n_leg, eps, att, k0, sig_c)
        pt = None
        if pattern is not None and pattern[0] != "isotropic":
            kind, pv, pa, pb = pattern
            pt = (kind, np.asarray(pv)[t], pa, pb)
        contrib, pwr, opl, valid = radiometry(
            p, q, fn, fa, f1, f2, path, n_leg=n_leg, att=att,
            k0=0.0 if k0 is None else k0, gam=gam, coherent=coherent,
            pattern_t=pt, taper_s=taper_s, rough=rough,
            d_phi_area=d_phi_area, diffuse=dif)
        b = twtt_bin(2.0 * opl / c, t0, dt)
        ok = valid & (b >= 0) & (b < n_samples)
        bins_out[t] = b
        valid_out[t] = valid
        seg = b
        if split_sides:
            dd = (q[0] - p[0]) * uct[t, 0] + (q[1] - p[1]) * uct[t, 1]
            dd = dd + (q[2] - p[2]) * uct[t, 2]
            seg = b + (dd > 0).astype(np.int64) * n_samples
        np.add.at(hist[t], seg[ok], contrib[ok])
        dropped[t] += np.sum(np.where(ok, 0.0, pwr))
    if split_sides:
        hist = hist.reshape(nT, 2, n_samples).transpose(0, 2, 1)
    if return_bins:
        return hist, dropped, bins_out, valid_out
    return hist, dropped
