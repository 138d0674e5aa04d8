"""Reference-signature entry points of the coherent facet kernel.

``coherent_cluttergram`` mirrors ``kernels/coherent.py:247-362`` and
``refracted_cluttergram`` mirrors ``kernels/multilayer.py:594-770``: same
positional/keyword arguments, same argument meaning, same ValueError cases
(raised before any device work), same return structure ``(out, dropped)``
as NumPy arrays.  Added keywords:

* ``precision``: ``"mixed"`` (default; float64 range/bin/phase, float32
  radiometry; field returned complex64 like the reference) or ``"fp64"``
  (plain float64; field returned complex128).
* ``device``: torch device of the GPU to run on (default: current).

``block_size`` and ``window_cull`` are accepted and ignored: the device
kernel always evaluates every (trace, facet) pair in 2-D facet tiles.  The
field equals the reference's culled sum exactly (culled facets are silent,
``geometry.py:27-38``); ``dropped`` counts the power of every out-of-window
or invalid pair, i.e. the reference's value with ``window_cull=False``.

There is no CPU path: without a GPU and the built HIP library these raise.
"""

import numpy as np
import torch

from . import device as _dv
from ._lib import BIN_INVALID  # noqa: F401  (re-exported for callers)


def _check_gpu():
    if not torch.cuda.is_available():
        raise RuntimeError("no GPU visible: the coherent facet kernel runs on "
                           "the MI355X only (no CPU fallback)")


def _reference_range(positions, centers):
    """Per-trace |p - mean(centres)| in float64 (``coherent.py:311-312``)."""
    m = np.asarray(centers, np.float64).mean(0)
    d = np.asarray(positions, np.float64) - m
    return np.sqrt((d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2])


def _rough_tuple(roughness, d_phi_area):
    if roughness is None:
        return None, None
    sigma, lc, phasors, n_terms, *rest = roughness
    acf = rest[0] if rest else "gaussian"
    if acf not in ("gaussian", "exponential"):
        raise ValueError(f"unknown roughness acf {acf!r}")
    if acf == "exponential" and not d_phi_area:
        raise ValueError("roughness acf='exponential' requires the "
                         "area-only D_Phi (d_phi_area / grazing fix)")
    return (float(sigma), float(lc), int(n_terms), acf), np.asarray(phasors)


def coherent_cluttergram(positions, u_ct, centers, normals, areas, e1, e2, *,
                         k, gamma, t0, dt, n_samples, c, split_sides=False,
                         interp_bins=False, pattern=None, roughness=None,
                         taper_s=None, d_phi_area=False, block_size=None,
                         window_cull=True, precision="mixed", device=None,
                         return_bins=False):
    """Binned coherent surface field for every trace (module docstring).

    Returns ``(field, dropped)``: field (T, n) or (T, n, 2) [left, right],
    complex64 (mixed) / complex128 (fp64); dropped (T,) float32/float64.
    ``return_bins`` adds the per-pair (T, F) floor bins (caller's facet
    order)."""
    n = len(centers)
    rough, phasors = _rough_tuple(roughness, d_phi_area)
    g = np.asarray(gamma, np.float64)
    gamma_facet = g.ndim > 0
    if gamma_facet and g.shape != (n,):
        raise ValueError(f"per-facet gamma shape {g.shape} != ({n},)")
    if precision not in ("mixed", "fp64"):
        raise ValueError("precision must be 'mixed' or 'fp64'")
    _check_gpu()
    pos = np.asarray(positions, np.float64)
    facets = _dv.DeviceFacets(centers, normals, areas, e1, e2,
                              gamma=g if gamma_facet else None,
                              phasors=phasors, device=device)
    pv = None if pattern is None else pattern[1]
    traces = _dv.DeviceTraces(pos, u_ct, _reference_range(pos, centers) if n else None,
                              pv, device=device)
    res = _dv.surface_field(
        facets, traces, k=k, gamma=0.0 if gamma_facet else float(gamma), t0=t0,
        dt=dt, n_samples=n_samples, c=c, precision=precision,
        split_sides=split_sides, interp_bins=interp_bins, pattern=pattern,
        roughness=rough, taper_s=taper_s, d_phi_area=d_phi_area,
        pair_bins=return_bins)
    torch.cuda.synchronize(facets.device)
    out, dropped = res[0], res[1]
    field = out.cpu().numpy()
    field = field[..., 0] + 1j * field[..., 1]
    drop = dropped.cpu().numpy()
    if precision == "mixed":
        field = field.astype(np.complex64)
        drop = drop.astype(np.float32)
    if split_sides:
        field = field.reshape(-1, 2, n_samples).transpose(0, 2, 1)
    if return_bins:
        return field, drop, facets.unpermute_pairs(res[2].cpu().numpy())
    return field, drop


def incoherent_cluttergram(positions, u_ct, centers, normals, areas, *, t0, dt,
                           n_samples, c, split_sides=False, pattern=None,
                           block_size=65536, precision="mixed", device=None,
                           return_bins=False):
    """Binned incoherent power for every trace (``kernels/incoherent.py:66-105``):
    per facet (A cos)^2 / r^4 [x g^4], out-of-window power dropped.

    Returns ``(power, dropped)``: power (T, n) or (T, n, 2) [left, right],
    float32 (mixed, like the reference) / float64 (fp64); dropped (T,).
    ``block_size`` is accepted and ignored (2-D facet tiles on the device)."""
    if precision not in ("mixed", "fp64"):
        raise ValueError("precision must be 'mixed' or 'fp64'")
    _check_gpu()
    n = len(centers)
    pos = np.asarray(positions, np.float64)
    z3 = np.zeros((n, 3))
    facets = _dv.DeviceFacets(centers, normals, areas, z3, z3, device=device)
    pv = None if pattern is None else pattern[1]
    traces = _dv.DeviceTraces(pos, u_ct, None, pv, device=device)
    res = _dv.incoherent_power(facets, traces, t0=t0, dt=dt, n_samples=n_samples, c=c,
                               precision=precision, split_sides=split_sides,
                               pattern=pattern, pair_bins=return_bins)
    torch.cuda.synchronize(facets.device)
    power = res[0].cpu().numpy()
    drop = res[1].cpu().numpy()
    if precision == "mixed":
        power = power.astype(np.float32)
        drop = drop.astype(np.float32)
    if split_sides:
        power = power.reshape(-1, 2, n_samples).transpose(0, 2, 1)
    if return_bins:
        return power, drop, facets.unpermute_pairs(res[2].cpu().numpy())
    return power, drop


def refracted_cluttergram(positions, u_ct, target, crossed, eps_leg, att_leg, *,
                          mode, t0, dt, n_samples, c, gamma=0.0, k0=None,
                          split_sides=False, pattern=None, roughness=None,
                          crossed_sigma=None, diffuse=None, taper_s=None,
                          d_phi_area=False, block_size=None,
                          refraction="sequential", pad_to=None,
                          joint_newton=6, joint_backtrack=4, window_cull=True,
                          precision="mixed", device=None, return_bins=False,
                          flat_stack=False):
    """Binned refracted-path contributions of one target interface.

    Same conventions as the reference (module docstring); returns
    ``(out, dropped)`` with out complex (coherent) or real power
    (incoherent), (T, n[, 2]).  ``flat_stack`` (joint only, an opt-in
    beyond the reference): False (default) runs the reference's joint solve
    (2 unknowns per crossed interface, damped Newton with its budgets; the
    stack, pads included, at most 32 deep); True or "auto" runs its
    two-unknown flat-stack form when every crossed interface below the
    surface is a horizontal plane (stacks up to 64 deep).  The flat-stack
    form converges on deep stacks where the reference's 6 x 4 budget can
    stop short, so a few far off-nadir lanes that the reference drops are
    valid there."""
    coherent = mode == "coherent"
    if mode not in ("coherent", "incoherent"):
        raise ValueError(f"unknown mode {mode!r}")
    if (roughness is not None or crossed_sigma is not None) and not coherent:
        raise ValueError("roughness requires coherent mode")
    if (taper_s is not None or d_phi_area) and not coherent:
        raise ValueError("the grazing fix requires coherent mode")
    g = np.asarray(gamma, np.float64)
    gamma_facet = g.ndim > 0
    if gamma_facet and not coherent:
        raise ValueError("per-facet gamma requires coherent mode")
    n = len(target.centers)
    if gamma_facet and g.shape != (n,):
        raise ValueError(f"per-facet gamma shape {g.shape} != ({n},)")
    rough, phasors = _rough_tuple(roughness, d_phi_area)
    dif_amp = dif_ph = n_exp = None
    if diffuse is not None:
        if not coherent:
            raise ValueError("diffuse channel requires coherent mode")
        dif_amp, dif_ph, n_exp = diffuse
        dif_amp = np.asarray(dif_amp, np.float64)
        if dif_amp.shape != (n,):
            raise ValueError(f"diffuse amp shape {dif_amp.shape} != ({n},)")
    if len(eps_leg) != len(crossed) + 1 or len(att_leg) != len(crossed) + 1:
        raise ValueError("eps_leg/att_leg need len(crossed) + 1 entries")
    if refraction not in ("sequential", "joint"):
        raise ValueError(f"unknown refraction solver {refraction!r}")
    joint = refraction == "joint"
    flat = joint and flat_stack is not False and all(_dv.is_horizontal(f) for f in crossed[1:])
    if joint and not flat and (pad_to or len(crossed)) > 32:
        raise ValueError("joint refraction supports stacks of at most 32 interfaces "
                         "(flat_stack=True takes horizontal stacks up to 64 deep)")
    if precision not in ("mixed", "fp64"):
        raise ValueError("precision must be 'mixed' or 'fp64'")
    _check_gpu()
    pos = np.asarray(positions, np.float64)
    tgt = _dv.DeviceFacets(target.centers, target.normals, target.areas,
                           target.e1, target.e2,
                           gamma=g if gamma_facet else None, phasors=phasors,
                           diffuse_amp=dif_amp, diffuse_phasors=dif_ph,
                           device=device)
    ifaces = _dv.DeviceInterfaces(crossed, crossed_sigma, device=device,
                                  pad_to=(pad_to or len(crossed)) if joint else None,
                                  z_platform=pos[:, 2].min())
    pv = None if pattern is None else pattern[1]
    traces = _dv.DeviceTraces(pos, u_ct, None, pv, device=device)
    res = _dv.refracted_field(
        tgt, traces, ifaces, eps_leg, att_leg, mode=mode, t0=t0, dt=dt,
        n_samples=n_samples, c=c, gamma=0.0 if gamma_facet else float(gamma),
        k0=k0, precision=precision, split_sides=split_sides, pattern=pattern,
        roughness=rough, taper_s=taper_s, d_phi_area=d_phi_area, n_exp=n_exp,
        refraction=refraction, joint_newton=joint_newton, joint_backtrack=joint_backtrack,
        pair_bins=return_bins, flat_stack=flat_stack)
    torch.cuda.synchronize(tgt.device)
    raw = res[0].cpu().numpy()
    drop = res[1].cpu().numpy()
    if coherent:
        out = raw[..., 0] + 1j * raw[..., 1]
        if precision == "mixed":
            out = out.astype(np.complex64)
    else:
        out = raw[..., 0].copy()
        if precision == "mixed":
            out = out.astype(np.float32)
    if precision == "mixed":
        drop = drop.astype(np.float32)
    if split_sides:
        out = out.reshape(-1, 2, n_samples).transpose(0, 2, 1)
    if return_bins:
        return out, drop, tgt.unpermute_pairs(res[2].cpu().numpy())
    return out, drop


def trace_diagnostics(positions, interfaces, eps, *, t0, dt, n_samples, c, device=None):
    """Per-trace output diagnostics on the device (``ss_trace_diagnostics``):
    ``nadir_twtt`` (T, L) -- the vertical two-way time through the stack with
    in-medium speeds via each interface's horizontally nearest facet
    (``simulate.py:398-412``) -- and the surface's first return
    (``output.py:77-82``): ``first_return_twtt`` (T,) and its facet centre
    ``first_return_xyz`` (T, 3), local frame.  A trace with no facet in the
    window reports facet 0 at +inf, as the reference."""
    _check_gpu()
    pos = np.asarray(positions, np.float64)
    traces = _dv.DeviceTraces(pos, np.zeros_like(pos), None, None, device=device)
    nadir = np.empty((len(pos), len(interfaces)))
    prev = pos[:, 2].copy()
    acc = np.zeros(len(pos))
    first = None
    for j, f in enumerate(interfaces):
        fac = _dv.DeviceFacets(f.centers, f.normals, f.areas, f.e1, f.e2, device=device)
        d = _dv.trace_diagnostics(fac, traces, t0=t0, dt=dt, n_samples=n_samples, c=c)
        z = d["nadir_z"].cpu().numpy()
        acc = acc + np.sqrt(eps[j]) * (prev - z)
        nadir[:, j] = 2.0 * acc / c
        prev = z
        if j == 0:
            fi = d["first_index"].cpu().numpy()
            ftw = d["first_twtt"].cpu().numpy()
            xyz = d["first_xyz"].cpu().numpy()
            none = fi < 0
            xyz[none] = np.asarray(f.centers, np.float64)[0]
            first = (ftw, xyz)
    return nadir, first[0], first[1]
