"""Along-track post-processing on the device: presum, unfocused SAR,
multilook and focused SAR (backprojection), restating the reference's
``processing.py:1-348`` on ``SimResult`` (no xarray).

Same functions, arguments, conventions and errors:

* ``presum(res, n)`` -- coherent non-overlapping boxcar of ``n`` traces
  (the field summed, a trailing remainder dropped), ``:143-171``;
* ``unfocused_sar(res, aperture_m=None, n_traces=None)`` -- the same boxcar
  as a stride-1 moving window (``T - n + 1`` outputs), with the Doppler
  aliasing warning, ``:174-221``;
* ``multilook(res, n)`` -- incoherent non-overlapping power average (any
  field dropped), ``:224-243``;
* ``focused_sar(res, aperture_m, window="hann")`` -- straight-track
  time-domain backprojection through air of a 2-D (slow_time, twtt) field,
  ``:270-348``.

presum / unfocused_sar / focused_sar refuse incoherent results.  Every
other slow-time quantity (positions, dropped power, nadir delays, first
returns) is averaged over the same windows; each step is appended to the
JSON list ``attrs["processing"]``.  The sums run on the GPU
(``ss_slow_time_boxcar``, ``ss_focused_sar``); there is no CPU path.
Added keyword: ``precision`` ("fp64" default, or "mixed") for focused_sar;
``device``.
"""

import json
import warnings
from dataclasses import replace

import numpy as np
import torch

from . import device as _dv

C_DEFAULT = 299792458.0


def _gpu(device):
    if not torch.cuda.is_available():
        raise RuntimeError("no GPU visible: along-track processing runs on the MI355X only "
                           "(no CPU fallback)")
    return torch.device("cuda", torch.cuda.current_device()) if device is None \
        else torch.device(device)


def _record(attrs, step):
    a = dict(attrs)
    steps = json.loads(a.get("processing", "[]"))
    steps.append(step)
    a["processing"] = json.dumps(steps)
    return a


def _require_coherent(res, op):
    if res.attrs.get("mode") != "coherent" or res.field is None:
        raise ValueError(
            f"{op}() needs a coherent-mode result with a complex `field` (got mode="
            f"{res.attrs.get('mode')!r}). Coherent summation of power is not meaningful; "
            "for incoherent speckle reduction use multilook(res, n) instead.")


def _config(res):
    try:
        return json.loads(res.attrs["config"])
    except (KeyError, ValueError):
        return {}


def _speed_of_light(res):
    return float(_config(res).get("radar", {}).get("c", C_DEFAULT))


def _wavelength(res):
    rc = _config(res).get("radar", {})
    if rc.get("f0"):
        return float(rc.get("c", C_DEFAULT)) / float(rc["f0"])
    return None


def _along_track(pos):
    seg = np.linalg.norm(np.diff(np.asarray(pos, np.float64), axis=0), axis=1)
    return np.concatenate([[0.0], np.cumsum(seg)])


def _mean_spacing(pos):
    pos = np.asarray(pos, np.float64)
    if pos.shape[0] < 2:
        return 0.0
    return float(np.linalg.norm(np.diff(pos, axis=0), axis=1).mean())


def _average(x, n, stride, n_out):
    """Window means of a host slow-time array (None passes through)."""
    if x is None:
        return None
    x = np.asarray(x)
    return np.stack([x[k * stride:k * stride + n].mean(axis=0) for k in range(n_out)])


def _complex_boxcar(field, n, stride, n_out, dev):
    f = np.asarray(field)
    x = torch.as_tensor(np.stack([f.real, f.imag], -1).astype(np.float64), device=dev)
    y = _dv.slow_time_boxcar(x, n, stride, n_out)
    torch.cuda.synchronize(dev)
    yh = y.cpu().numpy()
    return (yh[..., 0] + 1j * yh[..., 1]).astype(np.complex64)


def _reduce(res, n, stride, n_out, step, dev, coherent):
    if coherent:
        field = _complex_boxcar(res.field, n, stride, n_out, dev)
        power = None
    else:
        p = torch.as_tensor(np.asarray(res.power, np.float64), device=dev)
        y = _dv.slow_time_boxcar(p, n, stride, n_out, scale=1.0 / n)
        torch.cuda.synchronize(dev)
        field, power = None, y.cpu().numpy().astype(np.float32)
    attrs = _record(res.attrs, step)
    if not coherent:
        attrs["mode"] = "incoherent"
    avg = lambda x: _average(x, n, stride, n_out)  # noqa: E731
    return replace(res, field=field, power_data=power,
                   dropped_power=avg(res.dropped_power), nadir_twtt=avg(res.nadir_twtt),
                   positions=avg(res.positions), attrs=attrs,
                   first_return_twtt=avg(res.first_return_twtt),
                   first_return_xyz=avg(res.first_return_xyz),
                   first_return_llh=avg(res.first_return_llh))


def presum(res, n, device=None):
    """Coherent non-overlapping boxcar sum of ``n`` adjacent traces."""
    _require_coherent(res, "presum")
    n = int(n)
    if n < 1:
        raise ValueError("presum count n must be >= 1")
    T = res.field.shape[0]
    n_out = T // n
    if n_out < 1:
        raise ValueError(f"presum n={n} exceeds slow_time length {T}")
    step = {"op": "presum", "n": n, "n_traces_in": T, "n_traces_out": n_out,
            "convention": "non-overlapping coherent boxcar (CReSIS presum); "
                          "decimates along-track sampling by n"}
    return _reduce(res, n, n, n_out, step, _gpu(device), True)


def _doppler_guard(spacing, wavelength, aperture_m, r_ref, op):
    """Warn when the trace spacing under-samples the coherent Doppler history
    of the aperture: spacing > lambda / (4 sin theta_max) (``:246-267``)."""
    if r_ref <= 0:
        return
    theta = np.arctan((aperture_m / 2.0) / r_ref)
    st = np.sin(theta)
    if st <= 0:
        return
    crit = wavelength / (4.0 * st)
    if spacing > crit:
        warnings.warn(
            f"{op}: along-track trace spacing {spacing:.3f} m exceeds lambda/(4 sin "
            f"theta_max) = {crit:.3f} m for a {aperture_m:.0f} m aperture at range "
            f"{r_ref:.0f} m (half-angle {np.degrees(theta):.1f} deg): coherent (point/"
            "specular) returns will alias in Doppler. Densify the track or shrink the "
            "aperture.")


def unfocused_sar(res, aperture_m=None, n_traces=None, device=None):
    """Unfocused SAR: stride-1 coherent moving sum along track."""
    _require_coherent(res, "unfocused_sar")
    spacing = _mean_spacing(res.positions)
    if n_traces is None:
        if aperture_m is None:
            raise ValueError("give unfocused_sar n_traces or aperture_m")
        n = max(1, int(round(aperture_m / spacing))) if spacing > 0 else 1
    else:
        n = int(n_traces)
    if n < 1:
        raise ValueError("unfocused window must be >= 1 trace")
    T = res.field.shape[0]
    if n > T:
        raise ValueError(f"unfocused window n={n} exceeds slow_time length {T}")
    lam = _wavelength(res)
    if aperture_m is not None and lam is not None and res.nadir_twtt is not None:
        r_ref = float(np.nanmedian(res.nadir_twtt)) * _speed_of_light(res) / 2.0
        _doppler_guard(spacing, lam, aperture_m, r_ref, "unfocused_sar")
    n_out = T - n + 1
    step = {"op": "unfocused_sar", "n_traces": n, "aperture_m": aperture_m,
            "mean_spacing_m": spacing, "n_traces_in": T, "n_traces_out": n_out,
            "convention": "stride-1 coherent moving boxcar (CSARP_standard-like "
                          "unfocused SAR); preserves along-track sampling"}
    return _reduce(res, n, 1, n_out, step, _gpu(device), True)


def multilook(res, n, device=None):
    """Incoherent non-overlapping average of power over ``n`` traces."""
    n = int(n)
    if n < 1:
        raise ValueError("multilook count n must be >= 1")
    T = np.asarray(res.power).shape[0]
    n_out = T // n
    if n_out < 1:
        raise ValueError(f"multilook n={n} exceeds slow_time length {T}")
    step = {"op": "multilook", "n": n, "n_traces_in": T, "n_traces_out": n_out,
            "convention": "non-overlapping incoherent power average; "
                          "speckle contrast ~ 1/sqrt(n)"}
    return _reduce(res, n, n, n_out, step, _gpu(device), False)


def focused_sar(res, aperture_m, *, window="hann", precision="fp64", wavelength=None,
                device=None):
    """Straight-track time-domain backprojection focusing through air.

    Needs a 2-D (slow_time, twtt) coherent field: select a side / layer
    first.  ``wavelength`` defaults to c / radar.f0 of the run's config."""
    _require_coherent(res, "focused_sar")
    if res.field.ndim != 2:
        raise ValueError("focused_sar needs a 2-D (slow_time, twtt) field; select a single "
                         f"side/layer first (got shape {res.field.shape})")
    lam = wavelength if wavelength is not None else _wavelength(res)
    if lam is None:
        raise ValueError("focused_sar needs the wavelength (radar.f0 in the config, or "
                         "wavelength=)")
    if window not in ("hann", "none", "rect", None):
        raise ValueError(f"unknown window {window!r} (use 'hann' or 'none')")
    dev = _gpu(device)
    c = _speed_of_light(res)
    k = 2.0 * np.pi / lam
    twtt = np.asarray(res.twtt, np.float64)
    s = _along_track(res.positions)
    spacing = _mean_spacing(res.positions)
    _doppler_guard(spacing, lam, aperture_m, float((c * twtt / 2.0).min()), "focused_sar")
    f = np.asarray(res.field)
    x = torch.as_tensor(np.stack([f.real, f.imag], -1).astype(np.float64), device=dev)
    y = _dv.focused_sar(x, torch.as_tensor(twtt, device=dev), torch.as_tensor(s, device=dev),
                        c=c, k=k, aperture_m=aperture_m, window=window, precision=precision)
    torch.cuda.synchronize(dev)
    yh = y.cpu().numpy()
    out = (yh[..., 0] + 1j * yh[..., 1]).astype(np.complex64)
    n_ap = int(np.round(aperture_m / spacing)) + 1 if spacing > 0 else f.shape[0]
    step = {"op": "focused_sar", "aperture_m": aperture_m, "window": window,
            "mean_spacing_m": spacing, "n_aperture_traces_approx": n_ap,
            "model": "straight-track time-domain backprojection through air "
                     "(surface-referenced); in-ice focusing out of scope"}
    return replace(res, field=out, attrs=_record(res.attrs, step))
