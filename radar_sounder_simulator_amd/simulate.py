"""simulate(): the front door of the hot path.

Follows the reference orchestrator (``simulate.py:172-412``) step for step:

* single interface: the surface coherent kernel, gamma = fresnel_normal of
  the first two media, roughness arguments with the wavenumber of the medium
  above, grazing fix, antenna pattern, waveform (``:172-226``); in
  incoherent mode the power-summing kernel instead (``:186-193``);
* several interfaces: interface 0 through the surface kernel (coherent or
  incoherent), interface j >= 1 through the refracted kernel in the run's
  mode -- sequential for j = 1, ``cfg.refraction`` for j > 1 with joint
  calls padded to power-of-two interface buckets -- per-layer gamma (or a
  per-facet gamma map), roughness, crossing roughness, diffuse channel
  (coherent only), then one waveform pass over the layer axis (linear) and
  per-layer nadir delays (``:263-412``).

Scenes: a ``PlanarScene`` (vertex lattices already in a local frame) or a
``geodesy.ProjectedScene`` (EPSG:3413 DEMs and a lat/lon/h nav track):
the latter is tessellated in the ENU frame centred on its DEM and its
first return is also reported in lat/lon (``output.py:77-82``), as the
reference does for its synthetic and DEM scenes.

The output is a ``SimResult`` (``output.py``); interface kinds dem / flat /
offset resolve as in ``layered.py:56-118`` (offsets shift along +z here).
"""

import warnings
from dataclasses import dataclass, field

import numpy as np

from . import antenna as _ant
from . import geodesy as _geo
from .config import DemInterface, FlatInterface, OffsetInterface, SimConfig
from .kernels import (coherent_cluttergram, incoherent_cluttergram, refracted_cluttergram,
                      trace_diagnostics)
from .output import SimResult
from .physics import fresnel_normal
from .roughness import n_terms_for, speckle_phasors, speckle_phasors_for_keys
from .scene import Facets, check_facet_size, grid_facets
from .waveform import apply_waveform


@dataclass
class PlanarScene:
    """Heights of one or more (ny+1, nx+1) vertex lattices at ``spacing`` m
    (``dems[0]`` is the surface; a bare DemInterface at position i takes
    ``dems[i]``), a track in the same frame, optional per-interface maps of
    per-facet gamma / diffuse amplitude (name -> (ny, nx) array) and roll."""

    dems: list
    spacing: float
    track: object
    gamma_maps: dict = field(default_factory=dict)
    diffuse_maps: dict = field(default_factory=dict)
    nav_roll: np.ndarray = None
    name: str = "planar"


def _pattern(rc, scene, track):
    a = rc.antenna
    if a.kind == "isotropic":
        return None
    ant = _ant.Antenna(**a.model_dump())
    roll = None
    if a.roll_source == "nav":
        roll = scene.nav_roll if scene.nav_roll is not None else np.zeros(len(track.positions))
    return _ant.pattern_args(ant, track.u_at, track.u_ct, roll)


def warn_quantization_alias(rc):
    """Warn when a chirped coherent run has the envelope-quantisation alias
    in band (``simulate.py:71-95`` of the reference): binning quantises the
    envelope delay to dt, which plants noise at the aliased carrier
    f_a = |f0 - round(f0 dt)/dt|; the compressed pulse passes it when
    f_a < B/2."""
    wf = rc.waveform
    if wf.kind != "chirp" or wf.interp_bins:
        return
    f_alias = abs(rc.f0 - round(rc.f0 * rc.dt) / rc.dt)
    if f_alias < wf.bandwidth / 2.0:
        warnings.warn(
            f"chirped run with the envelope-quantization alias in band "
            f"(|f0 - round(f0*dt)/dt| = {f_alias / 1e6:.1f} MHz < B/2 = "
            f"{wf.bandwidth / 2e6:.1f} MHz): expect a nonphysical off-nadir floor; "
            "set waveform.interp_bins=true or choose dt so the alias falls out of band")


def _check_facet_sizes(ifaces, positions, lam, eps):
    """Fresnel-zone facet-size check of every interface with the in-medium
    wavelength above it (``simulate.py:210-213, 300-304``)."""
    for j, f in enumerate(ifaces):
        lam_j = lam / np.sqrt(eps[j])
        min_range = max(float(positions[:, 2].min() - f.centers[:, 2].max()), lam_j)
        check_facet_size(f, lam_j, min_range)


def _roughness_args(iface, facets, k_local, seed):
    rc = iface.roughness
    if rc is None or rc.sigma_m == 0.0:
        return None
    edge = max(float(np.linalg.norm(facets.e1, axis=1).max()),
               float(np.linalg.norm(facets.e2, axis=1).max()))
    if rc.corr_length_m > edge:
        # simulate.py:157-162: roughness above the facet scale is facet tilt
        warnings.warn(f"interface {iface.name or '?'}: roughness corr_length_m = "
                      f"{rc.corr_length_m:g} m exceeds the facet size ({edge:.1f} m); "
                      "roughness at scales above the facet is really facet tilt -- "
                      "put it in the DEM")
    n_terms = n_terms_for((2.0 * k_local * rc.sigma_m) ** 2)
    ph = (speckle_phasors_for_keys(facets.phase_keys, seed)
          if facets.phase_keys is not None else speckle_phasors(len(facets), seed))
    if rc.acf != "gaussian":
        return (rc.sigma_m, rc.corr_length_m, ph, n_terms, rc.acf)
    return (rc.sigma_m, rc.corr_length_m, ph, n_terms)


def joint_pad_to(j, n_max):
    """Power-of-two padding bucket (``simulate.py:253-260``)."""
    b = 1
    while b < j:
        b *= 2
    return min(b, n_max)


def _lattice_facets(scene, heights, frame):
    if frame is not None:
        return _geo.build_facets(heights, scene.transform, frame)
    return grid_facets(heights, scene.spacing)


def build_interfaces(scene, cfg, frame=None):
    """Facets per configured interface, top-down (``frame``: the local frame
    of a projected scene, else None)."""
    out, names = [], []
    for i, ic in enumerate(cfg.interfaces):
        names.append(ic.name or f"interface_{i}")
        if isinstance(ic, DemInterface):
            if ic.path is not None:
                raise ValueError("DEM files are not read by this library: give the "
                                 "interface heights in PlanarScene.dems")
            if i >= len(scene.dems):
                raise ValueError(f"interface {i} ({ic.name or 'unnamed'}) needs a DEM: the "
                                 "scene provides none")
            out.append(_lattice_facets(scene, scene.dems[i], frame))
        elif isinstance(ic, FlatInterface):
            ref = np.asarray(scene.dems[0], np.float64)
            out.append(_lattice_facets(scene, np.full_like(ref, ic.elevation), frame))
        else:
            out.append(None)
    for i, ic in enumerate(cfg.interfaces):
        if isinstance(ic, OffsetInterface):
            r = ic.reference if isinstance(ic.reference, int) else names.index(ic.reference)
            src = out[r]
            if src is None:
                raise ValueError(f"offset interface {i} references an unresolved offset")
            c = src.centers.copy()
            c[:, 2] += ic.offset
            out[i] = Facets(c, src.normals.copy(), src.areas.copy(), src.e1.copy(),
                            src.e2.copy(), src.cell.copy(), src.grid_shape,
                            None if src.phase_keys is None else src.phase_keys.copy())
    return out, names


def nadir_twtt_layers(pos, interfaces, eps, c):
    """Per-layer nadir two-way time via the horizontally nearest facet centre
    per interface (``simulate.py:398-412``), O(T F) per interface on the
    host; ``simulate`` uses the device form (``kernels.trace_diagnostics``)."""
    nadir = np.empty((pos.shape[0], len(interfaces)))
    prev = pos[:, 2].copy()
    acc = np.zeros(pos.shape[0])
    for j, f in enumerate(interfaces):
        d2 = ((f.centers[None, :, :2] - pos[:, None, :2]) ** 2).sum(-1)
        z = f.centers[d2.argmin(axis=1), 2]
        acc = acc + np.sqrt(eps[j]) * (prev - z)
        nadir[:, j] = 2.0 * acc / c
        prev = z
    return nadir


def simulate(scene, sim_config: SimConfig, device=None):
    """Run the hot path on ``scene`` (``PlanarScene`` or
    ``geodesy.ProjectedScene``); returns a ``SimResult``."""
    cfg = sim_config
    coherent = cfg.mode == "coherent"
    rc = cfg.radar
    if coherent and rc.f0 is None:
        raise ValueError("coherent mode requires radar.f0 to be set")
    if len(cfg.interfaces) > 1 and rc.waveform.interp_bins:
        raise ValueError("waveform.interp_bins is not yet supported for multilayer runs")
    frame = None
    if isinstance(scene, _geo.ProjectedScene):
        frame = _geo.LocalFrame.centered_on(scene)
        track = _geo.nav_to_frame(scene.nav_llh, frame, scene.nav_roll)
    else:
        track = scene.track
    ifaces, names = build_interfaces(scene, cfg, frame)
    eps = [m.eps_r for m in cfg.media]
    att = [m.attenuation_db_per_km for m in cfg.media]
    k0 = 2.0 * np.pi / rc.wavelength if coherent else None
    if coherent:
        warn_quantization_alias(rc)
        _check_facet_sizes(ifaces, track.positions, rc.wavelength, eps)
    window = dict(t0=rc.t0, dt=rc.dt, n_samples=rc.n_samples, c=rc.c,
                  split_sides=cfg.split_sides, pattern=_pattern(rc, scene, track),
                  precision=cfg.precision, device=device)
    gfx = cfg.grazing_fix
    gfx_kw = dict(taper_s=gfx.s_eff if gfx else None, d_phi_area=gfx is not None)
    sig_all = np.array([(ic.roughness.sigma_m if ic.roughness else 0.0)
                        for ic in cfg.interfaces])
    gmaps = getattr(scene, "gamma_maps", None) or {}
    dmaps = getattr(scene, "diffuse_maps", None) or {}
    if (gmaps or dmaps) and not coherent:
        raise ValueError("gamma_maps / diffuse_maps require coherent mode")
    unknown = (set(gmaps) | set(dmaps)) - set(names)
    if unknown:
        raise ValueError(f"maps for unknown interfaces: {unknown}")
    if names[0] in dmaps:
        raise ValueError("diffuse_maps on the top interface is not wired (the surface "
                         "uses the surface kernel)")
    outs, drops = [], []
    for j, target in enumerate(ifaces):
        gamma_j = fresnel_normal(eps[j], eps[j + 1])
        if names[j] in gmaps:
            gamma_j = np.asarray(gmaps[names[j]], np.float64).ravel()
        rough = (_roughness_args(cfg.interfaces[j], target, k0 * np.sqrt(eps[j]),
                                 (cfg.roughness_seed, j)) if coherent else None)
        if j == 0 and coherent:
            out, drop = coherent_cluttergram(
                track.positions, track.u_ct, target.centers, target.normals,
                target.areas, target.e1, target.e2, k=k0, gamma=gamma_j,
                interp_bins=rc.waveform.interp_bins, roughness=rough, **gfx_kw, **window)
        elif j == 0:
            out, drop = incoherent_cluttergram(
                track.positions, track.u_ct, target.centers, target.normals,
                target.areas, **window)
        else:
            refr = cfg.refraction if j > 1 else "sequential"
            n_max = len(ifaces) - 1
            crossed_sig = sig_all[:j] if coherent and sig_all[:j].any() else None
            dif = None
            if names[j] in dmaps:
                ph = (speckle_phasors_for_keys(target.phase_keys,
                                               (cfg.roughness_seed, 1000 + j))
                      if target.phase_keys is not None else
                      speckle_phasors(len(target), (cfg.roughness_seed, 1000 + j)))
                dif = (np.asarray(dmaps[names[j]], np.float64).ravel(), ph,
                       float(cfg.diffuse_exponent))
            out, drop = refracted_cluttergram(
                track.positions, track.u_ct, target, ifaces[:j], eps[:j + 1], att[:j + 1],
                mode=cfg.mode, gamma=gamma_j if coherent else 0.0, k0=k0, refraction=refr,
                pad_to=joint_pad_to(j, n_max) if refr == "joint" else None,
                flat_stack=refr == "joint" and cfg.joint_form == "flat_stack",
                roughness=rough, crossed_sigma=crossed_sig, diffuse=dif,
                **(gfx_kw if coherent else {}), **window)
        outs.append(out)
        drops.append(drop)
    single = len(ifaces) == 1
    out = outs[0] if single else np.stack(outs, axis=-1)
    out = apply_waveform(out, rc, cfg.mode)
    dropped = drops[0] if single else np.stack(drops, axis=-1)
    nadir, first_twtt, first_xyz = trace_diagnostics(
        track.positions, ifaces, eps, t0=rc.t0, dt=rc.dt, n_samples=rc.n_samples, c=rc.c,
        device=device)
    first_llh = frame.local_to_llh(first_xyz) if frame is not None else None
    return SimResult.build(field=out if coherent else None,
                           power=None if coherent else out, dropped=dropped,
                           nadir_twtt=nadir, t0=rc.t0, dt=rc.dt, positions=track.positions,
                           layers=None if single else names, config=cfg,
                           first_return_twtt=first_twtt, first_return_xyz=first_xyz,
                           first_return_llh=first_llh)
