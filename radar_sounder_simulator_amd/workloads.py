"""Seeded synthetic workloads of the benchmark configurations.

Every workload is a planar scene (metres, x east / y north / z up) with a
straight level track along +x centred over the scene, built from seeded
Gaussian-ACF height lattices (``scene.gaussian_surface``; the multilayer
scene and the uneven mix from keyed ones, ``scene.SeededLattice``, which
the device draws from their seeds).  Frozen choices
the configuration text leaves open are recorded here:

* sampling: 48 MHz (dt = 1/48 MHz) wherever the rate is not given; window
  start t0 = 2 (h - 100 m)/c so the surface nadir return lands ~32 bins in;
* pulse length 20 us (the reference's 60 MHz instrument preset) with the
  analytic Hann compressed pulse unless stated (config 3: Tukey(0.2));
* "cos^2 antenna pattern": one-way POWER cos^2(theta) off the nadir
  boresight, i.e. field gain cos(theta), as a ``tabulated`` pattern
  (181 samples over 0..90 deg, zero beyond);
* "-10 dB bed reflectivity": field coefficient -10^(-10/20);
* trace spacing 25 m for the 256-trace surface+bed scene;
* exponential sub-facet roughness implies the grazing fix (taper
  s_eff = 0.05 + area-only D_Phi), as the reference's config validator
  requires (``config.py:360-365``);
* multilayer scene (2048 x 1024 at 30 m): a gently rough surface (sigma
  1 m, l 300 m); the 20 firn sublayers are 5 m thick over the top 100 m,
  each with n = 1 + 0.845 rho at its top (rho(z) = 917 - 567 exp(-z/30 m),
  kg/m^3), ice below (rho 917); every interface is a target, the firn ones
  with their normal-incidence Fresnel coefficient; bed at mean depth
  3000 m (sigma 400 m stays clear of the 1500 m horizon), -10 dB;
  10 dB/km one-way in firn and ice; 1000 traces at 30 m; 60 MHz
  sampling; deeper layers (j > 1) through the joint solve padded to
  power-of-two stacks as the reference's simulate() does.
"""

from dataclasses import dataclass, field

import numpy as np

from . import antenna as _ant
from .physics import C, fresnel_normal
from .roughness import n_terms_for, speckle_phasors
from .scene import (LatticeFacets, SeededLattice, eps_kovacs, gaussian_surface, grid_facets,
                    straight_track)
from .waveform import Waveform


@dataclass
class Layer:
    """One target interface of a workload."""

    name: str
    facets: object
    gamma: float
    roughness: tuple = None       # (sigma, l, phasors, n_terms, acf)
    crossed: list = field(default_factory=list)
    eps_leg: list = None
    att_leg: list = None
    refraction: str = "sequential"  # or "joint" (padded to pad_to planes)
    pad_to: int = None
    lattice: object = None          # scene.Lattice the facets tessellate, if any
    flat_stack: bool = False        # joint: the two-unknown flat-stack form (opt-in)


@dataclass
class Workload:
    name: str
    track: object
    layers: list
    k0: float
    t0: float
    dt: float
    n_samples: int
    c: float
    pattern: tuple = None
    taper_s: float = None
    d_phi_area: bool = False
    waveform: Waveform = None
    split_sides: bool = False
    meta: dict = field(default_factory=dict)

    def pairs(self):
        return sum(len(L.facets) for L in self.layers) * len(self.track.positions)


def config1(n=128, n_traces=64, seed=1, n_samples=1024):
    """128 x 128 facets at 20 m, sigma 2 m / l 200 m Gaussian heights, ice
    eps 3.15 (Fresnel only), 20 km altitude, 64 traces at 10 m, isotropic,
    60 MHz / 15 MHz chirp, 1024 samples at 48 MHz."""
    rng = np.random.default_rng(seed)
    dx = 20.0
    h = 2.0 * gaussian_surface(n + 1, dx, 200.0, rng)
    surf = grid_facets(h, dx)
    alt = 20000.0
    track = straight_track(n_traces, 10.0, alt)
    f0 = 60e6
    k0 = 2.0 * np.pi * f0 / C
    dt = 1.0 / 48e6
    return Workload(
        name="cfg1_surface_fp64", track=track,
        layers=[Layer("surface", surf, fresnel_normal(1.0, 3.15))],
        k0=k0, t0=2.0 * (alt - 100.0) / C, dt=dt, n_samples=n_samples, c=C,
        waveform=Waveform(kind="chirp", bandwidth=15e6, pulse_length=20e-6),
        meta=dict(f0=f0, facets=f"{n}x{n}@{dx}m", seed=seed))


def config2(n=512, n_traces=256, seed=2, n_samples=4096, trace_spacing=25.0):
    """Surface + bed with refraction: 512^2 facets at 25 m (surface sigma 3 m
    / l 300 m; bed 2000 m deep, sigma 300 m / l 2 km, independent seed),
    bed -10 dB, ice eps 3.15 with 10 dB/km, exponential sub-facet roughness
    (surface 0.2 m, bed 0.5 m, l 20 m) + grazing fix, 256 traces at 20 km,
    cos^2 pattern, 60 MHz / 15 MHz chirp, 4096 samples at 48 MHz."""
    rng_s = np.random.default_rng([seed, 0])
    rng_b = np.random.default_rng([seed, 1])
    dx = 25.0
    hs = 3.0 * gaussian_surface(n + 1, dx, 300.0, rng_s)
    hb = -2000.0 + 300.0 * gaussian_surface(n + 1, dx, 2000.0, rng_b)
    surf = grid_facets(hs, dx)
    bed = grid_facets(hb, dx)
    alt = 20000.0
    track = straight_track(n_traces, trace_spacing, alt)
    f0 = 60e6
    k0 = 2.0 * np.pi * f0 / C
    eps_ice = 3.15
    k_ice = k0 * np.sqrt(eps_ice)
    nt_s = n_terms_for((2.0 * k0 * 0.2) ** 2)
    nt_b = n_terms_for((2.0 * k_ice * 0.5) ** 2)
    rough_s = (0.2, 20.0, speckle_phasors(len(surf), (seed, 0)), nt_s, "exponential")
    rough_b = (0.5, 20.0, speckle_phasors(len(bed), (seed, 1)), nt_b, "exponential")
    ant = _ant.cos_power_table(2.0)
    pattern = _ant.pattern_args(ant, track.u_at, track.u_ct)
    layers = [
        Layer("surface", surf, fresnel_normal(1.0, eps_ice), rough_s),
        Layer("bed", bed, -10.0 ** (-10.0 / 20.0), rough_b, crossed=[surf],
              eps_leg=[1.0, eps_ice], att_leg=[0.0, 10.0]),
    ]
    return Workload(
        name="cfg2_surface_bed", track=track, layers=layers, k0=k0,
        t0=2.0 * (alt - 100.0) / C, dt=1.0 / 48e6, n_samples=n_samples, c=C,
        pattern=pattern, taper_s=0.05, d_phi_area=True,
        waveform=Waveform(kind="chirp", bandwidth=15e6, pulse_length=20e-6),
        meta=dict(f0=f0, facets=f"{n}x{n}@{dx}m x2", seed=seed,
                  n_terms=(nt_s, nt_b)))


def _pad_bucket(j, n_max):
    """Power-of-two joint padding bucket of a stack of j interfaces, at
    most n_max = the scene's interface count - 1 (``simulate.py:253-260``)."""
    b = 1
    while b < j:
        b *= 2
    return min(b, n_max)


def config3(nx=2048, ny=1024, n_traces=1000, seed=3, n_samples=8192, n_firn=20,
            firn_depth=100.0, trace_spacing=30.0, horizons=(500.0, 1000.0, 1500.0),
            flat_stack=True):
    """Multilayer scene: 2048 x 1024 facets at 30 m; firn rho(z) = 917 -
    567 exp(-z/30 m) in 20 horizontal sublayers with n = 1 + 0.845 rho; three
    horizontal horizons at 500/1000/1500 m (-60 dB); rough bed (sigma
    400 m, l 3 km); 1000 traces at 20 km, cos^2 pattern, 150 MHz centre,
    30 MHz chirp with a Tukey(0.2) pedestal, 8192 samples.  Layers: surface,
    the n_firn sublayer bases, the horizons, the bed (all targets); a layer
    j >= 2 takes the joint solve over its crossed stack, in its flat-stack
    form when ``flat_stack`` (every crossed interface below the surface is
    horizontal here), else the reference's general form."""
    dx = 30.0
    # seeded lattices: the device draws and tessellates them (ResidentScene);
    # the host builds facets only where an oracle reads them
    surf_lat = SeededLattice(ny, nx, dx, sigma=1.0, corr_len=300.0, seed=(seed, 0))
    bed_lat = SeededLattice(ny, nx, dx, sigma=400.0, corr_len=3000.0, seed=(seed, 1),
                            mean=-3000.0)
    surf, bed = LatticeFacets(surf_lat), LatticeFacets(bed_lat)
    planes = {}

    def plane(z):
        lat = SeededLattice(ny, nx, dx, mean=z)
        f = LatticeFacets(lat)
        planes[id(f)] = lat
        return f

    tops = np.arange(n_firn) * (firn_depth / n_firn)
    rho = 917.0 - 567.0 * np.exp(-tops / 30.0)
    eps = [1.0] + list(eps_kovacs(rho)) + [float(eps_kovacs(917.0))]
    zs = [-(k + 1) * firn_depth / n_firn for k in range(n_firn)] + [-h for h in horizons]
    alt = 20000.0
    track = straight_track(n_traces, trace_spacing, alt)
    f0 = 150e6
    k0 = 2.0 * np.pi * f0 / C
    ant = _ant.cos_power_table(2.0)
    pattern = _ant.pattern_args(ant, track.u_at, track.u_ct)
    layers = [Layer("surface", surf, fresnel_normal(1.0, eps[1]), lattice=surf_lat)]
    crossed = [surf]
    eps_leg, att_leg = [1.0, eps[1]], [0.0, 10.0]
    n_media = len(eps) - 1                     # media below the air
    n_max = len(zs) + 1                        # interfaces - 1
    for i, z in enumerate(zs):
        e_above = eps[min(i + 1, n_media)]
        e_below = eps[min(i + 2, n_media)]
        if i < n_firn:
            name, gam = f"firn{i + 1}", fresnel_normal(e_above, e_below)
        else:
            name, gam = f"horizon{int(horizons[i - n_firn])}", -10.0 ** (-60.0 / 20.0)
        j = len(layers)
        f = plane(z)
        layers.append(Layer(name, f, gam, crossed=list(crossed), eps_leg=list(eps_leg),
                            att_leg=list(att_leg), lattice=planes[id(f)],
                            refraction="joint" if j > 1 else "sequential",
                            pad_to=_pad_bucket(len(crossed), n_max) if j > 1 else None,
                            flat_stack=flat_stack and j > 1))
        crossed.append(layers[-1].facets)
        eps_leg.append(e_below)
        att_leg.append(10.0)
    layers.append(Layer("bed", bed, -10.0 ** (-10.0 / 20.0), crossed=list(crossed),
                        eps_leg=list(eps_leg), att_leg=list(att_leg), refraction="joint",
                        lattice=bed_lat,
                        pad_to=_pad_bucket(len(crossed), n_max), flat_stack=flat_stack))
    dt = 1.0 / 60e6
    return Workload(
        name="cfg3_multilayer", track=track, layers=layers, k0=k0,
        t0=2.0 * (alt - 100.0) / C, dt=dt, n_samples=n_samples, c=C, pattern=pattern,
        waveform=Waveform(kind="chirp", bandwidth=30e6, pulse_length=20e-6, window="tukey",
                          tukey_alpha=0.2),
        meta=dict(f0=f0, facets=f"{nx}x{ny}@{dx}m x{len(layers)}", seed=seed,
                  n_layers=len(layers)))
