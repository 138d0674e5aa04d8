"""Simulation output container (the reference's ``output.py:47-197`` Dataset
without xarray): coherent runs carry ``field`` (T, n[, 2][, L]) complex and
``power`` = |field|^2; incoherent runs carry ``power`` only (``field`` is
None); plus ``dropped_power``, per-layer ``nadir_twtt``, the fast-time axis
and provenance attributes; ``combine`` sums layers (coherent fields add
before squaring, incoherent powers add); ``save``/``load`` round-trip
through a NumPy ``.npz`` with a JSON attribute record (NetCDF needs
h5netcdf, which this library does not use).

The reference's ``first_return_*`` diagnostics (``output.py:77-82``, a
(T, F, 3) float64 host array there, ~50 GB at the 2.1 M-facet scale) are
computed on the device (``kernels.trace_diagnostics``):
``first_return_twtt`` and the facet centre ``first_return_xyz`` in the
local frame; for a projected (geodetic) scene also ``first_return_lat`` /
``first_return_lon`` (the facet centre through the frame's ENU -> WGS84
conversion, as the reference).
"""

import json
from dataclasses import dataclass, field

import numpy as np

VERSION = "0.1"


@dataclass
class SimResult:
    field: np.ndarray
    dropped_power: np.ndarray
    nadir_twtt: np.ndarray
    twtt: np.ndarray
    positions: np.ndarray
    layer: list = None
    attrs: dict = field(default_factory=dict)
    first_return_twtt: np.ndarray = None
    first_return_xyz: np.ndarray = None
    power_data: np.ndarray = None          # incoherent runs
    first_return_llh: np.ndarray = None    # projected scenes: (T, 3) lat, lon, h

    @property
    def power(self):
        if self.field is None:
            return self.power_data
        return np.abs(self.field) ** 2

    @property
    def first_return_lat(self):
        return None if self.first_return_llh is None else self.first_return_llh[:, 0]

    @property
    def first_return_lon(self):
        return None if self.first_return_llh is None else self.first_return_llh[:, 1]

    @classmethod
    def build(cls, *, field=None, power=None, dropped, nadir_twtt, t0, dt, positions,
              layers=None, config=None, first_return_twtt=None, first_return_xyz=None,
              first_return_llh=None):
        if (field is None) == (power is None):
            raise ValueError("give exactly one of field (coherent) and power (incoherent)")
        n = (field if field is not None else power).shape[1]
        attrs = {"soundersim_amd_version": VERSION,
                 "mode": "coherent" if field is not None else "incoherent"}
        if config is not None:
            attrs["config"] = config.model_dump_json()
        nad = nadir_twtt if layers is not None else nadir_twtt[:, 0]
        return cls(field, dropped, nad, t0 + dt * np.arange(n), np.asarray(positions),
                   list(layers) if layers is not None else None, attrs,
                   first_return_twtt, first_return_xyz,
                   None if power is None else _power_dtype(power),
                   first_return_llh)

    def sel_layer(self, name):
        j = self.layer.index(name)
        return (self.field if self.field is not None else self.power_data)[..., j]


def _power_dtype(power):
    """Incoherent power as float32 (the reference's dtype) unless it arrives
    as float64 (the FP64 tier), which is kept."""
    p = np.asarray(power)
    return p if p.dtype == np.float64 else p.astype(np.float32)


def combine(res, dim="layer"):
    """Total power over layers: |sum of fields|^2 (coherent), sum of powers
    (incoherent)."""
    if dim != "layer" or res.layer is None:
        raise ValueError("combine over 'layer' needs a multilayer result")
    if res.field is None:
        return res.power_data.sum(axis=-1)
    return np.abs(res.field.sum(axis=-1)) ** 2


_OPTIONAL = ("first_return_twtt", "first_return_xyz", "first_return_llh", "power_data")


def save(res, path):
    extra = {k: getattr(res, k) for k in _OPTIONAL if getattr(res, k) is not None}
    if res.field is not None:
        extra["field"] = res.field
    np.savez(path, dropped_power=res.dropped_power,
             nadir_twtt=res.nadir_twtt, twtt=res.twtt, positions=res.positions,
             meta=np.frombuffer(json.dumps({"layer": res.layer, "attrs": res.attrs}).encode(),
                                dtype=np.uint8), **extra)


def load(path):
    with np.load(path, allow_pickle=False) as z:
        meta = json.loads(z["meta"].tobytes().decode())
        opt = {k: (z[k] if k in z else None) for k in _OPTIONAL}
        return SimResult(z["field"] if "field" in z else None, z["dropped_power"],
                         z["nadir_twtt"], z["twtt"], z["positions"], meta["layer"],
                         meta["attrs"], **opt)
